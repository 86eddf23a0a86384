"""The selector's cost model: oracle pinning (CPU), the HIP forward / backward /
training kernels against the double-precision oracle (GPU), and what the
trained selector gains over the analytic model on held-out shapes (CPU,
from the committed timing database and model)."""
import ctypes as C
import os

import numpy as np
import pytest

import numc_amd as nc
from numc_amd import autotune as at
from oracle import mlp_oracle as mo

DIN, H, DOUT = 18, 64, 1


def rand_problem(rng, B, dout=DOUT, din=DIN, hidden=H):
    n = din * hidden + hidden + hidden * hidden + hidden + hidden * dout + dout
    p = (rng.standard_normal(n) * 0.3).astype(np.float32)
    x = rng.standard_normal((B, din)).astype(np.float32)
    t = rng.standard_normal((B, dout)).astype(np.float32)
    m = (rng.random((B, dout)) < 0.7).astype(np.float32)
    return p, x, t, m


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    din, hidden, dout = 5, 7, 3
    p, x, t, m = rand_problem(rng, 6, dout, din, hidden)
    p = p.astype(np.float64)
    loss, g = mo.loss_grad(p, x, t, m, din, hidden, dout)
    eps = 1e-6
    for i in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = eps
        lp, _ = mo.loss_grad(p + e, x, t, m, din, hidden, dout)
        lm, _ = mo.loss_grad(p - e, x, t, m, din, hidden, dout)
        assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-6 * max(1.0, abs(g[i]))


@pytest.mark.gpu
@pytest.mark.parametrize("masked", [False, True])
def test_gpu_gradients_match_oracle(ctx, masked):
    L = at._setup()
    rng = np.random.default_rng(1)
    B = 300
    p, x, t, m = rand_problem(rng, B)
    net = L.numc_mlp_create(ctx.h, DIN, H, DOUT, 3)
    assert net
    nc.check(L.numc_mlp_set_params(net, at.fptr(p), p.size), "set_params")
    g = np.zeros(p.size, np.float32)
    loss = C.c_double()
    nc.check(L.numc_mlp_loss_grad(net, at.fptr(x), at.fptr(t), at.fptr(m) if masked else None, B, C.byref(loss),
                                  at.fptr(g)), "loss_grad")
    want_loss, want_g = mo.loss_grad(p, x, t, m if masked else None, DIN, H, DOUT)
    assert loss.value == pytest.approx(want_loss, rel=1e-5)
    assert np.max(np.abs(g - want_g)) <= 1e-4 * np.max(np.abs(want_g))
    y = np.zeros((B, DOUT), np.float32)
    nc.check(L.numc_mlp_forward(net, at.fptr(x), B, at.fptr(y)), "forward")
    want_y, _ = mo.forward(p, x, DIN, H, DOUT)
    assert np.max(np.abs(y - want_y)) <= 1e-5 * max(1.0, np.max(np.abs(want_y)))
    L.numc_mlp_free(net)


@pytest.mark.gpu
def test_gpu_training_fits_and_is_deterministic(ctx):
    L = at._setup()
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (4000, DIN)).astype(np.float32)
    t = (np.sin(2 * x[:, :1]) + x[:, 1:2] * x[:, 2:3]).astype(np.float32)
    params = []
    for _ in range(2):
        net = L.numc_mlp_create(ctx.h, DIN, H, DOUT, 11)
        losses = np.zeros(30)
        nc.check(L.numc_mlp_train(net, at.fptr(x), at.fptr(t), None, len(x), 30, 256, 3e-3, 1e-4, 5,
                                  losses.ctypes.data_as(C.POINTER(C.c_double))), "train")
        n = L.numc_mlp_param_count(net)
        p = np.zeros(n, np.float32)
        nc.check(L.numc_mlp_get_params(net, at.fptr(p), n), "get_params")
        params.append(p)
        L.numc_mlp_free(net)
        assert losses[-1] < 0.1 * losses[0]
    assert np.array_equal(params[0], params[1]), "same seed must give the same weights"


def test_host_selector_matches_oracle_forward():
    """numc_selector_predict (host C++) runs the saved weights like the oracle"""
    if not os.path.exists(at.DEFAULT_MODEL):
        pytest.skip("no trained selector committed")
    L = at._setup()
    nc.check(L.numc_selector_load(at.DEFAULT_MODEL.encode()), "load")
    params, din, hidden, dout = at.load_model_params(at.DEFAULT_MODEL)
    buf = (C.c_float * 16)()
    for fam, M, N, K, fl in at.sample_shapes(5, 50):
        nc.check(L.numc_selector_predict(fam, M, N, K, fl, buf, 16), "predict")
        X = np.array([at.features(fam, M, N, K, fl, c) for c in range(nc.config_count(fam))])
        want, _ = mo.forward(params, X, din, hidden, dout)
        got = np.array([buf[c] for c in range(nc.config_count(fam))])
        assert np.max(np.abs(got - want[:, 0])) <= 1e-4


def test_selector_regret_on_held_out_shapes():
    """The learned selector is within 5% of exhaustive autotuning on the
    held-out shapes and beats the analytic model it replaces."""
    if not (os.path.exists(at.DEFAULT_MODEL) and os.path.exists(at.DEFAULT_DB)):
        pytest.skip("no timing database / trained selector committed")
    r = at.evaluate(at.DEFAULT_DB, at.DEFAULT_MODEL)
    assert r["shapes"] >= 1000
    assert r["mlp"]["mean_regret"] <= 0.05
    assert r["mlp"]["total_time_ratio"] <= 1.05
    assert r["mlp"]["mean_regret"] < r["analytic"]["mean_regret"]
    # numc_matmul's rule: time the model's close candidates (all of them for
    # calls under 2^-5 ms) once per shape -- p95 within 5 %, worst within 1.5x
    fb = r["mlp_with_call_timing"]
    assert fb["mean_regret"] <= r["mlp"]["mean_regret"]
    assert fb["p95_regret"] <= 0.05 and fb["max_regret"] <= 0.5
    assert 0 < fb["shapes_timed"] < r["shapes"]
