"""The benchmark workloads the other GPU tests do not reach, at their own sizes,
checked against the oracle through the C ABI, plus the autotuner itself.

  * The uneven batch: 4096 independent problems, M, N, K log-uniform integers
    in [16, 1024] (the shapes bench.py times), through numc_matmul_batched in
    float32, split-BF16 and int8 -> int32.  Integers: every problem's row and
    column checksums mod 2^32 against exact numpy int64 sums, and whole
    problems against the oracle on a sample.  Floats: sampled entries of every
    problem against the FP64 oracle under the method's componentwise bound.
    The method is numc's own packed-vs-naive sweep
    (tests/matmul/test_matmul.c:170-350) at the batch's scale.
  * FP32 8192^3 with the FP8 method on N(0,1) and ill-scaled inputs (rows of A
    and columns of B scaled by 2^u, u integer uniform in [-20, 20]), sampled
    entries against the oracle's sums of the quantised operands; plain FP32
    and split-BF16 on the ill-scaled set against the FP64 oracle.
  * The learned selector's regret on held-out shapes, re-timed on this GPU by
    the autotuner: mean <= 5 % and below the analytic model it replaces.
  * numc_autotune_shape (finite positive timings for every configuration of
    every family and layout) and numc_selector_tune (the pinned winner is
    what numc_matmul then runs, with results the oracle accepts).
"""
import ctypes as C
import math

import numpy as np
import pytest

import bench
import numc_amd as nc
import oracle

pytestmark = pytest.mark.gpu
U = 2.0 ** -24
U64 = 2.0 ** -53
BATCH_SEED = 4   # bench.py's batch seed (its --seed default 3, plus 1)


def _autotune_sigs():
    L = nc.lib()
    L.numc_autotune_shape.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.POINTER(C.c_float), C.c_int]
    L.numc_autotune_shape.restype = C.c_int
    L.numc_selector_choice.argtypes = [C.c_int] * 5
    L.numc_selector_choice.restype = C.c_int
    L.numc_selector_heuristic.argtypes = [C.c_int] * 4
    L.numc_selector_heuristic.restype = C.c_int
    return L


# ---- the uneven batch -------------------------------------------------------------

def _batch(ctx, dims, dt, odt):
    As, Bs, Os = [], [], []
    for i, (M, N, K) in enumerate(dims):
        A, B = ctx.array((M, K), dt), ctx.array((K, N), dt)
        if dt == nc.INT8:
            A.fill_random(nc.DIST_INT, BATCH_SEED, 2 * i, -128, 127)
            B.fill_random(nc.DIST_INT, BATCH_SEED, 2 * i + 1, -128, 127)
        else:
            A.fill_random(nc.DIST_UNIFORM, BATCH_SEED, 2 * i, -1, 1)
            B.fill_random(nc.DIST_UNIFORM, BATCH_SEED, 2 * i + 1, -1, 1)
        As.append(A)
        Bs.append(B)
        Os.append(ctx.array((M, N), odt, zeros=True))
    return As, Bs, Os


@pytest.mark.parametrize("method,bound_extra", [(nc.METHOD_PLAIN, 0.0), (nc.METHOD_SPLIT_BF16, 3 * 2.0 ** -16)],
                         ids=["f32", "split_bf16"])
def test_uneven_batch_4096_float(method, bound_extra):
    dims = bench.batch_shapes(4096, BATCH_SEED)
    rng = np.random.default_rng(40)
    with nc.Ctx() as ctx:
        As, Bs, Os = _batch(ctx, dims, nc.FLOAT32, nc.FLOAT32)
        print("uneven batch 4096 float method %d" % method, flush=True)
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, method), "numc_matmul_batched")
        worst = 0.0
        for i, (M, N, K) in enumerate(dims):
            a, b, c = As[i].numpy(), Bs[i].numpy(), Os[i].numpy()
            rows = rng.integers(0, M, 16)
            cols = rng.integers(0, N, 16)
            rows[0], cols[0] = M - 1, N - 1          # the last tile's edge
            ref, absr = oracle.fref(a, b, rows, cols)
            err = float(np.max(np.abs(c[rows, cols].astype(np.float64) - ref) / absr))
            assert err <= bound_extra + (K + 2) * U, (i, M, N, K, err)
            worst = max(worst, err)
        print("worst componentwise error %.3e" % worst, flush=True)


def test_uneven_batch_4096_int8_to_int32():
    dims = bench.batch_shapes(4096, BATCH_SEED)
    sample = set(np.random.default_rng(41).choice(4096, 64, replace=False).tolist())
    with nc.Ctx() as ctx:
        As, Bs, Os = _batch(ctx, dims, nc.INT8, nc.INT32)
        print("uneven batch 4096 int8 -> int32", flush=True)
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, nc.METHOD_AUTO, -1, nc.OPT_WIDEN), "numc_matmul_batched")
        for i in range(len(dims)):
            a, b, c = As[i].numpy(), Bs[i].numpy(), Os[i].numpy().astype(np.int64)
            a64, b64 = a.astype(np.int64), b.astype(np.int64)
            m = (1 << 32) - 1
            assert np.array_equal(c.sum(axis=0) & m, (a64.sum(axis=0) @ b64) & m), i
            assert np.array_equal(c.sum(axis=1) & m, (a64 @ b64.sum(axis=1)) & m), i
            if i in sample:
                assert np.array_equal(Os[i].numpy(), oracle.matmul_int(a, b, np.int32)), i


# ---- FP32 8192^3: FP8 and the ill-scaled set --------------------------------------------

N3 = 8192


def _config3_inputs(ctx, ill):
    A, B = ctx.array((N3, N3), nc.FLOAT32), ctx.array((N3, N3), nc.FLOAT32)
    A.fill_random(nc.DIST_NORMAL, 3, 0)
    B.fill_random(nc.DIST_NORMAL, 3, 1)
    a, b = A.numpy(), B.numpy()
    if ill:
        # bench.py's ill-scaled set: exact power-of-two row / column scales
        rng = np.random.default_rng(3)
        u = rng.integers(-20, 20, N3, endpoint=True)
        v = rng.integers(-20, 20, N3, endpoint=True)
        a = np.ldexp(a, u[:, None]).astype(np.float32)
        b = np.ldexp(b, v[None, :]).astype(np.float32)
        A, B = ctx.from_numpy(a), ctx.from_numpy(b)
    A.to_device()
    B.to_device()
    return A, B, a, b


def _sample(seed, nrows=128, ncols=128, n=2048):
    rng = np.random.default_rng(seed)
    rows = rng.choice(N3, nrows, replace=False)
    cols = rng.choice(N3, ncols, replace=False)
    return rows, cols, rng.integers(0, nrows, n), rng.integers(0, ncols, n)


@pytest.mark.parametrize("ill", [False, True], ids=["normal", "ill_scaled"])
def test_config3_fp8_8192(ill):
    with nc.Ctx() as ctx:
        A, B, a, b = _config3_inputs(ctx, ill)
        O = ctx.array((N3, N3), nc.FLOAT32, zeros=True)
        O.to_device()
        print("fp8 8192^3 ill=%s" % ill, flush=True)
        nc.check(nc.matmul(A, B, O, nc.METHOD_FP8_SCALED), "fp8 matmul")
        c = O.numpy()
    rows, cols, ri, ci = _sample(50 + ill)
    # whole rows of A and columns of B: each 32-k block carries its own scale
    qa = np.ascontiguousarray(oracle.fp8_quantize_blocks(a[rows]))
    qb = np.ascontiguousarray(oracle.fp8_quantize_blocks(b[:, cols].T).T)
    ref, absr = oracle.fref(qa, qb, ri, ci)
    got = c[rows[ri], cols[ci]].astype(np.float64)
    err = float(np.max(np.abs(got - ref) / absr))
    # the f8f6f4 MFMA's internal sum of 64 products (see test_fp8_configs)
    assert err <= -(-N3 // 64) * 2.0 ** -11 + (N3 + 2) * U, err
    # and end to end against the unquantised FP64 sums: e4m3 keeps 3 mantissa bits
    ref0, absr0 = oracle.fref(np.ascontiguousarray(a[rows]), np.ascontiguousarray(b[:, cols]), ri, ci)
    err0 = float(np.max(np.abs(got - ref0) / absr0))
    assert err0 <= 2 * 2.0 ** -4 + 2.0 ** -8, err0
    print("fp8 componentwise error %.3e vs quantised operands, %.3e vs FP64" % (err, err0), flush=True)


@pytest.mark.parametrize("method,bound", [(nc.METHOD_PLAIN, (N3 + 2) * U),
                                          (nc.METHOD_SPLIT_BF16, 3 * 2.0 ** -16 + (N3 + 2) * U)],
                         ids=["f32", "split_bf16"])
def test_config3_ill_scaled_8192(method, bound):
    with nc.Ctx() as ctx:
        A, B, a, b = _config3_inputs(ctx, True)
        O = ctx.array((N3, N3), nc.FLOAT32, zeros=True)
        O.to_device()
        print("method %d 8192^3 ill-scaled" % method, flush=True)
        nc.check(nc.matmul(A, B, O, method), "matmul")
        c = O.numpy()
    rows, cols, ri, ci = _sample(60 + method)
    ref, absr = oracle.fref(np.ascontiguousarray(a[rows]), np.ascontiguousarray(b[:, cols]), ri, ci)
    err = float(np.max(np.abs(c[rows[ri], cols[ci]].astype(np.float64) - ref) / absr))
    assert err <= bound, err


# ---- the learned selector's regret, re-timed on this GPU -------------------------------

def _heldout_subset(n, max_flop=2e11):
    from numc_amd import autotune
    shapes = autotune.sample_shapes()
    out = []
    for idx in range(autotune.N_TRAIN, autotune.N_TRAIN + autotune.N_HOLDOUT):
        fam, M, N, K, fl = shapes[idx]
        if 2.0 * M * N * K <= max_flop:
            out.append((idx, fam, M, N, K, fl))
        if len(out) == n:
            break
    return out


def test_selector_regret_heldout_retimed():
    L = _autotune_sigs()
    assert L.numc_selector_active() == 1, "the trained selector does not match the compiled configurations"
    buf = (C.c_float * 32)()
    cand = (C.c_int * 32)()
    L.numc_selector_candidates.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_int), C.c_int]
    reg_mlp, reg_ana, reg_call = [], [], []
    with nc.Ctx() as ctx:
        for idx, fam, M, N, K, fl in _heldout_subset(160):
            ncfg = nc.config_count(fam)
            reps = 5 if 2.0 * M * N * K < 1e9 else 2
            best = [math.inf] * ncfg
            for _ in range(2):   # min of two timings per configuration
                nc.check(L.numc_autotune_shape(ctx.h, fam, M, N, K, fl, reps, buf, 32), "numc_autotune_shape")
                best = [min(best[i], buf[i]) for i in range(ncfg)]
            t_best = min(best)
            pick = L.numc_selector_choice(fam, M, N, K, fl)
            ana = L.numc_selector_heuristic(fam, M, N, K)
            assert 0 <= pick < ncfg and 0 <= ana < ncfg
            # numc_matmul's rule: time the close candidates once, keep the fastest
            nc_ = L.numc_selector_candidates(fam, M, N, K, fl, 1, cand, 32)
            assert nc_ >= 1 and cand[0] == pick
            call = min((cand[i] for i in range(nc_)), key=lambda c: best[c])
            reg_mlp.append(best[pick] / t_best - 1)
            reg_ana.append(best[ana] / t_best - 1)
            reg_call.append(best[call] / t_best - 1)
        nc.check(nc.lib().numc_autotune_release(ctx.h), "numc_autotune_release")
    m, a, c = float(np.mean(reg_mlp)), float(np.mean(reg_ana)), float(np.mean(reg_call))
    print("held-out regret (%d shapes): learned mean %.4f p95 %.4f; with call timing mean %.4f p95 %.4f max %.3f; "
          "analytic mean %.4f p95 %.4f" % (len(reg_mlp), m, np.percentile(reg_mlp, 95), c,
                                          np.percentile(reg_call, 95), max(reg_call), a, np.percentile(reg_ana, 95)),
          flush=True)
    assert m <= 0.05
    assert m < a
    assert c <= m and np.percentile(reg_call, 95) <= 0.05


# ---- the autotuner -------------------------------------------------------------------------

FAM_CASE = {  # family -> (input dtype, method)
    nc.FAMILY_F32: (nc.FLOAT32, nc.METHOD_PLAIN), nc.FAMILY_SPLIT_BF16: (nc.FLOAT32, nc.METHOD_SPLIT_BF16),
    nc.FAMILY_FP8: (nc.FLOAT32, nc.METHOD_FP8_SCALED), nc.FAMILY_F64: (nc.FLOAT64, nc.METHOD_AUTO),
    nc.FAMILY_I8: (nc.INT8, nc.METHOD_AUTO), nc.FAMILY_BF16: (nc.BFLOAT16, nc.METHOD_AUTO),
    nc.FAMILY_F16: (nc.FLOAT16, nc.METHOD_AUTO)}


@pytest.mark.parametrize("fam", range(nc.FAMILY_COUNT), ids=nc.FAMILY_NAMES)
def test_autotune_shape_times_every_config_and_layout(ctx, fam):
    L = _autotune_sigs()
    buf = (C.c_float * 32)()
    ncfg = nc.config_count(fam)
    for flags in range(8):
        for i in range(32):
            buf[i] = -1.0
        nc.check(L.numc_autotune_shape(ctx.h, fam, 200, 300, 257, flags, 2, buf, 32), "numc_autotune_shape")
        t = np.array([buf[i] for i in range(ncfg)])
        assert np.all(np.isfinite(t)) and np.all(t > 0), (fam, flags, t)
        assert all(buf[i] == -1.0 for i in range(ncfg, 32)), "wrote past the family's configurations"
    # too small an output buffer is an error, not an overrun
    assert L.numc_autotune_shape(ctx.h, fam, 64, 64, 64, 0, 1, buf, ncfg - 1) == nc.ERR_BOUNDS
    nc.check(nc.lib().numc_autotune_release(ctx.h), "numc_autotune_release")


def _operand(ctx, x, transposed, padded, dt):
    """an array view holding x (rows x cols) stored transposed and/or with 16
    elements of padding per stored row"""
    store = np.ascontiguousarray(x.T if transposed else x)
    if padded:
        store = np.concatenate([store, np.zeros((store.shape[0], 16), store.dtype)], axis=1)
    S = ctx.from_numpy(store, dt if dt in (nc.BFLOAT16,) else None)
    if padded:
        S = S.slice(1, 0, store.shape[1] - 16)
    if transposed:
        S.transpose()
    return S


def _check_family(fam, a, b, c):
    K = a.shape[1]
    if fam == nc.FAMILY_I8:
        assert np.array_equal(c, oracle.matmul_int(a, b))
        return
    if fam == nc.FAMILY_FP8:
        qa, qb = oracle.fp8_quantized_inputs(a, b)
        ref, absr = oracle.fref(qa, qb)
        assert oracle.err_metrics(c, ref, absr)[0] <= -(-K // 64) * 2.0 ** -11 + (K + 2) * U
        return
    if fam in (nc.FAMILY_BF16, nc.FAMILY_F16):
        av = oracle.bf16_to_f32(a).astype(np.float64).reshape(a.shape) if fam == nc.FAMILY_BF16 else a.astype(np.float64)
        bv = oracle.bf16_to_f32(b).astype(np.float64).reshape(b.shape) if fam == nc.FAMILY_BF16 else b.astype(np.float64)
        cv = oracle.bf16_to_f32(c).astype(np.float64).reshape(c.shape) if fam == nc.FAMILY_BF16 else c.astype(np.float64)
        ref, absr = av @ bv, np.abs(av) @ np.abs(bv)
        u16 = 2.0 ** -8 if fam == nc.FAMILY_BF16 else 2.0 ** -11
        assert np.all(np.abs(cv - ref) <= (K + 3) * U * absr + u16 * np.abs(ref) + 2.0 ** -25)
        return
    ref, absr = oracle.fref(a, b)
    err = oracle.err_metrics(c, ref, absr)[0]
    bound = {nc.FAMILY_F32: (K + 2) * U, nc.FAMILY_SPLIT_BF16: 3 * 2.0 ** -16 + (K + 2) * U,
             nc.FAMILY_F64: (K + 2) * U64}[fam]
    assert err <= bound, (fam, err)


@pytest.mark.parametrize("fam", range(nc.FAMILY_COUNT), ids=nc.FAMILY_NAMES)
def test_selector_tune_pins_the_winner_and_keeps_results(ctx, fam):
    L = _autotune_sigs()
    rng = np.random.default_rng(70 + fam)
    dt, method = FAM_CASE[fam]
    M, N, K = 190, 330, 515
    for flags in (0, 3, 4):
        best = L.numc_selector_tune(ctx.h, fam, M, N, K, flags, 2)
        assert 0 <= best < nc.config_count(fam), best
        assert L.numc_selector_choice(fam, M, N, K, flags) == best
        if dt == nc.INT8:
            a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
            b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
        elif dt == nc.BFLOAT16:
            a = oracle.split_bf16(rng.standard_normal((M, K)).astype(np.float32))[0].reshape(M, K)
            b = oracle.split_bf16(rng.standard_normal((K, N)).astype(np.float32))[0].reshape(K, N)
        else:
            npdt = {nc.FLOAT32: np.float32, nc.FLOAT64: np.float64, nc.FLOAT16: np.float16}[dt]
            a, b = rng.standard_normal((M, K)).astype(npdt), rng.standard_normal((K, N)).astype(npdt)
        A = _operand(ctx, a, flags & 1, flags & 4, dt)
        B = _operand(ctx, b, flags & 2, flags & 4, dt)
        O = ctx.array((M, N), dt, zeros=True)
        nc.check(nc.matmul(A, B, O, method), "matmul after tuning")
        _check_family(fam, a, b, O.numpy())
        nc.check(L.numc_selector_pin(fam, M, N, K, flags, -1), "unpin")
    nc.check(nc.lib().numc_autotune_release(ctx.h), "numc_autotune_release")


# ---- the direct (one-launch, in-place) configurations -------------------------------------

def _misaligned(ctx, x, dt, transposed):
    """a view holding x whose rows start one element past a 16-byte boundary
    and whose row stride is not a multiple of 16 bytes (stored transposed
    when asked: unit row stride, k stride of rows + 1 elements)"""
    store = np.ascontiguousarray(x.T if transposed else x)
    pad = np.zeros((store.shape[0], store.shape[1] + 1), store.dtype)
    pad[:, 1:] = store
    S = ctx.from_numpy(pad, dt if dt == nc.BFLOAT16 else None).slice(1, 1, store.shape[1] + 1)
    if transposed:
        S.transpose()
    return S


@pytest.mark.parametrize("fam", [nc.FAMILY_F32, nc.FAMILY_SPLIT_BF16, nc.FAMILY_F64, nc.FAMILY_I8, nc.FAMILY_BF16,
                                 nc.FAMILY_F16], ids=["f32", "split_bf16", "f64", "i8", "bf16", "f16"])
def test_direct_configs_misaligned_and_ragged(ctx, fam):
    """every direct configuration on operand views at element (not 16-byte)
    alignment, row-major and stored transposed, with K around the MFMA
    fragment length (tails read nothing past K) and M, N off the tile"""
    dt, method = FAM_CASE[fam]
    rng = np.random.default_rng(90 + fam)
    cfgs = [c for c in range(nc.config_count(fam)) if nc.config_get(fam, c)["direct"]]
    assert cfgs
    kf = nc.config_get(fam, cfgs[0])["kc"]
    for cfg in cfgs:
        for M, K, N in [(1, 1, 1), (33, kf - 1, 70), (70, kf + 1, 33), (65, 3 * kf + 5, 97), (5, 1000, 2)]:
            if dt == nc.INT8:
                a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
                b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
            elif dt == nc.BFLOAT16:
                a = oracle.split_bf16(rng.standard_normal((M, K)).astype(np.float32))[0].reshape(M, K)
                b = oracle.split_bf16(rng.standard_normal((K, N)).astype(np.float32))[0].reshape(K, N)
            else:
                npdt = {nc.FLOAT32: np.float32, nc.FLOAT64: np.float64, nc.FLOAT16: np.float16}[dt]
                a, b = rng.standard_normal((M, K)).astype(npdt), rng.standard_normal((K, N)).astype(npdt)
            for ta, tb in [(False, False), (True, True), (False, True)]:
                A, B = _misaligned(ctx, a, dt, ta), _misaligned(ctx, b, dt, tb)
                O = ctx.array((M, N), dt, zeros=True)
                nc.check(nc.matmul(A, B, O, method, cfg), "direct matmul")
                _check_family(fam, a, b, O.numpy())


def test_direct_int8_widen_and_uint8(ctx):
    """int8 -> int32 (the exact sum) and uint8 -> uint8 (mod 256) through the
    direct configurations; int16 inputs fall back to packed digit planes under
    the selector"""
    rng = np.random.default_rng(95)
    cfgs = [c for c in range(nc.config_count(nc.FAMILY_I8)) if nc.config_get(nc.FAMILY_I8, c)["direct"]]
    for cfg in cfgs:
        for M, K, N in [(40, 300, 70), (3, 5, 7), (129, 64, 31)]:
            a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
            b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
            O = ctx.array((M, N), nc.INT32, zeros=True)
            nc.check(nc.matmul(ctx.from_numpy(a), ctx.from_numpy(b), O, nc.METHOD_AUTO, cfg, nc.OPT_WIDEN), "widen")
            assert np.array_equal(O.numpy(), oracle.matmul_int(a, b, np.int32))
            ua, ub = a.view(np.uint8), b.view(np.uint8)
            O = ctx.array((M, N), nc.UINT8, zeros=True)
            nc.check(nc.matmul(ctx.from_numpy(ua), ctx.from_numpy(ub), O, nc.METHOD_AUTO, cfg), "uint8")
            assert np.array_equal(O.numpy(), oracle.matmul_int(ua, ub))
    a = rng.integers(-30000, 30000, (20, 40), dtype=np.int16)
    b = rng.integers(-30000, 30000, (40, 10), dtype=np.int16)
    O = ctx.array((20, 10), nc.INT16, zeros=True)
    nc.check(nc.matmul(ctx.from_numpy(a), ctx.from_numpy(b), O), "int16 under the selector")
    assert np.array_equal(O.numpy(), oracle.matmul_int(a, b))
