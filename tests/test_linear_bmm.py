"""numc_linear and numc_bmm, the GEMM consumers of numc's roadmap
(ROADMAP.md:139-140: numc_linear(x, w, bias, out) = x @ W + bias with the bias
folded into the GEMM store, numc_bmm(a, b, out) for (B,H,S,D) @ (B,H,D,S)).

CPU: the validation contract (numc_matmul's order: NULL -> dtype -> ndim ->
inner dimension -> output shape).  GPU: closed forms, then random cases of
every dtype against the oracle (bit-exact for integers, componentwise bound
for floats), with folded and unfolded leading axes, views and broadcasting.
"""
import zlib

import numpy as np
import pytest

import numc_amd as nc
import oracle

U = 2.0 ** -24
U64 = 2.0 ** -53
DT = {np.float32: nc.FLOAT32, np.float64: nc.FLOAT64, np.int8: nc.INT8, np.uint8: nc.UINT8,
      np.int16: nc.INT16, np.int32: nc.INT32, np.uint32: nc.UINT32, np.int64: nc.INT64}


def _arr(ctx, shape, dt=nc.FLOAT32):
    return ctx.array(tuple(shape), dt, zeros=True)


# ---- oracle pins (CPU) ---------------------------------------------------------

def test_oracle_bmm_matches_numpy_matmul_on_integers():
    """numpy.matmul accumulates integers in their dtype (mod 2^w): an
    independent check of the broadcast indexing and of the wraparound"""
    rng = np.random.default_rng(1)
    for sa, sb in [((2, 3, 5, 7), (2, 3, 7, 4)), ((1, 3, 5, 7), (2, 1, 7, 4)), ((5, 7), (3, 7, 4)),
                   ((4, 5, 7), (7, 2))]:
        a = rng.integers(-128, 127, sa, dtype=np.int8, endpoint=True)
        b = rng.integers(-128, 127, sb, dtype=np.int8, endpoint=True)
        got, _ = oracle.bmm(a, b)
        assert np.array_equal(got, np.matmul(a, b))


def test_oracle_linear_closed_forms():
    rng = np.random.default_rng(2)
    w = rng.integers(-100, 100, (6, 5)).astype(np.int8)
    bias = rng.integers(-128, 127, 5).astype(np.int8)
    zero, _ = oracle.linear(np.zeros((3, 6), np.int8), w, bias)
    assert np.array_equal(zero, np.broadcast_to(bias, (3, 5)))
    eye, _ = oracle.linear(np.eye(6, dtype=np.int8), w, bias)
    assert np.array_equal(eye, (w.astype(np.int64) + bias.astype(np.int64)).astype(np.int8))  # mod 2^8
    wf = rng.standard_normal((6, 5))
    bf = rng.standard_normal(5)
    ref, _ = oracle.linear(np.eye(6), wf, bf)
    assert np.allclose(ref, wf + bf, rtol=0, atol=1e-15)


# ---- validation (CPU) ------------------------------------------------------------

@pytest.mark.parametrize("case", [
    # x, w, bias (None: no bias), out, dtypes (x, w, bias, out), code
    ((2, 3), (3, 4), (4,), (2, 4), (nc.FLOAT32, nc.FLOAT32, nc.FLOAT64, nc.FLOAT32), nc.ERR_TYPE),
    ((2, 3), (3, 4), None, (2, 4), (nc.FLOAT32, nc.INT32, None, nc.FLOAT32), nc.ERR_TYPE),
    ((2, 3), (3, 4, 1), (4,), (2, 4), None, nc.ERR_SHAPE),    # w not 2-D
    ((2, 3), (3, 4), (1, 4), (2, 4), None, nc.ERR_SHAPE),     # bias not 1-D
    ((2, 3), (3, 4), (4,), (2, 1, 4), None, nc.ERR_SHAPE),    # out ndim
    ((2, 3), (5, 4), (4,), (2, 4), None, nc.ERR_SHAPE),       # inner
    ((2, 3), (3, 4), (5,), (2, 4), None, nc.ERR_SHAPE),       # bias length
    ((2, 3), (3, 4), (4,), (3, 4), None, nc.ERR_SHAPE),       # out leading axis
    ((2, 3), (3, 4), (4,), (2, 5), None, nc.ERR_SHAPE),       # out last axis
])
def test_linear_validation(ctx, case):
    sx, sw, sbias, so, dts, code = case
    dts = dts or (nc.FLOAT32,) * 4
    x, w, o = _arr(ctx, sx, dts[0]), _arr(ctx, sw, dts[1]), _arr(ctx, so, dts[3])
    bias = _arr(ctx, sbias, dts[2]) if sbias is not None else None
    assert nc.linear(x, w, bias, o) == code
    assert "linear:" in nc.last_error()[1]


def test_linear_null(ctx):
    a = _arr(ctx, (2, 2))
    L = nc.lib()
    assert L.numc_linear(None, a.h, a.h, a.h) == nc.ERR_NULL
    assert L.numc_linear(a.h, None, a.h, a.h) == nc.ERR_NULL
    assert L.numc_linear(a.h, a.h, a.h, None) == nc.ERR_NULL


@pytest.mark.parametrize("case", [
    ((2, 3), (3, 4), (2, 4), (nc.FLOAT32, nc.INT32, nc.FLOAT32), nc.ERR_TYPE),
    ((3,), (3, 4), (4,), None, nc.ERR_SHAPE),                 # 1-D operand
    ((2, 2, 3), (2, 3, 4), (2, 2, 4, 1), None, nc.ERR_SHAPE),  # out ndim
    ((2, 2, 3), (2, 5, 4), (2, 2, 4), None, nc.ERR_SHAPE),     # inner
    ((3, 2, 3), (2, 3, 4), (3, 2, 4), None, nc.ERR_SHAPE),     # batch 3 vs 2
    ((2, 2, 3), (1, 3, 4), (3, 2, 4), None, nc.ERR_SHAPE),     # out batch axis
    ((2, 2, 3), (2, 3, 4), (2, 2, 5), None, nc.ERR_SHAPE),     # out matrix shape
])
def test_bmm_validation(ctx, case):
    sa, sb, so, dts, code = case
    dts = dts or (nc.FLOAT32,) * 3
    a, b, o = _arr(ctx, sa, dts[0]), _arr(ctx, sb, dts[1]), _arr(ctx, so, dts[2])
    assert nc.bmm(a, b, o) == code
    assert "bmm:" in nc.last_error()[1]


def test_bmm_null(ctx):
    a = _arr(ctx, (2, 2))
    assert nc.lib().numc_bmm(None, a.h, a.h) == nc.ERR_NULL
    assert nc.lib().numc_bmm(a.h, a.h, None) == nc.ERR_NULL


# ---- GPU --------------------------------------------------------------------------

def _gen(rng, dt, shape):
    if np.issubdtype(dt, np.integer):
        info = np.iinfo(dt)
        return rng.integers(info.min, info.max, shape, dtype=dt, endpoint=True)
    return rng.standard_normal(shape).astype(dt)


def _check(got, want, den, K, dt, bias_added):
    if np.issubdtype(dt, np.integer):
        assert np.array_equal(got, want)
        return
    u = U64 if dt == np.float64 else U
    err = np.abs(got.astype(np.float64) - want)
    # the GEMM's componentwise bound, plus one rounding of acc + bias
    bound = (K + 3) * u * den + (2 * u * np.abs(want) if bias_added else 0.0)
    assert np.all(err <= bound), np.max(err - bound)


@pytest.mark.gpu
def test_linear_closed_forms(ctx):
    """x = 0 gives the bias on every row (also for K = 0); x = I gives w + bias"""
    rng = np.random.default_rng(3)
    w = rng.integers(-128, 127, (40, 33), dtype=np.int8, endpoint=True)
    bias = rng.integers(-128, 127, 33, dtype=np.int8, endpoint=True)
    for x in (np.zeros((7, 40), np.int8), np.eye(40, dtype=np.int8)):
        O = ctx.array((x.shape[0], 33), nc.INT8, zeros=True)
        nc.check(nc.linear(ctx.from_numpy(x), ctx.from_numpy(w), ctx.from_numpy(bias), O), "linear")
        want = np.broadcast_to(bias, (7, 33)) if not x.any() else \
            (w.astype(np.int64) + bias.astype(np.int64)).astype(np.int8)
        assert np.array_equal(O.numpy(), want)
    O = ctx.from_numpy(np.full((5, 33), 9, np.float32))
    bf = rng.standard_normal(33).astype(np.float32)
    nc.check(nc.linear(ctx.array((5, 0), nc.FLOAT32), ctx.array((0, 33), nc.FLOAT32), ctx.from_numpy(bf), O), "k=0")
    assert np.array_equal(O.numpy(), np.broadcast_to(bf, (5, 33)))


LIN_SHAPES = [((300, 200), 150), ((2, 3, 70), 90), ((129,), 65), ((4, 257, 64), 300)]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", list(DT), ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("with_bias", [True, False], ids=["bias", "nobias"])
def test_linear_random(ctx, dt, with_bias):
    rng = np.random.default_rng(zlib.crc32(repr((np.dtype(dt).str, with_bias)).encode()))
    for sx, n in LIN_SHAPES:
        k = sx[-1]
        x, w = _gen(rng, dt, sx), _gen(rng, dt, (k, n))
        bias = _gen(rng, dt, (n,)) if with_bias else None
        O = ctx.from_numpy(np.full(sx[:-1] + (n,), 3, dt))
        nc.check(nc.linear(ctx.from_numpy(x), ctx.from_numpy(w),
                           ctx.from_numpy(bias) if with_bias else None, O), "linear")
        want, den = oracle.linear(x, w, bias)
        _check(O.numpy(), want, den, k, dt, with_bias)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.int32], ids=lambda d: np.dtype(d).name)
def test_linear_views(ctx, dt):
    """x with a leading axis that does not fold (stepped), a transposed w
    (a torch-style (N, K) weight), a stepped bias, an output with gaps"""
    rng = np.random.default_rng(4)
    base_x = _gen(rng, dt, (6, 50, 40))
    wt = _gen(rng, dt, (30, 40))                        # (N, K) storage
    base_bias = _gen(rng, dt, (60,))
    X = ctx.from_numpy(base_x).slice(0, 0, 6, 2)        # (3, 50, 40), axis 0 stepped
    W = ctx.from_numpy(wt).transpose()                  # (40, 30)
    B = ctx.from_numpy(base_bias).slice(0, 0, 60, 2)    # (30,), stride 2
    O_full = ctx.from_numpy(np.full((3, 50, 31), 5, dt))
    O = O_full.slice(2, 0, 30)
    nc.check(nc.linear(X, W, B, O), "linear views")
    want, den = oracle.linear(base_x[::2], np.ascontiguousarray(wt.T), base_bias[::2])
    full = O_full.numpy()
    _check(full[..., :30], want, den, 40, dt, True)
    assert np.all(full[..., 30] == 5)


BMM_SHAPES = [((2, 3, 33, 40), (2, 3, 40, 50)),       # (B, H, S, D) @ (B, H, D, S)
              ((1, 3, 17, 20), (2, 1, 20, 9)),        # both broadcast
              ((31, 64), (4, 64, 70)),                # 2-D a against a batch
              ((5, 45, 30), (30, 12)),                # batch against a 2-D b
              ((1, 64, 64), (1, 64, 64))]             # one slice


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int8, np.uint16, np.int32], ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("shapes", BMM_SHAPES, ids=[str(s) for s in BMM_SHAPES])
def test_bmm_random(ctx, dt, shapes):
    rng = np.random.default_rng(zlib.crc32(repr((np.dtype(dt).str, shapes)).encode()))
    sa, sb = shapes
    a, b = _gen(rng, dt, sa), _gen(rng, dt, sb)
    want, den = oracle.bmm(a, b)
    O = ctx.from_numpy(np.full(want.shape, 3, dt))
    nc.check(nc.bmm(ctx.from_numpy(a), ctx.from_numpy(b), O), "bmm")
    _check(O.numpy(), want, den, sa[-1], dt, False)


@pytest.mark.gpu
def test_bmm_transposed_operand_and_empty(ctx):
    """b given as a transposed view (K^T of attention scores), and an empty batch"""
    rng = np.random.default_rng(5)
    q = rng.standard_normal((2, 4, 48, 32)).astype(np.float32)
    k = rng.standard_normal((2, 4, 48, 32)).astype(np.float32)
    Kt = ctx.from_numpy(k).transpose((0, 1, 3, 2))      # (2, 4, 32, 48) view
    O = ctx.array((2, 4, 48, 48), nc.FLOAT32, zeros=True)
    nc.check(nc.bmm(ctx.from_numpy(q), Kt, O), "bmm kt")
    want, den = oracle.bmm(q, np.ascontiguousarray(np.swapaxes(k, -1, -2)))
    _check(O.numpy(), want, den, 32, np.float32, False)
    O = ctx.array((0, 3, 5), nc.FLOAT32)
    nc.check(nc.bmm(ctx.array((0, 3, 4), nc.FLOAT32), ctx.array((0, 4, 5), nc.FLOAT32), O), "empty")


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int32, np.uint8, np.int8], ids=lambda d: np.dtype(d).name)
def test_linear_every_config(ctx, dt):
    """the biased store of every kernel configuration of the dtype's family,
    split-K ones included (their reduction adds the bias); K = 512 with an
    aligned x is the case where float32 and int8 kernels read x in place"""
    fam = {np.float32: nc.FAMILY_F32, np.float64: nc.FAMILY_F64}.get(dt, nc.FAMILY_I8)
    rng = np.random.default_rng(zlib.crc32(np.dtype(dt).str.encode()))
    for cfg in range(nc.config_count(fam)):
        if nc.config_get(fam, cfg)["direct"] and dt in (np.int32,):
            continue   # one int8 digit plane only (test_int_configs_full_range covers the refusal)
        for M, K, N in [(130, 257, 70), (1, 700, 33), (100, 512, 90)]:
            x, w, bias = _gen(rng, dt, (M, K)), _gen(rng, dt, (K, N)), _gen(rng, dt, (N,))
            O = ctx.array((M, N), DT[dt], zeros=True)
            nc.check(nc.linear(ctx.from_numpy(x), ctx.from_numpy(w), ctx.from_numpy(bias), O,
                               nc.METHOD_PLAIN if not np.issubdtype(dt, np.integer) else nc.METHOD_AUTO, cfg),
                     "linear cfg %d" % cfg)
            want, den = oracle.linear(x, w, bias)
            _check(O.numpy(), want, den, K, dt, True)


def test_linear_method_rejected(ctx):
    x, w, o = _arr(ctx, (2, 3)), _arr(ctx, (3, 4)), _arr(ctx, (2, 4))
    assert nc.linear(x, w, None, o, nc.METHOD_SPLIT_BF16) == nc.ERR_VALUE
    assert "linear:" in nc.last_error()[1]


@pytest.mark.gpu
def test_bmm_zero_k_and_many_slices(ctx):
    """K = 0 gives zeros in every slice (persistent batched kernel); 300
    small int8 slices wrap exactly"""
    O = ctx.from_numpy(np.full((6, 5, 7), 9, np.float32))
    nc.check(nc.bmm(ctx.array((6, 5, 0), nc.FLOAT32), ctx.array((6, 0, 7), nc.FLOAT32), O), "bmm k=0")
    assert np.all(O.numpy() == 0)
    rng = np.random.default_rng(11)
    a = rng.integers(-128, 127, (300, 9, 40), dtype=np.int8, endpoint=True)
    b = rng.integers(-128, 127, (300, 40, 11), dtype=np.int8, endpoint=True)
    O = ctx.array((300, 9, 11), nc.INT8, zeros=True)
    nc.check(nc.bmm(ctx.from_numpy(a), ctx.from_numpy(b), O), "bmm int8")
    assert np.array_equal(O.numpy(), np.matmul(a, b))   # numpy accumulates int8 mod 2^8


@pytest.mark.gpu
def test_bmm_transposed_a_and_strided_out(ctx):
    """a given as a transposed view, the output a view with gaps"""
    rng = np.random.default_rng(12)
    at = rng.standard_normal((4, 30, 20)).astype(np.float32)        # (batch, K, M) storage
    b = rng.standard_normal((4, 30, 25)).astype(np.float32)
    A = ctx.from_numpy(at).transpose((0, 2, 1))                     # (4, 20, 30)
    O_full = ctx.from_numpy(np.full((4, 20, 27), 5, np.float32))
    O = O_full.slice(2, 1, 26)                                      # (4, 20, 25)
    nc.check(nc.bmm(A, ctx.from_numpy(b), O), "bmm views")
    full = O_full.numpy()
    want, den = oracle.bmm(np.ascontiguousarray(np.swapaxes(at, 1, 2)), b)
    _check(full[:, :, 1:26], want, den, 30, np.float32, False)
    assert np.all(full[:, :, 0] == 5) and np.all(full[:, :, 26] == 5)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.int8], ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("b_transposed", [False, True], ids=["b-rows", "b-transposed"])
def test_bmm_operands_in_place(ctx, dt, b_transposed):
    """strided-batch bmm whose slices are 16-byte aligned with K = 64: float32
    and int8 read a's slices in place, float32 also b's when b is given as a
    transposed view (q @ k^T); exact sums against numpy's integer matmul"""
    rng = np.random.default_rng(31 + b_transposed)
    if dt == np.int8:
        a = rng.integers(-128, 127, (3, 4, 100, 64), dtype=np.int8, endpoint=True)
        kk = rng.integers(-128, 127, (3, 4, 90, 64), dtype=np.int8, endpoint=True)
    else:
        a = rng.integers(-4, 4, (3, 4, 100, 64), endpoint=True).astype(np.float32)
        kk = rng.integers(-4, 4, (3, 4, 90, 64), endpoint=True).astype(np.float32)
    bt = np.swapaxes(kk, 2, 3)                                        # (3, 4, 64, 90)
    if b_transposed:
        B = ctx.from_numpy(kk).transpose((0, 1, 3, 2))
    else:
        B = ctx.from_numpy(np.ascontiguousarray(bt))
    exact = np.matmul(a.astype(np.int64), bt.astype(np.int64))
    want = exact.astype(np.int8) if dt == np.int8 else exact.astype(np.float32)
    O = ctx.array(want.shape, DT[dt], zeros=True)
    A = ctx.from_numpy(a)
    fam = nc.FAMILY_I8 if dt == np.int8 else nc.FAMILY_F32
    ctx.profile(True)
    ctx.profile_read(0, fam)
    nc.check(nc.bmm(A, B, O), "bmm in place")
    packs = ctx.profile_read(0, fam)[1]
    ctx.profile(False)
    assert np.array_equal(O.numpy(), want)
    # which path ran: a is never packed; b is packed unless float32 reads it in place too
    assert packs == (0 if (b_transposed and dt == np.float32) else 1), packs
