"""numc_dot through the GPU GEMM (numc's src/reduction/dot.c:340-448).

CPU: the validation contract of numc's _check_dot (src/reduction/dispatch.h:
199-283).  GPU: numc's own dot cases (tests/reduction/test_dot_nd.c,
test_dot.c, committed as data in tests/golden/numc_known_answers.json), then
random cases of every shape class against the oracle: each N-D dot is cut
into its 2-D GEMMs here and every GEMM is checked with oracle.matmul_int
(bit-exact) or oracle.fref (componentwise bound).
"""
import json
import os
import zlib

import numpy as np
import pytest

import numc_amd as nc
import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "numc_known_answers.json")))
U = 2.0 ** -24
U64 = 2.0 ** -53


def _arr(ctx, shape, dtype=nc.FLOAT32):
    return ctx.array(tuple(shape), dtype, zeros=True)


# ---- validation (CPU) --------------------------------------------------------

@pytest.mark.parametrize("case", [
    # (a shape, b shape, out shape, dtypes, expected code)
    ((2, 3), (3, 2), (2, 2), (nc.FLOAT32, nc.FLOAT64, nc.FLOAT32), nc.ERR_TYPE),
    ((2, 3), (4, 2), (2, 2), None, nc.ERR_SHAPE),          # axis mismatch
    ((2, 3), (3,), (2, 1), None, nc.ERR_SHAPE),            # out ndim
    ((2, 3), (3, 5), (2, 4), None, nc.ERR_SHAPE),          # out shape
    ((2, 3), (4, 3, 5), (2, 5, 4), None, nc.ERR_SHAPE),    # out shape (batch axis order)
    ((), (4,), (4, 1), None, nc.ERR_SHAPE),                # 0-D: out ndim
    ((), (4, 2), (4, 3), None, nc.ERR_SHAPE),              # 0-D: out shape
    ((3,), (3,), (2,), None, nc.ERR_SHAPE),                # scalar out must have one element
])
def test_dot_validation(ctx, case):
    sa, sb, so, dts, code = case
    dts = dts or (nc.FLOAT32,) * 3
    a, b, o = _arr(ctx, sa, dts[0]), _arr(ctx, sb, dts[1]), _arr(ctx, so, dts[2])
    assert nc.dot(a, b, o) == code
    assert "dot:" in nc.last_error()[1]


def test_dot_null(ctx):
    a = _arr(ctx, (2, 2))
    assert nc.lib().numc_dot(None, a.h, a.h) == nc.ERR_NULL
    assert nc.lib().numc_dot(a.h, None, a.h) == nc.ERR_NULL
    assert nc.lib().numc_dot(a.h, a.h, None) == nc.ERR_NULL


# ---- reference by decomposition into 2-D GEMMs ----------------------------------

def want_dot(a, b):
    """(expected, |a|.|b| or None): numpy-dot index rules, arithmetic from the oracle"""
    if a.ndim == 0 or b.ndim == 0:
        s, other = (a, b) if a.ndim == 0 else (b, a)
        if np.issubdtype(other.dtype, np.integer):
            w = other.dtype.itemsize * 8
            sv = int(s)
            if w == 64:   # numc passes the scalar through double (dot.c:340-346)
                d = float(s)
                signed = np.issubdtype(other.dtype, np.signedinteger)
                sv = (-(1 << 63) if d >= 2.0 ** 63 else int(d)) if signed else (0 if d >= 2.0 ** 64 else int(d))
            prod = (other.astype(object) * sv) % (1 << w)   # low w bits
            return np.array(prod.tolist(), dtype=np.uint64).astype(other.dtype.str.replace("i", "u")).view(
                other.dtype).reshape(other.shape), None
        return (other * other.dtype.type(s)).astype(other.dtype), None
    k = a.shape[-1]
    a2 = np.ascontiguousarray(a).reshape(-1, k)
    if b.ndim == 1:
        b3 = np.ascontiguousarray(b).reshape(1, k, 1)
        out_shape = a.shape[:-1]
    else:
        b3 = np.ascontiguousarray(b).reshape(-1, k, b.shape[-1])
        out_shape = a.shape[:-1] + b.shape[:-2] + b.shape[-1:]
    res, den = [], []
    for p in range(b3.shape[0]):
        bp = np.ascontiguousarray(b3[p])
        if np.issubdtype(a.dtype, np.integer):
            res.append(oracle.matmul_int(a2, bp))
        else:
            r, d = oracle.fref(a2, bp)
            res.append(r)
            den.append(d)
    stack = np.stack(res)                         # (P, m, n)
    full = np.transpose(stack, (1, 0, 2))         # (m, P, n)
    exp = full.reshape(out_shape) if out_shape else full.reshape(())
    if den:
        return exp, np.transpose(np.stack(den), (1, 0, 2)).reshape(exp.shape)
    return exp, None


def run_dot(ctx, a, b, out_shape=None, views=None):
    A = views[0] if views else ctx.from_numpy(a)
    B = views[1] if views else ctx.from_numpy(b)
    want, den = want_dot(a, b)
    shape = out_shape if out_shape is not None else want.shape
    O = ctx.from_numpy(np.full(shape, 7, a.dtype))
    nc.check(nc.dot(A, B, O), "numc_dot")
    got = O.numpy().reshape(want.shape)
    return got, want, den


def assert_dot(got, want, den, K, dtype):
    if np.issubdtype(dtype, np.integer) or den is None:
        assert np.array_equal(got, want)
    else:
        u = U64 if dtype == np.float64 else U
        err = np.max(np.abs(got.astype(np.float64) - want) / np.where(den > 0, den, 1.0)) if got.size else 0.0
        assert err <= (K + 2) * u, err


# ---- GPU ---------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("idx", range(len(GOLD["dot"])))
def test_dot_numc_known_answers(ctx, idx):
    c = GOLD["dot"][idx]

    def data(spec, shape):
        n = int(np.prod(shape)) if shape else 1
        if isinstance(spec, str):
            return np.full(n, float(spec.split()[1]), np.float32).reshape(shape)
        return np.array(spec, np.float32).reshape(shape)
    a, b = data(c["a"], c["a_shape"]), data(c["b"], c["b_shape"])
    A, B = ctx.from_numpy(a), ctx.from_numpy(b)
    O = ctx.array(tuple(c["out_shape"]), nc.FLOAT32, zeros=True)
    nc.check(nc.dot(A, B, O), c["case"])
    want = data(c["expect"], c["out_shape"])
    assert np.array_equal(O.numpy(), want), c["source"]


SHAPES = [
    ((), (5, 7)), ((3, 4), ()),                       # 0-D with N-D
    ((300,), (300,)), ((1,), (1,)),                   # 1-D . 1-D
    ((37, 50), (50,)), ((2, 3, 40), (40,)),           # N-D . 1-D
    ((70, 90), (90, 33)), ((50,), (50, 20)),          # 2-D . 2-D, 1-D . 2-D
    ((6, 30), (4, 30, 9)), ((2, 3, 40), (5, 40, 6)),  # N-D . M-D (batched)
    ((2, 2, 3, 17), (3, 17, 4)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int8, np.uint16, np.int32, np.int64])
@pytest.mark.parametrize("shapes", SHAPES, ids=[str(s) for s in SHAPES])
def test_dot_random(ctx, dt, shapes):
    rng = np.random.default_rng(zlib.crc32(repr((np.dtype(dt).str, shapes)).encode()))
    sa, sb = shapes

    def gen(shape):
        if np.issubdtype(dt, np.integer):
            info = np.iinfo(dt)
            return rng.integers(info.min, info.max, shape, dtype=dt, endpoint=True)
        return rng.standard_normal(shape).astype(dt)
    a, b = gen(sa), gen(sb)
    got, want, den = run_dot(ctx, a, b)
    K = sa[-1] if sa else 1
    assert_dot(got, want, den, K, dt)


@pytest.mark.gpu
def test_dot_scalar_out_forms(ctx):
    """1-D . 1-D into a 0-D or a one-element 1-D output"""
    rng = np.random.default_rng(5)
    a = rng.integers(-100, 100, 257).astype(np.int32)
    b = rng.integers(-100, 100, 257).astype(np.int32)
    for shape in [(), (1,)]:
        O = ctx.array(shape, nc.INT32, zeros=True)
        nc.check(nc.dot(ctx.from_numpy(a), ctx.from_numpy(b), O), "dot")
        assert int(O.numpy().reshape(())) == int(np.dot(a.astype(np.int64), b.astype(np.int64)))


@pytest.mark.gpu
def test_dot_views_and_residency(ctx):
    """transposed / stepped views of a and b, an output with gaps, resident operands"""
    rng = np.random.default_rng(6)
    base_a = rng.standard_normal((40, 3, 24)).astype(np.float32)
    base_b = rng.standard_normal((2, 24, 30)).astype(np.float32)
    A = ctx.from_numpy(base_a).slice(2, 0, 24, 2)        # (40, 3, 12): k stepped
    B = ctx.from_numpy(base_b).slice(1, 0, 24, 2)        # (2, 12, 30)
    a, b = base_a[:, :, 0:24:2], base_b[:, 0:24:2, :]
    want, den = want_dot(a, b)
    O_full = ctx.from_numpy(np.full((40, 3, 2, 31), 7, np.float32))
    O = O_full.slice(3, 0, 30)                          # (40, 3, 2, 30) inside a wider buffer
    for x in (A, B, O_full):
        x.to_device()
    nc.check(nc.dot(A, B, O), "dot views")
    full = O_full.numpy()
    assert_dot(full[..., :30], want, den, 12, np.float32)
    assert np.all(full[..., 30] == 7)


@pytest.mark.gpu
def test_dot_zero_length_axis(ctx):
    """k = 0: every output is the empty sum"""
    A = ctx.array((4, 0), nc.FLOAT32)
    B = ctx.array((0, 3), nc.FLOAT32)
    O = ctx.from_numpy(np.full((4, 3), 5, np.float32))
    nc.check(nc.dot(A, B, O), "dot k=0")
    assert np.all(O.numpy() == 0)


@pytest.mark.gpu
def test_dot_scalar_64bit_round_trip(ctx):
    """the 0-D operand passes through double as numc converts it"""
    s = np.array((1 << 62) + 1, np.int64)                # not representable in double
    b = np.array([1, 2, 3], np.int64)
    O = ctx.array((3,), nc.INT64, zeros=True)
    nc.check(nc.dot(ctx.from_numpy(s), ctx.from_numpy(b), O), "dot")
    sd = int(float(s))                                   # 2^62
    assert O.numpy().tolist() == [(sd * v + (1 << 63)) % (1 << 64) - (1 << 63) for v in [1, 2, 3]]  # low 64 bits
