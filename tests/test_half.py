"""bfloat16 and float16 arrays (the "new dtypes" of numc's ROADMAP.md,
appended to NumcDType after FLOAT64).

CPU: sizes and codes, the seeded generator against the oracle's binary32
streams rounded by the oracle's bf16 conversion (pinned against torch) or
numpy's float16 cast, validation.  GPU: every configuration of both MFMA
families against the FP64 oracle with the bound of FP32 accumulation plus
one 16-bit rounding; numc_matmul_naive bit-exact against an emulation of its
binary32 FMA chain (the products of two 16-bit floats are exact in binary32);
numc_dot, numc_linear, numc_bmm and the pointer entry on 16-bit arrays.
"""
import zlib

import numpy as np
import pytest

import numc_amd as nc
import oracle

U = 2.0 ** -24
U16 = {nc.BFLOAT16: 2.0 ** -8, nc.FLOAT16: 2.0 ** -11}


def to16(x, dt):
    """binary32 values -> the dtype's numpy storage (bf16: uint16 bits)"""
    x = np.asarray(x, np.float32)
    if dt == nc.BFLOAT16:
        return oracle.split_bf16(x)[0].reshape(x.shape)
    return x.astype(np.float16)


def value(x, dt):
    """the dtype's numpy storage -> float64 values"""
    if dt == nc.BFLOAT16:
        return oracle.bf16_to_f32(x).astype(np.float64).reshape(np.shape(x))
    return np.asarray(x).astype(np.float64)


def rand16(rng, shape, dt, scale=1.0):
    return to16(rng.standard_normal(shape).astype(np.float32) * scale, dt)


HALVES = [nc.BFLOAT16, nc.FLOAT16]
IDS = ["bf16", "f16"]


# ---- CPU -----------------------------------------------------------------------

def test_codes_and_sizes(ctx):
    assert (nc.BFLOAT16, nc.FLOAT16) == (10, 11)
    for dt in HALVES:
        a = ctx.array((3, 5), dt)
        assert a.dtype == dt and a.strides == (10, 2)
    assert ctx.from_numpy(np.zeros(4, np.float16)).dtype == nc.FLOAT16
    assert ctx.from_numpy(np.zeros(4, np.uint16)).dtype == nc.UINT16
    assert ctx.from_numpy(np.zeros(4, np.uint16), nc.BFLOAT16).dtype == nc.BFLOAT16


@pytest.mark.parametrize("dt", HALVES, ids=IDS)
@pytest.mark.parametrize("dist", [nc.DIST_UNIFORM, nc.DIST_NORMAL], ids=["uniform", "normal"])
def test_generator_rounds_the_binary32_stream(ctx, dt, dist):
    """the 16-bit fill is the binary32 fill rounded to nearest even"""
    n = 70001
    a = ctx.array((n,), dt)
    a.fill_random(dist, 99, 3, -2.0, 3.0)
    f32 = oracle.fill_uniform(n, np.float32, 99, 3, -2.0, 3.0) if dist == nc.DIST_UNIFORM else \
        oracle.fill_normal(n, np.float32, 99, 3)
    assert np.array_equal(a.numpy(), to16(f32, dt))


def test_oracle_bf16_rounding_closed_forms():
    """ties to even, carries into the exponent, NaN stays NaN, inf stays inf"""
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 2.0 - 2.0 ** -9, np.inf, -np.inf, np.nan],
                 np.float32)
    b = to16(x, nc.BFLOAT16)
    assert b[:6].tolist() == [0x3F80, 0x3F80, 0x3F82, 0x4000, 0x7F80, 0xFF80]
    assert np.isnan(oracle.bf16_to_f32(b[6:]))[0]


def test_python_bf16_bits_matches_oracle():
    """numc_amd.bf16_bits (the Python helper) rounds as the oracle does, on
    random bits of every class (2-D input), and bf16_value inverts it"""
    rng = np.random.default_rng(5)
    x = rng.integers(0, 2 ** 32, (64, 257), dtype=np.uint64).astype(np.uint32).view(np.float32)
    got = nc.bf16_bits(x)
    want = oracle.split_bf16(x.ravel())[0].reshape(x.shape)
    fin = np.isfinite(x)
    assert got.shape == x.shape and np.array_equal(got[fin], want[fin])
    assert np.all(np.isnan(nc.bf16_value(got[np.isnan(x)])))
    assert np.array_equal(nc.bf16_bits(nc.bf16_value(got[fin])), got[fin])


def test_validation(ctx):
    a, b = ctx.array((2, 3), nc.BFLOAT16), ctx.array((3, 4), nc.FLOAT16)
    o = ctx.array((2, 4), nc.BFLOAT16)
    assert nc.matmul(a, b, o) == nc.ERR_TYPE                        # bf16 @ f16
    b2 = ctx.array((3, 4), nc.BFLOAT16)
    assert nc.matmul(a, b2, o, nc.METHOD_SPLIT_BF16) == nc.ERR_TYPE  # methods need float32
    # 8 packed configurations and 2 direct ones per 16-bit family
    assert nc.config_count(nc.FAMILY_BF16) == 10 and nc.config_count(nc.FAMILY_F16) == 10
    assert [nc.config_get(nc.FAMILY_BF16, i)["direct"] for i in range(10)] == [0] * 8 + [1] * 2


# ---- GPU -----------------------------------------------------------------------

def check_gemm(got, a16, b16, dt, K):
    a, b = value(a16, dt).astype(np.float32), value(b16, dt).astype(np.float32)
    ref, den = oracle.fref(a, b)
    err = np.abs(value(got, dt) - ref)
    bound = (K + 3) * U * den + U16[dt] * np.abs(ref) + 2.0 ** -25
    assert np.all(err <= bound), float(np.max(err - bound))


SHAPES = [(300, 333, 200), (64, 64, 64), (1, 517, 3), (257, 1040, 129)]


@pytest.mark.gpu
@pytest.mark.parametrize("dt", HALVES, ids=IDS)
def test_every_config(ctx, dt):
    fam = nc.FAMILY_BF16 if dt == nc.BFLOAT16 else nc.FAMILY_F16
    for cfg in range(nc.config_count(fam)):
        rng = np.random.default_rng(600 + cfg)
        for M, K, N in SHAPES:
            a, b = rand16(rng, (M, K), dt), rand16(rng, (K, N), dt)
            O = ctx.array((M, N), dt, zeros=True)
            nc.check(nc.matmul(ctx.from_numpy(a, dt), ctx.from_numpy(b, dt), O, nc.METHOD_AUTO, cfg), "matmul")
            check_gemm(O.numpy(), a, b, dt, K)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", HALVES, ids=IDS)
def test_strided_and_offset_views(ctx, dt):
    """transposed A (row-contiguous pack), B one element into its buffer
    (unaligned rows: element-wise edge path), a stepped view (lane-slot pack)"""
    rng = np.random.default_rng(7)
    at = rand16(rng, (520, 300), dt)                     # A stored K x M
    bb = rand16(rng, (520, 2 * 150 + 1), dt)
    A = ctx.from_numpy(at, dt).transpose()               # (300, 520)
    B = ctx.from_numpy(bb, dt).slice(1, 1, 301, 2)       # (520, 150), column step 2
    O = ctx.array((300, 150), dt, zeros=True)
    nc.check(nc.matmul(A, B, O), "matmul views")
    check_gemm(O.numpy(), np.ascontiguousarray(at.T), np.ascontiguousarray(bb[:, 1:301:2]), dt, 520)
    B2 = ctx.from_numpy(bb, dt).slice(1, 1, 301)         # (520, 300) starting one element in
    O2 = ctx.array((300, 300), dt, zeros=True)
    nc.check(nc.matmul(A, B2, O2), "matmul offset")
    check_gemm(O2.numpy(), np.ascontiguousarray(at.T), np.ascontiguousarray(bb[:, 1:301]), dt, 520)


@pytest.mark.gpu
@pytest.mark.parametrize("dt", HALVES, ids=IDS)
def test_naive_bit_exact(ctx, dt):
    """numc_matmul_naive: binary32 FMAs in k order, one rounding at the end"""
    rng = np.random.default_rng(8)
    M, K, N = 37, 91, 29
    a, b = rand16(rng, (M, K), dt), rand16(rng, (K, N), dt)
    O = ctx.array((M, N), dt, zeros=True)
    nc.check(nc.lib().numc_matmul_naive(ctx.from_numpy(a, dt).h, ctx.from_numpy(b, dt).h, O.h), "naive")
    af, bf = value(a, dt).astype(np.float32), value(b, dt).astype(np.float32)
    acc = np.zeros((M, N), np.float32)
    for k in range(K):
        acc = acc + af[:, k:k + 1] * bf[k:k + 1, :]   # the product is exact in binary32
    assert np.array_equal(O.numpy(), to16(acc, dt))


@pytest.mark.gpu
@pytest.mark.parametrize("dt", HALVES, ids=IDS)
def test_dot_linear_bmm(ctx, dt):
    rng = np.random.default_rng(zlib.crc32(b"h%d" % dt))
    # numc_dot: N-D . 1-D and a 0-D scalar
    a, v = rand16(rng, (3, 40, 50), dt), rand16(rng, (50,), dt)
    O = ctx.array((3, 40), dt, zeros=True)
    nc.check(nc.dot(ctx.from_numpy(a, dt), ctx.from_numpy(v, dt), O), "dot")
    check_gemm(O.numpy().reshape(120, 1), a.reshape(120, 50), v.reshape(50, 1), dt, 50)
    s = to16(np.float32(1.5), dt).reshape(())
    O = ctx.array((50,), dt, zeros=True)
    nc.check(nc.dot(ctx.from_numpy(s, dt), ctx.from_numpy(v, dt), O), "dot scalar")
    assert np.array_equal(O.numpy(), to16(value(v, dt).astype(np.float32) * np.float32(1.5), dt))
    # numc_linear with a bias
    x, w, bias = rand16(rng, (2, 70, 130), dt), rand16(rng, (130, 90), dt), rand16(rng, (90,), dt)
    O = ctx.array((2, 70, 90), dt, zeros=True)
    nc.check(nc.linear(ctx.from_numpy(x, dt), ctx.from_numpy(w, dt), ctx.from_numpy(bias, dt), O), "linear")
    xf, wf = value(x, dt).reshape(140, 130).astype(np.float32), value(w, dt).astype(np.float32)
    ref, den = oracle.fref(xf, wf)
    ref = ref + value(bias, dt)
    err = np.abs(value(O.numpy(), dt).reshape(140, 90) - ref)
    assert np.all(err <= (130 + 3) * U * den + 2 * U * np.abs(ref) + U16[dt] * np.abs(ref) + 2.0 ** -25)
    # numc_bmm with a broadcast batch axis
    q, k = rand16(rng, (2, 3, 33, 40), dt), rand16(rng, (1, 3, 40, 21), dt)
    O = ctx.array((2, 3, 33, 21), dt, zeros=True)
    nc.check(nc.bmm(ctx.from_numpy(q, dt), ctx.from_numpy(k, dt), O), "bmm")
    got = O.numpy()
    for i in range(2):
        for h in range(3):
            check_gemm(got[i, h], q[i, h], k[0, h], dt, 40)
    # equal slices at linear offsets (the strided-batch launch), a shared 2-D b
    q, k = rand16(rng, (4, 130, 72), dt), rand16(rng, (72, 140), dt)
    O = ctx.array((4, 130, 140), dt, zeros=True)
    nc.check(nc.bmm(ctx.from_numpy(q, dt), ctx.from_numpy(k, dt), O), "bmm strided")
    got = O.numpy()
    for i in range(4):
        check_gemm(got[i], q[i], k, dt, 72)


@pytest.mark.gpu
def test_pointer_entry_bf16(ctx):
    rng = np.random.default_rng(9)
    a, b = rand16(rng, (100, 70), nc.BFLOAT16), rand16(rng, (70, 50), nc.BFLOAT16)
    c = np.zeros((100, 50), np.uint16)
    nc.check(nc.gemm_pointers(ctx, a, b, c, dtype=nc.BFLOAT16), "numc_amd_gemm bf16")
    check_gemm(c, a, b, nc.BFLOAT16, 70)
