"""Pins the oracle (oracle/numc_oracle.c) before anything is checked against it.

Parity status of each oracle function and what pins it:
  orc_matmul_int      numc's known answers (fixture), brute force with Python
                      integers on small random shapes (all 8 integer dtypes,
                      strided views, wraparound), numc's own build in
                      oracle/_ref (numc_matmul_naive, bit-exact) when present.
  orc_matmul_fref     math.fsum / fractions on small shapes (exact sums),
                      known answers, identity/ones closed forms.
  orc_err_metrics     planted errors with known componentwise/normwise values.
  orc_check_matmul    numc's error-contract cases (fixture) and numc's build.
  orc_bf16_rne / split  torch's bfloat16 conversion (independent RNE), the
                      2^-17 representation bound, exhaustive exponent sweep.
  orc_e4m3_rne / to_f32 torch.float8_e4m3fn for every code and for RNE ties.
  orc_fp8_scale_exp   closed form: scaled maximum lies in [128, 256); the
                      numpy fp8_scale_exps and fp8_quantize_blocks element by
                      element against it and the scalar e4m3 encoder.
  generator           splitmix64 / xoshiro256** published first outputs.
No oracle function is left unpinned.
"""
import fractions
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "numc_known_answers.json")))
NPT = {"int8": np.int8, "int16": np.int16, "int32": np.int32, "int64": np.int64, "uint8": np.uint8,
       "uint16": np.uint16, "uint32": np.uint32, "uint64": np.uint64, "float32": np.float32,
       "float64": np.float64}
INT_TYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64]


def brute_int(a, b, out_dtype):
    w = np.dtype(out_dtype).itemsize * 8
    M, K = a.shape
    N = b.shape[1]
    out = np.zeros((M, N), out_dtype)
    signed = np.issubdtype(np.dtype(out_dtype), np.signedinteger)
    for i in range(M):
        for j in range(N):
            s = sum(int(a[i, k]) * int(b[k, j]) for k in range(K)) % (1 << w)
            if signed and s >= 1 << (w - 1):
                s -= 1 << w
            out[i, j] = s
    return out


@pytest.mark.parametrize("case", range(len(GOLD["closed_form"])))
def test_closed_forms(case):
    c = GOLD["closed_form"][case]
    for dt in c["dtypes"]:
        a = np.array(c["a"], NPT[dt])
        b = np.array(c["b"], NPT[dt])
        if np.issubdtype(a.dtype, np.integer):
            got = oracle.matmul_int(a, b)
        else:
            got, _ = oracle.fref(a, b)
        assert np.array_equal(got, np.array(c["expect"])), dt


@pytest.mark.parametrize("dt", INT_TYPES)
def test_matmul_int_brute_force(dt):
    rng = np.random.default_rng(7)
    info = np.iinfo(dt)
    for M, K, N in [(3, 5, 4), (7, 9, 2), (1, 13, 1)]:
        a = rng.integers(info.min, info.max, (M, K), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (K, N), dtype=dt, endpoint=True)
        assert np.array_equal(oracle.matmul_int(a, b), brute_int(a, b, dt))


def test_matmul_int_strided_and_widening():
    rng = np.random.default_rng(3)
    big = rng.integers(-128, 127, (20, 30), dtype=np.int8, endpoint=True)
    a = big[::2, 1::3]          # (10, 10) with both strides non-unit
    b = big.T[3:13, ::3]        # transposed view
    assert np.array_equal(oracle.matmul_int(a, b), brute_int(a, b, np.int8))
    assert np.array_equal(oracle.matmul_int(a, b, np.int32), brute_int(a, b, np.int32))


def test_int8_wraparound_is_mod_256():
    a = np.full((1, 4), 100, np.int8)
    b = np.full((4, 1), 100, np.int8)
    # 4 * 100 * 100 = 40000 = 156 * 256 + 64
    assert oracle.matmul_int(a, b)[0, 0] == 64


def test_fref_exact_sums():
    rng = np.random.default_rng(11)
    a = rng.standard_normal((6, 40)).astype(np.float32)
    b = rng.standard_normal((40, 5)).astype(np.float32)
    ref, absr = oracle.fref(a, b)
    for i in range(6):
        for j in range(5):
            prods = [float(a[i, k]) * float(b[k, j]) for k in range(40)]  # exact in binary64
            assert abs(ref[i, j] - math.fsum(prods)) <= 1e-15 * math.fsum(abs(p) for p in prods)
            assert absr[i, j] == pytest.approx(math.fsum(abs(p) for p in prods), rel=1e-15)
    # double-double for binary64 inputs: compare with exact rationals
    a64 = rng.standard_normal((3, 25))
    b64 = rng.standard_normal((25, 3))
    ref64, _ = oracle.fref(a64, b64)
    for i in range(3):
        for j in range(3):
            exact = sum(fractions.Fraction(a64[i, k]) * fractions.Fraction(b64[k, j]) for k in range(25))
            assert ref64[i, j] == float(exact)


def test_fref_selected_entries_match_full():
    rng = np.random.default_rng(5)
    a = rng.standard_normal((17, 33)).astype(np.float32)
    b = rng.standard_normal((33, 19)).astype(np.float32)
    full, fa = oracle.fref(a, b)
    rows = np.array([0, 16, 5, 3])
    cols = np.array([0, 18, 7, 2])
    sel, sa = oracle.fref(a, b, rows, cols)
    assert np.array_equal(sel, full[rows, cols]) and np.array_equal(sa, fa[rows, cols])
    # a transposed view gives the transposed reference
    ft, _ = oracle.fref(b.T, a.T)
    assert np.array_equal(ft, full.T)


def test_err_metrics_planted():
    ref = np.array([[1.0, -2.0], [4.0, 0.5]])
    absr = np.array([[2.0, 4.0], [8.0, 1.0]])
    c = ref.astype(np.float32).copy()
    c[1, 0] += 0.25   # componentwise 0.25/8, normwise 0.25/4
    m = oracle.err_metrics(c, ref, absr)
    assert m[0] == pytest.approx(0.25 / 8)
    assert m[1] == pytest.approx(0.25 / 4)
    assert m[2] == pytest.approx(0.25 / 4)
    assert m[3] == pytest.approx(0.25)


def test_error_contract_order():
    for e in GOLD["errors"]:
        codes = {"float32": 8, "float64": 9}
        have = [int(e[k] is not None) for k in ("a", "b", "out")]
        shp = [e[k] or [2, 2] for k in ("a", "b", "out")]
        got = oracle.check_matmul(shp[0], shp[1], shp[2], *[codes[d] for d in e["dtypes"]], have=have)
        assert got == e["code"], e["case"]


def test_bf16_against_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    # every exponent, random mantissas, plus exact ties
    exps = np.arange(-126, 128)
    x = (rng.uniform(1, 2, (exps.size, 64)) * np.exp2(exps)[:, None]).astype(np.float32).ravel()
    ties = ((np.arange(1, 4096, dtype=np.uint32) << 16) | 0x8000).view(np.float32)
    x = np.concatenate([x, -x, ties, -ties, np.float32([0.0, 1.0, -3.5])])
    hi, lo = oracle.split_bf16(x)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(hi, want)
    rem = x.astype(np.float64) - oracle.bf16_to_f32(hi).astype(np.float64)
    want_lo = torch.from_numpy(rem.astype(np.float32)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(lo, want_lo)
    normal = np.abs(x) > 2.0 ** -100
    resid = np.abs(x[normal] - oracle.bf16_to_f32(hi[normal]).astype(np.float64)
                   - oracle.bf16_to_f32(lo[normal]).astype(np.float64))
    assert np.all(resid <= 2.0 ** -17 * np.abs(x[normal]))


def test_e4m3_against_torch():
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    dec = oracle.e4m3_decode(codes)
    want = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
    both_nan = np.isnan(dec) & np.isnan(want)
    assert np.array_equal(dec[~both_nan], want[~both_nan]) and np.array_equal(np.isnan(dec), np.isnan(want))
    # round trip of every finite code, and RNE on midpoints between neighbours
    finite = codes[~np.isnan(dec)]
    assert np.array_equal(oracle.e4m3(oracle.e4m3_decode(finite)), finite)
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-440, 440, 4000), rng.uniform(-0.02, 0.02, 4000)]).astype(np.float32)
    mids = ((dec[1:126].astype(np.float64) + dec[2:127]) / 2).astype(np.float32)
    x = np.concatenate([x, mids, -mids])
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(oracle.e4m3(x), want)
    # saturation instead of NaN above the largest finite value
    assert oracle.e4m3(np.float32([500.0, -1e30]))[0] == 0x7E


def test_e4m3_array_equals_scalar_and_quantized_blocks():
    """the array encoder is the scalar one element by element (every finite
    bf16-representable magnitude plus random values), and the block quantizer
    scales each 32-k block's maximum into [128, 256) before rounding, block by
    block, with the last block of a ragged row short"""
    hi = np.arange(0, 1 << 16, dtype=np.uint32)
    x = np.concatenate([(hi << 16).view(np.float32),
                        np.random.default_rng(2).standard_normal(5000).astype(np.float32)])
    x = x[np.isfinite(x)]
    assert np.array_equal(oracle.e4m3_array(x), oracle.e4m3(x))
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((7, 77)) * 2.0 ** rng.integers(-20, 20, (7, 77))).astype(np.float32)
    a[2, 32:64] = 0.0                       # an all-zero block: scale 2^0
    q = oracle.fp8_quantize_blocks(a)
    assert q.shape == a.shape
    for r in range(a.shape[0]):
        for b0 in range(0, a.shape[1], 32):
            blk = a[r, b0:b0 + 32]
            m = np.max(np.abs(blk))
            e = oracle.fp8_scale_exp(m)
            if m > 0:
                assert 128.0 <= m * 2.0 ** e < 256.0
            else:
                assert e == 0
            want = oracle.e4m3_decode(oracle.e4m3(np.ldexp(blk, e).astype(np.float32))).astype(np.float64)
            assert np.array_equal(q[r, b0:b0 + 32], np.ldexp(want, -e))
    # columns of B are quantized as the rows of B^T
    qa, qb = oracle.fp8_quantized_inputs(a[:, :40], a[:, :40].T)
    assert np.array_equal(qb, qa.T)


def test_fp8_scale_exps_equal_scalar():
    """the numpy block exponents are the C function's on maxima across the whole
    float32 range (zero, subnormals, both clamps, powers of two and neighbours)"""
    m = np.concatenate([np.float32([0.0, 1e-45, 1e-40, 1.1754944e-38, 2.0 ** -126, 1e-30, 0.5, 1.0, 448.0,
                                    3e38, np.finfo(np.float32).max]),
                        np.float32(2.0) ** np.arange(-149, 128).astype(np.float32),
                        np.nextafter(np.float32(2.0) ** np.arange(-100, 100).astype(np.float32), np.float32(0)),
                        np.abs(np.random.default_rng(4).standard_normal(2000).astype(np.float32)) *
                        np.float32(2.0) ** np.random.default_rng(5).integers(-60, 60, 2000).astype(np.float32)])
    want = np.array([oracle.fp8_scale_exp(float(v)) for v in m])
    assert np.array_equal(oracle.fp8_scale_exps(m), want)


def test_fp8_scale_exp_closed_form():
    for m in [1e-30, 3e-5, 0.5, 1.0, 1.7, 127.9, 128.0, 255.9, 256.0, 3e20]:
        e = oracle.fp8_scale_exp(m)
        assert 128.0 <= m * 2.0 ** e < 256.0
    assert oracle.fp8_scale_exp(0.0) == 0


def test_generator_published_values():
    # splitmix64 from state 0: first output 0xE220A8397B1DCDAF (Vigna's reference)
    st = [0]

    def sm():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return z ^ (z >> 31)
    assert sm() == 0xE220A8397B1DCDAF
    # block 0 of (seed 0, stream 0) is seeded from splitmix64 state 0
    st[0] = 0
    s = [sm() for _ in range(4)]
    first = ((((s[1] * 5) << 7 | (s[1] * 5 & (2 ** 64 - 1)) >> 57) & (2 ** 64 - 1)) * 9) & (2 ** 64 - 1)
    assert oracle.xoshiro_block(0, 0, 0, 1)[0] == first


def test_generator_ranges():
    u = oracle.fill_uniform(200000, np.float32, 1, 0)
    assert u.min() >= -1.0 and u.max() < 1.0 and abs(u.mean()) < 0.01
    n = oracle.fill_normal(200000, np.float64, 1, 1)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1.0) < 0.01
    i = oracle.fill_int(200000, np.int32, 1, 2, -100, 100)
    assert i.min() == -100 and i.max() == 100
    i8 = oracle.fill_int(100000, np.int8, 1, 3, -128, 127)
    assert i8.min() == -128 and i8.max() == 127


# ---- numc's own build (oracle/_ref), when it was built here --------------------

def _ref_lib():
    import ctypes as C
    if not os.path.exists(oracle.REF_LIB):
        pytest.skip("oracle/_ref not built (numc sources absent)")
    L = C.CDLL(oracle.REF_LIB)  # RTLD_LOCAL: its numc_* stay private to this handle
    L.numc_ctx_create.restype = C.c_void_p
    L.numc_ctx_free.argtypes = [C.c_void_p]
    L.numc_array_create.restype = C.c_void_p
    L.numc_array_create.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), C.c_size_t, C.c_int]
    L.numc_array_write.argtypes = [C.c_void_p, C.c_void_p]
    L.numc_array_data.restype = C.c_void_p
    L.numc_array_data.argtypes = [C.c_void_p]
    L.numc_matmul.argtypes = [C.c_void_p] * 3
    L.numc_matmul_naive.argtypes = [C.c_void_p] * 3
    return L


def _ref_run(L, fn, a, b):
    import ctypes as C
    ctx = L.numc_ctx_create()
    try:
        code = oracle.CODE_OF[a.dtype]
        M, K = a.shape
        N = b.shape[1]
        arrs = []
        for shp, src in (((M, K), a), ((K, N), b), ((M, N), None)):
            h = L.numc_array_create(ctx, (C.c_size_t * 2)(*shp), 2, code)
            if src is not None:
                L.numc_array_write(h, np.ascontiguousarray(src).ctypes.data_as(C.c_void_p))
            arrs.append(h)
        assert getattr(L, fn)(*arrs) == 0
        buf = (C.c_char * (M * N * a.itemsize)).from_address(L.numc_array_data(arrs[2]))
        return np.frombuffer(buf, a.dtype).reshape(M, N).copy()
    finally:
        L.numc_ctx_free(ctx)


@pytest.mark.parametrize("dt", INT_TYPES)
def test_oracle_equals_numc_naive(dt):
    L = _ref_lib()
    rng = np.random.default_rng(2)
    info = np.iinfo(dt)
    lo, hi = (info.min, info.max) if info.bits <= 32 else (-2 ** 40, 2 ** 40)
    if not np.issubdtype(dt, np.signedinteger):
        lo = 0
    a = rng.integers(lo, hi, (37, 53), dtype=dt, endpoint=True)
    b = rng.integers(lo, hi, (53, 29), dtype=dt, endpoint=True)
    assert np.array_equal(_ref_run(L, "numc_matmul_naive", a, b), oracle.matmul_int(a, b))


@pytest.mark.parametrize("dt", [np.int16, np.int32, np.uint32, np.int64])
def test_oracle_equals_numc_packed(dt):
    L = _ref_lib()
    rng = np.random.default_rng(4)
    a = rng.integers(0, 100, (250, 300), dtype=dt)
    b = rng.integers(0, 100, (300, 260), dtype=dt)
    assert np.array_equal(_ref_run(L, "numc_matmul", a, b), oracle.matmul_int(a, b))


def test_numc_packed_int8_saturates_unlike_naive():
    """numc's packed int8 GEMM saturates per KC block (so the wraparound
    contract follows numc_matmul_naive); documented, not matched."""
    L = _ref_lib()
    rng = np.random.default_rng(9)
    a = rng.integers(-128, 127, (256, 1024), dtype=np.int8, endpoint=True)
    b = rng.integers(-128, 127, (1024, 256), dtype=np.int8, endpoint=True)
    naive = _ref_run(L, "numc_matmul_naive", a, b)
    assert np.array_equal(naive, oracle.matmul_int(a, b))
    assert np.mean(_ref_run(L, "numc_matmul", a, b) != naive) > 0.5


def test_oracle_float_close_to_numc():
    L = _ref_lib()
    a = oracle.fill_uniform(512 * 512, np.float32, 42, 0).reshape(512, 512)
    b = oracle.fill_uniform(512 * 512, np.float32, 42, 1).reshape(512, 512)
    ref, absr = oracle.fref(a, b)
    got = _ref_run(L, "numc_matmul", a, b)
    assert oracle.err_metrics(got, ref, absr)[0] < 1e-6
