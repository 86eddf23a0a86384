"""GPU parity of numc_matmul / numc_matmul_naive / numc_matmul_ex against the
oracle, through the C ABI.  Bars: bit-exact for every integer dtype; for
floats the componentwise error |C - R| / (|A||B|) against the FP64 (or
double-double) oracle stays under the bound stated in each test.

The cases mirror numc's tests/matmul/test_matmul_naive.c and test_matmul.c
(fixture tests/golden/numc_known_answers.json) and add random, strided,
empty and full-range cases plus every kernel configuration of every family.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import numc_amd as nc
import oracle

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "numc_known_answers.json")))
NPT = {"int8": np.int8, "int16": np.int16, "int32": np.int32, "int64": np.int64, "uint8": np.uint8,
       "uint16": np.uint16, "uint32": np.uint32, "uint64": np.uint64, "float32": np.float32,
       "float64": np.float64}
U = 2.0 ** -24   # binary32 unit roundoff
U64 = 2.0 ** -53


def gemm(ctx, a, b, out_dtype=None, method=0, config=-1, flags=0, naive=False):
    A = ctx.from_numpy(a)
    B = ctx.from_numpy(b)
    O = ctx.array((a.shape[0], b.shape[1]), nc.DTYPE_OF_NP[np.dtype(out_dtype or a.dtype)], zeros=True)
    if naive:
        code = nc.lib().numc_matmul_naive(A.h, B.h, O.h)
    else:
        code = nc.matmul(A, B, O, method, config, flags)
    nc.check(code, "matmul")
    return O.numpy()


def fill(shape, dt, expr):
    i = np.arange(int(np.prod(shape)))
    return eval(expr, {"i": i}).astype(dt).reshape(shape)


def cw_err(c, a, b):
    ref, absr = oracle.fref(a, b)
    return oracle.err_metrics(c.astype(np.float64) if c.dtype == np.float64 else c, ref, absr)[0]


# ---- numc's known answers ----------------------------------------------------

@pytest.mark.parametrize("naive", [False, True])
def test_closed_forms(ctx, naive):
    for c in GOLD["closed_form"]:
        for dt in c["dtypes"]:
            got = gemm(ctx, np.array(c["a"], NPT[dt]), np.array(c["b"], NPT[dt]), naive=naive)
            assert np.array_equal(got, np.array(c["expect"], NPT[dt])), (dt, naive)


def test_identity(ctx):
    for c in GOLD["identity"]:
        dt, n = NPT[c["dtype"]], c["n"]
        a = fill((n, n), dt, "i+1" if c["fill"] == "i+1" else "i%" + c["fill"].split("%")[1])
        assert np.array_equal(gemm(ctx, a, np.eye(n, dtype=dt)), a), c["dtype"]


def test_ones(ctx):
    for c in GOLD["ones"]:
        dt = NPT[c["dtype"]]
        got = gemm(ctx, np.ones((c["M"], c["K"]), dt), np.ones((c["K"], c["N"]), dt))
        assert np.all(got == c["K"])


@pytest.mark.parametrize("idx", range(len(GOLD["cross"])))
def test_cross_fills(ctx, idx):
    c = GOLD["cross"][idx]
    dt = NPT[c["dtype"]]
    a = fill((c["M"], c["K"]), dt, c.get("fill_a", "i%7+1"))
    b = fill((c["K"], c["N"]), dt, c.get("fill_b", "i%5+1"))
    got = gemm(ctx, a, b)
    naive = gemm(ctx, a, b, naive=True)
    if np.issubdtype(dt, np.integer):
        want = oracle.matmul_int(a, b)
        assert np.array_equal(got, want) and np.array_equal(naive, want)
    else:
        ref, _ = oracle.fref(a, b)
        assert np.max(np.abs(got - ref)) <= c["tol"]
        assert np.max(np.abs(got - naive)) <= c["tol"]


def test_int8_cross(ctx):
    c = GOLD["int8_cross"]
    a = fill((c["M"], c["K"]), np.int8, "i%2")
    b = fill((c["K"], c["N"]), np.int8, "i%2")
    assert np.array_equal(gemm(ctx, a, b), gemm(ctx, a, b, naive=True))
    assert np.array_equal(gemm(ctx, a, b), oracle.matmul_int(a, b))


@pytest.mark.parametrize("edge", GOLD["edges"], ids=lambda e: "%dx%dx%d-%s" % tuple(e[:4]))
def test_edges(ctx, edge):
    M, K, N, dtn, tol = edge
    dt = NPT[dtn]
    a = fill((M, K), dt, "i%7+1")
    b = fill((K, N), dt, "i%5+1")
    got = gemm(ctx, a, b)
    if np.issubdtype(dt, np.integer):
        assert np.array_equal(got, oracle.matmul_int(a, b))
    else:
        ref, _ = oracle.fref(a, b)
        assert np.max(np.abs(got - ref)) <= tol


# ---- every kernel configuration -----------------------------------------------

SHAPES = [(300, 333, 200), (64, 64, 64), (1, 517, 3), (257, 31, 129)]


def rand_f32(rng, shape):
    return rng.uniform(-1, 1, shape).astype(np.float32)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_F32)))
def test_f32_configs(ctx, cfg):
    rng = np.random.default_rng(100 + cfg)
    for M, K, N in SHAPES:
        a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
        got = gemm(ctx, a, b, config=cfg)
        # deterministic bound for K-term binary32 accumulation: gamma_K = K*u/(1-K*u)
        assert cw_err(got, a, b) <= (K + 2) * U, (cfg, M, K, N)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_F32)))
def test_f32_configs_strided(ctx, cfg):
    """every FP32 configuration on transposed / stepped operand views (both
    packed) and an output with gaps, ragged shapes"""
    rng = np.random.default_rng(150 + cfg)
    for M, K, N in [(67, 45, 93), (130, 257, 33), (1, 19, 300)]:
        at = rng.standard_normal((K, M)).astype(np.float32)       # A stored transposed
        bs = rng.standard_normal((K, 2 * N + 1)).astype(np.float32)
        A = ctx.from_numpy(at).transpose()
        B = ctx.from_numpy(bs).slice(1, 1, 2 * N + 1, 2)            # every other column
        O_full = ctx.from_numpy(np.full((M, N + 3), 7, np.float32))
        O = O_full.slice(1, 1, N + 1)
        nc.check(nc.matmul(A, B, O, nc.METHOD_PLAIN, cfg), "strided f32")
        full = O_full.numpy()
        a, b = np.ascontiguousarray(at.T), np.ascontiguousarray(bs[:, 1::2])
        assert cw_err(np.ascontiguousarray(full[:, 1:N + 1]), a, b) <= (K + 2) * U, (cfg, M, K, N)
        assert np.all(full[:, 0] == 7) and np.all(full[:, N + 1:] == 7)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_SPLIT_BF16)))
def test_split_bf16_configs(ctx, cfg):
    rng = np.random.default_rng(200 + cfg)
    for M, K, N in SHAPES:
        a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
        got = gemm(ctx, a, b, method=nc.METHOD_SPLIT_BF16, config=cfg)
        # per product: dropped lo*lo and the two 2^-17 representation residues
        # (<= 3 * 2^-16 |a||b|), plus binary32 accumulation
        assert cw_err(got, a, b) <= 3 * 2.0 ** -16 + (K + 2) * U, (cfg, M, K, N)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_FP8)))
def test_fp8_configs(ctx, cfg):
    rng = np.random.default_rng(300 + cfg)
    for M, K, N in SHAPES:
        a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
        got = gemm(ctx, a, b, method=nc.METHOD_FP8_SCALED, config=cfg)
        qa, qb = oracle.fp8_quantized_inputs(a, b)
        ref, absr = oracle.fref(qa, qb)   # exact operands of the FP8 products
        err = oracle.err_metrics(got, ref, absr)[0]
        # the f8f6f4 MFMA sums its 64 products internally with less than
        # binary32 precision: measured up to 2^-14.3 per instruction for
        # N(0,1) operands and 2^-11.5 when a row's magnitudes spread over 2^12
        assert err <= -(-K // 64) * 2.0 ** -11 + (K + 2) * U, (cfg, M, K, N, err)


def test_fp8_spread_magnitudes(ctx):
    """Operands whose magnitudes spread over 2^12 within each row and column:
    the case where the f8f6f4 MFMA's internal sum loses the most."""
    rng = np.random.default_rng(310)
    for M, K, N in [(256, 31, 256), (256, 64, 256), (128, 1000, 192)]:
        a = (rng.uniform(-1, 1, (M, K)) * 2.0 ** rng.integers(-12, 0, (M, K))).astype(np.float32)
        b = (rng.uniform(-1, 1, (K, N)) * 2.0 ** rng.integers(-12, 0, (K, N))).astype(np.float32)
        got = gemm(ctx, a, b, method=nc.METHOD_FP8_SCALED)
        qa, qb = oracle.fp8_quantized_inputs(a, b)
        ref, absr = oracle.fref(qa, qb)
        err = oracle.err_metrics(got, ref, absr)[0]
        assert err <= -(-K // 64) * 2.0 ** -11 + (K + 2) * U, (M, K, N, err)


def test_fp8_block_scales(ctx):
    """Magnitudes that change by 2^40 from one 32-k block to the next along each
    row of A and column of B, so that every output sums products from tiny and
    from huge blocks in equal measure: one scale per row (or column) would
    flush the tiny blocks to zero, a scale per 32-k block keeps them.  Checked
    against the quantised operands and end to end against the FP64 sums of the
    unquantised inputs (e4m3 keeps 3 mantissa bits: 2 * 2^-4 per product)."""
    rng = np.random.default_rng(320)
    for M, K, N in [(64, 128, 96), (200, 320, 130), (96, 1000, 64)]:
        blk = np.arange(K) // 32
        sa = np.where(blk % 2 == 0, -40.0, 0.0)     # A: even blocks tiny
        sb = np.where(blk % 2 == 0, 40.0, 0.0)      # B: even blocks huge
        a = np.ldexp(rng.uniform(-1, 1, (M, K)), sa[None, :].astype(int)).astype(np.float32)
        b = np.ldexp(rng.uniform(-1, 1, (K, N)), sb[:, None].astype(int)).astype(np.float32)
        got = gemm(ctx, a, b, method=nc.METHOD_FP8_SCALED)
        qa, qb = oracle.fp8_quantized_inputs(a, b)
        ref, absr = oracle.fref(qa, qb)
        err = oracle.err_metrics(got, ref, absr)[0]
        assert err <= -(-K // 64) * 2.0 ** -11 + (K + 2) * U, (M, K, N, err)
        err0 = cw_err(got, a, b)
        assert err0 <= 2 * 2.0 ** -4 + 2.0 ** -8, (M, K, N, err0)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_F64)))
def test_f64_configs(ctx, cfg):
    rng = np.random.default_rng(400 + cfg)
    for M, K, N in SHAPES:
        a, b = rng.standard_normal((M, K)), rng.standard_normal((K, N))
        got = gemm(ctx, a, b, config=cfg)
        assert cw_err(got, a, b) <= (K + 2) * U64, (cfg, M, K, N)


INTS = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32]


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_I8)))
@pytest.mark.parametrize("dt", INTS, ids=lambda d: np.dtype(d).name)
def test_int_configs_full_range(ctx, cfg, dt):
    rng = np.random.default_rng(500 + cfg)
    info = np.iinfo(dt)
    if nc.config_get(nc.FAMILY_I8, cfg)["direct"] and dt not in (np.int8, np.uint8):
        # the direct kernel reads one int8 digit plane; wider integers need the
        # packed digit planes, and naming a direct configuration is refused
        a = np.ones((4, 5), dt)
        A, B = ctx.from_numpy(a), ctx.from_numpy(a.T.copy())
        O = ctx.array((4, 4), nc.DTYPE_OF_NP[np.dtype(dt)], zeros=True)
        assert nc.matmul(A, B, O, nc.METHOD_AUTO, cfg) == nc.ERR_VALUE
        assert "one int8 digit plane" in nc.last_error()[1]
        return
    for M, K, N in SHAPES:
        a = rng.integers(info.min, info.max, (M, K), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (K, N), dtype=dt, endpoint=True)
        assert np.array_equal(gemm(ctx, a, b, config=cfg), oracle.matmul_int(a, b)), (cfg, M, K, N)


@pytest.mark.parametrize("dt", [np.int64, np.uint64], ids=lambda d: np.dtype(d).name)
def test_int64(ctx, dt):
    rng = np.random.default_rng(6)
    info = np.iinfo(dt)
    for M, K, N in [(33, 70, 45), (1, 1, 1), (64, 64, 64)]:
        a = rng.integers(info.min, info.max, (M, K), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (K, N), dtype=dt, endpoint=True)
        assert np.array_equal(gemm(ctx, a, b), oracle.matmul_int(a, b))
        assert np.array_equal(gemm(ctx, a, b, naive=True), oracle.matmul_int(a, b))


@pytest.mark.parametrize("dt", INTS + [np.float32, np.float64], ids=lambda d: np.dtype(d).name)
def test_naive_random(ctx, dt):
    rng = np.random.default_rng(8)
    M, K, N = 45, 77, 39
    if np.issubdtype(dt, np.integer):
        info = np.iinfo(dt)
        a = rng.integers(info.min, info.max, (M, K), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (K, N), dtype=dt, endpoint=True)
        assert np.array_equal(gemm(ctx, a, b, naive=True), oracle.matmul_int(a, b))
    else:
        a, b = rng.standard_normal((M, K)).astype(dt), rng.standard_normal((K, N)).astype(dt)
        u = U if dt == np.float32 else U64
        assert cw_err(gemm(ctx, a, b, naive=True), a, b) <= (K + 2) * u


def test_widen_int8_to_int32(ctx):
    rng = np.random.default_rng(12)
    for dt, odt in [(np.int8, np.int32), (np.uint8, np.uint32), (np.int8, np.uint32)]:
        info = np.iinfo(dt)
        a = rng.integers(info.min, info.max, (130, 1000), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (1000, 70), dtype=dt, endpoint=True)
        got = gemm(ctx, a, b, out_dtype=odt, flags=nc.OPT_WIDEN)
        assert np.array_equal(got, oracle.matmul_int(a, b, odt)), (dt, odt)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_I8)))
def test_widen_int8_configs(ctx, cfg):
    """int8 -> int32 takes one digit plane (an int8 value is its own signed
    digit): every configuration, split-K ones included, and a sum that wraps
    mod 2^32 (all -128 entries, K = 2^17 + 40: K * 2^14 exceeds 2^31)."""
    rng = np.random.default_rng(620 + cfg)
    a = rng.integers(-128, 127, (130, 1000), dtype=np.int8, endpoint=True)
    b = rng.integers(-128, 127, (1000, 70), dtype=np.int8, endpoint=True)
    got = gemm(ctx, a, b, out_dtype=np.int32, flags=nc.OPT_WIDEN, config=cfg)
    assert np.array_equal(got, oracle.matmul_int(a, b, np.int32)), cfg
    K = (1 << 17) + 40
    a = np.full((33, K), -128, np.int8)
    b = np.full((K, 17), -128, np.int8)
    want = np.full((33, 17), np.int64(K) * 16384, np.int64).astype(np.uint32).view(np.int32)
    assert np.array_equal(oracle.matmul_int(a, b, np.int32), want)
    assert np.array_equal(gemm(ctx, a, b, out_dtype=np.int32, flags=nc.OPT_WIDEN, config=cfg), want), cfg


# ---- views, residency, empty shapes -----------------------------------------------

def test_strided_views(ctx):
    rng = np.random.default_rng(13)
    for dt in [np.float32, np.float64, np.int32, np.int8]:
        if np.issubdtype(dt, np.integer):
            base_a = rng.integers(-50, 50, (90, 70)).astype(dt)
            base_b = rng.integers(-50, 50, (80, 60)).astype(dt)
        else:
            base_a = rng.standard_normal((90, 70)).astype(dt)
            base_b = rng.standard_normal((80, 60)).astype(dt)
        A = ctx.from_numpy(base_a).transpose()          # (70, 90), strides (es, 70es)
        B = ctx.from_numpy(base_b).slice(0, 0, 80, 1)   # full view
        B = B.slice(1, 1, 60, 2)                        # (80, 30) every other column
        Bs = base_b[:, 1:60:2]
        Asub = A.slice(1, 0, 80)                        # (70, 80)
        O_full = ctx.from_numpy(np.full((70, 65), 7, dt))
        O = O_full.slice(1, 2, 62, 2)                   # (70, 30) with gaps
        nc.check(nc.matmul(Asub, B, O), "matmul strided")
        want_a = base_a.T[:, :80]
        full = O_full.numpy()
        got = full[:, 2:62:2]
        if np.issubdtype(dt, np.integer):
            assert np.array_equal(got, oracle.matmul_int(np.ascontiguousarray(want_a), np.ascontiguousarray(Bs)))
        else:
            assert cw_err(got, np.ascontiguousarray(want_a), np.ascontiguousarray(Bs)) <= 100 * U
        mask = np.ones(full.shape, bool)
        mask[:, 2:62:2] = False
        assert np.all(full[mask] == 7), "bytes between output columns must be left alone"


PACK_DTS = [np.int8, np.uint8, np.int16, np.uint32, np.float16, np.float32]


@pytest.mark.parametrize("dt", PACK_DTS, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("offset", [0, 1], ids=["aligned", "offset"])
def test_pack_interior_and_edge_tiles(ctx, dt, offset):
    """Operands large enough for whole interior pack tiles beside ragged edge
    tiles; offset 1 starts both operands one element into their buffers, so no
    row is 16-byte aligned and every 16-byte load of a tile is unaligned."""
    rng = np.random.default_rng(170 + offset)
    M, K, N = 96, 1040, 320
    if np.issubdtype(dt, np.integer):
        info = np.iinfo(dt)
        base_a = rng.integers(info.min, info.max, (M, K + offset), dtype=dt, endpoint=True)
        base_b = rng.integers(info.min, info.max, (K, N + offset), dtype=dt, endpoint=True)
    else:
        base_a = rng.standard_normal((M, K + offset)).astype(dt)
        base_b = rng.standard_normal((K, N + offset)).astype(dt)
    A = ctx.from_numpy(base_a).slice(1, offset, K + offset)
    B = ctx.from_numpy(base_b).slice(1, offset, N + offset)
    a, b = np.ascontiguousarray(base_a[:, offset:]), np.ascontiguousarray(base_b[:, offset:])
    methods = [nc.METHOD_PLAIN, nc.METHOD_SPLIT_BF16] if dt == np.float32 else [nc.METHOD_AUTO]
    for method in methods:
        O = ctx.array((M, N), nc.DTYPE_OF_NP[np.dtype(dt)], zeros=True)
        nc.check(nc.matmul(A, B, O, method), "matmul pack tiles")
        got = O.numpy()
        if np.issubdtype(dt, np.integer):
            assert np.array_equal(got, oracle.matmul_int(a, b)), (dt, offset)
        else:
            bound = (K + 2) * U + (3 * 2.0 ** -16 if method == nc.METHOD_SPLIT_BF16 else 0.0)
            if dt == np.float16:  # float16 values are exact in float32; the store rounds to 11 bits
                bound += 2.0 ** -11
                got, a, b = got.astype(np.float32), a.astype(np.float32), b.astype(np.float32)
            assert cw_err(got, a, b) <= bound, (method, offset)


EDGE_CASES = [(np.int8, nc.FAMILY_I8, nc.METHOD_AUTO), (np.uint8, nc.FAMILY_I8, nc.METHOD_AUTO),
              (np.int16, nc.FAMILY_I8, nc.METHOD_AUTO), (np.int32, nc.FAMILY_I8, nc.METHOD_AUTO),
              (np.float32, nc.FAMILY_F32, nc.METHOD_PLAIN),
              (np.float32, nc.FAMILY_SPLIT_BF16, nc.METHOD_SPLIT_BF16),
              (np.float16, nc.FAMILY_F16, nc.METHOD_AUTO)]


@pytest.mark.parametrize("dt,family,method", EDGE_CASES,
                         ids=["int8", "uint8", "int16", "int32", "f32", "split_bf16", "f16"])
def test_pack_partial_chunks_every_offset(ctx, dt, family, method):
    """The packs' partial 16-byte chunks (a row's last, cut chunk, read as the
    aligned chunks around it and shifted into place): rows starting at every
    byte offset within 16, with row lengths that leave every cut, in packed
    configurations (A k-contiguous, B row-contiguous), the last row ending at
    the end of its buffer."""
    cfg = next(i for i in range(nc.config_count(family)) if not nc.config_get(family, i)["direct"])
    rng = np.random.default_rng(404)
    es = np.dtype(dt).itemsize
    M, N = 33, 37
    for off in range(16 // es):
        for K in (1, 7, 13, 16, 17, 35):
            if np.issubdtype(dt, np.integer):
                info = np.iinfo(dt)
                base_a = rng.integers(info.min, info.max, (M, K + off), dtype=dt, endpoint=True)
                base_b = rng.integers(info.min, info.max, (K, N + off), dtype=dt, endpoint=True)
            else:
                base_a = rng.standard_normal((M, K + off)).astype(dt)
                base_b = rng.standard_normal((K, N + off)).astype(dt)
            A = ctx.from_numpy(base_a).slice(1, off, off + K)
            B = ctx.from_numpy(base_b).slice(1, off, off + N)
            a, b = np.ascontiguousarray(base_a[:, off:]), np.ascontiguousarray(base_b[:, off:])
            O = ctx.array((M, N), nc.DTYPE_OF_NP[np.dtype(dt)], zeros=True)
            nc.check(nc.matmul(A, B, O, method, cfg), "matmul partial chunks")
            got = O.numpy()
            if np.issubdtype(dt, np.integer):
                assert np.array_equal(got, oracle.matmul_int(a, b)), (off, K)
            else:
                bound = (K + 2) * U + (3 * 2.0 ** -16 if method == nc.METHOD_SPLIT_BF16 else 0.0)
                if dt == np.float16:  # float16 values are exact in float32; the store rounds to 11 bits
                    bound += 2.0 ** -11
                    got, a, b = got.astype(np.float32), a.astype(np.float32), b.astype(np.float32)
                assert cw_err(got, a, b) <= bound, (off, K)


def test_pointer_entry_host_and_device(ctx):
    """numc_amd_gemm (pointer-level entry for a host library keeping its own
    arrays) on host pointers with strided views, widening, and device pointers."""
    rng = np.random.default_rng(21)
    # float32, A transposed view, B every other column, C with gaps between columns
    base_a = rng.standard_normal((90, 70)).astype(np.float32)
    base_b = rng.standard_normal((80, 61)).astype(np.float32)
    a, b = base_a.T[:, :80], base_b[:, 1:61:2]
    full = np.full((70, 65), 7, np.float32)
    c = full[:, 2:62:2]
    for method, bound in [(nc.METHOD_PLAIN, 100 * U), (nc.METHOD_SPLIT_BF16, 3 * 2.0 ** -16 + 82 * U)]:
        full[:] = 7
        nc.check(nc.gemm_pointers(ctx, a, b, c, method=method), "numc_amd_gemm f32")
        assert cw_err(c.copy(), np.ascontiguousarray(a), np.ascontiguousarray(b)) <= bound
        mask = np.ones(full.shape, bool)
        mask[:, 2:62:2] = False
        assert np.all(full[mask] == 7), "bytes between output columns must be left alone"
    # every integer dtype up to 32 bits, exact wraparound; int8 -> int32 widening
    for dt in [np.int8, np.uint16, np.int32]:
        info = np.iinfo(dt)
        x = rng.integers(info.min, info.max, (33, 47), endpoint=True).astype(dt)
        y = rng.integers(info.min, info.max, (47, 29), endpoint=True).astype(dt)
        z = np.zeros((33, 29), dt)
        nc.check(nc.gemm_pointers(ctx, x, y, z), "numc_amd_gemm int")
        assert np.array_equal(z, oracle.matmul_int(x, y))
    x = rng.integers(-128, 127, (64, 300), endpoint=True).astype(np.int8)
    y = rng.integers(-128, 127, (300, 50), endpoint=True).astype(np.int8)
    z = np.zeros((64, 50), np.int32)
    nc.check(nc.gemm_pointers(ctx, x, y, z, flags=nc.OPT_WIDEN), "numc_amd_gemm widen")
    assert np.array_equal(z, x.astype(np.int64) @ y.astype(np.int64))
    # device pointers: operands already in HBM (here: arrays made resident)
    a2 = rng.uniform(-1, 1, (120, 96)).astype(np.float32)
    b2 = rng.uniform(-1, 1, (96, 72)).astype(np.float32)
    A, B = ctx.from_numpy(a2), ctx.from_numpy(b2)
    O = ctx.array((120, 72), nc.FLOAT32, zeros=True)
    for arr in (A, B, O):
        arr.to_device()
    L = nc.lib()
    da, db, do = (L.numc_array_device_data(arr.h) for arr in (A, B, O))
    opts = nc.MatmulOpts(nc.METHOD_PLAIN, -1, nc.OPT_DEVICE_POINTERS)
    nc.check(L.numc_amd_gemm(ctx.h, nc.FLOAT32, nc.FLOAT32, 120, 72, 96, da, 96 * 4, 4, db, 72 * 4, 4,
                             do, 72 * 4, 4, C.byref(opts)), "numc_amd_gemm device")
    assert cw_err(O.numpy(), a2, b2) <= 98 * U


def test_device_residency(ctx):
    rng = np.random.default_rng(14)
    a = rng.uniform(-1, 1, (200, 300)).astype(np.float32)
    b = rng.uniform(-1, 1, (300, 100)).astype(np.float32)
    A, B = ctx.from_numpy(a), ctx.from_numpy(b)
    O = ctx.array((200, 100), nc.FLOAT32, zeros=True)
    for x in (A, B, O):
        x.to_device()
        assert nc.lib().numc_array_on_device(x.h)
    nc.check(nc.matmul(A, B, O), "matmul resident")
    assert nc.lib().numc_array_on_device(O.h)
    got = O.numpy()                      # copies back, host current again
    assert not nc.lib().numc_array_on_device(O.h)
    assert cw_err(got, a, b) <= 302 * U
    # writing through the host pointer after numc_array_data is seen by the next call
    B.write(np.zeros_like(b))
    nc.check(nc.matmul(A, B, O), "matmul after write")
    assert np.all(O.numpy() == 0)


FILL_CASES = [(nc.FLOAT32, nc.DIST_UNIFORM, -2, 3), (nc.FLOAT64, nc.DIST_UNIFORM, -1, 1),
              (nc.INT32, nc.DIST_INT, -100, 100), (nc.INT8, nc.DIST_INT, -128, 127),
              (nc.UINT16, nc.DIST_INT, 0, 65535), (nc.BFLOAT16, nc.DIST_UNIFORM, -1, 1),
              (nc.FLOAT16, nc.DIST_UNIFORM, -1, 1), (nc.FLOAT32, nc.DIST_NORMAL, 0, 0),
              (nc.FLOAT64, nc.DIST_NORMAL, 0, 0), (nc.BFLOAT16, nc.DIST_NORMAL, 0, 0),
              (nc.FLOAT16, nc.DIST_NORMAL, 0, 0)]


@pytest.mark.parametrize("case", FILL_CASES, ids=lambda c: "%d-%d" % c[:2])
def test_device_fill_matches_host_fill(ctx, case):
    """a device-current array fills on the GPU with the host fill's code:
    identical uniform and integer values; normal values within one unit in
    the last place of the stored type, eight for float64 (the device computes
    log/sin/cos with its own binary64 library); the array stays device-current"""
    dt, dist, lo, hi = case
    n = 3 * 65536 + 1234                       # four streams, the last one partial
    D, H = ctx.array((n,), dt), ctx.array((n,), dt)
    D.to_device()
    D.fill_random(dist, 77, 4, lo, hi)
    assert nc.lib().numc_array_on_device(D.h)
    H.fill_random(dist, 77, 4, lo, hi)
    d, h = D.numpy(), H.numpy()
    if dist != nc.DIST_NORMAL:
        assert np.array_equal(d, h)
        return
    bits = {2: np.int16, 4: np.int32, 8: np.int64}[d.dtype.itemsize]
    di, hi_ = d.view(bits).astype(np.int64), h.view(bits).astype(np.int64)
    # float64 shows the libraries' differences directly (log, cos/sin near
    # their zeros, and the product: a few units in the last place); narrower
    # types round them away except next to a rounding boundary
    f64 = dt == nc.FLOAT64
    assert np.all(np.abs(di - hi_) <= (8 if f64 else 1))   # units in the last place
    assert np.mean(di == hi_) > (0.5 if f64 else 0.999)


def test_async_mode(ctx):
    ctx.set_sync(False)
    a = np.ones((64, 64), np.float32)
    A = ctx.from_numpy(a)
    O = ctx.array((64, 64), nc.FLOAT32, zeros=True)
    nc.check(nc.matmul(A, A, O), "async matmul")
    assert np.all(O.numpy() == 64)       # numc_array_data waits for the stream


def test_empty_shapes(ctx):
    for dt in [np.float32, np.int32, np.float64, np.int8]:
        z = gemm(ctx, np.zeros((5, 0), dt), np.zeros((0, 7), dt))
        assert z.shape == (5, 7) and np.all(z == 0)
        assert gemm(ctx, np.zeros((0, 4), dt), np.zeros((4, 3), dt)).shape == (0, 3)


def test_config1_f32_512(ctx):
    """f32 512^3, A, B uniform[-1,1) from the seeded generator, rel err <= 1e-5."""
    A = ctx.array((512, 512), nc.FLOAT32)
    B = ctx.array((512, 512), nc.FLOAT32)
    A.fill_random(nc.DIST_UNIFORM, 1, 0, -1, 1)
    B.fill_random(nc.DIST_UNIFORM, 1, 1, -1, 1)
    O = ctx.array((512, 512), nc.FLOAT32, zeros=True)
    nc.check(nc.lib().numc_matmul(A.h, B.h, O.h), "numc_matmul")
    a, b = A.numpy(), B.numpy()
    ref, absr = oracle.fref(a, b)
    m = oracle.err_metrics(O.numpy(), ref, absr)
    assert m[0] <= 1e-5 and m[1] <= 1e-5


def test_batched_uneven(ctx):
    rng = np.random.default_rng(15)
    shapes = [(17, 33, 65), (200, 16, 40), (1, 1000, 2), (128, 128, 128), (90, 250, 31)]
    As, Bs, Os, ref = [], [], [], []
    for M, K, N in shapes:
        a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
        As.append(ctx.from_numpy(a))
        Bs.append(ctx.from_numpy(b))
        Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
        ref.append((a, b))
    for method in [nc.METHOD_PLAIN, nc.METHOD_SPLIT_BF16]:
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, method), "batched")
        for (a, b), O in zip(ref, Os):
            assert cw_err(O.numpy(), a, b) <= 3 * 2.0 ** -16 + (a.shape[1] + 2) * U


# ---- batched (persistent kernel, cached HIP graph) ---------------------------------

def _logu(rng, lo, hi, n):
    return np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n))), lo, hi).astype(int)


def test_batched_many_uneven_and_graph_replay(ctx):
    rng = np.random.default_rng(16)
    n = 300
    Ms, Ns, Ks = _logu(rng, 1, 700, n), _logu(rng, 1, 700, n), _logu(rng, 1, 700, n)
    As, Bs, Os, ref = [], [], [], []
    for M, N, K in zip(Ms, Ns, Ks):
        a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
        As.append(ctx.from_numpy(a))
        Bs.append(ctx.from_numpy(b))
        Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
        ref.append((a, b))
    before = nc.lib().numc_ctx_batch_graph_replays(ctx.h)
    for rep in range(3):
        nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched")
        for (a, b), O in zip(ref, Os):
            assert cw_err(O.numpy(), a, b) <= (a.shape[1] + 2) * U
        # eager work between replays, and new host data for problem 0
        other = gemm(ctx, *ref[1])
        assert cw_err(other, *ref[1]) <= (ref[1][0].shape[1] + 2) * U
        a0 = rand_f32(rng, ref[0][0].shape)
        As[0].write(a0)
        ref[0] = (a0, ref[0][1])
    assert nc.lib().numc_ctx_batch_graph_replays(ctx.h) - before == 3


def test_batched_graph_cache_alternating_batches(ctx):
    """two batches called in turn both stay cached (a one-entry cache would
    re-plan and re-capture on every call); a fifth distinct batch evicts the
    least recently used one; profiling relaunches the kernels eagerly"""
    rng = np.random.default_rng(18)

    def make(n):
        As, Bs, Os, ref = [], [], [], []
        for M, N, K in zip(_logu(rng, 1, 300, n), _logu(rng, 1, 300, n), _logu(rng, 1, 300, n)):
            a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
            As.append(ctx.from_numpy(a))
            Bs.append(ctx.from_numpy(b))
            Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
            ref.append((a, b))
        return As, Bs, Os, ref

    batches = [make(n) for n in (5, 7, 3, 4, 6)]
    L = nc.lib()
    for As, Bs, Os, _ in batches[:2]:
        nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched")
    assert L.numc_ctx_batch_graphs_cached(ctx.h) == 2
    before = L.numc_ctx_batch_graph_replays(ctx.h)
    for rep in range(3):
        for As, Bs, Os, ref in batches[:2]:
            for O in Os:
                O.write(np.zeros(O.shape, np.float32))
            nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched")
            for (a, b), O in zip(ref, Os):
                assert cw_err(O.numpy(), a, b) <= (a.shape[1] + 2) * U
    # every call of both batches after the first was a replay of its own graph
    assert L.numc_ctx_batch_graph_replays(ctx.h) - before == 6
    for As, Bs, Os, _ in batches[2:]:
        nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched")
    assert L.numc_ctx_batch_graphs_cached(ctx.h) == 4
    # profiling: eager launches of the cached batch, each kernel timed
    As, Bs, Os, ref = batches[4]
    ctx.profile(True)
    nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched (profiled)")
    ms, n = ctx.profile_read(3, nc.FAMILY_F32)
    ctx.profile(False)
    assert n == 1 and ms > 0
    for (a, b), O in zip(ref, Os):
        assert cw_err(O.numpy(), a, b) <= (a.shape[1] + 2) * U


def _set_schedule(fam, direct, thresh):
    nc.check(nc.lib().numc_matmul_batched_set_schedule(fam, direct, thresh), "set schedule")


@pytest.mark.parametrize("direct", [1, 0], ids=["direct", "packed"])
@pytest.mark.parametrize("thresh", [0, 64, 1 << 30], ids=["all_large", "t64", "all_small"])
@pytest.mark.parametrize("fam,method,extra", [(nc.FAMILY_F32, nc.METHOD_PLAIN, 0.0),
                                              (nc.FAMILY_SPLIT_BF16, nc.METHOD_SPLIT_BF16, 3 * 2.0 ** -16)],
                         ids=["f32", "split_bf16"])
def test_batched_f32_schedules_views(ctx, direct, thresh, fam, method, extra):
    """both batched kernels and every tile class, FP32 and split-BF16, on
    uneven problems whose operands are row-major, stored transposed, or start
    one element past a 16-byte boundary; results against the oracle (split:
    plus the split's representation bound)"""
    rng = np.random.default_rng(19)
    before = nc.batched_schedule(fam)
    _set_schedule(fam, direct, thresh)
    try:
        As, Bs, Os, ref = [], [], [], []
        for t, (M, N, K) in enumerate(zip(_logu(rng, 1, 400, 24), _logu(rng, 1, 400, 24), _logu(rng, 1, 600, 24))):
            a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
            if t % 3 == 0:
                A, B = ctx.from_numpy(a), ctx.from_numpy(b)
            elif t % 3 == 1:   # stored transposed
                A, B = ctx.from_numpy(a.T.copy()).transpose(), ctx.from_numpy(b.T.copy()).transpose()
            else:              # one element past the aligned start, odd row strides
                pa = np.zeros((M, K + 1), np.float32)
                pa[:, 1:] = a
                pb = np.zeros((K, N + 1), np.float32)
                pb[:, 1:] = b
                A, B = ctx.from_numpy(pa).slice(1, 1, K + 1), ctx.from_numpy(pb).slice(1, 1, N + 1)
            As.append(A)
            Bs.append(B)
            Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
            ref.append((a, b))
        for rep in range(2):   # capture, then replay
            nc.check(nc.matmul_batched(ctx, As, Bs, Os, method), "batched")
            for (a, b), O in zip(ref, Os):
                assert cw_err(O.numpy(), a, b) <= extra + (a.shape[1] + 2) * U
    finally:
        _set_schedule(fam, *before)


@pytest.mark.parametrize("odt,flags", [(np.int32, nc.OPT_WIDEN), (np.int8, 0)], ids=["widen", "wrap"])
def test_batched_int8_direct_b_transposed(ctx, odt, flags):
    """the direct batched kernel on int8 operands with k-contiguous rows (B
    stored transposed), both tile classes; exact against the oracle"""
    rng = np.random.default_rng(20)
    before = nc.batched_schedule(nc.FAMILY_I8)
    _set_schedule(nc.FAMILY_I8, 1, 64)
    try:
        As, Bs, Os, ref = [], [], [], []
        for M, K, N in [(40, 300, 70), (128, 64, 128), (3, 5, 7), (260, 1000, 90), (65, 33, 200)]:
            a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
            b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
            As.append(ctx.from_numpy(a))
            Bs.append(ctx.from_numpy(b.T.copy()).transpose())
            Os.append(ctx.array((M, N), nc.DTYPE_OF_NP[np.dtype(odt)], zeros=True))
            ref.append(oracle.matmul_int(a, b, odt))
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, flags=flags), "batched int8")
        for want, O in zip(ref, Os):
            assert np.array_equal(O.numpy(), want)
    finally:
        _set_schedule(nc.FAMILY_I8, *before)


def test_split_direct_batch_bits(ctx):
    """the direct batched kernel splits float32 into bf16 hi/lo in registers
    (hardware round to nearest even) and adds lo*hi, hi*lo, hi*hi per 16-k
    fragment in k order, as the packed kernel does after the pack's software
    split: both give the same bits, ragged K included"""
    rng = np.random.default_rng(23)
    before = nc.batched_schedule(nc.FAMILY_SPLIT_BF16)
    _set_schedule(nc.FAMILY_SPLIT_BF16, 1, 32)
    try:
        shapes = [(40, 70, 300), (128, 64, 128), (3, 7, 5), (65, 200, 33), (96, 96, 1000)]
        As, Bs, Os, Ps = [], [], [], []
        for t, (M, N, K) in enumerate(shapes):
            a = rand_f32(rng, (M, K)) * np.float32(2.0 ** (t - 2))
            b = rand_f32(rng, (K, N))
            As.append(ctx.from_numpy(a))
            Bs.append(ctx.from_numpy(b))
            Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
            Ps.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, nc.METHOD_SPLIT_BF16), "batched split")
        packed = [i for i in range(nc.config_count(nc.FAMILY_SPLIT_BF16))
                  if not nc.config_get(nc.FAMILY_SPLIT_BF16, i)["direct"]
                  and nc.config_get(nc.FAMILY_SPLIT_BF16, i)["split_k"] == 1]
        for A, B, O, P in zip(As, Bs, Os, Ps):
            nc.check(nc.matmul(A, B, P, nc.METHOD_SPLIT_BF16, packed[0]), "packed split")
            assert np.array_equal(O.numpy().view(np.uint32), P.numpy().view(np.uint32))
    finally:
        _set_schedule(nc.FAMILY_SPLIT_BF16, *before)


def test_batched_tune_picks_a_schedule(ctx):
    """the tuner times the direct and packed kernels and keeps one; results
    under the kept schedule stay exact"""
    rng = np.random.default_rng(21)
    before = nc.batched_schedule(nc.FAMILY_F32)
    try:
        As, Bs, Os, ref = [], [], [], []
        for M, N, K in zip(_logu(rng, 16, 300, 40), _logu(rng, 16, 300, 40), _logu(rng, 16, 300, 40)):
            a, b = rand_f32(rng, (M, K)), rand_f32(rng, (K, N))
            As.append(ctx.from_numpy(a))
            Bs.append(ctx.from_numpy(b))
            Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
            ref.append((a, b))
        t = nc.matmul_batched_tune(ctx, As, Bs, Os, nc.METHOD_PLAIN, 0, 2)
        assert t >= 0
        d, t2 = nc.batched_schedule(nc.FAMILY_F32)
        assert t2 == t and d in (0, 1)
        nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched")
        for (a, b), O in zip(ref, Os):
            assert cw_err(O.numpy(), a, b) <= (a.shape[1] + 2) * U
    finally:
        _set_schedule(nc.FAMILY_F32, *before)


def test_batched_int8_widen_and_wrap(ctx):
    rng = np.random.default_rng(17)
    shapes = [(40, 300, 70), (128, 64, 128), (3, 5, 7), (260, 1000, 90)]
    for odt, flags in [(np.int32, nc.OPT_WIDEN), (np.int8, 0)]:
        As, Bs, Os, ref = [], [], [], []
        for M, K, N in shapes:
            a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
            b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
            As.append(ctx.from_numpy(a))
            Bs.append(ctx.from_numpy(b))
            Os.append(ctx.array((M, N), nc.DTYPE_OF_NP[np.dtype(odt)], zeros=True))
            ref.append(oracle.matmul_int(a, b, odt))
        nc.check(nc.matmul_batched(ctx, As, Bs, Os, flags=flags), "batched int8")
        for want, O in zip(ref, Os):
            assert np.array_equal(O.numpy(), want)


def test_batched_mixed_int8_uint8_widen(ctx):
    """int8 and uint8 problems widened to int32 in one batch: uint8 values
    above 127 must not be read as int8."""
    rng = np.random.default_rng(19)
    As, Bs, Os, ref = [], [], [], []
    for i, (M, K, N) in enumerate([(40, 70, 30), (64, 96, 80), (33, 250, 17)]):
        dt = np.int8 if i % 2 == 0 else np.uint8
        info = np.iinfo(dt)
        a = rng.integers(info.min, info.max, (M, K), dtype=dt, endpoint=True)
        b = rng.integers(info.min, info.max, (K, N), dtype=dt, endpoint=True)
        As.append(ctx.from_numpy(a))
        Bs.append(ctx.from_numpy(b))
        Os.append(ctx.array((M, N), nc.INT32, zeros=True))
        ref.append(a.astype(np.int64) @ b.astype(np.int64))
    nc.check(nc.matmul_batched(ctx, As, Bs, Os, flags=nc.OPT_WIDEN), "batched mixed")
    for want, O in zip(ref, Os):
        assert np.array_equal(O.numpy(), want.astype(np.int32))


@pytest.mark.parametrize("method", [nc.METHOD_PLAIN, nc.METHOD_SPLIT_BF16])
def test_batched_strided_operands(ctx, method):
    """Every pack walk of the batched kernel: transposed A (row-contiguous),
    transposed B (k-contiguous B^T), column-stepped views (lane slots)."""
    rng = np.random.default_rng(20)
    As, Bs, Os, ref = [], [], [], []
    for M, K, N in [(70, 90, 50), (130, 40, 200), (17, 300, 33)]:
        at = rng.uniform(-1, 1, (K, M)).astype(np.float32)     # A stored transposed
        bt = rng.uniform(-1, 1, (N, K)).astype(np.float32)     # B stored transposed
        bs = rng.uniform(-1, 1, (K, 2 * N)).astype(np.float32)  # B with column step 2
        A = ctx.from_numpy(at).transpose()
        B1 = ctx.from_numpy(bt).transpose()
        B2 = ctx.from_numpy(bs).slice(1, 0, 2 * N, 2)
        for B, b in ((B1, bt.T), (B2, bs[:, ::2])):
            As.append(A)
            Bs.append(B)
            Os.append(ctx.array((M, N), nc.FLOAT32, zeros=True))
            ref.append((np.ascontiguousarray(at.T), np.ascontiguousarray(b)))
    nc.check(nc.matmul_batched(ctx, As, Bs, Os, method), "batched strided")
    for (a, b), O in zip(ref, Os):
        K = a.shape[1]
        bound = (K + 2) * U if method == nc.METHOD_PLAIN else 3 * 2.0 ** -16 + (K + 2) * U
        assert cw_err(O.numpy(), a, b) <= bound


def test_batched_tune_keeps_results(ctx):
    """the schedule autotuner returns one of its thresholds and the batch
    gives the same (bit-exact) integer results under the tuned schedule"""
    rng = np.random.default_rng(22)
    As, Bs, Os, ref = [], [], [], []
    for M, K, N in [(40, 300, 70), (128, 64, 200), (3, 5, 7), (260, 100, 90), (90, 77, 65)]:
        a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
        b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
        As.append(ctx.from_numpy(a))
        Bs.append(ctx.from_numpy(b))
        Os.append(ctx.array((M, N), nc.INT32, zeros=True))
        ref.append(a.astype(np.int64) @ b.astype(np.int64))
    t = nc.matmul_batched_tune(ctx, As, Bs, Os, flags=nc.OPT_WIDEN, reps=2)
    assert t in (0, 32, 48, 64, 96, 128, 192, 256, 1 << 30)
    for O, want in zip(Os, ref):
        O.write(np.full(want.shape, -1, np.int32))   # results must be recomputed
    nc.check(nc.matmul_batched(ctx, As, Bs, Os, flags=nc.OPT_WIDEN), "batched after tune")
    for want, O in zip(ref, Os):
        assert np.array_equal(O.numpy(), want.astype(np.int32))


def test_batched_fallback_dtypes_and_empty(ctx):
    rng = np.random.default_rng(18)
    As, Bs, Os, ref = [], [], [], []
    for M, K, N in [(30, 40, 50), (0, 4, 3), (5, 0, 6)]:
        a, b = rng.standard_normal((M, K)), rng.standard_normal((K, N))
        As.append(ctx.from_numpy(a))
        Bs.append(ctx.from_numpy(b))
        Os.append(ctx.from_numpy(np.full((M, N), 3.0)))
        ref.append((a, b))
    nc.check(nc.matmul_batched(ctx, As, Bs, Os), "batched f64")
    assert cw_err(Os[0].numpy(), *ref[0]) <= 42 * U64
    assert Os[1].numpy().shape == (0, 3)
    assert np.all(Os[2].numpy() == 0)


# ---- A read in place (no pack of A) --------------------------------------------

@pytest.mark.parametrize("name", ["float32", "bfloat16", "float16", "int8"])
@pytest.mark.parametrize("offset", [0, 1], ids=["aligned", "offset"])
def test_a_read_in_place(ctx, name, offset):
    """float32 and int8: an A whose rows are k-contiguous and 16-byte aligned,
    with K a multiple of the configuration's k-chunk, is copied into LDS by the
    GEMM itself (no pack of A): every configuration, M ragged against every
    tile, A a view whose rows sit 320 elements apart.  offset 1 starts the view
    one element in, so its rows are unaligned and A is packed: both paths give
    the same bits.  bfloat16 and float16 always pack A; they run the same cases.
    Floats hold small integers, so every sum is exact in the FP32 accumulator
    and the one rounding to the output dtype is known (round to nearest even)."""
    rng = np.random.default_rng(700 + offset)
    M, K, N, W = 300, 256, 200, 320
    if name == "int8":
        fam, dt = nc.FAMILY_I8, nc.INT8
        base = rng.integers(-128, 127, (M, W), dtype=np.int8, endpoint=True)
        b = rng.integers(-128, 127, (K, N), dtype=np.int8, endpoint=True)
        a = base[:, offset:offset + K]
        want = oracle.matmul_int(np.ascontiguousarray(a), b)
        A, B = ctx.from_numpy(base), ctx.from_numpy(b)
    else:
        fam = {"float32": nc.FAMILY_F32, "bfloat16": nc.FAMILY_BF16, "float16": nc.FAMILY_F16}[name]
        dt = {"float32": nc.FLOAT32, "bfloat16": nc.BFLOAT16, "float16": nc.FLOAT16}[name]
        base_i = rng.integers(-4, 4, (M, W), endpoint=True)
        b_i = rng.integers(-4, 4, (K, N), endpoint=True)
        exact = (base_i[:, offset:offset + K] @ b_i).astype(np.float32)  # |sum| <= 4096: exact
        if name == "float32":
            A, B = ctx.from_numpy(base_i.astype(np.float32)), ctx.from_numpy(b_i.astype(np.float32))
            want = exact
        elif name == "float16":
            A, B = ctx.from_numpy(base_i.astype(np.float16)), ctx.from_numpy(b_i.astype(np.float16))
            want = exact.astype(np.float16)
        else:
            A = ctx.from_numpy(nc.bf16_bits(base_i.astype(np.float32)), dtype=nc.BFLOAT16)
            B = ctx.from_numpy(nc.bf16_bits(b_i.astype(np.float32)), dtype=nc.BFLOAT16)
            want = nc.bf16_bits(exact)
    Av = A.slice(1, offset, offset + K)
    for cfg in range(nc.config_count(fam)):
        O = ctx.array((M, N), dt, zeros=True)
        nc.check(nc.matmul(Av, B, O, nc.METHOD_AUTO, cfg), "matmul A in place")
        got = O.numpy()
        assert np.array_equal(got, want), (name, offset, cfg)


@pytest.mark.parametrize("name", ["float32", "int8"])
@pytest.mark.parametrize("a_off,b_off", [(0, 0), (1, 0), (0, 1), (1, 1)],
                         ids=["both-aligned", "a-offset", "b-offset", "both-offset"])
def test_b_read_in_place(ctx, name, a_off, b_off):
    """B stored transposed (a torch-style (N, K) weight passed as its transposed
    view): the rows of B^T are k-contiguous, and when they are 16-byte aligned
    with K a multiple of the k-chunk the float32 GEMM copies them in place
    too, with or without A in place (int8 reads only A in place; it runs the
    same cases).  An offset of one element makes that operand unaligned, so it
    is packed.  Every configuration; exact sums."""
    rng = np.random.default_rng(800 + 2 * a_off + b_off)
    M, K, N, W = 300, 256, 200, 320
    if name == "int8":
        fam, dt = nc.FAMILY_I8, nc.INT8
        base_a = rng.integers(-128, 127, (M, W), dtype=np.int8, endpoint=True)
        base_w = rng.integers(-128, 127, (N, W), dtype=np.int8, endpoint=True)
        a, w = base_a[:, a_off:a_off + K], base_w[:, b_off:b_off + K]
        want = oracle.matmul_int(np.ascontiguousarray(a), np.ascontiguousarray(w.T))
    else:
        fam, dt = nc.FAMILY_F32, nc.FLOAT32
        base_a = rng.integers(-4, 4, (M, W), endpoint=True).astype(np.float32)
        base_w = rng.integers(-4, 4, (N, W), endpoint=True).astype(np.float32)
        a, w = base_a[:, a_off:a_off + K], base_w[:, b_off:b_off + K]
        want = (a.astype(np.int64) @ w.T.astype(np.int64)).astype(np.float32)  # exact
    Av = ctx.from_numpy(base_a).slice(1, a_off, a_off + K)
    Bv = ctx.from_numpy(base_w).slice(1, b_off, b_off + K).transpose()   # (K, N), rows of B^T contiguous
    for cfg in range(nc.config_count(fam)):
        O = ctx.array((M, N), dt, zeros=True)
        nc.check(nc.matmul(Av, Bv, O, nc.METHOD_AUTO, cfg), "matmul B in place")
        assert np.array_equal(O.numpy(), want), (name, a_off, b_off, cfg)


@pytest.mark.parametrize("cfg", range(nc.config_count(nc.FAMILY_I8)))
def test_int8_a_in_place_b_any_strides(ctx, cfg):
    """int8 with A read in place and B a view strided in both dimensions (every
    other column): B goes through the generic one-plane digit pack, which
    must not touch plane masks when the call has none."""
    rng = np.random.default_rng(900 + cfg)
    M, K, N = 130, 256, 70
    a = rng.integers(-128, 127, (M, K), dtype=np.int8, endpoint=True)
    base_b = rng.integers(-128, 127, (K, 2 * N), dtype=np.int8, endpoint=True)
    B = ctx.from_numpy(base_b).slice(1, 0, 2 * N, 2)                # (K, N), strides (2N, 2)
    for odt, flags in [(np.int8, 0), (np.int32, nc.OPT_WIDEN)]:
        O = ctx.array((M, N), nc.DTYPE_OF_NP[np.dtype(odt)], zeros=True)
        nc.check(nc.matmul(ctx.from_numpy(a), B, O, nc.METHOD_AUTO, cfg, flags), "int8 strided B")
        want = oracle.matmul_int(a, np.ascontiguousarray(base_b[:, ::2]), odt)
        assert np.array_equal(O.numpy(), want), (cfg, odt)


def test_device_copy(ctx):
    """numc_device_copy: exact bytes for sizes below, at and past its grid
    stride (2048 blocks x 256 threads x 16 bytes), and its argument checks"""
    rng = np.random.default_rng(41)
    for nbytes in [16, 48, 4096 * 16 + 32, (1 << 23) + 16 * 3, 2048 * 256 * 16 * 4 + 16 * 5]:
        x = rng.integers(0, 256, nbytes, dtype=np.uint8)
        X = ctx.from_numpy(x)
        Y = ctx.array((nbytes,), nc.UINT8, zeros=True)
        X.to_device()
        Y.to_device()
        src, dst = nc.lib().numc_array_device_data(X.h), nc.lib().numc_array_device_data(Y.h)
        nc.check(nc.lib().numc_device_copy(ctx.h, dst, src, nbytes), "numc_device_copy")
        Y.to_host()
        assert np.array_equal(Y.numpy(), x), nbytes
    assert nc.lib().numc_device_copy(ctx.h, dst, src, 24) == nc.ERR_VALUE
    assert nc.lib().numc_device_copy(ctx.h, None, src, 16) == nc.ERR_NULL
    # refused before any launch: host memory, misalignment, a range past the allocation
    H = ctx.array((64,), nc.UINT8, zeros=True)        # host-current: its data pointer is (pinned) host memory
    assert nc.lib().numc_device_copy(ctx.h, nc.lib().numc_array_data(H.h), src, 16) == nc.ERR_VALUE
    assert "not device memory" in nc.last_error()[1]
    hostbuf = np.zeros(64, np.uint8)
    assert nc.lib().numc_device_copy(ctx.h, dst, hostbuf.ctypes.data, 16) == nc.ERR_VALUE
    assert nc.lib().numc_device_copy(ctx.h, dst + 4, src, 16) == nc.ERR_VALUE
    assert nc.lib().numc_device_copy(ctx.h, dst, src, 1 << 40) == nc.ERR_VALUE
    assert "past its allocation" in nc.last_error()[1]
    nc.check(nc.lib().numc_device_copy(ctx.h, dst, src, 16), "numc_device_copy after refusals")
