"""numc_amd_gemm_noalloc (the entry point that allocates nothing; its packed
kernels' workspace is the library's per-GPU scratch array) against
numc_amd_gemm on the same device operands, through the C ABI: the same
configuration gives the same bits (the kernels are deterministic), two calls
in a row with the second smaller (the scratch keeps the first call's
contents), and results under the selector's own choice against the oracle.
"""
import ctypes as C

import numpy as np
import pytest

import numc_amd as nc
import oracle

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


def _dev(ctx, x, dt=None):
    A = ctx.from_numpy(x, dt)
    A.to_device()
    return A


def _call(fn, ctx, A, B, O, dt, odt, M, N, K, es, eo, method, config, flags):
    opts = nc.MatmulOpts(method, config, flags | nc.OPT_DEVICE_POINTERS)
    L = nc.lib()
    pa, pb, po = (L.numc_array_device_data(x.h) for x in (A, B, O))
    head = (nc.lib().numc_ctx_stream(ctx.h),) if fn == "noalloc" else (ctx.h,)
    f = L.numc_amd_gemm_noalloc if fn == "noalloc" else L.numc_amd_gemm
    return f(*head, dt, odt, M, N, K, pa, K * es, es, pb, N * es, es, po, N * eo, eo, C.byref(opts))


CASES = [  # (name, input dtype, output dtype, numpy in, numpy out, method, flags, family)
    ("f32", nc.FLOAT32, nc.FLOAT32, np.float32, np.float32, nc.METHOD_PLAIN, 0, nc.FAMILY_F32),
    ("split_bf16", nc.FLOAT32, nc.FLOAT32, np.float32, np.float32, nc.METHOD_SPLIT_BF16, 0, nc.FAMILY_SPLIT_BF16),
    ("fp8", nc.FLOAT32, nc.FLOAT32, np.float32, np.float32, nc.METHOD_FP8_SCALED, 0, nc.FAMILY_FP8),
    ("f64", nc.FLOAT64, nc.FLOAT64, np.float64, np.float64, nc.METHOD_AUTO, 0, nc.FAMILY_F64),
    ("int32", nc.INT32, nc.INT32, np.int32, np.int32, nc.METHOD_AUTO, 0, nc.FAMILY_I8),
    ("int8_widen", nc.INT8, nc.INT32, np.int8, np.int32, nc.METHOD_AUTO, nc.OPT_WIDEN, nc.FAMILY_I8),
    ("bf16", nc.BFLOAT16, nc.BFLOAT16, np.uint16, np.uint16, nc.METHOD_AUTO, 0, nc.FAMILY_BF16),
    ("f16", nc.FLOAT16, nc.FLOAT16, np.float16, np.float16, nc.METHOD_AUTO, 0, nc.FAMILY_F16),
]


def _inputs(rng, npin, dt, M, N, K):
    if np.issubdtype(npin, np.integer) and dt != nc.BFLOAT16:
        info = np.iinfo(npin)
        return (rng.integers(info.min, info.max, (M, K), dtype=npin, endpoint=True),
                rng.integers(info.min, info.max, (K, N), dtype=npin, endpoint=True))
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    if dt == nc.BFLOAT16:
        return nc.bf16_bits(a), nc.bf16_bits(b)
    return a.astype(npin), b.astype(npin)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_numc_amd_gemm(ctx, case):
    name, dt, odt, npin, npout, method, flags, fam = case
    rng = np.random.default_rng(123)
    es, eo = np.dtype(npin).itemsize, np.dtype(npout).itemsize
    # every packed configuration (these use the scratch) and the selector's choice
    cfgs = [c for c in range(nc.config_count(fam)) if not nc.config_get(fam, c)["direct"]] + [-1]
    for cfg in cfgs:
        for M, N, K in [(700, 900, 1000), (130, 70, 257)]:   # the second call is smaller
            a, b = _inputs(rng, npin, dt, M, N, K)
            A, B = _dev(ctx, a, dt if dt == nc.BFLOAT16 else None), _dev(ctx, b, dt if dt == nc.BFLOAT16 else None)
            O1 = ctx.array((M, N), odt, zeros=True)
            O2 = ctx.array((M, N), odt, zeros=True)
            O1.to_device()
            O2.to_device()
            print("noalloc %s config %d %dx%dx%d" % (name, cfg, M, N, K), flush=True)
            nc.check(_call("noalloc", ctx, A, B, O1, dt, odt, M, N, K, es, eo, method, cfg, flags), "noalloc")
            ctx.synchronize()
            got = O1.numpy()
            if cfg >= 0:
                nc.check(_call("gemm", ctx, A, B, O2, dt, odt, M, N, K, es, eo, method, cfg, flags), "numc_amd_gemm")
                assert np.array_equal(got.view(np.uint8), O2.numpy().view(np.uint8)), (name, cfg, M, N, K)
            elif np.issubdtype(npout, np.integer) and dt != nc.BFLOAT16:
                assert np.array_equal(got, oracle.matmul_int(a, b, npout))
            elif dt == nc.BFLOAT16:
                av, bv = nc.bf16_value(a).astype(np.float64), nc.bf16_value(b).astype(np.float64)
                ref, den = av @ bv, np.abs(av) @ np.abs(bv)
                assert np.all(np.abs(nc.bf16_value(got) - ref) <= (K + 3) * U * den + 2.0 ** -8 * np.abs(ref))
            elif dt == nc.FLOAT16:  # float16 values are exact in float32; the store rounds to 11 bits
                ref, absr = oracle.fref(a.astype(np.float32), b.astype(np.float32))
                assert oracle.err_metrics(got.astype(np.float32), ref, absr)[0] <= (K + 2) * U + 2.0 ** -11
            elif method == nc.METHOD_FP8_SCALED:
                qa, qb = oracle.fp8_quantized_inputs(a, b)
                ref, absr = oracle.fref(qa, qb)
                assert oracle.err_metrics(got, ref, absr)[0] <= -(-K // 64) * 2.0 ** -11 + (K + 2) * U
            else:
                ref, absr = oracle.fref(a, b)
                bound = (K + 2) * (2.0 ** -53 if npin == np.float64 else U) + \
                    (3 * 2.0 ** -16 if method == nc.METHOD_SPLIT_BF16 else 0)
                assert oracle.err_metrics(got, ref, absr)[0] <= bound


def test_noalloc_rejects_workspace_beyond_scratch(ctx):
    """int32 operands of 4096 x 4096 pack four digit planes each (256 MiB),
    beyond the 64 MiB scratch: refused before any launch, with the reason"""
    n = 4096
    A = ctx.array((n, n), nc.INT32)
    A.to_device()
    O = ctx.array((n, n), nc.INT32)
    O.to_device()
    code = _call("noalloc", ctx, A, A, O, nc.INT32, nc.INT32, n, n, n, 4, 4, nc.METHOD_AUTO, -1, 0)
    assert code == nc.ERR_VALUE
    assert "scratch" in nc.last_error()[1]
    code = _call("noalloc", ctx, A, A, O, nc.INT32, nc.INT32, n, n, n, 4, 4, nc.METHOD_AUTO, 0, 0)
    assert code == nc.ERR_VALUE


def test_noalloc_null_and_types(ctx):
    opts = nc.MatmulOpts(0, -1, 0)
    L = nc.lib()
    assert L.numc_amd_gemm_noalloc(None, nc.FLOAT32, nc.FLOAT32, 4, 4, 4, None, 16, 4, None, 16, 4, None, 16, 4,
                                   C.byref(opts)) == nc.ERR_NULL
    x = ctx.array((4, 4), nc.FLOAT32)
    x.to_device()
    p = L.numc_array_device_data(x.h)
    assert L.numc_amd_gemm_noalloc(None, nc.FLOAT32, nc.FLOAT64, 4, 4, 4, p, 16, 4, p, 16, 4, p, 32, 8,
                                   C.byref(opts)) == nc.ERR_TYPE
    assert L.numc_amd_gemm_noalloc(None, nc.INT64, nc.INT64, 4, 4, 4, p, 32, 8, p, 32, 8, p, 32, 8,
                                   C.byref(opts)) == nc.ERR_TYPE
