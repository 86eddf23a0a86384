"""The benchmark configurations at their full sizes, on the GPU.

Full-size parity uses size-independent properties plus exact spot checks:
  * integer GEMMs: every column sum and every row sum of C equals, mod 2^w,
    (1^T A) B and A (B 1) computed exactly with numpy int64 arithmetic, and 64
    whole rows match the oracle bit for bit;
  * float GEMMs: 4096 sampled entries against the FP64 / double-double oracle
    under the componentwise bound of the method.
"""
import numpy as np
import pytest

import numc_amd as nc
import oracle

pytestmark = pytest.mark.gpu
U = 2.0 ** -24
U64 = 2.0 ** -53


def seeded(ctx, shape, dt, dist, seed, stream, lo=-1.0, hi=1.0):
    A = ctx.array(shape, dt)
    A.fill_random(dist, seed, stream, lo, hi)
    return A


def run(ctx, A, B, out_dt, method=0, flags=0):
    M, N = A.shape[0], B.shape[1]
    O = ctx.array((M, N), out_dt, zeros=True)
    nc.check(nc.matmul(A, B, O, method, -1, flags), "matmul")
    return O.numpy()


def int_checks(a, b, c, rows=64, seed=0):
    w = c.dtype.itemsize * 8
    mask = (1 << w) - 1 if w < 64 else None
    a64 = a.astype(np.int64)
    b64 = b.astype(np.int64)
    # (1^T A) B mod 2^w against the column sums of C
    colsum_a = a64.sum(axis=0)                    # exact (|sum| < 2^63)
    want_cols = (colsum_a @ b64) & mask
    got_cols = c.astype(np.int64).sum(axis=0) & mask
    assert np.array_equal(got_cols, want_cols), "column checksums differ"
    rowsum_b = b64.sum(axis=1)
    want_rows = (a64 @ rowsum_b) & mask
    got_rows = c.astype(np.int64).sum(axis=1) & mask
    assert np.array_equal(got_rows, want_rows), "row checksums differ"
    rng = np.random.default_rng(seed)
    sel = rng.choice(a.shape[0], rows, replace=False)
    want = oracle.matmul_int(np.ascontiguousarray(a[sel]), b, c.dtype)
    assert np.array_equal(c[sel], want), "sampled rows differ"


def float_sample_err(a, b, c, n=4096, seed=0):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, a.shape[0], n)
    cols = rng.integers(0, b.shape[1], n)
    ref, absr = oracle.fref(a, b, rows, cols)
    d = np.abs(c[rows, cols].astype(np.float64) - ref)
    return float(np.max(d / absr))


def test_config2_f64_2048(ctx):
    A = seeded(ctx, (2048, 2048), nc.FLOAT64, nc.DIST_NORMAL, 2, 0)
    B = seeded(ctx, (2048, 2048), nc.FLOAT64, nc.DIST_NORMAL, 2, 1)
    c = run(ctx, A, B, nc.FLOAT64)
    assert float_sample_err(A.numpy(), B.numpy(), c) <= 2050 * U64


def test_config2_int32_2048(ctx):
    A = seeded(ctx, (2048, 2048), nc.INT32, nc.DIST_INT, 2, 2, -100, 100)
    B = seeded(ctx, (2048, 2048), nc.INT32, nc.DIST_INT, 2, 3, -100, 100)
    int_checks(A.numpy(), B.numpy(), run(ctx, A, B, nc.INT32))


def test_config2_int8_4096_wraparound(ctx):
    A = seeded(ctx, (4096, 4096), nc.INT8, nc.DIST_INT, 2, 4, -128, 127)
    B = seeded(ctx, (4096, 4096), nc.INT8, nc.DIST_INT, 2, 5, -128, 127)
    int_checks(A.numpy(), B.numpy(), run(ctx, A, B, nc.INT8), rows=16)


def test_int32_full_range_2048(ctx):
    """all four digit planes in use"""
    A = seeded(ctx, (1024, 2048), nc.INT32, nc.DIST_INT, 7, 0, -2 ** 31, 2 ** 31 - 1)
    B = seeded(ctx, (2048, 512), nc.INT32, nc.DIST_INT, 7, 1, -2 ** 31, 2 ** 31 - 1)
    int_checks(A.numpy(), B.numpy(), run(ctx, A, B, nc.INT32), rows=8)


@pytest.mark.parametrize("method,bound", [(nc.METHOD_PLAIN, 8194 * U),
                                          (nc.METHOD_SPLIT_BF16, 3 * 2.0 ** -16 + 8194 * U)])
def test_config3_f32_8192_sampled(ctx, method, bound):
    A = seeded(ctx, (8192, 8192), nc.FLOAT32, nc.DIST_NORMAL, 3, 0)
    B = seeded(ctx, (8192, 8192), nc.FLOAT32, nc.DIST_NORMAL, 3, 1)
    c = run(ctx, A, B, nc.FLOAT32, method)
    assert float_sample_err(A.numpy(), B.numpy(), c, n=2048) <= bound


def test_bench_error_metric_matches_oracle(ctx):
    """bench.py's full-matrix error (from the library's FP64 GEMM) agrees with
    the CPU oracle's componentwise error on a size the oracle finishes."""
    import bench
    n = 384
    A = seeded(ctx, (n, n), nc.FLOAT32, nc.DIST_NORMAL, 5, 0)
    B = seeded(ctx, (n, n), nc.FLOAT32, nc.DIST_NORMAL, 5, 1)
    a, b = A.numpy(), B.numpy()
    outs = {"f32": run(ctx, A, B, nc.FLOAT32, nc.METHOD_PLAIN),
            "split_bf16": run(ctx, A, B, nc.FLOAT32, nc.METHOD_SPLIT_BF16)}
    got = bench.errors_vs_fp64(nc, ctx, a, b, outs, n)
    ref, absr = oracle.fref(a, b)
    for name, c in outs.items():
        want = oracle.err_metrics(c, ref, absr)
        assert want[0] > 0
        assert got[name]["componentwise"] == pytest.approx(want[0], rel=1e-6), name
        assert got[name]["normwise"] == pytest.approx(want[1], rel=1e-6), name
        assert got[name]["elementwise_rel"] == pytest.approx(want[2], rel=1e-6), name
