"""ctypes wrapper of oracle/libnumc_oracle.so -- the checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  Provenance of each
function is in numc_oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnumc_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libnumc_ref.so")
REF_BENCH = os.path.join(HERE, "_ref", "numc_ref_bench")

NP_DTYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32,
             np.uint64, np.float32, np.float64]
CODE_OF = {np.dtype(t): i for i, t in enumerate(NP_DTYPES)}

_P = C.c_void_p
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_fill_uniform.argtypes = [_P, C.c_int, C.c_size_t, C.c_uint64, C.c_uint64, C.c_double, C.c_double]
        L.orc_fill_normal.argtypes = [_P, C.c_int, C.c_size_t, C.c_uint64, C.c_uint64]
        L.orc_fill_int.argtypes = [_P, C.c_int, C.c_size_t, C.c_uint64, C.c_uint64, C.c_int64, C.c_int64]
        L.orc_matmul_int.argtypes = [C.c_int, C.c_int, _P, _P, _P] + [C.c_size_t] * 3 + [C.c_ssize_t] * 6
        L.orc_matmul_fref.argtypes = [C.c_int, _P, _P] + [C.c_size_t] * 3 + [C.c_ssize_t] * 4 + \
            [_P, _P, C.c_size_t, _P, _P]
        L.orc_err_metrics.argtypes = [C.c_int, _P, C.c_size_t, C.c_size_t, C.c_ssize_t, C.c_ssize_t, _P, _P,
                                      C.c_double, _P]
        L.orc_bf16_rne.argtypes = [C.c_float]
        L.orc_bf16_rne.restype = C.c_uint16
        L.orc_split_bf16.argtypes = [_P, C.c_size_t, _P, _P]
        L.orc_e4m3_rne.argtypes = [C.c_float]
        L.orc_e4m3_rne.restype = C.c_uint8
        L.orc_e4m3_rne_n.argtypes = [_P, C.c_size_t, _P]
        L.orc_e4m3_to_f32.argtypes = [C.c_uint8]
        L.orc_e4m3_to_f32.restype = C.c_float
        L.orc_fp8_scale_exp.argtypes = [C.c_float]
        L.orc_fp8_scale_exp.restype = C.c_int
        L.orc_check_matmul.argtypes = [C.c_int] * 9 + [_P, _P, _P]
        L.orc_rng_block.argtypes = [_P, C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_rng_next.argtypes = [_P]
        L.orc_rng_next.restype = C.c_uint64
        _L = L
    return _L


def _ptr(x):
    return x.ctypes.data_as(C.c_void_p)


def _estrides(x):
    return [s // x.itemsize for s in x.strides]


# ---- generator ---------------------------------------------------------------

def fill_uniform(n, dtype, seed, stream, lo=-1.0, hi=1.0):
    x = np.empty(n, dtype)
    lib().orc_fill_uniform(_ptr(x), CODE_OF[x.dtype], n, seed, stream, lo, hi)
    return x


def fill_normal(n, dtype, seed, stream):
    x = np.empty(n, dtype)
    lib().orc_fill_normal(_ptr(x), CODE_OF[x.dtype], n, seed, stream)
    return x


def fill_int(n, dtype, seed, stream, lo, hi):
    x = np.empty(n, dtype)
    lib().orc_fill_int(_ptr(x), CODE_OF[x.dtype], n, seed, stream, int(lo), int(hi))
    return x


def xoshiro_block(seed, stream, block, count):
    st = (C.c_uint64 * 4)()
    lib().orc_rng_block(st, seed, stream, block)
    return [lib().orc_rng_next(st) for _ in range(count)]


# ---- matmul references ---------------------------------------------------------

def matmul_int(a, b, out_dtype=None):
    """Exact sum of products mod 2^w (numc_matmul_naive semantics); a, b may be strided views."""
    out_dtype = np.dtype(out_dtype or a.dtype)
    M, K = a.shape
    N = b.shape[1]
    out = np.zeros((M, N), out_dtype)
    rsa, csa = _estrides(a)
    rsb, csb = _estrides(b)
    lib().orc_matmul_int(CODE_OF[a.dtype], CODE_OF[out_dtype], _ptr(a), _ptr(b), _ptr(out), M, K, N,
                         rsa, csa, rsb, csb, N, 1)
    return out


def fref(a, b, rows=None, cols=None):
    """FP64 (f32 inputs) / double-double (f64 inputs) reference and |A||B|.
    With rows/cols: only those entries (1-D results)."""
    M, K = a.shape
    N = b.shape[1]
    rsa, csa = _estrides(a)
    rsb, csb = _estrides(b)
    if rows is not None:
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        ref = np.empty(len(rows))
        absr = np.empty(len(rows))
        lib().orc_matmul_fref(CODE_OF[a.dtype], _ptr(a), _ptr(b), M, K, N, rsa, csa, rsb, csb, _ptr(ref),
                              _ptr(absr), len(rows), _ptr(rows), _ptr(cols))
        return ref, absr
    ref = np.empty((M, N))
    absr = np.empty((M, N))
    lib().orc_matmul_fref(CODE_OF[a.dtype], _ptr(a), _ptr(b), M, K, N, rsa, csa, rsb, csb, _ptr(ref),
                          _ptr(absr), 0, None, None)
    return ref, absr


def err_metrics(c, ref, absr, tau=1e-3):
    """[componentwise, normwise, elementwise(|r| > tau*max|r|), max abs] of c against ref."""
    c = np.ascontiguousarray(c)
    out = np.zeros(4)
    M, N = c.shape
    lib().orc_err_metrics(CODE_OF[c.dtype], _ptr(c), M, N, N, 1, _ptr(np.ascontiguousarray(ref)),
                          _ptr(np.ascontiguousarray(absr)), tau, _ptr(out))
    return out


def check_matmul(a_shape, b_shape, o_shape, dta=8, dtb=8, dto=8, have=(1, 1, 1)):
    sa = (C.c_size_t * 2)(*(list(a_shape) + [0, 0])[:2])
    sb = (C.c_size_t * 2)(*(list(b_shape) + [0, 0])[:2])
    so = (C.c_size_t * 2)(*(list(o_shape) + [0, 0])[:2])
    return lib().orc_check_matmul(have[0], have[1], have[2], dta, dtb, dto, len(a_shape), len(b_shape),
                                  len(o_shape), sa, sb, so)


# ---- number formats ----------------------------------------------------------------

def split_bf16(x):
    x = np.ascontiguousarray(x, np.float32).ravel()
    hi = np.empty(x.size, np.uint16)
    lo = np.empty(x.size, np.uint16)
    lib().orc_split_bf16(_ptr(x), x.size, _ptr(hi), _ptr(lo))
    return hi, lo


def bf16_to_f32(h):
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def e4m3(x):
    return np.array([lib().orc_e4m3_rne(float(v)) for v in np.ravel(x)], np.uint8)


def e4m3_decode(q):
    return np.array([lib().orc_e4m3_to_f32(int(v)) for v in np.ravel(q)], np.float32)


def fp8_scale_exp(m):
    return lib().orc_fp8_scale_exp(float(m))


def e4m3_array(x):
    """orc_e4m3_rne over every element of x (same shape, uint8 codes)"""
    x = np.ascontiguousarray(x, np.float32)
    q = np.empty(x.shape, np.uint8)
    lib().orc_e4m3_rne_n(_ptr(x), x.size, _ptr(q))
    return q


def _e4m3_table():
    return np.array([lib().orc_e4m3_to_f32(q) for q in range(256)], np.float64)


def fp8_scale_exps(m):
    """fp8_scale_exp over an array of maxima (numpy; the tests check it against
    the C function element by element)"""
    m = np.asarray(m, np.float32)
    _, q = np.frexp(m)                      # m = f * 2^q, f in [0.5, 1)
    e = np.clip(8 - q.astype(np.int64), -126, 126)
    return np.where(m > 0, e, 0)


def fp8_quantize_blocks(x):
    """Each row of x in blocks of 32 consecutive k (block b holds k 32b..32b+31):
    the block scaled by 2^e, e = fp8_scale_exp of the block's max |x|, rounded
    to e4m3, decoded and unscaled back to float64 -- the exact operand rows the
    FP8 method multiplies (every row of every 32-k packed fragment carries its
    own power-of-two scale).  Columns of B are the rows of B^T."""
    x = np.asarray(x, np.float32)
    R, K = x.shape
    nb = -(-K // 32)
    xp = np.zeros((R, nb * 32), np.float32)
    xp[:, :K] = x
    blocks = xp.reshape(R, nb, 32)
    e = fp8_scale_exps(np.max(np.abs(blocks), axis=2, initial=0.0))      # (R, nb)
    scaled = np.ldexp(blocks.astype(np.float64), e[:, :, None]).astype(np.float32)
    q = _e4m3_table()[e4m3_array(scaled)]
    return np.ldexp(q, -e[:, :, None]).reshape(R, nb * 32)[:, :K]


def fp8_quantized_inputs(a, b):
    """A with a scale per row per 32 k, B with a scale per column per 32 k, e4m3,
    decoded and unscaled back to float64: the exact operands the FP8 method
    multiplies."""
    return fp8_quantize_blocks(a), fp8_quantize_blocks(np.asarray(b).T).T


# ---- GEMM consumers of numc's roadmap (ROADMAP.md:139-140) --------------------
# numc names numc_linear (out = x @ W + bias) and numc_bmm (batched matmul) but
# implements neither; these restate the definitions on top of matmul_int / fref.
# Pinned in tests by closed forms (x = 0 gives the broadcast bias, x = I gives
# w + bias) and, for the batch broadcasting, by numpy.matmul on integers
# (numpy accumulates in the array dtype, i.e. mod 2^w).

def _wrap_add(x, y):
    """x + y in x's integer dtype, mod 2^w"""
    w = x.dtype.itemsize * 8
    u = (x.astype(object) + np.asarray(y).astype(object)) % (1 << w)
    return np.array(u.tolist(), dtype=np.uint64).astype(x.dtype.str.replace("i", "u")).view(x.dtype).reshape(x.shape)


def linear(x, w, bias=None):
    """(expected, |x||w| or None) of out = x @ w + bias for x (..., K), w (K, N)."""
    K, N = w.shape
    x2 = np.ascontiguousarray(x).reshape(-1, K)
    out_shape = x.shape[:-1] + (N,)
    if np.issubdtype(x.dtype, np.integer):
        r = matmul_int(x2, np.ascontiguousarray(w))
        if bias is not None:
            r = _wrap_add(r, np.broadcast_to(bias, r.shape))
        return r.reshape(out_shape), None
    ref, absr = fref(x2, np.ascontiguousarray(w))
    if bias is not None:
        ref = ref + bias.astype(np.float64)
    return ref.reshape(out_shape), absr.reshape(out_shape)


def bmm(a, b):
    """(expected, |a||b| or None) of numpy.matmul-style batched a @ b (both >= 2-D)."""
    batch = np.broadcast_shapes(a.shape[:-2], b.shape[:-2])
    M, K, N = a.shape[-2], a.shape[-1], b.shape[-1]
    ab = np.broadcast_to(a, batch + (M, K))
    bb = np.broadcast_to(b, batch + (K, N))
    integer = np.issubdtype(a.dtype, np.integer)
    out = np.zeros(batch + (M, N), a.dtype if integer else np.float64)
    den = None if integer else np.zeros(batch + (M, N))
    for idx in np.ndindex(*batch):
        x, y = np.ascontiguousarray(ab[idx]), np.ascontiguousarray(bb[idx])
        if integer:
            out[idx] = matmul_int(x, y)
        else:
            out[idx], den[idx] = fref(x, y)
    return out, den
