"""numc_amd -- Python view of the numc C ABI served by libnumc_amd.so.

The product is the C library (include/numc/*.h, include/numc_amd.h); this
module only loads it with ctypes and wraps contexts and arrays so tests and
bench.py can drive it.  It raises if the library is missing: there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnumc_amd.so")

# NumcDType (include/numc/dtype.h); numpy has no bfloat16, so BFLOAT16 arrays
# come and go as their uint16 bit patterns (bf16_bits / bf16_value convert)
INT8, INT16, INT32, INT64, UINT8, UINT16, UINT32, UINT64, FLOAT32, FLOAT64, BFLOAT16, FLOAT16 = range(12)
NP_DTYPES = [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32,
             np.uint64, np.float32, np.float64, np.uint16, np.float16]
DTYPE_OF_NP = {np.dtype(t): i for i, t in enumerate(NP_DTYPES[:10])}
DTYPE_OF_NP[np.dtype(np.float16)] = FLOAT16


def bf16_bits(x):
    """float32 values -> bfloat16 bit patterns (uint16), round to nearest even"""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    q = ((u >> 16) | np.where((u & 0x7FFFFF) != 0, np.uint64(0x40), np.uint64(0))).astype(np.uint16)
    return np.where(nan, q, r).astype(np.uint16)


def bf16_value(b):
    """bfloat16 bit patterns (uint16) -> float32 values"""
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)

# NumcErrorCode
ERR_NONE, ERR_NULL, ERR_MALLOC, ERR_SHAPE, ERR_TYPE, ERR_BOUNDS, ERR_VALUE = 0, -1, -2, -3, -4, -5, -6
ERR_INTERNAL = -99

# NumcMethod / NumcFamily / flags (include/numc_amd.h)
METHOD_AUTO, METHOD_PLAIN, METHOD_SPLIT_BF16, METHOD_FP8_SCALED = 0, 1, 2, 3
FAMILY_F32, FAMILY_SPLIT_BF16, FAMILY_FP8, FAMILY_F64, FAMILY_I8, FAMILY_BF16, FAMILY_F16 = range(7)
FAMILY_COUNT = 7
FAMILY_NAMES = ["f32", "split_bf16", "fp8", "f64", "i8", "bf16", "f16"]
OPT_WIDEN = 1
OPT_DEVICE_POINTERS = 2
DIST_UNIFORM, DIST_NORMAL, DIST_INT = 0, 1, 2


class NumcError_(C.Structure):
    _fields_ = [("code", C.c_int), ("msg", C.c_char_p)]


class Checkpoint(C.Structure):
    _fields_ = [("_block", C.c_void_p), ("_index", C.c_size_t)]


class Slice(C.Structure):
    _fields_ = [("axis", C.c_size_t), ("start", C.c_size_t), ("stop", C.c_size_t),
                ("step", C.c_size_t)]


class MatmulOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("config", C.c_int), ("flags", C.c_int)]


class KernelConfig(C.Structure):
    _fields_ = [("bm", C.c_int), ("bn", C.c_int), ("kc", C.c_int), ("waves_m", C.c_int),
                ("waves_n", C.c_int), ("split_k", C.c_int), ("waves_k", C.c_int), ("direct", C.c_int)]


_P = C.c_void_p
_SZ = C.c_size_t
_SIGS = {
    # error.h
    "numc_get_error": (NumcError_, []),
    "numc_set_error": (C.c_int, [C.c_int, C.c_char_p]),
    "numc_log_error": (None, [C.POINTER(NumcError_)]),
    # array.h
    "numc_ctx_create": (_P, []),
    "numc_ctx_free": (None, [_P]),
    "numc_ctx_checkpoint": (Checkpoint, [_P]),
    "numc_ctx_restore": (None, [_P, Checkpoint]),
    "numc_array_create": (_P, [_P, C.POINTER(_SZ), _SZ, C.c_int]),
    "numc_array_zeros": (_P, [_P, C.POINTER(_SZ), _SZ, C.c_int]),
    "numc_array_fill": (_P, [_P, C.POINTER(_SZ), _SZ, C.c_int, _P]),
    "numc_array_copy": (_P, [_P]),
    "numc_array_write": (None, [_P, _P]),
    "numc_array_reshape": (C.c_int, [_P, C.POINTER(_SZ), _SZ]),
    "numc_array_reshape_copy": (_P, [_P, C.POINTER(_SZ), _SZ]),
    "numc_array_transpose": (C.c_int, [_P, C.POINTER(_SZ)]),
    "numc_array_transpose_copy": (_P, [_P, C.POINTER(_SZ)]),
    "numc_array_slice": (_P, [_P, C.POINTER(Slice)]),
    "numc_array_is_contiguous": (C.c_bool, [_P]),
    "numc_array_contiguous": (C.c_int, [_P]),
    "numc_array_size": (_SZ, [_P]),
    "numc_array_capacity": (_SZ, [_P]),
    "numc_array_elem_size": (_SZ, [_P]),
    "numc_array_ndim": (_SZ, [_P]),
    "numc_array_shape": (None, [_P, C.POINTER(_SZ)]),
    "numc_array_strides": (None, [_P, C.POINTER(_SZ)]),
    "numc_array_dtype": (C.c_int, [_P]),
    "numc_array_data": (_P, [_P]),
    # math.h
    "numc_matmul": (C.c_int, [_P, _P, _P]),
    "numc_matmul_naive": (C.c_int, [_P, _P, _P]),
    # numc_amd.h
    "numc_matmul_ex": (C.c_int, [_P, _P, _P, C.POINTER(MatmulOpts)]),
    "numc_dot": (C.c_int, [_P, _P, _P]),
    "numc_linear": (C.c_int, [_P, _P, _P, _P]),
    "numc_linear_ex": (C.c_int, [_P, _P, _P, _P, C.POINTER(MatmulOpts)]),
    "numc_bmm": (C.c_int, [_P, _P, _P]),
    "numc_amd_gemm": (C.c_int, [_P, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t,
                                _P, C.c_size_t, C.c_size_t, _P, C.c_size_t, C.c_size_t,
                                _P, C.c_size_t, C.c_size_t, C.POINTER(MatmulOpts)]),
    "numc_amd_gemm_noalloc": (C.c_int, [_P, C.c_int, C.c_int, C.c_size_t, C.c_size_t, C.c_size_t,
                                        _P, C.c_size_t, C.c_size_t, _P, C.c_size_t, C.c_size_t,
                                        _P, C.c_size_t, C.c_size_t, C.POINTER(MatmulOpts)]),
    "numc_matmul_batched": (C.c_int, [_P, _SZ, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                      C.POINTER(MatmulOpts)]),
    "numc_matmul_batched_tune": (C.c_int, [_P, _SZ, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P),
                                           C.POINTER(MatmulOpts), C.c_int]),
    "numc_ctx_batch_graph_replays": (C.c_longlong, [_P]),
    "numc_ctx_batch_graphs_cached": (C.c_int, [_P]),
    "numc_matmul_batched_schedule": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "numc_matmul_batched_set_schedule": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "numc_autotune_release": (C.c_int, [_P]),
    "numc_device_copy": (C.c_int, [_P, _P, _P, _SZ]),
    "numc_selector_tune": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "numc_selector_pin": (C.c_int, [C.c_int] * 6),
    "numc_selector_active": (C.c_int, []),
    "numc_selector_candidates": (C.c_int, [C.c_int] * 6 + [C.POINTER(C.c_int), C.c_int]),
    "numc_array_to_device": (C.c_int, [_P]),
    "numc_array_to_host": (C.c_int, [_P]),
    "numc_array_on_device": (C.c_int, [_P]),
    "numc_array_device_data": (_P, [_P]),
    "numc_ctx_set_sync": (C.c_int, [_P, C.c_int]),
    "numc_ctx_synchronize": (C.c_int, [_P]),
    "numc_ctx_stream": (_P, [_P]),
    "numc_ctx_timer_mark": (C.c_int, [_P, C.c_int]),
    "numc_ctx_timer_elapsed": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "numc_ctx_profile": (C.c_int, [_P, C.c_int]),
    "numc_ctx_profile_read": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_longlong), C.c_int]),
    "numc_array_fill_random": (C.c_int, [_P, C.c_int, C.c_uint64, C.c_uint64, C.c_double,
                                         C.c_double]),
    "numc_kernel_config_count": (C.c_int, [C.c_int]),
    "numc_kernel_config_get": (C.c_int, [C.c_int, C.c_int, C.POINTER(KernelConfig)]),
    "numc_matmul_family": (C.c_int, [_P, _P, _P, C.POINTER(MatmulOpts)]),
}

_lib = None


def lib():
    """Load libnumc_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libnumc_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def last_error():
    e = lib().numc_get_error()
    return e.code, (e.msg.decode() if e.msg else "")


class NumcFailure(RuntimeError):
    def __init__(self, code, what):
        c, msg = last_error()
        super().__init__("%s failed with %d: %s" % (what, code, msg))
        self.code = code


def check(code, what):
    if code != 0:
        raise NumcFailure(code, what)
    return code


def _sizes(seq):
    arr = (C.c_size_t * max(1, len(seq)))(*seq)
    return arr


class Ctx:
    def __init__(self):
        self.h = lib().numc_ctx_create()
        if not self.h:
            raise NumcFailure(-2, "numc_ctx_create")

    def close(self):
        if self.h:
            lib().numc_ctx_free(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def array(self, shape, dtype, zeros=False):
        fn = lib().numc_array_zeros if zeros else lib().numc_array_create
        h = fn(self.h, _sizes(shape), len(shape), dtype)
        if not h:
            raise NumcFailure(-2, "numc_array_create")
        return Array(h, self)

    def from_numpy(self, x, dtype=None):
        """array with x's shape and elements; dtype overrides the numpy mapping
        (BFLOAT16 takes uint16 bit patterns)"""
        x = np.asarray(x)
        if not x.flags.c_contiguous:   # (ascontiguousarray would turn 0-D into 1-D)
            x = x.copy(order="C")
        dt = DTYPE_OF_NP[x.dtype] if dtype is None else dtype
        if np.dtype(NP_DTYPES[dt]).itemsize != x.dtype.itemsize:
            raise ValueError("from_numpy: %s elements for dtype %d" % (x.dtype, dt))
        a = self.array(x.shape, dt)
        lib().numc_array_write(a.h, x.ctypes.data_as(C.c_void_p))
        return a

    def set_sync(self, sync):
        check(lib().numc_ctx_set_sync(self.h, int(sync)), "numc_ctx_set_sync")

    def synchronize(self):
        check(lib().numc_ctx_synchronize(self.h), "numc_ctx_synchronize")

    def profile(self, enable=True):
        check(lib().numc_ctx_profile(self.h, int(enable)), "numc_ctx_profile")

    def profile_read(self, category, family, reset=True):
        ms, n = C.c_double(), C.c_longlong()
        check(lib().numc_ctx_profile_read(self.h, category, family, C.byref(ms), C.byref(n), int(reset)),
              "numc_ctx_profile_read")
        return ms.value, n.value

    def mark(self, slot):
        check(lib().numc_ctx_timer_mark(self.h, slot), "numc_ctx_timer_mark")

    def elapsed_ms(self, s0, s1):
        ms = C.c_float()
        check(lib().numc_ctx_timer_elapsed(self.h, s0, s1, C.byref(ms)), "numc_ctx_timer_elapsed")
        return ms.value


class Array:
    def __init__(self, h, ctx):
        self.h = h
        self.ctx = ctx

    @property
    def ndim(self):
        return lib().numc_array_ndim(self.h)

    @property
    def shape(self):
        n = self.ndim
        buf = (C.c_size_t * max(1, n))()
        lib().numc_array_shape(self.h, buf)
        return tuple(buf[i] for i in range(n))

    @property
    def strides(self):
        n = self.ndim
        buf = (C.c_size_t * max(1, n))()
        lib().numc_array_strides(self.h, buf)
        return tuple(buf[i] for i in range(n))

    @property
    def dtype(self):
        return lib().numc_array_dtype(self.h)

    def numpy(self):
        """Copy of the array's elements (honours the view's strides)."""
        ptr = lib().numc_array_data(self.h)
        shape, strides = self.shape, self.strides
        npdt = np.dtype(NP_DTYPES[self.dtype])
        if 0 in shape:
            return np.zeros(shape, npdt)
        if not shape:   # 0-D: one element
            return np.array(np.frombuffer((C.c_char * npdt.itemsize).from_address(ptr), npdt)[0])
        span = sum((s - 1) * st for s, st in zip(shape, strides)) + npdt.itemsize
        buf = (C.c_char * span).from_address(ptr)
        return np.ndarray(shape, npdt, buffer=buf, strides=strides).copy()

    def write(self, x):
        x = np.ascontiguousarray(x, dtype=NP_DTYPES[self.dtype])
        lib().numc_array_write(self.h, x.ctypes.data_as(C.c_void_p))

    def fill_random(self, dist, seed, stream, lo=-1.0, hi=1.0):
        check(lib().numc_array_fill_random(self.h, dist, seed, stream, lo, hi), "numc_array_fill_random")

    def to_device(self):
        check(lib().numc_array_to_device(self.h), "numc_array_to_device")

    def to_host(self):
        check(lib().numc_array_to_host(self.h), "numc_array_to_host")

    def transpose(self, axes=(1, 0)):
        """numc_array_transpose in place (axes: the new order, one entry per axis)"""
        check(lib().numc_array_transpose(self.h, _sizes(list(axes))), "numc_array_transpose")
        return self

    def slice(self, axis, start, stop, step=1):
        s = Slice(axis, start, stop, step)
        h = lib().numc_array_slice(self.h, C.byref(s))
        if not h:
            raise NumcFailure(-1, "numc_array_slice")
        return Array(h, self.ctx)


def matmul(a, b, out, method=METHOD_AUTO, config=-1, flags=0):
    opts = MatmulOpts(method, config, flags)
    return lib().numc_matmul_ex(a.h, b.h, out.h, C.byref(opts))


def dot(a, b, out):
    """numc_dot(a, b, out) -> numc error code"""
    return lib().numc_dot(a.h, b.h, out.h)


def linear(x, w, bias, out, method=METHOD_AUTO, config=-1):
    """numc_linear(x, w, bias, out): out = x @ w + bias (bias may be None) -> error code;
    a method or configuration goes through numc_linear_ex"""
    bh = bias.h if bias is not None else None
    if method == METHOD_AUTO and config < 0:
        return lib().numc_linear(x.h, w.h, bh, out.h)
    opts = MatmulOpts(method, config, 0)
    return lib().numc_linear_ex(x.h, w.h, bh, out.h, C.byref(opts))


def bmm(a, b, out):
    """numc_bmm(a, b, out): batched a @ b with broadcast batch axes -> error code"""
    return lib().numc_bmm(a.h, b.h, out.h)


def gemm_pointers(ctx, a, b, c, method=METHOD_AUTO, config=-1, flags=0, dtype=None):
    """numc_amd_gemm on numpy arrays (host pointers; byte strides taken from
    the arrays, so strided views of a, b and c work as they are).  dtype
    overrides the numpy mapping for all three (BFLOAT16: uint16 bit patterns)."""
    dt, odt = (DTYPE_OF_NP[a.dtype], DTYPE_OF_NP[c.dtype]) if dtype is None else (dtype, dtype)
    M, K = a.shape
    N = b.shape[1]
    opts = MatmulOpts(method, config, flags)
    return lib().numc_amd_gemm(ctx.h, dt, odt, M, N, K,
                               a.ctypes.data, a.strides[0], a.strides[1],
                               b.ctypes.data, b.strides[0], b.strides[1],
                               c.ctypes.data, c.strides[0], c.strides[1], C.byref(opts))


def matmul_batched_tune(ctx, As, Bs, Outs, method=METHOD_AUTO, flags=0, reps=3):
    n = len(As)
    pa = (C.c_void_p * n)(*[x.h for x in As])
    pb = (C.c_void_p * n)(*[x.h for x in Bs])
    po = (C.c_void_p * n)(*[x.h for x in Outs])
    opts = MatmulOpts(method, -1, flags)
    return lib().numc_matmul_batched_tune(ctx.h, n, pa, pb, po, C.byref(opts), reps)


def batched_schedule(family):
    """(direct, threshold) of the family's batched schedule"""
    d, t = C.c_int(), C.c_int()
    check(lib().numc_matmul_batched_schedule(family, C.byref(d), C.byref(t)), "numc_matmul_batched_schedule")
    return d.value, t.value


class BatchHandles:
    """the three handle arrays numc_matmul_batched takes, built once (a C caller
    keeps its arrays of NumcArray pointers the same way)"""

    def __init__(self, As, Bs, Outs):
        self.n = len(As)
        self.pa = (C.c_void_p * self.n)(*[x.h for x in As])
        self.pb = (C.c_void_p * self.n)(*[x.h for x in Bs])
        self.po = (C.c_void_p * self.n)(*[x.h for x in Outs])
        self.keep = (As, Bs, Outs)


def matmul_batched(ctx, As, Bs=None, Outs=None, method=METHOD_AUTO, config=-1, flags=0):
    """numc_matmul_batched over lists of arrays, or over a BatchHandles (As)"""
    h = As if isinstance(As, BatchHandles) else BatchHandles(As, Bs, Outs)
    opts = MatmulOpts(method, config, flags)
    return lib().numc_matmul_batched(ctx.h, h.n, h.pa, h.pb, h.po, C.byref(opts))


def config_count(family):
    return lib().numc_kernel_config_count(family)


def config_get(family, idx):
    c = KernelConfig()
    check(lib().numc_kernel_config_get(family, idx, C.byref(c)), "numc_kernel_config_get")
    return dict(bm=c.bm, bn=c.bn, kc=c.kc, waves_m=c.waves_m, waves_n=c.waves_n, split_k=c.split_k,
                waves_k=c.waves_k, direct=c.direct)
