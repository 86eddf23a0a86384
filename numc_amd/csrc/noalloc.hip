// noalloc.hip -- numc_amd_gemm_noalloc: the pointer-level GEMM of
// numc_amd_gemm (device operands) for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No counterpart in numc.
//
// The packed kernels' workspace (packed operand planes, split-K partial
// sums, integer plane masks, FP8 scales) comes from `scratch`, a file-scope
// device array of NUMC_NOALLOC_SCRATCH bytes that each GPU the library runs on
// holds until the process ends.  Host code never uses the array's name as an
// address: the first call on a device, with that device current, asks
// hipGetSymbolAddress for the device's own copy and keeps it per device.
// Every call on one GPU shares that GPU's scratch: calls that overlap on one
// GPU (two streams, two threads) race on it.  Whatever must start at zero is
// cleared by the kernels of each call (a fill kernel, no memset node).
#include "gemm.h"
#include "internal.h"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[NUMC_NOALLOC_SCRATCH];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch` (looked up once per device)
int scratch_for_current_device(unsigned char **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return numc_hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "gemm_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return numc_hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// bytes of scratch a configuration needs for p (workspace, then aux)
size_t scratch_need(const GemmProblem &p, const GemmConfig &c) {
  size_t aux = 0, ws = numc_gemm_workspace(p, c, &aux);
  return align256(ws) + aux;
}
} // namespace

extern "C" NUMC_API int numc_amd_gemm_noalloc(void *stream, int dtype, int out_dtype, size_t M, size_t N, size_t K,
                                              const void *a, size_t rsa, size_t csa, const void *b, size_t rsb,
                                              size_t csb, void *c, size_t rso, size_t cso, const NumcMatmulOpts *opts) {
  if ((M && K && !a) || (K && N && !b) || (M && N && !c))
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "gemm_noalloc: NULL pointer (a=%p b=%p c=%p)", a, b, c);
  const bool widen = opts && (opts->flags & NUMC_OPT_WIDEN);
  bool types_ok = dtype >= 0 && dtype < NUMC_DTYPE_COUNT && dtype == out_dtype;
  if (!types_ok && widen)
    types_ok = (dtype == NUMC_DTYPE_INT8 || dtype == NUMC_DTYPE_UINT8) &&
               (out_dtype == NUMC_DTYPE_INT32 || out_dtype == NUMC_DTYPE_UINT32);
  if (!types_ok) return NUMC_SET_ERROR(NUMC_ERR_TYPE, "gemm_noalloc: dtype mismatch (in=%d out=%d)", dtype, out_dtype);
  const size_t lim = (size_t)1 << 30;
  if (M >= lim || N >= lim || K >= lim) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "gemm_noalloc: dimensions must be below 2^30");
  const size_t ein = numc_dtype_size((NumcDType)dtype), eout = numc_dtype_size((NumcDType)out_dtype);
  if (rsa % ein || csa % ein || rsb % ein || csb % ein || rso % eout || cso % eout)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "gemm_noalloc: strides must be multiples of the element size");
  NumcMatmulOpts o = opts ? *opts : NumcMatmulOpts{NUMC_METHOD_AUTO, -1, 0};
  o.flags |= NUMC_OPT_DEVICE_POINTERS;
  const int family = numc_matmul_family_dt(dtype, &o);
  if (family < 0) return family;
  if (family >= NUMC_FAMILY_COUNT)
    return NUMC_SET_ERROR(NUMC_ERR_TYPE, "gemm_noalloc: 64-bit integers are not supported (no scratch-free kernel)");
  if (M == 0 || N == 0) return 0;
  unsigned char *ws = nullptr;
  int err = scratch_for_current_device(&ws);
  if (err) return err;

  GemmProblem p{};
  p.family = family;
  p.in_dt = (NumcDType)dtype;
  p.out_dt = (NumcDType)out_dtype;
  p.M = (int)M;
  p.N = (int)N;
  p.K = (int)K;
  p.rsa = (long long)(rsa / ein);
  p.csa = (long long)(csa / ein);
  p.rsb = (long long)(rsb / ein);
  p.csb = (long long)(csb / ein);
  p.rso = (long long)(rso / eout);
  p.cso = (long long)(cso / eout);
  p.A = a;
  p.B = b;
  p.C = c;
  const bool direct_ok = numc_direct_ok(p);
  // the configuration: the caller's, else the selector's choice (the exact
  // shape table, the learned model or the analytic model -- never a timing
  // run, which would need events), else any whose workspace fits the scratch
  int cfg_idx = o.config;
  if (cfg_idx < 0) {
    const int lay = (p.csa != 1 ? 1 : 0) | (p.csb != 1 ? 2 : 0) |
                    ((p.csa == 1 && p.rsa > p.K) || (p.csb == 1 && p.rsb > p.N) ? 4 : 0);
    cfg_idx = numc_select_config(family, p.M, p.N, p.K, lay, direct_ok);
    const GemmConfig *g = numc_gemm_config(family, cfg_idx);
    if (!g || scratch_need(p, *g) > NUMC_NOALLOC_SCRATCH) {
      cfg_idx = -1;
      for (int i = 0; i < numc_gemm_config_count(family) && cfg_idx < 0; i++) {
        const GemmConfig *ci = numc_gemm_config(family, i);
        if ((!ci->cfg.direct || direct_ok) && scratch_need(p, *ci) <= NUMC_NOALLOC_SCRATCH) cfg_idx = i;
      }
    }
  }
  const GemmConfig *cfg = cfg_idx >= 0 ? numc_gemm_config(family, cfg_idx) : nullptr;
  if (!cfg)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE,
                          "gemm_noalloc: no configuration's workspace fits the %d MiB scratch for %zu x %zu x %zu",
                          NUMC_NOALLOC_SCRATCH >> 20, M, N, K);
  if (cfg->cfg.direct && !direct_ok)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "gemm_noalloc: config %d reads the operands in place, which needs one "
                                          "int8 digit plane", cfg_idx);
  size_t aux = 0, need = numc_gemm_workspace(p, *cfg, &aux);
  if (align256(need) + aux > NUMC_NOALLOC_SCRATCH)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "gemm_noalloc: config %d needs %zu bytes of workspace, above the %d MiB scratch",
                          cfg_idx, align256(need) + aux, NUMC_NOALLOC_SCRATCH >> 20);
  // a context on the caller's stream, made of nothing but that stream
  NumcCtx view{};
  view.stream = (hipStream_t)stream;
  view.gpu_ready = true;
  view.synchronous = false;
  return numc_gemm_launch(&view, p, *cfg, ws, ws + align256(need));
}
