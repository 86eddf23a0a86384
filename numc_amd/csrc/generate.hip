// generate.hip -- the seeded fill on the GPU, for arrays whose device copy is
// current: one thread per 2^16-element block runs the block's xoshiro256**
// stream (rng.h, the code the host fill runs), so the array never crosses
// PCIe.  Uniform and integer fills are bit-identical to the host fill; the
// normal fill computes log, sin and cos with the device's binary64 library,
// which may differ from the host's in the last bit of binary64 (and so, on
// rare roundings, in the last bit of the stored value).
#include "internal.h"
#include "rng.h"

namespace {
__global__ void __launch_bounds__(64) k_fill_random(unsigned char *p, long long n, long long nb, NumcFillSpec f) {
  const long long b = blockIdx.x * 64LL + threadIdx.x;
  if (b < nb) numc_fill_block(p, (size_t)n, (size_t)b, f);
}
} // namespace

int numc_fill_random_device(NumcArray *arr, const NumcFillSpec &f) {
  NumcCtx *ctx = arr->ctx;
  int err = numc_gpu_init(ctx);
  if (err) return err;
  const long long n = (long long)arr->size;
  if (n == 0) return 0;
  unsigned char *d = numc_dev_ptr(arr);
  if (!d) return numc_get_error().code ? numc_get_error().code : NUMC_ERR_INTERNAL;
  const long long nb = (n + 65535) >> 16;
  hipLaunchKernelGGL(k_fill_random, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, ctx->stream, d, n, nb, f);
  HIP_TRY(hipGetLastError(), "fill_random launch");
  return numc_finish(ctx);
}
