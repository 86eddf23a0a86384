// batched.hip -- numc_matmul_batched: many independent GEMMs of uneven shape.
#include "gemm.h"
#include "internal.h"

extern "C" void numc_batch_cache_free(NumcCtx *) {}

int numc_matmul_one(const NumcArray *a, const NumcArray *b, NumcArray *out, const NumcMatmulOpts *opts,
                    bool finish); // matmul.cpp

extern "C" NUMC_API int numc_matmul_batched(NumcCtx *ctx, size_t count, NumcArray *const *a,
                                            NumcArray *const *b, NumcArray *const *out,
                                            const NumcMatmulOpts *opts) {
  if (!ctx || (count && (!a || !b || !out)))
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "matmul_batched: NULL argument");
  for (size_t i = 0; i < count; i++) {
    int err = numc_matmul_one(a[i], b[i], out[i], opts, false);
    if (err) return err;
  }
  return numc_finish(ctx);
}
