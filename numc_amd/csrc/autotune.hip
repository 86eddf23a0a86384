// autotune.hip -- exhaustive timing of every kernel configuration of a family
// on one shape (the autotuner's measurement primitive).
//
// numc has no runtime tuning: its blocking constants and small-problem
// thresholds are compile-time (src/intrinsics/gemm_avx2.h:21-64,
// src/matmul/matmul.c:241-256).  Here every configuration of the chosen
// family is run on device-resident scratch operands of the requested shape
// and layout, whole calls (pack + GEMM + split-K reduce) timed with HIP
// events on the context stream.  The sweep driver (numc_amd/autotune.py)
// turns these timings into the selector's training database.
#include "gemm.h"
#include "internal.h"

#include <algorithm>
#include <vector>

namespace {
__global__ void k_fill_pattern(unsigned char *p, long long n, int dtype) {
  long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= n) return;
  unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ULL;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ULL;
  h ^= h >> 32;
  switch (dtype) {
  case NUMC_DTYPE_FLOAT32: ((float *)p)[i] = (float)((h >> 40) * 0x1p-23) - 1.0f; break;
  case NUMC_DTYPE_FLOAT64: ((double *)p)[i] = (double)(h >> 11) * 0x1p-52 - 1.0; break;
  case NUMC_DTYPE_BFLOAT16: // sign, exponent 126 or 127 (magnitudes in [0.5, 2)), random mantissa
    ((uint16_t *)p)[i] = (uint16_t)(((h >> 8) & 0x8000u) | (0x3F00u + ((h >> 16) & 0xFFu)));
    break;
  case NUMC_DTYPE_FLOAT16: // sign, exponent 14 or 15 (magnitudes in [0.5, 2)), random mantissa
    ((uint16_t *)p)[i] = (uint16_t)(((h >> 8) & 0x8000u) | (0x3800u + ((h >> 16) & 0x7FFu)));
    break;
  default: ((int8_t *)p)[i] = (int8_t)(h & 0xFF); break;
  }
}

struct TuneBuffers {
  void *a = nullptr, *b = nullptr, *c = nullptr;
  size_t ab = 0, bb = 0, cb = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

int grow(void **p, size_t *have, size_t need, int dtype, hipStream_t st) {
  if (*have >= need) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  size_t want = need + need / 4 + 256;
  HIP_TRY(hipMalloc(p, want), "hipMalloc(autotune scratch)");
  *have = want;
  const size_t es = numc_dtype_size((NumcDType)dtype);
  const long long n = (long long)(want / es);
  hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (unsigned char *)*p, n, dtype);
  return 0;
}
} // namespace

static TuneBuffers *tune_buffers(NumcCtx *ctx) {
  if (!ctx->tune) ctx->tune = new TuneBuffers();
  return (TuneBuffers *)ctx->tune;
}

extern "C" NUMC_API int numc_autotune_release(NumcCtx *ctx) {
  if (!ctx) return NUMC_SET_ERROR(NUMC_ERR_NULL, "ctx is NULL");
  TuneBuffers *t = tune_buffers(ctx);
  if (ctx->gpu_ready) (void)hipStreamSynchronize(ctx->stream);
  if (t->a) (void)hipFree(t->a);
  if (t->b) (void)hipFree(t->b);
  if (t->c) (void)hipFree(t->c);
  if (t->e0) (void)hipEventDestroy(t->e0);
  if (t->e1) (void)hipEventDestroy(t->e1);
  delete t;
  ctx->tune = nullptr;
  return 0;
}

// flags: bit 0 A stored transposed (K x M), bit 1 B stored transposed (N x K),
// bit 2 leading dimensions padded by 16 elements (strided rows)
extern "C" NUMC_API int numc_autotune_shape(NumcCtx *ctx, int family, int M, int N, int K, int flags, int reps,
                                            float *ms_out, int ms_len) {
  if (!ctx || !ms_out) return NUMC_SET_ERROR(NUMC_ERR_NULL, "autotune: NULL argument");
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "autotune: bad family %d", family);
  if (M <= 0 || N <= 0 || K <= 0) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "autotune: empty shape");
  const int ncfg = numc_gemm_config_count(family);
  if (ms_len < ncfg) return NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "autotune: need room for %d timings", ncfg);
  int err = numc_gpu_init(ctx);
  if (err) return err;
  TuneBuffers *tb = tune_buffers(ctx);
  if (!tb->e0) {
    HIP_TRY(hipEventCreate(&tb->e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&tb->e1), "hipEventCreate");
  }
  const NumcDType dt = family == NUMC_FAMILY_F64    ? NUMC_DTYPE_FLOAT64
                       : family == NUMC_FAMILY_I8   ? NUMC_DTYPE_INT8
                       : family == NUMC_FAMILY_BF16 ? NUMC_DTYPE_BFLOAT16
                       : family == NUMC_FAMILY_F16  ? NUMC_DTYPE_FLOAT16
                                                    : NUMC_DTYPE_FLOAT32;
  const size_t es = numc_dtype_size(dt);
  const long long pad = (flags & 4) ? 16 : 0;
  GemmProblem p{};
  p.family = family;
  p.in_dt = p.out_dt = dt;
  p.M = M;
  p.N = N;
  p.K = K;
  // A: M x K view; stored K x M when transposed
  long long a_elems, b_elems;
  if (flags & 1) {
    p.rsa = 1;
    p.csa = M + pad;
    a_elems = (long long)K * (M + pad);
  } else {
    p.rsa = K + pad;
    p.csa = 1;
    a_elems = (long long)M * (K + pad);
  }
  if (flags & 2) {
    p.rsb = 1;
    p.csb = K + pad;
    b_elems = (long long)N * (K + pad);
  } else {
    p.rsb = N + pad;
    p.csb = 1;
    b_elems = (long long)K * (N + pad);
  }
  p.rso = N;
  p.cso = 1;
  if ((err = grow(&tb->a, &tb->ab, (size_t)a_elems * es, dt, ctx->stream))) return err;
  if ((err = grow(&tb->b, &tb->bb, (size_t)b_elems * es, dt, ctx->stream))) return err;
  if ((err = grow(&tb->c, &tb->cb, (size_t)M * N * es, dt, ctx->stream))) return err;
  p.A = tb->a;
  p.B = tb->b;
  p.C = tb->c;
  for (int i = 0; i < ncfg; i++) {
    const GemmConfig *c = numc_gemm_config(family, i);
    size_t aux = 0, need = numc_gemm_workspace(p, *c, &aux);
    if ((err = numc_ws_reserve(ctx, &ctx->ws, need))) return err;
    if ((err = numc_ws_reserve(ctx, &ctx->ws_aux, aux))) return err;
    if ((err = numc_gemm_launch(ctx, p, *c, ctx->ws.ptr, ctx->ws_aux.ptr))) return err; // warmup
    HIP_TRY(hipEventRecord(tb->e0, ctx->stream), "hipEventRecord");
    for (int r = 0; r < std::max(1, reps); r++)
      if ((err = numc_gemm_launch(ctx, p, *c, ctx->ws.ptr, ctx->ws_aux.ptr))) return err;
    HIP_TRY(hipEventRecord(tb->e1, ctx->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(tb->e1), "hipEventSynchronize");
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, tb->e0, tb->e1), "hipEventElapsedTime");
    ms_out[i] = ms / (float)std::max(1, reps);
  }
  return 0;
}

// Time candidate configurations on one call's own operands and return the
// fastest in *best: each candidate runs once to warm up, then once timed (five
// times when a run takes under 50 us, where launch noise is of the order of
// the differences).  The output is rewritten by every run; the caller's run of
// the winner writes the result.  The runs stay out of the kernel profile.
int numc_tune_call(NumcCtx *ctx, const GemmProblem &p, const int *cand, int n, int *best) {
  TuneBuffers *tb = tune_buffers(ctx);
  if (!tb->e0) {
    HIP_TRY(hipEventCreate(&tb->e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&tb->e1), "hipEventCreate");
  }
  void *prof = ctx->prof;
  ctx->prof = nullptr;
  int err = 0;
  float best_ms = 0.f;
  *best = cand[0];
  for (int i = 0; i < n && !err; i++) {
    const GemmConfig *c = numc_gemm_config(p.family, cand[i]);
    if (!c) continue;
    size_t aux = 0, need = numc_gemm_workspace(p, *c, &aux);
    if ((err = numc_ws_reserve(ctx, &ctx->ws, need)) || (err = numc_ws_reserve(ctx, &ctx->ws_aux, aux))) break;
    if ((err = numc_gemm_launch(ctx, p, *c, ctx->ws.ptr, ctx->ws_aux.ptr))) break;
    float ms = 0.f;
    for (int reps = 1; reps <= 5 && !err; reps += 4) {
      if (hipEventRecord(tb->e0, ctx->stream) != hipSuccess) err = NUMC_ERR_INTERNAL;
      for (int r = 0; r < reps && !err; r++) err = numc_gemm_launch(ctx, p, *c, ctx->ws.ptr, ctx->ws_aux.ptr);
      if (err) break;
      if (hipEventRecord(tb->e1, ctx->stream) != hipSuccess || hipEventSynchronize(tb->e1) != hipSuccess ||
          hipEventElapsedTime(&ms, tb->e0, tb->e1) != hipSuccess) {
        err = NUMC_SET_ERROR(NUMC_ERR_INTERNAL, "autotune: event timing failed");
        break;
      }
      ms /= (float)reps;
      if (ms >= 0.05f) break;
    }
    if (!err && (i == 0 || ms < best_ms)) {
      best_ms = ms;
      *best = cand[i];
    }
  }
  ctx->prof = prof;
  return err;
}
