// matmul.cpp -- numc_matmul and friends: validation, residency staging,
// family/config choice and launch.
//
// Mirrors numc's src/matmul/matmul.c:333-389 (entry points) and
// src/matmul/dispatch.h:20-59 (validation: NULL -> dtype -> ndim -> inner
// dimension -> output shape, same codes and messages).  Where numc passes only
// rsa, csa, rsb and rso to its kernels (matmul.c:359-363, 372-377), every
// stride of a, b and out is honoured here.
#include "gemm.h"
#include "internal.h"

#include <cmath>
#include <cstring>

static int check_matmul(const NumcArray *a, const NumcArray *b, const NumcArray *out, bool widen) {
  if (!a || !b || !out)
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "matmul: NULL pointer (a=%p b=%p out=%p)", (const void *)a,
                          (const void *)b, (const void *)out);
  bool types_ok = a->dtype == b->dtype && a->dtype == out->dtype;
  if (!types_ok && widen)
    types_ok = a->dtype == b->dtype &&
               ((a->dtype == NUMC_DTYPE_INT8 || a->dtype == NUMC_DTYPE_UINT8) &&
                (out->dtype == NUMC_DTYPE_INT32 || out->dtype == NUMC_DTYPE_UINT32));
  if (!types_ok)
    return NUMC_SET_ERROR(NUMC_ERR_TYPE, "matmul: dtype mismatch (a=%d b=%d out=%d)", a->dtype, b->dtype,
                          out->dtype);
  if (a->dim != 2 || b->dim != 2 || out->dim != 2)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "matmul: ndim mismatch (a.dim=%zu b.dim=%zu out.dim=%zu)", a->dim,
                          b->dim, out->dim);
  if (a->shape[1] != b->shape[0])
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE,
                          "matmul: shape mismatch (a.shape[1]=%zu b.shape[0]=%zu) expected a is (M, K) "
                          "and b is (K, N)",
                          a->shape[1], b->shape[0]);
  if (out->shape[0] != a->shape[0] || out->shape[1] != b->shape[1])
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE,
                          "matmul: shape mismatch (out.shape[0]=%zu out.shape[1]=%zu) expected a is(M, K) "
                          "and b is (K, N) and out is (M, N)",
                          out->shape[0], out->shape[1]);
  const size_t lim = (size_t)1 << 30;
  if (a->shape[0] >= lim || a->shape[1] >= lim || b->shape[1] >= lim)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "matmul: dimensions must be below 2^30");
  return 0;
}

// which kernel family runs (or -1 for the 64-bit integer kernel)
static int pick_family_dt(NumcDType dt, const NumcMatmulOpts *o) {
  int method = o ? o->method : NUMC_METHOD_AUTO;
  switch (dt) {
  case NUMC_DTYPE_FLOAT32:
    if (method == NUMC_METHOD_SPLIT_BF16) return NUMC_FAMILY_SPLIT_BF16;
    if (method == NUMC_METHOD_FP8_SCALED) return NUMC_FAMILY_FP8;
    if (method == NUMC_METHOD_AUTO || method == NUMC_METHOD_PLAIN) return NUMC_FAMILY_F32;
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: unknown method %d", method);
  case NUMC_DTYPE_FLOAT64:
  case NUMC_DTYPE_INT64:
  case NUMC_DTYPE_UINT64:
  default:
    if (method != NUMC_METHOD_AUTO && method != NUMC_METHOD_PLAIN)
      return NUMC_SET_ERROR(NUMC_ERR_TYPE, "matmul: method %d needs float32 operands", method);
    if (dt == NUMC_DTYPE_FLOAT64) return NUMC_FAMILY_F64;
    if (dt == NUMC_DTYPE_INT64 || dt == NUMC_DTYPE_UINT64) return 100;
    return NUMC_FAMILY_I8;
  }
}

static int pick_family(const NumcArray *a, const NumcMatmulOpts *o) { return pick_family_dt(a->dtype, o); }

// ---- residency staging -----------------------------------------------------

static int stage_in(NumcCtx *ctx, const NumcArray *x, const void **dptr) {
  unsigned char *d = numc_dev_ptr(x);
  if (!d) return numc_get_error().code ? numc_get_error().code : NUMC_ERR_INTERNAL;
  if (x->st->residency == RES_HOST) {
    size_t span = numc_view_span(x);
    if (span) HIP_TRY(hipMemcpyAsync(d, x->data, span, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
  }
  *dptr = d;
  return 0;
}

static int stage_out_begin(NumcCtx *ctx, const NumcArray *x, void **dptr) {
  unsigned char *d = numc_dev_ptr(x);
  if (!d) return numc_get_error().code ? numc_get_error().code : NUMC_ERR_INTERNAL;
  if (x->st->residency == RES_HOST) {
    size_t span = numc_view_span(x);
    // a view with gaps: bring the gaps over so the download keeps them
    if (span > x->size * x->elem_size)
      HIP_TRY(hipMemcpyAsync(d, x->data, span, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
  }
  *dptr = d;
  return 0;
}

static int stage_out_end(NumcCtx *ctx, const NumcArray *x, void *d) {
  if (x->st->residency == RES_HOST) {
    size_t span = numc_view_span(x);
    if (span) HIP_TRY(hipMemcpyAsync(x->data, d, span, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy D2H");
  }
  return 0;
}

// ---- config choice ------------------------------------------------------------

int numc_selector_choose(int family, int M, int N, int K, int flags); // selector.cpp

// Analytic default: estimated time = rounds of 256 CUs x per-block work at a
// tile-dependent efficiency, plus the split-K reduction traffic.
static int heuristic_config(int family, int M, int N, int K) {
  int best = 0;
  double best_t = 1e300;
  for (int i = 0; i < numc_gemm_config_count(family); i++) {
    const GemmConfig *c = numc_gemm_config(family, i);
    double nbm = std::ceil((double)M / c->cfg.bm), nbn = std::ceil((double)N / c->cfg.bn);
    double blocks = nbm * nbn * c->cfg.split_k;
    double rounds = std::ceil(blocks / 256.0);
    double area = (double)c->cfg.bm * c->cfg.bn;
    double eff = area >= 65536 ? 1.0 : area >= 32768 ? 0.9 : area >= 16384 ? 0.8 : 0.45;
    double kb = std::ceil((double)K / c->cfg.split_k / c->cfg.kc) * c->cfg.kc;
    double t = rounds * area * kb / eff + 2000.0 * rounds;
    if (c->cfg.split_k > 1) t += (double)M * N * c->cfg.split_k * 0.5 + 20000.0;
    if (t < best_t) {
      best_t = t;
      best = i;
    }
  }
  return best;
}

extern "C" NUMC_API int numc_selector_heuristic(int family, int M, int N, int K) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "bad family %d", family);
  return heuristic_config(family, M, N, K);
}

int numc_select_config(int family, int M, int N, int K, int flags) {
  int s = numc_selector_choose(family, M, N, K, flags);
  if (s >= 0 && s < numc_gemm_config_count(family)) return s;
  return heuristic_config(family, M, N, K);
}

extern "C" NUMC_API int numc_selector_choice(int family, int M, int N, int K, int flags) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "bad family %d", family);
  return numc_select_config(family, M, N, K, flags);
}

// ---- entry points ---------------------------------------------------------------

static inline long long es_stride(size_t bytes, size_t es) { return (long long)(bytes / es); }

// configuration choice, workspaces and launch of one problem whose operands
// are already on the device
static int launch_problem(NumcCtx *ctx, const GemmProblem &p, const NumcMatmulOpts *opts) {
  if (p.family == 100) return numc_i64_launch(ctx, p);
  // layout flags as the autotuner measures them: A stored K x M, B stored N x K, padded rows
  const int lay = (p.csa != 1 ? 1 : 0) | (p.csb != 1 ? 2 : 0) |
                  ((p.csa == 1 && p.rsa > p.K) || (p.csb == 1 && p.rsb > p.N) ? 4 : 0);
  int cfg_idx = opts && opts->config >= 0 ? opts->config : numc_select_config(p.family, p.M, p.N, p.K, lay);
  const GemmConfig *cfg = numc_gemm_config(p.family, cfg_idx);
  if (!cfg) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: no config %d for family %d", cfg_idx, p.family);
  size_t aux = 0, need = numc_gemm_workspace(p, *cfg, &aux);
  int err;
  if ((err = numc_ws_reserve(ctx, &ctx->ws, need))) return err;
  if ((err = numc_ws_reserve(ctx, &ctx->ws_aux, aux))) return err;
  return numc_gemm_launch(ctx, p, *cfg, ctx->ws.ptr, ctx->ws_aux.ptr);
}

int numc_matmul_one(const NumcArray *a, const NumcArray *b, NumcArray *out, const NumcMatmulOpts *opts,
                    bool finish) {
  const bool widen = opts && (opts->flags & NUMC_OPT_WIDEN);
  int err = check_matmul(a, b, out, widen);
  if (err) return err;
  int family = pick_family(a, opts);
  if (family < 0) return family;
  NumcCtx *ctx = out->ctx;
  if ((err = numc_gpu_init(ctx))) return err;

  GemmProblem p{};
  p.family = family;
  p.in_dt = a->dtype;
  p.out_dt = out->dtype;
  p.M = (int)a->shape[0];
  p.K = (int)a->shape[1];
  p.N = (int)b->shape[1];
  const size_t ein = a->elem_size, eout = out->elem_size;
  p.rsa = es_stride(a->strides[0], ein);
  p.csa = es_stride(a->strides[1], ein);
  p.rsb = es_stride(b->strides[0], ein);
  p.csb = es_stride(b->strides[1], ein);
  p.rso = es_stride(out->strides[0], eout);
  p.cso = es_stride(out->strides[1], eout);
  if (p.M == 0 || p.N == 0) return 0;

  if ((err = stage_in(ctx, a, &p.A))) return err;
  if ((err = stage_in(ctx, b, &p.B))) return err;
  if ((err = stage_out_begin(ctx, out, &p.C))) return err;

  if ((err = launch_problem(ctx, p, opts))) return err;
  if ((err = stage_out_end(ctx, out, p.C))) return err;
  return finish ? numc_finish(ctx) : 0;
}

static int run_matmul(const NumcArray *a, const NumcArray *b, NumcArray *out, const NumcMatmulOpts *opts) {
  return numc_matmul_one(a, b, out, opts, true);
}

extern "C" NUMC_API int numc_matmul(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  return run_matmul(a, b, out, nullptr);
}

extern "C" NUMC_API int numc_matmul_ex(const NumcArray *a, const NumcArray *b, NumcArray *out,
                                       const NumcMatmulOpts *opts) {
  return run_matmul(a, b, out, opts);
}

extern "C" NUMC_API int numc_matmul_family(const NumcArray *a, const NumcArray *b, const NumcArray *out,
                                           const NumcMatmulOpts *opts) {
  int err = check_matmul(a, b, out, opts && (opts->flags & NUMC_OPT_WIDEN));
  if (err) return err;
  return pick_family(a, opts);
}

extern "C" NUMC_API int numc_matmul_naive(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  int err = check_matmul(a, b, out, false);
  if (err) return err;
  NumcCtx *ctx = out->ctx;
  if ((err = numc_gpu_init(ctx))) return err;
  const int M = (int)a->shape[0], K = (int)a->shape[1], N = (int)out->shape[1];
  // numc_matmul_naive reads data as contiguous whatever the strides are
  // (matmul.c:333-343): use the same bytes
  auto dense_in = [&](const NumcArray *x, const void **d) -> int {
    const ArenaBlock *blk = x->st->block;
    if ((unsigned char *)x->data + x->size * x->elem_size > blk->host + blk->capacity)
      return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "matmul_naive: view reads past its storage (use a contiguous copy)");
    unsigned char *p = numc_dev_ptr(x);
    if (!p) return NUMC_ERR_INTERNAL;
    if (x->st->residency == RES_HOST && x->size)
      HIP_TRY(hipMemcpyAsync(p, x->data, x->size * x->elem_size, hipMemcpyHostToDevice, ctx->stream), "H2D");
    *d = p;
    return 0;
  };
  const void *da, *db;
  if ((err = dense_in(a, &da)) || (err = dense_in(b, &db))) return err;
  const ArenaBlock *ob = out->st->block;
  if ((unsigned char *)out->data + out->size * out->elem_size > ob->host + ob->capacity)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "matmul_naive: out view extends past its storage");
  unsigned char *dc = numc_dev_ptr(out);
  if (!dc) return NUMC_ERR_INTERNAL;
  if ((err = numc_naive_launch(ctx, a->dtype, da, db, dc, M, K, N))) return err;
  if (out->st->residency == RES_HOST && out->size)
    HIP_TRY(hipMemcpyAsync(out->data, dc, out->size * out->elem_size, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  return numc_finish(ctx);
}

// ---- pointer-level entry point ----------------------------------------------------
//
// For a host library that keeps its own arrays: numc's numc_matmul
// (src/matmul/matmul.c:347-389) would call this after its own _check_matmul
// with its arrays' data pointers, shapes and byte strides.

static size_t span_bytes(size_t rows, size_t cols, size_t rs, size_t cs, size_t es) {
  if (rows == 0 || cols == 0) return 0;
  return (rows - 1) * rs + (cols - 1) * cs + es;
}

extern "C" NUMC_API int numc_amd_gemm(NumcCtx *ctx, int dtype, int out_dtype, size_t M, size_t N, size_t K,
                                      const void *a, size_t rsa, size_t csa, const void *b, size_t rsb,
                                      size_t csb, void *c, size_t rso, size_t cso, const NumcMatmulOpts *opts) {
  if (!ctx || (M && K && !a) || (K && N && !b) || (M && N && !c))
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "amd_gemm: NULL pointer (ctx=%p a=%p b=%p c=%p)", (void *)ctx, a, b, c);
  const bool widen = opts && (opts->flags & NUMC_OPT_WIDEN);
  bool types_ok = dtype >= 0 && dtype <= NUMC_DTYPE_FLOAT64 && (dtype == out_dtype);
  if (!types_ok && widen)
    types_ok = (dtype == NUMC_DTYPE_INT8 || dtype == NUMC_DTYPE_UINT8) &&
               (out_dtype == NUMC_DTYPE_INT32 || out_dtype == NUMC_DTYPE_UINT32);
  if (!types_ok) return NUMC_SET_ERROR(NUMC_ERR_TYPE, "amd_gemm: dtype mismatch (in=%d out=%d)", dtype, out_dtype);
  const size_t lim = (size_t)1 << 30;
  if (M >= lim || N >= lim || K >= lim) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "amd_gemm: dimensions must be below 2^30");
  const size_t ein = numc_dtype_size((NumcDType)dtype), eout = numc_dtype_size((NumcDType)out_dtype);
  if (rsa % ein || csa % ein || rsb % ein || csb % ein || rso % eout || cso % eout)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "amd_gemm: strides must be multiples of the element size");
  int family = pick_family_dt((NumcDType)dtype, opts);
  if (family < 0) return family;
  if (M == 0 || N == 0) return 0;
  int err;
  if ((err = numc_gpu_init(ctx))) return err;

  GemmProblem p{};
  p.family = family;
  p.in_dt = (NumcDType)dtype;
  p.out_dt = (NumcDType)out_dtype;
  p.M = (int)M;
  p.N = (int)N;
  p.K = (int)K;
  p.rsa = es_stride(rsa, ein);
  p.csa = es_stride(csa, ein);
  p.rsb = es_stride(rsb, ein);
  p.csb = es_stride(csb, ein);
  p.rso = es_stride(rso, eout);
  p.cso = es_stride(cso, eout);
  const bool device = opts && (opts->flags & NUMC_OPT_DEVICE_POINTERS);
  if (device) {
    p.A = a;
    p.B = b;
    p.C = c;
  } else {
    // host operands: stage the spans the strided views cover
    const size_t sa = span_bytes(M, K, rsa, csa, ein), sb = span_bytes(K, N, rsb, csb, ein);
    const size_t sc = span_bytes(M, N, rso, cso, eout);
    const size_t oa = 0, ob = (sa + 255) & ~(size_t)255, oc = ob + ((sb + 255) & ~(size_t)255);
    if ((err = numc_ws_reserve(ctx, &ctx->ws_stage, oc + sc))) return err;
    unsigned char *d = (unsigned char *)ctx->ws_stage.ptr;
    if (sa) HIP_TRY(hipMemcpyAsync(d + oa, a, sa, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
    if (sb) HIP_TRY(hipMemcpyAsync(d + ob, b, sb, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
    // a view with gaps: bring the gaps over so the download keeps them
    if (sc > M * N * eout) HIP_TRY(hipMemcpyAsync(d + oc, c, sc, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
    p.A = d + oa;
    p.B = d + ob;
    p.C = d + oc;
  }
  if ((err = launch_problem(ctx, p, opts))) return err;
  if (!device) {
    const size_t sc = span_bytes(M, N, rso, cso, eout);
    HIP_TRY(hipMemcpyAsync(c, p.C, sc, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy D2H");
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return 0;
  }
  return numc_finish(ctx);
}
