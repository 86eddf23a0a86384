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
#include <cstdlib>
#include <cstring>
#include <vector>

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
    if (dt == NUMC_DTYPE_BFLOAT16) return NUMC_FAMILY_BF16;
    if (dt == NUMC_DTYPE_FLOAT16) return NUMC_FAMILY_F16;
    if (dt == NUMC_DTYPE_INT64 || dt == NUMC_DTYPE_UINT64) return 100;
    return NUMC_FAMILY_I8;
  }
}

static int pick_family(const NumcArray *a, const NumcMatmulOpts *o) { return pick_family_dt(a->dtype, o); }

int numc_matmul_family_dt(int dtype, const NumcMatmulOpts *opts) { return pick_family_dt((NumcDType)dtype, opts); }

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

int numc_selector_choose(int family, int M, int N, int K, int flags, bool allow_direct); // selector.cpp
void numc_selector_record(int family, int M, int N, int K, int flags, int cfg);
extern "C" NUMC_API int numc_selector_candidates(int family, int M, int N, int K, int flags, int allow_direct, int *out,
                                                 int max);

// timing the model's close candidates on a call (once per exact shape); the
// environment variable NUMC_AUTOTUNE=off keeps to the model's best guess
static bool call_tuning_enabled() {
  static const bool on = [] {
    const char *e = getenv("NUMC_AUTOTUNE");
    return !(e && strcmp(e, "off") == 0);
  }();
  return on;
}

// Analytic default (used without a trained cost model).  Direct
// configurations where K <= 2^kt and M*N <= 2^mt for the family -- numc's
// gemmsup rule (src/matmul/matmul.c:241-256) with thresholds measured on this
// GPU: the values minimise the mean regret over the 20000 training shapes of
// the autotuner's database (numc_amd/data/autotune_db.csv.gz) -- with 64x64
// tiles when they alone make 256 tiles, else 32x32.  Otherwise a packed
// configuration by rounds of 256 CUs x per-block work at a tile-dependent
// efficiency, plus the split-K reduction.
static const int kDirectLog2K[NUMC_FAMILY_COUNT] = {13, 12, -1, 14, 13, 12, 12};
static const int kDirectLog2MN[NUMC_FAMILY_COUNT] = {24, 24, -1, 22, 22, 22, 22};

static int heuristic_config(int family, int M, int N, int K, bool allow_direct) {
  if (allow_direct && kDirectLog2K[family] >= 0 && (long long)K <= (1LL << kDirectLog2K[family]) &&
      (long long)M * N <= (1LL << kDirectLog2MN[family])) {
    int small = -1, big = -1;
    for (int i = 0; i < numc_gemm_config_count(family); i++) {
      const GemmConfig *c = numc_gemm_config(family, i);
      if (!c->cfg.direct) continue;
      if (small < 0 || c->cfg.bm < numc_gemm_config(family, small)->cfg.bm) small = i;
      if (big < 0 || c->cfg.bm > numc_gemm_config(family, big)->cfg.bm) big = i;
    }
    if (big >= 0) {
      const int bm = numc_gemm_config(family, big)->cfg.bm;
      const long long tiles = ((long long)M + bm - 1) / bm * (((long long)N + bm - 1) / bm);
      return tiles >= 256 ? big : small;
    }
  }
  int best = 0;
  double best_t = 1e300;
  for (int i = 0; i < numc_gemm_config_count(family); i++) {
    const GemmConfig *c = numc_gemm_config(family, i);
    if (c->cfg.direct) continue;
    const double nbm = std::ceil((double)M / c->cfg.bm), nbn = std::ceil((double)N / c->cfg.bn);
    const double blocks = nbm * nbn * c->cfg.split_k;
    const double rounds = std::ceil(blocks / 256.0);
    const double area = (double)c->cfg.bm * c->cfg.bn;
    const double eff = area >= 65536 ? 1.0 : area >= 32768 ? 0.9 : area >= 16384 ? 0.8 : 0.45;
    const double kb = std::ceil((double)K / c->cfg.split_k / c->cfg.kc) * c->cfg.kc;
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
  return heuristic_config(family, M, N, K, true);
}

int numc_select_config(int family, int M, int N, int K, int flags, bool allow_direct) {
  int s = numc_selector_choose(family, M, N, K, flags, allow_direct);
  if (s >= 0 && s < numc_gemm_config_count(family)) return s;
  return heuristic_config(family, M, N, K, allow_direct);
}

extern "C" NUMC_API int numc_selector_choice(int family, int M, int N, int K, int flags) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "bad family %d", family);
  return numc_select_config(family, M, N, K, flags, true);
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
  const bool direct_ok = numc_direct_ok(p);
  int cfg_idx;
  if (opts && opts->config >= 0) {
    cfg_idx = opts->config;
  } else {
    // the learned choice; where the model cannot separate its best few
    // configurations, time them once on this call and keep the winner for the
    // shape (the exhaustive autotuner's rule, on the candidates only).  An
    // integer problem's digit-plane count is part of the shape's entry.
    const int planes = numc_problem_planes(p);
    const int key = lay | (planes > 1 ? planes << 4 : 0);
    int cand[32];
    const int nc = numc_selector_candidates(p.family, p.M, p.N, p.K, key, direct_ok, cand, 32);
    if (nc > 1 && call_tuning_enabled()) {
      int err = numc_tune_call(ctx, p, cand, nc, &cfg_idx);
      if (err) return err;
      numc_selector_record(p.family, p.M, p.N, p.K, key, cfg_idx);
    } else if (nc >= 1) {
      cfg_idx = cand[0];
    } else {
      cfg_idx = heuristic_config(p.family, p.M, p.N, p.K, direct_ok);
    }
  }
  const GemmConfig *cfg = numc_gemm_config(p.family, cfg_idx);
  if (!cfg) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: no config %d for family %d", cfg_idx, p.family);
  if (cfg->cfg.direct && !direct_ok)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE,
                          "matmul: config %d reads the operands in place, which needs one int8 digit plane "
                          "(int8 inputs, or a one-byte result)", cfg_idx);
  size_t aux = 0, need = numc_gemm_workspace(p, *cfg, &aux);
  int err;
  if ((err = numc_ws_reserve(ctx, &ctx->ws, need))) return err;
  if ((err = numc_ws_reserve(ctx, &ctx->ws_aux, aux))) return err;
  return numc_gemm_launch(ctx, p, *cfg, ctx->ws.ptr, ctx->ws_aux.ptr);
}

// one 2-D problem out = a @ b (+ bias[j] on every row when bias is given, a
// 1-D array of the output dtype, added in the GEMM's store)
static int matmul_bias(const NumcArray *a, const NumcArray *b, const NumcArray *bias, NumcArray *out,
                       const NumcMatmulOpts *opts, bool finish) {
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
  if (bias) {
    if ((err = stage_in(ctx, bias, &p.bias))) return err;
    p.bias_s = es_stride(bias->strides[0], eout);
  }
  if ((err = stage_out_begin(ctx, out, &p.C))) return err;

  if ((err = launch_problem(ctx, p, opts))) return err;
  if ((err = stage_out_end(ctx, out, p.C))) return err;
  return finish ? numc_finish(ctx) : 0;
}

int numc_matmul_one(const NumcArray *a, const NumcArray *b, NumcArray *out, const NumcMatmulOpts *opts,
                    bool finish) {
  return matmul_bias(a, b, nullptr, out, opts, finish);
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
  bool types_ok = dtype >= 0 && dtype < NUMC_DTYPE_COUNT && (dtype == out_dtype);
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

// ---- numc_dot ----------------------------------------------------------------------
//
// numc's numc_dot (src/reduction/dot.c:441-448; validation
// src/reduction/dispatch.h:199-283) on the GPU:
//   0-D with anything    out = other * scalar in the dtype's arithmetic, the
//                        scalar passing through double as numc's
//                        _dot_scalar_case -> numc_mul_scalar does
//                        (dot.c:340-346);
//   1-D . 1-D            inner product (a 1 x K by K x 1 GEMM);
//   N-D . 1-D, N-D . M-D sum over a's last axis and b's second-to-last:
//                        out = a.shape[:-1] + b.shape[:-2] + b.shape[-1:],
//                        one GEMM per run of a's leading indices that is
//                        linear in both a and out, and per index of b's
//                        leading axes; several of them go to the persistent
//                        batched kernel in one launch.
// Integers give the exact sum mod 2^w as numc's reduction and naive kernels
// do (dot.c:22-24, 43-69); floats compute in their own precision.  Where numc
// flattens a's leading axes with a->strides[0] as the row stride
// (dot.c:371-372, correct only for 2-D a) and divides by a zero-length last
// axis (dot.c:368), this follows the numpy semantics its checker describes.

static int check_dot(const NumcArray *a, const NumcArray *b, const NumcArray *out) {
  if (!a || !b || !out)
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "dot: NULL pointer (a=%p b=%p out=%p)", (const void *)a, (const void *)b,
                          (const void *)out);
  if (a->dtype != b->dtype || a->dtype != out->dtype)
    return NUMC_SET_ERROR(NUMC_ERR_TYPE, "dot: dtype mismatch (a=%d b=%d out=%d)", a->dtype, b->dtype, out->dtype);
  if (a->dim == 0 || b->dim == 0) {
    const NumcArray *other = a->dim == 0 ? b : a;
    if (out->dim != other->dim)
      return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "dot: output ndim mismatch (out.dim=%zu other.dim=%zu)", out->dim,
                            other->dim);
    for (size_t i = 0; i < out->dim; i++)
      if (out->shape[i] != other->shape[i])
        return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "dot: shape mismatch at axis %zu", i);
    return 0;
  }
  const size_t al = a->dim - 1, br = b->dim == 1 ? 0 : b->dim - 2;
  if (a->shape[al] != b->shape[br])
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "dot: axis mismatch (a.shape[%zu]=%zu b.shape[%zu]=%zu)", al,
                          a->shape[al], br, b->shape[br]);
  const size_t expected = (a->dim - 1) + (b->dim > 1 ? b->dim - 1 : 0);
  const bool scalar_out = expected == 0 && (out->dim == 0 || (out->dim == 1 && out->size == 1));
  if (!scalar_out && out->dim != expected)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "dot: output ndim mismatch (out.dim=%zu expected=%zu)", out->dim, expected);
  if (!scalar_out) {
    size_t o = 0;
    bool ok = true;
    for (size_t i = 0; i + 1 < a->dim; i++) ok = ok && out->shape[o++] == a->shape[i];
    if (b->dim > 1) {
      for (size_t i = 0; i + 2 < b->dim; i++) ok = ok && out->shape[o++] == b->shape[i];
      ok = ok && out->shape[o++] == b->shape[b->dim - 1];
    }
    if (!ok) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "dot: output shape mismatch");
  }
  return 0;
}

// the scalar of a 0-D operand in the dtype's representation after numc's
// round trip through double (exact except for 64-bit integers beyond 2^53;
// out-of-range results as the x86-64 conversion sequences give them)
static int dot_scalar_bits(const NumcArray *sc, uint64_t *bits) {
  unsigned char raw[8] = {0};
  if (sc->st->residency == RES_DEVICE) {
    unsigned char *d = numc_dev_ptr(sc);
    if (!d) return NUMC_ERR_INTERNAL;
    HIP_TRY(hipMemcpyAsync(raw, d, sc->elem_size, hipMemcpyDeviceToHost, sc->ctx->stream), "hipMemcpy D2H");
    HIP_TRY(hipStreamSynchronize(sc->ctx->stream), "hipStreamSynchronize");
  } else {
    if (sc->ctx->gpu_ready && !sc->ctx->synchronous)
      HIP_TRY(hipStreamSynchronize(sc->ctx->stream), "hipStreamSynchronize");
    memcpy(raw, sc->data, sc->elem_size);
  }
  double v = 0;
  switch (sc->dtype) {
  case NUMC_DTYPE_INT8: v = *(int8_t *)raw; break;
  case NUMC_DTYPE_INT16: v = *(int16_t *)raw; break;
  case NUMC_DTYPE_INT32: v = *(int32_t *)raw; break;
  case NUMC_DTYPE_INT64: v = (double)*(int64_t *)raw; break;
  case NUMC_DTYPE_UINT8: v = *(uint8_t *)raw; break;
  case NUMC_DTYPE_UINT16: v = *(uint16_t *)raw; break;
  case NUMC_DTYPE_UINT32: v = *(uint32_t *)raw; break;
  case NUMC_DTYPE_UINT64: v = (double)*(uint64_t *)raw; break;
  case NUMC_DTYPE_FLOAT32: v = *(float *)raw; break;
  case NUMC_DTYPE_FLOAT64: v = *(double *)raw; break;
  case NUMC_DTYPE_BFLOAT16: {
    const uint32_t u = (uint32_t)*(uint16_t *)raw << 16;
    float f;
    memcpy(&f, &u, 4);
    v = f;
    break;
  }
  case NUMC_DTYPE_FLOAT16: {
    _Float16 h;
    memcpy(&h, raw, 2);
    v = (double)h;
    break;
  }
  }
  switch (sc->dtype) {
  case NUMC_DTYPE_INT64:
    *bits = v >= 9223372036854775808.0 ? (uint64_t)1 << 63 : (uint64_t)(int64_t)v;
    break;
  case NUMC_DTYPE_UINT64:
    *bits = v >= 18446744073709551616.0 ? 0 : (uint64_t)v;
    break;
  case NUMC_DTYPE_FLOAT32:
  case NUMC_DTYPE_BFLOAT16:
  case NUMC_DTYPE_FLOAT16: { // 16-bit floats scale in binary32 (exact for the scalar), one rounding
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    *bits = u;
    break;
  }
  case NUMC_DTYPE_FLOAT64: memcpy(bits, &v, 8); break;
  case NUMC_DTYPE_UINT8: case NUMC_DTYPE_UINT16: case NUMC_DTYPE_UINT32: *bits = (uint64_t)v; break;
  default: *bits = (uint64_t)(int64_t)v; break; // int8/16/32: exact
  }
  return 0;
}

static int dot_scalar(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  const NumcArray *sc = a->dim == 0 ? a : b, *other = a->dim == 0 ? b : a;
  if (other->size == 0) return 0;
  if (other->dim > NUMC_EW_MAX_DIMS)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "dot: at most %d dimensions on the GPU path", NUMC_EW_MAX_DIMS);
  NumcCtx *ctx = out->ctx;
  int err = numc_gpu_init(ctx);
  if (err) return err;
  ScaleArgs s{};
  if ((err = dot_scalar_bits(sc, &s.scalar_bits))) return err;
  const void *src;
  void *dst;
  if ((err = stage_in(ctx, other, &src)) || (err = stage_out_begin(ctx, out, &dst))) return err;
  s.src = (const unsigned char *)src;
  s.dst = (unsigned char *)dst;
  s.n = (long long)other->size;
  s.ndim = (int)other->dim;
  s.dtype = other->dtype;
  for (size_t i = 0; i < other->dim; i++) {
    s.shape[i] = (long long)other->shape[i];
    s.ss[i] = (long long)(other->strides[i] / other->elem_size);
    s.ds[i] = (long long)(out->strides[i] / out->elem_size);
  }
  if ((err = numc_scale_launch(ctx, s))) return err;
  if ((err = stage_out_end(ctx, out, dst))) return err;
  return numc_finish(ctx);
}

// a 2-D view header sharing src's storage (no allocation)
static void view2(NumcArray *v, const NumcArray *src, size_t byte_off, size_t rows, size_t cols, size_t rstride,
                  size_t cstride) {
  memset(v, 0, sizeof *v);
  v->ctx = src->ctx;
  v->dtype = src->dtype;
  v->elem_size = src->elem_size;
  v->st = src->st;
  v->dim = 2;
  v->shape = v->shape_buf;
  v->strides = v->strides_buf;
  v->shape[0] = rows;
  v->shape[1] = cols;
  v->strides[0] = rstride;
  v->strides[1] = cstride;
  v->size = rows * cols;
  v->capacity = v->size;
  v->is_contiguous = cstride == v->elem_size && (rows <= 1 || rstride == cols * v->elem_size);
  v->data = (unsigned char *)src->data + byte_off;
}

static int dot_gemm(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  const size_t es = a->elem_size;
  const size_t k = a->shape[a->dim - 1];
  const size_t n = b->dim == 1 ? 1 : b->shape[b->dim - 1];
  const size_t na = a->dim - 1;                    // a's leading axes (out axes 0..na-1)
  const size_t nbb = b->dim >= 2 ? b->dim - 2 : 0; // b's batch axes (out axes na..na+nbb-1)
  const size_t csa = a->strides[a->dim - 1];
  const size_t rsb = b->strides[b->dim == 1 ? 0 : b->dim - 2];
  const size_t csb = b->dim == 1 ? es : b->strides[b->dim - 1];
  const size_t cso = b->dim == 1 ? es : out->strides[out->dim - 1];
  if (out->size == 0) return 0;

  // the trailing run of a's leading axes that is linear in a and in out
  size_t c = na;
  if (na > 0) {
    c = na - 1;
    while (c > 0 && a->strides[c - 1] == a->strides[c] * a->shape[c] &&
           out->strides[c - 1] == out->strides[c] * out->shape[c])
      c--;
  }
  size_t M = 1;
  for (size_t i = c; i < na; i++) M *= a->shape[i];
  const size_t rsa = na > 0 ? a->strides[na - 1] : k * es;
  const size_t rso = na > 0 ? out->strides[na - 1] : n * es;

  // every combination of the outer a axes and of b's batch axes is a problem
  size_t n_outer = 1, n_batch = 1;
  for (size_t i = 0; i < c; i++) n_outer *= a->shape[i];
  for (size_t i = 0; i < nbb; i++) n_batch *= b->shape[i];
  const size_t count = n_outer * n_batch;
  std::vector<NumcArray> views(3 * count);
  std::vector<NumcArray *> pa(count), pb(count), po(count);
  for (size_t t = 0; t < count; t++) {
    size_t oi = t / n_batch, bi = t % n_batch;
    size_t aoff = 0, boff = 0, ooff = 0;
    for (size_t i = c; i-- > 0;) { // outer a axes
      const size_t idx = oi % a->shape[i];
      oi /= a->shape[i];
      aoff += idx * a->strides[i];
      ooff += idx * out->strides[i];
    }
    for (size_t i = nbb; i-- > 0;) { // b's batch axes
      const size_t idx = bi % b->shape[i];
      bi /= b->shape[i];
      boff += idx * b->strides[i];
      ooff += idx * out->strides[na + i];
    }
    view2(&views[3 * t], a, aoff, M, k, rsa, csa);
    view2(&views[3 * t + 1], b, boff, k, n, rsb, csb);
    view2(&views[3 * t + 2], out, ooff, M, n, rso, cso);
    pa[t] = &views[3 * t];
    pb[t] = &views[3 * t + 1];
    po[t] = &views[3 * t + 2];
  }
  if (count == 1) return numc_matmul_one(pa[0], pb[0], po[0], nullptr, true);
  return numc_matmul_batched(out->ctx, count, pa.data(), pb.data(), po.data(), nullptr);
}

extern "C" NUMC_API int numc_dot(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  int err = check_dot(a, b, out);
  if (err) return err;
  if (a->dim == 0 || b->dim == 0) return dot_scalar(a, b, out);
  return dot_gemm(a, b, out);
}

// ===========================================================================
// numc_linear and numc_bmm: the GEMM consumers numc's roadmap names
// (ROADMAP.md:139-140).  numc has no implementation of either, so the
// validation below follows numc_matmul's order (NULL -> dtype -> ndim ->
// inner dimension -> output shape) and message style.
// ===========================================================================

static int check_linear(const NumcArray *x, const NumcArray *w, const NumcArray *bias, const NumcArray *out) {
  if (!x || !w || !out)
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "linear: NULL pointer (x=%p w=%p out=%p)", (const void *)x,
                          (const void *)w, (const void *)out);
  if (x->dtype != w->dtype || x->dtype != out->dtype || (bias && bias->dtype != out->dtype))
    return NUMC_SET_ERROR(NUMC_ERR_TYPE, "linear: dtype mismatch (x=%d w=%d bias=%d out=%d)", x->dtype, w->dtype,
                          bias ? (int)bias->dtype : -1, out->dtype);
  if (x->dim < 1 || w->dim != 2 || out->dim != x->dim || (bias && bias->dim != 1))
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE,
                          "linear: ndim mismatch (x.dim=%zu w.dim=%zu bias.dim=%zu out.dim=%zu) expected x "
                          "(..., K), w (K, N), bias (N), out (..., N)",
                          x->dim, w->dim, bias ? bias->dim : (size_t)1, out->dim);
  const size_t K = x->shape[x->dim - 1], N = w->shape[1];
  if (w->shape[0] != K)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "linear: shape mismatch (x.shape[-1]=%zu w.shape[0]=%zu)", K,
                          w->shape[0]);
  if (bias && bias->shape[0] != N)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "linear: shape mismatch (bias.shape[0]=%zu w.shape[1]=%zu)",
                          bias->shape[0], N);
  for (size_t i = 0; i + 1 < x->dim; i++)
    if (out->shape[i] != x->shape[i])
      return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "linear: shape mismatch (out.shape[%zu]=%zu x.shape[%zu]=%zu)", i,
                            out->shape[i], i, x->shape[i]);
  if (out->shape[out->dim - 1] != N)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "linear: shape mismatch (out.shape[-1]=%zu w.shape[1]=%zu)",
                          out->shape[out->dim - 1], N);
  return 0;
}

extern "C" NUMC_API int numc_linear_ex(const NumcArray *x, const NumcArray *w, const NumcArray *bias, NumcArray *out,
                                       const NumcMatmulOpts *opts) {
  int err = check_linear(x, w, bias, out);
  if (err) return err;
  if (opts && (opts->method != NUMC_METHOD_AUTO && opts->method != NUMC_METHOD_PLAIN))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "linear: method %d has no biased store (use AUTO or PLAIN)", opts->method);
  if (opts && (opts->flags & NUMC_OPT_WIDEN))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "linear: NUMC_OPT_WIDEN is not supported");
  if (out->size == 0) return 0;
  const size_t es = x->elem_size;
  const size_t K = x->shape[x->dim - 1], N = w->shape[1];
  const size_t na = x->dim - 1; // leading axes of x (and of out)
  const size_t csa = x->strides[x->dim - 1], cso = out->strides[out->dim - 1];
  // the trailing run of leading axes that is linear in x and in out folds into M
  size_t c = na;
  if (na > 0) {
    c = na - 1;
    while (c > 0 && x->strides[c - 1] == x->strides[c] * x->shape[c] &&
           out->strides[c - 1] == out->strides[c] * out->shape[c])
      c--;
  }
  size_t M = 1, n_outer = 1;
  for (size_t i = c; i < na; i++) M *= x->shape[i];
  for (size_t i = 0; i < c; i++) n_outer *= x->shape[i];
  const size_t rsa = na > 0 ? x->strides[na - 1] : K * es;
  const size_t rso = na > 0 ? out->strides[na - 1] : N * out->elem_size;
  NumcArray va, vo;
  for (size_t t = 0; t < n_outer; t++) {
    size_t oi = t, aoff = 0, ooff = 0;
    for (size_t i = c; i-- > 0;) {
      const size_t idx = oi % x->shape[i];
      oi /= x->shape[i];
      aoff += idx * x->strides[i];
      ooff += idx * out->strides[i];
    }
    view2(&va, x, aoff, M, K, rsa, csa);
    view2(&vo, out, ooff, M, N, rso, cso);
    if ((err = matmul_bias(&va, w, bias, &vo, opts, false))) return err;
  }
  return numc_finish(out->ctx);
}

extern "C" NUMC_API int numc_linear(const NumcArray *x, const NumcArray *w, const NumcArray *bias, NumcArray *out) {
  return numc_linear_ex(x, w, bias, out, nullptr);
}

static int check_bmm(const NumcArray *a, const NumcArray *b, const NumcArray *out) {
  if (!a || !b || !out)
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "bmm: NULL pointer (a=%p b=%p out=%p)", (const void *)a, (const void *)b,
                          (const void *)out);
  if (a->dtype != b->dtype || a->dtype != out->dtype)
    return NUMC_SET_ERROR(NUMC_ERR_TYPE, "bmm: dtype mismatch (a=%d b=%d out=%d)", a->dtype, b->dtype, out->dtype);
  const size_t nd = a->dim > b->dim ? a->dim : b->dim;
  if (a->dim < 2 || b->dim < 2 || out->dim != nd)
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE,
                          "bmm: ndim mismatch (a.dim=%zu b.dim=%zu out.dim=%zu) expected a (..., M, K), b (..., K, "
                          "N), out (..., M, N) with the broadcast batch axes",
                          a->dim, b->dim, out->dim);
  if (a->shape[a->dim - 1] != b->shape[b->dim - 2])
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "bmm: shape mismatch (a.shape[-1]=%zu b.shape[-2]=%zu)",
                          a->shape[a->dim - 1], b->shape[b->dim - 2]);
  // batch axes, aligned from the right: equal, or 1 in one operand (broadcast)
  for (size_t i = 0; i + 2 < nd; i++) {
    const size_t ia = i + a->dim, ib = i + b->dim; // index + nd - offset, kept unsigned
    const size_t da = ia >= nd ? a->shape[ia - nd] : 1, db = ib >= nd ? b->shape[ib - nd] : 1;
    if (da != db && da != 1 && db != 1)
      return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "bmm: batch axis %zu does not broadcast (a=%zu b=%zu)", i, da, db);
    if (out->shape[i] != (da == 1 ? db : da))
      return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "bmm: shape mismatch (out.shape[%zu]=%zu expected %zu)", i,
                            out->shape[i], da == 1 ? db : da);
  }
  if (out->shape[nd - 2] != a->shape[a->dim - 2] || out->shape[nd - 1] != b->shape[b->dim - 1])
    return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "bmm: shape mismatch (out.shape[-2:]=(%zu, %zu) expected (%zu, %zu))",
                          out->shape[nd - 2], out->shape[nd - 1], a->shape[a->dim - 2], b->shape[b->dim - 1]);
  return 0;
}

// The slice offset (elements) of operand x for the flattened batch index t of
// out, when it is t * S for every t (S = 0: one matrix for every slice).
// False when x's batch axes do not follow out's linearly (an axis
// broadcast inside a batched operand, or gaps between slices).
static bool batch_stride(const NumcArray *x, const NumcArray *out, long long *S) {
  const size_t nbo = out->dim - 2;
  long long inner = 1, s = 0;
  bool have = false;
  for (size_t i = nbo; i-- > 0;) {
    const long long n = (long long)out->shape[i];
    long long xs = 0;
    if (i + x->dim >= out->dim) {
      const size_t xi = i + x->dim - out->dim;
      if (x->shape[xi] != 1) xs = (long long)(x->strides[xi] / x->elem_size);
    }
    if (n > 1) {
      if (!have) {
        if (xs % inner) return false;
        s = xs / inner;
        have = true;
      } else if (xs != s * inner) {
        return false;
      }
    }
    inner *= n;
  }
  *S = s;
  return true;
}

// every slice of numc_bmm in one strided-batch GEMM (packs and MFMA kernel
// with blockIdx.z = slice); returns 1 when the operands do not qualify
static int bmm_strided(const NumcArray *a, const NumcArray *b, NumcArray *out, size_t count) {
  long long sa, sb, so;
  if (count < 2 || count > 65535 || !batch_stride(a, out, &sa) || !batch_stride(b, out, &sb) ||
      !batch_stride(out, out, &so) || so == 0)
    return 1;
  const int family = pick_family(a, nullptr);
  if (family < 0 || family == 100 || family == NUMC_FAMILY_FP8 || family == NUMC_FAMILY_SPLIT_BF16) return 1;
  const size_t es = a->elem_size;
  GemmProblem p{};
  p.family = family;
  p.in_dt = a->dtype;
  p.out_dt = out->dtype;
  p.M = (int)a->shape[a->dim - 2];
  p.K = (int)a->shape[a->dim - 1];
  p.N = (int)b->shape[b->dim - 1];
  if (p.K == 0) return 1;
  p.rsa = es_stride(a->strides[a->dim - 2], es);
  p.csa = es_stride(a->strides[a->dim - 1], es);
  p.rsb = es_stride(b->strides[b->dim - 2], es);
  p.csb = es_stride(b->strides[b->dim - 1], es);
  p.rso = es_stride(out->strides[out->dim - 2], out->elem_size);
  p.cso = es_stride(out->strides[out->dim - 1], out->elem_size);
  p.batch = (int)count;
  p.a_bs = sa;
  p.b_bs = sb;
  p.c_bs = so;
  const GemmConfig *cfg = numc_strided_batch_config(family, p.M, p.N);
  if (!cfg) return 1;
  NumcCtx *ctx = out->ctx;
  int err;
  if ((err = numc_gpu_init(ctx))) return err;
  if ((err = stage_in(ctx, a, &p.A)) || (err = stage_in(ctx, b, &p.B)) || (err = stage_out_begin(ctx, out, &p.C)))
    return err;
  size_t aux = 0, need = numc_gemm_workspace(p, *cfg, &aux);
  if ((err = numc_ws_reserve(ctx, &ctx->ws, need)) || (err = numc_ws_reserve(ctx, &ctx->ws_aux, aux))) return err;
  if ((err = numc_gemm_launch(ctx, p, *cfg, ctx->ws.ptr, ctx->ws_aux.ptr))) return err;
  if ((err = stage_out_end(ctx, out, p.C))) return err;
  return numc_finish(ctx);
}

extern "C" NUMC_API int numc_bmm(const NumcArray *a, const NumcArray *b, NumcArray *out) {
  int err = check_bmm(a, b, out);
  if (err) return err;
  const size_t nd = out->dim, nb = nd - 2;
  const size_t M = a->shape[a->dim - 2], K = a->shape[a->dim - 1], N = b->shape[b->dim - 1];
  size_t count = 1;
  for (size_t i = 0; i < nb; i++) count *= out->shape[i];
  if (count == 0 || M == 0 || N == 0) return 0;
  // equal slices at linear offsets: one strided-batch launch
  if ((err = bmm_strided(a, b, out, count)) != 1) return err;
  std::vector<NumcArray> views(3 * count);
  std::vector<NumcArray *> pa(count), pb(count), po(count);
  for (size_t t = 0; t < count; t++) {
    size_t r = t, aoff = 0, boff = 0, ooff = 0;
    for (size_t i = nb; i-- > 0;) { // out batch axis i; a's is i + a.dim - nd (if any), likewise b's
      const size_t idx = r % out->shape[i];
      r /= out->shape[i];
      ooff += idx * out->strides[i];
      if (i + a->dim >= nd) {
        const size_t ai = i + a->dim - nd;
        if (a->shape[ai] != 1) aoff += idx * a->strides[ai];
      }
      if (i + b->dim >= nd) {
        const size_t bi = i + b->dim - nd;
        if (b->shape[bi] != 1) boff += idx * b->strides[bi];
      }
    }
    view2(&views[3 * t], a, aoff, M, K, a->strides[a->dim - 2], a->strides[a->dim - 1]);
    view2(&views[3 * t + 1], b, boff, K, N, b->strides[b->dim - 2], b->strides[b->dim - 1]);
    view2(&views[3 * t + 2], out, ooff, M, N, out->strides[nd - 2], out->strides[nd - 1]);
    pa[t] = &views[3 * t];
    pb[t] = &views[3 * t + 1];
    po[t] = &views[3 * t + 2];
  }
  if (count == 1) return numc_matmul_one(pa[0], pb[0], po[0], nullptr, true);
  return numc_matmul_batched(out->ctx, count, pa.data(), pb.data(), po.data(), nullptr);
}
