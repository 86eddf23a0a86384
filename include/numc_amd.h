/* numc_amd.h -- MI355X additions to the numc C ABI.
 *
 * Nothing here replaces a numc entry point; each declaration is additive.
 * Where an addition has a counterpart in numc the comment names it. */
#ifndef NUMC_AMD_H
#define NUMC_AMD_H

#include "numc/array.h"
#include "numc/export.h"
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- float32 GEMM methods ------------------------------------------------ */
typedef enum {
  NUMC_METHOD_AUTO = 0,       /* = PLAIN for float32; the only choice for other dtypes */
  NUMC_METHOD_PLAIN = 1,      /* FP32 MFMA, FP32 accumulate */
  NUMC_METHOD_SPLIT_BF16 = 2, /* a = hi + lo (bf16 each), hi*hi + hi*lo + lo*hi, FP32 accumulate */
  NUMC_METHOD_FP8_SCALED = 3, /* OCP e4m3 with a power-of-two scale per 32 consecutive k of each row of A and
                                 column of B (the MFMA's E8M0 block scales), FP32 accumulate */
} NumcMethod;

typedef struct {
  int method;   /* NumcMethod */
  int config;   /* kernel configuration index for the family, -1 = let the selector choose */
  int flags;    /* NUMC_OPT_* */
} NumcMatmulOpts;

#define NUMC_OPT_WIDEN 1          /* allow int8/uint8 inputs with an int32/uint32 output */
#define NUMC_OPT_DEVICE_POINTERS 2 /* numc_amd_gemm: a, b, c are device addresses */

/* numc_matmul with a method and configuration (no counterpart in numc; the
 * default options give exactly numc_matmul).  With NUMC_OPT_WIDEN, a and b
 * may be INT8 (or UINT8) and out INT32 (or UINT32): the exact sum mod 2^32. */
NUMC_API int numc_matmul_ex(const NumcArray *a, const NumcArray *b, NumcArray *out,
                            const NumcMatmulOpts *opts);

/* Many independent problems out[i] = a[i] @ b[i] (no counterpart in numc;
 * numc_bmm and numc_dot's batched case run on it).  All problems share
 * one dtype (and method); shapes may differ.  Runs as one persistent kernel
 * launch, replayed from a cached HIP graph when the same arrays repeat. */
NUMC_API int numc_matmul_batched(NumcCtx *ctx, size_t count, NumcArray *const *a,
                                 NumcArray *const *b, NumcArray *const *out,
                                 const NumcMatmulOpts *opts);

/* Fused linear layer out = x @ w + bias, the numc_linear(x, w, bias, out) of
 * numc's roadmap (ROADMAP.md:139; numc has no implementation).  x is
 * (..., K), w (K, N), bias (N) or NULL, out (..., N), one dtype throughout.
 * The bias is added in the GEMM's store, on the accumulator before it is
 * written: floats round once more (fl(acc + bias)), integers keep the exact
 * sum mod 2^w.  Leading axes of x that are linear in x and out fold into the
 * GEMM's M; others run as separate GEMMs.  Errors as numc_matmul. */
NUMC_API int numc_linear(const NumcArray *x, const NumcArray *w, const NumcArray *bias, NumcArray *out);

/* numc_linear with a kernel configuration (opts->config, -1 = selector) and
 * method AUTO or PLAIN (others: NUMC_ERR_VALUE; no counterpart in numc). */
NUMC_API int numc_linear_ex(const NumcArray *x, const NumcArray *w, const NumcArray *bias, NumcArray *out,
                            const NumcMatmulOpts *opts);

/* Batched matrix multiply out[...] = a[...] @ b[...], the numc_bmm(a, b, out)
 * of numc's roadmap (ROADMAP.md:140; numc has no implementation).  a is
 * (..., M, K), b (..., K, N), both at least 2-D; the batch axes broadcast as
 * numpy.matmul's do (aligned from the right, equal or 1) and out is
 * (batch..., M, N).  When each operand's slices sit at offsets linear in the
 * batch index (or the operand is 2-D), all slices run as one strided-batch
 * GEMM launch; other layouts run in one persistent numc_matmul_batched launch
 * (a single slice runs as numc_matmul). */
NUMC_API int numc_bmm(const NumcArray *a, const NumcArray *b, NumcArray *out);

/* Pointer-level GEMM for a host library that keeps its own arrays: the call
 * numc's numc_matmul (src/matmul/matmul.c:347-389) makes to its dtype kernel
 * table (matmul.c:210-320, kernels taking data pointers and strides) becomes
 * this one.  c (M x N) = a (M x K) @ b (K x N); strides are in bytes, as in
 * numc's NumcArray (src/internal.h:35-43), and every one is honoured.  a, b
 * and c are host pointers (staged through the context's device workspace,
 * gaps of a strided c preserved), or device addresses with
 * NUMC_OPT_DEVICE_POINTERS.  Result contract, methods and errors as
 * numc_matmul_ex (dtype == out_dtype unless NUMC_OPT_WIDEN).  The context
 * comes from numc_ctx_create; no arrays of it are needed. */
NUMC_API int numc_amd_gemm(NumcCtx *ctx, int dtype, int out_dtype, size_t M, size_t N, size_t K,
                           const void *a, size_t rsa, size_t csa, const void *b, size_t rsb, size_t csb,
                           void *c, size_t rso, size_t cso, const NumcMatmulOpts *opts);

/* numc_amd_gemm for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * no counterpart in numc.  a, b and c are device addresses (as with
 * NUMC_OPT_DEVICE_POINTERS), the kernels run on `stream` (a hipStream_t; NULL
 * = the default stream) and return without synchronizing, and nothing is
 * allocated: the workspace of the packed kernels comes from a 64 MiB device
 * array the library holds on each GPU for the process's lifetime.  Problems
 * whose every configuration needs more (packed operands of about 16M
 * elements and up) fail with NUMC_ERR_VALUE; the selector's choice is used
 * without call-time timing.  Calls that overlap on one GPU (other streams or
 * threads) share that array and race.  64-bit integers are refused.  Result
 * contract, opts and errors otherwise as numc_amd_gemm. */
NUMC_API int numc_amd_gemm_noalloc(void *stream, int dtype, int out_dtype, size_t M, size_t N, size_t K,
                                   const void *a, size_t rsa, size_t csa, const void *b, size_t rsb, size_t csb,
                                   void *c, size_t rso, size_t cso, const NumcMatmulOpts *opts);

/* Autotune the batched schedule for this batch: times the whole batch under
 * each candidate schedule and keeps the fastest for the batch's dtype family
 * in later numc_matmul_batched calls.  A schedule is the kernel -- direct
 * (float32 plain or split-BF16, and int8 inputs: operands read where they
 * lie, no pack pass; each wave takes one 64x64, 64x32, 32x64 or 32x32 tile
 * over the whole K) or packed (one pack launch, then 128x128 / 64x64 tiles)
 * -- and the threshold above which a problem's M and N give it the larger
 * tile.  Returns the threshold. */
NUMC_API int numc_matmul_batched_tune(NumcCtx *ctx, size_t count, NumcArray *const *a,
                                      NumcArray *const *b, NumcArray *const *out,
                                      const NumcMatmulOpts *opts, int reps);
/* the family's current batched schedule (direct 1 / packed 0, threshold) */
NUMC_API int numc_matmul_batched_schedule(int family, int *direct, int *threshold);
NUMC_API int numc_matmul_batched_set_schedule(int family, int direct, int threshold);

/* number of times the context replayed one of its cached batched graphs */
NUMC_API long long numc_ctx_batch_graph_replays(NumcCtx *ctx);
/* graphs cached on the context: the last 4 distinct batches (by arrays,
 * shapes, strides and schedule), least recently used replaced first */
NUMC_API int numc_ctx_batch_graphs_cached(NumcCtx *ctx);

/* ---- residency -------------------------------------------------------------
 * By default the host copy is the current one: every GPU call uploads its
 * inputs and downloads its output, so pointers from numc_array_data behave
 * as in numc.  numc_array_to_device makes the device copy current for the
 * array's whole storage (views included); GEMMs then read and write HBM
 * directly until numc_array_to_host or numc_array_data brings it back. */
NUMC_API int numc_array_to_device(NumcArray *arr);
NUMC_API int numc_array_to_host(NumcArray *arr);
NUMC_API int numc_array_on_device(const NumcArray *arr);
/* device address of element 0 of the view (allocates the block's twin) */
NUMC_API void *numc_array_device_data(const NumcArray *arr);

/* ---- execution control ----------------------------------------------------- */
/* synchronous != 0 (default): calls return after the GPU finished.
 * 0: calls return after enqueueing on the context's stream; results are
 * ready after numc_ctx_synchronize (numc_array_data also waits). */
NUMC_API int numc_ctx_set_sync(NumcCtx *ctx, int synchronous);
NUMC_API int numc_ctx_synchronize(NumcCtx *ctx);
NUMC_API void *numc_ctx_stream(NumcCtx *ctx); /* hipStream_t of the context */
/* HIP events on the context stream, slots 0..15 */
NUMC_API int numc_ctx_timer_mark(NumcCtx *ctx, int slot);
NUMC_API int numc_ctx_timer_elapsed(NumcCtx *ctx, int slot0, int slot1, float *ms);

/* Kernel timing: with profiling on, HIP events are recorded on the context
 * stream around every pack (category 0), GEMM (1), split-K reduce (2) and
 * batched GEMM (3) launch; read returns the summed milliseconds and launch
 * count of one category and family (family NUMC_FAMILY_COUNT = other
 * kernels).  While profiling, numc_matmul_batched launches its cached batch's
 * kernels one by one instead of replaying the graph, so each is timed. */
NUMC_API int numc_ctx_profile(NumcCtx *ctx, int enable);
NUMC_API int numc_ctx_profile_read(NumcCtx *ctx, int category, int family, double *total_ms,
                                   long long *launches, int reset);

/* plain device-to-device copy by a simple vector kernel on the context stream
 * (the known-traffic reference used to calibrate HBM counters) */
NUMC_API int numc_device_copy(NumcCtx *ctx, void *dst, const void *src, size_t bytes);

/* ---- seeded inputs ----------------------------------------------------------
 * Fills a contiguous array with the documented generator: element i comes from
 * xoshiro256** block i>>16 seeded by splitmix64 from seed, stream and block.
 * dist: 0 uniform [lo,hi), 1 normal(0,1) (Box-Muller), 2 integer uniform
 * [lo,hi] by rejection.  16-bit floats round the binary32 value.  An array
 * whose device copy is current (numc_array_to_device) fills on the GPU and
 * stays device-current, with the same values (normal: within one unit in the
 * last place, a few for float64, where the device's log/sin/cos differ); any
 * other array fills on the host and leaves the host copy current. */
NUMC_API int numc_array_fill_random(NumcArray *arr, int dist, uint64_t seed,
                                    uint64_t stream, double lo, double hi);

/* ---- kernel configurations ------------------------------------------------ */
typedef enum {
  NUMC_FAMILY_F32 = 0,
  NUMC_FAMILY_SPLIT_BF16 = 1,
  NUMC_FAMILY_FP8 = 2,
  NUMC_FAMILY_F64 = 3,
  NUMC_FAMILY_I8 = 4, /* every integer dtype up to 32 bits (int8 digit planes) */
  NUMC_FAMILY_BF16 = 5, /* bfloat16 arrays: one bf16 MFMA product, FP32 accumulate */
  NUMC_FAMILY_F16 = 6,  /* float16 arrays: one fp16 MFMA product, FP32 accumulate */
  NUMC_FAMILY_COUNT = 7
} NumcFamily;

typedef struct {
  int bm, bn, kc;     /* block tile rows, cols, k-chunk (direct: k per MFMA fragment) */
  int waves_m, waves_n;
  int split_k;        /* number of K slices summed by a second launch (1 = none) */
  int waves_k;        /* waves of one block splitting K, summed in LDS (1 = none) */
  int direct;         /* 1: one launch reads A and B in place (any strides and
                         alignment, no pack and no workspace), numc's gemmsup
                         counterpart; 0: operands packed into fragment order */
} NumcKernelConfig;

NUMC_API int numc_kernel_config_count(int family);
NUMC_API int numc_kernel_config_get(int family, int index, NumcKernelConfig *out);
/* ---- autotuner and learned selector -----------------------------------------
 * numc_autotune_shape times every configuration of `family` on device scratch
 * operands of shape M x K @ K x N (flags: 1 A stored transposed, 2 B stored
 * transposed, 4 padded rows), whole calls, mean of `reps`, into ms_out.
 * The selector predicts log2(ms) per configuration with an MLP cost model
 * (every family); numc_matmul runs the argmin, except that where the model
 * cannot separate its best few configurations it times them once on the call
 * and keeps the winner for that exact shape (numc_selector_candidates;
 * environment NUMC_AUTOTUNE=off disables this).  Without a trained model (or
 * one trained for other configuration tables) an analytic model is used:
 * numc's gemmsup rule, direct configurations below measured K and M*N
 * thresholds, else a packed one by estimated rounds of work. */
NUMC_API int numc_autotune_shape(NumcCtx *ctx, int family, int M, int N, int K, int flags, int reps,
                                 float *ms_out, int ms_len);
NUMC_API int numc_autotune_release(NumcCtx *ctx);
NUMC_API int numc_selector_feature_count(void);
NUMC_API int numc_selector_features(int family, int M, int N, int K, int flags, int cfg, float *out);
NUMC_API int numc_selector_load(const char *path);
NUMC_API int numc_selector_unload(void);
NUMC_API int numc_selector_active(void);
NUMC_API int numc_selector_predict(int family, int M, int N, int K, int flags, float *out, int len);
/* exhaustive autotune of one exact shape on this device: its best
 * configuration is used for that shape from then on (returns the index) */
NUMC_API int numc_selector_tune(NumcCtx *ctx, int family, int M, int N, int K, int flags, int reps);
/* pin (cfg >= 0) or unpin (cfg < 0) the configuration of an exact shape */
NUMC_API int numc_selector_pin(int family, int M, int N, int K, int flags, int cfg);
/* The configurations a numc_matmul call would consider for this shape: 1 when
 * the choice is settled (an exact-shape entry, or a confident model), up to 3
 * whose predictions lie within 2^0.1 of the best, or all of them when the best
 * predicted call takes under 2^-5 ms; in predicted order into out (at most
 * max).  Returns the count, 0 without a trained model.  allow_direct = 0
 * leaves out the direct configurations (integers of more than one digit
 * plane). */
NUMC_API int numc_selector_candidates(int family, int M, int N, int K, int flags, int allow_direct, int *out,
                                      int max);
/* the analytic model used without a trained selector */
NUMC_API int numc_selector_heuristic(int family, int M, int N, int K);
/* the configuration numc_matmul would run for this shape: exact-shape table,
 * then the learned selector, then the analytic model */
NUMC_API int numc_selector_choice(int family, int M, int N, int K, int flags);

/* The cost model's network: 3 layers (din -> hidden -> hidden -> dout, ReLU),
 * trained and evaluated by hand-written HIP kernels on the context's stream. */
typedef struct NumcMlp NumcMlp;
NUMC_API NumcMlp *numc_mlp_create(NumcCtx *ctx, int din, int hidden, int dout, uint64_t seed);
NUMC_API void numc_mlp_free(NumcMlp *mlp);
NUMC_API size_t numc_mlp_param_count(const NumcMlp *mlp);
NUMC_API int numc_mlp_set_params(NumcMlp *mlp, const float *params, size_t count);
NUMC_API int numc_mlp_get_params(NumcMlp *mlp, float *params, size_t count);
NUMC_API int numc_mlp_forward(NumcMlp *mlp, const float *x, long long rows, float *y);
NUMC_API int numc_mlp_loss_grad(NumcMlp *mlp, const float *x, const float *t, const float *mask, int batch,
                                double *loss, float *grads);
NUMC_API int numc_mlp_train(NumcMlp *mlp, const float *x, const float *t, const float *mask, long long rows,
                            int epochs, int batch, float lr, float lr_final, uint64_t seed, double *loss_out);

/* family a call with these arguments would run (or a negative error code) */
NUMC_API int numc_matmul_family(const NumcArray *a, const NumcArray *b,
                                const NumcArray *out, const NumcMatmulOpts *opts);

#ifdef __cplusplus
}
#endif
#endif
