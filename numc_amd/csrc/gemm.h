// gemm.h -- host interface of the GEMM kernels (gemm.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "numc/numc.h"

struct NumcCtx;

// one GEMM on device-resident operands; strides in elements
struct GemmProblem {
  int family;            // NumcFamily
  NumcDType in_dt, out_dt;
  int M, N, K;
  const void *A;
  long long rsa, csa;
  const void *B;
  long long rsb, csb;
  void *C;
  long long rso, cso;
  const void *bias;      // optional: out[i][j] += bias[j * bias_s] (output dtype), fused into the store
  long long bias_s;
  // strided batch (0 or 1: one problem): slice z reads A + z*a_bs, B + z*b_bs
  // and writes C + z*c_bs (elements; 0 = the same matrix for every slice)
  int batch;
  long long a_bs, b_bs, c_bs;
};

struct GemmArgs;
typedef void (*GemmLaunchFn)(const GemmArgs &, dim3, hipStream_t);

struct GemmConfig {
  NumcKernelConfig cfg;
  int T; // rows per MFMA fragment
  GemmLaunchFn launch;
  GemmLaunchFn launch_bias; // same kernel with a per-column bias in the store (float families), or null
  // same kernel reading operands in place (GemmArgs::Araw / Braw), or null:
  // [A, B, A and B][plain store, bias store]
  GemmLaunchFn inplace[3][2];
  // direct configurations (cfg.direct): launch / launch_bias run the one-launch
  // kernel on the raw operands (GemmArgs::Araw, a_rs, a_ks, Braw, b_rs, b_ks)
};

struct PackArgs {
  const void *src;
  long long rs, ks; // element (r, k) at src[r*rs + k*ks]
  int R, K;
  int nRT, nKF;     // padded row-fragment and k-fragment counts
  unsigned char *dst;
  long long plane_bytes;
  unsigned char *sc;  // FP8: block scale bytes, [row tile][k-fragment pair][64]
  int *flags;         // integer digit planes: nonzero mask
  int mode;           // batched pack: 0 k-contiguous tiles, 1 row-contiguous tiles, 2 lane slots,
                      // 3 int8 byte tiles (k-contiguous), 4 int8 byte tiles (row-contiguous)
  long long src_zs, dst_zs; // strided batch: bytes between slices of the source and of the packed planes
  int tile_rt;        // int8 / 16-bit k-contiguous byte tiles: row tiles per block (1, 2 or 4)
};

// one problem of the persistent batched kernel (packed operands in a workspace)
struct BatchProblem {
  const unsigned char *A, *B;
  long long a_plane, b_plane;
  int nKF, nKC, M, N;
  void *C;
  long long rso, cso;
  int out_bytes;
};

// one problem of the persistent direct batched kernel (operands in place;
// B given as the rows of B^T: row j at B + j*b_rs, k stride b_ks; bytes)
struct BatchDirect {
  const unsigned char *A, *B;
  long long a_rs, a_ks, b_rs, b_ks;
  int M, N, K;
  void *C;
  long long rso, cso;
  int out_bytes;
};
int numc_batch_direct_occupancy(int family);
// The direct batched kernel's work queues, one per XCD: `counter` points at
// NUMC_BATCH_COUNTER_BYTES of device memory holding queue q's counter at int
// q * NUMC_BATCH_COUNTER_STRIDE (cleared by the launch) and the queues' item
// starts (NUMC_BATCH_QUEUES + 1 ints, set by the host) at int NUMC_BATCH_QSTART.
#define NUMC_BATCH_QUEUES 8
#define NUMC_BATCH_COUNTER_STRIDE 32
#define NUMC_BATCH_QSTART (NUMC_BATCH_QUEUES * NUMC_BATCH_COUNTER_STRIDE)
#define NUMC_BATCH_COUNTER_BYTES ((NUMC_BATCH_QSTART + NUMC_BATCH_QUEUES + 1) * 4 + 60)
// direct batched kernel (FP32, split-BF16 and int8 families; gemm_small.hip): tile classes of 64 or
// 32 rows by 64 or 32 columns
int numc_batch_direct_launch(int family, const BatchDirect *probs, const int4 *items, int nitems, int *counter,
                             int grid, hipStream_t st, NumcCtx *prof);

int numc_batch_class_shape(int family, int cls, int *bm, int *bn, int *kc);
int numc_batch_occupancy(int family);
struct BatchChoice {
  int cls;  // 0 = 128x128 tiles, 1 = 64x64
  float ms; // predicted time of the whole problem
};
BatchChoice numc_batch_choice(int family, int M, int N, int K, int thresh);
int numc_batch_threshold(int family);
void numc_batch_set_threshold(int family, int t);
// whether the family's batches run the direct kernel (no pack) or the packed one
bool numc_batch_direct(int family);
void numc_batch_set_direct(int family, bool direct);
int numc_batch_pack_blocks(int family, PackArgs &pk); // sets pk.mode
int numc_batch_launch(int family, const PackArgs *descs, const long long *prefix, const int *owner,
                      long long units, const BatchProblem *probs, const int4 *items, int nitems, int *counter,
                      int grid, hipStream_t st, NumcCtx *prof); // prof: time the kernels (eager launches only)

const GemmConfig *numc_gemm_config(int family, int idx);
// the direct (in-place, one-launch) configurations of a family (gemm_small.hip);
// numc_gemm_config lists them after the packed ones
const GemmConfig *numc_direct_configs(int family, int *count);
// whether a direct configuration can run problem p (one digit plane for
// integers, no strided batch; FP8 has none)
bool numc_direct_ok(const GemmProblem &p);
// configuration for a strided batch of equal M x N slices
const GemmConfig *numc_strided_batch_config(int family, int M, int N);
int numc_gemm_config_count(int family);
// device bytes needed in ctx->ws (and ctx->ws_aux) for a problem under a config
size_t numc_gemm_workspace(const GemmProblem &p, const GemmConfig &c, size_t *aux_bytes);
int numc_gemm_launch(NumcCtx *ctx, const GemmProblem &p, const GemmConfig &c, void *ws, void *aux);
int numc_naive_launch(NumcCtx *ctx, NumcDType dt, const void *a, const void *b, void *c, int M, int K, int N);

// out[i] = src[i] * scalar over an N-D strided view (element strides), in the
// dtype's own arithmetic: integers keep the low w bits of the product
#define NUMC_EW_MAX_DIMS 16
struct ScaleArgs {
  const unsigned char *src;
  unsigned char *dst;
  long long n;
  int ndim, dtype;
  long long shape[NUMC_EW_MAX_DIMS], ss[NUMC_EW_MAX_DIMS], ds[NUMC_EW_MAX_DIMS];
  uint64_t scalar_bits; // the scalar in the dtype's representation
};
int numc_scale_launch(NumcCtx *ctx, const ScaleArgs &a);
int numc_i64_launch(NumcCtx *ctx, const GemmProblem &p);
// bytes of device scratch behind numc_amd_gemm_noalloc (per GPU, for the
// process's lifetime)
#define NUMC_NOALLOC_SCRATCH (64 << 20)
// the kernel family of a dtype and method (matmul.cpp): a NumcFamily, 100 for
// 64-bit integers, or a negative error code
int numc_matmul_family_dt(int dtype, const NumcMatmulOpts *opts);
// int8 digit planes of an integer problem (1 for the other families)
int numc_problem_planes(const GemmProblem &p);
// time the candidate configurations on the problem's own operands (autotune.hip)
int numc_tune_call(NumcCtx *ctx, const GemmProblem &p, const int *cand, int n, int *best);
// heuristic / learned choice of config index for a family and shape (direct
// configurations only when allowed)
int numc_select_config(int family, int M, int N, int K, int flags, bool allow_direct);
