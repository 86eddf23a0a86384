/*
 * ref_bench.c -- times numc's own numc_matmul (built into _ref/libnumc_ref.so
 * from the numc sources) on inputs drawn with the oracle's documented
 * generator, so the bytes equal those the GPU path is given for the same
 * (seed, stream).  Method follows numc's bench/numc/bench_matmul_csv.c:93-109:
 * warmup, then the minimum per-iteration wall time; GFLOP/s = 2MNK / t.
 *
 * TEST INFRASTRUCTURE ONLY (the cpu_baseline leg of bench.py runs it).
 *
 * usage: numc_ref_bench DTYPE M K N DIST SEED WARMUP ITERS [LO HI]
 *   DTYPE: numc dtype code (8 = float32, 9 = float64, 2 = int32, 0 = int8 ...)
 *   DIST : u = uniform[LO,HI), n = normal(0,1), i = integer uniform [LO,HI]
 *        numc_ref_bench batch DTYPE SEED WARMUP ITERS LO HI  < shapes
 *   one numc_matmul per problem of a batch (numc has no batched GEMM); the
 *   shapes come as lines "M N K" on stdin, problem i's A and B are filled
 *   with streams 2i and 2i+1 (uniform [LO,HI) for floats, integers in
 *   [LO,HI] otherwise); the time of a pass over the whole batch is minimised.
 * prints one JSON object.
 */
#include <numc/array.h>
#include <numc/error.h>
#include <numc/math.h>

#include "numc_oracle.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* position-weighted sum of the 32-bit words, mod 2^64 (cheap to recompute
 * with numpy on the other side); n must be a multiple of 4 */
static uint64_t wsum(const void *p, size_t n) {
  const uint32_t *w = (const uint32_t *)p;
  uint64_t h = 0;
  for (size_t i = 0; i < n / 4; i++) h += (uint64_t)w[i] * (uint64_t)(i + 1);
  return h;
}

static void fill(void *dst, int dt, size_t n, char dist, uint64_t seed,
                 uint64_t stream, double lo, double hi) {
  if (dist == 'n') orc_fill_normal(dst, dt, n, seed, stream);
  else if (dist == 'i') orc_fill_int(dst, dt, n, seed, stream, (int64_t)lo, (int64_t)hi);
  else orc_fill_uniform(dst, dt, n, seed, stream, lo, hi);
}

static int run_batch(int argc, char **argv) {
  if (argc < 8) {
    fprintf(stderr, "usage: %s batch DTYPE SEED WARMUP ITERS LO HI < shapes\n", argv[0]);
    return 2;
  }
  int dt = atoi(argv[2]);
  uint64_t seed = strtoull(argv[3], NULL, 10);
  int warmup = atoi(argv[4]), iters = atoi(argv[5]);
  double lo = atof(argv[6]), hi = atof(argv[7]);
  size_t cap = 1024, n = 0;
  size_t (*dims)[3] = malloc(cap * sizeof *dims);
  size_t m, nn, k;
  while (scanf("%zu %zu %zu", &m, &nn, &k) == 3) {
    if (n == cap) dims = realloc(dims, (cap *= 2) * sizeof *dims);
    dims[n][0] = m; dims[n][1] = nn; dims[n][2] = k;
    n++;
  }
  NumcCtx *ctx = numc_ctx_create();
  NumcArray **a = malloc(n * sizeof *a), **b = malloc(n * sizeof *b), **c = malloc(n * sizeof *c);
  if (!ctx || !a || !b || !c) return 3;
  size_t es = orc_dtype_size(dt);
  uint64_t in_sum = 0;
  double flops = 0;
  for (size_t i = 0; i < n; i++) {
    size_t M = dims[i][0], N = dims[i][1], K = dims[i][2];
    size_t sa[2] = {M, K}, sb[2] = {K, N}, so[2] = {M, N};
    a[i] = numc_array_create(ctx, sa, 2, (NumcDType)dt);
    b[i] = numc_array_create(ctx, sb, 2, (NumcDType)dt);
    c[i] = numc_array_zeros(ctx, so, 2, (NumcDType)dt);
    if (!a[i] || !b[i] || !c[i]) return 4;
    char dist = (dt == 8 || dt == 9) ? 'u' : 'i';
    fill(numc_array_data(a[i]), dt, M * K, dist, seed, 2 * i, lo, hi);
    fill(numc_array_data(b[i]), dt, K * N, dist, seed, 2 * i + 1, lo, hi);
    /* byte sums of the inputs (checked against the GPU side's bytes) */
    const unsigned char *pa = numc_array_data(a[i]), *pb = numc_array_data(b[i]);
    for (size_t x = 0; x < M * K * es; x++) in_sum += pa[x];
    for (size_t x = 0; x < K * N * es; x++) in_sum += pb[x];
    flops += 2.0 * (double)M * (double)N * (double)K;
  }
  for (int w = 0; w < warmup; w++)
    for (size_t i = 0; i < n; i++)
      if (numc_matmul(a[i], b[i], c[i])) return 5;
  double best = 1e300;
  for (int it = 0; it < iters; it++) {
    double t0 = now_s();
    for (size_t i = 0; i < n; i++)
      if (numc_matmul(a[i], b[i], c[i])) return 5;
    double t = now_s() - t0;
    if (t < best) best = t;
  }
  int threads = 1;
#ifdef _OPENMP
  threads = omp_get_max_threads();
#endif
  printf("{\"dtype\": %d, \"problems\": %zu, \"flop\": %.6e, \"ms_min\": %.6f, \"gflops\": %.4f, "
         "\"threads\": %d, \"warmup\": %d, \"iters\": %d, \"in_byte_sum\": \"%llu\"}\n",
         dt, n, flops, best * 1e3, flops / best * 1e-9, threads, warmup, iters, (unsigned long long)in_sum);
  numc_ctx_free(ctx);
  free(a); free(b); free(c); free(dims);
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && strcmp(argv[1], "batch") == 0) return run_batch(argc, argv);
  if (argc < 9) {
    fprintf(stderr, "usage: %s DTYPE M K N DIST SEED WARMUP ITERS [LO HI]\n", argv[0]);
    return 2;
  }
  int dt = atoi(argv[1]);
  size_t M = strtoull(argv[2], NULL, 10), K = strtoull(argv[3], NULL, 10),
         N = strtoull(argv[4], NULL, 10);
  char dist = argv[5][0];
  uint64_t seed = strtoull(argv[6], NULL, 10);
  int warmup = atoi(argv[7]), iters = atoi(argv[8]);
  double lo = argc > 9 ? atof(argv[9]) : -1.0, hi = argc > 10 ? atof(argv[10]) : 1.0;

  NumcCtx *ctx = numc_ctx_create();
  if (!ctx) return 3;
  size_t sa[2] = {M, K}, sb[2] = {K, N}, so[2] = {M, N};
  NumcArray *a = numc_array_create(ctx, sa, 2, (NumcDType)dt);
  NumcArray *b = numc_array_create(ctx, sb, 2, (NumcDType)dt);
  NumcArray *c = numc_array_zeros(ctx, so, 2, (NumcDType)dt);
  if (!a || !b || !c) return 4;
  fill(numc_array_data(a), dt, M * K, dist, seed, 0, lo, hi);
  fill(numc_array_data(b), dt, K * N, dist, seed, 1, lo, hi);

  for (int w = 0; w < warmup; w++)
    if (numc_matmul(a, b, c)) return 5;
  double best = 1e300;
  for (int it = 0; it < iters; it++) {
    double t0 = now_s();
    if (numc_matmul(a, b, c)) return 5;
    double t = now_s() - t0;
    if (t < best) best = t;
  }
  int threads = 1;
#ifdef _OPENMP
  threads = omp_get_max_threads();
#endif
  double flops = 2.0 * (double)M * (double)N * (double)K;
  size_t es = orc_dtype_size(dt);
  printf("{\"dtype\": %d, \"M\": %zu, \"K\": %zu, \"N\": %zu, \"ms_min\": %.6f, "
         "\"gflops\": %.4f, \"threads\": %d, \"warmup\": %d, \"iters\": %d, "
         "\"a_wsum\": \"%llu\", \"b_wsum\": \"%llu\", \"c_wsum\": \"%llu\"}\n",
         dt, M, K, N, best * 1e3, flops / best * 1e-9, threads, warmup, iters,
         (unsigned long long)wsum(numc_array_data(a), M * K * es),
         (unsigned long long)wsum(numc_array_data(b), K * N * es),
         (unsigned long long)wsum(numc_array_data(c), M * N * es));
  numc_ctx_free(ctx);
  return 0;
}
