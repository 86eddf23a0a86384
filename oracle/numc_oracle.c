/*
 * numc_oracle.c -- CPU restatement of numc's matmul semantics (the checker).
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product; loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Function-by-function provenance is in numc_oracle.h.
 */
#include "numc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

size_t orc_dtype_size(int dt) {
  static const size_t sz[] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8};
  return (dt >= 0 && dt <= ORC_F64) ? sz[dt] : 0;
}

/* ------------------------------------------------------------------------ */
/* generator                                                                 */
/* ------------------------------------------------------------------------ */

uint64_t orc_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_rng_block(orc_rng *r, uint64_t seed, uint64_t stream, uint64_t block) {
  uint64_t st = seed ^ (stream * 0x9E3779B97F4A7C15ULL) ^ (block * 0xD1B54A32D192ED03ULL);
  for (int i = 0; i < 4; i++) r->s[i] = orc_splitmix64(&st);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t orc_rng_next(orc_rng *r) {
  uint64_t *s = r->s;
  uint64_t result = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

#define ORC_BLOCK_SHIFT 16
#define ORC_BLOCK ((size_t)1 << ORC_BLOCK_SHIFT)

static void store_real(void *dst, int dt, size_t i, double v) {
  if (dt == ORC_F32) ((float *)dst)[i] = (float)v;
  else ((double *)dst)[i] = v;
}

void orc_fill_uniform(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream,
                      double lo, double hi) {
  size_t nb = (n + ORC_BLOCK - 1) >> ORC_BLOCK_SHIFT;
#pragma omp parallel for schedule(static)
  for (size_t b = 0; b < nb; b++) {
    orc_rng r;
    orc_rng_block(&r, seed, stream, b);
    size_t end = (b + 1) * ORC_BLOCK < n ? (b + 1) * ORC_BLOCK : n;
    for (size_t i = b * ORC_BLOCK; i < end; i++) {
      uint64_t x = orc_rng_next(&r);
      if (dt == ORC_F32) {
        float u = (float)(x >> 40) * 0x1p-24f;
        ((float *)dst)[i] = (float)lo + ((float)hi - (float)lo) * u;
      } else {
        double u = (double)(x >> 11) * 0x1p-53;
        ((double *)dst)[i] = lo + (hi - lo) * u;
      }
    }
  }
}

void orc_fill_normal(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream) {
  size_t nb = (n + ORC_BLOCK - 1) >> ORC_BLOCK_SHIFT;
#pragma omp parallel for schedule(static)
  for (size_t b = 0; b < nb; b++) {
    orc_rng r;
    orc_rng_block(&r, seed, stream, b);
    size_t end = (b + 1) * ORC_BLOCK < n ? (b + 1) * ORC_BLOCK : n;
    for (size_t i = b * ORC_BLOCK; i < end; i += 2) {
      uint64_t x = orc_rng_next(&r), y = orc_rng_next(&r);
      double u1 = (double)((x >> 11) + 1) * 0x1p-53;
      double u2 = (double)(y >> 11) * 0x1p-53;
      double rad = sqrt(-2.0 * log(u1));
      double th = 6.283185307179586476925286766559 * u2;
      store_real(dst, dt, i, rad * cos(th));
      if (i + 1 < end) store_real(dst, dt, i + 1, rad * sin(th));
    }
  }
}

static void store_int(void *dst, int dt, size_t i, int64_t v) {
  switch (dt) {
  case ORC_I8: ((int8_t *)dst)[i] = (int8_t)v; break;
  case ORC_I16: ((int16_t *)dst)[i] = (int16_t)v; break;
  case ORC_I32: ((int32_t *)dst)[i] = (int32_t)v; break;
  case ORC_I64: ((int64_t *)dst)[i] = v; break;
  case ORC_U8: ((uint8_t *)dst)[i] = (uint8_t)v; break;
  case ORC_U16: ((uint16_t *)dst)[i] = (uint16_t)v; break;
  case ORC_U32: ((uint32_t *)dst)[i] = (uint32_t)v; break;
  case ORC_U64: ((uint64_t *)dst)[i] = (uint64_t)v; break;
  case ORC_F32: ((float *)dst)[i] = (float)v; break;
  case ORC_F64: ((double *)dst)[i] = (double)v; break;
  }
}

void orc_fill_int(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream,
                  int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)(hi - lo) + 1;   /* 1 .. 2^32 */
  uint64_t limit = ((1ULL << 32) / span) * span;
  size_t nb = (n + ORC_BLOCK - 1) >> ORC_BLOCK_SHIFT;
#pragma omp parallel for schedule(static)
  for (size_t b = 0; b < nb; b++) {
    orc_rng r;
    orc_rng_block(&r, seed, stream, b);
    size_t end = (b + 1) * ORC_BLOCK < n ? (b + 1) * ORC_BLOCK : n;
    for (size_t i = b * ORC_BLOCK; i < end; i++) {
      uint64_t x32;
      do {
        x32 = orc_rng_next(&r) >> 32;
      } while (x32 >= limit);
      store_int(dst, dt, i, lo + (int64_t)(x32 % span));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* validation: dispatch.h:20-59                                              */
/* ------------------------------------------------------------------------ */

int orc_check_matmul(int have_a, int have_b, int have_out, int dta, int dtb,
                     int dto, int nda, int ndb, int ndo, const size_t *sa,
                     const size_t *sb, const size_t *so) {
  if (!have_a || !have_b || !have_out) return -1;
  if (dta != dtb || dta != dto) return -4;
  if (nda != 2 || ndb != 2 || ndo != 2) return -3;
  if (sa[1] != sb[0]) return -3;
  if (so[0] != sa[0] || so[1] != sb[1]) return -3;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* exact integer matmul                                                      */
/* ------------------------------------------------------------------------ */

static inline uint64_t load_u64(const void *p, int dt, ptrdiff_t i) {
  switch (dt) {
  case ORC_I8: return (uint64_t)(int64_t)((const int8_t *)p)[i];
  case ORC_I16: return (uint64_t)(int64_t)((const int16_t *)p)[i];
  case ORC_I32: return (uint64_t)(int64_t)((const int32_t *)p)[i];
  case ORC_I64: return (uint64_t)((const int64_t *)p)[i];
  case ORC_U8: return ((const uint8_t *)p)[i];
  case ORC_U16: return ((const uint16_t *)p)[i];
  case ORC_U32: return ((const uint32_t *)p)[i];
  case ORC_U64: return ((const uint64_t *)p)[i];
  }
  return 0;
}

static inline void store_u64(void *p, int dt, ptrdiff_t i, uint64_t v) {
  switch (orc_dtype_size(dt)) {
  case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
  case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
  case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
  case 8: ((uint64_t *)p)[i] = v; break;
  }
}

/* Sum of products in Z/2^64 (two's complement wraparound of uint64), which
 * equals the exact integer sum reduced mod 2^64; storing the low w bits then
 * gives the exact sum mod 2^w -- the value numc_matmul_naive produces for
 * every integer dtype (kernel.h:18-62: 8/16-bit types accumulate in a wider
 * type and cast back each step, 32/64-bit types wrap natively). */
void orc_matmul_int(int dt_in, int dt_out, const void *a, const void *b, void *out,
                    size_t M, size_t K, size_t N,
                    ptrdiff_t rsa, ptrdiff_t csa, ptrdiff_t rsb, ptrdiff_t csb,
                    ptrdiff_t rso, ptrdiff_t cso) {
#pragma omp parallel
  {
    uint64_t *acc = (uint64_t *)malloc((N ? N : 1) * sizeof(uint64_t));
    uint64_t *brow = (uint64_t *)malloc((N ? N : 1) * sizeof(uint64_t));
#pragma omp for schedule(dynamic, 4)
    for (size_t i = 0; i < M; i++) {
      memset(acc, 0, N * sizeof(uint64_t));
      for (size_t k = 0; k < K; k++) {
        uint64_t aik = load_u64(a, dt_in, (ptrdiff_t)i * rsa + (ptrdiff_t)k * csa);
        if (!aik) continue;
        const ptrdiff_t bk = (ptrdiff_t)k * rsb;
        for (size_t j = 0; j < N; j++)
          acc[j] += aik * load_u64(b, dt_in, bk + (ptrdiff_t)j * csb);
      }
      for (size_t j = 0; j < N; j++)
        store_u64(out, dt_out, (ptrdiff_t)i * rso + (ptrdiff_t)j * cso, acc[j]);
    }
    free(acc);
    free(brow);
  }
}

/* ------------------------------------------------------------------------ */
/* floating point references                                                 */
/* ------------------------------------------------------------------------ */

static inline double load_real(const void *p, int dt, ptrdiff_t i) {
  return dt == ORC_F32 ? (double)((const float *)p)[i] : ((const double *)p)[i];
}

/* one entry of the reference */
static void fref_entry(int dt, const void *a, const void *b, size_t K, ptrdiff_t ao,
                       ptrdiff_t csa, ptrdiff_t bo, ptrdiff_t rsb, double *r,
                       double *ab) {
  double s = 0.0, c = 0.0, sa = 0.0;
  if (dt == ORC_F32) {
    /* fp32*fp32 is exact in binary64; only the accumulation rounds */
    for (size_t k = 0; k < K; k++) {
      double x = load_real(a, dt, ao + (ptrdiff_t)k * csa);
      double y = load_real(b, dt, bo + (ptrdiff_t)k * rsb);
      s += x * y;
      sa += fabs(x) * fabs(y);
    }
  } else {
    /* double-double: TwoProd via fma, TwoSum accumulation */
    for (size_t k = 0; k < K; k++) {
      double x = load_real(a, dt, ao + (ptrdiff_t)k * csa);
      double y = load_real(b, dt, bo + (ptrdiff_t)k * rsb);
      double p = x * y;
      double pe = fma(x, y, -p);
      double t = s + p;
      double bb = t - s;
      double e = (s - (t - bb)) + (p - bb);
      s = t;
      c += e + pe;
      sa += fabs(x) * fabs(y);
    }
    s = s + c;
  }
  *r = s;
  if (ab) *ab = sa;
}

void orc_matmul_fref(int dt, const void *a, const void *b, size_t M, size_t K,
                     size_t N, ptrdiff_t rsa, ptrdiff_t csa, ptrdiff_t rsb,
                     ptrdiff_t csb, double *ref, double *absr, size_t nsel,
                     const int64_t *rows, const int64_t *cols) {
  if (nsel > 0) {
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t s = 0; s < nsel; s++) {
      ptrdiff_t i = (ptrdiff_t)rows[s], j = (ptrdiff_t)cols[s];
      fref_entry(dt, a, b, K, i * rsa, csa, j * csb, rsb, &ref[s],
                 absr ? &absr[s] : NULL);
    }
    return;
  }
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (size_t i = 0; i < M; i++)
    for (size_t j = 0; j < N; j++)
      fref_entry(dt, a, b, K, (ptrdiff_t)i * rsa, csa, (ptrdiff_t)j * csb, rsb,
                 &ref[i * N + j], absr ? &absr[i * N + j] : NULL);
}

void orc_err_metrics(int dt, const void *c, size_t M, size_t N, ptrdiff_t rso,
                     ptrdiff_t cso, const double *ref, const double *absr,
                     double tau, double *out) {
  double maxr = 0.0;
  for (size_t i = 0; i < M * N; i++)
    if (fabs(ref[i]) > maxr) maxr = fabs(ref[i]);
  double cw = 0.0, nw = 0.0, ew = 0.0, ma = 0.0;
  for (size_t i = 0; i < M; i++)
    for (size_t j = 0; j < N; j++) {
      double v = load_real(c, dt, (ptrdiff_t)i * rso + (ptrdiff_t)j * cso);
      double r = ref[i * N + j];
      double d = fabs(v - r);
      if (d != d) d = INFINITY; /* NaN counts as unbounded error */
      if (d > ma) ma = d;
      double den = absr ? absr[i * N + j] : fabs(r);
      double q = den > 0 ? d / den : (d > 0 ? INFINITY : 0.0);
      if (q > cw) cw = q;
      if (fabs(r) > tau * maxr && fabs(r) > 0) {
        double e = d / fabs(r);
        if (e > ew) ew = e;
      }
    }
  nw = maxr > 0 ? ma / maxr : ma;
  out[0] = cw;
  out[1] = nw;
  out[2] = ew;
  out[3] = ma;
}

/* ------------------------------------------------------------------------ */
/* number formats                                                            */
/* ------------------------------------------------------------------------ */

static inline uint32_t f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static inline float u2f(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

/* binary32 -> bfloat16, round to nearest even (finite inputs; inf stays inf) */
uint16_t orc_bf16_rne(float x) {
  uint32_t u = f2u(x);
  if ((u & 0x7F800000u) == 0x7F800000u) /* inf / nan: truncate, keep quiet bit */
    return (uint16_t)((u >> 16) | ((u & 0x7FFFFF) ? 0x40 : 0));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

float orc_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* x = hi + lo + r with hi = bf16(x), lo = bf16(x - hi); x - hi is exact in
 * binary32, so |r| <= 2^-9 |x - hi| <= 2^-18 |x| for normal x. */
void orc_split_bf16(const float *x, size_t n, uint16_t *hi, uint16_t *lo) {
  for (size_t i = 0; i < n; i++) {
    uint16_t h = orc_bf16_rne(x[i]);
    float rem = x[i] - orc_bf16_to_f32(h);
    hi[i] = h;
    lo[i] = orc_bf16_rne(rem);
  }
}

/* binary32 -> OCP FP8 E4M3 (bias 7, no infinities, max finite 448 = 0x7E),
 * round to nearest even, saturating to +-448.  NaN maps to 0x7F. */
uint8_t orc_e4m3_rne(float x) {
  uint32_t u = f2u(x);
  uint8_t sign = (uint8_t)((u >> 24) & 0x80);
  uint32_t a = u & 0x7FFFFFFFu;
  if (a > 0x7F800000u) return (uint8_t)(sign | 0x7F);
  if (a >= 0x43E00000u) return (uint8_t)(sign | 0x7E); /* >= 448 (and inf): saturate */
  int e = (int)(a >> 23) - 127;
  uint32_t code;
  if (e < -6) {
    /* subnormal range of e4m3: value = m * 2^-9, m in 0..7 (8 -> min normal) */
    float q = u2f(a) * 512.0f;              /* exact: power-of-two scaling */
    code = (uint32_t)rintf(q);              /* ties to even (default mode) */
  } else {
    uint32_t m = a & 0x7FFFFFu;
    uint32_t keep = m >> 20, rem = m & 0xFFFFFu;
    if (rem > 0x80000u || (rem == 0x80000u && (keep & 1u))) keep++;
    if (keep == 8) { keep = 0; e++; }
    code = ((uint32_t)(e + 7) << 3) | keep;
    if (code > 0x7E) code = 0x7E;
  }
  return (uint8_t)(sign | code);
}

void orc_e4m3_rne_n(const float *x, size_t n, uint8_t *q) {
  for (size_t i = 0; i < n; i++) q[i] = orc_e4m3_rne(x[i]);
}

float orc_e4m3_to_f32(uint8_t q) {
  int s = q >> 7, e = (q >> 3) & 0xF, m = q & 7;
  float v;
  if (e == 0xF && m == 7) return NAN;
  if (e == 0) v = ldexpf((float)m, -9);
  else v = ldexpf(1.0f + (float)m / 8.0f, e - 7);
  return s ? -v : v;
}

int orc_fp8_scale_exp(float maxabs) {
  if (!(maxabs > 0.0f)) return 0;
  int q;
  frexpf(maxabs, &q); /* maxabs = f * 2^q, f in [0.5, 1) */
  int e = 8 - q;
  if (e > 126) e = 126;
  if (e < -126) e = -126;
  return e;
}
