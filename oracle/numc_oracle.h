/*
 * numc_oracle.h -- CPU restatement of numc's matmul semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker.
 *
 * What each function restates (reference paths relative to the numc tree):
 *   orc_matmul_int     src/matmul/kernel.h:18-62 (numc_matmul_naive): the
 *                      exact sum of products reduced mod 2^w for every integer
 *                      dtype; strides follow src/matmul/matmul.c:359-363 but
 *                      all six strides are honoured.
 *   orc_matmul_fref    FP64 reference for float32 inputs (products exact in
 *                      binary64) and double-double reference for float64
 *                      inputs, plus the componentwise denominator |A|.|B|.
 *   orc_check_matmul   src/matmul/dispatch.h:20-59 (validation order/codes).
 *   orc_bf16_rne / orc_split_bf16 / orc_e4m3_* / orc_fp8_scale_exp
 *                      the number formats used by the split-BF16 and FP8
 *                      variants (not in numc; pinned by exhaustive and
 *                      closed-form checks in tests/test_oracle.py).
 *   orc_rng_*          the seeded input generator (splitmix64 -> xoshiro256**)
 *                      documented bit-exactly below.
 *
 * Parity status of every function is listed in the header of
 * tests/test_oracle.py.
 */
#ifndef NUMC_ORACLE_H
#define NUMC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dtype codes: identical numbering to numc's NumcDType (include/numc/dtype.h:21-32) */
enum {
  ORC_I8 = 0, ORC_I16, ORC_I32, ORC_I64,
  ORC_U8, ORC_U16, ORC_U32, ORC_U64,
  ORC_F32, ORC_F64
};

size_t orc_dtype_size(int dt);

/* ---- seeded generator ------------------------------------------------
 * Element i of a tensor filled with (seed, stream) comes from block
 * b = i >> 16.  The block's xoshiro256** state is four successive outputs of
 * splitmix64 started at  seed ^ (stream * 0x9E3779B97F4A7C15) ^ (b * 0xD1B54A32D192ED03).
 * Within a block, elements draw successive outputs in index order.
 *   uniform f32 [lo,hi):  u = (x >> 40) * 2^-24 ; v = lo + (hi-lo)*u  (float math)
 *   uniform f64 [lo,hi):  u = (x >> 11) * 2^-53 ; v = lo + (hi-lo)*u
 *   normal (f32/f64):     pairs of draws x,y -> u1 = ((x>>11)+1)*2^-53,
 *                         u2 = (y>>11)*2^-53, r = sqrt(-2 ln u1),
 *                         z0 = r cos(2 pi u2), z1 = r sin(2 pi u2) (double),
 *                         rounded to the dtype; block length is even.
 *   integer [lo,hi]:      span = hi-lo+1; x32 = x >> 32; rejection of
 *                         x32 >= (2^32 / span) * span; v = lo + x32 % span.
 */
uint64_t orc_splitmix64(uint64_t *state);
typedef struct { uint64_t s[4]; } orc_rng;
void orc_rng_block(orc_rng *r, uint64_t seed, uint64_t stream, uint64_t block);
uint64_t orc_rng_next(orc_rng *r);

void orc_fill_uniform(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream,
                      double lo, double hi);
void orc_fill_normal(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream);
void orc_fill_int(void *dst, int dt, size_t n, uint64_t seed, uint64_t stream,
                  int64_t lo, int64_t hi);

/* ---- validation ------------------------------------------------------- */
/* Returns 0 or the numc error code (-1 NULL, -4 TYPE, -3 SHAPE), in the
 * order NULL -> dtype -> ndim -> inner dim -> out shape. */
int orc_check_matmul(int have_a, int have_b, int have_out, int dta, int dtb,
                     int dto, int nda, int ndb, int ndo, const size_t *sa,
                     const size_t *sb, const size_t *so);

/* ---- exact integer matmul (mod 2^w) ---------------------------------------
 * strides are in ELEMENTS and may be any non-negative values.
 * dt_out may be wider than dt_in (widening int8 -> int32 mode); the sum is
 * then reduced mod 2^(8*sizeof(out)). */
void orc_matmul_int(int dt_in, int dt_out, const void *a, const void *b, void *out,
                    size_t M, size_t K, size_t N,
                    ptrdiff_t rsa, ptrdiff_t csa, ptrdiff_t rsb, ptrdiff_t csb,
                    ptrdiff_t rso, ptrdiff_t cso);

/* ---- floating point references ---------------------------------------------
 * ref[i*N+j]  = sum_k a_ik b_kj  (binary64 for f32 inputs, double-double
 *               rounded to binary64 for f64 inputs)
 * absr[i*N+j] = sum_k |a_ik| |b_kj|   (may be NULL)
 * rows/cols restrict the computation to selected entries when nsel > 0:
 * entry s is (rows[s], cols[s]) and results go to ref[s] / absr[s]. */
void orc_matmul_fref(int dt, const void *a, const void *b, size_t M, size_t K,
                     size_t N, ptrdiff_t rsa, ptrdiff_t csa, ptrdiff_t rsb,
                     ptrdiff_t csb, double *ref, double *absr, size_t nsel,
                     const int64_t *rows, const int64_t *cols);

/* componentwise / normwise error of c (dt F32 or F64, strided) against ref.
 * out[0] = max |c-r| / absr   (entries with absr == 0 count only if c != r)
 * out[1] = max |c-r| / max|r|
 * out[2] = max |c-r|/|r| over |r| > tau * max|r|
 * out[3] = max |c-r| */
void orc_err_metrics(int dt, const void *c, size_t M, size_t N, ptrdiff_t rso,
                     ptrdiff_t cso, const double *ref, const double *absr,
                     double tau, double *out);

/* ---- number formats -------------------------------------------------------- */
uint16_t orc_bf16_rne(float x);
float orc_bf16_to_f32(uint16_t h);
void orc_split_bf16(const float *x, size_t n, uint16_t *hi, uint16_t *lo);
uint8_t orc_e4m3_rne(float x);        /* OCP e4m3fn, saturating to +-448 */
void orc_e4m3_rne_n(const float *x, size_t n, uint8_t *q); /* the same over an array */
float orc_e4m3_to_f32(uint8_t q);
/* exponent e such that the row maximum times 2^e lies in [128, 256);
 * 0 for an all-zero row. */
int orc_fp8_scale_exp(float maxabs);

#ifdef __cplusplus
}
#endif
#endif
