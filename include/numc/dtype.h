/* Element types.  The enumerator order is part of the ABI and matches numc's
 * include/numc/dtype.h:21-32 through FLOAT64.  BFLOAT16 and FLOAT16 are the
 * "new dtypes" of numc's ROADMAP.md, appended so that numc's codes keep their
 * values; both are stored as their 16-bit IEEE-style bit patterns. */
#ifndef NUMC_DTYPE_H
#define NUMC_DTYPE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

typedef enum {
  NUMC_DTYPE_INT8,
  NUMC_DTYPE_INT16,
  NUMC_DTYPE_INT32,
  NUMC_DTYPE_INT64,
  NUMC_DTYPE_UINT8,
  NUMC_DTYPE_UINT16,
  NUMC_DTYPE_UINT32,
  NUMC_DTYPE_UINT64,
  NUMC_DTYPE_FLOAT32,
  NUMC_DTYPE_FLOAT64,
  NUMC_DTYPE_BFLOAT16, /* 1 sign, 8 exponent, 7 mantissa bits (the top half of a binary32) */
  NUMC_DTYPE_FLOAT16,  /* IEEE 754 binary16 */
} NumcDType;

#define NUMC_DTYPE_COUNT 12

static inline bool numc_dtype_is_float(NumcDType dt) {
  return dt == NUMC_DTYPE_FLOAT32 || dt == NUMC_DTYPE_FLOAT64 || dt == NUMC_DTYPE_BFLOAT16 ||
         dt == NUMC_DTYPE_FLOAT16;
}

static inline bool numc_dtype_is_signed(NumcDType dt) {
  return dt == NUMC_DTYPE_INT8 || dt == NUMC_DTYPE_INT16 ||
         dt == NUMC_DTYPE_INT32 || dt == NUMC_DTYPE_INT64 ||
         numc_dtype_is_float(dt);
}

/* bytes per element */
static inline size_t numc_dtype_size(NumcDType dt) {
  switch (dt) {
  case NUMC_DTYPE_INT8: case NUMC_DTYPE_UINT8: return 1;
  case NUMC_DTYPE_INT16: case NUMC_DTYPE_UINT16: case NUMC_DTYPE_BFLOAT16: case NUMC_DTYPE_FLOAT16: return 2;
  case NUMC_DTYPE_INT32: case NUMC_DTYPE_UINT32: case NUMC_DTYPE_FLOAT32: return 4;
  default: return 8;
  }
}

#endif
