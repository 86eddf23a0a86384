/* Element types.  The enumerator order is part of the ABI and matches numc's
 * include/numc/dtype.h:21-32 through FLOAT64.  BFLOAT16 and FLOAT16 are the
 * "new dtypes" of numc's ROADMAP.md, appended so that numc's codes keep their
 * values; both are stored as their 16-bit IEEE-style bit patterns. */
#ifndef NUMC_DTYPE_H
#define NUMC_DTYPE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

/* C element types of numc's dtypes (numc's sources instantiate their kernels
 * through these and the generators below). */
#ifdef __cplusplus
#define NUMC_ALIGNOF(type) alignof(type)
#else
#define NUMC_ALIGNOF(type) _Alignof(type)
#endif
#define NUMC_INT8 int8_t
#define NUMC_INT16 int16_t
#define NUMC_INT32 int32_t
#define NUMC_INT64 int64_t
#define NUMC_UINT8 uint8_t
#define NUMC_UINT16 uint16_t
#define NUMC_UINT32 uint32_t
#define NUMC_UINT64 uint64_t
#define NUMC_FLOAT32 float
#define NUMC_FLOAT64 double
/* storage of the two 16-bit float dtypes (their bit patterns) */
#define NUMC_BFLOAT16 uint16_t
#define NUMC_FLOAT16 uint16_t

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

/* X-macro generators over numc's own ten dtypes, X(enum, C type), in enum
 * order; the 16-bit float dtypes take part in none of them (their arithmetic
 * is not the storage type's). */
#define NUMC_X_INT8(X) X(NUMC_DTYPE_INT8, NUMC_INT8)
#define NUMC_X_INT16(X) X(NUMC_DTYPE_INT16, NUMC_INT16)
#define NUMC_X_INT32(X) X(NUMC_DTYPE_INT32, NUMC_INT32)
#define NUMC_X_INT64(X) X(NUMC_DTYPE_INT64, NUMC_INT64)
#define NUMC_X_UINT8(X) X(NUMC_DTYPE_UINT8, NUMC_UINT8)
#define NUMC_X_UINT16(X) X(NUMC_DTYPE_UINT16, NUMC_UINT16)
#define NUMC_X_UINT32(X) X(NUMC_DTYPE_UINT32, NUMC_UINT32)
#define NUMC_X_UINT64(X) X(NUMC_DTYPE_UINT64, NUMC_UINT64)
#define NUMC_X_FLOAT32(X) X(NUMC_DTYPE_FLOAT32, NUMC_FLOAT32)
#define NUMC_X_FLOAT64(X) X(NUMC_DTYPE_FLOAT64, NUMC_FLOAT64)

#define GENERATE_SIGNED_INT_NUMC_TYPES(X) NUMC_X_INT8(X) NUMC_X_INT16(X) NUMC_X_INT32(X) NUMC_X_INT64(X)
#define GENERATE_UNSIGNED_INT_NUMC_TYPES(X) NUMC_X_UINT8(X) NUMC_X_UINT16(X) NUMC_X_UINT32(X) NUMC_X_UINT64(X)
#define GENERATE_INT_NUMC_TYPES(X) GENERATE_SIGNED_INT_NUMC_TYPES(X) GENERATE_UNSIGNED_INT_NUMC_TYPES(X)
#define GENERATE_FLOAT_NUMC_TYPES(X) NUMC_X_FLOAT32(X) NUMC_X_FLOAT64(X)
#define GENERATE_NUMC_TYPES(X) GENERATE_INT_NUMC_TYPES(X) GENERATE_FLOAT_NUMC_TYPES(X)
#define GENERATE_INT8_INT16_NUMC_TYPES(X) NUMC_X_INT8(X) NUMC_X_INT16(X) NUMC_X_UINT8(X) NUMC_X_UINT16(X)
#define GENERATE_SIGNED_INT8_INT16_INT32_NUMC_TYPES(X) NUMC_X_INT8(X) NUMC_X_INT16(X) NUMC_X_INT32(X)
#define GENERATE_SIGNED_64BIT_NUMC_TYPES(X) NUMC_X_INT64(X)
#define GENERATE_INT32_NUMC_TYPES(X) NUMC_X_INT32(X) NUMC_X_UINT32(X)
#define GENERATE_INT64_NUMC_TYPES(X) NUMC_X_INT64(X) NUMC_X_UINT64(X)
#define GENERATE_64BIT_NUMC_TYPES(X) NUMC_X_INT64(X) NUMC_X_UINT64(X) NUMC_X_FLOAT64(X)
#define GENERATE_32BIT_NUMC_TYPES(X) NUMC_X_INT32(X) NUMC_X_UINT32(X) NUMC_X_FLOAT32(X)

static inline bool numc_dtype_is_float(NumcDType dt) {
  return dt == NUMC_DTYPE_FLOAT32 || dt == NUMC_DTYPE_FLOAT64 || dt == NUMC_DTYPE_BFLOAT16 ||
         dt == NUMC_DTYPE_FLOAT16;
}

static inline bool numc_dtype_is_signed(NumcDType dt) {
  return dt == NUMC_DTYPE_INT8 || dt == NUMC_DTYPE_INT16 ||
         dt == NUMC_DTYPE_INT32 || dt == NUMC_DTYPE_INT64 ||
         numc_dtype_is_float(dt);
}

/* alignment of an element (the 16-bit floats: that of their storage) */
static inline size_t numc_dtype_align(NumcDType dt) {
  switch (dt) {
  case NUMC_DTYPE_INT8: return NUMC_ALIGNOF(NUMC_INT8);
  case NUMC_DTYPE_UINT8: return NUMC_ALIGNOF(NUMC_UINT8);
  case NUMC_DTYPE_INT16: return NUMC_ALIGNOF(NUMC_INT16);
  case NUMC_DTYPE_UINT16: return NUMC_ALIGNOF(NUMC_UINT16);
  case NUMC_DTYPE_BFLOAT16: return NUMC_ALIGNOF(NUMC_BFLOAT16);
  case NUMC_DTYPE_FLOAT16: return NUMC_ALIGNOF(NUMC_FLOAT16);
  case NUMC_DTYPE_INT32: return NUMC_ALIGNOF(NUMC_INT32);
  case NUMC_DTYPE_UINT32: return NUMC_ALIGNOF(NUMC_UINT32);
  case NUMC_DTYPE_FLOAT32: return NUMC_ALIGNOF(NUMC_FLOAT32);
  case NUMC_DTYPE_INT64: return NUMC_ALIGNOF(NUMC_INT64);
  case NUMC_DTYPE_UINT64: return NUMC_ALIGNOF(NUMC_UINT64);
  default: return NUMC_ALIGNOF(NUMC_FLOAT64);
  }
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
