/* Contexts and arrays (numc's include/numc/array.h).
 *
 * A context owns every array made from it: host storage comes from a bump
 * arena, and each arena block has a device twin in HBM allocated the first
 * time an array in that block is used on the GPU, so checkpoint/restore and
 * numc_ctx_free release host and device memory together.  Arrays are opaque;
 * shapes are in elements and strides in bytes. */
#ifndef NUMC_ARRAY_H
#define NUMC_ARRAY_H

#include "numc/dtype.h"
#include "numc/export.h"
#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NumcCtx NumcCtx;
typedef struct NumcArray NumcArray;

typedef struct {
  size_t axis, start, stop, step;
} NumcSlice;

#define NUMC_SLICE(arr, ...) numc_array_slice((arr), &(NumcSlice){__VA_ARGS__})

/* Opaque arena position (numc's array.h NumcCheckpoint). */
typedef struct {
  void *_block;
  size_t _index;
} NumcCheckpoint;

/* -- context (array.h: numc_ctx_*) -- */
NUMC_API NumcCtx *numc_ctx_create(void);
NUMC_API void numc_ctx_free(NumcCtx *ctx);
NUMC_API NumcCheckpoint numc_ctx_checkpoint(NumcCtx *ctx);
NUMC_API void numc_ctx_restore(NumcCtx *ctx, NumcCheckpoint checkpoint);

/* -- creation -- */
NUMC_API NumcArray *numc_array_create(NumcCtx *ctx, const size_t *shape,
                                      size_t dim, NumcDType dtype);
NUMC_API NumcArray *numc_array_zeros(NumcCtx *ctx, const size_t *shape,
                                     size_t dim, NumcDType dtype);
NUMC_API NumcArray *numc_array_fill(NumcCtx *ctx, const size_t *shape,
                                    size_t dim, NumcDType dtype,
                                    const void *value);
NUMC_API NumcArray *numc_array_copy(const NumcArray *arr);
NUMC_API void numc_array_write(NumcArray *arr, const void *data);

/* -- shape manipulation (views share storage) -- */
NUMC_API int numc_array_reshape(NumcArray *arr, const size_t *new_shape,
                                size_t new_dim);
NUMC_API NumcArray *numc_array_reshape_copy(const NumcArray *arr,
                                            const size_t *new_shape,
                                            size_t new_dim);
NUMC_API int numc_array_transpose(NumcArray *arr, const size_t *axes);
NUMC_API NumcArray *numc_array_transpose_copy(const NumcArray *arr,
                                              const size_t *axes);
NUMC_API NumcArray *numc_array_slice(const NumcArray *arr, NumcSlice *slice);
NUMC_API bool numc_array_is_contiguous(NumcArray *arr);
NUMC_API int numc_array_contiguous(NumcArray *arr);

/* -- properties -- */
NUMC_API size_t numc_array_size(const NumcArray *arr);
NUMC_API size_t numc_array_capacity(const NumcArray *arr);
NUMC_API size_t numc_array_elem_size(const NumcArray *arr);
NUMC_API size_t numc_array_ndim(const NumcArray *arr);
NUMC_API void numc_array_shape(const NumcArray *arr, size_t *shape);
NUMC_API void numc_array_strides(const NumcArray *arr, size_t *strides);
NUMC_API NumcDType numc_array_dtype(const NumcArray *arr);
/* Host pointer to the data.  If the device copy is the current one (see
 * numc_array_to_device in numc_amd.h) it is copied back first, and the host
 * copy becomes the current one again, since the caller may write to it. */
NUMC_API void *numc_array_data(const NumcArray *arr);

#ifdef __cplusplus
}
#endif

#endif
