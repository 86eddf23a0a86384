/* GEMM entry points kept from numc's include/numc/math.h. */
#ifndef NUMC_MATH_H
#define NUMC_MATH_H

#include "numc/array.h"
#include "numc/export.h"

#ifdef __cplusplus
extern "C" {
#endif

/* out = a @ b for 2-D a (M x K), b (K x N), out (M x N), all one dtype.
 * Replaces numc's numc_matmul (include/numc/math.h:573, src/matmul/matmul.c:347):
 * same validation order and codes (NULL -> dtype -> ndim -> inner -> out shape),
 * computed on the GPU.  Floats compute in their own precision (FP32 MFMA for
 * float32, FP64 MFMA for float64); integers return the exact sum of products
 * reduced mod 2^w (numc_matmul_naive semantics).  Every byte stride of a, b
 * and out is honoured.  Synchronous: out holds the result on return. */
NUMC_API int numc_matmul(const NumcArray *a, const NumcArray *b,
                         NumcArray *out);

/* numc_dot (math.h:593, src/reduction/dot.c:441-448): numpy-style dot with
 * numc's validation (src/reduction/dispatch.h:199-283, same codes and
 * messages).  0-D operand: out = other * scalar; 1-D . 1-D: inner product;
 * N-D . 1-D and N-D . M-D: sum over a's last axis and b's second-to-last,
 * out shape a.shape[:-1] + b.shape[:-2] + b.shape[-1:].  Computed by the GPU
 * GEMM (several problems in one batched launch); integers give the exact sum
 * mod 2^w. */
NUMC_API int numc_dot(const NumcArray *a, const NumcArray *b, NumcArray *out);

/* numc_matmul_naive (math.h:603, matmul.c:333): same result contract; it
 * treats a, b and out as contiguous row-major (as numc does) and zeroes out
 * first.  Runs a simple one-thread-per-output GPU kernel. */
NUMC_API int numc_matmul_naive(const NumcArray *a, const NumcArray *b,
                               NumcArray *out);

#ifdef __cplusplus
}
#endif

#endif
