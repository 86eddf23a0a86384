// internal.h -- private structures of libnumc_amd (host side).
#pragma once

#include <hip/hip_runtime.h>

#include "numc/numc.h"

#include <cstddef>
#include <cstdint>

#define NUMC_INLINE_DIMS 8
#define NUMC_BLOCK_BYTES ((size_t)8 << 20) /* default arena block, as numc (internal.h:26) */
#define NUMC_DATA_ALIGN 256                 /* device-friendly; numc uses 32 */

// One arena block: host bytes plus a device twin of the same size, created
// the first time an array stored in the block is used on the GPU.  Offsets
// match: device address = dev + (host address - host).
struct ArenaBlock {
  ArenaBlock *next;
  size_t capacity, index;
  unsigned char *host;
  unsigned char *dev;
  bool pinned;
};

enum Residency { RES_HOST = 0, RES_DEVICE = 1 };

// Storage shared by an array and all of its views.
struct Storage {
  ArenaBlock *block;
  unsigned char *host; // first byte
  size_t bytes;
  int residency;
};

struct Workspace {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct NumcCtx {
  ArenaBlock *head, *current;
  // GPU state, created on first device use
  bool gpu_ready;
  int device;
  hipStream_t stream;
  bool synchronous;
  hipEvent_t events[16];
  bool events_made;
  Workspace ws;      // packed operands, partial sums
  Workspace ws_aux;  // flags, scales, batched tables
  Workspace ws_stage; // host operands of numc_amd_gemm
  void *batch_cache; // BatchCache (batched.hip)
  void *prof;        // Profiler (runtime.cpp), null unless profiling is on
  void *tune;        // autotuner scratch operands (autotune.hip)
};

// ---- kernel timing on the context stream (numc_ctx_profile) ----
enum ProfCat { PROF_PACK = 0, PROF_GEMM = 1, PROF_REDUCE = 2, PROF_BATCHED = 3, PROF_CATS = 4 };
void numc_prof_begin(NumcCtx *ctx, int cat, int family);
void numc_prof_end(NumcCtx *ctx);

struct NumcArray {
  NumcCtx *ctx;
  void *data;
  size_t *shape, *strides;
  size_t shape_buf[NUMC_INLINE_DIMS], strides_buf[NUMC_INLINE_DIMS];
  size_t dim, elem_size, size, capacity;
  bool is_contiguous;
  NumcDType dtype;
  Storage *st;
};

// ---- runtime helpers (runtime.cpp) ----
void *numc_arena_alloc(NumcCtx *ctx, size_t size, size_t align, ArenaBlock **blk);
int numc_gpu_init(NumcCtx *ctx);                       // 0 or error code (sets message)
int numc_hip_fail(hipError_t e, const char *what);     // maps a HIP error to a numc code
unsigned char *numc_dev_ptr(const NumcArray *arr);     // device twin of arr->data (nullptr on failure)
int numc_storage_host_current(Storage *st, NumcCtx *ctx);   // D2H if the device copy is current
// byte span [lo, hi) touched by a strided view, relative to arr->data (lo <= 0 never: strides >= 0)
size_t numc_view_span(const NumcArray *arr);
int numc_ws_reserve(NumcCtx *ctx, Workspace *ws, size_t bytes);
int numc_finish(NumcCtx *ctx); // synchronize when the context is synchronous

#define HIP_TRY(expr, what)                      \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return numc_hip_fail(_e, what); \
  } while (0)
