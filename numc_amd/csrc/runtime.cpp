// runtime.cpp -- contexts, arenas, arrays, errors and residency.
//
// Public behaviour follows numc's src/array_core.c, src/arena.h and
// src/error.c; the additions are the device twin of each arena block and the
// residency state of each storage (both private: NumcCtx and NumcArray are
// opaque in numc's ABI, include/numc/array.h:10-11).
#include "internal.h"
#include "rng.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

// ---------------------------------------------------------------------------
// errors (numc src/error.c:8-58)
// ---------------------------------------------------------------------------

static thread_local char g_err_buf[256];
static thread_local NumcError g_err = {0, nullptr};

extern "C" NUMC_API int numc_set_error_v(int code, const char *func, const char *file,
                                         int line, const char *fmt, ...) {
  (void)file;
  (void)line;
  char msg[192];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  snprintf(g_err_buf, sizeof g_err_buf, "%s: %s", func ? func : "(unknown)", msg);
  g_err.code = code;
  g_err.msg = g_err_buf;
  return code;
}

extern "C" NUMC_API int numc_set_error(int code, const char *msg) {
  return numc_set_error_v(code, "numc_set_error", __FILE__, __LINE__, "%s", msg ? msg : "");
}

extern "C" NUMC_API NumcError numc_get_error(void) { return g_err; }

extern "C" NUMC_API void numc_log_error(const NumcError *err) {
  if (!err || err->code == 0) return;
  fprintf(stderr, "[ERROR] numc:%s\n", err->msg ? err->msg : "(null)");
}

int numc_hip_fail(hipError_t e, const char *what) {
  int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? NUMC_ERR_MALLOC
                                                                           : NUMC_ERR_INTERNAL;
  return numc_set_error_v(code, what, __FILE__, __LINE__, "HIP error %d (%s)", (int)e,
                          hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// GPU presence (probed once per process)
// ---------------------------------------------------------------------------

static int gpu_count() {
  static int n = -1;
  if (n < 0) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    n = c;
  }
  return n;
}

// ---------------------------------------------------------------------------
// arena (numc src/arena.h:321-473): bump blocks, 8 MiB default, oversize
// requests get their own block, blocks after a checkpoint are released by
// restore.  Host blocks are pinned when a GPU is present so uploads and
// downloads are DMA transfers.
// ---------------------------------------------------------------------------

static ArenaBlock *block_new(size_t capacity) {
  ArenaBlock *b = (ArenaBlock *)calloc(1, sizeof(ArenaBlock));
  if (!b) return nullptr;
  b->capacity = capacity;
  if (gpu_count() > 0 &&
      hipHostMalloc((void **)&b->host, capacity, hipHostMallocDefault) == hipSuccess) {
    b->pinned = true;
  } else {
    b->host = (unsigned char *)aligned_alloc(NUMC_DATA_ALIGN,
                                             (capacity + NUMC_DATA_ALIGN - 1) & ~(size_t)(NUMC_DATA_ALIGN - 1));
    b->pinned = false;
  }
  if (!b->host) {
    free(b);
    return nullptr;
  }
  return b;
}

static void block_free(ArenaBlock *b) {
  if (b->dev) (void)hipFree(b->dev);
  if (b->pinned) (void)hipHostFree(b->host);
  else free(b->host);
  free(b);
}

void *numc_arena_alloc(NumcCtx *ctx, size_t size, size_t align, ArenaBlock **blk) {
  if (size == 0) size = 1;
  size_t need = size + align - 1;
  if (need < size) return nullptr;
  if (!ctx->current) {
    ArenaBlock *b = block_new(need > NUMC_BLOCK_BYTES ? need : NUMC_BLOCK_BYTES);
    if (!b) return nullptr;
    ctx->head = ctx->current = b;
  }
  for (;;) {
    ArenaBlock *b = ctx->current;
    uintptr_t p = (uintptr_t)(b->host + b->index);
    size_t pad = (size_t)((align - (p & (align - 1))) & (align - 1));
    if (b->index + pad + size <= b->capacity) {
      b->index += pad;
      void *r = b->host + b->index;
      b->index += size;
      if (blk) *blk = b;
      return r;
    }
    if (b->next) { // reuse blocks kept after a restore-to-earlier position
      ctx->current = b->next;
      continue;
    }
    size_t cap = b->capacity < NUMC_BLOCK_BYTES ? NUMC_BLOCK_BYTES : b->capacity;
    if (cap * 2 > cap) cap *= 2;
    if (cap < need) cap = need;
    ArenaBlock *nb = block_new(cap);
    if (!nb) return nullptr;
    b->next = nb;
    ctx->current = nb;
  }
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

extern "C" NUMC_API NumcCtx *numc_ctx_create(void) {
  NumcCtx *ctx = (NumcCtx *)calloc(1, sizeof(NumcCtx));
  if (!ctx) {
    NUMC_SET_ERROR(NUMC_ERR_MALLOC, "failed to allocate context (%zu bytes)", sizeof(NumcCtx));
    return nullptr;
  }
  ctx->synchronous = true;
  ctx->device = -1;
  return ctx;
}

int numc_gpu_init(NumcCtx *ctx) {
  if (ctx->gpu_ready) return 0;
  if (gpu_count() <= 0)
    return NUMC_SET_ERROR(NUMC_ERR_INTERNAL, "no HIP device is available (the GEMM runs only on the GPU)");
  HIP_TRY(hipGetDevice(&ctx->device), "hipGetDevice");
  HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
  ctx->gpu_ready = true;
  return 0;
}

int numc_finish(NumcCtx *ctx) {
  if (ctx->gpu_ready && ctx->synchronous)
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return 0;
}

extern "C" void numc_batch_cache_free(NumcCtx *ctx); // batched.hip
extern "C" NUMC_API int numc_autotune_release(NumcCtx *ctx); // autotune.hip

// ---------------------------------------------------------------------------
// kernel timing: events recorded on the context stream around each launch
// ---------------------------------------------------------------------------

namespace {
struct ProfRec {
  hipEvent_t start, end;
  int cat, family;
};
struct Profiler {
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  int open = -1;
  double total_ms[PROF_CATS][NUMC_FAMILY_COUNT + 1] = {};
  long long launches[PROF_CATS][NUMC_FAMILY_COUNT + 1] = {};
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  void drain() {
    for (auto &r : recs) {
      float ms = 0.f;
      (void)hipEventSynchronize(r.end);
      (void)hipEventElapsedTime(&ms, r.start, r.end);
      int f = r.family < 0 || r.family > NUMC_FAMILY_COUNT ? NUMC_FAMILY_COUNT : r.family;
      total_ms[r.cat][f] += ms;
      launches[r.cat][f] += 1;
      pool.push_back(r.start);
      pool.push_back(r.end);
    }
    recs.clear();
  }
  ~Profiler() {
    drain();
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
} // namespace

void numc_prof_begin(NumcCtx *ctx, int cat, int family) {
  Profiler *p = (Profiler *)ctx->prof;
  if (!p) return;
  if (p->recs.size() > 4096) p->drain();
  ProfRec r{p->get(), p->get(), cat, family};
  (void)hipEventRecord(r.start, ctx->stream);
  p->recs.push_back(r);
  p->open = (int)p->recs.size() - 1;
}

void numc_prof_end(NumcCtx *ctx) {
  Profiler *p = (Profiler *)ctx->prof;
  if (!p || p->open < 0) return;
  (void)hipEventRecord(p->recs[p->open].end, ctx->stream);
  p->open = -1;
}

extern "C" NUMC_API int numc_ctx_profile(NumcCtx *ctx, int enable) {
  if (!ctx) return NUMC_SET_ERROR(NUMC_ERR_NULL, "ctx is NULL");
  int err = numc_gpu_init(ctx);
  if (err) return err;
  if (enable && !ctx->prof) ctx->prof = new Profiler();
  if (!enable && ctx->prof) {
    (void)hipStreamSynchronize(ctx->stream);
    delete (Profiler *)ctx->prof;
    ctx->prof = nullptr;
  }
  return 0;
}

extern "C" NUMC_API int numc_ctx_profile_read(NumcCtx *ctx, int category, int family, double *total_ms,
                                              long long *launches, int reset) {
  if (!ctx || !total_ms || !launches) return NUMC_SET_ERROR(NUMC_ERR_NULL, "NULL argument");
  Profiler *p = (Profiler *)ctx->prof;
  if (!p) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "profiling is not enabled on this context");
  if (category < 0 || category >= PROF_CATS || family < 0 || family > NUMC_FAMILY_COUNT)
    return NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "no profile slot (%d, %d)", category, family);
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  p->drain();
  *total_ms = p->total_ms[category][family];
  *launches = p->launches[category][family];
  if (reset) {
    p->total_ms[category][family] = 0;
    p->launches[category][family] = 0;
  }
  return 0;
}

extern "C" NUMC_API void numc_ctx_free(NumcCtx *ctx) {
  if (!ctx) return;
  if (ctx->gpu_ready) (void)hipStreamSynchronize(ctx->stream);
  numc_batch_cache_free(ctx);
  if (ctx->tune) numc_autotune_release(ctx);
  if (ctx->prof) delete (Profiler *)ctx->prof;
  for (ArenaBlock *b = ctx->head; b;) {
    ArenaBlock *n = b->next;
    block_free(b);
    b = n;
  }
  if (ctx->ws.ptr) (void)hipFree(ctx->ws.ptr);
  if (ctx->ws_aux.ptr) (void)hipFree(ctx->ws_aux.ptr);
  if (ctx->ws_stage.ptr) (void)hipFree(ctx->ws_stage.ptr);
  if (ctx->events_made)
    for (auto &e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->gpu_ready) (void)hipStreamDestroy(ctx->stream);
  free(ctx);
}

extern "C" NUMC_API NumcCheckpoint numc_ctx_checkpoint(NumcCtx *ctx) {
  NumcCheckpoint cp = {nullptr, 0};
  if (ctx && ctx->current) {
    cp._block = ctx->current;
    cp._index = ctx->current->index;
  }
  return cp;
}

extern "C" NUMC_API void numc_ctx_restore(NumcCtx *ctx, NumcCheckpoint cp) {
  if (!ctx) return;
  if (ctx->gpu_ready) (void)hipStreamSynchronize(ctx->stream);
  ArenaBlock *keep = (ArenaBlock *)cp._block;
  ArenaBlock *b = keep ? keep->next : ctx->head;
  while (b) {
    ArenaBlock *n = b->next;
    block_free(b);
    b = n;
  }
  if (!keep) {
    ctx->head = ctx->current = nullptr;
    return;
  }
  keep->next = nullptr;
  keep->index = cp._index;
  ctx->current = keep;
}

extern "C" NUMC_API int numc_ctx_set_sync(NumcCtx *ctx, int synchronous) {
  if (!ctx) return NUMC_SET_ERROR(NUMC_ERR_NULL, "ctx is NULL");
  ctx->synchronous = synchronous != 0;
  return 0;
}

extern "C" NUMC_API int numc_ctx_synchronize(NumcCtx *ctx) {
  if (!ctx) return NUMC_SET_ERROR(NUMC_ERR_NULL, "ctx is NULL");
  if (ctx->gpu_ready) HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return 0;
}

extern "C" NUMC_API void *numc_ctx_stream(NumcCtx *ctx) {
  if (!ctx || numc_gpu_init(ctx)) return nullptr;
  return (void *)ctx->stream;
}

extern "C" NUMC_API int numc_ctx_timer_mark(NumcCtx *ctx, int slot) {
  if (!ctx) return NUMC_SET_ERROR(NUMC_ERR_NULL, "ctx is NULL");
  if (slot < 0 || slot >= 16) return NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "timer slot %d not in 0..15", slot);
  int err = numc_gpu_init(ctx);
  if (err) return err;
  if (!ctx->events_made) {
    for (auto &e : ctx->events) HIP_TRY(hipEventCreate(&e), "hipEventCreate");
    ctx->events_made = true;
  }
  HIP_TRY(hipEventRecord(ctx->events[slot], ctx->stream), "hipEventRecord");
  return 0;
}

extern "C" NUMC_API int numc_ctx_timer_elapsed(NumcCtx *ctx, int s0, int s1, float *ms) {
  if (!ctx || !ms) return NUMC_SET_ERROR(NUMC_ERR_NULL, "NULL argument");
  if (!ctx->events_made || s0 < 0 || s1 < 0 || s0 >= 16 || s1 >= 16)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "timer slots not recorded");
  HIP_TRY(hipEventSynchronize(ctx->events[s1]), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(ms, ctx->events[s0], ctx->events[s1]), "hipEventElapsedTime");
  return 0;
}

int numc_ws_reserve(NumcCtx *ctx, Workspace *ws, size_t bytes) {
  if (ws->bytes >= bytes) return 0;
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (ws->ptr) (void)hipFree(ws->ptr);
  ws->ptr = nullptr;
  ws->bytes = 0;
  size_t want = bytes + bytes / 4 + 4096;
  HIP_TRY(hipMalloc(&ws->ptr, want), "hipMalloc(workspace)");
  ws->bytes = want;
  return 0;
}

// ---------------------------------------------------------------------------
// arrays (numc src/array_core.c)
// ---------------------------------------------------------------------------

static bool strides_for(size_t *strides, const size_t *shape, size_t dim, size_t es) {
  if (dim == 0) return true;
  strides[dim - 1] = es;
  for (size_t i = dim - 1; i > 0; i--)
    if (__builtin_mul_overflow(strides[i], shape[i], &strides[i - 1])) return false;
  return true;
}

static int init_dims(NumcArray *arr, NumcCtx *ctx, size_t dim) {
  if (dim <= NUMC_INLINE_DIMS) {
    arr->shape = arr->shape_buf;
    arr->strides = arr->strides_buf;
    return 0;
  }
  arr->shape = (size_t *)numc_arena_alloc(ctx, dim * sizeof(size_t), alignof(size_t), nullptr);
  arr->strides = (size_t *)numc_arena_alloc(ctx, dim * sizeof(size_t), alignof(size_t), nullptr);
  if (!arr->shape || !arr->strides)
    return NUMC_SET_ERROR(NUMC_ERR_MALLOC, "failed to allocate strides/shape");
  return 0;
}

static bool dense(const NumcArray *arr) {
  size_t expect = arr->elem_size;
  for (size_t i = arr->dim; i-- > 0;) {
    if (arr->shape[i] != 1 && arr->strides[i] != expect) return false;
    expect *= arr->shape[i];
  }
  return true;
}

extern "C" NUMC_API NumcArray *numc_array_create(NumcCtx *ctx, const size_t *shape, size_t dim,
                                                 NumcDType dtype) {
  if (!ctx || (!shape && dim > 0)) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "array_create: NULL or invalid args (ctx=%p shape=%p dim=%zu)",
                   (void *)ctx, (const void *)shape, dim);
    return nullptr;
  }
  if ((unsigned)dtype >= NUMC_DTYPE_COUNT) {
    NUMC_SET_ERROR(NUMC_ERR_TYPE, "array_create: unknown dtype %d", (int)dtype);
    return nullptr;
  }
  size_t es = numc_dtype_size(dtype), size = 1, cap = 0;
  for (size_t i = 0; i < dim; i++)
    if (__builtin_mul_overflow(size, shape[i], &size)) {
      NUMC_SET_ERROR(NUMC_ERR_SHAPE, "array_create: element count overflow at dim %zu", i);
      return nullptr;
    }
  if (__builtin_mul_overflow(size, es, &cap)) {
    NUMC_SET_ERROR(NUMC_ERR_SHAPE, "array_create: byte-size overflow (size=%zu elem_size=%zu)", size, es);
    return nullptr;
  }
  NumcArray *arr = (NumcArray *)numc_arena_alloc(ctx, sizeof(NumcArray), alignof(NumcArray), nullptr);
  Storage *st = (Storage *)numc_arena_alloc(ctx, sizeof(Storage), alignof(Storage), nullptr);
  if (!arr || !st) {
    NUMC_SET_ERROR(NUMC_ERR_MALLOC, "array_create: failed to allocate header");
    return nullptr;
  }
  memset(arr, 0, sizeof *arr);
  arr->ctx = ctx;
  if (init_dims(arr, ctx, dim)) return nullptr;
  ArenaBlock *blk = nullptr;
  arr->data = numc_arena_alloc(ctx, cap, NUMC_DATA_ALIGN, &blk);
  if (!arr->data) {
    NUMC_SET_ERROR(NUMC_ERR_MALLOC, "array_create: failed to allocate %zu bytes of data", cap);
    return nullptr;
  }
  if (dim) {
    memcpy(arr->shape, shape, dim * sizeof(size_t));
    if (!strides_for(arr->strides, shape, dim, es)) {
      NUMC_SET_ERROR(NUMC_ERR_SHAPE, "array_create: stride computation overflow");
      return nullptr;
    }
  }
  st->block = blk;
  st->host = (unsigned char *)arr->data;
  st->bytes = cap;
  st->residency = RES_HOST;
  arr->st = st;
  arr->dim = dim;
  arr->is_contiguous = true;
  arr->elem_size = es;
  arr->size = size;
  arr->capacity = cap;
  arr->dtype = dtype;
  return arr;
}

int numc_storage_host_current(Storage *st, NumcCtx *ctx) {
  if (ctx->gpu_ready && !ctx->synchronous)
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (st->residency != RES_DEVICE) return 0;
  ArenaBlock *b = st->block;
  unsigned char *dev = b->dev + (st->host - b->host);
  HIP_TRY(hipMemcpyAsync(st->host, dev, st->bytes, hipMemcpyDeviceToHost, ctx->stream), "hipMemcpy D2H");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  st->residency = RES_HOST;
  return 0;
}

static void host_current(const NumcArray *arr) {
  if (arr && arr->st) (void)numc_storage_host_current(arr->st, arr->ctx);
}

unsigned char *numc_dev_ptr(const NumcArray *arr) {
  ArenaBlock *b = arr->st->block;
  if (!b->dev) {
    if (numc_gpu_init(arr->ctx)) return nullptr;
    hipError_t e = hipMalloc((void **)&b->dev, b->capacity);
    if (e != hipSuccess) {
      b->dev = nullptr;
      numc_hip_fail(e, "hipMalloc(arena block twin)");
      return nullptr;
    }
  }
  return b->dev + ((unsigned char *)arr->data - b->host);
}

size_t numc_view_span(const NumcArray *arr) {
  if (arr->size == 0) return 0;
  size_t span = arr->elem_size;
  for (size_t i = 0; i < arr->dim; i++) span += (arr->shape[i] - 1) * arr->strides[i];
  return span;
}

extern "C" NUMC_API bool numc_array_is_contiguous(NumcArray *arr) {
  size_t expect = arr->elem_size;
  for (size_t i = arr->dim; i-- > 0;) {
    if (arr->strides[i] != expect) return false;
    expect *= arr->shape[i];
  }
  return true;
}

static void copy_strided_to_dense(const NumcArray *src, unsigned char *dst) {
  if (src->size == 0) return;
  if (src->dim == 0 || dense(src)) {
    memcpy(dst, src->data, src->size * src->elem_size);
    return;
  }
  std::vector<size_t> coord(src->dim, 0);
  const unsigned char *base = (const unsigned char *)src->data;
  for (size_t i = 0; i < src->size; i++) {
    size_t off = 0;
    for (size_t d = 0; d < src->dim; d++) off += coord[d] * src->strides[d];
    memcpy(dst, base + off, src->elem_size);
    dst += src->elem_size;
    for (size_t d = src->dim; d-- > 0;) {
      if (++coord[d] < src->shape[d]) break;
      coord[d] = 0;
    }
  }
}

extern "C" NUMC_API int numc_array_contiguous(NumcArray *arr) {
  if (!arr) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "array_contiguous: arr is NULL");
    return -1;
  }
  if (numc_array_is_contiguous(arr)) return 0;
  host_current(arr);
  ArenaBlock *blk = nullptr;
  size_t bytes = arr->size * arr->elem_size;
  unsigned char *nd = (unsigned char *)numc_arena_alloc(arr->ctx, bytes, NUMC_DATA_ALIGN, &blk);
  Storage *st = (Storage *)numc_arena_alloc(arr->ctx, sizeof(Storage), alignof(Storage), nullptr);
  if (!nd || !st) {
    NUMC_SET_ERROR(NUMC_ERR_MALLOC, "array_contiguous: failed to allocate %zu bytes", bytes);
    return -1;
  }
  copy_strided_to_dense(arr, nd);
  st->block = blk;
  st->host = nd;
  st->bytes = bytes;
  st->residency = RES_HOST;
  arr->st = st;
  arr->data = nd;
  arr->capacity = bytes;
  strides_for(arr->strides, arr->shape, arr->dim, arr->elem_size);
  arr->is_contiguous = true;
  return 0;
}

extern "C" NUMC_API NumcArray *numc_array_zeros(NumcCtx *ctx, const size_t *shape, size_t dim,
                                                NumcDType dtype) {
  NumcArray *arr = numc_array_create(ctx, shape, dim, dtype);
  if (arr) memset(arr->data, 0, arr->capacity);
  return arr;
}

extern "C" NUMC_API NumcArray *numc_array_fill(NumcCtx *ctx, const size_t *shape, size_t dim,
                                               NumcDType dtype, const void *value) {
  if (!value) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "array_fill: value is NULL");
    return nullptr;
  }
  NumcArray *arr = numc_array_create(ctx, shape, dim, dtype);
  if (!arr) return nullptr;
  unsigned char *p = (unsigned char *)arr->data;
  size_t n = dim == 0 ? 1 : arr->size;
  for (size_t i = 0; i < n; i++) memcpy(p + i * arr->elem_size, value, arr->elem_size);
  return arr;
}

extern "C" NUMC_API void numc_array_write(NumcArray *arr, const void *data) {
  if (!arr || !data) return;
  host_current(arr);
  memcpy(arr->data, data, arr->capacity);
  arr->st->residency = RES_HOST;
}

extern "C" NUMC_API NumcArray *numc_array_copy(const NumcArray *arr) {
  if (!arr) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "array_copy: arr is NULL");
    return nullptr;
  }
  host_current(arr);
  NumcArray *c = numc_array_create(arr->ctx, arr->shape, arr->dim, arr->dtype);
  if (!c) return nullptr;
  copy_strided_to_dense(arr, (unsigned char *)c->data);
  return c;
}

// reshape: numc src/array_core.c:456-497 (same checks, codes and messages).
// One check more: a non-contiguous view is refused with NUMC_ERR_SHAPE, where
// numc would give it dense strides over the view's data (other elements).
extern "C" NUMC_API int numc_array_reshape(NumcArray *arr, const size_t *new_shape, size_t new_dim) {
  if (!arr || !new_shape || new_dim == 0) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "reshape: NULL or invalid args (arr=%p new_shape=%p new_dim=%zu)",
                   (void *)arr, (const void *)new_shape, new_dim);
    return -1;
  }
  size_t size = 1;
  for (size_t i = 0; i < new_dim; i++)
    if (__builtin_mul_overflow(size, new_shape[i], &size)) {
      NUMC_SET_ERROR(NUMC_ERR_SHAPE, "reshape: element count overflow at dim %zu", i);
      return -1;
    }
  if (size != arr->size) {
    NUMC_SET_ERROR(NUMC_ERR_SHAPE, "reshape: cannot reshape array of size %zu into size %zu", arr->size, size);
    return -1;
  }
  if (!dense(arr)) {
    NUMC_SET_ERROR(NUMC_ERR_SHAPE, "reshape: array is not contiguous");
    return -1;
  }
  if (new_dim > NUMC_INLINE_DIMS && arr->dim <= NUMC_INLINE_DIMS && init_dims(arr, arr->ctx, new_dim))
    return -1;
  arr->dim = new_dim;
  memcpy(arr->shape, new_shape, new_dim * sizeof(size_t));
  strides_for(arr->strides, new_shape, new_dim, arr->elem_size);
  arr->is_contiguous = numc_array_is_contiguous(arr);
  return 0;
}

// numc src/array_core.c:499-521: a dense copy, then reshape
extern "C" NUMC_API NumcArray *numc_array_reshape_copy(const NumcArray *arr, const size_t *new_shape,
                                                       size_t new_dim) {
  if (!arr || !new_shape || new_dim == 0) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "reshape_copy: NULL or invalid args");
    return nullptr;
  }
  NumcArray *c = numc_array_copy(arr);
  if (!c || numc_array_reshape(c, new_shape, new_dim) < 0) return nullptr;
  return c;
}

// transpose: numc src/array_core.c:523-560 (permutes shape and strides; no data moves)
extern "C" NUMC_API int numc_array_transpose(NumcArray *arr, const size_t *axes) {
  if (!arr || !axes) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "transpose: NULL arg (arr=%p axes=%p)", (void *)arr, (const void *)axes);
    return -1;
  }
  std::vector<size_t> sh(arr->dim), sd(arr->dim);
  uint64_t seen = 0;
  for (size_t i = 0; i < arr->dim; i++) {
    size_t ax = axes[i];
    if (ax >= arr->dim) {
      NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "transpose: axis %zu out of range (ndim=%zu)", ax, arr->dim);
      return -1;
    }
    if (ax < 64 && (seen & (1ULL << ax))) {
      NUMC_SET_ERROR(NUMC_ERR_SHAPE, "transpose: repeated axis %zu", ax);
      return -1;
    }
    if (ax < 64) seen |= 1ULL << ax;
    sh[i] = arr->shape[ax];
    sd[i] = arr->strides[ax];
  }
  memcpy(arr->shape, sh.data(), arr->dim * sizeof(size_t));
  memcpy(arr->strides, sd.data(), arr->dim * sizeof(size_t));
  arr->is_contiguous = numc_array_is_contiguous(arr);
  return 0;
}

// numc src/array_core.c:562-581: a dense copy, then transpose
extern "C" NUMC_API NumcArray *numc_array_transpose_copy(const NumcArray *arr, const size_t *axes) {
  if (!arr || !axes) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "transpose_copy: NULL arg");
    return nullptr;
  }
  NumcArray *c = numc_array_copy(arr);
  if (!c || numc_array_transpose(c, axes) < 0) return nullptr;
  return c;
}

// slice: numc src/array_core.c:583-650 (a view sharing the parent's data;
// start/stop/step checks and codes as numc's)
extern "C" NUMC_API NumcArray *numc_array_slice(const NumcArray *arr, NumcSlice *sl) {
  if (!arr || !sl) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "slice: NULL arg (arr=%p slice=%p)", (const void *)arr, (void *)sl);
    return nullptr;
  }
  if (sl->axis >= arr->dim) {
    NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "slice: axis %zu out of range (ndim=%zu)", sl->axis, arr->dim);
    return nullptr;
  }
  size_t n = arr->shape[sl->axis];
  if (n == 0) {
    NUMC_SET_ERROR(NUMC_ERR_SHAPE, "slice: cannot slice zero-sized dimension %zu", sl->axis);
    return nullptr;
  }
  if (sl->step == 0) sl->step = 1;
  if (sl->stop == 0 || sl->stop > n) sl->stop = n;
  if (sl->start >= n) {
    NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "slice: start %zu out of range (dim_size=%zu)", sl->start, n);
    return nullptr;
  }
  if (sl->start >= sl->stop) {
    NUMC_SET_ERROR(NUMC_ERR_VALUE, "slice: empty range (start=%zu >= stop=%zu)", sl->start, sl->stop);
    return nullptr;
  }
  NumcArray *v = (NumcArray *)numc_arena_alloc(arr->ctx, sizeof(NumcArray), alignof(NumcArray), nullptr);
  if (!v) {
    NUMC_SET_ERROR(NUMC_ERR_MALLOC, "slice: failed to allocate view header");
    return nullptr;
  }
  memset(v, 0, sizeof *v);
  v->ctx = arr->ctx;
  v->dim = arr->dim;
  v->elem_size = arr->elem_size;
  v->dtype = arr->dtype;
  v->st = arr->st;
  if (init_dims(v, arr->ctx, arr->dim)) return nullptr;
  memcpy(v->shape, arr->shape, arr->dim * sizeof(size_t));
  memcpy(v->strides, arr->strides, arr->dim * sizeof(size_t));
  v->data = (unsigned char *)arr->data + sl->start * arr->strides[sl->axis];
  v->shape[sl->axis] = (sl->stop - sl->start + sl->step - 1) / sl->step;
  v->strides[sl->axis] = arr->strides[sl->axis] * sl->step;
  size_t size = 1;
  for (size_t d = 0; d < v->dim; d++) size *= v->shape[d];
  v->size = size;
  v->capacity = size * v->elem_size;
  v->is_contiguous = false;
  return v;
}

extern "C" NUMC_API size_t numc_array_size(const NumcArray *arr) { return arr->size; }
extern "C" NUMC_API size_t numc_array_capacity(const NumcArray *arr) { return arr->capacity; }
extern "C" NUMC_API size_t numc_array_elem_size(const NumcArray *arr) { return arr->elem_size; }
extern "C" NUMC_API size_t numc_array_ndim(const NumcArray *arr) { return arr->dim; }
extern "C" NUMC_API void numc_array_shape(const NumcArray *arr, size_t *shape) {
  memcpy(shape, arr->shape, arr->dim * sizeof(size_t));
}
extern "C" NUMC_API void numc_array_strides(const NumcArray *arr, size_t *strides) {
  memcpy(strides, arr->strides, arr->dim * sizeof(size_t));
}
extern "C" NUMC_API NumcDType numc_array_dtype(const NumcArray *arr) { return arr->dtype; }

extern "C" NUMC_API void *numc_array_data(const NumcArray *arr) {
  if (!arr) return nullptr;
  host_current(arr);
  return arr->data;
}

// ---------------------------------------------------------------------------
// residency
// ---------------------------------------------------------------------------

extern "C" NUMC_API int numc_array_to_device(NumcArray *arr) {
  if (!arr) return NUMC_SET_ERROR(NUMC_ERR_NULL, "array_to_device: arr is NULL");
  Storage *st = arr->st;
  if (st->residency == RES_DEVICE) return 0;
  NumcCtx *ctx = arr->ctx;
  int err = numc_gpu_init(ctx);
  if (err) return err;
  ArenaBlock *b = st->block;
  if (!b->dev) {
    NumcArray probe = *arr;
    probe.data = st->host;
    if (!numc_dev_ptr(&probe)) return numc_get_error().code;
  }
  unsigned char *dev = b->dev + (st->host - b->host);
  HIP_TRY(hipMemcpyAsync(dev, st->host, st->bytes, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  st->residency = RES_DEVICE;
  return 0;
}

extern "C" NUMC_API int numc_array_to_host(NumcArray *arr) {
  if (!arr) return NUMC_SET_ERROR(NUMC_ERR_NULL, "array_to_host: arr is NULL");
  return numc_storage_host_current(arr->st, arr->ctx);
}

extern "C" NUMC_API int numc_array_on_device(const NumcArray *arr) {
  return arr && arr->st->residency == RES_DEVICE;
}

extern "C" NUMC_API void *numc_array_device_data(const NumcArray *arr) {
  if (!arr) {
    NUMC_SET_ERROR(NUMC_ERR_NULL, "array_device_data: arr is NULL");
    return nullptr;
  }
  return numc_dev_ptr(arr);
}

// ---------------------------------------------------------------------------
// plain device copy (16-byte vectors, grid-stride): the reference kernel of
// known traffic that HBM counters are calibrated against
// ---------------------------------------------------------------------------

__global__ void k_copy16(const uint4 *src, uint4 *dst, long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}

// [p, p + bytes) lies in one device allocation (not host memory, pinned or not)
static int check_device_range(const void *p, size_t bytes, const char *what) {
  hipPointerAttribute_t at{};
  hipError_t e = hipPointerGetAttributes(&at, p);
  if (e != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError(); // an unknown (host) pointer is not an error of later calls
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "device_copy: %s (%p) is not device memory", what, p);
  }
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
    (void)hipGetLastError();
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "device_copy: %s (%p) has no device allocation", what, p);
  }
  if ((const char *)p + bytes > (const char *)base + size)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "device_copy: %zu bytes at %s run past its allocation", bytes, what);
  return 0;
}

extern "C" NUMC_API int numc_device_copy(NumcCtx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx || !dst || !src) return NUMC_SET_ERROR(NUMC_ERR_NULL, "device_copy: NULL argument");
  if (bytes % 16) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "device_copy: size must be a multiple of 16");
  if (((uintptr_t)dst | (uintptr_t)src) & 15)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "device_copy: dst and src must be 16-byte aligned");
  int err = numc_gpu_init(ctx);
  if (err) return err;
  if ((err = check_device_range(src, bytes, "src")) || (err = check_device_range(dst, bytes, "dst"))) return err;
  long long n = (long long)(bytes / 16);
  unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 2048);
  if (n) hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)src, (uint4 *)dst, n);
  HIP_TRY(hipGetLastError(), "device_copy launch");
  return numc_finish(ctx);
}

// ---------------------------------------------------------------------------
// seeded inputs (rng.h): one xoshiro256** stream per 2^16-element block, so
// the fill is parallel and independent of the thread count.  A device-current
// array fills on the GPU (generate.hip) with the same code.
// ---------------------------------------------------------------------------

int numc_fill_random_device(NumcArray *arr, const NumcFillSpec &f); // generate.hip

extern "C" NUMC_API int numc_array_fill_random(NumcArray *arr, int dist, uint64_t seed, uint64_t stream,
                                               double lo, double hi) {
  if (!arr) return NUMC_SET_ERROR(NUMC_ERR_NULL, "fill_random: arr is NULL");
  if (!dense(arr)) return NUMC_SET_ERROR(NUMC_ERR_SHAPE, "fill_random: array must be contiguous");
  if (dist < 0 || dist > 2) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "fill_random: dist %d not in 0..2", dist);
  if (dist == 2 && (hi < lo || hi - lo >= 4294967296.0))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "fill_random: integer range must satisfy 0 <= hi-lo < 2^32");
  const NumcFillSpec f = numc_fill_spec(arr->dtype, dist, seed, stream, lo, hi);
  if (arr->st->residency == RES_DEVICE) return numc_fill_random_device(arr, f);
  const size_t n = arr->size, nb = (n + 65535) >> 16;
  void *p = arr->data;
  auto work = [&](size_t b0, size_t b1) {
    for (size_t b = b0; b < b1; b++) numc_fill_block(p, n, b, f);
  };
  unsigned nt = std::thread::hardware_concurrency();
  if (nt > 16) nt = 16;
  if (nt < 1 || nb < 4) nt = 1;
  std::vector<std::thread> th;
  size_t per = (nb + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    size_t b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
    if (b0 < b1) th.emplace_back(work, b0, b1);
  }
  for (auto &t : th) t.join();
  arr->st->residency = RES_HOST;
  return 0;
}
