// batched.hip -- numc_matmul_batched: many independent GEMMs of uneven shape
// in one persistent launch, replayed from a cached HIP graph.
//
// No numc counterpart: numc's ROADMAP.md lists numc_bmm as the batched GEMM
// client (numc_bmm in matmul.cpp runs on this); numc_dot's batched loop
// (src/reduction/dot.c:409-421) calls the single-problem kernel per batch and
// runs on this too.  For a new set of arrays the host builds:
//   * a pack descriptor per operand (all packed by one launch),
//   * a problem table (packed operand addresses, output view),
//   * a work list of (problem, tile row, tile col, tile class) items ordered
//     longest-first; a problem takes 128x128 tiles when M and N both exceed
//     the family's threshold, else 64x64 (numc_batch_choice, selector.cpp),
//     and numc_matmul_batched_tune picks that threshold by timing the batch,
// and captures fill-counter -> batched pack -> persistent GEMM into a graph.
// The graph, its tables and its workspace are cached on the context and
// replayed while the same arrays (same addresses, shapes, strides) come back;
// such a call stages host-resident operands, compares the key and launches
// the graph, with no planning.
#include "gemm.h"
#include "internal.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <vector>

int numc_matmul_one(const NumcArray *a, const NumcArray *b, NumcArray *out, const NumcMatmulOpts *opts,
                    bool finish); // matmul.cpp

// One work queue per XCD: every tile of a problem goes to one XCD's queue, so
// the tiles that re-read a problem's rows and columns (or packed panels)
// share that XCD's L2.  Problems are dealt longest-first to the least loaded
// queue; each queue runs its longest tiles first; a block or wave whose queue
// is empty takes from the next ones.  work: (predicted tile time, item with
// the problem in .x), reordered queue by queue; qstart: each queue's first item.
static void deal_to_queues(std::vector<std::pair<double, int4>> &work, size_t count,
                           int (&qstart)[NUMC_BATCH_QUEUES + 1]) {
  std::vector<double> ptime(count, 0.0);
  for (const auto &x : work) ptime[x.second.x] += x.first;
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ptime[x] > ptime[y]; });
  std::vector<int> bin(count, 0);
  double load[NUMC_BATCH_QUEUES] = {0};
  for (int i : order) {
    if (ptime[i] == 0.0) continue;
    const int b = (int)(std::min_element(load, load + NUMC_BATCH_QUEUES) - load);
    bin[i] = b;
    load[b] += ptime[i];
  }
  std::stable_sort(work.begin(), work.end(), [&](const std::pair<double, int4> &x, const std::pair<double, int4> &y) {
    const int bx = bin[x.second.x], by = bin[y.second.x];
    return bx != by ? bx < by : x.first > y.first;
  });
  for (int b = 0; b <= NUMC_BATCH_QUEUES; b++) qstart[b] = 0;
  for (const auto &x : work) qstart[bin[x.second.x] + 1]++;
  for (int b = 0; b < NUMC_BATCH_QUEUES; b++) qstart[b + 1] += qstart[b];
}

namespace {
// the launch a cached graph holds, kept to relaunch its kernels eagerly when
// the context profiles kernels (events cannot time kernels inside a graph)
struct BatchLaunch {
  int family = 0;
  const PackArgs *descs = nullptr;
  const long long *prefix = nullptr;
  const int *owner = nullptr;
  long long units = 0;
  const BatchProblem *probs = nullptr;
  const int4 *items = nullptr;
  int nitems = 0;
  int *counter = nullptr;
  int grid = 0;
  bool direct = false;               // the direct kernel on the operands as they lie
  const BatchDirect *dprobs = nullptr;
};

struct BatchEntry {
  std::vector<uint64_t> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *dev = nullptr; // tables and packed operands of this batch
  size_t dev_bytes = 0;
  BatchLaunch launch;
  long long last_use = 0;
  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dev) (void)hipFree(dev);
    exec = nullptr;
    graph = nullptr;
    dev = nullptr;
    dev_bytes = 0;
    key.clear();
  }
};

// the graphs of the last few distinct batches (least recently used goes first)
constexpr int kBatchCacheEntries = 4;
struct BatchCache {
  BatchEntry entries[kBatchCacheEntries];
  long long replays = 0, tick = 0;
  void release() {
    for (auto &e : entries) e.release();
  }
  BatchEntry *find(const std::vector<uint64_t> &key) {
    for (auto &e : entries)
      if (e.exec && e.key == key) return &e;
    return nullptr;
  }
  BatchEntry *victim() {
    BatchEntry *v = &entries[0];
    for (auto &e : entries) {
      if (!e.exec) return &e;
      if (e.last_use < v->last_use) v = &e;
    }
    return v;
  }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline long long cdivll(long long a, long long b) { return (a + b - 1) / b; }

int stage_in(NumcCtx *ctx, const NumcArray *x, const void **dptr) {
  unsigned char *d = numc_dev_ptr(x);
  if (!d) return numc_get_error().code ? numc_get_error().code : NUMC_ERR_INTERNAL;
  if (x->st->residency == RES_HOST) {
    size_t span = numc_view_span(x);
    if (span) HIP_TRY(hipMemcpyAsync(d, x->data, span, hipMemcpyHostToDevice, ctx->stream), "hipMemcpy H2D");
  }
  *dptr = d;
  return 0;
}
} // namespace

extern "C" void numc_batch_cache_free(NumcCtx *ctx) {
  if (!ctx || !ctx->batch_cache) return;
  BatchCache *c = (BatchCache *)ctx->batch_cache;
  c->release();
  delete c;
  ctx->batch_cache = nullptr;
}

extern "C" NUMC_API long long numc_ctx_batch_graph_replays(NumcCtx *ctx) {
  return ctx && ctx->batch_cache ? ((BatchCache *)ctx->batch_cache)->replays : 0;
}

extern "C" NUMC_API int numc_ctx_batch_graphs_cached(NumcCtx *ctx) {
  if (!ctx || !ctx->batch_cache) return 0;
  int n = 0;
  for (auto &e : ((BatchCache *)ctx->batch_cache)->entries) n += e.exec != nullptr;
  return n;
}

extern "C" NUMC_API int numc_matmul_batched(NumcCtx *ctx, size_t count, NumcArray *const *a,
                                            NumcArray *const *b, NumcArray *const *out,
                                            const NumcMatmulOpts *opts) {
  if (!ctx || (count && (!a || !b || !out)))
    return NUMC_SET_ERROR(NUMC_ERR_NULL, "matmul_batched: NULL argument");
  if (count == 0) return 0;
  if (count > (1u << 24)) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched: too many problems");
  int family = -1;
  for (size_t i = 0; i < count; i++) {
    int f = numc_matmul_family(a[i], b[i], out[i], opts);
    if (f < 0) return f;
    if (out[i]->ctx != ctx) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched: problem %zu belongs to another context", i);
    if (i && f != family) return NUMC_SET_ERROR(NUMC_ERR_TYPE, "matmul_batched: problems mix dtypes or methods");
    family = f;
  }
  const NumcDType in_dt = a[0]->dtype, out_dt = out[0]->dtype;
  for (size_t i = 1; i < count; i++)
    if (out[i]->dtype != out_dt) return NUMC_SET_ERROR(NUMC_ERR_TYPE, "matmul_batched: output dtypes differ");
  bool same_in = true; // int8 and uint8 inputs may share a family when widening
  for (size_t i = 1; i < count; i++) same_in = same_in && a[i]->dtype == in_dt;
  const bool batched_kernel =
      family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 ||
      (family == NUMC_FAMILY_I8 && same_in && in_dt == NUMC_DTYPE_INT8 &&
       (out_dt == NUMC_DTYPE_INT8 || out_dt == NUMC_DTYPE_INT32));
  int err = numc_gpu_init(ctx);
  if (err) return err;
  if (!batched_kernel) { // other dtypes: one launch sequence per problem on the same stream
    for (size_t i = 0; i < count; i++)
      if ((err = numc_matmul_one(a[i], b[i], out[i], opts, false))) return err;
    return numc_finish(ctx);
  }

  // ---- stage operands and describe the problems ----
  const int kf = family == NUMC_FAMILY_F32 ? 8 : family == NUMC_FAMILY_SPLIT_BF16 ? 16 : 32;
  const int planes = family == NUMC_FAMILY_SPLIT_BF16 ? 2 : 1;
  const size_t ein = a[0]->elem_size, eout = out[0]->elem_size;
  struct Prob {
    int M, N, K, cls, bm, bn, kc;
    float pred; // predicted time (orders the work queue)
    const void *A, *B;
    void *C;
    long long rsa, csa, rsb, csb, rso, cso;
    long long pla, plb, offA, offB;
    int nRTa, nRTb, nKF;
  };
  const int thresh = numc_batch_threshold(family);
  // the direct kernel reads one int8 digit plane: int8 inputs only
  bool direct = numc_batch_direct(family) &&
                (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 ||
                 (family == NUMC_FAMILY_I8 && same_in && in_dt == NUMC_DTYPE_INT8));
  // int8: only when every A row and every column of B is k-contiguous (B
  // stored transposed); other int8 batches run the packed kernel
  for (size_t i = 0; direct && family == NUMC_FAMILY_I8 && i < count; i++)
    direct = a[i]->strides[1] == ein && b[i]->strides[0] == ein;
  std::vector<Prob> P(count);
  for (size_t i = 0; i < count; i++) {
    Prob &p = P[i];
    p.M = (int)a[i]->shape[0];
    p.K = (int)a[i]->shape[1];
    p.N = (int)b[i]->shape[1];
    p.rsa = (long long)(a[i]->strides[0] / ein);
    p.csa = (long long)(a[i]->strides[1] / ein);
    p.rsb = (long long)(b[i]->strides[0] / ein);
    p.csb = (long long)(b[i]->strides[1] / ein);
    p.rso = (long long)(out[i]->strides[0] / eout);
    p.cso = (long long)(out[i]->strides[1] / eout);
    p.A = p.B = nullptr;
    p.C = nullptr;
    p.cls = p.bm = p.bn = p.kc = 0;
    p.pred = 0.f;
    if (p.M == 0 || p.N == 0) continue;
    if ((err = stage_in(ctx, a[i], &p.A)) || (err = stage_in(ctx, b[i], &p.B))) return err;
    unsigned char *dc = numc_dev_ptr(out[i]);
    if (!dc) return NUMC_ERR_INTERNAL;
    if (out[i]->st->residency == RES_HOST && numc_view_span(out[i]) > out[i]->size * eout)
      HIP_TRY(hipMemcpyAsync(dc, out[i]->data, numc_view_span(out[i]), hipMemcpyHostToDevice, ctx->stream), "H2D");
    p.C = dc;
  }

  // ---- cache key: everything the graph bakes in.  A call with the same
  // arrays as the cached graph replays it and skips the planning below ----
  std::vector<uint64_t> key;
  key.reserve(4 + count * 12);
  key.push_back((uint64_t)family);
  key.push_back((uint64_t)(uint32_t)thresh | ((uint64_t)direct << 32));
  key.push_back((uint64_t)out_dt);
  key.push_back(count);
  for (size_t i = 0; i < count; i++) {
    const Prob &p = P[i];
    uint64_t v[] = {(uint64_t)(uintptr_t)p.A, (uint64_t)(uintptr_t)p.B, (uint64_t)(uintptr_t)p.C,
                    (uint64_t)p.M, (uint64_t)p.N, (uint64_t)p.K, (uint64_t)p.rsa, (uint64_t)p.csa,
                    (uint64_t)p.rsb, (uint64_t)p.csb, (uint64_t)p.rso, (uint64_t)p.cso};
    key.insert(key.end(), v, v + 12);
  }
  if (!ctx->batch_cache) ctx->batch_cache = new BatchCache();
  BatchCache *cache = (BatchCache *)ctx->batch_cache;

  BatchEntry *entry = cache->find(key);
  if (!entry && direct) {
    // ---- plan of the direct kernel: tile class per problem, work items
    // longest first; the tables are all it needs (no packed operands) ----
    std::vector<BatchDirect> dp(count);
    std::vector<std::pair<double, int4>> work;
    for (size_t i = 0; i < count; i++) {
      const Prob &p = P[i];
      BatchDirect &d = dp[i];
      memset(&d, 0, sizeof d);
      if (!p.C) continue;
      d.A = (const unsigned char *)p.A;
      d.a_rs = p.rsa * (long long)ein;
      d.a_ks = p.csa * (long long)ein;
      d.B = (const unsigned char *)p.B;
      d.b_rs = p.csb * (long long)ein; // rows of B^T
      d.b_ks = p.rsb * (long long)ein;
      d.M = p.M;
      d.N = p.N;
      d.K = p.K;
      d.C = p.C;
      d.rso = p.rso;
      d.cso = p.cso;
      d.out_bytes = (int)eout;
      // each dimension above the threshold takes 64-wide tiles, else 32
      const int tm = p.M > thresh ? 64 : 32, tn = p.N > thresh ? 64 : 32;
      const int cls = (tm == 64 ? 2 : 0) | (tn == 64 ? 1 : 0);
      // a tile's time: its multiply-adds plus a fixed start-up, so long-K
      // tiles go first
      const double tile_time = (double)tm * tn * (p.K > 0 ? p.K : 1) + 2.0e5;
      const int nbm = (int)cdivll(p.M, tm), nbn = (int)cdivll(p.N, tn);
      for (int x = 0; x < nbm; x++)
        for (int y = 0; y < nbn; y++) work.push_back({tile_time, make_int4((int)i, x, y, cls)});
    }
    if (work.size() > (size_t)INT_MAX)
      return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched: too many tiles for one launch (%zu)", work.size());
    int qstart[NUMC_BATCH_QUEUES + 1];
    deal_to_queues(work, count, qstart);
    const int nitems = (int)work.size();
    if (nitems > 0) {
      std::vector<int4> items(work.size());
      for (size_t x = 0; x < work.size(); x++) items[x] = work[x].second;
      const size_t o_prob = 0;
      const size_t o_item = align_up(o_prob + dp.size() * sizeof(BatchDirect), 256);
      const size_t o_cnt = align_up(o_item + items.size() * sizeof(int4), 256);
      const size_t total = o_cnt + NUMC_BATCH_COUNTER_BYTES;
      BatchEntry *e = cache->victim();
      HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      e->release();
      HIP_TRY(hipMalloc(&e->dev, total), "hipMalloc(batched tables)");
      e->dev_bytes = total;
      unsigned char *base = (unsigned char *)e->dev;
      HIP_TRY(hipMemcpyAsync(base + o_prob, dp.data(), dp.size() * sizeof(BatchDirect), hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipMemcpyAsync(base + o_item, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream), "H2D");
      // the queue starts follow the counters (which the graph clears)
      HIP_TRY(hipMemcpyAsync(base + o_cnt + NUMC_BATCH_QSTART * sizeof(int), qstart, sizeof qstart,
                             hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      int cus = 256;
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
      BatchLaunch &L = e->launch;
      L = BatchLaunch();
      L.family = family;
      L.direct = true;
      L.dprobs = (const BatchDirect *)(base + o_prob);
      L.items = (const int4 *)(base + o_item);
      L.nitems = nitems;
      L.counter = (int *)(base + o_cnt);
      L.grid = std::max(1, std::min(nitems, numc_batch_direct_occupancy(family) * cus));
      HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      int lerr = numc_batch_direct_launch(family, L.dprobs, L.items, L.nitems, L.counter, L.grid, ctx->stream, nullptr);
      hipGraph_t graph = nullptr;
      hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
      if (lerr) {
        if (graph) (void)hipGraphDestroy(graph);
        e->release();
        return lerr;
      }
      if (ce != hipSuccess) {
        e->release();
        return numc_hip_fail(ce, "hipStreamEndCapture");
      }
      e->graph = graph;
      ce = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
      if (ce != hipSuccess) {
        e->release();
        return numc_hip_fail(ce, "hipGraphInstantiate");
      }
      e->key = key;
      entry = e;
    }
  }
  if (!entry && !direct) {
    // ---- plan: tile class and packed layout per problem ----
    long long packed = 0;
    for (size_t i = 0; i < count; i++) {
      Prob &p = P[i];
      if (!p.C) continue;
      const BatchChoice ch = numc_batch_choice(family, p.M, p.N, p.K, thresh);
      p.cls = ch.cls;
      p.pred = ch.ms;
      numc_batch_class_shape(family, p.cls, &p.bm, &p.bn, &p.kc);
      const long long kp = cdivll(p.K > 0 ? p.K : 1, p.kc) * p.kc;
      p.nKF = (int)(kp / kf);
      p.nRTa = (int)(cdivll(p.M, p.bm) * p.bm / 32);
      p.nRTb = (int)(cdivll(p.N, p.bn) * p.bn / 32);
      p.pla = (long long)p.nRTa * p.nKF * 1024;
      p.plb = (long long)p.nRTb * p.nKF * 1024;
      p.offA = packed;
      packed += p.pla * planes;
      p.offB = packed;
      packed += p.plb * planes;
    }
    // ---- tables: work items longest-first ----
    std::vector<size_t> order(count);
    std::iota(order.begin(), order.end(), 0);
    // longest predicted tile first (a problem's time spread over its tiles)
    auto tile_time = [&](size_t i) {
      const Prob &p = P[i];
      const double tiles = (double)cdivll(p.M, p.bm ? p.bm : 1) * cdivll(p.N, p.bn ? p.bn : 1);
      return p.pred / (tiles > 0 ? tiles : 1.0);
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return tile_time(x) > tile_time(y); });
    std::vector<std::pair<double, int4>> work;
    std::vector<PackArgs> descs;
    std::vector<long long> prefix(1, 0);
    std::vector<BatchProblem> probs(count);
    for (size_t i = 0; i < count; i++) {
      const Prob &p = P[i];
      if (!p.C) continue;
      PackArgs da{}, db{};
      da.src = p.A; da.rs = p.rsa; da.ks = p.csa; da.R = p.M; da.K = p.K;
      da.nRT = p.nRTa; da.nKF = p.nKF; da.plane_bytes = p.pla;
      db.src = p.B; db.rs = p.csb; db.ks = p.rsb; db.R = p.N; db.K = p.K;
      db.nRT = p.nRTb; db.nKF = p.nKF; db.plane_bytes = p.plb;
      const int ba = numc_batch_pack_blocks(family, da), bb = numc_batch_pack_blocks(family, db);
      descs.push_back(da);
      prefix.push_back(prefix.back() + ba);
      descs.push_back(db);
      prefix.push_back(prefix.back() + bb);
    }
    for (size_t oi = 0; oi < count; oi++) {
      const size_t i = order[oi];
      const Prob &p = P[i];
      if (!p.C) continue;
      const int nbm = (int)cdivll(p.M, p.bm), nbn = (int)cdivll(p.N, p.bn);
      for (int x = 0; x < nbm; x++)
        for (int y = 0; y < nbn; y++) work.push_back({tile_time(i), make_int4((int)i, x, y, p.cls)});
    }
    if (work.size() > (size_t)INT_MAX || prefix.back() > (long long)INT_MAX)
      return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched: too many tiles for one launch (%zu work items, %lld pack blocks)",
                            work.size(), prefix.back());
    int qstart[NUMC_BATCH_QUEUES + 1];
    deal_to_queues(work, count, qstart);
    std::vector<int4> items(work.size());
    for (size_t x = 0; x < work.size(); x++) items[x] = work[x].second;
    const int ndesc = (int)descs.size();
    const int nitems = (int)items.size();

    if (nitems > 0) {
      // owner[b]: the operand (descriptor) that pack block b works on
      std::vector<int> owner((size_t)prefix.back());
      for (int d = 0; d < ndesc; d++)
        std::fill(owner.begin() + prefix[d], owner.begin() + prefix[d + 1], d);
      // device layout: descs | prefix | owner | probs | items | counter | packed operands
      size_t o_desc = 0;
      size_t o_pref = align_up(o_desc + descs.size() * sizeof(PackArgs), 256);
      size_t o_own = align_up(o_pref + prefix.size() * sizeof(long long), 256);
      size_t o_prob = align_up(o_own + owner.size() * sizeof(int), 256);
      size_t o_item = align_up(o_prob + probs.size() * sizeof(BatchProblem), 256);
      size_t o_cnt = align_up(o_item + items.size() * sizeof(int4), 256);
      size_t o_pack = align_up(o_cnt + NUMC_BATCH_COUNTER_BYTES, 256);
      size_t total = o_pack + (size_t)packed + 256;
      BatchEntry *e = cache->victim();
      HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
      e->release();
      HIP_TRY(hipMalloc(&e->dev, total), "hipMalloc(batched workspace)");
      e->dev_bytes = total;
      unsigned char *base = (unsigned char *)e->dev;
      // fill in device addresses now that the workspace exists
      size_t di = 0;
      for (size_t i = 0; i < count; i++) {
        const Prob &p = P[i];
        BatchProblem &bp = probs[i];
        memset(&bp, 0, sizeof bp);
        if (!p.C) continue;
        descs[di].dst = base + o_pack + p.offA;
        descs[di + 1].dst = base + o_pack + p.offB;
        di += 2;
        bp.A = base + o_pack + p.offA;
        bp.B = base + o_pack + p.offB;
        bp.a_plane = p.pla;
        bp.b_plane = p.plb;
        bp.nKF = p.nKF;
        bp.nKC = (int)(cdivll(p.K > 0 ? p.K : 1, p.kc));
        bp.M = p.M;
        bp.N = p.N;
        bp.C = p.C;
        bp.rso = p.rso;
        bp.cso = p.cso;
        bp.out_bytes = (int)eout;
      }
      HIP_TRY(hipMemcpyAsync(base + o_desc, descs.data(), descs.size() * sizeof(PackArgs), hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipMemcpyAsync(base + o_pref, prefix.data(), prefix.size() * sizeof(long long), hipMemcpyHostToDevice, ctx->stream), "H2D");
      if (!owner.empty())
        HIP_TRY(hipMemcpyAsync(base + o_own, owner.data(), owner.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipMemcpyAsync(base + o_prob, probs.data(), probs.size() * sizeof(BatchProblem), hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipMemcpyAsync(base + o_item, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipMemcpyAsync(base + o_cnt + NUMC_BATCH_QSTART * sizeof(int), qstart, sizeof qstart,
                             hipMemcpyHostToDevice, ctx->stream), "H2D");
      HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");

      int cus = 256;
      (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
      const int grid = std::max(1, std::min(nitems, numc_batch_occupancy(family) * cus));
      BatchLaunch &L = e->launch;
      L = BatchLaunch();
      L.family = family;
      L.descs = (const PackArgs *)(base + o_desc);
      L.prefix = (const long long *)(base + o_pref);
      L.owner = (const int *)(base + o_own);
      L.units = prefix.back();
      L.probs = (const BatchProblem *)(base + o_prob);
      L.items = (const int4 *)(base + o_item);
      L.nitems = nitems;
      L.counter = (int *)(base + o_cnt);
      L.grid = grid;
      HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      int lerr = numc_batch_launch(L.family, L.descs, L.prefix, L.owner, L.units, L.probs, L.items, L.nitems,
                                   L.counter, L.grid, ctx->stream, nullptr);
      hipGraph_t graph = nullptr;
      hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
      if (lerr) {
        if (graph) (void)hipGraphDestroy(graph);
        e->release();
        return lerr;
      }
      if (ce != hipSuccess) {
        e->release();
        return numc_hip_fail(ce, "hipStreamEndCapture");
      }
      e->graph = graph;
      ce = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
      if (ce != hipSuccess) {
        e->release();
        return numc_hip_fail(ce, "hipGraphInstantiate");
      }
      e->key = key;
      entry = e;
    }
  }

  if (entry) {
    entry->last_use = ++cache->tick;
    if (ctx->prof) {
      // profiling: the same kernels launched eagerly, each timed on the stream
      const BatchLaunch &L = entry->launch;
      if (L.direct)
        err = numc_batch_direct_launch(L.family, L.dprobs, L.items, L.nitems, L.counter, L.grid, ctx->stream, ctx);
      else
        err = numc_batch_launch(L.family, L.descs, L.prefix, L.owner, L.units, L.probs, L.items, L.nitems, L.counter,
                                L.grid, ctx->stream, ctx);
      if (err) return err;
    } else {
      HIP_TRY(hipGraphLaunch(entry->exec, ctx->stream), "hipGraphLaunch");
      cache->replays++;
    }
  }
  for (size_t i = 0; i < count; i++) {
    const Prob &p = P[i];
    if (!p.C || out[i]->st->residency != RES_HOST) continue;
    size_t span = numc_view_span(out[i]);
    if (span) HIP_TRY(hipMemcpyAsync(out[i]->data, p.C, span, hipMemcpyDeviceToHost, ctx->stream), "D2H");
  }
  return numc_finish(ctx);
}

// Autotune the batch schedule: time the whole batch (graph replays) under each
// candidate schedule -- the direct kernel (FP32, split-BF16 and int8 batches) or the
// packed one, and the threshold above which a problem takes the larger tile
// class -- and keep the fastest for the family.  Returns the chosen threshold
// (numc_matmul_batched_schedule reports the kernel).
extern "C" NUMC_API int numc_matmul_batched_tune(NumcCtx *ctx, size_t count, NumcArray *const *a, NumcArray *const *b,
                                                 NumcArray *const *out, const NumcMatmulOpts *opts, int reps) {
  if (!ctx || !count || !a || !b || !out) return NUMC_SET_ERROR(NUMC_ERR_NULL, "matmul_batched_tune: NULL argument");
  const int family = numc_matmul_family(a[0], b[0], out[0], opts);
  if (family < 0) return family;
  if (family >= NUMC_FAMILY_COUNT) return NUMC_SET_ERROR(NUMC_ERR_TYPE, "matmul_batched_tune: no batched kernel for this dtype");
  int err = numc_gpu_init(ctx);
  if (err) return err;
  static const int cand[] = {0, 32, 48, 64, 96, 128, 192, 256, 1 << 30};
  constexpr int kCand = (int)(sizeof cand / sizeof cand[0]);
  const bool can_direct = family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 || family == NUMC_FAMILY_I8;
  const int nrep = std::max(2, reps);
  std::vector<hipEvent_t> ev((size_t)nrep + 1, nullptr);
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) err = NUMC_SET_ERROR(NUMC_ERR_INTERNAL, "matmul_batched_tune: hipEventCreate");
  // one schedule's time: its graph is built by an untimed call and replayed
  // once more untimed, then nrep calls are timed one by one; the fastest
  // counts (a call that met a clock change or a host delay only loses)
  auto time_schedule = [&](bool direct, int t, float &ms) -> int {
    numc_batch_set_direct(family, direct);
    numc_batch_set_threshold(family, t);
    int e;
    for (int w = 0; w < 2; w++)
      if ((e = numc_matmul_batched(ctx, count, a, b, out, opts))) return e;
    (void)hipEventRecord(ev[0], ctx->stream);
    for (int r = 0; r < nrep; r++) {
      if ((e = numc_matmul_batched(ctx, count, a, b, out, opts))) return e;
      (void)hipEventRecord(ev[(size_t)r + 1], ctx->stream);
    }
    if (hipEventSynchronize(ev[(size_t)nrep]) != hipSuccess)
      return NUMC_SET_ERROR(NUMC_ERR_INTERNAL, "matmul_batched_tune: hipEventSynchronize");
    ms = 0.f;
    for (int r = 0; r < nrep; r++) {
      float x = 0.f;
      (void)hipEventElapsedTime(&x, ev[(size_t)r], ev[(size_t)r + 1]);
      if (r == 0 || x < ms) ms = x;
    }
    return 0;
  };
  int best = numc_batch_threshold(family);
  bool best_direct = numc_batch_direct(family);
  // the clock first: the current schedule for about 20 ms, untimed
  for (int w = 0; w < 12 && !err; w++) err = numc_matmul_batched(ctx, count, a, b, out, opts);
  // every schedule once, then the three fastest again, best of both rounds
  struct Cand { bool direct; int t; float ms; };
  std::vector<Cand> res;
  for (int d = can_direct ? 1 : 0; d >= 0 && !err; d--)
    for (int i = 0; i < kCand && !err; i++) {
      float ms = 0.f;
      if (!(err = time_schedule(d != 0, cand[i], ms))) res.push_back({d != 0, cand[i], ms});
    }
  std::stable_sort(res.begin(), res.end(), [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
  if (res.size() > 3) res.resize(3);
  for (auto &c : res) {
    float ms = 0.f;
    if (err || (err = time_schedule(c.direct, c.t, ms))) break;
    c.ms = std::min(c.ms, ms);
  }
  if (!err && !res.empty()) {
    const Cand *w = &res[0];
    for (const auto &c : res)
      if (c.ms < w->ms) w = &c;
    best = w->t;
    best_direct = w->direct;
  }
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  numc_batch_set_direct(family, best_direct);
  numc_batch_set_threshold(family, best);
  return err ? err : best;
}

extern "C" NUMC_API int numc_matmul_batched_schedule(int family, int *direct, int *threshold) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT || !direct || !threshold)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched_schedule: bad argument");
  *direct = numc_batch_direct(family) ? 1 : 0;
  *threshold = numc_batch_threshold(family);
  return 0;
}

extern "C" NUMC_API int numc_matmul_batched_set_schedule(int family, int direct, int threshold) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched_set_schedule: bad family %d", family);
  if (direct && family != NUMC_FAMILY_F32 && family != NUMC_FAMILY_SPLIT_BF16 && family != NUMC_FAMILY_I8)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul_batched_set_schedule: no direct batched kernel for family %d", family);
  numc_batch_set_direct(family, direct != 0);
  numc_batch_set_threshold(family, threshold);
  return 0;
}
