// selector.cpp -- learned tile selector: picks a kernel configuration per
// (family, shape, layout) with the MLP cost model trained on the autotuner's
// timings (numc_amd/data/selector.bin), falling back to the analytic model
// in matmul.cpp when no trained model matches the compiled configuration
// tables.  numc's counterpart is its compile-time threshold table
// (src/matmul/matmul.c:241-256) choosing gemmsup vs packed.
//
// Model input (numc_selector_features; numc_amd/autotune.py builds its
// training rows with it):
//   log2 M/16, log2 N/16, log2 K/16, family one-hot (7), transposed A,
//   transposed B, padded rows, log2 bm/8, log2 bn/8, log2 kc/8,
//   log2(waves)/3, log2(split)/3, log2(blocks)/16, log2(rounds of 256)/8,
//   log2(waves splitting K)/2, direct (operands read in place by one launch),
//   A read in place by the packed kernel, B read in place by the packed kernel
// Model output: log2 of the predicted call time in ms.
#include "gemm.h"
#include "internal.h"

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace {
constexpr int kDin = 24;
// families the cost model covers: all of them
constexpr int kFam = NUMC_FAMILY_COUNT;
constexpr int kFileVersion = 2;

struct HostMlp {
  int din = 0, H = 0, dout = 0;
  std::vector<float> p;
  float forward(const float *x) const {
    std::vector<float> h1(H), h2(H);
    const float *W0 = p.data(), *b0 = W0 + (size_t)din * H, *W1 = b0 + H, *b1 = W1 + (size_t)H * H, *W2 = b1 + H,
                *b2 = W2 + (size_t)H * dout;
    for (int j = 0; j < H; j++) {
      float s = b0[j];
      for (int k = 0; k < din; k++) s = std::fma(x[k], W0[k * H + j], s);
      h1[j] = s > 0 ? s : 0;
    }
    for (int j = 0; j < H; j++) {
      float s = b1[j];
      for (int k = 0; k < H; k++) s = std::fma(h1[k], W1[k * H + j], s);
      h2[j] = s > 0 ? s : 0;
    }
    float s = b2[0];
    for (int k = 0; k < H; k++) s = std::fma(h2[k], W2[k * dout], s);
    return s;
  }
};

std::mutex g_mu;
HostMlp g_model;
bool g_tried = false, g_loaded = false;
struct ShapeKey {
  int family, flags, M, N, K;
  bool operator==(const ShapeKey &o) const {
    return family == o.family && flags == o.flags && M == o.M && N == o.N && K == o.K;
  }
};
struct ShapeKeyHash {
  size_t operator()(const ShapeKey &k) const {
    uint64_t h = 1469598103934665603ULL;
    for (int v : {k.family, k.flags, k.M, k.N, k.K}) h = (h ^ (uint32_t)v) * 1099511628211ULL;
    return (size_t)h;
  }
};
std::unordered_map<ShapeKey, int, ShapeKeyHash> g_cache;
// exact shapes whose best configuration was measured (exhaustive autotune):
// consulted before the model
std::unordered_map<ShapeKey, int, ShapeKeyHash> g_tuned;
bool g_tuned_tried = false;

std::string default_path() {
  Dl_info info;
  if (dladdr((void *)&numc_select_config, &info) && info.dli_fname) {
    std::string so = info.dli_fname;
    size_t s = so.rfind('/');
    std::string dir = s == std::string::npos ? "." : so.substr(0, s);
    return dir + "/../data/selector.bin";
  }
  return "numc_amd/data/selector.bin";
}

// file: "NMLS", version 2, din, hidden, dout, nfam, per family {ncfg, ncfg x 8 ints
// (NumcKernelConfig)}, float params
bool load_file(const char *path, HostMlp *m, char *why, size_t whylen) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    snprintf(why, whylen, "cannot open %s", path);
    return false;
  }
  char magic[4];
  int hdr[5];
  bool ok = fread(magic, 1, 4, f) == 4 && memcmp(magic, "NMLS", 4) == 0 && fread(hdr, sizeof(int), 5, f) == 5 &&
            hdr[0] == kFileVersion && hdr[1] == kDin && hdr[3] == 1 && hdr[4] == kFam;
  if (!ok) {
    snprintf(why, whylen, "bad header");
    fclose(f);
    return false;
  }
  for (int fam = 0; fam < kFam && ok; fam++) {
    int n = 0;
    ok = fread(&n, sizeof(int), 1, f) == 1 && n == numc_gemm_config_count(fam);
    for (int i = 0; ok && i < n; i++) {
      int d[8];
      ok = fread(d, sizeof(int), 8, f) == 8;
      const NumcKernelConfig &c = numc_gemm_config(fam, i)->cfg;
      ok = ok && d[0] == c.bm && d[1] == c.bn && d[2] == c.kc && d[3] == c.waves_m && d[4] == c.waves_n &&
           d[5] == c.split_k && d[6] == c.waves_k && d[7] == c.direct;
    }
  }
  if (!ok) {
    snprintf(why, whylen, "configuration tables differ from the trained model");
    fclose(f);
    return false;
  }
  m->din = hdr[1];
  m->H = hdr[2];
  m->dout = hdr[3];
  size_t np = (size_t)m->din * m->H + m->H + (size_t)m->H * m->H + m->H + (size_t)m->H * m->dout + m->dout;
  m->p.resize(np);
  ok = fread(m->p.data(), sizeof(float), np, f) == np;
  fclose(f);
  if (!ok) snprintf(why, whylen, "truncated parameters");
  return ok;
}

std::string default_tuned_path() {
  std::string p = default_path();
  return p.substr(0, p.rfind('/')) + "/tuned.csv";
}

// lines "family,M,N,K,flags,cfg"; '#' starts a comment
void load_tuned(const char *path) {
  FILE *f = fopen(path, "r");
  if (!f) return;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int fam, M, N, K, fl, cfg;
    if (sscanf(line, "%d,%d,%d,%d,%d,%d", &fam, &M, &N, &K, &fl, &cfg) == 6 && cfg >= 0 &&
        cfg < numc_gemm_config_count(fam))
      g_tuned[ShapeKey{fam, fl, M, N, K}] = cfg;
  }
  fclose(f);
}

void ensure_tuned() {
  if (g_tuned_tried) return;
  g_tuned_tried = true;
  const char *env = getenv("NUMC_SELECTOR");
  if (env && strcmp(env, "off") == 0) return;
  load_tuned(default_tuned_path().c_str());
}

void ensure_loaded() {
  ensure_tuned();
  if (g_tried) return;
  g_tried = true;
  const char *env = getenv("NUMC_SELECTOR");
  if (env && strcmp(env, "off") == 0) return;
  std::string path = env && *env ? env : default_path();
  char why[256];
  g_loaded = load_file(path.c_str(), &g_model, why, sizeof why);
}
} // namespace

// features of (family, shape, layout flags, config) -- see the header comment
extern "C" NUMC_API int numc_selector_features(int family, int M, int N, int K, int flags, int cfg, float *out) {
  const GemmConfig *c = numc_gemm_config(family, cfg);
  if (!c || !out) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_features: no config %d in family %d", cfg, family);
  const double m = M > 0 ? M : 1, n = N > 0 ? N : 1, k = K > 0 ? K : 1;
  const double blocks = std::ceil(m / c->cfg.bm) * std::ceil(n / c->cfg.bn) * c->cfg.split_k;
  int x = 0;
  out[x++] = (float)(std::log2(m) / 16);
  out[x++] = (float)(std::log2(n) / 16);
  out[x++] = (float)(std::log2(k) / 16);
  for (int f = 0; f < kFam; f++) out[x++] = f == family ? 1.f : 0.f;
  out[x++] = (flags & 1) ? 1.f : 0.f;
  out[x++] = (flags & 2) ? 1.f : 0.f;
  out[x++] = (flags & 4) ? 1.f : 0.f;
  out[x++] = (float)(std::log2((double)c->cfg.bm) / 8);
  out[x++] = (float)(std::log2((double)c->cfg.bn) / 8);
  out[x++] = (float)(std::log2((double)c->cfg.kc) / 8);
  out[x++] = (float)(std::log2((double)c->cfg.waves_m * c->cfg.waves_n) / 3);
  out[x++] = (float)(std::log2((double)c->cfg.split_k) / 3);
  out[x++] = (float)(std::log2(blocks) / 16);
  out[x++] = (float)(std::log2(std::ceil(blocks / 256.0)) / 8);
  out[x++] = (float)(std::log2((double)c->cfg.waves_k) / 2);
  out[x++] = c->cfg.direct ? 1.f : 0.f;
  // whether the packed kernel reads an operand in place (16-byte aligned
  // k-contiguous rows and K a multiple of the k-chunk; arrays start aligned)
  const long long es = family == NUMC_FAMILY_F64 ? 8 : family == NUMC_FAMILY_I8 ? 1 : family >= NUMC_FAMILY_BF16 ? 2 : 4;
  const long long pad = (flags & 4) ? 16 : 0;
  const bool rows16 = (((long long)K + pad) * es) % 16 == 0;
  const bool fam_ip = !c->cfg.direct && (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_I8) && K % c->cfg.kc == 0;
  out[x++] = fam_ip && c->inplace[0][0] && !(flags & 1) && rows16 ? 1.f : 0.f;
  out[x++] = fam_ip && c->inplace[1][0] && (flags & 2) && rows16 ? 1.f : 0.f;
  return x == kDin ? 0 : NUMC_SET_ERROR(NUMC_ERR_INTERNAL, "selector_features: %d features, expected %d", x, kDin);
}

extern "C" NUMC_API int numc_selector_feature_count(void) { return kDin; }

// (re)load a trained selector; path NULL = the default next to the library
extern "C" NUMC_API int numc_selector_load(const char *path) {
  std::lock_guard<std::mutex> lk(g_mu);
  HostMlp m;
  char why[256];
  std::string p = path ? path : default_path();
  if (!load_file(p.c_str(), &m, why, sizeof why)) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_load: %s", why);
  g_model = m;
  g_loaded = true;
  g_tried = true;
  g_cache.clear();
  return 0;
}

// exhaustive autotune of one shape on this device; the best configuration is
// used for that exact shape from then on (returns its index)
extern "C" NUMC_API int numc_autotune_shape(NumcCtx *ctx, int family, int M, int N, int K, int flags, int reps,
                                            float *ms_out, int ms_len);
extern "C" NUMC_API int numc_selector_tune(NumcCtx *ctx, int family, int M, int N, int K, int flags, int reps) {
  float ms[32];
  int err = numc_autotune_shape(ctx, family, M, N, K, flags, reps, ms, 32);
  if (err) return err;
  int best = 0;
  for (int i = 1; i < numc_gemm_config_count(family); i++)
    if (ms[i] < ms[best]) best = i;
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_tuned();
  g_tuned[ShapeKey{family, flags, M, N, K}] = best;
  return best;
}

// pin a configuration for an exact shape (cfg < 0 removes the pin)
extern "C" NUMC_API int numc_selector_pin(int family, int M, int N, int K, int flags, int cfg) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT || cfg >= numc_gemm_config_count(family))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_pin: bad family/config");
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_tuned();
  if (cfg < 0) g_tuned.erase(ShapeKey{family, flags, M, N, K});
  else g_tuned[ShapeKey{family, flags, M, N, K}] = cfg;
  return 0;
}

extern "C" NUMC_API int numc_selector_unload(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_loaded = false;
  g_tried = true;
  g_cache.clear();
  g_tuned.clear();
  g_tuned_tried = true;
  return 0;
}

extern "C" NUMC_API int numc_selector_active(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  return g_loaded ? 1 : 0;
}

// predicted log2(ms) of every config of a family (host inference)
extern "C" NUMC_API int numc_selector_predict(int family, int M, int N, int K, int flags, float *out, int len) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  if (!g_loaded) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_predict: no trained selector loaded");
  if (family < 0 || family >= kFam) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_predict: family %d not modelled", family);
  const int n = numc_gemm_config_count(family);
  if (len < n) return NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "selector_predict: need %d slots", n);
  float x[kDin];
  for (int i = 0; i < n; i++) {
    numc_selector_features(family, M, N, K, flags, i, x);
    out[i] = g_model.forward(x);
  }
  return 0;
}

// Candidates for one call: the settled choice (an exact-shape entry, or the
// model's best when it is confident); where the model's predictions of its
// best few configurations lie within 2^0.1 of each other, those (at most 3)
// in predicted order for the caller to time once; and every configuration
// when the best predicted call takes under 2^-5 ms, where run-to-run noise is
// of the order of the differences and timing them all costs little.  At most
// `max`.  Returns the count, 0 without a trained model.  flags bits 0-2 are
// the layout (what the model sees); higher bits only separate the
// exact-shape entries.
extern "C" NUMC_API int numc_selector_candidates(int family, int M, int N, int K, int flags, int allow_direct, int *out,
                                                 int max) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT || !out || max < 1)
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "selector_candidates: bad argument");
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  auto tt = g_tuned.find(ShapeKey{family, flags, M, N, K});
  if (tt != g_tuned.end() && (allow_direct || !numc_gemm_config(family, tt->second)->cfg.direct)) {
    out[0] = tt->second;
    return 1;
  }
  if (!g_loaded) return 0;
  const int n = numc_gemm_config_count(family);
  float x[kDin];
  std::vector<std::pair<float, int>> pred;
  for (int i = 0; i < n; i++) {
    if (!allow_direct && numc_gemm_config(family, i)->cfg.direct) continue;
    numc_selector_features(family, M, N, K, flags & 7, i, x);
    pred.push_back({g_model.forward(x), i});
  }
  if (pred.empty()) return 0;
  std::sort(pred.begin(), pred.end());
  int k = 1;
  if (pred[0].first < -5.f) {
    k = std::min((int)pred.size(), max);
  } else {
    while (k < (int)pred.size() && k < std::min(max, 3) && pred[k].first - pred[0].first <= 0.1f) k++;
  }
  for (int i = 0; i < k; i++) out[i] = pred[i].second;
  return k;
}

// remember the measured winner of an exact shape (flags as in candidates)
void numc_selector_record(int family, int M, int N, int K, int flags, int cfg) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_tuned();
  g_tuned[ShapeKey{family, flags, M, N, K}] = cfg;
}

int numc_selector_choose(int family, int M, int N, int K, int flags, bool allow_direct) {
  std::lock_guard<std::mutex> lk(g_mu);
  ensure_loaded();
  // the cache keys carry whether direct configurations were allowed (bit 8)
  const ShapeKey key{family, flags | (allow_direct ? 0 : 256), M, N, K};
  auto tt = g_tuned.find(ShapeKey{family, flags, M, N, K});
  if (tt != g_tuned.end() && (allow_direct || !numc_gemm_config(family, tt->second)->cfg.direct)) return tt->second;
  if (!g_loaded || family >= kFam) return -1;
  auto it = g_cache.find(key);
  if (it != g_cache.end()) return it->second;
  const int n = numc_gemm_config_count(family);
  float x[kDin];
  int best = -1;
  float bt = 0;
  for (int i = 0; i < n; i++) {
    if (!allow_direct && numc_gemm_config(family, i)->cfg.direct) continue;
    numc_selector_features(family, M, N, K, flags & 7, i, x);
    float t = g_model.forward(x);
    if (best < 0 || t < bt) {
      best = i;
      bt = t;
    }
  }
  if (g_cache.size() > 65536) g_cache.clear();
  g_cache[key] = best;
  return best;
}

// Tile class of one problem of the persistent batched kernel and its
// estimated time (padded multiply-adds; it orders the work queue
// longest-first).  A problem takes 128x128 tiles when both M and N exceed the
// family's threshold, else 64x64.  The threshold is 64 until
// numc_matmul_batched_tune measures a better one on a real batch: the cost
// model, trained on single calls, favours small tiles that keep one problem's
// GPU busy, which is not what a full batch needs.
static int g_batch_thresh[NUMC_FAMILY_COUNT] = {64, 64, 64, 64, 64, 64, 64};

// FP32, split-BF16 and int8 batches run the direct kernel until
// numc_matmul_batched_tune measures otherwise; the other families have only
// the packed one
static bool g_batch_direct[NUMC_FAMILY_COUNT] = {true, true, false, false, true, false, false};

bool numc_batch_direct(int family) {
  std::lock_guard<std::mutex> lk(g_mu);
  return family >= 0 && family < NUMC_FAMILY_COUNT && g_batch_direct[family];
}

void numc_batch_set_direct(int family, bool direct) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (family >= 0 && family < NUMC_FAMILY_COUNT)
    g_batch_direct[family] =
        direct && (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 || family == NUMC_FAMILY_I8);
}

int numc_batch_threshold(int family) {
  std::lock_guard<std::mutex> lk(g_mu);
  return family >= 0 && family < NUMC_FAMILY_COUNT ? g_batch_thresh[family] : 64;
}

void numc_batch_set_threshold(int family, int t) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (family >= 0 && family < NUMC_FAMILY_COUNT) g_batch_thresh[family] = t;
}

BatchChoice numc_batch_choice(int family, int M, int N, int K, int thresh) {
  (void)family;
  BatchChoice r;
  r.cls = (M > thresh && N > thresh) ? 0 : 1;
  const int t = r.cls == 0 ? 128 : 64;
  r.ms = (float)(std::ceil(M / (double)t) * std::ceil(N / (double)t) * t * t * (double)(K > 0 ? K : 1) * 1e-9);
  return r;
}
