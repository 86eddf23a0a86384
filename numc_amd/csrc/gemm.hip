// gemm.hip -- pack kernels, the packed-fragment MFMA GEMM, split-K reduce,
// the naive kernel and the 64-bit integer kernel, plus their launch tables.
//
// Replaces numc's packed GEMM (src/intrinsics/gemm_avx2.h, micro-kernels in
// src/intrinsics/gemm_ukernel_*_avx2.S) and its small-problem path
// (gemmsup_avx2.h) for every dtype; numc_matmul_naive's kernel
// (src/matmul/kernel.h:18-62) gets a direct GPU counterpart.
#include "gemm.h"
#include "gemm_kernels.h"
#include "internal.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

static inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }
static inline unsigned blocks_for(long long n) { return (unsigned)cdiv(n, 256); }

// ===========================================================================
// number formats (device)
// ===========================================================================

// the FP8 block exponent: maxabs * 2^ex lies in [128, 256) (0 for an
// all-zero block), clamped to +-126 so that the E8M0 byte 127 - ex is finite
__device__ inline int fp8_scale_exp(float maxabs) {
  if (!(maxabs > 0.0f)) return 0;
  int q;
  frexpf(maxabs, &q);
  int e = 8 - q;
  return e > 126 ? 126 : (e < -126 ? -126 : e);
}

// ===========================================================================
// small utility kernels
// ===========================================================================

__global__ void k_fill_u32(uint32_t *p, long long n, uint32_t v) {
  long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i < n) p[i] = v;
}

// out[i*rso + j*cso] = bias[j*bias_s], or 0 without a bias, for an M x N
// strided view (element size es): the product of an empty sum
__global__ void k_zero_strided(unsigned char *c, int M, int N, long long rso, long long cso, int es,
                               const unsigned char *bias, long long bias_s) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)M * N) return;
  long long i = t / N, j = t % N;
  unsigned char *p = c + (i * rso + j * cso) * es;
  const unsigned char *q = bias ? bias + j * bias_s * es : nullptr;
  for (int b = 0; b < es; b++) p[b] = q ? q[b] : 0;
}

// ===========================================================================
// pack kernels: one thread writes one 16-byte lane slot of a fragment in
// every plane.  Thread t -> (rt, kf, lane) = (t / 64 / nKF, t / 64 % nKF, t % 64),
// so the destination offset of plane 0 is t*16.
// ===========================================================================

template <class S> __device__ inline S ld(const void *p, long long i) { return NUMC_GLOBAL(const S, p)[i]; }

// a vector load from an address aligned only to its elements: global vector
// loads on gfx950 take any byte alignment (a chunk that straddles two cache
// lines costs one more line), and the compiler emits the same
// global_load_dwordx4 for this copy as for an aligned load
typedef unsigned int u32x4_any __attribute__((ext_vector_type(4), aligned(1)));
template <class V> __device__ __attribute__((always_inline)) inline V ldu(const void *p) {
  static_assert(sizeof(V) == 16, "16-byte chunks");
  const u32x4_any v = *NUMC_GLOBAL(const u32x4_any, p);
  V r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// the same load marked non-temporal (streamed past the caches): the float
// packs read each source byte once, and their sources then no longer push
// the packed planes the GEMM reads next out of the L2 and the MALL (FP8
// 8192^2 pack 0.135 -> 0.105 ms; split-BF16 0.177 -> 0.162 ms; the 16-bit
// float packs, 8192^2 bf16 0.099 -> 0.087 ms).  The int8 and integer digit
// packs keep cached loads: their tiles of short rows share lines (int8
// batch pack 0.25 -> 0.28 ms with streamed loads), and the int32 digit pack
// is one round of blocks, where a streamed load's longer trip shows (0.0148
// -> 0.0164 ms).
template <class V> __device__ __attribute__((always_inline)) inline V ldu_stream(const void *p) {
  static_assert(sizeof(V) == 16, "16-byte chunks");
  const u32x4_any v = __builtin_nontemporal_load(NUMC_GLOBAL(const u32x4_any, p));
  V r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// One 16-byte chunk of an edge tile (16 / sizeof(S) values from src, n_ok
// of them inside the matrix, zero past them).  A whole chunk is one vector
// load; a partial one is read as the 16-byte-aligned chunks that hold its
// first and its last valid byte (an aligned chunk never crosses a page, so
// both lie in pages the matrix occupies); a chunk with no valid byte loads
// nothing.  A thread issues the loads of all its chunks first (edge_issue),
// then shifts each into place and clears the bytes past its valid ones
// (edge_finish): element loads would each wait for the one before.
typedef unsigned int u32x4_al __attribute__((ext_vector_type(4)));
struct EdgeChunk {
  u32x4_al lo, hi;
  int s, nb; // byte offset of the chunk in lo:hi, valid bytes
};

template <class S>
__device__ __attribute__((always_inline)) inline EdgeChunk edge_issue(const S *src, long long n_ok) {
  constexpr int EPC = 16 / sizeof(S);
  const int nb = n_ok <= 0 ? 0 : n_ok >= EPC ? 16 : (int)n_ok * (int)sizeof(S);
  const unsigned long long p = (unsigned long long)src;
  EdgeChunk e;
  e.nb = nb;
  e.s = nb == 16 ? 0 : (int)(p & 15);
  e.lo = u32x4_al{0u, 0u, 0u, 0u};
  e.hi = u32x4_al{0u, 0u, 0u, 0u};
  if (nb == 16) {
    e.lo = *NUMC_GLOBAL(const u32x4_any, src);
  } else if (nb > 0) {
    const unsigned long long c0 = p - (unsigned long long)e.s;
    e.lo = *NUMC_GLOBAL(const u32x4_al, c0);
    if (e.s + nb > 16) e.hi = *NUMC_GLOBAL(const u32x4_al, c0 + 16);
  }
  return e;
}

__device__ __attribute__((always_inline)) inline uint4 edge_finish(const EdgeChunk &e) {
  typedef unsigned long long u64;
  // lo:hi as four 64-bit words; the chunk is bytes s .. s+15 of them, taken
  // with 64-bit funnel shifts of whole words (written out rather than as an
  // array indexed by s: the compiler turns such a select into a scratch
  // store and a load at a variable offset)
  const u64 L0 = ((u64)e.lo.y << 32) | e.lo.x, L1 = ((u64)e.lo.w << 32) | e.lo.z;
  const u64 H0 = ((u64)e.hi.y << 32) | e.hi.x, H1 = ((u64)e.hi.w << 32) | e.hi.z;
  const bool up = (e.s & 8) != 0;
  const u64 a = up ? L1 : L0, b = up ? H0 : L1, c = up ? H1 : H0;
  const int sh = (e.s & 7) * 8;
  u64 o0 = sh ? (a >> sh) | (b << (64 - sh)) : a;
  u64 o1 = sh ? (b >> sh) | (c << (64 - sh)) : b;
  // clear the bytes past the valid ones
  if (e.nb < 8) o0 &= (1ull << (8 * e.nb)) - 1ull;
  o1 = e.nb <= 8 ? 0ull : e.nb < 16 ? o1 & ((1ull << (8 * (e.nb - 8))) - 1ull) : o1;
  return make_uint4((uint32_t)o0, (uint32_t)(o0 >> 32), (uint32_t)o1, (uint32_t)(o1 >> 32));
}


// f32 -> f32 (T=32, 4 per lane, KF=8): unit t of a plane
__device__ inline void pack_unit_f32(const PackArgs &a, long long t) {
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long row = (fr / a.nKF) * 32 + (lane & 31);
  int k0 = kf * 8 + (lane >> 5) * 4;
  uint32_t v[4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    int k = k0 + e;
    v[e] = (row < a.R && k < a.K) ? __float_as_uint(ld<float>(a.src, row * a.rs + (long long)k * a.ks)) : 0u;
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(v[0], v[1], v[2], v[3]);
}

// f32 -> bf16 hi (plane 0) and lo (plane 1)  (T=32, 8 per lane, KF=16)
__device__ inline void pack_unit_split(const PackArgs &a, long long t) {
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long row = (fr / a.nKF) * 32 + (lane & 31);
  int k0 = kf * 16 + (lane >> 5) * 8;
  uint32_t hi[4], lo[4];
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    float x0 = 0.f, x1 = 0.f;
    if (row < a.R) {
      if (k0 + e < a.K) x0 = ld<float>(a.src, row * a.rs + (long long)(k0 + e) * a.ks);
      if (k0 + e + 1 < a.K) x1 = ld<float>(a.src, row * a.rs + (long long)(k0 + e + 1) * a.ks);
    }
    uint32_t h0 = f32_to_bf16_bits(x0), h1 = f32_to_bf16_bits(x1);
    uint32_t l0 = f32_to_bf16_bits(x0 - __uint_as_float(h0 << 16));
    uint32_t l1 = f32_to_bf16_bits(x1 - __uint_as_float(h1 << 16));
    hi[e / 2] = (h0 & 0xFFFFu) | (h1 << 16);
    lo[e / 2] = (l0 & 0xFFFFu) | (l1 << 16);
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
  *(uint4 *)(a.dst + a.plane_bytes + t * 16) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
}

// int8 -> one int8 plane (T=32, 16 per lane, KF=32); exact for int8 sources
__device__ inline void pack_unit_i8(const PackArgs &a, long long t) {
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long row = (fr / a.nKF) * 32 + (lane & 31);
  int k0 = kf * 32 + (lane >> 5) * 16;
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < 16; e++) {
    int k = k0 + e;
    uint32_t x = (row < a.R && k < a.K) ? (uint32_t)(uint8_t)ld<int8_t>(a.src, row * a.rs + (long long)k * a.ks) : 0u;
    w[e >> 2] |= x << (8 * (e & 3));
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(w[0], w[1], w[2], w[3]);
}

// strided batch: blockIdx.z selects the slice (a no-op for one problem)
__device__ __attribute__((always_inline)) inline void zslice(PackArgs &a) {
  a.src = (const unsigned char *)a.src + (long long)blockIdx.z * a.src_zs;
  a.dst += (long long)blockIdx.z * a.dst_zs;
}

__global__ void k_pack_f32(PackArgs a) {
  zslice(a);
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t < (long long)a.nRT * a.nKF * 64) pack_unit_f32(a, t);
}

__global__ void k_pack_split_bf16(PackArgs a) {
  zslice(a);
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t < (long long)a.nRT * a.nKF * 64) pack_unit_split(a, t);
}


// FP8 block scales.  Each row of each packed fragment -- 32 consecutive k of
// one row, 32-aligned -- gets its own power-of-two scale: the values are
// multiplied by 2^ex (ex = fp8_scale_exp of the block's max |x|, so that the
// block's largest magnitude lands in [128, 256)) and rounded to e4m3, and the
// GEMM's MFMA multiplies the block back by the E8M0 scale 127 - ex.  The two
// lane halves of a fragment row (lanes r and r + 32, 16 values each) combine
// their maxima with one cross-half exchange, so every caller runs all 64
// lanes of a wave together.  Scale bytes: [row tile][k-fragment pair][64],
// byte r + 32*(kf & 1) for row r of k-fragment kf -- the order in which the
// MFMA takes them (lane r scales fragment 2p, lane r + 32 fragment 2p + 1).
__device__ __attribute__((always_inline)) inline void fp8_block_unit(const PackArgs &a, long long rt, int kf, int lane,
                                                                     const float *x, unsigned char *dst16) {
  float m = 0.f;
#pragma unroll
  for (int e = 0; e < 16; e++) m = fmaxf(m, fabsf(x[e]));
  m = fmaxf(m, __shfl_xor(m, 32));
  const int ex = fp8_scale_exp(m);
  // v_cvt_pk_fp8_f32 rounds to nearest even into OCP e4m3.  Below 448 in
  // magnitude (a scaled block lies below 256) it matched a software RNE
  // encoder on every bf16-representable input, every e4m3 rounding tie with
  // its float neighbours and 4M random values.  (A non-finite input becomes a
  // NaN code, so the output row or column it reaches is NaN.)
  uint32_t w[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int lo = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[4 * q], ex), ldexpf(x[4 * q + 1], ex), 0, false);
    w[q] = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(x[4 * q + 2], ex), ldexpf(x[4 * q + 3], ex), lo, true);
  }
  *(uint4 *)dst16 = make_uint4(w[0], w[1], w[2], w[3]);
  if (lane < 32) a.sc[(rt * (a.nKF >> 1) + (kf >> 1)) * 64 + lane + 32 * (kf & 1)] = (unsigned char)(127 - ex);
}

// one 16-byte lane slot of an FP8 fragment, any strides (T=32, 16 per lane, KF=32)
__device__ inline void pack_unit_fp8(const PackArgs &a, long long t) {
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long rt = fr / a.nKF;
  long long row = rt * 32 + (lane & 31);
  int k0 = kf * 32 + (lane >> 5) * 16;
  float x[16];
#pragma unroll
  for (int e = 0; e < 16; e++) {
    int k = k0 + e;
    x[e] = (row < a.R && k < a.K) ? ld<float>(a.src, row * a.rs + (long long)k * a.ks) : 0.f;
  }
  fp8_block_unit(a, rt, kf, lane, x, a.dst + t * 16);
}

__global__ void k_pack_fp8(PackArgs a) {
  zslice(a);
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t < (long long)a.nRT * a.nKF * 64) pack_unit_fp8(a, t);
}

// ---- tiled packs for f32 rows with unit k-stride ------------------------
// A block stages a 32-row x 128-k tile through LDS: reads are whole 512-byte
// row segments (coalesced), writes are whole 1 KiB fragments.  MODE_PK: 0
// plain f32 (KF 8), 1 bf16 hi/lo (KF 16), 2 e4m3 with block scales (KF 32).
// LDS words one pack tile needs (every tiled pack below fits)
#define PACK_TILE_WORDS (128 * 36)

// tile `bid` of a k-contiguous operand: 32 rows x 128 k through LDS
template <int MODE_PK>
__device__ __attribute__((always_inline)) inline void pack_f32_tile_k(const PackArgs &a, long long bid, float *tile) {
  constexpr int TK = 128, PITCH = TK + 4;
  static_assert(32 * PITCH <= PACK_TILE_WORDS, "tile");
  const int nkt = (a.nKF * (MODE_PK == 0 ? 8 : MODE_PK == 1 ? 16 : 32) + TK - 1) / TK;
  const long long rt = bid / nkt;
  const int kt = (int)(bid % nkt);
  const int tid = threadIdx.x;
  const long long row0 = rt * 32;
  const int k0 = kt * TK;
  // 32 rows x 128 floats = 1024 float4, 4 per thread
  float4 v[4];
  const float *base = (const float *)a.src + row0 * a.rs + k0;
  if (row0 + 32 <= a.R && k0 + TK <= a.K) {
    // interior tile: every load in flight at once
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int idx = tid + it * 256;
      v[it] = ldu_stream<float4>(base + (idx >> 5) * a.rs + (idx & 31) * 4);
    }
  } else {
    EdgeChunk e[4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int idx = tid + it * 256, k = k0 + (idx & 31) * 4;
      const long long row = row0 + (idx >> 5);
      e[it] = edge_issue<float>((const float *)a.src + row * a.rs + k, row < a.R ? (long long)a.K - k : 0);
    }
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const uint4 q = edge_finish(e[it]);
      v[it] = make_float4(__uint_as_float(q.x), __uint_as_float(q.y), __uint_as_float(q.z), __uint_as_float(q.w));
    }
  }
#pragma unroll
  for (int it = 0; it < 4; it++) {
    const int idx = tid + it * 256;
    float *t = tile + (idx >> 5) * PITCH + (idx & 31) * 4;
    t[0] = v[it].x; t[1] = v[it].y; t[2] = v[it].z; t[3] = v[it].w;
  }
  __syncthreads();
  constexpr int KF = MODE_PK == 0 ? 8 : MODE_PK == 1 ? 16 : 32;
  constexpr int E = KF / 2;          // elements per lane
  constexpr int FRAGS = TK / KF;     // fragments in the tile
  constexpr int UNITS = FRAGS * 64;
  for (int u = tid; u < UNITS; u += 256) {
    const int f = u >> 6, lane = u & 63;
    const int kf = kt * FRAGS + f;
    if (kf >= a.nKF) continue;
    const int r = lane & 31;
    const float *t = tile + r * PITCH + f * KF + (lane >> 5) * E;
    const long long off = (rt * a.nKF + kf) * 1024 + lane * 16;
    if constexpr (MODE_PK == 0) {
      *(uint4 *)(a.dst + off) = make_uint4(__float_as_uint(t[0]), __float_as_uint(t[1]),
                                           __float_as_uint(t[2]), __float_as_uint(t[3]));
    } else if constexpr (MODE_PK == 1) {
      uint32_t hi[4], lo[4];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        float x0 = t[e], x1 = t[e + 1];
        uint32_t h0 = f32_to_bf16_bits(x0), h1 = f32_to_bf16_bits(x1);
        uint32_t l0 = f32_to_bf16_bits(x0 - __uint_as_float(h0 << 16));
        uint32_t l1 = f32_to_bf16_bits(x1 - __uint_as_float(h1 << 16));
        hi[e / 2] = (h0 & 0xFFFFu) | (h1 << 16);
        lo[e / 2] = (l0 & 0xFFFFu) | (l1 << 16);
      }
      *(uint4 *)(a.dst + off) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
      *(uint4 *)(a.dst + a.plane_bytes + off) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    } else {
      // (a wave's 64 units share f, so the lane halves meet in fp8_block_unit)
      fp8_block_unit(a, rt, kf, lane, t, a.dst + off);
    }
  }
}

template <int MODE_PK>
__global__ void __launch_bounds__(256) k_pack_f32_tiled(PackArgs a) {
  zslice(a);
  __shared__ float tile[PACK_TILE_WORDS];
  pack_f32_tile_k<MODE_PK>(a, blockIdx.x, tile);
}

// Same for sources whose rows are contiguous (unit row stride, e.g. B^T of a
// row-major B): a 128-row x 32-k tile, read as 512-byte runs along rows and
// transposed through LDS.
template <int MODE_PK>
__device__ __attribute__((always_inline)) inline void pack_f32_tile_rows(const PackArgs &a, long long bid, float *tile) {
  constexpr int TR = 128, TK = 32, PITCH = TK + 1;
  // row r of the tile sits at LDS row r + r / 32: the transposing writes (a
  // wave's lanes at rows 4l + j, one k) and the fragment reads (32
  // consecutive rows, one k) then both fall on 32 different banks
  static_assert((TR + TR / 32) * PITCH <= PACK_TILE_WORDS, "tile");
  constexpr int KF = MODE_PK == 0 ? 8 : MODE_PK == 1 ? 16 : 32;
  // consecutive blocks take neighbouring row blocks at one k range, so the
  // segments they read complete each other's source rows
  const int nrb = (a.nRT + 3) / 4;
  const long long rb = bid % nrb;   // block of 4 row fragments
  const int kt = (int)(bid / nrb);
  const int tid = threadIdx.x;
  const long long row0 = rb * TR;
  const int k0 = kt * TK;
  // 32 k x 128 rows = 1024 float4 along rows, 4 per thread
  float4 v[4];
  const float *base = (const float *)a.src + (long long)k0 * a.ks + row0;
  if (row0 + TR <= a.R && k0 + TK <= a.K) {
    // interior tile: every load in flight at once
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int idx = tid + it * 256;
      v[it] = ldu_stream<float4>(base + (long long)(idx >> 5) * a.ks + (idx & 31) * 4);
    }
  } else {
    EdgeChunk e[4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int idx = tid + it * 256, k = k0 + (idx >> 5), r4 = (idx & 31) * 4;
      e[it] = edge_issue<float>((const float *)a.src + (long long)k * a.ks + row0 + r4,
                                k < a.K ? (long long)a.R - (row0 + r4) : 0);
    }
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const uint4 q = edge_finish(e[it]);
      v[it] = make_float4(__uint_as_float(q.x), __uint_as_float(q.y), __uint_as_float(q.z), __uint_as_float(q.w));
    }
  }
#pragma unroll
  for (int it = 0; it < 4; it++) {
    const int idx = tid + it * 256;
    const int kk = idx >> 5, r4 = (idx & 31) * 4, lr = r4 + (r4 >> 5);
    tile[(lr + 0) * PITCH + kk] = v[it].x;
    tile[(lr + 1) * PITCH + kk] = v[it].y;
    tile[(lr + 2) * PITCH + kk] = v[it].z;
    tile[(lr + 3) * PITCH + kk] = v[it].w;
  }
  __syncthreads();
  constexpr int E = KF / 2;
  constexpr int FK = TK / KF;            // fragments along k in the tile
  constexpr int UNITS = 4 * FK * 64;
  for (int u = tid; u < UNITS; u += 256) {
    const int f = u >> 6, lane = u & 63;
    const int rl = f / FK, fk = f % FK;   // row fragment in tile, k fragment
    const long long rt = rb * 4 + rl;
    const int kf = kt * FK + fk;
    if (rt >= a.nRT || kf >= a.nKF) continue;
    const int r = rl * 32 + (lane & 31);
    const float *t = tile + (r + rl) * PITCH + fk * KF + (lane >> 5) * E;
    const long long off = (rt * a.nKF + kf) * 1024 + lane * 16;
    if constexpr (MODE_PK == 0) {
      *(uint4 *)(a.dst + off) = make_uint4(__float_as_uint(t[0]), __float_as_uint(t[1]),
                                           __float_as_uint(t[2]), __float_as_uint(t[3]));
    } else if constexpr (MODE_PK == 1) {
      uint32_t hi[4], lo[4];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        float x0 = t[e], x1 = t[e + 1];
        uint32_t h0 = f32_to_bf16_bits(x0), h1 = f32_to_bf16_bits(x1);
        uint32_t l0 = f32_to_bf16_bits(x0 - __uint_as_float(h0 << 16));
        uint32_t l1 = f32_to_bf16_bits(x1 - __uint_as_float(h1 << 16));
        hi[e / 2] = (h0 & 0xFFFFu) | (h1 << 16);
        lo[e / 2] = (l0 & 0xFFFFu) | (l1 << 16);
      }
      *(uint4 *)(a.dst + off) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
      *(uint4 *)(a.dst + a.plane_bytes + off) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    } else {
      fp8_block_unit(a, rt, kf, lane, t, a.dst + off);
    }
  }
}

template <int MODE_PK>
__global__ void __launch_bounds__(256) k_pack_f32_tiled_rows(PackArgs a) {
  zslice(a);
  __shared__ float tile[PACK_TILE_WORDS];
  pack_f32_tile_rows<MODE_PK>(a, blockIdx.x, tile);
}

// f64 -> f64 (T=16, 2 per lane, KF=8)
__global__ void k_pack_f64(PackArgs a) {
  zslice(a);
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)a.nRT * a.nKF * 64) return;
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long row = (fr / a.nKF) * 16 + (lane & 15);
  int k0 = kf * 8 + (lane >> 4) * 2;
  double v0 = 0.0, v1 = 0.0;
  if (row < a.R) {
    if (k0 < a.K) v0 = ld<double>(a.src, row * a.rs + (long long)k0 * a.ks);
    if (k0 + 1 < a.K) v1 = ld<double>(a.src, row * a.rs + (long long)(k0 + 1) * a.ks);
  }
  unsigned long long b0 = __double_as_longlong(v0), b1 = __double_as_longlong(v1);
  *(uint4 *)(a.dst + t * 16) = make_uint4((uint32_t)b0, (uint32_t)(b0 >> 32), (uint32_t)b1, (uint32_t)(b1 >> 32));
}

// integer -> P signed base-256 digit planes (T=32, 16 per lane, KF=32).
// x = sum_p d_p 256^p (mod 256^P) with d_p in [-128, 127]; the operand's
// mask gets a bit per plane that holds a nonzero digit.
// A wave ORs its nonzero-plane mask into its block's slot of the mask (one
// of NUMC_FLAG_SLOTS words on separate lines: same-address atomics serialise
// in L2).  The block reads its slot when it starts (flags_peek, the load in
// flight with the data loads) and a wave issues the atomic at its end only
// when it would set a new bit (flags_or_seen), so once the bits are set the
// remaining waves only read; a stale read only costs a redundant atomic, and
// the atomic returns nothing, so no wave waits on the mask at its end (with
// the read and the wait at the end, the int32 2048^2 digit pack took 0.020
// ms instead of 0.015).  A null mask reads as 0.
__device__ inline int flags_peek(const int *flags) {
  return flags ? __hip_atomic_load(flags + (blockIdx.x % NUMC_FLAG_SLOTS) * NUMC_FLAG_STRIDE, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT)
               : 0;
}
__device__ inline void flags_or_seen(int *flags, int mask, int seen) {
  if (mask & ~seen) atomicOr(flags + (blockIdx.x % NUMC_FLAG_SLOTS) * NUMC_FLAG_STRIDE, mask);
}

// The P signed base-256 digits of x (x = sum_p d_p 256^p mod 2^32, each d_p
// in [-128, 127]) into byte e of the lanes' plane words w[p].  32-bit
// arithmetic is enough: every product sum is taken mod 2^32, and carries
// past 2^32 drop out of the low P <= 4 digits.
template <int P>
__device__ __attribute__((always_inline)) inline void digit_bytes(uint32_t x, uint32_t (&w)[P][4], int e) {
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int d = (int)(int8_t)(x & 0xFFu);
    x = (x - (uint32_t)d) >> 8;
    w[p][e >> 2] |= (uint32_t)(uint8_t)d << (8 * (e & 3));
  }
}

// one lane slot t of every plane, any strides (the whole wave calls it: the
// plane mask takes a ballot)
template <class S, int P>
__device__ __attribute__((always_inline)) inline void pack_digits_unit(const PackArgs &a, long long t) {
  const int seen = flags_peek(a.flags);
  bool live = t < (long long)a.nRT * a.nKF * 64;
  uint32_t w[P][4];
#pragma unroll
  for (int p = 0; p < P; p++) w[p][0] = w[p][1] = w[p][2] = w[p][3] = 0;
  if (live) {
    int lane = (int)(t & 63);
    long long fr = t >> 6;
    int kf = (int)(fr % a.nKF);
    long long row = (fr / a.nKF) * 32 + (lane & 31);
    int k0 = kf * 32 + (lane >> 5) * 16;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      int k = k0 + e;
      const S x = (row < a.R && k < a.K) ? ld<S>(a.src, row * a.rs + (long long)k * a.ks) : (S)0;
      digit_bytes<P>((uint32_t)(long long)x, w, e);
    }
#pragma unroll
    for (int p = 0; p < P; p++)
      *(uint4 *)(a.dst + p * a.plane_bytes + t * 16) = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
  }
  int mask = 0;
#pragma unroll
  for (int p = 0; p < P; p++)
    if (__ballot((w[p][0] | w[p][1] | w[p][2] | w[p][3]) != 0)) mask |= 1 << p;
  if ((threadIdx.x & 63) == 0 && mask && a.flags) flags_or_seen(a.flags, mask, seen); // no masks: null
}

template <class S, int P>
__global__ void k_pack_digits(PackArgs a) {
  zslice(a);
  pack_digits_unit<S, P>(a, blockIdx.x * 256LL + threadIdx.x);
}

// ===========================================================================
// the packed-fragment MFMA GEMM
//
// Block tile BM x BN, WGM x WGN waves, each wave WM x WN = (TM x TN) MFMA
// tiles of T x T.  Per k-chunk (KC) the block's fragments of A and B (all
// planes) are loaded with one coalesced 16-byte load per lane per fragment,
// written linearly to LDS (double buffered, one barrier per chunk), and each
// wave reads its fragments back with conflict-free 16-byte LDS reads.
// MODE_SPLIT3 holds bf16 hi/lo planes of both operands and issues
// lo*hi + hi*lo + hi*hi per fragment pair; MODE_SINGLE walks a list of plane
// products concatenated along K (integer digit passes).
// ===========================================================================

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// s_waitcnt vmcnt(n) that leaves the LDS/export counters alone (gfx9 encoding:
// vmcnt[3:0] | expcnt[6:4] | lgkmcnt[11:8] | vmcnt[5:4] at [15:14])
#define WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))

// Workgroup barrier of the copy pipeline.  __syncthreads() (and even an
// LDS-only fence, since copies into LDS count as vector memory operations)
// waits for every outstanding global->LDS copy (vmcnt(0)) and would drain
// the pipeline at each chunk.  Here each wave has already waited for its own
// copies of the chunk about to be read (WAIT_VMCNT); this waits for its LDS
// reads (lgkmcnt(0): nobody overwrites a stage still being read), and the
// empty asm statements keep the compiler from moving memory accesses across
// the s_barrier.
__device__ __attribute__((always_inline)) inline void lds_barrier() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4)); // lgkmcnt(0), vmcnt/expcnt untouched
  __asm__ __volatile__("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ __volatile__("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Scheduling hint for one k step: spread its NDMA global->LDS copies and NRD
// fragment reads evenly over its NMF MFMAs (the compiler otherwise bunches
// the reads after the last MFMA, and the next k step waits for them).
// Masks: 0x008 MFMA, 0x020 vector-memory read, 0x100 LDS read.
template <int NDMA, int NRD, int NMF> struct Interleave {
  static constexpr int NMEM = NDMA + NRD;
  static constexpr int GAP = NMEM > 0 ? (NMF / (NMEM + 1) > 0 ? NMF / (NMEM + 1) : 1) : NMF;
  template <int I> __device__ static __attribute__((always_inline)) inline void step() {
    if constexpr (I < NMEM) {
      if constexpr (I < NDMA)
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      else
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if constexpr ((I + 1) * GAP <= NMF) __builtin_amdgcn_sched_group_barrier(0x008, GAP, 0);
      step<I + 1>();
    } else if constexpr (NMF > NMEM * GAP) {
      __builtin_amdgcn_sched_group_barrier(0x008, NMF - NMEM * GAP, 0);
    }
  }
  __device__ static __attribute__((always_inline)) inline void apply() { step<0>(); }
};

template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int NS>
struct TileShape {
  static constexpr int T = MF::T, KF = MF::KF;
  static constexpr int KFC = KC / KF;
  static constexpr int P = (MODE == MODE_SPLIT3) ? 2 : 1;
  static constexpr int FRC = P * (BM / T) * KFC + P * (BN / T) * KFC;
  // block-scaled operands: one scale byte per row per 32 k of each operand,
  // copied 256 bytes per wave instruction, every wave the same number of
  // copies (the area is rounded up to whole rounds of copies)
  static constexpr int NW = WGM * WGN;
  static constexpr int SCR = IsScaled<MF>::value ? (BM / T + BN / T) * KC : 0;
  static constexpr int SCB = (SCR / 256 + NW - 1) / NW * NW * 256;
  static constexpr int SV = FRC * 64 + SCB / 16;  // u32x4 elements per stage
  static constexpr int LDS_VEC = NS * SV;         // u32x4 elements of LDS
};

// One BM x BN output tile of one problem: k-chunks [c0, c1) of slice `slice`
// (the range is computed here), accumulated in registers and stored.
template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int EPI, int NS, int VAR>
__device__ __attribute__((always_inline)) inline void gemm_tile(const GemmArgs &g, const int bm, const int bn,
                                                                const int slice, u32x4 *lds) {
  constexpr int T = MF::T, KF = MF::KF, FR = MF::FR;
  constexpr int NW = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / T, TN = WN / T;
  constexpr int KFC = KC / KF;
  constexpr int P = (MODE == MODE_SPLIT3) ? 2 : 1;
  constexpr int RA = BM / T, RB = BN / T;
  constexpr int A_FR = P * RA * KFC, B_FR = P * RB * KFC, FRC = A_FR + B_FR;
  // Block-scaled families (FP8): each stage also holds the chunk's scale
  // bytes, A's [row tile][KC / 64 pairs][64] then B's likewise (SCA + SCB
  // bytes after the fragments).  A wave copies them 4 bytes per lane (global
  // loads into LDS place lane l's dword at +4l), 256 bytes per instruction,
  // LS instructions per wave; when the SCT / 256 copies do not split evenly
  // over the waves, the last round's spare copies repeat scale bytes into
  // the padding after SCT, where nobody reads them.
  constexpr bool SC = IsScaled<MF>::value;
  constexpr int SCA = SC ? RA * KC : 0, SCB = SC ? RB * KC : 0, SCT = SCA + SCB;
  static_assert(SCT % 256 == 0, "scale bytes per chunk must be whole 256-byte copies");
  constexpr int LS = (SCT / 256 + NW - 1) / NW;
  constexpr int SV = TileShape<MF, BM, BN, KC, WGM, WGN, MODE, NS>::SV; // u32x4 per stage
  static_assert(SV == FRC * 64 + LS * NW * 256 / 16, "stage layout");
  constexpr int LA = A_FR / NW, LB = B_FR / NW, LPW = LA + LB + LS;
  static_assert(A_FR % NW == 0 && B_FR % NW == 0, "fragments must split evenly over waves");
  static_assert(KFC % FR == 0 && TM >= 1 && TN >= 1, "bad tile");
  static_assert(NS >= 2 && LPW * (NS - 1) <= 63, "too many loads in flight for vmcnt");
  using Acc = typename MF::Acc;
  using Raw = typename MF::Raw;

  const int tid = threadIdx.x, lane = tid & 63;
  // the wave index is uniform: keep everything derived from it in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave / WGN, wn = wave % WGN;

  // plane products: product q uses plane (pa0 + q) of A and (pb0 - q) of B.
  // Integer digit passes: a launch per pass adds (sum_s << 8s) into C, or,
  // with g.passes > 1, the tile runs every pass itself, from the highest
  // down, shifting its accumulators 8 bits left between passes (Horner), and
  // its one store writes sum_s (sum_s << 8s) mod 2^32.
  constexpr bool FUSE = std::is_same<MF, MfI8>::value;
  const int npass = FUSE && g.passes > 1 ? g.passes : 1;
  int nprod = g.nprod, pa0 = g.pa[0], pb0 = g.pb[0];
  int ma = 0, mb = 0;
  if (g.plane_flags) {
    plane_masks(g.plane_flags, &ma, &mb);
    if (npass == 1) {
      nprod = int_products(g.shift_pass, ma, mb, &pa0, &pb0);
      if (nprod == 0 && g.shift_pass > 0) return; // this digit pass adds nothing
    }
  }

  // wave-uniform byte offsets of the fragments this wave brings into LDS
  long long uoffA[LA], uoffB[LB];
#pragma unroll
  for (int i = 0; i < LA; i++) {
    int f = wave + i * NW;
    int p = f / (RA * KFC), rt = (f / KFC) % RA, kf = f % KFC;
    uoffA[i] = p * g.a_plane + ((long long)(bm * RA + rt) * g.nKF + kf) * NUMC_FRAG_BYTES;
  }
#pragma unroll
  for (int i = 0; i < LB; i++) {
    int f = wave + i * NW;
    int p = f / (RB * KFC), rt = (f / KFC) % RB, kf = f % KFC;
    uoffB[i] = p * g.b_plane + ((long long)(bn * RB + rt) * g.nKF + kf) * NUMC_FRAG_BYTES;
  }
  const int loff = lane * 16;
  // A read in place: lane l's 16 bytes of a fragment are row (l & 31) of its
  // row tile, bytes 32*kf + 16*(l >> 5) of the k-chunk (the packed layout's
  // order), so the copies land in LDS exactly as the packed fragments do.
  // Rows past M repeat row M-1: their products reach no stored output.
  // B in place (VAR & 4) likewise, with the rows of B^T.
  constexpr bool ADIR = (VAR & 2) != 0, BDIR = (VAR & 4) != 0;
  static_assert(!(ADIR || BDIR) || P == 1, "operands are read in place only with one plane");
  const unsigned char *adir[ADIR ? LA : 1];
  const unsigned char *bdir[BDIR ? LB : 1];
  if constexpr (ADIR) {
#pragma unroll
    for (int i = 0; i < LA; i++) {
      const int f = wave + i * NW, rt = (f / KFC) % RA, kf = f % KFC;
      const long long row = min((long long)(bm * RA + rt) * 32 + (lane & 31), (long long)g.M - 1);
      adir[i] = g.Araw + row * g.a_rs + kf * 32 + (lane >> 5) * 16;
    }
  }
  if constexpr (BDIR) {
#pragma unroll
    for (int i = 0; i < LB; i++) {
      const int f = wave + i * NW, rt = (f / KFC) % RB, kf = f % KFC;
      const long long row = min((long long)(bn * RB + rt) * 32 + (lane & 31), (long long)g.N - 1);
      bdir[i] = g.Braw + row * g.b_rs + kf * 32 + (lane >> 5) * 16;
    }
  }

  // scale copies: lane byte o of copy s of the chunk's scale area reads row
  // tile o / KC of A (or, past SCA, of B), bytes (o % KC) of the chunk's KC
  // (the chunk's pairs are contiguous in each row tile's scales); spare
  // copies (o >= SCT) read the bytes of o % SCT again
  const unsigned char *sdir[SC ? LS : 1];
  if constexpr (SC) {
    const long long pairs = g.nKF >> 1;
#pragma unroll
    for (int i = 0; i < LS; i++) {
      const int o = ((wave + i * NW) * 256 + lane * 4) % SCT;
      if (o < SCA)
        sdir[i] = g.a_sc + ((long long)(bm * RA + o / KC) * pairs) * 64 + o % KC;
      else
        sdir[i] = g.b_sc + ((long long)(bn * RB + (o - SCA) / KC) * pairs) * 64 + (o - SCA) % KC;
    }
  }

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int v = 0; v < MF::NACC; v++) acc[i][j][v] = 0;

  // chunk c = product q * nKC + k-chunk kc: one 1 KiB global->LDS copy per
  // fragment, lane l's 16 bytes landing at slot + 16*l
  auto issue = [&](int c, int s) __attribute__((always_inline)) {
    const int q = c / g.nKC, kc = c - q * g.nKC;
    const long long ko = (long long)kc * KFC * NUMC_FRAG_BYTES;
    const unsigned char *pA = g.A + ko + (long long)(pa0 + q) * g.a_plane + loff;
    const unsigned char *pB = g.B + ko + (long long)(pb0 - q) * g.b_plane + loff;
    u32x4 *base = lds + s * SV;
    const long long kb = (long long)kc * KFC * 32; // in place: 32 bytes of each row per k-fragment
    if constexpr (ADIR) {
#pragma unroll
      for (int i = 0; i < LA; i++)
        __builtin_amdgcn_global_load_lds((global_void_t *)(adir[i] + kb),
                                         (lds_void_t *)(base + (wave + i * NW) * 64), 16, 0, 0);
    } else {
#pragma unroll
      for (int i = 0; i < LA; i++)
        __builtin_amdgcn_global_load_lds((global_void_t *)(pA + uoffA[i]),
                                         (lds_void_t *)(base + (wave + i * NW) * 64), 16, 0, 0);
    }
    if constexpr (BDIR) {
#pragma unroll
      for (int i = 0; i < LB; i++)
        __builtin_amdgcn_global_load_lds((global_void_t *)(bdir[i] + kb),
                                         (lds_void_t *)(base + (A_FR + wave + i * NW) * 64), 16, 0, 0);
    } else {
#pragma unroll
      for (int i = 0; i < LB; i++)
        __builtin_amdgcn_global_load_lds((global_void_t *)(pB + uoffB[i]),
                                         (lds_void_t *)(base + (A_FR + wave + i * NW) * 64), 16, 0, 0);
    }
    if constexpr (SC) {
      unsigned char *sbase = (unsigned char *)(base + FRC * 64);
#pragma unroll
      for (int i = 0; i < LS; i++)
        __builtin_amdgcn_global_load_lds((global_void_t *)(sdir[i] + (long long)kc * KC),
                                         (lds_void_t *)(sbase + (wave + i * NW) * 256), 4, 0, 0);
    }
  };
  // fragments of one k step, double buffered in registers (with their
  // scale bytes: lane l's byte of pair kf / 2 in each row tile)
  u32x4 af[2][P][TM][FR], bfr[2][P][TN][FR];
  int sca[2][SC ? TM : 1], scb[2][SC ? TN : 1];
  auto fload = [&](int s, int kf, int buf) __attribute__((always_inline)) {
    const u32x4 *base = lds + s * SV;
    if constexpr (SC) {
      const unsigned char *sl = (const unsigned char *)(base + FRC * 64) + (kf >> 1) * 64 + lane;
#pragma unroll
      for (int i = 0; i < TM; i++) sca[buf][i] = sl[(wm * TM + i) * KC];
#pragma unroll
      for (int j = 0; j < TN; j++) scb[buf][j] = sl[SCA + (wn * TN + j) * KC];
    }
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int r = 0; r < FR; r++)
          af[buf][p][i][r] = base[((p * RA + wm * TM + i) * KFC + kf + r) * 64 + lane];
#pragma unroll
    for (int p = 0; p < P; p++)
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int r = 0; r < FR; r++)
          bfr[buf][p][j][r] = base[(A_FR + (p * RB + wn * TN + j) * KFC + kf + r) * 64 + lane];
  };
  auto mfmas = [&](int buf) __attribute__((always_inline)) {
    if constexpr (MODE == MODE_SPLIT3 && (VAR & 1)) {
      // product-major: consecutive MFMAs go to different accumulators; each
      // accumulator still adds lo*hi, hi*lo, hi*hi in that order
#pragma unroll
      for (int q = 0; q < 3; q++)
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
            MF::mma(acc[i][j], af[buf][q == 0 ? 1 : 0][i], bfr[buf][q == 1 ? 1 : 0][j]);
    } else if constexpr (MODE == MODE_SPLIT3) {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
          MF::mma(acc[i][j], af[buf][1][i], bfr[buf][0][j]);
          MF::mma(acc[i][j], af[buf][0][i], bfr[buf][1][j]);
          MF::mma(acc[i][j], af[buf][0][i], bfr[buf][0][j]);
        }
    } else if constexpr (SC) {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) MF::mma(acc[i][j], af[buf][0][i], bfr[buf][0][j], sca[buf][i], scb[buf][j]);
    } else {
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) MF::mma(acc[i][j], af[buf][0][i], bfr[buf][0][j]);
    }
  };

  // The fragment reads run one k step ahead of the MFMAs, across chunk
  // boundaries too.  During the last k step of chunk c the wave waits for
  // its copies of chunk c+1 and meets the others at the barrier; after it,
  // chunk c+1 is visible to all and chunk c's stage is read by nobody (its
  // last fragments are already in registers), so that stage is refilled with
  // chunk c+NS and chunk c+1's first fragments are read while chunk c's last
  // MFMAs run.  Copies past the last chunk are clamped to it (into stages
  // nobody reads again), so NS-1 chunks are always in flight behind the one
  // being waited for, every wait takes a constant count and the steady-state
  // loop has no branches.
  constexpr int NKS = KFC / FR;
  constexpr int NRD = P * (TM + TN) * FR + (SC ? TM + TN : 0);            // LDS reads per k step
  constexpr int NMF = TM * TN * (MODE == MODE_SPLIT3 ? 3 : 1) * MF::NMMA; // MFMAs per k step
  for (int pass = npass - 1; pass >= 0; pass--) {
    if constexpr (FUSE) {
      if (npass > 1) {
        if (pass < npass - 1) {
#pragma unroll
          for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++)
#pragma unroll
              for (int v = 0; v < MF::NACC; v++) acc[i][j][v] = (int)((uint32_t)acc[i][j][v] << 8);
          __syncthreads(); // every wave is done reading the last pass's stages
        }
        if (g.plane_flags) nprod = int_products(pass, ma, mb, &pa0, &pb0);
      }
    }
    const int V = nprod * g.nKC;
    const int c0 = (int)((long long)slice * V / g.ksplit);
    const int c1 = (int)((long long)(slice + 1) * V / g.ksplit);
    if (c0 < c1) {
      const int clast = c1 - 1;
      int stage = 0;
      auto body = [&](auto b0, int c, auto is_last) __attribute__((always_inline)) {
        constexpr int B0 = decltype(b0)::value;
        const int next = stage == NS - 1 ? 0 : stage + 1;
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
          const int buf = (B0 + ks) & 1;
          if (ks + 1 < NKS) {
            fload(stage, (ks + 1) * FR, buf ^ 1);
            mfmas(buf);
            Interleave<0, NRD, NMF>::apply();
          } else if constexpr (!decltype(is_last)::value) {
            __builtin_amdgcn_sched_barrier(0);
            WAIT_VMCNT(LPW * (NS - 2));
            lds_barrier();
            issue(min(c + NS, clast), stage);
            fload(next, 0, buf ^ 1);
            mfmas(buf);
            Interleave<LPW, NRD, NMF>::apply();
          } else {
            mfmas(buf);
          }
        }
        stage = next;
      };
      using I0 = std::integral_constant<int, 0>;
      using I1 = std::integral_constant<int, 1>;
      using More = std::false_type;
      using Last = std::true_type;
#pragma unroll
      for (int u = 0; u < NS; u++) issue(min(c0 + u, clast), u);
      WAIT_VMCNT(LPW * (NS - 1));
      lds_barrier();
      fload(0, 0, 0);
      int c = c0;
      if constexpr (NKS % 2 == 0) {
        for (; c < clast; c++) body(I0(), c, More());
        body(I0(), c, Last());
      } else {
        // the register buffer alternates per chunk: two chunks per iteration
        for (; c + 1 < clast; c += 2) {
          body(I0(), c, More());
          body(I1(), c + 1, More());
        }
        if (c < clast) {
          body(I0(), c, More());
          body(I1(), c + 1, Last());
        } else {
          body(I0(), c, Last());
        }
      }
      // the clamped copies past the last chunk may still be in flight: let
      // them land before the stages are reused or the workgroup ends
      WAIT_VMCNT(0);
    }
  } // digit passes

  // epilogue (split-K slices store raw partial sums; their reduction stores)
  const int row0 = bm * BM + wm * WM, col0 = bn * BN + wn * WN;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int col = col0 + j * T + MF::col(lane);
      const auto bias = g.ksplit == 1 ? epi_bias<EPI>(g, col) : (typename EpiBias<EPI>::T)0; // per lane
#pragma unroll
      for (int v = 0; v < MF::NACC; v++) {
        const int row = row0 + i * T + MF::row(v, lane);
        if (row < g.M && col < g.N) {
          if (g.ksplit > 1)
            ((Raw *)g.partial)[((long long)slice * g.M + row) * g.N + col] = (Raw)acc[i][j][v];
          else
            epi_store<EPI, Raw>(g, row, col, (Raw)acc[i][j][v], bias);
        }
      }
    }
}

template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int EPI, int NS, int VAR>
__global__ void __launch_bounds__(WGM *WGN * 64, 1) k_gemm(const GemmArgs gin) {
  GemmArgs g = gin; // local copy: the tile code captures it by reference
  g.A += (long long)blockIdx.z * g.a_zs; // strided batch: slice blockIdx.z
  g.B += (long long)blockIdx.z * g.b_zs;
  g.Araw += (long long)blockIdx.z * g.a_zs; // (an operand read in place has its own slice stride in a_zs / b_zs)
  g.Braw += (long long)blockIdx.z * g.b_zs;
  g.C = (unsigned char *)g.C + (long long)blockIdx.z * g.c_zs;
  // NS stages of fragments, filled straight from HBM by global_load_lds
  __shared__ u32x4 lds[TileShape<MF, BM, BN, KC, WGM, WGN, MODE, NS>::LDS_VEC];
  // block -> tile: blocks with equal pid % 8 share an XCD, so give each XCD
  // a contiguous range of tiles, then walk tiles in groups of 8 block rows
  int pid = blockIdx.x;
  const int nblk = g.nbm * g.nbn;
  if ((nblk & 7) == 0) pid = (pid & 7) * (nblk >> 3) + (pid >> 3);
  const int per_group = 8 * g.nbn;
  const int first = (pid / per_group) * 8;
  const int gsz = min(g.nbm - first, 8);
  const int bm = first + (pid % per_group) % gsz;
  const int bn = (pid % per_group) / gsz;
  gemm_tile<MF, BM, BN, KC, WGM, WGN, MODE, EPI, NS, VAR>(g, bm, bn, blockIdx.y, lds);
}

// split-K: sum the slices in slice order, then the final epilogue
template <class Raw, int EPI>
__global__ void k_splitk_reduce(const GemmArgs g) {
  if (g.plane_flags && g.shift_pass > 0) {
    int pa0, pb0;
    int ma, mb;
    plane_masks(g.plane_flags, &ma, &mb);
    if (int_products(g.shift_pass, ma, mb, &pa0, &pb0) == 0) return;
  }
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)g.M * g.N) return;
  const Raw *p = (const Raw *)g.partial;
  Raw s = p[t];
  for (int z = 1; z < g.ksplit; z++) s += p[(long long)z * g.M * g.N + t];
  const int col = (int)(t % g.N);
  epi_store<EPI, Raw>(g, (int)(t / g.N), col, s, epi_bias<EPI>(g, col));
}

// ===========================================================================
// numc_matmul_naive counterpart: one thread per output, contiguous operands,
// k in order.  Floats use fused multiply-add (what numc's -O3 build of
// `out += a*b` contracts to); integers keep the exact sum mod 2^64 and store
// its low bits.
// ===========================================================================

template <class T, bool FLOAT>
__global__ void k_naive(const T *a, const T *b, T *c, int M, int K, int N) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)M * N) return;
  long long i = t / N, j = t % N;
  if constexpr (FLOAT) {
    T acc = 0;
    for (int k = 0; k < K; k++) acc = fma(a[i * K + k], b[(long long)k * N + j], acc);
    c[t] = acc;
  } else {
    unsigned long long acc = 0;
    for (int k = 0; k < K; k++)
      acc += (unsigned long long)(long long)a[i * K + k] * (unsigned long long)(long long)b[(long long)k * N + j];
    c[t] = (T)acc;
  }
}

// bfloat16 / float16 (no numc counterpart): binary32 fused multiply-adds of
// the widened values in k order, one rounding to nearest even at the store.
template <bool BF>
__global__ void k_naive_h16(const uint16_t *a, const uint16_t *b, uint16_t *c, int M, int K, int N) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)M * N) return;
  long long i = t / N, j = t % N;
  float acc = 0.f;
  for (int k = 0; k < K; k++) {
    const uint32_t x = a[i * K + k], y = b[(long long)k * N + j];
    acc = fmaf(BF ? bf16_bits_to_f32(x) : f16_bits_to_f32(x), BF ? bf16_bits_to_f32(y) : f16_bits_to_f32(y), acc);
  }
  c[t] = (uint16_t)(BF ? f32_to_bf16_bits(acc) : f32_to_f16_bits(acc));
}

// 64-bit integers: no MFMA; 16x16 output tile per block, k staged in LDS.
template <class T>
__global__ void k_gemm_i64(const T *a, const T *b, T *c, int M, int K, int N, long long rsa, long long csa,
                           long long rsb, long long csb, long long rso, long long cso, const T *bias,
                           long long bias_s) {
  __shared__ unsigned long long sa[16][17], sb[16][17];
  int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int row = blockIdx.y * 16 + ty, col = blockIdx.x * 16 + tx;
  unsigned long long acc = 0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    int ka = k0 + tx, kb = k0 + ty;
    sa[ty][tx] = (row < M && ka < K) ? (unsigned long long)a[row * rsa + ka * csa] : 0ull;
    sb[ty][tx] = (col < N && kb < K) ? (unsigned long long)b[kb * rsb + col * csb] : 0ull;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) acc += sa[ty][k] * sb[k][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    if (bias) acc += (unsigned long long)bias[col * bias_s];
    c[row * rso + col * cso] = (T)acc;
  }
}

// ===========================================================================
// persistent batched GEMM: every (problem, tile) work item of many uneven
// problems in one launch.  Blocks take items from an atomic counter; each item
// names its problem and its tile class (two shapes compiled in), so small
// problems get small tiles and the queue is ordered longest-first.
// ===========================================================================

template <class MF, int MODE, int EPI, int BM1, int BN1, int KC1, int NS1, int BM2, int BN2, int KC2, int NS2,
          int OCC>
__global__ void __launch_bounds__(256, OCC) k_gemm_batched(const BatchProblem *probs, const int4 *items, int nitems,
                                                          int *counter) {
  using S1 = TileShape<MF, BM1, BN1, KC1, 2, 2, MODE, NS1>;
  using S2 = TileShape<MF, BM2, BN2, KC2, 2, 2, MODE, NS2>;
  constexpr int LDSV = S1::LDS_VEC > S2::LDS_VEC ? S1::LDS_VEC : S2::LDS_VEC;
  static_assert(OCC * LDSV * 16 <= 160 * 1024, "OCC blocks must fit one CU's LDS");
  __shared__ u32x4 lds[LDSV];
  __shared__ int s_item;
  // per-XCD work queues (gemm.h): the block takes from its own XCD's queue
  // (block b runs on XCD b % 8), then from the others in turn; a queue found
  // empty stays empty, so after finding all of them empty the block is done
  // Thread 0 takes the next item's index while the current tile runs, so the
  // atomic's round trip is off the path from one tile to the next.
  (void)nitems;
  const int *qs = counter + NUMC_BATCH_QSTART;
  int q = blockIdx.x % NUMC_BATCH_QUEUES, left = NUMC_BATCH_QUEUES;
  int q0 = qs[q], qn = qs[q + 1] - q0;
  int nxt = 0;
  if (threadIdx.x == 0) nxt = atomicAdd(counter + q * NUMC_BATCH_COUNTER_STRIDE, 1);
  for (;;) {
    if (threadIdx.x == 0) s_item = nxt;
    __syncthreads(); // also: the previous tile's LDS reads are over
    const int it = s_item;
    __syncthreads(); // everyone read s_item before thread 0 overwrites it
    if (it >= qn) {
      if (--left == 0) return;
      q = (q + 1) % NUMC_BATCH_QUEUES;
      q0 = qs[q];
      qn = qs[q + 1] - q0;
      if (threadIdx.x == 0) nxt = atomicAdd(counter + q * NUMC_BATCH_COUNTER_STRIDE, 1);
      continue;
    }
    if (threadIdx.x == 0) nxt = atomicAdd(counter + q * NUMC_BATCH_COUNTER_STRIDE, 1);
    const int4 w = items[q0 + it];
    const BatchProblem pr = probs[w.x];
    GemmArgs g{};
    g.A = pr.A;
    g.B = pr.B;
    g.a_plane = pr.a_plane;
    g.b_plane = pr.b_plane;
    g.nKF = pr.nKF;
    g.nKC = pr.nKC;
    g.M = pr.M;
    g.N = pr.N;
    g.ksplit = 1;
    g.C = pr.C;
    g.rso = pr.rso;
    g.cso = pr.cso;
    g.out_bytes = pr.out_bytes;
    g.nprod = 1;
    if (w.w == 0)
      gemm_tile<MF, BM1, BN1, KC1, 2, 2, MODE, EPI, NS1, 0>(g, w.y, w.z, 0, lds);
    else
      gemm_tile<MF, BM2, BN2, KC2, 2, 2, MODE, EPI, NS2, 0>(g, w.y, w.z, 0, lds);
  }
}

// tile classes of the batched kernel per family: {BM, BN, KC}
static const int kBatchClass[3][2][3] = {
    {{128, 128, 16}, {64, 64, 32}},   // f32
    {{128, 128, 32}, {64, 64, 32}},   // split-BF16
    {{128, 128, 32}, {64, 64, 64}},   // int8 (one digit plane)
};

// resident blocks per CU of the batched kernel (its grid is this many per CU)
int numc_batch_occupancy(int family) {
  return family == NUMC_FAMILY_I8 ? 3 : family == NUMC_FAMILY_F32 ? 3 : 2;
}

int numc_batch_class_shape(int family, int cls, int *bm, int *bn, int *kc) {
  int f = family == NUMC_FAMILY_F32 ? 0 : family == NUMC_FAMILY_SPLIT_BF16 ? 1 : family == NUMC_FAMILY_I8 ? 2 : -1;
  if (f < 0 || cls < 0 || cls > 1) return -1;
  *bm = kBatchClass[f][cls][0];
  *bn = kBatchClass[f][cls][1];
  *kc = kBatchClass[f][cls][2];
  return 0;
}


// ===========================================================================
// configuration tables
// ===========================================================================

template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int EPI, int NS, int VAR>
static void launch_gemm(const GemmArgs &g, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((k_gemm<MF, BM, BN, KC, WGM, WGN, MODE, EPI, NS, VAR>), grid, dim3(WGM * WGN * 64), 0, s, g);
}

#define CFGP(MF, BM, BN, KC, WGM, WGN, MODE, EPI, SPLIT, NS, VAR) \
  {{BM, BN, KC, WGM, WGN, SPLIT, 1, 0}, MF::T, &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE, EPI, NS, VAR>, nullptr}
#define CFGS(MF, BM, BN, KC, WGM, WGN, MODE, EPI, SPLIT, NS) CFGP(MF, BM, BN, KC, WGM, WGN, MODE, EPI, SPLIT, NS, 0)
// plain FP32 / FP64 / integer configurations also compile a biased-store twin (EPI + 4)
#define CFGB(MF, BM, BN, KC, WGM, WGN, EPI, SPLIT, NS)                                                      \
  {{BM, BN, KC, WGM, WGN, SPLIT, 1, 0}, MF::T, &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI, NS, 0>, \
   &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI + 4, NS, 0>}

// ... and also the twins that read A, B or both in place (FP32 and int8; the
// 16-bit kernels consume A fast enough that the gathered copies slowed them
// by 10 % at 8192^3, more than the pack they save)
#define CFGIP(MF, BM, BN, KC, WGM, WGN, EPI, NS, V)                                              \
  {&launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI, NS, V>,                            \
   &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI + 4, NS, V>}
#define CFGD(MF, BM, BN, KC, WGM, WGN, EPI, SPLIT, NS)                                                      \
  {{BM, BN, KC, WGM, WGN, SPLIT, 1, 0}, MF::T, &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI, NS, 0>, \
   &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI + 4, NS, 0>,                                  \
   {CFGIP(MF, BM, BN, KC, WGM, WGN, EPI, NS, 2), CFGIP(MF, BM, BN, KC, WGM, WGN, EPI, NS, 4),             \
    CFGIP(MF, BM, BN, KC, WGM, WGN, EPI, NS, 6)}}
// A in place only: int8 (its kernel slowed more with both operands gathered
// than the B pack costs: 4096^3 with B^T rows contiguous 0.0995 ms against
// 0.0951 ms with B packed) and an FP32 configuration whose B-in-place twins
// would spill
#define CFGDA(MF, BM, BN, KC, WGM, WGN, EPI, SPLIT, NS)                                                     \
  {{BM, BN, KC, WGM, WGN, SPLIT, 1, 0}, MF::T, &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI, NS, 0>, \
   &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE_SINGLE, EPI + 4, NS, 0>,                                  \
   {CFGIP(MF, BM, BN, KC, WGM, WGN, EPI, NS, 2), {nullptr, nullptr}, {nullptr, nullptr}}}

// (family, tile, k-chunk, waves, split-K, LDS stages); stages x chunk bytes <= 160 KiB
static const GemmConfig kF32[] = {
    CFGD(MfF32, 256, 256, 32, 2, 4, EPI_F32, 1, 2),
    CFGDA(MfF32, 256, 256, 16, 2, 2, EPI_F32, 1, 4),
    CFGD(MfF32, 256, 128, 16, 2, 2, EPI_F32, 1, 4),
    CFGD(MfF32, 128, 256, 16, 2, 2, EPI_F32, 1, 4),
    CFGD(MfF32, 128, 128, 32, 2, 2, EPI_F32, 1, 4),
    CFGD(MfF32, 64, 64, 32, 2, 2, EPI_F32, 1, 4),
    CFGD(MfF32, 128, 128, 32, 2, 2, EPI_F32, 4, 4),
    CFGD(MfF32, 64, 64, 32, 2, 2, EPI_F32, 8, 4),
};
static const GemmConfig kSplit[] = {
    CFGP(MfBF16, 256, 256, 32, 4, 2, MODE_SPLIT3, EPI_F32, 1, 2, 1),
    CFGS(MfBF16, 256, 256, 32, 2, 4, MODE_SPLIT3, EPI_F32, 1, 2),
    CFGS(MfBF16, 256, 256, 16, 2, 2, MODE_SPLIT3, EPI_F32, 1, 4),
    CFGS(MfBF16, 256, 128, 32, 4, 2, MODE_SPLIT3, EPI_F32, 1, 3),
    CFGS(MfBF16, 128, 256, 32, 2, 4, MODE_SPLIT3, EPI_F32, 1, 3),
    CFGS(MfBF16, 128, 128, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1, 4),
    CFGS(MfBF16, 64, 64, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1, 4),
    CFGS(MfBF16, 128, 128, 32, 2, 2, MODE_SPLIT3, EPI_F32, 4, 4),
    CFGS(MfBF16, 64, 64, 32, 2, 2, MODE_SPLIT3, EPI_F32, 8, 4),
};
static const GemmConfig kFP8[] = {
    CFGS(MfFP8, 256, 128, 64, 2, 2, MODE_SINGLE, EPI_F32, 1, 4),
    CFGS(MfFP8, 256, 256, 64, 4, 2, MODE_SINGLE, EPI_F32, 1, 4),
    CFGS(MfFP8, 256, 256, 128, 4, 2, MODE_SINGLE, EPI_F32, 1, 2),
    CFGS(MfFP8, 256, 256, 64, 2, 2, MODE_SINGLE, EPI_F32, 1, 4),
    CFGS(MfFP8, 128, 128, 128, 2, 2, MODE_SINGLE, EPI_F32, 1, 4),
    CFGS(MfFP8, 64, 64, 128, 2, 2, MODE_SINGLE, EPI_F32, 1, 4),
    CFGS(MfFP8, 128, 128, 128, 2, 2, MODE_SINGLE, EPI_F32, 4, 4),
};
static const GemmConfig kF64[] = {
    CFGB(MfF64, 128, 128, 16, 2, 2, EPI_F64, 1, 4),
    CFGB(MfF64, 64, 64, 32, 2, 2, EPI_F64, 1, 4),
    CFGB(MfF64, 128, 64, 16, 2, 2, EPI_F64, 1, 4),
    CFGB(MfF64, 64, 64, 32, 2, 2, EPI_F64, 8, 4),
    CFGB(MfF64, 256, 128, 16, 4, 2, EPI_F64, 1, 3),
    CFGB(MfF64, 128, 128, 32, 2, 2, EPI_F64, 1, 2),
};
static const GemmConfig kI8[] = {
    CFGDA(MfI8, 128, 128, 64, 2, 2, EPI_INT, 1, 4),
    CFGDA(MfI8, 256, 128, 64, 2, 2, EPI_INT, 1, 4),
    CFGDA(MfI8, 64, 64, 128, 2, 2, EPI_INT, 1, 4),
    CFGDA(MfI8, 128, 128, 64, 2, 2, EPI_INT, 4, 4),
    CFGDA(MfI8, 256, 256, 64, 4, 2, EPI_INT, 1, 4),
    CFGDA(MfI8, 256, 256, 128, 4, 2, EPI_INT, 1, 2),
    // 8 waves on a 128 x 128 tile: two waves per SIMD where a problem has
    // one tile per CU (int8 2048^3: 10 % faster than 4 waves)
    CFGDA(MfI8, 128, 128, 64, 2, 4, EPI_INT, 1, 4),
};

// bfloat16 / float16 arrays: one MFMA product (fragments twice as dense in k
// as the split planes, so the k-chunks are twice as long)
static const GemmConfig kBF16[] = {
    CFGB(MfBF16, 256, 256, 64, 2, 4, EPI_BF16, 1, 2),
    CFGB(MfBF16, 256, 256, 32, 2, 4, EPI_BF16, 1, 4),
    CFGB(MfBF16, 256, 128, 64, 4, 2, EPI_BF16, 1, 2),
    CFGB(MfBF16, 128, 256, 64, 2, 4, EPI_BF16, 1, 2),
    CFGB(MfBF16, 128, 128, 64, 2, 2, EPI_BF16, 1, 3),
    CFGB(MfBF16, 64, 64, 64, 2, 2, EPI_BF16, 1, 4),
    CFGB(MfBF16, 128, 128, 64, 2, 2, EPI_BF16, 4, 3),
    CFGB(MfBF16, 64, 64, 64, 2, 2, EPI_BF16, 8, 4),
};
static const GemmConfig kF16[] = {
    CFGB(MfF16, 256, 256, 64, 2, 4, EPI_F16, 1, 2),
    CFGB(MfF16, 256, 256, 32, 2, 4, EPI_F16, 1, 4),
    CFGB(MfF16, 256, 128, 64, 4, 2, EPI_F16, 1, 2),
    CFGB(MfF16, 128, 256, 64, 2, 4, EPI_F16, 1, 2),
    CFGB(MfF16, 128, 128, 64, 2, 2, EPI_F16, 1, 3),
    CFGB(MfF16, 64, 64, 64, 2, 2, EPI_F16, 1, 4),
    CFGB(MfF16, 128, 128, 64, 2, 2, EPI_F16, 4, 3),
    CFGB(MfF16, 64, 64, 64, 2, 2, EPI_F16, 8, 4),
};

// strided batches of equal slices (numc_bmm): small tiles with two LDS
// stages, so that several blocks share a CU and one block's copies and
// stores overlap another's MFMAs (no split-K: the slices fill the GPU)
static const GemmConfig kStridedBatch[NUMC_FAMILY_COUNT][2] = {
    {CFGD(MfF32, 128, 128, 16, 2, 2, EPI_F32, 1, 2), CFGD(MfF32, 64, 64, 32, 2, 2, EPI_F32, 1, 2)},
    {CFGS(MfBF16, 128, 128, 16, 2, 2, MODE_SPLIT3, EPI_F32, 1, 2), CFGS(MfBF16, 64, 64, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1, 2)},
    {CFGS(MfFP8, 128, 128, 64, 2, 2, MODE_SINGLE, EPI_F32, 1, 2), CFGS(MfFP8, 64, 64, 128, 2, 2, MODE_SINGLE, EPI_F32, 1, 2)},
    {CFGB(MfF64, 64, 64, 16, 2, 2, EPI_F64, 1, 2), CFGB(MfF64, 64, 64, 32, 2, 2, EPI_F64, 1, 2)},
    {CFGDA(MfI8, 128, 128, 64, 2, 2, EPI_INT, 1, 2), CFGDA(MfI8, 64, 64, 64, 2, 2, EPI_INT, 1, 2)},
    {CFGB(MfBF16, 128, 128, 32, 2, 2, EPI_BF16, 1, 2), CFGB(MfBF16, 64, 64, 64, 2, 2, EPI_BF16, 1, 2)},
    {CFGB(MfF16, 128, 128, 32, 2, 2, EPI_F16, 1, 2), CFGB(MfF16, 64, 64, 64, 2, 2, EPI_F16, 1, 2)},
};

const GemmConfig *numc_strided_batch_config(int family, int M, int N) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return nullptr;
  return &kStridedBatch[family][M >= 128 && N >= 128 ? 0 : 1];
}

#undef CFGB
#undef CFGD
#undef CFGDA
#undef CFGIP

struct FamilyTable {
  const GemmConfig *cfg;
  int n;
};
static const FamilyTable kFamilies[NUMC_FAMILY_COUNT] = {
    {kF32, (int)(sizeof kF32 / sizeof kF32[0])},
    {kSplit, (int)(sizeof kSplit / sizeof kSplit[0])},
    {kFP8, (int)(sizeof kFP8 / sizeof kFP8[0])},
    {kF64, (int)(sizeof kF64 / sizeof kF64[0])},
    {kI8, (int)(sizeof kI8 / sizeof kI8[0])},
    {kBF16, (int)(sizeof kBF16 / sizeof kBF16[0])},
    {kF16, (int)(sizeof kF16 / sizeof kF16[0])},
};

// a family's packed configurations, then its direct ones
const GemmConfig *numc_gemm_config(int family, int idx) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT || idx < 0) return nullptr;
  if (idx < kFamilies[family].n) return &kFamilies[family].cfg[idx];
  int nd = 0;
  const GemmConfig *d = numc_direct_configs(family, &nd);
  idx -= kFamilies[family].n;
  return idx < nd ? &d[idx] : nullptr;
}

int numc_gemm_config_count(int family) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return 0;
  int nd = 0;
  numc_direct_configs(family, &nd);
  return kFamilies[family].n + nd;
}

extern "C" NUMC_API int numc_kernel_config_count(int family) { return numc_gemm_config_count(family); }

extern "C" NUMC_API int numc_kernel_config_get(int family, int index, NumcKernelConfig *out) {
  const GemmConfig *c = numc_gemm_config(family, index);
  if (!c || !out) return NUMC_SET_ERROR(NUMC_ERR_BOUNDS, "no config %d in family %d", index, family);
  *out = c->cfg;
  return 0;
}

// ===========================================================================
// host launch of one GEMM problem
// ===========================================================================


// packed bytes of one plane: rows padded to the block tile, k to the chunk
static void plane_geometry(const GemmConfig &c, int R, int K, int kf_elems, int block_rows, int *nRT, int *nKF,
                           long long *bytes) {
  long long rp = cdiv(R > 0 ? R : 1, block_rows) * block_rows;
  long long kp = cdiv(K > 0 ? K : 1, c.cfg.kc) * c.cfg.kc;
  *nRT = (int)(rp / c.T);
  *nKF = (int)(kp / kf_elems);
  *bytes = (long long)(*nRT) * (*nKF) * NUMC_FRAG_BYTES;
}

static int family_kf(int family) {
  switch (family) {
  case NUMC_FAMILY_F32: return MfF32::KF;
  case NUMC_FAMILY_SPLIT_BF16: return MfBF16::KF;
  case NUMC_FAMILY_FP8: return MfFP8::KF;
  case NUMC_FAMILY_F64: return MfF64::KF;
  case NUMC_FAMILY_BF16: return MfBF16::KF;
  case NUMC_FAMILY_F16: return MfF16::KF;
  default: return MfI8::KF;
  }
}

// Digit planes of an integer GEMM: one per output byte (the int32 MFMA sum is
// exact mod 2^32), except that an int8 input is its own single signed digit,
// whatever the output width.
static int digit_planes(NumcDType in, NumcDType out) {
  if (in == NUMC_DTYPE_INT8) return 1;
  size_t s = numc_dtype_size(out);
  return s >= 4 ? 4 : (int)s;
}

// Signed base-256 digits of the 16 values of one fragment lane into P planes
// (16 bytes per plane), and the planes' nonzero mask for the wave.
template <class S, int P>
__device__ __attribute__((always_inline)) inline void digits_unit(const PackArgs &a, const uint32_t *t, long long off,
                                                                  bool live, int seen) {
  uint32_t w[P][4];
#pragma unroll
  for (int p = 0; p < P; p++) w[p][0] = w[p][1] = w[p][2] = w[p][3] = 0;
  if (live) {
#pragma unroll
    for (int e = 0; e < 16; e++) digit_bytes<P>((uint32_t)(long long)(S)t[e], w, e);
#pragma unroll
    for (int p = 0; p < P; p++)
      *(uint4 *)(a.dst + p * a.plane_bytes + off) = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
  }
  int mask = 0;
#pragma unroll
  for (int p = 0; p < P; p++)
    if (__ballot((w[p][0] | w[p][1] | w[p][2] | w[p][3]) != 0)) mask |= 1 << p;
  if ((threadIdx.x & 63) == 0 && mask && a.flags) flags_or_seen(a.flags, mask, seen); // (seen: flags_peek)
}

// one 16-byte chunk of source values (EPC of them) into LDS as raw 32-bit
// words; (S)word gives the value back
// the values of a 16-byte chunk as 32-bit words; (S)word gives each back
template <class S>
__device__ __attribute__((always_inline)) inline void unpack_chunk(uint4 q, uint32_t *v) {
  constexpr int EPC = 16 / sizeof(S);
  union { uint4 q; S e[EPC]; } u;
  u.q = q;
#pragma unroll
  for (int e = 0; e < EPC; e++) v[e] = (uint32_t)u.e[e];
}

// integer digit planes, rows with unit k-stride: 32 rows x 128 k tiles
// staged through LDS so that global reads are 16-byte and coalesced
template <class S, int P>
__device__ __attribute__((always_inline)) inline void digits_tile_k(const PackArgs &a, long long bid, uint32_t *tile) {
  constexpr int TK = 128, PITCH = TK + 1, EPC = 16 / sizeof(S), CPR = TK / EPC;
  static_assert(32 * PITCH <= PACK_TILE_WORDS, "tile");
  constexpr int NL = 32 * CPR / 256; // 16-byte loads per thread
  static_assert(NL >= 1 && 32 * CPR % 256 == 0, "chunks");
  const int nkt = (a.nKF * 32 + TK - 1) / TK;
  const long long rt = bid / nkt;
  const int kt = (int)(bid % nkt);
  const int tid = threadIdx.x;
  const long long row0 = rt * 32;
  const int k0 = kt * TK;
  const S *base = (const S *)a.src + row0 * a.rs + k0;
  const int seen = flags_peek(a.flags);
  uint4 q[NL];
  if (row0 + 32 <= a.R && k0 + TK <= a.K) {
    // interior tile: every load in flight at once
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256;
      q[it] = ldu<uint4>(base + (idx / CPR) * a.rs + (idx % CPR) * EPC);
    }
  } else {
    EdgeChunk e[NL];
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256, r = idx / CPR, c = (idx % CPR) * EPC;
      e[it] = edge_issue<S>(base + r * a.rs + c, row0 + r < a.R ? (long long)a.K - (k0 + c) : 0);
    }
#pragma unroll
    for (int it = 0; it < NL; it++) q[it] = edge_finish(e[it]);
  }
#pragma unroll
  for (int it = 0; it < NL; it++) {
    const int idx = tid + it * 256, r = idx / CPR, c = (idx % CPR) * EPC;
    uint32_t v[EPC];
    unpack_chunk<S>(q[it], v);
#pragma unroll
    for (int e = 0; e < EPC; e++) tile[r * PITCH + c + e] = v[e];
  }
  __syncthreads();
  // 4 fragments x 64 lanes: one unit per thread
  const int f = tid >> 6, lane = tid & 63;
  const int kf = kt * (TK / 32) + f;
  uint32_t t[16];
#pragma unroll
  for (int e = 0; e < 16; e++) t[e] = tile[(lane & 31) * PITCH + f * 32 + (lane >> 5) * 16 + e];
  digits_unit<S, P>(a, t, (rt * a.nKF + kf) * 1024 + lane * 16, kf < a.nKF, seen);
}

template <class S, int P>
__global__ void __launch_bounds__(256) k_pack_digits_tiled(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  digits_tile_k<S, P>(a, blockIdx.x, tile);
}

// integer digit planes, rows contiguous (unit row stride): 128 rows x 32 k
template <class S, int P>
__device__ __attribute__((always_inline)) inline void digits_tile_rows(const PackArgs &a, long long bid, uint32_t *tile) {
  constexpr int TR = 128, TK = 32, PITCH = TK + 1, EPC = 16 / sizeof(S), CPK = TR / EPC;
  static_assert(TR * PITCH <= PACK_TILE_WORDS, "tile");
  // one k fragment (32 k) per tile; consecutive blocks take neighbouring
  // row blocks at one k range, so the segments they read complete each
  // other's source rows
  constexpr int NL = TK * CPK / 256; // 16-byte loads per thread
  static_assert(NL >= 1 && TK * CPK % 256 == 0, "chunks");
  const int nrb = (a.nRT + 3) / 4;
  const long long rb = bid % nrb;
  const int kt = (int)(bid / nrb);
  const int tid = threadIdx.x;
  const long long row0 = rb * TR;
  const int k0 = kt * TK;
  const S *base = (const S *)a.src + (long long)k0 * a.ks + row0;
  const int seen = flags_peek(a.flags);
  uint4 q[NL];
  if (row0 + TR <= a.R && k0 + TK <= a.K) {
    // interior tile: every load in flight at once
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256;
      q[it] = ldu<uint4>(base + (long long)(idx / CPK) * a.ks + (idx % CPK) * EPC);
    }
  } else {
    EdgeChunk e[NL];
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256, kk = idx / CPK, r = (idx % CPK) * EPC;
      e[it] = edge_issue<S>(base + (long long)kk * a.ks + r, k0 + kk < a.K ? (long long)a.R - (row0 + r) : 0);
    }
#pragma unroll
    for (int it = 0; it < NL; it++) q[it] = edge_finish(e[it]);
  }
#pragma unroll
  for (int it = 0; it < NL; it++) {
    const int idx = tid + it * 256, kk = idx / CPK, r = (idx % CPK) * EPC;
    uint32_t v[EPC];
    unpack_chunk<S>(q[it], v);
#pragma unroll
    for (int e = 0; e < EPC; e++) tile[(r + e) * PITCH + kk] = v[e];
  }
  __syncthreads();
  // 4 row fragments x 64 lanes: one unit per thread
  const int rl = tid >> 6, lane = tid & 63;
  const long long rt = rb * 4 + rl;
  uint32_t t[16];
#pragma unroll
  for (int e = 0; e < 16; e++) t[e] = tile[(rl * 32 + (lane & 31)) * PITCH + (lane >> 5) * 16 + e];
  digits_unit<S, P>(a, t, (rt * a.nKF + kt) * 1024 + lane * 16, rt < a.nRT, seen);
}

template <class S, int P>
__global__ void __launch_bounds__(256) k_pack_digits_tiled_rows(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  digits_tile_rows<S, P>(a, blockIdx.x, tile);
}

// blocks of one operand's digit pack and the walk it takes (mode 0: k-contiguous
// tiles, 1: row-contiguous tiles, 2: one thread per lane slot)
static long long digits_blocks(PackArgs &pk) {
  if (pk.ks == 1) {
    pk.mode = 0;
    return (long long)pk.nRT * ((pk.nKF * 32LL + 127) / 128);
  }
  if (pk.rs == 1) {
    pk.mode = 1;
    return ((pk.nRT + 3) / 4) * (long long)pk.nKF;
  }
  pk.mode = 2;
  return ((long long)pk.nRT * pk.nKF * 64 + 255) / 256;
}

// both operands' digit planes in one launch: blocks [0, na) pack A, the rest B
template <class S, int P>
__global__ void __launch_bounds__(256) k_pack_pair_digits(const PackArgs a, const PackArgs b, unsigned na) {
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  const bool isa = blockIdx.x < na;
  const PackArgs &x = isa ? a : b;
  const long long u = isa ? blockIdx.x : blockIdx.x - na;
  if (x.mode == 0) digits_tile_k<S, P>(x, u, tile);
  else if (x.mode == 1) digits_tile_rows<S, P>(x, u, tile);
  else pack_digits_unit<S, P>(x, u * 256 + threadIdx.x);
}

template <class S, int P> static void launch_pair_digits_p(PackArgs &pa, PackArgs &pb, hipStream_t st) {
  const long long na = digits_blocks(pa), nb = digits_blocks(pb);
  hipLaunchKernelGGL((k_pack_pair_digits<S, P>), dim3((unsigned)(na + nb)), dim3(256), 0, st, pa, pb, (unsigned)na);
}

template <class S> static void launch_pair_digits(int planes, PackArgs &pa, PackArgs &pb, hipStream_t st) {
  switch (planes) {
  case 1: launch_pair_digits_p<S, 1>(pa, pb, st); break;
  case 2: launch_pair_digits_p<S, 2>(pa, pb, st); break;
  default: launch_pair_digits_p<S, 4>(pa, pb, st); break;
  }
}

// ---------------------------------------------------------------------------
// one-byte operands with one digit plane (int8 and uint8 results): a digit is
// the byte itself, so a fragment lane is 16 source bytes, copied as is
// ---------------------------------------------------------------------------

// k-contiguous rows: a tile of 32*RT rows x 512/RT bytes of k (16
// fragments), RT = 1, 2 or 4 as a.tile_rt says (set on the host with the
// block count, so that short rows fill a block with more of them).  Each row
// segment is read as 512/RT contiguous bytes, staged in LDS (row pitch 16
// bytes more, so the 16-byte lane reads of consecutive rows fall on
// different banks), then written out as 16 whole fragments.
template <int RT, bool STREAM>
__device__ __attribute__((always_inline)) inline void bytes_tile_k_rt(const PackArgs &a, long long bid, uint32_t *tile) {
  constexpr int KB = 512 / RT, CPR = KB / 16, PW = CPR + 1, NF = KB / 32, ROWS = 32 * RT;
  static_assert(ROWS * PW * 4 <= PACK_TILE_WORDS, "tile");
  constexpr int NL = ROWS * CPR / 256; // 16-byte loads per thread
  uint4 *t = (uint4 *)tile;
  const int nkb = (a.nKF + NF - 1) / NF;
  const long long rb = bid / nkb; // the block's RT row tiles start at row tile rb * RT
  const int kb = (int)(bid % nkb);
  const int tid = threadIdx.x;
  const long long row0 = rb * ROWS;
  const unsigned char *base = (const unsigned char *)a.src + row0 * a.rs + (long long)kb * KB;
  const int seen = flags_peek(a.flags);
  uint4 v[NL];
  if (row0 + ROWS <= a.R && kb * KB + KB <= a.K) {
    // interior tile: every load in flight at once
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256;
      v[it] = STREAM ? ldu_stream<uint4>(base + (idx / CPR) * a.rs + (idx % CPR) * 16)
                     : ldu<uint4>(base + (idx / CPR) * a.rs + (idx % CPR) * 16);
    }
  } else {
    EdgeChunk e[NL];
#pragma unroll
    for (int it = 0; it < NL; it++) {
      const int idx = tid + it * 256, r = idx / CPR, c = idx % CPR;
      const int k = kb * KB + c * 16;
      e[it] = edge_issue<unsigned char>(base + r * a.rs + c * 16, row0 + r < a.R ? (long long)a.K - k : 0);
    }
#pragma unroll
    for (int it = 0; it < NL; it++) v[it] = edge_finish(e[it]);
  }
#pragma unroll
  for (int it = 0; it < NL; it++) {
    const int idx = tid + it * 256;
    t[(idx / CPR) * PW + idx % CPR] = v[it];
  }
  __syncthreads();
  bool any = false;
#pragma unroll
  for (int it = 0; it < 16 * 64 / 256; it++) {
    const int u = tid + it * 256, f = u >> 6, lane = u & 63;
    const int j = f / NF, fk = f % NF; // row tile j of the block, its k fragment fk
    const long long rt = rb * RT + j;
    const int kf = kb * NF + fk;
    if (rt < a.nRT && kf < a.nKF) {
      const uint4 v = t[(j * 32 + (lane & 31)) * PW + fk * 2 + (lane >> 5)];
      *(uint4 *)(a.dst + (rt * a.nKF + kf) * 1024 + lane * 16) = v;
      any |= (v.x | v.y | v.z | v.w) != 0;
    }
  }
  if (__ballot(any) && (threadIdx.x & 63) == 0 && a.flags) flags_or_seen(a.flags, 1, seen);
}

template <bool STREAM = false>
__device__ __attribute__((always_inline)) inline void bytes_tile_k(const PackArgs &a, long long bid, uint32_t *tile) {
  if (a.tile_rt == 4) bytes_tile_k_rt<4, STREAM>(a, bid, tile);
  else if (a.tile_rt == 2) bytes_tile_k_rt<2, STREAM>(a, bid, tile);
  else bytes_tile_k_rt<1, STREAM>(a, bid, tile);
}

__global__ void __launch_bounds__(256) k_pack_bytes_k(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  bytes_tile_k(a, blockIdx.x, tile);
}

// row-contiguous rows (B^T of a row-major B): a 128-row x 128-k byte tile.
// Each thread reads 16 consecutive rows at 4 consecutive k (four 16-byte
// loads), transposes the 4 x 16 bytes in registers and writes 16 dwords of
// 4 k each into LDS; each fragment lane then reads its 16 k of one row.
// The 16-byte column chunks of row r sit at chunk ^ ((r >> 4) & 7), so the
// eight 16-row groups that one wave writes fall on different banks.
__device__ __attribute__((always_inline)) inline void bytes_tile_rows(const PackArgs &a, long long bid, uint32_t *tile) {
  constexpr int TR = 128, TK = 128, PITCH = TK + 16; // bytes per LDS row
  static_assert(TR * PITCH / 4 <= PACK_TILE_WORDS, "tile");
  unsigned char *t8 = (unsigned char *)tile;
  const int nrb = (a.nRT + 3) / 4;
  const long long rb = bid % nrb;
  const int kt = (int)(bid / nrb);
  const long long row0 = rb * TR;
  const int k0 = kt * TK;
  const int tid = threadIdx.x;
  // thread -> (k quad, 16-row group): 32 quads x 8 groups
  const int kq = tid >> 3, rg = tid & 7;
  const int kk = k0 + kq * 4;
  const long long r = row0 + rg * 16;
  uint32_t w[4][4]; // [k offset][dword of 4 rows]
  uint4 v[4];
  const unsigned char *base = (const unsigned char *)a.src + (long long)k0 * a.ks + row0;
  const int seen = flags_peek(a.flags);
  if (row0 + TR <= a.R && k0 + TK <= a.K) {
    // interior tile: all four loads in flight at once
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = ldu<uint4>(base + (long long)(kq * 4 + q) * a.ks + rg * 16);
  } else {
    EdgeChunk e[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = kk + q;
      e[q] = edge_issue<unsigned char>((const unsigned char *)a.src + (long long)k * a.ks + r, k < a.K ? a.R - r : 0);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = edge_finish(e[q]);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    w[q][0] = v[q].x;
    w[q][1] = v[q].y;
    w[q][2] = v[q].z;
    w[q][3] = v[q].w;
  }
  // row rg*16 + 4g + j gets bytes j of w[0..3][g]: 4 consecutive k
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t d = ((w[0][g] >> (8 * j)) & 0xFFu) | (((w[1][g] >> (8 * j)) & 0xFFu) << 8) |
                         (((w[2][g] >> (8 * j)) & 0xFFu) << 16) | (((w[3][g] >> (8 * j)) & 0xFFu) << 24);
      *(uint32_t *)(t8 + (rg * 16 + 4 * g + j) * PITCH + (((kq >> 2) ^ rg) << 4) + (kq & 3) * 4) = d;
    }
  __syncthreads();
  // 4 row fragments x 4 k fragments x 64 lanes: four lane slots per thread
  bool any = false;
#pragma unroll
  for (int it = 0; it < 4; it++) {
    const int u = tid + it * 256;
    const int f = u >> 6, lane = u & 63;
    const int rl = f >> 2, fk = f & 3;
    const long long rt = rb * 4 + rl;
    const int kf = kt * 4 + fk;
    if (rt < a.nRT && kf < a.nKF) {
      const int row = rl * 32 + (lane & 31);
      const uint4 v = *(const uint4 *)(t8 + row * PITCH + (((fk * 2 + (lane >> 5)) ^ ((row >> 4) & 7)) << 4));
      *(uint4 *)(a.dst + (rt * a.nKF + kf) * 1024 + lane * 16) = v;
      any |= (v.x | v.y | v.z | v.w) != 0;
    }
  }
  if (__ballot(any) && (threadIdx.x & 63) == 0 && a.flags) flags_or_seen(a.flags, 1, seen);
}

__global__ void __launch_bounds__(256) k_pack_bytes_rows(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  bytes_tile_rows(a, blockIdx.x, tile);
}

// ---------------------------------------------------------------------------
// bfloat16 / float16 operands: raw 16-bit copies.  A fragment lane is 8
// values (16 bytes) of one row, so the byte geometry is that of the int8
// fragments: 32-byte row segments, lane halves of 16 bytes.
// ---------------------------------------------------------------------------

// k-contiguous rows: the int8 byte tile with K and the row stride in bytes
__global__ void __launch_bounds__(256) k_pack_raw16_k(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  a.K *= 2;
  a.rs *= 2;
  bytes_tile_k<true>(a, blockIdx.x, tile);
}

__device__ __attribute__((always_inline)) inline uint32_t u4_at(const uint4 &v, int i) {
  return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// row-contiguous rows (B^T of a row-major B): a 128-row x 64-k tile (4 k
// fragments).  Each thread reads 8 consecutive rows at 4 consecutive k (four
// 16-byte loads), regroups them into 4 k per row and writes 8 bytes per row
// into LDS; each fragment lane then reads its 8 k of one row.  The 16-byte
// column chunks of row r sit at chunk ^ ((r >> 3) & 7).
__device__ __attribute__((always_inline)) inline void raw16_tile_rows(const PackArgs &a, long long bid, uint32_t *tile) {
  constexpr int TR = 128, TK = 64, PITCH = TK * 2 + 16; // bytes per LDS row
  static_assert(TR * PITCH / 4 <= PACK_TILE_WORDS, "tile");
  unsigned char *t8 = (unsigned char *)tile;
  const int nrb = (a.nRT + 3) / 4;
  const long long rb = bid % nrb;
  const int kt = (int)(bid / nrb);
  const long long row0 = rb * TR;
  const int k0 = kt * TK;
  const int tid = threadIdx.x;
  const int rg = tid & 15, kq = tid >> 4; // 16 groups of 8 rows x 16 groups of 4 k
  const uint16_t *src = (const uint16_t *)a.src;
  const uint16_t *base = src + (long long)k0 * a.ks + row0;
  uint4 v[4];
  if (row0 + TR <= a.R && k0 + TK <= a.K) {
    // interior tile: all four loads in flight at once
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = ldu_stream<uint4>(base + (long long)(kq * 4 + q) * a.ks + rg * 8);
  } else {
    EdgeChunk e[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = k0 + kq * 4 + q;
      const long long r = row0 + rg * 8;
      e[q] = edge_issue<uint16_t>(src + (long long)k * a.ks + r, k < a.K ? a.R - r : 0);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = edge_finish(e[q]);
  }
#pragma unroll
  for (int j = 0; j < 8; j++) { // row rg*8 + j: the j-th 16-bit value of each load
    uint32_t h[4];
#pragma unroll
    for (int q = 0; q < 4; q++) h[q] = (u4_at(v[q], j >> 1) >> (16 * (j & 1))) & 0xFFFFu;
    const int row = rg * 8 + j;
    const int chunk = (kq >> 1) ^ (rg & 7);
    *(uint2 *)(t8 + row * PITCH + (chunk << 4) + (kq & 1) * 8) = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
  }
  __syncthreads();
  // 4 row fragments x 4 k fragments x 64 lanes: four lane slots per thread
#pragma unroll
  for (int it = 0; it < 4; it++) {
    const int u = tid + it * 256;
    const int f = u >> 6, lane = u & 63;
    const int rl = f >> 2, fk = f & 3;
    const long long rt = rb * 4 + rl;
    const int kf = kt * 4 + fk;
    if (rt < a.nRT && kf < a.nKF) {
      const int row = rl * 32 + (lane & 31);
      const int chunk = (fk * 2 + (lane >> 5)) ^ ((row >> 3) & 7);
      *(uint4 *)(a.dst + (rt * a.nKF + kf) * 1024 + lane * 16) = *(const uint4 *)(t8 + row * PITCH + (chunk << 4));
    }
  }
}

__global__ void __launch_bounds__(256) k_pack_raw16_rows(PackArgs a) {
  zslice(a);
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  raw16_tile_rows(a, blockIdx.x, tile);
}

// any other strides: one thread per fragment lane slot
__device__ inline void raw16_unit(const PackArgs &a, long long t) {
  const int lane = (int)(t & 63);
  const long long fr = t >> 6;
  const int kf = (int)(fr % a.nKF);
  const long long row = (fr / a.nKF) * 32 + (lane & 31);
  const int k0 = kf * 16 + (lane >> 5) * 8;
  const uint16_t *src = (const uint16_t *)a.src;
  uint32_t w[4] = {0, 0, 0, 0};
  if (row < a.R) {
#pragma unroll
    for (int e = 0; e < 8; e++)
      if (k0 + e < a.K) w[e >> 1] |= (uint32_t)src[row * a.rs + (long long)(k0 + e) * a.ks] << (16 * (e & 1));
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void __launch_bounds__(256) k_pack_raw16(PackArgs a) {
  zslice(a);
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t < (long long)a.nRT * a.nKF * 64) raw16_unit(a, t);
}

// ---------------------------------------------------------------------------
// batched pack: every operand of every problem in one launch.  Operand i owns
// blocks [prefix[i], prefix[i+1]) and owner[b] names the operand of block b
// (one load, where a search of prefix would take a chain of dependent ones);
// its `mode` (set on the host) says which tile walk it takes: 0 k-contiguous
// tiles, 1 row-contiguous tiles, 2 one thread per 16-byte lane slot (any
// other strides), 3/4 the byte tiles of int8 operands, 5/6/7 the raw 16-bit
// walks (k-contiguous, row-contiguous, other strides).
// ---------------------------------------------------------------------------
// block u of operand a, by the walk a.mode names
template <int FAM>
__device__ __attribute__((always_inline)) inline void pack_walk(const PackArgs &a, long long u, uint32_t *tile) {
  constexpr int MODE_PK = FAM == NUMC_FAMILY_F32 ? 0 : FAM == NUMC_FAMILY_FP8 ? 2 : 1;
  if constexpr (FAM == NUMC_FAMILY_BF16 || FAM == NUMC_FAMILY_F16) { // 16-bit values copied as they are
    if (a.mode == 5) {
      PackArgs b = a; // the k-contiguous rows as bytes
      b.K *= 2;
      b.rs *= 2;
      bytes_tile_k<true>(b, u, tile);
    } else if (a.mode == 6) {
      raw16_tile_rows(a, u, tile);
    } else {
      const long long t = u * 256 + threadIdx.x;
      if (t < (long long)a.nRT * a.nKF * 64) raw16_unit(a, t);
    }
    return;
  }
  if constexpr (FAM == NUMC_FAMILY_I8) {
    if (a.mode == 3) {
      bytes_tile_k(a, u, tile);
      return;
    }
    if (a.mode == 4) {
      bytes_tile_rows(a, u, tile);
      return;
    }
  }
  if (a.mode == 0) {
    if constexpr (FAM == NUMC_FAMILY_I8) digits_tile_k<int8_t, 1>(a, u, tile);
    else pack_f32_tile_k<MODE_PK>(a, u, (float *)tile);
  } else if (a.mode == 1) {
    if constexpr (FAM == NUMC_FAMILY_I8) digits_tile_rows<int8_t, 1>(a, u, tile);
    else pack_f32_tile_rows<MODE_PK>(a, u, (float *)tile);
  } else {
    const long long t = u * 256 + threadIdx.x;
    if (t < (long long)a.nRT * a.nKF * 64) {
      if constexpr (FAM == NUMC_FAMILY_F32) pack_unit_f32(a, t);
      else if constexpr (FAM == NUMC_FAMILY_SPLIT_BF16) pack_unit_split(a, t);
      else if constexpr (FAM == NUMC_FAMILY_FP8) pack_unit_fp8(a, t);
      else pack_unit_i8(a, t);
    }
  }
}

template <int FAM>
__global__ void __launch_bounds__(256) k_pack_batched(const PackArgs *descs, const long long *prefix,
                                                      const int *owner, long long units) {
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  const long long b = blockIdx.x;
  if (b >= units) return;
  const int lo = owner[b];
  const PackArgs a = descs[lo];
  pack_walk<FAM>(a, b - prefix[lo], tile);
}

// both operands of one problem in one launch (one launch gap less per call):
// blocks [0, na) pack A, the others B
template <int FAM>
__global__ void __launch_bounds__(256) k_pack_pair(const PackArgs a, const PackArgs b, unsigned na) {
  __shared__ uint32_t tile[PACK_TILE_WORDS];
  if (blockIdx.x < na) pack_walk<FAM>(a, blockIdx.x, tile);
  else pack_walk<FAM>(b, blockIdx.x - na, tile);
}

// blocks of bytes_tile_k for an operand, with its row tiles per block
// (pk.tile_rt): the fewest blocks, so that short rows fill a block with more
// of them
static long long bytes_k_blocks(PackArgs &pk) {
  long long best = -1;
  for (int rt = 1; rt <= 4; rt *= 2) {
    const long long n = cdiv((long long)pk.nRT, rt) * cdiv((long long)pk.nKF, 16 / rt);
    if (best < 0 || n < best) {
      best = n;
      pk.tile_rt = rt;
    }
  }
  return best;
}

// blocks the batched pack gives one operand, and the walk it takes
int numc_batch_pack_blocks(int family, PackArgs &pk) {
  if (family == NUMC_FAMILY_BF16 || family == NUMC_FAMILY_F16) { // 5/6/7: the raw 16-bit walks
    if (pk.ks == 1) {
      pk.mode = 5;
      return (int)bytes_k_blocks(pk);
    }
    if (pk.rs == 1) {
      pk.mode = 6;
      return (int)(cdiv(pk.nRT, 4) * cdiv(pk.nKF, 4));
    }
    pk.mode = 7;
    return (int)cdiv((long long)pk.nRT * pk.nKF * 64, 256);
  }
  const int kf = family == NUMC_FAMILY_F32 ? 8 : family == NUMC_FAMILY_SPLIT_BF16 ? 16 : 32;
  if (family == NUMC_FAMILY_I8 && pk.ks == 1) { // int8 operands, one plane: byte tiles
    pk.mode = 3;
    return (int)bytes_k_blocks(pk);
  }
  if (family == NUMC_FAMILY_I8 && pk.rs == 1) {
    pk.mode = 4;
    return (int)(cdiv(pk.nRT, 4) * cdiv(pk.nKF, 4));
  }
  if (pk.ks == 1) {
    pk.mode = 0;
    return (int)((long long)pk.nRT * cdiv((long long)pk.nKF * kf, 128));
  }
  if (pk.rs == 1) {
    pk.mode = 1;
    return (int)(cdiv(pk.nRT, 4) * cdiv((long long)pk.nKF * kf, 32));
  }
  pk.mode = 2;
  return (int)cdiv((long long)pk.nRT * pk.nKF * 64, 256);
}

template <class S, int P> static void launch_digits_p(PackArgs &pk, unsigned nb, unsigned nz, hipStream_t st) {
  if constexpr (sizeof(S) == 1 && P == 1) {
    if (pk.ks == 1) {
      const unsigned nbk = (unsigned)bytes_k_blocks(pk);
      hipLaunchKernelGGL(k_pack_bytes_k, dim3(nbk, 1, nz), dim3(256), 0, st, pk);
      return;
    }
    if (pk.rs == 1) {
      dim3 grid((unsigned)(cdiv(pk.nRT, 4) * cdiv(pk.nKF, 4)), 1, nz);
      hipLaunchKernelGGL(k_pack_bytes_rows, grid, dim3(256), 0, st, pk);
      return;
    }
  }
  if (pk.ks == 1) {
    dim3 grid((unsigned)((long long)pk.nRT * cdiv((long long)pk.nKF * 32, 128)), 1, nz);
    hipLaunchKernelGGL((k_pack_digits_tiled<S, P>), grid, dim3(256), 0, st, pk);
  } else if (pk.rs == 1) {
    dim3 grid((unsigned)(cdiv(pk.nRT, 4) * pk.nKF), 1, nz);
    hipLaunchKernelGGL((k_pack_digits_tiled_rows<S, P>), grid, dim3(256), 0, st, pk);
  } else {
    hipLaunchKernelGGL((k_pack_digits<S, P>), dim3(nb, 1, nz), dim3(256), 0, st, pk);
  }
}

template <class S>
static void launch_digits(int planes, PackArgs &pk, unsigned nb, unsigned nz, hipStream_t st) {
  switch (planes) {
  case 1: launch_digits_p<S, 1>(pk, nb, nz, st); break;
  case 2: launch_digits_p<S, 2>(pk, nb, nz, st); break;
  default: launch_digits_p<S, 4>(pk, nb, nz, st); break;
  }
}

static void launch_pack(int family, NumcDType in_dt, int planes, PackArgs &pk, unsigned nz, hipStream_t st) {
  unsigned nb = blocks_for((long long)pk.nRT * pk.nKF * 64);
  if (pk.ks == 1 && (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 || family == NUMC_FAMILY_FP8)) {
    // rows with unit k-stride: stage 32 x 128 tiles through LDS
    const int kf = family == NUMC_FAMILY_F32 ? 8 : family == NUMC_FAMILY_SPLIT_BF16 ? 16 : 32;
    const long long nkt = cdiv((long long)pk.nKF * kf, 128);
    dim3 grid((unsigned)((long long)pk.nRT * nkt), 1, nz);
    if (family == NUMC_FAMILY_F32) hipLaunchKernelGGL((k_pack_f32_tiled<0>), grid, dim3(256), 0, st, pk);
    else if (family == NUMC_FAMILY_SPLIT_BF16) hipLaunchKernelGGL((k_pack_f32_tiled<1>), grid, dim3(256), 0, st, pk);
    else hipLaunchKernelGGL((k_pack_f32_tiled<2>), grid, dim3(256), 0, st, pk);
    return;
  }
  if (pk.rs == 1 && (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 || family == NUMC_FAMILY_FP8)) {
    // rows contiguous: stage 128 x 32 tiles through LDS
    const long long nkt = cdiv((long long)pk.nKF * family_kf(family), 32);
    dim3 grid((unsigned)(cdiv(pk.nRT, 4) * nkt), 1, nz);
    if (family == NUMC_FAMILY_F32) hipLaunchKernelGGL((k_pack_f32_tiled_rows<0>), grid, dim3(256), 0, st, pk);
    else if (family == NUMC_FAMILY_SPLIT_BF16) hipLaunchKernelGGL((k_pack_f32_tiled_rows<1>), grid, dim3(256), 0, st, pk);
    else hipLaunchKernelGGL((k_pack_f32_tiled_rows<2>), grid, dim3(256), 0, st, pk);
    return;
  }
  if (family == NUMC_FAMILY_BF16 || family == NUMC_FAMILY_F16) { // raw 16-bit copies
    if (pk.ks == 1) {
      const unsigned nbk = (unsigned)bytes_k_blocks(pk);
      hipLaunchKernelGGL(k_pack_raw16_k, dim3(nbk, 1, nz), dim3(256), 0, st, pk);
    }
    else if (pk.rs == 1)
      hipLaunchKernelGGL(k_pack_raw16_rows, dim3((unsigned)(cdiv(pk.nRT, 4) * cdiv(pk.nKF, 4)), 1, nz), dim3(256), 0, st, pk);
    else
      hipLaunchKernelGGL(k_pack_raw16, dim3(nb, 1, nz), dim3(256), 0, st, pk);
    return;
  }
  switch (family) {
  case NUMC_FAMILY_F32: hipLaunchKernelGGL(k_pack_f32, dim3(nb, 1, nz), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_SPLIT_BF16: hipLaunchKernelGGL(k_pack_split_bf16, dim3(nb, 1, nz), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_FP8: hipLaunchKernelGGL(k_pack_fp8, dim3(nb, 1, nz), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_F64: hipLaunchKernelGGL(k_pack_f64, dim3(nb, 1, nz), dim3(256), 0, st, pk); break;
  default:
    switch (in_dt) {
    case NUMC_DTYPE_INT8: launch_digits<int8_t>(planes, pk, nb, nz, st); break;
    case NUMC_DTYPE_UINT8: launch_digits<uint8_t>(planes, pk, nb, nz, st); break;
    case NUMC_DTYPE_INT16: launch_digits<int16_t>(planes, pk, nb, nz, st); break;
    case NUMC_DTYPE_UINT16: launch_digits<uint16_t>(planes, pk, nb, nz, st); break;
    case NUMC_DTYPE_INT32: launch_digits<int32_t>(planes, pk, nb, nz, st); break;
    default: launch_digits<uint32_t>(planes, pk, nb, nz, st); break;
    }
  }
}

int numc_problem_planes(const GemmProblem &p) {
  return p.family == NUMC_FAMILY_I8 ? digit_planes(p.in_dt, p.out_dt) : 1;
}

bool numc_direct_ok(const GemmProblem &p) {
  if (p.batch > 1 || p.family == NUMC_FAMILY_FP8) return false;
  return p.family != NUMC_FAMILY_I8 || digit_planes(p.in_dt, p.out_dt) == 1;
}

size_t numc_gemm_workspace(const GemmProblem &p, const GemmConfig &c, size_t *aux_bytes) {
  if (c.cfg.direct) { // operands read in place: no workspace
    if (aux_bytes) *aux_bytes = 0;
    return 0;
  }
  int kf = family_kf(p.family), nRT, nKF;
  long long pa, pb;
  plane_geometry(c, p.M, p.K, kf, c.cfg.bm, &nRT, &nKF, &pa);
  plane_geometry(c, p.N, p.K, kf, c.cfg.bn, &nRT, &nKF, &pb);
  int planes = p.family == NUMC_FAMILY_SPLIT_BF16 ? 2 : (p.family == NUMC_FAMILY_I8 ? digit_planes(p.in_dt, p.out_dt) : 1);
  size_t raw = (p.family == NUMC_FAMILY_F64) ? 8 : 4;
  size_t partial = c.cfg.split_k > 1 ? (size_t)c.cfg.split_k * p.M * p.N * raw : 0;
  const size_t nz = p.batch > 1 ? (size_t)p.batch : 1;
  const size_t scales = p.family == NUMC_FAMILY_FP8 ? (size_t)(pa + pb) / 32 : 0; // FP8 block scales
  size_t bytes = (size_t)(pa + pb) * planes * nz + scales + partial + 4 * 256;
  if (aux_bytes) *aux_bytes = 2 * NUMC_FLAG_INTS * sizeof(int) + 512;
  return bytes;
}

int numc_gemm_launch(NumcCtx *ctx, const GemmProblem &p, const GemmConfig &c, void *ws, void *aux) {
  hipStream_t st = ctx->stream;
  const int family = p.family;
  const int kf = family_kf(family);
  const int planes = family == NUMC_FAMILY_SPLIT_BF16 ? 2 : (family == NUMC_FAMILY_I8 ? digit_planes(p.in_dt, p.out_dt) : 1);

  if (p.bias && (family == NUMC_FAMILY_SPLIT_BF16 || family == NUMC_FAMILY_FP8))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: a bias needs the plain method (family %d)", family);
  if (p.M == 0 || p.N == 0) return 0;
  if (p.K == 0) {
    long long n = (long long)p.M * p.N;
    const int es = (int)numc_dtype_size(p.out_dt);
    for (int z = 0; z < (p.batch > 1 ? p.batch : 1); z++)
      hipLaunchKernelGGL(k_zero_strided, dim3(blocks_for(n)), dim3(256), 0, st, (unsigned char *)p.C + z * p.c_bs * es,
                         p.M, p.N, p.rso, p.cso, es, (const unsigned char *)p.bias, p.bias_s);
    return 0;
  }

  if (c.cfg.direct) {
    // one launch on the operands as they lie (gemm_small.hip)
    if (!numc_direct_ok(p))
      return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: a direct configuration needs one digit plane and no batch");
    const size_t ein = numc_dtype_size(p.in_dt);
    GemmArgs g{};
    g.Araw = (const unsigned char *)p.A;
    g.a_rs = p.rsa * (long long)ein;
    g.a_ks = p.csa * (long long)ein;
    g.Braw = (const unsigned char *)p.B; // rows of B^T: column j at B + j*csb, k stride rsb
    g.b_rs = p.csb * (long long)ein;
    g.b_ks = p.rsb * (long long)ein;
    g.M = p.M;
    g.N = p.N;
    g.K = p.K;
    g.nbm = (int)cdiv(p.M, c.cfg.bm);
    g.nbn = (int)cdiv(p.N, c.cfg.bn);
    g.ksplit = 1;
    g.C = p.C;
    g.rso = p.rso;
    g.cso = p.cso;
    g.out_bytes = (int)numc_dtype_size(p.out_dt);
    g.bias = p.bias;
    g.bias_s = p.bias_s;
    g.nprod = 1;
    if ((long long)g.nbm * g.nbn > 0x7FFFFFFFLL) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: too many tiles");
    GemmLaunchFn fn = p.bias ? c.launch_bias : c.launch;
    if (!fn) return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: a bias needs the plain method (family %d)", family);
    numc_prof_begin(ctx, PROF_GEMM, family);
    fn(g, dim3((unsigned)((long long)g.nbm * g.nbn)), st);
    numc_prof_end(ctx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return numc_hip_fail(e, "direct gemm launch");
    return 0;
  }

  int nRTa, nRTb, nKF;
  long long pla, plb;
  plane_geometry(c, p.M, p.K, kf, c.cfg.bm, &nRTa, &nKF, &pla);
  plane_geometry(c, p.N, p.K, kf, c.cfg.bn, &nRTb, &nKF, &plb);
  // strided batch: slice z's packed planes follow slice z-1's
  const int nz = p.batch > 1 ? p.batch : 1;
  if (nz > 1 && (c.cfg.split_k > 1 || family == NUMC_FAMILY_FP8 || nz > 65535))
    return NUMC_SET_ERROR(NUMC_ERR_VALUE, "matmul: a strided batch needs a configuration without split-K");
  unsigned char *wsb = (unsigned char *)ws;
  unsigned char *Ap = wsb;
  unsigned char *Bp = Ap + pla * planes * nz;
  // FP8: the block scales (a byte per row of each packed fragment) follow the
  // planes; then the split-K partial sums
  unsigned char *Asc = Bp + plb * planes * nz;
  unsigned char *Bsc = Asc + (family == NUMC_FAMILY_FP8 ? pla / 32 : 0);
  unsigned char *after = Bsc + (family == NUMC_FAMILY_FP8 ? plb / 32 : 0);
  after = (unsigned char *)(((uintptr_t)after + 255) & ~(uintptr_t)255);
  void *partial = c.cfg.split_k > 1 ? (void *)after : nullptr;
  const size_t ein = numc_dtype_size(p.in_dt), eout = numc_dtype_size(p.out_dt);

  // aux: the digit plane masks (NUMC_FLAG_INTS ints per operand)
  int *flags = (int *)aux;

  PackArgs pk{};
  // one pack launch for both operands where the batched walks cover the
  // family.  A single int8 digit plane needs no plane masks (they only let
  // the GEMM skip empty plane products), so that case sets none.
  const bool pair = nz == 1 && (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_SPLIT_BF16 ||
                                family == NUMC_FAMILY_FP8 || family == NUMC_FAMILY_BF16 || family == NUMC_FAMILY_F16 ||
                                (family == NUMC_FAMILY_I8 && planes == 1 && p.in_dt == NUMC_DTYPE_INT8));
  // (the plane masks matter only with several digit planes)
  const bool masks = family == NUMC_FAMILY_I8 && !(planes == 1 && p.in_dt == NUMC_DTYPE_INT8);
  // A read in place by the GEMM (no pack of A): FP32 and single-plane int8,
  // A's rows k-contiguous and 16-byte aligned, K a multiple of the k-chunk
  // (every copy stays inside its row)
  // (a strided batch: every slice 16-byte aligned too)
  const bool inplace_ok = c.inplace[0][0] && !masks && p.K % c.cfg.kc == 0 &&
                          (family == NUMC_FAMILY_F32 || family == NUMC_FAMILY_I8);
  const bool adir = inplace_ok && p.csa == 1 && ((uintptr_t)p.A & 15) == 0 &&
                    ((long long)p.rsa * (long long)ein) % 16 == 0 &&
                    (nz == 1 || ((long long)p.a_bs * (long long)ein) % 16 == 0);
  // B likewise when the rows of B^T are k-contiguous (B stored transposed):
  // FP32 only
  const bool bdir = inplace_ok && c.inplace[1][0] && p.rsb == 1 && ((uintptr_t)p.B & 15) == 0 &&
                    ((long long)p.csb * (long long)ein) % 16 == 0 &&
                    (nz == 1 || ((long long)p.b_bs * (long long)ein) % 16 == 0);
  if (masks)
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(2 * NUMC_FLAG_INTS)), dim3(256), 0, st, (uint32_t *)flags,
                       (long long)(2 * NUMC_FLAG_INTS), 0u);

  // A: rows M, element (i, k) at A[i*rsa + k*csa]
  pk.src = p.A; pk.rs = p.rsa; pk.ks = p.csa; pk.R = p.M; pk.K = p.K;
  pk.nRT = nRTa; pk.nKF = nKF; pk.dst = Ap; pk.plane_bytes = pla;
  pk.sc = Asc; pk.flags = masks ? flags : nullptr;
  pk.src_zs = p.a_bs * (long long)ein; pk.dst_zs = pla * planes;
  // B^T: rows N, element (j, k) at B[k*rsb + j*csb]
  PackArgs pb = pk;
  pb.src = p.B; pb.rs = p.csb; pb.ks = p.rsb; pb.R = p.N; pb.K = p.K;
  pb.nRT = nRTb; pb.nKF = nKF; pb.dst = Bp; pb.plane_bytes = plb;
  pb.sc = Bsc; pb.flags = masks ? flags + NUMC_FLAG_INTS : nullptr;
  pb.src_zs = p.b_bs * (long long)ein; pb.dst_zs = plb * planes;
  // (each pack launch is one profiled pack)
  if (adir && bdir) {
    // nothing to pack
  } else if (adir) {
    numc_prof_begin(ctx, PROF_PACK, family);
    launch_pack(family, p.in_dt, planes, pb, (unsigned)nz, st);
  } else if (bdir) {
    numc_prof_begin(ctx, PROF_PACK, family);
    launch_pack(family, p.in_dt, planes, pk, (unsigned)nz, st);
  } else if (pair) {
    numc_prof_begin(ctx, PROF_PACK, family);
    const long long na = numc_batch_pack_blocks(family, pk), nb = numc_batch_pack_blocks(family, pb);
    const dim3 grid((unsigned)(na + nb));
    if (family == NUMC_FAMILY_F32) hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_F32>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
    else if (family == NUMC_FAMILY_SPLIT_BF16) hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_SPLIT_BF16>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
    else if (family == NUMC_FAMILY_FP8) hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_FP8>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
    else if (family == NUMC_FAMILY_BF16) hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_BF16>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
    else if (family == NUMC_FAMILY_F16) hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_F16>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
    else hipLaunchKernelGGL((k_pack_pair<NUMC_FAMILY_I8>), grid, dim3(256), 0, st, pk, pb, (unsigned)na);
  } else if (family == NUMC_FAMILY_I8 && nz == 1) {
    // wider integers (several digit planes): both operands in one launch too
    numc_prof_begin(ctx, PROF_PACK, family);
    switch (p.in_dt) {
    case NUMC_DTYPE_UINT8: launch_pair_digits<uint8_t>(planes, pk, pb, st); break;
    case NUMC_DTYPE_INT16: launch_pair_digits<int16_t>(planes, pk, pb, st); break;
    case NUMC_DTYPE_UINT16: launch_pair_digits<uint16_t>(planes, pk, pb, st); break;
    case NUMC_DTYPE_INT32: launch_pair_digits<int32_t>(planes, pk, pb, st); break;
    case NUMC_DTYPE_UINT32: launch_pair_digits<uint32_t>(planes, pk, pb, st); break;
    default: launch_pair_digits<int8_t>(planes, pk, pb, st); break;
    }
  } else {
    numc_prof_begin(ctx, PROF_PACK, family);
    launch_pack(family, p.in_dt, planes, pk, (unsigned)nz, st);
    numc_prof_end(ctx);
    numc_prof_begin(ctx, PROF_PACK, family);
    launch_pack(family, p.in_dt, planes, pb, (unsigned)nz, st);
  }
  numc_prof_end(ctx);

  GemmArgs g{};
  g.A = Ap;
  g.B = Bp;
  g.a_plane = pla;
  g.b_plane = plb;
  g.nKF = nKF;
  g.nKC = (int)(cdiv(p.K, c.cfg.kc));
  g.M = p.M;
  g.N = p.N;
  g.nbm = (int)cdiv(p.M, c.cfg.bm);
  g.nbn = (int)cdiv(p.N, c.cfg.bn);
  g.ksplit = c.cfg.split_k;
  g.C = p.C;
  g.rso = p.rso;
  g.cso = p.cso;
  g.out_bytes = (int)numc_dtype_size(p.out_dt);
  g.a_sc = Asc;
  g.b_sc = Bsc;
  g.partial = partial;
  g.nprod = 1;
  g.pa[0] = g.pb[0] = 0;
  g.bias = p.bias;
  g.bias_s = p.bias_s;
  g.a_zs = adir ? p.a_bs * (long long)ein : pla * planes;
  g.b_zs = bdir ? p.b_bs * (long long)ein : plb * planes;
  g.c_zs = p.c_bs * (long long)eout;
  g.Araw = (const unsigned char *)p.A;
  g.a_rs = (long long)p.rsa * (long long)ein;
  g.Braw = (const unsigned char *)p.B;
  g.b_rs = (long long)p.csb * (long long)ein;
  dim3 grid((unsigned)(g.nbm * g.nbn), (unsigned)g.ksplit, (unsigned)nz);
  const bool bias = p.bias != nullptr;
  // integer digit planes: one launch runs every pass (the tile shifts its
  // accumulators between passes), so the empty passes cost no launches
  g.passes = family == NUMC_FAMILY_I8 ? planes : 1;
  g.shift_pass = 0;
  g.plane_flags = masks ? flags : nullptr;
  numc_prof_begin(ctx, PROF_GEMM, family);
  GemmLaunchFn fn = c.launch;
  if (adir || bdir) fn = c.inplace[adir && bdir ? 2 : (bdir ? 1 : 0)][bias ? 1 : 0];
  else if (bias && c.launch_bias) fn = c.launch_bias;
  fn(g, grid, st);
  numc_prof_end(ctx);
  if (g.ksplit > 1) {
    unsigned nb = blocks_for((long long)p.M * p.N);
    switch (family) {
    case NUMC_FAMILY_F64:
      if (bias) hipLaunchKernelGGL((k_splitk_reduce<double, EPI_F64B>), dim3(nb), dim3(256), 0, st, g);
      else hipLaunchKernelGGL((k_splitk_reduce<double, EPI_F64>), dim3(nb), dim3(256), 0, st, g);
      break;
    case NUMC_FAMILY_FP8: hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F32>), dim3(nb), dim3(256), 0, st, g); break;
    case NUMC_FAMILY_BF16:
      if (bias) hipLaunchKernelGGL((k_splitk_reduce<float, EPI_BF16B>), dim3(nb), dim3(256), 0, st, g);
      else hipLaunchKernelGGL((k_splitk_reduce<float, EPI_BF16>), dim3(nb), dim3(256), 0, st, g);
      break;
    case NUMC_FAMILY_F16:
      if (bias) hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F16B>), dim3(nb), dim3(256), 0, st, g);
      else hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F16>), dim3(nb), dim3(256), 0, st, g);
      break;
    case NUMC_FAMILY_I8:
      if (bias) hipLaunchKernelGGL((k_splitk_reduce<int, EPI_INTB>), dim3(nb), dim3(256), 0, st, g);
      else hipLaunchKernelGGL((k_splitk_reduce<int, EPI_INT>), dim3(nb), dim3(256), 0, st, g);
      break;
    default:
      if (bias) hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F32B>), dim3(nb), dim3(256), 0, st, g);
      else hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F32>), dim3(nb), dim3(256), 0, st, g);
      break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "gemm launch");
  return 0;
}

int numc_naive_launch(NumcCtx *ctx, NumcDType dt, const void *a, const void *b, void *c, int M, int K, int N) {
  long long n = (long long)M * N;
  if (n == 0) return 0;
  dim3 grid(blocks_for(n)), blk(256);
  hipStream_t s = ctx->stream;
  switch (dt) {
  case NUMC_DTYPE_INT8: hipLaunchKernelGGL((k_naive<int8_t, false>), grid, blk, 0, s, (const int8_t *)a, (const int8_t *)b, (int8_t *)c, M, K, N); break;
  case NUMC_DTYPE_INT16: hipLaunchKernelGGL((k_naive<int16_t, false>), grid, blk, 0, s, (const int16_t *)a, (const int16_t *)b, (int16_t *)c, M, K, N); break;
  case NUMC_DTYPE_INT32: hipLaunchKernelGGL((k_naive<int32_t, false>), grid, blk, 0, s, (const int32_t *)a, (const int32_t *)b, (int32_t *)c, M, K, N); break;
  case NUMC_DTYPE_INT64: hipLaunchKernelGGL((k_naive<int64_t, false>), grid, blk, 0, s, (const int64_t *)a, (const int64_t *)b, (int64_t *)c, M, K, N); break;
  case NUMC_DTYPE_UINT8: hipLaunchKernelGGL((k_naive<uint8_t, false>), grid, blk, 0, s, (const uint8_t *)a, (const uint8_t *)b, (uint8_t *)c, M, K, N); break;
  case NUMC_DTYPE_UINT16: hipLaunchKernelGGL((k_naive<uint16_t, false>), grid, blk, 0, s, (const uint16_t *)a, (const uint16_t *)b, (uint16_t *)c, M, K, N); break;
  case NUMC_DTYPE_UINT32: hipLaunchKernelGGL((k_naive<uint32_t, false>), grid, blk, 0, s, (const uint32_t *)a, (const uint32_t *)b, (uint32_t *)c, M, K, N); break;
  case NUMC_DTYPE_UINT64: hipLaunchKernelGGL((k_naive<uint64_t, false>), grid, blk, 0, s, (const uint64_t *)a, (const uint64_t *)b, (uint64_t *)c, M, K, N); break;
  case NUMC_DTYPE_FLOAT32: hipLaunchKernelGGL((k_naive<float, true>), grid, blk, 0, s, (const float *)a, (const float *)b, (float *)c, M, K, N); break;
  case NUMC_DTYPE_FLOAT64: hipLaunchKernelGGL((k_naive<double, true>), grid, blk, 0, s, (const double *)a, (const double *)b, (double *)c, M, K, N); break;
  case NUMC_DTYPE_BFLOAT16: hipLaunchKernelGGL((k_naive_h16<true>), grid, blk, 0, s, (const uint16_t *)a, (const uint16_t *)b, (uint16_t *)c, M, K, N); break;
  case NUMC_DTYPE_FLOAT16: hipLaunchKernelGGL((k_naive_h16<false>), grid, blk, 0, s, (const uint16_t *)a, (const uint16_t *)b, (uint16_t *)c, M, K, N); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "naive launch");
  return 0;
}

// 16-bit float elements for k_scale (bits; the scalar comes as binary32)
struct Bf16Elem { uint16_t b; };
struct F16Elem { uint16_t b; };

template <class T> __device__ inline T scale_one(T x, uint64_t bits) {
  if constexpr (std::is_same_v<T, Bf16Elem>) {
    return Bf16Elem{(uint16_t)f32_to_bf16_bits(bf16_bits_to_f32(x.b) * __uint_as_float((uint32_t)bits))};
  } else if constexpr (std::is_same_v<T, F16Elem>) {
    return F16Elem{(uint16_t)f32_to_f16_bits(f16_bits_to_f32(x.b) * __uint_as_float((uint32_t)bits))};
  } else if constexpr (std::is_integral_v<T> && sizeof(T) == 8) {
    return (T)((uint64_t)x * bits); // low 64 bits
  } else if constexpr (std::is_integral_v<T>) {
    return (T)(uint32_t)((uint32_t)(int32_t)x * (uint32_t)bits); // low w bits
  } else if constexpr (sizeof(T) == 4) {
    return x * __uint_as_float((uint32_t)bits);
  } else {
    return x * __longlong_as_double((long long)bits);
  }
}

template <class T>
__global__ void k_scale(ScaleArgs a) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < a.n; i += (long long)gridDim.x * 256) {
    long long rem = i, so = 0, dof = 0;
    for (int d = a.ndim - 1; d >= 0; d--) {
      const long long q = rem / a.shape[d], r = rem - q * a.shape[d];
      so += r * a.ss[d];
      dof += r * a.ds[d];
      rem = q;
    }
    ((T *)a.dst)[dof] = scale_one<T>(((const T *)a.src)[so], a.scalar_bits);
  }
}

int numc_scale_launch(NumcCtx *ctx, const ScaleArgs &a) {
  if (a.n == 0) return 0;
  const unsigned nb = (unsigned)std::min<long long>(cdiv(a.n, 256), 65536);
  hipStream_t s = ctx->stream;
  switch (a.dtype) {
  case NUMC_DTYPE_INT8: hipLaunchKernelGGL(k_scale<int8_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_INT16: hipLaunchKernelGGL(k_scale<int16_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_INT32: hipLaunchKernelGGL(k_scale<int32_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_INT64: hipLaunchKernelGGL(k_scale<int64_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_UINT8: hipLaunchKernelGGL(k_scale<uint8_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_UINT16: hipLaunchKernelGGL(k_scale<uint16_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_UINT32: hipLaunchKernelGGL(k_scale<uint32_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_UINT64: hipLaunchKernelGGL(k_scale<uint64_t>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_FLOAT32: hipLaunchKernelGGL(k_scale<float>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_BFLOAT16: hipLaunchKernelGGL(k_scale<Bf16Elem>, dim3(nb), dim3(256), 0, s, a); break;
  case NUMC_DTYPE_FLOAT16: hipLaunchKernelGGL(k_scale<F16Elem>, dim3(nb), dim3(256), 0, s, a); break;
  default: hipLaunchKernelGGL(k_scale<double>, dim3(nb), dim3(256), 0, s, a); break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "scale launch");
  return 0;
}

int numc_i64_launch(NumcCtx *ctx, const GemmProblem &p) {
  if (p.M == 0 || p.N == 0) return 0;
  dim3 grid((unsigned)cdiv(p.N, 16), (unsigned)cdiv(p.M, 16));
  if (p.in_dt == NUMC_DTYPE_INT64)
    hipLaunchKernelGGL((k_gemm_i64<int64_t>), grid, dim3(256), 0, ctx->stream, (const int64_t *)p.A,
                       (const int64_t *)p.B, (int64_t *)p.C, p.M, p.K, p.N, p.rsa, p.csa, p.rsb, p.csb, p.rso, p.cso,
                       (const int64_t *)p.bias, p.bias_s);
  else
    hipLaunchKernelGGL((k_gemm_i64<uint64_t>), grid, dim3(256), 0, ctx->stream, (const uint64_t *)p.A,
                       (const uint64_t *)p.B, (uint64_t *)p.C, p.M, p.K, p.N, p.rsa, p.csa, p.rsb, p.csb, p.rso, p.cso,
                       (const uint64_t *)p.bias, p.bias_s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "i64 launch");
  return 0;
}

int numc_batch_launch(int family, const PackArgs *descs, const long long *prefix, const int *owner,
                      long long units, const BatchProblem *probs, const int4 *items, int nitems, int *counter,
                      int grid, hipStream_t st, NumcCtx *prof) {
  // the queue counters (the ints before the queue starts)
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (uint32_t *)counter, (long long)NUMC_BATCH_QSTART, 0u);
  const unsigned nb = (unsigned)units; // blocks (numc_batch_pack_blocks summed)
  if (nb) {
    if (prof) numc_prof_begin(prof, PROF_PACK, family);
    if (family == NUMC_FAMILY_F32) hipLaunchKernelGGL((k_pack_batched<NUMC_FAMILY_F32>), dim3(nb), dim3(256), 0, st, descs, prefix, owner, units);
    else if (family == NUMC_FAMILY_SPLIT_BF16) hipLaunchKernelGGL((k_pack_batched<NUMC_FAMILY_SPLIT_BF16>), dim3(nb), dim3(256), 0, st, descs, prefix, owner, units);
    else hipLaunchKernelGGL((k_pack_batched<NUMC_FAMILY_I8>), dim3(nb), dim3(256), 0, st, descs, prefix, owner, units);
    if (prof) numc_prof_end(prof);
  }
  if (prof) numc_prof_begin(prof, PROF_BATCHED, family);
  if (family == NUMC_FAMILY_F32)
    hipLaunchKernelGGL((k_gemm_batched<MfF32, MODE_SINGLE, EPI_F32, 128, 128, 16, 3, 64, 64, 32, 3, 3>), dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  else if (family == NUMC_FAMILY_SPLIT_BF16)
    hipLaunchKernelGGL((k_gemm_batched<MfBF16, MODE_SPLIT3, EPI_F32, 128, 128, 32, 2, 64, 64, 32, 4, 2>), dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  // int8: the 128 x 128 class in k-chunks of 32 with 6 stages, 3 blocks per CU.  Its
  // tiles average 4 k-chunks of 64; the deeper pipeline took the 4096-problem batch's
  // GEMM from 0.63 to 0.58 ms (k-chunks of 64 with 2 stages at 4 blocks per CU before)
  else
    hipLaunchKernelGGL((k_gemm_batched<MfI8, MODE_SINGLE, EPI_INT, 128, 128, 32, 6, 64, 64, 64, 4, 3>), dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  if (prof) numc_prof_end(prof);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "batched launch");
  return 0;
}
