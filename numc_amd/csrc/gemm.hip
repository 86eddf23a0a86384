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

// ===========================================================================
// number formats (device)
// ===========================================================================

__device__ inline uint32_t bf16_rne_bits(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0x7FFFFFu) ? 0x40u : 0u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

// binary32 -> OCP e4m3fn, round to nearest even, saturating to +-448
__device__ inline uint32_t e4m3_rne(float x) {
  uint32_t u = __float_as_uint(x);
  uint32_t sign = (u >> 24) & 0x80u;
  uint32_t a = u & 0x7FFFFFFFu;
  if (a > 0x7F800000u) return sign | 0x7Fu;
  if (a >= 0x43E00000u) return sign | 0x7Eu;
  int e = (int)(a >> 23) - 127;
  uint32_t code;
  if (e < -6) {
    code = (uint32_t)__builtin_rintf(__uint_as_float(a) * 512.0f);
  } else {
    uint32_t m = a & 0x7FFFFFu, keep = m >> 20, rem = m & 0xFFFFFu;
    if (rem > 0x80000u || (rem == 0x80000u && (keep & 1u))) keep++;
    if (keep == 8u) { keep = 0; e++; }
    code = ((uint32_t)(e + 7) << 3) | keep;
    if (code > 0x7Eu) code = 0x7Eu;
  }
  return sign | code;
}

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

// out[i*rso + j*cso] = 0 for an M x N strided view (element size es)
__global__ void k_zero_strided(unsigned char *c, int M, int N, long long rso, long long cso, int es) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)M * N) return;
  long long i = t / N, j = t % N;
  unsigned char *p = c + (i * rso + j * cso) * es;
  for (int b = 0; b < es; b++) p[b] = 0;
}

// ===========================================================================
// pack kernels: one thread writes one 16-byte lane slot of a fragment in
// every plane.  Thread t -> (rt, kf, lane) = (t / 64 / nKF, t / 64 % nKF, t % 64),
// so the destination offset of plane 0 is t*16.
// ===========================================================================

template <class S> __device__ inline S ld(const void *p, long long i) { return ((const S *)p)[i]; }

// f32 -> f32 (T=32, 4 per lane, KF=8)
__global__ void k_pack_f32(PackArgs a) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)a.nRT * a.nKF * 64) return;
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
__global__ void k_pack_split_bf16(PackArgs a) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)a.nRT * a.nKF * 64) return;
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
    uint32_t h0 = bf16_rne_bits(x0), h1 = bf16_rne_bits(x1);
    uint32_t l0 = bf16_rne_bits(x0 - __uint_as_float(h0 << 16));
    uint32_t l1 = bf16_rne_bits(x1 - __uint_as_float(h1 << 16));
    hi[e / 2] = (h0 & 0xFFFFu) | (h1 << 16);
    lo[e / 2] = (l0 & 0xFFFFu) | (l1 << 16);
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
  *(uint4 *)(a.dst + a.plane_bytes + t * 16) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
}

// f32 * 2^row_exp[row] -> e4m3  (T=32, 16 per lane, KF=32)
__global__ void k_pack_fp8(PackArgs a) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)a.nRT * a.nKF * 64) return;
  int lane = (int)(t & 63);
  long long fr = t >> 6;
  int kf = (int)(fr % a.nKF);
  long long row = (fr / a.nKF) * 32 + (lane & 31);
  int k0 = kf * 32 + (lane >> 5) * 16;
  int ex = row < a.R ? a.row_exp[row] : 0;
  uint32_t w[4] = {0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < 16; e++) {
    int k = k0 + e;
    float x = (row < a.R && k < a.K) ? ld<float>(a.src, row * a.rs + (long long)k * a.ks) : 0.f;
    w[e >> 2] |= e4m3_rne(ldexpf(x, ex)) << (8 * (e & 3));
  }
  *(uint4 *)(a.dst + t * 16) = make_uint4(w[0], w[1], w[2], w[3]);
}

// f64 -> f64 (T=16, 2 per lane, KF=8)
__global__ void k_pack_f64(PackArgs a) {
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
// x = sum_p d_p 256^p (mod 256^P) with d_p in [-128, 127]; flags[0] gets a
// bit per plane that holds a nonzero digit.
template <class S, int P>
__global__ void k_pack_digits(PackArgs a) {
  long long t = blockIdx.x * 256LL + threadIdx.x;
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
      long long x = (row < a.R && k < a.K) ? (long long)ld<S>(a.src, row * a.rs + (long long)k * a.ks) : 0;
      unsigned long long u = (unsigned long long)x;
#pragma unroll
      for (int p = 0; p < P; p++) {
        int8_t d = (int8_t)(u & 0xFFu);
        u = (u - (unsigned long long)(long long)d) >> 8;
        w[p][e >> 2] |= (uint32_t)(uint8_t)d << (8 * (e & 3));
      }
    }
#pragma unroll
    for (int p = 0; p < P; p++)
      *(uint4 *)(a.dst + p * a.plane_bytes + t * 16) = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
  }
  int mask = 0;
#pragma unroll
  for (int p = 0; p < P; p++)
    if (__ballot((w[p][0] | w[p][1] | w[p][2] | w[p][3]) != 0)) mask |= 1 << p;
  if ((threadIdx.x & 63) == 0 && mask) atomicOr(a.flags, mask);
}

// ---- FP8 scales: row maxima of |x| as float bits, then exponents ----------

// one block per row (rows with unit k-stride read coalesced)
__global__ void k_rowmax_rowwise(const float *src, long long rs, long long ks, int R, int K, uint32_t *bits) {
  int row = blockIdx.x;
  float m = 0.f;
  for (int k = threadIdx.x; k < K; k += 256) m = fmaxf(m, fabsf(src[row * rs + (long long)k * ks]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) bits[row] = __float_as_uint(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3])));
}

// 256 consecutive rows per block, a k-slice per blockIdx.y (rows with unit
// row-stride read coalesced); combined with atomicMax on the float bits
__global__ void k_rowmax_colwise(const float *src, long long rs, long long ks, int R, int K, int kslice,
                                 uint32_t *bits) {
  int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= R) return;
  int k0 = blockIdx.y * kslice, k1 = min(K, k0 + kslice);
  float m = 0.f;
  for (int k = k0; k < k1; k++) m = fmaxf(m, fabsf(src[row * rs + (long long)k * ks]));
  atomicMax(&bits[row], __float_as_uint(m));
}

__global__ void k_exps(const uint32_t *bits, int *ex, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ex[i] = fp8_scale_exp(__uint_as_float(bits[i]));
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

template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int EPI>
__global__ void __launch_bounds__(WGM *WGN * 64, 1) k_gemm(const GemmArgs gin) {
  const GemmArgs g = gin; // local copy: lambdas below capture it by reference
  constexpr int T = MF::T, KF = MF::KF, FR = MF::FR;
  constexpr int NW = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / T, TN = WN / T;
  constexpr int KFC = KC / KF;
  constexpr int P = (MODE == MODE_SPLIT3) ? 2 : 1;
  constexpr int RA = BM / T, RB = BN / T;
  constexpr int A_FR = P * RA * KFC, B_FR = P * RB * KFC, FRC = A_FR + B_FR;
  constexpr int LA = A_FR / NW, LB = B_FR / NW;
  static_assert(A_FR % NW == 0 && B_FR % NW == 0, "fragments must split evenly over waves");
  static_assert(KFC % FR == 0 && TM >= 1 && TN >= 1, "bad tile");
  using Acc = typename MF::Acc;
  using Raw = typename MF::Raw;

  __shared__ u32x4 lds[2 * FRC * 64];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;

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

  // plane products: product q uses plane (pa0 + q) of A and (pb0 - q) of B
  int nprod = g.nprod, pa0 = g.pa[0], pb0 = g.pb[0];
  if (g.plane_flags) {
    nprod = int_products(g.shift_pass, g.plane_flags[0], g.plane_flags[1], &pa0, &pb0);
    if (nprod == 0 && g.shift_pass > 0) return; // this digit pass adds nothing
  }
  const int V = nprod * g.nKC;
  const int c0 = (int)((long long)blockIdx.y * V / g.ksplit);
  const int c1 = (int)((long long)(blockIdx.y + 1) * V / g.ksplit);

  long long offA[LA], offB[LB];
#pragma unroll
  for (int i = 0; i < LA; i++) {
    int f = wave + i * NW;
    int p = f / (RA * KFC), rt = (f / KFC) % RA, kf = f % KFC;
    offA[i] = p * g.a_plane + ((long long)(bm * RA + rt) * g.nKF + kf) * NUMC_FRAG_BYTES + lane * 16;
  }
#pragma unroll
  for (int i = 0; i < LB; i++) {
    int f = wave + i * NW;
    int p = f / (RB * KFC), rt = (f / KFC) % RB, kf = f % KFC;
    offB[i] = p * g.b_plane + ((long long)(bn * RB + rt) * g.nKF + kf) * NUMC_FRAG_BYTES + lane * 16;
  }

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int v = 0; v < MF::NACC; v++) acc[i][j][v] = 0;

  u32x4 ra[LA], rb[LB];
  auto gload = [&](int c) __attribute__((always_inline)) {
    int q = c / g.nKC, kc = c - q * g.nKC;
    long long ko = (long long)kc * KFC * NUMC_FRAG_BYTES;
    const unsigned char *pA = g.A + ko + (long long)(pa0 + q) * g.a_plane;
    const unsigned char *pB = g.B + ko + (long long)(pb0 - q) * g.b_plane;
#pragma unroll
    for (int i = 0; i < LA; i++) ra[i] = *(const u32x4 *)(pA + offA[i]);
#pragma unroll
    for (int i = 0; i < LB; i++) rb[i] = *(const u32x4 *)(pB + offB[i]);
  };
  auto sstore = [&](int s) __attribute__((always_inline)) {
    u32x4 *base = lds + s * FRC * 64;
#pragma unroll
    for (int i = 0; i < LA; i++) base[(wave + i * NW) * 64 + lane] = ra[i];
#pragma unroll
    for (int i = 0; i < LB; i++) base[(A_FR + wave + i * NW) * 64 + lane] = rb[i];
  };
  auto compute = [&](int s) __attribute__((always_inline)) {
    const u32x4 *base = lds + s * FRC * 64;
#pragma unroll
    for (int kf = 0; kf < KFC; kf += FR) {
      u32x4 af[P][TM][FR], bfr[P][TN][FR];
#pragma unroll
      for (int p = 0; p < P; p++)
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int r = 0; r < FR; r++)
            af[p][i][r] = base[((p * RA + wm * TM + i) * KFC + kf + r) * 64 + lane];
#pragma unroll
      for (int p = 0; p < P; p++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
          for (int r = 0; r < FR; r++)
            bfr[p][j][r] = base[(A_FR + (p * RB + wn * TN + j) * KFC + kf + r) * 64 + lane];
      if constexpr (MODE == MODE_SPLIT3) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++) {
            MF::mma(acc[i][j], af[1][i], bfr[0][j]);
            MF::mma(acc[i][j], af[0][i], bfr[1][j]);
            MF::mma(acc[i][j], af[0][i], bfr[0][j]);
          }
      } else {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++) MF::mma(acc[i][j], af[0][i], bfr[0][j]);
      }
    }
  };

  if (c0 < c1) {
    gload(c0);
    sstore(0);
    __syncthreads();
  }
  for (int c = c0; c < c1; c++) {
    const int s = (c - c0) & 1;
    const bool more = c + 1 < c1;
    if (more) gload(c + 1);
    compute(s);
    if (more) sstore(s ^ 1);
    __syncthreads();
  }

  // epilogue
  const int row0 = bm * BM + wm * WM, col0 = bn * BN + wn * WN;
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int v = 0; v < MF::NACC; v++) {
        const int row = row0 + i * T + MF::row(v, lane), col = col0 + j * T + MF::col(lane);
        if (row < g.M && col < g.N) {
          if (g.ksplit > 1)
            ((Raw *)g.partial)[((long long)blockIdx.y * g.M + row) * g.N + col] = (Raw)acc[i][j][v];
          else
            epi_store<EPI, Raw>(g, row, col, (Raw)acc[i][j][v]);
        }
      }
}

// split-K: sum the slices in slice order, then the final epilogue
template <class Raw, int EPI>
__global__ void k_splitk_reduce(const GemmArgs g) {
  if (g.plane_flags && g.shift_pass > 0) {
    int pa0, pb0;
    if (int_products(g.shift_pass, g.plane_flags[0], g.plane_flags[1], &pa0, &pb0) == 0) return;
  }
  long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= (long long)g.M * g.N) return;
  const Raw *p = (const Raw *)g.partial;
  Raw s = p[t];
  for (int z = 1; z < g.ksplit; z++) s += p[(long long)z * g.M * g.N + t];
  epi_store<EPI, Raw>(g, (int)(t / g.N), (int)(t % g.N), s);
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

// 64-bit integers: no MFMA; 16x16 output tile per block, k staged in LDS.
template <class T>
__global__ void k_gemm_i64(const T *a, const T *b, T *c, int M, int K, int N, long long rsa, long long csa,
                           long long rsb, long long csb, long long rso, long long cso) {
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
  if (row < M && col < N) c[row * rso + col * cso] = (T)acc;
}

// ===========================================================================
// configuration tables
// ===========================================================================

template <class MF, int BM, int BN, int KC, int WGM, int WGN, int MODE, int EPI>
static void launch_gemm(const GemmArgs &g, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((k_gemm<MF, BM, BN, KC, WGM, WGN, MODE, EPI>), grid, dim3(WGM * WGN * 64), 0, s, g);
}

#define CFG(MF, BM, BN, KC, WGM, WGN, MODE, EPI, SPLIT) \
  {{BM, BN, KC, WGM, WGN, SPLIT}, MF::T, &launch_gemm<MF, BM, BN, KC, WGM, WGN, MODE, EPI>}

static const GemmConfig kF32[] = {
    CFG(MfF32, 256, 256, 16, 2, 2, MODE_SINGLE, EPI_F32, 1),
    CFG(MfF32, 128, 128, 32, 2, 2, MODE_SINGLE, EPI_F32, 1),
    CFG(MfF32, 256, 128, 16, 2, 2, MODE_SINGLE, EPI_F32, 1),
    CFG(MfF32, 128, 256, 16, 2, 2, MODE_SINGLE, EPI_F32, 1),
    CFG(MfF32, 64, 64, 32, 2, 2, MODE_SINGLE, EPI_F32, 1),
    CFG(MfF32, 128, 128, 32, 2, 2, MODE_SINGLE, EPI_F32, 4),
    CFG(MfF32, 64, 64, 32, 2, 2, MODE_SINGLE, EPI_F32, 8),
};
static const GemmConfig kSplit[] = {
    CFG(MfBF16, 256, 256, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1),
    CFG(MfBF16, 128, 128, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1),
    CFG(MfBF16, 256, 128, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1),
    CFG(MfBF16, 64, 64, 32, 2, 2, MODE_SPLIT3, EPI_F32, 1),
    CFG(MfBF16, 128, 128, 32, 2, 2, MODE_SPLIT3, EPI_F32, 4),
    CFG(MfBF16, 64, 64, 32, 2, 2, MODE_SPLIT3, EPI_F32, 8),
};
static const GemmConfig kFP8[] = {
    CFG(MfFP8, 256, 128, 64, 2, 2, MODE_SINGLE, EPI_FP8, 1),
    CFG(MfFP8, 128, 128, 128, 2, 2, MODE_SINGLE, EPI_FP8, 1),
    CFG(MfFP8, 64, 64, 128, 2, 2, MODE_SINGLE, EPI_FP8, 1),
    CFG(MfFP8, 128, 128, 128, 2, 2, MODE_SINGLE, EPI_FP8, 4),
};
static const GemmConfig kF64[] = {
    CFG(MfF64, 128, 128, 16, 2, 2, MODE_SINGLE, EPI_F64, 1),
    CFG(MfF64, 64, 64, 32, 2, 2, MODE_SINGLE, EPI_F64, 1),
    CFG(MfF64, 128, 64, 16, 2, 2, MODE_SINGLE, EPI_F64, 1),
    CFG(MfF64, 64, 64, 32, 2, 2, MODE_SINGLE, EPI_F64, 8),
};
static const GemmConfig kI8[] = {
    CFG(MfI8, 128, 128, 64, 2, 2, MODE_SINGLE, EPI_INT, 1),
    CFG(MfI8, 256, 128, 64, 2, 2, MODE_SINGLE, EPI_INT, 1),
    CFG(MfI8, 64, 64, 128, 2, 2, MODE_SINGLE, EPI_INT, 1),
    CFG(MfI8, 128, 128, 64, 2, 2, MODE_SINGLE, EPI_INT, 4),
};
#undef CFG

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
};

const GemmConfig *numc_gemm_config(int family, int idx) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return nullptr;
  if (idx < 0 || idx >= kFamilies[family].n) return nullptr;
  return &kFamilies[family].cfg[idx];
}

int numc_gemm_config_count(int family) {
  if (family < 0 || family >= NUMC_FAMILY_COUNT) return 0;
  return kFamilies[family].n;
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

static inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }
static inline unsigned blocks_for(long long n) { return (unsigned)cdiv(n, 256); }

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
  default: return MfI8::KF;
  }
}

static int digit_planes(NumcDType out) {
  size_t s = numc_dtype_size(out);
  return s >= 4 ? 4 : (int)s;
}

template <class S>
static void launch_digits(int planes, PackArgs &pk, unsigned nb, hipStream_t st) {
  switch (planes) {
  case 1: hipLaunchKernelGGL((k_pack_digits<S, 1>), dim3(nb), dim3(256), 0, st, pk); break;
  case 2: hipLaunchKernelGGL((k_pack_digits<S, 2>), dim3(nb), dim3(256), 0, st, pk); break;
  default: hipLaunchKernelGGL((k_pack_digits<S, 4>), dim3(nb), dim3(256), 0, st, pk); break;
  }
}

static void launch_pack(int family, NumcDType in_dt, int planes, PackArgs &pk, hipStream_t st) {
  unsigned nb = blocks_for((long long)pk.nRT * pk.nKF * 64);
  switch (family) {
  case NUMC_FAMILY_F32: hipLaunchKernelGGL(k_pack_f32, dim3(nb), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_SPLIT_BF16: hipLaunchKernelGGL(k_pack_split_bf16, dim3(nb), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_FP8: hipLaunchKernelGGL(k_pack_fp8, dim3(nb), dim3(256), 0, st, pk); break;
  case NUMC_FAMILY_F64: hipLaunchKernelGGL(k_pack_f64, dim3(nb), dim3(256), 0, st, pk); break;
  default:
    switch (in_dt) {
    case NUMC_DTYPE_INT8: launch_digits<int8_t>(planes, pk, nb, st); break;
    case NUMC_DTYPE_UINT8: launch_digits<uint8_t>(planes, pk, nb, st); break;
    case NUMC_DTYPE_INT16: launch_digits<int16_t>(planes, pk, nb, st); break;
    case NUMC_DTYPE_UINT16: launch_digits<uint16_t>(planes, pk, nb, st); break;
    case NUMC_DTYPE_INT32: launch_digits<int32_t>(planes, pk, nb, st); break;
    default: launch_digits<uint32_t>(planes, pk, nb, st); break;
    }
  }
}

static void launch_rowmax(const float *src, long long rs, long long ks, int R, int K, uint32_t *bits,
                          hipStream_t st) {
  if (R <= 0) return;
  if (ks == 1 || (rs != 1 && K >= 256)) {
    hipLaunchKernelGGL(k_rowmax_rowwise, dim3(R), dim3(256), 0, st, src, rs, ks, R, K, bits);
  } else {
    int kslice = 256;
    dim3 grid(blocks_for(R), (unsigned)cdiv(K > 0 ? K : 1, kslice));
    hipLaunchKernelGGL(k_rowmax_colwise, grid, dim3(256), 0, st, src, rs, ks, R, K, kslice, bits);
  }
}

size_t numc_gemm_workspace(const GemmProblem &p, const GemmConfig &c, size_t *aux_bytes) {
  int kf = family_kf(p.family), nRT, nKF;
  long long pa, pb;
  plane_geometry(c, p.M, p.K, kf, c.cfg.bm, &nRT, &nKF, &pa);
  plane_geometry(c, p.N, p.K, kf, c.cfg.bn, &nRT, &nKF, &pb);
  int planes = p.family == NUMC_FAMILY_SPLIT_BF16 ? 2 : (p.family == NUMC_FAMILY_I8 ? digit_planes(p.out_dt) : 1);
  size_t raw = (p.family == NUMC_FAMILY_F64) ? 8 : 4;
  size_t partial = c.cfg.split_k > 1 ? (size_t)c.cfg.split_k * p.M * p.N * raw : 0;
  size_t bytes = (size_t)(pa + pb) * planes + partial + 3 * 256;
  if (aux_bytes) *aux_bytes = 256 + (size_t)(p.M + p.N) * 8 + 512;
  return bytes;
}

int numc_gemm_launch(NumcCtx *ctx, const GemmProblem &p, const GemmConfig &c, void *ws, void *aux) {
  hipStream_t st = ctx->stream;
  const int family = p.family;
  const int kf = family_kf(family);
  const int planes = family == NUMC_FAMILY_SPLIT_BF16 ? 2 : (family == NUMC_FAMILY_I8 ? digit_planes(p.out_dt) : 1);

  if (p.M == 0 || p.N == 0) return 0;
  if (p.K == 0) {
    long long n = (long long)p.M * p.N;
    hipLaunchKernelGGL(k_zero_strided, dim3(blocks_for(n)), dim3(256), 0, st, (unsigned char *)p.C, p.M, p.N,
                       p.rso, p.cso, (int)numc_dtype_size(p.out_dt));
    return 0;
  }

  int nRTa, nRTb, nKF;
  long long pla, plb;
  plane_geometry(c, p.M, p.K, kf, c.cfg.bm, &nRTa, &nKF, &pla);
  plane_geometry(c, p.N, p.K, kf, c.cfg.bn, &nRTb, &nKF, &plb);
  unsigned char *wsb = (unsigned char *)ws;
  unsigned char *Ap = wsb;
  unsigned char *Bp = Ap + pla * planes;
  size_t raw = family == NUMC_FAMILY_F64 ? 8 : 4;
  void *partial = c.cfg.split_k > 1 ? (void *)(Bp + plb * planes) : nullptr;
  (void)raw;

  // aux: [0..1] digit plane masks, then FP8 row max bits / exponents
  int *flags = (int *)aux;
  uint32_t *bitsA = (uint32_t *)((unsigned char *)aux + 256);
  uint32_t *bitsB = bitsA + p.M;
  int *expA = (int *)(bitsB + p.N);
  int *expB = expA + p.M;

  PackArgs pk{};
  if (family == NUMC_FAMILY_FP8) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(p.M + p.N)), dim3(256), 0, st, bitsA, (long long)(p.M + p.N), 0u);
    launch_rowmax((const float *)p.A, p.rsa, p.csa, p.M, p.K, bitsA, st);
    launch_rowmax((const float *)p.B, p.csb, p.rsb, p.N, p.K, bitsB, st);
    hipLaunchKernelGGL(k_exps, dim3(blocks_for(p.M + p.N)), dim3(256), 0, st, bitsA, expA, p.M + p.N);
  }
  if (family == NUMC_FAMILY_I8)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, (uint32_t *)flags, 2LL, 0u);

  // A: rows M, element (i, k) at A[i*rsa + k*csa]
  pk.src = p.A; pk.rs = p.rsa; pk.ks = p.csa; pk.R = p.M; pk.K = p.K;
  pk.nRT = nRTa; pk.nKF = nKF; pk.dst = Ap; pk.plane_bytes = pla;
  pk.row_exp = expA; pk.flags = flags;
  numc_prof_begin(ctx, PROF_PACK, family);
  launch_pack(family, p.in_dt, planes, pk, st);
  // B^T: rows N, element (j, k) at B[k*rsb + j*csb]
  pk.src = p.B; pk.rs = p.csb; pk.ks = p.rsb; pk.R = p.N; pk.K = p.K;
  pk.nRT = nRTb; pk.nKF = nKF; pk.dst = Bp; pk.plane_bytes = plb;
  pk.row_exp = expB; pk.flags = flags + 1;
  launch_pack(family, p.in_dt, planes, pk, st);
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
  g.row_exp = expA;
  g.col_exp = expB;
  g.partial = partial;
  g.nprod = 1;
  g.pa[0] = g.pb[0] = 0;
  dim3 grid((unsigned)(g.nbm * g.nbn), (unsigned)g.ksplit);
  const int passes = family == NUMC_FAMILY_I8 ? planes : 1;
  for (int s = 0; s < passes; s++) {
    g.shift_pass = s;
    g.plane_flags = family == NUMC_FAMILY_I8 ? flags : nullptr;
    numc_prof_begin(ctx, PROF_GEMM, family);
    c.launch(g, grid, st);
    numc_prof_end(ctx);
    if (g.ksplit > 1) {
      unsigned nb = blocks_for((long long)p.M * p.N);
      switch (family) {
      case NUMC_FAMILY_F64: hipLaunchKernelGGL((k_splitk_reduce<double, EPI_F64>), dim3(nb), dim3(256), 0, st, g); break;
      case NUMC_FAMILY_FP8: hipLaunchKernelGGL((k_splitk_reduce<float, EPI_FP8>), dim3(nb), dim3(256), 0, st, g); break;
      case NUMC_FAMILY_I8: hipLaunchKernelGGL((k_splitk_reduce<int, EPI_INT>), dim3(nb), dim3(256), 0, st, g); break;
      default: hipLaunchKernelGGL((k_splitk_reduce<float, EPI_F32>), dim3(nb), dim3(256), 0, st, g); break;
      }
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
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "naive launch");
  return 0;
}

int numc_i64_launch(NumcCtx *ctx, const GemmProblem &p) {
  if (p.M == 0 || p.N == 0) return 0;
  dim3 grid((unsigned)cdiv(p.N, 16), (unsigned)cdiv(p.M, 16));
  if (p.in_dt == NUMC_DTYPE_INT64)
    hipLaunchKernelGGL((k_gemm_i64<int64_t>), grid, dim3(256), 0, ctx->stream, (const int64_t *)p.A,
                       (const int64_t *)p.B, (int64_t *)p.C, p.M, p.K, p.N, p.rsa, p.csa, p.rsb, p.csb, p.rso, p.cso);
  else
    hipLaunchKernelGGL((k_gemm_i64<uint64_t>), grid, dim3(256), 0, ctx->stream, (const uint64_t *)p.A,
                       (const uint64_t *)p.B, (uint64_t *)p.C, p.M, p.K, p.N, p.rsa, p.csa, p.rsb, p.csb, p.rso, p.cso);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "i64 launch");
  return 0;
}
