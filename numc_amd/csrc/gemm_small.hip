// gemm_small.hip -- the direct GEMM: one launch reads A and B where they lie
// (any strides, any element alignment), with no pack kernel, no workspace and
// no split-K reduction launch.
//
// numc's counterpart is its small-problem path: below a per-dtype M*K*N
// threshold numc_matmul calls gemmsup_<dt>_avx2 (src/matmul/matmul.c:241-256,
// 352-366; src/intrinsics/gemmsup_avx2.h:123-916), which multiplies the
// strided operands in place without packing.  Here the threshold is not a
// constant: the direct configurations sit in each family's configuration
// table beside the packed ones, and the selector (learned cost model,
// exhaustive autotune) picks between them per shape.
//
// Layout of the work.  A block owns a BM x BN output tile; its WK waves split
// K (contiguous ranges of MFMA fragments) and each wave computes the whole
// tile for its range, TM x TN MFMA tiles of T x T.  A lane loads the 16 bytes
// its MFMA operand needs straight from memory into registers, in the packed
// fragment order the other kernels read from LDS: row (lane mod T) of the
// tile and E consecutive k at (lane div T) * E.  Loads run PF fragments ahead
// of the MFMAs (a register ring).  At the end the waves' partial sums meet in
// LDS and are added in wave order (deterministic), and the store applies the
// family's epilogue (dtype, rounding, wraparound, optional bias).
//
// Edges: rows past M (columns past N) read row M-1 (N-1), whose products
// reach no stored output; k past K reads nothing and contributes zeros, in
// both operands, so no memory outside the views is touched.
#include "gemm.h"
#include "gemm_kernels.h"
#include "internal.h"

#include <type_traits>

namespace {

// a 16-byte vector type with byte alignment (global vector loads on gfx950
// take any byte alignment)
typedef unsigned int u32x4_b __attribute__((ext_vector_type(4), aligned(1)));

// a 16-byte load from an address aligned only to its elements
__device__ __attribute__((always_inline)) inline uint4 ldu16(const unsigned char *p) {
  const u32x4_b v = *NUMC_GLOBAL(const u32x4_b, p);
  return make_uint4(v.x, v.y, v.z, v.w);
}

// What one fragment lane slot holds: E source elements of type S (16 bytes of
// source per plane; the split method reads 32 bytes of float32 and makes a
// bf16 hi plane and lo plane of 16 bytes each).
template <class MF, int MODE> struct Direct;
template <> struct Direct<MfF32, MODE_SINGLE> {
  using S = float;
  static constexpr int E = 4, P = 1;
};
template <> struct Direct<MfBF16, MODE_SPLIT3> {
  using S = float;
  static constexpr int E = 8, P = 2;
};
template <> struct Direct<MfF64, MODE_SINGLE> {
  using S = double;
  static constexpr int E = 2, P = 1;
};
template <> struct Direct<MfI8, MODE_SINGLE> { // int8, or uint8 with a one-byte result (one digit plane)
  using S = uint8_t;
  static constexpr int E = 16, P = 1;
};
template <> struct Direct<MfBF16, MODE_SINGLE> {
  using S = uint16_t;
  static constexpr int E = 8, P = 1;
};
template <> struct Direct<MfF16, MODE_SINGLE> {
  using S = uint16_t;
  static constexpr int E = 8, P = 1;
};

// E elements (k0 .. k0+E-1) of one row starting at p (= the row's element k0),
// k stride ks bytes; zeros at k >= kend.  KC: the k stride is the element
// size, so a whole slot is one (or two) 16-byte loads.  FULL: every lane's
// E elements lie below kend (a whole k-fragment inside K), so the loads take
// no per-lane conditions (no divergent branches in the main loop).
template <class S, int E, bool KC, bool FULL = false>
__device__ __attribute__((always_inline)) inline void load_slot(const unsigned char *p, long long ks, int k0, int kend,
                                                                S (&v)[E]) {
  constexpr int NQ = E * (int)sizeof(S) / 16;
  if constexpr (FULL && !KC) {
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = *NUMC_GLOBAL(const S, p + (long long)e * ks);
  } else if (KC && (FULL || k0 + E <= kend)) {
    union {
      uint4 q[NQ];
      S e[E];
    } u;
#pragma unroll
    for (int q = 0; q < NQ; q++) u.q[q] = ldu16(p + 16 * q);
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = u.e[e];
  } else {
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = k0 + e < kend ? *NUMC_GLOBAL(const S, p + (long long)e * ks) : (S)0;
  }
}

// the slot's elements as MFMA operand registers (one u32x4 per plane)
template <class MF, int MODE, class S, int E>
__device__ __attribute__((always_inline)) inline void to_frag(const S (&v)[E], u32x4 *out, int stride) {
  if constexpr (MODE == MODE_SPLIT3) {
    // hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even by
    // v_cvt_pk_bf16_f32 (two elements per instruction): the packs' split
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    uint32_t hi[4], lo[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      const f32x2 x = {v[e], v[e + 1]};
      const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
      const f32x2 r = {x.x - __uint_as_float(h << 16), x.y - __uint_as_float(h & 0xFFFF0000u)};
      hi[e / 2] = h;
      lo[e / 2] = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
    }
    out[0] = u32x4{hi[0], hi[1], hi[2], hi[3]};
    out[stride] = u32x4{lo[0], lo[1], lo[2], lo[3]};
  } else {
    union {
      S e[E];
      u32x4 q;
    } u;
#pragma unroll
    for (int e = 0; e < E; e++) u.e[e] = v[e];
    out[0] = u.q;
  }
}

__global__ void k_fill_counter(int *c) {
  if (threadIdx.x == 0) *c = 0;
}

// the direct batched kernel's queue counters
__global__ void k_clear_queues(int *c) {
  if (threadIdx.x < NUMC_BATCH_QUEUES) c[threadIdx.x * NUMC_BATCH_COUNTER_STRIDE] = 0;
}

template <class Raw> __device__ __attribute__((always_inline)) inline Raw add_raw(Raw a, Raw b) {
  if constexpr (std::is_integral<Raw>::value) return (Raw)((uint32_t)a + (uint32_t)b); // mod 2^32
  else return a + b;
}

// LDS words (of the accumulator type) a direct tile's reduction takes: every
// wave's partial sums at once (SEQ false), or one wave's at a time (SEQ true)
template <class MF, int TM, int TN, int WK, bool SEQ> constexpr int direct_red_words() {
  return WK == 1 ? 1 : (SEQ ? 1 : WK) * TM * TN * MF::NACC * 64;
}

// One BM x BN output tile (bm, bn) of the direct GEMM, the whole K, computed
// by the block's WK waves.  The reduction adds the waves' partial sums in wave
// order through `red`: all at once, each wave then storing a share (SEQ
// false), or one wave at a time into wave 0, which stores (SEQ true: a
// quarter of the LDS, three more barriers).
template <class MF, int MODE, int TM, int TN, int WK, int PF, int EPI, bool AK, bool BK, bool SEQ>
__device__ __attribute__((always_inline)) inline void direct_tile(const GemmArgs &g, const int bm, const int bn,
                                                                  typename MF::Raw *red) {
  using D = Direct<MF, MODE>;
  using S = typename D::S;
  using Acc = typename MF::Acc;
  using Raw = typename MF::Raw;
  constexpr int T = MF::T, KF = MF::KF, E = D::E, P = D::P, NACC = MF::NACC;
  static_assert(KF * T == 64 * E, "a lane slot holds E elements");
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  // this wave's k-fragments [f0, f1) (with WK = 1 each wave has the whole K,
  // whichever wave of the block it is)
  const int nkf = (g.K + KF - 1) / KF;
  const int wk = WK == 1 ? 0 : w;
  const int f0 = (int)((long long)wk * nkf / WK), f1 = (int)((long long)(wk + 1) * nkf / WK);
  // the ring runs over the whole k-fragments; a last fragment that K cuts
  // (K % KF != 0) is multiplied after it, with per-lane masked loads
  const bool part = g.K % KF != 0 && f1 == nkf && f1 > f0;
  const int n = f1 - f0 - (part ? 1 : 0);
  const int ksub = (lane / T) * E;
  const unsigned char *pa[TM], *pb[TN];
#pragma unroll
  for (int i = 0; i < TM; i++) {
    const long long r = min((long long)(bm * TM + i) * T + lane % T, (long long)g.M - 1);
    pa[i] = g.Araw + r * g.a_rs + (long long)(f0 * KF + ksub) * g.a_ks;
  }
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const long long c = min((long long)(bn * TN + j) * T + lane % T, (long long)g.N - 1);
    pb[j] = g.Braw + c * g.b_rs + (long long)(f0 * KF + ksub) * g.b_ks;
  }
  const long long a_step = (long long)KF * g.a_ks, b_step = (long long)KF * g.b_ks;

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++)
#pragma unroll
      for (int v = 0; v < NACC; v++) acc[i][j][v] = 0;

  // register ring of PF fragment steps: [slot][plane][tile]
  u32x4 af[PF][P][TM], bfr[PF][P][TN];
  // step s's fragments into ring slot `slot`: FULL, a whole k-fragment inside
  // K (no per-lane conditions); otherwise the fragment K cuts, zeros past K
  auto load = [&](int s, int slot, auto full) __attribute__((always_inline)) {
    constexpr bool FULL = decltype(full)::value;
    const int k0 = (f0 + s) * KF + ksub;
#pragma unroll
    for (int i = 0; i < TM; i++) {
      S v[E];
      load_slot<S, E, AK, FULL>(pa[i] + s * a_step, g.a_ks, k0, g.K, v);
      to_frag<MF, MODE>(v, &af[slot][0][i], TM);
    }
#pragma unroll
    for (int j = 0; j < TN; j++) {
      S v[E];
      load_slot<S, E, BK, FULL>(pb[j] + s * b_step, g.b_ks, k0, g.K, v);
      to_frag<MF, MODE>(v, &bfr[slot][0][j], TN);
    }
  };
  auto mfmas = [&](int slot) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
        if constexpr (MODE == MODE_SPLIT3) {
          // lo*hi, hi*lo, hi*hi: the order of the packed split kernel
          MF::mma(acc[i][j], &af[slot][1][i], &bfr[slot][0][j]);
          MF::mma(acc[i][j], &af[slot][0][i], &bfr[slot][1][j]);
          MF::mma(acc[i][j], &af[slot][0][i], &bfr[slot][0][j]);
        } else {
          MF::mma(acc[i][j], &af[slot][0][i], &bfr[slot][0][j]);
        }
      }
  };

  // The ring's loads are unconditional: a load past the last whole fragment
  // repeats that fragment into a slot no MFMA reads, so the steady-state loop
  // has no branches and the compiler's waits count exactly the loads issued
  // after the one awaited.  The loop runs the whole groups of PF steps; the
  // rest (fewer than PF, already loaded) follow, then the cut fragment.
  using Full = std::true_type;
  if (n > 0) {
    const int nmain = n / PF * PF, rest = n - nmain;
#pragma unroll
    for (int s = 0; s < PF - 1; s++) load(min(s, n - 1), s, Full());
    for (int base = 0; base < nmain; base += PF) {
#pragma unroll
      for (int u = 0; u < PF; u++) {
        // (each step's loads stay in their step: the scheduler would sink
        // them towards their use PF - 1 steps later)
        __builtin_amdgcn_sched_barrier(0);
        load(min(base + u + PF - 1, n - 1), (u + PF - 1) % PF, Full());
        mfmas(u);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int u = 0; u < PF - 1; u++)
      if (u < rest) mfmas(u);
  }
  if (part) { // the fragment K cuts, in k order after the others
    load(n, 0, std::false_type());
    mfmas(0);
  }

  const int row0 = bm * TM * T, col0 = bn * TN * T;
  if constexpr (WK > 1 && SEQ) {
    // waves 1 .. WK-1 in turn hand their partial sums to wave 0
    constexpr int Q = TM * TN * NACC;
#pragma unroll 1
    for (int z = 1; z < WK; z++) {
      if (w == z) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
#pragma unroll
            for (int v = 0; v < NACC; v++) red[((i * TN + j) * NACC + v) * 64 + lane] = (Raw)acc[i][j][v];
      }
      __syncthreads();
      if (w == 0) {
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
          for (int j = 0; j < TN; j++)
#pragma unroll
            for (int v = 0; v < NACC; v++)
              acc[i][j][v] = add_raw<Raw>((Raw)acc[i][j][v], red[((i * TN + j) * NACC + v) * 64 + lane]);
      }
      __syncthreads();
    }
    static_assert(Q > 0, "tile");
  }
  if constexpr (WK == 1 || SEQ) {
    if (WK > 1 && w != 0) return;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++) {
        const int col = col0 + j * T + MF::col(lane);
        const auto bias = epi_bias<EPI>(g, col);
#pragma unroll
        for (int v = 0; v < NACC; v++) {
          const int row = row0 + i * T + MF::row(v, lane);
          if (row < g.M && col < g.N) epi_store<EPI, Raw>(g, row, col, (Raw)acc[i][j][v], bias);
        }
      }
  } else {
    // the waves' partial sums meet in LDS; wave w adds up and stores the
    // values [w*Q/WK, (w+1)*Q/WK) of every lane, in wave order
    constexpr int Q = TM * TN * NACC;
    static_assert(Q % WK == 0, "values split evenly over the waves");
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
      for (int j = 0; j < TN; j++)
#pragma unroll
        for (int v = 0; v < NACC; v++) red[((w * TM + i) * TN * NACC + j * NACC + v) * 64 + lane] = (Raw)acc[i][j][v];
    __syncthreads();
#pragma unroll
    for (int qq = 0; qq < Q / WK; qq++) {
      const int q = w * (Q / WK) + qq;
      Raw s = red[q * 64 + lane];
#pragma unroll
      for (int z = 1; z < WK; z++) s = add_raw<Raw>(s, red[(z * Q + q) * 64 + lane]);
      const int i = q / (TN * NACC), j = (q / NACC) % TN, v = q % NACC;
      const int row = row0 + i * T + MF::row(v, lane), col = col0 + j * T + MF::col(lane);
      if (row < g.M && col < g.N) epi_store<EPI, Raw>(g, row, col, s, epi_bias<EPI>(g, col));
    }
  }
}

template <class MF, int MODE, int TM, int TN, int WK, int PF, int EPI, bool AK, bool BK>
__global__ void __launch_bounds__(WK * 64) k_gemm_direct(const GemmArgs g) {
  __shared__ typename MF::Raw red[direct_red_words<MF, TM, TN, WK, false>()];
  const int bm = blockIdx.x / g.nbn, bn = blockIdx.x - bm * g.nbn;
  direct_tile<MF, MODE, TM, TN, WK, PF, EPI, AK, BK, false>(g, bm, bn, red);
}

// ---------------------------------------------------------------------------
// persistent batched GEMM on the operands as they lie: every (problem, tile)
// work item of many uneven problems in one launch, with no pack pass.  Each
// wave is on its own: it takes an item from an atomic counter and computes
// that tile over the whole K (no LDS, no barriers), so a wave's load pipeline
// runs the full length of its K and the waves of a block never wait for each
// other.  An item names its problem and tile class (64 or 32 rows, 64 or 32
// columns: each dimension above the schedule's threshold takes 64); the queue
// is ordered longest-first.  Each problem's k strides pick whole
// 16-byte loads or element loads.
// ---------------------------------------------------------------------------
template <class MF, int MODE, int EPI, int TM, int TN, int PF>
__device__ __attribute__((always_inline)) inline void batch_wave_item(const GemmArgs &g, int bm, int bn, bool ak,
                                                                      bool bk) {
  // int8: k-contiguous operands only (the host sends nothing else; k-strided
  // bytes would take 16 loads per lane slot)
  if constexpr (std::is_same<MF, MfI8>::value) {
    direct_tile<MF, MODE, TM, TN, 1, PF, EPI, true, true, false>(g, bm, bn, nullptr);
    return;
  }
  if (ak && bk) direct_tile<MF, MODE, TM, TN, 1, PF, EPI, true, true, false>(g, bm, bn, nullptr);
  else if (ak) direct_tile<MF, MODE, TM, TN, 1, PF, EPI, true, false, false>(g, bm, bn, nullptr);
  else if (bk) direct_tile<MF, MODE, TM, TN, 1, PF, EPI, false, true, false>(g, bm, bn, nullptr);
  else direct_tile<MF, MODE, TM, TN, 1, PF, EPI, false, false, false>(g, bm, bn, nullptr);
}

template <class MF, int MODE, int EPI, int OCC>
__global__ void __launch_bounds__(256, OCC) k_gemm_batched_direct(const BatchDirect *probs, const int4 *items,
                                                                   int nitems, int *counter) {
  constexpr long long es = sizeof(typename Direct<MF, MODE>::S);
  const int lane = threadIdx.x & 63;
  // queue q's items are items[qs[q], qs[q+1]); a wave takes from its own
  // XCD's queue (blocks are dealt to the 8 XCDs in turn, so block b runs on
  // XCD b % 8), then from the others in turn; a queue found empty stays empty
  // (its counter only grows), so after finding all of them empty the wave is
  // done
  (void)nitems;
  const int *qs = counter + NUMC_BATCH_QSTART;
  int q = blockIdx.x % NUMC_BATCH_QUEUES, left = NUMC_BATCH_QUEUES;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(counter + q * NUMC_BATCH_COUNTER_STRIDE, 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    const int q0 = qs[q];
    if (it >= qs[q + 1] - q0) {
      if (--left == 0) return;
      q = (q + 1) % NUMC_BATCH_QUEUES;
      continue;
    }
    const int4 w = items[q0 + it];
    const BatchDirect pr = probs[w.x];
    GemmArgs g{};
    g.Araw = pr.A;
    g.a_rs = pr.a_rs;
    g.a_ks = pr.a_ks;
    g.Braw = pr.B;
    g.b_rs = pr.b_rs;
    g.b_ks = pr.b_ks;
    g.M = pr.M;
    g.N = pr.N;
    g.K = pr.K;
    g.ksplit = 1;
    g.C = pr.C;
    g.rso = pr.rso;
    g.cso = pr.cso;
    g.out_bytes = pr.out_bytes;
    g.nprod = 1;
    const bool ak = pr.a_ks == es, bk = pr.b_ks == es;
    // class: bit 1 = 64 rows (else 32), bit 0 = 64 columns (else 32)
    switch (w.w) {
    case 3: batch_wave_item<MF, MODE, EPI, 2, 2, 3>(g, w.y, w.z, ak, bk); break;
    case 2: batch_wave_item<MF, MODE, EPI, 2, 1, 3>(g, w.y, w.z, ak, bk); break;
    case 1: batch_wave_item<MF, MODE, EPI, 1, 2, 3>(g, w.y, w.z, ak, bk); break;
    default: batch_wave_item<MF, MODE, EPI, 1, 1, 4>(g, w.y, w.z, ak, bk); break;
    }
  }
}

// the kernel for the operands' k strides: a unit k stride takes whole 16-byte loads
template <class MF, int MODE, int TM, int TN, int WK, int PF, int EPI>
void launch_direct(const GemmArgs &g, dim3 grid, hipStream_t st) {
  constexpr long long es = sizeof(typename Direct<MF, MODE>::S);
  const bool ak = g.a_ks == es, bk = g.b_ks == es;
  const dim3 blk(WK * 64);
  if (ak && bk) hipLaunchKernelGGL((k_gemm_direct<MF, MODE, TM, TN, WK, PF, EPI, true, true>), grid, blk, 0, st, g);
  else if (ak) hipLaunchKernelGGL((k_gemm_direct<MF, MODE, TM, TN, WK, PF, EPI, true, false>), grid, blk, 0, st, g);
  else if (bk) hipLaunchKernelGGL((k_gemm_direct<MF, MODE, TM, TN, WK, PF, EPI, false, true>), grid, blk, 0, st, g);
  else hipLaunchKernelGGL((k_gemm_direct<MF, MODE, TM, TN, WK, PF, EPI, false, false>), grid, blk, 0, st, g);
}

// {tile, fragment k, waves 1 x 1, no split-K, waves splitting K, direct}
#define DCFG(MF, MODE, TM, TN, WK, PF, EPI)                                                                   \
  {{TM * MF::T, TN * MF::T, MF::KF, 1, 1, 1, WK, 1}, MF::T, &launch_direct<MF, MODE, TM, TN, WK, PF, EPI>, \
   nullptr, {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}}}
// with the biased-store twin (EPI + 4)
#define DCFGB(MF, TM, TN, WK, PF, EPI)                                                                             \
  {{TM * MF::T, TN * MF::T, MF::KF, 1, 1, 1, WK, 1}, MF::T, &launch_direct<MF, MODE_SINGLE, TM, TN, WK, PF, EPI>, \
   &launch_direct<MF, MODE_SINGLE, TM, TN, WK, PF, EPI + 4>, {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}}}

const GemmConfig kDirF32[] = {DCFGB(MfF32, 1, 1, 4, 4, EPI_F32), DCFGB(MfF32, 2, 2, 4, 3, EPI_F32)};
const GemmConfig kDirSplit[] = {DCFG(MfBF16, MODE_SPLIT3, 1, 1, 4, 4, EPI_F32),
                                DCFG(MfBF16, MODE_SPLIT3, 2, 2, 4, 3, EPI_F32)};
const GemmConfig kDirF64[] = {DCFGB(MfF64, 2, 2, 4, 4, EPI_F64), DCFGB(MfF64, 4, 4, 2, 3, EPI_F64)};
const GemmConfig kDirI8[] = {DCFGB(MfI8, 1, 1, 4, 4, EPI_INT), DCFGB(MfI8, 2, 2, 4, 3, EPI_INT)};
const GemmConfig kDirBF16[] = {DCFGB(MfBF16, 1, 1, 4, 4, EPI_BF16), DCFGB(MfBF16, 2, 2, 4, 3, EPI_BF16)};
const GemmConfig kDirF16[] = {DCFGB(MfF16, 1, 1, 4, 4, EPI_F16), DCFGB(MfF16, 2, 2, 4, 3, EPI_F16)};
#undef DCFG
#undef DCFGB
} // namespace

// the direct configurations of a family (none for FP8, whose operands are
// quantised to e4m3 with their block scales by the pack)
const GemmConfig *numc_direct_configs(int family, int *count) {
  const GemmConfig *t = nullptr;
  int n = 0;
  switch (family) {
  case NUMC_FAMILY_F32: t = kDirF32, n = (int)(sizeof kDirF32 / sizeof kDirF32[0]); break;
  case NUMC_FAMILY_SPLIT_BF16: t = kDirSplit, n = (int)(sizeof kDirSplit / sizeof kDirSplit[0]); break;
  case NUMC_FAMILY_F64: t = kDirF64, n = (int)(sizeof kDirF64 / sizeof kDirF64[0]); break;
  case NUMC_FAMILY_I8: t = kDirI8, n = (int)(sizeof kDirI8 / sizeof kDirI8[0]); break;
  case NUMC_FAMILY_BF16: t = kDirBF16, n = (int)(sizeof kDirBF16 / sizeof kDirBF16[0]); break;
  case NUMC_FAMILY_F16: t = kDirF16, n = (int)(sizeof kDirF16 / sizeof kDirF16[0]); break;
  default: break;
  }
  *count = n;
  return t;
}

// resident blocks per CU of the direct batched kernel (its grid is this many
// per CU; the kernels' launch bounds ask the compiler for that many)
constexpr int kBatchDirectOcc[NUMC_FAMILY_COUNT] = {2, 2, 1, 1, 2, 1, 1};
int numc_batch_direct_occupancy(int family) {
  return family >= 0 && family < NUMC_FAMILY_COUNT ? kBatchDirectOcc[family] : 1;
}

int numc_batch_direct_launch(int family, const BatchDirect *probs, const int4 *items, int nitems, int *counter,
                             int grid, hipStream_t st, NumcCtx *prof) {
  hipLaunchKernelGGL(k_clear_queues, dim3(1), dim3(64), 0, st, counter);
  if (prof) numc_prof_begin(prof, PROF_BATCHED, family);
  if (family == NUMC_FAMILY_F32)
    hipLaunchKernelGGL((k_gemm_batched_direct<MfF32, MODE_SINGLE, EPI_F32, kBatchDirectOcc[NUMC_FAMILY_F32]>),
                       dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  else if (family == NUMC_FAMILY_SPLIT_BF16)
    hipLaunchKernelGGL((k_gemm_batched_direct<MfBF16, MODE_SPLIT3, EPI_F32, kBatchDirectOcc[NUMC_FAMILY_SPLIT_BF16]>),
                       dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  else
    hipLaunchKernelGGL((k_gemm_batched_direct<MfI8, MODE_SINGLE, EPI_INT, kBatchDirectOcc[NUMC_FAMILY_I8]>),
                       dim3(grid), dim3(256), 0, st, probs, items, nitems, counter);
  if (prof) numc_prof_end(prof);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return numc_hip_fail(e, "batched direct launch");
  return 0;
}
