// gemm_kernels.h -- device code shared by the GEMM translation units.
//
// Operand layout ("packed fragment order").  Every operand the MFMA kernels
// read has been rewritten by a pack kernel into 1 KiB fragments: 64 lanes x
// 16 bytes, exactly what one wave loads for one MFMA operand.  Lane l of a
// fragment holds row (l mod T) and the E = 16/elem consecutive k values
// starting at (l div T)*E, so a fragment covers T rows and KF = (64/T)*E k.
// Fragments of one plane are stored [row-fragment rt][k-fragment kf], so the
// k-run of one row panel is contiguous; rows are zero padded to the block
// tile and k to the k-chunk, which removes every bound check from the inner
// loop.  B is packed as B^T (its rows are the output columns), so A and B use
// the same code.  Planes (bf16 hi/lo, int8 digits) follow each other.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

#define NUMC_FRAG_BYTES 1024

// ---------------------------------------------------------------------------
// MFMA flavours.  T: rows per fragment; KF: k per fragment; FR: fragments of
// each operand consumed by one mma() call; NMMA: MFMA instructions per call.
// ---------------------------------------------------------------------------

struct MfF32 { // v_mfma_f32_32x32x2_f32 x4 per fragment
  static constexpr int T = 32, KF = 8, FR = 1, NACC = 16, NMMA = 4;
  using Acc = f32x16;
  using Raw = float;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b) {
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a[0].x), __uint_as_float(b[0].x), c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a[0].y), __uint_as_float(b[0].y), c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a[0].z), __uint_as_float(b[0].z), c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(a[0].w), __uint_as_float(b[0].w), c, 0, 0, 0);
  }
  __device__ static inline int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
  __device__ static inline int col(int lane) { return lane & 31; }
};

struct MfBF16 { // v_mfma_f32_32x32x16_bf16
  static constexpr int T = 32, KF = 16, FR = 1, NACC = 16, NMMA = 1;
  using Acc = f32x16;
  using Raw = float;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b) {
    bf16x8 av = __builtin_bit_cast(bf16x8, a[0]);
    bf16x8 bv = __builtin_bit_cast(bf16x8, b[0]);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
  }
  __device__ static inline int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
  __device__ static inline int col(int lane) { return lane & 31; }
};

struct MfF16 { // v_mfma_f32_32x32x16_f16 (fragment layout as MfBF16)
  static constexpr int T = 32, KF = 16, FR = 1, NACC = 16, NMMA = 1;
  using Acc = f32x16;
  using Raw = float;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b) {
    f16x8 av = __builtin_bit_cast(f16x8, a[0]);
    f16x8 bv = __builtin_bit_cast(f16x8, b[0]);
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, c, 0, 0, 0);
  }
  __device__ static inline int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
  __device__ static inline int col(int lane) { return lane & 31; }
};

struct MfI8 { // v_mfma_i32_32x32x32_i8
  static constexpr int T = 32, KF = 32, FR = 1, NACC = 16, NMMA = 1;
  using Acc = i32x16;
  using Raw = int;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b) {
    i32x4 av = __builtin_bit_cast(i32x4, a[0]);
    i32x4 bv = __builtin_bit_cast(i32x4, b[0]);
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, c, 0, 0, 0);
  }
  __device__ static inline int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
  __device__ static inline int col(int lane) { return lane & 31; }
};

// v_mfma_scale_f32_32x32x64_f8f6f4, both operands OCP e4m3 with E8M0 block
// scales (2^(s-127)).  The MFMA's 64 k are two packed fragments: a[0] holds
// k 0-31 (lane half h: k 16h..16h+15) and a[1] k 32-63.  Measured on gfx950:
// the scale in lane r (r < 32) multiplies k 0-31 of row r -- fragment a[0],
// bytes of both lane halves -- and the scale in lane r + 32 multiplies
// k 32-63 -- fragment a[1]; the same for the columns of B.  So a scale
// covers one row of one packed fragment: 32 consecutive k, 32-aligned.
struct MfFP8 {
  static constexpr int T = 32, KF = 32, FR = 2, NACC = 16, NMMA = 1;
  static constexpr bool SCALED = true;
  using Acc = f32x16;
  using Raw = float;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b, int sa, int sb) {
    i32x8 av = __builtin_shufflevector(__builtin_bit_cast(i32x4, a[0]), __builtin_bit_cast(i32x4, a[1]), 0, 1, 2, 3, 4, 5, 6, 7);
    i32x8 bv = __builtin_shufflevector(__builtin_bit_cast(i32x4, b[0]), __builtin_bit_cast(i32x4, b[1]), 0, 1, 2, 3, 4, 5, 6, 7);
    c = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, sa, 0, sb);
  }
  __device__ static inline int row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }
  __device__ static inline int col(int lane) { return lane & 31; }
};

struct MfF64 { // v_mfma_f64_16x16x4_f64 x2 per fragment
  static constexpr int T = 16, KF = 8, FR = 1, NACC = 4, NMMA = 2;
  using Acc = f64x4;
  using Raw = double;
  __device__ static inline void mma(Acc &c, const u32x4 *a, const u32x4 *b) {
    double a0 = __hiloint2double((int)a[0].y, (int)a[0].x), a1 = __hiloint2double((int)a[0].w, (int)a[0].z);
    double b0 = __hiloint2double((int)b[0].y, (int)b[0].x), b1 = __hiloint2double((int)b[0].w, (int)b[0].z);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c, 0, 0, 0);
  }
  // output element v of lane l is row (l / 16) + 4v (checked on gfx950)
  __device__ static inline int row(int v, int lane) { return (lane >> 4) + 4 * v; }
  __device__ static inline int col(int lane) { return lane & 15; }
};

// ---------------------------------------------------------------------------
// kernel arguments
// ---------------------------------------------------------------------------

enum GemmMode { MODE_SINGLE = 0, MODE_SPLIT3 = 1 };

// MFMA families whose operands carry per-block scales (MF::SCALED)
template <class MF, class = void> struct IsScaled { static constexpr bool value = false; };
template <class MF> struct IsScaled<MF, std::void_t<decltype(MF::SCALED)>> { static constexpr bool value = MF::SCALED; };
// EPI_BF16 / EPI_F16: FP32 sums rounded to nearest even into 16-bit floats.
// EPI_*B (= EPI_* + 4): the same outputs plus a per-column bias (numc_linear),
// separate kernels so that the plain stores keep their code and register budget.
enum GemmEpi {
  EPI_F32 = 0, EPI_F64 = 1, EPI_INT = 3,
  EPI_F32B = 4, EPI_F64B = 5, EPI_INTB = 7,
  EPI_BF16 = 8, EPI_F16 = 9, EPI_BF16B = 12, EPI_F16B = 13
};

// binary32 -> bfloat16 / binary16 bits, round to nearest even (NaN stays NaN)
__device__ inline uint32_t f32_to_bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0x7FFFFFu) ? 0x40u : 0u);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ inline uint32_t f32_to_f16_bits(float x) {
  return (uint32_t)__builtin_bit_cast(unsigned short, (_Float16)x);
}
__device__ inline float bf16_bits_to_f32(uint32_t b) { return __uint_as_float(b << 16); }
__device__ inline float f16_bits_to_f32(uint32_t b) { return (float)__builtin_bit_cast(_Float16, (unsigned short)b); }

struct GemmArgs {
  const unsigned char *A, *B; // packed operands
  long long a_plane, b_plane; // bytes per plane
  int nKF;                    // k-fragments per row panel (padded K / KF)
  int nKC;                    // k-chunks (padded K / KC)
  int M, N;
  int nbm, nbn;               // block grid
  int ksplit;                 // K slices (gridDim.y)
  // epilogue
  void *C;
  long long rso, cso;         // output strides in elements
  int out_bytes;              // integer output element size (1, 2, 4)
  int shift_pass;             // integer digit pass s (adds (sum << 8s))
  int passes;                 // > 1: every digit pass in this launch (int8 MFMA kernels), shift_pass 0
  const int *plane_flags;     // integer: [0] nonzero-plane mask of A, [1] of B (null: use pa/pb)
  const unsigned char *a_sc, *b_sc; // FP8: E8M0 block scales, [row tile][k-fragment pair][64 lanes]
  void *partial;              // split-K partial sums [ksplit][M][N] (raw accumulator type)
  int nprod;                  // MODE_SINGLE: products concatenated along K
  int pa[4], pb[4];           // plane of A / B for each product
  const void *bias;           // optional per-column bias (output dtype), added at the store
  long long bias_s;           // bias stride in elements
  long long a_zs, b_zs, c_zs; // strided batch: bytes between slices (blockIdx.z) of A, B (packed) and C
  const unsigned char *Araw;  // A read in place (kernels compiled with VAR & 2): k-contiguous rows,
  long long a_rs;             // a_rs bytes apart, 16-byte aligned, K a multiple of the k-chunk
  const unsigned char *Braw;  // B read in place (VAR & 4): the same for the rows of B^T
  long long b_rs;
  long long a_ks, b_ks;       // direct kernels: bytes between consecutive k of A / of B^T (any value)
  int K;                      // direct kernels: the inner dimension
};

// Nonzero-digit-plane masks of the integer packs: each operand's mask is
// spread over NUMC_FLAG_SLOTS words, one per 128-byte line, so that the
// pack's blocks OR into different L2 lines instead of queueing on one
// address; a reader ORs the slots.
#define NUMC_FLAG_SLOTS 32
#define NUMC_FLAG_STRIDE 32                                 // ints between slots
#define NUMC_FLAG_INTS (NUMC_FLAG_SLOTS * NUMC_FLAG_STRIDE) // ints per operand

// Both operands' masks in one vector load: lanes 0-31 read A's slots, lanes
// 32-63 B's, and a ballot per plane bit ORs them (call with the whole wave).
static_assert(NUMC_FLAG_SLOTS == 32, "one slot per lane of a half wave");
__device__ inline void plane_masks(const int *flags, int *ma, int *mb) {
  const int lane = threadIdx.x & 63;
  const int v = flags[(lane >> 5) * NUMC_FLAG_INTS + (lane & 31) * NUMC_FLAG_STRIDE];
  int a = 0, b = 0;
#pragma unroll
  for (int p = 0; p < 4; p++) { // at most 4 digit planes
    const unsigned long long bal = __ballot((v >> p) & 1);
    if (bal & 0xFFFFFFFFull) a |= 1 << p;
    if (bal >> 32) b |= 1 << p;
  }
  *ma = a;
  *mb = b;
}

// Products of integer digit pass s: (i, s - i) for i = pa0 .. pa0 + n - 1,
// restricted to planes holding a nonzero digit (masks from the pack kernels).
__device__ inline int int_products(int s, int maskA, int maskB, int *pa0, int *pb0) {
  int na = 32 - __clz(maskA), nb = 32 - __clz(maskB);
  int lo = max(0, s - nb + 1), hi = min(s, na - 1);
  *pa0 = lo;
  *pb0 = s - lo;
  return hi >= lo ? hi - lo + 1 : 0;
}

// ---------------------------------------------------------------------------
// epilogue helpers
// ---------------------------------------------------------------------------

// The bias term of column col as the store adds it (0 without a bias and
// for epilogues that take none).  Integer outputs: the low 32 bits, added by
// the first digit pass only.  Split-K slices store raw partial sums and their
// reduction adds the bias.
// Global (address space 1) views of generic pointers.  Operand and output
// pointers read from a problem table in memory are generic pointers, and
// generic (flat) accesses count on both the vector-memory and the LDS
// counters: the compiler then waits for every access in flight
// (vmcnt(0) lgkmcnt(0)) before using any load, which empties a load pipeline
// at every step.  Every operand and output here lives in device memory.
#define NUMC_GLOBAL(T, p) ((__attribute__((address_space(1))) T *)(p))

template <int EPI> struct EpiBias { using T = uint32_t; };
template <> struct EpiBias<EPI_F32B> { using T = float; };
template <> struct EpiBias<EPI_F64B> { using T = double; };
template <> struct EpiBias<EPI_BF16B> { using T = float; };
template <> struct EpiBias<EPI_F16B> { using T = float; };

template <int EPI>
__device__ inline typename EpiBias<EPI>::T epi_bias(const GemmArgs &g, int col) {
  using T = typename EpiBias<EPI>::T;
  if constexpr (EPI == EPI_F32B || EPI == EPI_F64B) {
    return col < g.N ? NUMC_GLOBAL(const T, g.bias)[(long long)col * g.bias_s] : (T)0;
  } else if constexpr (EPI == EPI_BF16B || EPI == EPI_F16B) {
    if (col >= g.N) return 0.f;
    const uint32_t b = NUMC_GLOBAL(const uint16_t, g.bias)[(long long)col * g.bias_s];
    return EPI == EPI_BF16B ? bf16_bits_to_f32(b) : f16_bits_to_f32(b);
  } else if constexpr (EPI == EPI_INTB) {
    if (g.shift_pass || col >= g.N) return 0u;
    const long long b = (long long)col * g.bias_s;
    return g.out_bytes == 1 ? (uint32_t)NUMC_GLOBAL(const uint8_t, g.bias)[b]
           : g.out_bytes == 2 ? (uint32_t)NUMC_GLOBAL(const uint16_t, g.bias)[b]
                              : NUMC_GLOBAL(const uint32_t, g.bias)[b];
  } else {
    return 0u;
  }
}

template <int EPI, class Raw>
__device__ inline void epi_store(const GemmArgs &g, int row, int col, Raw v, typename EpiBias<EPI>::T bias) {
  const long long off = (long long)row * g.rso + (long long)col * g.cso;
  if constexpr (EPI == EPI_F32) {
    NUMC_GLOBAL(float, g.C)[off] = (float)v;
  } else if constexpr (EPI == EPI_F32B) {
    NUMC_GLOBAL(float, g.C)[off] = (float)v + bias;
  } else if constexpr (EPI == EPI_F64) {
    NUMC_GLOBAL(double, g.C)[off] = (double)v;
  } else if constexpr (EPI == EPI_F64B) {
    NUMC_GLOBAL(double, g.C)[off] = (double)v + bias;
  } else if constexpr (EPI == EPI_BF16 || EPI == EPI_BF16B || EPI == EPI_F16 || EPI == EPI_F16B) {
    float r = (float)v;
    if constexpr (EPI == EPI_BF16B || EPI == EPI_F16B) r += bias;
    NUMC_GLOBAL(uint16_t, g.C)[off] = (uint16_t)(EPI == EPI_BF16 || EPI == EPI_BF16B ? f32_to_bf16_bits(r) : f32_to_f16_bits(r));
  } else {
    // integer digit pass s adds (sum << 8s); the first pass writes the output,
    // with the bias (mod 2^w) under EPI_INTB
    uint32_t x = (uint32_t)(int)v << (8 * g.shift_pass);
    if constexpr (EPI == EPI_INTB) x += bias;
    if (g.out_bytes == 1) {
      __attribute__((address_space(1))) uint8_t *p = NUMC_GLOBAL(uint8_t, g.C) + off;
      *p = (uint8_t)(g.shift_pass ? (uint32_t)*p + x : x);
    } else if (g.out_bytes == 2) {
      __attribute__((address_space(1))) uint16_t *p = NUMC_GLOBAL(uint16_t, g.C) + off;
      *p = (uint16_t)(g.shift_pass ? (uint32_t)*p + x : x);
    } else {
      __attribute__((address_space(1))) uint32_t *p = NUMC_GLOBAL(uint32_t, g.C) + off;
      *p = g.shift_pass ? *p + x : x;
    }
  }
}
