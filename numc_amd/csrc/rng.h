// rng.h -- the seeded input generator, one source for the host fill
// (runtime.cpp) and the device fill (generate.hip).
//
// Element i of an array comes from xoshiro256** block i >> 16, seeded by
// splitmix64 from (seed, stream, block), so blocks fill independently.
// dist 0: uniform [lo, hi) (binary32 arithmetic for 32- and 16-bit floats,
// binary64 otherwise); dist 1: normal(0,1) by Box-Muller, two values per pair
// of draws; dist 2: integers uniform in [lo, hi] by rejection on 32-bit draws.
// 16-bit floats round the binary32 value to nearest even.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "numc/dtype.h"

struct NumcXoshiro {
  uint64_t s[4];
  __host__ __device__ static uint64_t mix(uint64_t &st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  __host__ __device__ NumcXoshiro(uint64_t seed, uint64_t stream, uint64_t block) {
    uint64_t st = seed ^ (stream * 0x9E3779B97F4A7C15ULL) ^ (block * 0xD1B54A32D192ED03ULL);
    for (int w = 0; w < 4; w++) s[w] = mix(st);
  }
  __host__ __device__ static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  __host__ __device__ uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
};

// binary32 -> bfloat16 / binary16 bits, round to nearest even (NaN stays NaN)
__host__ __device__ inline uint16_t numc_f32_to_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0x7FFFFFu) ? 0x40u : 0u));
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline uint16_t numc_f32_to_f16(float x) {
  _Float16 h = (_Float16)x; // the compiler's conversion rounds to nearest even
  uint16_t b;
  memcpy(&b, &h, 2);
  return b;
}
// binary64 -> binary32 as a rounding of its own: the optimiser would fold
// (_Float16)(float)d into one binary64 -> binary16 rounding, which differs on ties
__host__ __device__ inline float numc_narrow(double v) {
#ifdef __HIP_DEVICE_COMPILE__
  float f = (float)v;
  __asm__ volatile("" : "+v"(f));
  return f;
#else
  volatile float f = (float)v;
  return f;
#endif
}

__host__ __device__ inline void numc_store_val(void *p, NumcDType dt, size_t i, double fv, int64_t iv, bool is_int) {
  switch (dt) {
  case NUMC_DTYPE_INT8: ((int8_t *)p)[i] = (int8_t)iv; break;
  case NUMC_DTYPE_INT16: ((int16_t *)p)[i] = (int16_t)iv; break;
  case NUMC_DTYPE_INT32: ((int32_t *)p)[i] = (int32_t)iv; break;
  case NUMC_DTYPE_INT64: ((int64_t *)p)[i] = iv; break;
  case NUMC_DTYPE_UINT8: ((uint8_t *)p)[i] = (uint8_t)iv; break;
  case NUMC_DTYPE_UINT16: ((uint16_t *)p)[i] = (uint16_t)iv; break;
  case NUMC_DTYPE_UINT32: ((uint32_t *)p)[i] = (uint32_t)iv; break;
  case NUMC_DTYPE_UINT64: ((uint64_t *)p)[i] = (uint64_t)iv; break;
  case NUMC_DTYPE_FLOAT32: ((float *)p)[i] = is_int ? (float)iv : (float)fv; break;
  case NUMC_DTYPE_FLOAT64: ((double *)p)[i] = is_int ? (double)iv : fv; break;
  case NUMC_DTYPE_BFLOAT16: ((uint16_t *)p)[i] = numc_f32_to_bf16(is_int ? (float)iv : (float)fv); break;
  case NUMC_DTYPE_FLOAT16: ((uint16_t *)p)[i] = numc_f32_to_f16(numc_narrow(is_int ? (double)iv : fv)); break;
  }
}

struct NumcFillSpec {
  NumcDType dt;
  int dist;
  uint64_t seed, stream;
  double lo, hi;
  int64_t ilo;
  uint64_t span, limit; // dist 2: hi - lo + 1 and the rejection limit
};

__host__ __device__ inline NumcFillSpec numc_fill_spec(NumcDType dt, int dist, uint64_t seed, uint64_t stream, double lo,
                                                       double hi) {
  NumcFillSpec f;
  f.dt = dt;
  f.dist = dist;
  f.seed = seed;
  f.stream = stream;
  f.lo = lo;
  f.hi = hi;
  f.ilo = (int64_t)lo;
  f.span = dist == 2 ? (uint64_t)((int64_t)hi - f.ilo) + 1 : 1;
  f.limit = dist == 2 ? ((1ULL << 32) / f.span) * f.span : 0;
  return f;
}

// elements [b << 16, min((b + 1) << 16, n)) of the array at p
__host__ __device__ inline void numc_fill_block(void *p, size_t n, size_t b, const NumcFillSpec &f) {
  // each multiply and add rounds on its own, as on the host: no fused
  // multiply-add on the device either
#pragma clang fp contract(off)
  NumcXoshiro r(f.seed, f.stream, b);
  size_t end = (b + 1) << 16;
  if (end > n) end = n;
  const NumcDType dt = f.dt;
  for (size_t i = b << 16; i < end;) {
    if (f.dist == 0) {
      uint64_t x = r.next();
      if (dt == NUMC_DTYPE_FLOAT32 || dt == NUMC_DTYPE_BFLOAT16 || dt == NUMC_DTYPE_FLOAT16) {
        float u = (float)(x >> 40) * 0x1p-24f;
        float v = (float)f.lo + ((float)f.hi - (float)f.lo) * u;
        if (dt == NUMC_DTYPE_FLOAT32) ((float *)p)[i] = v;
        else ((uint16_t *)p)[i] = dt == NUMC_DTYPE_BFLOAT16 ? numc_f32_to_bf16(v) : numc_f32_to_f16(v);
      } else {
        double u = (double)(x >> 11) * 0x1p-53;
        numc_store_val(p, dt, i, f.lo + (f.hi - f.lo) * u, 0, false);
      }
      i++;
    } else if (f.dist == 1) {
      uint64_t x = r.next(), y = r.next();
      double u1 = (double)((x >> 11) + 1) * 0x1p-53, u2 = (double)(y >> 11) * 0x1p-53;
      double rad = sqrt(-2.0 * log(u1)), th = 6.283185307179586476925286766559 * u2;
      numc_store_val(p, dt, i, rad * cos(th), 0, false);
      if (i + 1 < end) numc_store_val(p, dt, i + 1, rad * sin(th), 0, false);
      i += 2;
    } else {
      uint64_t x32;
      do x32 = r.next() >> 32;
      while (x32 >= f.limit);
      numc_store_val(p, dt, i, 0.0, f.ilo + (int64_t)(x32 % f.span), true);
      i++;
    }
  }
}
