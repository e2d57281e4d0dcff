// Saber polynomial-product helpers: negacyclic products as Toeplitz dot
// products over LDS-staged operands, and the 13/10-bit unpackers.
#pragma once
#include "common.hpp"

namespace dpkem {

// Toeplitz layout.  For a polynomial a, the reversed extended sequence is
//   r[p] = a[-p] for p in [-255, 0],  r[p] = -a[256 - p] for p in [1, 255],
// so that the negacyclic product c = a * s satisfies
//   c[t] = sum_k r[k - t] * s[k].
// Copy 1 stores r[p] at u16 offset p + 256 and copy 2 at p + 257, so that the
// pair (r[k - t], r[k - t + 1]) is one aligned dword in copy 1 when t is even
// and in copy 2 when t is odd.  A thread owns 8 consecutive outputs t0..t0+7
// (t0 % 8 == 0) and slides one 16-byte window per copy along k.

constexpr int kToepStride = 1040;  // u16 per polynomial (2 copies x 520)
typedef unsigned short us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t dot2(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b), c, false);
}

// Store coefficient q of polynomial P into both Toeplitz copies.
__device__ __forceinline__ void toep_put(uint16_t* T, int P, int q, uint16_t v) {
  uint16_t* r1 = T + P * kToepStride;
  uint16_t* r2 = r1 + 520;
  r1[256 - q] = v;
  r2[257 - q] = v;
  if (q) {
    uint16_t nv = (uint16_t)(0u - v);
    r1[512 - q] = nv;
    r2[513 - q] = nv;
  }
}

// acc[d] += sum_k r_P[k - (t0 + d)] * S[k],  d = 0..7, where r_P is the
// Toeplitz form of polynomial P in T and S a plain polynomial (16-byte aligned).
// The negacyclic product is commutative, so either factor may be the staged one.
__device__ __forceinline__ void toep_dot(uint32_t acc[8], const uint16_t* T, int P, const uint16_t* S, int t0) {
  const uint4* r1 = reinterpret_cast<const uint4*>(T + P * kToepStride);
  const uint4* r2 = reinterpret_cast<const uint4*>(T + P * kToepStride + 520);
  const uint4* sv = reinterpret_cast<const uint4*>(S);
  const int base = (256 - t0) >> 3;  // 16-byte index of the r[-t0] pair in each copy
  uint4 p1 = r1[base - 1];
  uint4 p2 = r2[base - 1];
#pragma unroll
  for (int kk = 0; kk < 256; kk += 8) {
    const uint4 c1 = r1[base + (kk >> 3)];
    const uint4 c2 = r2[base + (kk >> 3)];
    const uint4 s4 = sv[kk >> 3];
    const uint32_t w1[8] = {p1.x, p1.y, p1.z, p1.w, c1.x, c1.y, c1.z, c1.w};  // dwords -4..3
    const uint32_t w2[8] = {p2.x, p2.y, p2.z, p2.w, c2.x, c2.y, c2.z, c2.w};
    const uint32_t sp[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
    for (int kp = 0; kp < 4; ++kp) {
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        // pair for output t0+d at k = kk + 2kp: dword (kp - d/2) of copy (d & 1)
        const uint32_t w = (d & 1) ? w2[4 + kp - (d >> 1)] : w1[4 + kp - (d >> 1)];
        acc[d] = dot2(w, sp[kp], acc[d]);
      }
    }
    p1 = c1;
    p2 = c2;
  }
}

// 13-bit little-endian group: 8 coefficients from 13 bytes.
__device__ __forceinline__ void unpack13_group(uint16_t v[8], const uint8_t* b) {
  uint64_t lo = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) lo |= (uint64_t)b[i] << (8 * i);
  uint64_t hi = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) hi |= (uint64_t)b[8 + i] << (8 * i);
  // 104-bit value hi:lo
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int bit = 13 * i;
    uint64_t x = bit < 64 ? (lo >> bit) | (bit > 51 ? (hi << (64 - bit)) : 0) : (hi >> (bit - 64));
    v[i] = (uint16_t)(x & 0x1FFF);
  }
}

// 10-bit little-endian group: 4 coefficients from 5 bytes.
__device__ __forceinline__ void unpack10_group(uint16_t v[4], const uint8_t* b) {
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) x |= (uint64_t)b[i] << (8 * i);
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = (uint16_t)((x >> (10 * i)) & 0x3FF);
}

}  // namespace dpkem
