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

// ---------------------------------------------------------------- int8 MFMA form
//
// The same negacyclic products on v_mfma_i32_16x16x64_i8.  Per K-block j
// (one staged polynomial s_j) the product is C += T(s_j) B_j, where T(s) is
// the 256 x 256 negacyclic Toeplitz matrix, T[t][k] = s[t - k] for t >= k and
// -s[t - k + 256] otherwise, and B_j holds up to 16 columns of 256 int8 limbs.
// A 13-bit (or 10-bit) coefficient a splits into signed limbs
//   hi = (a + 128) >> 8,  lo = a - 256 hi  (lo in [-128, 127], hi in [0, 32])
// so a = 256 hi + lo exactly and the i32 sums need no bias correction.
//
// T(s) is held as the reversed sequence V[e] = T[t][t + e] (e = k - t in
// [-255, 255]): V[e] = s[-e] for e <= 0 and -s[256 - e] for e > 0.  Four
// copies, copy c holding V[e] at byte e + 256 + c, make every 16-byte run
// V[e0 .. e0 + 15] a 4-byte aligned read from copy (-e0) & 3.  The copies
// are 136 dwords apart (8 mod 32 banks): the lanes of one operand read use
// copy m & 3 and runs a dword apart within a copy, so each copy's runs fall
// in a bank window of their own.
//
// Tiles: M-tile a covers outputs t = 16a .. 16a + 15, K-chunk b covers
// k = 64b .. 64b + 63.  The A operand of tile (a, b) depends only on the
// diagonal delta = 16a - 64b (28 values), so it is read once per diagonal and
// used for up to four tiles.  Operand lanes: lane (q, m) = lane 16 q + m holds
// row m (A) or column m (B) at k = 16 q + 0..15 of the chunk; the result tile
// lane (g, n) holds rows 4 g .. 4 g + 3 of column n.

typedef int v4i8x4 __attribute__((ext_vector_type(4)));

constexpr int kVCopy = 544;                 // bytes per shifted copy (515 used)
constexpr int kVPoly = 4 * kVCopy;          // bytes per staged polynomial
constexpr int kToepDiagonals = 28;          // delta = 16 a - 64 b, a in [0, 16), b in [0, 4)

__host__ __device__ constexpr int toep8_delta(int di) { return 16 * di - 192; }
// M-tile served by diagonal di at K-chunk b, or -1
__host__ __device__ constexpr int toep8_tile(int di, int b) {
  return (toep8_delta(di) + 64 * b >= 0 && toep8_delta(di) + 64 * b < 256) ? (toep8_delta(di) + 64 * b) / 16 : -1;
}

// Store coefficient q (value v) of the staged polynomial into its four copies.
__host__ __device__ __forceinline__ void toep8_put(int8_t* V, int q, int v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    V[c * kVCopy + 256 - q + c] = (int8_t)v;
    if (q) V[c * kVCopy + 512 - q + c] = (int8_t)(-v);
  }
}

#if defined(__HIP__) || defined(__HIPCC__)
// The same four copies written by a whole wave at once: lane g holds
// coefficients 4g .. 4g + 3 of the polynomial, packed as v (byte k =
// s[4g + 3 - k], the ascending-address order of the copies) and nv (the
// negated bytes).  In copy c the aligned dwords at bytes 256 - 4g and
// 512 - 4g hold s[4g + c - i] (negated in the second) in byte i, i.e. bytes
// of this lane and of lane g - 1: one v_alignbyte each, with lane g - 1's
// values from a one-lane wave rotation (lane 0 takes lane 63's negated
// values: the wrap of the negacyclic sequence).  Lane 63 also fills the
// dword at byte 0 of each copy.  Bytes at 512 + c .. 515 + c of copy c
// (never read) receive junk.
__device__ __forceinline__ void toep8_put_wave(int8_t* V, int g, uint32_t v, uint32_t nv) {
  const uint32_t z = g == 63 ? nv : v;
  const uint32_t zm = (uint32_t)__builtin_amdgcn_mov_dpp((int)z, 0x13C, 0xF, 0xF, false);   // wave_ror:1
  const uint32_t nm = (uint32_t)__builtin_amdgcn_mov_dpp((int)nv, 0x13C, 0xF, 0xF, false);  // wave_ror:1
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    uint32_t* w = reinterpret_cast<uint32_t*>(V + c * kVCopy);
    w[64 - g] = __builtin_amdgcn_alignbyte(zm, v, 3 - c);
    w[128 - g] = __builtin_amdgcn_alignbyte(nm, nv, 3 - c);
    if (g == 63) w[0] = __builtin_amdgcn_alignbyte(v, v, 3 - c);
  }
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Signed limbs of four coefficients packed in pairs (u0 = x0 | x1 << 16,
// u1 = x2 | x3 << 16, each x < 2^13): the low limb of x is its low byte read
// as int8 (x - 256 hi), the high limb hi = (x + 128) >> 8; lo / hi return the
// four bytes in coefficient order.
__device__ __forceinline__ void limbs8x4(uint32_t u0, uint32_t u1, uint32_t* lo, uint32_t* hi) {
  const u16x2 r = {128, 128}, e = {8, 8};
  const uint32_t h0 = __builtin_bit_cast(uint32_t, (__builtin_bit_cast(u16x2, u0) + r) >> e);
  const uint32_t h1 = __builtin_bit_cast(uint32_t, (__builtin_bit_cast(u16x2, u1) + r) >> e);
  *lo = __builtin_amdgcn_perm(u1, u0, 0x06040200u);  // bytes 0, 2 of u0, then of u1
  *hi = __builtin_amdgcn_perm(h1, h0, 0x06040200u);
}
#endif

// Byte offset (within the staged polynomial) of the aligned 16-byte run
// V[e0 .. e0 + 15].
__host__ __device__ __forceinline__ int toep8_run(int e0) {
  const int c = (-e0) & 3;
  return c * kVCopy + e0 + 256 + c;
}

// e0 of lane (q, m) for diagonal di
__host__ __device__ __forceinline__ int toep8_e0(int di, int q, int m) { return 16 * q - m - toep8_delta(di); }

// Signed limbs of a coefficient below 2^13.
__host__ __device__ __forceinline__ void limbs8(uint32_t a, int8_t* lo, int8_t* hi) {
  const int h = (int)((a + 128) >> 8);
  *hi = (int8_t)h;
  *lo = (int8_t)((int)a - 256 * h);
}

// B-operand columns: output column n (0 .. L - 1 the matrix rows, L the
// inner product) has its low limbs in column n and its high limbs in column
// 8 + n.  Limb arrays are stored per K-block in slot order lo 0..L, hi 0..L;
// returns the slot of result / operand column nl, or -1 (unused column).
__host__ __device__ __forceinline__ int toep8_slot(int nl, int L, bool inner) {
  const int last = inner ? L : L - 1;
  if (nl <= last) return nl;
  if (nl >= 8 && nl - 8 <= last) return L + 1 + (nl - 8);
  return -1;
}

// acc[a] += T(s) B for one K-block: bop[b] is this lane's B operand of K-chunk
// b, V the staged polynomial, mfma(a, b, c) the matrix instruction.
template <class MFMA>
__device__ __forceinline__ void toep8_block(v4i8x4 acc[16], const int8_t* V, const v4i8x4 bop[4], int lane,
                                           MFMA&& mfma) {
  const int q = lane >> 4, m = lane & 15;
#pragma unroll
  for (int di = 0; di < kToepDiagonals; ++di) {
    const uint32_t* r = reinterpret_cast<const uint32_t*>(V + toep8_run(toep8_e0(di, q, m)));
    const v4i8x4 aop = {(int)r[0], (int)r[1], (int)r[2], (int)r[3]};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      constexpr int kUnused = -1;
      const int a = toep8_tile(di, b);
      if (a != kUnused) acc[a] = mfma(aop, bop[b], acc[a]);
    }
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
