// Keccak-f[1600] for gfx950, one state per lane held in VGPRs.
//
// Replaces the reference's per-thread local-memory permutation
// (FrodoKEM976-GPU/src/shake.cu:66-331, SABER-GPU/src/poly.cu:436-701) and
// its 25-thread shared-memory form (FrodoKEM976-GPU/src/shake.cu:380-604).
// The 25 lanes stay in registers for the whole sponge; every index is a
// compile-time constant so the state never spills to scratch.  64-bit
// rotates lower to v_alignbit_b32 pairs and the 3-input XOR / chi terms to
// v_bitop3_b32 on gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dpkem {

__constant__ static const uint64_t kKeccakRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

// 64-bit helpers on the two 32-bit halves.  gfx950 has a three-input
// boolean op (v_bitop3_b32, truth table over src0/src1/src2 = 0xF0/0xCC/0xAA)
// and a funnel shift (v_alignbit_b32); hipcc does not form either from plain
// C here, so they are called explicitly.
__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint64_t mk64(uint32_t lo, uint32_t hi) { return ((uint64_t)hi << 32) | lo; }

__device__ __forceinline__ uint64_t xor3(uint64_t a, uint64_t b, uint64_t c) {
  return mk64(__builtin_amdgcn_bitop3_b32(lo32(a), lo32(b), lo32(c), 0x96),
              __builtin_amdgcn_bitop3_b32(hi32(a), hi32(b), hi32(c), 0x96));
}
// a ^ (~b & c)
__device__ __forceinline__ uint64_t chi(uint64_t a, uint64_t b, uint64_t c) {
  return mk64(__builtin_amdgcn_bitop3_b32(lo32(a), lo32(b), lo32(c), 0xD2),
              __builtin_amdgcn_bitop3_b32(hi32(a), hi32(b), hi32(c), 0xD2));
}
// rotate left by a compile-time constant 0 < n < 64
template <int N>
__device__ __forceinline__ uint64_t rotl(uint64_t x) {
  static_assert(N > 0 && N < 64, "rotation");
  const uint32_t l = lo32(x), h = hi32(x);
  if constexpr (N == 32) {
    return mk64(h, l);
  } else if constexpr (N < 32) {
    return mk64(__builtin_amdgcn_alignbit(l, h, 32 - N), __builtin_amdgcn_alignbit(h, l, 32 - N));
  } else {
    return mk64(__builtin_amdgcn_alignbit(h, l, 64 - N), __builtin_amdgcn_alignbit(l, h, 64 - N));
  }
}

// One round: theta, rho+pi, chi, iota.  Lane (x, y) lives at index x + 5y.
// theta is folded as A ^ C[x-1] ^ rot(C[x+1], 1) (one three-input XOR per lane).
#define DPKEM_KECCAK_ROUND(A, RC)                                                          \
  do {                                                                                     \
    uint64_t C0 = xor3(xor3(A[0], A[5], A[10]), A[15], A[20]);                             \
    uint64_t C1 = xor3(xor3(A[1], A[6], A[11]), A[16], A[21]);                             \
    uint64_t C2 = xor3(xor3(A[2], A[7], A[12]), A[17], A[22]);                             \
    uint64_t C3 = xor3(xor3(A[3], A[8], A[13]), A[18], A[23]);                             \
    uint64_t C4 = xor3(xor3(A[4], A[9], A[14]), A[19], A[24]);                             \
    uint64_t R0 = rotl<1>(C0), R1 = rotl<1>(C1), R2 = rotl<1>(C2), R3 = rotl<1>(C3),        \
             R4 = rotl<1>(C4);                                                             \
    /* B[y][2x+3y] = rot(A[x][y] ^ D[x]); B is indexed x + 5y after pi */                   \
    uint64_t B0 = xor3(A[0], C4, R1);                                                      \
    uint64_t B1 = rotl<44>(xor3(A[6], C0, R2));                                            \
    uint64_t B2 = rotl<43>(xor3(A[12], C1, R3));                                           \
    uint64_t B3 = rotl<21>(xor3(A[18], C2, R4));                                           \
    uint64_t B4 = rotl<14>(xor3(A[24], C3, R0));                                           \
    uint64_t B5 = rotl<28>(xor3(A[3], C2, R4));                                            \
    uint64_t B6 = rotl<20>(xor3(A[9], C3, R0));                                            \
    uint64_t B7 = rotl<3>(xor3(A[10], C4, R1));                                            \
    uint64_t B8 = rotl<45>(xor3(A[16], C0, R2));                                           \
    uint64_t B9 = rotl<61>(xor3(A[22], C1, R3));                                           \
    uint64_t B10 = rotl<1>(xor3(A[1], C0, R2));                                            \
    uint64_t B11 = rotl<6>(xor3(A[7], C1, R3));                                            \
    uint64_t B12 = rotl<25>(xor3(A[13], C2, R4));                                          \
    uint64_t B13 = rotl<8>(xor3(A[19], C3, R0));                                           \
    uint64_t B14 = rotl<18>(xor3(A[20], C4, R1));                                          \
    uint64_t B15 = rotl<27>(xor3(A[4], C3, R0));                                           \
    uint64_t B16 = rotl<36>(xor3(A[5], C4, R1));                                           \
    uint64_t B17 = rotl<10>(xor3(A[11], C0, R2));                                          \
    uint64_t B18 = rotl<15>(xor3(A[17], C1, R3));                                          \
    uint64_t B19 = rotl<56>(xor3(A[23], C2, R4));                                          \
    uint64_t B20 = rotl<62>(xor3(A[2], C1, R3));                                           \
    uint64_t B21 = rotl<55>(xor3(A[8], C2, R4));                                           \
    uint64_t B22 = rotl<39>(xor3(A[14], C3, R0));                                          \
    uint64_t B23 = rotl<41>(xor3(A[15], C4, R1));                                          \
    uint64_t B24 = rotl<2>(xor3(A[21], C0, R2));                                           \
    A[0] = chi(B0, B1, B2) ^ (RC);                                                         \
    A[1] = chi(B1, B2, B3);                                                                \
    A[2] = chi(B2, B3, B4);                                                                \
    A[3] = chi(B3, B4, B0);                                                                \
    A[4] = chi(B4, B0, B1);                                                                \
    A[5] = chi(B5, B6, B7);                                                                \
    A[6] = chi(B6, B7, B8);                                                                \
    A[7] = chi(B7, B8, B9);                                                                \
    A[8] = chi(B8, B9, B5);                                                                \
    A[9] = chi(B9, B5, B6);                                                                \
    A[10] = chi(B10, B11, B12);                                                            \
    A[11] = chi(B11, B12, B13);                                                            \
    A[12] = chi(B12, B13, B14);                                                            \
    A[13] = chi(B13, B14, B10);                                                            \
    A[14] = chi(B14, B10, B11);                                                            \
    A[15] = chi(B15, B16, B17);                                                            \
    A[16] = chi(B16, B17, B18);                                                            \
    A[17] = chi(B17, B18, B19);                                                            \
    A[18] = chi(B18, B19, B15);                                                            \
    A[19] = chi(B19, B15, B16);                                                            \
    A[20] = chi(B20, B21, B22);                                                            \
    A[21] = chi(B21, B22, B23);                                                            \
    A[22] = chi(B22, B23, B24);                                                            \
    A[23] = chi(B23, B24, B20);                                                            \
    A[24] = chi(B24, B20, B21);                                                            \
  } while (0)

// Full permutation.  The round loop is kept rolled by two so that each
// inlined call site stays small (the sponges below call it from many places).
__device__ __forceinline__ void keccakf(uint64_t A[25]) {
#pragma unroll 2
  for (int r = 0; r < 24; ++r) {
    DPKEM_KECCAK_ROUND(A, kKeccakRC[r]);
  }
}

__device__ __forceinline__ void keccak_zero(uint64_t A[25]) {
#pragma unroll
  for (int i = 0; i < 25; ++i) A[i] = 0;
}

// Rates in 64-bit words.
constexpr int kRateShake128 = 21;  // 168 B
constexpr int kRateShake256 = 17;  // 136 B
constexpr int kRateSha3_256 = 17;  // 136 B
constexpr int kRateSha3_512 = 9;   // 72 B

// Absorb `nblocks` full rate blocks of 64-bit little-endian words from p.
template <int RW>
__device__ __forceinline__ void absorb_blocks(uint64_t A[25], const uint64_t* __restrict__ p, int nblocks) {
  for (int b = 0; b < nblocks; ++b) {
#pragma unroll
    for (int i = 0; i < RW; ++i) A[i] ^= p[b * RW + i];
    keccakf(A);
  }
}

// XOR the padding for a message whose last block holds `pos_bytes` bytes.
template <int RW>
__device__ __forceinline__ void pad_at(uint64_t A[25], int pos_bytes, uint8_t dsbyte) {
  // pos_bytes is a compile-time constant at every call site
#pragma unroll
  for (int i = 0; i < RW; ++i)
    if (i == pos_bytes / 8) A[i] ^= (uint64_t)dsbyte << (8 * (pos_bytes % 8));
  A[RW - 1] ^= 0x8000000000000000ULL;
}

}  // namespace dpkem
