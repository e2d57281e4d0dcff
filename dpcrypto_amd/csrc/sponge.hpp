// Lane-per-instance sponges over the VGPR Keccak permutation.
//
// Each lane runs one complete hash for its own instance: the message is
// streamed from that instance's slice of global memory in 64-bit words and
// the state never leaves registers.  Message lengths are compile-time
// constants at every call site (all KEM buffers are multiples of 8 bytes), so
// block counts and padding positions are resolved statically.
#pragma once
#include "common.hpp"
#include "keccak.hpp"

namespace dpkem {

constexpr uint64_t kPadEnd = 0x8000000000000000ULL;
constexpr uint64_t kDsSha3 = 0x06;
constexpr uint64_t kDsShake = 0x1F;

// Absorb NW message words from p with rate RW, pad with domain byte DS and
// run the final permutation (the state then holds the first output block).
template <int RW, int NW>
__device__ __forceinline__ void absorb_global(uint64_t A[25], const uint8_t* __restrict__ p, uint64_t ds) {
  constexpr int NB = NW / RW, REM = NW % RW;
  keccak_zero(A);
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int i = 0; i < RW; ++i) A[i] ^= ld64(p + 8 * (b * RW + i));
    keccakf(A);
  }
#pragma unroll
  for (int i = 0; i < REM; ++i) A[i] ^= ld64(p + 8 * (NB * RW + i));
  A[REM] ^= ds;
  A[RW - 1] ^= kPadEnd;
  keccakf(A);
}

// Absorb a short register-resident message of NW (< RW) words.
template <int RW, int NW>
__device__ __forceinline__ void absorb_regs(uint64_t A[25], const uint64_t* m, uint64_t ds) {
  static_assert(NW < RW, "single block");
  keccak_zero(A);
#pragma unroll
  for (int i = 0; i < NW; ++i) A[i] = m[i];
  A[NW] ^= ds;
  A[RW - 1] ^= kPadEnd;
  keccakf(A);
}

// SHA3-256 of NW global words -> 4 words.
template <int NW>
__device__ __forceinline__ void sha3_256_global(uint64_t out[4], const uint8_t* p) {
  uint64_t A[25];
  absorb_global<kRateSha3_256, NW>(A, p, kDsSha3);
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// SHA3-256 of NW register words -> 4 words.
template <int NW>
__device__ __forceinline__ void sha3_256_regs(uint64_t out[4], const uint64_t* m) {
  uint64_t A[25];
  absorb_regs<kRateSha3_256, NW>(A, m, kDsSha3);
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = A[i];
}

// SHA3-512 of 8 register words (64 bytes) -> 8 words.
__device__ __forceinline__ void sha3_512_64(uint64_t out[8], const uint64_t m[8]) {
  uint64_t A[25];
  absorb_regs<kRateSha3_512, 8>(A, m, kDsSha3);
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = A[i];
}

// SHAKE128 of a register message of NW words, squeezing OUTW words to dst.
template <int NW, int OUTW>
__device__ __forceinline__ void shake128_regs_to_global(uint8_t* __restrict__ dst, const uint64_t* m) {
  constexpr int RW = kRateShake128;
  constexpr int NB = OUTW / RW, REM = OUTW % RW;
  uint64_t A[25];
  absorb_regs<RW, NW>(A, m, kDsShake);
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int i = 0; i < RW; ++i) st64(dst + 8 * (b * RW + i), A[i]);
    keccakf(A);
  }
#pragma unroll
  for (int i = 0; i < REM; ++i) st64(dst + 8 * (NB * RW + i), A[i]);
}

}  // namespace dpkem
