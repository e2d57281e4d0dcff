// Plain one-workgroup-per-KEM comparator forms of both schemes.
//
// These follow the reference's execution structure (FrodoKEM976-GPU/src/
// cuda_kernel.cu:21-335, SABER-GPU/src/cuda_kernel.cu:19-299): one kernel
// launch per step, every intermediate (A, S, E, B, raw XOF streams, staged
// hash inputs) round-trips through HBM, products are schoolbook u16
// multiply-accumulate with one thread per output (no MFMA, no dot products,
// no Toeplitz staging), and there is no fusion.  Hashes use the same
// lane-per-instance VGPR Keccak as the batched form, so the comparison
// isolates what fusion, MFMA and the dot-product kernels buy.  Unlike the
// reference, each instance uses its own inputs and every buffer it
// accumulates into is initialised.
#include "common.hpp"
#include "keccak.hpp"
#include "launch.hpp"
#include "sponge.hpp"
#include "prof.hpp"

namespace dpkem {
namespace plain {

static inline int lanes(int n) { return (n + 63) / 64; }

// ------------------------------------------------------------ generic hashing (lane per instance)

// out = SHAKE128(in[0:32], 8 * OUTW bytes)
template <int OUTW>
__global__ void __launch_bounds__(64) shake128_seed32(int n, const uint8_t* in, size_t is, uint8_t* out, size_t os) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  uint64_t m[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) m[k] = ld64(in + (size_t)i * is + 8 * k);
  shake128_regs_to_global<4, OUTW>(out + (size_t)i * os, m);
}

// out = SHA3-256(in[0 : 8 NW])
template <int NW>
__global__ void __launch_bounds__(64) sha3_256_k(int n, const uint8_t* in, size_t is, uint8_t* out, size_t os) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  uint64_t h[4];
  sha3_256_global<NW>(h, in + (size_t)i * is);
#pragma unroll
  for (int k = 0; k < 4; ++k) st64(out + (size_t)i * os + 8 * k, h[k]);
}

__global__ void __launch_bounds__(64) sha3_512_64_k(int n, const uint8_t* in, size_t is, uint8_t* out, size_t os) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  uint64_t m[8], h[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) m[k] = ld64(in + (size_t)i * is + 8 * k);
  sha3_512_64(h, m);
#pragma unroll
  for (int k = 0; k < 8; ++k) st64(out + (size_t)i * os + 8 * k, h[k]);
}

// out = SHAKE256(in[0 : 8 NW], 8 * OUTW bytes), OUTW <= 17
template <int NW, int OUTW>
__global__ void __launch_bounds__(64) shake256_k(int n, const uint8_t* in, size_t is, uint8_t* out, size_t os) {
  static_assert(OUTW <= kRateShake256, "one output block");
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  uint64_t A[25];
  absorb_global<kRateShake256, NW>(A, in + (size_t)i * is, kDsShake);
#pragma unroll
  for (int k = 0; k < OUTW; ++k) st64(out + (size_t)i * os + 8 * k, A[k]);
}

// out = SHAKE256(dom || seed[0:24], 8 * OUTW bytes)
template <int OUTW>
__global__ void __launch_bounds__(64) shake256_dom_seed(int n, uint64_t dom, const uint8_t* seed, size_t ss,
                                                        uint8_t* out, size_t os) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const uint8_t* s = seed + (size_t)i * ss;
  uint64_t A[25];
  keccak_zero(A);
  const uint64_t s0 = ld64(s), s1 = ld64(s + 8), s2 = ld64(s + 16);
  A[0] = dom | (s0 << 8);
  A[1] = (s0 >> 56) | (s1 << 8);
  A[2] = (s1 >> 56) | (s2 << 8);
  A[3] = (s2 >> 56) | (kDsShake << 8);
  A[kRateShake256 - 1] ^= kPadEnd;
  keccakf(A);
  uint8_t* o = out + (size_t)i * os;
  constexpr int NB = OUTW / kRateShake256, REM = OUTW % kRateShake256;
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int k = 0; k < kRateShake256; ++k) st64(o + 8 * (b * kRateShake256 + k), A[k]);
    keccakf(A);
  }
#pragma unroll
  for (int k = 0; k < REM; ++k) st64(o + 8 * (NB * kRateShake256 + k), A[k]);
}

// byte copy, one workgroup per instance
__global__ void __launch_bounds__(256) copy_k(int n, const uint8_t* src, size_t ss, uint8_t* dst, size_t ds, int len) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int k = threadIdx.x; k < len; k += 256) dst[(size_t)i * ds + k] = src[(size_t)i * ss + k];
}

// ------------------------------------------------------------ Saber steps (256 threads per instance)

__global__ void __launch_bounds__(256) sab_unpack13(int n, const uint8_t* in, size_t is, uint16_t* out, int npoly) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const uint8_t* b = in + (size_t)i * is;
  for (int c = threadIdx.x; c < 256 * npoly; c += 256) {
    uint32_t v = 0;
    for (int k = 0; k < 13; ++k) v |= (uint32_t)((b[(13 * c + k) >> 3] >> ((13 * c + k) & 7)) & 1) << k;
    out[(size_t)i * 256 * npoly + c] = (uint16_t)v;
  }
}

__global__ void __launch_bounds__(256) sab_unpack10(int n, const uint8_t* in, size_t is, uint16_t* out) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const uint8_t* b = in + (size_t)i * is;
  for (int c = threadIdx.x; c < 768; c += 256) {
    uint32_t v = 0;
    for (int k = 0; k < 10; ++k) v |= (uint32_t)((b[(10 * c + k) >> 3] >> ((10 * c + k) & 7)) & 1) << k;
    out[(size_t)i * 768 + c] = (uint16_t)v;
  }
}

__global__ void __launch_bounds__(256) sab_cbd(int n, const uint8_t* sraw, uint16_t* s) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < 768; c += 256) {
    const uint32_t byte = sraw[(size_t)i * 768 + c];
    s[(size_t)i * 768 + c] = (uint16_t)(popc(byte & 0xF) - popc(byte >> 4));
  }
}

// mode 0: out[r] = sum_j A[r][j] s[j]; mode 1: A^T; mode 2: out = sum_j a[j] s[j] (inner)
__global__ void __launch_bounds__(256) sab_polymul(int n, const uint16_t* a, const uint16_t* s, uint16_t* out, int mode) {
  __shared__ uint16_t La[9 * 256];
  __shared__ uint16_t Ls[3 * 256];
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  const int na = mode == 2 ? 3 : 9;
  for (int c = t; c < na * 256; c += 256) La[c] = a[(size_t)i * na * 256 + c];
  for (int c = t; c < 768; c += 256) Ls[c] = s[(size_t)i * 768 + c];
  __syncthreads();
  const int rows = mode == 2 ? 1 : 3;
  for (int r = 0; r < rows; ++r) {
    uint16_t acc = 0;
    for (int j = 0; j < 3; ++j) {
      const uint16_t* p = La + 256 * (mode == 2 ? j : (mode == 1 ? 3 * j + r : 3 * r + j));
      const uint16_t* q = Ls + 256 * j;
      for (int k = 0; k <= t; ++k) acc = (uint16_t)(acc + p[t - k] * q[k]);
      for (int k = t + 1; k < 256; ++k) acc = (uint16_t)(acc - p[256 + t - k] * q[k]);
    }
    out[(size_t)i * rows * 256 + r * 256 + t] = acc;
  }
}

// (x + h1) >> 3, 10-bit pack: 768 coefficients -> 960 bytes at dst
__global__ void __launch_bounds__(256) sab_round_pack10(int n, const uint16_t* b, uint8_t* dst, size_t ds) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int g = threadIdx.x; g < 192; g += 256) {
    uint64_t x = 0;
    for (int k = 0; k < 4; ++k) x |= (uint64_t)(((b[(size_t)i * 768 + 4 * g + k] + 4u) >> 3) & 0x3FF) << (10 * k);
    for (int k = 0; k < 5; ++k) dst[(size_t)i * ds + 5 * g + k] = (uint8_t)(x >> (8 * k));
  }
}

// c_m = (v - m 2^9 + h1) >> 6, 4-bit pack -> ct[960:1088]
__global__ void __launch_bounds__(256) sab_cm(int n, const uint16_t* v, const uint8_t* m, size_t ms, uint8_t* ct) {
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n || t >= 128) return;
  uint32_t out = 0;
  for (int k = 0; k < 2; ++k) {
    const int c = 2 * t + k;
    const uint32_t mb = (m[(size_t)i * ms + c / 8] >> (c % 8)) & 1;
    out |= (((v[(size_t)i * 256 + c] - (mb << 9) + 4u) >> 6) & 0xF) << (4 * k);
  }
  ct[(size_t)i * kSaberCt + kSaberPolyVecCompBytes + t] = (uint8_t)out;
}

// m'_c = ((v + h2 - c_m 2^6) >> 9) & 1
__global__ void __launch_bounds__(256) sab_msg(int n, const uint16_t* v, const uint8_t* ct, uint8_t* m, size_t ms) {
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n || t >= 32) return;
  uint32_t byte = 0;
  for (int k = 0; k < 8; ++k) {
    const int c = 8 * t + k;
    const uint32_t cm = (ct[(size_t)i * kSaberCt + kSaberPolyVecCompBytes + c / 2] >> (4 * (c & 1))) & 0xF;
    byte |= (((v[(size_t)i * 256 + c] + 228u - (cm << 6)) >> 9) & 1u) << k;
  }
  m[(size_t)i * ms + t] = (uint8_t)byte;
}

__global__ void __launch_bounds__(256) sab_pack13(int n, const uint16_t* s, uint8_t* dst, size_t ds) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int g = threadIdx.x; g < 96; g += 256) {
    uint64_t lo = 0, hi = 0;
    for (int k = 0; k < 8; ++k) {
      const uint64_t v = s[(size_t)i * 768 + 8 * g + k] & 0x1FFF;
      const int bit = 13 * k;
      if (bit < 64) lo |= v << bit;
      if (bit + 13 > 64) hi |= bit >= 64 ? v << (bit - 64) : v >> (64 - bit);
    }
    for (int k = 0; k < 8; ++k) dst[(size_t)i * ds + 13 * g + k] = (uint8_t)(lo >> (8 * k));
    for (int k = 0; k < 5; ++k) dst[(size_t)i * ds + 13 * g + 8 + k] = (uint8_t)(hi >> (8 * k));
  }
}

__global__ void __launch_bounds__(256) verify_k(int n, const uint8_t* a, size_t as, const uint8_t* b, size_t bs, int len,
                                                uint32_t* fail) {
  __shared__ uint32_t d;
  const int i = blockIdx.x;
  if (i >= n) return;
  if (threadIdx.x == 0) d = 0;
  __syncthreads();
  uint32_t x = 0;
  for (int k = threadIdx.x; k < len; k += 256) x |= a[(size_t)i * as + k] ^ b[(size_t)i * bs + k];
  if (x) atomicOr(&d, 1u);
  __syncthreads();
  if (threadIdx.x == 0) fail[i] = d;
}

// K = SHA3-256(select(kr[0:32], z, fail) || hct)
__global__ void __launch_bounds__(64) sab_kdf(int n, const uint8_t* kr, const uint8_t* hct, const uint8_t* sk,
                                              const uint32_t* fail, uint8_t* ss) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  uint64_t m[8];
  const uint64_t mask = fail ? 0 - (uint64_t)(fail[i] != 0) : 0;
  for (int k = 0; k < 4; ++k) {
    const uint64_t z = sk ? ld64(sk + (size_t)i * kSaberSk + kSaberSk - 32 + 8 * k) : 0;
    m[k] = (ld64(kr + (size_t)i * 64 + 8 * k) & ~mask) | (z & mask);
    m[4 + k] = ld64(hct + (size_t)i * 32 + 8 * k);
  }
  uint64_t h[4];
  sha3_256_regs<8>(h, m);
  for (int k = 0; k < 4; ++k) st64(ss + (size_t)i * kSaberSs + 8 * k, h[k]);
}

// ------------------------------------------------------------ Frodo steps

__constant__ static const uint16_t kCdf[10] = {5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766};

// in place on cnt little-endian u16 values per instance
__global__ void __launch_bounds__(256) fr_sample(int n, uint16_t* r, size_t rs, int cnt) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < cnt; c += 256) {
    const uint32_t x = r[(size_t)i * rs + c];
    const uint32_t t = x >> 1, sg = x & 1;
    uint32_t e = 0;
    for (int j = 0; j < 10; ++j) e += ((uint32_t)kCdf[j] - t) >> 31;
    r[(size_t)i * rs + c] = (uint16_t)((e ^ (0u - sg)) + sg);
  }
}

// A (976 x 976 u16, row-major) for each instance: one thread per row
__global__ void __launch_bounds__(1024) fr_gen_a(int n, const uint8_t* pk, size_t ps, uint16_t* A) {
  const int i = blockIdx.x, row = threadIdx.x;
  if (i >= n || row >= kFrodoN) return;
  const uint8_t* p = pk + (size_t)i * ps;
  const uint64_t s0 = ld64(p), s1 = ld64(p + 8);
  uint64_t St[25];
  keccak_zero(St);
  St[0] = (uint64_t)row | (s0 << 16);
  St[1] = (s0 >> 48) | (s1 << 16);
  St[2] = (s1 >> 48) | (kDsShake << 16);
  St[kRateShake128 - 1] ^= kPadEnd;
  uint64_t* dst = reinterpret_cast<uint64_t*>(A + ((size_t)i * kFrodoN + row) * kFrodoN);
  for (int b = 0; b < 12; ++b) {
    keccakf(St);
#pragma unroll
    for (int k = 0; k < kRateShake128; ++k)
      if (b * kRateShake128 + k < kFrodoN / 4) dst[b * kRateShake128 + k] = St[k];
  }
}

// B = A S + E (S^T given as NBAR x N).  A is read along its rows, each row
// once, neighbouring lanes reading neighbouring entries: a wave takes 8
// rows at a time, lane l accumulates the 64 sums (row r0 + rr, column k)
// over entries 4 l .. 4 l + 3 (+ 256 m) of those rows, and a butterfly over
// the wave then leaves lane L with sum L (each step sends the half of its
// sums a lane does not keep).
__global__ void __launch_bounds__(1024) fr_as(int n, const uint16_t* A, const uint16_t* st, size_t sts, uint16_t* B) {
  __shared__ __align__(8) uint16_t Ls[kFrodoNbar * kFrodoN];
  static_assert(kFrodoN % 8 == 0 && kFrodoN % 4 == 0, "8-row groups, 4-entry reads");
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  const uint16_t* e = st + kFrodoNbar * kFrodoN;  // E follows S^T
  for (int c = t; c < kFrodoNbar * kFrodoN; c += 1024) Ls[c] = st[(size_t)i * sts + c];
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  const uint16_t* Ai = A + (size_t)i * kFrodoN * kFrodoN;
  for (int r0 = 8 * wave; r0 < kFrodoN; r0 += 8 * 16) {
    uint32_t acc[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) acc[q] = 0;
    for (int j = 4 * lane; j < kFrodoN; j += 256) {
      uint32_t s[kFrodoNbar][4];
#pragma unroll
      for (int k = 0; k < kFrodoNbar; ++k) {
        const uint2 w = *reinterpret_cast<const uint2*>(Ls + k * kFrodoN + j);
        s[k][0] = w.x & 0xFFFF, s[k][1] = w.x >> 16, s[k][2] = w.y & 0xFFFF, s[k][3] = w.y >> 16;
      }
#pragma unroll
      for (int rr = 0; rr < 8; ++rr) {
        const uint2 w = *reinterpret_cast<const uint2*>(Ai + (size_t)(r0 + rr) * kFrodoN + j);
        const uint32_t x[4] = {w.x & 0xFFFF, w.x >> 16, w.y & 0xFFFF, w.y >> 16};
#pragma unroll
        for (int k = 0; k < kFrodoNbar; ++k)
          acc[8 * rr + k] += x[0] * s[k][0] + x[1] * s[k][1] + x[2] * s[k][2] + x[3] * s[k][3];
      }
    }
#pragma unroll
    for (int m = 64; m > 1; m >>= 1) {
      const int d = m >> 1;
      const bool up = (lane & d) != 0;
#pragma unroll
      for (int q = 0; q < d; ++q) {
        const uint32_t keep = up ? acc[q + d] : acc[q];
        const uint32_t send = up ? acc[q] : acc[q + d];
        acc[q] = keep + (uint32_t)__shfl_xor((int)send, d, 64);
      }
    }
    const int row = r0 + (lane >> 3), k = lane & 7;
    B[(size_t)i * kFrodoNbar * kFrodoN + row * kFrodoNbar + k] =
        (uint16_t)(acc[0] + e[(size_t)i * sts + row * kFrodoNbar + k]);
  }
}

// B' = S' A + E': thread per column
__global__ void __launch_bounds__(1024) fr_sa(int n, const uint16_t* A, const uint16_t* sp, size_t sps, uint16_t* Bp) {
  __shared__ uint16_t Ls[kFrodoNbar * kFrodoN];
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  for (int c = t; c < kFrodoNbar * kFrodoN; c += 1024) Ls[c] = sp[(size_t)i * sps + c];
  __syncthreads();
  if (t >= kFrodoN) return;
  uint16_t acc[kFrodoNbar];
  for (int k = 0; k < kFrodoNbar; ++k) acc[k] = sp[(size_t)i * sps + kFrodoNbar * kFrodoN + k * kFrodoN + t];  // E'
  const uint16_t* a = A + (size_t)i * kFrodoN * kFrodoN + t;
  for (int r = 0; r < kFrodoN; ++r) {
    const uint16_t x = a[(size_t)r * kFrodoN];
#pragma unroll
    for (int k = 0; k < kFrodoNbar; ++k) acc[k] = (uint16_t)(acc[k] + Ls[k * kFrodoN + r] * x);
  }
  for (int k = 0; k < kFrodoNbar; ++k) Bp[(size_t)i * kFrodoNbar * kFrodoN + k * kFrodoN + t] = acc[k];
}

// big-endian 16-bit pack / unpack
__global__ void __launch_bounds__(256) fr_pack(int n, const uint16_t* v, size_t vs, int cnt, uint8_t* dst, size_t ds) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < cnt; c += 256) {
    const uint16_t x = v[(size_t)i * vs + c];
    dst[(size_t)i * ds + 2 * c] = (uint8_t)(x >> 8);
    dst[(size_t)i * ds + 2 * c + 1] = (uint8_t)x;
  }
}
__global__ void __launch_bounds__(256) fr_unpack(int n, const uint8_t* src, size_t ss, int cnt, uint16_t* v, size_t vs) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < cnt; c += 256)
    v[(size_t)i * vs + c] = (uint16_t)((src[(size_t)i * ss + 2 * c] << 8) | src[(size_t)i * ss + 2 * c + 1]);
}
// little-endian u16 from bytes (sk's S^T) and to bytes
__global__ void __launch_bounds__(256) fr_le_load(int n, const uint8_t* src, size_t ss, int cnt, uint16_t* v, size_t vs) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < cnt; c += 256)
    v[(size_t)i * vs + c] = (uint16_t)(src[(size_t)i * ss + 2 * c] | (src[(size_t)i * ss + 2 * c + 1] << 8));
}
__global__ void __launch_bounds__(256) fr_le_store(int n, const uint16_t* v, size_t vs, int cnt, uint8_t* dst, size_t ds) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int c = threadIdx.x; c < cnt; c += 256) {
    dst[(size_t)i * ds + 2 * c] = (uint8_t)v[(size_t)i * vs + c];
    dst[(size_t)i * ds + 2 * c + 1] = (uint8_t)(v[(size_t)i * vs + c] >> 8);
  }
}

// C = S' B + E'' + Encode(mu): thread per output (64)
__global__ void __launch_bounds__(64) fr_sb(int n, const uint16_t* sp, size_t sps, const uint16_t* B, const uint8_t* mu,
                                            size_t ms, uint16_t* C) {
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  const int r = t >> 3, c = t & 7;
  uint16_t acc = sp[(size_t)i * sps + 2 * kFrodoNbar * kFrodoN + t];  // E''
  for (int k = 0; k < kFrodoN; ++k)
    acc = (uint16_t)(acc + sp[(size_t)i * sps + r * kFrodoN + k] * B[(size_t)i * kFrodoNbar * kFrodoN + k * kFrodoNbar + c]);
  const uint8_t* m = mu + (size_t)i * ms + 3 * r;
  const uint32_t t24 = m[0] | ((uint32_t)m[1] << 8) | ((uint32_t)m[2] << 16);
  C[(size_t)i * 64 + t] = (uint16_t)(acc + (((t24 >> (3 * c)) & 7u) << 13));
}

// mu' = Decode(C - B' S): thread per output
__global__ void __launch_bounds__(64) fr_bs_decode(int n, const uint16_t* Bp, const uint16_t* st, const uint16_t* C,
                                                   uint8_t* mu, size_t ms) {
  __shared__ uint16_t W[64];
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= n) return;
  const int r = t >> 3, c = t & 7;
  uint16_t acc = 0;
  for (int k = 0; k < kFrodoN; ++k)
    acc = (uint16_t)(acc + Bp[(size_t)i * kFrodoNbar * kFrodoN + r * kFrodoN + k] *
                               st[(size_t)i * kFrodoNbar * kFrodoN + c * kFrodoN + k]);
  W[t] = (uint16_t)(C[(size_t)i * 64 + t] - acc);
  __syncthreads();
  if (t < 8) {
    uint32_t word = 0;
    for (int j = 0; j < 8; ++j) word |= ((((uint32_t)W[8 * t + j] + (1u << 12)) >> 13) & 7u) << (3 * j);
    uint8_t* m = mu + (size_t)i * ms + 3 * t;
    m[0] = (uint8_t)word;
    m[1] = (uint8_t)(word >> 8);
    m[2] = (uint8_t)(word >> 16);
  }
}

// Fin[15744:15768] = fail ? s : k
__global__ void __launch_bounds__(64) fr_select(int n, const uint8_t* k, size_t ks, const uint8_t* sk, const uint32_t* fail,
                                                uint8_t* fin, size_t fs) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const uint64_t mask = 0 - (uint64_t)(fail[i] != 0);
  for (int w = 0; w < 3; ++w)
    st64(fin + (size_t)i * fs + kFrodoCt + 8 * w,
         (ld64(k + (size_t)i * ks + 8 * w) & ~mask) | (ld64(sk + (size_t)i * kFrodoSk + 8 * w) & mask));
}

}  // namespace plain

// ------------------------------------------------------------ sequencing

using namespace plain;
#define L64(kern, ...) DPKEM_LAUNCH("plain", kern, dim3(lanes(n)), dim3(64), st, n, __VA_ARGS__)
#define LWG(kern, thr, ...) DPKEM_LAUNCH("plain", kern, dim3(n), dim3(thr), st, n, __VA_ARGS__)

size_t plain_scratch_per_inst(int scheme) {
  if (scheme == 1) return 16384;  // Saber (see carve below)
  return (size_t)kFrodoN * kFrodoN * 2 + 131072;
}

namespace {
struct Cut {
  uint8_t* p;
  size_t per;  // bytes per instance of the current array
  uint8_t* take(int n, size_t bytes) {
    uint8_t* r = p;
    p += ((size_t)n * bytes + 255) & ~(size_t)255;
    return r;
  }
};
}  // namespace

void plain_saber_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch) {
  Cut c{scratch, 0};
  uint8_t* araw = c.take(n, 3744);
  uint8_t* sraw = c.take(n, 768);
  uint16_t* A = reinterpret_cast<uint16_t*>(c.take(n, 9 * 512));
  uint16_t* s = reinterpret_cast<uint16_t*>(c.take(n, 3 * 512));
  uint16_t* b = reinterpret_cast<uint16_t*>(c.take(n, 3 * 512));
  uint16_t* bb = reinterpret_cast<uint16_t*>(c.take(n, 3 * 512));
  uint16_t* v = reinterpret_cast<uint16_t*>(c.take(n, 512));
  uint8_t* buf = c.take(n, 64);
  uint8_t* kr = c.take(n, 64);
  uint8_t* hct = c.take(n, 32);
  uint8_t* ct2 = c.take(n, kSaberCt);
  uint32_t* fail = reinterpret_cast<uint32_t*>(c.take(n, 4));
  // IND-CPA encryption of message `m` (stride ms) with noise seed kr+32 under pk (stride ps) into ct
  auto encrypt = [&](const uint8_t* pk, size_t ps, const uint8_t* m, size_t ms, uint8_t* ct) {
    L64(shake128_seed32<96>, kr + 32, (size_t)64, sraw, (size_t)768);
    L64(shake128_seed32<468>, pk + kSaberPolyVecCompBytes, ps, araw, (size_t)3744);
    LWG(sab_unpack13, 256, araw, (size_t)3744, A, 9);
    LWG(sab_cbd, 256, sraw, s);
    LWG(sab_polymul, 256, A, s, b, 0);
    LWG(sab_round_pack10, 256, b, ct, (size_t)kSaberCt);
    LWG(sab_unpack10, 256, pk, ps, bb);
    LWG(sab_polymul, 256, bb, s, v, 2);
    LWG(sab_cm, 256, v, m, ms, ct);
  };
  if (op == 0) {  // keygen: i0 coins -> o0 pk, o1 sk
    uint8_t* pk = o0;
    uint8_t* sk = o1;
    L64(shake128_seed32<4>, i0, (size_t)96, pk + kSaberPolyVecCompBytes, (size_t)kSaberPk);
    L64(shake128_seed32<468>, pk + kSaberPolyVecCompBytes, (size_t)kSaberPk, araw, (size_t)3744);
    L64(shake128_seed32<96>, i0 + 32, (size_t)96, sraw, (size_t)768);
    LWG(sab_unpack13, 256, araw, (size_t)3744, A, 9);
    LWG(sab_cbd, 256, sraw, s);
    LWG(sab_polymul, 256, A, s, b, 1);
    LWG(sab_round_pack10, 256, b, pk, (size_t)kSaberPk);
    LWG(sab_pack13, 256, s, sk, (size_t)kSaberSk);
    LWG(copy_k, 256, pk, (size_t)kSaberPk, sk + kSaberPolyVecBytes, (size_t)kSaberSk, kSaberPk);
    L64(sha3_256_k<kSaberPk / 8>, pk, (size_t)kSaberPk, sk + kSaberSk - 64, (size_t)kSaberSk);
    LWG(copy_k, 256, i0 + 64, (size_t)96, sk + kSaberSk - 32, (size_t)kSaberSk, 32);
  } else if (op == 1) {  // enc: i0 pk, i1 coins -> o0 ct, o1 ss
    L64(sha3_256_k<4>, i1, (size_t)32, buf, (size_t)64);
    L64(sha3_256_k<kSaberPk / 8>, i0, (size_t)kSaberPk, buf + 32, (size_t)64);
    L64(sha3_512_64_k, buf, (size_t)64, kr, (size_t)64);
    encrypt(i0, kSaberPk, buf, 64, o0);
    L64(sha3_256_k<kSaberCt / 8>, o0, (size_t)kSaberCt, hct, (size_t)32);
    L64(sab_kdf, kr, hct, (const uint8_t*)nullptr, (const uint32_t*)nullptr, o1);
  } else {  // dec: i0 ct, i1 sk -> o0 ss
    LWG(sab_unpack13, 256, i1, (size_t)kSaberSk, s, 3);
    LWG(sab_unpack10, 256, i0, (size_t)kSaberCt, bb);
    LWG(sab_polymul, 256, bb, s, v, 2);
    LWG(sab_msg, 256, v, i0, buf, (size_t)64);
    LWG(copy_k, 256, i1 + kSaberSk - 64, (size_t)kSaberSk, buf + 32, (size_t)64, 32);
    L64(sha3_512_64_k, buf, (size_t)64, kr, (size_t)64);
    encrypt(i1 + kSaberPolyVecBytes, kSaberSk, buf, 64, ct2);
    LWG(verify_k, 256, i0, (size_t)kSaberCt, ct2, (size_t)kSaberCt, kSaberCt, fail);
    L64(sha3_256_k<kSaberCt / 8>, i0, (size_t)kSaberCt, hct, (size_t)32);
    L64(sab_kdf, kr, hct, i1, fail, o0);
  }
}

void plain_frodo_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch) {
  constexpr int M = kFrodoN * kFrodoNbar;
  constexpr size_t kSp = (size_t)(2 * M + 64) * 2;  // S' | E' | E'' as u16 (also keygen S^T | E)
  Cut c{scratch, 0};
  uint16_t* A = reinterpret_cast<uint16_t*>(c.take(n, (size_t)kFrodoN * kFrodoN * 2));
  uint16_t* sp = reinterpret_cast<uint16_t*>(c.take(n, kSp));
  uint16_t* B = reinterpret_cast<uint16_t*>(c.take(n, M * 2));
  uint16_t* Bp = reinterpret_cast<uint16_t*>(c.take(n, M * 2));
  uint16_t* Bpp = reinterpret_cast<uint16_t*>(c.take(n, M * 2));
  uint16_t* S = reinterpret_cast<uint16_t*>(c.take(n, M * 2));
  uint16_t* C = reinterpret_cast<uint16_t*>(c.take(n, 128));
  uint16_t* CC = reinterpret_cast<uint16_t*>(c.take(n, 128));
  uint8_t* g2in = c.take(n, 48);
  uint8_t* g2out = c.take(n, 48);
  uint8_t* fin = c.take(n, kFrodoCt + 24);
  uint8_t* ct2 = c.take(n, kFrodoCt);
  uint32_t* fail = reinterpret_cast<uint32_t*>(c.take(n, 4));
  // S' E' E'' from seedSE (stride 48 in g2out), then B'' = S' A + E' and C = S' B + E'' + enc(mu)
  auto encrypt = [&](const uint8_t* pk, size_t ps, const uint8_t* mu, size_t ms, uint16_t* bp_out, uint16_t* c_out) {
    L64(shake256_dom_seed<(2 * M + 64) / 4>, (uint64_t)0x96, g2out, (size_t)48, (uint8_t*)sp, kSp);
    LWG(fr_sample, 256, sp, kSp / 2, 2 * M + 64);
    LWG(fr_gen_a, 1024, pk, ps, A);
    LWG(fr_sa, 1024, A, sp, kSp / 2, bp_out);
    LWG(fr_unpack, 256, pk + kFrodoSeedA, ps, M, B, (size_t)M);
    LWG(fr_sb, 64, sp, kSp / 2, B, mu, ms, c_out);
  };
  if (op == 0) {  // keygen: i0 coins -> o0 pk, o1 sk
    uint8_t* pk = o0;
    uint8_t* sk = o1;
    L64((shake256_k<2, 2>), i0 + 48, (size_t)64, pk, (size_t)kFrodoPk);
    L64(shake256_dom_seed<2 * M / 4>, (uint64_t)0x5F, i0 + 24, (size_t)64, (uint8_t*)sp, kSp);
    LWG(fr_sample, 256, sp, kSp / 2, 2 * M);
    LWG(fr_gen_a, 1024, pk, (size_t)kFrodoPk, A);
    LWG(fr_as, 1024, A, sp, kSp / 2, B);
    LWG(fr_pack, 256, B, (size_t)M, M, pk + kFrodoSeedA, (size_t)kFrodoPk);
    LWG(copy_k, 256, i0, (size_t)64, sk, (size_t)kFrodoSk, kFrodoSec);
    LWG(copy_k, 256, pk, (size_t)kFrodoPk, sk + kFrodoSec, (size_t)kFrodoSk, kFrodoPk);
    LWG(fr_le_store, 256, sp, kSp / 2, M, sk + kFrodoSec + kFrodoPk, (size_t)kFrodoSk);
    L64((shake256_k<kFrodoPk / 8, 3>), pk, (size_t)kFrodoPk, sk + kFrodoSk - kFrodoSec, (size_t)kFrodoSk);
  } else if (op == 1) {  // enc: i0 pk, i1 mu -> o0 ct, o1 ss
    L64((shake256_k<kFrodoPk / 8, 3>), i0, (size_t)kFrodoPk, g2in, (size_t)48);
    LWG(copy_k, 256, i1, (size_t)24, g2in + 24, (size_t)48, 24);
    L64((shake256_k<6, 6>), g2in, (size_t)48, g2out, (size_t)48);
    encrypt(i0, kFrodoPk, i1, 24, Bp, C);
    LWG(fr_pack, 256, Bp, (size_t)M, M, o0, (size_t)kFrodoCt);
    LWG(fr_pack, 256, C, (size_t)64, 64, o0 + 2 * M, (size_t)kFrodoCt);
    LWG(copy_k, 256, o0, (size_t)kFrodoCt, fin, (size_t)(kFrodoCt + 24), kFrodoCt);
    LWG(copy_k, 256, g2out + 24, (size_t)48, fin + kFrodoCt, (size_t)(kFrodoCt + 24), 24);
    L64((shake256_k<(kFrodoCt + 24) / 8, 3>), fin, (size_t)(kFrodoCt + 24), o1, (size_t)kFrodoSs);
  } else {  // dec: i0 ct, i1 sk -> o0 ss
    LWG(fr_unpack, 256, i0, (size_t)kFrodoCt, M, Bp, (size_t)M);
    LWG(fr_unpack, 256, i0 + 2 * M, (size_t)kFrodoCt, 64, C, (size_t)64);
    LWG(fr_le_load, 256, i1 + kFrodoSec + kFrodoPk, (size_t)kFrodoSk, M, S, (size_t)M);
    LWG(fr_bs_decode, 64, Bp, S, C, g2in + 24, (size_t)48);
    LWG(copy_k, 256, i1 + kFrodoSk - kFrodoSec, (size_t)kFrodoSk, g2in, (size_t)48, 24);
    L64((shake256_k<6, 6>), g2in, (size_t)48, g2out, (size_t)48);
    encrypt(i1 + kFrodoSec, kFrodoSk, g2in + 24, 48, Bpp, CC);
    LWG(fr_pack, 256, Bpp, (size_t)M, M, ct2, (size_t)kFrodoCt);
    LWG(fr_pack, 256, CC, (size_t)64, 64, ct2 + 2 * M, (size_t)kFrodoCt);
    LWG(verify_k, 256, i0, (size_t)kFrodoCt, ct2, (size_t)kFrodoCt, kFrodoCt, fail);
    LWG(copy_k, 256, i0, (size_t)kFrodoCt, fin, (size_t)(kFrodoCt + 24), kFrodoCt);
    L64(fr_select, g2out + 24, (size_t)48, i1, fail, fin, (size_t)(kFrodoCt + 24));
    L64((shake256_k<(kFrodoCt + 24) / 8, 3>), fin, (size_t)(kFrodoCt + 24), o0, (size_t)kFrodoSs);
  }
}

}  // namespace dpkem
