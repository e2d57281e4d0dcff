// Keccak-f[1600] split over a lane pair, for latency-bound serial sponges.
//
// A hash chain (SHAKE stream, H(pk), ...) is inherently serial, and at the
// batch sizes of interest (4096 instances = 64 waves) one state per lane
// leaves most SIMDs idle, so the chain's wall time is the per-wave
// instruction count of a permutation.  Here each 64-bit Keccak lane word is
// bit-interleaved: lane 2k holds its even bits and lane 2k+1 its odd bits
// (25 x 32-bit per lane).  In that representation
//   rot64(w, 2s)     = (rot32(e, s),     rot32(o, s))
//   rot64(w, 2s + 1) = (rot32(o, s + 1), rot32(e, s))
// so XOR / AND / NOT stay local and an odd rotation is one partner exchange
// (DPP quad_perm [1,0,3,2]) plus a rotate whose amount depends on the lane
// parity.  A round is ~109 VALU ops per lane instead of 180, i.e. ~1.65x
// lower latency per permutation.
//
// Everything is written over a lane-vector type V so that the identical
// round code runs on the device (V = one lane, partner via DPP) and in a
// host emulation of a lane pair used by the tests.
#pragma once
#include <stdint.h>

#if defined(__HIP__) || defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DPKEM_HD __host__ __device__
#else
#define DPKEM_HD
#endif

namespace dpkem {
namespace k2 {

// Bit-interleaved round constants: {even half, odd half}.
constexpr uint32_t compress_even(uint64_t x) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= (uint32_t)((x >> (2 * i)) & 1) << i;
  return r;
}
constexpr uint64_t kRCs[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

// Lane-pair round over an abstract lane vector.  Ops required from OPS:
//   xor3(a,b,c), chi(a,b,c) = a ^ (~b & c), rotc<N>(x) (uniform rotate left by N),
//   rotv(x, sh) (rotate left by the per-lane shift operand sh),
//   swap(x) (partner lane's value).
// sh[k] is the per-lane shift operand that rotates left by (k - p), where p
// is the lane parity (0 even bits, 1 odd bits); it is computed once per
// permutation by OPS::shift(p, k).
template <class OPS, class V>
inline __attribute__((always_inline)) DPKEM_HD void round2(V a[25], const V* sh, const V& rc) {
  using O = OPS;
  V C0 = O::xor3(O::xor3(a[0], a[5], a[10]), a[15], a[20]);
  V C1 = O::xor3(O::xor3(a[1], a[6], a[11]), a[16], a[21]);
  V C2 = O::xor3(O::xor3(a[2], a[7], a[12]), a[17], a[22]);
  V C3 = O::xor3(O::xor3(a[3], a[8], a[13]), a[18], a[23]);
  V C4 = O::xor3(O::xor3(a[4], a[9], a[14]), a[19], a[24]);
  // rot64(C, 1): partner half rotated by (1 - p)
  const V r1 = sh[1];  // rotate by (1 - p): odd offset 2s+1 with s = 0
  V R0 = O::rotv(O::swap(C0), r1), R1 = O::rotv(O::swap(C1), r1), R2 = O::rotv(O::swap(C2), r1),
    R3 = O::rotv(O::swap(C3), r1), R4 = O::rotv(O::swap(C4), r1);
  // D[x] = C[x-1] ^ rot(C[x+1], 1);   T = A[x][y] ^ D[x]
#define K2_EVEN(dst, src, c, r, N) V dst = O::template rotc<N>(O::xor3(src, c, r))
#define K2_ODD(dst, src, c, r, S) V dst = O::rotv(O::swap(O::xor3(src, c, r)), sh[(S) + 1])
  V B0 = O::xor3(a[0], C4, R1);
  K2_EVEN(B1, a[6], C0, R2, 22);   // 44
  K2_ODD(B2, a[12], C1, R3, 21);   // 43
  K2_ODD(B3, a[18], C2, R4, 10);   // 21
  K2_EVEN(B4, a[24], C3, R0, 7);   // 14
  K2_EVEN(B5, a[3], C2, R4, 14);   // 28
  K2_EVEN(B6, a[9], C3, R0, 10);   // 20
  K2_ODD(B7, a[10], C4, R1, 1);    // 3
  K2_ODD(B8, a[16], C0, R2, 22);   // 45
  K2_ODD(B9, a[22], C1, R3, 30);   // 61
  K2_ODD(B10, a[1], C0, R2, 0);    // 1
  K2_EVEN(B11, a[7], C1, R3, 3);   // 6
  K2_ODD(B12, a[13], C2, R4, 12);  // 25
  K2_EVEN(B13, a[19], C3, R0, 4);  // 8
  K2_EVEN(B14, a[20], C4, R1, 9);  // 18
  K2_ODD(B15, a[4], C3, R0, 13);   // 27
  K2_EVEN(B16, a[5], C4, R1, 18);  // 36
  K2_EVEN(B17, a[11], C0, R2, 5);  // 10
  K2_ODD(B18, a[17], C1, R3, 7);   // 15
  K2_EVEN(B19, a[23], C2, R4, 28); // 56
  K2_EVEN(B20, a[2], C1, R3, 31);  // 62
  K2_ODD(B21, a[8], C2, R4, 27);   // 55
  K2_ODD(B22, a[14], C3, R0, 19);  // 39
  K2_ODD(B23, a[15], C4, R1, 20);  // 41
  K2_EVEN(B24, a[21], C0, R2, 1);  // 2
#undef K2_EVEN
#undef K2_ODD
  a[0] = O::xor2(O::chi(B0, B1, B2), rc);
  a[1] = O::chi(B1, B2, B3);
  a[2] = O::chi(B2, B3, B4);
  a[3] = O::chi(B3, B4, B0);
  a[4] = O::chi(B4, B0, B1);
  a[5] = O::chi(B5, B6, B7);
  a[6] = O::chi(B6, B7, B8);
  a[7] = O::chi(B7, B8, B9);
  a[8] = O::chi(B8, B9, B5);
  a[9] = O::chi(B9, B5, B6);
  a[10] = O::chi(B10, B11, B12);
  a[11] = O::chi(B11, B12, B13);
  a[12] = O::chi(B12, B13, B14);
  a[13] = O::chi(B13, B14, B10);
  a[14] = O::chi(B14, B10, B11);
  a[15] = O::chi(B15, B16, B17);
  a[16] = O::chi(B16, B17, B18);
  a[17] = O::chi(B17, B18, B19);
  a[18] = O::chi(B18, B19, B15);
  a[19] = O::chi(B19, B15, B16);
  a[20] = O::chi(B20, B21, B22);
  a[21] = O::chi(B21, B22, B23);
  a[22] = O::chi(B22, B23, B24);
  a[23] = O::chi(B23, B24, B20);
  a[24] = O::chi(B24, B20, B21);
}

// Even / odd bit extraction of a 64-bit word into 32 bits (lane's half), and
// the inverse spread, as plain integer code (used on host and device).
inline __attribute__((always_inline)) DPKEM_HD uint32_t squeeze_bits16(uint32_t x) {
  // x holds 16 bits at even positions 0,2,..,30 -> compact into bits 0..15
  x &= 0x55555555u;
  x = (x ^ (x >> 1)) & 0x33333333u;
  x = (x ^ (x >> 2)) & 0x0F0F0F0Fu;
  x = (x ^ (x >> 4)) & 0x00FF00FFu;
  x = (x ^ (x >> 8)) & 0x0000FFFFu;
  return x;
}
inline __attribute__((always_inline)) DPKEM_HD uint32_t spread_bits16(uint32_t x) {
  // bits 0..15 -> even positions 0,2,..,30
  x &= 0x0000FFFFu;
  x = (x ^ (x << 8)) & 0x00FF00FFu;
  x = (x ^ (x << 4)) & 0x0F0F0F0Fu;
  x = (x ^ (x << 2)) & 0x33333333u;
  x = (x ^ (x << 1)) & 0x55555555u;
  return x;
}
// Half of the 64-bit word w that lane parity p holds.
inline __attribute__((always_inline)) DPKEM_HD uint32_t deint(uint64_t w, uint32_t p) {
  const uint32_t lo = (uint32_t)w >> p, hi = (uint32_t)(w >> 32) >> p;
  return squeeze_bits16(lo) | (squeeze_bits16(hi) << 16);
}
// 32-bit half (low if p == 0, high if p == 1) of the word whose even half is
// e and odd half is o.  Device form: one v_perm interleaves the bytes of the
// two 16-bit pieces ([e0, o0, e1, o1]), then three delta swaps interleave
// nibbles, bit pairs and bits inside each byte (an outer perfect shuffle).
inline __attribute__((always_inline)) DPKEM_HD uint32_t reint_half(uint32_t e, uint32_t o, uint32_t p) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t x = __builtin_amdgcn_perm(o, e, p ? 0x07030602u : 0x05010400u);
  uint32_t t = ((x >> 4) ^ x) & 0x00F000F0u;
  x ^= t ^ (t << 4);
  t = ((x >> 2) ^ x) & 0x0C0C0C0Cu;
  x ^= t ^ (t << 2);
  t = ((x >> 1) ^ x) & 0x22222222u;
  return x ^ t ^ (t << 1);
#else
  const uint32_t sh = 16 * p;
  return spread_bits16(e >> sh) | (spread_bits16(o >> sh) << 1);
#endif
}
inline __attribute__((always_inline)) DPKEM_HD uint64_t reint_full(uint32_t e, uint32_t o) {
  return (uint64_t)reint_half(e, o, 0) | ((uint64_t)reint_half(e, o, 1) << 32);
}

}  // namespace k2
}  // namespace dpkem

#if defined(__HIP__) || defined(__HIPCC__)
namespace dpkem {
namespace k2 {

// Round constant r as {even half, odd half}: one scalar load per round pair.
struct RcPair {
  uint32_t e, o;
};
#define K2_RC(r) {compress_even(kRCs[r]), compress_even(kRCs[r] >> 1)}
__constant__ static const RcPair kRC2[24] = {
    K2_RC(0),  K2_RC(1),  K2_RC(2),  K2_RC(3),  K2_RC(4),  K2_RC(5),  K2_RC(6),  K2_RC(7),
    K2_RC(8),  K2_RC(9),  K2_RC(10), K2_RC(11), K2_RC(12), K2_RC(13), K2_RC(14), K2_RC(15),
    K2_RC(16), K2_RC(17), K2_RC(18), K2_RC(19), K2_RC(20), K2_RC(21), K2_RC(22), K2_RC(23)};
#undef K2_RC

struct DevOps {
  static __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
  }
  static __device__ __forceinline__ uint32_t chi(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xD2);
  }
  static __device__ __forceinline__ uint32_t xor2(uint32_t a, uint32_t b) { return a ^ b; }
  template <int N>
  static __device__ __forceinline__ uint32_t rotc(uint32_t x) {
    return __builtin_amdgcn_alignbit(x, x, 32 - N);
  }
  // rotate left by (k - p): alignbit's shift operand is 32 - k + p (low 5 bits used)
  static __device__ __forceinline__ uint32_t shift(uint32_t p, uint32_t k) { return 32u - k + p; }
  static __device__ __forceinline__ uint32_t rotv(uint32_t x, uint32_t sh) { return __builtin_amdgcn_alignbit(x, x, sh); }
  // value held by the partner lane of the pair (quad_perm [1,0,3,2])
  static __device__ __forceinline__ uint32_t swap(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  }
};

// Full permutation on this lane's half of the state (p = lane parity).
// Both halves of the round constant are read at the uniform index r (scalar
// loads, waited on with lgkmcnt) and the lane's half is selected with a mask:
// a per-lane table index would become a vector load whose vmcnt wait also
// waits for every store the sponge still has in flight.  The table address
// is pinned in SGPRs before the loop; otherwise the compiler rebuilds it
// (s_getpc and 64-bit adds) in every iteration, scalar instructions that
// take issue slots of the lone chain wave.
__device__ __forceinline__ void keccakf2(uint32_t a[25], uint32_t p) {
  typedef const __attribute__((address_space(4))) RcPair ConstRc;
  uint32_t sh[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) sh[k] = DevOps::shift(p, k);  // only the used entries survive
  const uint32_t odd = 0u - p;
  ConstRc* rct = (ConstRc*)kRC2;
  asm volatile("" : "+s"(rct));
#pragma unroll 4
  for (int r = 0; r < 24; ++r) {
    const uint32_t re = rct[r].e, ro = rct[r].o;
    const uint32_t rc = re ^ ((re ^ ro) & odd);
    round2<DevOps>(a, sh, rc);
  }
}

// The same permutation fully unrolled: round constants become immediates (no
// scalar loads or loop branches on the chain), at ~20 KB of code per call
// site, so it is used only in the block loops that carry a whole serial
// chain (the sliced squeezes, the staged absorbs, Frodo's noise squeeze):
// 2-6% off those launches.
__device__ __forceinline__ void keccakf2_unrolled(uint32_t a[25], uint32_t p) {
  uint32_t sh[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) sh[k] = DevOps::shift(p, k);
  const uint32_t odd = 0u - p;
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    const uint32_t re = compress_even(kRCs[r]), ro = compress_even(kRCs[r] >> 1);
    round2<DevOps>(a, sh, re ^ ((re ^ ro) & odd));
  }
}

__device__ __forceinline__ void zero2(uint32_t a[25]) {
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = 0;
}

// This lane's half of a compile-time 64-bit constant.
__device__ __forceinline__ uint32_t half_const(uint64_t c, uint32_t p) {
  return p ? compress_even(c >> 1) : compress_even(c);
}

// 32-bit piece (low if p == 0, high if p == 1) of output word `mine`/partner.
__device__ __forceinline__ uint32_t out_half(uint32_t mine, uint32_t p) {
  const uint32_t other = DevOps::swap(mine);
  return p ? reint_half(other, mine, 1) : reint_half(mine, other, 0);
}
// Full 64-bit output word (both lanes of the pair get the same value).
__device__ __forceinline__ uint64_t out_full(uint32_t mine, uint32_t p) {
  const uint32_t other = DevOps::swap(mine);
  return p ? reint_full(other, mine) : reint_full(mine, other);
}

// Absorb NW message words from global memory with rate RW (64-bit words),
// pad with domain byte ds and run the final permutation.
// The words of block b + 1 (or of the final partial block) are loaded while
// block b is permuted, so a block's loads never wait out a full memory
// latency on the chain's critical path.
template <int RW, int NW>
__device__ __forceinline__ void absorb2_global(uint32_t a[25], uint32_t p, const uint8_t* __restrict__ src,
                                               uint64_t ds) {
  constexpr int NB = NW / RW, REM = NW % RW;
  const uint64_t* w = reinterpret_cast<const uint64_t*>(src);
  zero2(a);
  uint64_t nxt[RW];
#pragma unroll
  for (int i = 0; i < RW; ++i)
    if (NB > 0 || i < REM) nxt[i] = w[i];  // block 0, or the whole message when shorter than a block
  for (int b = 0; b < NB; ++b) {
    uint64_t cur[RW];
#pragma unroll
    for (int i = 0; i < RW; ++i) cur[i] = nxt[i];
    if (b + 1 < NB) {
#pragma unroll
      for (int i = 0; i < RW; ++i) nxt[i] = w[(b + 1) * RW + i];
    } else {
#pragma unroll
      for (int i = 0; i < REM; ++i) nxt[i] = w[NB * RW + i];
    }
#pragma unroll
    for (int i = 0; i < RW; ++i) a[i] ^= deint(cur[i], p);
    keccakf2(a, p);
  }
#pragma unroll
  for (int i = 0; i < REM; ++i) a[i] ^= deint(nxt[i], p);
  a[REM] ^= half_const(ds, p);
  a[RW - 1] ^= half_const(0x8000000000000000ULL, p);
  keccakf2(a, p);
}

// Message staging by a helper wave.  A lane-pair sponge spends ~22 VALU ops
// per absorbed 64-bit word on deint, about one permutation's worth over an
// eight-block message; a wave on another SIMD can do that conversion ahead
// of it.  The helper writes word i of block b, half p of the sponge's
// instance slot s (s = 2 k + p), to msg[(b RW + i) SLOTS + s] and after each
// block raises the LDS count `ready`; the sponge waits for a block's count
// before reading it.  Every wave of a workgroup is resident from its start
// and the helper publishes each block once after bounded work, so the waits
// end; their iteration bound only makes the loops' ends unconditional.
//
// A wait that reaches its bound gives up: the wave goes on with whatever the
// LDS holds (in-bounds reads, garbage results) and adds to wait_expired, a
// counter of this translation unit that the host reads and clears
// (saber_wait_expired, frodo_wait_expired) and reports as a failed launch.  wait_spin_limit is
// the bound (host-settable; 0 makes every wait give up at once, which is how
// the failure path is exercised without a faulty kernel).
static __device__ uint32_t wait_spin_limit = 1u << 22;
static __device__ uint32_t wait_expired = 0;

__device__ __forceinline__ void lds_wait_count(const uint32_t* cnt, uint32_t target) {
  const uint32_t limit = wait_spin_limit;
  for (uint32_t spin = 0; limit != 0; ++spin) {
    if (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) >= target) return;
    if (spin >= limit) break;
    __builtin_amdgcn_s_sleep(1);
  }
  __hip_atomic_fetch_add(&wait_expired, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Helper side: the NW message words at w for slot s with half p; lane 0 of
// the wave publishes.  Block b + 1's words are loaded while block b is
// converted (uniform branches only, so the loads stay in flight).
template <int RW, int NW, int SLOTS>
__device__ __forceinline__ void stage2_message(uint32_t* msg, uint32_t* ready, uint32_t p, int s,
                                               const uint64_t* __restrict__ w, bool lane0) {
  constexpr int NB = NW / RW, REM = NW % RW;
  uint64_t cur[RW];
#pragma unroll
  for (int i = 0; i < RW; ++i)
    if (NB > 0 || i < REM) cur[i] = w[i];
  for (int b = 0; b < NB; ++b) {
    uint64_t nxt[RW];
    if (b + 1 < NB) {
#pragma unroll
      for (int i = 0; i < RW; ++i) nxt[i] = w[(b + 1) * RW + i];
    } else {
#pragma unroll
      for (int i = 0; i < REM; ++i) nxt[i] = w[NB * RW + i];
    }
#pragma unroll
    for (int i = 0; i < RW; ++i) msg[(b * RW + i) * SLOTS + s] = deint(cur[i], p);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane0) __hip_atomic_fetch_add(ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
    for (int i = 0; i < RW; ++i) cur[i] = nxt[i];
  }
  if (REM > 0) {
#pragma unroll
    for (int i = 0; i < REM; ++i) msg[(NB * RW + i) * SLOTS + s] = deint(cur[i], p);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane0) __hip_atomic_fetch_add(ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// Sponge side: absorb2_global over the staged halves of slot s.
template <int RW, int NW, int SLOTS>
__device__ __forceinline__ void absorb2_staged(uint32_t a[25], uint32_t p, const uint32_t* msg, const uint32_t* ready,
                                               int s, uint64_t ds) {
  constexpr int NB = NW / RW, REM = NW % RW;
  zero2(a);
  for (int b = 0; b < NB; ++b) {
    lds_wait_count(ready, (uint32_t)(b + 1));
#pragma unroll
    for (int i = 0; i < RW; ++i) a[i] ^= msg[(b * RW + i) * SLOTS + s];
    keccakf2_unrolled(a, p);
  }
  if (REM > 0) lds_wait_count(ready, (uint32_t)(NB + 1));
#pragma unroll
  for (int i = 0; i < REM; ++i) a[i] ^= msg[(NB * RW + i) * SLOTS + s];
  a[REM] ^= half_const(ds, p);
  a[RW - 1] ^= half_const(0x8000000000000000ULL, p);
  keccakf2(a, p);
}

// Absorb a register-resident message of NW (< RW) words.
template <int RW, int NW>
__device__ __forceinline__ void absorb2_regs(uint32_t a[25], uint32_t p, const uint64_t* m, uint64_t ds) {
  static_assert(NW < RW, "single block");
  zero2(a);
#pragma unroll
  for (int i = 0; i < NW; ++i) a[i] = deint(m[i], p);
  a[NW] ^= half_const(ds, p);
  a[RW - 1] ^= half_const(0x8000000000000000ULL, p);
  keccakf2(a, p);
}

// Absorb NW (< RW) words given as this lane's halves (e.g. the halves of a
// previous digest, which need no conversion).
template <int RW, int NW>
__device__ __forceinline__ void absorb2_halves(uint32_t a[25], uint32_t p, const uint32_t* h, uint64_t ds) {
  static_assert(NW < RW, "single block");
  zero2(a);
#pragma unroll
  for (int i = 0; i < NW; ++i) a[i] = h[i];
  a[NW] ^= half_const(ds, p);
  a[RW - 1] ^= half_const(0x8000000000000000ULL, p);
  keccakf2(a, p);
}

// Squeeze OUTW words (rate RW) from a state that already holds the first
// output block, writing this lane's halves: dst[2 w + p] = half of word w.
template <int RW, int OUTW>
__device__ __forceinline__ void squeeze2_sliced(uint32_t a[25], uint32_t p, uint32_t* __restrict__ dst) {
  constexpr int NB = OUTW / RW, REM = OUTW % RW;
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int i = 0; i < RW; ++i) dst[2 * (b * RW + i) + p] = a[i];
    keccakf2_unrolled(a, p);
  }
#pragma unroll
  for (int i = 0; i < REM; ++i) dst[2 * (NB * RW + i) + p] = a[i];
}

// squeeze2_sliced with a wave-uniform run-time length: kernels whose waves
// squeeze streams of different lengths share this one call site, i.e. one
// copy of the unrolled permutation in the instruction cache instead of one
// per stream.
template <int RW>
__device__ __forceinline__ void squeeze2_sliced_rt(uint32_t a[25], uint32_t p, uint32_t* __restrict__ dst, int outw) {
  const int nb = outw / RW, rem = outw - nb * RW;
  for (int b = 0; b < nb; ++b) {
#pragma unroll
    for (int i = 0; i < RW; ++i) dst[2 * (b * RW + i) + p] = a[i];
    keccakf2_unrolled(a, p);
  }
#pragma unroll
  for (int i = 0; i < RW; ++i)
    if (i < rem) dst[2 * (nb * RW + i) + p] = a[i];
}

// absorb2_staged with a wave-uniform run-time message length (nw words),
// for the same reason.
template <int RW, int SLOTS>
__device__ __forceinline__ void absorb2_staged_rt(uint32_t a[25], uint32_t p, const uint32_t* msg,
                                                  const uint32_t* ready, int s, int nw, uint64_t ds) {
  const int nb = nw / RW, rem = nw - nb * RW;
  zero2(a);
  for (int b = 0; b < nb; ++b) {
    lds_wait_count(ready, (uint32_t)(b + 1));
#pragma unroll
    for (int i = 0; i < RW; ++i) a[i] ^= msg[(b * RW + i) * SLOTS + s];
    keccakf2_unrolled(a, p);
  }
  if (rem > 0) lds_wait_count(ready, (uint32_t)(nb + 1));
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    if (i < rem) a[i] ^= msg[(nb * RW + i) * SLOTS + s];
    if (i == rem) a[i] ^= half_const(ds, p);
  }
  a[RW - 1] ^= half_const(0x8000000000000000ULL, p);
  keccakf2(a, p);
}

// Store full 64-bit output words: the even lane writes the low 32 bits and the
// odd lane the high 32 bits of each word.
template <int NW>
__device__ __forceinline__ void store2_words(const uint32_t* h, uint32_t p, uint8_t* __restrict__ dst) {
#pragma unroll
  for (int i = 0; i < NW; ++i) reinterpret_cast<uint32_t*>(dst)[2 * i + p] = out_half(h[i], p);
}

// Rebuild a word stream written by squeeze2_sliced (plain integer code; used
// by the kernels that consume GenMatrix / GenSecret streams).
__device__ __forceinline__ uint64_t sliced_word(const uint32_t* src, int w) {
  return reint_full(src[2 * w], src[2 * w + 1]);
}

}  // namespace k2
}  // namespace dpkem
#endif
