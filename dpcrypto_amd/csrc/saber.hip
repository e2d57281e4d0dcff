// Saber-family batched kernels for gfx950: LightSaber (l = 2), Saber (l = 3)
// and FireSaber (l = 4); every kernel is a template over the rank l.
//
// Reference path replaced (benlwk/DPCrypto SABER-GPU/):
//   hashing      sha3_256_gpu / sha3_512_gpu / GenSecret_gpu (<<<1,BATCH>>>)
//                and the 25-thread shake128_gpu GenMatrix   src/poly.cu:897-1093
//   CBD          cbd_gpu (mu = 10 / 8 / 6 branches)          src/poly.cu:815-891
//   products     DoDP4Av7 / DoDP4Av7_inner (+ operand repackers)  src/poly.cu:14-206
//   rounding     post_process{,2,3}                          src/poly.cu:360-382
//   packers      POLVECp2BS / BS2POLVECp / BS2POLVECq / POLT2BS / msg  src/pack.cu:8-211
//   FO glue      copysk / verify_gpu / cmov_gpu              src/pack.cu:215-253
//   drivers      saber_enc / saber_dec                       src/cuda_kernel.cu:19-299
//   parameters   LightSaber / Saber / FireSaber              include/params.h:10-19
// Keygen, absent from the reference, follows the published scheme.
//
// Design: each op is three launches.
//   hash1  lane-pair Keccak (two lanes per state, keccak2.hpp), split in two
//          waves per 32 instances so GenMatrix (10 / 23 / 40 permutations
//          for l = 2 / 3 / 4) runs beside the message / secret chain;
//   poly   one workgroup of 32 (l + 1) threads per instance: unpacking, CBD,
//          the negacyclic products as LDS-staged Toeplitz dot products
//          (v_dot2_u32_u16, 8 outputs per thread), rounding and packing;
//   hash2  lane-pair closing hashes (H(pk) or H(ct) and K).
#include "common.hpp"
#include "keccak.hpp"
#include "sponge.hpp"
#include "keccak2.hpp"
#include "prof.hpp"
#include "saber_poly.hpp"
#include "launch.hpp"

#include <algorithm>

namespace dpkem {

// Sizes of the Saber parameter set of rank L (mu, eT as published).
template <int L>
struct SaberP {
  static_assert(L >= 2 && L <= 4, "LightSaber, Saber, FireSaber");
  static constexpr int mu = L == 2 ? 10 : L == 3 ? 8 : 6;
  static constexpr int et = L == 2 ? 3 : L == 3 ? 4 : 6;
  static constexpr uint32_t h2 = 256 - (1u << (10 - et - 1)) + 4;
  static constexpr int vecq = 416 * L;  // POLVECq bytes (13 bits)
  static constexpr int vecp = 320 * L;  // POLVECp bytes (10 bits)
  static constexpr int pk = vecp + 32;
  static constexpr int sk = vecq + pk + 64;
  static constexpr int ct = vecp + 32 * et;
  static constexpr int matrix_bytes = L * vecq;   // GenMatrix stream
  static constexpr int secret_bytes = 32 * L * mu;  // GenSecret stream (256 mu / 8 per polynomial)
  static constexpr int threads = 64;              // product workgroup: one wave (int8 MFMA)
};
static_assert(SaberP<3>::pk == kSaberPk && SaberP<3>::sk == kSaberSk && SaberP<3>::ct == kSaberCt, "Saber sizes");
static_assert(SaberP<3>::matrix_bytes == kSaberGenMatrixBytes && SaberP<3>::secret_bytes == kSaberGenSecretBytes,
              "Saber streams");
static_assert(SaberP<2>::pk == 672 && SaberP<2>::sk == 1568 && SaberP<2>::ct == 736, "LightSaber sizes");
static_assert(SaberP<4>::pk == 1312 && SaberP<4>::sk == 3040 && SaberP<4>::ct == 1472, "FireSaber sizes");

// ----------------------------------------------------------------- hash1
// Lane-pair sponges (keccak2.hpp): lanes 2k / 2k+1 hold the even / odd bit
// halves of instance k's state, 32 instances per wave.  Role 0 is the
// GenMatrix chain, role 1 the message / secret chain.  GenMatrix / GenSecret
// streams leave in the sliced layout (half of word w at u32 index 2w + p).

using k2::absorb2_global;
using k2::absorb2_halves;
using k2::absorb2_regs;
using k2::squeeze2_sliced;
using k2::store2_words;

// The bodies below take the instance group `vb` (G instances, lane pair k
// of the wave serving instance vb * G + k) and the role explicitly so that
// the pipelined chain kernels can run several of them in one launch.
#define SAB_PAIR_SETUP                                   \
  const int lane = threadIdx.x & 63;                     \
  const uint32_t p = lane & 1;                           \
  const int inst = vb * G + (lane >> 1);                 \
  if ((lane >> 1) >= G || inst >= n) return;

template <int L, int G = 32>
__device__ __forceinline__ void keygen_hash1_body(int vb, int role, int n, const uint8_t* __restrict__ coins,
                                                  uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                  uint8_t* __restrict__ araw, uint8_t* __restrict__ sraw) {
  using P = SaberP<L>;
  SAB_PAIR_SETUP
  const uint8_t* c = coins + (size_t)inst * kSaberKeygenCoins;
  uint32_t a[25];
  if (role == 0) {
    uint64_t m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) m[i] = ld64(c + 8 * i);
    absorb2_regs<kRateShake128, 4>(a, p, m, kDsShake);  // seedA = SHAKE128(seedA_raw, 32)
    uint32_t h[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = a[i];
    store2_words<4>(h, p, pk + (size_t)inst * P::pk + P::vecp);
    store2_words<4>(h, p, sk + (size_t)inst * P::sk + P::vecq + P::vecp);
    absorb2_halves<kRateShake128, 4>(a, p, h, kDsShake);  // GenMatrix = SHAKE128(seedA)
    squeeze2_sliced<kRateShake128, P::matrix_bytes / 8>(
        a, p, reinterpret_cast<uint32_t*>(araw + (size_t)inst * P::matrix_bytes));
  } else {
    uint64_t m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) m[i] = ld64(c + 32 + 8 * i);
    absorb2_regs<kRateShake128, 4>(a, p, m, kDsShake);  // GenSecret stream
    squeeze2_sliced<kRateShake128, P::secret_bytes / 8>(
        a, p, reinterpret_cast<uint32_t*>(sraw + (size_t)inst * P::secret_bytes));
    uint32_t* z = reinterpret_cast<uint32_t*>(sk + (size_t)inst * P::sk + P::sk - 32);
    const uint32_t* zc = reinterpret_cast<const uint32_t*>(c + 64);
#pragma unroll
    for (int i = 0; i < 4; ++i) z[2 * i + p] = zc[2 * i + p];
  }
}

// Encaps' m = SHA3-256(m_raw) for instance inst: this lane's halves into
// h[0..3], the bytes into mbuf.
__device__ __forceinline__ void enc_m(int inst, uint32_t p, const uint8_t* __restrict__ coins, uint8_t* __restrict__ mbuf,
                                      uint32_t h[4]) {
  uint32_t a[25];
  uint64_t raw[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) raw[i] = ld64(coins + (size_t)inst * kSaberEncCoins + 8 * i);
  absorb2_regs<kRateSha3_256, 4>(a, p, raw, kDsSha3);
#pragma unroll
  for (int i = 0; i < 4; ++i) h[i] = a[i];
  store2_words<4>(h, p, mbuf + (size_t)inst * 32);
}

// Encaps' kr = G(m || H(pk)) from the halves g[0..7] (m, H(pk)) into krbuf;
// a then holds the GenSecret sponge of kr[32:64], its first block ready.
__device__ __forceinline__ void enc_g_absorb(int inst, uint32_t p, const uint32_t g[8], uint8_t* __restrict__ krbuf,
                                             uint32_t a[25]) {
  absorb2_halves<kRateSha3_512, 8>(a, p, g, kDsSha3);
  uint32_t kr[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) kr[i] = a[i];
  store2_words<8>(kr, p, krbuf + (size_t)inst * 64);
  absorb2_halves<kRateShake128, 4>(a, p, kr + 4, kDsShake);
}

// enc_g_absorb, then the GenSecret stream.
template <int L>
__device__ __forceinline__ void enc_g_chain(int inst, uint32_t p, const uint32_t g[8], uint8_t* __restrict__ sraw,
                                            uint8_t* __restrict__ krbuf) {
  using P = SaberP<L>;
  uint32_t a[25];
  enc_g_absorb(inst, p, g, krbuf, a);
  squeeze2_sliced<kRateShake128, P::secret_bytes / 8>(
      a, p, reinterpret_cast<uint32_t*>(sraw + (size_t)inst * P::secret_bytes));
}

template <int L, int G = 32>
__device__ __forceinline__ void enc_hash1_body(int vb, int role, int n, const uint8_t* __restrict__ pk,
                                               const uint8_t* __restrict__ coins, uint8_t* __restrict__ araw,
                                               uint8_t* __restrict__ sraw, uint8_t* __restrict__ mbuf,
                                               uint8_t* __restrict__ krbuf) {
  using P = SaberP<L>;
  SAB_PAIR_SETUP
  const uint8_t* pkp = pk + (size_t)inst * P::pk;
  uint32_t a[25];
  if (role == 0) {
    uint64_t seed[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) seed[i] = ld64(pkp + P::vecp + 8 * i);
    absorb2_regs<kRateShake128, 4>(a, p, seed, kDsShake);
    squeeze2_sliced<kRateShake128, P::matrix_bytes / 8>(
        a, p, reinterpret_cast<uint32_t*>(araw + (size_t)inst * P::matrix_bytes));
  } else {
    uint32_t g[8];
    enc_m(inst, p, coins, mbuf, g);
    absorb2_global<kRateSha3_256, P::pk / 8>(a, p, pkp, kDsSha3);  // H(pk)
#pragma unroll
    for (int i = 0; i < 4; ++i) g[4 + i] = a[i];
    enc_g_chain<L>(inst, p, g, sraw, krbuf);
  }
}

// Decaps' message chain for instance inst: kr = G(m' || H(pk)) (H(pk) from
// sk), then the GenSecret stream from kr[32:64].  m points at the instance's
// 32 bytes of m' (global memory or LDS).
// hpk: the instance's H(pk) words from sk (sk[sk - 64 .. sk - 32)), loaded by
// the caller.
template <int L>
__device__ __forceinline__ void load_hpk(int inst, const uint8_t* __restrict__ sk, uint64_t hpk[4]) {
  using P = SaberP<L>;
#pragma unroll
  for (int i = 0; i < 4; ++i) hpk[i] = ld64(sk + (size_t)inst * P::sk + P::sk - 64 + 8 * i);
}

// dec_msg_absorb leaves a holding the GenSecret sponge of kr[32:64] (first
// block ready); dec_msg_chain also squeezes the stream.  kq (optional):
// K-hat' halves also go to kq[i * ks + slot] in LDS, for a wave of the same
// workgroup that hashes K; the lane of slot 0 then raises the LDS count
// kq_ready.
__device__ __forceinline__ void dec_msg_absorb(int inst, uint32_t p, const uint64_t hpk[4], const uint8_t* m,
                                               uint8_t* __restrict__ krbuf, uint32_t a[25], uint32_t* kq = nullptr,
                                               int ks = 0, int slot = 0, uint32_t* kq_ready = nullptr) {
  uint64_t buf[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    buf[i] = ld64(m + 8 * i);  // m'
    buf[4 + i] = hpk[i];       // H(pk) from sk
  }
  absorb2_regs<kRateSha3_512, 8>(a, p, buf, kDsSha3);
  uint32_t kr[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) kr[i] = a[i];
  if (kq) {
#pragma unroll
    for (int i = 0; i < 4; ++i) kq[i * ks + slot] = kr[i];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (slot == 0) __hip_atomic_fetch_add(kq_ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  store2_words<8>(kr, p, krbuf + (size_t)inst * 64);
  absorb2_halves<kRateShake128, 4>(a, p, kr + 4, kDsShake);
}

template <int L>
__device__ __forceinline__ void dec_msg_chain(int inst, uint32_t p, const uint64_t hpk[4], const uint8_t* m,
                                              uint8_t* __restrict__ sraw, uint8_t* __restrict__ krbuf,
                                              uint32_t* kq = nullptr, int ks = 0, int slot = 0,
                                              uint32_t* kq_ready = nullptr) {
  using P = SaberP<L>;
  uint32_t a[25];
  dec_msg_absorb(inst, p, hpk, m, krbuf, a, kq, ks, slot, kq_ready);
  squeeze2_sliced<kRateShake128, P::secret_bytes / 8>(
      a, p, reinterpret_cast<uint32_t*>(sraw + (size_t)inst * P::secret_bytes));
}

template <int L, int G = 32>
__device__ __forceinline__ void dec_hash1_body(int vb, int role, int n, const uint8_t* __restrict__ sk,
                                               const uint8_t* __restrict__ mbuf, uint8_t* __restrict__ araw,
                                               uint8_t* __restrict__ sraw, uint8_t* __restrict__ krbuf) {
  using P = SaberP<L>;
  SAB_PAIR_SETUP
  const uint8_t* s = sk + (size_t)inst * P::sk;
  uint32_t a[25];
  if (role == 0) {
    uint64_t seed[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) seed[i] = ld64(s + P::vecq + P::vecp + 8 * i);
    absorb2_regs<kRateShake128, 4>(a, p, seed, kDsShake);
    squeeze2_sliced<kRateShake128, P::matrix_bytes / 8>(
        a, p, reinterpret_cast<uint32_t*>(araw + (size_t)inst * P::matrix_bytes));
  } else {
    uint64_t hpk[4];
    load_hpk<L>(inst, sk, hpk);
    dec_msg_chain<L>(inst, p, hpk, mbuf + (size_t)inst * 32, sraw, krbuf);
  }
}

// ----------------------------------------------------------------- hash2

// H(pk) into sk.
template <int L, int G = 32>
__device__ __forceinline__ void keygen_hash2_body(int vb, int n, const uint8_t* __restrict__ pk,
                                                  uint8_t* __restrict__ sk) {
  using P = SaberP<L>;
  SAB_PAIR_SETUP
  uint32_t a[25];
  absorb2_global<kRateSha3_256, P::pk / 8>(a, p, pk + (size_t)inst * P::pk, kDsSha3);
  store2_words<4>(a, p, sk + (size_t)inst * P::sk + P::sk - 64);
}

// K = SHA3-256(kr[0:32] || H(ct)); for decaps kr[0:32] is replaced by z when
// the re-encryption differs (constant-time select on the mismatch flag).
// ABS absorbs H(ct) of instance inst into a (absorb2_global over ct, or the
// LDS-staged halves in the pipelined chain).
template <int L, int G, class ABS>
__device__ __forceinline__ void hash2_core(int vb, int n, ABS&& absorb_ct, const uint8_t* __restrict__ krbuf,
                                           const uint8_t* __restrict__ sk, const uint32_t* __restrict__ fail,
                                           uint8_t* __restrict__ ss) {
  using P = SaberP<L>;
  SAB_PAIR_SETUP
  uint32_t a[25], m[8];
  uint64_t k[4], z[4];  // loaded before H(ct) so that their latency is off the chain
  const uint64_t mask = fail ? 0 - (uint64_t)(fail[inst] != 0) : 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    k[i] = ld64(krbuf + (size_t)inst * 64 + 8 * i);
    z[i] = sk ? ld64(sk + (size_t)inst * P::sk + P::sk - 32 + 8 * i) : 0;
  }
  absorb_ct(a, p, inst, lane);
#pragma unroll
  for (int i = 0; i < 4; ++i) m[4 + i] = a[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) m[i] = k2::deint((k[i] & ~mask) | (z[i] & mask), p);
  absorb2_halves<kRateSha3_256, 8>(a, p, m, kDsSha3);
  store2_words<4>(a, p, ss + (size_t)inst * kSaberSs);
}

template <int L, int G = 32>
__device__ __forceinline__ void hash2_body(int vb, int n, const uint8_t* __restrict__ ct,
                                           const uint8_t* __restrict__ krbuf, const uint8_t* __restrict__ sk,
                                           const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss) {
  using P = SaberP<L>;
  hash2_core<L, G>(
      vb, n,
      [&](uint32_t* a, uint32_t p, int inst, int) {
        absorb2_global<kRateSha3_256, P::ct / 8>(a, p, ct + (size_t)inst * P::ct, kDsSha3);
      },
      krbuf, sk, fail, ss);
}

// ----------------------------------------------------------------- hash kernels
// hash1: blockDim = 128, wave 0 runs role 0 and wave 1 role 1 for the same 32
// instances (role_sel >= 0: that role only, 64 threads).  hash2: 64 threads
// per 32 instances.

template <int L>
__global__ void __launch_bounds__(128) sab_keygen_hash1(int role_sel, int n, const uint8_t* __restrict__ coins,
                                                        uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                        uint8_t* __restrict__ araw, uint8_t* __restrict__ sraw) {
  keygen_hash1_body<L>(blockIdx.x, role_sel >= 0 ? role_sel : (int)(threadIdx.x >> 6), n, coins, pk, sk, araw, sraw);
}

template <int L>
__global__ void __launch_bounds__(128) sab_enc_hash1(int role_sel, int n, const uint8_t* __restrict__ pk,
                                                     const uint8_t* __restrict__ coins, uint8_t* __restrict__ araw,
                                                     uint8_t* __restrict__ sraw, uint8_t* __restrict__ mbuf,
                                                     uint8_t* __restrict__ krbuf) {
  enc_hash1_body<L>(blockIdx.x, role_sel >= 0 ? role_sel : (int)(threadIdx.x >> 6), n, pk, coins, araw, sraw, mbuf,
                    krbuf);
}

template <int L>
__global__ void __launch_bounds__(128) sab_dec_hash1(int role_sel, int n, const uint8_t* __restrict__ sk,
                                                     const uint8_t* __restrict__ mbuf, uint8_t* __restrict__ araw,
                                                     uint8_t* __restrict__ sraw, uint8_t* __restrict__ krbuf) {
  dec_hash1_body<L>(blockIdx.x, role_sel >= 0 ? role_sel : (int)(threadIdx.x >> 6), n, sk, mbuf, araw, sraw, krbuf);
}

template <int L>
__global__ void __launch_bounds__(64) sab_keygen_hash2(int n, const uint8_t* __restrict__ pk, uint8_t* __restrict__ sk) {
  keygen_hash2_body<L>(blockIdx.x, n, pk, sk);
}

template <int L>
__global__ void __launch_bounds__(64) sab_hash2(int n, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ krbuf,
                                                const uint8_t* __restrict__ sk, const uint32_t* __restrict__ fail,
                                                uint8_t* __restrict__ ss) {
  hash2_body<L>(blockIdx.x, n, ct, krbuf, sk, fail, ss);
}

// Pipelined keygen -> encaps -> decaps chain (saber_pipe_launch).  Stages of
// different ops that read disjoint inputs share a launch:
//   x0  four waves per 32 instances: wave 0 computes keygen's seedA and
//       stores it into pk and sk; after one workgroup barrier wave 0 runs
//       keygen's GenMatrix, wave 1 keygen's GenSecret, wave 2 encaps'
//       GenMatrix from the seedA bytes of pk and wave 3 decaps' GenMatrix
//       from those of sk (the same bytes, handed over through LDS);
//   x1  wave 0: encaps' H(pk), then G and GenSecret; wave 1: keygen's H(pk) into sk; wave 2: encaps' m, handed
//       to wave 0 through LDS, then encaps' and decaps' GenMatrix streams
//       rewritten as plain words (plain_matrix_streams); wave 3
//       deinterleaves pk into LDS ahead of wave 0 (stage2_message);
//   x2  (after the encaps products, below sab_dec_inner) decaps' inner
//       product and message chain beside encaps' K and decaps' H(ct), and
//       decaps' K for both outcomes of the re-encryption's compare, so the
//       re-encryption launch only selects one (sab_poly<kSabReenc>).
// Every op still computes all of its own hashes; only their order changes.
template <int L>
__global__ void __launch_bounds__(256) sab_pipe_x0(int n, const uint8_t* __restrict__ kc, uint8_t* __restrict__ pk,
                                                   uint8_t* __restrict__ sk, uint8_t* __restrict__ araw_k,
                                                   uint8_t* __restrict__ sraw_k, uint8_t* __restrict__ araw_e,
                                                   uint8_t* __restrict__ araw_d) {
  using P = SaberP<L>;
  __shared__ uint32_t seed_h[4][64];  // wave 0's halves of seedA, by lane
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, vb = blockIdx.x;
  const uint32_t p = lane & 1;
  const int inst = vb * 32 + (lane >> 1);
  const bool live = inst < n;  // every block holds at least one live instance
  uint32_t a[25], h[4];
  if (wave == 0 && live) {
    const uint8_t* c = kc + (size_t)inst * kSaberKeygenCoins;
    uint64_t m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) m[i] = ld64(c + 8 * i);
    absorb2_regs<kRateShake128, 4>(a, p, m, kDsShake);  // seedA = SHAKE128(seedA_raw, 32)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      h[i] = a[i];
      seed_h[i][lane] = a[i];
    }
    store2_words<4>(h, p, pk + (size_t)inst * P::pk + P::vecp);
    store2_words<4>(h, p, sk + (size_t)inst * P::sk + P::vecq + P::vecp);
  }
  __syncthreads();
  if (!live) return;
  if (wave == 1) {
    keygen_hash1_body<L>(vb, 1, n, kc, pk, sk, nullptr, sraw_k);
    return;
  }
  // GenMatrix = SHAKE128(seedA): keygen's (wave 0), encaps' from the seedA
  // bytes of pk (wave 2) and decaps' from those of sk (wave 3); the bytes
  // wave 0 just stored are handed over through LDS, so the two waves need no
  // global round trip before their first permutation
  if (wave != 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = seed_h[i][lane];
  }
  absorb2_halves<kRateShake128, 4>(a, p, h, kDsShake);
  uint8_t* dst = wave == 0 ? araw_k : wave == 2 ? araw_e : araw_d;
  squeeze2_sliced<kRateShake128, P::matrix_bytes / 8>(
      a, p, reinterpret_cast<uint32_t*>(dst + (size_t)inst * P::matrix_bytes));
}

// Encaps' and decaps' GenMatrix streams of this workgroup's 32 instances,
// written in the sliced layout by sab_pipe_x0, rewritten in place as plain
// 64-bit words (their products run after this launch and then skip the
// re-interleave).  One wave, one instance at a time, 8-byte accesses.
template <int L>
__device__ __forceinline__ void plain_matrix_streams(int n, int lane, uint8_t* __restrict__ araw_e,
                                                     uint8_t* __restrict__ araw_d) {
  constexpr int MW = SaberP<L>::matrix_bytes / 8, J = (MW + 63) / 64;
  const int first = (int)blockIdx.x * 32, cnt = min(32, n - first);
#pragma unroll 1
  for (int q = 0; q < 2 * cnt; ++q) {
    uint2* w = reinterpret_cast<uint2*>((q < cnt ? araw_e : araw_d) + (size_t)(first + q % cnt) * SaberP<L>::matrix_bytes);
    uint2 v[J];
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (lane + 64 * j < MW) v[j] = w[lane + 64 * j];
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (lane + 64 * j < MW) w[lane + 64 * j] = make_uint2(k2::reint_half(v[j].x, v[j].y, 0), k2::reint_half(v[j].x, v[j].y, 1));
  }
}

template <int L>
struct X1Lds {
  static constexpr int kPkBlocks = (SaberP<L>::pk / 8 + kRateSha3_256 - 1) / kRateSha3_256;
  uint32_t msg[kPkBlocks * kRateSha3_256 * 64];  // pk halves for encaps' H(pk)
  uint32_t mh[4][64];                            // wave 2's halves of m, by lane
  uint32_t ready, m_ready;
};

template <int L>
__global__ void __launch_bounds__(256) sab_pipe_x1(int n, const uint8_t* __restrict__ pk,
                                                   const uint8_t* __restrict__ ecoins, uint8_t* __restrict__ sk,
                                                   uint8_t* __restrict__ sraw_e, uint8_t* __restrict__ mbuf_e,
                                                   uint8_t* __restrict__ krbuf_e, uint8_t* __restrict__ araw_e,
                                                   uint8_t* __restrict__ araw_d) {
  using P = SaberP<L>;
  constexpr int PW = P::pk / 8;
  __shared__ X1Lds<L> S;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t p = lane & 1;
  const int inst = blockIdx.x * 32 + (lane >> 1);
  const bool live = inst < n;
  if (threadIdx.x == 0) {
    S.ready = 0;
    S.m_ready = 0;
  }
  __syncthreads();  // the only workgroup barrier
  uint32_t g[8];
  if (wave == 3) {
    // helper: deinterleaves pk into LDS ahead of encaps' H(pk) (wave 0)
    const uint64_t* w = reinterpret_cast<const uint64_t*>(pk + (size_t)min(inst, n - 1) * P::pk);
    k2::stage2_message<kRateSha3_256, PW, 64>(S.msg, &S.ready, p, lane, w, lane == 0);
  } else if (live) {
    if (wave == 0) {
      uint32_t a[25];
      k2::absorb2_staged<kRateSha3_256, PW, 64>(a, p, S.msg, &S.ready, lane, kDsSha3);  // encaps' H(pk)
#pragma unroll
      for (int i = 0; i < 4; ++i) g[4 + i] = a[i];
    } else if (wave == 1) {
      keygen_hash2_body<L>(blockIdx.x, n, pk, sk);
    } else {
      uint32_t h[4];
      enc_m(inst, p, ecoins, mbuf_e, h);
#pragma unroll
      for (int i = 0; i < 4; ++i) S.mh[i][lane] = h[i];
    }
  }
  if (wave == 2) {  // m handed to wave 0 through LDS (an LDS count, so that wave 0 never waits for wave 1)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_fetch_add(&S.m_ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    plain_matrix_streams<L>(n, lane, araw_e, araw_d);
  }
  if (!live || wave != 0) return;
  k2::lds_wait_count(&S.m_ready, 1u);
#pragma unroll
  for (int i = 0; i < 4; ++i) g[i] = S.mh[i][lane];
  enc_g_chain<L>(inst, p, g, sraw_e, krbuf_e);
}

// ----------------------------------------------------------------- products

enum SabMode { kSabKeygen = 0, kSabEnc = 1, kSabReenc = 2 };

struct SabPolyArgs {
  int n;
  const uint8_t* araw;   // GenMatrix stream per instance
  const uint8_t* sraw;   // GenSecret stream per instance
  const uint8_t* pk;     // keygen: out (b part); enc: in; reenc: pk inside sk
  size_t pk_stride;      // bytes between instances' pk
  const uint8_t* mbuf;   // enc / reenc: message, 32 B / instance
  uint8_t* out;          // keygen: pk; enc: ct; reenc: unused
  uint8_t* sk;           // keygen: sk out
  const uint8_t* ct_in;  // reenc: received ct to compare with
  uint32_t* fail;        // reenc: per-instance mismatch flag (or nullptr)
  // reenc in the pipelined chain: K for both outcomes of the compare (64 B
  // per instance, K then the implicit-rejection K), selected into ss
  const uint8_t* kk;
  uint8_t* ss;
  int force_reject;      // test facility: treat every re-encryption as differing
};

constexpr int kARow = 264;  // u16 row stride of plain polynomials in the dot2 products (spreads LDS banks)

// Products on int8 MFMA (saber_poly.hpp, toep8_*): one 64-thread wave per
// instance.  The secret (s or s', small) is the staged Toeplitz factor; the
// B operand holds, per K-block j, the signed limbs of the matrix polynomials
// (A[i][j], or A[j][i] for keygen's A^T s) in columns i / 8 + i and of b[j]
// (inner product) in columns L / 8 + L.  Per instance: L x 64 matrix
// instructions, 28 L Toeplitz operand reads.
//
// The K-blocks stream through LDS one at a time (the Toeplitz copies of s_j,
// the limbs of block j and its raw bytes, ~6-7 KB): block j + 1's global
// words are loaded into registers while block j is unpacked and multiplied,
// and the small footprint lets all of a 4096-instance batch be resident at
// once instead of in two rounds.
// Slot strides of the B-operand limbs and the limb sums are padded off a
// multiple of the LDS bank count: the sixteen lanes of one access read or
// write sixteen different slots at the same k.
constexpr int kLbSlot = 272;    // bytes per B-operand slot (256 used)
constexpr int kPartSlot = 264;  // u16 per limb-sum slot (256 used)
constexpr int kPolyWords = 52;  // 64-bit words of one 13-bit packed polynomial (416 B)

template <int L>
struct SabPolyLds {
  union {
    struct {
      alignas(16) int8_t V[kVPoly];                    // Toeplitz copies of s_j
      alignas(16) int8_t Lb[2 * (L + 1) * kLbSlot];    // B-operand limbs of K-block j: [slot][k]
      alignas(16) uint64_t mstage[L * kPolyWords];     // matrix polynomials of K-block j (bytes)
      alignas(16) uint64_t sstage[SaberP<L>::mu * 4];  // s_j's CBD bytes (32 mu)
      alignas(16) uint32_t bstage[80 + 4];             // b_j (320 B, 10-bit), + slack for the last group's read
    } k;
    struct {
      alignas(16) uint16_t part[2 * (L + 1) * kPartSlot];  // limb sums per slot (mod 2^16)
      alignas(8) uint8_t outb[SaberP<L>::ct + 64];
    } out;
  } u;
};

// Ordering of LDS accesses between the lanes of one wave: a wave's LDS
// operations complete in issue order, so wavefront-scope fences (no wait
// instructions) and a code-motion barrier suffice; in particular no wait for
// the global loads still in flight.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Up to 5 bytes at off (two aligned dword reads).
__device__ __forceinline__ uint64_t lds_bytes64(const uint8_t* base, int off) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(base + (off & ~3));
  return ((((uint64_t)w[1] << 32) | w[0]) >> (8 * (off & 3)));
}

// One instance's product on one wave (tid = its lane) with its own LDS S.
// A_PLAIN: the matrix stream holds plain 64-bit words (sab_pipe_x1 rewrote
// it), not the sliced halves of the squeeze.
template <int MODE, int L, bool A_PLAIN = false>
__device__ __forceinline__ void sab_poly_body(const SabPolyArgs& a, SabPolyLds<L>& S, const int tid, const int inst) {
  using P = SaberP<L>;
  constexpr int NT = 64;
  constexpr bool HAS_B = MODE != kSabKeygen;
  constexpr int MW = L * kPolyWords, MPT = (MW + NT - 1) / NT;  // sliced matrix words per K-block
  constexpr int SW = P::mu * 4;                                // sliced secret words per polynomial
  constexpr int CW = MODE == kSabReenc ? P::ct / 4 : 0, CPT = (CW + NT - 1) / NT;

  const uint2* ar = reinterpret_cast<const uint2*>(a.araw + (size_t)inst * P::matrix_bytes);
  const uint2* sr = reinterpret_cast<const uint2*>(a.sraw + (size_t)inst * P::secret_bytes);
  const uint32_t* bp = HAS_B ? reinterpret_cast<const uint32_t*>(a.pk + (size_t)inst * a.pk_stride) : nullptr;
  // Word t of K-block j's matrix polynomials: polynomial t / 52 of the block
  // is output column t / 52, i.e. A[col][j] (encaps) or A[j][col] (keygen).
  uint2 mv[MPT], sv = {0, 0};
  uint32_t bv[2] = {0, 0};
  auto fetch = [&](int j) {
#pragma unroll
    for (int r = 0; r < MPT; ++r) {
      const int t = tid + r * NT;
      if (t < MW) {
        const int col = t / kPolyWords, w = t - kPolyWords * col;
        const int poly = MODE == kSabKeygen ? j * L + col : col * L + j;
        mv[r] = ar[kPolyWords * poly + w];
      }
    }
    if (tid < SW) sv = sr[SW * j + tid];
    if (HAS_B) {
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (tid + r * NT < 80) bv[r] = bp[80 * j + tid + r * NT];
    }
  };
  uint32_t cv[CPT > 0 ? CPT : 1];  // re-encryption: the received ciphertext, compared at the end
  if (MODE == kSabReenc) {
    const uint32_t* ct = reinterpret_cast<const uint32_t*>(a.ct_in + (size_t)inst * P::ct);
#pragma unroll
    for (int r = 0; r < CPT; ++r)
      if (tid + r * NT < CW) cv[r] = ct[tid + r * NT];
  }
  fetch(0);

  const int slot = toep8_slot(tid & 15, L, HAS_B);
  v4i8x4 acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = v4i8x4{0, 0, 0, 0};

#pragma unroll 1
  for (int j = 0; j < L; ++j) {
    // 1. this block's bytes into LDS (sliced words rebuilt), then block j + 1's loads
#pragma unroll
    for (int r = 0; r < MPT; ++r)
      if (tid + r * NT < MW)
        S.u.k.mstage[tid + r * NT] = A_PLAIN ? (uint64_t)mv[r].x | ((uint64_t)mv[r].y << 32)
                                             : k2::reint_full(mv[r].x, mv[r].y);
    if (tid < SW) S.u.k.sstage[tid] = k2::reint_full(sv.x, sv.y);
    if (HAS_B) {
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (tid + r * NT < 80) S.u.k.bstage[tid + r * NT] = bv[r];
    }
    if (j + 1 < L) fetch(j + 1);
    wave_sync();

    // 2. CBD of s_j (mu bits per coefficient: HW(low mu/2) - HW(high mu/2))
    //    into the Toeplitz copies; the matrix polynomials and b_j (groups of 4
    //    13- or 10-bit values, one group per lane per polynomial) into the limb slots
    {
      const int g = tid;  // coefficients 4g .. 4g + 3 from mu / 2 bytes
      const uint64_t v = lds_bytes64(reinterpret_cast<const uint8_t*>(S.u.k.sstage), g * (P::mu / 2));
      uint32_t pv = 0, nv = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t x = (uint32_t)(v >> (P::mu * k)) & ((1u << P::mu) - 1);
        const int half = P::mu / 2;
        const int c = (int)popc(x & ((1u << half) - 1)) - (int)popc(x >> half);
        pv |= ((uint32_t)c & 0xFF) << (8 * (3 - k));
        nv |= ((uint32_t)(-c) & 0xFF) << (8 * (3 - k));
      }
      toep8_put_wave(S.u.k.V, g, pv, nv);
    }
    {
      // coefficients 4q .. 4q + 3 (q = lane) of each matrix polynomial: the
      // 52 bits at bit 52 q of its 416 bytes, i.e. byte 6.5 q, shift 0 or 4
      // (both fixed per lane), read as three dwords and aligned
      const uint8_t* mb = reinterpret_cast<const uint8_t*>(S.u.k.mstage);
      const int q = tid, off = (52 * q) >> 3, sh = off & 3;
      const uint32_t bs = (52 * q) & 7;
      const uint32_t* d = reinterpret_cast<const uint32_t*>(mb + (off & ~3));
#pragma unroll
      for (int col = 0; col < L; ++col) {
        const uint32_t d0 = d[104 * col], d1 = d[104 * col + 1], d2 = d[104 * col + 2];
        const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, sh), w1 = __builtin_amdgcn_alignbyte(d2, d1, sh);
        const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, bs), hi = w1 >> bs;
        const uint32_t x0 = lo & 0x1FFF, x1 = (lo >> 13) & 0x1FFF;
        const uint32_t x2 = __builtin_amdgcn_alignbit(hi, lo, 26) & 0x1FFF, x3 = (hi >> 7) & 0x1FFF;
        uint32_t l4, h4;
        limbs8x4(x0 | (x1 << 16), x2 | (x3 << 16), &l4, &h4);
        *reinterpret_cast<uint32_t*>(S.u.k.Lb + col * kLbSlot + 4 * q) = l4;
        *reinterpret_cast<uint32_t*>(S.u.k.Lb + (L + 1 + col) * kLbSlot + 4 * q) = h4;
      }
    }
    if (HAS_B) {
      const int g = tid;  // b_j coefficients 4g .. 4g + 3 (10-bit fields of 5 bytes)
      const uint64_t x = lds_bytes64(reinterpret_cast<const uint8_t*>(S.u.k.bstage), 5 * g);
      const uint32_t w0 = (uint32_t)x, w1 = (uint32_t)(x >> 32);
      uint32_t lo, hi;
      limbs8x4((w0 & 0x3FF) | (((w0 >> 10) & 0x3FF) << 16),
               ((w0 >> 20) & 0x3FF) | ((__builtin_amdgcn_alignbit(w1, w0, 30) & 0x3FF) << 16), &lo, &hi);
      int8_t* blk = S.u.k.Lb + 4 * g;
      *reinterpret_cast<uint32_t*>(blk + L * kLbSlot) = lo;
      *reinterpret_cast<uint32_t*>(blk + (2 * L + 1) * kLbSlot) = hi;
    }
    wave_sync();

    // keygen: sk[416 j .. 416 j + 416) = s_j in 13 bits (two's complement),
    // read back from Toeplitz copy 0 (s[q] at byte 256 - q)
    if (MODE == kSabKeygen && tid < 32) {
      uint64_t lo = 0, hi = 0;
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const uint64_t v = (uint64_t)(int64_t)S.u.k.V[256 - (8 * tid + d)] & 0x1FFF;
        const int bit = 13 * d;
        if (bit < 64) lo |= v << bit;
        if (bit + 13 > 64) hi |= bit >= 64 ? v << (bit - 64) : v >> (64 - bit);
      }
      uint8_t* o = a.sk + (size_t)inst * P::sk + 416 * j + 13 * tid;
#pragma unroll
      for (int i = 0; i < 8; ++i) o[i] = (uint8_t)(lo >> (8 * i));
#pragma unroll
      for (int i = 0; i < 5; ++i) o[8 + i] = (uint8_t)(hi >> (8 * i));
    }

    // 3. products: result tile t, lane (g, n) = rows 16 t + 4 g + 0..3 of column n
    v4i8x4 bop[4];
    const int8_t* colp = S.u.k.Lb + (slot < 0 ? 0 : slot) * kLbSlot + 16 * (tid >> 4);
#pragma unroll
    for (int bb = 0; bb < 4; ++bb)
      bop[bb] = slot >= 0 ? *reinterpret_cast<const v4i8x4*>(colp + 64 * bb) : v4i8x4{0, 0, 0, 0};
    toep8_block(acc, S.u.k.V, bop, tid, [](v4i8x4 x, v4i8x4 y, v4i8x4 z) {
      return __builtin_amdgcn_mfma_i32_16x16x64_i8(x, y, z, 0, 0, 0);
    });
    wave_sync();  // the next block (or the limb sums) overwrite V / Lb / the stages
  }

  if (slot >= 0) {
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      uint2 w;
      w.x = ((uint32_t)acc[t][0] & 0xFFFF) | ((uint32_t)acc[t][1] << 16);
      w.y = ((uint32_t)acc[t][2] & 0xFFFF) | ((uint32_t)acc[t][3] << 16);
      *reinterpret_cast<uint2*>(S.u.out.part + slot * kPartSlot + 16 * t + 4 * (tid >> 4)) = w;
    }
  }
  wave_sync();

  // 4. rounding + packing into the LDS output image: rows 0..L-1 are b' (or
  //    keygen's b), row L the inner product v'
  uint8_t* outb = S.u.out.outb;
  {
    // rows 0..L-1: b' (keygen's b) = (x + h1) >> 3 in 10 bits, where
    // x = lo + 256 hi (the limb sums): only bits 3..12 of x + 4 are kept, so
    // the sums are combined and rounded mod 2^16 two coefficients per
    // instruction (16-bit lanes); lane q packs coefficients 4q .. 4q + 3 of
    // each row into the 5 bytes at 5q
    const int q = tid;
    const u16x2 k4 = {4, 4}, k3 = {3, 3}, k8 = {8, 8};
#pragma unroll
    for (int row = 0; row < L; ++row) {
      const uint2 lo = *reinterpret_cast<const uint2*>(S.u.out.part + row * kPartSlot + 4 * q);
      const uint2 hi = *reinterpret_cast<const uint2*>(S.u.out.part + (L + 1 + row) * kPartSlot + 4 * q);
      const u16x2 x01 = __builtin_bit_cast(u16x2, lo.x) + (__builtin_bit_cast(u16x2, hi.x) << k8);
      const u16x2 x23 = __builtin_bit_cast(u16x2, lo.y) + (__builtin_bit_cast(u16x2, hi.y) << k8);
      const uint32_t v01 = __builtin_bit_cast(uint32_t, (x01 + k4) >> k3) & 0x03FF03FFu;
      const uint32_t v23 = __builtin_bit_cast(uint32_t, (x23 + k4) >> k3) & 0x03FF03FFu;
      const uint32_t w = (v01 & 0x3FF) | ((v01 >> 16) << 10) | ((v23 & 0x3FF) << 20) | (v23 >> 16 << 30);
      uint8_t* o = outb + row * 320 + 5 * q;
      o[0] = (uint8_t)w;
      o[1] = (uint8_t)(w >> 8);
      o[2] = (uint8_t)(w >> 16);
      o[3] = (uint8_t)(w >> 24);
      o[4] = (uint8_t)(v23 >> 18);
    }
  }
  if (HAS_B && tid < 32) {
    // row L: c_m = (v' - m 2^9 + h1) >> (10 - eT) -> eT bits, 8 coefficients
    // per lane; byte grp of m holds the bits of coefficients 8 grp .. 8 grp + 7
    const int grp = tid;
    const uint16_t* plo = S.u.out.part + L * kPartSlot + 8 * grp;
    const uint16_t* phi = S.u.out.part + (2 * L + 1) * kPartSlot + 8 * grp;
    const uint32_t mbyte = a.mbuf[(size_t)inst * 32 + grp];
    uint64_t packed = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const uint32_t v = (uint32_t)plo[d] + 256u * (uint32_t)phi[d];
      const uint32_t mb = (mbyte >> d) & 1;
      const uint64_t c = ((v - (mb << 9) + 4u) >> (10 - P::et)) & ((1u << P::et) - 1);
      packed |= c << (P::et * d);
    }
    uint8_t* o = outb + P::vecp + P::et * grp;
#pragma unroll
    for (int i = 0; i < P::et; ++i) o[i] = (uint8_t)(packed >> (8 * i));
  }
  wave_sync();

  // 5. emit
  const uint32_t* ob = reinterpret_cast<const uint32_t*>(outb);
  if (MODE == kSabKeygen) {
    uint8_t* pk = a.out + (size_t)inst * P::pk;
    uint8_t* sk = a.sk + (size_t)inst * P::sk;
    for (int i = tid; i < P::vecp / 4; i += NT) {
      reinterpret_cast<uint32_t*>(pk)[i] = ob[i];
      reinterpret_cast<uint32_t*>(sk + P::vecq)[i] = ob[i];
    }
  } else if (MODE == kSabEnc) {
    uint32_t* ct = reinterpret_cast<uint32_t*>(a.out + (size_t)inst * P::ct);
    for (int i = tid; i < P::ct / 4; i += NT) ct[i] = ob[i];
  } else {
    uint32_t d = 0;
#pragma unroll
    for (int r = 0; r < CPT; ++r)
      if (tid + r * NT < CW) d |= cv[r] ^ ob[tid + r * NT];
    const bool differ = __builtin_amdgcn_ballot_w64(d != 0) != 0 || a.force_reject != 0;
    if (a.fail && tid == 0) a.fail[inst] = differ ? 1u : 0u;
    if (a.kk && tid < kSaberSs / 4) {  // constant-time select of the precomputed K
      const uint32_t mask = differ ? 0xFFFFFFFFu : 0u;
      const uint32_t* kk = reinterpret_cast<const uint32_t*>(a.kk + (size_t)inst * 64);
      reinterpret_cast<uint32_t*>(a.ss + (size_t)inst * kSaberSs)[tid] =
          (kk[tid] & ~mask) | (kk[kSaberSs / 4 + tid] & mask);
    }
  }
}

// IPW instances per workgroup, one wave each (instances-per-workgroup is a
// tuning knob: it changes how many workgroups the dispatcher places and how
// the waves spread over CUs, not the arithmetic).
template <int MODE, int L, int IPW = 1, bool A_PLAIN = false>
__global__ void __launch_bounds__(64 * IPW) __attribute__((amdgpu_waves_per_eu(4))) sab_poly(SabPolyArgs a) {
  __shared__ __align__(16) SabPolyLds<L> S[IPW];
  const int w = IPW == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int inst = (int)blockIdx.x * IPW + w;
  if (inst >= a.n) return;
  sab_poly_body<MODE, L, A_PLAIN>(a, S[w], IPW == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63), inst);
}

// Decaps inner product v = b' . s and message recovery
// m'_t = ((v + h2 - c_m 2^(10 - eT)) >> 9) & 1.  Two instances per 64-thread
// block; s (from sk) is the Toeplitz-staged factor, b' (from ct) the plain one.
template <int L>
struct SabInnerLds {
  uint16_t Ts[2][L * kToepStride];
  uint16_t Bp[2][L * kARow];
};

template <int L>
__global__ void __launch_bounds__(64) sab_dec_inner(int n, const uint8_t* __restrict__ ct,
                                                    const uint8_t* __restrict__ sk, uint8_t* __restrict__ mbuf) {
  using P = SaberP<L>;
  __shared__ __align__(16) SabInnerLds<L> S;
  const int half = threadIdx.x >> 5, t = threadIdx.x & 31;
  const int inst = blockIdx.x * 2 + half;
  const bool live = inst < n;
  const uint8_t* c = ct + (size_t)(live ? inst : 0) * P::ct;
  const uint8_t* s = sk + (size_t)(live ? inst : 0) * P::sk;
  if (live) {
    #pragma unroll 1
    for (int g = t; g < 64 * L; g += 32) {  // b' from ct (10-bit)
      uint8_t bytes[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) bytes[i] = c[5 * g + i];
      uint16_t v[4];
      unpack10_group(v, bytes);
      uint2 q;
      q.x = v[0] | ((uint32_t)v[1] << 16);
      q.y = v[2] | ((uint32_t)v[3] << 16);
      *reinterpret_cast<uint2*>(S.Bp[half] + (g >> 6) * kARow + (g & 63) * 4) = q;
    }
    #pragma unroll 1
    for (int g = t; g < 32 * L; g += 32) {  // s from sk (13-bit)
      uint8_t bytes[13];
#pragma unroll
      for (int i = 0; i < 13; ++i) bytes[i] = s[13 * g + i];
      uint16_t v[8];
      unpack13_group(v, bytes);
#pragma unroll
      for (int i = 0; i < 8; ++i) toep_put(S.Ts[half], g >> 5, (g & 31) * 8 + i, v[i]);
    }
  }
  __syncthreads();
  if (!live) return;
  uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int t0 = 8 * t;
#pragma unroll 1
  for (int j = 0; j < L; ++j) toep_dot(acc, S.Ts[half], j, S.Bp[half] + j * kARow, t0);
  uint64_t cm = 0;  // eT bits for each of t0..t0+7
#pragma unroll
  for (int i = 0; i < P::et; ++i) cm |= (uint64_t)c[P::vecp + P::et * t + i] << (8 * i);
  uint32_t byte = 0;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const uint32_t cmd = (uint32_t)(cm >> (P::et * d)) & ((1u << P::et) - 1);
    byte |= (((acc[d] + P::h2 - (cmd << (10 - P::et))) >> 9) & 1u) << d;
  }
  mbuf[(size_t)inst * 32 + t] = (uint8_t)byte;
}

// Decaps' inner product v = b'^T s of one instance on a whole wave, on the
// int8 MFMA Toeplitz tiles of the products (saber_poly.hpp): per K-block j,
// T(s_j) (s from sk, 13-bit two's complement, small) is the staged factor and
// b'_j (from ct, 10-bit) the B operand, its low limbs in column 0 and high
// limbs in column 8; then m'_t = ((v_t + h2 - c_m,t 2^(10 - eT)) >> 9) & 1,
// as in sab_dec_inner.  The instance's raw bytes (s, b', c_m) arrive in
// registers (InnerRaw, loaded one instance ahead by the caller) and pass
// through LDS once; all accesses are by the one wave, so LDS ordering needs
// no workgroup barrier (wave_sync).
template <int L>
struct InnerRaw {
  static constexpr int SW = 104 * L, BW = 80 * L, CW = 8 * SaberP<L>::et;  // dwords of s, b', c_m
  uint32_t s[(SW + 63) / 64], b[(BW + 63) / 64], c;
};

template <int L>
__device__ __forceinline__ void load_inner_raw(int lane, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ sk,
                                               InnerRaw<L>& r) {
  using R = InnerRaw<L>;
  const uint32_t* sg = reinterpret_cast<const uint32_t*>(sk);
  const uint32_t* cg = reinterpret_cast<const uint32_t*>(ct);
#pragma unroll
  for (int i = 0; i < (R::SW + 63) / 64; ++i)
    if (lane + 64 * i < R::SW) r.s[i] = sg[lane + 64 * i];
#pragma unroll
  for (int i = 0; i < (R::BW + 63) / 64; ++i)
    if (lane + 64 * i < R::BW) r.b[i] = cg[lane + 64 * i];
  if (lane < R::CW) r.c = cg[R::BW + lane];
}

template <int L>
struct InnerLds {
  union {
    struct {
      alignas(16) int8_t V[kVPoly];                       // Toeplitz copies of s_j
      alignas(16) int8_t Lb[2 * kLbSlot];                 // b'_j limbs: slot 0 low, slot 1 high
      alignas(16) uint32_t sb[InnerRaw<L>::SW + 4];       // s (13-bit), + slack for the last window
      alignas(16) uint32_t bb[InnerRaw<L>::BW + 4];       // b' (10-bit) + slack
      alignas(16) uint32_t cb[InnerRaw<L>::CW];           // c_m (eT-bit)
    } k;
    alignas(16) int32_t part[2][256];                     // column sums: low, high limbs
  } u;
};

template <int L>
__device__ __forceinline__ uint32_t dec_inner_mfma(int lane, const InnerRaw<L>& r, InnerLds<L>& W) {
  using P = SaberP<L>;
  using R = InnerRaw<L>;
#pragma unroll
  for (int i = 0; i < (R::SW + 63) / 64; ++i)
    if (lane + 64 * i < R::SW) W.u.k.sb[lane + 64 * i] = r.s[i];
#pragma unroll
  for (int i = 0; i < (R::BW + 63) / 64; ++i)
    if (lane + 64 * i < R::BW) W.u.k.bb[lane + 64 * i] = r.b[i];
  if (lane < R::CW) W.u.k.cb[lane] = r.c;
  v4i8x4 acc[16];
#pragma unroll
  for (int t = 0; t < 16; ++t) acc[t] = v4i8x4{0, 0, 0, 0};
  const int slot = toep8_slot(lane & 15, 0, true);
  wave_sync();
#pragma unroll 1
  for (int j = 0; j < L; ++j) {
    {
      // s_j coefficients 4g .. 4g + 3 (g = lane): the 13-bit fields at bit 52 g
      const int bit = 52 * lane, off = 416 * j + (bit >> 3);
      const uint32_t* d = reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(W.u.k.sb) + (off & ~3));
      const uint32_t d0 = d[0], d1 = d[1], d2 = d[2];
      const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, off & 3), w1 = __builtin_amdgcn_alignbyte(d2, d1, off & 3);
      const uint32_t sh = bit & 7;  // 0 or 4
      const uint32_t x[4] = {(w0 >> sh) & 0x1FFF, (w0 >> (sh + 13)) & 0x1FFF,
                             __builtin_amdgcn_alignbit(w1, w0, sh + 26) & 0x1FFF, (w1 >> (sh + 7)) & 0x1FFF};
      uint32_t pv = 0, nv = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int v = (int)(x[k] << 19) >> 19;  // 13-bit two's complement
        pv |= ((uint32_t)v & 0xFF) << (8 * (3 - k));
        nv |= ((uint32_t)(-v) & 0xFF) << (8 * (3 - k));
      }
      toep8_put_wave(W.u.k.V, lane, pv, nv);
      // b'_j coefficients 4g .. 4g + 3: 10-bit fields of 5 bytes
      const uint64_t y = lds_bytes64(reinterpret_cast<const uint8_t*>(W.u.k.bb), 320 * j + 5 * lane);
      const uint32_t y0 = (uint32_t)y, y1 = (uint32_t)(y >> 32);
      uint32_t lo, hi;
      limbs8x4((y0 & 0x3FF) | (((y0 >> 10) & 0x3FF) << 16),
               ((y0 >> 20) & 0x3FF) | ((__builtin_amdgcn_alignbit(y1, y0, 30) & 0x3FF) << 16), &lo, &hi);
      *reinterpret_cast<uint32_t*>(W.u.k.Lb + 4 * lane) = lo;
      *reinterpret_cast<uint32_t*>(W.u.k.Lb + kLbSlot + 4 * lane) = hi;
    }
    wave_sync();
    v4i8x4 bop[4];
    const int8_t* colp = W.u.k.Lb + (slot < 0 ? 0 : slot) * kLbSlot + 16 * (lane >> 4);
#pragma unroll
    for (int bq = 0; bq < 4; ++bq)
      bop[bq] = slot >= 0 ? *reinterpret_cast<const v4i8x4*>(colp + 64 * bq) : v4i8x4{0, 0, 0, 0};
    toep8_block(acc, W.u.k.V, bop, lane, [](v4i8x4 x, v4i8x4 y, v4i8x4 z) {
      return __builtin_amdgcn_mfma_i32_16x16x64_i8(x, y, z, 0, 0, 0);
    });
    wave_sync();
  }
  uint64_t cm = 0;  // eT bits for each of 8 lane .. 8 lane + 7 (read before the sums overwrite cb)
  if (lane < 32) {
    const uint8_t* cbb = reinterpret_cast<const uint8_t*>(W.u.k.cb);
#pragma unroll
    for (int i = 0; i < P::et; ++i) cm |= (uint64_t)cbb[P::et * lane + i] << (8 * i);
  }
  wave_sync();
  if (slot >= 0) {
#pragma unroll
    for (int t = 0; t < 16; ++t)
      *reinterpret_cast<int4*>(&W.u.part[slot][16 * t + 4 * (lane >> 4)]) =
          make_int4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
  }
  wave_sync();
  uint32_t byte = 0;
  if (lane < 32) {
#pragma unroll
    for (int dd = 0; dd < 8; ++dd) {
      const uint32_t v = (uint32_t)W.u.part[0][8 * lane + dd] + 256u * (uint32_t)W.u.part[1][8 * lane + dd];
      const uint32_t cmd = (uint32_t)(cm >> (P::et * dd)) & ((1u << P::et) - 1);
      byte |= (((v + P::h2 - (cmd << (10 - P::et))) >> 9) & 1u) << dd;
    }
  }
  wave_sync();  // the next instance overwrites the sums
  return byte;
}

// x2 of the pipelined chain: one workgroup of twelve waves per 16
// instances.  A workgroup's waves go to the CU's four SIMDs in a repeating
// order of period four (measured on MI355X: waves w and w + 4 always share a
// SIMD), so the roles are laid out by w mod 4 and the two long hash chains
// each get a SIMD of their own from the start:
//   w = 0          encaps' K = SHA3-256(kr[0:32] || H(ct));
//   w = 1          decaps' H(ct), handed to wave 3 through LDS;
//   w = 3          deinterleaves both chains' copy of ct into LDS block by
//                  block (stage2_message: lanes 0..31 for encaps' K, 32..63
//                  for decaps' H(ct)), so the two chains absorb without
//                  conversion work; then, once decaps' H(ct) (wave 1) and
//                  K-hat' (the chain wave, after G) are posted, one
//                  permutation on all 64 lanes for both of decaps' possible
//                  K: SHA3-256(K-hat' || H(ct)) on lanes 0..31 and the
//                  implicit-rejection SHA3-256(z || H(ct)) on lanes 32..63,
//                  into kk_d (64 B per instance);
//   w = 2, 6, 10, 7, 11  decaps' inner products on int8 MFMA, instances w',
//                  w' + 5, ... (w' = 0 .. 4) one after the other (three waves
//                  per SIMD hide the Toeplitz operand reads' latency); m'
//                  goes to LDS and to mbuf_d (the re-encryption reads it).
//                  Wave kX2ChainWave then waits until all five have posted
//                  (an LDS counter, not a workgroup barrier, so the hash
//                  waves never wait) and runs decaps' message chain (G,
//                  GenSecret) for the 16 instances on lane pairs;
//   w = 4, 5, 8, 9 leave at once (they only keep the layout).
// Twelve waves (three per SIMD) leave each wave up to 168 registers: the
// MFMA accumulators of the inner products and the hash chains' prefetched
// blocks fit without spills.
constexpr int kX2Group = 16;
constexpr int kX2InnerWaves = 5;
constexpr int kX2ChainWave = 2;
constexpr int kX2HelperWave = 3;
constexpr int kX2Threads = 768;

template <int L>
struct X2Lds {
  static constexpr int kCtBlocks = (SaberP<L>::ct / 8 + kRateSha3_256 - 1) / kRateSha3_256;
  InnerLds<L> w[kX2InnerWaves];
  uint32_t msg[2][kCtBlocks * kRateSha3_256 * 2 * kX2Group];  // ct halves for encaps' K / decaps' H(ct)
  alignas(8) uint8_t mq[kX2Group][32];
  uint32_t hq[4][2 * kX2Group];  // decaps' H(ct) halves by lane-pair slot (wave 1 -> the K wave)
  uint32_t kq[4][2 * kX2Group];  // decaps' K-hat' halves by slot (the chain wave -> the K wave)
  uint32_t posted, ready, hc_ready, kq_ready;
};

template <int L>
__global__ void __launch_bounds__(kX2Threads) sab_pipe_x2(int n, const uint8_t* __restrict__ sk,
                                                          const uint8_t* __restrict__ ct, uint8_t* __restrict__ mbuf_d,
                                                          uint8_t* __restrict__ sraw_d, uint8_t* __restrict__ krbuf_d,
                                                          const uint8_t* __restrict__ krbuf_e, uint8_t* __restrict__ ss,
                                                          uint8_t* __restrict__ kk_d) {
  using P = SaberP<L>;
  constexpr int CW = P::ct / 8, SL = 2 * kX2Group;
  __shared__ X2Lds<L> S;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = blockIdx.x * kX2Group;
  if (threadIdx.x == 0) {
    S.posted = 0;
    S.ready = 0;
    S.hc_ready = 0;
    S.kq_ready = 0;
  }
  __syncthreads();  // the only workgroup barrier: every wave passes it before any work
  if ((wave & 3) < 2) {
    auto staged = [&](uint32_t* a, uint32_t p, int, int ln) {
      k2::absorb2_staged<kRateSha3_256, CW, SL>(a, p, S.msg[wave], &S.ready, ln, kDsSha3);
    };
    if (wave == 0) {
      hash2_core<L, kX2Group>(blockIdx.x, n, staged, krbuf_e, nullptr, nullptr, ss);
    } else if (wave == 1) {  // decaps' H(ct), handed to the K wave through LDS
      const uint32_t p = lane & 1;
      const int k = lane >> 1;
      if (k < kX2Group && base + k < n) {
        uint32_t a[25];
        staged(a, p, base + k, lane);
#pragma unroll
        for (int i = 0; i < 4; ++i) S.hq[i][lane] = a[i];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_fetch_add(&S.hc_ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return;
  }
  if (wave == kX2HelperWave) {
    {
      const int k = (lane & 31) >> 1, inst = min(base + k, n - 1);  // absent instances stage a copy (unused)
      const uint64_t* w = reinterpret_cast<const uint64_t*>(ct + (size_t)inst * P::ct);
      k2::stage2_message<kRateSha3_256, CW, SL>(S.msg[lane >> 5], &S.ready, lane & 1, lane & 31, w, lane == 0);
    }
    // then decaps' K for both outcomes of the re-encryption's compare, once
    // H(ct) and K-hat' are posted: lanes 0..31 K = SHA3-256(K-hat' || H(ct)),
    // lanes 32..63 the implicit-rejection K = SHA3-256(z || H(ct)); the
    // re-encryption launch picks one per instance in constant time
    const int s = lane & 31, k = s >> 1, inst = base + k;
    const uint32_t p = lane & 1;
    const bool rej = lane >= 32;
    uint64_t z[4];
    const uint8_t* zs = sk + (size_t)min(inst, n - 1) * P::sk + P::sk - 32;
#pragma unroll
    for (int i = 0; i < 4; ++i) z[i] = ld64(zs + 8 * i);
    k2::lds_wait_count(&S.hc_ready, 1u);
    k2::lds_wait_count(&S.kq_ready, 1u);
    uint32_t m[8], a[25];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      m[i] = rej ? k2::deint(z[i], p) : S.kq[i][s];
      m[4 + i] = S.hq[i][s];
    }
    k2::absorb2_halves<kRateSha3_256, 8>(a, p, m, kDsSha3);
    if (inst < n) store2_words<4>(a, p, kk_d + (size_t)inst * 64 + (rej ? 32 : 0));
    return;
  }
  const int iw = wave == 2 ? 0 : wave == 6 ? 1 : wave == 10 ? 2 : wave == 7 ? 3 : 4;  // inner waves 2, 6, 10, 7, 11
  uint64_t hpk[4];  // the chain wave's H(pk) words, loaded before the inner products
  if (wave == kX2ChainWave) load_hpk<L>(min(base + (lane >> 1), n - 1), sk, hpk);
  InnerRaw<L> raw;  // the next instance's bytes, loaded while the current one is multiplied
  {
    const int ci = min(base + iw, n - 1);
    load_inner_raw<L>(lane, ct + (size_t)ci * P::ct, sk + (size_t)ci * P::sk, raw);
  }
#pragma unroll 1
  for (int k = iw; k < kX2Group; k += kX2InnerWaves) {
    const int inst = base + k;
    const bool live = inst < n;
    const InnerRaw<L> cur = raw;
    if (k + kX2InnerWaves < kX2Group) {
      const int ci = min(inst + kX2InnerWaves, n - 1);
      load_inner_raw<L>(lane, ct + (size_t)ci * P::ct, sk + (size_t)ci * P::sk, raw);
    }
    const uint32_t byte = dec_inner_mfma<L>(lane, cur, S.w[iw]);
    if (live && lane < 32) {
      S.mq[k][lane] = (uint8_t)byte;
      mbuf_d[(size_t)inst * 32 + lane] = (uint8_t)byte;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) __hip_atomic_fetch_add(&S.posted, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (wave != kX2ChainWave) return;
  // each inner wave posts once after bounded work (lds_wait_count)
  k2::lds_wait_count(&S.posted, (uint32_t)kX2InnerWaves);
  const uint32_t p = lane & 1;
  const int k = lane >> 1;
  if (k >= kX2Group || base + k >= n) return;
  dec_msg_chain<L>(base + k, p, hpk, S.mq[k], sraw_d, krbuf_d, &S.kq[0][0], 2 * kX2Group, lane, &S.kq_ready);
}

// ----------------------------------------------------------------- independent ops in shared launches
// Independent keygen, encaps and decaps tasks of one rank (a plan's tasks
// with no instance-wise dependency between them) run as ONE sequence of
// three launches instead of three sequences of three or four:
//   mix_x0    256 threads per workgroup.  Workgroups [0, gke) serve 32
//             keygen and 32 encaps instances each: wave 0 keygen's seedA and
//             GenMatrix (the longest chain, 1 + 23 permutations for l = 3),
//             wave 1 keygen's GenSecret, wave 2 encaps' GenMatrix, wave 3
//             encaps' m, H(pk), G and GenSecret.  Workgroups [gke, gke + gd)
//             serve 32 decaps instances: wave 0 GenMatrix; wave 1 H(ct),
//             then decaps' K for both outcomes of the re-encryption's compare
//             (the implicit-rejection K at once, K from K-hat' once the
//             message chain posts it); waves 2 and 3 the inner products on
//             int8 MFMA (16 instances each), then wave 2 decaps' message
//             chain (G, GenSecret) for all 32 on lane pairs.
//   mix_poly  keygen's, encaps' and the re-encryption's products in one grid;
//             the re-encryption selects decaps' K in constant time.
//   mix_x1    keygen's H(pk) into sk beside encaps' K.
// Every op still computes all of its own hashes from its own inputs.  A
// launch of 256 such workgroups places one per CU and a workgroup's four
// waves on the CU's four SIMDs, so at 4096 instances of each op every chain
// of mix_x0 has a SIMD to itself; launched on separate streams the same
// chains share SIMDs and take longer.
constexpr int kMixGroup = 32;

struct SabMixArgs {
  int nk, ne, nd, gke;  // instances per op; workgroups of keygen + encaps
  const uint8_t* kc;
  uint8_t *pk_k, *sk_k, *araw_k, *sraw_k;
  const uint8_t *pk_e, *ec;
  uint8_t *ct_e, *ss_e, *araw_e, *sraw_e, *mbuf_e, *krbuf_e;
  const uint8_t *ct_d, *sk_d;
  uint8_t *ss_d, *araw_d, *sraw_d, *mbuf_d, *krbuf_d, *kk_d;
  // rolling pipeline only (sab_roll_x0): the closing hashes of the two
  // previous steps' batches -- keygen's H(pk) into sk_t for the keys pk_t of
  // one step earlier, encaps' K into ss_t for the ciphertexts ct_t (and kr
  // in krbuf_t) of two steps earlier; all n = nk = ne = nd instances
  const uint8_t* pk_t;
  uint8_t* sk_t;
  const uint8_t *ct_t, *krbuf_t;
  uint8_t* ss_t;
};

template <int L>
struct MixDecLds {
  InnerLds<L> w[2];
  alignas(8) uint8_t mq[kMixGroup][32];
  uint32_t kq[4][2 * kMixGroup];  // decaps' K-hat' halves by lane (wave 2's chain -> wave 1)
  uint32_t posted, kq_ready;
};

// Every squeeze of mix_x0 (keygen's, encaps' and decaps' GenMatrix, keygen's,
// encaps' and decaps' GenSecret) goes through one call site with a run-time
// length, so the kernel holds one copy of the unrolled permutation: with a
// copy per stream the four waves of a CU ran four ~20 KB loops through the
// instruction cache.  Each wave first brings its sponge to its first output
// block (absorb phase, role by role), then squeezes.
//
// ROLL (sab_roll_x0, rolling pipeline): keygen's GenSecret moves to wave 3 of
// the decaps workgroup, after its inner products (about 14 + 6 of the 24
// permutation-times of the longest chain), and wave 1 of the keygen / encaps
// workgroup runs the previous batches' closing hashes instead (keygen's
// H(pk), 8 permutations, then encaps' H(ct) and K, 10): measured per-role
// times leave every wave within the keygen GenMatrix chain's length, so the
// closing hashes need no launch of their own.
template <int L, bool ROLL>
__device__ __forceinline__ void x0_body(const SabMixArgs& m, MixDecLds<L>& S) {
  using P = SaberP<L>;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t p = lane & 1;
  const bool dgrp = (int)blockIdx.x >= m.gke;
  const int g = dgrp ? (int)blockIdx.x - m.gke : (int)blockIdx.x;
  const int inst = g * kMixGroup + (lane >> 1);
  uint32_t a[25];
  uint8_t* dst;  // this wave's stream (sliced layout)
  int outw;      // its length in 64-bit words
  if (!dgrp) {
    if (wave == 0) {  // keygen: seedA = SHAKE128(seedA_raw, 32) into pk and sk, then GenMatrix
      if (inst >= m.nk) return;
      const uint8_t* c = m.kc + (size_t)inst * kSaberKeygenCoins;
      uint64_t w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = ld64(c + 8 * i);
      absorb2_regs<kRateShake128, 4>(a, p, w, kDsShake);
      uint32_t h[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) h[i] = a[i];
      store2_words<4>(h, p, m.pk_k + (size_t)inst * P::pk + P::vecp);
      store2_words<4>(h, p, m.sk_k + (size_t)inst * P::sk + P::vecq + P::vecp);
      absorb2_halves<kRateShake128, 4>(a, p, h, kDsShake);
      dst = m.araw_k + (size_t)inst * P::matrix_bytes;
      outw = P::matrix_bytes / 8;
    } else if (ROLL && wave == 1) {  // the closing hashes of the previous batches
      if (inst >= m.nk) return;
      absorb2_global<kRateSha3_256, P::pk / 8>(a, p, m.pk_t + (size_t)inst * P::pk, kDsSha3);  // keygen's H(pk)
      store2_words<4>(a, p, m.sk_t + (size_t)inst * P::sk + P::sk - 64);
      uint64_t kw[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) kw[i] = ld64(m.krbuf_t + (size_t)inst * 64 + 8 * i);
      absorb2_global<kRateSha3_256, P::ct / 8>(a, p, m.ct_t + (size_t)inst * P::ct, kDsSha3);  // encaps' H(ct)
      uint32_t mm[8];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        mm[i] = k2::deint(kw[i], p);
        mm[4 + i] = a[i];
      }
      absorb2_halves<kRateSha3_256, 8>(a, p, mm, kDsSha3);  // K = SHA3-256(kr[0:32] || H(ct))
      store2_words<4>(a, p, m.ss_t + (size_t)inst * kSaberSs);
      return;
    } else if (wave == 1) {  // keygen: GenSecret stream, z into sk
      if (inst >= m.nk) return;
      const uint8_t* c = m.kc + (size_t)inst * kSaberKeygenCoins;
      uint64_t w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = ld64(c + 32 + 8 * i);
      absorb2_regs<kRateShake128, 4>(a, p, w, kDsShake);
      uint32_t* z = reinterpret_cast<uint32_t*>(m.sk_k + (size_t)inst * P::sk + P::sk - 32);
      const uint32_t* zc = reinterpret_cast<const uint32_t*>(c + 64);
#pragma unroll
      for (int i = 0; i < 4; ++i) z[2 * i + p] = zc[2 * i + p];
      dst = m.sraw_k + (size_t)inst * P::secret_bytes;
      outw = P::secret_bytes / 8;
    } else if (wave == 2) {  // encaps: GenMatrix from the seedA bytes of pk
      if (inst >= m.ne) return;
      uint64_t w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = ld64(m.pk_e + (size_t)inst * P::pk + P::vecp + 8 * i);
      absorb2_regs<kRateShake128, 4>(a, p, w, kDsShake);
      dst = m.araw_e + (size_t)inst * P::matrix_bytes;
      outw = P::matrix_bytes / 8;
    } else {  // encaps: m, H(pk), G, then the GenSecret stream
      if (inst >= m.ne) return;
      uint32_t gw[8];
      enc_m(inst, p, m.ec, m.mbuf_e, gw);
      absorb2_global<kRateSha3_256, P::pk / 8>(a, p, m.pk_e + (size_t)inst * P::pk, kDsSha3);  // H(pk)
#pragma unroll
      for (int i = 0; i < 4; ++i) gw[4 + i] = a[i];
      enc_g_absorb(inst, p, gw, m.krbuf_e, a);
      dst = m.sraw_e + (size_t)inst * P::secret_bytes;
      outw = P::secret_bytes / 8;
    }
  } else {
    const int base = g * kMixGroup, n = m.nd;
    if (threadIdx.x == 0) {
      S.posted = 0;
      S.kq_ready = 0;
    }
    __syncthreads();  // the only workgroup barrier (decaps workgroups)
    if (wave == 1) {  // H(ct), then decaps' K for both outcomes of the re-encryption's compare
      if (inst >= n) return;  // lane pairs of absent instances (lane 0 is always present)
      uint64_t z[4];
      const uint8_t* zs = m.sk_d + (size_t)inst * P::sk + P::sk - 32;
#pragma unroll
      for (int i = 0; i < 4; ++i) z[i] = ld64(zs + 8 * i);
      uint32_t hc[4], mm[8];
      absorb2_global<kRateSha3_256, P::ct / 8>(a, p, m.ct_d + (size_t)inst * P::ct, kDsSha3);  // H(ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        hc[i] = a[i];
        mm[i] = k2::deint(z[i], p);
        mm[4 + i] = hc[i];
      }
      absorb2_halves<kRateSha3_256, 8>(a, p, mm, kDsSha3);  // implicit-rejection K = SHA3-256(z || H(ct))
      store2_words<4>(a, p, m.kk_d + (size_t)inst * 64 + 32);
      k2::lds_wait_count(&S.kq_ready, 1u);
#pragma unroll
      for (int i = 0; i < 4; ++i) mm[i] = S.kq[i][lane];
      absorb2_halves<kRateSha3_256, 8>(a, p, mm, kDsSha3);  // K = SHA3-256(K-hat' || H(ct))
      store2_words<4>(a, p, m.kk_d + (size_t)inst * 64);
      return;
    }
    if (wave == 0) {  // decaps: GenMatrix from the seedA bytes of sk
      if (inst >= n) return;
      uint64_t w[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = ld64(m.sk_d + (size_t)inst * P::sk + P::vecq + P::vecp + 8 * i);
      absorb2_regs<kRateShake128, 4>(a, p, w, kDsShake);
      dst = m.araw_d + (size_t)inst * P::matrix_bytes;
      outw = P::matrix_bytes / 8;
    } else {
      // waves 2 and 3: inner products of instances iw, iw + 2, ... (the next
      // instance's bytes loaded while the current one is multiplied); then
      // wave 2 runs decaps' message chain (G, GenSecret) for all 32
      const int iw = wave - 2;
      uint64_t hpk[4];
      if (wave == 2) load_hpk<L>(min(inst, n - 1), m.sk_d, hpk);
      InnerRaw<L> raw;
      {
        const int ci = min(base + iw, n - 1);
        load_inner_raw<L>(lane, m.ct_d + (size_t)ci * P::ct, m.sk_d + (size_t)ci * P::sk, raw);
      }
#pragma unroll 1
      for (int k = iw; k < kMixGroup && base + k < n; k += 2) {
        const int ki = base + k;
        const InnerRaw<L> cur = raw;
        if (k + 2 < kMixGroup) {
          const int ci = min(ki + 2, n - 1);
          load_inner_raw<L>(lane, m.ct_d + (size_t)ci * P::ct, m.sk_d + (size_t)ci * P::sk, raw);
        }
        const uint32_t byte = dec_inner_mfma<L>(lane, cur, S.w[iw]);
        if (lane < 32) {
          S.mq[k][lane] = (uint8_t)byte;
          m.mbuf_d[(size_t)ki * 32 + lane] = (uint8_t)byte;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_fetch_add(&S.posted, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (ROLL && wave == 3) {  // keygen: GenSecret stream, z into sk
        if (inst >= m.nk) return;
        const uint8_t* c = m.kc + (size_t)inst * kSaberKeygenCoins;
        uint64_t w[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = ld64(c + 32 + 8 * i);
        absorb2_regs<kRateShake128, 4>(a, p, w, kDsShake);
        uint32_t* z = reinterpret_cast<uint32_t*>(m.sk_k + (size_t)inst * P::sk + P::sk - 32);
        const uint32_t* zc = reinterpret_cast<const uint32_t*>(c + 64);
#pragma unroll
        for (int i = 0; i < 4; ++i) z[2 * i + p] = zc[2 * i + p];
        dst = m.sraw_k + (size_t)inst * P::secret_bytes;
        outw = P::secret_bytes / 8;
      } else {
        if (wave != 2) return;
        k2::lds_wait_count(&S.posted, 2u);  // each inner wave posts once after bounded work
        if (inst >= n) return;
        dec_msg_absorb(inst, p, hpk, S.mq[lane >> 1], m.krbuf_d, a, &S.kq[0][0], 2 * kMixGroup, lane, &S.kq_ready);
        dst = m.sraw_d + (size_t)inst * P::secret_bytes;
        outw = P::secret_bytes / 8;
      }
    }
  }
  k2::squeeze2_sliced_rt<kRateShake128>(a, p, reinterpret_cast<uint32_t*>(dst), outw);
}

template <int L>
__global__ void __launch_bounds__(256) sab_mix_x0(SabMixArgs m) {
  __shared__ MixDecLds<L> S;
  x0_body<L, false>(m, S);
}

template <int L>
__global__ void __launch_bounds__(256) sab_roll_x0(SabMixArgs m) {
  __shared__ MixDecLds<L> S;
  // In the overlapped pipeline this launch shares its SIMDs with the
  // previous step's products: its waves (chains, and the helpers the chains
  // wait on) take issue priority, and the product waves fill the gaps.
  __builtin_amdgcn_s_setprio(3);
  x0_body<L, true>(m, S);
}

// Products of the three ops: workgroups [0, bk) keygen, [bk, bk + be)
// encaps, the rest the re-encryption; IPW one-wave instances per workgroup.
// Workgroup b of the products grid (IPW one-wave instances; w this wave's
// instance slot, tid its lane).
template <int L, int IPW>
__device__ __forceinline__ void mix_poly_wg(const SabPolyArgs& ak, const SabPolyArgs& ae, const SabPolyArgs& ad,
                                            int bk, int be, SabPolyLds<L>* S, int b, int w, int tid) {
  if (b < bk) {
    const int inst = b * IPW + w;
    if (inst < ak.n) sab_poly_body<kSabKeygen, L>(ak, S[w], tid, inst);
    return;
  }
  b -= bk;
  if (b < be) {
    const int inst = b * IPW + w;
    if (inst < ae.n) sab_poly_body<kSabEnc, L>(ae, S[w], tid, inst);
    return;
  }
  b -= be;
  const int inst = b * IPW + w;
  if (inst < ad.n) sab_poly_body<kSabReenc, L>(ad, S[w], tid, inst);
}

template <int L, int IPW>
__global__ void __launch_bounds__(64 * IPW) __attribute__((amdgpu_waves_per_eu(4)))
sab_mix_poly(SabPolyArgs ak, SabPolyArgs ae, SabPolyArgs ad, int bk, int be) {
  __shared__ __align__(16) SabPolyLds<L> S[IPW];
  const int w = IPW == 1 ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tid = IPW == 1 ? (int)threadIdx.x : (int)(threadIdx.x & 63);
  mix_poly_wg<L, IPW>(ak, ae, ad, bk, be, S, blockIdx.x, w, tid);
}

// One overlapped rolling step in one launch: workgroups [0, hash_wgs) are
// sab_roll_x0's (this step's chains and closing hashes, at issue priority),
// the rest the previous step's products, four one-wave instances each.  The
// dispatcher places the chain workgroups first, one per CU, and the product
// workgroups fill the remaining wave slots, so the products of step k - 1
// run beside the chains of step k without a second stream or an event.
template <int L>
__global__ void __launch_bounds__(256) sab_roll_fused(SabMixArgs m, int hash_wgs, SabPolyArgs ak, SabPolyArgs ae,
                                                      SabPolyArgs ad, int bk, int be) {
  __shared__ __align__(16) union FusedLds {
    MixDecLds<L> h;
    SabPolyLds<L> p[4];
  } U;
  if ((int)blockIdx.x < hash_wgs) {
    __builtin_amdgcn_s_setprio(3);
    x0_body<L, true>(m, U.h);
    return;
  }
  mix_poly_wg<L, 4>(ak, ae, ad, bk, be, U.p, (int)blockIdx.x - hash_wgs,
                    __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), (int)(threadIdx.x & 63));
}

// Closing hashes of 32 keygen and 32 encaps instances per workgroup: wave 0
// keygen's H(pk) into sk, wave 1 encaps' H(ct) and K; waves 2 and 3
// deinterleave pk and ct into LDS block by block ahead of them
// (stage2_message), so the two chains absorb without conversion work, through
// one shared absorb call site (one copy of the unrolled permutation).
template <int L>
struct MixX1Lds {
  static constexpr int kPkBlocks = (SaberP<L>::pk / 8 + kRateSha3_256 - 1) / kRateSha3_256;
  static constexpr int kCtBlocks = (SaberP<L>::ct / 8 + kRateSha3_256 - 1) / kRateSha3_256;
  uint32_t pkm[kPkBlocks * kRateSha3_256 * 2 * kMixGroup];
  uint32_t ctm[kCtBlocks * kRateSha3_256 * 2 * kMixGroup];
  uint32_t pk_ready, ct_ready;
};

template <int L>
__global__ void __launch_bounds__(256) sab_mix_x1(SabMixArgs m) {
  using P = SaberP<L>;
  __shared__ MixX1Lds<L> S;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const uint32_t p = lane & 1;
  const int base = (int)blockIdx.x * kMixGroup, inst = base + (lane >> 1);
  if (threadIdx.x == 0) {
    S.pk_ready = 0;
    S.ct_ready = 0;
  }
  __syncthreads();  // the only workgroup barrier
  if (wave >= 2) {  // helpers; absent instances stage a copy of the last one (unused)
    const bool kg = wave == 2;
    const int n = kg ? m.nk : m.ne;
    if (base >= n) return;
    const uint8_t* src = kg ? m.pk_k + (size_t)min(inst, n - 1) * P::pk : m.ct_e + (size_t)min(inst, n - 1) * P::ct;
    if (kg)
      k2::stage2_message<kRateSha3_256, P::pk / 8, 2 * kMixGroup>(S.pkm, &S.pk_ready, p, lane,
                                                                  reinterpret_cast<const uint64_t*>(src), lane == 0);
    else
      k2::stage2_message<kRateSha3_256, P::ct / 8, 2 * kMixGroup>(S.ctm, &S.ct_ready, p, lane,
                                                                  reinterpret_cast<const uint64_t*>(src), lane == 0);
    return;
  }
  const bool kg = wave == 0;
  if (inst >= (kg ? m.nk : m.ne)) return;
  uint32_t a[25];
  uint64_t kw[4];  // encaps' kr[0:32], loaded before H(ct) so that its latency is off the chain
  if (!kg) {
#pragma unroll
    for (int i = 0; i < 4; ++i) kw[i] = ld64(m.krbuf_e + (size_t)inst * 64 + 8 * i);
  }
  k2::absorb2_staged_rt<kRateSha3_256, 2 * kMixGroup>(a, p, kg ? S.pkm : S.ctm, kg ? &S.pk_ready : &S.ct_ready, lane,
                                                      kg ? P::pk / 8 : P::ct / 8, kDsSha3);
  if (kg) {
    store2_words<4>(a, p, m.sk_k + (size_t)inst * P::sk + P::sk - 64);  // H(pk)
    return;
  }
  uint32_t mm[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    mm[i] = k2::deint(kw[i], p);
    mm[4 + i] = a[i];
  }
  absorb2_halves<kRateSha3_256, 8>(a, p, mm, kDsSha3);  // K = SHA3-256(kr[0:32] || H(ct))
  store2_words<4>(a, p, m.ss_e + (size_t)inst * kSaberSs);
}

// ----------------------------------------------------------------- host launchers
// (declared in launch.hpp; all run on the given stream)

static inline int blocks_for(int n, int per) { return (n + per - 1) / per; }

// Profiler names: Saber keeps the plain kernel names, the other ranks add ",l<L>".
template <int L>
const char* kname(const char* base, const char* l2, const char* l4) {
  return L == 3 ? base : L == 2 ? l2 : l4;
}
#define SAB_NAME(base) kname<L>(base, base ",l2", base ",l4")

// The product launch with `ipw` instances per workgroup (1, 2 or 4).
template <int MODE, int L, bool A_PLAIN = false>
static void poly_launch(const char* name, hipStream_t st, int n, const SabPolyArgs& a, int ipw) {
  if (ipw == 4) DPKEM_LAUNCH(name, (sab_poly<MODE, L, 4, A_PLAIN>), dim3(blocks_for(n, 4)), dim3(256), st, a);
  else if (ipw == 2) DPKEM_LAUNCH(name, (sab_poly<MODE, L, 2, A_PLAIN>), dim3(blocks_for(n, 2)), dim3(128), st, a);
  else DPKEM_LAUNCH(name, (sab_poly<MODE, L, 1, A_PLAIN>), dim3(n), dim3(64), st, a);
}

hipError_t saber_wait_expired(uint32_t* count, bool clear) {
  hipError_t err = hipMemcpyFromSymbol(count, HIP_SYMBOL(k2::wait_expired), sizeof(uint32_t), 0, hipMemcpyDeviceToHost);
  if (err == hipSuccess && clear && *count) {
    const uint32_t zero = 0;
    err = hipMemcpyToSymbol(HIP_SYMBOL(k2::wait_expired), &zero, sizeof zero, 0, hipMemcpyHostToDevice);
  }
  return err;
}

hipError_t saber_set_wait_spin_limit(uint32_t limit) {
  return hipMemcpyToSymbol(HIP_SYMBOL(k2::wait_spin_limit), &limit, sizeof limit, 0, hipMemcpyHostToDevice);
}

// Scratch per instance of one op run as a mixed sequence (the largest,
// decaps: GenMatrix and GenSecret streams, m', kr, both candidate K).
size_t saber_scratch_per_inst(int l) {
  const size_t m = l == 2 ? SaberP<2>::matrix_bytes : l == 3 ? SaberP<3>::matrix_bytes : SaberP<4>::matrix_bytes;
  const size_t s = l == 2 ? SaberP<2>::secret_bytes : l == 3 ? SaberP<3>::secret_bytes : SaberP<4>::secret_bytes;
  return m + s + 32 + 64 + 64;
}

// dst[i] = src for i < n (bytes per row a multiple of 4).
__global__ void __launch_bounds__(256) broadcast_rows(int n, const uint32_t* __restrict__ src, int words,
                                                      uint32_t* __restrict__ dst) {
  const size_t total = (size_t)n * words;
  for (size_t x = (size_t)blockIdx.x * 256 + threadIdx.x; x < total; x += (size_t)gridDim.x * 256)
    dst[x] = src[x % words];
}

void broadcast_rows_launch(hipStream_t st, int n, const uint8_t* src, size_t bytes, uint8_t* dst) {
  if (n <= 0) return;
  const int words = (int)(bytes / 4);
  const size_t total = (size_t)n * words;
  const int blocks = (int)std::min<size_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(broadcast_rows, dim3(blocks), dim3(256), 0, st, n, reinterpret_cast<const uint32_t*>(src), words,
                     reinterpret_cast<uint32_t*>(dst));
}

// Pipelined keygen -> encaps -> decaps of the same n instances on one stream
// (see sab_pipe_x0 .. sab_pipe_x2).  Per instance the serial Keccak path drops
// from 24 + 8 + 23 + 10 + 23 + 10 to 24 + 14 + 10 + 1 permutations (all
// three GenMatrix streams start from the seedA bytes of pk / sk as soon as
// keygen has written them; keygen's H(pk), encaps' m and K and decaps' H(ct)
// run beside other stages, decaps' inner product beside encaps' K, decaps'
// K in the re-encryption launch), and ten launches become six.  scratch:
// saber_pipe_scratch_per_inst(l) bytes per instance: three GenMatrix streams,
// two GenSecret, two m and two kr buffers (keygen's GenSecret reuses decaps'
// buffer, which the stream order makes safe) and decaps' two candidate K,
// plus 256-byte alignment of ten arrays per launch.
template <int L>
static size_t pipe_scratch() {
  using P = SaberP<L>;
  return 3 * (size_t)P::matrix_bytes + 2 * ((size_t)P::secret_bytes + 32 + 64) + 64;
}

size_t saber_pipe_scratch_per_inst(int l) {
  return l == 2 ? pipe_scratch<2>() : l == 4 ? pipe_scratch<4>() : pipe_scratch<3>();
}

template <int L>
static void pipe_l(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk, uint8_t* ct,
                   uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int ipw, int force_reject) {
  using P = SaberP<L>;
  size_t off = 0;
  auto take = [&](size_t per) {
    uint8_t* q = scratch + off;
    off += ((size_t)n * per + 255) & ~(size_t)255;
    return q;
  };
  uint8_t* araw_k = take(P::matrix_bytes);
  uint8_t* araw_d = take(P::matrix_bytes);
  uint8_t* sraw_d = take(P::secret_bytes);  // keygen's GenSecret, then decaps'
  uint8_t* araw_e = take(P::matrix_bytes);
  uint8_t* sraw_e = take(P::secret_bytes);
  uint8_t* mbuf_e = take(32);
  uint8_t* mbuf_d = take(32);
  uint8_t* krbuf_e = take(64);
  uint8_t* krbuf_d = take(64);
  uint8_t* kk_d = take(64);
  const int nb = blocks_for(n, 32);
  const dim3 g(nb);
  DPKEM_LAUNCH(SAB_NAME("sab_pipe_x0"), sab_pipe_x0<L>, g, dim3(256), st, n, kc, pk, sk, araw_k, sraw_d, araw_e,
               araw_d);
  SabPolyArgs ak{n, araw_k, sraw_d, nullptr, 0, nullptr, pk, sk, nullptr, nullptr};
  poly_launch<kSabKeygen, L>(SAB_NAME("sab_poly<kSabKeygen>"), st, n, ak, ipw);
  DPKEM_LAUNCH(SAB_NAME("sab_pipe_x1"), sab_pipe_x1<L>, g, dim3(256), st, n, pk, ec, sk, sraw_e, mbuf_e, krbuf_e,
               araw_e, araw_d);
  SabPolyArgs ae{n, araw_e, sraw_e, pk, (size_t)P::pk, mbuf_e, ct, nullptr, nullptr, nullptr};
  poly_launch<kSabEnc, L, true>(SAB_NAME("sab_poly<kSabEnc>"), st, n, ae, ipw);
  DPKEM_LAUNCH(SAB_NAME("sab_pipe_x2"), sab_pipe_x2<L>, dim3(blocks_for(n, kX2Group)), dim3(kX2Threads), st, n, sk, ct,
               mbuf_d, sraw_d, krbuf_d, krbuf_e, ss, kk_d);
  SabPolyArgs ad{n,       araw_d, sraw_d, sk + P::vecq, (size_t)P::sk, mbuf_d, nullptr, nullptr, ct, nullptr,
                 kk_d,    ss2,    force_reject};
  poly_launch<kSabReenc, L, true>(SAB_NAME("sab_reenc"), st, n, ad, ipw);
}

void saber_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int l, int ipw, int force_reject) {
  if (n <= 0) return;
  if (l == 2) pipe_l<2>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, ipw, force_reject);
  else if (l == 4) pipe_l<4>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, ipw, force_reject);
  else pipe_l<3>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, ipw, force_reject);
}

// Independent keygen / encaps / decaps tasks of one rank in one launch
// sequence (sab_mix_x0, sab_mix_poly, sab_mix_x1).  Scratch: eleven arrays,
// each 256-byte aligned: keygen's GenMatrix and GenSecret streams; encaps'
// streams, m and kr; decaps' streams, m', kr and its two candidate K.
template <int L>
static size_t mix_scratch(size_t nk, size_t ne, size_t nd) {
  using P = SaberP<L>;
  auto r = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t M = P::matrix_bytes, S = P::secret_bytes;
  return r(nk * M) + r(nk * S) + r(ne * M) + r(ne * S) + r(ne * 32) + r(ne * 64) + r(nd * M) + r(nd * S) +
         r(nd * 32) + r(nd * 64) + r(nd * 64);
}

size_t saber_mix_scratch_bytes(int l, size_t nk, size_t ne, size_t nd) {
  return l == 2 ? mix_scratch<2>(nk, ne, nd) : l == 4 ? mix_scratch<4>(nk, ne, nd) : mix_scratch<3>(nk, ne, nd);
}

template <int L, int IPW>
static void mix_poly_launch(const char* name, hipStream_t st, const SabPolyArgs& ak, const SabPolyArgs& ae,
                            const SabPolyArgs& ad) {
  const int bk = blocks_for(ak.n, IPW), be = blocks_for(ae.n, IPW), bd = blocks_for(ad.n, IPW);
  DPKEM_LAUNCH(name, (sab_mix_poly<L, IPW>), dim3(bk + be + bd), dim3(64 * IPW), st, ak, ae, ad, bk, be);
}

// The eleven scratch arrays of a mixed sequence, in mix_scratch order.
struct MixScratch {
  uint8_t *araw_k, *sraw_k, *araw_e, *sraw_e, *mbuf_e, *krbuf_e, *araw_d, *sraw_d, *mbuf_d, *krbuf_d, *kk_d;
};

template <int L>
static MixScratch mix_carve(uint8_t* scratch, int nk, int ne, int nd) {
  using P = SaberP<L>;
  size_t off = 0;
  auto take = [&](size_t per, int n) {
    uint8_t* q = scratch + off;
    off += ((size_t)n * per + 255) & ~(size_t)255;
    return q;
  };
  MixScratch x;
  x.araw_k = take(P::matrix_bytes, nk);
  x.sraw_k = take(P::secret_bytes, nk);
  x.araw_e = take(P::matrix_bytes, ne);
  x.sraw_e = take(P::secret_bytes, ne);
  x.mbuf_e = take(32, ne);
  x.krbuf_e = take(64, ne);
  x.araw_d = take(P::matrix_bytes, nd);
  x.sraw_d = take(P::secret_bytes, nd);
  x.mbuf_d = take(32, nd);
  x.krbuf_d = take(64, nd);
  x.kk_d = take(64, nd);
  return x;
}

template <int L>
static void mix_poly_any(const char* name, hipStream_t st, const SabPolyArgs& ak, const SabPolyArgs& ae,
                         const SabPolyArgs& ad, int ipw) {
  if (ipw == 4) mix_poly_launch<L, 4>(name, st, ak, ae, ad);
  else if (ipw == 2) mix_poly_launch<L, 2>(name, st, ak, ae, ad);
  else mix_poly_launch<L, 1>(name, st, ak, ae, ad);
}

// The products of a mixed sequence over the arrays of m.
template <int L>
static void mix_poly_args(const SabMixArgs& m, SabPolyArgs& ak, SabPolyArgs& ae, SabPolyArgs& ad) {
  using P = SaberP<L>;
  ak = SabPolyArgs{m.nk, m.araw_k, m.sraw_k, nullptr, 0, nullptr, m.pk_k, m.sk_k, nullptr, nullptr};
  ae = SabPolyArgs{m.ne, m.araw_e, m.sraw_e, m.pk_e, (size_t)P::pk, m.mbuf_e, m.ct_e, nullptr, nullptr, nullptr};
  ad = SabPolyArgs{m.nd,    m.araw_d, m.sraw_d, m.sk_d + P::vecq, (size_t)P::sk, m.mbuf_d, nullptr, nullptr, m.ct_d,
                   nullptr, m.kk_d,   m.ss_d,   0};
}

template <int L>
static void mix_products(hipStream_t st, const SabMixArgs& m, int ipw) {
  SabPolyArgs ak, ae, ad;
  mix_poly_args<L>(m, ak, ae, ad);
  mix_poly_any<L>(SAB_NAME("sab_mix_poly"), st, ak, ae, ad, ipw);
}

template <int L>
static void mix_l(hipStream_t st, const SaberMixDesc& d, uint8_t* scratch, int ipw) {
  const int nk = d.nk, ne = d.ne, nd = d.nd;
  const MixScratch x = mix_carve<L>(scratch, nk, ne, nd);
  SabMixArgs m{};
  m.nk = nk;
  m.ne = ne;
  m.nd = nd;
  m.gke = blocks_for(std::max(nk, ne), kMixGroup);
  m.kc = d.kc;
  m.pk_k = d.pk_k;
  m.sk_k = d.sk_k;
  m.araw_k = x.araw_k;
  m.sraw_k = x.sraw_k;
  m.pk_e = d.pk_e;
  m.ec = d.ec;
  m.ct_e = d.ct_e;
  m.ss_e = d.ss_e;
  m.araw_e = x.araw_e;
  m.sraw_e = x.sraw_e;
  m.mbuf_e = x.mbuf_e;
  m.krbuf_e = x.krbuf_e;
  m.ct_d = d.ct_d;
  m.sk_d = d.sk_d;
  m.ss_d = d.ss_d;
  m.araw_d = x.araw_d;
  m.sraw_d = x.sraw_d;
  m.mbuf_d = x.mbuf_d;
  m.krbuf_d = x.krbuf_d;
  m.kk_d = x.kk_d;
  const int gd = blocks_for(nd, kMixGroup);
  if (m.gke + gd == 0) return;
  DPKEM_LAUNCH(SAB_NAME("sab_mix_x0"), sab_mix_x0<L>, dim3(m.gke + gd), dim3(256), st, m);
  mix_products<L>(st, m, ipw);
  if (nk > 0 || ne > 0)
    DPKEM_LAUNCH(SAB_NAME("sab_mix_x1"), sab_mix_x1<L>, dim3(m.gke), dim3(256), st, m);
}

// One step of the rolling pipeline (see saber_roll_launch in launch.hpp):
// sab_roll_x0, then the mixed products.
template <int L>
static SabMixArgs roll_args(const RollDesc& d) {
  const int n = d.n;
  const MixScratch xk = mix_carve<L>(d.scratch_k, n, n, n);
  const MixScratch xe = mix_carve<L>(d.scratch_e, n, n, n);
  const MixScratch xd = mix_carve<L>(d.scratch_d, n, n, n);
  SabMixArgs m{};
  m.nk = m.ne = m.nd = n;
  m.gke = blocks_for(n, kMixGroup);
  m.kc = d.kc;
  m.pk_k = d.pk_k;
  m.sk_k = d.sk_k;
  m.araw_k = xk.araw_k;
  m.sraw_k = xk.sraw_k;
  m.pk_e = d.pk_e;
  m.ec = d.ec;
  m.ct_e = d.ct_e;
  m.araw_e = xe.araw_e;
  m.sraw_e = xe.sraw_e;
  m.mbuf_e = xe.mbuf_e;
  m.krbuf_e = xe.krbuf_e;
  m.ct_d = d.ct_d;
  m.sk_d = d.sk_d;
  m.ss_d = d.ss_d;
  m.araw_d = xd.araw_d;
  m.sraw_d = xd.sraw_d;
  m.mbuf_d = xd.mbuf_d;
  m.krbuf_d = xd.krbuf_d;
  m.kk_d = xd.kk_d;
  m.pk_t = d.pk_e;         // keygen of one step earlier: the keys encaps uses now
  m.sk_t = d.sk_e;
  m.ct_t = d.ct_d;         // encaps of two steps earlier: the ciphertexts decaps uses now
  m.krbuf_t = xd.krbuf_e;  // its kr, in that slot's encaps scratch
  m.ss_t = d.ss_e_prev;
  return m;
}

template <int L>
static void roll_l(hipStream_t st, const RollDesc& d, int ipw) {
  const SabMixArgs m = roll_args<L>(d);
  DPKEM_LAUNCH(SAB_NAME("sab_roll_x0"), sab_roll_x0<L>, dim3(2 * m.gke), dim3(256), st, m);
  mix_products<L>(st, m, ipw);
}

// hash: this step's chains (or none); prod: the step whose products run
// beside them (or none).
template <int L>
static void roll_fused_l(hipStream_t st, const RollDesc* hash, const RollDesc* prod) {
  const SabMixArgs m = hash ? roll_args<L>(*hash) : SabMixArgs{};
  const int hash_wgs = hash ? 2 * m.gke : 0;
  SabPolyArgs ak{}, ae{}, ad{};
  int bk = 0, be = 0, bd = 0;
  if (prod) {
    mix_poly_args<L>(roll_args<L>(*prod), ak, ae, ad);
    bk = blocks_for(ak.n, 4);
    be = blocks_for(ae.n, 4);
    bd = blocks_for(ad.n, 4);
  }
  if (hash_wgs + bk + be + bd == 0) return;
  DPKEM_LAUNCH(SAB_NAME("sab_roll_fused"), sab_roll_fused<L>, dim3(hash_wgs + bk + be + bd), dim3(256), st, m,
               hash_wgs, ak, ae, ad, bk, be);
}

void saber_roll_launch(hipStream_t st, int l, const RollDesc& d, int ipw) {
  if (d.n <= 0) return;
  if (l == 2) roll_l<2>(st, d, ipw);
  else if (l == 4) roll_l<4>(st, d, ipw);
  else roll_l<3>(st, d, ipw);
}

void saber_roll_fused_launch(hipStream_t st, int l, const RollDesc* hash, const RollDesc* prod) {
  if (hash && hash->n <= 0) hash = nullptr;
  if (prod && prod->n <= 0) prod = nullptr;
  if (l == 2) roll_fused_l<2>(st, hash, prod);
  else if (l == 4) roll_fused_l<4>(st, hash, prod);
  else roll_fused_l<3>(st, hash, prod);
}

void saber_mix_launch(hipStream_t st, int l, const SaberMixDesc& d, uint8_t* scratch, int ipw) {
  if (d.nk < 0 || d.ne < 0 || d.nd < 0) return;
  if (l == 2) mix_l<2>(st, d, scratch, ipw);
  else if (l == 4) mix_l<4>(st, d, scratch, ipw);
  else mix_l<3>(st, d, scratch, ipw);
}

// Chained keygen -> encaps -> decaps of the same n Saber (l = 3) instances on
// one stream.  Encaps reads seedA from the pk keygen just wrote and decaps
// from the copy in sk, so both GenMatrix streams equal keygen's: the chain
// expands A once (23 of 98 serial Keccak permutations per instance saved
// twice) and the encaps / decaps hash kernels run only their message roles.
// scratch holds kSaberChainScratchPerInst bytes per instance; the buffers are
// reused from op to op, which the stream order makes safe.
void saber_chain_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                        uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int ipw) {
  constexpr int L = 3;
  using P = SaberP<L>;
  if (n <= 0) return;
  size_t off = 0;
  auto take = [&](size_t per) {
    uint8_t* q = scratch + off;
    off += ((size_t)n * per + 255) & ~(size_t)255;
    return q;
  };
  uint8_t* araw = take(P::matrix_bytes);
  uint8_t* sraw = take(P::secret_bytes);
  uint8_t* mbuf = take(32);
  uint8_t* krbuf = take(64);
  uint32_t* fail = reinterpret_cast<uint32_t*>(take(4));
  const dim3 g(blocks_for(n, 32)), b64(64), b128(128);
  // keygen
  DPKEM_LAUNCH("sab_keygen_hash1", sab_keygen_hash1<L>, g, b128, st, -1, n, kc, pk, sk, araw, sraw);
  SabPolyArgs ak{n, araw, sraw, nullptr, 0, nullptr, pk, sk, nullptr, nullptr};
  poly_launch<kSabKeygen, L>("sab_poly<kSabKeygen>", st, n, ak, ipw);
  DPKEM_LAUNCH("sab_keygen_hash2", sab_keygen_hash2<L>, g, b64, st, n, pk, sk);
  // encaps: message role only, A from keygen
  DPKEM_LAUNCH("sab_enc_hash1", sab_enc_hash1<L>, g, b64, st, 1, n, pk, ec, araw, sraw, mbuf, krbuf);
  SabPolyArgs ae{n, araw, sraw, pk, (size_t)P::pk, mbuf, ct, nullptr, nullptr, nullptr};
  poly_launch<kSabEnc, L>("sab_poly<kSabEnc>", st, n, ae, ipw);
  DPKEM_LAUNCH("sab_hash2", sab_hash2<L>, g, b64, st, n, ct, krbuf, nullptr, nullptr, ss);
  // decaps: message role only, A from keygen
  DPKEM_LAUNCH("sab_dec_inner", sab_dec_inner<L>, dim3(blocks_for(n, 2)), b64, st, n, ct, sk, mbuf);
  DPKEM_LAUNCH("sab_dec_hash1", sab_dec_hash1<L>, g, b64, st, 1, n, sk, mbuf, araw, sraw, krbuf);
  SabPolyArgs ad{n, araw, sraw, sk + P::vecq, (size_t)P::sk, mbuf, nullptr, nullptr, ct, fail};
  poly_launch<kSabReenc, L>("sab_poly<kSabReenc>", st, n, ad, ipw);
  DPKEM_LAUNCH("sab_hash2", sab_hash2<L>, g, b64, st, n, ct, krbuf, sk, fail, ss2);
}

}  // namespace dpkem
