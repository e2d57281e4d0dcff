// FrodoKEM-976-SHAKE batched kernels for gfx950.
//
// Reference path replaced (benlwk/DPCrypto FrodoKEM976-GPU/):
//   shake256_gpu (25-thread, shared memory)        src/shake.cu:495-604
//   shake128_ser  A-row generation into HBM        src/shake.cu:619-645
//   sample_n_gpu{,_small}                          src/sample_pack.cu:13-46
//   as_plus_e_gpu / DoDP4Av5s / sa_plus_e_gpu      src/matrix_kernel.cu:13-241
//   mul_add_sb_plus_e_gpu / mul_bs_gpu             src/matrix_kernel.cu:245-278
//   pack / unpack / key_encode / key_decode        src/sample_pack.cu:55-198, src/encode_decode.cu
//   copy_* / add / sub / ct_verify / ct_select     src/util.cu
//   host drivers                                   src/cuda_kernel.cu:21-335
//
// Per op the engine runs
//   pre    SHAKE256 chains (pkh, G2, noise stream) on lane pairs (two lanes
//          per state, keccak2.hpp) with the CDF sampler fused into the
//          squeeze; secrets and errors leave as int8;
//   fused  one 1024-thread workgroup per instance; lane i squeezes row i of A
//          (SHAKE128(LE16(i) || seedA)) in VGPRs and feeds it straight into
//          v_mfma_i32_16x16x64_i8 as two byte limbs, so A never reaches memory:
//            keygen  B  = A S + E       (reduction along the lane's own row)
//            enc/dec B' = S' A + E'     (reduction across lanes: the row data is
//                                        transposed by selector MFMAs first)
//          followed by the epilogue (pack, V = S' B + E'', encode, compare);
//   post   lane-pair SHAKE256 (pkh for keygen, ss for enc/dec).
//
// Limb split: a = 256 * hi + lo with hi = (int8)(a >> 8) and
// lo = (int8)((a & 0xFF) ^ 0x80), so a == 256 hi + lo + 128 (mod 2^16);
// the 128 * sum(S) term is added once per output.  With |S| <= 10 and 976
// terms every i32 partial sum stays below 2^21 in magnitude.
#include <string.h>

#include <algorithm>

#include "common.hpp"
#include "keccak.hpp"
#include "frodo_sampler.hpp"
#include "sponge.hpp"
#include "keccak2.hpp"
#include "prof.hpp"
#include "launch.hpp"

namespace dpkem {

typedef int v4i __attribute__((ext_vector_type(4)));

// ----------------------------------------------------------------- parameter sets
//
// FrodoKEM-976 is the reference's set (FrodoKEM976-GPU/include/params.h:9-24);
// Frodo-640 and Frodo-1344 are the other two round-3 SHAKE sets.  Each differs
// in n, log2 q (D), message bits per entry (B), secret length, the SHAKE of
// the serial chains and the CDF table; A rows are SHAKE128 in all three.
// threads x passes covers the n rows of A in the fused kernels (one row per
// lane and pass); the shapes keep 14-16 waves per CU resident.
template <int N>
struct FrodoSet;
template <>
struct FrodoSet<640> {
  static constexpr int n = 640, logq = 15, b = 2, sec = 16, chain_rate = kRateShake128, threads = 320, passes = 2;
};
template <>
struct FrodoSet<976> {
  static constexpr int n = 976, logq = 16, b = 3, sec = 24, chain_rate = kRateShake256, threads = 512, passes = 2;
};
template <>
struct FrodoSet<1344> {
  static constexpr int n = 1344, logq = 16, b = 4, sec = 32, chain_rate = kRateShake256, threads = 448, passes = 3;
};

// Sampler magnitude tables (frodo_sampler.hpp), staged into LDS by the chains.
__device__ const SampleLut<640> kSampleLut640 = make_sample_lut<640>();
__device__ const SampleLut<976> kSampleLut976 = make_sample_lut<976>();
__device__ const SampleLut<1344> kSampleLut1344 = make_sample_lut<1344>();
template <int N>
__device__ __forceinline__ const uint8_t* sample_lut_global() {
  return N == 640 ? kSampleLut640.e : N == 976 ? kSampleLut976.e : kSampleLut1344.e;
}

template <int N>
struct FrodoP : FrodoSet<N> {
  using S = FrodoSet<N>;
  static constexpr int mat = S::n * kFrodoNbar;                  // entries of S / E / S' / E' / B / B'
  static constexpr int pk = kFrodoSeedA + S::logq * S::n;        // seedA || pack_D(B)
  static constexpr int sk = S::sec + pk + 2 * mat + S::sec;      // s || pk || S^T (LE u16) || pkh
  static constexpr int ct = S::logq * (S::n + kFrodoNbar);       // pack_D(B') || pack_D(C)
  static constexpr int kcoins = 2 * S::sec + 16;                 // s || seedSE || z
  static constexpr int ecoins = S::sec;                          // mu
  static constexpr int ls = S::sec / 8;                          // secret length in 64-bit words
  static constexpr int ct_c = S::logq * S::n;                    // offset of C in ct
  static constexpr int sk_s = S::sec + pk;                       // offset of S^T in sk
  static constexpr int sk_pkh = sk - S::sec;                     // offset of pkh in sk
  static constexpr int units = S::n / 16;                        // MFMA K-units of 16 columns per row
  static constexpr int row_blocks = (S::n + 83) / 84;            // SHAKE128 blocks per row of A
  static constexpr int srow = 16 * units + 32;                   // LDS row stride of int8 S^T (keygen)
  static constexpr int sprow = S::threads * S::passes;           // LDS row stride of int8 S' (enc / dec)
  static constexpr uint32_t qmask = (1u << S::logq) - 1;
  static_assert(S::n % 16 == 0 && row_blocks % 4 == 0, "row generation works in groups of four blocks");
  static_assert(S::threads % 64 == 0 && sprow >= S::n && sprow - S::threads < S::n, "every pass has rows");
  static_assert(pk % 8 == 0 && sk % 8 == 0 && ct % 8 == 0 && ct_c % 4 == 0 && kcoins % 8 == 0,
                "rows and fields stay 8-byte aligned");
  static_assert(S::b * 64 == 8 * S::sec, "mu fills the 8 x 8 message matrix");
};
// Launch shape of the fused kernels: TH lanes per workgroup cover the n rows
// of A in `passes` passes of one row per lane.  FrodoKEM-976's default is
// 512 x 2: two workgroups per CU, so that one's prologue and epilogue (S'
// staging, B' / V / C) overlap the other's row generation (measured 8-9%
// faster than 1024 x 1 at 8192 instances); 1024 x 1 (one row per lane, one
// workgroup per CU) is the alternative shape the tuning knob selects.
template <class P, int TH>
struct FrodoShape {
  static constexpr int passes = (P::n + TH - 1) / TH;
  static constexpr int sprow = TH * passes;  // LDS row stride of int8 S' (enc / dec)
  static_assert(TH % 64 == 0 && sprow >= P::n && sprow - TH < P::n, "every pass has rows");
};
constexpr int kFrodoAltThreads = 1024;  // the FrodoKEM-976 alternative shape

using F976 = FrodoP<976>;
static_assert(F976::pk == kFrodoPk && F976::sk == kFrodoSk && F976::ct == kFrodoCt &&
              F976::kcoins == kFrodoKeygenCoins, "FrodoKEM976-GPU/include/api.h:7-10");

// Task tables.  The fused kernels and the decaps B'S kernel take their
// work as a table of tasks (FrMt, a kernel argument): per task its instance
// count, arrays and scratch; a launch over one array is a table of one task.
// Several independent tasks of one op and parameter set -- uneven batch
// sizes, each with its own arrays -- can so share one launch sequence
// (frodo_mt_launch).  The fused kernels are persistent: a grid of at most
// the resident workgroups walks the concatenated work list (work item = task
// instance) with a fixed stride.
struct FrMtTask {
  int n, first, first_group;
  size_t key_stride;  // bytes between the key rows of i1 (decaps' sk; 0 = one key for all)
  const uint8_t* i0;
  const uint8_t* i1;
  uint8_t* o0;
  uint8_t* o1;
  uint32_t *s8, *e8, *epp8, *fail, *state;
  uint8_t *kbuf, *mubuf;
};
struct FrMt {
  int ntasks, total, groups;
  FrMtTask t[kFrodoMtTasks];
};

// threadIdx.x behind an empty asm statement, read at the top of each work
// item: values derived from it (selectors, operand offsets) are then
// computed per item instead of being hoisted out of the item loop and held
// in registers across the whole row generation.
__device__ __forceinline__ int fresh_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ int mt_task_of_group(const FrMt& M, int g) {
  int t = 0;
  while (t + 1 < M.ntasks && g >= M.t[t + 1].first_group) ++t;
  return t;
}
__device__ __forceinline__ int mt_task_of_item(const FrMt& M, int i) {
  int t = 0;
  while (t + 1 < M.ntasks && i >= M.t[t + 1].first) ++t;
  return t;
}


// FrodoKEM-976 names kept for the AES rows and the shared-key kernels
constexpr int kFN = kFrodoN, kFNB = kFrodoNbar;
constexpr int kFMat = F976::mat;          // 7808
constexpr int kFCtC1 = F976::ct_c;        // 15616: offset of C in ct
constexpr int kFUnits = F976::units;      // 61

// Entry idx of a D-bit, most-significant-bit-first packed stream (D = 16:
// big-endian u16).  Reads only the bytes the entry occupies.
template <int D>
__device__ __forceinline__ uint32_t unpack_at(const uint8_t* __restrict__ base, int idx) {
  if constexpr (D == 16) {
    return ((uint32_t)base[2 * idx] << 8) | base[2 * idx + 1];
  } else {
    static_assert(D == 15, "D = 15 or 16");
    const int bit = 15 * idx, byte = bit >> 3, sh = bit & 7;
    uint32_t v = ((uint32_t)base[byte] << 16) | ((uint32_t)base[byte + 1] << 8);
    if (sh > 1) v |= base[byte + 2];
    return (v >> (9 - sh)) & 0x7FFFu;
  }
}

// Output word w (bytes 4w .. 4w + 3, as stored little-endian) of the 15-bit
// packed stream of src[0 .. count): the stream's bits 32 w .. 32 w + 31 span
// at most four entries.
template <class W>
__device__ __forceinline__ uint32_t pack15_word(const W* src, int count, int w) {
  const int bit = 32 * w, e = bit / 15, off = bit % 15;
  uint64_t acc = 0;
#pragma unroll
  for (int t = 0; t < 4; ++t) acc = (acc << 15) | (e + t < count ? ((uint64_t)src[e + t] & 0x7FFF) : 0);
  return __builtin_bswap32((uint32_t)(acc >> (28 - off)));
}

// ----------------------------------------------------------------- chains (lane pairs)
//
// The serial chains are latency-bound single waves; in a multi-stream plan
// they run beside other chunks' fused kernels, whose VALU-saturated waves
// may share their SIMD.  Chain waves raise their issue priority so that the
// SIMD issues for them first and the fused waves fill the gaps.
__device__ __forceinline__ void chain_priority() { __builtin_amdgcn_s_setprio(3); }
//
// Serial SHAKE chains (SHAKE256 for Frodo-976 / -1344, SHAKE128 for
// Frodo-640) run on lane pairs (keccak2.hpp): lanes 2k / 2k+1 hold the even
// / odd bit halves of instance k's state, 32 instances per 64-thread block.
// Squeezed noise words are rebuilt as 32-bit halves, so each lane samples two
// of the four entries of every word.

using k2::absorb2_global;
using k2::absorb2_halves;
using k2::absorb2_regs;
using k2::store2_words;

// The chain bodies take the 32-instance group vb explicitly, so that the
// pipelined chain kernels can run several of them in one launch.
#define FR_PAIR_SETUP                                   \
  const int lane = threadIdx.x & 63;                    \
  const uint32_t p = lane & 1;                          \
  const int inst = vb * 32 + (lane >> 1);               \
  if (inst >= n) return;

// Stage the sampler table in `lds` (32 KB) with the 64 lanes of one wave and
// order the copy before the wave's table reads (a wave's LDS operations
// complete in issue order, so a wavefront-scope fence suffices).
template <class P>
__device__ __forceinline__ const uint8_t* stage_sample_lut(uint8_t* lds, int lane) {
  const uint4* src = reinterpret_cast<const uint4*>(sample_lut_global<P::n>());
  uint4* dst = reinterpret_cast<uint4*>(lds);
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {  // 2048 16-byte rows: 32 per lane in four rounds
    uint4 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = src[(c * 8 + i) * 64 + lane];
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[(c * 8 + i) * 64 + lane] = v[i];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return lds;
}

// Two samples from the 32-bit half x of a squeezed word (its two 16-bit
// draws) given their table magnitudes e (low / high 16 bits): the int8 pair
// (returned, low 16 bits) and the little-endian u16 pair (le).  The sign is
// applied per 16-bit half.
typedef unsigned short fr_u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t sample2(uint32_t x, uint32_t e, uint32_t* le) {
  const fr_u16x2 sm = __builtin_bit_cast(fr_u16x2, x & 0x00010001u);
  const fr_u16x2 ev = __builtin_bit_cast(fr_u16x2, e);
  const uint32_t v = __builtin_bit_cast(uint32_t, (ev ^ ((fr_u16x2){0, 0} - sm)) + sm);
  *le = v;
  return __builtin_amdgcn_perm(v, v, 0x0C0C0200u);  // bytes 0 and 2: the int8 pair
}

// The CNT words of one squeezed block: this lane's halves, then all table
// reads of the block (so their LDS latency is paid once per block), then
// f(w, half, magnitudes) per word.
template <int CNT, class F>
__device__ __forceinline__ void sample_block(const uint32_t a[25], uint32_t p, const uint8_t* lut, int w0, F&& f) {
  uint32_t x[CNT], e[CNT];
#pragma unroll
  for (int i = 0; i < CNT; ++i) x[i] = k2::out_half(a[i], p);
#pragma unroll
  for (int i = 0; i < CNT; ++i) e[i] = (uint32_t)lut[(x[i] >> 1) & 0x7FFF] | ((uint32_t)lut[x[i] >> 17] << 16);
#pragma unroll
  for (int i = 0; i < CNT; ++i) f(w0 + i, x[i], e[i]);
}

// SHAKE(dom || seed[sec bytes]) on a lane pair, squeezing OUTW words; f(w,
// half, e) receives this lane's 32-bit half (low half for the even lane, high
// for the odd lane) of every output word w and the sampler table's
// magnitudes of its two 16-bit draws (sample_block).
template <class P, int OUTW, class F>
__device__ __forceinline__ void shake_seed_squeeze2(uint32_t p, uint64_t dom, const uint64_t seed[P::ls],
                                                    const uint8_t* lut, F&& f) {
  constexpr int RW = P::chain_rate, LS = P::ls;
  uint64_t m[LS + 1];
  m[0] = dom | (seed[0] << 8);
#pragma unroll
  for (int i = 1; i < LS; ++i) m[i] = (seed[i - 1] >> 56) | (seed[i] << 8);
  m[LS] = (seed[LS - 1] >> 56) | (kDsShake << 8);
  uint32_t a[25];
  k2::zero2(a);
#pragma unroll
  for (int i = 0; i <= LS; ++i) a[i] = k2::deint(m[i], p);
  a[RW - 1] ^= k2::half_const(kPadEnd, p);
  k2::keccakf2(a, p);
  constexpr int NB = OUTW / RW, REM = OUTW % RW;
  for (int b = 0; b < NB; ++b) {
    sample_block<RW>(a, p, lut, b * RW, f);
    k2::keccakf2_unrolled(a, p);
  }
  if constexpr (REM > 0) sample_block<REM>(a, p, lut, NB * RW, f);
}

// The same squeeze split over two waves of a workgroup: the chain wave only
// permutes and publishes each squeezed block (its lanes' halves) into one of
// two LDS slots; a sampler wave on another SIMD rebuilds each lane's 32-bit
// half from the pair's two halves, reads the table and runs f.  `ready`
// counts published blocks, `done` the blocks the sampler has read (a slot is
// refilled only once the block two before it is done).  Both waits are the
// bounded LDS waits of keccak2.hpp.
template <class P>
struct FrRing {
  uint32_t w[2][P::chain_rate][64];  // [slot][word][lane]
  uint32_t ready, done;
};

template <class P, int CNT>
__device__ __forceinline__ void ring_publish(FrRing<P>* ring, const uint32_t a[25], int b, int lane) {
  if (b >= 2) k2::lds_wait_count(&ring->done, (uint32_t)(b - 1));
#pragma unroll
  for (int i = 0; i < CNT; ++i) ring->w[b & 1][i][lane] = a[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (lane == 0) __hip_atomic_fetch_add(&ring->ready, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <class P, int OUTW>
__device__ __forceinline__ void shake_seed_squeeze2_ring(uint32_t p, uint64_t dom, const uint64_t seed[P::ls],
                                                         FrRing<P>* ring, int lane) {
  constexpr int RW = P::chain_rate, LS = P::ls;
  uint64_t m[LS + 1];
  m[0] = dom | (seed[0] << 8);
#pragma unroll
  for (int i = 1; i < LS; ++i) m[i] = (seed[i - 1] >> 56) | (seed[i] << 8);
  m[LS] = (seed[LS - 1] >> 56) | (kDsShake << 8);
  uint32_t a[25];
  k2::zero2(a);
#pragma unroll
  for (int i = 0; i <= LS; ++i) a[i] = k2::deint(m[i], p);
  a[RW - 1] ^= k2::half_const(kPadEnd, p);
  k2::keccakf2(a, p);
  constexpr int NB = OUTW / RW, REM = OUTW % RW;
  static_assert(REM > 0, "the last block is a partial one");
  for (int b = 0; b < NB; ++b) {
    ring_publish<P, RW>(ring, a, b, lane);
    k2::keccakf2_unrolled(a, p);
  }
  ring_publish<P, REM>(ring, a, NB, lane);
}

// The same with the domain byte and the length (outw words, not a multiple
// of the rate) known at run time, so that chain waves squeezing streams of
// different lengths share one call site: one copy of the unrolled
// permutation in the instruction cache.  The final, partial block is
// published whole; its sampler reads only the first outw % rate words.
template <class P>
__device__ __forceinline__ void shake_seed_squeeze2_ring_rt(uint32_t p, uint64_t dom, const uint64_t seed[P::ls],
                                                            FrRing<P>* ring, int lane, int outw) {
  constexpr int RW = P::chain_rate, LS = P::ls;
  uint64_t m[LS + 1];
  m[0] = dom | (seed[0] << 8);
#pragma unroll
  for (int i = 1; i < LS; ++i) m[i] = (seed[i - 1] >> 56) | (seed[i] << 8);
  m[LS] = (seed[LS - 1] >> 56) | (kDsShake << 8);
  uint32_t a[25];
  k2::zero2(a);
#pragma unroll
  for (int i = 0; i <= LS; ++i) a[i] = k2::deint(m[i], p);
  a[RW - 1] ^= k2::half_const(kPadEnd, p);
  k2::keccakf2(a, p);
  const int nb = outw / RW;
  for (int b = 0; b < nb; ++b) {
    ring_publish<P, RW>(ring, a, b, lane);
    k2::keccakf2_unrolled(a, p);
  }
  ring_publish<P, RW>(ring, a, nb, lane);
}

template <class P, int CNT, class F>
__device__ __forceinline__ void ring_sample_block(FrRing<P>* ring, uint32_t p, const uint8_t* lut, int b, int lane,
                                                  F&& f) {
  k2::lds_wait_count(&ring->ready, (uint32_t)(b + 1));
  uint32_t x[CNT], e[CNT];
#pragma unroll
  for (int i = 0; i < CNT; ++i) {
    const uint32_t mine = ring->w[b & 1][i][lane], other = ring->w[b & 1][i][lane ^ 1];
    x[i] = p ? k2::reint_half(other, mine, 1) : k2::reint_half(mine, other, 0);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the slot's reads are done: the chain may refill it
  if (lane == 0) __hip_atomic_fetch_add(&ring->done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
  for (int i = 0; i < CNT; ++i) e[i] = (uint32_t)lut[(x[i] >> 1) & 0x7FFF] | ((uint32_t)lut[x[i] >> 17] << 16);
  const int w0 = b * P::chain_rate;
#pragma unroll
  for (int i = 0; i < CNT; ++i) f(w0 + i, x[i], e[i]);
}

template <class P, int OUTW, class F>
__device__ __forceinline__ void sample_ring(uint32_t p, FrRing<P>* ring, int lane, const uint8_t* lut, F&& f) {
  constexpr int RW = P::chain_rate, NB = OUTW / RW, REM = OUTW % RW;
  for (int b = 0; b < NB; ++b) ring_sample_block<P, RW>(ring, p, lut, b, lane, f);
  ring_sample_block<P, REM>(ring, p, lut, NB, lane, f);
}

// Roles of a chain body: the whole chain on one wave, or split into the
// chain wave and its sampler wave (FrRing).
enum FrRole { kFrWhole = 0, kFrChain = 1, kFrSampler = 2 };

// ----------------------------------------------------------------- keygen chains

// seedA = SHAKE(z, 16); S^T || E = sample(SHAKE(0x5F || seedSE, 32 n)).
// Writes pk[0:16], sk[0:sec] = s, sk[sec:sec+16] = seedA, the sk S^T part (LE
// u16) and the int8 copies of S^T ([k][j]) and E ([i][k]) for the fused kernel.
// OWN_LUT = false: the sampler reads the table at `lut`, staged by the caller.
template <class P, int ROLE = kFrWhole, bool OWN_LUT = true>
__device__ __forceinline__ void keygen_pre_body(int vb, int n, const uint8_t* __restrict__ coins,
                                                uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                uint32_t* __restrict__ s8, uint32_t* __restrict__ e8,
                                                FrRing<P>* ring = nullptr, const uint8_t* lut = nullptr) {
  if constexpr (ROLE != kFrChain && OWN_LUT) {
    __shared__ __align__(16) uint8_t lut_lds[32768];
    lut = stage_sample_lut<P>(lut_lds, threadIdx.x & 63);
  }
  FR_PAIR_SETUP
  constexpr int RW = P::chain_rate, LS = P::ls, OUTW = 2 * P::mat / 4;
  const uint8_t* c = coins + (size_t)inst * P::kcoins;
  uint8_t* pp = pk + (size_t)inst * P::pk;
  uint8_t* q = sk + (size_t)inst * P::sk;
  if constexpr (ROLE != kFrChain) {  // seedA (off the noise chain when a sampler wave runs it)
    const uint64_t z[2] = {ld64(c + 2 * P::sec), ld64(c + 2 * P::sec + 8)};
    uint32_t a[25];
    absorb2_regs<RW, 2>(a, p, z, kDsShake);
    store2_words<2>(a, p, pp);
    store2_words<2>(a, p, q + P::sec);
    const uint32_t* sc = reinterpret_cast<const uint32_t*>(c);
#pragma unroll
    for (int i = 0; i < LS; ++i) reinterpret_cast<uint32_t*>(q)[2 * i + p] = sc[2 * i + p];
  }
  uint8_t* sd = reinterpret_cast<uint8_t*>(s8 + (size_t)inst * (P::mat / 4)) + 2 * p;
  uint8_t* ed = reinterpret_cast<uint8_t*>(e8 + (size_t)inst * (P::mat / 4)) + 2 * p;
  uint8_t* skS = q + P::sk_s + 4 * p;
  auto f = [&](int w, uint32_t x, uint32_t e) {
    uint32_t le;
    const uint32_t b = sample2(x, e, &le);
    if (w < P::mat / 4) {
      *reinterpret_cast<uint16_t*>(sd + 4 * w) = (uint16_t)b;
      *reinterpret_cast<uint32_t*>(skS + 8 * w) = le;
    } else {
      *reinterpret_cast<uint16_t*>(ed + 4 * (w - P::mat / 4)) = (uint16_t)b;
    }
  };
  if constexpr (ROLE == kFrSampler) {
    sample_ring<P, OUTW>(p, ring, lane, lut, f);
  } else {
    uint64_t seed[LS];
#pragma unroll
    for (int i = 0; i < LS; ++i) seed[i] = ld64(c + P::sec + 8 * i);
    if constexpr (ROLE == kFrChain) shake_seed_squeeze2_ring<P, OUTW>(p, 0x5F, seed, ring, lane);
    else shake_seed_squeeze2<P, OUTW>(p, 0x5F, seed, lut, f);
  }
}

template <class P>
__device__ __forceinline__ void ring_init(FrRing<P>* ring) {
  if (threadIdx.x == 0) {
    ring->ready = 0;
    ring->done = 0;
  }
  __syncthreads();
}

// One 128-thread block per 32 instances: wave 0 runs the noise chain, wave 1
// seedA and the sampling (FrRing).
template <class P>
__device__ __forceinline__ void keygen_pre_split(int vb, int n, const uint8_t* __restrict__ coins,
                                                 uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                 uint32_t* __restrict__ s8, uint32_t* __restrict__ e8) {
  __shared__ FrRing<P> ring;
  ring_init(&ring);
  chain_priority();  // the sampler too: the chain waits on it for ring slots
  if ((threadIdx.x >> 6) == 0) {
    keygen_pre_body<P, kFrChain>(vb, n, coins, pk, sk, s8, e8, &ring);
  } else {
    keygen_pre_body<P, kFrSampler>(vb, n, coins, pk, sk, s8, e8, &ring);
  }
}

template <class P>
__global__ void __launch_bounds__(128) frodo_keygen_pre(int n, const uint8_t* __restrict__ coins,
                                                        uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                        uint32_t* __restrict__ s8, uint32_t* __restrict__ e8) {
  keygen_pre_split<P>(blockIdx.x, n, coins, pk, sk, s8, e8);
}

// pkh = SHAKE(pk, sec); sk = s || pk || S^T || pkh (pk already copied by the fused kernel).
template <class P>
__device__ __forceinline__ void keygen_post_body(int vb, int n, const uint8_t* __restrict__ pk, uint8_t* __restrict__ sk) {
  FR_PAIR_SETUP
  uint32_t a[25];
  absorb2_global<P::chain_rate, P::pk / 8>(a, p, pk + (size_t)inst * P::pk, kDsShake);
  store2_words<P::ls>(a, p, sk + (size_t)inst * P::sk + P::sk_pkh);
}

template <class P>
__global__ void __launch_bounds__(64) frodo_keygen_post(int n, const uint8_t* __restrict__ pk, uint8_t* __restrict__ sk) {
  chain_priority();
  keygen_post_body<P>(blockIdx.x, n, pk, sk);
}

// ----------------------------------------------------------------- enc / dec chains

// enc: pkh = SHAKE(pk, sec), or read from `pkh` rows when the caller holds it
// (a key store keeps each key's pkh); mu = coins.   dec: pkh from sk;
// mu = recovered mu'.
// G2 = SHAKE(pkh || mu, 2 sec) = seedSE || k; S' || E' || E'' from
// SHAKE(0x96 || seedSE, 32 n + 128).  k goes to kbuf[0:sec].
template <class P, bool DEC>
__device__ __forceinline__ void encdec_seed(int inst, uint32_t p, const uint8_t* __restrict__ keys, size_t key_stride,
                                            const uint8_t* __restrict__ mu, size_t mu_stride,
                                            uint8_t* __restrict__ kbuf, const uint8_t* __restrict__ pkh,
                                            uint64_t seed[P::ls]) {
  constexpr int RW = P::chain_rate, LS = P::ls;
  const uint8_t* key = keys + (size_t)inst * key_stride;
  uint32_t a[25], g[2 * LS];
  if (DEC || pkh) {
    const uint8_t* h = DEC ? key + P::sk_pkh : pkh + (size_t)inst * P::sec;
#pragma unroll
    for (int i = 0; i < LS; ++i) g[i] = k2::deint(ld64(h + 8 * i), p);
  } else {
    absorb2_global<RW, P::pk / 8>(a, p, key, kDsShake);
#pragma unroll
    for (int i = 0; i < LS; ++i) g[i] = a[i];
  }
  const uint8_t* m = mu + (size_t)inst * mu_stride;
#pragma unroll
  for (int i = 0; i < LS; ++i) g[LS + i] = k2::deint(ld64(m + 8 * i), p);
  absorb2_halves<RW, 2 * LS>(a, p, g, kDsShake);
  store2_words<LS>(a + LS, p, kbuf + (size_t)inst * 64);
#pragma unroll
  for (int i = 0; i < LS; ++i) seed[i] = k2::out_full(a[i], p);
}

// The chain (G2, then the noise squeeze) and / or the sampling; OWN_LUT as
// in keygen_pre_body.
template <class P, bool DEC, int ROLE = kFrWhole, bool OWN_LUT = true>
__device__ __forceinline__ void encdec_pre_body(int vb, int n, const uint8_t* __restrict__ keys, size_t key_stride,
                                                const uint8_t* __restrict__ mu, size_t mu_stride,
                                                uint8_t* __restrict__ kbuf, uint32_t* __restrict__ s8,
                                                uint32_t* __restrict__ e8, uint32_t* __restrict__ epp8,
                                                const uint8_t* __restrict__ pkh, FrRing<P>* ring = nullptr,
                                                const uint8_t* lut = nullptr) {
  if constexpr (ROLE != kFrChain && OWN_LUT) {
    __shared__ __align__(16) uint8_t lut_lds[32768];  // staged by the one wave of a workgroup that samples
    lut = stage_sample_lut<P>(lut_lds, threadIdx.x & 63);
  }
  FR_PAIR_SETUP
  constexpr int LS = P::ls, OUTW = (2 * P::mat + 64) / 4;
  uint8_t* sd = reinterpret_cast<uint8_t*>(s8 + (size_t)inst * (P::mat / 4)) + 2 * p;
  uint8_t* ed = reinterpret_cast<uint8_t*>(e8 + (size_t)inst * (P::mat / 4)) + 2 * p;
  uint8_t* pd = reinterpret_cast<uint8_t*>(epp8 + (size_t)inst * 16) + 2 * p;
  auto f = [&](int w, uint32_t x, uint32_t e) {
    uint32_t le;
    const uint16_t b = (uint16_t)sample2(x, e, &le);
    if (w < P::mat / 4) *reinterpret_cast<uint16_t*>(sd + 4 * w) = b;
    else if (w < 2 * P::mat / 4) *reinterpret_cast<uint16_t*>(ed + 4 * (w - P::mat / 4)) = b;
    else *reinterpret_cast<uint16_t*>(pd + 4 * (w - 2 * P::mat / 4)) = b;
  };
  if constexpr (ROLE == kFrSampler) {
    sample_ring<P, OUTW>(p, ring, lane, lut, f);
  } else {
    uint64_t seed[LS];
    encdec_seed<P, DEC>(inst, p, keys, key_stride, mu, mu_stride, kbuf, pkh, seed);
    if constexpr (ROLE == kFrChain) shake_seed_squeeze2_ring<P, OUTW>(p, 0x96, seed, ring, lane);
    else shake_seed_squeeze2<P, OUTW>(p, 0x96, seed, lut, f);
  }
}

// One 128-thread block per 32 instances: wave 0 runs G2 and the noise chain,
// wave 1 the sampling (FrRing).
template <class P, bool DEC>
__device__ __forceinline__ void encdec_pre_split(int vb, int n, const uint8_t* __restrict__ keys, size_t key_stride,
                                                 const uint8_t* __restrict__ mu, size_t mu_stride,
                                                 uint8_t* __restrict__ kbuf, uint32_t* __restrict__ s8,
                                                 uint32_t* __restrict__ e8, uint32_t* __restrict__ epp8,
                                                 const uint8_t* __restrict__ pkh) {
  __shared__ FrRing<P> ring;
  ring_init(&ring);
  chain_priority();  // the sampler too: the chain waits on it for ring slots
  if ((threadIdx.x >> 6) == 0) {
    encdec_pre_body<P, DEC, kFrChain>(vb, n, keys, key_stride, mu, mu_stride, kbuf, s8, e8, epp8, pkh, &ring);
  } else {
    encdec_pre_body<P, DEC, kFrSampler>(vb, n, keys, key_stride, mu, mu_stride, kbuf, s8, e8, epp8, pkh, &ring);
  }
}

template <class P, bool DEC>
__global__ void __launch_bounds__(128) frodo_encdec_pre(int n, const uint8_t* __restrict__ keys, size_t key_stride,
                                                        const uint8_t* __restrict__ mu, size_t mu_stride,
                                                        uint8_t* __restrict__ kbuf, uint32_t* __restrict__ s8,
                                                        uint32_t* __restrict__ e8, uint32_t* __restrict__ epp8,
                                                        const uint8_t* __restrict__ pkh) {
  encdec_pre_split<P, DEC>(blockIdx.x, n, keys, key_stride, mu, mu_stride, kbuf, s8, e8, epp8, pkh);
}

// ss = SHAKE(ct || kbar, sec); for decaps kbar = fail ? s : k'.  The full
// blocks hold only ct words (ss_absorb); the last block holds the remaining
// ct words, kbar and the padding (ss_finish).  The pipelined chain absorbs
// decaps' ct while the re-encryption runs and keeps the state per lane in
// `state` (25 words per lane, word-major: state[i * 2n + 2 inst + p]).
template <class P>
struct FrodoSsShape {
  static constexpr int RW = P::chain_rate, LS = P::ls;
  static constexpr int CTW = P::ct / 8;
  static constexpr int NB = CTW / RW;        // full blocks of ct words
  static constexpr int REM = CTW - NB * RW;  // ct words in the last block
  static_assert(REM + LS <= RW - 1, "kbar and the padding share the final block");
};

template <class P>
__device__ __forceinline__ void ss_absorb(uint32_t a[25], uint32_t p, const uint8_t* __restrict__ c) {
  using Q = FrodoSsShape<P>;
  k2::zero2(a);
  for (int b = 0; b < Q::NB; ++b) {
#pragma unroll
    for (int i = 0; i < Q::RW; ++i) a[i] ^= k2::deint(ld64(c + 8 * (b * Q::RW + i)), p);
    k2::keccakf2(a, p);
  }
}

template <class P>
__device__ __forceinline__ void ss_finish(uint32_t a[25], uint32_t p, int inst, const uint8_t* __restrict__ c,
                                          const uint8_t* __restrict__ kbuf, const uint8_t* __restrict__ sk,
                                          size_t sk_stride, const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss,
                                          int force_reject = 0) {
  using Q = FrodoSsShape<P>;
#pragma unroll
  for (int i = 0; i < Q::REM; ++i) a[i] ^= k2::deint(ld64(c + 8 * (Q::NB * Q::RW + i)), p);
  uint64_t mask = 0;
  if (fail) mask = 0 - (uint64_t)(fail[inst] != 0 || force_reject != 0);
#pragma unroll
  for (int i = 0; i < Q::LS; ++i) {
    const uint64_t k = ld64(kbuf + (size_t)inst * 64 + 8 * i);
    const uint64_t s = sk ? ld64(sk + (size_t)inst * sk_stride + 8 * i) : 0;
    a[Q::REM + i] ^= k2::deint((k & ~mask) | (s & mask), p);
  }
  a[Q::REM + Q::LS] ^= k2::half_const(kDsShake, p);
  a[Q::RW - 1] ^= k2::half_const(kPadEnd, p);
  k2::keccakf2(a, p);
  store2_words<Q::LS>(a, p, ss + (size_t)inst * P::sec);
}

template <class P>
__device__ __forceinline__ void ss_body(int vb, int n, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ kbuf,
                                        const uint8_t* __restrict__ sk, size_t sk_stride,
                                        const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss) {
  FR_PAIR_SETUP
  const uint8_t* c = ct + (size_t)inst * P::ct;
  uint32_t a[25];
  ss_absorb<P>(a, p, c);
  ss_finish<P>(a, p, inst, c, kbuf, sk, sk_stride, fail, ss);
}

template <class P>
__global__ void __launch_bounds__(64) frodo_ss(int n, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ kbuf,
                                               const uint8_t* __restrict__ sk, size_t sk_stride,
                                               const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss) {
  chain_priority();
  ss_body<P>(blockIdx.x, n, ct, kbuf, sk, sk_stride, fail, ss);
}

// decaps: W = C - B' S, mu' = Decode(W) (B bits per entry, rounded).  256
// threads per instance, each over columns k = tid, tid + 256, ... of all
// eight rows of B' and S^T (neighbouring lanes read neighbouring entries),
// holding the 64 sums (i, j) in registers; a butterfly over the wave then
// leaves lane L with sum L (each step sends the half a lane does not keep),
// and the four waves add theirs in LDS.
// Task arrays: i0 = ct, i1 = sk rows key_stride bytes apart, mubuf.  One
// block per work item.
template <class P>
__global__ void __launch_bounds__(256) frodo_dec_mu(const FrMt M) {
  __shared__ uint32_t W[64];
  if ((int)blockIdx.x >= M.total) return;
  const FrMtTask& task = M.t[mt_task_of_item(M, blockIdx.x)];
  const int inst = blockIdx.x - task.first, tid = threadIdx.x;
  const uint8_t* __restrict__ ct = task.i0;
  const uint8_t* __restrict__ sk = task.i1;
  const size_t sk_stride = task.key_stride;
  uint8_t* __restrict__ mubuf = task.mubuf;
  if (tid < 64) W[tid] = 0;
  __syncthreads();
  const uint8_t* c = ct + (size_t)inst * P::ct;
  const uint8_t* sT = sk + (size_t)inst * sk_stride + P::sk_s;
  uint32_t acc[64];
#pragma unroll
  for (int t = 0; t < 64; ++t) acc[t] = 0;
  for (int k = tid; k < P::n; k += 256) {
    uint32_t bp[8], s[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      bp[r] = unpack_at<P::logq>(c, r * P::n + k);
      s[r] = sT[2 * (r * P::n + k)] | ((uint32_t)sT[2 * (r * P::n + k) + 1] << 8);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[8 * i + j] += bp[i] * s[j];
  }
  const int lane = tid & 63;
#pragma unroll
  for (int m = 64; m > 1; m >>= 1) {
    const int d = m >> 1;  // partner distance = half the values still held
    const bool up = (lane & d) != 0;
#pragma unroll
    for (int t = 0; t < d; ++t) {
      const uint32_t keep = up ? acc[t + d] : acc[t];
      const uint32_t send = up ? acc[t] : acc[t + d];
      acc[t] = keep + (uint32_t)__shfl_xor((int)send, d, 64);
    }
  }
  atomicAdd(&W[lane], acc[0]);
  __syncthreads();
  if (tid < 8) {
    constexpr int SH = P::logq - P::b;
    uint32_t word = 0;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int idx = 8 * tid + jj;
      const uint32_t w = (unpack_at<P::logq>(c + P::ct_c, idx) - W[idx]) & P::qmask;
      word |= (((w + (1u << (SH - 1))) >> SH) & ((1u << P::b) - 1)) << (P::b * jj);
    }
    uint8_t* m = mubuf + (size_t)inst * 32 + P::b * tid;
#pragma unroll
    for (int t = 0; t < P::b; ++t) m[t] = (uint8_t)(word >> (8 * t));
  }
}

// ----------------------------------------------------------------- fused A generation + MFMA

// Byte limbs of four u16 entries held in one 64-bit word: lo = the low
// bytes (entry k's at byte k) biased by 0x80, hi = the high bytes.  v_perm
// selects bytes of {d1, d0} (selector 0-3: d0's bytes, 4-7: d1's).
__device__ __forceinline__ void limb_split(uint64_t w, uint32_t& lo, uint32_t& hi) {
  const uint32_t d0 = (uint32_t)w, d1 = (uint32_t)(w >> 32);
  lo = __builtin_amdgcn_perm(d1, d0, 0x06040200u) ^ 0x80808080u;
  hi = __builtin_amdgcn_perm(d1, d0, 0x07050301u);
}

// Low bytes of the four lanes of an MFMA result, packed (element i at byte i).
__device__ __forceinline__ int pack_low_bytes(const v4i t) {
  return (int)(__builtin_amdgcn_perm((uint32_t)t[1], (uint32_t)t[0], 0x0C0C0400u) |
               __builtin_amdgcn_perm((uint32_t)t[3], (uint32_t)t[2], 0x04000C0Cu));
}

__device__ __forceinline__ void unit_limbs(const uint64_t w[4], v4i& lo, v4i& hi) {
  uint32_t l[4], h[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) limb_split(w[m], l[m], h[m]);
  lo = v4i{(int)l[0], (int)l[1], (int)l[2], (int)l[3]};
  hi = v4i{(int)h[0], (int)h[1], (int)h[2], (int)h[3]};
}

// Row generation: each block of 21 words (84 entries of A) is cut into
// MFMA units of 4 words (16 columns).  84 is not a multiple of 16, so up to
// three words carry into the next block; the carry count follows the block
// index mod 4 and is a template parameter so all register indices are static.
template <int C, class UNIT>
__device__ __forceinline__ void consume_block(const uint64_t A[25], uint64_t pend[3], UNIT&& unit) {
  constexpr int TOTAL = kRateShake128 + C, NU = TOTAL / 4, CO = TOTAL % 4;
#pragma unroll
  for (int v = 0; v < NU; ++v) {
    uint64_t w[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int idx = 4 * v + m;
      w[m] = idx < C ? pend[idx] : A[idx - C];
    }
    unit(w);
  }
#pragma unroll
  for (int m = 0; m < CO; ++m) pend[m] = A[kRateShake128 - CO + m];
}

// Row `row` of A = SHAKE128(LE16(row) || seedA), in units of 16 columns; the
// last units may run past n (the caller drops them).
template <int BLOCKS, class UNIT>
__device__ __forceinline__ void gen_row(int row, uint64_t s0, uint64_t s1, UNIT&& unit) {
  uint64_t A[25];
  keccak_zero(A);
  A[0] = (uint64_t)(row & 0xFFFF) | (s0 << 16);
  A[1] = (s0 >> 48) | (s1 << 16);
  A[2] = (s1 >> 48) | (kDsShake << 16);
  A[kRateShake128 - 1] ^= kPadEnd;
  uint64_t pend[3] = {0, 0, 0};
  static_assert(BLOCKS % 4 == 0, "carry pattern repeats every four blocks");
  for (int t = 0; t < BLOCKS; t += 4) {
    keccakf(A);
    consume_block<0>(A, pend, unit);
    keccakf(A);
    consume_block<1>(A, pend, unit);
    keccakf(A);
    consume_block<2>(A, pend, unit);
    keccakf(A);
    consume_block<3>(A, pend, unit);
  }
}


// ----------------------------------------------------------------- FrodoKEM-976-AES rows
// A[i][j..j+7] = AES-128_seedA(LE16(i) || LE16(j) || 0^12) as little-endian
// u16; lane = row as for SHAKE128, two blocks per 16-column MFMA unit.  One
// T-table Te[x] = 2S(x) | S(x) << 8 | S(x) << 16 | 3S(x) << 24 (the MixColumns
// column of input row 0; rows 1-3 are its byte rotations) is replicated 32
// times in LDS, entry x of copy c at dword 32 x + c: lane l reads copy l % 32,
// so the 32 lanes of a half-wave hit 32 different banks.  The S-box is
// computed from x^254 in GF(2^8) when a workgroup starts; round keys are
// wave-uniform and live in scalar registers.
constexpr int kAesCopies = 32;
constexpr int kAesLdsWords = 256 * kAesCopies + 44;  // tables + 11 round keys

__device__ __forceinline__ uint32_t gf_mul2(uint32_t a) { return ((a << 1) ^ ((a >> 7) * 0x1Bu)) & 0xFFu; }

__device__ __forceinline__ uint32_t gf_mul(uint32_t a, uint32_t b) {
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    r ^= (b & 1u) ? a : 0u;
    a = gf_mul2(a);
    b >>= 1;
  }
  return r;
}

__device__ uint32_t aes_sbox(uint32_t x) {
  uint32_t inv = 1, base = x;  // x^254 = x^-1 (0 -> 0)
#pragma unroll
  for (int e = 254; e; e >>= 1) {
    if (e & 1) inv = gf_mul(inv, base);
    base = gf_mul(base, base);
  }
  uint32_t r = inv;
#pragma unroll
  for (int k = 1; k < 5; ++k) r ^= ((inv << k) | (inv >> (8 - k))) & 0xFFu;
  return r ^ 0x63u;
}

// Build the replicated T-table and expand seedA (16 bytes) into T[256*32 ...].
__device__ void aes_setup(uint32_t* T, const uint8_t* seed_a, int tid, int nthreads) {
  for (int x = tid; x < 256; x += nthreads) {
    const uint32_t sx = aes_sbox((uint32_t)x), m2 = gf_mul2(sx);
    T[x * kAesCopies] = m2 | (sx << 8) | (sx << 16) | ((m2 ^ sx) << 24);
  }
  __syncthreads();
  for (int t = tid; t < 256 * kAesCopies; t += nthreads)
    if (t & (kAesCopies - 1)) T[t] = T[t & ~(kAesCopies - 1)];
  if (tid == 0) {
    uint32_t* rk = T + 256 * kAesCopies;
    for (int i = 0; i < 4; ++i)
      rk[i] = seed_a[4 * i] | ((uint32_t)seed_a[4 * i + 1] << 8) | ((uint32_t)seed_a[4 * i + 2] << 16) |
              ((uint32_t)seed_a[4 * i + 3] << 24);
    uint32_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
      uint32_t t = rk[i - 1];
      if (i % 4 == 0) {
        t = (t >> 8) | (t << 24);  // RotWord (bytes are little-endian)
        uint32_t sub = 0;
        for (int b = 0; b < 4; ++b) sub |= ((T[((t >> (8 * b)) & 0xFF) * kAesCopies] >> 8) & 0xFFu) << (8 * b);
        t = sub ^ rcon;
        rcon = gf_mul2(rcon);
      }
      rk[i] = rk[i - 4] ^ t;
    }
  }
  __syncthreads();
}

struct AesRows {
  const uint8_t* tb;  // T-table base in bytes
  uint32_t c4;        // this lane's copy offset in bytes
  uint32_t rk[44];

  __device__ __forceinline__ uint32_t te(uint32_t s, int byte) const {
    const uint32_t x = __builtin_amdgcn_ubfe(s, 8 * byte, 8);
    return *reinterpret_cast<const uint32_t*>(tb + ((x << 7) | c4));
  }
  __device__ __forceinline__ static uint32_t rotl8(uint32_t v, int k) {
    return k ? __builtin_amdgcn_alignbit(v, v, 32 - 8 * k) : v;
  }
  // one block: in = (w0, 0, 0, 0) -> out[4]
  __device__ __forceinline__ void encrypt(uint32_t w0, uint32_t out[4]) const {
    uint32_t s0 = w0 ^ rk[0], s1 = rk[1], s2 = rk[2], s3 = rk[3];
#pragma unroll
    for (int r = 1; r < 10; ++r) {
      const uint32_t t0 = te(s0, 0) ^ rotl8(te(s1, 1), 1) ^ rotl8(te(s2, 2), 2) ^ rotl8(te(s3, 3), 3) ^ rk[4 * r];
      const uint32_t t1 = te(s1, 0) ^ rotl8(te(s2, 1), 1) ^ rotl8(te(s3, 2), 2) ^ rotl8(te(s0, 3), 3) ^ rk[4 * r + 1];
      const uint32_t t2 = te(s2, 0) ^ rotl8(te(s3, 1), 1) ^ rotl8(te(s0, 2), 2) ^ rotl8(te(s1, 3), 3) ^ rk[4 * r + 2];
      const uint32_t t3 = te(s3, 0) ^ rotl8(te(s0, 1), 1) ^ rotl8(te(s1, 2), 2) ^ rotl8(te(s2, 3), 3) ^ rk[4 * r + 3];
      s0 = t0; s1 = t1; s2 = t2; s3 = t3;
    }
    // last round: S(x) is byte 1 of Te[x]
    const uint32_t st[4] = {s0, s1, s2, s3};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t b0 = (te(st[c], 0) >> 8) & 0xFFu;
      const uint32_t b1 = te(st[(c + 1) & 3], 1) & 0xFF00u;
      const uint32_t b2 = (te(st[(c + 2) & 3], 2) << 8) & 0xFF0000u;
      const uint32_t b3 = (te(st[(c + 3) & 3], 3) << 16) & 0xFF000000u;
      out[c] = (b0 | b1 | b2 | b3) ^ rk[40 + c];
    }
  }
  // row i: 61 units of 16 columns (= 976), each from blocks j and j + 8
  template <class UNIT>
  __device__ __forceinline__ void operator()(int row, UNIT&& unit) const {
    for (int u = 0; u < kFUnits; ++u) {
      uint32_t a[4], b[4];
      encrypt((uint32_t)(row & 0xFFFF) | ((uint32_t)(16 * u) << 16), a);
      encrypt((uint32_t)(row & 0xFFFF) | ((uint32_t)(16 * u + 8) << 16), b);
      const uint64_t w[4] = {a[0] | ((uint64_t)a[1] << 32), a[2] | ((uint64_t)a[3] << 32),
                             b[0] | ((uint64_t)b[1] << 32), b[2] | ((uint64_t)b[3] << 32)};
      unit(w);
    }
  }
};

__device__ __forceinline__ void aes_rows_init(AesRows& g, const uint32_t* T, int lane) {
  g.tb = reinterpret_cast<const uint8_t*>(T);
  g.c4 = (uint32_t)(lane & (kAesCopies - 1)) * 4u;
#pragma unroll
  for (int i = 0; i < 44; ++i) g.rk[i] = __builtin_amdgcn_readfirstlane(T[256 * kAesCopies + i]);
}

// keygen: B = A S + E.  pk[16:] and sk[sec+16:] receive pack_D(B).
// AES: FrodoKEM-976-AES rows instead of SHAKE128 rows.  Lanes take rows
// wave * 64 + lane + pass * threads; a wave whose rows all lie past n in the
// last pass skips it.
// Task arrays: o0 = pk, o1 = sk, s8 / e8 = the chains' int8 S^T and E.
template <class P, bool AES, bool PERSIST, int TH>
__global__ void __launch_bounds__(TH) frodo_as_fused(const FrMt M) {
  static_assert(!AES || P::n == 976, "AES rows are defined for FrodoKEM-976");
  using SH = FrodoShape<P, TH>;
  constexpr int T = TH, SROW = P::srow, N = P::n;
  __shared__ __align__(16) int8_t Sl[9 * SROW];  // rows 0..7 = S^T, row 8 = zeros
  __shared__ int colsum[8];
  __shared__ __align__(16) uint32_t aes_lds[AES ? kAesLdsWords : 1];
  __shared__ __align__(16) uint16_t Bst[P::logq < 16 ? P::mat : 1];  // D < 16: B staged for packing
  for (int item = blockIdx.x;; item += (int)gridDim.x) {
  if (item >= M.total) break;
  const int tid = PERSIST ? fresh_tid() : (int)threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const FrMtTask& task = M.t[mt_task_of_item(M, item)];
  const int inst = item - task.first;
  uint8_t* __restrict__ pk = task.o0;
  uint8_t* __restrict__ sk = task.o1;
  const uint32_t* __restrict__ s8 = task.s8;
  const int8_t* __restrict__ e8 = reinterpret_cast<const int8_t*>(task.e8);
  __syncthreads();  // every wave is done with the previous item's LDS
  const uint32_t* s = s8 + (size_t)inst * (P::mat / 4);
  uint32_t* Sd = reinterpret_cast<uint32_t*>(Sl);
  for (int x = tid; x < 9 * SROW / 4; x += T) {
    const int k = x / (SROW / 4), j4 = x % (SROW / 4);
    Sd[x] = (k < 8 && j4 < N / 4) ? s[k * (N / 4) + j4] : 0u;
  }
  if (tid < 8) colsum[tid] = 0;
  __syncthreads();
  for (int x = tid; x < P::mat / 4; x += T) {
    const int k = x / (N / 4), j4 = x % (N / 4);
    const uint32_t v = Sd[k * (SROW / 4) + j4];
    const int sum = (int8_t)v + (int8_t)(v >> 8) + (int8_t)(v >> 16) + (int8_t)(v >> 24);
    atomicAdd(&colsum[k], sum);
  }
  __syncthreads();

  const uint8_t* p = pk + (size_t)inst * P::pk;
  const uint64_t s0 = ld64(p), s1 = ld64(p + 8);
  AesRows aesg;
  if constexpr (AES) {
    aes_setup(aes_lds, p, tid, T);
    aes_rows_init(aesg, aes_lds, lane);
  }
  const int q = lane >> 4, nn = lane & 15;
  // B operand: lane (q, nn) carries S^T column 4r + nn%4 when nn/4 == q, else zeros
  const bool mine = (nn >> 2) == q;
  const int boff0 = mine ? (nn & 3) * SROW : 8 * SROW;
  const int boff1 = mine ? (4 + (nn & 3)) * SROW : 8 * SROW;
  const int g = lane >> 4;
  const int8_t* e = e8 + (size_t)inst * P::mat;
  uint8_t* pb = pk + (size_t)inst * P::pk + kFrodoSeedA;
  uint8_t* sb = sk + (size_t)inst * P::sk + P::sec + kFrodoSeedA;
  for (int pass = 0; pass < SH::passes; ++pass) {
    const int row0 = pass * T + wave * 64;
    if (row0 >= N) break;
    v4i acc_lo0 = {0, 0, 0, 0}, acc_lo1 = {0, 0, 0, 0}, acc_hi0 = {0, 0, 0, 0}, acc_hi1 = {0, 0, 0, 0};
    int u = 0;
    auto unit = [&](const uint64_t w[4]) {
      if (u < P::units) {
        v4i lo, hi;
        unit_limbs(w, lo, hi);
        const v4i b0 = *reinterpret_cast<const v4i*>(Sl + boff0 + 16 * u);
        const v4i b1 = *reinterpret_cast<const v4i*>(Sl + boff1 + 16 * u);
        acc_lo0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(lo, b0, acc_lo0, 0, 0, 0);
        acc_lo1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(lo, b1, acc_lo1, 0, 0, 0);
        acc_hi0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(hi, b0, acc_hi0, 0, 0, 0);
        acc_hi1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(hi, b1, acc_hi1, 0, 0, 0);
      }
      ++u;
    };
    if constexpr (AES) aesg(row0 + lane, unit);
    else gen_row<P::row_blocks>(row0 + lane, s0, s1, unit);

    // D layout: lane (g, nn), element i = C[4g + i][nn].  C row m comes from
    // A-operand lane 16 * (nn / 4) + m; the column is 4r + nn % 4.
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const v4i lo = r ? acc_lo1 : acc_lo0, hi = r ? acc_hi1 : acc_hi0;
      const int col = 4 * r + (nn & 3);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int grow = row0 + 16 * (nn >> 2) + 4 * g + i;
        if (grow < N) {
          const uint32_t v = (uint32_t)lo[i] + 256u * (uint32_t)hi[i] + 128u * (uint32_t)colsum[col] +
                             (uint32_t)(int32_t)e[grow * kFrodoNbar + col];
          const int idx = grow * kFrodoNbar + col;
          if constexpr (P::logq == 16) {
            const uint16_t be = (uint16_t)(((v & 0xFF) << 8) | ((v >> 8) & 0xFF));
            *reinterpret_cast<uint16_t*>(pb + 2 * idx) = be;
            *reinterpret_cast<uint16_t*>(sb + 2 * idx) = be;
          } else {
            Bst[idx] = (uint16_t)(v & P::qmask);
          }
        }
      }
    }
  }
  if constexpr (P::logq < 16) {
    __syncthreads();
    for (int w = tid; w < P::logq * P::mat / 32; w += T) {
      const uint32_t word = pack15_word(Bst, P::mat, w);
      reinterpret_cast<uint32_t*>(pb)[w] = word;
      reinterpret_cast<uint32_t*>(sb)[w] = word;
    }
  }
  if (!PERSIST) break;  // one work item per workgroup
  }  // item
}

// enc / dec: B' = S' A + E', then V = S' B + E'' and C = V + Encode(mu).
// enc writes ct; dec compares with the received ct and raises fail[inst].
// Task arrays.  enc: i0 = pk, i1 = coins (mu), o0 = ct.  dec: i0 = ct,
// i1 = sk (its pk part is used), mubuf = mu', fail.  Both: s8 / e8 / epp8 =
// the chains' int8 S', E', E''.
template <class P, bool DEC, bool AES, bool PERSIST, int TH>
__global__ void __launch_bounds__(TH) frodo_sa_fused(const FrMt M) {
  static_assert(!AES || P::n == 976, "AES rows are defined for FrodoKEM-976");
  using SH = FrodoShape<P, TH>;
  constexpr int T = TH, SP = SH::sprow, N = P::n;
  __shared__ __align__(16) uint32_t Bacc[P::mat];
  __shared__ __align__(16) int8_t Sp[8 * SP];  // S' rows, zero for i >= n
  __shared__ int rowsum[8];
  __shared__ uint32_t V[64];
  __shared__ uint32_t diff;
  __shared__ __align__(16) uint32_t aes_lds[AES ? kAesLdsWords : 1];
  for (int item = blockIdx.x;; item += (int)gridDim.x) {
  if (item >= M.total) break;
  const int tid = PERSIST ? fresh_tid() : (int)threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const FrMtTask& task = M.t[mt_task_of_item(M, item)];
  const int inst = item - task.first;
  const uint8_t* __restrict__ pkb = DEC ? task.i1 + P::sec : task.i0;
  const size_t pk_stride = DEC ? (size_t)P::sk : (size_t)P::pk;
  const uint32_t* __restrict__ s8 = task.s8;
  const int8_t* __restrict__ e8 = reinterpret_cast<const int8_t*>(task.e8);
  const int8_t* __restrict__ epp8 = reinterpret_cast<const int8_t*>(task.epp8);
  const uint8_t* __restrict__ mu = DEC ? task.mubuf : task.i1;
  const size_t mu_stride = DEC ? (size_t)32 : (size_t)P::ecoins;
  uint8_t* __restrict__ ct_out = task.o0;
  const uint8_t* __restrict__ ct_in = task.i0;
  uint32_t* __restrict__ fail = task.fail;
  __syncthreads();  // every wave is done with the previous item's LDS
  const uint32_t* s = s8 + (size_t)inst * (P::mat / 4);
  uint32_t* Sd = reinterpret_cast<uint32_t*>(Sp);
  for (int x = tid; x < 8 * SP / 4; x += T) {
    const int k = x / (SP / 4), i4 = x % (SP / 4);
    Sd[x] = i4 < N / 4 ? s[k * (N / 4) + i4] : 0u;
  }
  for (int x = tid; x < P::mat; x += T) Bacc[x] = 0;
  if (tid < 8) rowsum[tid] = 0;
  if (tid < 64) V[tid] = 0;
  if (tid == 0) diff = 0;
  __syncthreads();
  for (int x = tid; x < 8 * SP / 4; x += T) {
    const uint32_t v = Sd[x];
    const int sum = (int8_t)v + (int8_t)(v >> 8) + (int8_t)(v >> 16) + (int8_t)(v >> 24);
    atomicAdd(&rowsum[x / (SP / 4)], sum);
  }

  const uint8_t* pk = pkb + (size_t)inst * pk_stride;
  const uint64_t s0 = ld64(pk), s1 = ld64(pk + 8);
  AesRows aesg;
  if constexpr (AES) {
    aes_setup(aes_lds, pk, tid, T);
    aes_rows_init(aesg, aes_lds, lane);
  }
  const int g = lane >> 4, nn = lane & 15;
  // transpose selectors: B operand of transpose q is one-hot(byte nn) on lanes of group q
  const int onehot = 1 << (8 * (nn & 3));
  v4i sel[4];
#pragma unroll
  for (int qq = 0; qq < 4; ++qq)
#pragma unroll
    for (int d = 0; d < 4; ++d) sel[qq][d] = (g == qq && d == (nn >> 2)) ? onehot : 0;
  const v4i zero = {0, 0, 0, 0};

  for (int pass = 0; pass < SH::passes; ++pass) {
    const int row0 = pass * T + wave * 64;
    if (row0 >= N) break;
    // product A operand: lane (g, m = nn) holds S'[m][rows row0 + 16 q + 4 g + i] in dword q, byte i
    v4i sop = {0, 0, 0, 0};
    if (nn < 8) {
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) sop[qq] = (int)Sd[nn * (SP / 4) + (row0 + 16 * qq + 4 * g) / 4];
    }
    int u = 0;
    auto unit = [&](const uint64_t w[4]) {
      if (u < P::units) {
        v4i lm[2];
        unit_limbs(w, lm[0], lm[1]);
        v4i prod[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          // t_q: lane (g, nn) element i = limb of row 16 q + 4 g + i at column 16 u + nn
          v4i packed;
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) {
            const v4i t = __builtin_amdgcn_mfma_i32_16x16x64_i8(lm[h], sel[qq], zero, 0, 0, 0);
            packed[qq] = pack_low_bytes(t);
          }
          prod[h] = __builtin_amdgcn_mfma_i32_16x16x64_i8(sop, packed, zero, 0, 0, 0);
        }
        const int col = 16 * u + nn;
        if (g < 2 && col < N) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            atomicAdd(&Bacc[(4 * g + i) * N + col], (uint32_t)prod[0][i] + 256u * (uint32_t)prod[1][i]);
        }
      }
      ++u;
    };
    if constexpr (AES) aesg(row0 + lane, unit);
    else gen_row<P::row_blocks>(row0 + lane, s0, s1, unit);
  }
  __syncthreads();

  // B' = acc + 128 rowsum + E'  (k-major, 8 x n)
  const int8_t* ep = e8 + (size_t)inst * P::mat;
  const uint8_t* cin = DEC ? ct_in + (size_t)inst * P::ct : nullptr;
  uint8_t* cout = DEC ? nullptr : ct_out + (size_t)inst * P::ct;
  uint32_t d = 0;
  for (int x = tid; x < P::mat; x += T) {
    const uint32_t v = (Bacc[x] + 128u * (uint32_t)rowsum[x / N] + (uint32_t)(int32_t)ep[x]) & P::qmask;
    if constexpr (P::logq == 16) {
      const uint16_t be = (uint16_t)((v << 8) | (v >> 8));
      if (DEC) d |= (uint32_t)(*reinterpret_cast<const uint16_t*>(cin + 2 * x) ^ be);
      else *reinterpret_cast<uint16_t*>(cout + 2 * x) = be;
    } else {
      Bacc[x] = v;  // packed below
    }
  }
  // V = S' B (B = pack_D(B) in pk, n x 8): 64 outputs, T / 64 slices of rows
  {
    constexpr int NSL = T / 64, CH = (N + NSL - 1) / NSL;
    const int o = tid & 63, i = o >> 3, j = o & 7, slice = tid >> 6;
    const int k1 = min(N, slice * CH + CH);
    uint32_t acc = 0;
    for (int k = slice * CH; k < k1; ++k)
      acc += (uint32_t)(int32_t)Sp[i * SP + k] * unpack_at<P::logq>(pk + kFrodoSeedA, k * kFrodoNbar + j);
    atomicAdd(&V[o], acc);
  }
  __syncthreads();
  if (tid < 64) {
    // B message bits of entry tid, least significant first from the bytes of mu
    const uint8_t* m = mu + (size_t)inst * mu_stride;
    const int bit = P::b * tid, sh = bit & 7;
    uint32_t bits = m[bit >> 3];
    if (sh + P::b > 8) bits |= (uint32_t)m[(bit >> 3) + 1] << 8;
    const uint32_t enc = ((bits >> sh) & ((1u << P::b) - 1)) << (P::logq - P::b);
    const uint32_t cv = (V[tid] + (uint32_t)(int32_t)epp8[(size_t)inst * 64 + tid] + enc) & P::qmask;
    if constexpr (P::logq == 16) {
      const uint16_t be = (uint16_t)((cv << 8) | (cv >> 8));
      if (DEC) d |= (uint32_t)(*reinterpret_cast<const uint16_t*>(cin + P::ct_c + 2 * tid) ^ be);
      else *reinterpret_cast<uint16_t*>(cout + P::ct_c + 2 * tid) = be;
    } else {
      V[tid] = cv;
    }
  }
  if constexpr (P::logq < 16) {
    // pack_15(B') || pack_15(C): ct words 0 .. ct_c / 4 from B', then C
    __syncthreads();
    constexpr int WB = P::ct_c / 4, WC = (P::ct - P::ct_c) / 4;
    for (int w = tid; w < WB + WC; w += T) {
      const uint32_t word = w < WB ? pack15_word(Bacc, P::mat, w) : pack15_word(V, 64, w - WB);
      if (DEC) d |= reinterpret_cast<const uint32_t*>(cin)[w] ^ word;
      else reinterpret_cast<uint32_t*>(cout)[w] = word;
    }
  }
  if (DEC) {
    if (d) atomicOr(&diff, 1u);
    __syncthreads();
    if (tid == 0) fail[inst] = diff;
  }
  if (!PERSIST) break;  // one work item per workgroup
  }  // item
}


// enc / dec over groups of up to G instances that share one public key (same
// seedA, same B): the group's A is generated once and multiplied by the S'
// of every member.  Two members fill the 16 rows of one product MFMA
// (2 x nbar = 16), so G = 4 costs two product MFMAs per limb beside the four
// shared transposes.  A 4-member group needs 158.9 KB of LDS (one workgroup
// per CU, as the 125-VGPR budget allows anyway); groups of 1 to 3 leave
// slots empty.  Group b holds instances first .. first+cnt-1:
// groups[b] = (first, cnt), or, with groups == nullptr, (G b, min(G, n - G b)).
// Members are independent KEM instances; only A, B and the pk bytes are common.
template <bool DEC, int G>
__global__ void __launch_bounds__(1024) frodo_sa_grouped(int ngroups, const uint2* __restrict__ groups, int n,
                                                         const uint8_t* __restrict__ pkb, size_t pk_stride,
                                                         const uint32_t* __restrict__ s8, const int8_t* __restrict__ e8,
                                                         const int8_t* __restrict__ epp8, const uint8_t* __restrict__ mu,
                                                         size_t mu_stride, uint8_t* __restrict__ ct_out,
                                                         const uint8_t* __restrict__ ct_in, uint32_t* __restrict__ fail) {
  static_assert(G % 2 == 0, "two members per product MFMA");
  __shared__ __align__(16) uint32_t Bacc[G * kFMat];
  __shared__ __align__(16) int8_t Sp[G * 8 * 1024];  // member j, row k at (8 j + k) * 1024; zero past 976
  __shared__ int rowsum[G * 8];
  __shared__ uint32_t V[G * 64];
  __shared__ uint32_t diff[G];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  if (b >= ngroups) return;
  const int first = groups ? (int)groups[b].x : G * b;
  const int cnt = groups ? (int)groups[b].y : min(G, n - G * b);
  uint32_t* Sd = reinterpret_cast<uint32_t*>(Sp);
  for (int x = tid; x < G * 8 * 256; x += 1024) {
    const int j = x >> 11, k = (x >> 8) & 7, i4 = x & 255;
    Sd[x] = (j < cnt && i4 < kFN / 4) ? s8[(size_t)(first + j) * (kFMat / 4) + k * (kFN / 4) + i4] : 0u;
  }
  for (int x = tid; x < G * kFMat; x += 1024) Bacc[x] = 0;
  if (tid < G * 8) rowsum[tid] = 0;
  if (tid < G * 64) V[tid] = 0;
  if (tid < G) diff[tid] = 0;
  __syncthreads();
  for (int x = tid; x < G * 8 * 256; x += 1024) {
    const uint32_t v = Sd[x];
    const int sum = (int8_t)v + (int8_t)(v >> 8) + (int8_t)(v >> 16) + (int8_t)(v >> 24);
    atomicAdd(&rowsum[x >> 8], sum);
  }

  const uint8_t* pk = pkb + (size_t)first * pk_stride;
  const uint64_t s0 = ld64(pk), s1 = ld64(pk + 8);
  const int g = lane >> 4, nn = lane & 15;
  // product p, A operand: lane (g, nn) holds S' row (nn & 7) of member 2p + nn / 8
  // at wave rows 16 q + 4 g + i (dword q, byte i)
  v4i sop[G / 2];
#pragma unroll
  for (int pp = 0; pp < G / 2; ++pp)
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
      sop[pp][qq] = (int)Sd[((2 * pp + (nn >> 3)) * 8 + (nn & 7)) * 256 + (wave * 64 + 16 * qq + 4 * g) / 4];
  const int onehot = 1 << (8 * (nn & 3));
  v4i sel[4];
#pragma unroll
  for (int qq = 0; qq < 4; ++qq)
#pragma unroll
    for (int d = 0; d < 4; ++d) sel[qq][d] = (g == qq && d == (nn >> 2)) ? onehot : 0;
  const v4i zero = {0, 0, 0, 0};

  int u = 0;
  gen_row<F976::row_blocks>(wave * 64 + lane, s0, s1, [&](const uint64_t w[4]) {
    if (u < kFUnits) {
      v4i lm[2];
      unit_limbs(w, lm[0], lm[1]);
      v4i packed[2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
          const v4i t = __builtin_amdgcn_mfma_i32_16x16x64_i8(lm[h], sel[qq], zero, 0, 0, 0);
          packed[h][qq] = pack_low_bytes(t);
        }
      const int col = 16 * u + nn;
#pragma unroll
      for (int pp = 0; pp < G / 2; ++pp) {
        const v4i p0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(sop[pp], packed[0], zero, 0, 0, 0);
        const v4i p1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(sop[pp], packed[1], zero, 0, 0, 0);
        // D row 4 g + i = member 2 pp + g / 2, S' row 4 (g & 1) + i
        const int j = 2 * pp + (g >> 1);
        if (j < cnt && col < kFN) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            atomicAdd(&Bacc[j * kFMat + (4 * (g & 1) + i) * kFN + col], (uint32_t)p0[i] + 256u * (uint32_t)p1[i]);
        }
      }
    }
    ++u;
  });
  __syncthreads();

  // B' = acc + 128 rowsum + E' per member
  uint32_t dj[G];
#pragma unroll
  for (int j = 0; j < G; ++j) dj[j] = 0;
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if (j >= cnt) break;
    const size_t inst = (size_t)(first + j);
    const int8_t* ep = e8 + inst * kFMat;
    for (int x = tid; x < kFMat; x += 1024) {
      const uint32_t v = (Bacc[j * kFMat + x] + 128u * (uint32_t)rowsum[j * 8 + x / kFN] + (uint32_t)(int32_t)ep[x]) & 0xFFFF;
      const uint16_t be = (uint16_t)((v << 8) | (v >> 8));
      if (DEC) dj[j] |= (uint32_t)(*reinterpret_cast<const uint16_t*>(ct_in + inst * kFrodoCt + 2 * x) ^ be);
      else *reinterpret_cast<uint16_t*>(ct_out + inst * kFrodoCt + 2 * x) = be;
    }
  }
  // V = S' B for every member (B read once per k)
  {
    const int o = tid & 63, i = o >> 3, jj = o & 7, slice = tid >> 6;
    uint32_t acc[G];
#pragma unroll
    for (int j = 0; j < G; ++j) acc[j] = 0;
    for (int k = slice * 61; k < slice * 61 + 61; ++k) {
      const uint8_t* bp = pk + kFrodoSeedA + 2 * (k * kFNB + jj);
      const uint32_t bv = ((uint32_t)bp[0] << 8) | bp[1];
#pragma unroll
      for (int j = 0; j < G; ++j) acc[j] += (uint32_t)(int32_t)Sp[(j * 8 + i) * 1024 + k] * bv;
    }
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < cnt) atomicAdd(&V[j * 64 + o], acc[j]);
  }
  __syncthreads();
  if (tid < 64) {
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j >= cnt) break;
      const size_t inst = (size_t)(first + j);
      const uint8_t* m = mu + inst * mu_stride + 3 * (tid >> 3);
      const uint32_t t24 = m[0] | ((uint32_t)m[1] << 8) | ((uint32_t)m[2] << 16);
      const uint32_t enc = ((t24 >> (3 * (tid & 7))) & 7u) << 13;
      const uint32_t cv = (V[j * 64 + tid] + (uint32_t)(int32_t)epp8[inst * 64 + tid] + enc) & 0xFFFF;
      const uint16_t be = (uint16_t)((cv << 8) | (cv >> 8));
      if (DEC) dj[j] |= (uint32_t)(*reinterpret_cast<const uint16_t*>(ct_in + inst * kFrodoCt + kFCtC1 + 2 * tid) ^ be);
      else *reinterpret_cast<uint16_t*>(ct_out + inst * kFrodoCt + kFCtC1 + 2 * tid) = be;
    }
  }
  if (DEC) {
#pragma unroll
    for (int j = 0; j < G; ++j) atomicOr(&diff[j], dj[j] != 0 ? 1u : 0u);
    __syncthreads();
    if (tid < cnt) fail[first + tid] = diff[tid];
  }
}

// ----------------------------------------------------------------- pipelined keygen -> encaps -> decaps
//
// For keygen -> encaps -> decaps of the same instances (a plan chain), the
// serial chains of different ops that read disjoint inputs share launches:
//   x1   wave 0: keygen's pkh into sk; wave 1: encaps' pkh, G2 and noise
//        chain, wave 2 its sampling (both ops only need pk, complete
//        after the keygen product);
//   x2   wave 0: encaps' ss; wave 1: decaps' G2 and noise chain (after
//        dec_mu), wave 3 its sampling; wave 2: decaps' SHAKE over the
//        full ct blocks, kept in `state`;
//   ssd  decaps' last ss block (ct tail, kbar chosen by the FO compare).
// Every op still computes all of its own hashes; only their order changes.
template <class P>
__global__ void __launch_bounds__(192) frodo_pipe_x1(int n, const uint8_t* __restrict__ pk, uint8_t* __restrict__ sk,
                                                     const uint8_t* __restrict__ ecoins, uint8_t* __restrict__ kbuf_e,
                                                     uint32_t* __restrict__ s8, uint32_t* __restrict__ e8,
                                                     uint32_t* __restrict__ epp8) {
  __shared__ FrRing<P> ring;
  ring_init(&ring);
  const int wave = threadIdx.x >> 6;
  chain_priority();
  if (wave == 0) {
    keygen_post_body<P>(blockIdx.x, n, pk, sk);
  } else if (wave == 1) {
    encdec_pre_body<P, false, kFrChain>(blockIdx.x, n, pk, (size_t)P::pk, ecoins, (size_t)P::ecoins, kbuf_e, s8, e8,
                                        epp8, nullptr, &ring);
  } else {
    encdec_pre_body<P, false, kFrSampler>(blockIdx.x, n, pk, (size_t)P::pk, ecoins, (size_t)P::ecoins, kbuf_e, s8, e8,
                                          epp8, nullptr, &ring);
  }
}

template <class P>
__global__ void __launch_bounds__(256) frodo_pipe_x2(int n, const uint8_t* __restrict__ ct, const uint8_t* __restrict__ sk,
                                                     const uint8_t* __restrict__ kbuf_e, uint8_t* __restrict__ ss_e,
                                                     const uint8_t* __restrict__ mubuf, uint8_t* __restrict__ kbuf_d,
                                                     uint32_t* __restrict__ s8, uint32_t* __restrict__ e8,
                                                     uint32_t* __restrict__ epp8, uint32_t* __restrict__ state) {
  __shared__ FrRing<P> ring;
  ring_init(&ring);
  const int wave = threadIdx.x >> 6, vb = blockIdx.x;
  chain_priority();
  if (wave == 0) {
    ss_body<P>(vb, n, ct, kbuf_e, nullptr, 0, nullptr, ss_e);
  } else if (wave == 1) {
    encdec_pre_body<P, true, kFrChain>(vb, n, sk, (size_t)P::sk, mubuf, 32, kbuf_d, s8, e8, epp8, nullptr, &ring);
  } else if (wave == 3) {
    encdec_pre_body<P, true, kFrSampler>(vb, n, sk, (size_t)P::sk, mubuf, 32, kbuf_d, s8, e8, epp8, nullptr, &ring);
  } else {
    FR_PAIR_SETUP
    uint32_t a[25];
    ss_absorb<P>(a, p, ct + (size_t)inst * P::ct);
#pragma unroll
    for (int i = 0; i < 25; ++i) state[(size_t)i * 2 * n + 2 * inst + p] = a[i];
  }
}

// decaps' ss from the absorbed ct state: the last block (remaining ct
// words, kbar, padding) and the squeeze.
template <class P>
__device__ __forceinline__ void ssd_body(int vb, int n, const uint8_t* __restrict__ ct,
                                         const uint32_t* __restrict__ state, const uint8_t* __restrict__ kbuf_d,
                                         const uint8_t* __restrict__ sk, size_t sk_stride,
                                         const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss,
                                         int force_reject) {
  FR_PAIR_SETUP
  uint32_t a[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) a[i] = state[(size_t)i * 2 * n + 2 * inst + p];
  ss_finish<P>(a, p, inst, ct + (size_t)inst * P::ct, kbuf_d, sk, sk_stride, fail, ss, force_reject);
}

// force_reject: test facility, every instance takes the implicit-rejection
// branch (ss from s instead of k).
template <class P>
__global__ void __launch_bounds__(64) frodo_pipe_ssd(int n, const uint8_t* __restrict__ ct,
                                                     const uint32_t* __restrict__ state, const uint8_t* __restrict__ kbuf_d,
                                                     const uint8_t* __restrict__ sk, size_t sk_stride,
                                                     const uint32_t* __restrict__ fail, uint8_t* __restrict__ ss,
                                                     int force_reject) {
  chain_priority();
  ssd_body<P>(blockIdx.x, n, ct, state, kbuf_d, sk, sk_stride, fail, ss, force_reject);
}

// decaps' chains before the re-encryption, one 192-thread block per 32
// instances: wave 0 runs G2 and the S' / E' / E'' squeeze (encdec_pre_body),
// wave 2 samples it (FrRing), wave 1 absorbs the full blocks of ct for ss (ss_absorb) and leaves the
// sponge state in `state`, so that after the re-encryption only ss's last
// block remains (ssd_body) instead of the whole ct absorb.
template <class P>
__device__ __forceinline__ void dec_pre_body(int vb, int n, const uint8_t* __restrict__ ct,
                                             const uint8_t* __restrict__ sk, size_t sk_stride,
                                             const uint8_t* __restrict__ mubuf, uint8_t* __restrict__ kbuf,
                                             uint32_t* __restrict__ s8, uint32_t* __restrict__ e8,
                                             uint32_t* __restrict__ epp8, uint32_t* __restrict__ state) {
  __shared__ FrRing<P> ring;
  ring_init(&ring);
  const int wave = threadIdx.x >> 6;
  chain_priority();
  if (wave == 0) {
    encdec_pre_body<P, true, kFrChain>(vb, n, sk, sk_stride, mubuf, 32, kbuf, s8, e8, epp8, nullptr, &ring);
  } else if (wave == 2) {
    encdec_pre_body<P, true, kFrSampler>(vb, n, sk, sk_stride, mubuf, 32, kbuf, s8, e8, epp8, nullptr, &ring);
  } else {
    FR_PAIR_SETUP
    uint32_t a[25];
    ss_absorb<P>(a, p, ct + (size_t)inst * P::ct);
#pragma unroll
    for (int i = 0; i < 25; ++i) state[(size_t)i * 2 * n + 2 * inst + p] = a[i];
  }
}

template <class P>
__global__ void __launch_bounds__(192) frodo_dec_pre(int n, const uint8_t* __restrict__ ct,
                                                     const uint8_t* __restrict__ sk, size_t sk_stride,
                                                     const uint8_t* __restrict__ mubuf, uint8_t* __restrict__ kbuf,
                                                     uint32_t* __restrict__ s8, uint32_t* __restrict__ e8,
                                                     uint32_t* __restrict__ epp8, uint32_t* __restrict__ state) {
  dec_pre_body<P>(blockIdx.x, n, ct, sk, sk_stride, mubuf, kbuf, s8, e8, epp8, state);
}

// ----------------------------------------------------------------- rolling keygen -> encaps -> decaps
//
// One step of a rolling pipeline over three batch slots: keygen of slot r,
// encaps of slot r - 1 (to the keys keygen wrote one step earlier), decaps
// of slot r - 2 (of the ciphertexts encaps wrote one step earlier).  Run
// chained per batch, each chain launch holds one chain wave per 32
// instances, 256 for a batch of 8192 on 1,024 SIMDs, and lasts as long as
// its longest chain; the chains of the three ops here are independent, so
// they all go into ONE launch, one 448-thread workgroup per 32 instances of
// each op:
//   wave 0  keygen's noise chain (slot r), then the closing hash of slot
//           r - 2's encaps: ss = SHAKE(ct || k)
//   wave 1  encaps' pkh, G2 and noise chain (slot r - 1)
//   wave 2  decaps' G2 and noise chain (slot r - 2, mu' from frodo_dec_mu)
//   wave 3  decaps' SHAKE over the full ct blocks (state, finished after the
//           re-encryption by frodo_pipe_ssd), then the closing hash of slot
//           r - 1's keygen: pkh = SHAKE(pk) into its sk
//   waves 4 / 5 / 6  the samplers of waves 1 / 2 / 0 (FrRing), wave 6 also
//           keygen's seedA
// Waves w and w + 4 of a workgroup share a SIMD: each SIMD runs one chain
// (wave 0: 230 + 117 permutations for FrodoKEM-976, wave 1: 347, wave 2:
// 232, wave 3: 116 + 115) and at most one sampler.  The three noise
// squeezes share one call site (shake_seed_squeeze2_ring_rt).  Every op
// still computes all of its own hashes.
struct FrRollArgs {
  int n;
  // keygen, slot r
  const uint8_t* kc;
  uint8_t *pk_k, *sk_k;
  uint32_t *s8_k, *e8_k;
  // encaps, slot r - 1; that slot's keygen closing hash writes sk_e
  const uint8_t *pk_e, *ec;
  uint8_t *sk_e, *kbuf_e;
  uint32_t *s8_e, *e8_e, *epp8_e;
  // decaps, slot r - 2; that slot's encaps closing hash reads kbuf_ed, writes ss_ed
  const uint8_t *ct_d, *sk_d, *mubuf_d, *kbuf_ed;
  uint8_t *kbuf_d, *ss_ed;
  uint32_t *s8_d, *e8_d, *epp8_d, *state_d;
};

template <class P>
__global__ void __launch_bounds__(448) frodo_roll_x0(const FrRollArgs A) {
  constexpr int RW = P::chain_rate, LS = P::ls;
  constexpr int OUTW_K = 2 * P::mat / 4, OUTW_E = (2 * P::mat + 64) / 4;
  static_assert(OUTW_K % RW > 0 && OUTW_E % RW > 0, "each noise stream ends in a partial block");
  __shared__ FrRing<P> ring[3];
  __shared__ __align__(16) uint8_t lut[32768];
  {
    const uint4* src = reinterpret_cast<const uint4*>(sample_lut_global<P::n>());
    uint4* dst = reinterpret_cast<uint4*>(lut);
    for (int i = threadIdx.x; i < 32768 / 16; i += 448) dst[i] = src[i];
    if (threadIdx.x < 3) {
      ring[threadIdx.x].ready = 0;
      ring[threadIdx.x].done = 0;
    }
  }
  __syncthreads();
  chain_priority();  // the samplers too: the chains wait on them for ring slots
  const int wave = threadIdx.x >> 6, vb = blockIdx.x, n = A.n;
  if (wave < 3) {
    const int lane = threadIdx.x & 63;
    const uint32_t p = lane & 1;
    const int inst = vb * 32 + (lane >> 1);
    if (inst < n) {
      uint64_t seed[LS];
      uint64_t dom = 0x96;
      int outw = OUTW_E;
      if (wave == 0) {
        const uint8_t* c = A.kc + (size_t)inst * P::kcoins + P::sec;
#pragma unroll
        for (int i = 0; i < LS; ++i) seed[i] = ld64(c + 8 * i);
        dom = 0x5F;
        outw = OUTW_K;
      } else if (wave == 1) {
        encdec_seed<P, false>(inst, p, A.pk_e, (size_t)P::pk, A.ec, (size_t)P::ecoins, A.kbuf_e, nullptr, seed);
      } else {
        encdec_seed<P, true>(inst, p, A.sk_d, (size_t)P::sk, A.mubuf_d, 32, A.kbuf_d, nullptr, seed);
      }
      shake_seed_squeeze2_ring_rt<P>(p, dom, seed, &ring[wave], lane, outw);
    }
    if (wave == 0) ss_body<P>(vb, n, A.ct_d, A.kbuf_ed, nullptr, 0, nullptr, A.ss_ed);
  } else if (wave == 3) {
    const int lane = threadIdx.x & 63;
    const uint32_t p = lane & 1;
    const int inst = vb * 32 + (lane >> 1);
    if (inst < n) {
      uint32_t a[25];
      ss_absorb<P>(a, p, A.ct_d + (size_t)inst * P::ct);
#pragma unroll
      for (int i = 0; i < 25; ++i) A.state_d[(size_t)i * 2 * n + 2 * inst + p] = a[i];
    }
    keygen_post_body<P>(vb, n, A.pk_e, A.sk_e);
  } else if (wave == 4) {
    encdec_pre_body<P, false, kFrSampler, false>(vb, n, A.pk_e, (size_t)P::pk, A.ec, (size_t)P::ecoins, A.kbuf_e,
                                                 A.s8_e, A.e8_e, A.epp8_e, nullptr, &ring[1], lut);
  } else if (wave == 5) {
    encdec_pre_body<P, true, kFrSampler, false>(vb, n, A.sk_d, (size_t)P::sk, A.mubuf_d, 32, A.kbuf_d, A.s8_d, A.e8_d,
                                                A.epp8_d, nullptr, &ring[2], lut);
  } else {
    keygen_pre_body<P, kFrSampler, false>(vb, n, A.kc, A.pk_k, A.sk_k, A.s8_k, A.e8_k, &ring[0], lut);
  }
}

// ----------------------------------------------------------------- host launchers

static inline int blocks_for(int n, int per) { return (n + per - 1) / per; }

// profiler names: FrodoKEM-976 keeps the plain kernel name, the others carry ",640" / ",1344"
template <class P>
static const char* fr_name(const char* n976, const char* n640, const char* n1344) {
  return P::n == 976 ? n976 : P::n == 640 ? n640 : n1344;
}
#define FR_NAME(base) fr_name<P>(base, base ",640", base ",1344")

// Resident workgroups of a kernel on the current device (0 on error).
template <class K>
static int resident_blocks(K kernel, int threads) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess)
    return 0;
  return cus * per_cu;
}

// Grid of a fused launch over `total` work items: one workgroup per item,
// or for the persistent form at most the resident workgroups (looked up
// once per kernel).
template <class P, int OP, bool AES, bool PERSIST, int TH>
static dim3 fused_grid(int total) {
  if (!PERSIST) return dim3(total);
  static const int res = OP == 0   ? resident_blocks(frodo_as_fused<P, AES, true, TH>, TH)
                         : OP == 1 ? resident_blocks(frodo_sa_fused<P, false, AES, true, TH>, TH)
                                   : resident_blocks(frodo_sa_fused<P, true, AES, true, TH>, TH);
  return dim3(res > 0 ? std::min(res, total) : total);
}

// A task table entry over one set of arrays, with the chains' scratch.
static FrMtTask task_of(int n, const uint8_t* i0, const uint8_t* i1, uint8_t* o0, uint8_t* o1, uint32_t* s8,
                        uint32_t* e8, uint32_t* epp8 = nullptr, uint8_t* kbuf = nullptr, uint8_t* mubuf = nullptr,
                        uint32_t* fail = nullptr, uint32_t* state = nullptr, size_t key_stride = 0) {
  FrMtTask t;
  memset(&t, 0, sizeof t);
  t.n = n;
  t.key_stride = key_stride;
  t.i0 = i0;
  t.i1 = i1;
  t.o0 = o0;
  t.o1 = o1;
  t.s8 = s8;
  t.e8 = e8;
  t.epp8 = epp8;
  t.kbuf = kbuf;
  t.mubuf = mubuf;
  t.fail = fail;
  t.state = state;
  return t;
}

static FrMt table_of(const FrMtTask* tasks, int ntasks) {
  FrMt M;
  memset(&M, 0, sizeof M);
  M.ntasks = ntasks;
  for (int k = 0; k < ntasks; ++k) {
    M.t[k] = tasks[k];
    M.t[k].first = M.total;
    M.t[k].first_group = M.groups;
    M.total += tasks[k].n;
    M.groups += blocks_for(tasks[k].n, 32);
  }
  return M;
}

template <class P, bool AES, int TH>
static void launch_as_fused_t(hipStream_t st, const FrMt& M, bool persist) {
  const char* name = AES ? "frodo_as_fused<aes>" : FR_NAME("frodo_as_fused");
  if (persist)
    DPKEM_LAUNCH(name, (frodo_as_fused<P, AES, true, TH>), (fused_grid<P, 0, AES, true, TH>(M.total)), dim3(TH), st, M);
  else
    DPKEM_LAUNCH(name, (frodo_as_fused<P, AES, false, TH>), (fused_grid<P, 0, AES, false, TH>(M.total)), dim3(TH), st,
                 M);
}
template <class P, bool DEC, bool AES, int TH>
static void launch_sa_fused_t(hipStream_t st, const FrMt& M, bool persist) {
  const char* name = DEC ? (AES ? "frodo_sa_fused<true,aes>" : FR_NAME("frodo_sa_fused<true>"))
                         : (AES ? "frodo_sa_fused<false,aes>" : FR_NAME("frodo_sa_fused<false>"));
  if (persist)
    DPKEM_LAUNCH(name, (frodo_sa_fused<P, DEC, AES, true, TH>), (fused_grid<P, DEC ? 2 : 1, AES, true, TH>(M.total)),
                 dim3(TH), st, M);
  else
    DPKEM_LAUNCH(name, (frodo_sa_fused<P, DEC, AES, false, TH>), (fused_grid<P, DEC ? 2 : 1, AES, false, TH>(M.total)),
                 dim3(TH), st, M);
}
// threads: 0 (or the set's own) = the set's default shape; kFrodoAltThreads
// selects FrodoKEM-976's alternative
template <class P, bool AES>
static void launch_as_fused(hipStream_t st, const FrMt& M, bool persist = false, int threads = 0) {
  if constexpr (P::n == 976) {
    if (threads == kFrodoAltThreads) return launch_as_fused_t<P, AES, kFrodoAltThreads>(st, M, persist);
  }
  launch_as_fused_t<P, AES, P::threads>(st, M, persist);
}
template <class P, bool DEC, bool AES>
static void launch_sa_fused(hipStream_t st, const FrMt& M, bool persist = false, int threads = 0) {
  if constexpr (P::n == 976) {
    if (threads == kFrodoAltThreads) return launch_sa_fused_t<P, DEC, AES, kFrodoAltThreads>(st, M, persist);
  }
  launch_sa_fused_t<P, DEC, AES, P::threads>(st, M, persist);
}
template <class P>
static void launch_dec_mu(hipStream_t st, const FrMt& M) {
  DPKEM_LAUNCH(FR_NAME("frodo_dec_mu"), frodo_dec_mu<P>, dim3(M.total), dim3(256), st, M);
}

template <class P, bool AES>
static void keygen_t(hipStream_t st, int n, const uint8_t* coins, uint8_t* pk, uint8_t* sk, uint32_t* s8, uint32_t* e8,
                     int threads) {
  DPKEM_LAUNCH(FR_NAME("frodo_keygen_pre"), frodo_keygen_pre<P>, dim3(blocks_for(n, 32)), dim3(128), st, n, coins, pk,
               sk, s8, e8);
  const FrMtTask t = task_of(n, coins, nullptr, pk, sk, s8, e8);
  launch_as_fused<P, AES>(st, table_of(&t, 1), false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_keygen_post"), frodo_keygen_post<P>, dim3(blocks_for(n, 32)), dim3(64), st, n, pk, sk);
}

template <class P, bool AES>
static void enc_t(hipStream_t st, int n, const uint8_t* pk, const uint8_t* coins, uint8_t* ct, uint8_t* ss, uint32_t* s8,
                  uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, const uint8_t* pkh, int threads) {
  DPKEM_LAUNCH(FR_NAME("frodo_encdec_pre<false>"), (frodo_encdec_pre<P, false>), dim3(blocks_for(n, 32)), dim3(128), st,
               n, pk, (size_t)P::pk, coins, (size_t)P::ecoins, kbuf, s8, e8, epp8, pkh);
  const FrMtTask t = task_of(n, pk, coins, ct, ss, s8, e8, epp8, kbuf);
  launch_sa_fused<P, false, AES>(st, table_of(&t, 1), false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_ss"), frodo_ss<P>, dim3(blocks_for(n, 32)), dim3(64), st, n, ct, kbuf, nullptr,
               (size_t)0, nullptr, ss);
}

template <class P, bool AES>
static void dec_t(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, uint8_t* ss, uint32_t* s8, uint32_t* e8,
                  uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail, uint32_t* state, int threads) {
  const FrMtTask t = task_of(n, ct, sk, ss, nullptr, s8, e8, epp8, kbuf, mubuf, fail, state, (size_t)P::sk);
  const FrMt M = table_of(&t, 1);
  launch_dec_mu<P>(st, M);
  DPKEM_LAUNCH(FR_NAME("frodo_dec_pre"), frodo_dec_pre<P>, dim3(blocks_for(n, 32)), dim3(192), st, n, ct, sk,
               (size_t)P::sk, mubuf, kbuf, s8, e8, epp8, state);
  launch_sa_fused<P, true, AES>(st, M, false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_pipe_ssd"), frodo_pipe_ssd<P>, dim3(blocks_for(n, 32)), dim3(64), st, n, ct, state, kbuf,
               sk, (size_t)P::sk, fail, ss, 0);
}

void frodo_keygen_launch(hipStream_t st, int n, const uint8_t* coins, uint8_t* pk, uint8_t* sk, uint32_t* s8,
                         uint32_t* e8, bool aes, int set, int threads) {
  if (n <= 0) return;
  if (set == 640) keygen_t<FrodoP<640>, false>(st, n, coins, pk, sk, s8, e8, threads);
  else if (set == 1344) keygen_t<FrodoP<1344>, false>(st, n, coins, pk, sk, s8, e8, threads);
  else if (aes) keygen_t<F976, true>(st, n, coins, pk, sk, s8, e8, threads);
  else keygen_t<F976, false>(st, n, coins, pk, sk, s8, e8, threads);
}

void frodo_enc_launch(hipStream_t st, int n, const uint8_t* pk, const uint8_t* coins, uint8_t* ct, uint8_t* ss,
                      uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, const uint8_t* pkh, bool aes, int set,
                      int threads) {
  if (n <= 0) return;
  if (set == 640) enc_t<FrodoP<640>, false>(st, n, pk, coins, ct, ss, s8, e8, epp8, kbuf, pkh, threads);
  else if (set == 1344) enc_t<FrodoP<1344>, false>(st, n, pk, coins, ct, ss, s8, e8, epp8, kbuf, pkh, threads);
  else if (aes) enc_t<F976, true>(st, n, pk, coins, ct, ss, s8, e8, epp8, kbuf, pkh, threads);
  else enc_t<F976, false>(st, n, pk, coins, ct, ss, s8, e8, epp8, kbuf, pkh, threads);
}

void frodo_dec_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, uint8_t* ss, uint32_t* s8,
                      uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail, uint32_t* state,
                      bool aes, int set, int threads) {
  if (n <= 0) return;
  if (set == 640) dec_t<FrodoP<640>, false>(st, n, ct, sk, ss, s8, e8, epp8, kbuf, mubuf, fail, state, threads);
  else if (set == 1344) dec_t<FrodoP<1344>, false>(st, n, ct, sk, ss, s8, e8, epp8, kbuf, mubuf, fail, state, threads);
  else if (aes) dec_t<F976, true>(st, n, ct, sk, ss, s8, e8, epp8, kbuf, mubuf, fail, state, threads);
  else dec_t<F976, false>(st, n, ct, sk, ss, s8, e8, epp8, kbuf, mubuf, fail, state, threads);
}


// Pipelined keygen -> encaps -> decaps of the same n instances on one stream
// (frodo_pipe_x1 / x2 / ssd).  Per instance the serial SHAKE path loses
// keygen's pkh (115 permutations for FrodoKEM-976), encaps' ss (116) and all
// but the last block of decaps' ss; ten launches become eight.  scratch:
// frodo_pipe_scratch_per_inst(set) bytes per instance.
size_t frodo_pipe_scratch_per_inst(int set) { return 16 * (size_t)set + 3 * 64 + 32 + 4 + 200; }

template <class P, bool AES>
static void pipe_t(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk, uint8_t* ct,
                   uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int threads, int force_reject) {
  size_t off = 0;
  auto take = [&](size_t per) {
    uint8_t* q = scratch + off;
    off += ((size_t)n * per + 255) & ~(size_t)255;
    return q;
  };
  uint32_t* s8 = reinterpret_cast<uint32_t*>(take(P::mat));  // keygen's, then encaps', then decaps'
  uint32_t* e8 = reinterpret_cast<uint32_t*>(take(P::mat));
  uint32_t* epp8 = reinterpret_cast<uint32_t*>(take(64));
  uint8_t* kbuf_e = take(64);
  uint8_t* kbuf_d = take(64);
  uint8_t* mubuf = take(32);
  uint32_t* fail = reinterpret_cast<uint32_t*>(take(4));
  uint32_t* state = reinterpret_cast<uint32_t*>(take(200));
  const dim3 g(blocks_for(n, 32));
  DPKEM_LAUNCH(FR_NAME("frodo_keygen_pre"), frodo_keygen_pre<P>, g, dim3(128), st, n, kc, pk, sk, s8, e8);
  const FrMtTask tk = task_of(n, kc, nullptr, pk, sk, s8, e8);
  launch_as_fused<P, AES>(st, table_of(&tk, 1), false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_pipe_x1"), frodo_pipe_x1<P>, g, dim3(192), st, n, pk, sk, ec, kbuf_e, s8, e8, epp8);
  const FrMtTask te = task_of(n, pk, ec, ct, ss, s8, e8, epp8, kbuf_e);
  launch_sa_fused<P, false, AES>(st, table_of(&te, 1), false, threads);
  const FrMtTask td = task_of(n, ct, sk, ss2, nullptr, s8, e8, epp8, kbuf_d, mubuf, fail, state, (size_t)P::sk);
  const FrMt Md = table_of(&td, 1);
  launch_dec_mu<P>(st, Md);
  DPKEM_LAUNCH(FR_NAME("frodo_pipe_x2"), frodo_pipe_x2<P>, g, dim3(256), st, n, ct, sk, kbuf_e, ss, mubuf, kbuf_d, s8,
               e8, epp8, state);
  launch_sa_fused<P, true, AES>(st, Md, false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_pipe_ssd"), frodo_pipe_ssd<P>, g, dim3(64), st, n, ct, state, kbuf_d, sk, (size_t)P::sk,
               fail, ss2, force_reject);
}

void frodo_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int set, bool aes, int threads,
                       int force_reject) {
  if (n <= 0) return;
  if (set == 640) pipe_t<FrodoP<640>, false>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, threads, force_reject);
  else if (set == 1344) pipe_t<FrodoP<1344>, false>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, threads, force_reject);
  else if (aes) pipe_t<F976, true>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, threads, force_reject);
  else pipe_t<F976, false>(st, n, kc, ec, pk, sk, ct, ss, ss2, scratch, threads, force_reject);
}

// The bounded LDS waits of this file's kernels (the noise chains' ring
// hand-overs, keccak2.hpp lds_wait_count) count their give-ups in this
// translation unit's own device counter, and take their bound from its own
// copy of the limit.
hipError_t frodo_wait_expired(uint32_t* count, bool clear) {
  hipError_t err = hipMemcpyFromSymbol(count, HIP_SYMBOL(k2::wait_expired), sizeof(uint32_t), 0, hipMemcpyDeviceToHost);
  if (err == hipSuccess && clear && *count) {
    const uint32_t zero = 0;
    err = hipMemcpyToSymbol(HIP_SYMBOL(k2::wait_expired), &zero, sizeof zero, 0, hipMemcpyHostToDevice);
  }
  return err;
}

hipError_t frodo_set_wait_spin_limit(uint32_t limit) {
  return hipMemcpyToSymbol(HIP_SYMBOL(k2::wait_spin_limit), &limit, sizeof limit, 0, hipMemcpyHostToDevice);
}

// Rolling pipeline: one slot's scratch is s8 | e8 | epp8 | kbuf_e | kbuf_d |
// mubuf | fail | state, each array 256-aligned.  A slot carries one op per
// step, so its s8 / e8 / epp8 serve keygen, encaps and decaps in turn;
// kbuf_e lives on from encaps' chain to the closing hash one step later.
size_t frodo_roll_scratch_bytes(int set, size_t n) { return n * frodo_pipe_scratch_per_inst(set) + 8 * 256; }

struct FrSlot {
  uint32_t *s8, *e8, *epp8;
  uint8_t *kbuf_e, *kbuf_d, *mubuf;
  uint32_t *fail, *state;
};

template <class P>
static FrSlot fr_slot(uint8_t* base, int n) {
  size_t off = 0;
  auto take = [&](size_t per) {
    uint8_t* q = base + off;
    off += ((size_t)n * per + 255) & ~(size_t)255;
    return q;
  };
  FrSlot s;
  s.s8 = reinterpret_cast<uint32_t*>(take(P::mat));
  s.e8 = reinterpret_cast<uint32_t*>(take(P::mat));
  s.epp8 = reinterpret_cast<uint32_t*>(take(64));
  s.kbuf_e = take(64);
  s.kbuf_d = take(64);
  s.mubuf = take(32);
  s.fail = reinterpret_cast<uint32_t*>(take(4));
  s.state = reinterpret_cast<uint32_t*>(take(200));
  return s;
}

// Step: decaps' mu' (slot r - 2), every chain (frodo_roll_x0), the keygen /
// encaps / re-encryption products, decaps' last ss block.
template <class P, bool AES>
static void roll_t(hipStream_t st, const RollDesc& d, int threads) {
  const int n = d.n;
  const FrSlot K = fr_slot<P>(d.scratch_k, n), E = fr_slot<P>(d.scratch_e, n), D = fr_slot<P>(d.scratch_d, n);
  const FrMtTask td =
      task_of(n, d.ct_d, d.sk_d, d.ss_d, nullptr, D.s8, D.e8, D.epp8, D.kbuf_d, D.mubuf, D.fail, D.state, (size_t)P::sk);
  const FrMt Md = table_of(&td, 1);
  launch_dec_mu<P>(st, Md);
  FrRollArgs a;
  a.n = n;
  a.kc = d.kc;
  a.pk_k = d.pk_k;
  a.sk_k = d.sk_k;
  a.s8_k = K.s8;
  a.e8_k = K.e8;
  a.pk_e = d.pk_e;
  a.ec = d.ec;
  a.sk_e = d.sk_e;
  a.kbuf_e = E.kbuf_e;
  a.s8_e = E.s8;
  a.e8_e = E.e8;
  a.epp8_e = E.epp8;
  a.ct_d = d.ct_d;
  a.sk_d = d.sk_d;
  a.mubuf_d = D.mubuf;
  a.kbuf_ed = D.kbuf_e;
  a.kbuf_d = D.kbuf_d;
  a.ss_ed = d.ss_e_prev;
  a.s8_d = D.s8;
  a.e8_d = D.e8;
  a.epp8_d = D.epp8;
  a.state_d = D.state;
  const dim3 g(blocks_for(n, 32));
  DPKEM_LAUNCH(FR_NAME("frodo_roll_x0"), frodo_roll_x0<P>, g, dim3(448), st, a);
  const FrMtTask tk = task_of(n, d.kc, nullptr, d.pk_k, d.sk_k, K.s8, K.e8);
  launch_as_fused<P, AES>(st, table_of(&tk, 1), false, threads);
  const FrMtTask te = task_of(n, d.pk_e, d.ec, d.ct_e, nullptr, E.s8, E.e8, E.epp8, E.kbuf_e);
  launch_sa_fused<P, false, AES>(st, table_of(&te, 1), false, threads);
  launch_sa_fused<P, true, AES>(st, Md, false, threads);
  DPKEM_LAUNCH(FR_NAME("frodo_pipe_ssd"), frodo_pipe_ssd<P>, g, dim3(64), st, n, d.ct_d, D.state, D.kbuf_d, d.sk_d,
               (size_t)P::sk, D.fail, d.ss_d, 0);
}

void frodo_roll_launch(hipStream_t st, int set, bool aes, const RollDesc& d, int threads) {
  if (d.n <= 0) return;
  if (set == 640) roll_t<FrodoP<640>, false>(st, d, threads);
  else if (set == 1344) roll_t<FrodoP<1344>, false>(st, d, threads);
  else if (aes) roll_t<F976, true>(st, d, threads);
  else roll_t<F976, false>(st, d, threads);
}

// Encaps / decaps where instances share public keys in groups (see
// frodo_sa_grouped).  pk_stride / sk_stride = 0: one key for all n.
void frodo_enc_grouped_launch(hipStream_t st, int n, const uint8_t* pk, size_t pk_stride, const uint8_t* coins,
                              uint8_t* ct, uint8_t* ss, uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf,
                              const uint2* groups, int ngroups, const uint8_t* pkh) {
  if (n <= 0) return;
  if (!groups) ngroups = blocks_for(n, kFrodoGroup);
  DPKEM_LAUNCH("frodo_encdec_pre<false>", (frodo_encdec_pre<F976, false>), dim3(blocks_for(n, 32)), dim3(128), st, n,
               pk, pk_stride, coins, (size_t)kFrodoEncCoins, kbuf, s8, e8, epp8, pkh);
  DPKEM_LAUNCH("frodo_sa_grouped<false>", (frodo_sa_grouped<false, kFrodoGroup>), dim3(ngroups), dim3(1024), st,
               ngroups, groups, n, pk, pk_stride, s8, reinterpret_cast<const int8_t*>(e8),
               reinterpret_cast<const int8_t*>(epp8), coins, (size_t)kFrodoEncCoins, ct, nullptr, nullptr);
  DPKEM_LAUNCH("frodo_ss", frodo_ss<F976>, dim3(blocks_for(n, 32)), dim3(64), st, n, ct, kbuf, nullptr, (size_t)0,
               nullptr, ss);
}

void frodo_dec_grouped_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, size_t sk_stride, uint8_t* ss,
                              uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail,
                              uint32_t* state, const uint2* groups, int ngroups) {
  if (n <= 0) return;
  if (!groups) ngroups = blocks_for(n, kFrodoGroup);
  {
    using P = F976;
    const FrMtTask t = task_of(n, ct, sk, ss, nullptr, s8, e8, epp8, kbuf, mubuf, fail, state, sk_stride);
    launch_dec_mu<P>(st, table_of(&t, 1));
  }
  DPKEM_LAUNCH("frodo_dec_pre", frodo_dec_pre<F976>, dim3(blocks_for(n, 32)), dim3(192), st, n, ct, sk, sk_stride,
               mubuf, kbuf, s8, e8, epp8, state);
  DPKEM_LAUNCH("frodo_sa_grouped<true>", (frodo_sa_grouped<true, kFrodoGroup>), dim3(ngroups), dim3(1024), st,
               ngroups, groups, n, sk + kFrodoSec, sk_stride, s8, reinterpret_cast<const int8_t*>(e8),
               reinterpret_cast<const int8_t*>(epp8), mubuf, (size_t)32, nullptr, ct, fail);
  DPKEM_LAUNCH("frodo_pipe_ssd", frodo_pipe_ssd<F976>, dim3(blocks_for(n, 32)), dim3(64), st, n, ct, state, kbuf, sk,
               sk_stride, fail, ss, 0);
}

// ----------------------------------------------------------------- multi-task launches
//
// Independent tasks of one op and parameter set in one launch sequence: the
// chain kernels map their 32-instance groups onto the tasks' groups (each
// task's instances padded to whole groups) and the fused / B'S kernels take
// the whole task table.

// keygen: seedA, S / E chains | enc: pkh, G2, S' / E' / E'' chains
template <class P, int OP>
__global__ void __launch_bounds__(128) frodo_mt_pre(const FrMt M) {
  const FrMtTask& T = M.t[mt_task_of_group(M, blockIdx.x)];
  const int vb = blockIdx.x - T.first_group;
  if constexpr (OP == 0) keygen_pre_split<P>(vb, T.n, T.i0, T.o0, T.o1, T.s8, T.e8);
  else encdec_pre_split<P, false>(vb, T.n, T.i0, (size_t)P::pk, T.i1, (size_t)P::ecoins, T.kbuf, T.s8, T.e8, T.epp8,
                                  nullptr);
}

// decaps: G2 and the S' / E' / E'' chains beside ss's ct absorb
template <class P>
__global__ void __launch_bounds__(192) frodo_mt_dec_pre(const FrMt M) {
  const FrMtTask& T = M.t[mt_task_of_group(M, blockIdx.x)];
  dec_pre_body<P>(blockIdx.x - T.first_group, T.n, T.i0, T.i1, (size_t)P::sk, T.mubuf, T.kbuf, T.s8, T.e8, T.epp8,
                  T.state);
}

// keygen: pkh into sk | enc: ss = SHAKE(ct || k) | dec: ss's last block
template <class P, int OP>
__global__ void __launch_bounds__(64) frodo_mt_post(const FrMt M) {
  chain_priority();
  const FrMtTask& T = M.t[mt_task_of_group(M, blockIdx.x)];
  const int vb = blockIdx.x - T.first_group;
  if constexpr (OP == 0) keygen_post_body<P>(vb, T.n, T.o0, T.o1);
  else if constexpr (OP == 1) ss_body<P>(vb, T.n, T.o0, T.kbuf, nullptr, 0, nullptr, T.o1);
  else ssd_body<P>(vb, T.n, T.i0, T.state, T.kbuf, T.i1, (size_t)P::sk, T.fail, T.o0, 0);
}

template <class P, bool AES>
static void mt_t(hipStream_t st, int op, const FrodoMtDesc* d, int ntasks, bool persist, int threads) {
  FrMtTask tasks[kFrodoMtTasks];
  for (int k = 0; k < ntasks; ++k) {
    const int n = d[k].n;
    // per-task scratch, carved as the single-task launchers carve it
    size_t off = 0;
    auto take = [&](size_t per) {
      uint8_t* q = d[k].scratch + off;
      off += ((size_t)n * per + 255) & ~(size_t)255;
      return q;
    };
    uint32_t* s8 = reinterpret_cast<uint32_t*>(take(P::mat));
    uint32_t* e8 = reinterpret_cast<uint32_t*>(take(P::mat));
    uint32_t* epp8 = reinterpret_cast<uint32_t*>(take(64));
    uint8_t* kbuf = take(64);
    uint8_t* mubuf = take(32);
    uint32_t* fail = reinterpret_cast<uint32_t*>(take(4));
    uint32_t* state = reinterpret_cast<uint32_t*>(take(200));
    tasks[k] = task_of(n, d[k].i0, d[k].i1, d[k].o0, d[k].o1, s8, e8, epp8, kbuf, mubuf, fail, state, (size_t)P::sk);
  }
  const FrMt M = table_of(tasks, ntasks);
  if (M.total == 0) return;
  const dim3 groups(M.groups);
  if (op == 0) {
    DPKEM_LAUNCH(FR_NAME("frodo_mt_pre<keygen>"), (frodo_mt_pre<P, 0>), groups, dim3(128), st, M);
    launch_as_fused<P, AES>(st, M, persist, threads);
    DPKEM_LAUNCH(FR_NAME("frodo_mt_post<keygen>"), (frodo_mt_post<P, 0>), groups, dim3(64), st, M);
  } else if (op == 1) {
    DPKEM_LAUNCH(FR_NAME("frodo_mt_pre<enc>"), (frodo_mt_pre<P, 1>), groups, dim3(128), st, M);
    launch_sa_fused<P, false, AES>(st, M, persist, threads);
    DPKEM_LAUNCH(FR_NAME("frodo_mt_post<enc>"), (frodo_mt_post<P, 1>), groups, dim3(64), st, M);
  } else {
    launch_dec_mu<P>(st, M);
    DPKEM_LAUNCH(FR_NAME("frodo_mt_dec_pre"), frodo_mt_dec_pre<P>, groups, dim3(192), st, M);
    launch_sa_fused<P, true, AES>(st, M, persist, threads);
    DPKEM_LAUNCH(FR_NAME("frodo_mt_post<dec>"), (frodo_mt_post<P, 2>), groups, dim3(64), st, M);
  }
}

void frodo_mt_launch(hipStream_t st, int op, int set, bool aes, const FrodoMtDesc* tasks, int ntasks, bool persist,
                     int threads) {
  if (ntasks <= 0 || ntasks > kFrodoMtTasks) return;
  if (set == 640) mt_t<FrodoP<640>, false>(st, op, tasks, ntasks, persist, threads);
  else if (set == 1344) mt_t<FrodoP<1344>, false>(st, op, tasks, ntasks, persist, threads);
  else if (aes) mt_t<F976, true>(st, op, tasks, ntasks, persist, threads);
  else mt_t<F976, false>(st, op, tasks, ntasks, persist, threads);
}

}  // namespace dpkem
