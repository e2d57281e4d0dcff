// Host-side launcher declarations shared by the engine and the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dpkem {

// ---- batched form (saber.hip, frodo.hip)
// Saber family: l = 2 (LightSaber), 3 (Saber), 4 (FireSaber).  One op of n
// instances runs as a mixed sequence (saber_mix_launch below) over
// saber_scratch_per_inst(l) bytes of scratch per instance.
size_t saber_scratch_per_inst(int l);

// Bounded LDS waits of the Saber-family kernels (keccak2.hpp lds_wait_count)
// on the current device: how many gave up since the last clear, and their
// spin bound.
hipError_t saber_wait_expired(uint32_t* count, bool clear);
// One wave of `perms` back-to-back lane-pair permutations (latency probe);
// out: 64 words.
void saber_probe_perm_launch(hipStream_t st, int perms, uint32_t* out);
hipError_t saber_set_wait_spin_limit(uint32_t limit);
// The same for the FrodoKEM kernels (frodo.hip: the noise chains' ring
// hand-overs); each kernel file keeps its own counter and bound.
hipError_t frodo_wait_expired(uint32_t* count, bool clear);
hipError_t frodo_set_wait_spin_limit(uint32_t limit);

// dst[i] = src for n rows of `bytes` bytes (a multiple of 4).
void broadcast_rows_launch(hipStream_t st, int n, const uint8_t* src, size_t bytes, uint8_t* dst);
// Chained Saber keygen -> enc -> dec of the same instances on one stream,
// sharing keygen's expanded matrix (saber.hip); scratch:
// kSaberChainScratchPerInst bytes per instance.
void saber_chain_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                        uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int ipw = 1);
constexpr size_t kSaberChainScratchPerInst = 3744 + 768 + 32 + 64 + 4;
// Pipelined keygen -> encaps -> decaps of the same instances on one stream
// (saber.hip): stages of different ops that read disjoint inputs share
// launches, every op computes all of its own hashes.  Saber family, rank l.
size_t saber_pipe_scratch_per_inst(int l);
// force_reject (test facility): every decapsulation of the chain takes the
// implicit-rejection branch.
void saber_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int l, int ipw = 1,
                       int force_reject = 0);
// Independent keygen (nk instances), encaps (ne) and decaps (nd) tasks of one
// rank l as one sequence of three launches (saber.hip, sab_mix_*); any count
// may be 0.  Arrays in the roles of the single-op launchers.
struct SaberMixDesc {
  int nk, ne, nd;
  const uint8_t* kc;
  uint8_t *pk_k, *sk_k;
  const uint8_t *pk_e, *ec;
  uint8_t *ct_e, *ss_e;
  const uint8_t *ct_d, *sk_d;
  uint8_t* ss_d;
};
size_t saber_mix_scratch_bytes(int l, size_t nk, size_t ne, size_t nd);
void saber_mix_launch(hipStream_t st, int l, const SaberMixDesc& d, uint8_t* scratch, int ipw = 1);
// One step of a rolling keygen -> encaps -> decaps pipeline over batch slots
// of n instances: keygen of slot r, encaps of slot r - 1 (to the keys keygen
// wrote one step earlier), decaps of slot r - 2 (of the ciphertexts encaps
// wrote one step earlier), and the closing hashes of the previous steps:
// keygen's pkh into slot r - 1's sk, encaps' shared secret into slot r - 2's
// ss.  Each slot has scratch of its own (saber_mix_scratch_bytes(l, n, n, n)
// or frodo_roll_scratch_bytes(set, n) bytes): scratch_k / _e / _d are those
// of slots r, r - 1, r - 2.
struct RollDesc {
  int n;
  const uint8_t* kc;  // slot r
  uint8_t *pk_k, *sk_k;
  const uint8_t *pk_e, *ec;  // slot r - 1
  uint8_t *ct_e, *sk_e;
  const uint8_t *ct_d, *sk_d;  // slot r - 2
  uint8_t *ss_d, *ss_e_prev;   // decaps' ss, and encaps' ss of the same slot
  uint8_t *scratch_k, *scratch_e, *scratch_d;
};
// Saber family (saber.hip: sab_roll_x0, then the mixed products).
void saber_roll_launch(hipStream_t st, int l, const RollDesc& d, int ipw = 1);
// Overlapped rolling step in one launch (saber.hip: sab_roll_fused): the
// chains of step `hash` and the products of step `prod` (either may be
// null), the products four instances per workgroup.
void saber_roll_fused_launch(hipStream_t st, int l, const RollDesc* hash, const RollDesc* prod);

// set: 976 (FrodoKEM-976), 640 or 1344 (Frodo-640 / Frodo-1344, SHAKE);
// aes: FrodoKEM-976-AES (A rows from AES-128) instead of -SHAKE.  Scratch
// per instance: s8 / e8 n * 8 bytes each (frodo_scratch_per_inst).
// threads: workgroup shape of the fused kernels (0 = the set's default,
// FrodoKEM-976 also 512: rows in two passes, two workgroups per CU).
void frodo_keygen_launch(hipStream_t st, int n, const uint8_t* coins, uint8_t* pk, uint8_t* sk, uint32_t* s8,
                         uint32_t* e8, bool aes = false, int set = 976, int threads = 0);
// pkh: nullptr (hash each pk) or one sec-byte pkh row per instance
void frodo_enc_launch(hipStream_t st, int n, const uint8_t* pk, const uint8_t* coins, uint8_t* ct, uint8_t* ss,
                      uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, const uint8_t* pkh = nullptr,
                      bool aes = false, int set = 976, int threads = 0);
// Pipelined keygen -> encaps -> decaps of the same instances on one stream
// (frodo.hip); set / aes as above.
size_t frodo_pipe_scratch_per_inst(int set);
void frodo_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int set, bool aes, int threads = 0,
                       int force_reject = 0);
// FrodoKEM (frodo.hip: frodo_dec_mu, frodo_roll_x0 -- every chain of the
// step in one launch --, the three fused products, frodo_pipe_ssd); set /
// aes / threads as above.
size_t frodo_roll_scratch_bytes(int set, size_t n);
void frodo_roll_launch(hipStream_t st, int set, bool aes, const RollDesc& d, int threads = 0);
// Groups of up to kFrodoGroup instances sharing one public key expand A once;
// groups[b] = (first, count) or nullptr for consecutive groups of
// kFrodoGroup; key strides may be 0 (one key for all n).
constexpr int kFrodoGroup = 4;
void frodo_enc_grouped_launch(hipStream_t st, int n, const uint8_t* pk, size_t pk_stride, const uint8_t* coins,
                              uint8_t* ct, uint8_t* ss, uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf,
                              const uint2* groups, int ngroups, const uint8_t* pkh = nullptr);
// state: 200 bytes per instance (decaps' ss sponge state after the ct absorb)
void frodo_dec_grouped_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, size_t sk_stride, uint8_t* ss,
                              uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail,
                              uint32_t* state, const uint2* groups, int ngroups);
void frodo_dec_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, uint8_t* ss, uint32_t* s8,
                      uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail, uint32_t* state,
                      bool aes = false, int set = 976, int threads = 0);

// Independent tasks of one Frodo op (0 keygen, 1 enc, 2 dec) and parameter
// set as ONE launch sequence (frodo.hip, multi-task launches): per task its
// instance count, arrays in the roles of the single-task launchers (keygen:
// i0 coins, o0 pk, o1 sk; enc: i0 pk, i1 coins, o0 ct, o1 ss; dec: i0 ct,
// i1 sk, o0 ss) and frodo_mt_scratch_bytes(set, n) bytes of scratch.
constexpr int kFrodoMtTasks = 16;
struct FrodoMtDesc {
  int n;
  const uint8_t* i0;
  const uint8_t* i1;
  uint8_t* o0;
  uint8_t* o1;
  uint8_t* scratch;
};
// persist: the fused kernel runs as a persistent grid of resident workgroups
// walking all tasks' instances, instead of one workgroup per instance.
void frodo_mt_launch(hipStream_t st, int op, int set, bool aes, const FrodoMtDesc* tasks, int ntasks,
                     bool persist = true, int threads = 0);

// ---- plain comparator form (plain.hip); op 0 keygen, 1 enc, 2 dec with the
// same output / input roles as the batched launchers.
size_t plain_scratch_per_inst(int scheme);
void plain_saber_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch);
void plain_frodo_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch);

// Scratch bytes per instance of a FrodoKEM op in the batched form:
// s8 (8 n) | e8 (8 n) | epp8 64 | kbuf 64 | mubuf 32 | fail 4 | state 200 (+ padding)
constexpr size_t frodo_scratch_per_inst(int set) { return 16 * (size_t)set + 64 + 64 + 32 + 16 + 200; }
// scratch of one multi-task entry of n instances (seven arrays, each 256-aligned)
constexpr size_t frodo_mt_scratch_bytes(int set, size_t n) { return n * frodo_scratch_per_inst(set) + 7 * 256; }

}  // namespace dpkem
