/*
 * dpkem -- batched FrodoKEM-976-SHAKE and Saber (l=3) on AMD Instinct MI355X.
 *
 * C ABI of dpcrypto_amd/libdpkem.so.  Plain pointers and sizes only.
 *
 * All arrays are instance-major and contiguous (instance i of `pk` starts at
 * pk + i * pk_bytes), exactly as the reference lays out its batches
 * (FrodoKEM976-GPU/src/cuda_kernel.cu:30-46, SABER-GPU/src/cuda_kernel.cu:35-57).
 * Wire formats are the published round-3 byte layouts.
 *
 * Return codes: 0 on success, a negative DPKEM_E* code on error; nothing
 * aborts.  Decapsulation returns 0 on implicit rejection too (the shared
 * secret is then derived from the secret rejection key, as the schemes
 * specify).
 */
#ifndef DPKEM_H
#define DPKEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- constants */
#define DPKEM_FRODO976 0     /* FrodoKEM-976-SHAKE */
#define DPKEM_SABER 1        /* Saber, l = 3 */
#define DPKEM_FRODO976_AES 2 /* FrodoKEM-976-AES: as DPKEM_FRODO976 (same sizes and coins) with A rows from
                                AES-128 (named at FrodoKEM976-GPU/include/api.h:12, its AES declared at
                                FrodoKEM976-GPU/include/gen_a_mat_mul.cuh:6); engine, plans and KAT only,
                                batched form only */
#define DPKEM_LIGHTSABER 3   /* LightSaber (l = 2): pk 672, sk 1568, ct 736 B */
#define DPKEM_FIRESABER 4    /* FireSaber (l = 4): pk 1312, sk 3040, ct 1472 B
                                (both parameter sets at SABER-GPU/include/params.h:10-19, coins as Saber;
                                engine, plans and KAT only, batched form only) */
#define DPKEM_FRODO640 5     /* Frodo-640-SHAKE: pk 9616, sk 19888, ct 9720, ss 16 B; keygen coins
                                s(16) || seedSE(16) || z(16) = 48 B, encaps coins mu = 16 B */
#define DPKEM_FRODO1344 6    /* Frodo-1344-SHAKE: pk 21520, sk 43088, ct 21632, ss 32 B; keygen coins
                                s(32) || seedSE(32) || z(16) = 80 B, encaps coins mu = 32 B
                                (the round-3 FrodoKEM sets beside the reference's FrodoKEM-976,
                                FrodoKEM976-GPU/include/params.h:9-24; engine, plans and KAT only,
                                batched form only) */

#define DPKEM_MEM_HOST 0   /* pointers are host memory (pageable or pinned) */
#define DPKEM_MEM_DEVICE 1 /* pointers are device memory on the engine's GPU */

#define DPKEM_FORM_BATCHED 0 /* fused / MFMA / Toeplitz kernels (default) */
#define DPKEM_FORM_PLAIN 1   /* one-workgroup-per-KEM multi-launch comparator */

#define DPKEM_OK 0
#define DPKEM_EINVAL -1    /* bad argument (NULL pointer, unknown scheme, n < 0) */
#define DPKEM_ENOMEM -2    /* device or host allocation failed */
#define DPKEM_EDEVICE -3   /* HIP runtime error */
#define DPKEM_EBUDGET -4   /* request exceeds the engine's HBM budget */
#define DPKEM_EBUSY -5     /* serving queue full (try again) */
#define DPKEM_ENOTFOUND -6 /* unknown request / key handle */
#define DPKEM_ECANCELED -7 /* serving request dropped before it was admitted */

typedef struct dpkem_engine dpkem_engine;

/* ---------------------------------------------------------------- sizes */
/* Byte sizes per instance for `scheme`.  Any output pointer may be NULL.
 * keygen_coins: Frodo 64 = s || seedSE || z ; Saber 96 = seedA_raw || seed_s || z.
 * enc_coins:    Frodo 24 = mu ; Saber 32 = m_raw.
 * Replaces the size macros of FrodoKEM976-GPU/include/api.h:7-10 and
 * SABER-GPU/include/params.h:42-48. */
int dpkem_sizes(int scheme, size_t *pk, size_t *sk, size_t *ct, size_t *ss, size_t *keygen_coins,
                size_t *enc_coins);

/* ---------------------------------------------------------------- engine */
/* Bind an engine to HIP device `device` with an HBM budget for its scratch
 * (0 = 25% of free device memory).  One engine owns one HIP stream. */
int dpkem_engine_create(dpkem_engine **out, int device, size_t hbm_budget_bytes);
int dpkem_engine_destroy(dpkem_engine *e);
/* Run subsequent work on a caller-provided HIP stream (hipStream_t; NULL is
 * the null stream) when own == 0, or on the engine's own stream when own != 0.
 * Work already queued on the previous stream is waited for first. */
int dpkem_engine_set_stream(dpkem_engine *e, void *hip_stream, int own);
/* Select the kernel form used by subsequent calls (DPKEM_FORM_*). */
int dpkem_engine_set_form(dpkem_engine *e, int form);
/* Kernel tuning knobs: launch shapes only, never results (the autotuner in
 * dpcrypto_amd/autotune.py sweeps them with a plan's streams and chunks).
 * Plans take the engine's setting when they are created; a server
 * (dpkem_server_create, which has no engine) keeps the defaults.  No
 * counterpart in the reference, whose launch shapes are compile-time
 * constants (SABER-GPU/src/cuda_kernel.cu:82-87, DoDP4Av7 /
 * MatVecMul_gpu_shared at one workgroup per instance).
 * DPKEM_TUNE_SABER_POLY_IPW: instances per workgroup of the Saber-family
 * product kernel, one wave each: 1 (default), 2 or 4.
 * DPKEM_TUNE_FRODO_PERSIST: in a plan's multi-task Frodo launches
 * (DPKEM_PLAN_NO_MERGE), the fused A-generation + MFMA kernel runs as a
 * persistent grid of resident workgroups walking every task's instances (1,
 * default) or as one workgroup per instance (0).
 * DPKEM_TUNE_FRODO_THREADS: workgroup shape of the FrodoKEM-976 fused
 * kernels (SHAKE and AES): 512 lanes over the rows of A in two passes, two
 * workgroups per CU (default; also 0), or 1024 lanes with one row each, one
 * workgroup per CU; the other Frodo sets keep their shapes. */
#define DPKEM_TUNE_SABER_POLY_IPW 1
#define DPKEM_TUNE_FRODO_PERSIST 2
#define DPKEM_TUNE_FRODO_THREADS 3
int dpkem_engine_set_tuning(dpkem_engine *e, int knob, int value);
/* Block until all work submitted on the engine's stream has finished.
 * Returns DPKEM_EDEVICE (with a message in dpkem_engine_last_error) when a
 * bounded wait inside a kernel gave up since the last check: the results of
 * the launches since then are invalid.  Host-memory calls check the same
 * after their final copy, and dpkem_server_drain for the server. */
int dpkem_engine_sync(dpkem_engine *e);
/* Last HIP error string seen by this engine (static storage). */
const char *dpkem_engine_last_error(dpkem_engine *e);
/* Per-kernel timing: when on, every kernel launch is bracketed by HIP events
 * on the engine's stream.  kernel_stats waits for the stream, then writes a
 * JSON object {"kernel": [launches, total_ms], ...} into buf (reset != 0
 * clears the counters).  Replaces the reference's cudaEvent pairs around
 * whole drivers (FrodoKEM976-GPU/src/cuda_kernel.cu:55,89-92) and its
 * MEAS_MV / MEAS_IP phase events (SABER-GPU/src/cuda_kernel.cu:82-87, DoDP4Av7 / MatVecMul_gpu_shared
 * at one workgroup per instance). */
int dpkem_engine_set_profiling(dpkem_engine *e, int on);
int dpkem_engine_kernel_stats(dpkem_engine *e, char *buf, size_t len, int reset);
/* Latency bound of the serial hash chains, measured: one workgroup of one
 * wave runs `perms` back-to-back lane-pair Keccak-f[1600] permutations (the
 * form the chain kernels squeeze with, 32 states in lane pairs) on the
 * engine's stream; *ms receives the launch's HIP-event time.  Two calls with
 * different perms give the per-permutation latency of a lone wave, the floor
 * of a chain launch's critical path.  No counterpart in the reference; used
 * by the benchmark's roofline report.  perms in [1, 100000]. */
int dpkem_probe_perm_latency(dpkem_engine *e, int perms, float *ms);

/* ---------------------------------------------------------------- batched KEM
 * Explicit-coins entry points over arrays of n instances.  `mem` says where
 * the pointers live.  With DPKEM_MEM_DEVICE the call is asynchronous on the
 * engine's stream (call dpkem_engine_sync before reading results); with
 * DPKEM_MEM_HOST it stages through device scratch and returns when the
 * outputs are in host memory.
 *
 * Replace the reference's hidden-randomness batch drivers
 *   frodoKEMKeypair / frodoKEMEncap / frodoKEMDecap  FrodoKEM976-GPU/include/cuda_kernel.cuh:7-9
 *   saber_enc / saber_dec                            SABER-GPU/include/cuda_kernel.cuh:6-7
 * and the per-instance prototypes crypto_kem_keypair / enc / dec declared at
 * FrodoKEM976-GPU/include/api.h:14-18. */
int dpkem_keypair(dpkem_engine *e, int scheme, size_t n, uint8_t *pk, uint8_t *sk, const uint8_t *coins,
                  int mem);
int dpkem_enc(dpkem_engine *e, int scheme, size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk,
              const uint8_t *coins, int mem);
int dpkem_dec(dpkem_engine *e, int scheme, size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk,
              int mem);

/* n encapsulations to ONE public key (pk: one row) / n decapsulations with ONE
 * secret key (sk: one row); other arrays as in dpkem_enc / dpkem_dec.
 * Results equal dpkem_enc / dpkem_dec with the key repeated n times.  Frodo
 * expands the public matrix A once per group of 4 instances instead of once
 * per instance (A is ~97% of a Frodo op's Keccak work); Saber broadcasts the
 * key and runs its ordinary kernels.  Batched form only.  No counterpart in
 * the reference, whose drivers take one key per instance
 * (FrodoKEM976-GPU/src/cuda_kernel.cu:108-212). */
int dpkem_enc_shared_key(dpkem_engine *e, int scheme, size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk,
                         const uint8_t *coins, int mem);
int dpkem_dec_shared_key(dpkem_engine *e, int scheme, size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk,
                         int mem);

/* ---------------------------------------------------------------- plans (graph-captured task sets)
 * A plan records KEM tasks over device-resident arrays, optionally chained
 * instance-wise (task `after` produces what this task consumes, e.g.
 * keygen -> enc -> dec), splits each task into `chunks` instance ranges,
 * spreads them over `nstreams` internal streams and captures everything
 * into one HIP graph; dpkem_plan_launch replays it on the engine's stream.
 * Output / input roles per op are those of dpkem_keypair / enc / dec:
 *   KEYGEN: o0 = pk, o1 = sk, i0 = coins        ENC: o0 = ct, o1 = ss, i0 = pk, i1 = coins
 *   DEC:    o0 = ss, i0 = ct, i1 = sk
 * Replaces the reference's strictly sequential per-op launch sequences
 * (FrodoKEM976-GPU/src/cuda_kernel.cu:64-84,159-189,282-323). */
#define DPKEM_OP_KEYGEN 0
#define DPKEM_OP_ENC 1
#define DPKEM_OP_DEC 2
typedef struct dpkem_plan dpkem_plan;
int dpkem_plan_create(dpkem_engine *e, dpkem_plan **out, int nstreams, int chunks);
int dpkem_plan_add(dpkem_plan *p, int scheme, int op, size_t n, uint8_t *o0, uint8_t *o1, const uint8_t *i0,
                   const uint8_t *i1, int after, int *task_id);
/* Options (before dpkem_plan_build).  DPKEM_PLAN_SHARE_MATRIX: a Saber
 * keygen -> enc -> dec chain over the same arrays (enc reads keygen's pk,
 * dec reads enc's ct and keygen's sk) expands the public matrix once per
 * instance instead of three times; results are identical.  Off by default:
 * it applies only when one caller runs all three ops on the same instances
 * (self-tests, known-answer runs), never across parties. */
#define DPKEM_PLAN_SHARE_MATRIX 1u
/* By default a keygen -> enc -> dec chain of one scheme over the same arrays
 * runs pipelined: per chunk one launch sequence in which serial stages of
 * different ops that read disjoint inputs share launches (Saber family: the
 * three GenMatrix streams together, keygen's H(pk), encaps' K, decaps' H(ct);
 * Frodo: keygen's pkh beside encaps' chain, encaps' ss and decaps' ct absorb
 * beside decaps' chain).  Every op still computes all of its own hashes and
 * results are identical.  DPKEM_PLAN_NO_PIPELINE runs such chains as three
 * dependent tasks instead. */
#define DPKEM_PLAN_NO_PIPELINE 2u
/* DPKEM_PLAN_NO_GRAPH: dpkem_plan_launch issues the plan's launches directly
 * (same streams, events and kernels as the graph) instead of replaying a
 * captured graph, so that the engine's profiler (dpkem_engine_set_profiling)
 * can time each kernel of a plan.  Slower to launch; for measurement. */
#define DPKEM_PLAN_NO_GRAPH 4u
/* Independent Frodo tasks of one scheme and op (no `after`, not in a
 * keygen -> enc -> dec chain), e.g. uneven batches of a mixed workload, run
 * by default as ONE multi-task launch sequence -- all their chains in one
 * launch each and one persistent fused kernel walking every task's
 * instances -- instead of one sequence per task (groups of up to 16 tasks;
 * a task alone keeps the per-task path).  Results are identical.
 * DPKEM_PLAN_NO_MERGE keeps one launch sequence per task. */
#define DPKEM_PLAN_NO_MERGE 8u
/* DPKEM_PLAN_FORCE_REJECT (test facility): in pipelined keygen -> encaps ->
 * decaps chains every decapsulation takes the implicit-rejection branch, as
 * if its re-encryption had differed from the ciphertext (Saber family: ss =
 * SHA3-256(z || H(c)); Frodo: ss = SHAKE(c || s)), so that the constant-time
 * select of a chain that never sees a forged ciphertext can be checked.
 * No counterpart in the reference. */
#define DPKEM_PLAN_FORCE_REJECT 16u
int dpkem_plan_set_options(dpkem_plan *p, unsigned flags);
int dpkem_plan_build(dpkem_plan *p);
int dpkem_plan_launch(dpkem_plan *p);
int dpkem_plan_info(dpkem_plan *p, size_t *subtasks, size_t *scratch_bytes);
int dpkem_plan_destroy(dpkem_plan *p);

/* ---------------------------------------------------------------- rolling pipelines
 * A rolling keygen -> encaps -> decaps pipeline for a stream of batches of
 * one scheme (any DPKEM_* scheme): three batch
 * slots of n instances, each bound to device arrays (roles as in
 * dpkem_keypair / enc / dec: keygen coins, encaps coins, pk, sk, ct, encaps'
 * ss, decaps' ss2).  Step k (k = 0, 1, ...) generates the keys of slot
 * k mod 3, encapsulates to the keys of slot (k - 1) mod 3 (generated by step
 * k - 1) and decapsulates the ciphertexts of slot (k - 2) mod 3 (encapsulated
 * by step k - 1); it also finishes the previous steps' batches (keygen's
 * H(pk) into sk, encaps' K into ss) in the same launches.  A Saber-family
 * step is two kernel launches (all hash chains of the three ops, then all
 * products); a FrodoKEM step six (decaps' mu', all serial SHAKE chains of
 * the three ops and the closing hashes in one launch, the keygen, encaps and
 * re-encryption products, decaps' last ss block).
 * After step k (k >= 2) every output of slot (k - 2) mod 3 -- pk, sk, ct, ss,
 * ss2 -- is complete, and stays so until step k + 1 starts that slot's next
 * keygen; slots of steps before k = 2 hold garbage.  Each op recomputes
 * everything from its own inputs, exactly as dpkem_keypair / enc / dec do,
 * and the outputs are byte-identical to theirs.  Steps are graph replays (one
 * graph per step mod 3) on the engine's stream.  No counterpart in the
 * reference, whose drivers run one op on one batch at a time
 * (SABER-GPU/src/cuda_kernel.cu:19-299, FrodoKEM976-GPU/src/cuda_kernel.cu:21-335). */
typedef struct dpkem_roll dpkem_roll;
int dpkem_roll_create(dpkem_engine *e, int scheme, size_t n, dpkem_roll **out);
int dpkem_roll_bind(dpkem_roll *r, int slot, const uint8_t *kcoins, const uint8_t *ecoins, uint8_t *pk,
                    uint8_t *sk, uint8_t *ct, uint8_t *ss, uint8_t *ss2);
/* Overlapped form (DPKEM_ROLL_OVERLAP; Saber family; slots 0-5 bound): step
 * k generates the keys of slot k mod 6, encapsulates to slot (k - 2) mod 6
 * and decapsulates slot (k - 4) mod 6, so that step k's hash chains depend on
 * nothing step k - 1's products write; each step is one launch on the
 * engine's stream holding step k's chains and step k - 1's products (four
 * instances per workgroup; DPKEM_TUNE_SABER_POLY_IPW does not apply).  After
 * step k (k >= 5) every output of slot (k - 5) mod 6 is complete on the
 * engine's stream, until step k + 1 starts that slot's next keygen.
 * Launches are issued directly (no graph). */
#define DPKEM_ROLL_OVERLAP 256u
/* flags: 0, DPKEM_PLAN_NO_GRAPH (issue each step's launches directly, so
 * that the engine's profiler can time them), or DPKEM_ROLL_OVERLAP.  Takes the engine's current
 * DPKEM_TUNE_SABER_POLY_IPW (Saber family) or DPKEM_TUNE_FRODO_THREADS
 * (FrodoKEM) setting. */
int dpkem_roll_build(dpkem_roll *r, unsigned flags);
int dpkem_roll_step(dpkem_roll *r);
int dpkem_roll_info(dpkem_roll *r, uint64_t *steps, size_t *scratch_bytes);
int dpkem_roll_destroy(dpkem_roll *r);

/* ---------------------------------------------------------------- serving loop
 * Continuous batching over a device-resident key / ciphertext store.
 * Requests go in one at a time (or in arrays); a dispatcher thread groups
 * pending requests by (scheme, op) and admits the group with the oldest
 * request into any free in-flight lane, as a batch capped by that lane's
 * share of the HBM budget.  Lanes run on their own streams, so batches
 * admitted while others run overlap on the device.  Keygen returns a key
 * handle, encaps a ciphertext handle; decaps names both.  Only coins enter
 * and shared secrets leave host memory.  No counterpart in the reference,
 * whose drivers take one fixed batch per call
 * (FrodoKEM976-GPU/src/cuda_kernel.cu:21-335). */
typedef struct dpkem_server dpkem_server;
typedef struct {
  uint64_t id;          /* caller's id, echoed in the completion */
  int32_t scheme, op;   /* DPKEM_FRODO976 / DPKEM_SABER, DPKEM_OP_* */
  uint32_t key;         /* ENC, DEC: key handle from a KEYGEN completion */
  uint32_t ct;          /* DEC: ciphertext handle from an ENC completion */
  double t_arrival_us;  /* optional arrival time on dpkem_now_us()'s clock (0 = submit time) */
  uint8_t coins[96];    /* KEYGEN: keygen coins; ENC: encaps coins */
} dpkem_request;
typedef struct {
  uint64_t id;
  int32_t status;       /* DPKEM_OK or an error (ENOTFOUND: stale handle, EBUDGET: store full) */
  int32_t scheme, op;
  uint32_t key;         /* KEYGEN: new key handle */
  uint32_t ct;          /* ENC: new ciphertext handle */
  uint32_t reserved;
  double t_submit_us;   /* arrival (or submit) time */
  double t_admit_us;    /* batch launched on a lane (queue wait = t_admit_us - t_submit_us) */
  double t_done_us;     /* completion observed by the dispatcher */
  uint8_t ss[32];       /* ENC, DEC: shared secret (24 bytes for Frodo) */
} dpkem_completion;
typedef struct {
  int device;
  int form;                 /* DPKEM_FORM_BATCHED, or DPKEM_FORM_PLAIN for the comparator */
  int lanes;                /* in-flight batches (1..16) */
  size_t hbm_budget_bytes;  /* store + lanes (0 = 25% of free device memory) */
  size_t key_slots[2];      /* per scheme (index = DPKEM_FRODO976 / DPKEM_SABER) */
  size_t ct_slots[2];
  size_t max_batch[2];      /* per-batch cap (0 = as large as the budget allows); 1 = one request per launch */
  int reserve_lanes[2];     /* lanes only this scheme may use, so that a burst of long Frodo batches cannot
                               hold every lane while Saber requests wait (sum <= lanes; 0, 0 = all shared) */
  int group_by_key;         /* nonzero: Frodo encaps / decaps batches are ordered by key handle and members
                               sharing a key expand the public matrix once per group of 4 */
} dpkem_server_config;
double dpkem_now_us(void);
int dpkem_server_create(dpkem_server **out, const dpkem_server_config *cfg);
int dpkem_server_submit(dpkem_server *s, const dpkem_request *reqs, size_t n);
int dpkem_server_poll(dpkem_server *s, dpkem_completion *out, size_t cap, size_t *got);
int dpkem_server_drain(dpkem_server *s);
/* Drop every request not yet admitted to a lane; each gets a completion with
 * status DPKEM_ECANCELED.  Batches already in flight finish normally.
 * dpkem_server_destroy does this first. */
int dpkem_server_cancel(dpkem_server *s, size_t *cancelled);
/* kind: 0 = key handle, 1 = ciphertext handle */
int dpkem_server_release(dpkem_server *s, int scheme, int kind, uint32_t handle);
/* kind: 0 = pk, 1 = sk, 2 = ct; copies the stored bytes to host memory */
int dpkem_server_read(dpkem_server *s, int scheme, int kind, uint32_t handle, uint8_t *out);
int dpkem_server_stats(dpkem_server *s, char *buf, size_t len);
int dpkem_server_destroy(dpkem_server *s);

/* ---------------------------------------------------------------- allocation-free entry point
 * No counterpart in the reference (whose drivers cudaMalloc and cudaFree on
 * every call, FrodoKEM976-GPU/src/cuda_kernel.cu:30-46,96-104).  For callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  One keygen / encaps /
 * decaps (op = DPKEM_OP_*) of n instances of `scheme` in the batched form,
 * the same kernels and results as dpkem_keypair / dpkem_enc / dpkem_dec with
 * DPKEM_MEM_DEVICE, launched asynchronously on `hip_stream` (a hipStream_t;
 * NULL = the null stream) of GPU `device`.  Array roles as in dpkem_plan_add
 * (KEYGEN: o0 = pk, o1 = sk, i0 = coins; ENC: o0 = ct, o1 = ss, i0 = pk,
 * i1 = coins; DEC: o0 = ss, i0 = ct, i1 = sk), all device memory on `device`,
 * 8-byte aligned.  Kernel scratch comes from one static device array of
 * DPKEM_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); the call
 * makes no hipMalloc / hipFree.  n above what that array holds
 * ((DPKEM_NOALLOC_SCRATCH_BYTES - 4096) / per-instance scratch: Saber 13,796,
 * LightSaber 26,212, FireSaber 8,737, Frodo-976 4,160, Frodo-640 6,241,
 * Frodo-1344 3,048 instances) returns DPKEM_EBUDGET before anything runs;
 * split larger batches.  All calls on one GPU share its scratch array: calls
 * that overlap there (other streams or threads) race and must be serialised
 * by the caller.  The first call on a device also looks up the array's
 * device address there. */
#define DPKEM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int dpkem_run_noalloc(int device, void *hip_stream, int scheme, int op, size_t n, uint8_t *o0, uint8_t *o1,
                      const uint8_t *i0, const uint8_t *i1);

/* ---------------------------------------------------------------- randomness
 * PQClean-shaped wrappers draw coins from a randombytes callback.  The
 * default reads the OS entropy source (getrandom). */
typedef void (*dpkem_randombytes_fn)(uint8_t *out, size_t len, void *ctx);
int dpkem_set_randombytes(dpkem_randombytes_fn fn, void *ctx);

/* NIST SP 800-90A AES-256 CTR_DRBG without derivation function (the
 * generator behind the NIST PQC known-answer files) and KAT coin streams.
 * No counterpart in the reference, which hard-codes its randomness
 * (FrodoKEM976-GPU/src/cuda_kernel.cu:51, SABER-GPU/src/cuda_kernel.cu:62)
 * and only declares AES (FrodoKEM976-GPU/include/aes.h:59-67). */
typedef struct {
  uint8_t key[32];
  uint8_t v[16];
  uint64_t reseed_counter;
} dpkem_drbg;
int dpkem_aes256_ecb(const uint8_t key[32], const uint8_t *in, uint8_t *out, size_t nblocks);
int dpkem_drbg_init(dpkem_drbg *d, const uint8_t entropy[48], const uint8_t *personalization /* 48 B or NULL */);
int dpkem_drbg_bytes(dpkem_drbg *d, uint8_t *out, size_t len);
/* a dpkem_randombytes_fn over ctx = dpkem_drbg* (for dpkem_set_randombytes) */
void dpkem_randombytes_drbg(uint8_t *out, size_t len, void *ctx);
/* The 48-byte per-count seeds of a KAT file: the DRBG seeded with bytes
 * 0, 1, ..., 47, then one 48-byte draw per count. */
int dpkem_kat_seeds(size_t n, uint8_t *seeds);
/* Keygen and encaps coins of n KAT counts: per count, the DRBG seeded with
 * that count's seed, drawn with the scheme's randombytes call pattern
 * (FrodoKEM: 2 sec + 16 B, then sec B, sec = 16 / 24 / 32 for Frodo-640 / -976 / -1344;
 * Saber family: 3 x 32 B, then 32 B). */
int dpkem_kat_coins(int scheme, size_t n, const uint8_t *seeds, uint8_t *kcoins, uint8_t *ecoins);

/* crypto_kem_{keypair,enc,dec} over arrays of n instances (host memory),
 * on a process-wide default engine bound to device 0.
 * Replace PQCLEAN_FRODOKEM976*_crypto_kem_* (FrodoKEM976-GPU/include/api.h:14-18). */
int dpkem_frodo976_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_frodo976_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_frodo976_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_frodo976aes_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_frodo976aes_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_frodo976aes_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_lightsaber_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_lightsaber_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_lightsaber_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_firesaber_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_firesaber_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_firesaber_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_frodo640_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_frodo640_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_frodo640_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_frodo1344_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_frodo1344_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_frodo1344_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int dpkem_saber_crypto_kem_keypair(size_t n, uint8_t *pk, uint8_t *sk);
int dpkem_saber_crypto_kem_enc(size_t n, uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int dpkem_saber_crypto_kem_dec(size_t n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);

/* ---------------------------------------------------------------- reference-shaped drivers
 * Same signatures and batch size (768 instances, host arrays) as the
 * reference's benchmark drivers, so a caller of those can link this library
 * instead.  Randomness comes from the randombytes callback; unlike the
 * reference, every instance uses its own inputs, and saber_dec returns the
 * shared secrets in h_k.  `mode` 0 selects the batched form, 1 the plain form
 * (the reference's DP / INT32 switch, SABER-GPU/main.cu:25-33).
 * Keys are INPUTS here: saber_enc reads h_pk and saber_dec reads h_sk (h_pk
 * is unused: the public key is part of the secret key), so the caller fills
 * them first with dpkem_saber_crypto_kem_keypair.  The reference instead
 * overwrites them inside the drivers with its built-in test vector
 * (SABER-GPU/src/cuda_kernel.cu:59-60,198-199) and its main.cu passes them
 * uninitialised; a caller relinked against this library must add the
 * keypair call, or it encapsulates to whatever bytes the buffers hold.
 *   frodoKEMKeypair/Encap/Decap  FrodoKEM976-GPU/include/cuda_kernel.cuh:7-9 (BLOCK = 768, :2)
 *   saber_enc / saber_dec        SABER-GPU/include/cuda_kernel.cuh:6-7 (BATCH = 768, :1)
 * Return 0 on success (the reference returns void). */
#define DPKEM_REF_BATCH 768
int frodoKEMKeypair(uint8_t *h_pk, uint8_t *h_sk);
int frodoKEMEncap(uint8_t *ct, uint8_t *ss, const uint8_t *pk);
int frodoKEMDecap(uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
int saber_enc(uint8_t mode, uint8_t *h_k, uint8_t *h_pk, uint8_t *h_c);
int saber_dec(uint8_t mode, uint8_t *h_c, uint8_t *h_pk, uint8_t *h_sk, uint8_t *h_k);

#ifdef __cplusplus
}
#endif
#endif /* DPKEM_H */
