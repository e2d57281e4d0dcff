// Host harness for the plan bookkeeping (dpcrypto_amd/csrc/plan.cpp) and the
// engine's scratch carving (dpcrypto_amd/csrc/engine.cpp), built for the CPU
// against a fake HIP runtime (hip/hip_runtime.h beside this file) and run
// under AddressSanitizer / UBSan.
//
// The kernel launchers (launch.hpp) are replaced by stubs that do what the
// real ones do to memory on the host side of the interface: each one reads
// every input row and writes every output row of its n instances and writes
// the whole scratch region its real counterpart carves.  A plan that sizes or
// offsets a scratch region wrongly, hands a launcher a short array, reuses a
// destroyed stream or event, waits during a capture on an event recorded
// outside it, or ends a capture with unjoined work, therefore fails here (an
// ASan report or a fake-runtime abort).  Prints "OK" when every plan shape
// builds, replays and is destroyed cleanly.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../../dpcrypto_amd/csrc/engine_internal.hpp"

namespace {

struct Sz {
  size_t pk, sk, ct, ss, kc, ec;
};
const Sz kSz[7] = {{15632, 31296, 15744, 24, 64, 24}, {992, 2304, 1088, 32, 96, 32},
                   {15632, 31296, 15744, 24, 64, 24}, {672, 1568, 736, 32, 96, 32},
                   {1312, 3040, 1472, 32, 96, 32},    {9616, 19888, 9720, 16, 48, 16},
                   {21520, 43088, 21632, 32, 80, 32}};
volatile unsigned g_sink = 0;
long g_launches = 0;

void rd(const void* p, size_t n) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  unsigned s = 0;
  for (size_t i = 0; i < n; i += 61) s += b[i];
  if (n) s += b[n - 1];
  g_sink = g_sink + s;
}
void wr(void* p, size_t n) {
  if (n) memset(p, 0xA5, n);
}
void launch_on(hipStream_t st) {
  fakehip::op_on(st);
  ++g_launches;
}
// Scratch regions written since the last reset: the sub-tasks of one plan
// may run at the same time, so their scratch must not overlap.
std::vector<std::pair<uintptr_t, uintptr_t>> g_regions;
void scr(void* p, size_t n) {
  if (!n) return;
  const uintptr_t a = (uintptr_t)p, b = a + n;
  for (auto& r : g_regions)
    if (a < r.second && r.first < b) fakehip::fail("two launches of one plan share scratch");
  g_regions.push_back({a, b});
  wr(p, n);
}
int saber_l(int scheme) { return scheme == 3 ? 2 : scheme == 4 ? 4 : 3; }
int scheme_of_l(int l) { return l == 2 ? 3 : l == 4 ? 4 : 1; }
int frodo_scheme(int set) { return set == 640 ? 5 : set == 1344 ? 6 : 0; }
size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }
// bytes a launcher carves from `scratch` for arrays of per-instance sizes `per`
size_t carve(int n, std::initializer_list<size_t> per) {
  size_t off = 0;
  for (size_t p : per) off += a256((size_t)n * p);
  return off;
}

}  // namespace

namespace dpkem {

// ---- batched Saber family (per-array pointers carved by the engine)
size_t saber_scratch_per_inst(int l) {
  const size_t m = 416 * (size_t)l * l, s = 32 * (size_t)l * (l == 2 ? 10 : l == 3 ? 8 : 6);
  return m + s + 32 + 64 + 64;
}
void broadcast_rows_launch(hipStream_t st, int n, const uint8_t* src, size_t bytes, uint8_t* dst) {
  launch_on(st);
  rd(src, bytes);
  wr(dst, (size_t)n * bytes);
}
// chains over whole-chunk scratch regions (the launchers carve them themselves)
void saber_chain_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                        uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int) {
  const Sz& z = kSz[1];
  launch_on(st);
  rd(kc, n * z.kc);
  rd(ec, n * z.ec);
  wr(pk, n * z.pk);
  wr(sk, n * z.sk);
  wr(ct, n * z.ct);
  wr(ss, n * z.ss);
  wr(ss2, n * z.ss);
  scr(scratch, carve(n, {3744, 768, 32, 64, 4}));
}
size_t saber_pipe_scratch_per_inst(int l) {
  const size_t m = 416 * (size_t)l * l, s = 32 * (size_t)l * (l == 2 ? 10 : l == 3 ? 8 : 6);
  return 3 * m + 2 * s + 2 * 32 + 2 * 64 + 64;
}
void saber_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int l, int, int) {
  const Sz& z = kSz[scheme_of_l(l)];
  const size_t m = 416 * (size_t)l * l, s = 32 * (size_t)l * (l == 2 ? 10 : l == 3 ? 8 : 6);
  launch_on(st);
  rd(kc, n * z.kc);
  rd(ec, n * z.ec);
  wr(pk, n * z.pk);
  wr(sk, n * z.sk);
  wr(ct, n * z.ct);
  wr(ss, n * z.ss);
  wr(ss2, n * z.ss);
  scr(scratch, carve(n, {m, m, m, s, s, 32, 32, 64, 64, 64}));
}
size_t saber_mix_scratch_bytes(int l, size_t nk, size_t ne, size_t nd) {
  const size_t m = 416 * (size_t)l * l, s = 32 * (size_t)l * (l == 2 ? 10 : l == 3 ? 8 : 6);
  return a256(nk * m) + a256(nk * s) + a256(ne * m) + a256(ne * s) + a256(ne * 32) + a256(ne * 64) + a256(nd * m) +
         a256(nd * s) + a256(nd * 32) + a256(nd * 64) + a256(nd * 64);
}
void saber_mix_launch(hipStream_t st, int l, const SaberMixDesc& d, uint8_t* scratch, int) {
  const Sz& z = kSz[scheme_of_l(l)];
  if (d.nk < 0 || d.ne < 0 || d.nd < 0 || d.nk + d.ne + d.nd == 0) fakehip::fail("mixed launch with no instances");
  launch_on(st);
  rd(d.kc, d.nk * z.kc);
  wr(d.pk_k, d.nk * z.pk);
  wr(d.sk_k, d.nk * z.sk);
  rd(d.pk_e, d.ne * z.pk);
  rd(d.ec, d.ne * z.ec);
  wr(d.ct_e, d.ne * z.ct);
  wr(d.ss_e, d.ne * z.ss);
  rd(d.ct_d, d.nd * z.ct);
  rd(d.sk_d, d.nd * z.sk);
  wr(d.ss_d, d.nd * z.ss);
  scr(scratch, saber_mix_scratch_bytes(l, d.nk, d.ne, d.nd));
}
// A rolling step's accesses: its three slots' arrays in the roles of the
// three ops, and each slot's scratch of `per` bytes written in full.
void roll_access(hipStream_t st, const Sz& z, const RollDesc& d, size_t per) {
  const size_t n = (size_t)d.n;
  if (d.n <= 0) fakehip::fail("rolling step with no instances");
  launch_on(st);
  rd(d.kc, n * z.kc);
  wr(d.pk_k, n * z.pk);
  wr(d.sk_k, n * z.sk);
  rd(d.pk_e, n * z.pk);
  rd(d.ec, n * z.ec);
  wr(d.ct_e, n * z.ct);
  wr(d.sk_e, n * z.sk);
  rd(d.ct_d, n * z.ct);
  rd(d.sk_d, n * z.sk);
  wr(d.ss_d, n * z.ss);
  wr(d.ss_e_prev, n * z.ss);
  // the three slots' scratch: distinct regions (a step is one launch
  // sequence of its own: only its own regions are compared)
  g_regions.clear();
  scr(d.scratch_k, per);
  scr(d.scratch_e, per);
  scr(d.scratch_d, per);
}
void saber_roll_launch(hipStream_t st, int l, const RollDesc& d, int) {
  roll_access(st, kSz[scheme_of_l(l)], d, saber_mix_scratch_bytes(l, (size_t)d.n, (size_t)d.n, (size_t)d.n));
}
// One launch with a step's chains and another step's products: both steps'
// arrays and scratch, which must not overlap (they run at the same time).
void saber_roll_fused_launch(hipStream_t st, int l, const RollDesc* hash, const RollDesc* prod) {
  if (!hash && !prod) fakehip::fail("fused rolling launch with nothing to do");
  const size_t per = saber_mix_scratch_bytes(l, (size_t)(hash ? hash->n : prod->n), (size_t)(hash ? hash->n : prod->n),
                                             (size_t)(hash ? hash->n : prod->n));
  std::vector<std::pair<uintptr_t, uintptr_t>> mine;
  if (hash) {
    roll_access(st, kSz[scheme_of_l(l)], *hash, per);
    mine = g_regions;
  }
  if (prod) {
    roll_access(st, kSz[scheme_of_l(l)], *prod, per);
    for (auto& a : mine)
      for (auto& b : g_regions)
        if (a.first < b.second && b.first < a.second) fakehip::fail("a fused step's chains and products share scratch");
  }
}
// ---- Frodo
void frodo_keygen_launch(hipStream_t st, int n, const uint8_t* coins, uint8_t* pk, uint8_t* sk, uint32_t* s8,
                         uint32_t* e8, bool, int set, int) {
  const Sz& z = kSz[frodo_scheme(set)];
  launch_on(st);
  rd(coins, n * z.kc);
  wr(pk, n * z.pk);
  wr(sk, n * z.sk);
  scr(s8, (size_t)n * 8 * set);
  scr(e8, (size_t)n * 8 * set);
}
void frodo_enc_launch(hipStream_t st, int n, const uint8_t* pk, const uint8_t* coins, uint8_t* ct, uint8_t* ss,
                      uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, const uint8_t* pkh, bool, int set,
                      int) {
  const Sz& z = kSz[frodo_scheme(set)];
  launch_on(st);
  rd(pk, n * z.pk);
  rd(coins, n * z.ec);
  if (pkh) rd(pkh, n * z.ss);
  wr(ct, n * z.ct);
  wr(ss, n * z.ss);
  scr(s8, (size_t)n * 8 * set);
  scr(e8, (size_t)n * 8 * set);
  scr(epp8, (size_t)n * 64);
  scr(kbuf, (size_t)n * 64);
}
void frodo_dec_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, uint8_t* ss, uint32_t* s8,
                      uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail, uint32_t* state,
                      bool, int set, int) {
  const Sz& z = kSz[frodo_scheme(set)];
  launch_on(st);
  rd(ct, n * z.ct);
  rd(sk, n * z.sk);
  wr(ss, n * z.ss);
  scr(s8, (size_t)n * 8 * set);
  scr(e8, (size_t)n * 8 * set);
  scr(epp8, (size_t)n * 64);
  scr(kbuf, (size_t)n * 64);
  scr(mubuf, (size_t)n * 32);
  scr(fail, (size_t)n * 4);
  scr(state, (size_t)n * 200);
}
size_t frodo_pipe_scratch_per_inst(int set) { return 16 * (size_t)set + 3 * 64 + 32 + 4 + 200; }
void frodo_pipe_launch(hipStream_t st, int n, const uint8_t* kc, const uint8_t* ec, uint8_t* pk, uint8_t* sk,
                       uint8_t* ct, uint8_t* ss, uint8_t* ss2, uint8_t* scratch, int set, bool, int, int) {
  const Sz& z = kSz[frodo_scheme(set)];
  launch_on(st);
  rd(kc, n * z.kc);
  rd(ec, n * z.ec);
  wr(pk, n * z.pk);
  wr(sk, n * z.sk);
  wr(ct, n * z.ct);
  wr(ss, n * z.ss);
  wr(ss2, n * z.ss);
  scr(scratch, carve(n, {8 * (size_t)set, 8 * (size_t)set, 64, 64, 64, 32, 4, 200}));
}
size_t frodo_roll_scratch_bytes(int set, size_t n) { return n * frodo_pipe_scratch_per_inst(set) + 8 * 256; }
void frodo_roll_launch(hipStream_t st, int set, bool, const RollDesc& d, int) {
  roll_access(st, kSz[frodo_scheme(set)], d, frodo_roll_scratch_bytes(set, (size_t)d.n));
}
void frodo_enc_grouped_launch(hipStream_t st, int n, const uint8_t* pk, size_t pk_stride, const uint8_t* coins,
                              uint8_t* ct, uint8_t* ss, uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf,
                              const uint2*, int, const uint8_t*) {
  const Sz& z = kSz[0];
  launch_on(st);
  rd(pk, pk_stride ? n * pk_stride : z.pk);
  rd(coins, n * z.ec);
  wr(ct, n * z.ct);
  wr(ss, n * z.ss);
  scr(s8, (size_t)n * 7808);
  scr(e8, (size_t)n * 7808);
  scr(epp8, (size_t)n * 64);
  scr(kbuf, (size_t)n * 64);
}
void frodo_dec_grouped_launch(hipStream_t st, int n, const uint8_t* ct, const uint8_t* sk, size_t sk_stride, uint8_t* ss,
                              uint32_t* s8, uint32_t* e8, uint32_t* epp8, uint8_t* kbuf, uint8_t* mubuf, uint32_t* fail,
                              uint32_t* state, const uint2*, int) {
  const Sz& z = kSz[0];
  launch_on(st);
  rd(ct, n * z.ct);
  rd(sk, sk_stride ? n * sk_stride : z.sk);
  wr(ss, n * z.ss);
  scr(s8, (size_t)n * 7808);
  scr(e8, (size_t)n * 7808);
  scr(epp8, (size_t)n * 64);
  scr(kbuf, (size_t)n * 64);
  scr(mubuf, (size_t)n * 32);
  scr(fail, (size_t)n * 4);
  scr(state, (size_t)n * 200);
}
void frodo_mt_launch(hipStream_t st, int op, int set, bool, const FrodoMtDesc* d, int ntasks, bool, int) {
  if (ntasks < 1 || ntasks > kFrodoMtTasks) fakehip::fail("multi-task launch with a bad task count");
  const Sz& z = kSz[frodo_scheme(set)];
  launch_on(st);
  for (int k = 0; k < ntasks; ++k) {
    const size_t n = (size_t)d[k].n;
    if (op == 0) {
      rd(d[k].i0, n * z.kc);
      wr(d[k].o0, n * z.pk);
      wr(d[k].o1, n * z.sk);
    } else if (op == 1) {
      rd(d[k].i0, n * z.pk);
      rd(d[k].i1, n * z.ec);
      wr(d[k].o0, n * z.ct);
      wr(d[k].o1, n * z.ss);
    } else {
      rd(d[k].i0, n * z.ct);
      rd(d[k].i1, n * z.sk);
      wr(d[k].o0, n * z.ss);
    }
    scr(d[k].scratch, carve((int)n, {8 * (size_t)set, 8 * (size_t)set, 64, 64, 32, 4, 200}));
  }
}
// ---- plain comparator
size_t plain_scratch_per_inst(int scheme) { return scheme == 1 ? 16384 : 2 * 1905152 + 65536; }
void plain_saber_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch) {
  launch_on(st);
  (void)op, (void)o0, (void)o1, (void)i0, (void)i1;
  scr(scratch, (size_t)n * plain_scratch_per_inst(1));
}
void plain_frodo_launch(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
                        uint8_t* scratch) {
  launch_on(st);
  (void)op, (void)o0, (void)o1, (void)i0, (void)i1;
  scr(scratch, (size_t)n * plain_scratch_per_inst(0));
}
hipError_t saber_wait_expired(uint32_t* count, bool) {
  *count = 0;
  return hipSuccess;
}
hipError_t saber_set_wait_spin_limit(uint32_t) { return hipSuccess; }
hipError_t frodo_wait_expired(uint32_t* count, bool) {
  *count = 0;
  return hipSuccess;
}
hipError_t frodo_set_wait_spin_limit(uint32_t) { return hipSuccess; }
void saber_probe_perm_launch(hipStream_t st, int, uint32_t* out) {
  launch_on(st);
  wr(out, 64 * sizeof(uint32_t));
}

}  // namespace dpkem

namespace {

struct Arr {
  std::vector<uint8_t*> keep;
  uint8_t* get(size_t bytes) {
    // exactly `bytes` (rounded to 8 for the 8-byte alignment rule), so that
    // ASan sees any access past the array
    uint8_t* p = static_cast<uint8_t*>(aligned_alloc(8, (bytes + 7) & ~(size_t)7));
    memset(p, 1, bytes);
    keep.push_back(p);
    return p;
  }
  ~Arr() {
    for (auto p : keep) free(p);
  }
};

int check(int rc, const char* what) {
  if (rc != DPKEM_OK) {
    fprintf(stderr, "FAIL %s: rc %d\n", what, rc);
    exit(1);
  }
  return rc;
}

// keygen -> enc -> dec of one scheme over shared arrays (a chain)
void add_chain(dpkem_plan* p, Arr& A, int scheme, size_t n) {
  const Sz& z = kSz[scheme];
  uint8_t *kc = A.get(n * z.kc), *ec = A.get(n * z.ec), *pk = A.get(n * z.pk), *sk = A.get(n * z.sk);
  uint8_t *ct = A.get(n * z.ct), *ss = A.get(n * z.ss), *ss2 = A.get(n * z.ss);
  int t0, t1;
  check(dpkem_plan_add(p, scheme, DPKEM_OP_KEYGEN, n, pk, sk, kc, nullptr, -1, &t0), "add keygen");
  check(dpkem_plan_add(p, scheme, DPKEM_OP_ENC, n, ct, ss, pk, ec, t0, &t1), "add enc");
  check(dpkem_plan_add(p, scheme, DPKEM_OP_DEC, n, ss2, nullptr, ct, sk, t1, nullptr), "add dec");
}

int add_op(dpkem_plan* p, Arr& A, int scheme, int op, size_t n, int after = -1) {
  const Sz& z = kSz[scheme];
  int t = -1;
  if (op == DPKEM_OP_KEYGEN)
    check(dpkem_plan_add(p, scheme, op, n, A.get(n * z.pk), A.get(n * z.sk), A.get(n * z.kc), nullptr, after, &t),
          "add keygen");
  else if (op == DPKEM_OP_ENC)
    check(dpkem_plan_add(p, scheme, op, n, A.get(n * z.ct), A.get(n * z.ss), A.get(n * z.pk), A.get(n * z.ec), after,
                         &t),
          "add enc");
  else
    check(dpkem_plan_add(p, scheme, op, n, A.get(n * z.ss), nullptr, A.get(n * z.ct), A.get(n * z.sk), after, &t),
          "add dec");
  return t;
}

void run(dpkem_engine* e, int nstreams, int chunks, unsigned flags, int rebuilds,
         void (*fill)(dpkem_plan*, Arr&)) {
  Arr A;
  dpkem_plan* p = nullptr;
  check(dpkem_plan_create(e, &p, nstreams, chunks), "plan_create");
  fill(p, A);
  if (flags) check(dpkem_plan_set_options(p, flags), "set_options");
  for (int b = 0; b < rebuilds; ++b) {
    g_regions.clear();
    check(dpkem_plan_build(p), "plan_build");
    for (int r = 0; r < 3; ++r) {
      g_regions.clear();
      check(dpkem_plan_launch(p), "plan_launch");
    }
  }
  check(dpkem_plan_destroy(p), "plan_destroy");
}

}  // namespace

int main() {
  dpkem_engine* e = nullptr;
  check(dpkem_engine_create(&e, 0, (size_t)8 << 30), "engine_create");
  const unsigned kFlagSets[] = {0u, DPKEM_PLAN_NO_GRAPH, DPKEM_PLAN_NO_PIPELINE, DPKEM_PLAN_SHARE_MATRIX,
                                DPKEM_PLAN_NO_MERGE, DPKEM_PLAN_NO_MERGE | DPKEM_PLAN_NO_GRAPH,
                                DPKEM_PLAN_FORCE_REJECT};
  const int kSchedules[][2] = {{1, 1}, {2, 3}, {4, 4}, {8, 1}, {3, 5}, {16, 8}};
  int plans = 0;
  for (unsigned flags : kFlagSets) {
    for (const auto& sc : kSchedules) {
      // chains of every scheme, small and ragged
      run(e, sc[0], sc[1], flags, 1, [](dpkem_plan* p, Arr& A) {
        add_chain(p, A, DPKEM_SABER, 300);
        add_chain(p, A, DPKEM_SABER, 1);
        add_chain(p, A, DPKEM_LIGHTSABER, 33);
        add_chain(p, A, DPKEM_FIRESABER, 19);
        add_chain(p, A, DPKEM_FRODO976, 10);
        add_chain(p, A, DPKEM_FRODO976_AES, 3);
        add_chain(p, A, DPKEM_FRODO640, 2);
        add_chain(p, A, DPKEM_FRODO1344, 2);
      });
      // an uneven mixed workload: independent tasks of every scheme and op, 20 of
      // them Frodo-976 decaps (more than one merged group), plus dependants of
      // independent tasks with other sizes (joins on whole producers)
      run(e, sc[0], sc[1], flags, 2, [](dpkem_plan* p, Arr& A) {
        std::mt19937 rng(7);
        for (int k = 0; k < 40; ++k) {
          const int scheme = (int)(rng() % 7), op = (int)(rng() % 3);
          add_op(p, A, scheme, op, 1 + rng() % 70);
        }
        for (int k = 0; k < 20; ++k) add_op(p, A, DPKEM_FRODO976, DPKEM_OP_DEC, 1 + rng() % 9);
        const int t = add_op(p, A, DPKEM_FRODO976, DPKEM_OP_KEYGEN, 12);
        add_op(p, A, DPKEM_FRODO976, DPKEM_OP_ENC, 12, t);
        add_op(p, A, DPKEM_FRODO976, DPKEM_OP_ENC, 5, t);
        const int u = add_op(p, A, DPKEM_SABER, DPKEM_OP_KEYGEN, 40);
        add_op(p, A, DPKEM_SABER, DPKEM_OP_DEC, 40, u);
      });
      plans += 2;
    }
  }
  // rolling pipelines of every scheme, with and without graphs, several
  // steps and a rebuild
  for (int scheme : {DPKEM_SABER, DPKEM_LIGHTSABER, DPKEM_FIRESABER, DPKEM_FRODO976, DPKEM_FRODO976_AES,
                     DPKEM_FRODO640, DPKEM_FRODO1344}) {
    for (unsigned flags : {0u, DPKEM_PLAN_NO_GRAPH}) {
      for (size_t n : {(size_t)1, (size_t)37, (size_t)300}) {
        Arr A;
        const Sz& z = kSz[scheme];
        dpkem_roll* r = nullptr;
        check(dpkem_roll_create(e, scheme, n, &r), "roll_create");
        for (int k = 0; k < 3; ++k)
          check(dpkem_roll_bind(r, k, A.get(n * z.kc), A.get(n * z.ec), A.get(n * z.pk), A.get(n * z.sk),
                                A.get(n * z.ct), A.get(n * z.ss), A.get(n * z.ss)),
                "roll_bind");
        for (int b = 0; b < 2; ++b) {
          check(dpkem_roll_build(r, flags), "roll_build");
          for (int st = 0; st < 5; ++st) {
            g_regions.clear();
            check(dpkem_roll_step(r), "roll_step");
          }
        }
        check(dpkem_roll_destroy(r), "roll_destroy");
        ++plans;
      }
    }
  }
  // the overlapped form: six slots, hash launches and products on two
  // streams of the pipeline, every step joined by the engine's stream
  for (int scheme : {DPKEM_SABER, DPKEM_LIGHTSABER, DPKEM_FIRESABER}) {
    for (size_t n : {(size_t)1, (size_t)37, (size_t)300}) {
      Arr A;
      const Sz& z = kSz[scheme];
      dpkem_roll* r = nullptr;
      check(dpkem_roll_create(e, scheme, n, &r), "roll_create");
      for (int k = 0; k < 6; ++k)
        check(dpkem_roll_bind(r, k, A.get(n * z.kc), A.get(n * z.ec), A.get(n * z.pk), A.get(n * z.sk),
                              A.get(n * z.ct), A.get(n * z.ss), A.get(n * z.ss)),
              "roll_bind");
      for (int b = 0; b < 2; ++b) {
        check(dpkem_roll_build(r, DPKEM_ROLL_OVERLAP), "roll_build overlap");
        for (int st = 0; st < 13; ++st) check(dpkem_roll_step(r), "roll_step overlap");
      }
      check(dpkem_roll_destroy(r), "roll_destroy");
      ++plans;
    }
  }
  // the plain form (no chains, no merging)
  check(dpkem_engine_set_form(e, DPKEM_FORM_PLAIN), "set_form");
  run(e, 4, 2, 0, 1, [](dpkem_plan* p, Arr& A) {
    add_chain(p, A, DPKEM_SABER, 20);
    add_op(p, A, DPKEM_FRODO976, DPKEM_OP_DEC, 3);
    add_op(p, A, DPKEM_FRODO976, DPKEM_OP_DEC, 4);
  });
  check(dpkem_engine_destroy(e), "engine_destroy");
  printf("OK %d plans, %ld launches, %ld graphs\n", plans + 1, g_launches, fakehip::st().graphs);
  return 0;
}
