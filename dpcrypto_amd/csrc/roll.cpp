// Rolling keygen -> encaps -> decaps pipelines (dpkem_roll_*): three batch
// slots of one scheme; one step is, for the Saber family, sab_roll_x0 + the
// mixed products (saber_roll_launch), for FrodoKEM decaps' mu', every chain
// in one launch, the three fused products and decaps' last ss block
// (frodo_roll_launch).  Each step mod 3 is captured once into a HIP graph
// and replayed on the engine's stream.
//
// Overlapped form (Saber family, DPKEM_ROLL_OVERLAP, six slots): step k
// generates the keys of slot k, encapsulates to slot k - 2 and decapsulates
// slot k - 4 (mod 6).  Step k's chains then read nothing that step k - 1's
// products write and write nothing they read (keygen's pk of slot k - 2 and
// the ciphertexts of slot k - 4 came from the products of step k - 2; step
// k - 1's products touch slots k - 1, k - 3 and k - 5), so one launch per
// step (sab_roll_fused) carries step k's chains and step k - 1's products:
// the chain waves (one per SIMD, latency-bound) share the chip with the
// product waves, on the engine's stream, with no second stream or event.
#include <hip/hip_runtime.h>
#include <stdio.h>

#include <mutex>
#include <new>

#include "engine_internal.hpp"

using namespace dpkem;

struct dpkem_roll {
  dpkem_engine* e = nullptr;
  int scheme = 0, l = 3, ipw = 1;
  int set = 0, threads = 0;  // FrodoKEM: parameter set (0 = Saber family), fused-kernel shape
  bool aes = false;
  size_t n = 0;
  struct Slot {
    const uint8_t *kc = nullptr, *ec = nullptr;
    uint8_t *pk = nullptr, *sk = nullptr, *ct = nullptr, *ss = nullptr, *ss2 = nullptr;
  } slot[6];
  bool bound[6] = {false, false, false, false, false, false};
  int nslots = 3;  // 6 in the overlapped form
  uint8_t* scratch = nullptr;
  size_t per_slot = 0;
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
  uint64_t steps = 0;
  unsigned flags = 0;
  bool built = false;
};

namespace {

int rank_of(int scheme) {
  return scheme == DPKEM_SABER ? 3 : scheme == DPKEM_LIGHTSABER ? 2 : scheme == DPKEM_FIRESABER ? 4 : 0;
}
int set_of(int scheme) {
  return scheme == DPKEM_FRODO976 || scheme == DPKEM_FRODO976_AES ? 976
         : scheme == DPKEM_FRODO640                              ? 640
         : scheme == DPKEM_FRODO1344                             ? 1344
                                                                 : 0;
}

void release(dpkem_roll* r) {
  r->built = false;
  for (int k = 0; k < 3; ++k) {
    if (r->exec[k]) (void)hipGraphExecDestroy(r->exec[k]);
    if (r->graph[k]) (void)hipGraphDestroy(r->graph[k]);
    r->exec[k] = nullptr;
    r->graph[k] = nullptr;
  }
  if (r->scratch) (void)hipFree(r->scratch);
  r->scratch = nullptr;
}

// Arrays of step `phase`: keygen of slot phase, encaps of slot phase - 1 and
// decaps of slot phase - 2 (mod 3), or in the overlapped form slots phase,
// phase - 2, phase - 4 (mod 6).
RollDesc step_desc(const dpkem_roll* r, int phase) {
  const int S = r->nslots, lag = S == 6 ? 2 : 1;
  const int ks = phase, es = (phase + S - lag) % S, ds = (phase + S - 2 * lag) % S;
  const auto &K = r->slot[ks], &E = r->slot[es], &Dd = r->slot[ds];
  RollDesc d{};
  d.n = (int)r->n;
  d.kc = K.kc;
  d.pk_k = K.pk;
  d.sk_k = K.sk;
  d.pk_e = E.pk;
  d.ec = E.ec;
  d.ct_e = E.ct;
  d.sk_e = E.sk;
  d.ct_d = Dd.ct;
  d.sk_d = Dd.sk;
  d.ss_d = Dd.ss2;
  d.ss_e_prev = Dd.ss;
  d.scratch_k = r->scratch + (size_t)ks * r->per_slot;
  d.scratch_e = r->scratch + (size_t)es * r->per_slot;
  d.scratch_d = r->scratch + (size_t)ds * r->per_slot;
  return d;
}

void issue(dpkem_roll* r, hipStream_t st, int phase) {
  const RollDesc d = step_desc(r, phase);
  if (r->set) frodo_roll_launch(st, r->set, r->aes, d, r->threads);
  else saber_roll_launch(st, r->l, d, r->ipw);
}

}  // namespace

extern "C" {

int dpkem_roll_create(dpkem_engine* e, int scheme, size_t n, dpkem_roll** out) {
  if (!e || !out || (!rank_of(scheme) && !set_of(scheme)) || n == 0 || n > (size_t)(1 << 24) || e->form != DPKEM_FORM_BATCHED)
    return DPKEM_EINVAL;
  dpkem_roll* r = new (std::nothrow) dpkem_roll();
  if (!r) return DPKEM_ENOMEM;
  r->e = e;
  r->scheme = scheme;
  r->l = rank_of(scheme);
  r->set = set_of(scheme);
  r->aes = scheme == DPKEM_FRODO976_AES;
  r->n = n;
  *out = r;
  return DPKEM_OK;
}

int dpkem_roll_bind(dpkem_roll* r, int slot, const uint8_t* kcoins, const uint8_t* ecoins, uint8_t* pk, uint8_t* sk,
                    uint8_t* ct, uint8_t* ss, uint8_t* ss2) {
  if (!r || r->built || slot < 0 || slot > 5) return DPKEM_EINVAL;
  const void* ptrs[7] = {kcoins, ecoins, pk, sk, ct, ss, ss2};
  for (const void* q : ptrs)
    if (!q || ((uintptr_t)q & 7)) return DPKEM_EINVAL;
  r->slot[slot] = {kcoins, ecoins, pk, sk, ct, ss, ss2};
  r->bound[slot] = true;
  return DPKEM_OK;
}

int dpkem_roll_build(dpkem_roll* r, unsigned flags) {
  if (!r || (flags & ~(DPKEM_PLAN_NO_GRAPH | DPKEM_ROLL_OVERLAP))) return DPKEM_EINVAL;
  const bool overlap = (flags & DPKEM_ROLL_OVERLAP) != 0;
  if (overlap && r->set) return DPKEM_EINVAL;  // Saber family only
  for (int k = 0; k < (overlap ? 6 : 3); ++k)
    if (!r->bound[k]) return DPKEM_EINVAL;
  dpkem_engine* e = r->e;
  std::lock_guard<std::mutex> lock(e->mu);
  release(r);
  r->flags = flags;
  r->nslots = overlap ? 6 : 3;
  r->ipw = e->saber_ipw;
  r->threads = e->frodo_threads;
  r->per_slot = ((r->set ? frodo_roll_scratch_bytes(r->set, r->n) : saber_mix_scratch_bytes(r->l, r->n, r->n, r->n)) +
                 255) &
                ~(size_t)255;
  if (r->nslots * r->per_slot > e->budget) {
    snprintf(e->last_err, sizeof e->last_err, "rolling pipeline scratch %zu B exceeds HBM budget %zu B",
             r->nslots * r->per_slot, e->budget);
    return DPKEM_EBUDGET;
  }
  hipError_t err;
#define ROLL_TRY(call)                          \
  do {                                          \
    err = (call);                               \
    if (err != hipSuccess) {                    \
      int rc = record_hip_error(e, err, #call); \
      release(r);                               \
      return rc;                                \
    }                                           \
  } while (0)
  ROLL_TRY(hipSetDevice(e->device));
  ROLL_TRY(hipMalloc(&r->scratch, r->nslots * r->per_slot));
  if (overlap) {  // one launch per step, issued directly on the engine's stream
  } else if (!(flags & DPKEM_PLAN_NO_GRAPH)) {
    for (int ph = 0; ph < 3; ++ph) {
      hipStream_t cap;
      ROLL_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
      if (err != hipSuccess) {
        (void)hipStreamDestroy(cap);
        int rc = record_hip_error(e, err, "hipStreamBeginCapture");
        release(r);
        return rc;
      }
      Profiler* saved = tls_profiler;
      tls_profiler = nullptr;  // no timing events inside the graph
      issue(r, cap, ph);
      tls_profiler = saved;
      const hipError_t lerr = hipGetLastError();
      err = hipStreamEndCapture(cap, &r->graph[ph]);
      (void)hipStreamDestroy(cap);
      if (err != hipSuccess || lerr != hipSuccess) {
        int rc = record_hip_error(e, err != hipSuccess ? err : lerr, "rolling step capture");
        release(r);
        return rc;
      }
      ROLL_TRY(hipGraphInstantiate(&r->exec[ph], r->graph[ph], nullptr, nullptr, 0));
    }
  }
#undef ROLL_TRY
  r->steps = 0;
  r->built = true;
  return DPKEM_OK;
}

int dpkem_roll_step(dpkem_roll* r) {
  if (!r || !r->built) return DPKEM_EINVAL;
  dpkem_engine* e = r->e;
  std::lock_guard<std::mutex> lock(e->mu);
  hipError_t err = hipSetDevice(e->device);
  if (err != hipSuccess) return record_hip_error(e, err, "hipSetDevice");
  const int phase = (int)(r->steps % (uint64_t)r->nslots);
  if (r->nslots == 6) {  // this step's chains beside the previous step's products
    const RollDesc cur = step_desc(r, phase), prev = step_desc(r, (phase + 5) % 6);
    tls_profiler = e->profiling ? &e->prof : nullptr;
    saber_roll_fused_launch(e->stream, r->l, &cur, r->steps > 0 ? &prev : nullptr);
    tls_profiler = nullptr;
    err = hipGetLastError();
  } else if (r->flags & DPKEM_PLAN_NO_GRAPH) {
    tls_profiler = e->profiling ? &e->prof : nullptr;
    issue(r, e->stream, phase);
    tls_profiler = nullptr;
    err = hipGetLastError();
  } else {
    err = hipGraphLaunch(r->exec[phase], e->stream);
  }
  if (err != hipSuccess) return record_hip_error(e, err, "rolling step");
  ++r->steps;
  return DPKEM_OK;
}

int dpkem_roll_info(dpkem_roll* r, uint64_t* steps, size_t* scratch_bytes) {
  if (!r) return DPKEM_EINVAL;
  if (steps) *steps = r->steps;
  if (scratch_bytes) *scratch_bytes = r->nslots * r->per_slot;
  return DPKEM_OK;
}

int dpkem_roll_destroy(dpkem_roll* r) {
  if (!r) return DPKEM_EINVAL;
  {
    std::lock_guard<std::mutex> lock(r->e->mu);
    (void)hipSetDevice(r->e->device);
    (void)hipStreamSynchronize(r->e->stream);
    release(r);
  }
  delete r;
  return DPKEM_OK;
}

}  // extern "C"
