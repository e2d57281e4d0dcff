// Serving loop: continuous batching of keygen / encaps / decaps requests
// against a device-resident key and ciphertext store.
//
// The reference has no serving path at all (its drivers take one fixed
// batch of 768 instances per call, FrodoKEM976-GPU/src/cuda_kernel.cu:21-335,
// SABER-GPU/src/cuda_kernel.cu:19-299).  Here requests arrive one by one
// (dpkem_server_submit); a dispatcher thread groups pending requests by
// (scheme, op), and whenever one of `lanes` in-flight slots is free it admits
// the oldest group as a batch sized by the lane's share of the HBM budget.
// Each lane owns a HIP stream, device scratch and pinned staging, so batches
// admitted while others run overlap on the device.  Public keys, secret keys
// and ciphertexts never leave HBM: keygen returns a key handle, encaps a
// ciphertext handle; only coins go in and shared secrets come out.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <vector>

#include "engine_internal.hpp"
#include "launch.hpp"

using namespace dpkem;

namespace {

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SchemeDims {
  size_t pk, sk, ct, ss, kc, ec;
};
const SchemeDims kDims[2] = {{15632, 31296, 15744, 24, 64, 24}, {992, 2304, 1088, 32, 96, 32}};

// Copy `bytes` per instance between a slab (slot-indexed) and a dense batch.
// gather: dst[i] = slab[idx[i]];  scatter: slab[idx[i]] = src[i].
__global__ void __launch_bounds__(256) slab_copy(int n, const uint32_t* __restrict__ idx, const uint8_t* __restrict__ src,
                                                 uint8_t* __restrict__ dst, size_t bytes, int gather) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const size_t s = gather ? (size_t)idx[i] * bytes : (size_t)i * bytes;
  const size_t d = gather ? (size_t)i * bytes : (size_t)idx[i] * bytes;
  const uint32_t* sp = reinterpret_cast<const uint32_t*>(src + s);
  uint32_t* dp = reinterpret_cast<uint32_t*>(dst + d);
  for (size_t k = threadIdx.x; k < bytes / 4; k += 256) dp[k] = sp[k];
}

// dst[idx[i]] (24-byte rows) = the pkh field of dense Frodo sk row i.
__global__ void __launch_bounds__(64) pkh_scatter(int n, const uint32_t* __restrict__ idx, const uint8_t* __restrict__ sk,
                                                  uint8_t* __restrict__ pkh_slab) {
  const int i = blockIdx.x * 8 + threadIdx.x / 8, w = threadIdx.x % 8;  // 8 rows per 64-thread block
  if (i >= n || w >= 6) return;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(sk + (size_t)i * 31296 + 31272);
  reinterpret_cast<uint32_t*>(pkh_slab + (size_t)idx[i] * 24)[w] = src[w];
}

struct SlotPool {
  std::vector<uint32_t> free_list;
  std::vector<uint8_t> used;
  void init(size_t n) {
    free_list.resize(n);
    used.assign(n, 0);
    for (size_t i = 0; i < n; ++i) free_list[i] = (uint32_t)(n - 1 - i);
  }
  bool take(uint32_t* out) {
    if (free_list.empty()) return false;
    *out = free_list.back();
    free_list.pop_back();
    used[*out] = 1;
    return true;
  }
  bool give(uint32_t h) {
    if (h >= used.size() || !used[h]) return false;
    used[h] = 0;
    free_list.push_back(h);
    return true;
  }
  bool valid(uint32_t h) const { return h < used.size() && used[h]; }
};

struct Pending {
  dpkem_request r;
  double t_submit;
};

struct Lane {
  int only_scheme = -1;         // >= 0: reserved for this scheme
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint8_t* scratch = nullptr;   // device: kernel scratch + dense I/O + index arrays
  size_t scratch_bytes = 0;
  uint8_t* host = nullptr;      // pinned: coins / indices in, shared secrets out
  size_t host_bytes = 0;
  bool busy = false;
  int scheme = 0, op = 0;
  double t_admit = 0;
  std::vector<Pending> batch;
  std::vector<uint32_t> new_handles;
  uint8_t* host_ss = nullptr;
};

}  // namespace



struct dpkem_server {
  int device = 0;
  int form = DPKEM_FORM_BATCHED;
  bool group_by_key = false;
  size_t budget = 0;
  size_t max_batch[2] = {0, 0};
  // device-resident store
  uint8_t* pk_slab[2] = {nullptr, nullptr};
  uint8_t* sk_slab[2] = {nullptr, nullptr};
  uint8_t* ct_slab[2] = {nullptr, nullptr};
  uint8_t* pkh_slab = nullptr;  // Frodo: pkh = SHAKE256(pk) of every key, so encaps need not rehash pk
  SlotPool keys[2], cts[2];
  std::vector<Lane> lanes;
  // queues
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Pending> pending[2][3];
  size_t queue_cap = 1 << 20;
  std::vector<dpkem_completion> done;
  size_t inflight_requests = 0;
  std::condition_variable drained;
  std::atomic<bool> stop{false};
  std::thread worker;
  char last_err[256] = {0};
  // statistics
  uint64_t batches = 0, requests = 0, max_seen_batch = 0;
  uint64_t key_groups = 0, grouped_requests = 0;  // Frodo members launched in same-key groups
  uint64_t failed_batches = 0;
  // host-side fault injection for tests of the failure path (environment
  // variable DPKEM_SERVER_INJECT_FAULTS="L,E": the first L batch launches
  // report a launch error before any HIP call, and the completion events of
  // the next E batches read as failed); zero in normal use
  int inject_launch_faults = 0, inject_event_faults = 0;
};

namespace {

int init_lane(dpkem_server* s, Lane& L, size_t scratch_bytes, size_t host_bytes) {
  if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess) return DPKEM_EDEVICE;
  if (hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) return DPKEM_EDEVICE;
  if (hipMalloc(&L.scratch, scratch_bytes) != hipSuccess) return DPKEM_ENOMEM;
  if (hipHostMalloc(&L.host, host_bytes, hipHostMallocDefault) != hipSuccess) return DPKEM_ENOMEM;
  L.scratch_bytes = scratch_bytes;
  L.host_bytes = host_bytes;
  (void)s;
  return DPKEM_OK;
}

size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device bytes a lane needs for a batch of m instances of `scheme` (any op).
size_t lane_bytes(int scheme, int form, size_t m) {
  const SchemeDims& d = kDims[scheme];
  return chunk_scratch_bytes(scheme, form, m) + a256(m * d.pk) + a256(m * d.sk) + a256(m * d.ct) + a256(m * d.ss) +
         a256(m * 96) + 2 * a256(m * 4) + a256(m * 8) + a256(m * 24) + 4096;
}

// Launch one admitted batch on lane L (called with s->mu NOT held).  Returns
// the first HIP error of the launch sequence; *ngroups_out receives the
// number of same-key groups launched (0 when the batch is not grouped).
hipError_t launch_batch(dpkem_server* s, Lane& L, bool inject_fault, int* ngroups_out) {
  *ngroups_out = 0;
  if (inject_fault) return hipErrorLaunchFailure;
  const int sc = L.scheme, op = L.op;
  const SchemeDims& d = kDims[sc];
  const int m = (int)L.batch.size();
  // layout of the lane's device buffer
  uint8_t* base = L.scratch;
  size_t off = a256(chunk_scratch_bytes(sc, s->form, (size_t)m));
  auto take = [&](size_t b) { uint8_t* p = base + off; off += a256(b); return p; };
  uint8_t* d_pk = take((size_t)m * d.pk);
  uint8_t* d_sk = take((size_t)m * d.sk);
  uint8_t* d_ct = take((size_t)m * d.ct);
  uint8_t* d_ss = take((size_t)m * d.ss);
  uint8_t* d_coins = take((size_t)m * 96);
  uint32_t* d_idx0 = reinterpret_cast<uint32_t*>(take((size_t)m * 4));
  uint32_t* d_idx1 = reinterpret_cast<uint32_t*>(take((size_t)m * 4));
  uint2* d_groups = reinterpret_cast<uint2*>(take((size_t)m * 8));
  uint8_t* d_pkh = take((size_t)m * 24);
  // pinned staging: coins | idx0 | idx1 | groups | ss (the first four laid out as on the device)
  uint8_t* h = L.host;
  uint8_t* h_coins = h;
  uint32_t* h_idx0 = reinterpret_cast<uint32_t*>(h + a256((size_t)m * 96));
  uint32_t* h_idx1 = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(h_idx0) + a256((size_t)m * 4));
  uint2* h_groups = reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(h_idx1) + a256((size_t)m * 4));
  L.host_ss = reinterpret_cast<uint8_t*>(h_groups) + a256((size_t)m * 8);
  const size_t cw = op == kKeygen ? d.kc : d.ec;
  for (int i = 0; i < m; ++i) {
    const dpkem_request& r = L.batch[i].r;
    if (op != kDec) memcpy(h_coins + (size_t)i * cw, r.coins, cw);
    h_idx0[i] = op == kKeygen ? L.new_handles[i] : r.key;
    h_idx1[i] = op == kEnc ? L.new_handles[i] : r.ct;
  }
  // same-key groups of up to kFrodoGroup consecutive members (the batch was
  // ordered by key at admission)
  int ngroups = 0;
  const bool grouped = s->group_by_key && sc == DPKEM_FRODO976 && op != kKeygen;
  if (grouped) {
    for (int i = 0; i < m;) {
      int j = i + 1;
      while (j < m && j - i < kFrodoGroup && L.batch[j].r.key == L.batch[i].r.key) ++j;
      h_groups[ngroups++] = make_uint2((uint32_t)i, (uint32_t)(j - i));
      i = j;
    }
    *ngroups_out = ngroups;
  }
  hipStream_t st = L.stream;
  hipError_t first = hipSuccess;
  auto note = [&](hipError_t e) {
    if (first == hipSuccess && e != hipSuccess) first = e;
  };
  const size_t stage = a256((size_t)m * 96) + 2 * a256((size_t)m * 4) + (grouped ? a256((size_t)m * 8) : 0);
  // coins, index arrays and groups are laid out identically on both sides: one copy
  note(hipMemcpyAsync(d_coins, h_coins, stage, hipMemcpyHostToDevice, st));
  const dim3 g(m), b(256);
  if (op == kKeygen) {
    launch_chunk_on(st, s->form, sc, op, m, d_pk, d_sk, d_coins, nullptr, base);
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx0, d_pk, s->pk_slab[sc], d.pk, 0);
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx0, d_sk, s->sk_slab[sc], d.sk, 0);
    if (sc == DPKEM_FRODO976)
      hipLaunchKernelGGL(pkh_scatter, dim3((m + 7) / 8), dim3(64), 0, st, m, d_idx0, d_sk, s->pkh_slab);
  } else if (op == kEnc) {
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx0, s->pk_slab[sc], d_pk, d.pk, 1);
    const uint8_t* pkh = nullptr;
    if (sc == DPKEM_FRODO976) {
      hipLaunchKernelGGL(slab_copy, g, dim3(64), 0, st, m, d_idx0, s->pkh_slab, d_pkh, (size_t)24, 1);
      pkh = d_pkh;
    }
    if (grouped) launch_frodo_grouped(st, op, m, d_ct, d_ss, d_pk, d_coins, base, d_groups, ngroups, pkh);
    else launch_chunk_on(st, s->form, sc, op, m, d_ct, d_ss, d_pk, d_coins, base, pkh);
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx1, d_ct, s->ct_slab[sc], d.ct, 0);
    note(hipGetLastError());
    note(hipMemcpyAsync(L.host_ss, d_ss, (size_t)m * d.ss, hipMemcpyDeviceToHost, st));
  } else {
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx0, s->sk_slab[sc], d_sk, d.sk, 1);
    hipLaunchKernelGGL(slab_copy, g, b, 0, st, m, d_idx1, s->ct_slab[sc], d_ct, d.ct, 1);
    if (grouped) launch_frodo_grouped(st, op, m, d_ss, nullptr, d_ct, d_sk, base, d_groups, ngroups);
    else launch_chunk_on(st, s->form, sc, op, m, d_ss, nullptr, d_ct, d_sk, base);
    note(hipGetLastError());
    note(hipMemcpyAsync(L.host_ss, d_ss, (size_t)m * d.ss, hipMemcpyDeviceToHost, st));
  }
  note(hipGetLastError());
  note(hipEventRecord(L.done, st));
  return first;
}

// Retire lane L's batch (s->mu held).  status DPKEM_OK: every request
// completes with its result.  Otherwise every request completes with that
// status and no result, and the key / ciphertext slots its admission took are
// returned to their pools (after the lane's stream has stopped using them).
void retire_lane(dpkem_server* s, Lane& L, int status) {
  const double t = now_us();
  const SchemeDims& d = kDims[L.scheme];
  if (status != DPKEM_OK) {
    (void)hipStreamSynchronize(L.stream);
    for (uint32_t h : L.new_handles) {
      if (L.op == kKeygen) s->keys[L.scheme].give(h);
      else if (L.op == kEnc) s->cts[L.scheme].give(h);
    }
    s->failed_batches++;
  }
  for (size_t i = 0; i < L.batch.size(); ++i) {
    dpkem_completion c;
    memset(&c, 0, sizeof c);
    c.id = L.batch[i].r.id;
    c.status = status;
    c.scheme = L.scheme;
    c.op = L.op;
    c.t_submit_us = L.batch[i].r.t_arrival_us > 0 ? L.batch[i].r.t_arrival_us : L.batch[i].t_submit;
    c.t_admit_us = L.t_admit;
    c.t_done_us = t;
    if (status == DPKEM_OK) {
      if (L.op == kKeygen) c.key = L.new_handles[i];
      if (L.op == kEnc) c.ct = L.new_handles[i];
      if (L.op != kKeygen) memcpy(c.ss, L.host_ss + i * d.ss, d.ss);
    }
    s->done.push_back(c);
  }
  s->inflight_requests -= L.batch.size();
  L.batch.clear();
  L.new_handles.clear();
  L.busy = false;
}

// Choose the (scheme, op) group whose oldest request arrived first, among the
// schemes lane `only` admits (-1: both).
bool pick_group(dpkem_server* s, int only, int* sc, int* op) {
  double best = 0;
  bool found = false;
  for (int a = 0; a < 2; ++a) {
    if (only >= 0 && only != a) continue;
    for (int b = 0; b < 3; ++b) {
      if (s->pending[a][b].empty()) continue;
      const Pending& P = s->pending[a][b].front();
      const double t = P.r.t_arrival_us > 0 ? P.r.t_arrival_us : P.t_submit;
      if (!found || t < best) {
        best = t;
        *sc = a;
        *op = b;
        found = true;
      }
    }
  }
  return found;
}

void worker_loop(dpkem_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (true) {
    // 1. retire finished lanes: the completion event has fired (results
    // valid) or reports an error (the batch failed)
    for (auto& L : s->lanes) {
      if (!L.busy) continue;
      hipError_t q = hipEventQuery(L.done);
      if (q == hipErrorNotReady) continue;
      if (q == hipSuccess && s->inject_event_faults > 0) {
        s->inject_event_faults--;
        q = hipErrorLaunchFailure;
      }
      retire_lane(s, L, q == hipSuccess ? DPKEM_OK : DPKEM_EDEVICE);
    }
    if (s->inflight_requests == 0) s->drained.notify_all();
    // 2. admit new batches into free lanes
    bool launched = false;
    for (auto& L : s->lanes) {
      if (L.busy) continue;
      int sc, op;
      if (!pick_group(s, L.only_scheme, &sc, &op)) continue;
      auto& q = s->pending[sc][op];
      const size_t m = std::min(q.size(), s->max_batch[sc]);
      L.scheme = sc;
      L.op = op;
      L.batch.clear();
      L.new_handles.clear();
      for (size_t i = 0; i < m; ++i) {
        const Pending& P = q.front();
        uint32_t hnd = 0;
        bool ok = true;
        if (op == kKeygen) ok = s->keys[sc].take(&hnd);
        else if (op == kEnc) ok = s->keys[sc].valid(P.r.key) && s->cts[sc].take(&hnd);
        else ok = s->keys[sc].valid(P.r.key) && s->cts[sc].valid(P.r.ct);
        if (!ok) {  // no free slot or stale handle: fail this request now
          dpkem_completion c;
          memset(&c, 0, sizeof c);
          c.id = P.r.id;
          c.status = (op == kDec || (op == kEnc && !s->keys[sc].valid(P.r.key))) ? DPKEM_ENOTFOUND : DPKEM_EBUDGET;
          c.scheme = sc;
          c.op = op;
          c.t_submit_us = P.r.t_arrival_us > 0 ? P.r.t_arrival_us : P.t_submit;
          c.t_done_us = c.t_admit_us = now_us();
          s->done.push_back(c);
          s->inflight_requests--;
          q.pop_front();
          continue;
        }
        L.batch.push_back(P);
        L.new_handles.push_back(hnd);
        q.pop_front();
      }
      if (L.batch.empty()) continue;
      if (s->group_by_key && sc == DPKEM_FRODO976 && op != kKeygen) {
        // order members by key so that same-key requests sit next to each other
        std::vector<size_t> order(L.batch.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b2) { return L.batch[a].r.key < L.batch[b2].r.key; });
        std::vector<Pending> nb;
        std::vector<uint32_t> nh;
        nb.reserve(order.size());
        nh.reserve(order.size());
        for (size_t i : order) {
          nb.push_back(L.batch[i]);
          nh.push_back(L.new_handles[i]);
        }
        L.batch.swap(nb);
        L.new_handles.swap(nh);
      }
      L.busy = true;
      L.t_admit = now_us();
      s->batches++;
      s->requests += L.batch.size();
      s->max_seen_batch = std::max<uint64_t>(s->max_seen_batch, L.batch.size());
      const bool inject = s->inject_launch_faults > 0;
      if (inject) s->inject_launch_faults--;
      lk.unlock();
      int ngroups = 0;
      const hipError_t err = launch_batch(s, L, inject, &ngroups);
      lk.lock();
      if (ngroups) {
        s->key_groups += (uint64_t)ngroups;
        s->grouped_requests += (uint64_t)L.batch.size();
      }
      if (err != hipSuccess) {
        snprintf(s->last_err, sizeof s->last_err, "batch launch: %s", hipGetErrorString(err));
        retire_lane(s, L, DPKEM_EDEVICE);
      }
      launched = true;
    }
    if (s->stop && s->inflight_requests == 0) break;
    bool any_busy = false;
    for (auto& L : s->lanes) any_busy |= L.busy;
    if (!launched) {
      if (any_busy) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      } else {
        s->cv.wait_for(lk, std::chrono::milliseconds(5));
      }
    }
  }
}

void destroy_all(dpkem_server* s) {
  for (auto& L : s->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    if (L.scratch) (void)hipFree(L.scratch);
    if (L.host) (void)hipHostFree(L.host);
  }
  for (int i = 0; i < 2; ++i) {
    if (s->pk_slab[i]) (void)hipFree(s->pk_slab[i]);
    if (s->sk_slab[i]) (void)hipFree(s->sk_slab[i]);
    if (s->ct_slab[i]) (void)hipFree(s->ct_slab[i]);
  }
  if (s->pkh_slab) (void)hipFree(s->pkh_slab);
}

}  // namespace

extern "C" {

double dpkem_now_us(void) { return now_us(); }

int dpkem_server_create(dpkem_server** out, const dpkem_server_config* cfg) {
  if (!out || !cfg || cfg->lanes < 1 || cfg->lanes > 16) return DPKEM_EINVAL;
  if (cfg->reserve_lanes[0] < 0 || cfg->reserve_lanes[1] < 0 ||
      cfg->reserve_lanes[0] + cfg->reserve_lanes[1] > cfg->lanes)
    return DPKEM_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || cfg->device < 0 || cfg->device >= count) return DPKEM_EDEVICE;
  if (cfg->form != DPKEM_FORM_BATCHED && cfg->form != DPKEM_FORM_PLAIN) return DPKEM_EINVAL;
  dpkem_server* s = new (std::nothrow) dpkem_server();
  if (!s) return DPKEM_ENOMEM;
  s->device = cfg->device;
  s->form = cfg->form;
  if (const char* inj = getenv("DPKEM_SERVER_INJECT_FAULTS")) {
    if (sscanf(inj, "%d,%d", &s->inject_launch_faults, &s->inject_event_faults) < 1) s->inject_launch_faults = 0;
  }
  s->group_by_key = cfg->group_by_key != 0 && cfg->form == DPKEM_FORM_BATCHED;
  (void)hipSetDevice(s->device);
  size_t budget = cfg->hbm_budget_bytes;
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
    budget = fr / 4;
  }
  s->budget = budget;
  // store
  size_t store = 0;
  for (int sc = 0; sc < 2; ++sc) {
    const SchemeDims& d = kDims[sc];
    store += cfg->key_slots[sc] * (d.pk + d.sk + (sc == DPKEM_FRODO976 ? 24 : 0)) + cfg->ct_slots[sc] * d.ct;
  }
  if (store >= budget) {
    delete s;
    return DPKEM_EBUDGET;
  }
  // per-lane batch caps: the rest of the budget split over the lanes
  const size_t per_lane = (budget - store) / (size_t)cfg->lanes;
  for (int sc = 0; sc < 2; ++sc) {
    size_t cap = cfg->max_batch[sc] ? cfg->max_batch[sc] : (size_t)1 << 16;
    while (cap > 1 && lane_bytes(sc, s->form, cap) > per_lane) cap /= 2;
    if (lane_bytes(sc, s->form, cap) > per_lane) {
      delete s;
      return DPKEM_EBUDGET;
    }
    s->max_batch[sc] = cap;
  }
  int rc = DPKEM_OK;
  for (int sc = 0; sc < 2 && rc == DPKEM_OK; ++sc) {
    const SchemeDims& d = kDims[sc];
    s->keys[sc].init(cfg->key_slots[sc]);
    s->cts[sc].init(cfg->ct_slots[sc]);
    if (cfg->key_slots[sc]) {
      if (hipMalloc(&s->pk_slab[sc], cfg->key_slots[sc] * d.pk) != hipSuccess) rc = DPKEM_ENOMEM;
      if (rc == DPKEM_OK && hipMalloc(&s->sk_slab[sc], cfg->key_slots[sc] * d.sk) != hipSuccess) rc = DPKEM_ENOMEM;
    }
    if (rc == DPKEM_OK && cfg->ct_slots[sc] && hipMalloc(&s->ct_slab[sc], cfg->ct_slots[sc] * d.ct) != hipSuccess)
      rc = DPKEM_ENOMEM;
    if (rc == DPKEM_OK && sc == DPKEM_FRODO976 && cfg->key_slots[sc] &&
        hipMalloc(&s->pkh_slab, cfg->key_slots[sc] * 24) != hipSuccess)
      rc = DPKEM_ENOMEM;
  }
  const size_t lane_dev = std::max(lane_bytes(0, s->form, s->max_batch[0]), lane_bytes(1, s->form, s->max_batch[1]));
  const size_t lane_host =
      a256(std::max(s->max_batch[0], s->max_batch[1]) * (96 + 8 + 8 + 32)) + 4096 + 4 * 256;
  s->lanes.resize(cfg->lanes);
  for (int i = 0; i < cfg->lanes; ++i)
    s->lanes[i].only_scheme = i < cfg->reserve_lanes[0] ? 0 : i < cfg->reserve_lanes[0] + cfg->reserve_lanes[1] ? 1 : -1;
  for (auto& L : s->lanes)
    if (rc == DPKEM_OK) rc = init_lane(s, L, lane_dev, lane_host);
  if (rc != DPKEM_OK) {
    destroy_all(s);
    delete s;
    return rc;
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return DPKEM_OK;
}

int dpkem_server_submit(dpkem_server* s, const dpkem_request* reqs, size_t n) {
  if (!s || (!reqs && n)) return DPKEM_EINVAL;
  const double t = now_us();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    size_t queued = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) queued += s->pending[a][b].size();
    if (queued + n > s->queue_cap) return DPKEM_EBUSY;
    for (size_t i = 0; i < n; ++i)
      if (reqs[i].scheme < 0 || reqs[i].scheme > 1 || reqs[i].op < 0 || reqs[i].op > 2) return DPKEM_EINVAL;
    for (size_t i = 0; i < n; ++i) s->pending[reqs[i].scheme][reqs[i].op].push_back({reqs[i], t});
    s->inflight_requests += n;
  }
  s->cv.notify_one();
  return DPKEM_OK;
}

int dpkem_server_poll(dpkem_server* s, dpkem_completion* out, size_t cap, size_t* got) {
  if (!s || !got || (!out && cap)) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  const size_t m = std::min(cap, s->done.size());
  std::copy(s->done.begin(), s->done.begin() + (ptrdiff_t)m, out);
  s->done.erase(s->done.begin(), s->done.begin() + (ptrdiff_t)m);
  *got = m;
  return DPKEM_OK;
}

int dpkem_server_drain(dpkem_server* s) {
  if (!s) return DPKEM_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.notify_one();
  s->drained.wait(lk, [&] { return s->inflight_requests == 0; });
  // a bounded wait inside a kernel that gave up invalidates what ran since the last check
  uint32_t saber = 0, frodo = 0;
  (void)hipSetDevice(s->device);
  if (saber_wait_expired(&saber, true) != hipSuccess || frodo_wait_expired(&frodo, true) != hipSuccess || saber ||
      frodo) {
    snprintf(s->last_err, sizeof s->last_err, "a bounded LDS wait gave up inside a %s kernel",
             frodo && !saber ? "FrodoKEM" : "Saber-family");
    return DPKEM_EDEVICE;
  }
  return DPKEM_OK;
}

int dpkem_server_cancel(dpkem_server* s, size_t* cancelled) {
  if (!s) return DPKEM_EINVAL;
  size_t k = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const double t = now_us();
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 3; ++b) {
        for (const Pending& P : s->pending[a][b]) {
          dpkem_completion c;
          memset(&c, 0, sizeof c);
          c.id = P.r.id;
          c.status = DPKEM_ECANCELED;
          c.scheme = a;
          c.op = b;
          c.t_submit_us = P.r.t_arrival_us > 0 ? P.r.t_arrival_us : P.t_submit;
          c.t_done_us = c.t_admit_us = t;
          s->done.push_back(c);
        }
        k += s->pending[a][b].size();
        s->pending[a][b].clear();
      }
    s->inflight_requests -= k;
    if (s->inflight_requests == 0) s->drained.notify_all();
  }
  s->cv.notify_one();
  if (cancelled) *cancelled = k;
  return DPKEM_OK;
}

int dpkem_server_release(dpkem_server* s, int scheme, int kind, uint32_t handle) {
  if (!s || scheme < 0 || scheme > 1) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  const bool ok = kind == 0 ? s->keys[scheme].give(handle) : s->cts[scheme].give(handle);
  return ok ? DPKEM_OK : DPKEM_ENOTFOUND;
}

// Copy a stored object to host memory: kind 0 = pk, 1 = sk, 2 = ct.
int dpkem_server_read(dpkem_server* s, int scheme, int kind, uint32_t handle, uint8_t* out) {
  if (!s || !out || scheme < 0 || scheme > 1 || kind < 0 || kind > 2) return DPKEM_EINVAL;
  const SchemeDims& d = kDims[scheme];
  const uint8_t* src;
  size_t len;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (kind < 2 ? !s->keys[scheme].valid(handle) : !s->cts[scheme].valid(handle)) return DPKEM_ENOTFOUND;
    src = kind == 0 ? s->pk_slab[scheme] + (size_t)handle * d.pk
          : kind == 1 ? s->sk_slab[scheme] + (size_t)handle * d.sk
                      : s->ct_slab[scheme] + (size_t)handle * d.ct;
    len = kind == 0 ? d.pk : kind == 1 ? d.sk : d.ct;
  }
  (void)hipSetDevice(s->device);
  return hipMemcpy(out, src, len, hipMemcpyDeviceToHost) == hipSuccess ? DPKEM_OK : DPKEM_EDEVICE;
}

int dpkem_server_stats(dpkem_server* s, char* buf, size_t len) {
  if (!s || !buf || !len) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  const int w = snprintf(buf, len,
                         "{\"batches\": %llu, \"requests\": %llu, \"max_batch_seen\": %llu, \"lanes\": %zu, "
                         "\"max_batch_frodo976\": %zu, \"max_batch_saber\": %zu, \"hbm_budget\": %zu, "
                         "\"frodo_key_groups\": %llu, \"frodo_grouped_requests\": %llu, \"failed_batches\": %llu}",
                         (unsigned long long)s->batches, (unsigned long long)s->requests,
                         (unsigned long long)s->max_seen_batch, s->lanes.size(), s->max_batch[0], s->max_batch[1],
                         s->budget, (unsigned long long)s->key_groups, (unsigned long long)s->grouped_requests,
                         (unsigned long long)s->failed_batches);
  return (w > 0 && (size_t)w < len) ? DPKEM_OK : DPKEM_EINVAL;
}

int dpkem_server_destroy(dpkem_server* s) {
  if (!s) return DPKEM_EINVAL;
  (void)dpkem_server_cancel(s, nullptr);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_one();
  if (s->worker.joinable()) s->worker.join();
  destroy_all(s);
  delete s;
  return DPKEM_OK;
}

}  // extern "C"
