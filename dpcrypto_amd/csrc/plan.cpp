// Graph-captured task sets ("plans").
//
// The reference launches 15-31 kernels per op, one op at a time, on the
// legacy default stream (FrodoKEM976-GPU/src/cuda_kernel.cu:64-84, 159-189,
// 282-323; SABER-GPU/src/cuda_kernel.cu:68-133, 213-273).  A plan instead
// records a whole set of KEM tasks on device-resident arrays -- possibly
// different schemes and ops, with instance-wise dependencies such as
// keygen -> encaps -> decaps -- splits each task into instance chunks, spreads
// the chunks over several streams, and captures everything once into a HIP
// graph.  Replaying the graph issues all launches at once; the serial hash
// chains of one chunk (few waves, latency-bound) then run beside the
// throughput-bound kernels of other chunks.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "engine_internal.hpp"

using namespace dpkem;

namespace {

struct Task {
  int scheme, op;
  size_t n;
  uint8_t *o0, *o1;
  const uint8_t *i0, *i1;
  int after;  // task index this one depends on, or -1
};

enum ChainKind { kNoChain = 0, kShareChain = 1, kPipeChain = 2 };

struct Sub {
  int task;
  int chain = -1;          // >= 0: Saber keygen/enc/dec chain (tasks task, chain, chain2) in one launch sequence
  int chain2 = -1;
  int chain_kind = kNoChain;
  size_t begin, count;
  int stream;
  size_t scratch_off;
  hipEvent_t done = nullptr;
  std::vector<int> deps;   // sub indices
  std::vector<int> mt;     // non-empty: independent Frodo tasks of one op run as one multi-task launch sequence
  std::vector<size_t> mt_off;  // their scratch offsets
  int mix[3] = {-1, -1, -1};  // independent Saber-family keygen / enc / dec tasks in one mixed launch sequence
  bool has_mix() const { return mix[0] >= 0 || mix[1] >= 0 || mix[2] >= 0; }
};

}  // namespace

struct dpkem_plan {
  dpkem_engine* e = nullptr;
  int nstreams = 1, chunks = 1, form = DPKEM_FORM_BATCHED, saber_ipw = 1, frodo_persist = 1, frodo_threads = 0;
  unsigned flags = 0;
  std::vector<Task> tasks;
  std::vector<Sub> subs;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> join;
  hipEvent_t fork = nullptr;
  uint8_t* scratch = nullptr;
  size_t scratch_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool built = false;  // dpkem_plan_build succeeded (graph instantiated, or DPKEM_PLAN_NO_GRAPH)
};

namespace {

void release(dpkem_plan* p) {
  p->built = false;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  for (auto& s : p->subs) {
    if (s.done) (void)hipEventDestroy(s.done);
  }
  for (auto ev : p->join) (void)hipEventDestroy(ev);
  p->join.clear();
  if (p->fork) (void)hipEventDestroy(p->fork);
  p->fork = nullptr;
  for (auto st : p->streams) (void)hipStreamDestroy(st);
  p->streams.clear();
  if (p->scratch) (void)hipFree(p->scratch);
  p->scratch = nullptr;
  p->subs.clear();
}

size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

int saber_rank_of(int scheme) {
  return scheme == DPKEM_SABER ? 3 : scheme == DPKEM_LIGHTSABER ? 2 : scheme == DPKEM_FIRESABER ? 4 : 0;
}
// FrodoKEM parameter set (n) of a Frodo scheme, 0 otherwise
int frodo_set_of(int scheme) {
  return scheme == DPKEM_FRODO640 ? 640
         : scheme == DPKEM_FRODO1344 ? 1344
         : (scheme == DPKEM_FRODO976 || scheme == DPKEM_FRODO976_AES) ? 976
                                                                       : 0;
}

// Keygen t0 -> enc t1 -> dec t2 of one scheme over the same arrays (t1
// consumes t0's pk, t2 consumes t1's ct and t0's sk): each chunk runs the
// three ops as one launch sequence on one stream, either pipelined (default:
// serial stages of different ops that read disjoint inputs share launches)
// or, with DPKEM_PLAN_SHARE_MATRIX and Saber l = 3, expanding the public
// matrix once.  t1's and t2's sub-tasks are folded into t0's, so their
// dependants wait for whole chains.  Returns the chain kind.
int find_chain(const dpkem_plan* p, int t0, int* t1, int* t2) {
  const Task& K = p->tasks[t0];
  if (p->form != DPKEM_FORM_BATCHED || (!saber_rank_of(K.scheme) && !frodo_set_of(K.scheme)) || K.op != kKeygen ||
      K.after != -1)
    return kNoChain;
  const int kind = (p->flags & DPKEM_PLAN_SHARE_MATRIX) && K.scheme == DPKEM_SABER ? kShareChain
                   : (p->flags & DPKEM_PLAN_NO_PIPELINE)                           ? kNoChain
                                                                                   : kPipeChain;
  if (kind == kNoChain) return kNoChain;
  for (size_t a = t0 + 1; a < p->tasks.size(); ++a) {
    const Task& E = p->tasks[a];
    if (E.after != t0 || E.scheme != K.scheme || E.op != kEnc || E.n != K.n || E.i0 != K.o0) continue;
    for (size_t b = a + 1; b < p->tasks.size(); ++b) {
      const Task& Dd = p->tasks[b];
      if (Dd.after == (int)a && Dd.scheme == K.scheme && Dd.op == kDec && Dd.n == K.n && Dd.i0 == E.o0 &&
          Dd.i1 == K.o1) {
        *t1 = (int)a;
        *t2 = (int)b;
        return kind;
      }
    }
  }
  return kNoChain;
}

// Independent Frodo tasks of one (scheme, op) -- no `after`, not part of a
// keygen -> enc -> dec chain -- share one multi-task launch sequence when
// there are at least two of them (groups of up to kFrodoMtTasks, in task
// order).  Returns, per task, the index of its group in `groups`, or -1.
std::vector<int> merge_groups(const dpkem_plan* p, const std::vector<char>& in_chain,
                              std::vector<std::vector<int>>* groups) {
  std::vector<int> group_of(p->tasks.size(), -1);
  if (p->form != DPKEM_FORM_BATCHED || (p->flags & DPKEM_PLAN_NO_MERGE)) return group_of;
  for (size_t t = 0; t < p->tasks.size(); ++t) {
    const Task& T = p->tasks[t];
    if (group_of[t] >= 0 || in_chain[t] || T.after != -1 || !frodo_set_of(T.scheme)) continue;
    std::vector<int> g{(int)t};
    for (size_t u = t + 1; u < p->tasks.size() && g.size() < (size_t)kFrodoMtTasks; ++u) {
      const Task& U = p->tasks[u];
      if (group_of[u] < 0 && !in_chain[u] && U.after == -1 && U.scheme == T.scheme && U.op == T.op) g.push_back((int)u);
    }
    if (g.size() < 2) continue;
    for (int u : g) group_of[u] = (int)groups->size();
    groups->push_back(g);
  }
  return group_of;
}

// Independent Saber-family tasks of one scheme (no `after`, no dependants,
// not in a keygen -> enc -> dec chain) share one mixed launch sequence, at
// most one task per op each (saber_mix_launch: all their hash chains in one
// launch, all their products in one, the closing hashes in one).  A task that
// others depend on keeps its chunked launches, so that its dependants can
// start chunk by chunk.  Groups form in task order; returns, per task, the
// index of its group in `mixes`, or -1.
std::vector<int> mix_groups(const dpkem_plan* p, const std::vector<char>& in_chain,
                            std::vector<std::array<int, 3>>* mixes) {
  std::vector<int> mix_of(p->tasks.size(), -1);
  if (p->form != DPKEM_FORM_BATCHED || (p->flags & DPKEM_PLAN_NO_MERGE)) return mix_of;
  std::vector<char> has_dependant(p->tasks.size(), 0);
  for (const Task& T : p->tasks)
    if (T.after >= 0) has_dependant[T.after] = 1;
  auto eligible = [&](size_t t) {
    const Task& T = p->tasks[t];
    return mix_of[t] < 0 && !in_chain[t] && T.after == -1 && !has_dependant[t] && saber_rank_of(T.scheme) != 0;
  };
  for (size_t t = 0; t < p->tasks.size(); ++t) {
    const Task& T = p->tasks[t];
    if (!eligible(t)) continue;
    std::array<int, 3> g{-1, -1, -1};
    g[T.op] = (int)t;
    for (size_t u = t + 1; u < p->tasks.size(); ++u) {
      const Task& U = p->tasks[u];
      if (eligible(u) && U.scheme == T.scheme && g[U.op] < 0) g[U.op] = (int)u;
    }
    for (int u : g)
      if (u >= 0) mix_of[u] = (int)mixes->size();
    mixes->push_back(g);
  }
  return mix_of;
}

// Split every task into chunks and assign streams and dependencies.
void schedule(dpkem_plan* p) {
  p->subs.clear();
  std::vector<std::vector<int>> task_subs(p->tasks.size());
  std::vector<int> chain_of(p->tasks.size(), -1);
  std::vector<int> absorbed(p->tasks.size(), -1);  // task -> chain root it was folded into
  // tasks that take part in a keygen -> enc -> dec chain run in the chain's launches
  std::vector<char> in_chain(p->tasks.size(), 0);
  for (size_t t = 0; t < p->tasks.size(); ++t) {
    int t1 = -1, t2 = -1;
    if (!in_chain[t] && find_chain(p, (int)t, &t1, &t2)) in_chain[t] = in_chain[t1] = in_chain[t2] = 1;
  }
  std::vector<std::vector<int>> groups;
  const std::vector<int> group_of = merge_groups(p, in_chain, &groups);
  std::vector<std::array<int, 3>> mixes;
  const std::vector<int> mix_of = mix_groups(p, in_chain, &mixes);
  int chains = 0;
  size_t off = 0;
  for (size_t t = 0; t < p->tasks.size(); ++t) {
    if (absorbed[t] >= 0) {
      task_subs[t] = task_subs[absorbed[t]];
      chain_of[t] = chain_of[absorbed[t]];
      continue;
    }
    const Task& T = p->tasks[t];
    if (mix_of[t] >= 0) {
      const std::array<int, 3>& g = mixes[mix_of[t]];
      int first = (int)t;
      for (int u : g)
        if (u >= 0) first = std::min(first, u);
      if (first != (int)t) {  // a later member: runs in the group's sub-task
        task_subs[t] = task_subs[first];
        chain_of[t] = chain_of[first];
        continue;
      }
      chain_of[t] = chains++;
      Sub s;
      s.task = (int)t;
      s.begin = 0;
      s.count = T.n;
      s.stream = chain_of[t] % p->nstreams;
      s.scratch_off = off;
      size_t cnt[3] = {0, 0, 0};
      for (int k = 0; k < 3; ++k) {
        s.mix[k] = g[k];
        if (g[k] >= 0) cnt[k] = p->tasks[g[k]].n;
      }
      off += round_up(saber_mix_scratch_bytes(saber_rank_of(T.scheme), cnt[0], cnt[1], cnt[2]));
      task_subs[t].push_back((int)p->subs.size());
      p->subs.push_back(s);
      continue;
    }
    if (group_of[t] >= 0) {
      const std::vector<int>& g = groups[group_of[t]];
      if (g[0] != (int)t) {  // a later member: runs in the group's sub-task
        task_subs[t] = task_subs[g[0]];
        chain_of[t] = chain_of[g[0]];
        continue;
      }
      chain_of[t] = chains++;
      Sub s;
      s.task = (int)t;
      s.begin = 0;
      s.count = T.n;
      s.stream = chain_of[t] % p->nstreams;
      s.scratch_off = off;
      s.mt = g;
      for (int u : g) {
        s.mt_off.push_back(off);
        off += round_up(frodo_mt_scratch_bytes(frodo_set_of(T.scheme), p->tasks[u].n));
      }
      task_subs[t].push_back((int)p->subs.size());
      p->subs.push_back(s);
      continue;
    }
    int t1 = -1, t2 = -1;
    if (const int kind = find_chain(p, (int)t, &t1, &t2)) {
      absorbed[t1] = absorbed[t2] = (int)t;
      chain_of[t] = chains++;
      const size_t k = std::max<size_t>(1, std::min<size_t>((size_t)p->chunks, T.n));
      const size_t per = (T.n + k - 1) / k;
      for (size_t c = 0, b = 0; b < T.n; ++c, b += per) {
        Sub s;
        s.task = (int)t;
        s.chain = t1;
        s.chain2 = t2;
        s.chain_kind = kind;
        s.begin = b;
        s.count = std::min(per, T.n - b);
        s.stream = (int)((c + (size_t)chain_of[t]) % (size_t)p->nstreams);
        s.scratch_off = off;
        const size_t per_inst = kind == kShareChain          ? kSaberChainScratchPerInst
                                : saber_rank_of(T.scheme) ? saber_pipe_scratch_per_inst(saber_rank_of(T.scheme))
                                                          : frodo_pipe_scratch_per_inst(frodo_set_of(T.scheme));
        off += round_up(s.count * per_inst + 16 * 256);
        task_subs[t].push_back((int)p->subs.size());
        p->subs.push_back(s);
      }
      continue;
    }
    chain_of[t] = T.after >= 0 ? chain_of[T.after] : chains++;
    const size_t k = std::max<size_t>(1, std::min<size_t>((size_t)p->chunks, T.n));
    const size_t per = (T.n + k - 1) / k;
    for (size_t c = 0, b = 0; b < T.n; ++c, b += per) {
      Sub s;
      s.task = (int)t;
      s.begin = b;
      s.count = std::min(per, T.n - b);
      // chunk c of a dependency chain stays on one stream; independent chains
      // start one stream further on, so their chunks spread over all streams
      s.stream = (int)((c + (size_t)chain_of[t]) % (size_t)p->nstreams);
      s.scratch_off = off;
      off += round_up(chunk_scratch_bytes(T.scheme, p->form, s.count));
      if (T.after >= 0) {
        const Task& A = p->tasks[T.after];
        const auto& prev = task_subs[T.after];
        if (A.n == T.n && prev.size() == (size_t)std::max<size_t>(1, std::min<size_t>((size_t)p->chunks, A.n)) &&
            c < prev.size()) {
          s.deps.push_back(prev[c]);  // instance-aligned chunk
        } else {
          s.deps = prev;  // join on the whole producer
        }
      }
      task_subs[t].push_back((int)p->subs.size());
      p->subs.push_back(s);
    }
  }
  p->scratch_bytes = off;
}

// Issue every sub-task: origin -> fork -> per-stream sub-task chains ->
// join -> origin.  Recorded once into the graph, or run on each launch of a
// DPKEM_PLAN_NO_GRAPH plan.
bool issue(dpkem_plan* p, hipStream_t origin) {
  bool ok = hipEventRecord(p->fork, origin) == hipSuccess;
  for (auto st : p->streams) ok = ok && hipStreamWaitEvent(st, p->fork, 0) == hipSuccess;
  for (auto& s : p->subs) {
    const Task& T = p->tasks[s.task];
    hipStream_t st = p->streams[s.stream];
    for (int d : s.deps)
      if (p->subs[d].stream != s.stream) ok = ok && hipStreamWaitEvent(st, p->subs[d].done, 0) == hipSuccess;
    if (s.has_mix()) {
      SaberMixDesc d{};
      if (s.mix[0] >= 0) {
        const Task& K = p->tasks[s.mix[0]];
        d.nk = (int)K.n;
        d.kc = K.i0;
        d.pk_k = K.o0;
        d.sk_k = K.o1;
      }
      if (s.mix[1] >= 0) {
        const Task& E = p->tasks[s.mix[1]];
        d.ne = (int)E.n;
        d.pk_e = E.i0;
        d.ec = E.i1;
        d.ct_e = E.o0;
        d.ss_e = E.o1;
      }
      if (s.mix[2] >= 0) {
        const Task& Dd = p->tasks[s.mix[2]];
        d.nd = (int)Dd.n;
        d.ct_d = Dd.i0;
        d.sk_d = Dd.i1;
        d.ss_d = Dd.o0;
      }
      saber_mix_launch(st, saber_rank_of(T.scheme), d, p->scratch + s.scratch_off, p->saber_ipw);
      ok = ok && hipGetLastError() == hipSuccess;
      ok = ok && hipEventRecord(s.done, st) == hipSuccess;
      continue;
    }
    if (!s.mt.empty()) {
      FrodoMtDesc d[kFrodoMtTasks];
      for (size_t k = 0; k < s.mt.size(); ++k) {
        const Task& U = p->tasks[s.mt[k]];
        d[k] = {(int)U.n, U.i0, U.i1, U.o0, U.o1, p->scratch + s.mt_off[k]};
      }
      frodo_mt_launch(st, T.op, frodo_set_of(T.scheme), T.scheme == DPKEM_FRODO976_AES, d, (int)s.mt.size(),
                      p->frodo_persist != 0, p->frodo_threads);
      ok = ok && hipGetLastError() == hipSuccess;
      ok = ok && hipEventRecord(s.done, st) == hipSuccess;
      continue;
    }
    if (s.chain >= 0) {
      const Task& E = p->tasks[s.chain];
      const Task& Dd = p->tasks[s.chain2];
      const size_t b = s.begin;
      size_t ko[2], ki[2], eo[2], ei[2];
      io_sizes(T.scheme, kKeygen, ko, ki);
      io_sizes(T.scheme, kEnc, eo, ei);
      const uint8_t* kc = T.i0 + b * ki[0];
      const uint8_t* ec = E.i1 + b * ei[1];
      uint8_t *pk = T.o0 + b * ko[0], *sk = T.o1 + b * ko[1];
      uint8_t *ct = E.o0 + b * eo[0], *ss = E.o1 + b * eo[1], *ss2 = Dd.o0 + b * eo[1];
      if (s.chain_kind == kShareChain)
        saber_chain_launch(st, (int)s.count, kc, ec, pk, sk, ct, ss, ss2, p->scratch + s.scratch_off, p->saber_ipw);
      else if (saber_rank_of(T.scheme))
        saber_pipe_launch(st, (int)s.count, kc, ec, pk, sk, ct, ss, ss2, p->scratch + s.scratch_off,
                          saber_rank_of(T.scheme), p->saber_ipw, (p->flags & DPKEM_PLAN_FORCE_REJECT) != 0);
      else
        frodo_pipe_launch(st, (int)s.count, kc, ec, pk, sk, ct, ss, ss2, p->scratch + s.scratch_off,
                          frodo_set_of(T.scheme), T.scheme == DPKEM_FRODO976_AES, p->frodo_threads,
                          (p->flags & DPKEM_PLAN_FORCE_REJECT) != 0);
      ok = ok && hipGetLastError() == hipSuccess;
      ok = ok && hipEventRecord(s.done, st) == hipSuccess;
      continue;
    }
    size_t out[2], in[2];
    io_sizes(T.scheme, T.op, out, in);
    uint8_t* o0 = out[0] ? T.o0 + s.begin * out[0] : nullptr;
    uint8_t* o1 = out[1] ? T.o1 + s.begin * out[1] : nullptr;
    const uint8_t* i0 = in[0] ? T.i0 + s.begin * in[0] : nullptr;
    const uint8_t* i1 = in[1] ? T.i1 + s.begin * in[1] : nullptr;
    launch_chunk_on(st, p->form, T.scheme, T.op, (int)s.count, o0, o1, i0, i1, p->scratch + s.scratch_off, nullptr,
                    p->saber_ipw, p->frodo_threads);
    ok = ok && hipGetLastError() == hipSuccess;
    ok = ok && hipEventRecord(s.done, st) == hipSuccess;
  }
  for (int k = 0; k < p->nstreams; ++k) {
    ok = ok && hipEventRecord(p->join[k], p->streams[k]) == hipSuccess;
    ok = ok && hipStreamWaitEvent(origin, p->join[k], 0) == hipSuccess;
  }
  return ok;
}

}  // namespace

extern "C" {

int dpkem_plan_create(dpkem_engine* e, dpkem_plan** out, int nstreams, int chunks) {
  if (!e || !out || nstreams < 1 || nstreams > 16 || chunks < 1 || chunks > 64) return DPKEM_EINVAL;
  dpkem_plan* p = new (std::nothrow) dpkem_plan();
  if (!p) return DPKEM_ENOMEM;
  p->e = e;
  p->nstreams = nstreams;
  p->chunks = chunks;
  p->form = e->form;
  p->saber_ipw = e->saber_ipw;
  p->frodo_persist = e->frodo_persist;
  p->frodo_threads = e->frodo_threads;
  *out = p;
  return DPKEM_OK;
}

int dpkem_plan_add(dpkem_plan* p, int scheme, int op, size_t n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
                   const uint8_t* i1, int after, int* task_id) {
  if (!p || p->built || scheme < DPKEM_FRODO976 || scheme > DPKEM_FRODO1344 ||
      (scheme >= DPKEM_FRODO976_AES && p->form != DPKEM_FORM_BATCHED) || op < 0 || op > 2 || n == 0 ||
      n > (size_t)INT32_MAX)
    return DPKEM_EINVAL;
  if (after >= (int)p->tasks.size() || after < -1) return DPKEM_EINVAL;
  size_t out[2], in[2];
  io_sizes(scheme, op, out, in);
  uint8_t* outs[2] = {o0, o1};
  const uint8_t* ins[2] = {i0, i1};
  for (int k = 0; k < 2; ++k) {
    if ((out[k] && (!outs[k] || ((uintptr_t)outs[k] & 7))) || (in[k] && (!ins[k] || ((uintptr_t)ins[k] & 7))))
      return DPKEM_EINVAL;
  }
  p->tasks.push_back({scheme, op, n, o0, o1, i0, i1, after});
  if (task_id) *task_id = (int)p->tasks.size() - 1;
  return DPKEM_OK;
}

int dpkem_plan_set_options(dpkem_plan* p, unsigned flags) {
  if (!p || p->built ||
      (flags & ~(DPKEM_PLAN_SHARE_MATRIX | DPKEM_PLAN_NO_PIPELINE | DPKEM_PLAN_NO_GRAPH | DPKEM_PLAN_NO_MERGE |
                 DPKEM_PLAN_FORCE_REJECT)))
    return DPKEM_EINVAL;
  p->flags = flags;
  return DPKEM_OK;
}

int dpkem_plan_build(dpkem_plan* p) {
  if (!p || p->tasks.empty()) return DPKEM_EINVAL;
  dpkem_engine* e = p->e;
  std::lock_guard<std::mutex> lock(e->mu);
  release(p);
  schedule(p);
  if (p->scratch_bytes > e->budget) {
    snprintf(e->last_err, sizeof e->last_err, "plan scratch %zu B exceeds HBM budget %zu B", p->scratch_bytes, e->budget);
    return DPKEM_EBUDGET;
  }
  hipError_t err;
#define PLAN_TRY(call)                                              \
  do {                                                              \
    err = (call);                                                   \
    if (err != hipSuccess) {                                        \
      int rc = record_hip_error(e, err, #call);                     \
      release(p);                                                   \
      return rc;                                                    \
    }                                                               \
  } while (0)
  PLAN_TRY(hipSetDevice(e->device));
  PLAN_TRY(hipMalloc(&p->scratch, std::max<size_t>(p->scratch_bytes, 256)));
  p->streams.resize(p->nstreams, nullptr);
  for (auto& st : p->streams) PLAN_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  PLAN_TRY(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
  p->join.resize(p->nstreams, nullptr);
  for (auto& ev : p->join) PLAN_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  for (auto& s : p->subs) PLAN_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));

  if (p->flags & DPKEM_PLAN_NO_GRAPH) {  // launched directly by dpkem_plan_launch
    p->built = true;
    return DPKEM_OK;
  }
  // capture: origin -> fork -> per-stream sub-task chains -> join -> origin
  hipStream_t cap;
  PLAN_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    (void)hipStreamDestroy(cap);
    int rc = record_hip_error(e, err, "hipStreamBeginCapture");
    release(p);
    return rc;
  }
  Profiler* saved = tls_profiler;
  tls_profiler = nullptr;  // no timing events inside the graph
  bool ok = issue(p, cap);
  tls_profiler = saved;
  err = hipStreamEndCapture(cap, &p->graph);
  (void)hipStreamDestroy(cap);
  if (err != hipSuccess || !ok) {
    int rc = record_hip_error(e, err != hipSuccess ? err : hipErrorUnknown, "graph capture");
    release(p);
    return rc;
  }
  PLAN_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
#undef PLAN_TRY
  p->built = true;
  return DPKEM_OK;
}

int dpkem_plan_launch(dpkem_plan* p) {
  if (!p || !p->built) return DPKEM_EINVAL;
  dpkem_engine* e = p->e;
  std::lock_guard<std::mutex> lock(e->mu);
  hipError_t err = hipSetDevice(e->device);
  if (err != hipSuccess) return record_hip_error(e, err, "hipSetDevice");
  if (p->flags & DPKEM_PLAN_NO_GRAPH) {
    // the same launches, events and streams as the graph, issued now (the
    // engine's profiler, when on, times each kernel)
    tls_profiler = e->profiling ? &e->prof : nullptr;
    const bool ok = issue(p, e->stream);
    tls_profiler = nullptr;
    return ok ? DPKEM_OK : record_hip_error(e, hipGetLastError(), "plan launch");
  }
  err = hipGraphLaunch(p->exec, e->stream);
  return err == hipSuccess ? DPKEM_OK : record_hip_error(e, err, "hipGraphLaunch");
}

int dpkem_plan_info(dpkem_plan* p, size_t* subtasks, size_t* scratch_bytes) {
  if (!p) return DPKEM_EINVAL;
  if (subtasks) *subtasks = p->subs.size();
  if (scratch_bytes) *scratch_bytes = p->scratch_bytes;
  return DPKEM_OK;
}

int dpkem_plan_destroy(dpkem_plan* p) {
  if (!p) return DPKEM_EINVAL;
  {
    std::lock_guard<std::mutex> lock(p->e->mu);
    (void)hipSetDevice(p->e->device);
    (void)hipStreamSynchronize(p->e->stream);
    release(p);
  }
  delete p;
  return DPKEM_OK;
}

}  // extern "C"
