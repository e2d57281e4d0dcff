// Host engine and C ABI for the batched KEM kernels.
//
// Replaces the reference's host drivers (FrodoKEM976-GPU/src/cuda_kernel.cu:21-335,
// SABER-GPU/src/cuda_kernel.cu:19-299), which allocate and free device memory
// on every call, run on the legacy default stream with synchronous copies and
// report nothing.  Here an engine owns one stream and a grow-only scratch
// arena bounded by an HBM budget; batches larger than the budget allows are
// cut into chunks; host pointers are staged through device buffers; every HIP
// error becomes a return code.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>
#include <sys/random.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "engine_internal.hpp"

using namespace dpkem;

namespace {

struct SchemeSizes {
  size_t pk, sk, ct, ss, kcoins, ecoins;
};

const SchemeSizes kSizes[7] = {
    {15632, 31296, 15744, 24, 64, 24},  // FrodoKEM-976-SHAKE
    {992, 2304, 1088, 32, 96, 32},      // Saber
    {15632, 31296, 15744, 24, 64, 24},  // FrodoKEM-976-AES
    {672, 1568, 736, 32, 96, 32},       // LightSaber
    {1312, 3040, 1472, 32, 96, 32},     // FireSaber
    {9616, 19888, 9720, 16, 48, 16},    // Frodo-640-SHAKE
    {21520, 43088, 21632, 32, 80, 32},  // Frodo-1344-SHAKE
};

bool valid_scheme(int scheme) { return scheme >= DPKEM_FRODO976 && scheme <= DPKEM_FRODO1344; }
bool saber_family(int scheme) {
  return scheme == DPKEM_SABER || scheme == DPKEM_LIGHTSABER || scheme == DPKEM_FIRESABER;
}
int saber_rank(int scheme) { return scheme == DPKEM_LIGHTSABER ? 2 : scheme == DPKEM_FIRESABER ? 4 : 3; }
// FrodoKEM parameter set (n) of a Frodo scheme
int frodo_set(int scheme) { return scheme == DPKEM_FRODO640 ? 640 : scheme == DPKEM_FRODO1344 ? 1344 : 976; }
// schemes with no plain (one-workgroup-per-KEM) comparator form
bool batched_only(int scheme) { return scheme != DPKEM_FRODO976 && scheme != DPKEM_SABER; }

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

struct Carver {
  uint8_t* base;
  size_t off = 0;
  uint8_t* take(size_t bytes) {
    uint8_t* p = base + off;
    off += align_up(bytes);
    return p;
  }
};

}  // namespace

namespace {

int fail_hip(dpkem_engine* e, hipError_t err, const char* what) {
  if (err == hipSuccess) return DPKEM_OK;
  if (e) snprintf(e->last_err, sizeof e->last_err, "%s: %s", what, hipGetErrorString(err));
  return err == hipErrorOutOfMemory ? DPKEM_ENOMEM : DPKEM_EDEVICE;
}

#define HIP_TRY(e, call)                                  \
  do {                                                    \
    hipError_t _err = (call);                             \
    if (_err != hipSuccess) return fail_hip(e, _err, #call); \
  } while (0)

// After the engine's stream has drained: did a bounded wait inside a kernel
// give up (its launch then wrote invalid results)?  Clears the record.
int check_waits(dpkem_engine* e) {
  uint32_t saber = 0, frodo = 0;
  HIP_TRY(e, dpkem::saber_wait_expired(&saber, true));
  HIP_TRY(e, dpkem::frodo_wait_expired(&frodo, true));
  if (saber || frodo) {
    snprintf(e->last_err, sizeof e->last_err,
             "a bounded LDS wait gave up inside a %s kernel (%u lanes); results of launches since the last check "
             "are invalid",
             saber ? "Saber-family" : "FrodoKEM", saber + frodo);
    return DPKEM_EDEVICE;
  }
  return DPKEM_OK;
}

int ensure_arena(dpkem_engine* e, size_t bytes) {
  if (bytes <= e->arena_bytes) return DPKEM_OK;
  if (bytes > e->budget) {
    snprintf(e->last_err, sizeof e->last_err, "scratch of %zu B exceeds HBM budget %zu B", bytes, e->budget);
    return DPKEM_EBUDGET;
  }
  if (e->arena) {
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    HIP_TRY(e, hipFree(e->arena));
    e->arena = nullptr;
    e->arena_bytes = 0;
  }
  HIP_TRY(e, hipMalloc(&e->arena, bytes));
  e->arena_bytes = bytes;
  return DPKEM_OK;
}

}  // namespace

namespace dpkem {

int record_hip_error(dpkem_engine* e, hipError_t err, const char* what) { return fail_hip(e, err, what); }

// Bytes of device scratch one instance needs in the given form, excluding staged I/O.
size_t scratch_per_inst(int scheme, int form) {
  if (form == DPKEM_FORM_PLAIN) return align_up(plain_scratch_per_inst(scheme));
  return saber_family(scheme) ? align_up(saber_scratch_per_inst(saber_rank(scheme)))
                              : align_up(frodo_scratch_per_inst(frodo_set(scheme)));
}

// I/O byte sizes per instance: two outputs and two inputs (0 if unused).
void io_sizes(int scheme, int op, size_t out[2], size_t in[2]) {
  const SchemeSizes& s = kSizes[scheme];
  if (op == kKeygen) {
    out[0] = s.pk; out[1] = s.sk; in[0] = s.kcoins; in[1] = 0;
  } else if (op == kEnc) {
    out[0] = s.ct; out[1] = s.ss; in[0] = s.pk; in[1] = s.ecoins;
  } else {
    out[0] = s.ss; out[1] = 0; in[0] = s.ct; in[1] = s.sk;
  }
}

// Launch one chunk with all pointers on the device.
void launch_chunk_on(hipStream_t st, int form, int scheme, int op, int n, uint8_t* o0, uint8_t* o1,
                     const uint8_t* i0, const uint8_t* i1, uint8_t* scratch, const uint8_t* frodo_pkh, int saber_ipw,
                     int frodo_threads) {
  Carver c{scratch};
  if (form == DPKEM_FORM_PLAIN) {
    if (scheme == DPKEM_SABER) plain_saber_launch(st, op, n, o0, o1, i0, i1, scratch);
    else plain_frodo_launch(st, op, n, o0, o1, i0, i1, scratch);
    return;
  }
  if (saber_family(scheme)) {
    // one op as a mixed launch sequence (saber_mix_launch): its hash chains
    // in one launch, its products in one, keygen's / encaps' closing hashes
    // in one (decaps: two launches; its K for both compare outcomes is
    // hashed in the first, the re-encryption selects one)
    SaberMixDesc d{};
    if (op == kKeygen) {
      d.nk = n;
      d.kc = i0;
      d.pk_k = o0;
      d.sk_k = o1;
    } else if (op == kEnc) {
      d.ne = n;
      d.pk_e = i0;
      d.ec = i1;
      d.ct_e = o0;
      d.ss_e = o1;
    } else {
      d.nd = n;
      d.ct_d = i0;
      d.sk_d = i1;
      d.ss_d = o0;
    }
    saber_mix_launch(st, saber_rank(scheme), d, scratch, saber_ipw);
  } else {
    const int set = frodo_set(scheme);
    uint32_t* s8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 8 * set));
    uint32_t* e8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 8 * set));
    uint32_t* epp8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 64));
    uint8_t* kbuf = c.take((size_t)n * 64);
    uint8_t* mubuf = c.take((size_t)n * 32);
    uint32_t* fail = reinterpret_cast<uint32_t*>(c.take((size_t)n * 4));
    uint32_t* state = reinterpret_cast<uint32_t*>(c.take((size_t)n * 200));
    const bool aes = scheme == DPKEM_FRODO976_AES;
    if (op == kKeygen) frodo_keygen_launch(st, n, i0, o0, o1, s8, e8, aes, set, frodo_threads);
    else if (op == kEnc) frodo_enc_launch(st, n, i0, i1, o0, o1, s8, e8, epp8, kbuf, frodo_pkh, aes, set, frodo_threads);
    else frodo_dec_launch(st, n, i0, i1, o0, s8, e8, epp8, kbuf, mubuf, fail, state, aes, set, frodo_threads);
  }
}

void launch_frodo_grouped(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
                          const uint8_t* i1, uint8_t* scratch, const uint2* groups, int ngroups,
                          const uint8_t* frodo_pkh) {
  Carver c{scratch};
  uint32_t* s8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 7808));
  uint32_t* e8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 7808));
  uint32_t* epp8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 64));
  uint8_t* kbuf = c.take((size_t)n * 64);
  uint8_t* mubuf = c.take((size_t)n * 32);
  uint32_t* fail = reinterpret_cast<uint32_t*>(c.take((size_t)n * 4));
  uint32_t* state = reinterpret_cast<uint32_t*>(c.take((size_t)n * 200));
  if (op == kEnc)
    frodo_enc_grouped_launch(st, n, i0, kSizes[DPKEM_FRODO976].pk, i1, o0, o1, s8, e8, epp8, kbuf, groups, ngroups,
                             frodo_pkh);
  else
    frodo_dec_grouped_launch(st, n, i0, i1, kSizes[DPKEM_FRODO976].sk, o0, s8, e8, epp8, kbuf, mubuf, fail, state,
                             groups, ngroups);
}

// Upper bound on the arena bytes the chunk carver uses for n instances
// (per-array alignment padding included).
size_t chunk_scratch_bytes(int scheme, int form, size_t n) { return n * scratch_per_inst(scheme, form) + 16 * kAlign; }

}  // namespace dpkem

namespace {

// Encaps / decaps of a chunk whose key input (ENC: pk, DEC: sk) is one row
// shared by all instances.  FrodoKEM-976-SHAKE: groups of kFrodoGroup
// instances expand A once.  The other schemes: the key row is broadcast into
// `bcast` and the ordinary kernels run.
void launch_chunk_shared(hipStream_t st, int scheme, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
                         const uint8_t* i1, uint8_t* scratch, uint8_t* bcast) {
  if (scheme != DPKEM_FRODO976) {
    if (op == kEnc) {
      broadcast_rows_launch(st, n, i0, kSizes[scheme].pk, bcast);
      launch_chunk_on(st, DPKEM_FORM_BATCHED, scheme, op, n, o0, o1, bcast, i1, scratch);
    } else {
      broadcast_rows_launch(st, n, i1, kSizes[scheme].sk, bcast);
      launch_chunk_on(st, DPKEM_FORM_BATCHED, scheme, op, n, o0, o1, i0, bcast, scratch);
    }
    return;
  }
  Carver c{scratch};
  uint32_t* s8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 7808));
  uint32_t* e8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 7808));
  uint32_t* epp8 = reinterpret_cast<uint32_t*>(c.take((size_t)n * 64));
  uint8_t* kbuf = c.take((size_t)n * 64);
  uint8_t* mubuf = c.take((size_t)n * 32);
  uint32_t* fail = reinterpret_cast<uint32_t*>(c.take((size_t)n * 4));
  uint32_t* state = reinterpret_cast<uint32_t*>(c.take((size_t)n * 200));
  if (op == kEnc) frodo_enc_grouped_launch(st, n, i0, 0, i1, o0, o1, s8, e8, epp8, kbuf, nullptr, 0);
  else frodo_dec_grouped_launch(st, n, i0, i1, 0, o0, s8, e8, epp8, kbuf, mubuf, fail, state, nullptr, 0);
}

// shared_in: -1, or the index of the input that is one row for all n
// instances (ENC: 0 = pk, DEC: 1 = sk).  form_override: -1 (the engine's
// form) or the DPKEM_FORM_* this call runs in, without changing the engine's.
int run_op(dpkem_engine* e, int scheme, int op, size_t n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
           const uint8_t* i1, int mem, int shared_in = -1, int form_override = -1) {
  if (!e || !valid_scheme(scheme)) return DPKEM_EINVAL;
  if (form_override != -1 && form_override != DPKEM_FORM_BATCHED && form_override != DPKEM_FORM_PLAIN)
    return DPKEM_EINVAL;
  if (mem != DPKEM_MEM_HOST && mem != DPKEM_MEM_DEVICE) return DPKEM_EINVAL;
  if (n == 0) return DPKEM_OK;
  size_t out[2], in[2];
  io_sizes(scheme, op, out, in);
  uint8_t* outs[2] = {o0, o1};
  const uint8_t* ins[2] = {i0, i1};
  for (int k = 0; k < 2; ++k) {
    if (out[k] && !outs[k]) return DPKEM_EINVAL;
    if (in[k] && !ins[k]) return DPKEM_EINVAL;
    if (mem == DPKEM_MEM_DEVICE && ((out[k] && ((uintptr_t)outs[k] & 7)) || (in[k] && ((uintptr_t)ins[k] & 7)))) {
      snprintf(e->last_err, sizeof e->last_err, "device pointers must be 8-byte aligned");
      return DPKEM_EINVAL;
    }
  }
  std::lock_guard<std::mutex> lock(e->mu);
  const int form = form_override >= 0 ? form_override : e->form;
  if (batched_only(scheme) && form != DPKEM_FORM_BATCHED) {
    snprintf(e->last_err, sizeof e->last_err, "this parameter set runs in the batched form only");
    return DPKEM_EINVAL;
  }
  if (shared_in >= 0 && form != DPKEM_FORM_BATCHED) {
    snprintf(e->last_err, sizeof e->last_err, "shared-key ops run in the batched form only");
    return DPKEM_EINVAL;
  }
  HIP_TRY(e, hipSetDevice(e->device));
  size_t io_per = 0;
  for (int k = 0; k < 2; ++k) {
    if (mem == DPKEM_MEM_HOST) io_per += align_up(out[k]) + (k == shared_in ? 0 : align_up(in[k]));
  }
  // Saber broadcasts the shared key into a dense array
  if (shared_in >= 0 && scheme != DPKEM_FRODO976) io_per += align_up(in[shared_in]);
  const size_t per = scratch_per_inst(scheme, form) + io_per;
  // a shared key staged from host memory is one row per chunk
  const size_t fixed = 24 * kAlign + (shared_in >= 0 && mem == DPKEM_MEM_HOST ? align_up(in[shared_in]) : 0);
  size_t chunk = std::min(n, (size_t)INT32_MAX);
  while (chunk > 1 && chunk * per + fixed > e->budget) chunk = (chunk + 1) / 2;
  if (chunk * per + fixed > e->budget) return DPKEM_EBUDGET;
  int rc = ensure_arena(e, chunk * per + fixed);
  if (rc) return rc;
  for (size_t done = 0; done < n; done += chunk) {
    const int c = (int)std::min(chunk, n - done);
    uint8_t* dout[2];
    const uint8_t* din[2];
    Carver io{e->arena + align_up(chunk_scratch_bytes(scheme, form, chunk))};
    for (int k = 0; k < 2; ++k) {
      const bool one = k == shared_in;
      if (mem == DPKEM_MEM_DEVICE) {
        dout[k] = out[k] ? outs[k] + done * out[k] : nullptr;
        din[k] = in[k] ? ins[k] + (one ? 0 : done * in[k]) : nullptr;
      } else {
        dout[k] = out[k] ? io.take((size_t)c * out[k]) : nullptr;
        const size_t rows = one ? 1 : (size_t)c;
        uint8_t* d = in[k] ? io.take(rows * in[k]) : nullptr;
        if (d) HIP_TRY(e, hipMemcpyAsync(d, ins[k] + (one ? 0 : done * in[k]), rows * in[k], hipMemcpyHostToDevice,
                                         e->stream));
        din[k] = d;
      }
    }
    tls_profiler = e->profiling ? &e->prof : nullptr;
    if (shared_in >= 0) {
      uint8_t* bcast = scheme != DPKEM_FRODO976 ? io.take((size_t)c * in[shared_in]) : nullptr;
      launch_chunk_shared(e->stream, scheme, op, c, dout[0], dout[1], din[0], din[1], e->arena, bcast);
    } else {
      launch_chunk_on(e->stream, form, scheme, op, c, dout[0], dout[1], din[0], din[1], e->arena, nullptr,
                      e->saber_ipw, e->frodo_threads);
    }
    tls_profiler = nullptr;
    HIP_TRY(e, hipGetLastError());
    if (mem == DPKEM_MEM_HOST) {
      for (int k = 0; k < 2; ++k)
        if (out[k])
          HIP_TRY(e, hipMemcpyAsync(outs[k] + done * out[k], dout[k], (size_t)c * out[k], hipMemcpyDeviceToHost, e->stream));
      HIP_TRY(e, hipStreamSynchronize(e->stream));
      if ((rc = check_waits(e)) != DPKEM_OK) return rc;
    }
  }
  return DPKEM_OK;
}

// ---- randomness for the PQClean-shaped and reference-shaped wrappers
std::mutex g_rng_mu;
dpkem_randombytes_fn g_rng = nullptr;
void* g_rng_ctx = nullptr;

void os_randombytes(uint8_t* out, size_t len, void*) {
  while (len > 0) {
    ssize_t r = getrandom(out, len, 0);
    if (r <= 0) continue;
    out += r;
    len -= (size_t)r;
  }
}

void randombytes(uint8_t* out, size_t len) {
  std::lock_guard<std::mutex> lock(g_rng_mu);
  (g_rng ? g_rng : os_randombytes)(out, len, g_rng_ctx);
}

std::mutex g_default_mu;
dpkem_engine* g_default = nullptr;

dpkem_engine* default_engine() {
  std::lock_guard<std::mutex> lock(g_default_mu);
  if (!g_default && dpkem_engine_create(&g_default, 0, 0) != DPKEM_OK) g_default = nullptr;
  return g_default;
}

int wrapped(int scheme, int op, size_t n, uint8_t* o0, uint8_t* o1, const uint8_t* i0, const uint8_t* i1,
            int form = DPKEM_FORM_BATCHED) {
  dpkem_engine* e = default_engine();
  if (!e) return DPKEM_EDEVICE;
  const SchemeSizes& s = kSizes[scheme];
  std::vector<uint8_t> coins;
  if (op == kKeygen) {
    coins.resize(n * s.kcoins);
    randombytes(coins.data(), coins.size());
    i0 = coins.data();
  } else if (op == kEnc) {
    coins.resize(n * s.ecoins);
    randombytes(coins.data(), coins.size());
    i1 = coins.data();
  }
  return run_op(e, scheme, op, n, o0, o1, i0, i1, DPKEM_MEM_HOST, -1, form);
}

}  // namespace

// =================================================================== C ABI
extern "C" {

int dpkem_sizes(int scheme, size_t* pk, size_t* sk, size_t* ct, size_t* ss, size_t* kc, size_t* ec) {
  if (!valid_scheme(scheme)) return DPKEM_EINVAL;
  const SchemeSizes& s = kSizes[scheme];
  if (pk) *pk = s.pk;
  if (sk) *sk = s.sk;
  if (ct) *ct = s.ct;
  if (ss) *ss = s.ss;
  if (kc) *kc = s.kcoins;
  if (ec) *ec = s.ecoins;
  return DPKEM_OK;
}

int dpkem_engine_create(dpkem_engine** out, int device, size_t budget) {
  if (!out) return DPKEM_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return DPKEM_EDEVICE;
  dpkem_engine* e = new (std::nothrow) dpkem_engine();
  if (!e) return DPKEM_ENOMEM;
  e->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return DPKEM_EDEVICE;
  }
  e->stream = e->own_stream;
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)8 << 30;
    budget = free_b / 4;
  }
  e->budget = budget;
  // bound of the kernels' LDS waits (DPKEM_WAIT_SPIN_LIMIT overrides it, e.g.
  // 0 to exercise the give-up path); the device value is shared by the process
  uint32_t limit = 1u << 22;
  if (const char* v = getenv("DPKEM_WAIT_SPIN_LIMIT")) limit = (uint32_t)strtoul(v, nullptr, 0);
  if (dpkem::saber_set_wait_spin_limit(limit) != hipSuccess || dpkem::frodo_set_wait_spin_limit(limit) != hipSuccess) {
    (void)hipStreamDestroy(e->own_stream);
    delete e;
    return DPKEM_EDEVICE;
  }
  *out = e;
  return DPKEM_OK;
}

int dpkem_engine_destroy(dpkem_engine* e) {
  if (!e) return DPKEM_EINVAL;
  {
    std::lock_guard<std::mutex> lock(e->mu);
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    if (e->arena) (void)hipFree(e->arena);
    if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
    for (auto& r : e->prof.pending) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.stop);
    }
    for (auto ev : e->prof.pool) (void)hipEventDestroy(ev);
  }
  delete e;
  return DPKEM_OK;
}

int dpkem_engine_set_stream(dpkem_engine* e, void* s, int own) {
  if (!e) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  hipStream_t next = own ? e->own_stream : reinterpret_cast<hipStream_t>(s);
  if (next == e->stream) return DPKEM_OK;
  HIP_TRY(e, hipSetDevice(e->device));
  HIP_TRY(e, hipStreamSynchronize(e->stream));  // the arena may still be in use there
  e->stream = next;
  return DPKEM_OK;
}

int dpkem_engine_set_form(dpkem_engine* e, int form) {
  if (!e || (form != DPKEM_FORM_BATCHED && form != DPKEM_FORM_PLAIN)) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  e->form = form;
  return DPKEM_OK;
}

int dpkem_engine_set_tuning(dpkem_engine* e, int knob, int value) {
  if (!e) return DPKEM_EINVAL;
  if (knob == DPKEM_TUNE_SABER_POLY_IPW && (value == 1 || value == 2 || value == 4)) {
    std::lock_guard<std::mutex> lock(e->mu);
    e->saber_ipw = value;
    return DPKEM_OK;
  }
  if (knob == DPKEM_TUNE_FRODO_PERSIST && (value == 0 || value == 1)) {
    std::lock_guard<std::mutex> lock(e->mu);
    e->frodo_persist = value;
    return DPKEM_OK;
  }
  if (knob == DPKEM_TUNE_FRODO_THREADS && (value == 0 || value == 1024 || value == 512)) {
    std::lock_guard<std::mutex> lock(e->mu);
    e->frodo_threads = value == 512 ? 0 : value;
    return DPKEM_OK;
  }
  return DPKEM_EINVAL;
}

int dpkem_engine_sync(dpkem_engine* e) {
  if (!e) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  HIP_TRY(e, hipSetDevice(e->device));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  return check_waits(e);
}

const char* dpkem_engine_last_error(dpkem_engine* e) { return e ? e->last_err : "no engine"; }

int dpkem_engine_set_profiling(dpkem_engine* e, int on) {
  if (!e) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  e->profiling = on != 0;
  return DPKEM_OK;
}

int dpkem_probe_perm_latency(dpkem_engine* e, int perms, float* ms) {
  if (!e || !ms || perms < 1 || perms > 100000) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  HIP_TRY(e, hipSetDevice(e->device));
  uint32_t* out = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t err = hipMalloc(&out, 64 * sizeof(uint32_t));
  if (err == hipSuccess) err = hipEventCreate(&t0);
  if (err == hipSuccess) err = hipEventCreate(&t1);
  if (err == hipSuccess) err = hipEventRecord(t0, e->stream);
  if (err == hipSuccess) {
    saber_probe_perm_launch(e->stream, perms, out);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipEventRecord(t1, e->stream);
  if (err == hipSuccess) err = hipEventSynchronize(t1);
  if (err == hipSuccess) err = hipEventElapsedTime(ms, t0, t1);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (out) (void)hipFree(out);
  if (err != hipSuccess) return record_hip_error(e, err, "dpkem_probe_perm_latency");
  return DPKEM_OK;
}

int dpkem_engine_kernel_stats(dpkem_engine* e, char* buf, size_t len, int reset) {
  if (!e || !buf || len == 0) return DPKEM_EINVAL;
  std::lock_guard<std::mutex> lock(e->mu);
  HIP_TRY(e, hipSetDevice(e->device));
  HIP_TRY(e, hipStreamSynchronize(e->stream));
  for (auto& r : e->prof.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess) {
      auto& s = e->kstats[r.name];
      s.first += 1;
      s.second += ms;
    }
    e->prof.pool.push_back(r.start);
    e->prof.pool.push_back(r.stop);
  }
  e->prof.pending.clear();
  std::string out = "{";
  for (auto& kv : e->kstats) {
    char item[256];
    snprintf(item, sizeof item, "%s\"%s\": [%ld, %.6f]", out.size() > 1 ? ", " : "", kv.first.c_str(),
             kv.second.first, kv.second.second);
    out += item;
  }
  out += "}";
  if (reset) e->kstats.clear();
  if (out.size() + 1 > len) return DPKEM_EINVAL;
  memcpy(buf, out.c_str(), out.size() + 1);
  return DPKEM_OK;
}

int dpkem_keypair(dpkem_engine* e, int scheme, size_t n, uint8_t* pk, uint8_t* sk, const uint8_t* coins, int mem) {
  return run_op(e, scheme, kKeygen, n, pk, sk, coins, nullptr, mem);
}

int dpkem_enc(dpkem_engine* e, int scheme, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
              const uint8_t* coins, int mem) {
  return run_op(e, scheme, kEnc, n, ct, ss, pk, coins, mem);
}

int dpkem_dec(dpkem_engine* e, int scheme, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk, int mem) {
  return run_op(e, scheme, kDec, n, ss, nullptr, ct, sk, mem);
}

int dpkem_enc_shared_key(dpkem_engine* e, int scheme, size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk,
                         const uint8_t* coins, int mem) {
  return run_op(e, scheme, kEnc, n, ct, ss, pk, coins, mem, 0);
}

int dpkem_dec_shared_key(dpkem_engine* e, int scheme, size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk,
                         int mem) {
  return run_op(e, scheme, kDec, n, ss, nullptr, ct, sk, mem, 1);
}

int dpkem_set_randombytes(dpkem_randombytes_fn fn, void* ctx) {
  std::lock_guard<std::mutex> lock(g_rng_mu);
  g_rng = fn;
  g_rng_ctx = ctx;
  return DPKEM_OK;
}

int dpkem_frodo976_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_FRODO976, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_frodo976_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_FRODO976, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_frodo976_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_FRODO976, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_frodo976aes_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_FRODO976_AES, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_frodo976aes_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_FRODO976_AES, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_frodo976aes_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_FRODO976_AES, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_lightsaber_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_LIGHTSABER, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_lightsaber_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_LIGHTSABER, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_lightsaber_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_LIGHTSABER, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_firesaber_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_FIRESABER, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_firesaber_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_FIRESABER, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_firesaber_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_FIRESABER, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_frodo640_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_FRODO640, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_frodo640_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_FRODO640, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_frodo640_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_FRODO640, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_frodo1344_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_FRODO1344, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_frodo1344_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_FRODO1344, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_frodo1344_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_FRODO1344, kDec, n, ss, nullptr, ct, sk);
}
int dpkem_saber_crypto_kem_keypair(size_t n, uint8_t* pk, uint8_t* sk) {
  return wrapped(DPKEM_SABER, kKeygen, n, pk, sk, nullptr, nullptr);
}
int dpkem_saber_crypto_kem_enc(size_t n, uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  return wrapped(DPKEM_SABER, kEnc, n, ct, ss, pk, nullptr);
}
int dpkem_saber_crypto_kem_dec(size_t n, uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  return wrapped(DPKEM_SABER, kDec, n, ss, nullptr, ct, sk);
}

int frodoKEMKeypair(uint8_t* h_pk, uint8_t* h_sk) {
  if (!h_pk || !h_sk) return DPKEM_EINVAL;
  return wrapped(DPKEM_FRODO976, kKeygen, DPKEM_REF_BATCH, h_pk, h_sk, nullptr, nullptr);
}
int frodoKEMEncap(uint8_t* ct, uint8_t* ss, const uint8_t* pk) {
  if (!ct || !ss || !pk) return DPKEM_EINVAL;
  return wrapped(DPKEM_FRODO976, kEnc, DPKEM_REF_BATCH, ct, ss, pk, nullptr);
}
int frodoKEMDecap(uint8_t* ss, const uint8_t* ct, const uint8_t* sk) {
  if (!ss || !ct || !sk) return DPKEM_EINVAL;
  return wrapped(DPKEM_FRODO976, kDec, DPKEM_REF_BATCH, ss, nullptr, ct, sk);
}
int saber_enc(uint8_t mode, uint8_t* h_k, uint8_t* h_pk, uint8_t* h_c) {
  if (!h_k || !h_pk || !h_c || mode > 1) return DPKEM_EINVAL;
  return wrapped(DPKEM_SABER, kEnc, DPKEM_REF_BATCH, h_c, h_k, h_pk, nullptr,
                 mode == 0 ? DPKEM_FORM_BATCHED : DPKEM_FORM_PLAIN);
}
int saber_dec(uint8_t mode, uint8_t* h_c, uint8_t* h_pk, uint8_t* h_sk, uint8_t* h_k) {
  (void)h_pk;  // the public key is embedded in the secret key
  if (!h_c || !h_sk || !h_k || mode > 1) return DPKEM_EINVAL;
  return wrapped(DPKEM_SABER, kDec, DPKEM_REF_BATCH, h_k, nullptr, h_c, h_sk,
                 mode == 0 ? DPKEM_FORM_BATCHED : DPKEM_FORM_PLAIN);
}

}  // extern "C"
