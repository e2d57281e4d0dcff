// Allocation-free keygen / encaps / decaps (dpkem_run_noalloc).
//
// The engine's entry points take their kernel scratch (GenMatrix / GenSecret
// streams, noise and message buffers, FO flags) from an arena they grow with
// hipMalloc.  Callers that cannot allocate -- code under HIP graph capture,
// tight serving loops, harnesses that forbid it -- use this entry point
// instead: the same kernels, launched exactly as dpkem_keypair / enc / dec
// launch them for device-resident arrays, with their scratch carved from one
// file-scope device array of fixed size.
//
// Each GPU holds its own copy of `scratch` (at its own device address) until
// the process ends.  The host-side name `scratch` is not a device address;
// its address on a device comes from hipGetSymbolAddress, taken once per
// device with that device current, on the first call that runs there.
// Every call on one GPU shares that GPU's copy and leaves its contents
// behind; the kernels write every scratch byte they read, so nothing has to
// start at zero.  Calls that overlap on one GPU (different streams or host
// threads) race on the scratch: callers serialise them.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_internal.hpp"

namespace {

static __device__ __align__(256) uint8_t scratch[DPKEM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices] = {nullptr};

// The device address of `scratch` on `device` (nullptr on a HIP error).
uint8_t* scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device]) return g_ptr[device];
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return nullptr;
  if (prev != device && hipSetDevice(device) != hipSuccess) return nullptr;
  void* p = nullptr;
  const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (err != hipSuccess || p == nullptr) return nullptr;
  g_ptr[device] = static_cast<uint8_t*>(p);
  return g_ptr[device];
}

}  // namespace

namespace dpkem {

size_t noalloc_max_instances(int scheme) {
  if (scheme < DPKEM_FRODO976 || scheme > DPKEM_FRODO1344) return 0;
  const size_t per = scratch_per_inst(scheme, DPKEM_FORM_BATCHED);
  const size_t fixed = chunk_scratch_bytes(scheme, DPKEM_FORM_BATCHED, 0);
  return (DPKEM_NOALLOC_SCRATCH_BYTES - fixed) / per;
}

}  // namespace dpkem

extern "C" int dpkem_run_noalloc(int device, void* hip_stream, int scheme, int op, size_t n, uint8_t* o0, uint8_t* o1,
                                 const uint8_t* i0, const uint8_t* i1) {
  using namespace dpkem;
  if (scheme < DPKEM_FRODO976 || scheme > DPKEM_FRODO1344 || op < kKeygen || op > kDec) return DPKEM_EINVAL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return DPKEM_EDEVICE;
  if (device < 0 || device >= count || device >= kMaxDevices) return DPKEM_EINVAL;
  if (n == 0) return DPKEM_OK;
  if (n > noalloc_max_instances(scheme)) return DPKEM_EBUDGET;
  size_t out[2], in[2];
  io_sizes(scheme, op, out, in);
  uint8_t* outs[2] = {o0, o1};
  const uint8_t* ins[2] = {i0, i1};
  for (int k = 0; k < 2; ++k) {
    if ((out[k] && (!outs[k] || ((uintptr_t)outs[k] & 7))) || (in[k] && (!ins[k] || ((uintptr_t)ins[k] & 7))))
      return DPKEM_EINVAL;
  }
  uint8_t* s = scratch_on(device);
  if (!s) return DPKEM_EDEVICE;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return DPKEM_EDEVICE;
  if (prev != device && hipSetDevice(device) != hipSuccess) return DPKEM_EDEVICE;
  (void)hipGetLastError();  // start from a clean per-thread error state
  launch_chunk_on(static_cast<hipStream_t>(hip_stream), DPKEM_FORM_BATCHED, scheme, op, (int)n, o0, o1, i0, i1, s);
  const hipError_t err = hipGetLastError();
  if (prev != device) (void)hipSetDevice(prev);
  return err == hipSuccess ? DPKEM_OK : DPKEM_EDEVICE;
}
