// Engine internals shared by engine.cpp (batched ops, C ABI), plan.cpp
// (graph-captured task sets) and server.cpp (serving loop).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "../../include/dpkem.h"
#include "launch.hpp"
#include "prof.hpp"

struct dpkem_engine {
  int device = 0;
  size_t budget = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int form = DPKEM_FORM_BATCHED;
  int saber_ipw = 1;  // DPKEM_TUNE_SABER_POLY_IPW
  int frodo_persist = 1;  // DPKEM_TUNE_FRODO_PERSIST
  int frodo_threads = 0;  // DPKEM_TUNE_FRODO_THREADS (0 = the set's default shape, 1024 = FrodoKEM-976's other)
  uint8_t* arena = nullptr;
  size_t arena_bytes = 0;
  char last_err[256] = {0};
  std::mutex mu;
  bool profiling = false;
  dpkem::Profiler prof;
  std::map<std::string, std::pair<long, double>> kstats;  // name -> (launches, total ms)
};


namespace dpkem {

enum Op { kKeygen = 0, kEnc = 1, kDec = 2 };

// Per-instance scratch of the current form (bytes, 256-aligned).
size_t scratch_per_inst(int scheme, int form);
// Upper bound of the scratch one chunk of n instances carves.
size_t chunk_scratch_bytes(int scheme, int form, size_t n);
// I/O byte sizes per instance: outputs out[0..1], inputs in[0..1] (0 = unused).
void io_sizes(int scheme, int op, size_t out[2], size_t in[2]);
// Launch one chunk (all pointers on the device) on stream st using `scratch`.
// Frodo encaps / decaps of a chunk whose instances come in same-key groups:
// groups[b] = (first, count <= kFrodoGroup), instances of a group carry
// identical keys (i0 = pk rows for ENC, i1 = sk rows for DEC).
// frodo_pkh (ENC only): nullptr, or one cached 24-byte pkh row per instance.
void launch_frodo_grouped(hipStream_t st, int op, int n, uint8_t* o0, uint8_t* o1, const uint8_t* i0,
                          const uint8_t* i1, uint8_t* scratch, const uint2* groups, int ngroups,
                          const uint8_t* frodo_pkh = nullptr);
// saber_ipw: Saber-family product instances per workgroup (DPKEM_TUNE_SABER_POLY_IPW).
void launch_chunk_on(hipStream_t st, int form, int scheme, int op, int n, uint8_t* o0, uint8_t* o1,
                     const uint8_t* i0, const uint8_t* i1, uint8_t* scratch, const uint8_t* frodo_pkh = nullptr,
                     int saber_ipw = 1, int frodo_threads = 0);
int record_hip_error(dpkem_engine* e, hipError_t err, const char* what);

}  // namespace dpkem
