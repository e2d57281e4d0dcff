// A host-only stand-in for the HIP runtime API that dpcrypto_amd/csrc/plan.cpp
// uses, for the plan-bookkeeping harness (plan_harness.cpp).  Device memory is
// host memory; streams, events and graphs are records that check the rules of
// stream capture the way the runtime enforces them:
//   - work issued during a capture must go to a stream that has joined it
//     (the origin, or a stream that waited on an event recorded in it);
//   - an event waited on during a capture must have been recorded in it;
//   - at the end of the capture every joined stream's last work must have
//     been waited on by the origin (no unjoined work);
//   - destroyed streams / events / graphs are never used again.
// Any violation aborts with a message, so the harness fails loudly.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <set>
#include <vector>

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
  hipErrorUnknown = 999,
} hipError_t;

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint2 {
  unsigned x, y;
};

namespace fakehip {
struct Event;
struct Stream {
  bool alive = true;
  bool captured = false;
  long last_op = 0;      // sequence number of the stream's last operation in the capture
  long joined_upto = 0;  // last_op that the origin has waited for
};
struct Event {
  bool alive = true;
  bool recorded_in_capture = false;
  Stream* stream = nullptr;
  long op = 0;  // sequence number of the recording
};
struct State {
  std::set<Stream*> streams;
  std::set<Event*> events;
  Stream* origin = nullptr;  // capture in progress when non-null
  long seq = 0;
  long graphs = 0, launches_captured = 0;
  // destroyed records stay allocated until the process ends, so that a use
  // after destruction is detected above instead of reading freed memory
  ~State() {
    for (Stream* x : streams) delete x;
    for (Event* x : events) delete x;
  }
};
inline State& st() {
  static State s;
  return s;
}
[[noreturn]] inline void fail(const char* what) {
  fprintf(stderr, "fake HIP: %s\n", what);
  abort();
}
inline void check_stream(Stream* s) {
  if (s == nullptr) return;  // the null stream
  if (!st().streams.count(s) || !s->alive) fail("use of a destroyed or unknown stream");
}
// Record one operation (launch / copy / record) on stream s.
inline void op_on(Stream* s) {
  check_stream(s);
  if (st().origin) {
    if (!s || !s->captured) fail("work issued during a capture on a stream outside it");
    s->last_op = ++st().seq;
    st().launches_captured++;
  }
}
}  // namespace fakehip

typedef fakehip::Stream* hipStream_t;
typedef fakehip::Event* hipEvent_t;
typedef struct fake_graph* hipGraph_t;
typedef struct fake_graph_exec* hipGraphExec_t;
typedef enum { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1 } hipStreamCaptureMode;
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u

inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "fake"; }
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), n);
}
inline hipError_t hipFree(void* p) {
  free(p);
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  if (fakehip::st().origin) fakehip::fail("stream created during a capture");
  *s = new fakehip::Stream();
  fakehip::st().streams.insert(*s);
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  fakehip::check_stream(s);
  if (fakehip::st().origin == s) fakehip::fail("capturing stream destroyed");
  s->alive = false;  // kept allocated so that later use is detected
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t s) {
  fakehip::check_stream(s);
  if (fakehip::st().origin) fakehip::fail("stream synchronized during a capture");
  return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  *e = new fakehip::Event();
  fakehip::st().events.insert(*e);
  return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t* e) { return hipEventCreateWithFlags(e, 0); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
  if (!fakehip::st().events.count(e) || !e->alive) fakehip::fail("destroy of a destroyed or unknown event");
  e->alive = false;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (!fakehip::st().events.count(e) || !e->alive) fakehip::fail("record of a destroyed or unknown event");
  fakehip::op_on(s);
  e->stream = s;
  e->recorded_in_capture = fakehip::st().origin != nullptr;
  e->op = s ? s->last_op : 0;
  return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  using namespace fakehip;
  if (!st().events.count(e) || !e->alive) fail("wait on a destroyed or unknown event");
  check_stream(s);
  if (st().origin) {
    if (!e->recorded_in_capture) fail("wait during a capture on an event recorded outside it");
    if (!s) fail("null stream waits during a capture");
    s->captured = true;  // joins the capture
    if (s == st().origin && e->stream) e->stream->joined_upto = std::max(e->stream->joined_upto, e->op);
    s->last_op = ++st().seq;
  }
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  using namespace fakehip;
  check_stream(s);
  if (st().origin) fail("nested capture");
  st().origin = s;
  for (Stream* x : st().streams) x->captured = false, x->last_op = 0, x->joined_upto = 0;
  s->captured = true;
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  using namespace fakehip;
  if (st().origin != s) fail("end of a capture on another stream");
  for (Stream* x : st().streams)
    if (x != s && x->captured && x->last_op > x->joined_upto) fail("capture ended with unjoined work");
  st().origin = nullptr;
  st().graphs++;
  *g = reinterpret_cast<hipGraph_t>(new char[1]);
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, void*, void*, size_t) {
  if (!g) fakehip::fail("instantiate of a null graph");
  *x = reinterpret_cast<hipGraphExec_t>(new char[1]);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
  if (!x) fakehip::fail("launch of a null graph");
  fakehip::op_on(s);
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete[] reinterpret_cast<char*>(g);
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete[] reinterpret_cast<char*>(x);
  return hipSuccess;
}

// engine.cpp also needs these
inline hipError_t hipGetDeviceCount(int* n) {
  *n = 1;
  return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* free_b, size_t* total_b) {
  *free_b = (size_t)64 << 30;
  *total_b = (size_t)64 << 30;
  return hipSuccess;
}
typedef enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 } hipMemcpyKind;
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  fakehip::op_on(st);
  memmove(d, s, n);
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) {
  if (fakehip::st().origin) fakehip::fail("event synchronized during a capture");
  return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 0.f;
  return hipSuccess;
}
