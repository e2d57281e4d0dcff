// FrodoKEM error sampler tables, shared by the device code (frodo.hip) and
// the host test that checks them (tests/host/sampler_emul.cpp).
//
// The CDF sampler (FrodoKEM976-GPU/src/sample_pack.cu:13-32) maps a 16-bit
// draw r to magnitude e = #{j : cdf[j] < r >> 1} and sign r & 1.  e depends
// on the 15 bits r >> 1 only, so the serial chains read it from a 32 KB table
// (one byte per value) built at compile time as the step function
// e = k on cdf[k-1] < t <= cdf[k].
#pragma once
#include <stdint.h>

namespace dpkem {

// CDF tables without their final entry: FrodoKEM976-GPU/include/params.h:24
// for 976, the round-3 specification's tables for 640 / 1344.
template <int N>
struct FrodoCdf;
template <>
struct FrodoCdf<640> {
  static constexpr int ncdf = 12;
  static constexpr uint16_t v[ncdf] = {4643, 13363, 20579, 25843, 29227, 31145,
                                       32103, 32525, 32689, 32745, 32762, 32766};
};
template <>
struct FrodoCdf<976> {
  static constexpr int ncdf = 10;
  static constexpr uint16_t v[ncdf] = {5638, 15915, 23689, 28571, 31116, 32217, 32613, 32731, 32760, 32766};
};
template <>
struct FrodoCdf<1344> {
  static constexpr int ncdf = 6;
  static constexpr uint16_t v[ncdf] = {9142, 23462, 30338, 32361, 32725, 32765};
};

template <int N>
struct alignas(16) SampleLut {
  uint8_t e[32768];
};
template <int N>
constexpr SampleLut<N> make_sample_lut() {
  using C = FrodoCdf<N>;
  SampleLut<N> l{};
  int t = 0;
  for (int k = 0; k <= C::ncdf; ++k) {
    const int hi = k < C::ncdf ? C::v[k] : 32767;
    for (; t <= hi; ++t) l.e[t] = (uint8_t)k;
  }
  return l;
}

}  // namespace dpkem
