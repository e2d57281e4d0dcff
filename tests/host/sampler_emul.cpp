// The Frodo sampler magnitude tables of dpcrypto_amd/csrc/frodo_sampler.hpp,
// applied as the device does (e = table[r >> 1], value = (e ^ -s) + s with
// s = r & 1, mod 2^16), must equal the oracle's CDF sampler on all 65,536
// draws of each parameter set.  Prints "OK" on success.
#include <cstdint>
#include <cstdio>

#include "../../dpcrypto_amd/csrc/frodo_sampler.hpp"

extern "C" void orc_frodo_n_sample(int n, uint16_t* s, size_t count);

template <int N>
static int check() {
  static const dpkem::SampleLut<N> lut = dpkem::make_sample_lut<N>();
  static uint16_t want[65536];
  for (uint32_t r = 0; r < 65536; ++r) want[r] = (uint16_t)r;
  orc_frodo_n_sample(N, want, 65536);
  int bad = 0;
  for (uint32_t r = 0; r < 65536; ++r) {
    const uint32_t e = lut.e[r >> 1], s = r & 1;
    const uint16_t got = (uint16_t)((e ^ (0u - s)) + s);
    if (got != want[r] && bad++ < 4) std::printf("n=%d r=%u got %u want %u\n", N, r, got, want[r]);
  }
  return bad;
}

int main() {
  const int bad = check<640>() + check<976>() + check<1344>();
  std::printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
