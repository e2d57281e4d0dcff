// Host emulation of the lane-pair Keccak (dpcrypto_amd/csrc/keccak2.hpp):
// the same round template runs over a two-lane vector whose swap() exchanges
// the lanes, and must equal the oracle's Keccak-f[1600] on the full state.
// Also checks the bit (de)interleave helpers.  Prints "OK" on success.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../dpcrypto_amd/csrc/keccak2.hpp"

extern "C" void orc_keccak_f1600(uint64_t*);

struct V2 {
  uint32_t v[2];
};
struct HostOps {
  static V2 xor3(V2 a, V2 b, V2 c) { return {{a.v[0] ^ b.v[0] ^ c.v[0], a.v[1] ^ b.v[1] ^ c.v[1]}}; }
  static V2 chi(V2 a, V2 b, V2 c) { return {{a.v[0] ^ (~b.v[0] & c.v[0]), a.v[1] ^ (~b.v[1] & c.v[1])}}; }
  static V2 xor2(V2 a, V2 b) { return {{a.v[0] ^ b.v[0], a.v[1] ^ b.v[1]}}; }
  static uint32_t rol(uint32_t x, uint32_t n) { n &= 31; return n ? (x << n) | (x >> (32 - n)) : x; }
  template <int N>
  static V2 rotc(V2 x) { return {{rol(x.v[0], N), rol(x.v[1], N)}}; }
  // shift operand as on the device: alignbit(x, x, s) == rotate right by s == rotate left by 32 - s
  static V2 rotv(V2 x, V2 s) { return {{rol(x.v[0], 32 - (s.v[0] & 31)), rol(x.v[1], 32 - (s.v[1] & 31))}}; }
  static V2 swap(V2 x) { return {{x.v[1], x.v[0]}}; }
};

int main() {
  using namespace dpkem::k2;
  int bad = 0;
  uint64_t st[25], ref[25];
  for (int i = 0; i < 25; ++i) st[i] = ref[i] = 0x9E3779B97F4A7C15ULL * (i + 3) ^ ((uint64_t)i << 50);
  // lane-pair state
  V2 a[25];
  for (int i = 0; i < 25; ++i) a[i] = {{deint(st[i], 0), deint(st[i], 1)}};
  V2 sh[32];
  for (int k = 0; k < 32; ++k) sh[k] = {{32u - k + 0u, 32u - k + 1u}};
  for (int rep = 0; rep < 3; ++rep) {
    for (int r = 0; r < 24; ++r) {
      V2 rc = {{compress_even(kRCs[r]), compress_even(kRCs[r] >> 1)}};
      round2<HostOps>(a, sh, rc);
    }
    orc_keccak_f1600(ref);
  }
  for (int i = 0; i < 25; ++i)
    if (reint_full(a[i].v[0], a[i].v[1]) != ref[i]) ++bad;
  // helpers: deint / reint round trip on assorted words
  for (uint64_t w = 1; w < (1ULL << 63); w = w * 3 + 0x1234567) {
    uint32_t e = deint(w, 0), o = deint(w, 1);
    if (reint_full(e, o) != w) ++bad;
    if ((uint64_t)reint_half(e, o, 0) != (w & 0xFFFFFFFFu) || (uint64_t)reint_half(e, o, 1) != (w >> 32)) ++bad;
  }
  printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
