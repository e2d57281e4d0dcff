// Host emulation of the Saber Toeplitz product helpers
// (dpcrypto_amd/csrc/saber_poly.hpp): toep_put + toep_dot over all 32
// thread groups must equal the oracle's negacyclic product, for small signed
// and for arbitrary 16-bit second factors; the 13/10-bit unpackers must
// match a bit-by-bit extraction.  Prints "OK" on success.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#define __host__
#define __device__
#define __forceinline__ inline
struct uint4 { uint32_t x, y, z, w; };
typedef unsigned short hus2 __attribute__((ext_vector_type(2)));
static inline uint32_t fake_udot2(hus2 a, hus2 b, uint32_t c, bool) {
  return c + (uint32_t)a.x * b.x + (uint32_t)a.y * b.y;
}
#define __builtin_amdgcn_udot2 fake_udot2
#include "../../dpcrypto_amd/csrc/saber_poly.hpp"

extern "C" void orc_saber_polymul_acc(const uint16_t*, const uint16_t*, uint16_t*);

int main() {
  alignas(16) static uint16_t T[3 * dpkem::kToepStride];
  alignas(16) static uint16_t S[256];
  uint16_t a[256], s[256], ref[256];
  srand(1);
  int bad = 0;
  for (int trial = 0; trial < 40; ++trial) {
    memset(T, 0xAB, sizeof T);
    for (int i = 0; i < 256; ++i) {
      a[i] = (uint16_t)((rand() % 9) - 4);      // staged factor: small secret
      s[i] = (uint16_t)(rand() & 0x1FFF);       // plain factor: 13-bit
      if (trial & 1) { a[i] = (uint16_t)rand(); s[i] = (uint16_t)rand(); }
    }
    for (int q = 0; q < 256; ++q) dpkem::toep_put(T, 1, q, a[q]);
    memcpy(S, s, sizeof s);
    memset(ref, 0, sizeof ref);
    orc_saber_polymul_acc(a, s, ref);
    for (int g = 0; g < 32; ++g) {
      uint32_t acc[8] = {0};
      dpkem::toep_dot(acc, T, 1, S, 8 * g);
      for (int d = 0; d < 8; ++d)
        if ((uint16_t)acc[d] != ref[8 * g + d]) bad++;
    }
  }
  uint8_t b[13];
  for (int rep = 0; rep < 100; ++rep) {
    for (int i = 0; i < 13; ++i) b[i] = (uint8_t)rand();
    uint16_t v[8];
    dpkem::unpack13_group(v, b);
    for (int i = 0; i < 8; ++i) {
      uint32_t x = 0;
      for (int k = 0; k < 13; ++k) x |= (uint32_t)((b[(13 * i + k) / 8] >> ((13 * i + k) % 8)) & 1) << k;
      if (x != v[i]) bad++;
    }
    uint16_t w[4];
    dpkem::unpack10_group(w, b);
    for (int i = 0; i < 4; ++i) {
      uint32_t x = 0;
      for (int k = 0; k < 10; ++k) x |= (uint32_t)((b[(10 * i + k) / 8] >> ((10 * i + k) % 8)) & 1) << k;
      if (x != w[i]) bad++;
    }
  }
  printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
