// Host emulation of the int8 MFMA Saber products (saber_poly.hpp: toep8_*):
// the staged Toeplitz copies, the signed limb split, the B-operand slots and
// the diagonal tile walk of toep8_block are run for all 64 lanes of a wave,
// each matrix instruction is replayed as the 16 x 16 x 64 int8 product it
// stands for, and the combined sums are compared with the oracle's negacyclic
// products for the three product shapes (keygen A^T s, encaps A s' with the
// inner product b . s').  Prints "OK" on success.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#define __host__
#define __device__
#define __forceinline__ inline
struct uint4 { uint32_t x, y, z, w; };
typedef unsigned short hus2 __attribute__((ext_vector_type(2)));
static inline uint32_t fake_udot2(hus2 a, hus2 b, uint32_t c, bool) { return c + (uint32_t)a.x * b.x + (uint32_t)a.y * b.y; }
#define __builtin_amdgcn_udot2 fake_udot2
#include "../../dpcrypto_amd/csrc/saber_poly.hpp"

using dpkem::v4i8x4;
extern "C" void orc_saber_polymul_acc(const uint16_t*, const uint16_t*, uint16_t*);

struct Call { v4i8x4 a, b; };

// D[m][n] += sum over (q, t) of A(lane 16q+m).byte t * B(lane 16q+n).byte t;
// lane (g, n) element i holds D[4g + i][n].
static void mfma_wave(v4i8x4 acc[64], const v4i8x4 A[64], const v4i8x4 B[64]) {
  int32_t D[16][16] = {};
  for (int m = 0; m < 16; ++m)
    for (int n = 0; n < 16; ++n) {
      int32_t s = 0;
      for (int q = 0; q < 4; ++q)
        for (int t = 0; t < 16; ++t) {
          const int8_t a = (int8_t)(((uint32_t)A[16 * q + m][t / 4]) >> (8 * (t % 4)));
          const int8_t b = (int8_t)(((uint32_t)B[16 * q + n][t / 4]) >> (8 * (t % 4)));
          s += a * b;
        }
      D[m][n] = s;
    }
  for (int l = 0; l < 64; ++l)
    for (int i = 0; i < 4; ++i) acc[l][i] += D[4 * (l >> 4) + i][l & 15];
}

static int run(int L, int mode, int trial) {  // mode 0 keygen (A^T s), 1 encaps (A s', b . s')
  const bool inner = mode != 0;
  std::vector<uint16_t> s(L * 256), A(L * L * 256), bb(L * 256);
  for (auto& x : s) x = (uint16_t)((rand() % 11) - 5);
  for (auto& x : A) x = (uint16_t)(trial == 0 ? 0x1FFF : rand() & 0x1FFF);
  for (auto& x : bb) x = (uint16_t)(trial == 0 ? 0x3FF : rand() & 0x3FF);
  std::vector<int8_t> V(L * dpkem::kVPoly, 0x55), Lb(L * 2 * (L + 1) * 256, 0);
  for (int j = 0; j < L; ++j)
    for (int c = 0; c < 256; ++c) dpkem::toep8_put(&V[j * dpkem::kVPoly], c, (int16_t)s[j * 256 + c]);
  auto lb = [&](int j, int slot, int k) -> int8_t& { return Lb[((j * 2 * (L + 1)) + slot) * 256 + k]; };
  for (int r = 0; r < L; ++r)
    for (int c = 0; c < L; ++c) {
      const int out = mode == 0 ? c : r, j = mode == 0 ? r : c;
      for (int k = 0; k < 256; ++k) dpkem::limbs8(A[(r * L + c) * 256 + k], &lb(j, out, k), &lb(j, L + 1 + out, k));
    }
  if (inner)
    for (int j = 0; j < L; ++j)
      for (int k = 0; k < 256; ++k) dpkem::limbs8(bb[j * 256 + k], &lb(j, L, k), &lb(j, 2 * L + 1, k));
  v4i8x4 acc[16][64] = {};
  for (int j = 0; j < L; ++j) {
    // per lane: record the operands toep8_block hands to the matrix instruction
    std::vector<std::vector<Call>> calls(64);
    for (int lane = 0; lane < 64; ++lane) {
      const int q = lane >> 4, slot = dpkem::toep8_slot(lane & 15, L, inner);
      v4i8x4 bop[4];
      for (int b = 0; b < 4; ++b) {
        bop[b] = v4i8x4{0, 0, 0, 0};
        if (slot >= 0) memcpy(&bop[b], &lb(j, slot, 64 * b + 16 * q), 16);
      }
      v4i8x4 dummy[16] = {};
      dpkem::toep8_block(dummy, &V[j * dpkem::kVPoly], bop, lane, [&](v4i8x4 a, v4i8x4 b, v4i8x4 c) {
        calls[lane].push_back({a, b});
        return c;
      });
    }
    // replay in instruction order; the target tile follows the same walk
    size_t idx = 0;
    for (int di = 0; di < dpkem::kToepDiagonals; ++di)
      for (int b = 0; b < 4; ++b) {
        const int a = dpkem::toep8_tile(di, b);
        if (a < 0) continue;
        v4i8x4 Aop[64], Bop[64];
        for (int l = 0; l < 64; ++l) { Aop[l] = calls[l][idx].a; Bop[l] = calls[l][idx].b; }
        mfma_wave(acc[a], Aop, Bop);
        ++idx;
      }
    for (int l = 0; l < 64; ++l)
      if (calls[l].size() != idx) return 1000;
  }
  // combine limbs and compare with the oracle
  int bad = 0;
  const int ncol = inner ? L + 1 : L;
  for (int n = 0; n < ncol; ++n) {
    uint16_t ref[256] = {};
    for (int j = 0; j < L; ++j) {
      const uint16_t* x = n < L ? &A[(mode == 0 ? j * L + n : n * L + j) * 256] : &bb[j * 256];
      orc_saber_polymul_acc(x, &s[j * 256], ref);
    }
    for (int t = 0; t < 256; ++t) {
      const int a = t >> 4, r = t & 15, g = r >> 2, i = r & 3;
      const int32_t lo = acc[a][16 * g + n][i], hi = acc[a][16 * g + 8 + n][i];
      if ((uint16_t)(lo + 256 * hi) != ref[t]) ++bad;
    }
  }
  return bad;
}

int main() {
  srand(7);
  int bad = 0;
  for (int L = 2; L <= 4; ++L)
    for (int mode = 0; mode < 2; ++mode)
      for (int trial = 0; trial < 3; ++trial) {
        const int b = run(L, mode, trial);
        if (b) printf("L=%d mode=%d trial=%d: %d mismatches\n", L, mode, trial, b);
        bad += b;
      }
  // limb split covers the whole 13-bit range exactly
  for (uint32_t a = 0; a < 8192; ++a) {
    int8_t lo, hi;
    dpkem::limbs8(a, &lo, &hi);
    if (256 * hi + lo != (int)a || hi < 0 || hi > 32) ++bad;
  }
  printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad != 0;
}
