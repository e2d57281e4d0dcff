/*
 * AES-128 encryption (FIPS-197) for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Used by FrodoKEM-976-AES, whose matrix A is AES-128 under seedA of the
 * blocks (row, column, 0, ..., 0); the reference only declares the AES it
 * would need (FrodoKEM976-GPU/include/aes.h:33-39,
 * FrodoKEM976-GPU/include/gen_a_mat_mul.cuh:6).  Byte-oriented and
 * table-free: the S-box is evaluated from exponent / logarithm tables of the
 * generator 3 of GF(2^8)* and the affine map.
 */
#include <string.h>

#include "oracle.h"

static uint8_t mul2(uint8_t a) { return (uint8_t)((a << 1) ^ ((a >> 7) * 0x1B)); }

static uint8_t exp3[256], log3[256];

/* filled once at library load, before any thread can call in */
__attribute__((constructor)) static void build_logs(void) {
  uint8_t v = 1;
  for (int i = 0; i < 255; ++i) {
    exp3[i] = v;
    log3[v] = (uint8_t)i;
    v ^= mul2(v); /* v *= 3 */
  }
}

static uint8_t sbox_of(uint8_t x) {
  uint8_t inv = x ? exp3[(255 - log3[x]) % 255] : 0;
  uint8_t r = inv;
  for (int k = 1; k < 5; ++k) r ^= (uint8_t)((inv << k) | (inv >> (8 - k)));
  return (uint8_t)(r ^ 0x63);
}

void orc_aes128_ecb(const uint8_t key[16], const uint8_t *in, uint8_t *out, size_t nblocks) {
  uint8_t rk[176];
  memcpy(rk, key, 16);
  uint8_t rcon = 1;
  for (int i = 16; i < 176; i += 4) {
    uint8_t t[4] = {rk[i - 4], rk[i - 3], rk[i - 2], rk[i - 1]};
    if (i % 16 == 0) {
      uint8_t t0 = t[0];
      t[0] = (uint8_t)(sbox_of(t[1]) ^ rcon);
      t[1] = sbox_of(t[2]);
      t[2] = sbox_of(t[3]);
      t[3] = sbox_of(t0);
      rcon = mul2(rcon);
    }
    for (int k = 0; k < 4; ++k) rk[i + k] = (uint8_t)(rk[i - 16 + k] ^ t[k]);
  }
  for (size_t b = 0; b < nblocks; ++b) {
    uint8_t s[16];
    for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(in[16 * b + i] ^ rk[i]);
    for (int r = 1; r <= 10; ++r) {
      uint8_t t[16];
      for (int c = 0; c < 4; ++c) /* SubBytes, ShiftRows: row j moves left by j columns */
        for (int j = 0; j < 4; ++j) t[4 * c + j] = sbox_of(s[4 * ((c + j) % 4) + j]);
      if (r < 10)
        for (int c = 0; c < 4; ++c) { /* MixColumns with the matrix (2 3 1 1) */
          uint8_t a[4], m[4];
          for (int j = 0; j < 4; ++j) {
            a[j] = t[4 * c + j];
            m[j] = mul2(a[j]);
          }
          for (int j = 0; j < 4; ++j)
            t[4 * c + j] = (uint8_t)(m[j] ^ m[(j + 1) % 4] ^ a[(j + 1) % 4] ^ a[(j + 2) % 4] ^ a[(j + 3) % 4]);
        }
      for (int i = 0; i < 16; ++i) s[i] = (uint8_t)(t[i] ^ rk[16 * r + i]);
    }
    memcpy(out + 16 * b, s, 16);
  }
}
