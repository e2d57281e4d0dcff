/*
 * FIPS 202 restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Follows the permutation and sponge as published in FIPS 202; the
 * reference carries its own copies at FrodoKEM976-GPU/src/shake.cu:66-377
 * (KeccakF1600_StatePermute, keccak_absorb_ser, keccak_squeezeblocks_ser)
 * and SABER-GPU/src/poly.cu:436-776.  Pinned by tests against Python's
 * hashlib at every (rate, input length, output length) the KEMs use.
 */
#include <string.h>

#include "oracle.h"

static const uint64_t kRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL,
    0x8000000080008000ULL, 0x000000000000808bULL, 0x0000000080000001ULL,
    0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008aULL,
    0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL,
    0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
    0x000000000000800aULL, 0x800000008000000aULL, 0x8000000080008081ULL,
    0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};

/* rotation offset of lane (x, y), indexed x + 5y */
static const unsigned kRho[25] = {0,  1,  62, 28, 27, 36, 44, 6,  55, 20, 3,  10, 43,
                                  25, 39, 41, 45, 15, 21, 8,  18, 2,  61, 56, 14};

static inline uint64_t rotl64(uint64_t v, unsigned n) {
  return n ? (v << n) | (v >> (64 - n)) : v;
}

void orc_keccak_f1600(uint64_t a[25]) {
  uint64_t b[25], c[5], d[5];
  for (int round = 0; round < 24; ++round) {
    for (int x = 0; x < 5; ++x) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
    for (int x = 0; x < 5; ++x) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
    /* theta + rho + pi: B[y, 2x+3y] = rot(A[x,y] ^ D[x], r[x,y]) */
    for (int y = 0; y < 5; ++y)
      for (int x = 0; x < 5; ++x) {
        int nx = y, ny = (2 * x + 3 * y) % 5;
        b[nx + 5 * ny] = rotl64(a[x + 5 * y] ^ d[x], kRho[x + 5 * y]);
      }
    /* chi */
    for (int y = 0; y < 5; ++y)
      for (int x = 0; x < 5; ++x)
        a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
    /* iota */
    a[0] ^= kRC[round];
  }
}

static void sponge(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen, size_t rate,
                   uint8_t pad) {
  uint64_t st[25];
  uint8_t block[200];
  memset(st, 0, sizeof st);
  /* absorb full blocks */
  while (inlen >= rate) {
    for (size_t i = 0; i < rate / 8; ++i) {
      uint64_t w = 0;
      for (int k = 0; k < 8; ++k) w |= (uint64_t)in[8 * i + k] << (8 * k);
      st[i] ^= w;
    }
    orc_keccak_f1600(st);
    in += rate;
    inlen -= rate;
  }
  /* final padded block */
  memset(block, 0, sizeof block);
  memcpy(block, in, inlen);
  block[inlen] ^= pad;
  block[rate - 1] ^= 0x80;
  for (size_t i = 0; i < rate / 8; ++i) {
    uint64_t w = 0;
    for (int k = 0; k < 8; ++k) w |= (uint64_t)block[8 * i + k] << (8 * k);
    st[i] ^= w;
  }
  /* squeeze */
  while (outlen > 0) {
    orc_keccak_f1600(st);
    size_t take = outlen < rate ? outlen : rate;
    for (size_t i = 0; i < take; ++i) out[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
    out += take;
    outlen -= take;
  }
}

void orc_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
  sponge(out, outlen, in, inlen, 168, 0x1F);
}
void orc_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
  sponge(out, outlen, in, inlen, 136, 0x1F);
}
void orc_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen) {
  sponge(out, 32, in, inlen, 136, 0x06);
}
void orc_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen) {
  sponge(out, 64, in, inlen, 72, 0x06);
}

/* ---------------- PRNG ---------------- */
static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void orc_prng_seed(orc_prng *p, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) p->s[i] = splitmix64(&x);
}

uint64_t orc_prng_next(orc_prng *p) {
  uint64_t *s = p->s;
  uint64_t result = rotl64(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return result;
}

void orc_prng_bytes(orc_prng *p, uint8_t *out, size_t len) {
  while (len > 0) {
    uint64_t v = orc_prng_next(p);
    size_t take = len < 8 ? len : 8;
    for (size_t i = 0; i < take; ++i) out[i] = (uint8_t)(v >> (8 * i));
    out += take;
    len -= take;
  }
}

void orc_prng_fill(uint64_t seed, uint8_t *out, size_t len) {
  orc_prng p;
  orc_prng_seed(&p, seed);
  orc_prng_bytes(&p, out, len);
}
