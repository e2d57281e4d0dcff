/*
 * Saber l=3 (round 3) restatement for the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Reference anchors (benlwk/DPCrypto, SABER-GPU/):
 *   parameters / sizes             include/params.h:6-48
 *   h1 = 4, h2 = 228               src/poly.cu:11-12
 *   encaps sequence                src/cuda_kernel.cu:68-133
 *   decaps sequence                src/cuda_kernel.cu:213-273
 *   negacyclic MatVec / InnerProd  src/poly.cu:71-206 (DP), 209-358 (plain)
 *   rounding                       src/poly.cu:360-382
 *   CBD (mu = 8)                   src/poly.cu:815-866
 *   GenMatrix / GenSecret          src/poly.cu:897-1052
 *   bit packers (13/10/4/1 bit)    src/pack.cu:8-211
 *   FO compare / cmov              src/pack.cu:215-253
 * Keygen is absent from the reference (include/cuda_kernel.cuh:6-7); it is
 * restated from the published scheme (b = round(A^T s), pk = b || seedA,
 * sk = s || pk || H(pk) || z).  The reference's implicit rejection never
 * fires (src/pack.cu:245-253); the published constant-time cmov is used.
 */
#include <string.h>

#include "oracle.h"

#define SN 256
#define SL 3
#define POLYBYTES 416         /* 13 * 256 / 8 */
#define POLYVECBYTES 1248     /* 3 * 416 */
#define POLYCOMPBYTES 320     /* 10 * 256 / 8 */
#define POLYVECCOMPBYTES 960  /* 3 * 320 */
#define SCALEBYTES 128        /* 4 * 256 / 8 */
#define H1 4
#define H2 228

void orc_saber_pack13(uint8_t *out, const uint16_t *in, size_t npoly) {
  /* little-endian bit stream of 13-bit fields */
  size_t nbits = 13 * SN * npoly;
  memset(out, 0, nbits / 8);
  for (size_t i = 0; i < SN * npoly; ++i) {
    uint32_t v = in[i] & 0x1FFF;
    for (int b = 0; b < 13; ++b)
      if (v >> b & 1) out[(13 * i + b) / 8] |= (uint8_t)(1u << ((13 * i + b) % 8));
  }
}

void orc_saber_unpack13(uint16_t *out, const uint8_t *in, size_t npoly) {
  for (size_t i = 0; i < SN * npoly; ++i) {
    uint16_t v = 0;
    for (int b = 0; b < 13; ++b) v |= (uint16_t)(((in[(13 * i + b) / 8] >> ((13 * i + b) % 8)) & 1) << b);
    out[i] = v;
  }
}

void orc_saber_pack10(uint8_t *out, const uint16_t *in, size_t npoly) {
  size_t nbits = 10 * SN * npoly;
  memset(out, 0, nbits / 8);
  for (size_t i = 0; i < SN * npoly; ++i) {
    uint32_t v = in[i] & 0x3FF;
    for (int b = 0; b < 10; ++b)
      if (v >> b & 1) out[(10 * i + b) / 8] |= (uint8_t)(1u << ((10 * i + b) % 8));
  }
}

void orc_saber_unpack10(uint16_t *out, const uint8_t *in, size_t npoly) {
  for (size_t i = 0; i < SN * npoly; ++i) {
    uint16_t v = 0;
    for (int b = 0; b < 10; ++b) v |= (uint16_t)(((in[(10 * i + b) / 8] >> ((10 * i + b) % 8)) & 1) << b);
    out[i] = v;
  }
}

static void pack4(uint8_t *out, const uint16_t *in) {
  for (int i = 0; i < SN / 2; ++i) out[i] = (uint8_t)((in[2 * i] & 0xF) | ((in[2 * i + 1] & 0xF) << 4));
}
static void unpack4(uint16_t *out, const uint8_t *in) {
  for (int i = 0; i < SN / 2; ++i) {
    out[2 * i] = in[i] & 0xF;
    out[2 * i + 1] = in[i] >> 4;
  }
}
static void msg_to_poly(uint16_t *out, const uint8_t m[32]) {
  for (int i = 0; i < SN; ++i) out[i] = (m[i / 8] >> (i % 8)) & 1;
}
static void poly_to_msg(uint8_t m[32], const uint16_t *in) {
  memset(m, 0, 32);
  for (int i = 0; i < SN; ++i) m[i / 8] |= (uint8_t)((in[i] & 1) << (i % 8));
}

void orc_saber_gen_matrix(uint16_t a[3 * 3 * 256], const uint8_t seed[32]) {
  uint8_t buf[SL * POLYVECBYTES];
  orc_shake128(buf, sizeof buf, seed, 32);
  orc_saber_unpack13(a, buf, SL * SL);
}

void orc_saber_gen_secret(uint16_t s[3 * 256], const uint8_t seed[32]) {
  uint8_t buf[SL * SN]; /* mu * n / 8 = 256 bytes per polynomial */
  orc_shake128(buf, sizeof buf, seed, 32);
  for (int i = 0; i < SL * SN; ++i) {
    uint8_t byte = buf[i];
    int lo = __builtin_popcount(byte & 0x0F), hi = __builtin_popcount(byte >> 4);
    s[i] = (uint16_t)(lo - hi);
  }
}

/* res += a * b in Z_{2^16}[x] / (x^256 + 1) */
void orc_saber_polymul_acc(const uint16_t *a, const uint16_t *b, uint16_t *res) {
  uint16_t tmp[SN];
  memset(tmp, 0, sizeof tmp);
  for (int i = 0; i < SN; ++i)
    for (int j = 0; j < SN; ++j) {
      uint16_t p = (uint16_t)(a[i] * b[j]);
      int k = i + j;
      if (k < SN)
        tmp[k] = (uint16_t)(tmp[k] + p);
      else
        tmp[k - SN] = (uint16_t)(tmp[k - SN] - p);
    }
  for (int i = 0; i < SN; ++i) res[i] = (uint16_t)(res[i] + tmp[i]);
}

void orc_saber_matvec(uint16_t *res, const uint16_t *a, const uint16_t *s, int transpose) {
  memset(res, 0, sizeof(uint16_t) * SL * SN);
  for (int i = 0; i < SL; ++i)
    for (int j = 0; j < SL; ++j) {
      const uint16_t *aij = transpose ? a + (j * SL + i) * SN : a + (i * SL + j) * SN;
      orc_saber_polymul_acc(aij, s + j * SN, res + i * SN);
    }
}

static void inner_prod(uint16_t *res, const uint16_t *b, const uint16_t *s) {
  memset(res, 0, sizeof(uint16_t) * SN);
  for (int j = 0; j < SL; ++j) orc_saber_polymul_acc(b + j * SN, s + j * SN, res);
}

static void indcpa_keypair(uint8_t *pk, uint8_t *sk_s, const uint8_t seed_a_raw[32],
                           const uint8_t seed_s[32]) {
  uint16_t a[SL * SL * SN], s[SL * SN], b[SL * SN];
  uint8_t seed_a[32];
  orc_shake128(seed_a, 32, seed_a_raw, 32);
  orc_saber_gen_matrix(a, seed_a);
  orc_saber_gen_secret(s, seed_s);
  orc_saber_matvec(b, a, s, 1);
  for (int i = 0; i < SL * SN; ++i) b[i] = (uint16_t)((b[i] + H1) >> 3);
  orc_saber_pack13(sk_s, s, SL);
  orc_saber_pack10(pk, b, SL);
  memcpy(pk + POLYVECCOMPBYTES, seed_a, 32);
}

void orc_saber_indcpa_enc(uint8_t *ct, const uint8_t m[32], const uint8_t seed_sp[32], const uint8_t *pk) {
  uint16_t a[SL * SL * SN], sp[SL * SN], bp[SL * SN], b[SL * SN], vp[SN], mp[SN];
  orc_saber_gen_matrix(a, pk + POLYVECCOMPBYTES);
  orc_saber_gen_secret(sp, seed_sp);
  orc_saber_matvec(bp, a, sp, 0);
  for (int i = 0; i < SL * SN; ++i) bp[i] = (uint16_t)((bp[i] + H1) >> 3);
  orc_saber_pack10(ct, bp, SL);
  orc_saber_unpack10(b, pk, SL);
  inner_prod(vp, b, sp);
  msg_to_poly(mp, m);
  /* bits 6..9 of (v' - m*2^9 + h1) are all that POLT keeps, so u16 wrap is exact */
  for (int i = 0; i < SN; ++i) vp[i] = (uint16_t)((uint16_t)(vp[i] - (mp[i] << 9) + H1) >> 6);
  pack4(ct + POLYVECCOMPBYTES, vp);
}

static void indcpa_dec(uint8_t m[32], const uint8_t *sk_s, const uint8_t *ct) {
  uint16_t s[SL * SN], b[SL * SN], v[SN], cm[SN];
  orc_saber_unpack13(s, sk_s, SL);
  orc_saber_unpack10(b, ct, SL);
  inner_prod(v, b, s);
  unpack4(cm, ct + POLYVECCOMPBYTES);
  for (int i = 0; i < SN; ++i) v[i] = (uint16_t)((uint16_t)(v[i] + H2 - (cm[i] << 6)) >> 9);
  poly_to_msg(m, v);
}

int orc_saber_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) {
  indcpa_keypair(pk, sk, coins, coins + 32);
  memcpy(sk + POLYVECBYTES, pk, ORC_SABER_PK);
  orc_sha3_256(sk + ORC_SABER_SK - 64, pk, ORC_SABER_PK);
  memcpy(sk + ORC_SABER_SK - 32, coins + 64, 32);
  return 0;
}

int orc_saber_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *m_raw) {
  uint8_t buf[64], kr[64];
  orc_sha3_256(buf, m_raw, 32);
  orc_sha3_256(buf + 32, pk, ORC_SABER_PK);
  orc_sha3_512(kr, buf, 64);
  orc_saber_indcpa_enc(ct, buf, kr + 32, pk);
  orc_sha3_256(kr + 32, ct, ORC_SABER_CT);
  orc_sha3_256(ss, kr, 64);
  return 0;
}

int orc_saber_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) {
  const uint8_t *pk = sk + POLYVECBYTES;
  uint8_t buf[64], kr[64], cmp[ORC_SABER_CT];
  indcpa_dec(buf, sk, ct);
  memcpy(buf + 32, sk + ORC_SABER_SK - 64, 32);
  orc_sha3_512(kr, buf, 64);
  orc_saber_indcpa_enc(cmp, buf, kr + 32, pk);
  uint8_t diff = 0;
  for (int i = 0; i < ORC_SABER_CT; ++i) diff |= (uint8_t)(ct[i] ^ cmp[i]);
  uint8_t mask = (uint8_t)(-(int)((uint8_t)(diff | (uint8_t)-diff) >> 7)); /* 0xFF on mismatch */
  orc_sha3_256(kr + 32, ct, ORC_SABER_CT);
  const uint8_t *z = sk + ORC_SABER_SK - 32;
  for (int i = 0; i < 32; ++i) kr[i] = (uint8_t)((kr[i] & ~mask) | (z[i] & mask));
  orc_sha3_256(ss, kr, 64);
  return 0;
}
