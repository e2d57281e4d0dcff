/*
 * Saber (round 3) restatement for the oracle (TEST INFRASTRUCTURE ONLY):
 * LightSaber (l = 2, mu = 10, eT = 3), Saber (l = 3, mu = 8, eT = 4) and
 * FireSaber (l = 4, mu = 6, eT = 6), parameter sets as at
 * SABER-GPU/include/params.h:10-19 and the CBD branches at
 * SABER-GPU/src/poly.cu:817-891.
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
#define SLMAX 4
#define POLYBYTES 416      /* 13 * 256 / 8 */
#define POLYCOMPBYTES 320  /* 10 * 256 / 8 */
#define H1 4

typedef struct {
  int l, mu, et;
  int h2;                        /* 2^(ep-2) - 2^(ep-eT-1) + 2^(eq-ep-1) */
  size_t pk, sk, ct, vecq, vecp, scale;
} params;

static params params_of(int l) {
  params p;
  p.l = l;
  p.mu = l == 2 ? 10 : l == 3 ? 8 : 6;
  p.et = l == 2 ? 3 : l == 3 ? 4 : 6;
  p.h2 = 256 - (1 << (10 - p.et - 1)) + 4;
  p.vecq = (size_t)l * POLYBYTES;
  p.vecp = (size_t)l * POLYCOMPBYTES;
  p.scale = (size_t)p.et * SN / 8;
  p.pk = p.vecp + 32;
  p.sk = p.vecq + p.pk + 64;
  p.ct = p.vecp + p.scale;
  return p;
}

/* little-endian bit stream of `bits`-bit fields */
static void pack_bits(uint8_t *out, const uint16_t *in, size_t count, int bits) {
  memset(out, 0, (size_t)bits * count / 8);
  for (size_t i = 0; i < count; ++i) {
    uint32_t v = in[i] & ((1u << bits) - 1);
    for (int b = 0; b < bits; ++b)
      if (v >> b & 1) out[(bits * i + b) / 8] |= (uint8_t)(1u << ((bits * i + b) % 8));
  }
}

static void unpack_bits(uint16_t *out, const uint8_t *in, size_t count, int bits) {
  for (size_t i = 0; i < count; ++i) {
    uint16_t v = 0;
    for (int b = 0; b < bits; ++b) v |= (uint16_t)(((in[(bits * i + b) / 8] >> ((bits * i + b) % 8)) & 1) << b);
    out[i] = v;
  }
}

void orc_saber_pack13(uint8_t *out, const uint16_t *in, size_t npoly) { pack_bits(out, in, SN * npoly, 13); }
void orc_saber_unpack13(uint16_t *out, const uint8_t *in, size_t npoly) { unpack_bits(out, in, SN * npoly, 13); }
void orc_saber_pack10(uint8_t *out, const uint16_t *in, size_t npoly) { pack_bits(out, in, SN * npoly, 10); }
void orc_saber_unpack10(uint16_t *out, const uint8_t *in, size_t npoly) { unpack_bits(out, in, SN * npoly, 10); }

static void msg_to_poly(uint16_t *out, const uint8_t m[32]) {
  for (int i = 0; i < SN; ++i) out[i] = (m[i / 8] >> (i % 8)) & 1;
}
static void poly_to_msg(uint8_t m[32], const uint16_t *in) {
  memset(m, 0, 32);
  for (int i = 0; i < SN; ++i) m[i / 8] |= (uint8_t)((in[i] & 1) << (i % 8));
}

void orc_saber_l_gen_matrix(int l, uint16_t *a, const uint8_t seed[32]) {
  uint8_t buf[SLMAX * SLMAX * POLYBYTES];
  orc_shake128(buf, (size_t)l * l * POLYBYTES, seed, 32);
  unpack_bits(a, buf, (size_t)l * l * SN, 13);
}

/* CBD with parameter mu: coefficient i takes bits [mu i, mu i + mu) of the
 * SHAKE128 stream, s = HW(low mu/2 bits) - HW(high mu/2 bits). */
void orc_saber_l_gen_secret(int l, uint16_t *s, const uint8_t seed[32]) {
  const params p = params_of(l);
  uint8_t buf[SLMAX * SN * 10 / 8];
  orc_shake128(buf, (size_t)l * SN * p.mu / 8, seed, 32);
  const int half = p.mu / 2;
  for (int i = 0; i < l * SN; ++i) {
    uint32_t v = 0;
    for (int b = 0; b < p.mu; ++b) v |= (uint32_t)((buf[(p.mu * i + b) / 8] >> ((p.mu * i + b) % 8)) & 1) << b;
    s[i] = (uint16_t)(__builtin_popcount(v & ((1u << half) - 1)) - __builtin_popcount(v >> half));
  }
}

void orc_saber_gen_matrix(uint16_t a[3 * 3 * 256], const uint8_t seed[32]) { orc_saber_l_gen_matrix(3, a, seed); }
void orc_saber_gen_secret(uint16_t s[3 * 256], const uint8_t seed[32]) { orc_saber_l_gen_secret(3, s, seed); }

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

static void matvec(int l, uint16_t *res, const uint16_t *a, const uint16_t *s, int transpose) {
  memset(res, 0, sizeof(uint16_t) * l * SN);
  for (int i = 0; i < l; ++i)
    for (int j = 0; j < l; ++j) {
      const uint16_t *aij = transpose ? a + (j * l + i) * SN : a + (i * l + j) * SN;
      orc_saber_polymul_acc(aij, s + j * SN, res + i * SN);
    }
}

void orc_saber_matvec(uint16_t *res, const uint16_t *a, const uint16_t *s, int transpose) {
  matvec(3, res, a, s, transpose);
}

static void inner_prod(int l, uint16_t *res, const uint16_t *b, const uint16_t *s) {
  memset(res, 0, sizeof(uint16_t) * SN);
  for (int j = 0; j < l; ++j) orc_saber_polymul_acc(b + j * SN, s + j * SN, res);
}

static void indcpa_keypair(const params *p, uint8_t *pk, uint8_t *sk_s, const uint8_t seed_a_raw[32],
                           const uint8_t seed_s[32]) {
  uint16_t a[SLMAX * SLMAX * SN], s[SLMAX * SN], b[SLMAX * SN];
  uint8_t seed_a[32];
  orc_shake128(seed_a, 32, seed_a_raw, 32);
  orc_saber_l_gen_matrix(p->l, a, seed_a);
  orc_saber_l_gen_secret(p->l, s, seed_s);
  matvec(p->l, b, a, s, 1);
  for (int i = 0; i < p->l * SN; ++i) b[i] = (uint16_t)((b[i] + H1) >> 3);
  pack_bits(sk_s, s, (size_t)p->l * SN, 13);
  pack_bits(pk, b, (size_t)p->l * SN, 10);
  memcpy(pk + p->vecp, seed_a, 32);
}

static void indcpa_enc(const params *p, uint8_t *ct, const uint8_t m[32], const uint8_t seed_sp[32],
                       const uint8_t *pk) {
  uint16_t a[SLMAX * SLMAX * SN], sp[SLMAX * SN], bp[SLMAX * SN], b[SLMAX * SN], vp[SN], mp[SN];
  orc_saber_l_gen_matrix(p->l, a, pk + p->vecp);
  orc_saber_l_gen_secret(p->l, sp, seed_sp);
  matvec(p->l, bp, a, sp, 0);
  for (int i = 0; i < p->l * SN; ++i) bp[i] = (uint16_t)((bp[i] + H1) >> 3);
  pack_bits(ct, bp, (size_t)p->l * SN, 10);
  unpack_bits(b, pk, (size_t)p->l * SN, 10);
  inner_prod(p->l, vp, b, sp);
  msg_to_poly(mp, m);
  /* only bits (10 - eT) .. 9 of (v' - m 2^9 + h1) are kept, so u16 wrap is exact */
  for (int i = 0; i < SN; ++i) vp[i] = (uint16_t)((uint16_t)(vp[i] - (mp[i] << 9) + H1) >> (10 - p->et));
  pack_bits(ct + p->vecp, vp, SN, p->et);
}

void orc_saber_indcpa_enc(uint8_t *ct, const uint8_t m[32], const uint8_t seed_sp[32], const uint8_t *pk) {
  const params p = params_of(3);
  indcpa_enc(&p, ct, m, seed_sp, pk);
}

static void indcpa_dec(const params *p, uint8_t m[32], const uint8_t *sk_s, const uint8_t *ct) {
  uint16_t s[SLMAX * SN], b[SLMAX * SN], v[SN], cm[SN];
  unpack_bits(s, sk_s, (size_t)p->l * SN, 13);
  unpack_bits(b, ct, (size_t)p->l * SN, 10);
  inner_prod(p->l, v, b, s);
  unpack_bits(cm, ct + p->vecp, SN, p->et);
  for (int i = 0; i < SN; ++i) v[i] = (uint16_t)((uint16_t)(v[i] + p->h2 - (cm[i] << (10 - p->et))) >> 9);
  poly_to_msg(m, v);
}

int orc_saber_l_keypair(int l, uint8_t *pk, uint8_t *sk, const uint8_t *coins) {
  if (l < 2 || l > SLMAX) return -1;
  const params p = params_of(l);
  indcpa_keypair(&p, pk, sk, coins, coins + 32);
  memcpy(sk + p.vecq, pk, p.pk);
  orc_sha3_256(sk + p.sk - 64, pk, p.pk);
  memcpy(sk + p.sk - 32, coins + 64, 32);
  return 0;
}

int orc_saber_l_enc(int l, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *m_raw) {
  if (l < 2 || l > SLMAX) return -1;
  const params p = params_of(l);
  uint8_t buf[64], kr[64];
  orc_sha3_256(buf, m_raw, 32);
  orc_sha3_256(buf + 32, pk, p.pk);
  orc_sha3_512(kr, buf, 64);
  indcpa_enc(&p, ct, buf, kr + 32, pk);
  orc_sha3_256(kr + 32, ct, p.ct);
  orc_sha3_256(ss, kr, 64);
  return 0;
}

int orc_saber_l_dec(int l, uint8_t *ss, const uint8_t *ct, const uint8_t *sk) {
  if (l < 2 || l > SLMAX) return -1;
  const params p = params_of(l);
  const uint8_t *pk = sk + p.vecq;
  uint8_t buf[64], kr[64], cmp[1472];
  indcpa_dec(&p, buf, sk, ct);
  memcpy(buf + 32, sk + p.sk - 64, 32);
  orc_sha3_512(kr, buf, 64);
  indcpa_enc(&p, cmp, buf, kr + 32, pk);
  uint8_t diff = 0;
  for (size_t i = 0; i < p.ct; ++i) diff |= (uint8_t)(ct[i] ^ cmp[i]);
  uint8_t mask = (uint8_t)(-(int)((uint8_t)(diff | (uint8_t)-diff) >> 7)); /* 0xFF on mismatch */
  orc_sha3_256(kr + 32, ct, p.ct);
  const uint8_t *z = sk + p.sk - 32;
  for (int i = 0; i < 32; ++i) kr[i] = (uint8_t)((kr[i] & ~mask) | (z[i] & mask));
  orc_sha3_256(ss, kr, 64);
  return 0;
}

int orc_saber_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) { return orc_saber_l_keypair(3, pk, sk, coins); }
int orc_saber_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *m_raw) {
  return orc_saber_l_enc(3, ct, ss, pk, m_raw);
}
int orc_saber_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) { return orc_saber_l_dec(3, ss, ct, sk); }
