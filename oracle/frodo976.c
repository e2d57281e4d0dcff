/*
 * FrodoKEM-976-SHAKE (round 3) restatement for the oracle
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Reference anchors (benlwk/DPCrypto, FrodoKEM976-GPU/):
 *   parameters, CDF table          include/params.h:9-24
 *   wire sizes                     include/api.h:7-10
 *   keygen sequence                src/cuda_kernel.cu:64-84
 *   encaps sequence                src/cuda_kernel.cu:159-191
 *   decaps sequence                src/cuda_kernel.cu:282-323
 *   CDF sampler                    src/sample_pack.cu:13-46
 *   16-bit pack / unpack           src/sample_pack.cu:55-198
 *   key encode / decode            src/encode_decode.cu:13-52
 *   A*S+E, S'*A+E', S'*B+E'', B'*S src/matrix_kernel.cu:13-37, 75-111, 245-278
 *   constant-time verify / select  src/util.cu:193-228
 * The reference's batched kernels do not keep instances apart and derive the
 * A-row seed from an uninitialised byte; the published layout
 * A[i] = SHAKE128(LE16(i) || seedA) is used here instead.
 *
 * FrodoKEM-976-AES (orc_frodo_aes_*) differs only in A: entries j..j+7 of
 * row i are AES-128_seedA(LE16(i) || LE16(j) || 0^12) read as little-endian
 * u16 (the variant named at FrodoKEM976-GPU/include/api.h:12; its AES is
 * declared but not defined, FrodoKEM976-GPU/include/gen_a_mat_mul.cuh:6).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define N ORC_FRODO_N
#define NBAR ORC_FRODO_NBAR
#define LEN_SEC 24  /* len(s) = len(seedSE) = len(k) = len(pkh) = len(ss) */
#define LEN_SEEDA 16
#define LEN_MU 24

static const uint16_t kCdf[11] = {5638,  15915, 23689, 28571, 31116, 32217,
                                  32613, 32731, 32760, 32766, 32767};

void orc_frodo_sample(uint16_t *s, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    uint16_t t = (uint16_t)(s[i] >> 1);
    uint16_t sign = (uint16_t)(s[i] & 1);
    uint16_t e = 0;
    for (int j = 0; j < 10; ++j) e = (uint16_t)(e + ((uint16_t)(kCdf[j] - t) >> 15));
    s[i] = (uint16_t)(((uint16_t)(-sign) ^ e) + sign);
  }
}

void orc_frodo_aes_gen_a_row(uint16_t row[N], const uint8_t seed_a[16], uint16_t i) {
  uint8_t in[2 * N], buf[2 * N];
  memset(in, 0, sizeof in);
  for (int j = 0; j < N; j += 8) {
    in[2 * j] = (uint8_t)i;
    in[2 * j + 1] = (uint8_t)(i >> 8);
    in[2 * j + 2] = (uint8_t)j;
    in[2 * j + 3] = (uint8_t)(j >> 8);
  }
  orc_aes128_ecb(seed_a, in, buf, N / 8);
  for (int j = 0; j < N; ++j) row[j] = (uint16_t)(buf[2 * j] | (buf[2 * j + 1] << 8));
}

void orc_frodo_gen_a_row(uint16_t row[N], const uint8_t seed_a[16], uint16_t i) {
  uint8_t in[2 + LEN_SEEDA];
  uint8_t buf[2 * N];
  in[0] = (uint8_t)i;
  in[1] = (uint8_t)(i >> 8);
  memcpy(in + 2, seed_a, LEN_SEEDA);
  orc_shake128(buf, sizeof buf, in, sizeof in);
  for (int j = 0; j < N; ++j) row[j] = (uint16_t)(buf[2 * j] | (buf[2 * j + 1] << 8));
}

void orc_frodo_pack16(uint8_t *out, const uint16_t *in, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    out[2 * i] = (uint8_t)(in[i] >> 8);
    out[2 * i + 1] = (uint8_t)in[i];
  }
}

void orc_frodo_unpack16(uint16_t *out, const uint8_t *in, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (uint16_t)((in[2 * i] << 8) | in[2 * i + 1]);
}

void orc_frodo_encode(uint16_t out[64], const uint8_t mu[24]) {
  for (int w = 0; w < 8; ++w) {
    uint32_t t = (uint32_t)mu[3 * w] | ((uint32_t)mu[3 * w + 1] << 8) | ((uint32_t)mu[3 * w + 2] << 16);
    for (int j = 0; j < 8; ++j) out[8 * w + j] = (uint16_t)(((t >> (3 * j)) & 7) << 13);
  }
}

void orc_frodo_decode(uint8_t mu[24], const uint16_t in[64]) {
  for (int w = 0; w < 8; ++w) {
    uint32_t t = 0;
    for (int j = 0; j < 8; ++j) {
      uint32_t v = (((uint32_t)in[8 * w + j] + (1u << 12)) >> 13) & 7;
      t |= v << (3 * j);
    }
    mu[3 * w] = (uint8_t)t;
    mu[3 * w + 1] = (uint8_t)(t >> 8);
    mu[3 * w + 2] = (uint8_t)(t >> 16);
  }
}

static void gen_row(uint16_t row[N], const uint8_t seed_a[16], uint16_t i, int aes) {
  if (aes) orc_frodo_aes_gen_a_row(row, seed_a, i);
  else orc_frodo_gen_a_row(row, seed_a, i);
}

static void as_plus_e(uint16_t *out, const uint16_t *st, const uint16_t *e, const uint8_t seed_a[16], int aes) {
  uint16_t row[N];
  for (int i = 0; i < N; ++i) {
    gen_row(row, seed_a, (uint16_t)i, aes);
    for (int k = 0; k < NBAR; ++k) {
      uint16_t acc = e[i * NBAR + k];
      for (int j = 0; j < N; ++j) acc = (uint16_t)(acc + row[j] * st[k * N + j]);
      out[i * NBAR + k] = acc;
    }
  }
}

void orc_frodo_as_plus_e(uint16_t *out, const uint16_t *st, const uint16_t *e, const uint8_t seed_a[16]) {
  as_plus_e(out, st, e, seed_a, 0);
}

static void sa_plus_e(uint16_t *out, const uint16_t *sp, const uint16_t *ep, const uint8_t seed_a[16], int aes) {
  uint16_t row[N];
  memcpy(out, ep, sizeof(uint16_t) * N * NBAR);
  for (int i = 0; i < N; ++i) {
    gen_row(row, seed_a, (uint16_t)i, aes);
    for (int k = 0; k < NBAR; ++k) {
      uint16_t s = sp[k * N + i];
      for (int j = 0; j < N; ++j) out[k * N + j] = (uint16_t)(out[k * N + j] + s * row[j]);
    }
  }
}

void orc_frodo_sa_plus_e(uint16_t *out, const uint16_t *sp, const uint16_t *ep, const uint8_t seed_a[16]) {
  sa_plus_e(out, sp, ep, seed_a, 0);
}

/* V (NBAR x NBAR) = S' (NBAR x N) * B (N x NBAR) + E'' */
static void sb_plus_e(uint16_t *v, const uint16_t *b, const uint16_t *sp, const uint16_t *epp) {
  for (int i = 0; i < NBAR; ++i)
    for (int j = 0; j < NBAR; ++j) {
      uint16_t acc = epp[i * NBAR + j];
      for (int k = 0; k < N; ++k) acc = (uint16_t)(acc + sp[i * N + k] * b[k * NBAR + j]);
      v[i * NBAR + j] = acc;
    }
}

/* W (NBAR x NBAR) = B' (NBAR x N) * S, S given transposed (NBAR x N) */
static void mul_bs(uint16_t *w, const uint16_t *bp, const uint16_t *st) {
  for (int i = 0; i < NBAR; ++i)
    for (int j = 0; j < NBAR; ++j) {
      uint16_t acc = 0;
      for (int k = 0; k < N; ++k) acc = (uint16_t)(acc + bp[i * N + k] * st[j * N + k]);
      w[i * NBAR + j] = acc;
    }
}

static void le_bytes_to_u16(uint16_t *out, const uint8_t *in, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (uint16_t)(in[2 * i] | (in[2 * i + 1] << 8));
}

static int frodo_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins, int aes) {
  const uint8_t *s = coins, *seed_se = coins + LEN_SEC, *z = coins + 2 * LEN_SEC;
  uint8_t se_in[1 + LEN_SEC];
  uint8_t *r = malloc(4 * N * NBAR);
  uint16_t *st = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t *e = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t *b = malloc(sizeof(uint16_t) * N * NBAR);
  if (!r || !st || !e || !b) {
    free(r); free(st); free(e); free(b);
    return -1;
  }
  /* seedA = SHAKE256(z, 16) */
  orc_shake256(pk, LEN_SEEDA, z, LEN_SEEDA);
  /* S^T || E from SHAKE256(0x5F || seedSE) */
  se_in[0] = 0x5F;
  memcpy(se_in + 1, seed_se, LEN_SEC);
  orc_shake256(r, 4 * N * NBAR, se_in, sizeof se_in);
  le_bytes_to_u16(st, r, N * NBAR);
  le_bytes_to_u16(e, r + 2 * N * NBAR, N * NBAR);
  orc_frodo_sample(st, N * NBAR);
  orc_frodo_sample(e, N * NBAR);
  as_plus_e(b, st, e, pk, aes);
  orc_frodo_pack16(pk + LEN_SEEDA, b, N * NBAR);
  /* sk = s || pk || S^T (LE) || SHAKE256(pk, 24) */
  memcpy(sk, s, LEN_SEC);
  memcpy(sk + LEN_SEC, pk, ORC_FRODO_PK);
  for (int i = 0; i < N * NBAR; ++i) {
    sk[LEN_SEC + ORC_FRODO_PK + 2 * i] = (uint8_t)st[i];
    sk[LEN_SEC + ORC_FRODO_PK + 2 * i + 1] = (uint8_t)(st[i] >> 8);
  }
  orc_shake256(sk + ORC_FRODO_SK - LEN_SEC, LEN_SEC, pk, ORC_FRODO_PK);
  free(r); free(st); free(e); free(b);
  return 0;
}

/* Shared by encaps and the decaps re-encryption: from (seedSE, mu, pk) build
 * B' (NBAR x N) and C (NBAR x NBAR). */
static int frodo_encrypt(uint16_t *bp, uint16_t *c, const uint8_t *seed_se, const uint8_t *mu,
                         const uint8_t *pk, int aes) {
  uint8_t se_in[1 + LEN_SEC];
  const size_t rlen = 2 * (2 * N + NBAR) * NBAR;
  uint8_t *r = malloc(rlen);
  uint16_t *sp = malloc(sizeof(uint16_t) * (2 * N + NBAR) * NBAR);
  uint16_t *b = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t v[NBAR * NBAR], enc[NBAR * NBAR];
  if (!r || !sp || !b) {
    free(r); free(sp); free(b);
    return -1;
  }
  se_in[0] = 0x96;
  memcpy(se_in + 1, seed_se, LEN_SEC);
  orc_shake256(r, rlen, se_in, sizeof se_in);
  le_bytes_to_u16(sp, r, (2 * N + NBAR) * NBAR);
  orc_frodo_sample(sp, (2 * N + NBAR) * NBAR);
  /* sp = S' (NBAR x N) | E' (NBAR x N) | E'' (NBAR x NBAR) */
  sa_plus_e(bp, sp, sp + N * NBAR, pk, aes);
  orc_frodo_unpack16(b, pk + LEN_SEEDA, N * NBAR);
  sb_plus_e(v, b, sp, sp + 2 * N * NBAR);
  orc_frodo_encode(enc, mu);
  for (int i = 0; i < NBAR * NBAR; ++i) c[i] = (uint16_t)(v[i] + enc[i]);
  free(r); free(sp); free(b);
  return 0;
}

static int frodo_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu, int aes) {
  uint8_t g2in[LEN_SEC + LEN_MU], g2out[2 * LEN_SEC];
  uint16_t *bp = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t c[NBAR * NBAR];
  uint8_t *fin = malloc(ORC_FRODO_CT + LEN_SEC);
  if (!bp || !fin) {
    free(bp); free(fin);
    return -1;
  }
  orc_shake256(g2in, LEN_SEC, pk, ORC_FRODO_PK); /* pkh */
  memcpy(g2in + LEN_SEC, mu, LEN_MU);
  orc_shake256(g2out, 2 * LEN_SEC, g2in, sizeof g2in); /* seedSE || k */
  if (frodo_encrypt(bp, c, g2out, mu, pk, aes)) {
    free(bp); free(fin);
    return -1;
  }
  orc_frodo_pack16(ct, bp, N * NBAR);
  orc_frodo_pack16(ct + 2 * N * NBAR, c, NBAR * NBAR);
  memcpy(fin, ct, ORC_FRODO_CT);
  memcpy(fin + ORC_FRODO_CT, g2out + LEN_SEC, LEN_SEC);
  orc_shake256(ss, LEN_SEC, fin, ORC_FRODO_CT + LEN_SEC);
  free(bp); free(fin);
  return 0;
}

static int frodo_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk, int aes) {
  const uint8_t *sk_s = sk, *pk = sk + LEN_SEC, *sk_st = sk + LEN_SEC + ORC_FRODO_PK;
  const uint8_t *pkh = sk + ORC_FRODO_SK - LEN_SEC;
  uint16_t *st = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t *bp = malloc(sizeof(uint16_t) * N * NBAR);
  uint16_t *bbp = malloc(sizeof(uint16_t) * N * NBAR);
  uint8_t *fin = malloc(ORC_FRODO_CT + LEN_SEC);
  uint16_t c[NBAR * NBAR], w[NBAR * NBAR], cc[NBAR * NBAR];
  uint8_t g2in[LEN_SEC + LEN_MU], g2out[2 * LEN_SEC];
  if (!st || !bp || !bbp || !fin) {
    free(st); free(bp); free(bbp); free(fin);
    return -1;
  }
  le_bytes_to_u16(st, sk_st, N * NBAR);
  orc_frodo_unpack16(bp, ct, N * NBAR);
  orc_frodo_unpack16(c, ct + 2 * N * NBAR, NBAR * NBAR);
  mul_bs(w, bp, st);
  for (int i = 0; i < NBAR * NBAR; ++i) w[i] = (uint16_t)(c[i] - w[i]);
  memcpy(g2in, pkh, LEN_SEC);
  orc_frodo_decode(g2in + LEN_SEC, w);
  orc_shake256(g2out, 2 * LEN_SEC, g2in, sizeof g2in);
  if (frodo_encrypt(bbp, cc, g2out, g2in + LEN_SEC, pk, aes)) {
    free(st); free(bp); free(bbp); free(fin);
    return -1;
  }
  /* constant-time equality of (B', C) with the re-encryption */
  uint16_t diff = 0;
  for (int i = 0; i < N * NBAR; ++i) diff |= (uint16_t)(bp[i] ^ bbp[i]);
  for (int i = 0; i < NBAR * NBAR; ++i) diff |= (uint16_t)(c[i] ^ cc[i]);
  uint8_t mask = (uint8_t)(-(int)((diff | (uint16_t)-diff) >> 15)); /* 0xFF on mismatch */
  memcpy(fin, ct, ORC_FRODO_CT);
  for (int i = 0; i < LEN_SEC; ++i)
    fin[ORC_FRODO_CT + i] = (uint8_t)((g2out[LEN_SEC + i] & ~mask) | (sk_s[i] & mask));
  orc_shake256(ss, LEN_SEC, fin, ORC_FRODO_CT + LEN_SEC);
  free(st); free(bp); free(bbp); free(fin);
  return 0;
}

int orc_frodo_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) { return frodo_keypair(pk, sk, coins, 0); }
int orc_frodo_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) { return frodo_enc(ct, ss, pk, mu, 0); }
int orc_frodo_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) { return frodo_dec(ss, ct, sk, 0); }
int orc_frodo_aes_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) { return frodo_keypair(pk, sk, coins, 1); }
int orc_frodo_aes_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) {
  return frodo_enc(ct, ss, pk, mu, 1);
}
int orc_frodo_aes_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) { return frodo_dec(ss, ct, sk, 1); }
