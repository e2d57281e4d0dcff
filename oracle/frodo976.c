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
 * Frodo-640 and Frodo-1344 (orc_frodo_n_*, round-3 parameter sets, not in
 * the reference) share the algorithm with other sizes: n = 640 / 1344,
 * D = 15 / 16 bits, B = 2 / 4 bits per message entry, 16 / 32-byte secrets,
 * SHAKE128 / SHAKE256, and their own CDF tables.
 *
 * FrodoKEM-976-AES (orc_frodo_aes_*) differs only in A: entries j..j+7 of
 * row i are AES-128_seedA(LE16(i) || LE16(j) || 0^12) read as little-endian
 * u16 (the variant named at FrodoKEM976-GPU/include/api.h:12; its AES is
 * declared but not defined, FrodoKEM976-GPU/include/gen_a_mat_mul.cuh:6).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NBAR ORC_FRODO_NBAR
#define NMAX 1344
#define LEN_SEEDA 16

typedef struct {
  int n, logq, b, len_sec, shake128, aes;
  const uint16_t *cdf;
  int cdf_len;
  size_t pk, sk, ct;
} fparams;

/* Frodo-976 (FrodoKEM976-GPU/include/params.h:24) */
static const uint16_t kCdf976[11] = {5638,  15915, 23689, 28571, 31116, 32217,
                                     32613, 32731, 32760, 32766, 32767};
/* Frodo-640 and Frodo-1344 (round-3 specification, Table of error distributions) */
static const uint16_t kCdf640[13] = {4643,  13363, 20579, 25843, 29227, 31145, 32103,
                                     32525, 32689, 32745, 32762, 32766, 32767};
static const uint16_t kCdf1344[7] = {9142, 23462, 30338, 32361, 32725, 32765, 32767};

static fparams fparams_of(int n, int aes) {
  fparams p;
  p.n = n;
  p.aes = aes;
  p.logq = n == 640 ? 15 : 16;
  p.b = n == 640 ? 2 : n == 976 ? 3 : 4;
  p.len_sec = n == 640 ? 16 : n == 976 ? 24 : 32;
  p.shake128 = n == 640;
  p.cdf = n == 640 ? kCdf640 : n == 976 ? kCdf976 : kCdf1344;
  p.cdf_len = n == 640 ? 13 : n == 976 ? 11 : 7;
  p.pk = LEN_SEEDA + (size_t)p.logq * n * NBAR / 8;
  p.sk = p.len_sec + p.pk + 2 * (size_t)n * NBAR + p.len_sec;
  p.ct = (size_t)p.logq * (n * NBAR + NBAR * NBAR) / 8;
  return p;
}

static void shake(const fparams *p, uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen) {
  if (p->shake128) orc_shake128(out, outlen, in, inlen);
  else orc_shake256(out, outlen, in, inlen);
}

static void sample(const fparams *p, uint16_t *s, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    uint16_t t = (uint16_t)(s[i] >> 1);
    uint16_t sign = (uint16_t)(s[i] & 1);
    uint16_t e = 0;
    for (int j = 0; j < p->cdf_len - 1; ++j) e = (uint16_t)(e + ((uint16_t)(p->cdf[j] - t) >> 15));
    s[i] = (uint16_t)(((uint16_t)(-sign) ^ e) + sign);
  }
}

void orc_frodo_sample(uint16_t *s, size_t n) {
  const fparams p = fparams_of(976, 0);
  sample(&p, s, n);
}

void orc_frodo_n_sample(int n, uint16_t *s, size_t count) {
  const fparams p = fparams_of(n, 0);
  sample(&p, s, count);
}

static void gen_row_shake(uint16_t *row, int n, const uint8_t seed_a[16], uint16_t i) {
  uint8_t in[2 + LEN_SEEDA];
  uint8_t buf[2 * NMAX];
  in[0] = (uint8_t)i;
  in[1] = (uint8_t)(i >> 8);
  memcpy(in + 2, seed_a, LEN_SEEDA);
  orc_shake128(buf, 2 * (size_t)n, in, sizeof in);
  for (int j = 0; j < n; ++j) row[j] = (uint16_t)(buf[2 * j] | (buf[2 * j + 1] << 8));
}

void orc_frodo_aes_gen_a_row(uint16_t row[ORC_FRODO_N], const uint8_t seed_a[16], uint16_t i) {
  enum { N = ORC_FRODO_N };
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

void orc_frodo_gen_a_row(uint16_t row[ORC_FRODO_N], const uint8_t seed_a[16], uint16_t i) {
  gen_row_shake(row, ORC_FRODO_N, seed_a, i);
}

static void gen_row(const fparams *p, uint16_t *row, const uint8_t seed_a[16], uint16_t i) {
  if (p->aes) orc_frodo_aes_gen_a_row(row, seed_a, i);
  else gen_row_shake(row, p->n, seed_a, i);
}

/* D-bit entries, most significant bit first (D = 16: big-endian u16) */
static void pack_d(uint8_t *out, const uint16_t *in, size_t count, int d) {
  memset(out, 0, (size_t)d * count / 8);
  for (size_t i = 0; i < count; ++i)
    for (int b = 0; b < d; ++b)
      if ((in[i] >> (d - 1 - b)) & 1) out[(d * i + b) / 8] |= (uint8_t)(0x80u >> ((d * i + b) % 8));
}

static void unpack_d(uint16_t *out, const uint8_t *in, size_t count, int d) {
  for (size_t i = 0; i < count; ++i) {
    uint16_t v = 0;
    for (int b = 0; b < d; ++b) v = (uint16_t)((v << 1) | ((in[(d * i + b) / 8] >> (7 - (d * i + b) % 8)) & 1));
    out[i] = v;
  }
}

void orc_frodo_pack16(uint8_t *out, const uint16_t *in, size_t n) { pack_d(out, in, n, 16); }
void orc_frodo_unpack16(uint16_t *out, const uint8_t *in, size_t n) { unpack_d(out, in, n, 16); }

/* B message bits per entry, taken least significant first from the bytes of mu */
static void encode(const fparams *p, uint16_t out[64], const uint8_t *mu) {
  for (int i = 0; i < 64; ++i) {
    uint32_t k = 0;
    for (int l = 0; l < p->b; ++l) k |= (uint32_t)((mu[(p->b * i + l) / 8] >> ((p->b * i + l) % 8)) & 1) << l;
    out[i] = (uint16_t)(k << (p->logq - p->b));
  }
}

static void decode(const fparams *p, uint8_t *mu, const uint16_t in[64]) {
  memset(mu, 0, (size_t)p->b * 8);
  const uint32_t qmask = (1u << p->logq) - 1;
  for (int i = 0; i < 64; ++i) {
    uint32_t v = ((((uint32_t)in[i] & qmask) + (1u << (p->logq - p->b - 1))) >> (p->logq - p->b)) & ((1u << p->b) - 1);
    for (int l = 0; l < p->b; ++l) mu[(p->b * i + l) / 8] |= (uint8_t)(((v >> l) & 1) << ((p->b * i + l) % 8));
  }
}

void orc_frodo_encode(uint16_t out[64], const uint8_t mu[24]) {
  const fparams p = fparams_of(976, 0);
  encode(&p, out, mu);
}
void orc_frodo_decode(uint8_t mu[24], const uint16_t in[64]) {
  const fparams p = fparams_of(976, 0);
  decode(&p, mu, in);
}

static void as_plus_e(const fparams *p, uint16_t *out, const uint16_t *st, const uint16_t *e, const uint8_t seed_a[16]) {
  uint16_t row[NMAX];
  for (int i = 0; i < p->n; ++i) {
    gen_row(p, row, seed_a, (uint16_t)i);
    for (int k = 0; k < NBAR; ++k) {
      uint16_t acc = e[i * NBAR + k];
      for (int j = 0; j < p->n; ++j) acc = (uint16_t)(acc + row[j] * st[k * p->n + j]);
      out[i * NBAR + k] = acc;
    }
  }
}

void orc_frodo_as_plus_e(uint16_t *out, const uint16_t *st, const uint16_t *e, const uint8_t seed_a[16]) {
  const fparams p = fparams_of(976, 0);
  as_plus_e(&p, out, st, e, seed_a);
}

static void sa_plus_e(const fparams *p, uint16_t *out, const uint16_t *sp, const uint16_t *ep, const uint8_t seed_a[16]) {
  uint16_t row[NMAX];
  memcpy(out, ep, sizeof(uint16_t) * p->n * NBAR);
  for (int i = 0; i < p->n; ++i) {
    gen_row(p, row, seed_a, (uint16_t)i);
    for (int k = 0; k < NBAR; ++k) {
      uint16_t s = sp[k * p->n + i];
      for (int j = 0; j < p->n; ++j) out[k * p->n + j] = (uint16_t)(out[k * p->n + j] + s * row[j]);
    }
  }
}

void orc_frodo_sa_plus_e(uint16_t *out, const uint16_t *sp, const uint16_t *ep, const uint8_t seed_a[16]) {
  const fparams p = fparams_of(976, 0);
  sa_plus_e(&p, out, sp, ep, seed_a);
}

/* V (NBAR x NBAR) = S' (NBAR x N) * B (N x NBAR) + E'' */
static void sb_plus_e(const fparams *p, uint16_t *v, const uint16_t *b, const uint16_t *sp, const uint16_t *epp) {
  for (int i = 0; i < NBAR; ++i)
    for (int j = 0; j < NBAR; ++j) {
      uint16_t acc = epp[i * NBAR + j];
      for (int k = 0; k < p->n; ++k) acc = (uint16_t)(acc + sp[i * p->n + k] * b[k * NBAR + j]);
      v[i * NBAR + j] = acc;
    }
}

/* W (NBAR x NBAR) = B' (NBAR x N) * S, S given transposed (NBAR x N) */
static void mul_bs(const fparams *p, uint16_t *w, const uint16_t *bp, const uint16_t *st) {
  for (int i = 0; i < NBAR; ++i)
    for (int j = 0; j < NBAR; ++j) {
      uint16_t acc = 0;
      for (int k = 0; k < p->n; ++k) acc = (uint16_t)(acc + bp[i * p->n + k] * st[j * p->n + k]);
      w[i * NBAR + j] = acc;
    }
}

static void le_bytes_to_u16(uint16_t *out, const uint8_t *in, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (uint16_t)(in[2 * i] | (in[2 * i + 1] << 8));
}

static void mask_q(const fparams *p, uint16_t *v, size_t n) {
  const uint16_t m = (uint16_t)((1u << p->logq) - 1);
  for (size_t i = 0; i < n; ++i) v[i] &= m;
}

static int frodo_keypair(const fparams *p, uint8_t *pk, uint8_t *sk, const uint8_t *coins) {
  const int n = p->n, ls = p->len_sec;
  const uint8_t *s = coins, *seed_se = coins + ls, *z = coins + 2 * ls;
  uint8_t se_in[1 + 32];
  uint8_t *r = malloc(4 * (size_t)n * NBAR);
  uint16_t *st = calloc((size_t)n * NBAR, sizeof(uint16_t));
  uint16_t *e = malloc(sizeof(uint16_t) * n * NBAR);
  uint16_t *b = malloc(sizeof(uint16_t) * n * NBAR);
  if (!r || !st || !e || !b) {
    free(r); free(st); free(e); free(b);
    return -1;
  }
  /* seedA = SHAKE(z, 16) */
  shake(p, pk, LEN_SEEDA, z, LEN_SEEDA);
  /* S^T || E from SHAKE(0x5F || seedSE) */
  se_in[0] = 0x5F;
  memcpy(se_in + 1, seed_se, ls);
  shake(p, r, 4 * (size_t)n * NBAR, se_in, 1 + ls);
  le_bytes_to_u16(st, r, n * NBAR);
  le_bytes_to_u16(e, r + 2 * n * NBAR, n * NBAR);
  sample(p, st, n * NBAR);
  sample(p, e, n * NBAR);
  as_plus_e(p, b, st, e, pk);
  mask_q(p, b, n * NBAR);
  pack_d(pk + LEN_SEEDA, b, n * NBAR, p->logq);
  /* sk = s || pk || S^T (LE) || SHAKE(pk, len_sec) */
  memcpy(sk, s, ls);
  memcpy(sk + ls, pk, p->pk);
  for (int i = 0; i < n * NBAR; ++i) {
    sk[ls + p->pk + 2 * i] = (uint8_t)st[i];
    sk[ls + p->pk + 2 * i + 1] = (uint8_t)(st[i] >> 8);
  }
  shake(p, sk + p->sk - ls, ls, pk, p->pk);
  free(r); free(st); free(e); free(b);
  return 0;
}

/* Shared by encaps and the decaps re-encryption: from (seedSE, mu, pk) build
 * B' (NBAR x N) and C (NBAR x NBAR), reduced mod q. */
static int frodo_encrypt(const fparams *p, uint16_t *bp, uint16_t *c, const uint8_t *seed_se, const uint8_t *mu,
                         const uint8_t *pk) {
  const int n = p->n;
  uint8_t se_in[1 + 32];
  const size_t rlen = 2 * (2 * (size_t)n + NBAR) * NBAR;
  uint8_t *r = malloc(rlen);
  uint16_t *sp = malloc(sizeof(uint16_t) * (2 * n + NBAR) * NBAR);
  uint16_t *b = malloc(sizeof(uint16_t) * n * NBAR);
  uint16_t v[NBAR * NBAR], enc[NBAR * NBAR];
  if (!r || !sp || !b) {
    free(r); free(sp); free(b);
    return -1;
  }
  se_in[0] = 0x96;
  memcpy(se_in + 1, seed_se, p->len_sec);
  shake(p, r, rlen, se_in, 1 + p->len_sec);
  le_bytes_to_u16(sp, r, (2 * n + NBAR) * NBAR);
  sample(p, sp, (2 * n + NBAR) * NBAR);
  /* sp = S' (NBAR x N) | E' (NBAR x N) | E'' (NBAR x NBAR) */
  sa_plus_e(p, bp, sp, sp + n * NBAR, pk);
  unpack_d(b, pk + LEN_SEEDA, n * NBAR, p->logq);
  sb_plus_e(p, v, b, sp, sp + 2 * n * NBAR);
  encode(p, enc, mu);
  for (int i = 0; i < NBAR * NBAR; ++i) c[i] = (uint16_t)(v[i] + enc[i]);
  mask_q(p, bp, n * NBAR);
  mask_q(p, c, NBAR * NBAR);
  free(r); free(sp); free(b);
  return 0;
}

static int frodo_enc(const fparams *p, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) {
  const int n = p->n, ls = p->len_sec;
  uint8_t g2in[64], g2out[64];
  uint16_t *bp = malloc(sizeof(uint16_t) * n * NBAR);
  uint16_t c[NBAR * NBAR];
  uint8_t *fin = malloc(p->ct + ls);
  if (!bp || !fin) {
    free(bp); free(fin);
    return -1;
  }
  shake(p, g2in, ls, pk, p->pk); /* pkh */
  memcpy(g2in + ls, mu, ls);     /* len(mu) = B * 64 / 8 = len_sec */
  shake(p, g2out, 2 * ls, g2in, 2 * ls); /* seedSE || k */
  if (frodo_encrypt(p, bp, c, g2out, mu, pk)) {
    free(bp); free(fin);
    return -1;
  }
  pack_d(ct, bp, n * NBAR, p->logq);
  pack_d(ct + p->logq * n * NBAR / 8, c, NBAR * NBAR, p->logq);
  memcpy(fin, ct, p->ct);
  memcpy(fin + p->ct, g2out + ls, ls);
  shake(p, ss, ls, fin, p->ct + ls);
  free(bp); free(fin);
  return 0;
}

static int frodo_dec(const fparams *p, uint8_t *ss, const uint8_t *ct, const uint8_t *sk) {
  const int n = p->n, ls = p->len_sec;
  const uint8_t *sk_s = sk, *pk = sk + ls, *sk_st = sk + ls + p->pk;
  const uint8_t *pkh = sk + p->sk - ls;
  uint16_t *st = malloc(sizeof(uint16_t) * n * NBAR);
  uint16_t *bp = malloc(sizeof(uint16_t) * n * NBAR);
  uint16_t *bbp = malloc(sizeof(uint16_t) * n * NBAR);
  uint8_t *fin = malloc(p->ct + ls);
  uint16_t c[NBAR * NBAR], w[NBAR * NBAR], cc[NBAR * NBAR];
  uint8_t g2in[64], g2out[64];
  if (!st || !bp || !bbp || !fin) {
    free(st); free(bp); free(bbp); free(fin);
    return -1;
  }
  le_bytes_to_u16(st, sk_st, n * NBAR);
  unpack_d(bp, ct, n * NBAR, p->logq);
  unpack_d(c, ct + p->logq * n * NBAR / 8, NBAR * NBAR, p->logq);
  mul_bs(p, w, bp, st);
  for (int i = 0; i < NBAR * NBAR; ++i) w[i] = (uint16_t)(c[i] - w[i]);
  memcpy(g2in, pkh, ls);
  decode(p, g2in + ls, w);
  shake(p, g2out, 2 * ls, g2in, 2 * ls);
  if (frodo_encrypt(p, bbp, cc, g2out, g2in + ls, pk)) {
    free(st); free(bp); free(bbp); free(fin);
    return -1;
  }
  /* constant-time equality of (B', C) with the re-encryption */
  uint16_t diff = 0;
  for (int i = 0; i < n * NBAR; ++i) diff |= (uint16_t)(bp[i] ^ bbp[i]);
  for (int i = 0; i < NBAR * NBAR; ++i) diff |= (uint16_t)(c[i] ^ cc[i]);
  uint8_t mask = (uint8_t)(-(int)((diff | (uint16_t)-diff) >> 15)); /* 0xFF on mismatch */
  memcpy(fin, ct, p->ct);
  for (int i = 0; i < ls; ++i) fin[p->ct + i] = (uint8_t)((g2out[ls + i] & ~mask) | (sk_s[i] & mask));
  shake(p, ss, ls, fin, p->ct + ls);
  free(st); free(bp); free(bbp); free(fin);
  return 0;
}

#define FP(n, aes) const fparams P = fparams_of(n, aes)
int orc_frodo_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) { FP(976, 0); return frodo_keypair(&P, pk, sk, coins); }
int orc_frodo_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) { FP(976, 0); return frodo_enc(&P, ct, ss, pk, mu); }
int orc_frodo_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) { FP(976, 0); return frodo_dec(&P, ss, ct, sk); }
int orc_frodo_aes_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins) { FP(976, 1); return frodo_keypair(&P, pk, sk, coins); }
int orc_frodo_aes_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) {
  FP(976, 1);
  return frodo_enc(&P, ct, ss, pk, mu);
}
int orc_frodo_aes_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk) { FP(976, 1); return frodo_dec(&P, ss, ct, sk); }
static int valid_n(int n) { return n == 640 || n == 976 || n == 1344; }
int orc_frodo_n_keypair(int n, uint8_t *pk, uint8_t *sk, const uint8_t *coins) {
  if (!valid_n(n)) return -1;
  FP(n, 0);
  return frodo_keypair(&P, pk, sk, coins);
}
int orc_frodo_n_enc(int n, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu) {
  if (!valid_n(n)) return -1;
  FP(n, 0);
  return frodo_enc(&P, ct, ss, pk, mu);
}
int orc_frodo_n_dec(int n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk) {
  if (!valid_n(n)) return -1;
  FP(n, 0);
  return frodo_dec(&P, ss, ct, sk);
}
