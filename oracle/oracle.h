/*
 * CPU oracle for the batched KEM engine -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a from-scratch C restatement of FIPS 202 (Keccak / SHAKE /
 * SHA3), FrodoKEM-976-SHAKE (round 3) and Saber l=3 (round 3) with explicit
 * coins.  It exists to check the HIP engine byte-for-byte.  Only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg may load it;
 * the product path (dpcrypto_amd/libdpkem.so) never links or calls it.
 *
 * Reference anchors (files in benlwk/DPCrypto):
 *   Frodo parameters / sizes      FrodoKEM976-GPU/include/params.h:9-24, include/api.h:7-10
 *   Frodo keygen / enc / dec flow FrodoKEM976-GPU/src/cuda_kernel.cu:21-335
 *   Saber parameters / sizes      SABER-GPU/include/params.h:6-48
 *   Saber enc / dec flow          SABER-GPU/src/cuda_kernel.cu:19-299
 * Where the reference deviates from the published scheme (per-instance
 * inputs ignored, Saber keygen absent, broken implicit rejection), the
 * published round-3 algorithm is followed.
 */
#ifndef DPKEM_ORACLE_H
#define DPKEM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- FIPS 202 ---------------- */
void orc_keccak_f1600(uint64_t st[25]);
void orc_shake128(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void orc_shake256(uint8_t *out, size_t outlen, const uint8_t *in, size_t inlen);
void orc_sha3_256(uint8_t out[32], const uint8_t *in, size_t inlen);
void orc_sha3_512(uint8_t out[64], const uint8_t *in, size_t inlen);

/* ---------------- seeded PRNG (xoshiro256** seeded by SplitMix64) ---------------- */
typedef struct { uint64_t s[4]; } orc_prng;
void orc_prng_seed(orc_prng *p, uint64_t seed);
uint64_t orc_prng_next(orc_prng *p);
/* Fill `len` bytes with little-endian bytes of successive outputs. */
void orc_prng_bytes(orc_prng *p, uint8_t *out, size_t len);
/* Convenience for bindings: seed, then fill. */
void orc_prng_fill(uint64_t seed, uint8_t *out, size_t len);

/* ---------------- FrodoKEM-976-SHAKE ---------------- */
#define ORC_FRODO_N 976
#define ORC_FRODO_NBAR 8
#define ORC_FRODO_PK 15632
#define ORC_FRODO_SK 31296
#define ORC_FRODO_CT 15744
#define ORC_FRODO_SS 24
#define ORC_FRODO_KEYGEN_COINS 64 /* s(24) || seedSE(24) || z(16) */
#define ORC_FRODO_ENC_COINS 24    /* mu */

/* FrodoKEM-976-AES: same sizes and coins; A rows from AES-128 (orc_frodo_aes_gen_a_row). */
int orc_frodo_aes_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins);
int orc_frodo_aes_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu);
int orc_frodo_aes_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
void orc_frodo_aes_gen_a_row(uint16_t row[ORC_FRODO_N], const uint8_t seed_a[16], uint16_t i);
void orc_aes128_ecb(const uint8_t key[16], const uint8_t *in, uint8_t *out, size_t nblocks);
/* Frodo-640 / Frodo-1344 (SHAKE), selected by n = 640 or 1344 (976 also
 * accepted): pk 16 + D*n, sk len_sec + pk + 16n + len_sec, ct D*(n + 8),
 * ss len_sec; coins s || seedSE || z (2 len_sec + 16) and mu (len_sec). */
int orc_frodo_n_keypair(int n, uint8_t *pk, uint8_t *sk, const uint8_t *coins);
int orc_frodo_n_enc(int n, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu);
int orc_frodo_n_dec(int n, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
void orc_frodo_n_sample(int n, uint16_t *s, size_t count);

int orc_frodo_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins);
int orc_frodo_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *mu);
int orc_frodo_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk);

/* building blocks, exported for unit tests */
void orc_frodo_sample(uint16_t *s, size_t n);
void orc_frodo_gen_a_row(uint16_t row[ORC_FRODO_N], const uint8_t seed_a[16], uint16_t i);
void orc_frodo_pack16(uint8_t *out, const uint16_t *in, size_t n);
void orc_frodo_unpack16(uint16_t *out, const uint8_t *in, size_t n);
void orc_frodo_encode(uint16_t out[64], const uint8_t mu[24]);
void orc_frodo_decode(uint8_t mu[24], const uint16_t in[64]);
/* out (N x NBAR) = A * S + E, with S given transposed (NBAR x N). */
void orc_frodo_as_plus_e(uint16_t *out, const uint16_t *st, const uint16_t *e, const uint8_t seed_a[16]);
/* out (NBAR x N) = S' * A + E'. */
void orc_frodo_sa_plus_e(uint16_t *out, const uint16_t *sp, const uint16_t *ep, const uint8_t seed_a[16]);

/* ---------------- Saber (l = 3) ---------------- */
#define ORC_SABER_PK 992
#define ORC_SABER_SK 2304
#define ORC_SABER_CT 1088
#define ORC_SABER_SS 32
#define ORC_SABER_KEYGEN_COINS 96 /* seedA_raw(32) || seed_s(32) || z(32) */
#define ORC_SABER_ENC_COINS 32    /* m_raw */

/* Saber family by module rank l: 2 = LightSaber, 3 = Saber, 4 = FireSaber.
 * pk = 320 l + 32, sk = 416 l + pk + 64, ct = 320 l + 32 eT (eT = 3, 4, 6). */
int orc_saber_l_keypair(int l, uint8_t *pk, uint8_t *sk, const uint8_t *coins);
int orc_saber_l_enc(int l, uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *m_raw);
int orc_saber_l_dec(int l, uint8_t *ss, const uint8_t *ct, const uint8_t *sk);
void orc_saber_l_gen_matrix(int l, uint16_t *a, const uint8_t seed[32]);
void orc_saber_l_gen_secret(int l, uint16_t *s, const uint8_t seed[32]);
int orc_saber_keypair(uint8_t *pk, uint8_t *sk, const uint8_t *coins);
int orc_saber_enc(uint8_t *ct, uint8_t *ss, const uint8_t *pk, const uint8_t *m_raw);
int orc_saber_dec(uint8_t *ss, const uint8_t *ct, const uint8_t *sk);

/* building blocks, exported for unit tests */
void orc_saber_gen_matrix(uint16_t a[3 * 3 * 256], const uint8_t seed[32]);
void orc_saber_gen_secret(uint16_t s[3 * 256], const uint8_t seed[32]);
void orc_saber_polymul_acc(const uint16_t *a, const uint16_t *b, uint16_t *res);
void orc_saber_matvec(uint16_t *res, const uint16_t *a, const uint16_t *s, int transpose);
void orc_saber_pack13(uint8_t *out, const uint16_t *in, size_t npoly);
void orc_saber_unpack13(uint16_t *out, const uint8_t *in, size_t npoly);
void orc_saber_pack10(uint8_t *out, const uint16_t *in, size_t npoly);
void orc_saber_unpack10(uint16_t *out, const uint8_t *in, size_t npoly);
/* IND-CPA encryption with explicit message and noise seed (ct = 1088 B). */
void orc_saber_indcpa_enc(uint8_t *ct, const uint8_t m[32], const uint8_t seed_sp[32], const uint8_t *pk);

/* ---------------- batched / threaded drivers (CPU baseline) ---------------- */
/* scheme: 0 = Frodo-976-SHAKE, 1 = Saber, 2 = Frodo-976-AES, 3 = LightSaber, 4 = FireSaber.
 * op: 0 keygen, 1 enc, 2 dec.
 * Runs `n` independent instances over `threads` POSIX threads. */
int orc_batch(int scheme, int op, size_t n, int threads,
              uint8_t *a, uint8_t *b, const uint8_t *c, const uint8_t *d);

#ifdef __cplusplus
}
#endif
#endif
