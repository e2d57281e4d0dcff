// AES-256 CTR_DRBG (NIST SP 800-90A, no derivation function) and the
// known-answer-test coin streams built on it.
//
// The NIST PQC KAT files were produced by seeding this DRBG with a 48-byte
// seed per test count and calling randombytes() exactly as the scheme's
// reference code does.  The reference here hard-codes its randomness instead
// (FrodoKEM976-GPU/src/cuda_kernel.cu:51, SABER-GPU/src/cuda_kernel.cu:62) and
// only declares AES (FrodoKEM976-GPU/include/aes.h:59-67), so this is new
// host code: it turns per-count seeds into the explicit coins that
// dpkem_keypair / dpkem_enc take, so a whole KAT file is one GPU batch.
//
// AES-256 follows FIPS-197: the S-box is computed at first use from the
// multiplicative inverse in GF(2^8) and the affine map, not tabulated here.
#include <string.h>

#include <mutex>

#include "../../include/dpkem.h"

namespace {

uint8_t g_sbox[256];
std::once_flag g_sbox_once;

uint8_t xtime(uint8_t a) { return (uint8_t)((a << 1) ^ ((a & 0x80) ? 0x1B : 0)); }

uint8_t gmul(uint8_t a, uint8_t b) {
  uint8_t r = 0;
  while (b) {
    if (b & 1) r ^= a;
    a = xtime(a);
    b >>= 1;
  }
  return r;
}

void build_sbox() {
  for (int x = 0; x < 256; ++x) {
    // inverse by exponentiation: x^254 (0 maps to 0)
    uint8_t inv = 1, base = (uint8_t)x;
    for (int e = 254; e; e >>= 1) {
      if (e & 1) inv = gmul(inv, base);
      base = gmul(base, base);
    }
    if (x == 0) inv = 0;
    uint8_t s = inv;
    for (int k = 1; k <= 4; ++k) s ^= (uint8_t)((inv << k) | (inv >> (8 - k)));
    g_sbox[x] = s ^ 0x63;
  }
}

struct Aes256 {
  uint8_t rk[15][16];
  explicit Aes256(const uint8_t key[32]) {
    std::call_once(g_sbox_once, build_sbox);
    uint8_t w[60][4];
    memcpy(w, key, 32);
    uint8_t rcon = 1;
    for (int i = 8; i < 60; ++i) {
      uint8_t t[4] = {w[i - 1][0], w[i - 1][1], w[i - 1][2], w[i - 1][3]};
      if (i % 8 == 0) {
        const uint8_t t0 = t[0];
        t[0] = (uint8_t)(g_sbox[t[1]] ^ rcon);
        t[1] = g_sbox[t[2]];
        t[2] = g_sbox[t[3]];
        t[3] = g_sbox[t0];
        rcon = xtime(rcon);
      } else if (i % 8 == 4) {
        for (auto& b : t) b = g_sbox[b];
      }
      for (int j = 0; j < 4; ++j) w[i][j] = w[i - 8][j] ^ t[j];
    }
    memcpy(rk, w, sizeof rk);
  }
  void encrypt(const uint8_t in[16], uint8_t out[16]) const {
    uint8_t s[16];
    for (int i = 0; i < 16; ++i) s[i] = in[i] ^ rk[0][i];
    for (int r = 1; r <= 14; ++r) {
      uint8_t t[16];
      // SubBytes + ShiftRows (state is column-major: byte 4c + row)
      for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) t[4 * c + row] = g_sbox[s[4 * ((c + row) & 3) + row]];
      if (r != 14) {  // MixColumns
        for (int c = 0; c < 4; ++c) {
          uint8_t* a = t + 4 * c;
          const uint8_t a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], all = a0 ^ a1 ^ a2 ^ a3;
          a[0] = a0 ^ all ^ xtime(a0 ^ a1);
          a[1] = a1 ^ all ^ xtime(a1 ^ a2);
          a[2] = a2 ^ all ^ xtime(a2 ^ a3);
          a[3] = a3 ^ all ^ xtime(a3 ^ a0);
        }
      }
      for (int i = 0; i < 16; ++i) s[i] = t[i] ^ rk[r][i];
    }
    memcpy(out, s, 16);
  }
};

void incr_v(uint8_t v[16]) {  // 128-bit big-endian increment
  for (int j = 15; j >= 0; --j)
    if (++v[j] != 0) break;
}

// CTR_DRBG_Update with AES-256 (seedlen = 48 bytes).
void drbg_update(dpkem_drbg* d, const uint8_t* provided) {
  uint8_t temp[48];
  const Aes256 aes(d->key);
  for (int i = 0; i < 3; ++i) {
    incr_v(d->v);
    aes.encrypt(d->v, temp + 16 * i);
  }
  if (provided)
    for (int i = 0; i < 48; ++i) temp[i] ^= provided[i];
  memcpy(d->key, temp, 32);
  memcpy(d->v, temp + 32, 16);
}

}  // namespace

extern "C" {

int dpkem_aes256_ecb(const uint8_t key[32], const uint8_t* in, uint8_t* out, size_t nblocks) {
  if (!key || (nblocks && (!in || !out))) return DPKEM_EINVAL;
  const Aes256 aes(key);
  for (size_t b = 0; b < nblocks; ++b) aes.encrypt(in + 16 * b, out + 16 * b);
  return DPKEM_OK;
}

int dpkem_drbg_init(dpkem_drbg* d, const uint8_t entropy[48], const uint8_t* personalization) {
  if (!d || !entropy) return DPKEM_EINVAL;
  uint8_t seed[48];
  memcpy(seed, entropy, 48);
  if (personalization)
    for (int i = 0; i < 48; ++i) seed[i] ^= personalization[i];
  memset(d->key, 0, 32);
  memset(d->v, 0, 16);
  drbg_update(d, seed);
  d->reseed_counter = 1;
  return DPKEM_OK;
}

int dpkem_drbg_bytes(dpkem_drbg* d, uint8_t* out, size_t len) {
  if (!d || (len && !out)) return DPKEM_EINVAL;
  {
    const Aes256 aes(d->key);
    uint8_t block[16];
    while (len > 0) {
      incr_v(d->v);
      aes.encrypt(d->v, block);
      const size_t k = len < 16 ? len : 16;
      memcpy(out, block, k);
      out += k;
      len -= k;
    }
  }
  drbg_update(d, nullptr);
  d->reseed_counter++;
  return DPKEM_OK;
}

void dpkem_randombytes_drbg(uint8_t* out, size_t len, void* ctx) {
  (void)dpkem_drbg_bytes(static_cast<dpkem_drbg*>(ctx), out, len);
}

int dpkem_kat_seeds(size_t n, uint8_t* seeds) {
  if (n && !seeds) return DPKEM_EINVAL;
  uint8_t entropy[48];
  for (int i = 0; i < 48; ++i) entropy[i] = (uint8_t)i;
  dpkem_drbg d;
  dpkem_drbg_init(&d, entropy, nullptr);
  for (size_t i = 0; i < n; ++i) dpkem_drbg_bytes(&d, seeds + 48 * i, 48);
  return DPKEM_OK;
}

int dpkem_kat_coins(int scheme, size_t n, const uint8_t* seeds, uint8_t* kcoins, uint8_t* ecoins) {
  if (scheme < DPKEM_FRODO976 || scheme > DPKEM_FRODO1344 || (n && (!seeds || !kcoins || !ecoins)))
    return DPKEM_EINVAL;
  const bool frodo = scheme == DPKEM_FRODO976 || scheme == DPKEM_FRODO976_AES || scheme == DPKEM_FRODO640 ||
                     scheme == DPKEM_FRODO1344;
  // FrodoKEM secret length: 16 / 24 / 32 bytes for Frodo-640 / -976 / -1344
  const size_t sec = scheme == DPKEM_FRODO640 ? 16 : scheme == DPKEM_FRODO1344 ? 32 : 24;
  for (size_t i = 0; i < n; ++i) {
    dpkem_drbg d;
    dpkem_drbg_init(&d, seeds + 48 * i, nullptr);
    if (frodo) {
      // keypair: randombytes(s || seedSE || z, 2 sec + 16); encaps: randombytes(mu, sec)
      dpkem_drbg_bytes(&d, kcoins + (2 * sec + 16) * i, 2 * sec + 16);
      dpkem_drbg_bytes(&d, ecoins + sec * i, sec);
    } else {
      // keypair: randombytes(seedA, 32); randombytes(seed_s, 32); randombytes(z, 32);
      // encaps: randombytes(m, 32)
      for (int k = 0; k < 3; ++k) dpkem_drbg_bytes(&d, kcoins + 96 * i + 32 * k, 32);
      dpkem_drbg_bytes(&d, ecoins + 32 * i, 32);
    }
  }
  return DPKEM_OK;
}

}  // extern "C"
