/*
 * Threaded batch driver over the oracle (TEST INFRASTRUCTURE ONLY): used by
 * the test suite for large parity batches and by the benchmark's CPU
 * baseline leg.  Instances are independent, so each thread takes a strided
 * slice of the batch.
 *
 * Buffer roles per (scheme, op), all instance-major and contiguous:
 *   keygen: a = pk[n],  b = sk[n],  c = coins[n]
 *   enc:    a = ct[n],  b = ss[n],  c = pk[n],  d = coins[n]
 *   dec:    a = ss[n],  c = ct[n],  d = sk[n]
 */
#include <pthread.h>
#include <stdlib.h>

#include "oracle.h"

typedef struct {
  int scheme, op;
  size_t begin, end;
  uint8_t *a, *b;
  const uint8_t *c, *d;
  int rc;
} job_t;

static void sizes(int scheme, size_t *pk, size_t *sk, size_t *ct, size_t *ss, size_t *kc, size_t *ec) {
  if (scheme == 0 || scheme == 2) { /* FrodoKEM-976-SHAKE / -AES */
    *pk = ORC_FRODO_PK; *sk = ORC_FRODO_SK; *ct = ORC_FRODO_CT; *ss = ORC_FRODO_SS;
    *kc = ORC_FRODO_KEYGEN_COINS; *ec = ORC_FRODO_ENC_COINS;
  } else if (scheme == 5 || scheme == 6) { /* Frodo-640 / Frodo-1344 (SHAKE) */
    const size_t n = scheme == 5 ? 640 : 1344, d = scheme == 5 ? 15 : 16, ls = scheme == 5 ? 16 : 32;
    *pk = 16 + d * n; *sk = ls + *pk + 16 * n + ls; *ct = d * (n + 8); *ss = ls;
    *kc = 2 * ls + 16; *ec = ls; /* s || seedSE || z(16) */
  } else {
    const size_t l = scheme == 1 ? 3 : scheme == 3 ? 2 : 4, et = l == 2 ? 3 : l == 3 ? 4 : 6;
    *pk = 320 * l + 32; *sk = 416 * l + *pk + 64; *ct = 320 * l + 32 * et; *ss = ORC_SABER_SS;
    *kc = ORC_SABER_KEYGEN_COINS; *ec = ORC_SABER_ENC_COINS;
  }
}

static void *run(void *arg) {
  job_t *j = arg;
  size_t pk, sk, ct, ss, kc, ec;
  sizes(j->scheme, &pk, &sk, &ct, &ss, &kc, &ec);
  for (size_t i = j->begin; i < j->end; ++i) {
    int rc;
    const int l = j->scheme == 1 ? 3 : j->scheme == 3 ? 2 : 4; /* Saber family rank */
    const int fn = j->scheme == 5 ? 640 : 1344;                 /* Frodo-640 / -1344 */
    if (j->scheme == 5 || j->scheme == 6)
      rc = j->op == 0   ? orc_frodo_n_keypair(fn, j->a + i * pk, j->b + i * sk, j->c + i * kc)
           : j->op == 1 ? orc_frodo_n_enc(fn, j->a + i * ct, j->b + i * ss, j->c + i * pk, j->d + i * ec)
                        : orc_frodo_n_dec(fn, j->a + i * ss, j->c + i * ct, j->d + i * sk);
    else if (j->op == 0)
      rc = j->scheme == 0   ? orc_frodo_keypair(j->a + i * pk, j->b + i * sk, j->c + i * kc)
           : j->scheme == 2 ? orc_frodo_aes_keypair(j->a + i * pk, j->b + i * sk, j->c + i * kc)
                            : orc_saber_l_keypair(l, j->a + i * pk, j->b + i * sk, j->c + i * kc);
    else if (j->op == 1)
      rc = j->scheme == 0   ? orc_frodo_enc(j->a + i * ct, j->b + i * ss, j->c + i * pk, j->d + i * ec)
           : j->scheme == 2 ? orc_frodo_aes_enc(j->a + i * ct, j->b + i * ss, j->c + i * pk, j->d + i * ec)
                            : orc_saber_l_enc(l, j->a + i * ct, j->b + i * ss, j->c + i * pk, j->d + i * ec);
    else
      rc = j->scheme == 0   ? orc_frodo_dec(j->a + i * ss, j->c + i * ct, j->d + i * sk)
           : j->scheme == 2 ? orc_frodo_aes_dec(j->a + i * ss, j->c + i * ct, j->d + i * sk)
                            : orc_saber_l_dec(l, j->a + i * ss, j->c + i * ct, j->d + i * sk);
    if (rc) j->rc = rc;
  }
  return NULL;
}

int orc_batch(int scheme, int op, size_t n, int threads, uint8_t *a, uint8_t *b, const uint8_t *c,
              const uint8_t *d) {
  if (scheme < 0 || scheme > 6 || op < 0 || op > 2 || threads < 1) return -1;
  if ((size_t)threads > n) threads = n ? (int)n : 1;
  pthread_t *tid = calloc((size_t)threads, sizeof *tid);
  job_t *jobs = calloc((size_t)threads, sizeof *jobs);
  if (!tid || !jobs) {
    free(tid); free(jobs);
    return -1;
  }
  size_t per = (n + (size_t)threads - 1) / (size_t)threads;
  int rc = 0;
  for (int t = 0; t < threads; ++t) {
    size_t lo = (size_t)t * per, hi = lo + per;
    if (lo > n) lo = n;
    if (hi > n) hi = n;
    jobs[t] = (job_t){scheme, op, lo, hi, a, b, c, d, 0};
    if (threads == 1) run(&jobs[t]);
    else if (pthread_create(&tid[t], NULL, run, &jobs[t])) rc = -1;
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  for (int t = 0; t < threads; ++t)
    if (jobs[t].rc) rc = jobs[t].rc;
  free(tid); free(jobs);
  return rc;
}
