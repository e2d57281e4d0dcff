"""Frodo-640 and Frodo-1344 (SHAKE) in the oracle.

The reference implements only FrodoKEM-976; the other two round-3 parameter
sets share its algorithm with n = 640 / 1344, D = 15 / 16 bits, B = 2 / 4
message bits per entry, 16 / 32-byte secrets, SHAKE128 / SHAKE256 and their
own CDF tables.  No known-answer file for them is in the container, so the
oracle is pinned by:

- the sampler: over all 2^16 inputs its distribution must be the rounded
  continuous Gaussian of the specification's sigma (2.8 / 2.3 / 1.4) to
  within 1e-4 total variation -- one CDF entry off by a few counts fails;
- numpy/hashlib restatements of keygen and encaps (15-bit MSB-first packing,
  B-bit encoding, SHAKE128 for Frodo-640, SHAKE128 A rows everywhere) at
  every output byte;
- decaps(encaps) round trips and implicit rejection on tampered ciphertexts.
"""
import hashlib

import numpy as np
import pytest
from scipy.stats import norm

PARAMS = {
    # scheme: n, D, B, len_sec, shake, sigma, CDF table
    5: (640, 15, 2, 16, hashlib.shake_128, 2.8,
        [4643, 13363, 20579, 25843, 29227, 31145, 32103, 32525, 32689, 32745, 32762, 32766, 32767]),
    6: (1344, 16, 4, 32, hashlib.shake_256, 1.4, [9142, 23462, 30338, 32361, 32725, 32765, 32767]),
}


@pytest.mark.parametrize("n,sigma", [(640, 2.8), (976, 2.3), (1344, 1.4)])
def test_sampler_is_the_rounded_gaussian(oracle, n, sigma):
    e = oracle.frodo_n_sample(n, np.arange(65536, dtype=np.uint16)).view(np.int16).astype(np.int64)
    k = np.arange(-20, 21)
    target = norm.cdf((k + 0.5) / sigma) - norm.cdf((k - 0.5) / sigma)
    got = np.bincount(e + 20, minlength=41) / 65536
    assert 0.5 * np.abs(got - target).sum() < 1e-4
    assert np.abs(got - target).max() < 2e-5
    # the low bit of the input is the sign: e(2t) = -e(2t + 1) for nonzero e
    assert np.array_equal(e[0::2], -e[1::2])


def _sample_np(r, cdf):
    t = (r >> 1).astype(np.int64)
    mag = np.searchsorted(np.array(cdf[:-1]), t, side="left")  # #{j: CDF[j] < t}
    return np.where(r & 1, -mag, mag).astype(np.int64)


def _a_matrix(seed_a, n):
    # A rows come from SHAKE128 in every parameter set (as in Frodo-976)
    rows = [np.frombuffer(hashlib.shake_128(i.to_bytes(2, "little") + seed_a).digest(2 * n), "<u2")
            for i in range(n)]
    return np.stack(rows).astype(np.int64)


def _pack(v, d):
    bits = np.unpackbits(v.astype(">u2").view(np.uint8)).reshape(-1, 16)[:, 16 - d:]
    return np.packbits(bits.reshape(-1)).tobytes()


def _encode(mu, b, logq):
    bits = np.unpackbits(np.frombuffer(mu, np.uint8), bitorder="little").reshape(64, b)
    return ((bits * (1 << np.arange(b))).sum(1) << (logq - b)).reshape(8, 8)


@pytest.mark.parametrize("scheme", [5, 6])
def test_keygen_and_encaps_vs_numpy_restatement(oracle, scheme):
    n, d, b, ls, shake, _, cdf = PARAMS[scheme]
    q = 1 << d
    coins = oracle.prng_bytes(600 + scheme, 2 * ls + 16)
    pk, sk = oracle.keypair(scheme, coins[None])
    pkb, skb = pk[0].tobytes(), sk[0].tobytes()
    s, seed_se, z = coins[:ls].tobytes(), coins[ls:2 * ls].tobytes(), coins[2 * ls:].tobytes()
    seed_a = shake(z).digest(16)
    r = np.frombuffer(shake(b"\x5f" + seed_se).digest(32 * n), "<u2")
    st = _sample_np(r[:8 * n], cdf).reshape(8, n)
    e = _sample_np(r[8 * n:], cdf).reshape(n, 8)
    A = _a_matrix(seed_a, n)
    B = (A @ st.T + e) % q
    assert len(pkb) == 16 + d * n and pkb == seed_a + _pack(B, d)
    assert skb == s + pkb + (st % 65536).astype("<u2").tobytes() + shake(pkb).digest(ls)

    mu = oracle.prng_bytes(700 + scheme, ls).tobytes()
    ct, ss = oracle.enc(scheme, pk, np.frombuffer(mu, np.uint8)[None])
    g2 = shake(shake(pkb).digest(ls) + mu).digest(2 * ls)
    r = np.frombuffer(shake(b"\x96" + g2[:ls]).digest(2 * (16 * n + 64)), "<u2")
    smp = _sample_np(r, cdf)
    sp, ep, epp = smp[:8 * n].reshape(8, n), smp[8 * n:16 * n].reshape(8, n), smp[16 * n:].reshape(8, 8)
    Bp = (sp @ A + ep) % q
    C = (sp @ B + epp + _encode(mu, b, d)) % q
    exp_ct = _pack(Bp, d) + _pack(C, d)
    assert len(exp_ct) == d * (n + 8) and ct[0].tobytes() == exp_ct
    assert ss[0].tobytes() == shake(exp_ct + g2[ls:]).digest(ls)


@pytest.mark.parametrize("scheme", [5, 6])
def test_roundtrip_and_implicit_rejection(oracle, scheme):
    n, d, b, ls, shake, _, _ = PARAMS[scheme]
    m = 3
    pk, sk = oracle.keypair(scheme, oracle.prng_bytes(800 + scheme, m * (2 * ls + 16)).reshape(m, -1), threads=3)
    ct, ss = oracle.enc(scheme, pk, oracle.prng_bytes(900 + scheme, m * ls).reshape(m, -1), threads=3)
    assert (oracle.dec(scheme, ct, sk, threads=3) == ss).all()
    bad = ct.copy()
    bad[0, 3] ^= 0x01          # inside B'
    bad[1, -1] ^= 0x80         # inside C (last entry)
    bad[2, d * n // 8] ^= 0x40  # first byte of C
    ss_bad = oracle.dec(scheme, bad, sk, threads=3)
    for i in range(m):
        assert ss_bad[i].tobytes() == shake(bad[i].tobytes() + sk[i, :ls].tobytes()).digest(ls)

