"""FrodoKEM-976-AES in the oracle: AES-128 and the A rows it generates.

Pinned by: the FIPS-197 AES-128 example; OpenSSL's AES-128 (libcrypto via
ctypes) on random keys and blocks and on the A-row block layout
(AES-128_seedA(LE16(i) || LE16(j) || 0^12) -> A[i][j..j+7], little-endian);
numpy restatements of keygen and encaps built on OpenSSL's AES; round trip
and implicit rejection.  No FrodoKEM-AES known-answer file is in the
container, so end-to-end bytes stay pinned by the restatements only.
"""
import hashlib

import numpy as np
import pytest

from openssl_aes import aes_ecb

FRODO_AES = 2

ossl = aes_ecb()
needs_openssl = pytest.mark.skipif(ossl is None, reason="libcrypto not found")


def test_aes128_fips197_example(oracle):
    key = bytes(range(16))
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert oracle.aes128_ecb(key, pt).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


@needs_openssl
def test_aes128_matches_openssl(oracle):
    rng = np.random.default_rng(17)
    for _ in range(20):
        key = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
        data = rng.integers(0, 256, 16 * 9, dtype=np.uint8).tobytes()
        assert oracle.aes128_ecb(key, data) == ossl(key, data)


def _row_blocks(i):
    blk = np.zeros((122, 8), np.uint16)
    blk[:, 0] = i
    blk[:, 1] = np.arange(0, 976, 8)
    return blk.astype("<u2").tobytes()


def _frodo_aes_a(seed):
    return np.stack([np.frombuffer(ossl(seed, _row_blocks(i)), "<u2") for i in range(976)]).astype(np.int64)


@needs_openssl
def test_frodo_aes_rows_vs_openssl(oracle):
    seed = bytes(range(100, 116))
    for i in (0, 1, 255, 256, 975):
        exp = np.frombuffer(ossl(seed, _row_blocks(i)), "<u2")
        assert np.array_equal(oracle.frodo_aes_gen_a_row(seed, i), exp)
    assert not np.array_equal(oracle.frodo_aes_gen_a_row(seed, 3), oracle.frodo_gen_a_row(seed, 3))


@needs_openssl
def test_frodo_aes_keygen_and_encaps_vs_numpy_restatement(oracle):
    from test_oracle import _frodo_sample_np
    coins = oracle.prng_bytes(41, 64)
    pk, sk = oracle.keypair(FRODO_AES, coins[None])
    pkb, skb = pk[0].tobytes(), sk[0].tobytes()
    s, seed_se, z = coins[:24].tobytes(), coins[24:48].tobytes(), coins[48:].tobytes()
    seed_a = hashlib.shake_256(z).digest(16)
    r = np.frombuffer(hashlib.shake_256(b"\x5f" + seed_se).digest(31232), "<u2")
    st = _frodo_sample_np(r[:7808]).reshape(8, 976)
    e = _frodo_sample_np(r[7808:]).reshape(976, 8)
    A = _frodo_aes_a(seed_a)
    B = (A @ st.T.astype(np.int64) + e) % 65536
    assert pkb == seed_a + B.astype(">u2").tobytes()
    assert skb == s + pkb + st.astype("<u2").tobytes() + hashlib.shake_256(pkb).digest(24)
    mu = oracle.prng_bytes(42, 24)
    ct, ss = oracle.enc(FRODO_AES, pk, mu[None])
    pkh = hashlib.shake_256(pkb).digest(24)
    g2 = hashlib.shake_256(pkh + mu.tobytes()).digest(48)
    r = np.frombuffer(hashlib.shake_256(b"\x96" + g2[:24]).digest(31360), "<u2")
    smp = _frodo_sample_np(r)
    sp, ep, epp = smp[:7808].reshape(8, 976), smp[7808:15616].reshape(8, 976), smp[15616:].reshape(8, 8)
    Bp = (sp.astype(np.int64) @ A + ep) % 65536
    bits = np.unpackbits(mu, bitorder="little").reshape(64, 3)
    enc = ((bits * [1, 2, 4]).sum(1) << 13).reshape(8, 8)
    C = (sp.astype(np.int64) @ B + epp + enc) % 65536
    exp_ct = Bp.astype(">u2").tobytes() + C.astype(">u2").tobytes()
    assert ct[0].tobytes() == exp_ct
    assert ss[0].tobytes() == hashlib.shake_256(exp_ct + g2[24:]).digest(24)


def test_frodo_aes_roundtrip_and_tamper(oracle):
    n = 2
    coins = oracle.prng_bytes(51, n * 64).reshape(n, 64)
    pk, sk = oracle.keypair(FRODO_AES, coins, threads=2)
    mu = oracle.prng_bytes(52, n * 24).reshape(n, 24)
    ct, ss = oracle.enc(FRODO_AES, pk, mu, threads=2)
    assert (oracle.dec(FRODO_AES, ct, sk, threads=2) == ss).all()
    bad = ct.copy(); bad[:, 200] ^= 0x08
    ss_bad = oracle.dec(FRODO_AES, bad, sk, threads=2)
    for i in range(n):
        assert ss_bad[i].tobytes() == hashlib.shake_256(bad[i].tobytes() + sk[i, :24].tobytes()).digest(24)
    # same coins, different A: the SHAKE variant gives a different public key
    pk0, _ = oracle.keypair(0, coins[:1])
    assert pk0[0, :16].tobytes() == pk[0, :16].tobytes() and not np.array_equal(pk0[0], pk[0])
