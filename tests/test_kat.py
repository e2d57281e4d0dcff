"""AES-256 CTR_DRBG and NIST KAT (.rsp) generation.

Pinned by: the FIPS-197 AES-256 example vector; OpenSSL's AES-256 (libcrypto
through ctypes) on random keys and blocks; the first two per-count seeds that
open every NIST PQC KAT file (DRBG seeded with bytes 0..47); and, for the
coin streams, the split-call property of the DRBG (every randombytes() call
ends with an update, so three 32-byte draws differ from one 96-byte draw
after the first 32 bytes).  KAT responses computed by the oracle round-trip
through the .rsp format.  The official .rsp files for these two schemes are
not in the container, so their contents stay unpinned beyond this.
"""
import numpy as np
import pytest

from dpcrypto_amd import kat
import dpcrypto_amd as D

# count 0 and count 1 seeds of every NIST PQC KAT file
KAT_SEED0 = "061550234D158C5EC95595FE04EF7A25767F2E24CC2BC479D09D86DC9ABCFDE7056A8C266F9EF97ED08541DBD2E1FFA1"
KAT_SEED1 = "D81C4D8D734FCBFBEADE3D3F8A039FAA2A2C9957E835AD55B22E75BF57BB556AC81ADDE6AEEB4A5A875C3BFCADFA958F"


def test_aes256_fips197_example():
    key = bytes(range(32))
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert kat.aes256_ecb(key, pt).hex() == "8ea2b7ca516745bfeafc49904b496089"


def test_aes256_matches_openssl():
    from openssl_aes import aes_ecb
    ref = aes_ecb()
    if ref is None:
        pytest.skip("libcrypto not found")
    rng = np.random.default_rng(11)
    for _ in range(20):
        key = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        data = rng.integers(0, 256, 16 * 7, dtype=np.uint8).tobytes()
        assert kat.aes256_ecb(key, data) == ref(key, data)


def test_kat_seed_stream_matches_published_seeds():
    s = kat.seeds(2)
    assert bytes(s[0]).hex().upper() == KAT_SEED0
    assert bytes(s[1]).hex().upper() == KAT_SEED1


def test_every_draw_ends_with_an_update():
    seed = bytes(range(48))
    one = kat.Drbg(seed).bytes(96)
    d = kat.Drbg(seed)
    three = d.bytes(32) + d.bytes(32) + d.bytes(32)
    assert one[:32] == three[:32] and one[32:] != three[32:]
    # personalization is XORed into the seed material
    pers = bytes([0xA5] * 48)
    assert kat.Drbg(seed, pers).bytes(32) == kat.Drbg(bytes(a ^ 0xA5 for a in seed)).bytes(32)


@pytest.mark.parametrize("scheme", [D.FRODO976, D.SABER, D.FRODO976_AES, D.LIGHTSABER, D.FIRESABER, D.FRODO640,
                                    D.FRODO1344])
def test_kat_coins_follow_the_schemes_call_pattern(scheme):
    sd = kat.seeds(3)
    kc, ec = kat.coins(scheme, sd)
    # FrodoKEM: one draw of s || seedSE || z (2 sec + 16 bytes), one of mu (sec)
    frodo_sec = {D.FRODO976: 24, D.FRODO976_AES: 24, D.FRODO640: 16, D.FRODO1344: 32}
    for i in range(3):
        d = kat.Drbg(bytes(sd[i]))
        if scheme in frodo_sec:
            sec = frodo_sec[scheme]
            want_k, want_e = d.bytes(2 * sec + 16), d.bytes(sec)
        else:
            want_k = d.bytes(32) + d.bytes(32) + d.bytes(32)
            want_e = d.bytes(32)
        assert bytes(kc[i]) == want_k and bytes(ec[i]) == want_e


@pytest.mark.parametrize("scheme", [D.FRODO976, D.SABER])
def test_rsp_roundtrip_with_the_oracle(oracle, scheme):
    sd = kat.seeds(2)
    ops = (lambda sc, kc: oracle.keypair(sc, kc, threads=2), lambda sc, pk, ec: oracle.enc(sc, pk, ec, threads=2),
           lambda sc, ct, sk: oracle.dec(sc, ct, sk, threads=2))
    pk, sk, ct, ss = kat.compute(scheme, sd, *ops)
    text = kat.format_rsp(scheme, sd, pk, sk, ct, ss)
    recs = kat.parse_rsp(text)
    assert [r["count"] for r in recs] == [0, 1]
    s = D.SIZES[scheme]
    for i, r in enumerate(recs):
        assert r["seed"] == bytes(sd[i])
        assert len(r["pk"]) == s["pk"] and len(r["sk"]) == s["sk"] and len(r["ct"]) == s["ct"]
        assert r["pk"] == bytes(pk[i]) and r["ss"] == bytes(ss[i])
