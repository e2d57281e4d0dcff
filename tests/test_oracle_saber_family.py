"""LightSaber (l=2) and FireSaber (l=4) in the oracle, beside Saber (l=3).

Parameter sets (l, mu, eT) = (2, 10, 3), (3, 8, 4), (4, 6, 6); sizes, CBD
and the eT-bit message packing follow the published round-3 scheme
(SABER-GPU/include/params.h:10-19, SABER-GPU/src/poly.cu:817-891).  Pinned
by: hashlib SHAKE128 / SHA3 streams with numpy restatements of the CBD
(bit-level, independent of the oracle's loop), of the keygen relation
pk.b = pack10((A^T s + h1) >> 3), of encaps (b' and c_m with eT bits) and of
decaps' implicit rejection; the l = 3 case of the same code is pinned by the
reference's tv.h in test_oracle.py.  No LightSaber / FireSaber vectors exist
in the reference, so end-to-end bytes are pinned by these restatements only.
"""
import hashlib

import numpy as np
import pytest

SCHEME = {2: 3, 3: 1, 4: 4}  # oracle scheme id by rank l
PARAMS = {2: (10, 3), 3: (8, 4), 4: (6, 6)}


def _bits(data: bytes) -> np.ndarray:
    return np.unpackbits(np.frombuffer(data, np.uint8), bitorder="little")


def _unpack(data: bytes, count: int, width: int) -> np.ndarray:
    b = _bits(data)[: count * width].reshape(count, width).astype(np.int64)
    return (b << np.arange(width)).sum(1)


def _pack(v, width: int) -> bytes:
    v = np.asarray(v, np.int64) & ((1 << width) - 1)
    bits = ((v[:, None] >> np.arange(width)) & 1).astype(np.uint8).ravel()
    return np.packbits(bits, bitorder="little").tobytes()


def _cbd(seed: bytes, l: int) -> np.ndarray:
    mu = PARAMS[l][0]
    stream = hashlib.shake_128(seed).digest(l * 256 * mu // 8)
    v = _unpack(stream, l * 256, mu)
    half = mu // 2
    lo = np.array([bin(x & ((1 << half) - 1)).count("1") for x in v])
    hi = np.array([bin(x >> half).count("1") for x in v])
    return lo - hi


def _negacyclic(a, b):
    c = np.convolve(a, b)
    out = c[:256].copy()
    out[: len(c) - 256] -= c[256:]
    return out


def _matvec(A, s, l, transpose):
    out = np.zeros((l, 256), np.int64)
    for i in range(l):
        for j in range(l):
            out[i] += _negacyclic(A[j][i] if transpose else A[i][j], s[j])
    return out


def _sizes(l):
    et = PARAMS[l][1]
    pk = 320 * l + 32
    return pk, 416 * l + pk + 64, 320 * l + 32 * et


@pytest.mark.parametrize("l", [2, 3, 4])
def test_gen_secret_cbd_vs_bit_level_restatement(oracle, l):
    seed = bytes(range(7, 39))
    got = oracle.saber_l_gen_secret(l, seed).astype(np.int16).astype(np.int64)
    assert np.array_equal(got, _cbd(seed, l))
    mu = PARAMS[l][0]
    assert got.min() >= -mu // 2 and got.max() <= mu // 2


@pytest.mark.parametrize("l", [2, 4])
def test_gen_matrix_vs_hashlib(oracle, l):
    seed = bytes(range(50, 82))
    got = oracle.saber_l_gen_matrix(l, seed)
    exp = _unpack(hashlib.shake_128(seed).digest(l * l * 416), l * l * 256, 13)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("l", [2, 4])
def test_keygen_and_encaps_vs_numpy_restatement(oracle, l):
    sc = SCHEME[l]
    mu, et = PARAMS[l]
    pkn, skn, ctn = _sizes(l)
    coins = oracle.prng_bytes(60 + l, 96)
    pk, sk = oracle.keypair(sc, coins[None])
    pk, sk = pk[0].tobytes(), sk[0].tobytes()
    assert len(pk) == pkn and len(sk) == skn
    seed_a = hashlib.shake_128(coins[:32].tobytes()).digest(32)
    A = _unpack(hashlib.shake_128(seed_a).digest(l * l * 416), l * l * 256, 13).reshape(l, l, 256)
    s = _cbd(coins[32:64].tobytes(), l).reshape(l, 256)
    b = (_matvec(A, s, l, transpose=True) + 4) % 8192 >> 3
    assert pk == _pack(b.ravel(), 10) + seed_a
    assert sk == _pack(s.ravel() % 8192, 13) + pk + hashlib.sha3_256(pk).digest() + coins[64:].tobytes()
    # encaps
    m_raw = oracle.prng_bytes(70 + l, 32)
    ct, ss = oracle.enc(sc, np.frombuffer(pk, np.uint8)[None], m_raw[None])
    m = hashlib.sha3_256(m_raw.tobytes()).digest()
    kr = hashlib.sha3_512(m + hashlib.sha3_256(pk).digest()).digest()
    sp = _cbd(kr[32:], l).reshape(l, 256)
    bp = (_matvec(A, sp, l, transpose=False) + 4) % 8192 >> 3
    bb = _unpack(pk, l * 256, 10).reshape(l, 256)
    vp = sum(_negacyclic(bb[j], sp[j]) for j in range(l))
    mbits = _bits(m).astype(np.int64)
    cm = ((vp - (mbits << 9) + 4) % 65536) >> (10 - et)
    exp_ct = _pack(bp.ravel(), 10) + _pack(cm, et)
    assert len(exp_ct) == ctn and ct[0].tobytes() == exp_ct
    assert ss[0].tobytes() == hashlib.sha3_256(kr[:32] + hashlib.sha3_256(exp_ct).digest()).digest()


@pytest.mark.parametrize("l", [2, 4])
def test_roundtrip_and_implicit_rejection(oracle, l):
    sc = SCHEME[l]
    n = 4
    coins = oracle.prng_bytes(80 + l, n * 96).reshape(n, 96)
    pk, sk = oracle.keypair(sc, coins, threads=2)
    m = oracle.prng_bytes(90 + l, n * 32).reshape(n, 32)
    ct, ss = oracle.enc(sc, pk, m, threads=2)
    assert np.array_equal(oracle.dec(sc, ct, sk, threads=2), ss)
    bad = ct.copy()
    bad[:, -1] ^= 0x10  # in the c_m part
    ss_bad = oracle.dec(sc, bad, sk, threads=2)
    for i in range(n):
        z = sk[i, -32:].tobytes()
        assert ss_bad[i].tobytes() == hashlib.sha3_256(z + hashlib.sha3_256(bad[i].tobytes()).digest()).digest()
