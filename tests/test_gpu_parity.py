"""GPU parity: the HIP engine against the CPU oracle, byte for byte.

Every case calls the C ABI of dpcrypto_amd/libdpkem.so (through the ctypes
binding) on seeded coins and compares pk / sk / ct / ss with the oracle on
the same coins.  Sizes are chosen so the oracle finishes in seconds; the
full bench sizes are covered by size-independent properties (decaps of
encaps recovers the shared secret; tampering triggers implicit rejection).
"""
import hashlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRODO, SABER = 0, 1


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def eng(torch_cuda):
    import dpcrypto_amd as D
    e = D.Engine(0)
    yield e
    e.close()


def _coins(oracle, seed, n, width):
    return oracle.prng_bytes(seed, n * width).reshape(n, width)


def _eq(a, b, what):
    a = np.asarray(a); b = np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.nonzero((a != b).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} instances differ, first {bad[:8].tolist()}"


# ------------------------------------------------------------------ Saber
@pytest.mark.parametrize("n", [1, 63, 130])
def test_saber_keygen_matches_oracle(eng, oracle, n):
    c = _coins(oracle, 0x44504B454D000001 + n, n, 96)
    pk, sk = eng.keypair(SABER, c)
    opk, osk = oracle.keypair(SABER, c)
    _eq(pk, opk, "pk")
    _eq(sk, osk, "sk")


@pytest.mark.parametrize("n", [1, 65, 200])
def test_saber_enc_dec_match_oracle(eng, oracle, n):
    c = _coins(oracle, 77 + n, n, 96)
    opk, osk = oracle.keypair(SABER, c)
    m = _coins(oracle, 78 + n, n, 32)
    ct, ss = eng.enc(SABER, opk, m)
    oct_, oss = oracle.enc(SABER, opk, m)
    _eq(ct, oct_, "ct")
    _eq(ss, oss, "ss")
    ss2 = eng.dec(SABER, oct_, osk)
    _eq(ss2, oss, "dec ss")


def test_saber_implicit_rejection(eng, oracle):
    n = 70
    c = _coins(oracle, 90, n, 96)
    pk, sk = oracle.keypair(SABER, c)
    ct, ss = oracle.enc(SABER, pk, _coins(oracle, 91, n, 32))
    bad = ct.copy()
    bad[np.arange(n), (np.arange(n) * 37) % 1088] ^= 1  # one flipped bit per instance
    got = eng.dec(SABER, bad, sk)
    _eq(got, oracle.dec(SABER, bad, sk), "rejected ss")
    for i in range(0, n, 17):
        z = sk[i, 2272:].tobytes()
        assert got[i].tobytes() == hashlib.sha3_256(z + hashlib.sha3_256(bad[i].tobytes()).digest()).digest()


def test_saber_tv_fixture(eng, oracle):
    """The reference's own test vector (SABER-GPU/include/tv.h) through the GPU."""
    import json, os
    tv = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "saber_tv.json")))
    pk = np.frombuffer(bytes.fromhex(tv["pk"]), np.uint8)[None]
    sk = np.frombuffer(bytes.fromhex(tv["sk_tv"]), np.uint8)[None]
    m_raw = np.frombuffer(bytes.fromhex(tv["buf_tv"])[:32], np.uint8)[None]
    ct, ss = eng.enc(SABER, pk, m_raw)
    assert hashlib.sha3_256(ct.tobytes()).hexdigest().startswith("dd05a10fcf07f1b2")
    assert ss.tobytes().hex().startswith("c842a1e43b3a3fe8")
    assert (eng.dec(SABER, ct, sk) == ss).all()


def test_saber_device_memory_and_full_batch(eng, oracle, torch_cuda):
    """Device-resident arrays at the 4096-instance bench size: round trip holds,
    and a seeded sample of instances matches the oracle."""
    torch = torch_cuda
    n = 4096
    c = torch.from_numpy(_coins(oracle, 0x44504B454D000001, n, 96)).cuda()
    m = torch.from_numpy(_coins(oracle, 0x44504B454D000011, n, 32)).cuda()
    pk, sk = eng.keypair(SABER, c)
    ct, ss = eng.enc(SABER, pk, m)
    ss2 = eng.dec(SABER, ct, sk)
    eng.sync()
    assert torch.equal(ss, ss2)
    idx = np.arange(0, n, 97)
    opk, osk = oracle.keypair(SABER, c.cpu().numpy()[idx])
    _eq(pk.cpu().numpy()[idx], opk, "pk sample")
    _eq(sk.cpu().numpy()[idx], osk, "sk sample")
    oct_, oss = oracle.enc(SABER, opk, m.cpu().numpy()[idx])
    _eq(ct.cpu().numpy()[idx], oct_, "ct sample")
    _eq(ss.cpu().numpy()[idx], oss, "ss sample")


# ------------------------------------------------------------------ Frodo
@pytest.mark.parametrize("n", [1, 5])
def test_frodo_keygen_matches_oracle(eng, oracle, n):
    c = _coins(oracle, 0x44504B454D000000 + n, n, 64)
    pk, sk = eng.keypair(FRODO, c)
    opk, osk = oracle.keypair(FRODO, c, threads=4)
    _eq(pk, opk, "pk")
    _eq(sk, osk, "sk")


@pytest.mark.parametrize("n", [1, 6])
def test_frodo_enc_dec_match_oracle(eng, oracle, n):
    c = _coins(oracle, 500 + n, n, 64)
    opk, osk = oracle.keypair(FRODO, c, threads=4)
    mu = _coins(oracle, 600 + n, n, 24)
    ct, ss = eng.enc(FRODO, opk, mu)
    oct_, oss = oracle.enc(FRODO, opk, mu, threads=4)
    _eq(ct, oct_, "ct")
    _eq(ss, oss, "ss")
    _eq(eng.dec(FRODO, oct_, osk), oss, "dec ss")


def test_frodo_implicit_rejection(eng, oracle):
    n = 4
    c = _coins(oracle, 700, n, 64)
    pk, sk = oracle.keypair(FRODO, c, threads=4)
    ct, ss = oracle.enc(FRODO, pk, _coins(oracle, 701, n, 24), threads=4)
    bad = ct.copy()
    bad[0, 3] ^= 0x10      # B' entry
    bad[1, 15700] ^= 0x01  # C entry
    bad[2, 15743] ^= 0x80  # last byte of C
    got = eng.dec(FRODO, bad, sk)
    _eq(got, oracle.dec(FRODO, bad, sk, threads=4), "rejected ss")
    assert (got[3] == ss[3]).all()  # untouched instance decapsulates normally
    for i in range(3):
        exp = hashlib.shake_256(bad[i].tobytes() + sk[i, :24].tobytes()).digest(24)
        assert got[i].tobytes() == exp


def test_frodo_device_roundtrip_batch(eng, oracle, torch_cuda):
    """A 256-instance device-resident batch: decaps(encaps) recovers every shared
    secret and the keygen of a sample of instances equals the oracle's."""
    torch = torch_cuda
    n = 256
    c = torch.from_numpy(_coins(oracle, 0x44504B454D000002, n, 64)).cuda()
    mu = torch.from_numpy(_coins(oracle, 0x44504B454D000012, n, 24)).cuda()
    pk, sk = eng.keypair(FRODO, c)
    ct, ss = eng.enc(FRODO, pk, mu)
    ss2 = eng.dec(FRODO, ct, sk)
    eng.sync()
    assert torch.equal(ss, ss2)
    idx = np.array([0, 131, 255])
    opk, osk = oracle.keypair(FRODO, c.cpu().numpy()[idx], threads=3)
    _eq(pk.cpu().numpy()[idx], opk, "pk sample")
    oct_, oss = oracle.enc(FRODO, opk, mu.cpu().numpy()[idx], threads=3)
    _eq(ct.cpu().numpy()[idx], oct_, "ct sample")


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_shaped_drivers(eng, oracle, mode):
    """frodoKEMKeypair/Encap/Decap and saber_enc/saber_dec with the reference's
    signatures (768-instance host arrays) and a seeded randombytes callback:
    every Saber instance, and a sample of Frodo instances, equal the oracle
    run on the same coins.  `mode` is saber_enc/saber_dec's DP / INT32 switch
    (0 = batched form, 1 = plain form; SABER-GPU/main.cu:25-33)."""
    import ctypes
    import dpcrypto_amd as D
    lib = D.load_library()
    state = {"seed": 12345 + 100 * mode}
    drawn = []

    @ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)
    def rb(out, length, ctx):
        buf = oracle.prng_bytes(state["seed"], length)
        state["seed"] += 1
        drawn.append(buf)
        ctypes.memmove(out, buf.ctypes.data, length)

    lib.dpkem_set_randombytes.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.dpkem_set_randombytes(ctypes.cast(rb, ctypes.c_void_p), None)
    try:
        N = 768
        # the reference's drivers take keys as inputs: fill them with the keypair wrapper first
        pk = np.zeros((N, 992), np.uint8); sk = np.zeros((N, 2304), np.uint8)
        lib.dpkem_saber_crypto_kem_keypair.argtypes = [ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
        assert lib.dpkem_saber_crypto_kem_keypair(N, pk.ctypes.data, sk.ctypes.data) == 0
        k = np.zeros((N, 32), np.uint8); c = np.zeros((N, 1088), np.uint8); k2 = np.zeros((N, 32), np.uint8)
        lib.saber_enc.argtypes = [ctypes.c_uint8] + [ctypes.c_void_p] * 3
        lib.saber_dec.argtypes = [ctypes.c_uint8] + [ctypes.c_void_p] * 4
        assert lib.saber_enc(mode, k.ctypes.data, pk.ctypes.data, c.ctypes.data) == 0
        assert lib.saber_dec(mode, c.ctypes.data, pk.ctypes.data, sk.ctypes.data, k2.ctypes.data) == 0
        assert len(drawn) == 2
        opk, osk = oracle.keypair(SABER, drawn[0].reshape(N, 96), threads=8)
        _eq(pk, opk, "saber pk"); _eq(sk, osk, "saber sk")
        oct_, oss = oracle.enc(SABER, opk, drawn[1].reshape(N, 32), threads=8)
        _eq(c, oct_, "saber_enc ct"); _eq(k, oss, "saber_enc k"); _eq(k2, oss, "saber_dec k")
        if mode == 1:
            return  # the Frodo drivers have no mode switch: checked once
        drawn.clear()
        fpk = np.zeros((N, 15632), np.uint8); fsk = np.zeros((N, 31296), np.uint8)
        fct = np.zeros((N, 15744), np.uint8); fss = np.zeros((N, 24), np.uint8); fss2 = np.zeros((N, 24), np.uint8)
        lib.frodoKEMKeypair.argtypes = [ctypes.c_void_p] * 2
        lib.frodoKEMEncap.argtypes = [ctypes.c_void_p] * 3
        lib.frodoKEMDecap.argtypes = [ctypes.c_void_p] * 3
        assert lib.frodoKEMKeypair(fpk.ctypes.data, fsk.ctypes.data) == 0
        assert lib.frodoKEMEncap(fct.ctypes.data, fss.ctypes.data, fpk.ctypes.data) == 0
        assert lib.frodoKEMDecap(fss2.ctypes.data, fct.ctypes.data, fsk.ctypes.data) == 0
        assert (fss == fss2).all()
        assert len(drawn) == 2
        idx = np.array([0, 1, 100, 383, 384, 511, 700, 767])
        opk, osk = oracle.keypair(FRODO, drawn[0].reshape(N, 64)[idx], threads=8)
        _eq(fpk[idx], opk, "frodo pk"); _eq(fsk[idx], osk, "frodo sk")
        oct_, oss = oracle.enc(FRODO, opk, drawn[1].reshape(N, 24)[idx], threads=8)
        _eq(fct[idx], oct_, "frodoKEMEncap ct"); _eq(fss[idx], oss, "frodoKEMEncap ss")
    finally:
        lib.dpkem_set_randombytes(None, None)


# ------------------------------------------------------------------ plain comparator form
@pytest.fixture(scope="module")
def plain_eng(torch_cuda):
    import dpcrypto_amd as D
    e = D.Engine(0, form=D.FORM_PLAIN)
    yield e
    e.close()


@pytest.mark.parametrize("scheme,n", [(SABER, 70), (FRODO, 3)])
def test_plain_form_matches_oracle(plain_eng, oracle, scheme, n):
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    c = _coins(oracle, 800 + scheme, n, s["kcoins"])
    m = _coins(oracle, 900 + scheme, n, s["ecoins"])
    pk, sk = plain_eng.keypair(scheme, c)
    opk, osk = oracle.keypair(scheme, c, threads=4)
    _eq(pk, opk, "pk")
    _eq(sk, osk, "sk")
    ct, ss = plain_eng.enc(scheme, opk, m)
    oct_, oss = oracle.enc(scheme, opk, m, threads=4)
    _eq(ct, oct_, "ct")
    _eq(ss, oss, "ss")
    _eq(plain_eng.dec(scheme, oct_, osk), oss, "dec ss")
    bad = oct_.copy(); bad[:, 7] ^= 2
    _eq(plain_eng.dec(scheme, bad, osk), oracle.dec(scheme, bad, osk, threads=4), "rejected ss")


def test_perm_latency_probe(eng):
    """dpkem_probe_perm_latency (the serial chains' measured latency floor in
    the bench's roofline): a longer run of back-to-back permutations takes
    longer, the per-permutation latency is microseconds, and out-of-range
    counts are rejected."""
    import dpcrypto_amd as D
    print("probe: 64 and 576 permutations on one wave", flush=True)
    t_short = eng.probe_perm_latency(64)
    t_long = eng.probe_perm_latency(576)
    assert 0.0 < t_short < t_long
    per_us = (t_long - t_short) * 1e3 / 512
    assert 1.0 < per_us < 50.0, per_us
    for bad in (0, 100001):
        with pytest.raises(D.DpkemError):
            eng.probe_perm_latency(bad)
