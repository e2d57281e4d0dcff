"""FrodoKEM-976-AES on the GPU (A rows from AES-128 in LDS T-tables) against
the oracle: keygen, encaps, decaps, implicit rejection, a device batch,
KAT coins and the shared-key entry points."""
import numpy as np
import pytest

import dpcrypto_amd as D

pytestmark = pytest.mark.gpu
AES = D.FRODO976_AES


@pytest.fixture(scope="module")
def eng():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = D.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("n", [1, 3])
def test_frodo_aes_ops_equal_oracle(eng, oracle, n):
    kc = oracle.prng_bytes(7100 + n, n * 64).reshape(n, 64)
    pk, sk = eng.keypair(AES, kc)
    opk, osk = oracle.keypair(AES, kc, threads=4)
    assert np.array_equal(pk, opk) and np.array_equal(sk, osk)
    ec = oracle.prng_bytes(7200 + n, n * 24).reshape(n, 24)
    ct, ss = eng.enc(AES, pk, ec)
    oct_, oss = oracle.enc(AES, opk, ec, threads=4)
    assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
    assert np.array_equal(eng.dec(AES, ct, sk), oss)
    bad = ct.copy(); bad[:, 7] ^= 0x02
    assert np.array_equal(eng.dec(AES, bad, sk), oracle.dec(AES, bad, osk, threads=4))
    # the SHAKE variant with the same coins differs (A differs)
    pk0, _ = eng.keypair(D.FRODO976, kc[:1])
    assert np.array_equal(pk0[0, :16], pk[0, :16]) and not np.array_equal(pk0[0], pk[0])


def test_frodo_aes_device_batch(eng, oracle):
    import torch
    n = 64
    kc = oracle.prng_bytes(7300, n * 64).reshape(n, 64)
    ec = oracle.prng_bytes(7301, n * 24).reshape(n, 24)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    pk, sk = eng.keypair(AES, torch.from_numpy(kc).cuda())
    ct, ss = eng.enc(AES, pk, torch.from_numpy(ec).cuda())
    ss2 = eng.dec(AES, ct, sk)
    torch.cuda.synchronize()
    assert torch.equal(ss, ss2)
    for i in (0, n - 1):
        opk, _ = oracle.keypair(AES, kc[i:i + 1])
        oct_, oss = oracle.enc(AES, opk, ec[i:i + 1])
        assert np.array_equal(pk[i].cpu().numpy(), opk[0])
        assert np.array_equal(ct[i].cpu().numpy(), oct_[0]) and np.array_equal(ss[i].cpu().numpy(), oss[0])


def test_frodo_aes_kat_and_shared_key(eng, oracle):
    from dpcrypto_amd import kat
    sd = kat.seeds(3)
    gpu = kat.format_rsp(AES, sd, *kat.compute(AES, sd, eng.keypair, eng.enc, eng.dec))
    ops = (lambda sc, kc: oracle.keypair(sc, kc, threads=3), lambda sc, pk, ec: oracle.enc(sc, pk, ec, threads=3),
           lambda sc, ct, sk: oracle.dec(sc, ct, sk, threads=3))
    assert gpu == kat.format_rsp(AES, sd, *kat.compute(AES, sd, *ops))
    kc = oracle.prng_bytes(7400, 64).reshape(1, 64)
    pk, sk = oracle.keypair(AES, kc)
    ec = oracle.prng_bytes(7401, 5 * 24).reshape(5, 24)
    ct, ss = eng.enc_shared_key(AES, pk, ec)
    oct_, oss = oracle.enc(AES, np.repeat(pk, 5, axis=0), ec, threads=5)
    assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
    assert np.array_equal(eng.dec_shared_key(AES, ct, sk), oss)


@pytest.mark.parametrize("scheme", [0, 2])
def test_frodo976_fused_shape_knob_keeps_every_byte(oracle, scheme):
    """DPKEM_TUNE_FRODO_THREADS = 1024 (one row per lane, one workgroup per
    CU, beside the default 512 x 2) changes the fused kernels' launch shape only: keygen / encaps /
    decaps of FrodoKEM-976-SHAKE and -AES equal the oracle, a tampered
    ciphertext included, and the knob rejects other values."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dpcrypto_amd as D
    n = 5
    e = D.Engine(0)
    try:
        with pytest.raises(D.DpkemError):
            e.set_tuning(D.TUNE_FRODO_THREADS, 256)
        e.set_tuning(D.TUNE_FRODO_THREADS, 1024)
        kc = oracle.prng_bytes(8800 + scheme, n * 64).reshape(n, 64)
        mu = oracle.prng_bytes(8900 + scheme, n * 24).reshape(n, 24)
        pk, sk = e.keypair(scheme, kc)
        opk, osk = oracle.keypair(scheme, kc, threads=4)
        assert np.array_equal(pk, opk) and np.array_equal(sk, osk)
        ct, ss = e.enc(scheme, pk, mu)
        oct_, oss = oracle.enc(scheme, opk, mu, threads=4)
        assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
        bad = oct_.copy()
        bad[2, 100] ^= 1
        assert np.array_equal(e.dec(scheme, bad, osk), oracle.dec(scheme, bad, osk, threads=4))
        assert np.array_equal(e.dec(scheme, oct_, osk), oss)
    finally:
        e.close()
