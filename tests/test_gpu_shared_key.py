"""Many encapsulations to / decapsulations with ONE key (dpkem_*_shared_key).

Every result must equal the ordinary per-instance op with the key repeated,
checked against the oracle: group sizes that leave 1..3 empty member slots,
a tampered ciphertext inside a group (implicit rejection must not leak to
its group neighbours), and a device-resident batch.
"""
import numpy as np
import pytest

import dpcrypto_amd as D

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = D.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("scheme", [D.FRODO976, D.SABER])
@pytest.mark.parametrize("n", [1, 3, 4, 5, 9])
def test_shared_key_equals_oracle(eng, oracle, scheme, n):
    s = D.SIZES[scheme]
    kc = oracle.prng_bytes(700 + scheme, s["kcoins"]).reshape(1, -1)
    pk, sk = oracle.keypair(scheme, kc, threads=1)
    ec = oracle.prng_bytes(800 + n, n * s["ecoins"]).reshape(n, -1)
    oct_, oss = oracle.enc(scheme, np.repeat(pk, n, axis=0), ec, threads=8)
    ct, ss = eng.enc_shared_key(scheme, pk, ec)
    assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
    # tamper one ciphertext in the middle of a group
    bad = ct.copy()
    t = n // 2
    bad[t, 100] ^= 0x01
    want = oracle.dec(scheme, bad, np.repeat(sk, n, axis=0), threads=8)
    got = eng.dec_shared_key(scheme, bad, sk)
    assert np.array_equal(got, want)
    ok = [i for i in range(n) if i != t]
    assert np.array_equal(got[ok], oss[ok]) and not np.array_equal(got[t], oss[t])


@pytest.mark.parametrize("scheme,n", [(D.FRODO976, 70), (D.SABER, 1000)])
def test_shared_key_device_batch_equals_per_instance_path(eng, oracle, scheme, n):
    import torch
    s = D.SIZES[scheme]
    kc = oracle.prng_bytes(900 + scheme, s["kcoins"]).reshape(1, -1)
    pk, sk = oracle.keypair(scheme, kc, threads=1)
    ec = oracle.prng_bytes(901, n * s["ecoins"]).reshape(n, -1)
    dpk, dsk, dec_ = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (pk, sk, ec))
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    ct, ss = eng.enc_shared_key(scheme, dpk, dec_)
    ct2, ss2 = eng.enc(scheme, dpk.repeat(n, 1), dec_)
    ss3 = eng.dec_shared_key(scheme, ct, dsk)
    torch.cuda.synchronize()
    assert torch.equal(ct, ct2) and torch.equal(ss, ss2) and torch.equal(ss3, ss)
    # spot-check two instances against the oracle
    for i in (0, n - 1):
        oc, os_ = oracle.enc(scheme, pk, ec[i:i + 1], threads=1)
        assert np.array_equal(ct[i].cpu().numpy(), oc[0]) and np.array_equal(ss[i].cpu().numpy(), os_[0])
