"""Frodo-640 and Frodo-1344 (SHAKE) on the GPU against the oracle.

Covers keygen / encaps / decaps at ragged batch sizes (Frodo-640 packs 15-bit
entries, Frodo-1344 runs three row passes per workgroup), implicit rejection
with flips in B' and in C, a device-resident batch with a full round trip,
KAT coins, the shared-key entry points and a graph-captured chained plan.
"""
import numpy as np
import pytest

import dpcrypto_amd as D

pytestmark = pytest.mark.gpu
SCHEMES = [D.FRODO640, D.FRODO1344]


@pytest.fixture(scope="module")
def eng():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    e = D.Engine(0)
    yield e
    e.close()


def _coins(oracle, seed, n, width):
    return oracle.prng_bytes(seed, n * width).reshape(n, width)


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("n", [1, 5, 34])
def test_ops_equal_oracle(eng, oracle, scheme, n):
    s = D.SIZES[scheme]
    print(f"frodo set {scheme} n={n}", flush=True)
    kc = _coins(oracle, 9000 + 10 * scheme + n, n, s["kcoins"])
    pk, sk = eng.keypair(scheme, kc)
    opk, osk = oracle.keypair(scheme, kc, threads=8)
    assert pk.shape == (n, s["pk"]) and np.array_equal(pk, opk) and np.array_equal(sk, osk)
    ec = _coins(oracle, 9100 + 10 * scheme + n, n, s["ecoins"])
    ct, ss = eng.enc(scheme, pk, ec)
    oct_, oss = oracle.enc(scheme, opk, ec, threads=8)
    assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
    assert np.array_equal(eng.dec(scheme, ct, sk), oss)
    bad = ct.copy()
    d = 15 if scheme == D.FRODO640 else 16
    c_off = d * (s["ct"] // d - 8)  # first byte of C
    bad[:, 3] ^= 0x01           # in B'
    bad[::2, c_off] ^= 0x40     # first entry of C on every other instance
    bad[1::3, -1] ^= 0x80       # last entry of C
    ss_bad = eng.dec(scheme, bad, sk)
    assert np.array_equal(ss_bad, oracle.dec(scheme, bad, osk, threads=8))
    assert not (ss_bad == oss).all(axis=1).any()


@pytest.mark.parametrize("scheme", SCHEMES)
def test_device_batch_roundtrip(eng, oracle, scheme):
    import torch
    s = D.SIZES[scheme]
    n = 300
    kc = _coins(oracle, 9200 + scheme, n, s["kcoins"])
    ec = _coins(oracle, 9300 + scheme, n, s["ecoins"])
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    pk, sk = eng.keypair(scheme, torch.from_numpy(kc).cuda())
    ct, ss = eng.enc(scheme, pk, torch.from_numpy(ec).cuda())
    ss2 = eng.dec(scheme, ct, sk)
    torch.cuda.synchronize()
    assert torch.equal(ss, ss2)
    idx = [0, 151, n - 1]
    opk, osk = oracle.keypair(scheme, kc[idx], threads=3)
    oct_, oss = oracle.enc(scheme, opk, ec[idx], threads=3)
    assert np.array_equal(pk[idx].cpu().numpy(), opk) and np.array_equal(sk[idx].cpu().numpy(), osk)
    assert np.array_equal(ct[idx].cpu().numpy(), oct_) and np.array_equal(ss[idx].cpu().numpy(), oss)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_kat_and_shared_key(eng, oracle, scheme):
    from dpcrypto_amd import kat
    s = D.SIZES[scheme]
    sd = kat.seeds(3)
    gpu = kat.format_rsp(scheme, sd, *kat.compute(scheme, sd, eng.keypair, eng.enc, eng.dec))
    ops = (lambda sc, kc: oracle.keypair(sc, kc, threads=3), lambda sc, pk, ec: oracle.enc(sc, pk, ec, threads=3),
           lambda sc, ct, sk: oracle.dec(sc, ct, sk, threads=3))
    assert gpu == kat.format_rsp(scheme, sd, *kat.compute(scheme, sd, *ops))
    pk, sk = oracle.keypair(scheme, _coins(oracle, 9400 + scheme, 1, s["kcoins"]))
    ec = _coins(oracle, 9401 + scheme, 6, s["ecoins"])
    ct, ss = eng.enc_shared_key(scheme, pk, ec)
    oct_, oss = oracle.enc(scheme, np.repeat(pk, 6, axis=0), ec, threads=6)
    assert np.array_equal(ct, oct_) and np.array_equal(ss, oss)
    assert np.array_equal(eng.dec_shared_key(scheme, ct, sk), oss)


def test_chained_plan_of_both_sets(eng, oracle):
    import torch
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    plan = eng.plan(nstreams=2, chunks=2)
    bufs = {}
    for k, scheme in enumerate(SCHEMES):
        s = D.SIZES[scheme]
        n = 7 + 4 * k
        kc = _coins(oracle, 9500 + scheme, n, s["kcoins"])
        ec = _coins(oracle, 9600 + scheme, n, s["ecoins"])
        b = dict(kc=kc, ec=ec, dkc=torch.from_numpy(kc).cuda(), dec=torch.from_numpy(ec).cuda())
        for name, w in (("pk", s["pk"]), ("sk", s["sk"]), ("ct", s["ct"]), ("ss", s["ss"]), ("ss2", s["ss"])):
            b[name] = torch.empty((n, w), dtype=torch.uint8, device="cuda")
        t0 = plan.keypair(scheme, b["pk"], b["sk"], b["dkc"])
        t1 = plan.enc(scheme, b["ct"], b["ss"], b["pk"], b["dec"], after=t0)
        plan.dec(scheme, b["ss2"], b["ct"], b["sk"], after=t1)
        bufs[scheme] = b
    plan.build()
    for rep in range(3):
        for b in bufs.values():
            b["ct"].zero_(); b["ss2"].zero_()
        plan.launch()
        torch.cuda.synchronize()
        for scheme, b in bufs.items():
            opk, osk = oracle.keypair(scheme, b["kc"], threads=8)
            oct_, oss = oracle.enc(scheme, opk, b["ec"], threads=8)
            assert np.array_equal(b["pk"].cpu().numpy(), opk), (scheme, rep)
            assert np.array_equal(b["ct"].cpu().numpy(), oct_), (scheme, rep)
            assert np.array_equal(b["ss2"].cpu().numpy(), oss), (scheme, rep)
    plan.close()
