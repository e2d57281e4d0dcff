"""The rolling keygen -> encaps -> decaps pipeline (dpkem_roll_*) against the
oracle: every output of every batch of every scheme (Saber family, FrodoKEM
sets), at ragged sizes, graph-replayed and issued directly, and
decapsulations of tampered ciphertexts (implicit rejection) inside the
pipeline."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FRODO976, SABER, FRODO976_AES, LIGHTSABER, FIRESABER, FRODO640, FRODO1344 = 0, 1, 2, 3, 4, 5, 6


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
    e.set_stream(torch_cuda.cuda.current_stream().cuda_stream)
    yield e
    e.set_tuning(D.TUNE_SABER_POLY_IPW, 1)
    e.set_tuning(D.TUNE_FRODO_THREADS, 0)
    e.close()


def _slots(torch, oracle, scheme, n, seed, count=3):
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    slots, ref = [], []
    for k in range(count):
        kc = oracle.prng_bytes(seed + 10 * k, n * s["kcoins"]).reshape(n, -1)
        ec = oracle.prng_bytes(seed + 10 * k + 1, n * s["ecoins"]).reshape(n, -1)
        opk, osk = oracle.keypair(scheme, kc, threads=8)
        oct_, oss = oracle.enc(scheme, opk, ec, threads=8)
        ref.append((opk, osk, oct_, oss))
        sl = {"kc": torch.from_numpy(kc).cuda(), "ec": torch.from_numpy(ec).cuda()}
        for name, w in (("pk", s["pk"]), ("sk", s["sk"]), ("ct", s["ct"]), ("ss", s["ss"]), ("ss2", s["ss"])):
            sl[name] = torch.zeros((n, w), dtype=torch.uint8, device="cuda")
        slots.append(sl)
    return slots, ref


# shape: the Saber product's instances per workgroup, or the Frodo fused
# kernels' workgroup size (0 = the set's default, 1024 = FrodoKEM-976's other)
@pytest.mark.parametrize("scheme,n,shape,graph", [
    (SABER, 300, 1, True), (SABER, 37, 4, True), (SABER, 1, 2, True), (SABER, 65, 1, False),
    (LIGHTSABER, 130, 4, True), (LIGHTSABER, 33, 1, False),
    (FIRESABER, 97, 1, True), (FIRESABER, 5, 4, False),
    (FRODO976, 300, 0, True), (FRODO976, 37, 1024, True), (FRODO976, 1, 0, False),
    (FRODO976_AES, 33, 0, True), (FRODO640, 40, 0, True), (FRODO640, 3, 0, False),
    (FRODO1344, 29, 0, True),
])
def test_roll_every_output_of_every_batch(eng, oracle, torch_cuda, scheme, n, shape, graph):
    import bench
    torch = torch_cuda
    print(f"roll: scheme {scheme}, n {n}, shape {shape}, graph {graph}", flush=True)
    slots, ref = _slots(torch, oracle, scheme, n, 5000 + n)
    frodo = scheme in (FRODO976, FRODO976_AES, FRODO640, FRODO1344)
    r = bench.build_roll(eng, scheme, slots, 1 if frodo else shape, graph=graph, threads=shape if frodo else 0)
    try:
        for step in range(7):
            for t in ("pk", "sk", "ct", "ss", "ss2"):
                slots[step % 3][t].zero_()
            r.step()
            torch.cuda.synchronize()
            if step >= 2:
                c = (step - 2) % 3
                for t, v in zip(("pk", "sk", "ct", "ss", "ss2"), ref[c] + (ref[c][3],)):
                    np.testing.assert_array_equal(slots[c][t].cpu().numpy(), v,
                                                  err_msg=f"{t} of slot {c} after step {step}")
        assert r.info()["steps"] == 7
    finally:
        r.close()


def _rejection_key(scheme, sk_row, ct_row):
    """The implicit-rejection shared secret of a decapsulation of ct_row:
    Saber family SHA3-256(z || SHA3-256(c)) with z the last 32 bytes of sk;
    FrodoKEM SHAKE(c || s) with s the first ss-length bytes of sk (SHAKE128
    for Frodo-640, SHAKE256 otherwise)."""
    if scheme in (SABER, LIGHTSABER, FIRESABER):
        return hashlib.sha3_256(sk_row[-32:] + hashlib.sha3_256(ct_row).digest()).digest()
    ln = {FRODO640: 16, FRODO1344: 32}.get(scheme, 24)
    h = hashlib.shake_128 if scheme == FRODO640 else hashlib.shake_256
    return h(ct_row + sk_row[:ln]).digest(ln)


@pytest.mark.parametrize("scheme,n,ipw", [(SABER, 300, 1), (SABER, 37, 4), (SABER, 1, 2), (LIGHTSABER, 65, 2),
                                           (FIRESABER, 33, 1), (FIRESABER, 97, 4)])
def test_roll_overlapped_every_output_of_every_batch(eng, oracle, torch_cuda, scheme, n, ipw):
    """The six-slot overlapped form (each step's chains beside the previous
    step's products in one launch): a slot's outputs are cleared
    before the step of its keygen, and after step k >= 5 every output of
    slot (k - 5) % 6 equals the oracle."""
    import bench
    torch = torch_cuda
    print(f"roll overlapped: scheme {scheme}, n {n}, ipw {ipw}", flush=True)
    slots, ref = _slots(torch, oracle, scheme, n, 6000 + n, count=6)
    r = bench.build_roll(eng, scheme, slots, ipw, overlap=True)
    try:
        for step in range(13):
            for t in ("pk", "sk", "ct", "ss", "ss2"):
                slots[step % 6][t].zero_()
            r.step()
            torch.cuda.synchronize()
            if step >= 5:
                c = (step - 5) % 6
                for t, v in zip(("pk", "sk", "ct", "ss", "ss2"), ref[c] + (ref[c][3],)):
                    np.testing.assert_array_equal(slots[c][t].cpu().numpy(), v,
                                                  err_msg=f"{t} of slot {c} after step {step}")
        assert r.info()["steps"] == 13
    finally:
        r.close()


@pytest.mark.parametrize("scheme", [SABER, LIGHTSABER, FIRESABER, FRODO976, FRODO640])
def test_roll_decapsulates_tampered_ciphertexts_by_implicit_rejection(eng, oracle, torch_cuda, scheme):
    """Between the step that encapsulates slot 0 and the step that
    decapsulates it, every third ciphertext of the slot gets one bit flipped:
    those decapsulations must return the implicit-rejection key of the
    flipped c, the others the oracle's shared secret."""
    import bench
    torch = torch_cuda
    n = 70
    slots, ref = _slots(torch, oracle, scheme, n, 7000 + scheme)
    r = bench.build_roll(eng, scheme, slots, 1)
    try:
        r.step()  # keygen of slot 0
        r.step()  # encaps of slot 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(slots[0]["ct"].cpu().numpy(), ref[0][2])
        slots[0]["ct"][::3, 11] ^= 0x40
        r.step()  # decaps of slot 0
        torch.cuda.synchronize()
        ct = slots[0]["ct"].cpu().numpy()
        sk = ref[0][1]
        got = slots[0]["ss2"].cpu().numpy()
        for i in range(n):
            if i % 3 == 0:
                want = _rejection_key(scheme, sk[i].tobytes(), ct[i].tobytes())
                assert got[i].tobytes() == want, f"instance {i}: not the implicit-rejection key"
            else:
                assert np.array_equal(got[i], ref[0][3][i]), f"instance {i}: decaps differs from the oracle"
    finally:
        r.close()


def test_roll_rejects_misuse(eng, torch_cuda):
    import dpcrypto_amd as D
    from dpcrypto_amd.kem import DpkemError
    with pytest.raises(DpkemError):
        eng.roll(7, 4)  # no such scheme
    with pytest.raises(DpkemError):
        eng.roll(D.FRODO976, 0)  # no instances
    r = eng.roll(D.SABER, 4)
    with pytest.raises(DpkemError):
        r.build()  # slots not bound
    def slot(scheme, n):
        s = D.SIZES[scheme]
        return [torch_cuda.zeros((n, s[k]), dtype=torch_cuda.uint8, device="cuda")
                for k in ("kcoins", "ecoins", "pk", "sk", "ct", "ss", "ss")]

    for k in range(3):
        r.bind(k, *slot(D.SABER, 4))
    with pytest.raises(DpkemError):
        r.build(overlap=True)  # the overlapped form needs six slots
    r.close()
    f = eng.roll(D.FRODO976, 1)
    for k in range(6):
        f.bind(k, *slot(D.FRODO976, 1))
    with pytest.raises(DpkemError):
        f.build(overlap=True)  # Saber family only
    f.close()
