"""The rolling keygen -> encaps -> decaps pipeline (dpkem_roll_*) against the
oracle: every output of every batch of every Saber-family rank, at ragged
sizes, graph-replayed and issued directly, and decapsulations of tampered
ciphertexts (implicit rejection) inside the pipeline."""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SABER, LIGHTSABER, FIRESABER = 1, 3, 4


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
    e.close()


def _slots(torch, oracle, scheme, n, seed):
    import dpcrypto_amd as D
    s = D.SIZES[scheme]
    slots, ref = [], []
    for k in range(3):
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


@pytest.mark.parametrize("scheme,n,ipw,graph", [
    (SABER, 300, 1, True), (SABER, 37, 4, True), (SABER, 1, 2, True), (SABER, 65, 1, False),
    (LIGHTSABER, 130, 4, True), (LIGHTSABER, 33, 1, False),
    (FIRESABER, 97, 1, True), (FIRESABER, 5, 4, False),
])
def test_roll_every_output_of_every_batch(eng, oracle, torch_cuda, scheme, n, ipw, graph):
    import bench
    torch = torch_cuda
    print(f"roll: scheme {scheme}, n {n}, ipw {ipw}, graph {graph}", flush=True)
    slots, ref = _slots(torch, oracle, scheme, n, 5000 + n)
    r = bench.build_roll(eng, scheme, slots, ipw, graph=graph)
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


@pytest.mark.parametrize("scheme", [SABER, LIGHTSABER, FIRESABER])
def test_roll_decapsulates_tampered_ciphertexts_by_implicit_rejection(eng, oracle, torch_cuda, scheme):
    """Between the step that encapsulates slot 0 and the step that
    decapsulates it, every third ciphertext of the slot gets one bit flipped:
    those decapsulations must return SHA3-256(z || SHA3-256(c)) with the
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
                want = hashlib.sha3_256(sk[i].tobytes()[-32:] + hashlib.sha3_256(ct[i].tobytes()).digest()).digest()
                assert got[i].tobytes() == want, f"instance {i}: not the implicit-rejection key"
            else:
                assert np.array_equal(got[i], ref[0][3][i]), f"instance {i}: decaps differs from the oracle"
    finally:
        r.close()


def test_roll_rejects_misuse(eng, torch_cuda):
    import dpcrypto_amd as D
    from dpcrypto_amd.kem import DpkemError
    with pytest.raises(DpkemError):
        eng.roll(D.FRODO976, 4)  # Saber family only
    r = eng.roll(D.SABER, 4)
    with pytest.raises(DpkemError):
        r.build()  # slots not bound
    r.close()
