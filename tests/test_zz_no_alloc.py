"""dpkem_run_noalloc: the allocation-free variant of dpkem_keypair / enc / dec.

The entry point takes its kernel scratch from a static device array instead
of the engine's arena.  On the GPU its outputs must equal the engine's
DPKEM_MEM_DEVICE entry points byte for byte, on two calls in a row (the
second smaller, so the first call's scratch contents are left behind), for
every op of Saber and FrodoKEM-976; n beyond the static array is rejected
before anything runs.
"""
import ctypes

import numpy as np
import pytest

import dpcrypto_amd as D

pytestmark = pytest.mark.gpu

EBUDGET, EINVAL = -4, -1
# (DPKEM_NOALLOC_SCRATCH_BYTES - 4096) / per-instance scratch, per scheme
MAX_N = {D.FRODO976: 4160, D.SABER: 13796, D.FRODO976_AES: 4160, D.LIGHTSABER: 26212, D.FIRESABER: 8737,
         D.FRODO640: 6241, D.FRODO1344: 3048}


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _lib():
    lib = D.load_library()
    vp = ctypes.c_void_p
    lib.dpkem_run_noalloc.argtypes = [ctypes.c_int, vp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, vp, vp, vp, vp]
    return lib


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _noalloc(torch, scheme, op, n, o0, o1, i0, i1):
    stream = torch.cuda.current_stream()
    rc = _lib().dpkem_run_noalloc(0, ctypes.c_void_p(stream.cuda_stream), scheme, op, n, _p(o0), _p(o1), _p(i0),
                                  _p(i1))
    assert rc == 0, rc
    torch.cuda.synchronize()


@pytest.mark.parametrize("scheme,sizes", [(D.SABER, (1000, 37)), (D.FRODO976, (40, 7))])
def test_noalloc_equals_the_engine_on_two_calls(torch_cuda, oracle, scheme, sizes):
    torch = torch_cuda
    s = D.SIZES[scheme]
    eng = D.Engine(0)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    dev = torch.device("cuda:0")
    try:
        for call, n in enumerate(sizes):
            print(f"no-alloc: scheme {scheme}, call {call}, n={n}", flush=True)
            kc = torch.from_numpy(oracle.prng_bytes(9100 + 10 * scheme + call, n * s["kcoins"]).reshape(n, -1)).to(dev)
            ec = torch.from_numpy(oracle.prng_bytes(9200 + 10 * scheme + call, n * s["ecoins"]).reshape(n, -1)).to(dev)
            # the existing entry points (engine arena)
            pk, sk = eng.keypair(scheme, kc)
            ct, ss = eng.enc(scheme, pk, ec)
            ss2 = eng.dec(scheme, ct, sk)
            eng.sync()
            # the allocation-free one, into fresh outputs
            npk = torch.zeros_like(pk); nsk = torch.zeros_like(sk)
            nct = torch.zeros_like(ct); nss = torch.zeros_like(ss); nss2 = torch.zeros_like(ss2)
            _noalloc(torch, scheme, D.OP_KEYGEN, n, npk, nsk, kc, None)
            _noalloc(torch, scheme, D.OP_ENC, n, nct, nss, npk, ec)
            _noalloc(torch, scheme, D.OP_DEC, n, nss2, None, nct, nsk)
            for a, b, what in ((npk, pk, "pk"), (nsk, sk, "sk"), (nct, ct, "ct"), (nss, ss, "ss"), (nss2, ss2, "dec ss")):
                assert torch.equal(a, b), f"{what} differs from the engine's (call {call}, n={n})"
            assert torch.equal(nss, nss2)
            # and a tampered ciphertext takes the implicit-rejection path in both
            bad = nct.clone()
            bad[:, 5] ^= 1
            rej = torch.zeros_like(ss)
            _noalloc(torch, scheme, D.OP_DEC, n, rej, None, bad, nsk)
            assert torch.equal(rej, eng.dec(scheme, bad, nsk))
    finally:
        eng.close()


def test_noalloc_rejects_batches_beyond_its_scratch(torch_cuda):
    lib = _lib()
    for scheme, m in MAX_N.items():
        # no compute: rejected before anything is launched (pointers never read)
        assert lib.dpkem_run_noalloc(0, None, scheme, D.OP_DEC, m + 1, None, None, None, None) == EBUDGET
    assert lib.dpkem_run_noalloc(0, None, 9, D.OP_DEC, 1, None, None, None, None) == EINVAL
    assert lib.dpkem_run_noalloc(0, None, D.SABER, 3, 1, None, None, None, None) == EINVAL
    # a null output within the limit is an argument error, not a launch
    assert lib.dpkem_run_noalloc(0, None, D.SABER, D.OP_DEC, 4, None, None, None, None) == EINVAL
    assert lib.dpkem_run_noalloc(0, None, D.SABER, D.OP_DEC, 0, None, None, None, None) == 0
