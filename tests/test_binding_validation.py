"""Misuse of the Python binding is rejected before any pointer reaches the library.

CPU-only: the engine and plan objects are built without a GPU (no C call is
made), and every case must raise in the binding's own checks.  The library
receives bare base pointers, so an array of the wrong width, rank or row
count, or host memory recorded into a graph, would otherwise let the kernels
read or write outside the caller's buffers.
"""
import numpy as np
import pytest

import dpcrypto_amd as D
from dpcrypto_amd import kem as K

torch = pytest.importorskip("torch")


def _engine():
    e = K.Engine.__new__(K.Engine)
    e.lib = D.load_library()
    e.h = None
    e.device = 0
    return e


def _plan(e):
    p = K.Plan.__new__(K.Plan)
    p.eng, p.lib, p.h, p._keep = e, e.lib, None, []
    return p


def u8(*shape):
    return torch.zeros(shape, dtype=torch.uint8)


@pytest.mark.parametrize("bad,exc", [
    (u8(4 * 96), ValueError),                                 # 1-D
    (u8(4, 32), ValueError),                                  # encaps-width coins given to keygen
    (u8(4, 96, 1), ValueError),                               # 3-D
    (torch.zeros((4, 96), dtype=torch.int32), TypeError),     # wrong dtype
    (u8(96, 4).t(), ValueError),                              # right shape, not contiguous
    (u8(4, 96), TypeError),                                   # right shape, host tensor
])
def test_keypair_rejects_malformed_tensors(bad, exc):
    with pytest.raises(exc):
        _engine().keypair(D.SABER, bad)


def test_enc_checks_every_array_width():
    e = _engine()
    with pytest.raises(ValueError, match="pk"):
        e.enc(D.FRODO976, u8(2, 15631), u8(2, 24))
    with pytest.raises(ValueError, match="coins"):
        e.enc(D.FRODO976, u8(2, 15632), u8(2, 16))
    with pytest.raises(ValueError, match="sk"):
        e.dec(D.SABER, u8(2, 1088), u8(2, 2303))


def test_raw_entry_points_check_roles_and_rows():
    e = _engine()
    s = D.SIZES[D.SABER]
    with pytest.raises(ValueError, match="rows"):
        # 3 rows of coins for 4 instances
        e.keypair_into(D.SABER, 4, u8(4, s["pk"]), u8(4, s["sk"]), u8(3, s["kcoins"]))
    with pytest.raises(ValueError, match="ss"):
        e.dec_into(D.SABER, 4, u8(4, 24), u8(4, s["ct"]), u8(4, s["sk"]))


def test_plan_rejects_host_arrays_and_wrong_shapes():
    e = _engine()
    p = _plan(e)
    s = D.SIZES[D.SABER]
    host = np.zeros((4, s["kcoins"]), np.uint8)
    with pytest.raises(TypeError):
        p.add(D.SABER, K.OP_KEYGEN, 4, u8(4, s["pk"]), u8(4, s["sk"]), host)
    with pytest.raises(ValueError, match="sk"):
        p.add(D.SABER, K.OP_KEYGEN, 4, u8(4, s["pk"]), u8(4, 10), u8(4, s["kcoins"]))
    with pytest.raises(ValueError, match="rows"):
        p.add(D.SABER, K.OP_ENC, 8, u8(8, s["ct"]), u8(8, s["ss"]), u8(4, s["pk"]), u8(8, s["ecoins"]))
    with pytest.raises(ValueError, match="missing"):
        p.add(D.SABER, K.OP_DEC, 4, u8(4, s["ss"]), None, None, u8(4, s["sk"]))
    with pytest.raises(ValueError):
        p.add(D.SABER, K.OP_DEC, 4, u8(4, s["ss"]), u8(4, s["ss"]), u8(4, s["ct"]), u8(4, s["sk"]))
    with pytest.raises(ValueError):
        p.add(9, K.OP_DEC, 4)
    # a well-formed host tensor passes the shape checks and fails only for living on the host
    with pytest.raises(TypeError, match="GPU"):
        p.add(D.SABER, K.OP_KEYGEN, 4, u8(4, s["pk"]), u8(4, s["sk"]), u8(4, s["kcoins"]))


def test_host_numpy_inputs_are_reshaped_by_width():
    e = _engine()
    x = e._host(np.zeros(2 * 96, np.uint8), 96)
    assert x.shape == (2, 96)
    with pytest.raises(ValueError):
        e._host(np.zeros(95, np.uint8), 96)
