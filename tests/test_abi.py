"""The C-ABI library loads and exports every function include/dpkem.h declares.

CPU-only: no compute call is made (there is no GPU here), only symbol lookup
and the pure-host size query.
"""
import ctypes

import pytest

import dpcrypto_amd as D


def test_header_declares_entry_points():
    names = D.header_functions()
    for must in ["dpkem_keypair", "dpkem_enc", "dpkem_dec", "dpkem_engine_create",
                 "frodoKEMKeypair", "frodoKEMEncap", "frodoKEMDecap", "saber_enc", "saber_dec",
                 "dpkem_frodo976_crypto_kem_keypair", "dpkem_saber_crypto_kem_dec"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = D.load_library()
    missing = [n for n in D.header_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("scheme", [D.FRODO976, D.SABER])
def test_sizes_query(scheme):
    lib = D.load_library()
    vals = [ctypes.c_size_t() for _ in range(6)]
    assert lib.dpkem_sizes(scheme, *[ctypes.byref(v) for v in vals]) == 0
    s = D.SIZES[scheme]
    assert [v.value for v in vals] == [s["pk"], s["sk"], s["ct"], s["ss"], s["kcoins"], s["ecoins"]]
    assert lib.dpkem_sizes(7, *[None] * 6) == -1
