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


@pytest.mark.parametrize("scheme", [D.FRODO976, D.SABER, D.FRODO976_AES, D.LIGHTSABER, D.FIRESABER, D.FRODO640,
                                    D.FRODO1344])
def test_sizes_query(scheme):
    lib = D.load_library()
    vals = [ctypes.c_size_t() for _ in range(6)]
    assert lib.dpkem_sizes(scheme, *[ctypes.byref(v) for v in vals]) == 0
    s = D.SIZES[scheme]
    assert [v.value for v in vals] == [s["pk"], s["sk"], s["ct"], s["ss"], s["kcoins"], s["ecoins"]]
    assert lib.dpkem_sizes(7, *[None] * 6) == -1
    assert lib.dpkem_sizes(-1, *[None] * 6) == -1


def test_set_tuning_rejects_a_null_engine():
    lib = D.load_library()
    lib.dpkem_engine_set_tuning.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    assert lib.dpkem_engine_set_tuning(None, D.TUNE_SABER_POLY_IPW, 2) == -1


def test_server_struct_layouts_match_numpy(tmp_path):
    """dpkem_request / dpkem_completion / dpkem_server_config as compiled by the
    host C compiler have the layouts the Python binding uses."""
    import os
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "dpkem.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(dpkem_request), offsetof(dpkem_request, scheme),
         offsetof(dpkem_request, key), offsetof(dpkem_request, t_arrival_us), offsetof(dpkem_request, coins),
         sizeof(dpkem_completion), offsetof(dpkem_completion, ss));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(dpkem_completion, t_submit_us), offsetof(dpkem_completion, key),
         sizeof(dpkem_server_config), offsetof(dpkem_server_config, reserve_lanes),
         offsetof(dpkem_completion, t_admit_us), offsetof(dpkem_completion, t_done_us));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-I", os.path.join(root, "include"), "-o", str(exe)], check=True)
    a, b = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    a = list(map(int, a.split())); b = list(map(int, b.split()))
    R, C = D.REQ_DTYPE, D.COMP_DTYPE
    assert a == [R.itemsize, R.fields["scheme"][1], R.fields["key"][1], R.fields["t_arrival_us"][1],
                 R.fields["coins"][1], C.itemsize, C.fields["ss"][1]]
    assert b[:2] == [C.fields["t_submit_us"][1], C.fields["key"][1]]
    import ctypes
    from dpcrypto_amd.kem import _ServerConfig
    assert b[2] == ctypes.sizeof(_ServerConfig)
    assert b[3] == _ServerConfig.reserve_lanes.offset
    assert b[4:6] == [C.fields["t_admit_us"][1], C.fields["t_done_us"][1]]
