"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only the test suite, ``__graft_entry__.smoke()`` and the benchmark's CPU
baseline leg import this module, and only as the checker / baseline.  The
product path (``dpcrypto_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FRODO = 0
SABER = 1
FRODO_AES = 2
LIGHTSABER = 3
FIRESABER = 4
FRODO640 = 5
FRODO1344 = 6

SIZES = {
    FRODO: dict(pk=15632, sk=31296, ct=15744, ss=24, kcoins=64, ecoins=24),
    SABER: dict(pk=992, sk=2304, ct=1088, ss=32, kcoins=96, ecoins=32),
    FRODO_AES: dict(pk=15632, sk=31296, ct=15744, ss=24, kcoins=64, ecoins=24),
    LIGHTSABER: dict(pk=672, sk=1568, ct=736, ss=32, kcoins=96, ecoins=32),
    FIRESABER: dict(pk=1312, sk=3040, ct=1472, ss=32, kcoins=96, ecoins=32),
    FRODO640: dict(pk=9616, sk=19888, ct=9720, ss=16, kcoins=48, ecoins=16),
    FRODO1344: dict(pk=21520, sk=43088, ct=21632, ss=32, kcoins=80, ecoins=32),
}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only; no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def build_native(out_path: str) -> str:
    """Compile the oracle with -O3 -march=native for this host into out_path
    (used by the benchmark's CPU baseline on the machine it measures)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "native", f"NATIVE_OUT={out_path}"], check=True)
    return out_path


def use_library(path: str | None) -> None:
    """Route the functions of this module to another build of the oracle
    (e.g. build_native's), or back to liboracle.so with None."""
    global _lib
    _lib = _bind(path) if path else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(_LIB_PATH)
    return _lib


def _bind(path: str):
    """Load one build of the oracle and declare its argument types."""
    lib_ = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    lib_.orc_shake128.argtypes = [vp, sz, vp, sz]
    lib_.orc_shake256.argtypes = [vp, sz, vp, sz]
    lib_.orc_sha3_256.argtypes = [vp, vp, sz]
    lib_.orc_sha3_512.argtypes = [vp, vp, sz]
    lib_.orc_keccak_f1600.argtypes = [vp]
    lib_.orc_prng_fill.argtypes = [ctypes.c_uint64, vp, sz]
    lib_.orc_batch.argtypes = [ctypes.c_int, ctypes.c_int, sz, ctypes.c_int, vp, vp, vp, vp]
    lib_.orc_batch.restype = ctypes.c_int
    lib_.orc_frodo_sample.argtypes = [vp, sz]
    lib_.orc_frodo_n_sample.argtypes = [ctypes.c_int, vp, sz]
    lib_.orc_frodo_gen_a_row.argtypes = [vp, vp, ctypes.c_uint16]
    lib_.orc_frodo_aes_gen_a_row.argtypes = [vp, vp, ctypes.c_uint16]
    lib_.orc_aes128_ecb.argtypes = [vp, vp, vp, sz]
    lib_.orc_saber_l_gen_matrix.argtypes = [ctypes.c_int, vp, vp]
    lib_.orc_saber_l_gen_secret.argtypes = [ctypes.c_int, vp, vp]
    lib_.orc_frodo_pack16.argtypes = [vp, vp, sz]
    lib_.orc_frodo_unpack16.argtypes = [vp, vp, sz]
    lib_.orc_frodo_encode.argtypes = [vp, vp]
    lib_.orc_frodo_decode.argtypes = [vp, vp]
    lib_.orc_frodo_as_plus_e.argtypes = [vp, vp, vp, vp]
    lib_.orc_frodo_sa_plus_e.argtypes = [vp, vp, vp, vp]
    lib_.orc_saber_gen_matrix.argtypes = [vp, vp]
    lib_.orc_saber_gen_secret.argtypes = [vp, vp]
    lib_.orc_saber_polymul_acc.argtypes = [vp, vp, vp]
    lib_.orc_saber_matvec.argtypes = [vp, vp, vp, ctypes.c_int]
    lib_.orc_saber_pack13.argtypes = [vp, vp, sz]
    lib_.orc_saber_unpack13.argtypes = [vp, vp, sz]
    lib_.orc_saber_pack10.argtypes = [vp, vp, sz]
    lib_.orc_saber_unpack10.argtypes = [vp, vp, sz]
    lib_.orc_saber_indcpa_enc.argtypes = [vp, vp, vp, vp]
    return lib_


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _u8(b) -> np.ndarray:
    return np.ascontiguousarray(np.frombuffer(bytes(b), dtype=np.uint8)) if not isinstance(b, np.ndarray) else np.ascontiguousarray(b, dtype=np.uint8)


# ---------------- hashing ----------------
def shake128(data: bytes, outlen: int) -> bytes:
    src = _u8(data); out = np.zeros(outlen, np.uint8)
    lib().orc_shake128(_p(out), outlen, _p(src), len(src))
    return out.tobytes()


def shake256(data: bytes, outlen: int) -> bytes:
    src = _u8(data); out = np.zeros(outlen, np.uint8)
    lib().orc_shake256(_p(out), outlen, _p(src), len(src))
    return out.tobytes()


def sha3_256(data: bytes) -> bytes:
    src = _u8(data); out = np.zeros(32, np.uint8)
    lib().orc_sha3_256(_p(out), _p(src), len(src))
    return out.tobytes()


def sha3_512(data: bytes) -> bytes:
    src = _u8(data); out = np.zeros(64, np.uint8)
    lib().orc_sha3_512(_p(out), _p(src), len(src))
    return out.tobytes()


def keccak_f1600(state: np.ndarray) -> np.ndarray:
    st = np.ascontiguousarray(state, dtype=np.uint64).copy()
    lib().orc_keccak_f1600(_p(st))
    return st


def prng_bytes(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint8)
    lib().orc_prng_fill(ctypes.c_uint64(seed & (2**64 - 1)), _p(out), n)
    return out


# ---------------- batched KEM ----------------
def keypair(scheme: int, coins: np.ndarray, threads: int = 1):
    s = SIZES[scheme]
    coins = np.ascontiguousarray(coins, np.uint8).reshape(-1, s["kcoins"])
    n = coins.shape[0]
    pk = np.zeros((n, s["pk"]), np.uint8); sk = np.zeros((n, s["sk"]), np.uint8)
    rc = lib().orc_batch(scheme, 0, n, threads, _p(pk), _p(sk), _p(coins), None)
    if rc:
        raise RuntimeError(f"oracle keypair failed ({rc})")
    return pk, sk


def enc(scheme: int, pk: np.ndarray, coins: np.ndarray, threads: int = 1):
    s = SIZES[scheme]
    pk = np.ascontiguousarray(pk, np.uint8).reshape(-1, s["pk"])
    coins = np.ascontiguousarray(coins, np.uint8).reshape(-1, s["ecoins"])
    n = pk.shape[0]
    ct = np.zeros((n, s["ct"]), np.uint8); ss = np.zeros((n, s["ss"]), np.uint8)
    rc = lib().orc_batch(scheme, 1, n, threads, _p(ct), _p(ss), _p(pk), _p(coins))
    if rc:
        raise RuntimeError(f"oracle enc failed ({rc})")
    return ct, ss


def dec(scheme: int, ct: np.ndarray, sk: np.ndarray, threads: int = 1):
    s = SIZES[scheme]
    ct = np.ascontiguousarray(ct, np.uint8).reshape(-1, s["ct"])
    sk = np.ascontiguousarray(sk, np.uint8).reshape(-1, s["sk"])
    n = ct.shape[0]
    ss = np.zeros((n, s["ss"]), np.uint8)
    rc = lib().orc_batch(scheme, 2, n, threads, _p(ss), None, _p(ct), _p(sk))
    if rc:
        raise RuntimeError(f"oracle dec failed ({rc})")
    return ss


# ---------------- Frodo building blocks ----------------
def frodo_sample(r: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(r, np.uint16).copy()
    lib().orc_frodo_sample(_p(s), s.size)
    return s


def frodo_n_sample(n: int, r: np.ndarray) -> np.ndarray:
    """The error sampler of Frodo-<n> (640, 976 or 1344) over raw 16-bit words."""
    s = np.ascontiguousarray(r, np.uint16).copy()
    lib().orc_frodo_n_sample(n, _p(s), s.size)
    return s


def frodo_gen_a_row(seed_a: bytes, i: int) -> np.ndarray:
    row = np.zeros(976, np.uint16); sa = _u8(seed_a)
    lib().orc_frodo_gen_a_row(_p(row), _p(sa), i)
    return row


def frodo_aes_gen_a_row(seed_a: bytes, i: int) -> np.ndarray:
    row = np.zeros(976, np.uint16); sa = _u8(seed_a)
    lib().orc_frodo_aes_gen_a_row(_p(row), _p(sa), i)
    return row


def aes128_ecb(key: bytes, data: bytes) -> bytes:
    k = _u8(key); d = _u8(data); out = np.zeros(len(data), np.uint8)
    lib().orc_aes128_ecb(_p(k), _p(d), _p(out), len(data) // 16)
    return out.tobytes()


def saber_l_gen_matrix(l: int, seed: bytes) -> np.ndarray:
    a = np.zeros(l * l * 256, np.uint16); sd = _u8(seed)
    lib().orc_saber_l_gen_matrix(l, _p(a), _p(sd))
    return a


def saber_l_gen_secret(l: int, seed: bytes) -> np.ndarray:
    s = np.zeros(l * 256, np.uint16); sd = _u8(seed)
    lib().orc_saber_l_gen_secret(l, _p(s), _p(sd))
    return s


def frodo_pack16(v: np.ndarray) -> bytes:
    v = np.ascontiguousarray(v, np.uint16); out = np.zeros(2 * v.size, np.uint8)
    lib().orc_frodo_pack16(_p(out), _p(v), v.size)
    return out.tobytes()


def frodo_unpack16(b: bytes) -> np.ndarray:
    src = _u8(b); out = np.zeros(len(src) // 2, np.uint16)
    lib().orc_frodo_unpack16(_p(out), _p(src), out.size)
    return out


def frodo_encode(mu: bytes) -> np.ndarray:
    src = _u8(mu); out = np.zeros(64, np.uint16)
    lib().orc_frodo_encode(_p(out), _p(src))
    return out


def frodo_decode(v: np.ndarray) -> bytes:
    v = np.ascontiguousarray(v, np.uint16); out = np.zeros(24, np.uint8)
    lib().orc_frodo_decode(_p(out), _p(v))
    return out.tobytes()


def frodo_as_plus_e(st: np.ndarray, e: np.ndarray, seed_a: bytes) -> np.ndarray:
    st = np.ascontiguousarray(st, np.uint16); e = np.ascontiguousarray(e, np.uint16)
    out = np.zeros(976 * 8, np.uint16); sa = _u8(seed_a)
    lib().orc_frodo_as_plus_e(_p(out), _p(st), _p(e), _p(sa))
    return out


def frodo_sa_plus_e(sp: np.ndarray, ep: np.ndarray, seed_a: bytes) -> np.ndarray:
    sp = np.ascontiguousarray(sp, np.uint16); ep = np.ascontiguousarray(ep, np.uint16)
    out = np.zeros(976 * 8, np.uint16); sa = _u8(seed_a)
    lib().orc_frodo_sa_plus_e(_p(out), _p(sp), _p(ep), _p(sa))
    return out


# ---------------- Saber building blocks ----------------
def saber_gen_matrix(seed: bytes) -> np.ndarray:
    out = np.zeros(9 * 256, np.uint16); sd = _u8(seed)
    lib().orc_saber_gen_matrix(_p(out), _p(sd))
    return out


def saber_gen_secret(seed: bytes) -> np.ndarray:
    out = np.zeros(3 * 256, np.uint16); sd = _u8(seed)
    lib().orc_saber_gen_secret(_p(out), _p(sd))
    return out


def saber_polymul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16); b = np.ascontiguousarray(b, np.uint16)
    res = np.zeros(256, np.uint16)
    lib().orc_saber_polymul_acc(_p(a), _p(b), _p(res))
    return res


def saber_matvec(a: np.ndarray, s: np.ndarray, transpose: bool) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint16); s = np.ascontiguousarray(s, np.uint16)
    res = np.zeros(768, np.uint16)
    lib().orc_saber_matvec(_p(res), _p(a), _p(s), int(transpose))
    return res


def saber_pack13(v: np.ndarray) -> bytes:
    v = np.ascontiguousarray(v, np.uint16); npoly = v.size // 256
    out = np.zeros(416 * npoly, np.uint8)
    lib().orc_saber_pack13(_p(out), _p(v), npoly)
    return out.tobytes()


def saber_unpack13(b: bytes) -> np.ndarray:
    src = _u8(b); npoly = len(src) // 416
    out = np.zeros(256 * npoly, np.uint16)
    lib().orc_saber_unpack13(_p(out), _p(src), npoly)
    return out


def saber_pack10(v: np.ndarray) -> bytes:
    v = np.ascontiguousarray(v, np.uint16); npoly = v.size // 256
    out = np.zeros(320 * npoly, np.uint8)
    lib().orc_saber_pack10(_p(out), _p(v), npoly)
    return out.tobytes()


def saber_unpack10(b: bytes) -> np.ndarray:
    src = _u8(b); npoly = len(src) // 320
    out = np.zeros(256 * npoly, np.uint16)
    lib().orc_saber_unpack10(_p(out), _p(src), npoly)
    return out


def saber_indcpa_enc(m: bytes, seed_sp: bytes, pk: bytes) -> bytes:
    out = np.zeros(1088, np.uint8)
    mm, sd, p = _u8(m), _u8(seed_sp), _u8(pk)
    lib().orc_saber_indcpa_enc(_p(out), _p(mm), _p(sd), _p(p))
    return out.tobytes()
