"""NIST PQC known-answer files (.rsp) for FrodoKEM-976-SHAKE and Saber.

A KAT file lists, per count, a 48-byte seed and the pk, sk, ct and ss that
the scheme produces when randombytes() is the AES-256 CTR_DRBG seeded with
that seed.  `dpkem_kat_seeds` / `dpkem_kat_coins` (drbg.cpp) turn the seeds
into the explicit coins of dpkem_keypair / dpkem_enc, so all counts of a
file run as one GPU batch.

    python -m dpcrypto_amd.kat generate --scheme saber --count 100 > PQCkemKAT_2304.rsp
    python -m dpcrypto_amd.kat verify --scheme saber PQCkemKAT_2304.rsp

`verify` recomputes every count of an existing file on the GPU and reports
the first mismatching field.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from . import kem as K

NAMES = {K.FRODO976: "FrodoKEM-976-SHAKE", K.SABER: "Saber", K.FRODO976_AES: "FrodoKEM-976-AES",
         K.LIGHTSABER: "LightSaber", K.FIRESABER: "FireSaber", K.FRODO640: "FrodoKEM-640-SHAKE",
         K.FRODO1344: "FrodoKEM-1344-SHAKE"}


class Drbg:
    """AES-256 CTR_DRBG (SP 800-90A, no derivation function) in libdpkem."""

    class _State(ctypes.Structure):
        _fields_ = [("key", ctypes.c_uint8 * 32), ("v", ctypes.c_uint8 * 16), ("reseed_counter", ctypes.c_uint64)]

    def __init__(self, entropy: bytes, personalization: bytes | None = None):
        if len(entropy) != 48 or (personalization is not None and len(personalization) != 48):
            raise ValueError("entropy and personalization are 48 bytes")
        self.lib = K.load_library()
        self.st = Drbg._State()
        pers = ctypes.c_char_p(personalization) if personalization is not None else None
        K._check_rc(self.lib.dpkem_drbg_init(ctypes.byref(self.st), ctypes.c_char_p(entropy), pers), "drbg_init")

    def bytes(self, n: int) -> bytes:
        out = ctypes.create_string_buffer(n)
        K._check_rc(self.lib.dpkem_drbg_bytes(ctypes.byref(self.st), out, n), "drbg_bytes")
        return out.raw


def aes256_ecb(key: bytes, blocks: bytes) -> bytes:
    if len(key) != 32 or len(blocks) % 16:
        raise ValueError("32-byte key, whole 16-byte blocks")
    out = ctypes.create_string_buffer(len(blocks))
    K._check_rc(K.load_library().dpkem_aes256_ecb(ctypes.c_char_p(key), ctypes.c_char_p(blocks), out,
                                                 len(blocks) // 16), "aes256_ecb")
    return out.raw


def seeds(n: int) -> np.ndarray:
    out = np.zeros((n, 48), np.uint8)
    K._check_rc(K.load_library().dpkem_kat_seeds(n, out.ctypes.data_as(ctypes.c_void_p)), "kat_seeds")
    return out


def coins(scheme: int, seed_rows: np.ndarray):
    seed_rows = np.ascontiguousarray(seed_rows, dtype=np.uint8)
    n = seed_rows.shape[0]
    s = K.SIZES[scheme]
    kc = np.zeros((n, s["kcoins"]), np.uint8)
    ec = np.zeros((n, s["ecoins"]), np.uint8)
    K._check_rc(K.load_library().dpkem_kat_coins(scheme, n, seed_rows.ctypes.data_as(ctypes.c_void_p),
                                                 kc.ctypes.data_as(ctypes.c_void_p),
                                                 ec.ctypes.data_as(ctypes.c_void_p)), "kat_coins")
    return kc, ec


def compute(scheme: int, seed_rows: np.ndarray, keypair, enc, dec):
    """pk, sk, ct, ss of every count, with the given batched keypair / enc /
    dec callables (the GPU engine's, or the oracle's in tests)."""
    kc, ec = coins(scheme, seed_rows)
    pk, sk = keypair(scheme, kc)
    ct, ss = enc(scheme, pk, ec)
    ss2 = dec(scheme, ct, sk)
    if not np.array_equal(np.asarray(ss), np.asarray(ss2)):
        raise RuntimeError("decapsulation disagrees with encapsulation")
    return [np.asarray(x) for x in (pk, sk, ct, ss)]


def format_rsp(scheme: int, seed_rows, pk, sk, ct, ss) -> str:
    hx = lambda a: bytes(a).hex().upper()
    lines = [f"# {NAMES[scheme]}", ""]
    for i in range(len(seed_rows)):
        lines += [f"count = {i}", f"seed = {hx(seed_rows[i])}", f"pk = {hx(pk[i])}", f"sk = {hx(sk[i])}",
                  f"ct = {hx(ct[i])}", f"ss = {hx(ss[i])}", ""]
    return "\n".join(lines)


def parse_rsp(text: str):
    """[{count, seed, pk, sk, ct, ss}] with byte values as bytes."""
    out, cur = [], {}
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        k, _, v = (x.strip() for x in line.partition("="))
        if k == "count":
            if cur:
                out.append(cur)
            cur = {"count": int(v)}
        else:
            cur[k] = bytes.fromhex(v)
    if cur:
        out.append(cur)
    return out


def _engine_ops():
    eng = K.Engine(0)
    return eng, eng.keypair, eng.enc, eng.dec


def _main(argv=None):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["generate", "verify"])
    ap.add_argument("--scheme", choices=["frodo640", "frodo976", "frodo1344", "frodo976aes", "lightsaber", "saber",
                                         "firesaber"], required=True)
    ap.add_argument("--count", type=int, default=100)
    ap.add_argument("file", nargs="?")
    a = ap.parse_args(argv)
    scheme = {"frodo976": K.FRODO976, "frodo976aes": K.FRODO976_AES, "saber": K.SABER,
              "lightsaber": K.LIGHTSABER, "firesaber": K.FIRESABER, "frodo640": K.FRODO640,
              "frodo1344": K.FRODO1344}[a.scheme]
    eng, kp, en, de = _engine_ops()
    try:
        if a.cmd == "generate":
            sd = seeds(a.count)
            sys.stdout.write(format_rsp(scheme, sd, *compute(scheme, sd, kp, en, de)))
            return 0
        recs = parse_rsp(open(a.file).read())
        sd = np.array([np.frombuffer(r["seed"], np.uint8) for r in recs])
        got = compute(scheme, sd, kp, en, de)
        for i, r in enumerate(recs):
            for j, f in enumerate(("pk", "sk", "ct", "ss")):
                if bytes(got[j][i]) != r[f]:
                    print(f"count {r['count']}: {f} differs", file=sys.stderr)
                    return 1
        print(f"{len(recs)} counts match", file=sys.stderr)
        return 0
    finally:
        eng.close()


if __name__ == "__main__":
    sys.exit(_main())
