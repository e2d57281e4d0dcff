"""Python binding of the dpkem C ABI (include/dpkem.h) over ctypes.

The product path is the HIP library ``dpcrypto_amd/libdpkem.so``; this module
only marshals pointers.  Arrays may be numpy ``uint8`` arrays (host memory)
or torch ``uint8`` tensors on the engine's GPU (device memory, no copies).
There is no CPU fallback: if the library is missing the import of the engine
fails with an error that says how to build it.

Mirrors the reference's batched entry points (FrodoKEM976-GPU/include/
cuda_kernel.cuh:7-9, SABER-GPU/include/cuda_kernel.cuh:6-7) with explicit
coins per instance.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpkem.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "dpkem.h")

FRODO976 = 0
SABER = 1
MEM_HOST = 0
MEM_DEVICE = 1
FORM_BATCHED = 0
FORM_PLAIN = 1
# kernel tuning knobs (dpkem_engine_set_tuning)
TUNE_SABER_POLY_IPW = 1
TUNE_FRODO_PERSIST = 2
TUNE_FRODO_THREADS = 3

ERRORS = {
    -1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "EBUDGET", -5: "EBUSY", -6: "ENOTFOUND", -7: "ECANCELED",
}

FRODO976_AES = 2  # FrodoKEM-976-AES: Frodo-976 with A rows from AES-128 (batched form only)
LIGHTSABER = 3    # Saber family, rank 2 (batched form only)
FIRESABER = 4     # Saber family, rank 4 (batched form only)
FRODO640 = 5      # Frodo-640-SHAKE (batched form only)
FRODO1344 = 6     # Frodo-1344-SHAKE (batched form only)

SIZES = {
    FRODO976: dict(pk=15632, sk=31296, ct=15744, ss=24, kcoins=64, ecoins=24),
    SABER: dict(pk=992, sk=2304, ct=1088, ss=32, kcoins=96, ecoins=32),
    FRODO976_AES: dict(pk=15632, sk=31296, ct=15744, ss=24, kcoins=64, ecoins=24),
    LIGHTSABER: dict(pk=672, sk=1568, ct=736, ss=32, kcoins=96, ecoins=32),
    FIRESABER: dict(pk=1312, sk=3040, ct=1472, ss=32, kcoins=96, ecoins=32),
    FRODO640: dict(pk=9616, sk=19888, ct=9720, ss=16, kcoins=48, ecoins=16),
    FRODO1344: dict(pk=21520, sk=43088, ct=21632, ss=32, kcoins=80, ecoins=32),
}

_lib = None


class DpkemError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"dpkem error {code} ({ERRORS.get(code, '?')}) {what}".strip())
        self.code = code


def _check_rc(rc, what):
    if rc:
        raise DpkemError(rc, what)


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libdpkem.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C dpcrypto_amd`")
        lib = ctypes.CDLL(path)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        lib.dpkem_sizes.argtypes = [i] + [ctypes.POINTER(sz)] * 6
        lib.dpkem_engine_create.argtypes = [ctypes.POINTER(vp), i, sz]
        lib.dpkem_engine_destroy.argtypes = [vp]
        lib.dpkem_engine_set_stream.argtypes = [vp, vp, i]
        lib.dpkem_engine_set_form.argtypes = [vp, i]
        lib.dpkem_engine_set_tuning.argtypes = [vp, i, i]
        lib.dpkem_engine_sync.argtypes = [vp]
        lib.dpkem_engine_last_error.argtypes = [vp]
        lib.dpkem_engine_last_error.restype = ctypes.c_char_p
        lib.dpkem_keypair.argtypes = [vp, i, sz, vp, vp, vp, i]
        lib.dpkem_enc.argtypes = [vp, i, sz, vp, vp, vp, vp, i]
        lib.dpkem_dec.argtypes = [vp, i, sz, vp, vp, vp, i]
        lib.dpkem_engine_set_profiling.argtypes = [vp, i]
        lib.dpkem_plan_create.argtypes = [vp, ctypes.POINTER(vp), i, i]
        lib.dpkem_plan_add.argtypes = [vp, i, i, sz, vp, vp, vp, vp, i, ctypes.POINTER(i)]
        lib.dpkem_plan_set_options.argtypes = [vp, ctypes.c_uint]
        lib.dpkem_plan_build.argtypes = [vp]
        lib.dpkem_plan_launch.argtypes = [vp]
        lib.dpkem_plan_info.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        lib.dpkem_plan_destroy.argtypes = [vp]
        lib.dpkem_roll_create.argtypes = [vp, i, sz, ctypes.POINTER(vp)]
        lib.dpkem_roll_bind.argtypes = [vp, i, vp, vp, vp, vp, vp, vp, vp]
        lib.dpkem_roll_build.argtypes = [vp, ctypes.c_uint]
        lib.dpkem_roll_step.argtypes = [vp]
        lib.dpkem_roll_info.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(sz)]
        lib.dpkem_roll_destroy.argtypes = [vp]
        lib.dpkem_now_us.restype = ctypes.c_double
        lib.dpkem_server_create.argtypes = [ctypes.POINTER(vp), vp]
        lib.dpkem_server_submit.argtypes = [vp, vp, sz]
        lib.dpkem_server_poll.argtypes = [vp, vp, sz, ctypes.POINTER(sz)]
        lib.dpkem_server_drain.argtypes = [vp]
        lib.dpkem_server_cancel.argtypes = [vp, ctypes.POINTER(sz)]
        lib.dpkem_server_release.argtypes = [vp, i, i, ctypes.c_uint32]
        lib.dpkem_server_read.argtypes = [vp, i, i, ctypes.c_uint32, vp]
        lib.dpkem_server_stats.argtypes = [vp, ctypes.c_char_p, sz]
        lib.dpkem_server_destroy.argtypes = [vp]
        lib.dpkem_engine_kernel_stats.argtypes = [vp, ctypes.c_char_p, sz, i]
        lib.dpkem_probe_perm_latency.argtypes = [vp, i, ctypes.POINTER(ctypes.c_float)]
        lib.dpkem_enc_shared_key.argtypes = [vp, i, sz, vp, vp, vp, vp, i]
        lib.dpkem_dec_shared_key.argtypes = [vp, i, sz, vp, vp, vp, i]
        lib.dpkem_aes256_ecb.argtypes = [vp, vp, vp, sz]
        lib.dpkem_drbg_init.argtypes = [vp, vp, vp]
        lib.dpkem_drbg_bytes.argtypes = [vp, vp, sz]
        lib.dpkem_kat_seeds.argtypes = [sz, vp]
        lib.dpkem_kat_coins.argtypes = [i, sz, vp, vp, vp]
        _lib = lib
    return _lib


def header_functions(path: str = HEADER_PATH):
    """Names of the functions declared in include/dpkem.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def check_device_array(x, width, rows=None, device=None, what="array", min_rows=None, placement=True):
    """Raise unless `x` is a contiguous 2-D uint8 CUDA tensor of `width`
    bytes per row (exactly `rows` / at least `min_rows` rows, and on GPU
    `device`, when given).  The library receives only a base pointer, so a
    tensor of the wrong shape would make its kernels read or write past the
    tensor's end.  placement=False checks everything but where it lives."""
    if not _is_torch(x):
        raise TypeError(f"{what}: device-resident arrays must be torch CUDA tensors, got {type(x).__name__}")
    import torch
    if x.dtype != torch.uint8:
        raise TypeError(f"{what}: dtype must be torch.uint8, got {x.dtype}")
    if x.dim() != 2 or x.shape[1] != width:
        raise ValueError(f"{what}: expected shape (n, {width}), got {tuple(x.shape)}")
    if rows is not None and x.shape[0] != rows:
        raise ValueError(f"{what}: expected {rows} rows, got {x.shape[0]}")
    if min_rows is not None and x.shape[0] < min_rows:
        raise ValueError(f"{what}: {x.shape[0]} rows for {min_rows} instances")
    if not x.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if not placement:
        return
    if not x.is_cuda:
        raise TypeError(f"{what}: tensor must live on a GPU, got device {x.device}")
    if device is not None and x.device.index != device:
        raise ValueError(f"{what}: tensor is on cuda:{x.device.index}, the engine on cuda:{device}")


def check_role_arrays(scheme, op, n, arrays, device):
    """Check the (o0, o1, i0, i1) arrays of one op against the scheme's
    sizes: every shape first, then where each lives."""
    if scheme not in SIZES or op not in ROLES:
        raise ValueError(f"unknown scheme {scheme} or op {op}")
    todo = []
    for x, role in zip(arrays, ROLES[op]):
        if role is None:
            if x is not None:
                raise ValueError(f"op {op} takes no array in this position")
            continue
        if x is None:
            raise ValueError(f"{role}: missing")
        todo.append((x, role))
    for placement in (False, True):
        for x, role in todo:
            check_device_array(x, SIZES[scheme][role], device=device, what=role, min_rows=n, placement=placement)


def _ptr(x):
    if x is None:
        return None
    if _is_torch(x):
        return ctypes.c_void_p(x.data_ptr())
    return x.ctypes.data_as(ctypes.c_void_p)


class Engine:
    """One engine = one GPU, one HIP stream, one grow-only scratch arena."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, form: int = FORM_BATCHED):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.dpkem_engine_create(ctypes.byref(h), device, hbm_budget)
        if rc:
            raise DpkemError(rc, "engine_create")
        self.h = h
        self.device = device
        if form != FORM_BATCHED:
            self.set_form(form)

    def close(self):
        if getattr(self, "h", None):
            self.lib.dpkem_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise DpkemError(rc, f"{what}: {self.lib.dpkem_engine_last_error(self.h).decode(errors='replace')}")

    def set_form(self, form: int):
        self._check(self.lib.dpkem_engine_set_form(self.h, form), "set_form")

    def set_tuning(self, knob: int, value: int):
        """Set a launch-shape knob (TUNE_*) for later calls and plans created afterwards."""
        self._check(self.lib.dpkem_engine_set_tuning(self.h, knob, value), "set_tuning")

    def set_stream(self, stream_handle: int | None, own: bool = False):
        """Bind to a HIP stream handle (0 = null stream), or back to the engine's own."""
        rc = self.lib.dpkem_engine_set_stream(self.h, ctypes.c_void_p(stream_handle or 0), int(own))
        self._check(rc, "set_stream")

    def sync(self):
        self._check(self.lib.dpkem_engine_sync(self.h), "sync")

    def set_profiling(self, on: bool):
        self._check(self.lib.dpkem_engine_set_profiling(self.h, int(on)), "set_profiling")

    def probe_perm_latency(self, perms: int) -> float:
        """ms of one launch of `perms` back-to-back lane-pair permutations on
        one wave (the serial chains' latency floor, measured)."""
        ms = ctypes.c_float(0.0)
        self._check(self.lib.dpkem_probe_perm_latency(self.h, int(perms), ctypes.byref(ms)), "probe_perm_latency")
        return ms.value

    def kernel_stats(self, reset: bool = True) -> dict:
        """{kernel: (launches, total_ms)} from HIP events on the engine's stream."""
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(self.lib.dpkem_engine_kernel_stats(self.h, buf, len(buf), int(reset)), "kernel_stats")
        return {k: (v[0], v[1]) for k, v in json.loads(buf.value.decode()).items()}

    # -------- array helpers
    def _mem(self, *arrays):
        kinds = {_is_torch(a) for a in arrays if a is not None}
        if len(kinds) != 1:
            raise TypeError("mix of host and device arrays")
        if kinds.pop():
            import torch
            for a in arrays:
                if a is not None:
                    check_device_array(a, a.shape[-1] if a.dim() else 0, device=self.device)
            # order the engine's work after / before torch's on the same stream
            self.set_stream(torch.cuda.current_stream(arrays[0].device).cuda_stream)
            return MEM_DEVICE
        return MEM_HOST

    @staticmethod
    def _empty_like_kind(ref, n, width):
        if _is_torch(ref):
            import torch
            return torch.empty((n, width), dtype=torch.uint8, device=ref.device)
        return np.empty((n, width), np.uint8)

    def _host(self, x, width, what="array"):
        """Validate one input array: a torch tensor must already be a
        contiguous (rows, width) uint8 tensor on the engine's GPU (the kernels
        index it with that row width); host data is copied into a contiguous
        (rows, width) uint8 array."""
        if _is_torch(x):
            check_device_array(x, width, what=what, placement=False)
            return x
        return np.ascontiguousarray(x, dtype=np.uint8).reshape(-1, width)

    # -------- KEM
    def keypair(self, scheme: int, coins):
        s = SIZES[scheme]
        coins = self._host(coins, s["kcoins"], "coins")
        n = coins.shape[0]
        pk = self._empty_like_kind(coins, n, s["pk"])
        sk = self._empty_like_kind(coins, n, s["sk"])
        mem = self._mem(coins)
        self._check(self.lib.dpkem_keypair(self.h, scheme, n, _ptr(pk), _ptr(sk), _ptr(coins), mem), "keypair")
        return pk, sk

    def enc(self, scheme: int, pk, coins):
        s = SIZES[scheme]
        pk = self._host(pk, s["pk"], "pk"); coins = self._host(coins, s["ecoins"], "coins")
        n = pk.shape[0]
        if coins.shape[0] != n:
            raise ValueError("pk and coins batch sizes differ")
        ct = self._empty_like_kind(pk, n, s["ct"])
        ss = self._empty_like_kind(pk, n, s["ss"])
        mem = self._mem(pk, coins)
        self._check(self.lib.dpkem_enc(self.h, scheme, n, _ptr(ct), _ptr(ss), _ptr(pk), _ptr(coins), mem), "enc")
        return ct, ss

    def dec(self, scheme: int, ct, sk):
        s = SIZES[scheme]
        ct = self._host(ct, s["ct"], "ct"); sk = self._host(sk, s["sk"], "sk")
        n = ct.shape[0]
        if sk.shape[0] != n:
            raise ValueError("ct and sk batch sizes differ")
        ss = self._empty_like_kind(ct, n, s["ss"])
        mem = self._mem(ct, sk)
        self._check(self.lib.dpkem_dec(self.h, scheme, n, _ptr(ss), _ptr(ct), _ptr(sk), mem), "dec")
        return ss

    def enc_shared_key(self, scheme: int, pk, coins):
        """n encapsulations to one public key (pk: one row)."""
        s = SIZES[scheme]
        pk = self._host(pk, s["pk"], "pk"); coins = self._host(coins, s["ecoins"], "coins")
        if pk.shape[0] != 1:
            raise ValueError("pk must be one row")
        n = coins.shape[0]
        ct = self._empty_like_kind(coins, n, s["ct"])
        ss = self._empty_like_kind(coins, n, s["ss"])
        mem = self._mem(pk, coins)
        self._check(self.lib.dpkem_enc_shared_key(self.h, scheme, n, _ptr(ct), _ptr(ss), _ptr(pk), _ptr(coins), mem),
                    "enc_shared_key")
        return ct, ss

    def dec_shared_key(self, scheme: int, ct, sk):
        """n decapsulations with one secret key (sk: one row)."""
        s = SIZES[scheme]
        ct = self._host(ct, s["ct"], "ct"); sk = self._host(sk, s["sk"], "sk")
        if sk.shape[0] != 1:
            raise ValueError("sk must be one row")
        n = ct.shape[0]
        ss = self._empty_like_kind(ct, n, s["ss"])
        mem = self._mem(ct, sk)
        self._check(self.lib.dpkem_dec_shared_key(self.h, scheme, n, _ptr(ss), _ptr(ct), _ptr(sk), mem),
                    "dec_shared_key")
        return ss

    def roll(self, scheme: int, n: int) -> "Roll":
        """A rolling keygen -> encaps -> decaps pipeline over three batch
        slots of n instances (dpkem_roll_*; any scheme)."""
        return Roll(self, scheme, n)

    def plan(self, nstreams: int = 1, chunks: int = 1, share_matrix: bool = False, pipeline: bool = True,
             graph: bool = True, merge: bool = True, force_reject: bool = False) -> "Plan":
        return Plan(self, nstreams, chunks, share_matrix, pipeline, graph, merge, force_reject)

    # raw pointer entry points (device arrays preallocated by the caller; each
    # holds at least n rows of its role's width)
    def _check_roles(self, scheme, op, n, arrays):
        check_role_arrays(scheme, op, n, arrays, self.device)

    def keypair_into(self, scheme, n, pk, sk, coins):
        self._check_roles(scheme, OP_KEYGEN, n, (pk, sk, coins, None))
        self._check(self.lib.dpkem_keypair(self.h, scheme, n, _ptr(pk), _ptr(sk), _ptr(coins), MEM_DEVICE), "keypair")

    def enc_into(self, scheme, n, ct, ss, pk, coins):
        self._check_roles(scheme, OP_ENC, n, (ct, ss, pk, coins))
        self._check(self.lib.dpkem_enc(self.h, scheme, n, _ptr(ct), _ptr(ss), _ptr(pk), _ptr(coins), MEM_DEVICE), "enc")

    def dec_into(self, scheme, n, ss, ct, sk):
        self._check_roles(scheme, OP_DEC, n, (ss, None, ct, sk))
        self._check(self.lib.dpkem_dec(self.h, scheme, n, _ptr(ss), _ptr(ct), _ptr(sk), MEM_DEVICE), "dec")


OP_KEYGEN, OP_ENC, OP_DEC = 0, 1, 2
# per op, the SIZES key of each array role (o0, o1, i0, i1) of dpkem_plan_add
ROLES = {OP_KEYGEN: ("pk", "sk", "kcoins", None), OP_ENC: ("ct", "ss", "pk", "ecoins"),
         OP_DEC: ("ss", None, "ct", "sk")}
PLAN_SHARE_MATRIX = 1
PLAN_NO_PIPELINE = 2
PLAN_NO_GRAPH = 4
PLAN_NO_MERGE = 8
PLAN_FORCE_REJECT = 16  # test facility: pipelined chains' decapsulations take the rejection branch
ROLL_OVERLAP = 256  # rolling pipeline over six slots, each step's chains beside the previous step's products in one launch


class Plan:
    """A graph-captured set of KEM tasks over device arrays (dpkem_plan_*)."""

    def __init__(self, engine: Engine, nstreams: int = 1, chunks: int = 1, share_matrix: bool = False,
                 pipeline: bool = True, graph: bool = True, merge: bool = True, force_reject: bool = False):
        self.eng = engine
        self.lib = engine.lib
        h = ctypes.c_void_p()
        engine._check(self.lib.dpkem_plan_create(engine.h, ctypes.byref(h), nstreams, chunks), "plan_create")
        self.h = h
        self._keep = []  # arrays referenced by the graph stay alive with the plan
        flags = ((PLAN_SHARE_MATRIX if share_matrix else 0) | (0 if pipeline else PLAN_NO_PIPELINE)
                 | (0 if graph else PLAN_NO_GRAPH) | (0 if merge else PLAN_NO_MERGE)
                 | (PLAN_FORCE_REJECT if force_reject else 0))
        if flags:
            engine._check(self.lib.dpkem_plan_set_options(self.h, flags), "plan_set_options")

    def add(self, scheme, op, n, o0=None, o1=None, i0=None, i1=None, after=-1) -> int:
        """Record a task over device arrays.  A graph replays the recorded
        pointers later, so every array must be a torch CUDA tensor on the
        engine's GPU with at least n rows of its role's width (host arrays
        would be dereferenced by the GPU)."""
        check_role_arrays(scheme, op, n, (o0, o1, i0, i1), self.eng.device)
        tid = ctypes.c_int(-1)
        self._keep += [a for a in (o0, o1, i0, i1) if a is not None]
        rc = self.lib.dpkem_plan_add(self.h, scheme, op, n, _ptr(o0), _ptr(o1), _ptr(i0), _ptr(i1), after,
                                     ctypes.byref(tid))
        self.eng._check(rc, "plan_add")
        return tid.value

    def keypair(self, scheme, pk, sk, coins, after=-1):
        return self.add(scheme, OP_KEYGEN, pk.shape[0], pk, sk, coins, None, after)

    def enc(self, scheme, ct, ss, pk, coins, after=-1):
        return self.add(scheme, OP_ENC, ct.shape[0], ct, ss, pk, coins, after)

    def dec(self, scheme, ss, ct, sk, after=-1):
        return self.add(scheme, OP_DEC, ss.shape[0], ss, None, ct, sk, after)

    def build(self):
        self.eng._check(self.lib.dpkem_plan_build(self.h), "plan_build")
        return self

    def launch(self):
        self.eng._check(self.lib.dpkem_plan_launch(self.h), "plan_launch")

    def info(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        self.lib.dpkem_plan_info(self.h, ctypes.byref(a), ctypes.byref(b))
        return {"subtasks": a.value, "scratch_bytes": b.value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.dpkem_plan_destroy(self.h)
            self.h = None
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving loop
REQ_DTYPE = np.dtype([("id", "<u8"), ("scheme", "<i4"), ("op", "<i4"), ("key", "<u4"), ("ct", "<u4"),
                      ("t_arrival_us", "<f8"), ("coins", "u1", (96,))])
COMP_DTYPE = np.dtype([("id", "<u8"), ("status", "<i4"), ("scheme", "<i4"), ("op", "<i4"), ("key", "<u4"),
                       ("ct", "<u4"), ("reserved", "<u4"), ("t_submit_us", "<f8"), ("t_admit_us", "<f8"),
                       ("t_done_us", "<f8"),
                       ("ss", "u1", (32,))])


class _ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("form", ctypes.c_int), ("lanes", ctypes.c_int),
                ("hbm_budget_bytes", ctypes.c_size_t), ("key_slots", ctypes.c_size_t * 2),
                ("ct_slots", ctypes.c_size_t * 2), ("max_batch", ctypes.c_size_t * 2),
                ("reserve_lanes", ctypes.c_int * 2), ("group_by_key", ctypes.c_int)]


def now_us() -> float:
    """The serving loop's clock (steady clock, microseconds)."""
    return load_library().dpkem_now_us()


class Server:
    """Continuous-batching request loop over a device-resident key store (dpkem_server_*)."""

    def __init__(self, device=0, lanes=4, hbm_budget=0, key_slots=(1024, 4096), ct_slots=(4096, 16384),
                 max_batch=(0, 0), form=FORM_BATCHED, reserve_lanes=(0, 0), group_by_key=True):
        self.lib = load_library()
        cfg = _ServerConfig(device, form, lanes, hbm_budget, (ctypes.c_size_t * 2)(*key_slots),
                            (ctypes.c_size_t * 2)(*ct_slots), (ctypes.c_size_t * 2)(*max_batch),
                            (ctypes.c_int * 2)(*reserve_lanes), int(bool(group_by_key)))
        h = ctypes.c_void_p()
        rc = self.lib.dpkem_server_create(ctypes.byref(h), ctypes.byref(cfg))
        if rc:
            raise DpkemError(rc, "server_create")
        self.h = h

    def submit(self, reqs: np.ndarray) -> int:
        reqs = np.ascontiguousarray(reqs, dtype=REQ_DTYPE)
        return self.lib.dpkem_server_submit(self.h, reqs.ctypes.data_as(ctypes.c_void_p), len(reqs))

    def poll(self, cap: int = 65536) -> np.ndarray:
        out = np.empty(cap, COMP_DTYPE)
        got = ctypes.c_size_t()
        rc = self.lib.dpkem_server_poll(self.h, out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(got))
        if rc:
            raise DpkemError(rc, "server_poll")
        return out[: got.value]

    def drain(self):
        rc = self.lib.dpkem_server_drain(self.h)
        if rc:
            raise DpkemError(rc, "server_drain")

    def cancel(self) -> int:
        """Drop requests not yet admitted (their completions carry ECANCELED)."""
        k = ctypes.c_size_t()
        rc = self.lib.dpkem_server_cancel(self.h, ctypes.byref(k))
        if rc:
            raise DpkemError(rc, "server_cancel")
        return k.value

    def release(self, scheme, kind, handle):
        return self.lib.dpkem_server_release(self.h, scheme, kind, handle)

    def read(self, scheme, kind, handle) -> np.ndarray:
        size = SIZES[scheme][("pk", "sk", "ct")[kind]]
        out = np.empty(size, np.uint8)
        rc = self.lib.dpkem_server_read(self.h, scheme, kind, handle, out.ctypes.data_as(ctypes.c_void_p))
        if rc:
            raise DpkemError(rc, "server_read")
        return out

    def stats(self) -> dict:
        import json
        buf = ctypes.create_string_buffer(1024)
        self.lib.dpkem_server_stats(self.h, buf, len(buf))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.dpkem_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_requests(n: int) -> np.ndarray:
    return np.zeros(n, REQ_DTYPE)


class Roll:
    """A rolling keygen -> encaps -> decaps pipeline (dpkem_roll_*): three
    batch slots of n instances of one scheme.  Step k generates
    the keys of slot k % 3, encapsulates to those of slot (k - 1) % 3 and
    decapsulates the ciphertexts of slot (k - 2) % 3; after step k >= 2 every
    output of slot (k - 2) % 3 is complete until step k + 1."""

    def __init__(self, engine: Engine, scheme: int, n: int):
        self.eng = engine
        self.lib = engine.lib
        self.scheme = scheme
        self.n = n
        h = ctypes.c_void_p()
        engine._check(self.lib.dpkem_roll_create(engine.h, scheme, n, ctypes.byref(h)), "roll_create")
        self.h = h
        self._keep = {}

    def bind(self, slot: int, kc, ec, pk, sk, ct, ss, ss2):
        """Device arrays of one slot: keygen coins, encaps coins, pk, sk, ct,
        encaps' ss, decaps' ss (torch CUDA tensors on the engine's GPU, at
        least n rows of their role's width)."""
        n = self.n
        check_role_arrays(self.scheme, OP_KEYGEN, n, (pk, sk, kc, None), self.eng.device)
        check_role_arrays(self.scheme, OP_ENC, n, (ct, ss, pk, ec), self.eng.device)
        check_role_arrays(self.scheme, OP_DEC, n, (ss2, None, ct, sk), self.eng.device)
        self._keep[slot] = (kc, ec, pk, sk, ct, ss, ss2)
        self.eng._check(self.lib.dpkem_roll_bind(self.h, slot, _ptr(kc), _ptr(ec), _ptr(pk), _ptr(sk), _ptr(ct),
                                                 _ptr(ss), _ptr(ss2)), "roll_bind")
        return self

    def build(self, graph: bool = True, overlap: bool = False):
        """overlap: the six-slot form (slots 0-5 bound; step k: keygen of
        slot k, encaps of slot k - 2, decaps of slot k - 4, mod 6), one
        launch per step holding its chains and the previous step's products;
        after step k every output of slot (k - 5) % 6 is complete."""
        flags = (0 if graph else PLAN_NO_GRAPH) | (ROLL_OVERLAP if overlap else 0)
        self.eng._check(self.lib.dpkem_roll_build(self.h, flags), "roll_build")
        return self

    def step(self):
        self.eng._check(self.lib.dpkem_roll_step(self.h), "roll_step")

    def info(self) -> dict:
        steps, scr = ctypes.c_uint64(), ctypes.c_size_t()
        self.eng._check(self.lib.dpkem_roll_info(self.h, ctypes.byref(steps), ctypes.byref(scr)), "roll_info")
        return {"steps": steps.value, "scratch_bytes": scr.value}

    def close(self):
        if self.h:
            self.eng._check(self.lib.dpkem_roll_destroy(self.h), "roll_destroy")
            self.h = None
            self._keep = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
