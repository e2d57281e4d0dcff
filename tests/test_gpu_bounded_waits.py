"""A bounded wait that gives up inside a kernel is reported, not hidden.

Some Saber-family kernels, and the FrodoKEM noise chains with their sampler
waves, hand data between waves of a workgroup through LDS counts and wait
for them with a bounded spin (keccak2.hpp lds_wait_count; each kernel file
keeps its own give-up counter, and the engine reads both).
If a wait reached its bound, the wave would go on with unpublished LDS and
write wrong results; the kernel records that, and the engine's next sync
returns DPKEM_EDEVICE.  The give-up is forced from the host: in a fresh
process, DPKEM_WAIT_SPIN_LIMIT=0 makes every wait give up at once (the
kernels still read only in-bounds LDS and exit normally).  A second engine
with the default bound then runs cleanly and matches the oracle.
"""
import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = textwrap.dedent("""
    import os, sys
    import numpy as np, torch
    sys.path.insert(0, %r)
    import dpcrypto_amd as D
    from oracle import pyoracle as O
    n = 256
    kc = O.prng_bytes(4242, n * 96).reshape(n, 96)
    ec = O.prng_bytes(4243, n * 32).reshape(n, 32)

    def chain(eng):
        dev = torch.device("cuda:0")
        s = D.SIZES[D.SABER]
        a = {k: torch.empty((n, w), dtype=torch.uint8, device=dev)
             for k, w in (("pk", s["pk"]), ("sk", s["sk"]), ("ct", s["ct"]), ("ss", 32), ("ss2", 32))}
        k, e = torch.from_numpy(kc).to(dev), torch.from_numpy(ec).to(dev)
        p = eng.plan(1, 1)
        t0 = p.keypair(D.SABER, a["pk"], a["sk"], k)
        t1 = p.enc(D.SABER, a["ct"], a["ss"], a["pk"], e, after=t0)
        p.dec(D.SABER, a["ss2"], a["ct"], a["sk"], after=t1)
        p.build()
        p.launch()
        return p, a

    os.environ["DPKEM_WAIT_SPIN_LIMIT"] = "0"
    eng = D.Engine(0)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    p, a = chain(eng)
    try:
        eng.sync()
        print("NO_ERROR", flush=True)
    except D.DpkemError as exc:
        print("REPORTED", exc.code, str(exc)[:120], flush=True)
    fk = torch.from_numpy(O.prng_bytes(4244, 40 * 64).reshape(40, 64)).cuda()
    try:
        eng.keypair(D.FRODO976, fk[:1])  # eager FrodoKEM keygen: its noise chain hands blocks to a sampler wave
        eng.sync()
        print("FRODO_NO_ERROR", flush=True)
    except D.DpkemError as exc:
        print("FRODO_REPORTED", exc.code, " ".join(str(exc)[:160].split()), flush=True)
    p.close(); eng.close()
    del os.environ["DPKEM_WAIT_SPIN_LIMIT"]
    eng = D.Engine(0)
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    p, a = chain(eng)
    eng.sync()
    opk, osk = O.keypair(D.SABER, kc, threads=8)
    oct_, oss = O.enc(D.SABER, opk, ec, threads=8)
    ok = (np.array_equal(a["pk"].cpu().numpy(), opk) and np.array_equal(a["ct"].cpu().numpy(), oct_)
          and np.array_equal(a["ss"].cpu().numpy(), oss) and np.array_equal(a["ss2"].cpu().numpy(), oss))
    fpk, fsk = eng.keypair(D.FRODO976, fk)
    eng.sync()
    gpk, gsk = O.keypair(D.FRODO976, fk.cpu().numpy(), threads=8)
    ok = ok and np.array_equal(fpk.cpu().numpy(), gpk) and np.array_equal(fsk.cpu().numpy(), gsk)
    print("CLEAN_RUN_OK" if ok else "CLEAN_RUN_DIFFERS", flush=True)
    p.close(); eng.close()
""") % ROOT


def test_expired_wait_is_reported_by_sync():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("bounded waits: subprocess with DPKEM_WAIT_SPIN_LIMIT=0, then the default bound", flush=True)
    env = {k: v for k, v in os.environ.items() if k != "DPKEM_WAIT_SPIN_LIMIT"}
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stdout.split("\n")
    assert any(ln.startswith("REPORTED -3") for ln in lines), r.stdout
    assert any(ln.startswith("FRODO_REPORTED -3") and "FrodoKEM" in ln for ln in lines), r.stdout
    assert "CLEAN_RUN_OK" in lines, r.stdout
