"""GPU parity at the benchmark's own configurations, on the paths it times.

Every test builds exactly what bench.py times (its coin generation, its plan
builder, its schedules) at the configuration's full size and compares the
GPU's bytes with the CPU oracle on the same coins:

  C1  FrodoKEM-976: 64 keypairs plus one encaps / decaps each, every byte
      (eager entry points and the pipelined plan)
  C2  Saber l=3, batch 4096: the headline's rolling pipeline (three batch
      slots; each step keygen of one slot, encaps of the previous, decaps of
      the one before; product 1, 2 and 4 instances per workgroup; and the
      overlapped six-slot form, each step's chains beside the previous
      step's products in one launch): a slot's outputs cleared when its batch
      starts, unrelated work between steps, and every output of every
      instance of every batch equal to the oracle once complete (two steps
      later, five in the overlapped form); and the chained plan (one batch's keygen -> encaps ->
      decaps inside each step, one chunk, product 1 and 4 instances per
      workgroup), three graph replays with the outputs cleared and unrelated
      work in between; every instance of every replay
  C3  FrodoKEM-976, batch 8192: the headline's rolling pipeline (default and
      1024-lane fused shapes): a slot's outputs cleared when its batch
      starts, unrelated work between steps; two steps later decaps recovers
      every shared secret and 64 instances of the slot (0 and 8191 among
      them) equal the oracle on pk, sk, ct and both ss; and the chained
      pipelined plan on the 1x1 and 8x4 schedules and with the 1024-lane
      fused shape, checked the same way
  C4  16,384 instances in batches of U[1, 4096] (random scheme, ops 1:4:4),
      one autotuned graph: equal to the plain form everywhere, and sampled
      instances of every batch equal the oracle
  C5  the serving loop at 500k req/s for 0.5 s on a 32 GiB budget: every
      request completes without error and 160 sampled completions (all
      ops, both schemes) equal the oracle

The shared-secret check of every instance mirrors the reference's own
all-instance comparison (SABER-GPU/src/cuda_kernel.cu:284-297,
FrodoKEM976-GPU/main.cu:26-32).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FRODO, SABER = 0, 1
THREADS = 16


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
    e.close()


def _eq(a, b, what, idx=None):
    a = np.asarray(a); b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shapes {a.shape} vs {b.shape}"
    bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    where = bad if idx is None else np.asarray(idx)[bad]
    assert bad.size == 0, f"{what}: {bad.size} instances differ, first {where[:8].tolist()}"


def _host(t):
    return t.cpu().numpy()


def _check_chain(oracle, scheme, a, idx):
    """pk / sk / ct / ss / decaps ss of instances `idx` of a chain step against the oracle."""
    kc, ec = _host(a["kc"])[idx], _host(a["ec"])[idx]
    opk, osk = oracle.keypair(scheme, kc, threads=THREADS)
    oct_, oss = oracle.enc(scheme, opk, ec, threads=THREADS)
    _eq(_host(a["pk"])[idx], opk, "pk", idx)
    _eq(_host(a["sk"])[idx], osk, "sk", idx)
    _eq(_host(a["ct"])[idx], oct_, "ct", idx)
    _eq(_host(a["ss"])[idx], oss, "ss", idx)
    _eq(_host(a["ss2"])[idx], oss, "decaps ss", idx)


def _clear(torch, a):
    for k in ("pk", "sk", "ct", "ss", "ss2"):
        a[k].zero_()


# ------------------------------------------------------------------ C1
def test_c1_frodo_64_keypairs_then_one_encaps_decaps_each(eng, oracle, torch_cuda):
    import bench
    n = 64
    print(f"C1: FrodoKEM-976, {n} instances", flush=True)
    kc = oracle.prng_bytes(0x44504B454D000001, n * 64).reshape(n, 64)
    mu = oracle.prng_bytes(0x44504B454D000101, n * 24).reshape(n, 24)
    opk, osk = oracle.keypair(FRODO, kc, threads=THREADS)
    oct_, oss = oracle.enc(FRODO, opk, mu, threads=THREADS)
    # eager entry points over host arrays
    pk, sk = eng.keypair(FRODO, kc)
    _eq(pk, opk, "pk"); _eq(sk, osk, "sk")
    ct, ss = eng.enc(FRODO, pk, mu)
    _eq(ct, oct_, "ct"); _eq(ss, oss, "ss")
    _eq(eng.dec(FRODO, ct, sk), oss, "decaps ss")
    # the pipelined plan over device arrays
    a = bench.chain_arrays(torch_cuda, FRODO, n, 0)
    a["kc"].copy_(torch_cuda.from_numpy(kc)); a["ec"].copy_(torch_cuda.from_numpy(mu))
    p = bench.build_chain_plan(eng, FRODO, a, 1, 1)
    p.launch()
    torch_cuda.cuda.synchronize()
    p.close()
    _check_chain(oracle, FRODO, a, np.arange(n))


# ------------------------------------------------------------------ C2
@pytest.mark.parametrize("overlap,ipw", [(False, 1), (False, 2), (False, 4), (True, 1), (True, 4)])
def test_c2_saber_rolling_steps_at_4096_every_instance(eng, oracle, torch_cuda, overlap, ipw):
    import bench
    torch = torch_cuda
    n = 4096
    S, lag = (6, 5) if overlap else (3, 2)  # slots; steps from a batch's keygen step to the one it is complete after
    print(f"C2: Saber rolling pipeline, n={n} per slot, {S} slots, ipw={ipw}", flush=True)
    slots = bench.rolling_slots(torch, SABER, n, 0x44504B454D000001, S)
    ref = []
    for sl in slots:
        kc, ec = _host(sl["kc"]), _host(sl["ec"])
        opk, osk = oracle.keypair(SABER, kc, threads=THREADS)
        oct_, oss = oracle.enc(SABER, opk, ec, threads=THREADS)
        ref.append((opk, osk, oct_, oss))
    r = bench.build_roll(eng, SABER, slots, ipw, overlap=overlap)
    junk = torch.zeros((n, 96), dtype=torch.uint8, device="cuda")
    try:
        for step in range(lag + 7):
            # a slot's outputs are cleared when its batch starts (its keygen
            # step) and must all be complete `lag` steps later
            _clear(torch, slots[step % S])
            r.step()
            torch.cuda.synchronize()
            if step >= lag:
                c = (step - lag) % S
                sl, (opk, osk, oct_, oss) = slots[c], ref[c]
                _eq(_host(sl["pk"]), opk, f"pk (step {step}, slot {c})")
                _eq(_host(sl["sk"]), osk, f"sk (step {step}, slot {c})")
                _eq(_host(sl["ct"]), oct_, f"ct (step {step}, slot {c})")
                _eq(_host(sl["ss"]), oss, f"encaps ss (step {step}, slot {c})")
                _eq(_host(sl["ss2"]), oss, f"decaps ss (step {step}, slot {c})")
            # unrelated eager work between steps, on the same engine
            eng.keypair(SABER, junk[:100])
            eng.sync()
    finally:
        r.close()


@pytest.mark.parametrize("ipw", [1, 4])
def test_c2_saber_chained_plan_at_4096_every_instance(eng, oracle, torch_cuda, ipw):
    import bench
    torch = torch_cuda
    n = 4096
    print(f"C2: Saber headline plan, n={n}, 4 streams x 1 chunk, ipw={ipw}", flush=True)
    a = bench.chain_arrays(torch, SABER, n, 0x44504B454D000001)
    p = bench.build_chain_plan(eng, SABER, a, 4, 1, ipw)
    kc, ec = _host(a["kc"]), _host(a["ec"])
    opk, osk = oracle.keypair(SABER, kc, threads=THREADS)
    oct_, oss = oracle.enc(SABER, opk, ec, threads=THREADS)
    junk = torch.zeros((n, 96), dtype=torch.uint8, device="cuda")
    try:
        for rep in range(3):
            _clear(torch, a)
            p.launch()
            torch.cuda.synchronize()
            _eq(_host(a["pk"]), opk, f"pk (replay {rep})")
            _eq(_host(a["sk"]), osk, f"sk (replay {rep})")
            _eq(_host(a["ct"]), oct_, f"ct (replay {rep})")
            _eq(_host(a["ss"]), oss, f"ss (replay {rep})")
            _eq(_host(a["ss2"]), oss, f"decaps ss (replay {rep})")
            # unrelated eager work between replays, on the same engine and scratch
            eng.keypair(SABER, junk[:100])
            eng.sync()
    finally:
        p.close()
        eng.set_tuning(1, 1)


# ------------------------------------------------------------------ C3
@pytest.mark.parametrize("threads", [0, 1024])
def test_c3_frodo_rolling_steps_at_8192(eng, oracle, torch_cuda, threads):
    import bench
    torch = torch_cuda
    n = 8192
    print(f"C3: FrodoKEM-976 rolling pipeline, n={n} per slot, fused threads {threads or 'default'}", flush=True)
    slots = bench.rolling_slots(torch, FRODO, n, 0x44504B454D000002)
    rng = np.random.default_rng(11)
    idx = np.unique(np.concatenate([[0, 1, 4095, 4096, 8191], rng.choice(n, 59, replace=False)]))
    ref = []
    for sl in slots:
        opk, osk = oracle.keypair(FRODO, _host(sl["kc"])[idx], threads=THREADS)
        oct_, oss = oracle.enc(FRODO, opk, _host(sl["ec"])[idx], threads=THREADS)
        ref.append((opk, osk, oct_, oss))
    r = bench.build_roll(eng, FRODO, slots, graph=True, threads=threads)
    junk = torch.zeros((64, 64), dtype=torch.uint8, device="cuda")
    try:
        for step in range(5):
            _clear(torch, slots[step % 3])
            r.step()
            torch.cuda.synchronize()
            if step >= 2:
                c = (step - 2) % 3
                sl, (opk, osk, oct_, oss) = slots[c], ref[c]
                assert torch.equal(sl["ss"], sl["ss2"]), f"decaps lost a shared secret (step {step}, slot {c})"
                assert bool((sl["ss"] != 0).any(dim=1).all()), f"an instance of slot {c} left its ss unwritten"
                _eq(_host(sl["pk"])[idx], opk, f"pk (step {step}, slot {c})", idx)
                _eq(_host(sl["sk"])[idx], osk, f"sk (step {step}, slot {c})", idx)
                _eq(_host(sl["ct"])[idx], oct_, f"ct (step {step}, slot {c})", idx)
                _eq(_host(sl["ss"])[idx], oss, f"encaps ss (step {step}, slot {c})", idx)
                _eq(_host(sl["ss2"])[idx], oss, f"decaps ss (step {step}, slot {c})", idx)
            # unrelated eager work between steps, on the same engine
            eng.keypair(FRODO, junk[:2])
            eng.sync()
    finally:
        r.close()
        eng.set_tuning(3, 0)  # later plans of this engine: default fused shape


@pytest.mark.parametrize("schedule", [(1, 1, 512), (8, 4, 512), (1, 1, 1024)])
def test_c3_frodo_plan_at_8192(eng, oracle, torch_cuda, schedule):
    import bench
    torch = torch_cuda
    n = 8192
    print(f"C3: FrodoKEM-976 pipelined plan, n={n}, schedule {schedule}", flush=True)
    a = bench.chain_arrays(torch, FRODO, n, 0x44504B454D000002)
    p = bench.build_chain_plan(eng, FRODO, a, *schedule)
    eng.set_tuning(3, 0)  # later plans of this engine: default fused shape
    try:
        for rep in range(2):
            _clear(torch, a)
            p.launch()
            torch.cuda.synchronize()
            assert torch.equal(a["ss"], a["ss2"]), f"decaps lost a shared secret (replay {rep})"
            assert bool((a["ss"] != 0).any(dim=1).all()), "an instance left its shared secret unwritten"
    finally:
        p.close()
    rng = np.random.default_rng(3)
    idx = np.unique(np.concatenate([[0, 1, 4095, 4096, 8191], rng.choice(n, 59, replace=False)]))
    _check_chain(oracle, FRODO, a, idx)


# ------------------------------------------------------------------ C4
def test_c4_mixed_default_configuration(oracle, torch_cuda):
    import dpcrypto_amd as D
    from dpcrypto_amd import workloads as W
    torch = torch_cuda
    print("C4: 16384 instances in batches of U[1,4096], autotuned plan vs plain", flush=True)
    eng = D.Engine(0)
    pe = D.Engine(0, form=D.FORM_PLAIN)
    try:
        r = W.run_mixed(torch, eng, pe, reps=1, tune_budget_s=5.0, keep_io=True)
    finally:
        pe.close()
        eng.close()
    assert sum(b[2] for b in r["batches"]) == 16384
    assert r["outputs_identical"]
    rng = np.random.default_rng(4)
    checked = 0
    for sc, op, n, ins, outs in r["io"]:
        k = min(n, 6 if sc == FRODO else 24)
        idx = np.unique(np.concatenate([[0, n - 1], rng.choice(n, k, replace=False)]))
        o = [_host(t)[idx] for t in outs[0]]
        x = [_host(t)[idx] for t in ins]
        what = f"batch ({'frodo' if sc == FRODO else 'saber'}, op {op}, n={n})"
        if op == D.OP_KEYGEN:
            opk, osk = oracle.keypair(sc, x[0], threads=THREADS)
            _eq(o[0], opk, f"{what} pk", idx); _eq(o[1], osk, f"{what} sk", idx)
        elif op == D.OP_ENC:
            oct_, oss = oracle.enc(sc, x[0], x[1], threads=THREADS)
            _eq(o[0], oct_, f"{what} ct", idx); _eq(o[1], oss, f"{what} ss", idx)
        else:
            _eq(o[0], oracle.dec(sc, x[0], x[1], threads=THREADS), f"{what} ss", idx)
        checked += len(idx)
    assert checked >= 100


# ------------------------------------------------------------------ C5
def test_c5_serving_500k_req_per_s_on_32GiB(oracle, torch_cuda):
    from dpcrypto_amd import kem as K
    from dpcrypto_amd import workloads as W
    rate, seconds = 500_000, 0.5
    print(f"C5: serving at {rate} req/s for {seconds} s, 32 GiB budget", flush=True)
    rng = np.random.default_rng(0x44504B454D000005)
    peak = rate * seconds * 1.3 + 1024
    key_slots = (int(peak * 0.3 / 9) + 1024, int(peak * 0.7 / 9) + 4096)
    ct_slots = (int(peak * 0.3 * 4 / 9) + 1024, int(peak * 0.7 * 4 / 9) + 4096)
    srv = K.Server(0, lanes=4, hbm_budget=32 << 30, key_slots=key_slots, ct_slots=ct_slots, reserve_lanes=(1, 1),
                   group_by_key=True)
    try:
        pools = W._populate(srv, rng, {K.FRODO976: 64, K.SABER: 256})
        reqs, t = W._make_stream(rng, rate, seconds, pools)
        comp, t0 = W.drive(srv, reqs, t, drain_cap_s=5.0)
        row = W.summarize(comp, t0, len(reqs), rate)
        print(f"C5: {row}", flush=True)
        assert row["errors"] == 0 and row["cancelled"] == 0
        assert row["completed"] == row["offered"] == len(reqs) > 100_000
        comp = comp[np.argsort(comp["id"])]
        assert (comp["id"] == np.arange(len(reqs))).all()
        # 160 sampled completions: 20 of every (scheme, op) pair, plus any 40
        pick = []
        for sc in (K.FRODO976, K.SABER):
            for op in (0, 1, 2):
                ids = np.nonzero((reqs["scheme"] == sc) & (reqs["op"] == op))[0]
                pick += rng.choice(ids, min(20, len(ids)), replace=False).tolist()
        pick += rng.choice(len(reqs), 40, replace=False).tolist()
        sizes = K.SIZES
        for i in sorted(set(pick)):
            r, c = reqs[i], comp[i]
            sc, op = int(r["scheme"]), int(r["op"])
            s = sizes[sc]
            assert c["status"] == 0 and int(c["scheme"]) == sc and int(c["op"]) == op
            if op == 0:
                opk, osk = oracle.keypair(sc, r["coins"][: s["kcoins"]][None])
                assert (srv.read(sc, 0, int(c["key"])) == opk[0]).all(), f"request {i}: pk"
                assert (srv.read(sc, 1, int(c["key"])) == osk[0]).all(), f"request {i}: sk"
            elif op == 1:
                pk = srv.read(sc, 0, int(r["key"]))
                oct_, oss = oracle.enc(sc, pk[None], r["coins"][: s["ecoins"]][None])
                assert (srv.read(sc, 2, int(c["ct"])) == oct_[0]).all(), f"request {i}: ct"
                assert (c["ss"][: s["ss"]] == oss[0]).all(), f"request {i}: ss"
            else:
                sk = srv.read(sc, 1, int(r["key"]))
                ct = srv.read(sc, 2, int(r["ct"]))
                oss = oracle.dec(sc, ct[None], sk[None])
                assert (c["ss"][: s["ss"]] == oss[0]).all(), f"request {i}: decaps ss"
    finally:
        srv.close()
