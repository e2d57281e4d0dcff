"""Mixed-batch and serving workloads over the engine, plan and server APIs.

mixed:   16,384 instances cut into batches of U[1, 4096] instances, each batch
         a random scheme (FrodoKEM-976-SHAKE or Saber l=3) and a random op
         (keygen : encaps : decaps = 1 : 4 : 4).  The batched run is one
         graph-captured plan holding every batch as an independent task, its
         (streams, chunks) schedule picked by the autotuner; the comparator
         runs the same batches eagerly, one call per batch, in the plain
         one-workgroup-per-KEM form.  Outputs of both runs must be identical.

serving: open-loop Poisson arrivals at a fixed rate into the serving loop
         (dpkem_server_*): 70% Saber / 30% Frodo, ops 1 : 4 : 4, keys and
         ciphertexts resident in the server's HBM store.  Latency of a
         request = completion seen by the dispatcher - its scheduled arrival
         time, so a late submission counts against the server.  The
         comparator is the same server admitting one request per launch.

Both are driven from bench.py; neither is used by the library itself.
"""
from __future__ import annotations

import statistics
import time

import numpy as np

from . import kem as K

KEYGEN, ENC, DEC = K.OP_KEYGEN, K.OP_ENC, K.OP_DEC
OP_P = (1 / 9, 4 / 9, 4 / 9)


def _pct(v, q):
    if len(v) == 0:
        return None
    return float(np.percentile(np.asarray(v, dtype=np.float64), q * 100.0))


# ---------------------------------------------------------------- mixed batches

def mixed_batches(seed: int, total: int = 16384, max_batch: int = 4096):
    """[(scheme, op, n)] with n ~ U[1, max_batch] summing to `total`."""
    rng = np.random.default_rng(seed)
    out, left = [], total
    while left > 0:
        n = int(min(left, rng.integers(1, max_batch + 1)))
        out.append((int(rng.integers(0, 2)), int(rng.choice(3, p=OP_P)), n))
        left -= n
    return out


def run_mixed(torch, eng, plain_eng, seed: int = 0x44504B454D000004, total: int = 16384, max_batch: int = 4096,
              reps: int = 3, tune_budget_s: float = 30.0, progress=None, keep_io: bool = False):
    """keep_io: also return every batch's (scheme, op, n, inputs, (plan
    outputs, plain outputs)) under "io", for checks of individual instances."""
    from . import autotune
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(seed + 1)
    batches = mixed_batches(seed, total, max_batch)
    stream = torch.cuda.current_stream(dev)
    eng.set_stream(stream.cuda_stream)
    plain_eng.set_stream(stream.cuda_stream)

    def rnd(n, w):
        return torch.from_numpy(rng.integers(0, 256, (n, w), dtype=np.uint8)).to(dev)

    # key / ciphertext pools per scheme (untimed), big enough for any batch
    pools = {}
    for sc in (K.FRODO976, K.SABER):
        need = max([n for s, op, n in batches if s == sc and op != KEYGEN], default=0)
        if need == 0:
            continue
        s = K.SIZES[sc]
        pk = torch.empty((need, s["pk"]), dtype=torch.uint8, device=dev)
        sk = torch.empty((need, s["sk"]), dtype=torch.uint8, device=dev)
        ct = torch.empty((need, s["ct"]), dtype=torch.uint8, device=dev)
        ss = torch.empty((need, s["ss"]), dtype=torch.uint8, device=dev)
        eng.keypair_into(sc, need, pk, sk, rnd(need, s["kcoins"]))
        eng.enc_into(sc, need, ct, ss, pk, rnd(need, s["ecoins"]))
        pools[sc] = (pk, sk, ct)
    torch.cuda.synchronize()

    # per-batch inputs and two output sets (batched plan / plain comparator)
    jobs = []
    for sc, op, n in batches:
        s = K.SIZES[sc]
        if op == KEYGEN:
            ins = (rnd(n, s["kcoins"]),)
            shapes = ((n, s["pk"]), (n, s["sk"]))
        elif op == ENC:
            ins = (pools[sc][0][:n], rnd(n, s["ecoins"]))
            shapes = ((n, s["ct"]), (n, s["ss"]))
        else:
            ins = (pools[sc][2][:n], pools[sc][1][:n])
            shapes = ((n, s["ss"]),)
        outs = [tuple(torch.empty(sh, dtype=torch.uint8, device=dev) for sh in shapes) for _ in range(2)]
        jobs.append((sc, op, n, ins, outs))

    def add_all(p):
        for sc, op, n, ins, outs in jobs:
            o = outs[0]
            if op == KEYGEN:
                p.keypair(sc, o[0], o[1], ins[0])
            elif op == ENC:
                p.enc(sc, o[0], o[1], ins[0], ins[1])
            else:
                p.dec(sc, o[0], ins[0], ins[1])
        return p.build()

    def build(ns, ch, ipw=1, persist=1, merge=1, threads=512):
        eng.set_tuning(K.TUNE_SABER_POLY_IPW, ipw)  # the Saber product's instances per workgroup
        eng.set_tuning(K.TUNE_FRODO_PERSIST, persist)  # persistent fused kernel in merged Frodo launches
        eng.set_tuning(K.TUNE_FRODO_THREADS, threads)  # Frodo fused kernels' workgroup size
        return add_all(eng.plan(nstreams=ns, chunks=ch, merge=bool(merge)))

    if progress:
        progress(f"mixed: {len(batches)} batches, autotuning the plan")
    # coordinate descent: streams x chunks with the merged fused kernels
    # persistent and single-pass (the two interact: a persistent grid holds
    # every CU until it ends, so it decides how other streams' work
    # overlaps), then each remaining knob's other values at the best schedule
    best, tried = autotune.tune(torch, eng, f"mixed|seed={seed}|total={total}|ipw|persist|merge|threads", build,
                                use_cache=False, budget_s=tune_budget_s,
                                candidates=[c + (1, pe, 1, 512) for pe in (1, 0) for c in autotune.DEFAULT_CANDIDATES],
                                refine=lambda b: [(b[0], b[1], 2, b[3], 1, 512), (b[0], b[1], 4, b[3], 1, 512),
                                                  (b[0], b[1], 1, b[3], 0, 512), (b[0], b[1], 1, b[3], 1, 1024)])
    plan = build(*best)
    plan.launch()
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        plan.launch()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    plan.close()
    batched_ms = statistics.median(times)

    if progress:
        progress("mixed: plain comparator, one eager call per batch")

    def plain_pass():
        for sc, op, n, ins, outs in jobs:
            o = outs[1]
            if op == KEYGEN:
                plain_eng.keypair_into(sc, n, o[0], o[1], ins[0])
            elif op == ENC:
                plain_eng.enc_into(sc, n, o[0], o[1], ins[0], ins[1])
            else:
                plain_eng.dec_into(sc, n, o[0], ins[0], ins[1])

    plain_pass()  # warm (allocates the plain form's scratch)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    plain_pass()
    b.record(stream)
    b.synchronize()
    plain_ms = a.elapsed_time(b)
    identical = all(torch.equal(x, y) for _, _, _, _, outs in jobs for x, y in zip(outs[0], outs[1]))
    mix = {"frodo976": [0, 0, 0], "saber": [0, 0, 0]}
    for sc, op, n in batches:
        mix["frodo976" if sc == K.FRODO976 else "saber"][op] += n
    return {
        "workload": f"{total} instances in {len(batches)} batches of U[1,{max_batch}], random scheme, "
                    f"keygen:enc:dec = 1:4:4",
        "batches": [[("frodo976" if sc == K.FRODO976 else "saber"), ("keygen", "enc", "dec")[op], n]
                    for sc, op, n in batches],
        "instances_by_scheme_op": mix,
        "batched_ops_per_s": round(total / (batched_ms / 1e3), 1),
        "batched_ms": round(batched_ms, 3),
        "schedule": {"nstreams": best[0], "chunks": best[1], "saber_poly_ipw": best[2], "frodo_persistent": best[3],
                     "frodo_tasks_merged": best[4], "frodo_threads": best[5],
                     "tuning_ms": {"x".join(map(str, c)): round(v, 3) for c, v in tried.items()}},
        "unmerged_ms": (round(min(v for c, v in tried.items() if c[4] == 0), 3)
                        if any(c[4] == 0 for c in tried) else None),
        "plain_ops_per_s": round(total / (plain_ms / 1e3), 1),
        "plain_ms": round(plain_ms, 3),
        "speedup": round(plain_ms / batched_ms, 2),
        "outputs_identical": bool(identical),
        **({"io": jobs} if keep_io else {}),
    }


# ---------------------------------------------------------------- serving

def _make_stream(rng, rate: float, duration_s: float, pools, saber_frac: float = 0.7):
    n = int(rng.poisson(rate * duration_s))
    gaps = rng.exponential(1e6 / rate, n)
    t = np.cumsum(gaps)
    reqs = K.make_requests(n)
    reqs["id"] = np.arange(n, dtype=np.uint64)
    reqs["scheme"] = np.where(rng.random(n) < saber_frac, K.SABER, K.FRODO976)
    reqs["op"] = rng.choice(3, size=n, p=OP_P)
    reqs["coins"] = rng.integers(0, 256, (n, 96), dtype=np.uint8)
    for sc in (K.FRODO976, K.SABER):
        keys, cts = pools[sc]
        m = reqs["scheme"] == sc
        pick = rng.integers(0, len(keys), int(m.sum()))
        reqs["key"][m] = keys[pick]
        reqs["ct"][m] = cts[pick]
    return reqs, t


def _populate(server, rng, per_scheme):
    """Keygen `per_scheme[sc]` keys and one ciphertext under each; returns
    {scheme: (key handles, ct handles)} with ct i encapsulated under key i."""
    pools = {}
    rid = 1 << 40
    for sc in (K.FRODO976, K.SABER):
        m = per_scheme[sc]
        r = K.make_requests(m)
        r["id"] = rid + np.arange(m)
        r["scheme"], r["op"] = sc, KEYGEN
        r["coins"] = rng.integers(0, 256, (m, 96), dtype=np.uint8)
        server.submit(r)
        server.drain()
        c = server.poll(m + 16)
        if len(c) != m or (c["status"] != 0).any():
            raise RuntimeError("key population failed")
        keys = np.sort(c["key"]).astype(np.uint32)
        r = K.make_requests(m)
        r["id"] = rid + m + np.arange(m)
        r["scheme"], r["op"] = sc, ENC
        r["key"] = keys
        r["coins"] = rng.integers(0, 256, (m, 96), dtype=np.uint8)
        server.submit(r)
        server.drain()
        c = server.poll(m + 16)
        if len(c) != m or (c["status"] != 0).any():
            raise RuntimeError("ciphertext population failed")
        order = np.argsort(c["id"])
        cts = c["ct"][order].astype(np.uint32)
        pools[sc] = (keys, cts)
        # one decaps batch, so that every kernel has run once before timing
        r = K.make_requests(m)
        r["id"] = rid + 2 * m + np.arange(m)
        r["scheme"], r["op"] = sc, DEC
        r["key"], r["ct"] = keys, cts
        server.submit(r)
        server.drain()
        c = server.poll(m + 16)
        if len(c) != m or (c["status"] != 0).any():
            raise RuntimeError("warm-up decapsulation failed")
        rid += 3 * m
    return pools


def drive(server, reqs, t_offsets_us, drain_cap_s: float = 2.0):
    """Submit `reqs` open-loop at their scheduled times; collect completions.
    Requests still pending `drain_cap_s` after the last arrival are cancelled."""
    n = len(reqs)
    t0 = K.now_us() + 2000.0
    arr = t0 + t_offsets_us
    reqs["t_arrival_us"] = arr
    done, got, i = [], 0, 0
    deadline = (arr[-1] if n else t0) + drain_cap_s * 1e6
    while got < n:
        now = K.now_us()
        if i < n:
            j = int(np.searchsorted(arr, now, side="right"))
            if j > i:
                rc = server.submit(reqs[i:j])
                if rc:
                    raise K.DpkemError(rc, "server_submit")
                i = j
        c = server.poll(16384)
        if len(c):
            done.append(c)
            got += len(c)
        if now > deadline and i == n:
            server.cancel()
            server.drain()
            c = server.poll(n + 16)
            if len(c):
                done.append(c)
                got += len(c)
            break
        if i < n:
            wait = (arr[i] - K.now_us()) * 1e-6
            if wait > 2e-4:
                time.sleep(min(wait, 1e-3) - 1e-4)
        else:
            time.sleep(5e-5)
    comp = np.concatenate(done) if done else np.zeros(0, K.COMP_DTYPE)
    return comp, t0


def summarize(comp, t0, n_offered, rate):
    ok = comp[comp["status"] == 0]
    lat_ms = (ok["t_done_us"] - ok["t_submit_us"]) / 1e3
    span_s = (ok["t_done_us"].max() - t0) / 1e6 if len(ok) else float("nan")
    out = {
        "offered_req_per_s": rate,
        "offered": int(n_offered),
        "completed": int(len(ok)),
        "cancelled": int((comp["status"] == -7).sum()),
        "errors": int(((comp["status"] != 0) & (comp["status"] != -7)).sum()),
        "ops_per_s": round(len(ok) / span_s, 1) if len(ok) else 0.0,
        "p50_ms": _pct(lat_ms, 0.5),
        "p99_ms": _pct(lat_ms, 0.99),
    }
    wait_ms = (ok["t_admit_us"] - ok["t_submit_us"]) / 1e3
    service_ms = (ok["t_done_us"] - ok["t_admit_us"]) / 1e3
    out["queue_p50_ms"], out["service_p50_ms"] = _pct(wait_ms, 0.5), _pct(service_ms, 0.5)
    for sc, name in ((K.SABER, "saber"), (K.FRODO976, "frodo976")):
        m = ok["scheme"] == sc
        out[f"{name}_p50_ms"] = _pct(lat_ms[m], 0.5)
        out[f"{name}_p99_ms"] = _pct(lat_ms[m], 0.99)
        out[f"{name}_queue_p50_ms"] = _pct(wait_ms[m], 0.5)
        out[f"{name}_service_p50_ms"] = _pct(service_ms[m], 0.5)
    for k, v in list(out.items()):
        if isinstance(v, float) and k.endswith("_ms"):
            out[k] = round(v, 3)
    return out


def release_created(server, comp):
    for c in comp[comp["status"] == 0]:
        if c["op"] == KEYGEN:
            server.release(int(c["scheme"]), 0, int(c["key"]))
        elif c["op"] == ENC:
            server.release(int(c["scheme"]), 1, int(c["ct"]))


def run_serving(rates=(10_000, 50_000, 100_000, 200_000, 500_000), duration_s: float = 0.5, one_per_launch=False,
                seed: int = 0x44504B454D000005, hbm_budget: int = 32 << 30, lanes: int = 4, drain_cap_s: float = 2.0,
                progress=None, saber_frac: float = 0.7, reserve_lanes=None, group_by_key: bool = True):
    rng = np.random.default_rng(seed)
    peak = max(rates) * duration_s * 1.3 + 1024
    key_slots = (int(peak * 0.3 / 9) + 1024, int(peak * 0.7 / 9) + 4096)
    ct_slots = (int(peak * 0.3 * 4 / 9) + 1024, int(peak * 0.7 * 4 / 9) + 4096)
    # batched: one lane reserved per scheme, the rest shared
    server = K.Server(0, lanes=1 if one_per_launch else lanes, hbm_budget=hbm_budget, key_slots=key_slots,
                      ct_slots=ct_slots, max_batch=(1, 1) if one_per_launch else (0, 0),
                      reserve_lanes=reserve_lanes if reserve_lanes is not None
                      else (0, 0) if one_per_launch or lanes < 3 else (1, 1),
                      group_by_key=group_by_key and not one_per_launch)
    try:
        pools = _populate(server, rng, {K.FRODO976: 64, K.SABER: 256})
        rows = []
        for rate in rates:
            reqs, t = _make_stream(rng, rate, duration_s, pools, saber_frac)
            if progress:
                progress(f"serving{' (one request per launch)' if one_per_launch else ''}: "
                         f"{rate} req/s, {len(reqs)} requests")
            comp, t0 = drive(server, reqs, t, drain_cap_s)
            rows.append(summarize(comp, t0, len(reqs), rate))
            release_created(server, comp)
        st = server.stats()
    finally:
        server.close()
    return {"rows": rows, "server": st}


def _main(argv=None):
    """`python -m dpcrypto_amd.workloads serving [--rates ...]`: one serving
    measurement in this (fresh) process, printed as one JSON line."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["serving"])
    ap.add_argument("--rates", default="10000,50000,100000,200000,500000")
    ap.add_argument("--seconds", type=float, default=0.5)
    ap.add_argument("--one-per-launch", action="store_true")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--no-key-grouping", action="store_true")
    a = ap.parse_args(argv)

    def progress(msg):
        print(f"[serving] {msg}", file=sys.stderr, flush=True)

    r = run_serving(tuple(int(x) for x in a.rates.split(",")), a.seconds, one_per_launch=a.one_per_launch,
                    lanes=a.lanes, progress=progress, group_by_key=not a.no_key_grouping)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    _main()
