"""Graph-captured plans (dpkem_plan_*): results equal the oracle on every replay.

A plan captures chained keygen -> encaps -> decaps tasks (split into instance
chunks over several streams) or a mix of independent Frodo / Saber tasks
into one HIP graph.  Each test replays the graph several times, clearing the
outputs and running unrelated eager work between replays, and checks every
replay byte for byte against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRODO, SABER = 0, 1


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
    yield e
    e.close()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


FRODO_AES, LIGHTSABER, FIRESABER = 2, 3, 4


@pytest.mark.parametrize("scheme,n,nstreams,chunks,share,pipe", [
    (SABER, 300, 4, 4, False, True), (SABER, 300, 4, 4, False, False), (SABER, 300, 4, 4, True, True),
    (SABER, 97, 2, 3, True, True), (SABER, 97, 2, 3, False, True), (LIGHTSABER, 70, 2, 2, False, True),
    (FIRESABER, 45, 3, 2, False, True), (FRODO, 10, 3, 3, False, True), (FRODO, 10, 3, 3, False, False),
    (FRODO, 37, 2, 1, False, True), (FRODO_AES, 6, 2, 2, False, True),
    # tiny and ragged batches: 32-instance hash groups and 16-instance x2
    # groups with one or a few live instances, clamped loads past n
    (SABER, 1, 1, 1, False, True), (SABER, 17, 1, 1, False, True), (SABER, 50, 1, 3, False, True),
    (LIGHTSABER, 2, 1, 1, False, True), (FIRESABER, 19, 1, 1, False, True), (FRODO, 1, 1, 1, False, True),
    (FRODO, 3, 1, 2, False, True)])
def test_chained_plan_matches_oracle_on_every_replay(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, share,
                                                     pipe):
    _chain_replays(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, share, pipe, True)


@pytest.mark.parametrize("scheme,n,nstreams,chunks,share,ipw", [
    (SABER, 97, 2, 3, False, 4), (SABER, 50, 1, 1, False, 2), (SABER, 45, 1, 2, True, 4),
    (LIGHTSABER, 33, 1, 1, False, 4), (FIRESABER, 19, 1, 1, False, 2), (FIRESABER, 70, 2, 2, False, 4)])
def test_chained_plan_with_product_ipw_matches_oracle(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, share,
                                                      ipw):
    """The Saber product's instances-per-workgroup knob (2 or 4 waves per
    workgroup, ragged tails) leaves every byte unchanged; plans take the
    engine's setting when they are created."""
    import dpcrypto_amd as D
    eng.set_tuning(D.TUNE_SABER_POLY_IPW, ipw)
    try:
        _chain_replays(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, share, True, True)
    finally:
        eng.set_tuning(D.TUNE_SABER_POLY_IPW, 1)


@pytest.mark.parametrize("scheme,n,nstreams,chunks", [(SABER, 97, 2, 3), (LIGHTSABER, 33, 1, 1), (FRODO, 10, 3, 3)])
def test_plan_without_graph_matches_oracle(eng, oracle, torch_cuda, scheme, n, nstreams, chunks):
    """DPKEM_PLAN_NO_GRAPH issues the same launches directly on every
    dpkem_plan_launch (used to time a plan's kernels); the engine's profiler
    then sees every kernel of the pipelined chain."""
    _chain_replays(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, False, True, False)


def _chain_replays(eng, oracle, torch_cuda, scheme, n, nstreams, chunks, share, pipe, graph):
    import dpcrypto_amd as D
    torch = torch_cuda
    s = D.SIZES[scheme]
    kc = oracle.prng_bytes(1000 + n, n * s["kcoins"]).reshape(n, -1)
    ec = oracle.prng_bytes(2000 + n, n * s["ecoins"]).reshape(n, -1)
    opk, osk = oracle.keypair(scheme, kc, threads=4)
    oct_, oss = oracle.enc(scheme, opk, ec, threads=4)
    dkc, dec_ = _dev(torch, kc), _dev(torch, ec)
    pk = torch.empty((n, s["pk"]), dtype=torch.uint8, device="cuda")
    sk = torch.empty((n, s["sk"]), dtype=torch.uint8, device="cuda")
    ct = torch.empty((n, s["ct"]), dtype=torch.uint8, device="cuda")
    ss = torch.empty((n, s["ss"]), dtype=torch.uint8, device="cuda")
    ss2 = torch.empty((n, s["ss"]), dtype=torch.uint8, device="cuda")
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    print(f"plan scheme={scheme} n={n} {nstreams}x{chunks} share={share} pipe={pipe} graph={graph}", flush=True)
    plan = eng.plan(nstreams=nstreams, chunks=chunks, share_matrix=share, pipeline=pipe, graph=graph)
    t0 = plan.keypair(scheme, pk, sk, dkc)
    t1 = plan.enc(scheme, ct, ss, pk, dec_, after=t0)
    plan.dec(scheme, ss2, ct, sk, after=t1)
    plan.build()
    # a keygen -> enc -> dec chain runs as one launch sequence per chunk
    # (pipelined, or Saber with matrix sharing); otherwise three dependent tasks
    per_chunk = 1 if (share and scheme == SABER) or pipe else 3
    assert plan.info()["subtasks"] == per_chunk * min(chunks, n)
    for rep in range(3):
        for t in (pk, sk, ct, ss, ss2):
            t.zero_()
        scratch = torch.randn(1 << 20, device="cuda").sum()  # unrelated eager work
        torch.cuda.synchronize()
        if not graph:
            eng.set_profiling(rep == 1)
        plan.launch()
        torch.cuda.synchronize()
        if not graph and rep == 1:
            eng.set_profiling(False)
            names = set(eng.kernel_stats(reset=True))
            expect = "sab_pipe_x0" if scheme != FRODO else "frodo_pipe_x1"
            assert any(k.startswith(expect) for k in names), names
        assert float(scratch) == float(scratch)
        assert np.array_equal(pk.cpu().numpy(), opk), f"pk replay {rep}"
        assert np.array_equal(sk.cpu().numpy(), osk), f"sk replay {rep}"
        assert np.array_equal(ct.cpu().numpy(), oct_), f"ct replay {rep}"
        assert np.array_equal(ss.cpu().numpy(), oss), f"ss replay {rep}"
        assert np.array_equal(ss2.cpu().numpy(), oss), f"dec ss replay {rep}"
    plan.close()


def test_mixed_independent_plan(eng, oracle, torch_cuda):
    """Independent tasks of both schemes and all three ops in one graph."""
    import dpcrypto_amd as D
    torch = torch_cuda
    # key material produced eagerly, then consumed by the mixed plan
    n_s, n_f = 130, 4
    skc = oracle.prng_bytes(31, n_s * 96).reshape(n_s, 96)
    fkc = oracle.prng_bytes(32, n_f * 64).reshape(n_f, 64)
    spk, ssk = oracle.keypair(SABER, skc)
    fpk, fsk = oracle.keypair(FRODO, fkc, threads=4)
    sm = oracle.prng_bytes(33, n_s * 32).reshape(n_s, 32)
    fm = oracle.prng_bytes(34, n_f * 24).reshape(n_f, 24)
    sct, sss = oracle.enc(SABER, spk, sm)
    fct, fss = oracle.enc(FRODO, fpk, fm, threads=4)
    d = {k: _dev(torch, v) for k, v in dict(skc=skc, fkc=fkc, spk=spk, fpk=fpk, ssk=ssk, fsk=fsk, sm=sm, fm=fm,
                                             sct=sct, fct=fct).items()}
    out = {
        "spk": torch.empty_like(d["spk"]), "ssk": torch.empty_like(d["ssk"]),
        "fpk": torch.empty_like(d["fpk"]), "fsk": torch.empty_like(d["fsk"]),
        "sct": torch.empty_like(d["sct"]), "sss": torch.empty((n_s, 32), dtype=torch.uint8, device="cuda"),
        "fct": torch.empty_like(d["fct"]), "fss": torch.empty((n_f, 24), dtype=torch.uint8, device="cuda"),
        "sdss": torch.empty((n_s, 32), dtype=torch.uint8, device="cuda"),
        "fdss": torch.empty((n_f, 24), dtype=torch.uint8, device="cuda"),
    }
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    plan = eng.plan(nstreams=4, chunks=2)
    plan.keypair(SABER, out["spk"], out["ssk"], d["skc"])
    plan.keypair(FRODO, out["fpk"], out["fsk"], d["fkc"])
    plan.enc(SABER, out["sct"], out["sss"], d["spk"], d["sm"])
    plan.enc(FRODO, out["fct"], out["fss"], d["fpk"], d["fm"])
    plan.dec(SABER, out["sdss"], d["sct"], d["ssk"])
    plan.dec(FRODO, out["fdss"], d["fct"], d["fsk"])
    plan.build()
    for rep in range(2):
        for t in out.values():
            t.zero_()
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        exp = dict(spk=spk, ssk=ssk, fpk=fpk, fsk=fsk, sct=sct, sss=sss, fct=fct, fss=fss, sdss=sss, fdss=fss)
        for k, v in exp.items():
            assert np.array_equal(out[k].cpu().numpy(), v), f"{k} replay {rep}"
    plan.close()


FRODO640, FRODO1344 = 5, 6


@pytest.mark.parametrize("scheme,sizes,nstreams,persist,threads", [
    (FRODO, (1, 31, 33, 100, 2), 2, 1, 512), (FRODO, (40, 7), 1, 0, 512), (FRODO, (9, 70), 2, 1, 1024),
    (FRODO_AES, (3, 9), 2, 1, 512), (FRODO_AES, (4, 2), 1, 1, 1024), (FRODO640, (5, 33), 1, 1, 0),
    (FRODO1344, (2, 7), 1, 1, 0)])
def test_independent_frodo_tasks_merge_into_one_launch_sequence(eng, oracle, torch_cuda, scheme, sizes, nstreams,
                                                                 persist, threads):
    """Independent Frodo tasks of one op (uneven sizes, separate arrays) run
    as one multi-task launch sequence with a persistent fused kernel; every
    instance of every task equals the oracle on each of two replays, and the
    same plan built with DPKEM_PLAN_NO_MERGE (one sequence per task) agrees."""
    import dpcrypto_amd as D
    torch = torch_cuda
    s = D.SIZES[scheme]
    print(f"merged Frodo tasks: scheme {scheme}, sizes {sizes}, persistent {persist}, threads {threads}", flush=True)
    eng.set_tuning(D.TUNE_FRODO_PERSIST, persist)
    eng.set_tuning(D.TUNE_FRODO_THREADS, threads)
    kcs = [oracle.prng_bytes(7100 + 10 * scheme + k, n * s["kcoins"]).reshape(n, -1) for k, n in enumerate(sizes)]
    ecs = [oracle.prng_bytes(7200 + 10 * scheme + k, n * s["ecoins"]).reshape(n, -1) for k, n in enumerate(sizes)]
    ref = [oracle.keypair(scheme, kc, threads=8) for kc in kcs]
    refe = [oracle.enc(scheme, pk, ec, threads=8) for (pk, _), ec in zip(ref, ecs)]
    # decaps inputs: the oracle's ciphertexts, one tampered instance per task
    cts = []
    for (ct, _), n in zip(refe, sizes):
        c = ct.copy()
        c[n - 1, 11] ^= 4
        cts.append(c)
    refd = [oracle.dec(scheme, c, sk, threads=8) for c, (_, sk) in zip(cts, ref)]

    def run(merge):
        p = eng.plan(nstreams=nstreams, chunks=1, merge=merge)
        outs = []
        for k, n in enumerate(sizes):
            pk = torch.empty((n, s["pk"]), dtype=torch.uint8, device="cuda")
            sk = torch.empty((n, s["sk"]), dtype=torch.uint8, device="cuda")
            ct = torch.empty((n, s["ct"]), dtype=torch.uint8, device="cuda")
            ss = torch.empty((n, s["ss"]), dtype=torch.uint8, device="cuda")
            ss2 = torch.empty((n, s["ss"]), dtype=torch.uint8, device="cuda")
            kc, ec, rpk, rct, rsk = _dev(torch, kcs[k]), _dev(torch, ecs[k]), _dev(torch, ref[k][0]), \
                _dev(torch, cts[k]), _dev(torch, ref[k][1])
            p.keypair(scheme, pk, sk, kc)
            p.enc(scheme, ct, ss, rpk, ec)
            p.dec(scheme, ss2, rct, rsk)
            outs.append((pk, sk, ct, ss, ss2, kc, ec, rpk, rct, rsk))
        p.build()
        return p, outs

    for merge in (True, False):
        p, outs = run(merge)
        try:
            info = p.info()
            # merged: one sub-task per op; unmerged: one per task and op
            assert info["subtasks"] == (3 if merge else 3 * len(sizes)), info
            for rep in range(2):
                for o in outs:
                    for t in o[:5]:
                        t.zero_()
                p.launch()
                eng.sync()
                for k, o in enumerate(outs):
                    what = f"merge={merge} replay {rep} task {k} (n={sizes[k]})"
                    assert np.array_equal(o[0].cpu().numpy(), ref[k][0]), f"{what}: pk"
                    assert np.array_equal(o[1].cpu().numpy(), ref[k][1]), f"{what}: sk"
                    assert np.array_equal(o[2].cpu().numpy(), refe[k][0]), f"{what}: ct"
                    assert np.array_equal(o[3].cpu().numpy(), refe[k][1]), f"{what}: ss"
                    assert np.array_equal(o[4].cpu().numpy(), refd[k]), f"{what}: decaps ss"
                eng.keypair(SABER, _dev(torch, oracle.prng_bytes(5, 96 * 3).reshape(3, 96)))  # unrelated work
                eng.sync()
        finally:
            p.close()
    eng.set_tuning(D.TUNE_FRODO_PERSIST, 1)
    eng.set_tuning(D.TUNE_FRODO_THREADS, 0)


FRODO640, FRODO1344 = 5, 6


def _rejection_key(scheme, ct_row, sk_row):
    """The implicit-rejection shared secret of (ct, sk), from the published
    definitions: Saber family SHA3-256(z || SHA3-256(c)) with z the last 32
    bytes of sk; FrodoKEM SHAKE(c || s, len ss) with s the first len-ss bytes
    of sk (SHAKE128 for Frodo-640, SHAKE256 otherwise)."""
    import hashlib
    c, skb = ct_row.tobytes(), sk_row.tobytes()
    if scheme in (SABER, LIGHTSABER, FIRESABER):
        return hashlib.sha3_256(skb[-32:] + hashlib.sha3_256(c).digest()).digest()
    sec = {FRODO: 24, FRODO_AES: 24, FRODO640: 16, FRODO1344: 32}[scheme]
    h = hashlib.shake_128 if scheme == FRODO640 else hashlib.shake_256
    return h(c + skb[:sec]).digest(sec)


@pytest.mark.parametrize("scheme,n,nstreams,chunks", [
    (SABER, 97, 2, 3), (SABER, 17, 1, 1), (LIGHTSABER, 33, 1, 1), (FIRESABER, 20, 1, 2),
    (FRODO, 10, 2, 2), (FRODO640, 5, 1, 1), (FRODO1344, 3, 1, 1), (FRODO_AES, 3, 1, 1)])
def test_forced_rejection_in_pipelined_chain(eng, oracle, torch_cuda, scheme, n, nstreams, chunks):
    """DPKEM_PLAN_FORCE_REJECT: every decapsulation of a pipelined chain takes
    the implicit-rejection branch of its constant-time select (a chain never
    sees a forged ciphertext, so this branch is otherwise unreachable there).
    keygen and encaps still equal the oracle; decaps' ss equals the
    rejection key of (ct, sk) on every instance and every replay."""
    import dpcrypto_amd as D
    torch = torch_cuda
    s = D.SIZES[scheme]
    kc = oracle.prng_bytes(3000 + n, n * s["kcoins"]).reshape(n, -1)
    ec = oracle.prng_bytes(4000 + n, n * s["ecoins"]).reshape(n, -1)
    opk, osk = oracle.keypair(scheme, kc, threads=4)
    oct_, oss = oracle.enc(scheme, opk, ec, threads=4)
    want = np.stack([np.frombuffer(_rejection_key(scheme, oct_[i], osk[i]), np.uint8) for i in range(n)])
    assert not np.array_equal(want, oss)
    dkc, dec_ = _dev(torch, kc), _dev(torch, ec)
    a = {k: torch.empty((n, s[w]), dtype=torch.uint8, device="cuda")
         for k, w in (("pk", "pk"), ("sk", "sk"), ("ct", "ct"), ("ss", "ss"), ("ss2", "ss"))}
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    print(f"forced rejection: scheme={scheme} n={n} {nstreams}x{chunks}", flush=True)
    plan = eng.plan(nstreams=nstreams, chunks=chunks, force_reject=True)
    t0 = plan.keypair(scheme, a["pk"], a["sk"], dkc)
    t1 = plan.enc(scheme, a["ct"], a["ss"], a["pk"], dec_, after=t0)
    plan.dec(scheme, a["ss2"], a["ct"], a["sk"], after=t1)
    plan.build()
    assert plan.info()["subtasks"] == min(chunks, n)  # pipelined: one launch sequence per chunk
    try:
        for rep in range(2):
            for t in a.values():
                t.zero_()
            plan.launch()
            torch.cuda.synchronize()
            assert np.array_equal(a["pk"].cpu().numpy(), opk), f"pk replay {rep}"
            assert np.array_equal(a["sk"].cpu().numpy(), osk), f"sk replay {rep}"
            assert np.array_equal(a["ct"].cpu().numpy(), oct_), f"ct replay {rep}"
            assert np.array_equal(a["ss"].cpu().numpy(), oss), f"ss replay {rep}"
            assert np.array_equal(a["ss2"].cpu().numpy(), want), f"rejection ss replay {rep}"
    finally:
        plan.close()


@pytest.mark.parametrize("scheme,sizes,ipw,merge", [
    (SABER, (300, 257, 130), 1, True),
    (SABER, (300, 257, 130), 4, True),
    (SABER, (300, 257, 130), 1, False),
    (LIGHTSABER, (65, 33, 97), 2, True),
    (FIRESABER, (31, 70, 45), 4, True),
    (SABER, (0, 0, 77), 1, True),
    (SABER, (40, 0, 0), 1, True),
    (SABER, (0, 33, 0), 4, True),
    (FIRESABER, (1, 1, 1), 1, True),
])
def test_independent_saber_ops_share_one_mixed_launch_sequence(eng, oracle, torch_cuda, scheme, sizes, ipw, merge):
    """Independent keygen / encaps / decaps tasks of one Saber-family rank
    (unequal sizes, any of them absent) run as one mixed launch sequence
    (merge on) or one sequence per task (merge off); every output equals the
    oracle, decapsulations of tampered ciphertexts included (every third
    ciphertext has one bit flipped: implicit rejection)."""
    import dpcrypto_amd as D
    torch = torch_cuda
    nk, ne, nd = sizes
    s = D.SIZES[scheme]
    print(f"mixed Saber ops: scheme {scheme}, sizes {sizes}, ipw {ipw}, merge {merge}", flush=True)
    kc = oracle.prng_bytes(900 + nk, max(nk, 1) * s["kcoins"]).reshape(-1, s["kcoins"])[:nk]
    # keys for encaps and decaps from other coins
    m = max(ne, nd, 1)
    kc2 = oracle.prng_bytes(901 + m, m * s["kcoins"]).reshape(m, s["kcoins"])
    pk2, sk2 = oracle.keypair(scheme, kc2, threads=8)
    ec = oracle.prng_bytes(902 + m, m * s["ecoins"]).reshape(m, s["ecoins"])
    ct2, ss2 = oracle.enc(scheme, pk2, ec, threads=8)
    ct_t = ct2.copy()
    ct_t[::3, 7] ^= 0x10
    exp_dec = ss2.copy()
    for i in range(0, m, 3):
        exp_dec[i] = np.frombuffer(_rejection_key(scheme, ct_t[i], sk2[i]), dtype=np.uint8)
    okp, osk = oracle.keypair(scheme, kc, threads=8) if nk else (None, None)
    dev = {k: _dev(torch, v) for k, v in dict(kc=kc if nk else np.zeros((1, s["kcoins"]), np.uint8), pk2=pk2,
                                             ec=ec, ct=ct_t, sk2=sk2).items()}
    out = {"pk": torch.empty((max(nk, 1), s["pk"]), dtype=torch.uint8, device="cuda"),
           "sk": torch.empty((max(nk, 1), s["sk"]), dtype=torch.uint8, device="cuda"),
           "ct": torch.empty((m, s["ct"]), dtype=torch.uint8, device="cuda"),
           "ss": torch.empty((m, s["ss"]), dtype=torch.uint8, device="cuda"),
           "ss2": torch.empty((m, s["ss"]), dtype=torch.uint8, device="cuda")}
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    eng.set_tuning(D.TUNE_SABER_POLY_IPW, ipw)
    plan = eng.plan(nstreams=2, chunks=1, merge=merge)
    if nk:
        plan.keypair(scheme, out["pk"][:nk], out["sk"][:nk], dev["kc"][:nk])
    if ne:
        plan.enc(scheme, out["ct"][:ne], out["ss"][:ne], dev["pk2"][:ne], dev["ec"][:ne])
    if nd:
        plan.dec(scheme, out["ss2"][:nd], dev["ct"][:nd], dev["sk2"][:nd])
    plan.build()
    try:
        for rep in range(2):
            for t in out.values():
                t.zero_()
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
            if nk:
                np.testing.assert_array_equal(out["pk"][:nk].cpu().numpy(), okp, err_msg=f"pk, replay {rep}")
                np.testing.assert_array_equal(out["sk"][:nk].cpu().numpy(), osk, err_msg=f"sk, replay {rep}")
            if ne:
                np.testing.assert_array_equal(out["ct"][:ne].cpu().numpy(), ct2[:ne], err_msg=f"ct, replay {rep}")
                np.testing.assert_array_equal(out["ss"][:ne].cpu().numpy(), ss2[:ne], err_msg=f"ss, replay {rep}")
            if nd:
                np.testing.assert_array_equal(out["ss2"][:nd].cpu().numpy(), exp_dec[:nd],
                                              err_msg=f"decaps ss, replay {rep}")
    finally:
        plan.close()
        eng.set_tuning(D.TUNE_SABER_POLY_IPW, 1)
