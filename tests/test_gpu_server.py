"""Serving loop (dpkem_server_*): continuous batching over a device-resident store.

Each result must be the same when its request runs alone and when it runs
inside a mixed stream of other requests (both schemes, all ops, several
in-flight lanes), and equal to the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRODO, SABER = 0, 1
KEYGEN, ENC, DEC = 0, 1, 2


@pytest.fixture(scope="module")
def server():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dpcrypto_amd as D
    s = D.Server(0, lanes=4, hbm_budget=4 << 30, key_slots=(64, 512), ct_slots=(256, 2048))
    yield s
    s.close()


def _run(server, reqs, timeout_s=60.0):
    import time
    import dpcrypto_amd as D
    assert server.submit(reqs) == 0
    server.drain()
    out = server.poll(len(reqs) + 16)
    assert len(out) == len(reqs)
    return {int(c["id"]): c for c in out}


def _req(rid, scheme, op, coins=None, key=0, ct=0):
    import dpcrypto_amd as D
    r = D.make_requests(1)[0]
    r["id"], r["scheme"], r["op"], r["key"], r["ct"] = rid, scheme, op, key, ct
    if coins is not None:
        r["coins"][: len(coins)] = coins
    return r


def test_alone_and_in_mix_are_identical(server, oracle):
    import dpcrypto_amd as D
    kc = {SABER: oracle.prng_bytes(5001, 96), FRODO: oracle.prng_bytes(5002, 64)}
    ec = {SABER: oracle.prng_bytes(5003, 32), FRODO: oracle.prng_bytes(5004, 24)}
    alone = {}
    for sc in (SABER, FRODO):
        # keygen alone
        c = _run(server, np.array([_req(1, sc, KEYGEN, kc[sc])], D.REQ_DTYPE))[1]
        assert c["status"] == 0
        key = int(c["key"])
        pk, sk = server.read(sc, 0, key), server.read(sc, 1, key)
        opk, osk = oracle.keypair(sc, kc[sc][None], threads=2)
        assert np.array_equal(pk, opk[0]) and np.array_equal(sk, osk[0])
        # encaps alone
        c = _run(server, np.array([_req(2, sc, ENC, ec[sc], key=key)], D.REQ_DTYPE))[2]
        assert c["status"] == 0
        ctid = int(c["ct"])
        oct_, oss = oracle.enc(sc, opk, ec[sc][None], threads=2)
        assert np.array_equal(server.read(sc, 2, ctid), oct_[0])
        nss = D.SIZES[sc]["ss"]
        assert np.array_equal(c["ss"][:nss], oss[0])
        # decaps alone
        c = _run(server, np.array([_req(3, sc, DEC, key=key, ct=ctid)], D.REQ_DTYPE))[3]
        assert np.array_equal(c["ss"][:nss], oss[0])
        alone[sc] = (key, ctid, pk, sk, oct_[0], oss[0])

    # the same three requests per scheme inside a mixed stream
    rng = np.random.default_rng(9)
    reqs = []
    rid = 100
    for _ in range(600):
        sc = SABER if rng.random() < 0.7 else FRODO
        op = rng.choice([KEYGEN, ENC, DEC], p=[1 / 9, 4 / 9, 4 / 9])
        key, ctid = alone[sc][0], alone[sc][1]
        w = 96 if op == KEYGEN else 32
        reqs.append(_req(rid, sc, op, rng.integers(0, 256, w, dtype=np.uint8), key=key, ct=ctid))
        rid += 1
    probes = {}
    for sc in (SABER, FRODO):
        key, ctid = alone[sc][0], alone[sc][1]
        for op, coins in ((KEYGEN, kc[sc]), (ENC, ec[sc]), (DEC, None)):
            probes[rid] = (sc, op)
            reqs.insert(int(rng.integers(0, len(reqs))), _req(rid, sc, op, coins, key=key, ct=ctid))
            rid += 1
    done = _run(server, np.array(reqs, D.REQ_DTYPE))
    assert all(c["status"] == 0 for c in done.values())
    for pid, (sc, op) in probes.items():
        c = done[pid]
        key, ctid, pk, sk, oct0, oss0 = alone[sc]
        nss = D.SIZES[sc]["ss"]
        if op == KEYGEN:
            assert np.array_equal(server.read(sc, 0, int(c["key"])), pk)
            assert np.array_equal(server.read(sc, 1, int(c["key"])), sk)
        elif op == ENC:
            assert np.array_equal(server.read(sc, 2, int(c["ct"])), oct0)
            assert np.array_equal(c["ss"][:nss], oss0)
        else:
            assert np.array_equal(c["ss"][:nss], oss0)
    st = server.stats()
    assert st["batches"] >= 6 and st["max_batch_seen"] > 1  # requests were batched together
    # release everything created in the mix
    for c in done.values():
        if c["op"] == KEYGEN:
            server.release(int(c["scheme"]), 0, int(c["key"]))
        elif c["op"] == ENC:
            server.release(int(c["scheme"]), 1, int(c["ct"]))


def test_stale_handles_are_rejected(server):
    import dpcrypto_amd as D
    c = _run(server, np.array([_req(7, SABER, DEC, key=63, ct=255)], D.REQ_DTYPE))[7]
    assert c["status"] == -6  # ENOTFOUND: handle not allocated


def test_frodo_requests_grouped_by_key_equal_the_oracle(oracle):
    """Interleaved encaps / decaps for three Frodo keys: the server orders each
    batch by key and expands A once per same-key group; every result must
    still equal the oracle for its own key and coins."""
    import dpcrypto_amd as D
    srv = D.Server(0, lanes=2, hbm_budget=4 << 30, key_slots=(8, 8), ct_slots=(64, 8), group_by_key=True)
    try:
        keys, pks, sks = [], [], []
        for k in range(3):
            kc = oracle.prng_bytes(6000 + k, 64)
            c = _run(srv, np.array([_req(10 + k, FRODO, KEYGEN, kc)], D.REQ_DTYPE))[10 + k]
            assert c["status"] == 0
            keys.append(int(c["key"]))
            pk, sk = oracle.keypair(FRODO, kc[None], threads=1)
            pks.append(pk[0]); sks.append(sk[0])
        rng = np.random.default_rng(5)
        pick = rng.integers(0, 3, 22)
        coins = rng.integers(0, 256, (22, 24), dtype=np.uint8)
        reqs = np.array([_req(100 + i, FRODO, ENC, coins[i], key=keys[pick[i]]) for i in range(22)], D.REQ_DTYPE)
        done = _run(srv, reqs)
        oct_, oss = oracle.enc(FRODO, np.stack([pks[p] for p in pick]), coins, threads=8)
        cts = []
        for i in range(22):
            c = done[100 + i]
            assert c["status"] == 0
            assert np.array_equal(srv.read(FRODO, 2, int(c["ct"])), oct_[i])
            assert np.array_equal(c["ss"][:24], oss[i])
            cts.append(int(c["ct"]))
        reqs = np.array([_req(200 + i, FRODO, DEC, key=keys[pick[i]], ct=cts[i]) for i in range(22)], D.REQ_DTYPE)
        done = _run(srv, reqs)
        for i in range(22):
            assert np.array_equal(done[200 + i]["ss"][:24], oss[i])
        st = srv.stats()
        assert st["frodo_grouped_requests"] >= 44 and st["frodo_key_groups"] < st["frodo_grouped_requests"]
    finally:
        srv.close()


def test_failed_batches_complete_with_edevice_and_release_their_slots(oracle, monkeypatch):
    """A batch whose launch fails, or whose completion event reports an error,
    completes every request with DPKEM_EDEVICE and hands back the key slots
    its admission took; the server keeps serving.  The faults are injected on
    the host (DPKEM_SERVER_INJECT_FAULTS: the first two launches fail before
    any HIP call, the next batch's event reads as failed), so every kernel
    that does run exits normally."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import dpcrypto_amd as D
    monkeypatch.setenv("DPKEM_SERVER_INJECT_FAULTS", "2,1")
    s = D.Server(0, lanes=1, hbm_budget=2 << 30, key_slots=(8, 8), ct_slots=(8, 8), max_batch=(4, 4),
                 group_by_key=False)
    monkeypatch.delenv("DPKEM_SERVER_INJECT_FAULTS")
    try:
        kc = oracle.prng_bytes(6001, 16 * 96).reshape(16, 96)
        reqs = np.array([_req(i, SABER, KEYGEN, kc[i]) for i in range(16)], D.REQ_DTYPE)
        out = _run(s, reqs)
        st = np.array([out[i]["status"] for i in range(16)])
        # three failed batches of four, then one that succeeds
        assert (st == -3).sum() == 12 and (st == 0).sum() == 4, st.tolist()
        assert s.stats()["failed_batches"] == 3
        ok = [i for i in range(16) if st[i] == 0]
        opk, osk = oracle.keypair(SABER, kc[ok], threads=2)
        for j, i in enumerate(ok):
            assert np.array_equal(s.read(SABER, 0, int(out[i]["key"])), opk[j])
        # the failed batches returned their slots: 4 more keys fit into the 8 slots, the 9th does not
        more = np.array([_req(100 + i, SABER, KEYGEN, kc[i]) for i in range(5)], D.REQ_DTYPE)
        out2 = _run(s, more)
        st2 = sorted(int(out2[100 + i]["status"]) for i in range(5))
        assert st2 == [-4, 0, 0, 0, 0], st2
    finally:
        s.close()
