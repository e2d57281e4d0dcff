"""Mixed-batch plans and the serving loop's open-loop driver at small sizes.

The mixed run puts random Frodo / Saber batches of every op into one
autotuned graph and compares every output byte with the plain form run on
the same inputs.  The serving runs check that every offered request
completes without error under Poisson arrivals, and that cancelling a
backlog accounts for every request exactly once.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_mixed_batches_plan_equals_plain(torch_cuda):
    import dpcrypto_amd as D
    from dpcrypto_amd import workloads as W
    eng = D.Engine(0)
    pe = D.Engine(0, form=D.FORM_PLAIN)
    try:
        r = W.run_mixed(torch_cuda, eng, pe, seed=77, total=700, max_batch=160, reps=2, tune_budget_s=10.0)
    finally:
        pe.close()
        eng.close()
    assert r["outputs_identical"]
    assert sum(b[2] for b in r["batches"]) == 700
    assert r["batched_ops_per_s"] > 0 and r["plain_ops_per_s"] > 0


def test_serving_low_rate_completes_everything(torch_cuda):
    from dpcrypto_amd import workloads as W
    r = W.run_serving(rates=(2000, 20000), duration_s=0.2, lanes=4, hbm_budget=4 << 30)
    for row in r["rows"]:
        assert row["errors"] == 0 and row["cancelled"] == 0
        assert row["completed"] == row["offered"] > 0
        assert row["p99_ms"] is not None


def test_cancel_accounts_for_every_request(torch_cuda):
    import dpcrypto_amd as D
    s = D.Server(0, lanes=1, hbm_budget=2 << 30, key_slots=(8, 8), ct_slots=(8, 8), max_batch=(1, 1))
    try:
        n = 40
        r = D.make_requests(n)
        r["id"] = np.arange(n)
        r["scheme"], r["op"] = 0, 0  # Frodo keygens, one per launch
        r["coins"] = np.random.default_rng(3).integers(0, 256, (n, 96), dtype=np.uint8)
        assert s.submit(r) == 0
        k = s.cancel()
        s.drain()
        out = s.poll(n + 8)
        assert len(out) == n
        assert sorted(out["id"].tolist()) == list(range(n))
        st = out["status"]
        assert int((st == -7).sum()) == k
        # admitted requests finished normally, or failed for want of a key slot
        assert set(st[st != -7].tolist()) <= {0, -4}
    finally:
        s.close()
