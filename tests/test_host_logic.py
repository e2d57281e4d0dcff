"""Host-side logic of the device code, compiled for the CPU.

The lane-pair Keccak round template in dpcrypto_amd/csrc/keccak2.hpp is run
in a two-lane host emulation (partner exchange = swapping the two lanes) and
compared with the oracle's permutation; the Saber Toeplitz helpers in
saber_poly.hpp (the dot2 form and the int8 MFMA tile walk, whose matrix
instructions are replayed as 16 x 16 x 64 products) are compared with the
oracle's negacyclic product; the Frodo sampler's magnitude tables
(frodo_sampler.hpp) are compared with the oracle's sampler on every draw.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CXX = shutil.which("g++") or shutil.which("clang++")
# ext_vector_type (used by the dot-product helpers) needs clang
CLANGXX = next((c for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")) if c and os.path.exists(c)), None)


def _build_and_run(src, tmp_path, oracle, cxx=None):
    exe = str(tmp_path / "prog")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run([cxx or CXX, "-O1", "-std=c++17", "-I", os.path.join(HERE, "host", "include"),
                    os.path.join(HERE, "host", src), "-o", exe,
                    f"-L{odir}", "-loracle", f"-Wl,-rpath,{odir}"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    return out.returncode, out.stdout


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_lane_pair_keccak_matches_oracle(tmp_path, oracle):
    rc, out = _build_and_run("k2_emul.cpp", tmp_path, oracle)
    assert rc == 0 and out.strip() == "OK", out


@pytest.mark.skipif(CLANGXX is None, reason="no clang++")
def test_saber_toeplitz_product_matches_oracle(tmp_path, oracle):
    rc, out = _build_and_run("toep_emul.cpp", tmp_path, oracle, cxx=CLANGXX)
    assert rc == 0 and out.strip() == "OK", out


@pytest.mark.skipif(CLANGXX is None, reason="no clang++")
def test_saber_mfma_toeplitz_tiling_matches_oracle(tmp_path, oracle):
    rc, out = _build_and_run("toep8_emul.cpp", tmp_path, oracle, cxx=CLANGXX)
    assert rc == 0 and out.strip() == "OK", out


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_frodo_sampler_tables_match_the_oracle_on_every_draw(tmp_path, oracle):
    rc, out = _build_and_run("sampler_emul.cpp", tmp_path, oracle)
    assert rc == 0 and out.strip() == "OK", out


def test_mixed_batch_sizes_sum_to_the_total():
    from dpcrypto_amd import workloads as W
    for seed in (1, 2, 3):
        b = W.mixed_batches(seed, total=16384, max_batch=4096)
        assert sum(x[2] for x in b) == 16384
        assert all(1 <= x[2] <= 4096 and x[0] in (0, 1) and x[1] in (0, 1, 2) for x in b)


class _FakeEvent:
    clock = [0.0]

    def __init__(self, enable_timing=True):
        self.t = None

    def record(self, stream=None):
        self.t = _FakeEvent.clock[0]

    def synchronize(self):
        pass

    def elapsed_time(self, other):
        return other.t - self.t


class _FakeTorch:
    class cuda:
        Event = _FakeEvent

        @staticmethod
        def get_device_name(i):
            return "fake"

        @staticmethod
        def current_stream():
            return None

        @staticmethod
        def synchronize():
            pass


class _FakePlan:
    def __init__(self, cost):
        self.cost = cost

    def launch(self):
        _FakeEvent.clock[0] += self.cost

    def close(self):
        pass


def test_autotune_sweeps_schedules_then_refines_the_knobs():
    """Coordinate descent: the (streams, chunks) sweep picks (2, 1) at the
    default knob value, then refine() proposes the knob's other values at
    that schedule, and the best of all timed candidates wins."""
    from dpcrypto_amd import autotune
    cost = {(1, 1, 1): 5.0, (2, 1, 1): 3.0, (4, 1, 1): 4.0, (2, 1, 2): 2.5, (2, 1, 4): 2.0}
    built = []

    def build(*cand):
        built.append(cand)
        return _FakePlan(cost[cand])

    best, tried = autotune.tune(_FakeTorch, None, "test", build, candidates=[(1, 1, 1), (2, 1, 1), (4, 1, 1)],
                                use_cache=False, refine=lambda b: [(b[0], b[1], 2), (b[0], b[1], 4)])
    assert best == (2, 1, 4)
    assert tried == pytest.approx(cost)
    assert built == [(1, 1, 1), (2, 1, 1), (4, 1, 1), (2, 1, 2), (2, 1, 4)]


def test_autotune_times_at_least_one_candidate_within_a_spent_budget():
    from dpcrypto_amd import autotune
    best, tried = autotune.tune(_FakeTorch, None, "test", lambda *c: _FakePlan(1.0), candidates=[(1, 1), (2, 1)],
                                use_cache=False, budget_s=-1.0)
    assert best == (1, 1) and list(tried) == [(1, 1)]


def test_bench_roofline_arithmetic():
    """bench.roofline_from: achieved = algorithmic lane-ops per launch / the
    average launch time, against the VALU peak; a serial-chain launch also
    against its issue floor (critical permutations x 24 rounds x 108 ops x 4
    clocks at 2.4 GHz)."""
    import sys
    sys.path.insert(0, ROOT)
    import bench
    n = 4096
    r = bench.roofline_from({"sab_pipe_x0": (10, 10 * 0.2), "sab_poly<kSabEnc>": (10, 10 * 0.05)}, n)
    assert r["kernel"] == "sab_pipe_x0" and r["avg_launch_ms"] == pytest.approx(0.2)
    perms, macs = bench.KERNEL_WORK["sab_pipe_x0"]
    work = n * (perms * bench.KECCAK_LANE_OPS + macs / 2.0)
    assert r["algorithmic_lane_ops_per_launch"] == pytest.approx(work)
    assert r["frac"] == pytest.approx(work / 0.2e-3 / bench.VALU_PEAK, abs=1e-4)
    assert r["share_of_step"] == pytest.approx(0.8)
    floor_s = 24 * 24 * 108 * 4 / 2.4e9
    assert r["chain_issue_floor"]["floor_ms"] == pytest.approx(floor_s * 1e3, abs=1e-4)
    assert r["chain_issue_floor"]["frac"] == pytest.approx(floor_s / 0.2e-3, abs=1e-4)
    # a throughput kernel has no chain floor
    r2 = bench.roofline_from({"frodo_sa_fused<false>": (1, 12.0)}, 8192)
    assert r2["chain_issue_floor"] is None and 0 < r2["frac"] < 1


@pytest.mark.skipif(CXX is None, reason="no host C++ compiler")
def test_plan_bookkeeping_under_asan_with_a_capture_checking_fake_runtime(tmp_path):
    """plan.cpp, roll.cpp and engine.cpp built for the CPU against a fake HIP runtime
    (tests/host/plan_harness) with AddressSanitizer and UBSan: every plan
    shape -- chains of every scheme, uneven mixes with merged Frodo groups
    and dependants, every option flag, several schedules, rebuilds, the
    plain form -- and every rolling pipeline shape builds, replays and is
    destroyed without a memory error,
    without two launches of a plan sharing scratch, and within the rules of
    stream capture (no unjoined work, no waits on events from outside the
    capture, no use of destroyed streams or events)."""
    exe = str(tmp_path / "plan_harness")
    hdir = os.path.join(HERE, "host", "plan_harness")
    csrc = os.path.join(ROOT, "dpcrypto_amd", "csrc")
    r = subprocess.run([CXX, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-D__HIP_PLATFORM_AMD__", "-I", hdir, os.path.join(hdir, "plan_harness.cpp"),
                        os.path.join(csrc, "plan.cpp"), os.path.join(csrc, "roll.cpp"), os.path.join(csrc, "engine.cpp"),
                        "-o", exe],
                       capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in r.stderr:
        pytest.skip("sanitizer runtime unavailable: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.startswith("OK"), (out.stdout + out.stderr)[-3000:]
