"""Host logic of the job service: the library's scheduler (admission by
footprint in arrival order, slots per grid class, blocks with per-job
budgets, the divergence stop, completion), compiled with g++ together with a
stand-in engine that counts steps per slot instead of running a device
(tests/sched_harness.cpp), so it runs without a GPU; plus the statistics of
the generated workloads."""
import ctypes as ct
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import xpfc_cuda_amd.serving as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DONE, NONFINITE, REJECTED = 0, 1, 2


@pytest.fixture(scope="module")
def H(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("sched") / "libsched_harness.so")
    subprocess.run([gxx, "-std=c++17", "-O1", "-shared", "-fPIC", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "xpfc_cuda_amd", "csrc", "xpfc_sched.cpp"),
                    os.path.join(ROOT, "tests", "sched_harness.cpp"), "-o", out], check=True)
    L = ct.CDLL(out)
    vp, i, d, ll = ct.c_void_p, ct.c_int, ct.c_double, ct.c_longlong
    P = ct.POINTER
    for name, res, args in [
        ("h_create", vp, [i, P(i), P(i), i, d, i, i, P(i)]), ("h_destroy", None, [vp]),
        ("h_submit", i, [vp, ll, i, ll, d, ll, d]), ("h_poll", i, [vp]), ("h_idle", i, [vp]),
        ("h_diverge_at", None, [vp, ll, ll]),
        ("h_results", i, [vp, P(ll), P(i), P(ll), P(d), P(d), P(d), i]),
        ("h_log", i, [vp, P(ll), i]), ("h_stat", d, [vp, i])]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


class Sched:
    """One scheduler over `slots` (per class) with stand-in engines: every
    class in one lane, or lanes[k] for class k, whose blocks stay in flight
    for busy[l] polls."""

    def __init__(self, H, slots, budget=1e12, block=50, lookahead=1, busy_polls=2, lanes=None, busy=None):
        self.H = H
        arr = (ct.c_int * len(slots))(*slots)
        nl = 1 if lanes is None else max(lanes) + 1
        la = None if lanes is None else (ct.c_int * len(slots))(*lanes)
        busy = busy or [busy_polls] * nl
        self.h = H.h_create(len(slots), arr, la, nl, budget, block, lookahead, (ct.c_int * nl)(*busy))
        self.results = {}

    def submit(self, jid, cls, steps, fp=1.0, cells=1, arrival=0.0):
        return self.H.h_submit(self.h, jid, cls, steps, fp, cells, arrival)

    def drain(self, max_polls=100000):
        for _ in range(max_polls):
            self.H.h_poll(self.h)
            self._collect()
            if self.H.h_idle(self.h):
                return
        raise AssertionError("scheduler did not drain")

    def _collect(self):
        n = 256
        ids, st, dn = (ct.c_longlong * n)(), (ct.c_int * n)(), (ct.c_longlong * n)()
        F, ta, tf = (ct.c_double * n)(), (ct.c_double * n)(), (ct.c_double * n)()
        k = self.H.h_results(self.h, ids, st, dn, F, ta, tf, n)
        for j in range(k):
            self.results[ids[j]] = dict(status=st[j], done=dn[j], F=F[j], admitted=ta[j], finished=tf[j])

    def log(self):
        n = self.H.h_log(self.h, None, 0)
        buf = (ct.c_longlong * n)()
        self.H.h_log(self.h, buf, n)
        return [tuple(buf[i:i + 4]) for i in range(0, n, 4)]

    def stat(self, w):
        return self.H.h_stat(self.h, w)

    def close(self):
        self.H.h_destroy(self.h)


def test_every_job_runs_exactly_its_steps(H):
    s = Sched(H, [3, 3])
    steps = {i: 37 + 23 * i for i in range(9)}
    for i in range(9):
        assert s.submit(i, i % 2, steps[i]) == 0
    s.drain()
    assert {i: r["F"] for i, r in s.results.items()} == steps
    assert all(r["status"] == DONE and r["done"] == steps[i] for i, r in s.results.items())
    assert all(r["admitted"] <= r["finished"] for r in s.results.values())
    # never more than three jobs of a class at once
    running = {0: set(), 1: set()}
    for kind, a, b, c in s.log():
        if kind == 1:
            running[b].add(a)
            assert len(running[b]) <= 3
        elif kind == 3:
            running[b].discard(a)
    s.close()


def test_admission_respects_the_budget_in_arrival_order(H):
    # footprint 10 per job, budget 25: two resident at a time
    s = Sched(H, [8], budget=25.0)
    for i in range(5):
        s.submit(i, 0, 60, fp=10.0)
    s.drain()
    assert s.stat(0) == 20.0 and len(s.results) == 5
    starts = [a for kind, a, b, c in s.log() if kind == 1]
    assert starts == [0, 1, 2, 3, 4]
    s.close()


def test_a_job_that_does_not_fit_blocks_later_jobs(H):
    """Memory is granted in arrival order: the large job 1 waits for job 0 and
    the small job 2 (of another class, with room in its slots and in memory)
    waits behind it."""
    s = Sched(H, [2, 2], budget=100.0)
    s.submit(0, 0, 200, fp=60.0)
    s.submit(1, 0, 50, fp=60.0)
    s.submit(2, 1, 50, fp=10.0)
    s.drain()
    r = s.results
    assert r[1]["admitted"] >= r[0]["finished"]
    assert r[2]["admitted"] >= r[1]["admitted"]
    s.close()


def test_full_class_lets_other_classes_go(H):
    s = Sched(H, [1, 1])
    s.submit(0, 0, 200)
    s.submit(1, 0, 50)   # waits for class 0's only slot
    s.submit(2, 1, 50)   # runs at once in class 1
    s.drain()
    r = s.results
    assert r[2]["admitted"] < r[0]["finished"] <= r[1]["admitted"]
    s.close()


def test_slots_are_reused(H):
    s = Sched(H, [2])
    s.submit(0, 0, 120)
    s.submit(1, 0, 30)
    s.submit(2, 0, 75)
    s.drain()
    r = s.results
    assert [r[i]["F"] for i in range(3)] == [120, 30, 75]
    assert r[2]["admitted"] >= r[1]["finished"]
    starts = {a: c for kind, a, b, c in s.log() if kind == 1}
    assert starts[2] == starts[1]  # the freed slot
    s.close()


def test_blocks_carry_budgets_and_lengthen_only_while_no_job_waits(H):
    """Job 0 runs while job 1 waits for the only slot: blocks of 50 steps,
    then a 32-step block holding its last 30; job 1 alone gets four blocks at
    once (200 steps) and a last block of 128 holding its last 70."""
    s = Sched(H, [1], lookahead=4)
    s.submit(0, 0, 130)
    s.submit(1, 0, 470)
    s.drain()
    blocks = [(a, c) for kind, a, b, c in s.log() if kind == 2]
    assert blocks == [(50, 50), (50, 50), (32, 30), (200, 200), (200, 200), (128, 70)]
    assert {i: r["F"] for i, r in s.results.items()} == {0: 130, 1: 470}
    s.close()


def test_jobs_of_different_lengths_share_blocks(H):
    """Two running jobs with 70 and 30 steps left: one 50-step block advances
    them by 50 and 30; the shorter one finishes after one block."""
    s = Sched(H, [2])
    s.submit(0, 0, 70)
    s.submit(1, 0, 30)
    s.drain()
    blocks = [(a, b, c) for kind, a, b, c in s.log() if kind == 2]
    assert blocks[0] == (50, 2, 80)
    assert blocks[1] == (32, 1, 20)
    s.close()


def test_diverged_job_stops_early_and_frees_its_slot(H):
    s = Sched(H, [2])
    s.submit(0, 0, 300)
    s.submit(1, 0, 1000)
    s.submit(2, 0, 80)
    H.h_diverge_at(s.h, 1, 60)   # non-finite once it has taken 60 steps
    s.drain()
    r = s.results
    assert r[1]["status"] == NONFINITE and r[1]["done"] == 100  # stopped after its second block
    assert r[0]["status"] == r[2]["status"] == DONE
    assert r[0]["F"] == 300 and r[2]["F"] == 80
    assert r[2]["admitted"] >= r[1]["finished"]
    assert r[2]["finished"] < r[0]["finished"]
    s.close()


def test_jobs_that_can_never_run_are_rejected(H):
    """An unknown class, no steps, or a footprint beyond the whole budget is
    rejected at submission (it would otherwise block the queue forever);
    other jobs still run."""
    s = Sched(H, [2], budget=100.0)
    assert s.submit(0, 5, 10) != 0
    assert s.submit(1, 0, 0) != 0
    assert s.submit(2, 0, 10, fp=101.0) != 0
    assert s.submit(3, 0, 10, fp=50.0) == 0
    s.drain()
    assert s.stat(5) == 3
    assert {i: r["status"] for i, r in s.results.items()} == {0: REJECTED, 1: REJECTED, 2: REJECTED, 3: DONE}
    s.close()


def test_cell_updates_count_the_budgets(H):
    s = Sched(H, [2])
    s.submit(0, 0, 70, cells=1000)
    s.submit(1, 0, 30, cells=10)
    s.drain()
    assert s.stat(4) == 70 * 1000 + 30 * 10
    s.close()


def test_lanes_advance_independently(H):
    """Class 0 (lane 0) has blocks that stay in flight for 20 polls, class 1
    (lane 1) for one: the two 100-step jobs of lane 1 finish before lane 0's
    first block ends, and lane 1's blocks are launched while lane 0's is in
    flight."""
    s = Sched(H, [2, 2], lanes=[0, 1], busy=[20, 1])
    s.submit(0, 0, 500)
    s.submit(1, 1, 100)
    s.submit(2, 1, 100)
    s.drain()
    r = s.results
    assert {i: x["F"] for i, x in r.items()} == {0: 500, 1: 100, 2: 100}
    assert r[1]["finished"] < 20 and r[2]["finished"] < 20
    kinds = [k for k, a, b, c in s.log() if k in (2, 12)]
    first_end = kinds.index(2)
    assert kinds[:first_end + 3].count(12) >= 2 and kinds[0] == 2
    s.close()


def test_a_lane_mid_block_does_not_hold_other_lanes_admissions(H):
    """Job 1 (class 0) waits for lane 0's block to end; job 2 (class 1, lane
    1), queued behind it, is admitted at once."""
    s = Sched(H, [2, 1], lanes=[0, 1], busy=[10, 1])
    s.submit(0, 0, 100)
    H.h_poll(s.h)                 # job 0 admitted, lane 0's block in flight
    s.submit(1, 0, 50, arrival=1.0)
    s.submit(2, 1, 50, arrival=1.0)
    s.drain()
    r = s.results
    assert r[2]["admitted"] < r[1]["admitted"]
    assert r[1]["admitted"] >= 11   # after the block of 10 busy polls
    s.close()


def test_lookahead_is_per_lane(H):
    """Lookahead 4: lane 1's class is full with a job waiting, so its blocks
    stay at 50 steps while lane 0, with no job waiting, takes 200."""
    s = Sched(H, [1, 1], lanes=[0, 1], lookahead=4)
    s.submit(0, 0, 400)
    s.submit(1, 1, 100)
    s.submit(2, 1, 100)
    s.drain()
    b0 = [a for k, a, b, c in s.log() if k == 2]
    b1 = [a for k, a, b, c in s.log() if k == 12]
    assert b0[0] == 200 and b1[:2] == [50, 50]
    assert {i: x["F"] for i, x in s.results.items()} == {0: 400, 1: 100, 2: 100}
    s.close()


def test_poisson_stream_statistics():
    jobs = S.poisson_jobs(rate=20.0, duration=200.0, seed=3)
    n = len(jobs)
    assert abs(n - 4000) < 4 * math.sqrt(4000)
    gaps = np.diff([0.0] + [j.arrival for j in jobs])
    assert np.mean(gaps) == pytest.approx(0.05, rel=0.05)
    assert all(j.seed == j.jid for j in jobs)
    d2 = [j.side for j in jobs if j.dim == 2]
    d3 = [j.side for j in jobs if j.dim == 3]
    assert set(d2) == {128, 160, 192, 256, 320, 384, 512, 640, 768, 1024}
    assert set(d3) == {64, 80, 96, 128, 160, 192, 256}
    # nearest side in log2: a side takes the log-mass between the midpoints to
    # its neighbours (128 in [64, 256]: from log2(sqrt(96*128)) to log2(sqrt(128*160)))
    f = np.mean(np.array(d3) == 128)
    want = (0.5 * math.log2(160 / 96)) / 2.0
    assert f == pytest.approx(want, abs=0.03)
    # the power-of-two rounding of the first build stays available
    p2 = S.poisson_jobs(rate=20.0, duration=50.0, seed=3, rounding="pow2")
    assert {j.side for j in p2 if j.dim == 3} == {64, 128, 256}
    assert np.mean([j.side == 128 for j in p2 if j.dim == 3]) == pytest.approx(0.5, abs=0.08)
    assert all(500 <= j.steps <= 5000 for j in jobs)
    assert all(0.05 <= j.sigma <= 0.2 and 0.02 <= j.n0 <= 0.08 for j in jobs)


def test_batch_draws():
    jobs = S.batch_jobs(64, 2000, 0)
    assert len(jobs) == 64 and all(j.dim == 2 and j.steps == 2000 for j in jobs)
    assert {j.side for j in jobs} == {128, 256, 512, 1024}
    assert [j.seed for j in jobs] == list(range(64))
    assert all(0.05 <= j.sigma <= 0.2 and 0.02 <= j.n0 <= 0.08 for j in jobs)
