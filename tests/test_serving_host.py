"""Host logic of the job service (admission by footprint, slots, blocks,
completion) with a stand-in for the device context, so it runs without a GPU.
The stand-in advances a per-slot step counter when its queued block 'finishes'
(after a given number of polls), like a stream that is busy for a while."""
import math

import numpy as np
import pytest

import xpfc_cuda_amd.serving as S


class FakeCtx:
    log = []

    def __init__(self, params):
        self.n = len(params)
        self.params = list(params)
        self.active = list(range(self.n))
        self.steps = [0] * self.n
        self.busy_polls = 0
        self.queued = 0

    def set_active(self, inst):
        assert self.busy_polls == 0
        self.active = list(range(self.n)) if inst is None else list(inst)

    def set_params(self, s, p):
        assert self.busy_polls == 0 and s not in self.active
        self.params[s] = p

    def reset(self, s):
        self.steps[s] = 0

    def initialize_noise(self, s, amp, seed):
        assert s not in self.active
        FakeCtx.log.append(("init", s, seed))

    def step(self, n):
        assert self.busy_polls == 0 and 1 <= n <= S.BLOCK * 8
        FakeCtx.log.append(("step", n))
        for s in self.active:
            self.steps[s] += n
        self.busy_polls = 2

    def query(self):
        if self.busy_polls > 0:
            self.busy_polls -= 1
        return self.busy_polls == 0

    def probe_finite(self):
        return [True] * self.n

    def fe_calc(self, s):
        return float(self.steps[s]), 0.0

    def avg_calc(self, s):
        return self.params[s].psi0

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(S, "BatchContext", FakeCtx)
    # footprint: 1 unit per cell
    monkeypatch.setattr(S, "instance_bytes", lambda p: int(np.prod(p.n)))
    FakeCtx.log = []
    return FakeCtx


def _job(jid, side, steps, dim=2, t=0.0):
    return S.Job(jid=jid, dim=dim, side=side, sigma=0.1, n0=0.02 + 0.001 * jid, steps=steps,
                 seed=jid, arrival=t)


def test_every_job_runs_exactly_its_steps(fake):
    jobs = [_job(i, 64 if i % 2 else 128, 37 + 23 * i) for i in range(9)]
    svc = S.Service(budget_bytes=1e12, slots=lambda d, s: 3)
    rep = svc.run(jobs, realtime=False, idle_sleep=0)
    assert rep["jobs"] == 9
    for j in svc.done:
        assert j.F_ref == j.steps  # the stand-in reports the slot's step count
        assert j.mean == pytest.approx(j.n0)
        assert j.admitted <= j.finished
    # three slots per class: never more than three jobs of a class at once
    assert set(rep["classes"].values()) == {3}


def test_admission_respects_the_budget(fake):
    # 128^2 = 16384 units per job; a budget of 2.5 jobs admits two at a time
    jobs = [_job(i, 128, 60) for i in range(5)]
    svc = S.Service(budget_bytes=2.5 * 128 * 128, slots=lambda d, s: 8)
    rep = svc.run(jobs, realtime=False, idle_sleep=0)
    assert rep["jobs"] == 5
    assert svc.peak_resident == 2 * 128 * 128
    order = sorted(svc.done, key=lambda j: j.admitted)
    assert [j.jid for j in order] == [0, 1, 2, 3, 4]


def test_slots_are_reused_and_blocks_stop_at_completion(fake):
    jobs = [_job(0, 64, 120), _job(1, 64, 30), _job(2, 64, 75)]
    svc = S.Service(budget_bytes=1e12, slots=lambda d, s: 2)
    svc.run(jobs, realtime=False, idle_sleep=0)
    by = {j.jid: j for j in svc.done}
    assert by[0].F_ref == 120 and by[1].F_ref == 30 and by[2].F_ref == 75
    # job 2 waited for job 1's slot
    assert by[2].admitted >= by[1].finished
    inits = [e for e in FakeCtx.log if e[0] == "init"]
    assert [e[2] for e in inits] == [0, 1, 2]
    assert inits[2][1] == inits[1][1]  # the freed slot


def test_unsupported_grid_is_rejected(fake, monkeypatch):
    monkeypatch.setattr(S, "instance_bytes", lambda p: 0)
    svc = S.Service()
    with pytest.raises(ValueError):
        svc.submit(_job(0, 100, 10))


def test_poisson_stream_statistics():
    jobs = S.poisson_jobs(rate=20.0, duration=200.0, seed=3)
    n = len(jobs)
    assert abs(n - 4000) < 4 * math.sqrt(4000)
    gaps = np.diff([0.0] + [j.arrival for j in jobs])
    assert np.mean(gaps) == pytest.approx(0.05, rel=0.05)
    assert all(j.seed == j.jid for j in jobs)
    d2 = [j.side for j in jobs if j.dim == 2]
    d3 = [j.side for j in jobs if j.dim == 3]
    assert set(d2) <= {128, 256, 512, 1024} and set(d3) <= {64, 128, 256}
    # nearest power of two in log2: the end points get half the mass of the inner sides
    f = np.mean(np.array(d3) == 128)
    assert f == pytest.approx(0.5, abs=0.04)
    assert all(500 <= j.steps <= 5000 for j in jobs)
    assert all(0.05 <= j.sigma <= 0.2 and 0.02 <= j.n0 <= 0.08 for j in jobs)


def test_diverged_job_stops_early_and_frees_its_slot(fake, monkeypatch):
    seeds = {}

    def init(self, s, amp, seed):
        seeds[s] = seed
        FakeCtx.log.append(("init", s, seed))

    # the job with seed 1 'diverges' once it has taken 60 steps
    monkeypatch.setattr(FakeCtx, "initialize_noise", init)
    monkeypatch.setattr(FakeCtx, "probe_finite",
                        lambda self: [not (seeds.get(s) == 1 and self.steps[s] >= 60)
                                      for s in range(self.n)])
    jobs = [_job(0, 64, 300), _job(1, 64, 1000), _job(2, 64, 80)]
    svc = S.Service(budget_bytes=1e12, slots=lambda d, s: 2)
    rep = svc.run(jobs, realtime=False, idle_sleep=0)
    by = {j.jid: j for j in svc.done}
    assert rep["nonfinite_jobs"] == 1 and by[1].status == "nonfinite"
    assert by[0].status == by[2].status == "done"
    assert by[0].F_ref == 300 and by[2].F_ref == 80
    # job 2 took job 1's slot long before job 1's 1000 steps would have ended
    assert by[2].admitted >= by[1].finished
    assert by[2].finished < by[0].finished


def test_blocks_lengthen_only_while_no_job_waits(fake):
    """While a job waits for the class's only slot, blocks stay at BLOCK
    steps; once nothing waits, LOOKAHEAD blocks are handed over at once."""
    jobs = [_job(0, 64, 130), _job(1, 64, 470)]
    svc = S.Service(budget_bytes=1e12, slots=lambda d, s: 1, lookahead=4)
    svc.run(jobs, realtime=False, idle_sleep=0)
    steps = [e[1] for e in FakeCtx.log if e[0] == "step"]
    # job 0 runs while job 1 waits: 50, 50, 30; then job 1 alone: 200, 200, 70
    assert steps == [50, 50, 30, 200, 200, 70]
    assert {j.jid: j.F_ref for j in svc.done} == {0: 130, 1: 470}
    with pytest.raises(ValueError):
        S.Service(lookahead=0)


def test_job_larger_than_the_budget_is_rejected_at_submission(fake):
    """A job that could never be admitted must not block the queue forever."""
    svc = S.Service(budget_bytes=128 * 128 - 1, slots=lambda d, s: 2)
    with pytest.raises(ValueError):
        svc.submit(_job(0, 128, 10))
    svc.submit(_job(1, 64, 10))  # a job that fits is still accepted
    assert svc.queued == 1
