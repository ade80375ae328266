"""Serving: each job's result is the same alone and in the mix.

Jobs of three size classes arrive together into a service with two slots per
class and a memory budget of four jobs, so they wait for slots and memory, run
beside jobs of other sizes and lengths, and join classes mid-run.  Each job is
then run alone in a one-instance context: the final field must be identical
bit for bit, and so must the free energy and mean (the final diagnostics use
a fixed reduction shape, independent of the pass variants).
"""
import numpy as np
import pytest

from xpfc_cuda_amd import Context
from xpfc_cuda_amd import serving as S

pytestmark = pytest.mark.gpu


def _jobs():
    rng = np.random.default_rng(4)
    jobs = []
    sizes = [(2, 32), (2, 64), (3, 16)]
    for jid in range(10):
        dim, side = sizes[jid % 3]
        jobs.append(S.Job(jid=jid, dim=dim, side=side, sigma=float(rng.uniform(0.07, 0.2)),
                          n0=float(rng.uniform(0.02, 0.06)), steps=int(rng.integers(30, 140)),
                          seed=jid, arrival=0.001 * jid))
    # strongly crystallising: the explicit nonlinear term diverges within 137
    # steps (the FP64 oracle too); its NaNs must stay in its own slot
    jobs[4].sigma, jobs[4].n0, jobs[4].steps = 0.062, 0.0564, 137
    return jobs


@pytest.mark.parametrize("lookahead", [1, 4])
def test_alone_equals_mix(lookahead):
    print("serving lookahead=%d" % lookahead, flush=True)
    jobs = _jobs()
    budget = 4 * max(S.instance_bytes(S.job_params(j)) for j in jobs)
    svc = S.Service(budget_bytes=budget, slots=lambda d, s: 2, keep_fields=True,
                    lookahead=lookahead)
    try:
        rep = svc.run(jobs, realtime=False, idle_sleep=0.0)
    finally:
        svc.close()
    print("service", rep, flush=True)
    assert rep["jobs"] == len(jobs) and rep["nonfinite_jobs"] == 1
    assert svc.peak_resident <= budget
    for job in svc.done:
        print("job", job.jid, job.shape, job.steps, job.status, flush=True)
        with Context(S.job_params(job)) as c:
            c.initialize_noise(job.amplitude, job.seed)
            c.step(job.steps)
            alone = c.get_field()
            if job.jid == 4:
                # stopped at the end of the block it diverged in
                assert job.status == "nonfinite" and not np.all(np.isfinite(alone))
                assert not np.all(np.isfinite(job.field))
                continue
            F, fm = c.fe_calc()
            m = c.avg_calc()
        assert job.status == "done" and np.all(np.isfinite(alone))
        assert np.array_equal(job.field, alone), "job %d differs from its solo run" % job.jid
        assert job.F_ref == F and job.f_mean == fm and job.mean == m


def test_service_run_in_the_library():
    """xpfc_service_run: the library's own serving loop (arrival times, polling)
    over a few jobs of two classes, one of which diverges; the results come
    back through xpfc_service_results and match each job alone."""
    import ctypes as ct
    from xpfc_cuda_amd import lib
    from xpfc_cuda_amd.api import _Job, _JobResult
    print("service run in the library", flush=True)
    jobs = _jobs()[:6]
    svc = S.Service(slots=lambda d, s: 2, keep_fields=True)
    svc.warm(sorted({(j.dim, j.side) for j in jobs}))
    arr = (_Job * len(jobs))(*[svc._request(j, 0.01 * j.jid) for j in jobs])
    try:
        assert lib().xpfc_service_run(svc.handle, arr, len(jobs), 1) == 0
        out = (_JobResult * 16)()
        n = lib().xpfc_service_results(svc.handle, out, 16)
        assert n == len(jobs)
        res = {r.id: r for r in out[:n]}
        st = svc.stats()
        assert st["done"] + st["nonfinite"] == len(jobs) and st["nonfinite"] == 1
        for j in jobs:
            r = res[j.jid]
            assert r.t_arrival <= r.t_admitted <= r.t_finished
            with Context(S.job_params(j)) as c:
                c.initialize_noise(j.amplitude, j.seed)
                c.step(j.steps)
                if j.jid == 4:
                    assert r.status == 1 and r.steps_done < j.steps
                    continue
                F, _ = c.fe_calc()
            assert r.status == 0 and r.steps_done == j.steps and r.F_ref == F
    finally:
        svc.close()
