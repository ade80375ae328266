"""Continuous-batching service for XPFC simulation jobs (Python binding).

A job is (grid, sigma, n0, steps, seed).  The scheduling runs in the library
(``xpfc_service_*`` in the C ABI, xpfc_sched.cpp / xpfc_service.cpp): jobs are
admitted in arrival order into free slots of their grid class while their
device footprint (``xpfc_instance_bytes``) fits the HBM budget beside the
resident jobs; the running jobs of the classes of one lane advance together
in blocks of up to 50 steps, each block one captured graph with one launch
per pass over the lane's grids (a context of several grids), each job by
min(block, its remaining steps); lanes run on their own streams with their
own block cadence (by default every grid class is a lane of its own, so a
small grid's jobs are not held to a large grid's blocks); a job whose field
diverges stops at the end of its block; finished
jobs free their slot for the next admission.  Fields stay on the device from
admission to completion.  A job's result (final free energy, mean, kept field)
is what the same job gives alone.  This module builds the job requests and the
arrival stream, submits jobs at their arrival times and polls the library.

The reference runs one simulation per process (main.cu:20-435) and has no
service; this module has no counterpart there.
"""
from __future__ import annotations

import ctypes as ct
import math
import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from .api import (Params, XPFCError, XPFC_DCF_MAX, XPFC_FP32_MIXED, _Job, _JobResult,
                  _ServiceClass, _ServiceConfig, _ServiceStats, _check, instance_bytes, lib)

TWO_PI = 2.0 * math.pi
BLOCK = 50      # steps per scheduling block (the library's graph block)
# Blocks handed out at once while no queued job waits.  Measured on the 20
# jobs/s stream with one context per size class: more lookahead shortened the
# median and lengthened the tail, so the default keeps the tail.
LOOKAHEAD = 1
JOB_DONE, JOB_NONFINITE, JOB_REJECTED = 0, 1, 2
_STATUS = {JOB_DONE: "done", JOB_NONFINITE: "nonfinite", JOB_REJECTED: "rejected"}


def square_peaks() -> Tuple[Tuple[float, float, float], ...]:
    """Square-lattice two-peak kernel: k = 2 pi and 2 sqrt(2) pi, alpha = 1, rho*beta = 1."""
    return ((TWO_PI, 1.0, 1.0), (TWO_PI * math.sqrt(2.0), 1.0, 1.0))


@dataclass
class Job:
    """One simulation request."""
    jid: int
    dim: int
    side: int
    sigma: float
    n0: float
    steps: int
    seed: int
    amplitude: float = 0.01
    dx: float = 0.125
    dt: float = 0.5
    arrival: float = 0.0        # seconds from the service start
    # filled in by the service
    admitted: float = float("nan")
    finished: float = float("nan")
    F_ref: float = float("nan")
    f_mean: float = float("nan")
    mean: float = float("nan")
    status: str = "queued"      # queued, done, nonfinite, rejected
    steps_done: int = 0
    field: Optional[np.ndarray] = None

    @property
    def shape(self) -> Tuple[int, ...]:
        return (self.side,) * self.dim

    @property
    def cells(self) -> int:
        return self.side ** self.dim

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def job_params(job: Job, precision: int = XPFC_FP32_MIXED) -> Params:
    """Physics of a job: square-lattice kernel (max of the peaks), a = 1, b = 1/3."""
    n = job.shape
    return Params(n=n, L=tuple(job.dx * m for m in n), dt=job.dt, Mo=1.0, eps=0.0, a=1.0,
                  b=1.0 / 3.0, sigma=job.sigma, psi0=job.n0, peaks=square_peaks(),
                  dcf_mode=XPFC_DCF_MAX, precision=precision, t_sample=0)


def class_slots(cells_per_class: int = 1 << 25) -> Callable[[int, int], int]:
    """Slot policy: about `cells_per_class` cells per class launch, 2..64 instances."""
    def slots(dim: int, side: int) -> int:
        return int(max(2, min(64, cells_per_class // side ** dim)))
    return slots


default_slots = class_slots()


def lane_groups(spec: str):
    """'class', 'one', or ';'-separated groups of ','-separated classes
    ('2d_128,2d_256;3d_256'): the lanes argument of Service.  Classes of no
    group get a lane each."""
    if spec in ("class", "one"):
        return spec
    lane = {}
    for i, grp in enumerate(spec.split(";")):
        for name in grp.split(","):
            name = name.strip()
            if name:
                d, side = name.split("d_")
                lane[(int(d), int(side))] = i
    nxt = len(spec.split(";"))
    return lambda dim, side: lane.get((dim, side), nxt + 1000 * dim + side)


class Service:
    """The library's job service over the grid classes of `classes` (or of
    the jobs passed to run()/warm()).

    budget_bytes: device memory the resident jobs may occupy (sum of
    ``instance_bytes`` of the admitted jobs).  slots: function (dim, side) ->
    slots of that class.  keep_fields: download each job's final field.
    lookahead: blocks per launch while no job waits.  lanes: function (dim,
    side) -> lane number (classes with the same number share blocks), or
    "class" (a lane per class) or "one" (every class in one lane).
    """

    def __init__(self, budget_bytes: float = 200e9, precision: int = XPFC_FP32_MIXED,
                 slots: Callable[[int, int], int] = None, keep_fields: bool = False,
                 lookahead: int = LOOKAHEAD, classes: Optional[Iterable[Tuple[int, int]]] = None,
                 dx: float = 0.125, block: int = BLOCK, lanes="class"):
        if lookahead < 1:
            raise ValueError("lookahead must be >= 1")
        self.lookahead = int(lookahead)
        self.block = int(block)
        self.budget = float(budget_bytes)
        self.precision = precision
        self.slots_of = slots or default_slots
        if lanes == "class":
            self.lane_of = None
        elif lanes == "one":
            self.lane_of = lambda dim, side: 0
        elif callable(lanes):
            self.lane_of = lanes
        else:
            raise ValueError("lanes must be 'class', 'one' or a function (dim, side) -> int")
        self.keep_fields = keep_fields
        self.dx = dx
        self.keys: List[Tuple[int, int]] = sorted(set(classes or []))
        self.handle = None
        self.done: List[Job] = []
        self.queue_samples: List[int] = []
        self.t0 = 0.0
        self._jobs: Dict[int, Job] = {}
        self._final: dict = {k: 0 for k, _ in _ServiceStats._fields_}
        self._fp: Dict[Tuple[int, int], int] = {}

    # -- set-up ---------------------------------------------------------------
    def footprint(self, job: Job) -> int:
        k = (job.dim, job.side)
        if k not in self._fp:
            b = instance_bytes(job_params(job, self.precision))
            if b <= 0:
                raise ValueError("unsupported job grid %s" % (job.shape,))
            self._fp[k] = b
        return self._fp[k]

    def warm(self, keys: Sequence[Tuple[int, int]] = ()):
        """Create the service (and its device context) for the given (dim,
        side) classes and any declared before; jobs of other grids are then
        rejected."""
        if self.handle is not None:
            return
        self.keys = sorted(set(self.keys) | set(keys))
        if not self.keys:
            raise ValueError("no grid classes")
        arr = (_ServiceClass * len(self.keys))()
        for i, (dim, side) in enumerate(self.keys):
            arr[i].dim = dim
            for d in range(3):
                arr[i].n[d] = side if d < dim else 1
                arr[i].L[d] = side * self.dx if d < dim else 1.0
            arr[i].slots = int(self.slots_of(dim, side))
            arr[i].lane = i if self.lane_of is None else int(self.lane_of(dim, side))
        cfg = _ServiceConfig(precision=self.precision, device=0, budget_bytes=self.budget,
                             block_steps=self.block, lookahead=self.lookahead,
                             keep_fields=1 if self.keep_fields else 0, nclasses=len(self.keys),
                             classes=ct.cast(arr, ct.POINTER(_ServiceClass)))
        h = ct.c_void_p()
        _check(lib().xpfc_service_create(ct.byref(cfg), ct.byref(h)), "xpfc_service_create")
        self.handle = h
        self.t0 = time.perf_counter()

    @property
    def classes(self) -> Dict[str, int]:
        return {"%dd_%d" % k: int(self.slots_of(*k)) for k in self.keys}

    # -- requests -------------------------------------------------------------
    def _request(self, job: Job, arrival: float) -> _Job:
        r = _Job()
        r.id = job.jid
        r.params = job_params(job, self.precision).to_c()
        r.steps = job.steps
        r.amplitude = job.amplitude
        r.seed = job.seed
        r.arrival = arrival
        return r

    def submit(self, job: Job, arrival: float = -1.0):
        """Queue a job now (arrival < 0) or at the given service time; raises
        ValueError for a job that can never run (unknown grid class, steps < 1,
        footprint beyond the budget)."""
        if job.steps < 1:
            raise ValueError("job %d: steps must be >= 1" % job.jid)
        if self.footprint(job) > self.budget:
            raise ValueError("job %d: device footprint %d bytes exceeds the budget of %d bytes"
                             % (job.jid, self.footprint(job), int(self.budget)))
        if self.handle is None:
            self.warm([(job.dim, job.side)])
        r = self._request(job, arrival)
        st = lib().xpfc_service_submit(self.handle, ct.byref(r))
        if st != 0:
            self._collect()
            raise ValueError("job %d rejected by the service (%s)" % (job.jid, lib().xpfc_last_error().decode()))
        self._jobs[job.jid] = job
        job.status = "queued"

    def poll(self) -> bool:
        r = lib().xpfc_service_poll(self.handle)
        if r < 0:
            _check(-r, "xpfc_service_poll")
        self._collect()
        return r > 0

    def idle(self) -> bool:
        return bool(lib().xpfc_service_idle(self.handle))

    def _collect(self):
        buf = (_JobResult * 64)()
        while True:
            n = lib().xpfc_service_results(self.handle, buf, 64)
            if n < 0:
                _check(-n, "xpfc_service_results")
            for r in buf[:n]:
                job = self._jobs.pop(int(r.id), None)
                if job is None:
                    continue
                job.status = _STATUS[r.status]
                job.F_ref, job.f_mean, job.mean = r.F_ref, r.f_mean, r.mean
                job.steps_done = int(r.steps_done)
                job.admitted, job.finished = r.t_admitted, r.t_finished
                job.arrival = r.t_arrival
                if self.keep_fields and r.status != JOB_REJECTED:
                    f = np.empty(job.cells, dtype=np.float64)
                    got = lib().xpfc_service_field(self.handle, job.jid, f.ctypes.data_as(ct.POINTER(ct.c_double)),
                                                   job.cells)
                    if got < 0:
                        _check(int(-got), "xpfc_service_field")
                    job.field = f.reshape(job.shape)
                self.done.append(job)
            if n < 64:
                break

    def stats(self) -> dict:
        if self.handle is None:  # closed: the figures it ended with
            return dict(self._final)
        st = _ServiceStats()
        _check(lib().xpfc_service_get_stats(self.handle, ct.byref(st)), "xpfc_service_get_stats")
        return {k: getattr(st, k) for k, _ in _ServiceStats._fields_}

    @property
    def peak_resident(self) -> float:
        return self.stats()["peak_resident_bytes"]

    # -- the serving loop -----------------------------------------------------
    def run(self, jobs: Sequence[Job], realtime: bool = True, idle_sleep: float = 1e-4,
            progress: Optional[Callable[[str], None]] = None) -> dict:
        """Serve `jobs` (by arrival): realtime submits each job at its arrival
        time on the service clock, otherwise all at once.  The library admits,
        runs and finishes them; this loop only submits and polls.  progress:
        called with a status line every 10 s."""
        self.warm(sorted({(j.dim, j.side) for j in jobs}))
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        total = len(pending)
        start = len(self.done)
        wall0 = time.perf_counter()
        last_q = last_p = -1e9
        while len(self.done) - start < total:
            t = time.perf_counter() - self.t0
            while pending and (not realtime or pending[0].arrival <= time.perf_counter() - wall0):
                job = pending.popleft()
                # on the service clock: the scheduled arrival, or now
                self.submit(job, arrival=job.arrival + (wall0 - self.t0) if realtime else -1.0)
            if t - last_q >= 0.1:
                self.queue_samples.append(int(self.stats()["queued"]))
                last_q = t
            if progress and t - last_p >= 10.0:
                last_p = t
                s = self.stats()
                progress("t=%.0fs done=%d queued=%d running=%d resident=%.1fGB"
                         % (t, len(self.done) - start, s["queued"], s["running"], s["resident_bytes"] / 1e9))
            if not self.poll():
                if pending and realtime:
                    time.sleep(min(idle_sleep, max(0.0, pending[0].arrival - (time.perf_counter() - wall0))))
                else:
                    time.sleep(idle_sleep)
        return self.report(time.perf_counter() - wall0, self.done[start:])

    def report(self, wall: float, done: Optional[List[Job]] = None) -> dict:
        done = self.done if done is None else done
        ran = [j for j in done if j.status != "rejected"]
        lat = np.array([j.latency for j in ran])
        s = self.stats()
        return {
            "jobs": len(done),
            "nonfinite_jobs": sum(j.status == "nonfinite" for j in done),
            "rejected_jobs": sum(j.status == "rejected" for j in done),
            "wall_s": wall,
            "p50_s": float(np.percentile(lat, 50)) if lat.size else float("nan"),
            "p99_s": float(np.percentile(lat, 99)) if lat.size else float("nan"),
            "mean_s": float(lat.mean()) if lat.size else float("nan"),
            "cell_updates_per_s": s["cell_updates"] / wall if wall > 0 else 0.0,
            "queue_max": int(s["queue_max"]),
            "queue_mean": float(np.mean(self.queue_samples)) if self.queue_samples else 0.0,
            "peak_resident_gb": s["peak_resident_bytes"] / 1e9,
            "blocks": int(s["blocks"]),
            "classes": self.classes,
            "by_class": self._by_class(ran),
        }

    @staticmethod
    def _by_class(ran: List[Job]) -> dict:
        """Per grid class: jobs, median / 99th-percentile latency and the
        median wait for admission, in seconds."""
        out = {}
        for key in sorted({(j.dim, j.side) for j in ran}):
            js = [j for j in ran if (j.dim, j.side) == key]
            lat = np.array([j.latency for j in js])
            wait = np.array([j.admitted - j.arrival for j in js])
            out["%dd_%d" % key] = {"jobs": len(js), "p50_s": round(float(np.percentile(lat, 50)), 4),
                                   "p99_s": round(float(np.percentile(lat, 99)), 4),
                                   "wait_p50_s": round(float(np.percentile(wait, 50)), 4)}
        return out

    def close(self):
        if self.handle is not None:
            self._final = self.stats()
            lib().xpfc_service_destroy(self.handle)
            self.handle = None


def batch_jobs(ninst: int = 64, steps: int = 2000, seed: int = 0,
               sides=(128, 256, 512, 1024)) -> List[Job]:
    """The batched workload: `ninst` independent 2-D square-lattice instances,
    side drawn uniformly from `sides`, sigma ~ U[0.05, 0.2], n0 ~ U[0.02, 0.08],
    noise amplitude 0.01 seeded by the instance id, `steps` steps each."""
    rng = np.random.default_rng(seed)
    side = rng.choice(list(sides), size=ninst)
    sig = rng.uniform(0.05, 0.2, ninst)
    n0 = rng.uniform(0.02, 0.08, ninst)
    return [Job(jid=i, dim=2, side=int(side[i]), sigma=float(sig[i]), n0=float(n0[i]),
                steps=steps, seed=i) for i in range(ninst)]


def grid_sides(lo: int, hi: int, rounding: str = "smooth") -> List[int]:
    """The grid sides in [lo, hi] a job's side is rounded to: every extent the
    library compiles (2^k, 3*2^k, 5*2^k) for "smooth", powers of two for "pow2"."""
    if rounding not in ("smooth", "pow2"):
        raise ValueError("rounding must be 'smooth' or 'pow2'")
    out = []
    for f in ((1, 3, 5) if rounding == "smooth" else (1,)):
        v = 4 * f
        while v <= hi:
            if v >= lo and v >= 8:
                out.append(v)
            v *= 2
    return sorted(out)


def log_uniform_side(rng: np.random.Generator, lo: int, hi: int, rounding: str = "smooth") -> int:
    """Side log-uniform in [lo, hi], rounded to the nearest grid side of
    `grid_sides(lo, hi, rounding)` in log2."""
    e = rng.uniform(math.log2(lo), math.log2(hi))
    sides = grid_sides(lo, hi, rounding)
    return int(min(sides, key=lambda v: abs(math.log2(v) - e)))


def poisson_jobs(rate: float, duration: float, seed: int = 0, frac_3d: float = 0.5,
                 side2=(128, 1024), side3=(64, 256), steps=(500, 5000), sigma=(0.05, 0.2),
                 n0=(0.02, 0.08), rounding: str = "smooth") -> List[Job]:
    """Poisson arrivals at `rate` jobs/s over `duration` s; job j draws its own
    size (log-uniform side rounded as `log_uniform_side`), sigma, n0 and step
    count; its noise seed is its id."""
    rng = np.random.default_rng(seed)
    jobs, t, jid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        dim = 3 if rng.random() < frac_3d else 2
        side = log_uniform_side(rng, *(side3 if dim == 3 else side2), rounding=rounding)
        jobs.append(Job(jid=jid, dim=dim, side=side, sigma=float(rng.uniform(*sigma)),
                        n0=float(rng.uniform(*n0)), steps=int(rng.integers(steps[0], steps[1] + 1)),
                        seed=jid, arrival=t))
        jid += 1
    return jobs
