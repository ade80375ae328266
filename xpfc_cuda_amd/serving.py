"""Continuous-batching service for XPFC simulation jobs.

A job is (grid, sigma, n0, steps, seed).  Jobs of one grid and precision share a
size-class context (``BatchContext``) with a fixed number of instance slots; a
job is admitted when its class has a free slot and its device footprint
(``xpfc_instance_bytes``) fits the HBM budget beside the jobs already resident.
Its field stays on the device from admission to completion.  Every size class
advances its running jobs together, in blocks of up to 50 steps (captured
hipGraphs inside the library), or up to `lookahead` such blocks at once while
no queued job waits for the class, on its own stream, so the classes' work
overlaps on the GPU; the host polls the classes without blocking and hands
each one its next block as soon as its previous block has finished, admitting
queued jobs into freed slots at that moment.  A job's result (final free energy, mean,
optionally the field) is what the same job gives alone: the instances of a
context never exchange data.

The reference runs one simulation per process (main.cu:20-435) and has no
service; this module has no counterpart there.  It only calls the C ABI.
"""
from __future__ import annotations

import heapq
import math
import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from .api import (BatchContext, Params, XPFCError, XPFC_DCF_MAX, XPFC_ERR_NONFINITE,
                  XPFC_FP32_MIXED, instance_bytes)

TWO_PI = 2.0 * math.pi
BLOCK = 50      # steps per scheduling block (the library's graph block)
# Blocks handed at once to a class that no queued job waits for.  Measured on
# the 20 jobs/s stream (60 s windows): 1 -> p50 0.71 s, p99 4.7 s; 2 -> 0.47 s,
# 6.2 s; 4 -> 0.29 s, 7.7 s; 8 -> 0.21 s, 9.7 s.  Longer runs of the small
# classes take GPU share from the largest class, whose queue sets the tail, so
# the default keeps the tail.
LOOKAHEAD = 1


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
    status: str = "queued"      # queued, running, done, nonfinite
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


class SizeClass:
    """The context of one (dim, side) class and the jobs in its slots."""

    def __init__(self, dim: int, side: int, nslots: int, template: Job, precision: int):
        self.key = (dim, side)
        self.nslots = nslots
        self.precision = precision
        p = job_params(template, precision)
        self.ctx = BatchContext([p] * nslots)
        self.ctx.set_active([])
        self.free = list(range(nslots - 1, -1, -1))
        self.running: Dict[int, Job] = {}      # slot -> job
        self.remaining: Dict[int, int] = {}    # slot -> steps left
        self.busy = False
        self.active_key: Tuple[int, ...] = ()
        self.cell_updates = 0

    def close(self):
        self.ctx.close()


class Service:
    """Admission, slots and the polling loop.

    budget_bytes: device memory the resident jobs may occupy (sum of
    ``instance_bytes`` of the admitted jobs).  slots: function (dim, side) ->
    slots of that size class.  keep_fields: download each job's final field.
    lookahead: blocks handed to a class at once while no queued job waits for
    it (fewer host round trips; a diverged job is then stopped up to that many
    blocks later).
    """

    def __init__(self, budget_bytes: float = 200e9, precision: int = XPFC_FP32_MIXED,
                 slots: Callable[[int, int], int] = default_slots, keep_fields: bool = False,
                 lookahead: int = LOOKAHEAD):
        if lookahead < 1:
            raise ValueError("lookahead must be >= 1")
        self.lookahead = int(lookahead)
        self.budget = float(budget_bytes)
        self.precision = precision
        self.slots_of = slots
        self.keep_fields = keep_fields
        self.classes: Dict[Tuple[int, int], SizeClass] = {}
        self.queues: Dict[Tuple[int, int], deque] = {}  # per size class, arrival order
        self.queued = 0
        self.resident = 0
        self.peak_resident = 0
        self.done: List[Job] = []
        self.queue_samples: List[int] = []
        self._fp: Dict[Tuple[int, int], int] = {}
        self.t0 = 0.0

    def footprint(self, job: Job) -> int:
        k = (job.dim, job.side)
        if k not in self._fp:
            b = instance_bytes(job_params(job, self.precision))
            if b <= 0:
                raise ValueError("unsupported job grid %s" % (job.shape,))
            self._fp[k] = b
        return self._fp[k]

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def submit(self, job: Job):
        if job.steps < 1:
            raise ValueError("job %d: steps must be >= 1" % job.jid)
        fp = self.footprint(job)  # rejects unsupported grids at submission
        if fp > self.budget:
            raise ValueError("job %d: device footprint %d bytes exceeds the budget of %d bytes"
                             % (job.jid, fp, int(self.budget)))
        self.queues.setdefault((job.dim, job.side), deque()).append(job)
        self.queued += 1

    def warm(self, keys: Sequence[Tuple[int, int]]):
        """Create the size-class contexts and tune their passes ahead of the
        first job (otherwise the first job of a class pays for both)."""
        for dim, side in keys:
            c = self._class(Job(jid=-1, dim=dim, side=side, sigma=0.1, n0=0.05, steps=1, seed=0))
            c.ctx.autotune()

    # -- per class -------------------------------------------------------------
    def _class(self, job: Job) -> SizeClass:
        k = (job.dim, job.side)
        c = self.classes.get(k)
        if c is None:
            c = SizeClass(job.dim, job.side, self.slots_of(job.dim, job.side), job, self.precision)
            self.classes[k] = c
        return c

    def _complete(self, c: SizeClass):
        finished = [s for s, r in c.remaining.items() if r == 0]
        for slot in finished:
            job = c.running.pop(slot)
            del c.remaining[slot]
            if job.status != "nonfinite":
                try:
                    job.F_ref, job.f_mean = c.ctx.fe_calc(slot)
                    job.status = "done"
                except XPFCError as e:
                    if e.status != XPFC_ERR_NONFINITE:
                        raise
                    job.status = "nonfinite"
            job.mean = c.ctx.avg_calc(slot)
            if self.keep_fields:
                job.field = c.ctx.get_field(slot)
            job.finished = self.now()
            c.free.append(slot)
            self.resident -= self.footprint(job)
            self.done.append(job)
        if finished:  # a freed slot leaves the active list before it is reused
            c.active_key = tuple(sorted(c.running))
            c.ctx.set_active(list(c.active_key))

    def _admit(self):
        """Admit queued jobs, in arrival order over all classes, into idle
        classes with free slots.  Memory is granted in arrival order: once a
        job waits for memory, no later job is admitted before it."""
        heap = [(q[0].arrival, k) for k, q in self.queues.items() if q]
        heapq.heapify(heap)
        while heap:
            _, k = heapq.heappop(heap)
            q = self.queues[k]
            job = q[0]
            c = self._class(job)
            if c.busy or not c.free:
                continue  # this class waits for its block or a slot
            fp = self.footprint(job)
            if self.resident + fp > self.budget:
                return
            q.popleft()
            self.queued -= 1
            self._start(c, job, fp)
            if q:
                heapq.heappush(heap, (q[0].arrival, k))

    def _start(self, c: SizeClass, job: Job, fp: int):
        slot = c.free.pop()
        c.ctx.set_params(slot, job_params(job, self.precision))
        c.ctx.reset(slot)
        c.ctx.initialize_noise(slot, job.amplitude, job.seed)
        c.running[slot] = job
        c.remaining[slot] = job.steps
        job.status = "running"
        job.admitted = self.now()
        self.resident += fp
        self.peak_resident = max(self.peak_resident, self.resident)

    def _launch(self, c: SizeClass):
        if not c.running:
            return
        slots = tuple(sorted(c.running))
        if slots != c.active_key:
            c.ctx.set_active(list(slots))
            c.active_key = slots
        cap = BLOCK if self.queues.get(c.key) else BLOCK * self.lookahead
        n = min(cap, min(c.remaining.values()))
        c.ctx.step(n)
        for s in slots:
            c.remaining[s] -= n
        c.cell_updates += n * len(slots) * c.running[slots[0]].cells
        c.busy = True

    def poll(self) -> bool:
        """One pass over the classes; True if any work was handed out."""
        worked = False
        for c in list(self.classes.values()):
            if c.busy:
                if not c.ctx.query():
                    continue
                c.busy = False
                # the explicit nonlinear term diverges for strongly
                # crystallising parameters: such a job stops at the end of
                # the block it diverged in and its slot is reused
                ok = c.ctx.probe_finite()
                for slot, job in c.running.items():
                    if not ok[slot]:
                        job.status = "nonfinite"
                        c.remaining[slot] = 0
                self._complete(c)
                worked = True
        self._admit()
        for c in self.classes.values():
            if not c.busy and c.running:
                self._launch(c)
                worked = True
        return worked

    def run(self, jobs: Sequence[Job], realtime: bool = True, idle_sleep: float = 1e-4,
            progress: Optional[Callable[[str], None]] = None) -> dict:
        """Serve `jobs` (sorted by arrival).  realtime: submit each job at its
        arrival time; otherwise submit all at once.  progress: called with a
        status line every 10 s."""
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        self.t0 = time.perf_counter()
        total = len(pending)
        last_q = -1.0
        last_p = 0.0
        while len(self.done) < total:
            t = self.now()
            if progress and t - last_p >= 10.0:
                last_p = t
                progress("t=%.0fs done=%d queued=%d running=%d resident=%.1fGB"
                         % (t, len(self.done), self.queued,
                            sum(len(c.running) for c in self.classes.values()), self.resident / 1e9))
            while pending and (not realtime or pending[0].arrival <= t):
                self.submit(pending.popleft())
            if t - last_q >= 0.1:
                self.queue_samples.append(self.queued)
                last_q = t
            if not self.poll():
                if pending and realtime:
                    time.sleep(min(idle_sleep, max(0.0, pending[0].arrival - self.now())))
                else:
                    time.sleep(idle_sleep)
        wall = self.now()
        return self.report(wall)

    def report(self, wall: float) -> dict:
        lat = np.array([j.latency for j in self.done])
        cu = sum(c.cell_updates for c in self.classes.values())
        return {
            "jobs": len(self.done),
            "nonfinite_jobs": sum(j.status == "nonfinite" for j in self.done),
            "wall_s": wall,
            "p50_s": float(np.percentile(lat, 50)) if lat.size else float("nan"),
            "p99_s": float(np.percentile(lat, 99)) if lat.size else float("nan"),
            "mean_s": float(lat.mean()) if lat.size else float("nan"),
            "cell_updates_per_s": cu / wall if wall > 0 else 0.0,
            "queue_max": int(max(self.queue_samples)) if self.queue_samples else 0,
            "queue_mean": float(np.mean(self.queue_samples)) if self.queue_samples else 0.0,
            "peak_resident_gb": self.peak_resident / 1e9,
            "classes": {"%dd_%d" % k: c.nslots for k, c in sorted(self.classes.items())},
        }

    def close(self):
        for c in self.classes.values():
            c.close()
        self.classes.clear()


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


def log_uniform_side(rng: np.random.Generator, lo: int, hi: int) -> int:
    """Side log-uniform in [lo, hi], rounded to the nearest power of two in log2."""
    e = rng.uniform(math.log2(lo), math.log2(hi))
    return int(2 ** int(round(e)))


def poisson_jobs(rate: float, duration: float, seed: int = 0, frac_3d: float = 0.5,
                 side2=(128, 1024), side3=(64, 256), steps=(500, 5000), sigma=(0.05, 0.2),
                 n0=(0.02, 0.08)) -> List[Job]:
    """Poisson arrivals at `rate` jobs/s over `duration` s; job j draws its own
    size, sigma, n0 and step count; its noise seed is its id."""
    rng = np.random.default_rng(seed)
    jobs, t, jid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        dim = 3 if rng.random() < frac_3d else 2
        side = log_uniform_side(rng, *(side3 if dim == 3 else side2))
        jobs.append(Job(jid=jid, dim=dim, side=side, sigma=float(rng.uniform(*sigma)),
                        n0=float(rng.uniform(*n0)), steps=int(rng.integers(steps[0], steps[1] + 1)),
                        seed=jid, arrival=t))
        jid += 1
    return jobs
