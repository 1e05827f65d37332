"""Continuous-batching request loop: cavity solves arrive over time, are admitted by device-memory footprint
while slots and memory are free, and advance together as one ensemble (one launch per kernel for every
admitted solve, the step replayed as a HIP graph). A solve's fields stay on the device from admission to
completion or preemption; the ensemble is re-formed whenever a solve joins or leaves, between steps.

Policy:
  * arrivals wait in a FIFO queue; the head is admitted when the ensemble has a free slot (max_members) and
    its footprint (cav_footprint, the exact device bytes of one solver) fits in the device-memory budget
    still free; a request whose footprint exceeds the whole budget is rejected at once;
  * admitted solves step in chunks of at most `chunk` steps, a chunk never running past the earliest
    completion, so a finished solve leaves (and frees its memory) right after its last step;
  * new solves are created and initialised on the host side of the loop between chunks; their set-up
    kernels run on their own streams;
  * with a quantum (steps), a full device is time-sliced: when the queue head cannot be admitted, solves
    that have run a whole quantum since their admission are preempted, longest-running first, until the head
    fits. A preempted solve's whole state is snapshotted to host memory (Solver.snapshot) and its device
    memory freed; it rejoins the queue at the tail and resumes from the snapshot (Solver.restore) when it is
    admitted again. Without a quantum a solve keeps its slot until it completes.

Latency of a request is arrival -> completion (its last step done and its report read), queueing
included. Each solve's result is bit-for-bit what it gives when run alone: the batched kernels give every
member the same blocks and reduction order as a solo run, and a restored snapshot continues the solve
exactly.

The reference has no serving loop; the request mix follows the serving workload (Poisson arrivals, sizes
32^3..192^3 with anisotropic variants, Re log-uniform in [100, 1e4], 100..2000 steps, mixed-precision FDM
at rtol 1e-10).
"""
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import workloads as W
from .python_api import Ensemble, Solver, footprint


@dataclass
class Request:
    rid: int
    settings: dict
    steps: int
    arrival: float               # seconds after the start of the run
    seed: int
    admitted: float = None
    finished: float = None
    report: dict = None
    error: str = ''
    done: int = 0
    bytes: int = 0
    snap: object = None          # host snapshot of a preempted solve
    preemptions: int = 0
    run: int = 0                 # steps since its latest admission

    @property
    def cells(self):
        return self.settings['nx'] * self.settings['ny'] * self.settings['nz']


def poisson_requests(duration, rate=20.0, seed=4, sizes=(32, 64, 96, 128, 192), steps=(100, 2000),
                     re_range=(100.0, 1e4), fdm_rtol=1e-10):
    """Arrivals of a Poisson process (rate per second) over `duration` seconds; size n^3 from `sizes`,
    with probability 1/2 one random axis halved; Re log-uniform; step count uniform."""
    rng = np.random.default_rng(seed)
    out, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            return out
        n = int(rng.choice(sizes))
        dims = [n, n, n]
        if rng.random() < 0.5:
            dims[int(rng.integers(0, 3))] //= 2
        re = float(np.exp(rng.uniform(np.log(re_range[0]), np.log(re_range[1]))))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        s = W.settings(dims[0], dims[1], dims[2], re=re, lid_shaped=1, adaptive_dt=1, cfl_target=0.5,
                       fdm_rtol=fdm_rtol)
        out.append(Request(k, s, nsteps, t, seed * 100000 + k))
        k += 1


@dataclass
class ServeStats:
    completed: list = field(default_factory=list)
    rejected: list = field(default_factory=list)
    unfinished: list = field(default_factory=list)
    preemptions: int = 0
    step_ms: list = field(default_factory=list)
    step_cells: list = field(default_factory=list)
    queue_depth: list = field(default_factory=list)     # (time, waiting, running)
    wall: float = 0.0

    def summary(self):
        lat = np.array([r.finished - r.arrival for r in self.completed]) if self.completed else np.zeros(0)
        wait = np.array([r.admitted - r.arrival for r in self.completed]) if self.completed else np.zeros(0)
        steps = np.array(self.step_ms) if self.step_ms else np.zeros(0)
        cell_steps = float(sum(self.step_cells))
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else None
        return {
            'completed': len(self.completed), 'rejected': len(self.rejected), 'unfinished': len(self.unfinished),
            'latency_s_p50': pct(lat, 50), 'latency_s_p99': pct(lat, 99),
            'queue_wait_s_p50': pct(wait, 50), 'queue_wait_s_p99': pct(wait, 99),
            'step_ms_p50': pct(steps, 50), 'step_ms_p99': pct(steps, 99),
            'mcell_steps_per_s': cell_steps / self.wall / 1e6 if self.wall > 0 else 0.0,
            'gpu_mcell_steps_per_s': cell_steps / (steps.sum() / 1e3) / 1e6 if steps.size else 0.0,
            'max_waiting': max((q[1] for q in self.queue_depth), default=0),
            'max_running': max((q[2] for q in self.queue_depth), default=0),
            'preemptions': self.preemptions,
            'wall_s': self.wall,
        }


class Server:
    """Continuous batching of cavity solves on the current GPU.

    budget_bytes: device memory the admitted solves may hold together; max_members: ensemble size cap;
    chunk: most steps between two admission points; quantum: steps after which a solve may be preempted for
    a waiting one (None: never); precision: the FDM mode every solve runs in;
    on_finish(request, solver): called with the live solver before its memory is released; clock and sleep
    measure and pass the time (a virtual clock replays an arrival pattern at any speed)."""

    def __init__(self, budget_bytes, max_members=64, chunk=20, precision='fdm_fp32', noise_sigma=1e-3,
                 on_finish=None, clock=time.perf_counter, sleep=time.sleep, quantum=None):
        self.budget = int(budget_bytes)
        self.max_members = int(max_members)
        self.chunk = int(chunk)
        self.quantum = None if quantum is None else max(1, int(quantum))
        self.precision = precision
        self.sigma = noise_sigma
        self.on_finish = on_finish
        self.clock = clock
        self.sleep = sleep

    def _victims(self, head, active, used):
        """Solves to preempt so that the queue head fits: those that ran a whole quantum, longest first, and
        only if together they make room for it (none when it fits already or cannot fit)."""
        if not head.bytes:
            head.bytes = footprint(head.settings, self.precision)
        if head.bytes > self.budget:
            return []
        fits = lambda n, u: n < self.max_members and u + head.bytes <= self.budget
        if fits(len(active), used):
            return []
        ready = sorted((a for a in active if a[0].run >= self.quantum), key=lambda a: -a[0].run)
        out, n, u = [], len(active), used
        for r, g in ready:
            out.append((r, g))
            n -= 1
            u -= r.bytes
            if fits(n, u):
                return out
        return []

    def run(self, requests, duration=None, drain=True):
        """Serve `requests` (sorted by arrival). Stops admitting at `duration` seconds (None: never) and, with
        drain, finishes what is running; returns ServeStats."""
        stats = ServeStats()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        active = []                      # [(request, solver)]
        used = 0
        ensemble = None
        t0 = self.clock()
        try:
            while True:
                now = self.clock() - t0
                closing = duration is not None and now >= duration
                while pending and pending[0].arrival <= now and not closing:
                    queue.append(pending.popleft())
                changed = False
                if self.quantum is not None and queue and not closing:
                    victims = self._victims(queue[0], active, used)
                    if victims:
                        if ensemble is not None:
                            ensemble.close()
                            ensemble = None
                        for r, g in victims:
                            r.snap = g.snapshot()
                            g.close()
                            used -= r.bytes
                            r.preemptions += 1
                            stats.preemptions += 1
                            queue.append(r)
                        active = [(r, g) for r, g in active if all(r is not q for q, _ in victims)]
                        changed = True
                while queue and len(active) < self.max_members and not closing:
                    r = queue[0]
                    if not r.bytes:
                        r.bytes = footprint(r.settings, self.precision)
                    if r.bytes > self.budget:
                        queue.popleft()
                        r.error = 'footprint exceeds the device-memory budget'
                        stats.rejected.append(r)
                        continue
                    if used + r.bytes > self.budget:
                        break
                    queue.popleft()
                    if ensemble is not None:
                        ensemble.close()
                        ensemble = None
                    g = Solver(r.settings, handle=True, precision=self.precision)
                    if r.snap is not None:
                        g.restore(r.snap)
                        r.snap = None
                    else:
                        g.initialize(*W.noise(r.settings, r.seed, self.sigma))
                        r.admitted = self.clock() - t0
                    r.run = 0
                    active.append((r, g))
                    used += r.bytes
                    changed = True
                stats.queue_depth.append((now, len(queue), len(active)))
                if not active:
                    if closing or (not pending and not queue):
                        break
                    if pending and not queue:
                        wait = pending[0].arrival - (self.clock() - t0)
                        if wait > 0:
                            self.sleep(min(wait, 0.05))
                    continue
                if ensemble is None or changed:
                    if ensemble is not None:
                        ensemble.close()
                    ensemble = Ensemble([g for _, g in active])
                k = min(self.chunk, min(r.steps - r.done for r, _ in active))
                _, per, errors = ensemble.time_steps(k)
                cells = sum(r.cells for r, _ in active)
                stats.step_ms.extend(per.tolist())
                stats.step_cells.extend([cells] * k)
                leaving = []
                for (r, g), err in zip(active, errors):
                    r.done += k
                    r.run += k
                    if err:
                        r.error = err
                    if err or r.done >= r.steps:
                        leaving.append((r, g))
                if leaving:
                    ensemble.close()
                    ensemble = None
                    for r, g in leaving:
                        r.report = g.report()
                        if self.on_finish is not None:
                            self.on_finish(r, g)
                        g.close()
                        used -= r.bytes
                        r.finished = self.clock() - t0
                        stats.completed.append(r)
                    active = [(r, g) for r, g in active if all(r is not q for q, _ in leaving)]
                if closing and not drain:
                    break
        finally:
            stats.wall = self.clock() - t0
            if ensemble is not None:
                ensemble.close()
            for r, g in active:
                stats.unfinished.append(r)
                g.close()
            stats.unfinished.extend(queue)
            stats.unfinished.extend(pending)
        return stats
