"""Continuous-batching request loop: cavity solves arrive over time, are admitted by device-memory footprint
while slots and memory are free, and advance together as one ensemble (one launch per kernel for every
admitted solve, the step replayed as a HIP graph). A solve's fields stay on the device from admission to
completion or preemption; the ensemble is re-formed whenever a solve joins or leaves, between steps.

Policy:
  * arrivals wait in a queue. policy='fifo': the head is admitted when the ensemble has a free slot
    (max_members) and its footprint (cav_footprint, the exact device bytes of one solver) fits in the
    device-memory budget still free, and nothing behind it passes it. policy='smallest': at every admission
    point the waiting solves are taken in order of their remaining work (cells x steps), each admitted if it
    fits, so a large solve that does not fit does not block small ones (large solves can wait long under
    overload). A request whose footprint exceeds the whole budget is rejected at once;
  * an admitted solve's memory is reserved at once and its solver is created and initialised (host FDM
    set-up, device arena, the initial projection) on a worker thread, off the stepping path; it joins the
    ensemble at the first chunk boundary after it is ready;
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
included. The summary gives it for the completed requests and, censored at the end of the run (a lower
bound for every request still waiting or running then), for every offered request. Each solve's result is bit-for-bit what it gives when run alone: the batched kernels give every
member the same blocks and reduction order as a solo run, and a restored snapshot continues the solve
exactly.

The reference has no serving loop; the request mix follows the serving workload (Poisson arrivals, sizes
32^3..192^3 with anisotropic variants, Re log-uniform in [100, 1e4], 100..2000 steps, mixed-precision FDM
at rtol 1e-10).
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor, wait as wait_futures, FIRST_COMPLETED
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
    ready: float = None          # when its solver had been created and it joined the ensemble
    future: object = None        # solver being prepared on the worker thread

    @property
    def work(self):
        return self.cells * (self.steps - self.done)

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
    queue_depth: list = field(default_factory=list)     # (time, waiting, admitted: running or being prepared)
    wall: float = 0.0
    reforms: int = 0
    reform_ms: float = 0.0          # host time re-forming the ensemble (plan, stream, graph capture)

    def summary(self):
        lat = np.array([r.finished - r.arrival for r in self.completed]) if self.completed else np.zeros(0)
        # every offered request: completed ones by their latency, the rest censored at the end of the run
        offered = [r for r in self.completed + self.unfinished if r.arrival <= self.wall]
        cens = np.array([(r.finished if r.finished is not None else self.wall) - r.arrival for r in offered])
        n_cens = sum(1 for r in offered if r.finished is None)
        wait = np.array([r.admitted - r.arrival for r in self.completed]) if self.completed else np.zeros(0)
        steps = np.array(self.step_ms) if self.step_ms else np.zeros(0)
        cell_steps = float(sum(self.step_cells))
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else None
        return {
            'completed': len(self.completed), 'rejected': len(self.rejected), 'unfinished': len(self.unfinished),
            'latency_s_p50': pct(lat, 50), 'latency_s_p99': pct(lat, 99),
            'offered': len(offered), 'censored': n_cens,
            'latency_all_offered_s_p50_lower_bound': pct(cens, 50),
            'latency_all_offered_s_p99_lower_bound': pct(cens, 99),
            'queue_wait_s_p50': pct(wait, 50), 'queue_wait_s_p99': pct(wait, 99),
            'step_ms_p50': pct(steps, 50), 'step_ms_p99': pct(steps, 99),
            'mcell_steps_per_s': cell_steps / self.wall / 1e6 if self.wall > 0 else 0.0,
            'gpu_mcell_steps_per_s': cell_steps / (steps.sum() / 1e3) / 1e6 if steps.size else 0.0,
            'max_waiting': max((q[1] for q in self.queue_depth), default=0),
            'max_running': max((q[2] for q in self.queue_depth), default=0),
            'preemptions': self.preemptions,
            'ensemble_reforms': self.reforms, 'ensemble_reform_ms_total': self.reform_ms,
            'wall_s': self.wall,
        }


class Server:
    """Continuous batching of cavity solves on the current GPU.

    budget_bytes: device memory the admitted solves may hold together; max_members: ensemble size cap;
    chunk: most steps between two admission points; quantum: steps after which a solve may be preempted for
    a waiting one (None: never); precision: the FDM mode every solve runs in; policy: 'fifo' or 'smallest'
    (see the module notes); prepare_workers: threads creating admitted solvers (0: on the stepping path);
    on_finish(request, solver): called with the live solver before its memory is released; clock and sleep
    measure and pass the time (a virtual clock replays an arrival pattern at any speed)."""

    def __init__(self, budget_bytes, max_members=64, chunk=20, precision='fdm_fp32', noise_sigma=1e-3,
                 on_finish=None, clock=time.perf_counter, sleep=time.sleep, quantum=None, policy='fifo',
                 prepare_workers=1):
        if policy not in ('fifo', 'smallest'):
            raise ValueError(f'unknown admission policy {policy!r}')
        self.budget = int(budget_bytes)
        self.max_members = int(max_members)
        self.chunk = int(chunk)
        self.quantum = None if quantum is None else max(1, int(quantum))
        self.precision = precision
        self.sigma = noise_sigma
        self.on_finish = on_finish
        self.clock = clock
        self.sleep = sleep
        self.policy = policy
        self.workers = max(0, int(prepare_workers))

    def _victims(self, head, active, used, busy=0):
        """Solves to preempt so that the queue head fits: those that ran a whole quantum, longest first, and
        only if together they make room for it (none when it fits already or cannot fit). used: bytes held
        by admitted solves (running or being prepared); busy: slots held by solves being prepared."""
        if not head.bytes:
            head.bytes = footprint(head.settings, self.precision)
        if head.bytes > self.budget:
            return []
        fits = lambda n, u: n + busy < self.max_members and u + head.bytes <= self.budget
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

    def _admissions(self, queue, busy, used):
        """(admit, reject): the waiting requests to admit now, in order, given `busy` slots and `used` bytes
        already held, and those that can never fit the budget. 'fifo' stops at the first request that does
        not fit; 'smallest' takes them by remaining work and passes over those that do not fit."""
        admit, reject = [], []
        order = sorted(queue, key=lambda r: r.work) if self.policy == 'smallest' else list(queue)
        for r in order:
            if busy + len(admit) >= self.max_members:
                break
            if not r.bytes:
                r.bytes = footprint(r.settings, self.precision)
            if r.bytes > self.budget:
                reject.append(r)
                continue
            if used + r.bytes > self.budget:
                if self.policy == 'fifo':
                    break
                continue
            admit.append(r)
            used += r.bytes
        return admit, reject

    def _prepare(self, r):
        """The solver of an admitted request: created, then restored from its snapshot or initialised."""
        g = Solver(r.settings, handle=True, precision=self.precision)
        try:
            if r.snap is not None:
                g.restore(r.snap)
            else:
                g.initialize(*W.noise(r.settings, r.seed, self.sigma))
        except Exception:
            g.close()
            raise
        return g

    def run(self, requests, duration=None, drain=True, progress=None, progress_every=30.0):
        """Serve `requests` (sorted by arrival). Stops admitting at `duration` seconds (None: never) and, with
        drain, finishes what is running; returns ServeStats. progress(now, completed, waiting, running), if given,
        is called every progress_every seconds of the loop's clock."""
        stats = ServeStats()
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue = deque()
        active = []                      # [(request, solver)]
        preparing = []                   # requests whose solver is being created
        used = 0
        ensemble = None
        pool = ThreadPoolExecutor(max_workers=self.workers) if self.workers else None
        t0 = self.clock()
        next_report = progress_every

        def drop_ensemble():
            nonlocal ensemble
            if ensemble is not None:
                ensemble.close()
                ensemble = None

        try:
            while True:
                now = self.clock() - t0
                closing = duration is not None and now >= duration
                while pending and pending[0].arrival <= now and not closing:
                    queue.append(pending.popleft())
                changed = False
                if self.quantum is not None and queue and not closing:
                    head = min(queue, key=lambda r: r.work) if self.policy == 'smallest' else queue[0]
                    victims = self._victims(head, active, used, len(preparing))
                    if victims:
                        drop_ensemble()
                        for r, g in victims:
                            r.snap = g.snapshot()
                            g.close()
                            used -= r.bytes
                            r.preemptions += 1
                            stats.preemptions += 1
                            queue.append(r)
                        active = [(r, g) for r, g in active if all(r is not q for q, _ in victims)]
                        changed = True
                # admission: reserve memory and a slot, prepare the solver
                admit, reject = self._admissions(queue, len(active) + len(preparing), used) if not closing else ([], [])
                for r in reject:
                    queue.remove(r)
                    r.error = 'footprint exceeds the device-memory budget'
                    stats.rejected.append(r)
                for r in admit:
                    queue.remove(r)
                    used += r.bytes
                    if r.snap is None:
                        r.admitted = self.clock() - t0
                    if pool is not None:
                        r.future = pool.submit(self._prepare, r)
                    preparing.append(r)
                # solvers that are ready join the ensemble (on the stepping path when there is no worker)
                still = []
                for r in preparing:
                    if pool is not None and not r.future.done():
                        still.append(r)
                        continue
                    try:
                        g = r.future.result() if pool is not None else self._prepare(r)
                    except Exception as e:          # creation failed: the request ends with the error
                        r.error = str(e)
                        used -= r.bytes
                        r.finished = self.clock() - t0
                        stats.completed.append(r)
                        continue
                    r.future = None
                    r.snap = None
                    r.run = 0
                    r.ready = self.clock() - t0
                    active.append((r, g))
                    changed = True
                preparing = still
                stats.queue_depth.append((now, len(queue), len(active) + len(preparing)))
                if progress is not None and now >= next_report:
                    progress(now, len(stats.completed), len(queue), len(active) + len(preparing))
                    next_report = now + progress_every
                if not active:
                    if preparing:
                        wait_futures([r.future for r in preparing], timeout=0.05, return_when=FIRST_COMPLETED)
                        continue
                    if closing or (not pending and not queue):
                        break
                    if pending and not queue:
                        wait = pending[0].arrival - (self.clock() - t0)
                        if wait > 0:
                            self.sleep(min(wait, 0.05))
                    continue
                if ensemble is None or changed:
                    drop_ensemble()
                    h0 = time.perf_counter()
                    ensemble = Ensemble([g for _, g in active])
                    stats.reforms += 1
                    stats.reform_ms += (time.perf_counter() - h0) * 1e3
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
                    drop_ensemble()
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
            drop_ensemble()
            for r, g in active:
                stats.unfinished.append(r)
                g.close()
            for r in preparing:
                if r.future is not None:
                    try:
                        r.future.result().close()
                    except Exception:
                        pass
                    r.future = None
                stats.unfinished.append(r)
            if pool is not None:
                pool.shutdown(wait=True)
            stats.unfinished.extend(queue)
            stats.unfinished.extend(pending)
        return stats
