"""Continuous batching: solves that arrive while others run are admitted between steps, leave when done, and
each result is bit for bit the result of the same solve run alone."""
import numpy as np
import pytest

from conftest import FIELDS
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu


def _requests():
    from les_impulse3d_64_amd.serving import Request
    specs = [  # (nx, ny, nz, Re, steps, arrival s)
        (16, 16, 12, 300.0, 9, 0.0),
        (24, 12, 16, 2000.0, 5, 0.0),
        (12, 20, 10, 150.0, 12, 0.01),
        (20, 20, 20, 5000.0, 7, 0.02),
        (16, 8, 24, 800.0, 4, 0.03),
        (32, 16, 16, 100.0, 6, 10.0),       # arrives after the others have left the queue
    ]
    out = []
    for k, (nx, ny, nz, re, steps, t) in enumerate(specs):
        s = W.settings(nx, ny, nz, re=re, lid_shaped=k % 2, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
        out.append(Request(k, s, steps, t, 500 + k))
    return out


def test_results_same_alone_and_in_the_mix():
    from les_impulse3d_64_amd.python_api import Solver
    from les_impulse3d_64_amd.serving import Server
    print('case: serving mix', flush=True)
    results = {}

    def keep(r, g):
        results[r.rid] = {f: g.field64(f) for f in FIELDS}

    clock_t = [0.0]

    def clock():              # a virtual clock: 5 ms per call, so arrivals interleave with chunks
        clock_t[0] += 0.005
        return clock_t[0]

    reqs = _requests()
    stats = Server(budget_bytes=64 << 30, max_members=3, chunk=3, precision='fdm_fp32', on_finish=keep,
                   clock=clock, sleep=lambda dt: None).run(reqs)
    assert len(stats.completed) == len(reqs) and not stats.rejected and not stats.unfinished
    assert max(q[2] for q in stats.queue_depth) == 3
    assert any(q[1] > 0 for q in stats.queue_depth)          # something waited for a slot
    assert len({r.admitted for r in reqs}) > 2               # solves joined at different times
    for r in reqs:
        assert r.error == '' and r.report['steps'] == r.steps and r.report['phi_rel'] <= 1e-10
        g = Solver(r.settings, handle=True, precision='fdm_fp32')
        g.initialize(*W.noise(r.settings, r.seed, 1e-3))
        rep = g.steps(r.steps)
        assert rep == r.report, (r.rid, rep, r.report)
        for f in FIELDS:
            assert np.array_equal(g.field64(f), results[r.rid][f]), (r.rid, f)
        g.close()


def test_admission_by_footprint():
    from les_impulse3d_64_amd.python_api import footprint
    from les_impulse3d_64_amd.serving import Server
    print('case: serving admission', flush=True)
    reqs = _requests()[:4]
    sizes = [footprint(r.settings, 'fdm_fp32') for r in reqs]
    budget = max(sizes) + min(sizes) - 1        # never room for the two largest together
    big = _requests()[5]
    big.settings = W.settings(64, 64, 64, re=100.0)
    big.arrival = 0.0
    stats = Server(budget_bytes=budget, max_members=8, chunk=2).run(reqs + [big])
    assert [r.rid for r in stats.rejected] == [big.rid]
    assert len(stats.completed) == len(reqs)
    for r in stats.completed:
        assert r.bytes == footprint(r.settings, 'fdm_fp32')


def test_preemption_keeps_results_same_as_alone():
    """Two slots, a quantum of 2 steps: waiting solves preempt running ones (snapshot to host memory, slot
    freed, resumed later from the snapshot), and every result is still bit for bit the solo run."""
    from les_impulse3d_64_amd.python_api import Solver
    from les_impulse3d_64_amd.serving import Server
    print('case: serving with preemption', flush=True)
    results = {}

    def keep(r, g):
        results[r.rid] = {f: g.field64(f) for f in FIELDS}

    clock_t = [0.0]

    def clock():
        clock_t[0] += 0.005
        return clock_t[0]

    reqs = _requests()
    stats = Server(budget_bytes=64 << 30, max_members=2, chunk=2, quantum=2, precision='fdm_fp32',
                   on_finish=keep, clock=clock, sleep=lambda dt: None).run(reqs)
    assert len(stats.completed) == len(reqs) and not stats.unfinished
    assert stats.preemptions > 0 and stats.summary()['preemptions'] == stats.preemptions
    assert sum(r.preemptions for r in reqs) == stats.preemptions
    assert max(q[2] for q in stats.queue_depth) == 2
    for r in reqs:
        assert r.error == '' and r.report['steps'] == r.steps and r.snap is None
        g = Solver(r.settings, handle=True, precision='fdm_fp32')
        g.initialize(*W.noise(r.settings, r.seed, 1e-3))
        rep = g.steps(r.steps)
        assert rep == r.report, (r.rid, rep, r.report)
        for f in FIELDS:
            assert np.array_equal(g.field64(f), results[r.rid][f]), (r.rid, f)
        g.close()
