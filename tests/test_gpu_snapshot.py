"""Whole-state snapshots: a solver restored from a snapshot (host or device memory) continues bit for bit as
the solver it was taken from, the warm-start history, the lagged dt rates and the statistics sums included;
a snapshot of another configuration, mode or size is refused."""
import numpy as np
import pytest

from conftest import FIELDS
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu

PROBES = ((.5, .5), (.8, .8), (.2, .2), (.9, .1))


def _settings(**kw):
    s = W.settings(24, 24, 16, re=800.0, stretch=1.5, lid_shaped=1, adaptive_dt=1, cfl_target=0.5,
                   fdm_rtol=1e-10, **kw)
    s['probes'] = PROBES
    return s


def _continue(g, steps, kick=None):
    """Steps with statistics weighted by dt and a kick halfway; everything the run leaves behind."""
    out = {'probes': [], 'lines': []}
    for k in range(steps):
        rep = g.step()
        out['probes'].append(g.probes())
        out['lines'].append(g.add_statistics(rep['dt']))
        if kick is not None and k == steps // 2:
            g.kick(kick, 0.05, 1)
    out['report'] = g.report()
    out['fields'] = {f: g.field64(f) for f in FIELDS}
    out['stats'] = g.statistics()
    out['passes'] = g.pass_counters()
    return out


def _same(a, b):
    assert a['report'] == b['report']
    assert a['passes'] == b['passes']
    for f in FIELDS:
        assert np.array_equal(a['fields'][f], b['fields'][f]), f
    for k in a['stats']:
        assert np.array_equal(a['stats'][k], b['stats'][k]), k
    assert all(np.array_equal(x, y) for x, y in zip(a['probes'], b['probes']))
    assert all(np.array_equal(x, y) for x, y in zip(a['lines'], b['lines']))


@pytest.mark.parametrize('mode', ['fp64', 'fdm_bf16x3', 'field_fp32'])
def test_restored_solver_continues_bit_for_bit(hip_lib, mode):
    from les_impulse3d_64_amd.python_api import Solver
    s = _settings()
    print(f'case: snapshot continuation, {mode}', flush=True)
    a = Solver(s, handle=True, precision=mode)
    a.initialize(*W.noise(s, 7, 1e-3))
    first = _continue(a, 5, kick=W.kick_noise(s, 70))          # history, lagged rates and statistics in use
    assert a.statistics()['uu'].max() > 0
    snap = a.snapshot()
    assert snap.dtype == np.uint8 and snap.size == a.snapshot_bytes()
    kick = W.kick_noise(s, 71)
    ref = _continue(a, 7, kick=kick)
    b = Solver(s, handle=True, precision=mode)                  # a fresh solver, never initialised
    b.restore(snap)
    assert b.report() == first['report']
    got = _continue(b, 7, kick=kick)
    _same(ref, got)
    print(f"{mode}: {ref['report']['steps']} steps, passes {ref['passes']}", flush=True)
    a.close()
    b.close()


def test_device_resident_snapshot_and_rollback(hip_lib):
    from les_impulse3d_64_amd.python_api import DeviceBuffer, Solver
    s = _settings()
    print('case: device-resident snapshot', flush=True)
    a = Solver(s, handle=True, precision='fdm_fp32')
    a.initialize(*W.noise(s, 8, 1e-3))
    _continue(a, 4)
    dev = DeviceBuffer(a.snapshot_bytes())
    a.snapshot(out=dev)
    host = a.snapshot()
    assert np.array_equal(dev.numpy(), host)
    ref = _continue(a, 5)
    a.restore(dev)                                              # roll the same solver back
    again = _continue(a, 5)
    _same(ref, again)
    b = Solver(s, handle=True, precision='fdm_fp32')
    b.restore(dev)
    _same(ref, _continue(b, 5))
    dev.close()
    a.close()
    b.close()


def test_restore_refuses_foreign_snapshots(hip_lib):
    from les_impulse3d_64_amd.python_api import Solver
    s = _settings()
    print('case: snapshot refusals', flush=True)
    a = Solver(s, handle=True)
    a.initialize(*W.noise(s, 9, 1e-3))
    a.step()
    snap = a.snapshot()
    other = Solver(_settings(dt_max=0.002), handle=True)
    with pytest.raises(RuntimeError, match='another configuration'):
        other.restore(snap)
    mixed = Solver(s, handle=True, precision='fdm_fp32')
    with pytest.raises(RuntimeError, match='precision mode'):
        mixed.restore(snap)
    with pytest.raises(RuntimeError, match='size mismatch'):
        a.restore(snap[:-8].copy())
    bad = snap.copy()
    bad[:8] = 0
    with pytest.raises(RuntimeError, match='not a cavity snapshot'):
        a.restore(bad)
    with pytest.raises(RuntimeError, match='needs'):
        a.snapshot(out=np.empty(100, dtype=np.uint8))
    before = a.field64('u')
    a.step()                                                    # the refused calls left the solver usable
    assert not np.array_equal(before, a.field64('u'))
    for g in (a, other, mixed):
        g.close()


@pytest.mark.parametrize('mode', ['fp64', 'fdm_fp32'])
def test_clock_change_forms_the_first_guess_again(hip_lib, mode):
    """A step's last gradient forms the next step's first warm-start guess for the clock it expects. When the host
    moves the clock in between, that step forms its guess again: it matches, bit for bit, a solver restored from
    a snapshot, which always forms the guess itself, with the same clock change."""
    from les_impulse3d_64_amd.python_api import Solver
    s = _settings()
    print(f'case: clock change between steps, {mode}', flush=True)
    a = Solver(s, handle=True, precision=mode)
    a.initialize(*W.noise(s, 10, 1e-3))
    _continue(a, 4)                                             # three history entries and the lagged rates
    snap = a.snapshot()
    b = Solver(s, handle=True, precision=mode)
    b.restore(snap)
    rep = a.report()
    assert rep['steps'] == 4 and rep['dt'] > 0
    for g in (a, b):
        g.set_clock(rep['time'] + 0.37 * rep['dt'], rep['steps'], rep['dt'])
    moved = _continue(a, 3)
    _same(moved, _continue(b, 3))
    # the comparison sees the guess: the same steps on the unmoved clock (a guess extrapolated to other times)
    # end on other bits
    c = Solver(s, handle=True, precision=mode)
    c.restore(snap)
    still = _continue(c, 3)
    assert any(not np.array_equal(moved['fields'][f], still['fields'][f]) for f in FIELDS)
    for g in (a, b, c):
        g.close()
