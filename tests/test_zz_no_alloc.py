"""The no-allocation batch entry point (cav_batch_steps_noalloc) against the ensemble entry points it is a
variant of: the same members stepped either way give the same reports and fields bit for bit, on two calls in
a row, the second with fewer and smaller cavities (the plan array keeps the first call's plan, which the second
rewrites), in FP64 and with the FP32 FDM forced through the redo path."""
import numpy as np
import pytest

from conftest import FIELDS
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    W.settings(48, 32, 40, re=800.0, stretch=1.8, lid_shaped=1, cfl_target=0.5, fdm_rtol=1e-10),
    W.settings(32, 32, 24, re=400.0, stretch=1.7, lid_shaped=0, cfl_target=0.5),
    W.settings(24, 16, 20, re=2000.0, stretch=1.2, lid_shaped=1, cfl_target=0.5),
    W.settings(16, 24, 13, re=150.0, lid_shaped=1, adaptive_dt=0, dt_fixed=2e-3),
    W.settings(64, 48, 32, re=3000.0, stretch=2.0, lid_shaped=0, cfl_target=0.5, fdm_rtol=1e-10),
]


def _make(settings, seed, precision):
    from les_impulse3d_64_amd.python_api import Solver
    g = Solver(settings, handle=True, precision=precision)
    g.initialize(*W.noise(settings, seed, 1e-2))
    return g


def _same(a, b, label):
    for f in FIELDS:
        assert np.array_equal(a.field64(f), b.field64(f)), (label, f)


@pytest.mark.parametrize('precision,passes', [('fp64', 0), ('fdm_fp32', 1), ('field_fp32', 0)])
def test_noalloc_batch_equals_ensemble(precision, passes):
    from les_impulse3d_64_amd.python_api import Ensemble, batch_steps_noalloc
    print('case: no-allocation batch vs ensemble', precision, passes, flush=True)
    ens = [_make(s, 60 + k, precision) for k, s in enumerate(CASES)]
    noa = [_make(s, 60 + k, precision) for k, s in enumerate(CASES)]
    for g in ens + noa:
        if passes:
            g.set_enqueued_passes(passes)          # one pass per projection: FP32 stalls and is redone
    # first call: all five cavities, 7 steps
    with Ensemble(ens) as e:
        if passes:
            e.set_enqueued_passes(passes)
        r_e, err_e = e.steps(7)
    r_n, err_n = batch_steps_noalloc(noa, 7)
    assert err_e == err_n == [''] * len(CASES)
    assert r_e == r_n
    for k in range(len(CASES)):
        _same(ens[k], noa[k], k)
        assert ens[k].pass_counters() == noa[k].pass_counters()
    if passes:
        assert all(g.redone_steps() > 0 for g in noa)
    # second call: the two smallest cavities only, 5 more steps
    pick = [3, 2]
    with Ensemble([ens[k] for k in pick]) as e:
        if passes:
            e.set_enqueued_passes(passes)
        r_e, err_e = e.steps(5)
    r_n, err_n = batch_steps_noalloc([noa[k] for k in pick], 5)
    assert err_e == err_n == ['', ''] and r_e == r_n
    for k in pick:
        _same(ens[k], noa[k], k)
        assert noa[k].report()['steps'] == 12
    for k in (0, 1, 4):
        assert noa[k].report()['steps'] == 7
    for g in ens + noa:
        g.close()


def test_noalloc_refusals():
    from les_impulse3d_64_amd.python_api import batch_steps_noalloc
    print('case: no-allocation refusals', flush=True)
    a = _make(CASES[2], 1, 'fp64')
    b = _make(CASES[3], 2, 'fdm_fp32')
    with pytest.raises(RuntimeError, match='precision mode'):
        batch_steps_noalloc([a, b], 1)
    with pytest.raises(RuntimeError, match='appears twice'):
        batch_steps_noalloc([a, a], 1)
    lib = a.lib
    assert lib.cav_batch_steps_noalloc(None, 257, 1, None, None) == 1
    assert b'between 1 and 256' in lib.cav_last_error()
    assert batch_steps_noalloc([a], 0)[0][0]['steps'] == 0
    a.close()
    b.close()
