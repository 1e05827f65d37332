"""The autotuner's launch choices (stage tile and z chunks, gradient z chunks, transform layout per axis) change
only the time: a cavity stepped with any of them has exactly the fields of the same cavity with the built-in
choices. The tuner keeps its pick per configuration (a second cavity of the same configuration gets the same
choices), and an ensemble tuned for stepping together still matches its members stepped alone."""
import numpy as np
import pytest

from conftest import FIELDS
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu

KNOB_SETS = [
    dict(stage_tile=1),
    dict(stage_tile=1, stage_target=64, stage_minz=16),
    dict(stage_tile=0, stage_target=4000, stage_minz=5),
    dict(stage_target=40, stage_minz=32),
    dict(grad_target=100000, grad_minz=5),
    dict(grad_target=30, grad_minz=16),
    dict(transform_layout=0b101010),        # every axis by whole columns
    dict(transform_layout=0b010101),        # every axis by chunks
    dict(transform_layout=0b011001, stage_tile=1, grad_target=50),
    dict(grad_tile=1),
    dict(grad_tile=2, grad_target=300, grad_minz=16),
    dict(grad_tile=3, grad_target=40),
]


def _make(settings, precision, seed=3):
    from les_impulse3d_64_amd.python_api import Solver
    g = Solver(settings, handle=True, precision=precision)
    g.initialize(*W.noise(settings, seed, 1e-2))
    return g


@pytest.mark.parametrize('precision', ['fp64', 'fdm_fp32', 'fdm_bf16x3', 'field_fp32'])
def test_launch_choices_give_identical_fields(precision):
    from les_impulse3d_64_amd.python_api import set_autotune
    print('case: launch choices', precision, flush=True)
    s = W.settings(48, 32, 40, re=900.0, stretch=1.8, lid_shaped=1, cfl_target=0.5, fdm_rtol=1e-10)
    set_autotune(False)
    try:
        ref = _make(s, precision)
        assert all(v == 0 for v in ref.tuning().values())
        ref_rep = ref.steps(4)
        want = {f: ref.field64(f) for f in FIELDS}
        for knobs in KNOB_SETS:
            g = _make(s, precision)
            g.set_tuning(**knobs)
            rep = g.steps(4)
            assert rep == ref_rep, (knobs, rep, ref_rep)
            for f in FIELDS:
                assert np.array_equal(g.field64(f), want[f]), (knobs, f)
            g.close()
        ref.close()
    finally:
        set_autotune(True)


def test_autotune_picks_caches_and_keeps_results():
    from les_impulse3d_64_amd.python_api import set_autotune
    print('case: autotune cache', flush=True)
    s = W.settings(64, 48, 40, re=700.0, stretch=1.5, lid_shaped=0, cfl_target=0.5)
    set_autotune(True)
    a = _make(s, 'fp64')
    b = _make(s, 'fp64')
    ta, tb = a.tuning(), b.tuning()
    print('tuned:', ta, flush=True)
    assert ta == tb                                   # the second creation reads the cache
    assert ta['stage_target'] > 0 and ta['grad_target'] > 0 and ta['stage_minz'] > 0
    set_autotune(False)
    try:
        c = _make(s, 'fp64')
    finally:
        set_autotune(True)
    assert all(v == 0 for v in c.tuning().values())
    ra, rc = a.steps(3), c.steps(3)
    assert ra == rc
    for f in FIELDS:
        assert np.array_equal(a.field64(f), c.field64(f)), f
    for g in (a, b, c):
        g.close()


def test_tuned_ensemble_matches_members_alone():
    from les_impulse3d_64_amd.python_api import Ensemble
    print('case: tuned ensemble', flush=True)
    cases = [W.ensemble_member(i) for i in (0, 6, 11, 24)]
    alone = []
    for s, noise in cases:
        from les_impulse3d_64_amd.python_api import Solver
        g = Solver(s, handle=True, precision='fdm_fp32')
        g.initialize(*noise)
        alone.append(g)
    batch = []
    for s, noise in cases:
        from les_impulse3d_64_amd.python_api import Solver
        g = Solver(s, handle=True, precision='fdm_fp32')
        g.initialize(*noise)
        batch.append(g)
    reps_alone = [g.steps(4) for g in alone]
    with Ensemble(batch) as e:
        e.autotune()
        tuned = [g.tuning() for g in batch]
        print('ensemble tuning:', tuned, flush=True)
        reps, errs = e.steps(4)
    assert errs == [''] * 4 and reps == reps_alone
    for k in range(4):
        for f in FIELDS:
            assert np.array_equal(batch[k].field64(f), alone[k].field64(f)), (k, f)
    for g in alone + batch:
        g.close()


def test_tuning_refusals():
    print('case: tuning refusals', flush=True)
    g = _make(W.settings(16, 16, 12), 'fp64')
    with pytest.raises(RuntimeError, match='tile'):
        g.set_tuning(stage_tile=2)
    with pytest.raises(RuntimeError, match='>= 0'):
        g.set_tuning(grad_target=-1)
    with pytest.raises(RuntimeError, match='layout bits'):
        g.set_tuning(transform_layout=64)
    with pytest.raises(RuntimeError, match='gradient tile'):
        g.set_tuning(grad_tile=4)
    g.close()
