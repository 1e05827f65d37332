"""Batched stepping: an ensemble of cavities with different sizes, grids, Reynolds numbers and lids,
advanced by shared launches, gives each member exactly the fields it has when stepped alone.

Every kernel gives a member the same blocks, tiles and reduction order whether it runs alone or in a batch,
so the comparison is bit for bit. The graph-replayed step and the kernel-by-kernel step are also identical,
and a member stepped alone agrees with the FP64 host oracle to the parity bar.
"""
import numpy as np
import pytest

from conftest import FIELDS, rel_err
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [
    W.settings(32, 32, 24, re=400.0, stretch=1.7, lid_shaped=1, cfl_target=0.5),
    W.settings(24, 16, 20, re=2000.0, stretch=1.2, lid_shaped=0, cfl_target=0.5),
    W.settings(16, 24, 13, re=150.0, lid_shaped=1, adaptive_dt=0, dt_fixed=2e-3),     # odd nz: dense z axis
    W.settings(48, 48, 32, re=3000.0, stretch=2.0, lid_shaped=1, cfl_target=0.5, fdm_rtol=1e-10),
]


def _make(settings, seed, precision):
    from les_impulse3d_64_amd.python_api import Solver
    g = Solver(settings, handle=True, precision=precision)
    g.initialize(*W.noise(settings, seed, 1e-2))
    return g


@pytest.mark.parametrize('precision', ['fp64', 'fdm_bf16x3', 'field_fp32'])
def test_ensemble_matches_members_alone(precision):
    from les_impulse3d_64_amd.python_api import Ensemble
    print('case: ensemble', precision, flush=True)
    alone = [_make(s, 10 + k, precision) for k, s in enumerate(CASES)]
    batch = [_make(s, 10 + k, precision) for k, s in enumerate(CASES)]
    steps = 6
    solo_reports = [g.steps(steps) for g in alone]
    with Ensemble(batch) as e:
        reports, errors = e.steps(steps)
    assert errors == [''] * len(CASES)
    for k in range(len(CASES)):
        for key in ('steps', 'time', 'dt', 'cfl', 'phi_rel', 'dt_reason', 'phi_passes'):
            assert reports[k][key] == solo_reports[k][key], (k, key, reports[k][key], solo_reports[k][key])
        for f in FIELDS:
            a, b = batch[k].field64(f), alone[k].field64(f)
            assert np.array_equal(a, b), (k, f, rel_err(a, b))
        assert batch[k].pass_counters() == alone[k].pass_counters()
    # members keep working alone afterwards, and stay in step with their twins
    r1, r2 = batch[0].step(), alone[0].step()
    assert r1 == r2 and np.array_equal(batch[0].field64('u'), alone[0].field64('u'))
    for g in alone + batch:
        g.close()


def test_ensemble_member_matches_oracle():
    from les_impulse3d_64_amd.python_api import Ensemble
    from oracle.oracle import Oracle
    print('case: ensemble vs oracle', flush=True)
    s = CASES[1]
    noise = W.noise(s, 7, 1e-2)
    o = Oracle(s)
    o.initialize(*noise)
    gs = [_make(c, 20 + k, 'fp64') for k, c in enumerate(CASES)]
    gs[1].initialize(*noise)
    with Ensemble(gs) as e:
        for _ in range(5):
            e.steps(1)
            o.step()
    worst = {f: rel_err(gs[1].field64(f), o.field(f)) for f in FIELDS}
    assert max(worst.values()) <= 1e-11, worst
    for g in gs:
        g.close()


def test_graph_and_plain_launches_agree():
    print('case: graph vs plain', flush=True)
    s = CASES[0]
    a = _make(s, 3, 'fp64')
    b = _make(s, 3, 'fp64')
    b.set_graph(False)
    ra, rb = a.steps(5), b.steps(5)
    assert ra == rb
    for f in FIELDS:
        assert np.array_equal(a.field64(f), b.field64(f)), f
    a.close()
    b.close()


def test_ensemble_errors_are_per_member():
    from les_impulse3d_64_amd.python_api import Ensemble
    print('case: ensemble error isolation', flush=True)
    gs = [_make(c, 30 + k, 'fp64') for k, c in enumerate(CASES[:3])]
    twin = _make(CASES[2], 32, 'fp64')
    u = gs[1].field64('u')
    u[2, 3, 4] = np.nan
    gs[1].set_field64('u', u)
    with Ensemble(gs) as e:
        reports, errors = e.steps(2)
    assert 'NaN' in errors[1] and errors[0] == '' and errors[2] == ''
    twin.steps(2)
    for f in FIELDS:
        assert np.array_equal(gs[2].field64(f), twin.field64(f)), f
    # the failed member recovers once its fields are set again
    gs[1].initialize(*W.noise(CASES[1], 5, 1e-2))
    gs[1].step()
    for g in gs + [twin]:
        g.close()


def test_mixed_modes_rejected_and_member_destroy_guarded():
    from les_impulse3d_64_amd.python_api import Ensemble
    print('case: ensemble guards', flush=True)
    a = _make(CASES[0], 1, 'fp64')
    b = _make(CASES[1], 2, 'fdm_fp32')
    with pytest.raises(RuntimeError, match='precision mode'):
        Ensemble([a, b])
    b.set_precision('fp64')
    e = Ensemble([a, b])
    assert a.lib.cav_destroy(a._h) != 0 and 'ensemble' in a.lib.cav_last_error().decode()
    e.close()
    a.close()
    b.close()


@pytest.mark.parametrize('precision', ['fp64', 'fdm_fp32', 'field_fp32'])
def test_stalled_steps_are_redone_identically(precision):
    """A step enqueues the usual number of refinement passes; a projection needing more stops the step with
    the state untouched and the step is redone with every pass. Forcing one pass per step makes the redo
    path run and must not change a bit."""
    print('case: redo path', precision, flush=True)
    s = CASES[3]
    a = _make(s, 41, precision)
    b = _make(s, 41, precision)
    b.set_enqueued_passes(1)
    ra, rb = a.steps(6), b.steps(6)
    assert ra == rb
    for f in FIELDS:
        assert np.array_equal(a.field64(f), b.field64(f)), f
    assert a.pass_counters() == b.pass_counters()
    if precision != 'fp64':
        assert b.redone_steps() >= 1          # the cold start needs two FP32 passes
    assert a.redone_steps() == 0
    a.close()
    b.close()


def test_a_stall_adds_a_spare_pass():
    """After its first stall a cavity enqueues one pass more, so later projections that need it run in the
    replay: with one pass forced, the FP32 cold start stalls once and the other 29 steps (1 to 2 passes
    each) run without a redo, bit for bit as with the default count."""
    print('case: spare pass after a stall', flush=True)
    s = CASES[3]
    a = _make(s, 43, 'fdm_fp32')
    b = _make(s, 43, 'fdm_fp32')
    b.set_enqueued_passes(1)
    ra, rb = a.steps(30), b.steps(30)
    assert ra == rb
    for f in FIELDS:
        assert np.array_equal(a.field64(f), b.field64(f)), f
    projections, passes = a.pass_counters()
    assert passes > projections                   # some projections took two passes
    assert a.redone_steps() == 0
    assert b.redone_steps() == 1
    a.close()
    b.close()


def test_ensemble_redo_catches_up():
    from les_impulse3d_64_amd.python_api import Ensemble
    print('case: ensemble redo', flush=True)
    alone = [_make(c, 50 + k, 'fdm_fp32') for k, c in enumerate(CASES)]
    batch = [_make(c, 50 + k, 'fdm_fp32') for k, c in enumerate(CASES)]
    for g in alone:
        g.steps(20)
    with Ensemble(batch) as e:
        e.set_enqueued_passes(1)
        reports, errors = e.steps(20)
    assert errors == [''] * len(CASES)
    for k in range(len(CASES)):
        assert reports[k] == alone[k].report()
        for f in FIELDS:
            assert np.array_equal(batch[k].field64(f), alone[k].field64(f)), (k, f)
        assert batch[k].redone_steps() > 0
    for g in alone + batch:
        g.close()


def test_configured_64_member_ensemble_matches_members_alone():
    """The 64-cavity ensemble workload itself (workloads.ensemble_member: every size in {32, 48, 64, 96} per
    axis, stretch U[1, 2], Re log-U[100, 5000], both lids, FP32 FDM at rtol 1e-10), stepped together:
    every member's reports and fields equal its solo run bit for bit."""
    from les_impulse3d_64_amd.python_api import Ensemble, Solver
    print('case: configured 64-member ensemble vs solo', flush=True)
    steps = 3
    members = [W.ensemble_member(i) for i in range(64)]
    sizes = {n for s, _ in members for n in (s['nx'], s['ny'], s['nz'])}
    assert sizes == {32, 48, 64, 96} and {s['lid_shaped'] for s, _ in members} == {0, 1}
    batch = []
    for s, noise in members:
        g = Solver(s, handle=True, precision='fdm_fp32')
        g.initialize(*noise)
        batch.append(g)
    with Ensemble(batch) as e:
        reports, errors = e.steps(steps)
    assert errors == [''] * 64, errors
    for i, (s, noise) in enumerate(members):
        g = Solver(s, handle=True, precision='fdm_fp32')
        g.initialize(*noise)
        rep = g.steps(steps)
        assert rep == reports[i], (i, rep, reports[i])
        assert rep['phi_rel'] <= 1e-10
        for f in FIELDS:
            assert np.array_equal(g.field64(f), batch[i].field64(f)), (i, f)
        g.close()
        batch[i].close()
    print(f'configured ensemble: 64 members equal to their solo runs after {steps} steps', flush=True)


def test_configured_ensemble_members_match_oracle():
    """Two configured members (32^3 uniform lid, 48x32x32 regularised lid) stepped inside an FP64 ensemble of
    eight configured members agree with the FP64 oracle to 1e-11 over 5 steps."""
    from les_impulse3d_64_amd.python_api import Ensemble, Solver
    from oracle.oracle import Oracle
    print('case: configured ensemble members vs oracle', flush=True)
    pick = [6, 2, 12, 17, 23, 34, 35, 56]
    gs, orcs = [], {}
    for i in pick:
        s, noise = W.ensemble_member(i)
        g = Solver(s, handle=True, precision='fp64')
        g.initialize(*noise)
        gs.append(g)
        if i in (6, 2):
            o = Oracle(s)
            o.initialize(*noise)
            orcs[i] = o
    with Ensemble(gs) as e:
        for _ in range(5):
            e.steps(1)
            for o in orcs.values():
                o.step()
    for i, o in orcs.items():
        g = gs[pick.index(i)]
        worst = {f: rel_err(g.field64(f), o.field(f)) for f in FIELDS}
        print(f'member {i}: worst {max(worst.values()):.2e}', flush=True)
        assert max(worst.values()) <= 1e-11, (i, worst)
        r, ro = g.report(), o.report()
        assert r['steps'] == ro['steps'] and r['dt_reason'] == ro['dt_reason']
        assert r['dt'] == pytest.approx(ro['dt'], rel=1e-12)
        o.close()
    for g in gs:
        g.close()
