"""Parity of the HIP solver (through the C-ABI) with the FP64 host oracle on the same inputs.

Bar: after one step every field within 1e-12 of the oracle (max|a-b|/max|b|, validate_gpu.py:83); over
the 200-step 32^3 parity run within 1e-9. The results are not bit-identical: the stencil kernels multiply
by host-built reciprocal and coefficient tables where the oracle divides (about one ulp per operation), the
FDM transforms accumulate in MFMA order and the reductions sum in tree order.
"""
import numpy as np
import pytest

from conftest import FIELDS, rel_err
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu


def _pair(settings, noise=None, handle=True):
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    g = Solver(settings, handle=handle)
    o = Oracle(settings)
    v, w = noise if noise is not None else (None, None)
    g.initialize(v, w)
    o.initialize(v, w)
    return g, o


def _compare(g, o, tol, label=''):
    worst = {}
    for f in FIELDS:
        worst[f] = rel_err(g.field64(f), o.field(f))
    bad = {k: v for k, v in worst.items() if v > tol}
    assert not bad, f'{label} fields above {tol:g}: {bad} (all: {worst})'
    return worst


def _reports_match(rg, ro):
    assert rg['steps'] == ro['steps']
    assert rg['dt_reason'] == ro['dt_reason']
    assert rg['phi_passes'] == ro['phi_passes']
    assert abs(rg['time'] - ro['time']) <= 1e-14 * max(1.0, abs(ro['time']))
    assert abs(rg['dt'] - ro['dt']) <= 1e-13 * abs(ro['dt'])


def test_initialize_projection():
    s, noise, _ = W.c0_parity()
    print('case: c0 initialize', flush=True)
    g, o = _pair(s, noise)
    _compare(g, o, 1e-12, 'after initialize')
    # walls exactly zero (validate_gpu.py:48)
    u, v, w = g.field64('u'), g.field64('v'), g.field64('w')
    assert np.abs(u[..., [0, -1]]).max() == 0 and np.abs(v[:, [0, -1]]).max() == 0 and np.abs(w[[0, -1]]).max() == 0
    g.close()


def test_one_step_c0():
    s, noise, _ = W.c0_parity()
    print('case: c0 one step', flush=True)
    g, o = _pair(s, noise)
    rg, ro = g.step(), o.step()
    _reports_match(rg, ro)
    _compare(g, o, 1e-12, 'c0 step 1')
    g.close()


def test_c0_200_steps():
    s, noise, nsteps = W.c0_parity()
    print('case: c0 200 steps', flush=True)
    g, o = _pair(s, noise)
    for k in range(nsteps):
        rg, ro = g.step(), o.step()
        if k in (0, 9, 99):
            _compare(g, o, 1e-12 if k == 0 else 1e-10, f'c0 step {k + 1}')
    _reports_match(rg, ro)
    worst = _compare(g, o, 1e-9, 'c0 step 200')
    print('c0 200-step worst relative differences:', worst, flush=True)
    pg, po = g.pass_counters(), o.pass_counters()
    assert pg == po
    g.close()


@pytest.mark.parametrize('case', [
    dict(nx=16, ny=16, nz=12, stretch=2.0, lid_shaped=1, impulse_order=4, dt_lag=1),
    dict(nx=16, ny=24, nz=10, stretch=1.5, lid_shaped=0, impulse_order=4, dt_lag=0),
    dict(nx=12, ny=12, nz=7, stretch=1.0, lid_shaped=1, impulse_order=2, dt_lag=1),
    dict(nx=20, ny=14, nz=9, stretch=3.0, lid_shaped=1, impulse_order=4, dt_lag=1, impulse_order4_ramp=3.0,
         phi_extrapolate=0),
])
def test_ten_steps_adaptive(case):
    case = dict(case)
    nx, ny, nz, stretch = case.pop('nx'), case.pop('ny'), case.pop('nz'), case.pop('stretch')
    s = W.settings(nx, ny, nz, re=400.0, stretch=stretch, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10, **case)
    print('case:', s, flush=True)
    g, o = _pair(s, W.noise(s, 7, 1e-2))
    for k in range(10):
        rg, ro = g.step(), o.step()
        _reports_match(rg, ro)
        _compare(g, o, 1e-11, f'step {k + 1}')
    g.close()


@pytest.mark.parametrize('case', [
    dict(nx=6, ny=6, nz=5, stretch=1.0, adaptive_dt=0, impulse_order=4, lid_shaped=1),
    dict(nx=6, ny=10, nz=5, stretch=2.0, adaptive_dt=1, impulse_order=4, lid_shaped=0),
    dict(nx=8, ny=8, nz=6, stretch=1.5, adaptive_dt=0, impulse_order=2, lid_shaped=1),
    dict(nx=10, ny=6, nz=9, stretch=2.5, adaptive_dt=1, impulse_order=4, lid_shaped=1, dt_lag=0),
])
def test_smallest_grids(case):
    """The smallest grids the library accepts (nx, ny even >= 6, nz >= 5): a plane is less than one 32 x 8
    column tile, every cell touches a wall or lies next to one, and the wide stencil's reach crosses the box."""
    case = dict(case)
    nx, ny, nz, stretch = case.pop('nx'), case.pop('ny'), case.pop('nz'), case.pop('stretch')
    extra = dict(dt_fixed=2e-3) if case['adaptive_dt'] == 0 else dict(cfl_target=0.5)
    s = W.settings(nx, ny, nz, re=300.0, stretch=stretch, fdm_rtol=1e-10, **extra, **case)
    print('case:', s, flush=True)
    g, o = _pair(s, W.noise(s, 11, 1e-2))
    for k in range(10):
        rg, ro = g.step(), o.step()
        _reports_match(rg, ro)
        _compare(g, o, 1e-11, f'step {k + 1}')
    g.close()


@pytest.mark.parametrize('grid', [
    dict(nx=24, ny=24, nz=16, stretch=2.0, lid_shaped=1),
    dict(nx=32, ny=20, nz=11, stretch=1.5, lid_shaped=0, impulse_order=2),   # no centreline (nx != ny)
])
def test_kick_and_outputs(grid):
    grid = dict(grid)
    nx, ny, nz = grid.pop('nx'), grid.pop('ny'), grid.pop('nz')
    s = W.settings(nx, ny, nz, re=800.0, adaptive_dt=1, cfl_target=0.5, probes=W.PROBES8, fdm_rtol=1e-10, **grid)
    print('case: kick/outputs', s, flush=True)
    g, o = _pair(s, W.noise(s, 11, 1e-2))
    for _ in range(3):
        g.step(), o.step()
    kn = W.kick_noise(s, 2001)
    g.kick(kn, 0.05, 2)
    o.kick(kn, 0.05, 2)
    _compare(g, o, 1e-11, 'after kick')
    for _ in range(2):
        g.step(), o.step()
    _compare(g, o, 1e-11, 'after kick + 2 steps')
    dg, do = g.diagnostics(), o.diagnostics()
    for k in ('kinetic_energy', 'enstrophy', 'umax', 'cfl_rate', 'gauge_ratio'):
        assert abs(dg[k] - do[k]) <= 1e-10 * abs(do[k]), (k, dg[k], do[k])
    assert dg['div_rms'] < 1e-9 and do['div_rms'] < 1e-9
    assert (dg['gauge_x'], dg['gauge_y']) == (do['gauge_x'], do['gauge_y'])
    np.testing.assert_allclose(g.probes(), o.probes().astype(np.float32), rtol=1e-6, atol=1e-7)
    lg, lo = g.add_statistics(0.25), o.add_statistics(0.25)
    if nx == ny:
        assert rel_err(lg, lo) < 1e-11
    else:
        assert lg is None and lo is None
    sg, so = g.statistics(), o.statistics()
    for k in sg:
        assert rel_err(sg[k], so[k]) < 1e-11, k
    mg, mo = g.midplane(), o.midplane()
    assert rel_err(mg, mo) < 1e-6
    qg, cg = g.q_volume(True)
    qo, co = o.q_volume(True)
    assert rel_err(qg, qo) < 1e-6 and rel_err(cg, co) < 1e-6
    g.close()


def test_singleton_abi_matches_handle():
    """The reference boundary (solver_* on one process-wide solver) gives the handle path's numbers."""
    from les_impulse3d_64_amd.python_api import Solver
    s, noise, _ = W.c0_parity()
    s = dict(s, nz=16)
    print('case: singleton', flush=True)
    a = Solver(s)                     # solver_* entry points
    b = Solver(s, handle=True)
    a.initialize(*W.noise(s, 0, 1e-3))
    b.initialize(*W.noise(s, 0, 1e-3))
    for _ in range(3):
        ra, rb = a.step(), b.step()
    assert ra == rb
    for f in FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    # float32 field round trip through set_field (float32 I/O of the reference ABI)
    u32 = a.field('u')
    a.set_field('u', u32)
    assert np.array_equal(a.field('u'), u32)
    a.close()
    b.close()


def test_error_paths():
    from les_impulse3d_64_amd.python_api import Solver
    s = W.settings(12, 12, 8, re=100.0)
    print('case: errors', flush=True)
    with pytest.raises(RuntimeError, match='even'):
        Solver(dict(s, nx=13), handle=True)
    with pytest.raises(RuntimeError, match='nz >= 5'):
        Solver(dict(s, nz=4), handle=True)
    with pytest.raises(RuntimeError, match='IMPULSE_ORDER'):
        Solver(dict(s, impulse_order=3), handle=True)
    g = Solver(s, handle=True)
    g.initialize(*W.noise(s, 1, 1e-2))
    u = g.field64('u')
    u[2, 3, 4] = np.nan
    g.set_field64('u', u)
    with pytest.raises(RuntimeError, match='NaN'):
        g.step()
    g.close()
    g = Solver(dict(s, fdm_max_passes=0), handle=True)
    with pytest.raises(RuntimeError, match='after 0 passes'):
        g.initialize(*W.noise(s, 1, 1e-2))
    g.close()


@pytest.mark.parametrize('mode', ['fdm_fp32', 'fdm_bf16x3'])
def test_mixed_precision_fdm(mode):
    """FP32 / split-BF16 transforms with FP64 refinement: every projection reaches FDM_RTOL = 1e-10 (the
    relative post-projection divergence) and the fields stay within 1e-7 of the FP64 oracle over 10 steps."""
    s = dict(W.settings(32, 32, 24, re=1000.0, stretch=2.0, lid_shaped=1, adaptive_dt=1, cfl_target=0.5), fdm_rtol=1e-10)
    print('case: mixed', mode, flush=True)
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    noise = W.noise(s, 2, 1e-2)
    g = Solver(s, handle=True, precision=mode)
    o = Oracle(s)
    g.initialize(*noise)
    o.initialize(*noise)
    for k in range(10):
        rg, ro = g.step(), o.step()
        assert rg['phi_rel'] <= 1e-10 and rg['dt_reason'] == ro['dt_reason']
    worst = _compare(g, o, 1e-7, f'{mode} after 10 steps')
    proj, passes = g.pass_counters()
    print(f'{mode}: worst field difference vs FP64 oracle {max(worst.values()):.2e}, '
          f'{passes / proj:.2f} FDM passes per projection', flush=True)
    assert passes / proj <= 4.0
    g.close()


@pytest.mark.parametrize('mode', ['fdm_fp32', 'fdm_bf16x3'])
def test_mixed_fdm_restarts_from_the_rhs(mode):
    """A warm-start guess worse than none (its residual at least the right-hand side) restarts the projection from
    zero, the first pass on the rhs, which the float-transform modes read from their float residual array. A
    potential the host sets to noise far above the field forces it: the steps still reach FDM_RTOL and stay
    within 1e-7 of the FP64 oracle given the same potential."""
    s = dict(W.settings(32, 32, 24, re=1000.0, stretch=2.0, lid_shaped=1, adaptive_dt=1, cfl_target=0.5), fdm_rtol=1e-10)
    print('case: mixed restart from the rhs', mode, flush=True)
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    noise = W.noise(s, 3, 1e-2)
    g = Solver(s, handle=True, precision=mode)
    o = Oracle(s)
    g.initialize(*noise)
    o.initialize(*noise)
    for _ in range(3):
        g.step()
        o.step()
    phi = g.field64('phi')
    garbage = np.random.default_rng(5).standard_normal(phi.shape) * (1e3 * np.abs(phi).max())
    g.set_field64('phi', garbage)
    o.set_field('phi', garbage)
    p0 = g.pass_counters()
    for _ in range(2):
        rg, ro = g.step(), o.step()
        assert rg['phi_rel'] <= 1e-10 and rg['dt_reason'] == ro['dt_reason']
    worst = _compare(g, o, 1e-7, f'{mode} after a potential of noise')
    p1 = g.pass_counters()
    print(f'{mode}: worst field difference vs FP64 oracle {max(worst.values()):.2e}, '
          f'{(p1[1] - p0[1]) / (p1[0] - p0[0]):.2f} FDM passes per projection after the restart', flush=True)
    g.close()


@pytest.mark.parametrize('mode', ['fp64', 'fdm_fp32', 'fdm_bf16x3'])
def test_symmetric_and_dense_fdm_agree(mode):
    """The even/odd half transforms (symmetric grids) and the dense transforms give the same step: both
    against the oracle, and the split is off on an axis with an odd cell count."""
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    s = W.settings(24, 32, 17, re=500.0, stretch=2.0, lid_shaped=1, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
    print('case: layouts', mode, flush=True)
    noise = W.noise(s, 9, 1e-2)
    runs = {}
    for sym in (True, False):
        g = Solver(s, handle=True, precision=mode)
        g.set_fdm_layout(sym)
        assert g.fdm_layout() == (3 if sym else 0)           # x and y mirror; nz = 17 is odd
        g.initialize(*noise)
        for _ in range(3):
            g.step()
        runs[sym] = {f: g.field64(f) for f in FIELDS}
        g.close()
    o = Oracle(s)
    o.initialize(*noise)
    o.steps(3)
    tol = 1e-11 if mode == 'fp64' else 1e-8
    for f in FIELDS:
        ref = o.field(f)
        assert rel_err(runs[True][f], ref) < tol, (f, rel_err(runs[True][f], ref))
        assert rel_err(runs[False][f], ref) < tol, (f, rel_err(runs[False][f], ref))


@pytest.mark.parametrize('mode', ['fp64', 'fdm_fp32', 'fdm_bf16x3'])
def test_long_axis_even_odd_transforms(mode):
    """Axes up to 256 cells take the whole-column even/odd transform, longer ones the chunked one: on a
    320 x 48 x 12 grid (x chunked, y and z whole-column) the even/odd step agrees with the dense step and
    with the oracle."""
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    s = W.settings(320, 48, 12, re=800.0, stretch=1.6, lid_shaped=0, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
    print('case: long axis', mode, flush=True)
    noise = W.noise(s, 13, 1e-2)
    runs = {}
    for sym in (True, False):
        g = Solver(s, handle=True, precision=mode)
        g.set_fdm_layout(sym)
        assert g.fdm_layout() == (7 if sym else 0)
        g.initialize(*noise)
        for _ in range(2):
            g.step()
        runs[sym] = {f: g.field64(f) for f in FIELDS}
        g.close()
    o = Oracle(s)
    o.initialize(*noise)
    o.steps(2)
    tol = 1e-11 if mode == 'fp64' else 1e-8
    for f in FIELDS:
        ref = o.field(f)
        assert rel_err(runs[True][f], ref) < tol, (f, rel_err(runs[True][f], ref))
        assert rel_err(runs[False][f], ref) < tol, (f, rel_err(runs[False][f], ref))


_WHOLE_COLUMN_REF = {}


@pytest.mark.parametrize('mode', ['fp64', 'fdm_fp32', 'fdm_bf16x3'])
@pytest.mark.parametrize('shape', [(192, 192, 12), (12, 192, 192), (192, 12, 192)])
def test_whole_column_transforms(mode, shape):
    """An axis of at most 256 cells with at least 36864 columns takes the persistent whole-column even/odd
    transform: the z, x and y axes in turn (the short axis keeps the chunked one). The even/odd step agrees
    with the dense step and with the oracle."""
    from les_impulse3d_64_amd.python_api import Solver
    from oracle.oracle import Oracle
    s = W.settings(*shape, re=600.0, stretch=1.8, lid_shaped=1, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
    print('case: whole-column', shape, mode, flush=True)
    noise = W.noise(s, 21, 1e-2)
    runs = {}
    for sym in (True, False):
        g = Solver(s, handle=True, precision=mode)
        g.set_fdm_layout(sym)
        assert g.fdm_layout() == (7 if sym else 0)
        g.initialize(*noise)
        for _ in range(2):
            g.step()
        runs[sym] = {f: g.field64(f) for f in FIELDS}
        g.close()
    if shape not in _WHOLE_COLUMN_REF:
        o = Oracle(s)
        o.initialize(*noise)
        o.steps(2)
        _WHOLE_COLUMN_REF[shape] = {f: o.field(f) for f in FIELDS}
        o.close()
    tol = 1e-11 if mode == 'fp64' else 1e-8
    for f in FIELDS:
        ref = _WHOLE_COLUMN_REF[shape][f]
        assert rel_err(runs[True][f], ref) < tol, (f, rel_err(runs[True][f], ref))
        assert rel_err(runs[False][f], ref) < tol, (f, rel_err(runs[False][f], ref))
