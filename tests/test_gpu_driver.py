"""The run driver over the HIP library: the reference's own GPU checks (validate_gpu.py:37-58) driven the way
its script drives them, and a short driver run with kicks, probes and statistics against the same driver
over the FP64 oracle."""
import numpy as np
import pytest

from conftest import FIELDS, rel_err
from les_impulse3d_64_amd import driver as D
from les_impulse3d_64_amd.python_api import Solver
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def test_validate_check1_finite_divergence_free_walls(hip_lib):
    """validate_gpu.py:37-48: 60 steps on the TEST mesh with an impulsive lid: every field finite,
    div_rms < 1e-5, and no flow through the walls (exactly zero)."""
    s = D.settings(MESH='TEST', LID_RAMP_T=0.0)
    sim = Solver(D.native_settings(s))
    run = D.Run(sim, s)
    sim.initialize(*D.initial_noise(sim, s))
    for _ in range(60):
        rep = D.advance(run, log=None)
    d = sim.diagnostics()
    f = {k: sim.field(k) for k in FIELDS}
    sim.close()
    print(f'check 1: div_rms {d["div_rms"]:.2e}, KE {d["kinetic_energy"]:.4g}, t {rep["time"]:.4g}, '
          f'passes {rep["phi_passes"]}', flush=True)
    assert all(np.isfinite(a).all() for a in f.values())
    assert d['div_rms'] < 1e-5
    assert np.abs(f['u'][..., [0, -1]]).max() == 0
    assert np.abs(f['v'][:, [0, -1]]).max() == 0
    assert np.abs(f['w'][[0, -1]]).max() == 0


def test_validate_check2_cfl_limited_dt(hip_lib):
    """validate_gpu.py:50-58: without dt lag and with DT_MAX = 1, a CFL-limited step has dt * rate = 0.4."""
    s = D.settings(MESH='TEST', LID_RAMP_T=0.0, DT_MAX=1.0, DT_LAG=False)
    sim = Solver(D.native_settings(s))
    run = D.Run(sim, s)
    sim.initialize(*D.initial_noise(sim, s))
    for _ in range(20):
        D.advance(run, log=None)
    rate = sim.diagnostics()['cfl_rate']
    rep = D.advance(run, log=None)
    sim.close()
    print(f'check 2: dt {rep["dt"]:.4g} ({rep["dt_reason"]}), dt*rate {rep["dt"] * rate:.6f}', flush=True)
    assert rep['dt_reason'] == 'CFL'                       # the case the reference's check covers
    assert abs(rep['dt'] * rate - 0.4) < 1e-10              # FP64 throughout (the reference allows 1e-4)


def test_driver_run_matches_oracle(hip_lib, tmp_path):
    """A 12-step driver run on a 24 x 24 x 12 stretched cavity with two kicks, probes and statistics from the
    start: the GPU and the oracle agree on every field, the probes, the statistics and the centreline window
    mean; a checkpoint written from the GPU run restarts the oracle run exactly."""
    s = D.settings(MESH='T', MESH_OPTIONS={'T': (24, 12)}, RE=800.0, LID_RAMP_T=0.01, GRID_MAX_ASPECT=2.0,
                   KICK_AMP=0.01, KICK_T=0.004, KICK_EVERY=0.006, KICK_COUNT=2, KICK_CELLS=4,
                   STATS_START_T=0.0, STATS_WINDOW_T=0.02, FDM_RTOL=1e-11,
                   PROBES=((.5, .5), (.8, .8), (.2, .2), (.9, .1), (.1, .9), (.5, .1), (.5, .9), (.1, .5)))
    g = Solver(D.native_settings(s), handle=True)
    o = Oracle(D.native_settings(s))
    rg, ro = D.Run(g, s), D.Run(o, s)
    g.initialize(*D.initial_noise(g, s))
    o.initialize(*D.initial_noise(o, s))
    kicks = []
    for _ in range(12):
        a = D.advance(rg, log=kicks.append)
        b = D.advance(ro, log=None)
        assert a['steps'] == b['steps'] and a['dt_reason'] == b['dt_reason']
        assert abs(a['time'] - b['time']) <= 1e-12 * b['time']
    assert len(kicks) == 2
    worst = max(rel_err(g.field64(f), o.field(f)) for f in FIELDS)
    assert worst < 1e-9, worst
    assert rel_err(np.array(rg.probe_v, dtype=np.float64), np.array(ro.probe_v)) < 1e-6      # float32 probes
    sg, so = g.statistics(), o.statistics()
    assert max(rel_err(sg[k], so[k]) for k in D.STAT_NAMES) < 1e-9
    assert rel_err(rg.window.mean(), ro.window.mean()) < 1e-9

    path = str(tmp_path / 'gpu.npz')
    D.save_checkpoint(rg, path)
    o2 = Oracle(D.native_settings(s))
    r2 = D.Run(o2, s)
    o2.initialize()
    D.load_checkpoint(r2, path, log=None)
    for f in FIELDS:
        assert np.array_equal(o2.field(f), g.field64(f)), f
    g.close()
    o.close()
    o2.close()
