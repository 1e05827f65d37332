"""Host logic of the run driver (les_impulse3d_64_amd/driver.py), on the CPU.

The driver is exercised over the FP64 oracle, which has the Solver surface: the settings mapping and its
refusals (main.py:79-98), the trailing centreline window against direct integration (main.py:101-127), the
order of work in advance() (main.py:344-364), the checkpoint round trip (main.py:159-191) and the files of
a short run (main.py:367-438).
"""
import csv
import os

import numpy as np
import pytest

from les_impulse3d_64_amd import driver as D
from oracle.oracle import Oracle


def small(**changes):
    """The TEST mesh (32 x 32 x 8) with a short clock, no lid ramp."""
    base = dict(MESH='TEST', LID_RAMP_T=0.0, RE=400.0, DT_MAX=0.004, PRINT_EVERY=2, CENTERLINE_EVERY_STEPS=3,
                CHECKPOINT_EVERY=4, STATS_START_T=0.0, STATS_WINDOW_T=0.01, KICK_AMP=0.0)
    base.update(changes)
    return D.settings(**base)


# ---------------------------------------------------------------------------------------------- settings
def test_native_settings_mapping():
    s = D.settings(MESH='COARSE', LID_PROFILE='uniform', ADAPTIVE_DT=False, GRID_MAX_ASPECT=None, RE=500.0)
    n = D.native_settings(s)
    assert (n['nx'], n['ny'], n['nz']) == (96, 96, 24)
    assert n['lid_shaped'] == 0 and n['adaptive_dt'] == 0 and n['dt_lag'] == 1 and n['grid_stretch'] == 1
    assert n['grid_max_aspect'] == 0.0                     # None means no cap
    assert n['nu'] == 1.0 / 500.0 and n['lz'] == 1.0 and n['diffusion_limit'] == 2.0
    assert n['probes'] == tuple(s['PROBES'])
    assert all(isinstance(n[k], int) for k in ('lid_shaped', 'adaptive_dt', 'dt_lag', 'phi_extrapolate'))


@pytest.mark.parametrize('change, named', [
    (dict(LES_ON=True), 'LES_ON = False'), (dict(WALL_MODEL_ON=True), 'WALL_MODEL_ON = False'),
    (dict(TWO_D=True), 'TWO_D = False'), (dict(Z_WALLS=False), 'Z_WALLS = True'),
    (dict(ADVECTION_SCHEME='weno'), "ADVECTION_SCHEME = 'eulerian_impulse'"),
    (dict(IMPULSE_FORM='simple'), "IMPULSE_FORM = 'consistent'"),
    (dict(LID_PROFILE='parabolic'), "LID_PROFILE 'regularised' or 'uniform'")])
def test_native_settings_refuses(change, named):
    with pytest.raises(ValueError, match='Eulerian impulse') as e:
        D.native_settings(D.settings(**change))
    assert named in str(e.value)


def test_two_refusals_are_both_named():
    with pytest.raises(ValueError) as e:
        D.native_settings(D.settings(LES_ON=True, TWO_D=True))
    assert 'LES_ON = False' in str(e.value) and 'TWO_D = False' in str(e.value)


# ---------------------------------------------------------------------------------------- centreline window
def window_mean_direct(pieces, length):
    """Mean of the piecewise-constant signal over [max(start, t_end - length), t_end] by direct integration."""
    t_end = pieces[-1][1]
    lo = max(pieces[0][0], t_end - length)
    num = 0.0
    den = 0.0
    for b, e, v in pieces:
        a, z = max(b, lo), min(e, t_end)
        if z > a:
            num = num + (z - a) * v
            den += z - a
    return num / den, den


def test_window_matches_direct_integration():
    rng = np.random.default_rng(7)
    w = D.CentrelineWindow(1.3)
    pieces = []
    t = 0.25
    for _ in range(60):
        dt = float(rng.uniform(0.01, 0.2))
        v = rng.standard_normal((5, 4))
        w.add(t, t + dt, v)
        pieces.append((t, t + dt, v))
        t += dt
        ref, span = window_mean_direct(pieces, 1.3)
        assert abs(w.weight - span) < 1e-12
        assert np.allclose(w.mean(), ref, rtol=0, atol=1e-12)
        # retained pieces: contiguous, ending now, starting at the cutoff once the window is full
        ent = w.entries
        assert ent[-1][1] == t
        assert all(abs(a[1] - b[0]) < 1e-15 for a, b in zip(ent, ent[1:]))
        assert ent[0][0] >= t - 1.3 - 1e-12


def test_window_constant_and_empty():
    w = D.CentrelineWindow(0.5)
    assert w.mean() is None
    for k in range(20):
        w.add(0.1 * k, 0.1 * (k + 1), np.full((5, 3), 2.5))
    assert np.allclose(w.mean(), 2.5, rtol=0, atol=1e-14)
    assert abs(w.weight - 0.5) < 1e-12


def test_window_piece_exactly_at_cutoff_is_dropped_whole():
    w = D.CentrelineWindow(1.0)
    w.add(0.0, 1.0, np.ones((5, 1)))
    w.add(1.0, 2.0, 3 * np.ones((5, 1)))          # cutoff 1.0: the first piece leaves entirely
    assert w.entries[0][0] == 1.0 and len(w.entries) == 1
    assert np.allclose(w.mean(), 3.0)


# ------------------------------------------------------------------------------------------- advance order
def test_advance_probes_kicks_statistics():
    s = small(KICK_AMP=0.01, KICK_T=0.005, KICK_EVERY=0.004, KICK_COUNT=2, KICK_CELLS=4, STATS_START_T=0.01)
    sim = Oracle(D.native_settings(s))
    run = D.Run(sim, s)
    sim.initialize(*D.initial_noise(sim, s))
    logs = []
    times = [0.0]
    for _ in range(8):
        rep = D.advance(run, log=logs.append)
        times.append(rep['time'])
    kicks_due = [k for k in range(2) for a, b in zip(times, times[1:]) if a < 0.005 + 0.004 * k <= b]
    assert len(logs) == len(kicks_due) == 2
    assert logs[0].startswith('[kick 1/2]') and logs[1].startswith('[kick 2/2]')
    assert len(run.probe_t) == 8 and run.probe_t == times[1:]
    assert all(p.shape == (8,) for p in run.probe_v)           # 4 probes, u and v each
    # statistics weight: the time after STATS_START_T, split into the steps that crossed it
    assert abs(run.stats_weight - (times[-1] - 0.01)) < 1e-15
    first = run.window.entries[0]
    assert first[0] >= 0.01 - 1e-15
    sim.close()


def test_kick_noise_and_initial_noise_generators():
    s = small(INIT_PERTURB=0.02, PERTURB_SEED=11)
    sim = Oracle(D.native_settings(s))
    v, w = D.initial_noise(sim, s)
    rng = np.random.RandomState(11)
    assert np.array_equal(v, 0.02 * rng.uniform(-1, 1, sim.shapes['v']))
    assert np.array_equal(w, 0.02 * rng.uniform(-1, 1, sim.shapes['w']))
    k2 = D.kick_noise(sim, s, 2)
    rng = np.random.default_rng(11 + 1001)
    for a, c in zip(k2, 'uvw'):
        assert np.array_equal(a, rng.standard_normal(sim.shapes[c]))
    assert D.initial_noise(sim, dict(s, INIT_PERTURB=0.0)) == (None, None)
    sim.close()


# --------------------------------------------------------------------------------------------- checkpoint
def test_checkpoint_round_trip(tmp_path):
    s = small()
    a = Oracle(D.native_settings(s))
    ra = D.Run(a, s)
    a.initialize(*D.initial_noise(a, s))
    for _ in range(5):
        D.advance(ra)
    path = str(tmp_path / 'ck.npz')
    D.save_checkpoint(ra, path)
    assert not os.path.exists(path + '.tmp.npz')

    b = Oracle(D.native_settings(s))
    rb = D.Run(b, s)
    b.initialize()
    D.load_checkpoint(rb, path, log=None)
    for f in D.FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    ra_rep, rb_rep = a.report(), b.report()
    assert (ra_rep['time'], ra_rep['steps'], ra_rep['dt']) == (rb_rep['time'], rb_rep['steps'], rb_rep['dt'])
    sa, sb = a.statistics(), b.statistics()
    assert all(np.array_equal(sa[k], sb[k]) for k in D.STAT_NAMES)
    assert rb.stats_weight == ra.stats_weight
    assert [e[:2] for e in rb.window.entries] == [e[:2] for e in ra.window.entries]
    # the reloaded window re-adds the pieces, so its running sum rounds differently from the one that had
    # pieces trimmed off
    assert np.allclose(rb.window.mean(), ra.window.mean(), rtol=1e-12, atol=0)
    assert rb.probe_t == ra.probe_t and all(np.array_equal(x, y) for x, y in zip(rb.probe_v, ra.probe_v))

    # continuing: the restarted run equals the original run with its fields set again (the history and
    # the lagged rates are cleared by set_field in both)
    for f in D.FIELDS:
        a.set_field64(f, a.field64(f))
    for _ in range(3):
        D.advance(ra)
        D.advance(rb)
    for f in D.FIELDS:
        assert np.array_equal(a.field(f), b.field(f)), f
    a.close()
    b.close()


def test_checkpoint_mesh_mismatch(tmp_path):
    s = small()
    a = Oracle(D.native_settings(s))
    ra = D.Run(a, s)
    a.initialize()
    path = str(tmp_path / 'ck.npz')
    D.save_checkpoint(ra, path)
    s2 = small(MESH_OPTIONS={'TEST': (16, 8)})
    b = Oracle(D.native_settings(s2))
    with pytest.raises(ValueError, match='does not match'):
        D.load_checkpoint(D.Run(b, s2), path, log=None)
    a.close()
    b.close()


def test_float32_checkpoint_loads(tmp_path):
    """A checkpoint holding float32 velocity and impulse (the reference driver's precision) is widened."""
    s = small()
    a = Oracle(D.native_settings(s))
    ra = D.Run(a, s)
    a.initialize(*D.initial_noise(a, s))
    D.advance(ra)
    path = str(tmp_path / 'ck.npz')
    D.save_checkpoint(ra, path)
    with np.load(path) as d:
        data = {k: d[k] for k in d.files}
    for f in ('u', 'v', 'w', 'mx', 'my', 'mz'):
        data[f] = data[f].astype(np.float32)
    p32 = str(tmp_path / 'ck32.npz')
    np.savez(p32, **data)
    b = Oracle(D.native_settings(s))
    D.load_checkpoint(D.Run(b, s), p32, log=None)
    for f in ('u', 'v', 'w', 'mx', 'my', 'mz'):
        assert np.array_equal(b.field(f), a.field(f).astype(np.float32).astype(np.float64))
    assert np.array_equal(b.field('phi'), a.field('phi'))
    a.close()
    b.close()


# ----------------------------------------------------------------------------------------------- run loop
def test_main_writes_outputs(tmp_path):
    s = small(OUT_DIR=str(tmp_path / 'out'), N_STEPS=6, T_END=1e9)
    sim = Oracle(D.native_settings(s))
    lines = []
    rep = D.main(s, solver=sim, log=lines.append)
    assert rep['steps'] == 6
    out = tmp_path / 'out'
    with open(out / 'history.csv') as fh:
        rows = list(csv.reader(fh))
    assert rows[0] == D.HISTORY_COLUMNS
    assert [int(r[0]) for r in rows[1:]] == [2, 4, 6]
    with open(out / 'probes.csv') as fh:
        rows = list(csv.reader(fh))
    assert rows[0] == ['t', 'u_0.5_0.5', 'v_0.5_0.5', 'u_0.8_0.8', 'v_0.8_0.8', 'u_0.2_0.2', 'v_0.2_0.2',
                       'u_0.9_0.1', 'v_0.9_0.1']
    assert len(rows) == 7
    with open(out / 'centerlines.csv') as fh:
        rows = list(csv.reader(fh))
    assert len(rows) == 1 + 32 and all(len(r) == 9 for r in rows)
    assert all(q != '' for q in rows[1])                       # statistics from t = 0: the means are there
    assert (out / s['CHECKPOINT']).is_file()
    assert lines[-1].startswith('Done: ')

    # restart: the history file is appended, the clock continues
    s2 = dict(s, N_STEPS=8)
    sim2 = Oracle(D.native_settings(s2))
    rep2 = D.main(s2, restart=str(out / s['CHECKPOINT']), solver=sim2, log=None)
    assert rep2['steps'] == 8
    with open(out / 'history.csv') as fh:
        rows = list(csv.reader(fh))
    assert [int(r[0]) for r in rows[1:]] == [2, 4, 6, 8]
    sim.close()
    sim2.close()
