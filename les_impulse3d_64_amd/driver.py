"""Run driver of the cavity solver: the reference front end's run semantics over libcavity_solver.so.

What is kept from the reference driver (main.py), because scripts and checkpoints depend on it:
- the settings names and defaults (main.py:19-76) and their mapping to Config, including the refusal of
  features the Eulerian-impulse solver does not have (native_settings, main.py:79-98);
- the noise generators of the initial condition and the kicks (main.py:130-143);
- the order of work after each step: probes, any kick that fell due inside the step, then the mid-span
  statistics with the part of the step after STATS_START_T as weight (advance, main.py:344-364);
- the trailing-window time mean of the centreline samples (CentrelineWindow, main.py:101-127);
- the checkpoint content and its atomic replace (main.py:159-191), and the CSV files history.csv,
  probes.csv and centerlines.csv (main.py:194-227, 379-417).

What differs: checkpoints store the velocity and impulse in FP64 (this solver's precision), so a restart
loses nothing; a float32 checkpoint written by the reference driver still loads. Plots and Q-isosurface
images are not produced. `main` takes the settings and the restart path as arguments (the command line
`--restart <file>` still works when this module is run as a script), and can run on an independent
solver handle in any precision mode.

Nothing here computes on the fields: every number comes from the library.
"""
import csv
import os
import sys
import time
from collections import deque

import numpy as np

from .python_api import Solver

# The reference driver's settings (main.py:19-76), same names, same defaults.
DEFAULTS = dict(
    L=1.0, U_LID=1.0, RE=22000.0, LZ_L=1, Z_WALLS=True, LID_PROFILE='regularised', LID_RAMP_T=0.3,
    MESH='FINE', MESH_OPTIONS={'TEST': (32, 8), 'VALID': (64, 4), 'COARSE': (96, 24), 'MEDIUM': (256, 4),
                               'FINE': (128, 128)},
    GRID_STRETCH=True, GRID_DX_WALL=0.001, GRID_MAX_ASPECT=4,
    REINIT_R=0.98, IMPULSE_ORDER=4, IMPULSE_ORDER4_RAMP=1, PHI_EXTRAPOLATE=True, FDM_RTOL=1.0e-9,
    FDM_MAX_PASSES=8,
    ADAPTIVE_DT=True, DT_FIXED=0.002, CFL_TARGET=0.4, DT_MAX=0.004, STRETCH_DIFFUSION_LIMIT=2.0, DT_LAG=True,
    INIT_PERTURB=0.005, PERTURB_SEED=3, KICK_T=100, KICK_COUNT=3, KICK_EVERY=50.0, KICK_AMP=0.005,
    KICK_CELLS=20,
    T_END=350.0, N_STEPS=6000000, STATS_START_T=45.0, STATS_WINDOW_T=100.0, PRINT_EVERY=100,
    PLOT_EVERY_STEPS=5000, CENTERLINE_EVERY_STEPS=5000, CHECKPOINT_EVERY=5000, STRUCTURE_EVERY_STEPS=5000,
    Q_ISOLEVEL=None, Q_AUTO_PERCENTILE=99.0, Q_COLOR='w', OUT_DIR='cavity_output',
    CHECKPOINT='cavity_checkpoint.npz', PROBES=((0.5, 0.5), (0.8, 0.8), (0.2, 0.2), (0.9, 0.1)),
    ADVECTION_SCHEME='eulerian_impulse', IMPULSE_FORM='consistent', LES_ON=False, WALL_MODEL_ON=False,
    TWO_D=False,
)
DEFAULTS['NU'] = DEFAULTS['U_LID'] * DEFAULTS['L'] / DEFAULTS['RE']

HISTORY_COLUMNS = ['step', 't', 'dt', 'dt_reason', 'KE', 'enstrophy', 'umax', 'div_rms', 'CFL', 'phi_passes',
                   'gauge_ratio', 'gauge_x', 'gauge_y', 'w_rms_mid', 'w_max_mid']
STAT_NAMES = ('u', 'v', 'uu', 'vv', 'ww', 'uv')
FIELDS = ('u', 'v', 'w', 'mx', 'my', 'mz', 'phi')


def settings(**changes):
    """The default settings with `changes` applied; NU follows RE, U_LID and L unless given."""
    s = dict(DEFAULTS)
    s.update(changes)
    if 'NU' not in changes:
        s['NU'] = s['U_LID'] * s['L'] / s['RE']
    return s


def native_settings(s):
    """Solver settings (Config field names) from driver settings (main.py:79-98). Raises ValueError naming
    every setting that asks for a feature outside the Eulerian-impulse scheme with walls on all sides."""
    needs = []
    if s['ADVECTION_SCHEME'] != 'eulerian_impulse':
        needs.append("ADVECTION_SCHEME = 'eulerian_impulse'")
    if s['IMPULSE_FORM'] != 'consistent':
        needs.append("IMPULSE_FORM = 'consistent'")
    for flag in ('LES_ON', 'WALL_MODEL_ON', 'TWO_D'):
        if s[flag]:
            needs.append(f'{flag} = False')
    if not s['Z_WALLS']:
        needs.append('Z_WALLS = True')
    if s['LID_PROFILE'] not in ('regularised', 'uniform'):
        needs.append("LID_PROFILE 'regularised' or 'uniform'")
    if needs:
        raise ValueError('The native solver implements the Eulerian impulse scheme only; set ' + ', '.join(needs))
    n, nz = s['MESH_OPTIONS'][s['MESH']]
    flags = dict(lid_shaped=s['LID_PROFILE'] == 'regularised', adaptive_dt=s['ADAPTIVE_DT'], dt_lag=s['DT_LAG'],
                 phi_extrapolate=s['PHI_EXTRAPOLATE'], grid_stretch=s['GRID_STRETCH'])
    out = {k: int(bool(v)) for k, v in flags.items()}
    out.update(nx=n, ny=n, nz=int(nz), impulse_order=int(s['IMPULSE_ORDER']),
               fdm_max_passes=int(s['FDM_MAX_PASSES']), length=s['L'], u_lid=s['U_LID'], nu=s['NU'],
               lz=s['LZ_L'] * s['L'], lid_ramp_t=s['LID_RAMP_T'], reinit_r=s['REINIT_R'],
               impulse_order4_ramp=s['IMPULSE_ORDER4_RAMP'], grid_dx_wall=s['GRID_DX_WALL'],
               grid_max_aspect=s['GRID_MAX_ASPECT'] or 0.0, dt_fixed=s['DT_FIXED'], dt_max=s['DT_MAX'],
               cfl_target=s['CFL_TARGET'], diffusion_limit=s['STRETCH_DIFFUSION_LIMIT'], fdm_rtol=s['FDM_RTOL'],
               probes=tuple(s['PROBES']))
    return out


class CentrelineWindow:
    """Time mean of piecewise-constant samples over the trailing `length` time units (main.py:101-127).

    Each sample holds over [begin, end). After an add ending at t, the retained pieces cover
    [max(first begin, t - length), t]; a piece straddling t - length keeps only its part after it."""

    def __init__(self, length):
        self.length = float(length)
        self._pieces = deque()          # [begin, end, sample]
        self.total = 0.0                # sum over pieces of (end - begin) * sample
        self.weight = 0.0               # sum over pieces of (end - begin)

    @property
    def entries(self):
        return [tuple(p) for p in self._pieces]

    def add(self, begin, end, sample):
        self._pieces.append([begin, end, sample])
        self.total = self.total + (end - begin) * sample
        self.weight += end - begin
        cutoff = end - self.length
        while self._pieces and self._pieces[0][0] < cutoff:
            b, e, v = self._pieces[0]
            cut = min(e, cutoff)
            self.total = self.total - (cut - b) * v
            self.weight -= cut - b
            if e > cutoff:                       # straddles the cutoff: keep the part inside the window
                self._pieces[0][0] = cutoff
                break
            self._pieces.popleft()

    def mean(self):
        return None if self.weight <= 0 else self.total / self.weight


def initial_noise(sim, s):
    """(v, w) noise of the initial condition: uniform in [-1, 1) times INIT_PERTURB * U_LID from
    RandomState(PERTURB_SEED), v drawn before w (main.py:130-137)."""
    if s['INIT_PERTURB'] <= 0:
        return None, None
    rng = np.random.RandomState(s['PERTURB_SEED'])
    scale = s['INIT_PERTURB'] * s['U_LID']
    return tuple(scale * rng.uniform(-1, 1, sim.shapes[c]) for c in ('v', 'w'))


def kick_noise(sim, s, number):
    """Standard-normal u, v, w noise of kick `number` (1-based) from default_rng(PERTURB_SEED + 999 + number)
    (main.py:140-143)."""
    rng = np.random.default_rng(s['PERTURB_SEED'] + 1000 + number - 1)
    return [rng.standard_normal(sim.shapes[c]) for c in ('u', 'v', 'w')]


class Run:
    """Front-end state beside the solver: statistics weight, centreline window, probe time series."""

    def __init__(self, sim, s):
        self.sim, self.s = sim, s
        self.stats_weight = 0.0
        self.window = CentrelineWindow(s['STATS_WINDOW_T'])
        self.probe_t, self.probe_v = [], []


def advance(run, log=print):
    """One step; then the probes, every kick whose time falls in (t_old, t_new], and the mid-span statistics
    weighted by the part of the step after STATS_START_T (main.py:344-364). Returns the step's report."""
    sim, s = run.sim, run.s
    t_old = sim.report()['time']
    rep = sim.step()
    t = rep['time']
    run.probe_t.append(t)
    run.probe_v.append(sim.probes())
    if s['KICK_AMP'] > 0:
        due = [k for k in range(int(s['KICK_COUNT'])) if t_old < s['KICK_T'] + k * s['KICK_EVERY'] <= t]
        for k in due:
            ke0 = sim.diagnostics()['kinetic_energy']
            sim.kick(kick_noise(sim, s, k + 1), s['KICK_AMP'], max(1, int(s['KICK_CELLS']) // 2))
            ke1 = sim.diagnostics()['kinetic_energy']
            if log:
                log(f'[kick {k + 1}/{s["KICK_COUNT"]}] t={t:.3f}: added {s["KICK_AMP"]:.3g} U_lid rms smooth noise '
                    f'(~{s["KICK_CELLS"]}-cell eddies), KE {ke0:.4g} -> {ke1:.4g}')
    if t > s['STATS_START_T']:
        weight = t - max(t_old, s['STATS_START_T'])
        run.stats_weight += weight
        line = sim.add_statistics(weight)
        if line is not None:                      # the centreline needs nx == ny
            run.window.add(t - weight, t, line)
    return rep


def _fp64_field(sim, name):
    get = getattr(sim, 'field64', None)
    return get(name) if get is not None else np.asarray(sim.field(name), dtype=np.float64)


def save_checkpoint(run, path):
    """Write the run to `path` (npz, replaced atomically): the seven fields, clock, statistics, centreline
    window and probe history under the reference driver's keys (main.py:159-172)."""
    sim = run.sim
    rep = sim.report()
    pieces = run.window.entries
    data = {name: _fp64_field(sim, name) for name in FIELDS}
    data.update(t=rep['time'], steps=rep['steps'], last_dt=rep['dt'], stat_w=run.stats_weight,
                mesh=np.array([sim.nx, sim.ny, sim.nz]),
                line_times=np.array([(b, e) for b, e, _ in pieces], dtype=np.float64).reshape(-1, 2),
                line_values=np.array([v for _, _, v in pieces], dtype=np.float64).reshape(-1, 5, sim.nx),
                probe_t=np.array(run.probe_t, dtype=np.float64),
                probe_hist=np.array(run.probe_v).reshape(len(run.probe_t), 2 * sim.config.n_probes))
    for k, v in sim.statistics().items():
        data['stat_' + k] = v
    tmp = path + '.tmp.npz'
    np.savez(tmp, **data)
    os.replace(tmp, path)


def load_checkpoint(run, path, log=print):
    """Restore a run from a checkpoint (main.py:175-191). Fields stored in FP64 are restored exactly; float32
    fields (a checkpoint of the reference driver) are widened. As in the reference, setting the fields
    clears the solver's potential history and lagged dt rates."""
    sim = run.sim
    with np.load(path, allow_pickle=False) as d:
        mesh = tuple(int(x) for x in d['mesh'])
        if mesh != (sim.nx, sim.ny, sim.nz):
            raise ValueError(f'checkpoint mesh {mesh} does not match {(sim.nx, sim.ny, sim.nz)}')
        set64 = getattr(sim, 'set_field64', None)
        for name in FIELDS:
            a = d[name]
            if a.dtype == np.float64 and set64 is not None:
                set64(name, a)
            else:
                sim.set_field(name, a)
        sim.set_clock(float(d['t']), int(d['steps']), float(d['last_dt']) if 'last_dt' in d else 0.0)
        if 'stat_u' in d:
            sim.set_statistics({k: d['stat_' + k] for k in STAT_NAMES})
            run.stats_weight = float(d['stat_w'])
        if 'line_times' in d:
            for (b, e), v in zip(d['line_times'], d['line_values']):
                run.window.add(float(b), float(e), np.array(v))
        if 'probe_t' in d:
            run.probe_t, run.probe_v = list(d['probe_t']), list(d['probe_hist'])
        t, steps = float(d['t']), int(d['steps'])
    if log:
        log(f'Restarted from {path}: t={t:.6g}, step={steps}')


def write_probes(run, path):
    """probes.csv: time, then u and v at each probe (main.py:194-199)."""
    names = [f'{c}_{px:g}_{py:g}' for px, py in run.s['PROBES'][:8] for c in 'uv']
    with open(path, 'w', newline='') as fh:
        out = csv.writer(fh)
        out.writerow(['t'] + names)
        for t, v in zip(run.probe_t, run.probe_v):
            out.writerow([f'{t:.8g}'] + [f'{float(x):.6g}' for x in v])


def centerline_values(sim, s):
    """Instantaneous u(y) at x = L/2 and v(x) at y = L/2 on the mid-span plane, over U_LID: the average of
    the two z layers about the mid-span (one layer when nz is odd) (main.py:202-210)."""
    u, v = sim.field('u'), sim.field('v')
    k2 = sim.nz // 2
    k1 = k2 - 1 if sim.nz % 2 == 0 else k2
    half = sim.nx // 2
    uc = 0.5 * (u[k1, :, half] + u[k2, :, half]) / s['U_LID']
    vc = 0.5 * (v[k1, sim.ny // 2, :] + v[k2, sim.ny // 2, :]) / s['U_LID']
    return uc, vc


def write_centerlines(run, path):
    """centerlines.csv: instantaneous centrelines plus the window mean and rms (main.py:213-227)."""
    sim, s = run.sim, run.s
    uc, vc = centerline_values(sim, s)
    mean = run.window.mean()
    y, x = sim.grid['y'][1] / s['L'], sim.grid['x'][1] / s['L']
    blank = ''
    if mean is not None:
        um, vm, uu, vv, ww = mean
        urms, vrms, wrms = (np.sqrt(np.maximum(a, 0)) for a in (uu - um * um, vv - vm * vm, ww))
    with open(path, 'w', newline='') as fh:
        out = csv.writer(fh)
        out.writerow(['y/L', 'u/U (instant, x=L/2)', 'u_mean/U', 'u_rms/U', 'x/L', 'v/U (instant, y=L/2)',
                      'v_mean/U', 'v_rms/U', 'w_rms/U'])
        for j in range(len(y)):
            left = [um[j], urms[j]] if mean is not None else [blank, blank]
            right = [vm[j], vrms[j], wrms[j]] if mean is not None else [blank, blank, blank]
            row = [y[j], uc[j]] + left + [x[j], vc[j]] + right
            out.writerow([q if isinstance(q, str) else f'{q:.7g}' for q in row])


def history_row(sim, s, rep):
    """One history.csv row (main.py:398-414): report, diagnostics, CFL and the mid-span w statistics."""
    d = sim.diagnostics()
    wmid = sim.midplane()[2]
    w_rms = float(np.sqrt(np.mean(np.square(wmid, dtype=np.float64)))) / s['U_LID']
    w_max = float(np.abs(wmid).max()) / s['U_LID']
    cfl = rep['dt'] * d['cfl_rate']
    row = [rep['steps'], f'{rep["time"]:.8g}', f'{rep["dt"]:.6g}', rep['dt_reason'], f'{d["kinetic_energy"]:.8g}',
           f'{d["enstrophy"]:.8g}', f'{d["umax"]:.6g}', f'{d["div_rms"]:.3g}', f'{cfl:.4g}', rep['phi_passes'],
           f'{d["gauge_ratio"]:.5g}', f'{d["gauge_x"]:.4f}', f'{d["gauge_y"]:.4f}', f'{w_rms:.5g}', f'{w_max:.5g}']
    return row, d, cfl, w_rms


def main(s=None, restart=None, solver=None, log=print):
    """The run loop (main.py:367-438): advance until T_END or N_STEPS, a history row every PRINT_EVERY
    steps, centrelines every CENTERLINE_EVERY_STEPS, a checkpoint every CHECKPOINT_EVERY steps; the
    checkpoint, probes and centrelines are written again on the way out, also after an error.
    `solver` may be a ready Solver built from native_settings(s) (for example on its own handle or in a
    mixed precision mode); by default the process-wide solver is created. Returns the final report."""
    s = dict(DEFAULTS if s is None else s)
    if restart is None and '--restart' in sys.argv:
        restart = sys.argv[sys.argv.index('--restart') + 1]
    os.makedirs(s['OUT_DIR'], exist_ok=True)
    sim = solver if solver is not None else Solver(native_settings(s))
    run = Run(sim, s)
    if restart:
        sim.initialize()
        load_checkpoint(run, restart, log=log)
    else:
        sim.initialize(*initial_noise(sim, s))
    gx = np.diff(sim.grid['x'][0])
    if log:
        log(f'Lid-driven cavity (MI355X HIP) Re={s["RE"]:g} mesh={s["MESH"]} ({sim.nx}x{sim.ny}x{sim.nz}), '
            f'NU={s["NU"]:.4g}, wall cell {gx[0]:.4g}, centre cell {gx[len(gx) // 2]:.4g}, '
            f'impulse order {s["IMPULSE_ORDER"]}')
    hist_path = os.path.join(s['OUT_DIR'], 'history.csv')
    fresh = not (restart and os.path.exists(hist_path))
    ckpt = os.path.join(s['OUT_DIR'], s['CHECKPOINT'])
    wall0 = time.perf_counter()
    rep = sim.report()
    steps0 = rep['steps']
    with open(hist_path, 'w' if fresh else 'a', newline='') as hist:
        writer = csv.writer(hist)
        if fresh:
            writer.writerow(HISTORY_COLUMNS)
        try:
            while rep['time'] < s['T_END'] and rep['steps'] < s['N_STEPS']:
                rep = advance(run, log=log)
                n = rep['steps']
                if n % s['PRINT_EVERY'] == 0:
                    row, d, cfl, w_rms = history_row(sim, s, rep)
                    writer.writerow(row)
                    hist.flush()
                    if log:
                        rate = (n - steps0) / (time.perf_counter() - wall0)
                        log(f'step {n:7d} t={rep["time"]:9.4f} dt={rep["dt"]:.3e} ({rep["dt_reason"]}) CFL={cfl:.3f} '
                            f'KE={d["kinetic_energy"]:.5f} Z={d["enstrophy"]:.3f} umax={d["umax"]:.3f} '
                            f'div={d["div_rms"]:.2e} phi_passes={rep["phi_passes"]} gauge={d["gauge_ratio"]:.2f}'
                            f'@({d["gauge_x"]:.3f},{d["gauge_y"]:.3f}) w_rms(mid)={w_rms:.2e} {rate:.2f} steps/s')
                if s['CENTERLINE_EVERY_STEPS'] and n % s['CENTERLINE_EVERY_STEPS'] == 0:
                    write_centerlines(run, os.path.join(s['OUT_DIR'], 'centerlines.csv'))
                if s['CHECKPOINT_EVERY'] and n % s['CHECKPOINT_EVERY'] == 0:
                    save_checkpoint(run, ckpt)
        finally:
            save_checkpoint(run, ckpt)
            write_probes(run, os.path.join(s['OUT_DIR'], 'probes.csv'))
            write_centerlines(run, os.path.join(s['OUT_DIR'], 'centerlines.csv'))
            rep = sim.report()
            if log:
                log(f'Done: t={rep["time"]:.4f}, steps={rep["steps"]}, wall={time.perf_counter() - wall0:.1f}s. '
                    f'Outputs in {s["OUT_DIR"]}/')
            if solver is None:
                sim.close()
    return rep


if __name__ == '__main__':
    main()
