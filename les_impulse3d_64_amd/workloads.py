"""The benchmark and parity workloads as solver settings plus seeded host inputs.

Each workload is a settings dict for python_api.make_config (the names of Config, python_api.py:11-18) and
a recipe for its FP64 host arrays. Defaults follow the reference driver (main.py:19-76 mapped by
native_settings, main.py:79-98): L = 1, U_lid = 1, nu = U_lid L / Re, gauge retention 0.98, 4th-order
Lamb vector with a 1-cell ramp, warm-started projections, at most 8 FDM passes, dt_max 0.004, diffusion
limit 2, lagged dt, no lid ramp.

A stretch factor f maps to grid_stretch = 1, grid_dx_wall = 1e-4, grid_max_aspect = f: the aspect cap
binds, so largest/smallest cell = f per axis (backend.cu:672-687). f = 1 gives the uniform grid.
Noise is numpy.random.default_rng(seed).standard_normal(shape) * sigma on the v then w face arrays.
"""
import numpy as np

BASE = dict(length=1.0, u_lid=1.0, lz=1.0, lid_ramp_t=0.0, reinit_r=0.98, impulse_order=4, impulse_order4_ramp=1.0,
            phi_extrapolate=1, fdm_max_passes=8, dt_max=0.004, diffusion_limit=2.0, dt_lag=1, cfl_target=0.4,
            dt_fixed=0.002, fdm_rtol=1e-9, adaptive_dt=1, grid_stretch=0, grid_dx_wall=1e-4, grid_max_aspect=0.0,
            lid_shaped=1)

PROBES8 = ((.5, .5), (.8, .8), (.2, .2), (.9, .1), (.1, .9), (.5, .1), (.5, .9), (.1, .5))


def settings(nx, ny=None, nz=None, re=100.0, stretch=1.0, **changes):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    s = dict(BASE, nx=int(nx), ny=int(ny), nz=int(nz), nu=BASE['u_lid'] * BASE['length'] / re)
    if stretch > 1.0:
        s.update(grid_stretch=1, grid_dx_wall=1e-4, grid_max_aspect=float(stretch))
    s.update(changes)
    return s


def noise(settings_, seed, sigma):
    """(v_noise, w_noise) for solver_initialize, shapes [z, y, x] of the v and w faces."""
    nx, ny, nz = settings_['nx'], settings_['ny'], settings_['nz']
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((nz, ny + 1, nx)) * sigma
    w = rng.standard_normal((nz + 1, ny, nx)) * sigma
    return v, w


def kick_noise(settings_, seed):
    """Standard-normal noise on the u, v, w face arrays for solver_kick."""
    nx, ny, nz = settings_['nx'], settings_['ny'], settings_['nz']
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape) for shape in ((nz, ny, nx + 1), (nz, ny + 1, nx), (nz + 1, ny, nx))]


def smooth_periodic(a, axis):
    """One periodic 1-2-1 pass along array axis `axis` (numpy order [z, y, x]), as the kick smooths its
    input (backend.cu:544-552)."""
    return 0.25 * np.roll(a, 1, axis) + 0.5 * a + 0.25 * np.roll(a, -1, axis)


def kick_blobs(settings_, seed, grid, peak=0.05, radius=0.05, passes=1):
    """Inputs of one Gaussian-blob velocity kick: (noise [u, v, w], amp, passes) for solver_kick.

    The blob is exp(-|x - c|^2 / radius^2) on every component, sampled at that component's face positions
    (grid: {'x': (faces, centres), 'y': ..., 'z': ...} as Solver.grid holds it). The centre c is drawn
    from default_rng(seed), uniform in the middle half of the box on each axis. solver_kick smooths each
    component (passes x periodic 1-2-1 along z, y, x) and rescales it to rms amp * u_lid before the
    projection (backend.cu:958-978); amp is chosen so the largest component of the smoothed, rescaled blob
    peaks at `peak`, the smoothing computed here on the host exactly as the kick applies it."""
    rng = np.random.default_rng(seed)
    fx, cx = grid['x']
    fy, cy = grid['y']
    fz, cz = grid['z']
    lengths = (fx[-1], fy[-1], fz[-1])
    centre = [lo * L + rng.uniform(0.25, 0.75) * L for lo, L in zip((0.0, 0.0, 0.0), lengths)]
    noise, ratio = [], 0.0
    for comp in range(3):
        xs = fx if comp == 0 else cx
        ys = fy if comp == 1 else cy
        zs = fz if comp == 2 else cz
        d2 = ((zs[:, None, None] - centre[2]) ** 2 + (ys[None, :, None] - centre[1]) ** 2
              + (xs[None, None, :] - centre[0]) ** 2)
        g = np.exp(-d2 / (radius * radius))
        noise.append(np.ascontiguousarray(g))
        sm = g
        for _ in range(max(1, passes)):
            for axis in (0, 1, 2):              # z, y, x
                sm = smooth_periodic(sm, axis)
        ratio = max(ratio, np.abs(sm).max() / np.sqrt((sm * sm).mean()))
    amp = peak / (ratio * settings_['u_lid'])
    return noise, float(amp), passes


KICK_STEPS = (500, 1000, 1500)
KICK_SEEDS = (2001, 2002, 2003)


def c0_parity():
    """32^3 uniform, Re 100, regularised lid, fixed dt 2e-3, sigma 1e-3 seed 0, 200 steps."""
    s = settings(32, re=100.0, lid_shaped=1, adaptive_dt=0, dt_fixed=2e-3, fdm_rtol=1e-9)
    return s, noise(s, 0, 1e-3), 200


def c1_stats():
    """128^3 stretched 1.5, Re 1000, uniform lid, CFL 0.5, sigma 1e-3 seed 1, 1000 steps, 8 probes."""
    s = settings(128, re=1000.0, stretch=1.5, lid_shaped=0, adaptive_dt=1, cfl_target=0.5, probes=PROBES8)
    return s, noise(s, 1, 1e-3), 1000


def c2_mixed(n=256):
    """256^3 stretched 2.0, Re 3200, regularised lid, CFL 0.5, sigma 1e-2 seed 2, FDM rtol 1e-10, 2000 steps,
    kicks after steps 500, 1000, 1500 (seeds 2001..2003)."""
    s = settings(n, re=3200.0, stretch=2.0, lid_shaped=1, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
    return s, noise(s, 2, 1e-2), 2000


def ensemble_member(i):
    """Cavity i of the 64-cavity ensemble: sizes from {32, 48, 64, 96}, stretch U[1, 2], Re log-U[100, 5000],
    lid regularised or uniform by coin flip, sigma 1e-3."""
    rng = np.random.default_rng(3000 + i)
    nx, ny, nz = (int(rng.choice([32, 48, 64, 96])) for _ in range(3))
    stretch = float(rng.uniform(1.0, 2.0))
    re = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
    lid = int(rng.integers(0, 2))
    s = settings(nx, ny, nz, re=re, stretch=stretch, lid_shaped=lid, adaptive_dt=1, cfl_target=0.5, fdm_rtol=1e-10)
    return s, noise(s, 3000 + i, 1e-3)
