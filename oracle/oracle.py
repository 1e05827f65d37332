"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the checker.
The oracle restates the reference solver (backend.cu) on the host in FP64; see oracle.cpp's header.
The wrapper mirrors les_impulse3d_64_amd.python_api.Solver, with every field in float64.
"""
import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

from les_impulse3d_64_amd.python_api import Config, Report, Diagnostics, DT_REASONS, FIELD_NUMBER, make_config

HERE = Path(__file__).resolve().parent
# ORACLE_LIB: another build of the same oracle (the sanitizer build of tests/test_host_asan.py)
LIB = Path(os.environ['ORACLE_LIB']) if os.environ.get('ORACLE_LIB') else HERE / 'liboracle.so'
_lib = None


def build():
    subprocess.run(['make', '-s', '-C', str(HERE)], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.is_file():
        build()
    lib = ct.CDLL(str(LIB))
    vp, d, i = ct.c_void_p, ct.c_double, ct.c_int
    lib.orc_last_error.restype = ct.c_char_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [vp]
    lib.orc_create_mode.restype = vp
    lib.orc_create_mode.argtypes = [vp, ct.c_int]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_destroy.restype = None
    sig = {
        'orc_initialize': [vp, vp, vp], 'orc_step': [vp, vp], 'orc_steps': [vp, i, vp],
        'orc_kick': [vp, vp, vp, vp, d, i], 'orc_report': [vp, vp], 'orc_diagnostics': [vp, vp],
        'orc_grid': [vp, i, vp, vp, vp, vp], 'orc_get_field': [vp, i, vp], 'orc_set_field': [vp, i, vp],
        'orc_set_clock': [vp, d, ct.c_longlong, d], 'orc_probes': [vp, vp], 'orc_add_statistics': [vp, d, vp],
        'orc_get_statistics': [vp, vp], 'orc_set_statistics': [vp, vp], 'orc_midplane': [vp, vp],
        'orc_q_volume': [vp, i, vp, vp], 'orc_pass_counters': [vp, vp, vp], 'orc_viscous_rate': [vp, vp],
        'orc_inviscid': [vp] + [vp] * 9, 'orc_rk3_stage': [vp, i, vp, vp, vp, vp, d, d, d, d, d, i, vp],
        'orc_divergence': [vp, vp, vp, vp, vp], 'orc_residual': [vp, vp, vp, vp],
        'orc_pressure_pass': [vp, vp, vp], 'orc_solve_pressure': [vp, vp, vp, vp, vp, vp],
        'orc_subtract_gradient': [vp, i, vp, vp, vp], 'orc_eigen': [vp, i, vp, vp, vp],
        'orc_inv_lambda': [vp, vp], 'orc_cfl_max': [vp, vp], 'orc_smooth': [vp, vp, i, i, i, i],
        'orc_axis_faces': [vp, i, d, vp], 'orc_guess_phi': [vp, i, vp, vp, vp, d, d, vp, vp],
        'orc_timing': [vp, vp, vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ct.c_int
    lib.orc_reduce.argtypes = [vp, ct.c_longlong, i]
    lib.orc_reduce.restype = d
    _lib = lib
    return lib


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(ct.c_void_p)


def f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 host solver with the Solver surface; every field is float64."""

    def __init__(self, settings, precision='fp64'):
        """precision 'fp64': the plain FP64 solver; 'field_fp32': the reference's own precision (float velocity,
        impulse, grid and stencils; FP64 potential and refinement; float transforms)."""
        if precision not in ('fp64', 'field_fp32'):
            raise ValueError(f'oracle precision {precision!r}: fp64 or field_fp32')
        self.lib = load()
        self.config = make_config(settings)
        self.precision = precision
        h = self.lib.orc_create_mode(ct.byref(self.config), 1 if precision == 'field_fp32' else 0)
        if not h:
            raise RuntimeError(self.lib.orc_last_error().decode())
        self.h = h
        nx, ny, nz = self.config.nx, self.config.ny, self.config.nz
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shapes = {'u': (nz, ny, nx + 1), 'v': (nz, ny + 1, nx), 'w': (nz + 1, ny, nx), 'phi': (nz, ny, nx)}
        self.shapes.update(mx=self.shapes['u'], my=self.shapes['v'], mz=self.shapes['w'])
        self.grid = {}
        self.widths, self.spacings = {}, {}
        for axis, (name, n) in enumerate((('x', nx), ('y', ny), ('z', nz))):
            faces, centres, widths, spacings = np.empty(n + 1), np.empty(n), np.empty(n), np.empty(n + 1)
            self.call('grid', axis, ptr(faces), ptr(centres), ptr(widths), ptr(spacings))
            self.grid[name] = (faces, centres)
            self.widths[name], self.spacings[name] = widths, spacings

    def call(self, name, *args):
        if getattr(self.lib, 'orc_' + name)(self.h, *args):
            raise RuntimeError(self.lib.orc_last_error().decode())

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self, v_noise=None, w_noise=None):
        v, w = f64(v_noise), f64(w_noise)
        self.call('initialize', ptr(v), ptr(w))

    def step(self):
        r = Report()
        self.call('step', ct.byref(r))
        return self._as_dict(r)

    def steps(self, n):
        r = Report()
        self.call('steps', int(n), ct.byref(r))
        return self._as_dict(r)

    def report(self):
        r = Report()
        self.call('report', ct.byref(r))
        return self._as_dict(r)

    @staticmethod
    def _as_dict(r):
        d = {name: getattr(r, name) for name, _ in Report._fields_}
        d['dt_reason'] = DT_REASONS[r.dt_reason]
        return d

    def kick(self, noise, amplitude, passes):
        n = [f64(a) for a in noise]
        self.call('kick', ptr(n[0]), ptr(n[1]), ptr(n[2]), float(amplitude), int(passes))

    def diagnostics(self):
        d = Diagnostics()
        self.call('diagnostics', ct.byref(d))
        return {name: getattr(d, name) for name, _ in Diagnostics._fields_}

    def field(self, name):
        out = np.empty(self.shapes[name])
        self.call('get_field', FIELD_NUMBER[name], ptr(out))
        return out

    field64 = field

    def set_field(self, name, values):
        a = f64(values)
        assert a.shape == self.shapes[name]
        self.call('set_field', FIELD_NUMBER[name], ptr(a))

    set_field64 = set_field

    def set_clock(self, time, steps, last_dt):
        self.call('set_clock', float(time), int(steps), float(last_dt))

    def probes(self):
        out = np.zeros(2 * self.config.n_probes)
        self.call('probes', ptr(out))
        return out

    def add_statistics(self, weight):
        if self.nx != self.ny:
            self.call('add_statistics', float(weight), None)
            return None
        line = np.empty((5, self.nx))
        self.call('add_statistics', float(weight), ptr(line))
        return line

    def statistics(self):
        sums = np.empty((6, self.ny, self.nx))
        self.call('get_statistics', ptr(sums))
        return dict(zip(('u', 'v', 'uu', 'vv', 'ww', 'uv'), sums))

    def set_statistics(self, sums):
        a = np.ascontiguousarray([sums[k] for k in ('u', 'v', 'uu', 'vv', 'ww', 'uv')], dtype=np.float64)
        self.call('set_statistics', ptr(a))

    def midplane(self):
        out = np.empty((4, self.ny, self.nx))
        self.call('midplane', ptr(out))
        return out

    def q_volume(self, colour_w=True):
        q = np.empty((self.nz, self.ny, self.nx))
        colour = np.empty_like(q)
        self.call('q_volume', int(bool(colour_w)), ptr(q), ptr(colour))
        return q, colour

    def pass_counters(self):
        a, b = ct.c_longlong(), ct.c_longlong()
        self.call('pass_counters', ct.byref(a), ct.byref(b))
        return a.value, b.value

    def timing(self):
        """(seconds spent in the FDM transforms since create, OpenMP threads)"""
        xs, nt = ct.c_double(), ct.c_int()
        self.call('timing', ct.byref(xs), ct.byref(nt))
        return xs.value, nt.value

    def viscous_rate(self):
        r = ct.c_double()
        self.call('viscous_rate', ct.byref(r))
        return r.value

    # ---- kernel-level entry points
    def inviscid(self, u, v, w, mx, my, mz):
        outs = [np.empty(self.shapes[c]) for c in 'uvw']
        args = [ptr(f64(a)) for a in (u, v, w, mx, my, mz)]
        keep = [f64(a) for a in (u, v, w, mx, my, mz)]
        args = [ptr(a) for a in keep]
        self.call('inviscid', *args, *[ptr(o) for o in outs])
        return outs

    def rk3_stage(self, comp, m, m0, vel, I, nu, lid, dt, a0, b, has_base):
        out = np.empty(self.shapes['uvw'[comp]])
        keep = [f64(a) for a in (m, m0, vel, I)]
        self.call('rk3_stage', int(comp), *[ptr(a) for a in keep], float(nu), float(lid), float(dt), float(a0),
                  float(b), int(has_base), ptr(out))
        return out

    def divergence(self, u, v, w):
        out = np.empty(self.shapes['phi'])
        keep = [f64(a) for a in (u, v, w)]
        self.call('divergence', *[ptr(a) for a in keep], ptr(out))
        return out

    def residual(self, phi, rhs):
        out = np.empty(self.shapes['phi'])
        a, b = f64(phi), f64(rhs)
        self.call('residual', ptr(a), ptr(b), ptr(out))
        return out

    def pressure_pass(self, r, phi=None):
        out = np.zeros(self.shapes['phi']) if phi is None else f64(phi).copy()
        a = f64(r)
        self.call('pressure_pass', ptr(a), ptr(out))
        return out

    def solve_pressure(self, rhs, guess=None):
        out = np.empty(self.shapes['phi'])
        a, g = f64(rhs), f64(guess)
        rel, passes = ct.c_double(), ct.c_int()
        self.call('solve_pressure', ptr(a), ptr(g), ptr(out), ct.byref(rel), ct.byref(passes))
        return out, rel.value, passes.value

    def guess_phi(self, times, scales, data, tau, scale):
        """guess_phi (backend.cu:857-873) over a history of len(times) entries, oldest first: data[k] is entry
        k's dphi (phi's shape). Returns the guess, or None when no guess is made."""
        n = len(times)
        t, sc = f64(np.asarray(times, dtype=float)), f64(np.asarray(scales, dtype=float))
        d = f64(np.asarray(data, dtype=float).reshape(n, -1)) if n else np.zeros(1)
        out = np.zeros(self.shapes['phi'])
        used = ct.c_int()
        self.call('guess_phi', n, ptr(t), ptr(sc), ptr(d), float(tau), float(scale), ptr(out), ct.byref(used))
        return out if used.value else None

    def subtract_gradient(self, comp, m, phi):
        out = np.empty(self.shapes['uvw'[comp]])
        a, b = f64(m), f64(phi)
        self.call('subtract_gradient', int(comp), ptr(a), ptr(b), ptr(out))
        return out

    def eigen(self, axis):
        n = (self.nx, self.ny, self.nz)[axis]
        S, lam, zm = np.empty((n, n)), np.empty(n), ct.c_int()
        self.call('eigen', int(axis), ptr(S), ptr(lam), ct.byref(zm))
        return S, lam, zm.value

    def inv_lambda(self):
        out = np.empty(self.shapes['phi'])
        self.call('inv_lambda', ptr(out))
        return out

    def cfl_max(self):
        r = ct.c_double()
        self.call('cfl_max', ct.byref(r))
        return r.value


def reduce(x, op):
    lib = load()
    a = f64(np.ravel(x))
    return lib.orc_reduce(ptr(a), a.size, int(op))


def smooth(a, axis):
    """One periodic 1-2-1 pass along axis (0 x, 1 y, 2 z) of a [z, y, x] array."""
    lib = load()
    src = f64(a)
    out = np.empty_like(src)
    sz, sy, sx = src.shape
    lib.orc_smooth(ptr(src), ptr(out), int(axis), sx, sy, sz)
    return out


def axis_faces(settings, n, length):
    lib = load()
    config = make_config(settings)
    out = np.empty(n + 1)
    if lib.orc_axis_faces(ct.byref(config), int(n), float(length), ptr(out)):
        raise RuntimeError(lib.orc_last_error().decode())
    return out
