"""ctypes binding of libcavity_solver.so (the MI355X HIP solver).

Mirrors the reference binding (python_api.py:1-161): the same three records, field by field, the same
`Solver` surface (initialize, step, report, kick, diagnostics, field, set_field, set_clock, probes,
add_statistics, statistics, set_statistics, midplane, q_volume, close) and the same error behaviour
(a nonzero return raises RuntimeError with solver_last_error's text).

Additions: `Solver(..., handle=True)` runs on an independent handle (cav_* API) instead of the
process-wide solver, so several solvers can live side by side; `field64`/`set_field64` move the FP64
state without rounding; `set_precision` selects the FDM precision; `steps(n)` advances n steps in one call.

Nothing here computes: the library does all the work on the GPU and fails loudly if it is missing.
"""
import ctypes as ct
from pathlib import Path

import numpy as np


class Config(ct.Structure):
    """python_api.py:11-18 / backend.cu:584-589 (288 bytes)."""
    _fields_ = ([(name, ct.c_int) for name in
                 'nx ny nz lid_shaped impulse_order adaptive_dt dt_lag phi_extrapolate fdm_max_passes '
                 'grid_stretch n_probes'.split()]
                + [(name, ct.c_double) for name in
                   'length u_lid nu lz lid_ramp_t reinit_r impulse_order4_ramp grid_dx_wall grid_max_aspect '
                   'dt_fixed dt_max cfl_target diffusion_limit fdm_rtol'.split()]
                + [('probe_x', ct.c_double * 8), ('probe_y', ct.c_double * 8)])


class Report(ct.Structure):
    """python_api.py:21-23 / backend.cu:590-594 (48 bytes)."""
    _fields_ = [('steps', ct.c_longlong), ('time', ct.c_double), ('dt', ct.c_double), ('cfl', ct.c_double),
                ('phi_rel', ct.c_double), ('dt_reason', ct.c_int), ('phi_passes', ct.c_int)]


class Diagnostics(ct.Structure):
    """python_api.py:26-28 / backend.cu:595-597 (64 bytes)."""
    _fields_ = [(name, ct.c_double) for name in
                'kinetic_energy enstrophy umax div_rms cfl_rate gauge_ratio gauge_x gauge_y'.split()]


DT_REASONS = ('fixed_DT', 'DT_MAX', 'CFL', 'viscous')
FIELD_NUMBER = {'u': 0, 'v': 1, 'w': 2, 'mx': 3, 'my': 4, 'mz': 5, 'phi': 6}
PRECISION = {'fp64': 0, 'fdm_fp32': 1, 'fdm_bf16x3': 2, 'field_fp32': 3}

LIBRARY_NAME = 'libcavity_solver.so'
_lib_cache = {}


def library_path():
    return Path(__file__).resolve().parent / LIBRARY_NAME


def load_library(path=None):
    """Load (once per path) the HIP library; FileNotFoundError if it has not been built."""
    path = Path(path) if path is not None else library_path()
    key = str(path)
    if key in _lib_cache:
        return _lib_cache[key]
    if not path.is_file():
        raise FileNotFoundError(f'Build the solver first (python -c "import __graft_entry__ as g; g.build()"). '
                                f'Missing: {path}')
    lib = ct.CDLL(key)
    lib.solver_last_error.restype = ct.c_char_p
    lib.cav_last_error.restype = ct.c_char_p
    lib.solver_set_clock.argtypes = [ct.c_double, ct.c_longlong, ct.c_double]
    lib.solver_kick.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_int]
    lib.solver_add_statistics.argtypes = [ct.c_double, ct.c_void_p]
    vp = ct.c_void_p
    lib.cav_set_clock.argtypes = [vp, ct.c_double, ct.c_longlong, ct.c_double]
    lib.cav_kick.argtypes = [vp, vp, vp, vp, ct.c_double, ct.c_int]
    lib.cav_add_statistics.argtypes = [vp, ct.c_double, vp]
    for name in ('cav_destroy', 'cav_initialize', 'cav_step', 'cav_steps', 'cav_report', 'cav_diagnostics',
                 'cav_grid', 'cav_get_field', 'cav_set_field', 'cav_get_field64', 'cav_set_field64', 'cav_probes',
                 'cav_get_statistics', 'cav_set_statistics', 'cav_midplane', 'cav_q_volume', 'cav_set_precision',
                 'cav_pass_counters', 'cav_time_steps', 'cav_set_kernel_timing', 'cav_kernel_times',
                 'cav_set_fdm_layout', 'cav_fdm_layout'):
        getattr(lib, name).restype = ct.c_int
    lib.cav_destroy.argtypes = [vp]
    lib.cav_initialize.argtypes = [vp, vp, vp]
    lib.cav_step.argtypes = [vp, vp]
    lib.cav_steps.argtypes = [vp, ct.c_int, vp]
    lib.cav_report.argtypes = [vp, vp]
    lib.cav_diagnostics.argtypes = [vp, vp]
    lib.cav_grid.argtypes = [vp, ct.c_int, vp, vp]
    lib.cav_get_field.argtypes = [vp, ct.c_int, vp]
    lib.cav_set_field.argtypes = [vp, ct.c_int, vp]
    lib.cav_get_field64.argtypes = [vp, ct.c_int, vp]
    lib.cav_set_field64.argtypes = [vp, ct.c_int, vp]
    lib.cav_probes.argtypes = [vp, vp]
    lib.cav_get_statistics.argtypes = [vp, vp]
    lib.cav_set_statistics.argtypes = [vp, vp]
    lib.cav_midplane.argtypes = [vp, vp]
    lib.cav_q_volume.argtypes = [vp, ct.c_int, vp, vp]
    lib.cav_set_precision.argtypes = [vp, ct.c_int]
    lib.cav_pass_counters.argtypes = [vp, vp, vp]
    lib.cav_set_fdm_layout.argtypes = [vp, ct.c_int]
    lib.cav_fdm_layout.argtypes = [vp, vp]
    lib.cav_time_steps.argtypes = [vp, ct.c_int, vp]
    lib.cav_set_kernel_timing.argtypes = [vp, ct.c_int]
    lib.cav_kernel_times.argtypes = [vp, ct.c_int, vp, vp, vp, vp]
    lib.cav_create.argtypes = [vp, vp]
    lib.cav_footprint.argtypes = [vp, ct.c_int, vp]
    lib.cav_set_graph.argtypes = [vp, ct.c_int]
    lib.cav_redone_steps.argtypes = [vp, vp]
    lib.cav_set_enqueued_passes.argtypes = [vp, ct.c_int]
    lib.cav_ensemble_set_enqueued_passes.argtypes = [vp, ct.c_int]
    lib.cav_ensemble_create.argtypes = [vp, ct.c_int, vp]
    lib.cav_ensemble_destroy.argtypes = [vp]
    lib.cav_ensemble_size.argtypes = [vp, vp]
    lib.cav_ensemble_steps.argtypes = [vp, ct.c_int, vp, vp]
    lib.cav_ensemble_time_steps.argtypes = [vp, ct.c_int, vp, vp, vp]
    lib.cav_ensemble_set_graph.argtypes = [vp, ct.c_int]
    lib.cav_ensemble_set_kernel_timing.argtypes = [vp, ct.c_int]
    lib.cav_ensemble_kernel_times.argtypes = [vp, ct.c_int, vp, vp, vp, vp]
    lib.cav_snapshot_bytes.argtypes = [vp, vp]
    lib.cav_snapshot.argtypes = [vp, vp, ct.c_longlong]
    lib.cav_restore.argtypes = [vp, vp, ct.c_longlong]
    lib.cav_device_alloc.argtypes = [ct.c_longlong, vp]
    lib.cav_device_free.argtypes = [vp]
    lib.cav_memcpy.argtypes = [vp, vp, ct.c_longlong]
    lib.cav_error_text.restype = ct.c_char_p
    lib.cav_error_text.argtypes = [ct.c_int]
    lib.cav_batch_steps_noalloc.argtypes = [vp, ct.c_int, ct.c_int, vp, vp]
    lib.cav_set_autotune.argtypes = [ct.c_int]
    lib.cav_autotune.argtypes = [vp]
    lib.cav_ensemble_autotune.argtypes = [vp]
    lib.cav_get_tuning.argtypes = [vp, vp, ct.c_int]
    lib.cav_set_tuning.argtypes = [vp, vp, ct.c_int]
    _lib_cache[key] = lib
    return lib


def pointer(a):
    """C pointer to a NumPy array's data (None stays None = NULL)."""
    return None if a is None else a.ctypes.data_as(ct.c_void_p)


class DeviceBuffer:
    """Device memory from the library's own HIP runtime, e.g. a snapshot kept on the device."""

    def __init__(self, nbytes, library=None):
        self.lib = load_library(library)
        self.nbytes = int(nbytes)
        self.ptr = ct.c_void_p()
        if self.lib.cav_device_alloc(self.nbytes, ct.byref(self.ptr)):
            raise RuntimeError(self.lib.cav_last_error().decode())

    def numpy(self):
        """A host copy of the contents."""
        out = np.empty(self.nbytes, dtype=np.uint8)
        if self.lib.cav_memcpy(pointer(out), self.ptr, self.nbytes):
            raise RuntimeError(self.lib.cav_last_error().decode())
        return out

    def close(self):
        if self.ptr:
            self.lib.cav_device_free(self.ptr)
            self.ptr = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _buffer(a, writeable=False):
    """Address and size in bytes of a snapshot buffer: a contiguous uint8 NumPy array or a DeviceBuffer."""
    if isinstance(a, DeviceBuffer):
        if not a.ptr:
            raise ValueError('device buffer is closed')
        return a.ptr, a.nbytes
    if not isinstance(a, np.ndarray) or not a.flags['C_CONTIGUOUS'] or (writeable and not a.flags['WRITEABLE']):
        raise ValueError('snapshot buffer must be a contiguous NumPy array' + (' (writeable)' if writeable else ''))
    return a.ctypes.data_as(ct.c_void_p), int(a.nbytes)


def make_config(settings):
    """Config record from a settings dict, as the reference binding fills it (python_api.py:59-66)."""
    config = Config()
    for name, _ in Config._fields_:
        if name in settings:
            setattr(config, name, settings[name])
    probes = list(settings.get('probes', ()))[:8]
    config.n_probes = len(probes)
    for k, (px, py) in enumerate(probes):
        config.probe_x[k], config.probe_y[k] = px, py
    return config


class Solver:
    """The native solver. Create one, call initialize(), then step() in a loop.

    handle=False (default) drives the process-wide solver through the reference's solver_* entry points;
    handle=True gives this object its own independent solver (cav_* entry points)."""

    def __init__(self, settings, library=None, handle=False, precision=None):
        self.lib = load_library(library)
        self.config = make_config(settings)
        self._h = None
        if handle:
            h = ct.c_void_p()
            if self.lib.cav_create(ct.byref(self.config), ct.byref(h)):
                raise RuntimeError(self.lib.cav_last_error().decode())
            self._h = h
        else:
            self.call('create', ct.byref(self.config))
        nx, ny, nz = self.config.nx, self.config.ny, self.config.nz
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shapes = {'u': (nz, ny, nx + 1), 'v': (nz, ny + 1, nx), 'w': (nz + 1, ny, nx), 'phi': (nz, ny, nx)}
        self.shapes.update(mx=self.shapes['u'], my=self.shapes['v'], mz=self.shapes['w'])
        self._read_grid()
        if precision is not None:
            self.set_precision(precision)

    def _read_grid(self):
        self.grid = {}
        for axis, (name, n) in enumerate((('x', self.nx), ('y', self.ny), ('z', self.nz))):
            faces, centres = np.empty(n + 1), np.empty(n)
            self.call('grid', axis, pointer(faces), pointer(centres))
            self.grid[name] = (faces, centres)

    def call(self, name, *args):
        if self._h is not None:
            if getattr(self.lib, 'cav_' + name)(self._h, *args):
                raise RuntimeError(self.lib.cav_last_error().decode())
        elif getattr(self.lib, 'solver_' + name)(*args):
            raise RuntimeError(self.lib.solver_last_error().decode())

    def initialize(self, v_noise=None, w_noise=None):
        v = None if v_noise is None else np.ascontiguousarray(v_noise, dtype=np.float64)
        w = None if w_noise is None else np.ascontiguousarray(w_noise, dtype=np.float64)
        if v is not None:
            assert v.size == int(np.prod(self.shapes['v'])), 'v_noise has the wrong size'
        if w is not None:
            assert w.size == int(np.prod(self.shapes['w'])), 'w_noise has the wrong size'
        self.call('initialize', pointer(v), pointer(w))

    def step(self):
        r = Report()
        self.call('step', ct.byref(r))
        return self._as_dict(r)

    def steps(self, n):
        """n steps in one library call (handle or singleton); report after the last."""
        r = Report()
        if self._h is not None:
            self.call('steps', int(n), ct.byref(r))
        else:
            for _ in range(int(n)):
                self.call('step', ct.byref(r))
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
        n = [np.ascontiguousarray(a, dtype=np.float64) for a in noise]
        for a, c in zip(n, 'uvw'):
            assert a.size == int(np.prod(self.shapes[c])), f'kick noise for {c} has the wrong size'
        self.call('kick', pointer(n[0]), pointer(n[1]), pointer(n[2]), float(amplitude), int(passes))

    def diagnostics(self):
        d = Diagnostics()
        self.call('diagnostics', ct.byref(d))
        return {name: getattr(d, name) for name, _ in Diagnostics._fields_}

    def field(self, name):
        out = np.empty(self.shapes[name], dtype=np.float64 if name == 'phi' else np.float32)
        self.call('get_field', FIELD_NUMBER[name], pointer(out))
        return out

    def set_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64 if name == 'phi' else np.float32)
        assert a.shape == self.shapes[name], f'{name}: shape {a.shape}, expected {self.shapes[name]}'
        self.call('set_field', FIELD_NUMBER[name], pointer(a))

    def field64(self, name):
        out = np.empty(self.shapes[name], dtype=np.float64)
        self.call('get_field64', FIELD_NUMBER[name], pointer(out))
        return out

    def set_field64(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        assert a.shape == self.shapes[name], f'{name}: shape {a.shape}, expected {self.shapes[name]}'
        self.call('set_field64', FIELD_NUMBER[name], pointer(a))

    def set_precision(self, mode):
        """'fp64' (default), 'fdm_fp32', 'fdm_bf16x3' or 'field_fp32' (the reference's float fields; entering or
        leaving it redoes the set-up on the float-rounded grid, carrying u and m over)."""
        self.call('set_precision', PRECISION[mode] if isinstance(mode, str) else int(mode))
        self._read_grid()

    def set_fdm_layout(self, symmetric):
        """symmetric=True (default): even/odd half transforms on axes whose grid mirrors; False: dense."""
        self.call('set_fdm_layout', int(bool(symmetric)))

    def fdm_layout(self):
        """Bit mask of the axes using the even/odd split (1 x, 2 y, 4 z)."""
        m = ct.c_int()
        self.call('fdm_layout', ct.byref(m))
        return m.value

    def pass_counters(self):
        a, b = ct.c_longlong(), ct.c_longlong()
        self.call('pass_counters', ct.byref(a), ct.byref(b))
        return a.value, b.value

    def redone_steps(self):
        """Steps redone with every refinement pass enqueued (handle solvers)."""
        n = ct.c_longlong()
        self.call('redone_steps', ct.byref(n))
        return n.value

    def set_enqueued_passes(self, passes):
        """Refinement passes a step enqueues before stalling (0: default); fewer make steps be redone."""
        self.call('set_enqueued_passes', int(passes))

    def _handle_only(self, what):
        if self._h is None:
            raise RuntimeError(what + ' needs a solver of its own (handle=True)')

    def snapshot_bytes(self):
        """Size of this solver's snapshot in bytes."""
        self._handle_only('snapshot')
        n = ct.c_longlong()
        self.call('snapshot_bytes', ct.byref(n))
        return n.value

    def snapshot(self, out=None):
        """The whole state (clock, lagged dt rates, warm-start history, u, m, phi, statistics sums) as a
        snapshot: into `out` (a uint8 NumPy array, or a DeviceBuffer for a device-resident copy) or into a new
        NumPy array. A solver of the same settings and precision continues bit for bit after restore()."""
        self._handle_only('snapshot')
        if out is None:
            out = np.empty(self.snapshot_bytes(), dtype=np.uint8)
        ptr, size = _buffer(out, writeable=True)
        self.call('snapshot', ptr, size)
        return out

    def restore(self, snap):
        """Take over the state of a snapshot (host or device memory)."""
        self._handle_only('restore')
        ptr, size = _buffer(snap)
        self.call('restore', ptr, size)

    def set_clock(self, time, steps, last_dt):
        self.call('set_clock', float(time), int(steps), float(last_dt))

    def probes(self):
        out = np.zeros(2 * self.config.n_probes, dtype=np.float32)
        self.call('probes', pointer(out))
        return out

    def add_statistics(self, weight):
        """Add the current mid-span field to the statistics; returns the centreline sample (u, v, uu, vv, ww)."""
        if self.nx != self.ny:
            self.call('add_statistics', float(weight), None)
            return None
        line = np.empty((5, self.nx))
        self.call('add_statistics', float(weight), pointer(line))
        return line

    def statistics(self):
        sums = np.empty((6, self.ny, self.nx))
        self.call('get_statistics', pointer(sums))
        return dict(zip(('u', 'v', 'uu', 'vv', 'ww', 'uv'), sums))

    def set_statistics(self, sums):
        a = np.ascontiguousarray([sums[k] for k in ('u', 'v', 'uu', 'vv', 'ww', 'uv')], dtype=np.float64)
        self.call('set_statistics', pointer(a))

    def midplane(self):
        out = np.empty((4, self.ny, self.nx), dtype=np.float32)
        self.call('midplane', pointer(out))
        return out

    def q_volume(self, colour_w=True):
        q = np.empty((self.nz, self.ny, self.nx), dtype=np.float32)
        colour = np.empty_like(q)
        self.call('q_volume', int(bool(colour_w)), pointer(q), pointer(colour))
        return q, colour

    # ---- timing hooks (handle solvers only)
    def time_steps(self, n):
        ms = ct.c_double()
        self.call('time_steps', int(n), ct.byref(ms))
        return ms.value

    def set_kernel_timing(self, on):
        self.call('set_kernel_timing', int(bool(on)))

    def kernel_times(self):
        return _kernel_times(lambda *a: self.call('kernel_times', *a))

    def set_graph(self, on):
        """Steps replay one captured HIP graph (default) or launch kernel by kernel."""
        self.call('set_graph', int(bool(on)))

    TUNE_KNOBS = ('stage_tile', 'stage_target', 'stage_minz', 'grad_target', 'grad_minz', 'transform_layout',
                  'grad_tile')

    def autotune(self):
        """Time the launch choices on this cavity now and keep the fastest (cached per configuration)."""
        self._handle_only('autotune')
        self.call('autotune')

    def tuning(self):
        """The cavity's launch choices, a dict of TUNE_KNOBS (0 = built-in default)."""
        self._handle_only('tuning')
        k = (ct.c_int * len(self.TUNE_KNOBS))()
        self.call('get_tuning', k, len(self.TUNE_KNOBS))
        return dict(zip(self.TUNE_KNOBS, list(k)))

    def set_tuning(self, **knobs):
        """Force launch choices (unnamed knobs keep their value)."""
        self._handle_only('set_tuning')
        cur = self.tuning()
        for name, v in knobs.items():
            if name not in cur:
                raise KeyError(name)
            cur[name] = int(v)
        k = (ct.c_int * len(self.TUNE_KNOBS))(*[cur[n] for n in self.TUNE_KNOBS])
        self.call('set_tuning', k, len(self.TUNE_KNOBS))

    def close(self):
        if self._h is not None:
            if self._h:
                self.lib.cav_destroy(self._h)
            self._h = ct.c_void_p()
        else:
            self.call('destroy')

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _kernel_times(call):
    cap = 64
    names = (ct.c_char_p * cap)()
    ms = (ct.c_double * cap)()
    launches = (ct.c_longlong * cap)()
    count = ct.c_int()
    call(cap, names, ms, launches, ct.byref(count))
    return {names[k].decode(): (ms[k], launches[k]) for k in range(count.value)}


class Ensemble:
    """Handle solvers advanced together: every kernel of a step runs once for all members, and the step
    replays as one HIP graph. Members keep working on their own between ensemble calls; close the ensemble
    before closing a member."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if not self.solvers or any(s._h is None for s in self.solvers):
            raise ValueError('an ensemble needs handle solvers (Solver(..., handle=True))')
        self.lib = self.solvers[0].lib
        handles = (ct.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ct.c_void_p()
        self._call(self.lib.cav_ensemble_create, handles, len(self.solvers), ct.byref(h))
        self._h = h

    def _call(self, fn, *args):
        if fn(*args):
            raise RuntimeError(self.lib.cav_last_error().decode())

    def steps(self, n):
        """n steps of every member. Returns (reports, errors): one report dict per member, and the error
        text per member ('' when it stepped cleanly)."""
        k = len(self.solvers)
        reports = (Report * k)()
        status = (ct.c_int * k)()
        self._call(self.lib.cav_ensemble_steps, self._h, int(n), reports, status)
        return ([Solver._as_dict(r) for r in reports],
                [self.lib.cav_error_text(status[m]).decode() for m in range(k)])

    def time_steps(self, n):
        """(total ms, per-step ms array, error texts) of n steps timed with HIP events."""
        k = len(self.solvers)
        total = ct.c_double()
        per = np.empty(int(n))
        status = (ct.c_int * k)()
        self._call(self.lib.cav_ensemble_time_steps, self._h, int(n), ct.byref(total), pointer(per), status)
        return total.value, per, [self.lib.cav_error_text(status[m]).decode() for m in range(k)]

    def set_graph(self, on):
        self._call(self.lib.cav_ensemble_set_graph, self._h, int(bool(on)))

    def autotune(self):
        """Tune the members' launch choices for stepping together (cached per member list)."""
        self._call(self.lib.cav_ensemble_autotune, self._h)

    def set_enqueued_passes(self, passes):
        self._call(self.lib.cav_ensemble_set_enqueued_passes, self._h, int(passes))

    def set_kernel_timing(self, on):
        self._call(self.lib.cav_ensemble_set_kernel_timing, self._h, int(bool(on)))

    def kernel_times(self):
        return _kernel_times(lambda *a: self._call(self.lib.cav_ensemble_kernel_times, self._h, *a))

    def close(self):
        if getattr(self, '_h', None):
            self.lib.cav_ensemble_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def set_autotune(on, library=None):
    """Tune every cavity as it is created (default on); off: cavities created later keep the built-in launch
    choices."""
    lib = load_library(library)
    if lib.cav_set_autotune(int(bool(on))):
        raise RuntimeError(lib.cav_last_error().decode())


def batch_steps_noalloc(solvers, n):
    """n steps of handle solvers stepped together with no device allocation (cav_batch_steps_noalloc): the
    results of an Ensemble of the same solvers, through a launch plan kept in the library's fixed device
    array (at most 256 solvers per call). Returns (reports, errors) as Ensemble.steps does."""
    solvers = list(solvers)
    if not solvers or any(s._h is None for s in solvers):
        raise ValueError('batch_steps_noalloc needs handle solvers (Solver(..., handle=True))')
    lib = solvers[0].lib
    k = len(solvers)
    handles = (ct.c_void_p * k)(*[s._h.value for s in solvers])
    reports = (Report * k)()
    status = (ct.c_int * k)()
    if lib.cav_batch_steps_noalloc(handles, k, int(n), reports, status):
        raise RuntimeError(lib.cav_last_error().decode())
    return ([Solver._as_dict(r) for r in reports], [lib.cav_error_text(status[m]).decode() for m in range(k)])


def footprint(settings, precision='fp64', library=None):
    """Device bytes one solver of these settings holds."""
    lib = load_library(library)
    config = make_config(settings)
    out = ct.c_longlong()
    mode = PRECISION[precision] if isinstance(precision, str) else int(precision)
    if lib.cav_footprint(ct.byref(config), mode, ct.byref(out)):
        raise RuntimeError(lib.cav_last_error().decode())
    return out.value
