"""The C-ABI boundary without a GPU: the library loads, exports every function include/cavity_solver.h
declares (the 18 reference exports of backend.cu:980-1195 among them), and the ctypes records have the
reference byte layouts (python_api.py:11-28)."""
import ctypes as ct
import re
import subprocess

from conftest import ROOT
from les_impulse3d_64_amd.python_api import Config, Report, Diagnostics, make_config

HEADER = ROOT / 'include' / 'cavity_solver.h'
LIB = ROOT / 'les_impulse3d_64_amd' / 'libcavity_solver.so'

REFERENCE_EXPORTS = ['solver_create', 'solver_destroy', 'solver_initialize', 'solver_step', 'solver_kick',
                     'solver_report', 'solver_diagnostics', 'solver_grid', 'solver_get_field', 'solver_set_field',
                     'solver_set_clock', 'solver_probes', 'solver_add_statistics', 'solver_get_statistics',
                     'solver_set_statistics', 'solver_midplane', 'solver_q_volume', 'solver_last_error']


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b((?:solver|cav)_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_reference_exports():
    names = declared_functions()
    for n in REFERENCE_EXPORTS:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    assert LIB.is_file(), 'build first: python -c "import __graft_entry__ as g; g.build()"'
    out = subprocess.run(['nm', '-D', '--defined-only', str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if ' T ' in line}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_without_gpu():
    lib = ct.CDLL(str(LIB))
    lib.solver_last_error.restype = ct.c_char_p
    assert lib.solver_last_error() == b''
    # a call that fails before touching the device: no solver yet
    r = Report()
    assert lib.solver_report(ct.byref(r)) == 1
    assert b'no solver' in lib.solver_last_error()


def test_record_layouts():
    assert ct.sizeof(Config) == 288 and ct.sizeof(Report) == 48 and ct.sizeof(Diagnostics) == 64
    assert Config.nx.offset == 0 and Config.n_probes.offset == 40 and Config.length.offset == 48
    assert Config.fdm_rtol.offset == 152 and Config.probe_x.offset == 160 and Config.probe_y.offset == 224
    assert Report.phi_passes.offset == 44 and Report.dt_reason.offset == 40


def test_header_layout_matches_ctypes():
    """Compile a probe against the header and compare sizeof/offsetof with the ctypes records."""
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "cavity_solver.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(CavConfig), sizeof(CavReport), sizeof(CavDiagnostics),
         offsetof(CavConfig, n_probes), offsetof(CavConfig, length), offsetof(CavConfig, fdm_rtol),
         offsetof(CavConfig, probe_x), offsetof(CavConfig, probe_y), offsetof(CavReport, phi_passes));
  return 0;
}
'''
    import tempfile
    import os
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, 'probe.c')
        exe = os.path.join(d, 'probe')
        open(c, 'w').write(src)
        subprocess.run(['gcc', '-I', str(ROOT / 'include'), c, '-o', exe], check=True)
        vals = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert vals == [288, 48, 64, 40, 48, 152, 160, 224, 44]


def test_make_config_mirrors_reference_binding():
    c = make_config({'nx': 32, 'ny': 32, 'nz': 8, 'nu': 1e-3, 'probes': [(0.5, 0.5), (0.2, 0.8)] * 6})
    assert c.n_probes == 8 and c.probe_y[1] == 0.8 and c.nu == 1e-3 and c.lid_shaped == 0


def test_grid_size_limit_without_gpu():
    """cav_footprint validates a configuration on the host: the face arrays of a role (viewed as one buffer
    with 32-bit byte offsets) must stay below 4 GiB together. 562^3 is the largest accepted cube; 564^3 and
    1024 x 1024 x 200 (accepted by the reference, backend.cu:1018-1020) are refused with the stated message."""
    import pytest
    from les_impulse3d_64_amd import workloads as W
    from les_impulse3d_64_amd.python_api import footprint

    def faces_bytes(nx, ny, nz):
        return 8 * ((nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)) + 512

    b = footprint(W.settings(562), 'fp64')
    assert faces_bytes(562, 562, 562) < 2 ** 32 and b > 7 * faces_bytes(562, 562, 562) - 7 * 512
    assert footprint(W.settings(562, 562, 563), 'fp64') > b
    for dims in ((564, 564, 564), (600, 600, 600), (1024, 1024, 200)):
        assert faces_bytes(*dims) >= 2 ** 32
        with pytest.raises(RuntimeError, match='face arrays of a field must stay below 4 GiB'):
            footprint(W.settings(*dims), 'fp64')
    with pytest.raises(RuntimeError, match='at most 1024 cells per axis'):
        footprint(W.settings(8, 8, 1026), 'fp64')


def test_noalloc_batch_refuses_more_than_its_plan_holds():
    """cav_batch_steps_noalloc checks the member count against its fixed plan array before touching the device."""
    lib = ct.CDLL(str(LIB))
    lib.cav_last_error.restype = ct.c_char_p
    lib.cav_batch_steps_noalloc.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
    for n in (0, 257, 10000):
        assert lib.cav_batch_steps_noalloc(None, n, 1, None, None) == 1
        assert b'between 1 and 256 cavities per call' in lib.cav_last_error()
