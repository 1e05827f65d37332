"""Host code under AddressSanitizer (CPU only): the library's set-up routines (csrc/setup.cpp, built with g++
and -fsanitize=address into a throwaway program) and the oracle (oracle/Makefile target `asan`, loaded into a
child interpreter with the sanitizer runtime preloaded) run small cases and must finish with no report."""
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / 'les_impulse3d_64_amd' / 'csrc'


def _libasan():
    gxx = shutil.which('g++')
    if not gxx:
        return None
    p = subprocess.run([gxx, '-print-file-name=libasan.so'], capture_output=True, text=True).stdout.strip()
    return p if p and os.path.isfile(p) else None


pytestmark = pytest.mark.skipif(_libasan() is None, reason='g++ with libasan is not available')


def test_setup_routines_under_asan():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, 'setup_asan')
        cmd = ['g++', '-std=c++17', '-O1', '-g', '-fno-omit-frame-pointer', '-fsanitize=address',
               '-fsanitize=undefined', '-D__HIP_PLATFORM_AMD__', '-I/opt/rocm/include', f'-I{CSRC}',
               str(ROOT / 'tests' / 'host_asan' / 'setup_asan.cpp'), str(CSRC / 'setup.cpp'), '-o', exe]
        b = subprocess.run(cmd, capture_output=True, text=True)
        assert b.returncode == 0, b.stderr[-3000:]
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:abort_on_error=0',
                   UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1')
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0 and 'setup asan ok' in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
        assert 'AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr, r.stderr[-4000:]


ORACLE_SCRIPT = r'''
import numpy as np
assert 'libasan' in open('/proc/self/maps').read(), 'the sanitizer runtime is not loaded'
from les_impulse3d_64_amd import workloads as W
from oracle.oracle import Oracle
for nx, ny, nz, order, stretch in ((6, 6, 5, 4, 1.0), (10, 8, 7, 4, 1.8), (12, 12, 9, 2, 1.5)):
    s = W.settings(nx, ny, nz, re=400.0, stretch=stretch, lid_shaped=1, adaptive_dt=1, cfl_target=0.5,
                   fdm_rtol=1e-10, impulse_order=order)
    o = Oracle(s)
    o.initialize(*W.noise(s, 3, 1e-2))
    o.steps(3)
    o.kick(W.kick_noise(s, 4), 0.05, 1)
    o.step()
    o.diagnostics()
    o.add_statistics(0.01)
    o.probes()
    o.midplane()
    o.q_volume(1)
    for f in ('u', 'v', 'w', 'mx', 'my', 'mz', 'phi'):
        o.set_field(f, o.field(f))
    o.step()
    o.close()
print('oracle asan ok')
'''


def test_oracle_under_asan():
    lib = _libasan()
    b = subprocess.run(['make', '-s', '-C', str(ROOT / 'oracle'), 'asan'], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, LD_PRELOAD=lib, ASAN_OPTIONS='detect_leaks=0:abort_on_error=0', OMP_NUM_THREADS='2',
               ORACLE_LIB=str(ROOT / 'oracle' / '_asan' / 'liboracle.so'), PYTHONPATH=str(ROOT))
    r = subprocess.run([sys.executable, '-c', ORACLE_SCRIPT], capture_output=True, text=True, env=env, cwd=str(ROOT),
                       timeout=900)
    assert r.returncode == 0 and 'oracle asan ok' in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
    assert 'AddressSanitizer' not in r.stderr, r.stderr[-4000:]
