"""The benchmark configurations at their full sizes, through properties that do not depend on the size (the
oracle would take minutes per step there): every projection reaches its tolerance, wall faces stay exactly
zero, probes read the centred velocity of their cells, the mid-span statistics obey Cauchy-Schwarz, a
Gaussian-blob kick leaves m - u unchanged (backend.cu:958-978 adds the same change to both), and the
split-BF16 FDM matches the FP64 FDM."""
import numpy as np
import pytest

from conftest import rel_err
from les_impulse3d_64_amd import workloads as W

pytestmark = pytest.mark.gpu


def _walls_zero(g):
    u, v, w = g.field64('u'), g.field64('v'), g.field64('w')
    return (np.abs(u[..., [0, -1]]).max() == 0 and np.abs(v[:, [0, -1]]).max() == 0
            and np.abs(w[[0, -1]]).max() == 0)


def test_c1_128_statistics_and_probes(hip_lib):
    """128^3, stretch 1.5, Re 1000, uniform lid, CFL 0.5, 8 probes, statistics weighted by dt every step."""
    from les_impulse3d_64_amd.python_api import Solver
    s, noise, _ = W.c1_stats()
    print('case: c1 128^3, 25 steps', flush=True)
    g = Solver(s, handle=True)
    g.initialize(*noise)
    fx, fy = g.grid['x'][0], g.grid['y'][0]
    cells = [(int(np.searchsorted(fx, px * s['length'], side='right')) - 1,
              int(np.searchsorted(fy, py * s['length'], side='right')) - 1) for px, py in W.PROBES8]
    total_w = 0.0
    for k in range(25):
        rep = g.step()
        assert rep['phi_rel'] <= s['fdm_rtol'] and rep['phi_passes'] >= 0, rep
        assert rep['dt_reason'] in ('CFL', 'viscous', 'DT_MAX'), rep
        probes = g.probes()
        if k % 8 == 0:
            u, v = g.field64('u'), g.field64('v')
            z = s['nz'] // 2
            want = np.array([[0.5 * (u[z, cy, cx] + u[z, cy, cx + 1]), 0.5 * (v[z, cy, cx] + v[z, cy + 1, cx])]
                             for cx, cy in cells]).ravel()
            np.testing.assert_allclose(probes, want.astype(np.float32), rtol=1e-6, atol=1e-9)
        line = g.add_statistics(rep['dt'])
        total_w += rep['dt']
        assert np.isfinite(line).all()
    assert _walls_zero(g)
    d = g.diagnostics()
    print(f"c1: t={rep['time']:.4g} KE={d['kinetic_energy']:.4g} div_rms={d['div_rms']:.2e}", flush=True)
    assert d['div_rms'] < 1e-8
    st = g.statistics()
    # (sum w a)^2 <= (sum w)(sum w a^2) and (sum w a b)^2 <= (sum w a^2)(sum w b^2), to rounding
    for a, aa in (('u', 'uu'), ('v', 'vv')):
        assert (st[aa] >= 0).all()
        assert (st[a] ** 2 <= total_w * st[aa] * (1 + 1e-12) + 1e-300).all(), a
    assert (st['ww'] >= 0).all()
    assert (st['uv'] ** 2 <= st['uu'] * st['vv'] * (1 + 1e-12) + 1e-300).all()
    g.close()


def test_c2_256_kick_and_mixed_fdm(hip_lib):
    """256^3, stretch 2.0, Re 3200, regularised lid, FDM rtol 1e-10, FP64 and split-BF16 FDM side by side: two
    steps, a Gaussian-blob kick (seed 2001: amplitude 0.05, radius 0.05), two steps. Both reach the tolerance
    at every projection; the kick keeps m - u and adds a divergence-free, nonzero velocity; the two modes'
    fields agree."""
    from les_impulse3d_64_amd.python_api import Solver
    s, noise, _ = W.c2_mixed(256)
    print('case: c2 256^3 fp64 + bf16x3 with a kick', flush=True)
    runs = {}
    for mode in ('fp64', 'fdm_bf16x3'):
        g = Solver(s, handle=True, precision=mode)
        g.initialize(*noise)
        for _ in range(2):
            rep = g.step()
            assert rep['phi_rel'] <= 1e-10, (mode, rep)
        u0 = {c: g.field64(c) for c in ('u', 'v', 'w')}
        gap0 = {c: g.field64(mc) - u0[c] for c, mc in (('u', 'mx'), ('v', 'my'), ('w', 'mz'))}
        g.kick(*W.kick_blobs(s, W.KICK_SEEDS[0], g.grid))
        u1 = {c: g.field64(c) for c in ('u', 'v', 'w')}
        for c, mc in (('u', 'mx'), ('v', 'my'), ('w', 'mz')):
            gap1 = g.field64(mc) - u1[c]
            scale = np.abs(g.field64(mc)).max()
            assert np.abs(gap1 - gap0[c]).max() <= 1e-13 * scale, (mode, c)
        added = max(np.abs(u1[c] - u0[c]).max() for c in u0)
        assert added > 1e-3, (mode, added)
        d = g.diagnostics()
        assert d['div_rms'] < 1e-8, (mode, d['div_rms'])
        assert _walls_zero(g)
        for _ in range(2):
            rep = g.step()
            assert rep['phi_rel'] <= 1e-10, (mode, rep)
        runs[mode] = {c: g.field64(c) for c in ('u', 'v', 'w', 'phi')}
        print(f"c2 {mode}: t={rep['time']:.4g} passes={rep['phi_passes']} phi_rel={rep['phi_rel']:.2e}", flush=True)
        g.close()
    for c in ('u', 'v', 'w'):
        assert rel_err(runs['fdm_bf16x3'][c], runs['fp64'][c]) < 1e-8, c


def test_c2_256_modes_agree_over_200_steps(hip_lib):
    """The four precision modes over 200 steps of the 256^3 workload with a blob kick after step 100: every
    projection reaches rtol 1e-10 in every mode, and the FP32 and split-BF16 FDM velocities stay within 1e-7
    of the FP64 ones (the refinement bounds each projection's error by the tolerance; the run is laminar). The
    FP32-field mode (the reference's precision) rounds the velocity and impulse to float at every stage: its
    velocity stays within 1e-4 of FP64's, its divergence is float rounding (div_rms below 1e-5), and its fields
    are float values."""
    from les_impulse3d_64_amd.python_api import Solver
    s, noise, _ = W.c2_mixed(256)
    print('case: c2 256^3, four modes, 200 steps with a kick', flush=True)
    runs = {}
    for mode in ('fp64', 'fdm_fp32', 'fdm_bf16x3', 'field_fp32'):
        g = Solver(s, handle=True, precision=mode)
        g.initialize(*noise)
        rep = g.steps(100)
        assert rep['phi_rel'] <= 1e-10, (mode, rep)
        g.kick(*W.kick_blobs(s, W.KICK_SEEDS[0], g.grid))
        rep = g.steps(100)
        assert rep['steps'] == 200 and rep['phi_rel'] <= 1e-10, (mode, rep)
        d = g.diagnostics()
        print(f"c2 {mode}: t={rep['time']:.4g} KE={d['kinetic_energy']:.4g} div_rms={d['div_rms']:.2e} "
              f"passes={g.pass_counters()}", flush=True)
        assert d['div_rms'] < (1e-5 if mode == 'field_fp32' else 1e-8), (mode, d['div_rms'])
        assert _walls_zero(g)
        runs[mode] = {c: g.field64(c) for c in ('u', 'v', 'w')}
        if mode == 'field_fp32':
            for c in ('u', 'v', 'w'):
                assert np.array_equal(runs[mode][c], runs[mode][c].astype(np.float32).astype(np.float64)), c
        g.close()
    for mode, bound in (('fdm_fp32', 1e-7), ('fdm_bf16x3', 1e-7), ('field_fp32', 1e-4)):
        worst = max(rel_err(runs[mode][c], runs['fp64'][c]) for c in ('u', 'v', 'w'))
        print(f'c2 {mode} vs fp64 after 200 steps: {worst:.2e}', flush=True)
        assert worst < bound, (mode, worst)
