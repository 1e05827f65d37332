"""Pins the oracle's grid, eigensystem, pressure solve and projection against closed forms, an independent
LAPACK eigensolver, a dense pseudo-inverse and invariants (CPU only).

Oracle functions covered here: tanh_faces/build_axis (backend.cu:643-699), axis_eigen/jacobi_eigen
(backend.cu:723-787), pressure_residual (345-359), pressure_pass (801-814), solve_pressure (816-847),
divergence/subtract_gradient/zero_walls/project (334-380, 849-854), reduce (424-443, 629-640).
"""
import numpy as np
import pytest

from conftest import rel_err
from les_impulse3d_64_amd import workloads as W
from oracle import oracle as O


def test_uniform_grid_is_linspace():
    s = W.settings(12, 12, 7)
    f = O.axis_faces(s, 12, 1.0)
    assert np.array_equal(f, np.array([i * (1.0 / 12) for i in range(12)] + [1.0]))
    o = O.Oracle(s)
    fz = o.grid['z'][0]
    assert fz[-1] == 7 * (1.0 / 7)       # z length is nz*(lz/nz) (backend.cu:1026)
    assert np.allclose(o.widths['x'], 1.0 / 12, rtol=1e-14)


@pytest.mark.parametrize('n,cap', [(16, 1.5), (32, 2.0), (48, 3.0), (64, 10.0)])
def test_stretched_grid_closed_form(n, cap):
    s = W.settings(n, stretch=cap)
    f = O.axis_faces(s, n, 1.0)
    assert f[0] == 0.0 and f[-1] == 1.0 and f[n // 2] == 0.5
    w = np.diff(f)
    assert np.all(w > 0)
    ratio = w.max() / w.min()
    assert cap * (1 - 1e-6) <= ratio <= cap * (1 + 1e-12)     # the aspect cap binds (60-step bisection)
    assert np.allclose(f + f[::-1], 1.0, atol=1e-14)            # symmetric about the centre
    # the faces are L/2 (1 + tanh(beta (2 xi - 1)) / tanh beta) for one beta: recover beta from f[1]
    xi = np.arange(n + 1) / n
    from scipy.optimize import brentq
    beta = brentq(lambda b: 0.5 * (1 + np.tanh(b * (2 / n - 1)) / np.tanh(b)) - f[1], 1e-6, 30)
    closed = 0.5 * (1 + np.tanh(beta * (2 * xi - 1)) / np.tanh(beta))
    closed[0], closed[-1], closed[n // 2] = 0.0, 1.0, 0.5
    assert np.allclose(f, closed, rtol=0, atol=1e-12)


def test_dx_wall_without_cap():
    s = W.settings(32, grid_stretch=1, grid_dx_wall=0.004, grid_max_aspect=0.0)
    f = O.axis_faces(s, 32, 1.0)
    assert abs((f[1] - f[0]) - 0.004) < 1e-12


def _operator_1d(widths, spacings):
    """K (n x n): Neumann second-difference with coefficients 1/spacing at interior faces."""
    n = len(widths)
    K = np.zeros((n, n))
    for i in range(n - 1):
        k = 1.0 / spacings[i + 1]
        K[i, i] -= k
        K[i + 1, i + 1] -= k
        K[i, i + 1] += k
        K[i + 1, i] += k
    return K


def test_uniform_eigen_closed_form():
    n = 16
    o = O.Oracle(W.settings(n, n, 8))
    S, lam, zm = o.eigen(0)
    h = 1.0 / n
    k = np.arange(n)
    expect = np.sort(-(4 / h ** 2) * np.sin(k * np.pi / (2 * n)) ** 2)
    assert np.allclose(np.sort(lam), expect, rtol=1e-12, atol=1e-9)
    assert lam[zm] == 0.0
    # each eigenvector is a DCT-II mode cos(k pi (i + 1/2) / n), up to sign and scale
    i = np.arange(n)
    for j in range(n):
        kk = int(round(np.arcsin(min(1.0, np.sqrt(max(-lam[j], 0) * h ** 2 / 4))) * 2 * n / np.pi)) if j != zm else 0
        mode = np.cos(kk * np.pi * (i + 0.5) / n)
        v = S[:, j]
        c = np.dot(v, mode) / np.dot(mode, mode)
        assert np.allclose(v, c * mode, atol=1e-10)


@pytest.mark.parametrize('axis', [0, 1, 2])
def test_stretched_eigen_generalised(axis):
    o = O.Oracle(W.settings(20, 14, 9, stretch=2.5))
    name = 'xyz'[axis]
    wdt, spc = o.widths[name], o.spacings[name]
    S, lam, zm = o.eigen(axis)
    K = _operator_1d(wdt, spc)
    Wm = np.diag(wdt)
    lam_full = lam.copy()
    lam_full[zm] = 0.0
    assert np.abs(K @ S - Wm @ S @ np.diag(lam_full)).max() < 1e-9 * np.abs(K).max()
    assert np.allclose(S.T @ Wm @ S, np.eye(len(wdt)), atol=1e-12)
    # eigenvalues against LAPACK on the symmetric form W^-1/2 K W^-1/2
    A = K / np.sqrt(np.outer(wdt, wdt))
    ref = np.linalg.eigvalsh(A)
    mine = np.sort(np.where(np.arange(len(lam)) == zm, 0.0, lam))
    ref = np.sort(ref)
    ref[np.argmax(ref)] = 0.0
    assert np.allclose(mine, np.sort(ref), rtol=1e-11, atol=1e-8)
    # the zero mode is the constant vector
    v = S[:, zm]
    assert np.allclose(v, v[0], rtol=1e-10)


def _dense_laplacian(o):
    """The 3-D pressure operator from Kronecker sums: L = Kx (x) Wy (x) Wz + Wx (x) Ky (x) Wz + Wx (x) Wy (x) Kz,
    in the [z, y, x] flattening (x fastest)."""
    Kx = _operator_1d(o.widths['x'], o.spacings['x'])
    Ky = _operator_1d(o.widths['y'], o.spacings['y'])
    Kz = _operator_1d(o.widths['z'], o.spacings['z'])
    Wx, Wy, Wz = (np.diag(o.widths[a]) for a in 'xyz')
    return np.kron(Kz, np.kron(Wy, Wx)) + np.kron(Wz, np.kron(Ky, Wx)) + np.kron(Wz, np.kron(Wy, Kx))


@pytest.fixture(scope='module')
def small():
    o = O.Oracle(W.settings(6, 8, 5, stretch=2.0, fdm_rtol=1e-12))
    return o, _dense_laplacian(o)


def test_residual_is_rhs_minus_operator(small):
    o, L = small
    rng = np.random.default_rng(1)
    phi, rhs = rng.standard_normal(o.shapes['phi']), rng.standard_normal(o.shapes['phi'])
    r = o.residual(phi, rhs)
    assert rel_err(r.ravel(), rhs.ravel() - L @ phi.ravel()) < 1e-13


def _compatible_rhs(o, seed):
    rng = np.random.default_rng(seed)
    f = [rng.standard_normal(o.shapes[c]) for c in 'uvw']
    f[0][..., [0, -1]] = 0
    f[1][:, [0, -1]] = 0
    f[2][[0, -1]] = 0
    return o.divergence(*f)


def test_divergence_integrates_to_zero(small):
    o, _ = small
    rhs = _compatible_rhs(o, 2)
    assert abs(rhs.sum()) < 1e-13 * np.abs(rhs).sum()


def test_fdm_pass_is_exact_inverse(small):
    """One FP64 pass from phi = 0 solves L phi = r for a compatible r."""
    o, L = small
    rhs = _compatible_rhs(o, 3)
    phi = o.pressure_pass(rhs)
    assert rel_err(L @ phi.ravel(), rhs.ravel()) < 1e-12


def test_solve_pressure_matches_pseudo_inverse(small):
    o, L = small
    rhs = _compatible_rhs(o, 4)
    phi, rel, passes = o.solve_pressure(rhs)
    ref = np.linalg.pinv(L) @ rhs.ravel()             # minimum norm = zero (unweighted) mean
    assert rel_err(phi.ravel(), ref) < 1e-10
    assert rel <= 1e-12 and passes >= 1
    assert abs(phi.mean()) < 1e-14 * np.abs(phi).max()
    # a guess that already solves it needs no pass; a useless guess is dropped
    _, rel0, p0 = o.solve_pressure(rhs, guess=phi)
    assert p0 == 0 and rel0 <= 1e-12
    _, rel1, p1 = o.solve_pressure(rhs, guess=-50 * phi)
    assert rel1 <= 1e-12 and p1 >= 1


def test_solve_pressure_zero_rhs(small):
    o, _ = small
    phi, rel, passes = o.solve_pressure(np.zeros(o.shapes['phi']))
    assert np.all(phi == 0) and rel == 0 and passes == 0


def test_projection_invariants():
    s = W.settings(12, 10, 8, stretch=2.0, fdm_rtol=1e-11)
    o = O.Oracle(s)
    v, w = W.noise(s, 5, 1.0)
    o.initialize(v, w)
    u = [o.field(c) for c in 'uvw']
    assert np.abs(u[0][..., [0, -1]]).max() == 0 and np.abs(u[1][:, [0, -1]]).max() == 0
    assert np.abs(u[2][[0, -1]]).max() == 0
    m = [o.field(c) for c in ('mx', 'my', 'mz')]
    rhs0 = o.divergence(*m)
    div = o.divergence(*u)
    rep = o.report()
    assert np.linalg.norm(div) / np.linalg.norm(rhs0) <= 1e-11
    assert abs(np.linalg.norm(div) / np.linalg.norm(rhs0) - rep['phi_rel']) < 1e-3 * max(rep['phi_rel'], 1e-16) + 1e-15
    # u = m - grad(phi) on interior faces with phi the stored potential
    phi = o.field('phi')
    ug = o.subtract_gradient(0, m[0], phi)
    assert np.array_equal(ug, u[0])
    # the projection removes an added gradient exactly (to the solver tolerance): set u := u + grad(chi),
    # then kick with zero noise, which is project(u)
    chi = np.random.default_rng(6).standard_normal(o.shapes['phi'])
    for k, c in enumerate('uvw'):
        o.set_field(c, u[k] - o.subtract_gradient(k, np.zeros(o.shapes[c]), chi))
    o.kick([np.zeros(o.shapes[c]) for c in 'uvw'], 0.0, 1)
    for k, c in enumerate('uvw'):
        assert rel_err(o.field(c), u[k]) < 1e-9
    # subtract_gradient of a linear potential gives its slope on interior faces
    xc = o.grid['x'][1]
    lin = np.broadcast_to(3.0 * xc, o.shapes['phi']).copy()
    g0 = o.subtract_gradient(0, np.zeros(o.shapes['u']), lin)
    assert np.allclose(g0[..., 1:-1], -3.0, rtol=1e-12) and np.all(g0[..., [0, -1]] == 0)


def test_reduce_matches_numpy():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(200001)
    assert abs(O.reduce(x, 1) - x.sum()) < 1e-11
    assert O.reduce(x, 0) == x.max()
    assert O.reduce(x, 2) == np.abs(x).max()
    assert abs(O.reduce(x, 3) - (x * x).sum()) < 1e-9
    y = x.copy()
    y[123456] = np.nan
    assert np.isnan(O.reduce(y, 0))          # NaN wins a max (backend.cu:426-428)
    assert O.reduce(np.array([-1.0, -2.0]), 0) == -1.0
