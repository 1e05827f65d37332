"""The projection warm start of the oracle (guess_phi, backend.cu:857-873), pinned by closed forms.

History entry k holds dphi_k = s_k f(t_k) for a per-cell quadratic f(t) = a + b t + c t^2. The straight line
through (t1, f(t1)) and (t2, f(t2)) takes at tau the value f(tau) - c (tau - t1)(tau - t2). So the guess for
(tau, scale) is scale times that value for the pair the rule picks: the entry nearest in time to tau (the later
entry on ties), then the next nearest whose time differs from it by more than 1e-12 relative. With one usable
time the guess is scale f(t1). Each case is built so that a wrong pair, a wrong tie-break or a dropped entry
scale gives a different value.
"""
import numpy as np
import pytest

from les_impulse3d_64_amd import workloads as W
from oracle import oracle as O

SCALE = 0.9


@pytest.fixture(scope='module')
def setup():
    o = O.Oracle(W.settings(6, 8, 5, stretch=1.5))
    rng = np.random.default_rng(5)
    shape = o.shapes['phi']
    a, b, c = (rng.standard_normal(shape) for _ in range(3))
    yield o, (a, b, c)
    o.close()


def _f(coef, t):
    a, b, c = coef
    return a + b * t + c * t * t


def _line(coef, t1, t2, tau):
    return _f(coef, tau) - coef[2] * (tau - t1) * (tau - t2)


def _history(coef, times, scales, garbage=()):
    """Entries s_k f(t_k); entries listed in garbage hold s_k (f(t_k) + 1) instead."""
    return [s * (_f(coef, t) + (1.0 if k in garbage else 0.0)) for k, (t, s) in enumerate(zip(times, scales))]


def _close(got, want):
    assert got is not None
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def _far(got, want):
    assert np.abs(got - want).max() > 1e-3


def test_extrapolates_from_the_two_nearest_entries(setup):
    o, coef = setup
    times, scales, tau = [0.1, 0.3, 0.45], [0.7, 1.3, 2.1], 0.5
    got = o.guess_phi(times, scales, _history(coef, times, scales), tau, SCALE)
    _close(got, SCALE * _line(coef, 0.45, 0.3, tau))
    _far(got, SCALE * _line(coef, 0.45, 0.1, tau))       # the oldest entry instead of the second nearest
    _far(got, SCALE * _f(coef, 0.45))                      # constant extrapolation


def test_interpolates_between_equally_near_entries(setup):
    o, coef = setup
    times, scales, tau = [0.125, 0.375, 0.5], [1.5, 0.25, 3.0], 0.25   # |t - tau| = 0.125 exactly for two
    got = o.guess_phi(times, scales, _history(coef, times, scales), tau, SCALE)
    _close(got, SCALE * _line(coef, 0.375, 0.125, tau))
    _far(got, SCALE * _line(coef, 0.375, 0.5, tau))


def test_later_entry_wins_a_tie_and_the_second_needs_another_time(setup):
    o, coef = setup
    # entries 0 and 1 share a time: the later one (1) is taken, entry 0 is skipped as the second point
    times, scales, tau = [0.25, 0.25, 0.5], [0.5, 1.75, 1.25], 0.3
    got = o.guess_phi(times, scales, _history(coef, times, scales, garbage=(0,)), tau, SCALE)
    _close(got, SCALE * _line(coef, 0.25, 0.5, tau))
    # the same with the later one spoiled: the guess changes
    bad = o.guess_phi(times, scales, _history(coef, times, scales, garbage=(1,)), tau, SCALE)
    _far(bad, SCALE * _line(coef, 0.25, 0.5, tau))


def test_one_usable_time_gives_the_scaled_entry(setup):
    o, coef = setup
    _close(o.guess_phi([0.4], [1.6], [1.6 * _f(coef, 0.4)], 0.55, SCALE), SCALE * _f(coef, 0.4))
    # two entries closer in time than 1e-12 relative count as one time: the nearer (later) entry alone
    times, scales = [0.5, 0.5 + 1e-14], [0.8, 1.9]
    got = o.guess_phi(times, scales, _history(coef, times, scales, garbage=(0,)), 0.6, SCALE)
    _close(got, SCALE * _f(coef, 0.5 + 1e-14))


def test_no_guess_without_history_or_when_disabled(setup):
    o, coef = setup
    assert o.guess_phi([], [], [], 0.3, SCALE) is None
    off = O.Oracle(W.settings(6, 8, 5, stretch=1.5, phi_extrapolate=0))
    try:
        assert off.guess_phi([0.1, 0.2], [1.0, 1.0], [_f(coef, 0.1), _f(coef, 0.2)], 0.3, SCALE) is None
    finally:
        off.close()
