"""Host-side parts of the serving loop: the seeded request mix and the statistics."""
import numpy as np

from les_impulse3d_64_amd.serving import Request, ServeStats, poisson_requests


def test_request_mix():
    reqs = poisson_requests(30.0, rate=20.0, seed=4)
    assert 450 < len(reqs) < 750                                  # Poisson(600)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert np.all(gaps > 0) and abs(gaps.mean() - 0.05) < 0.01
    halved = 0
    for r in reqs:
        dims = sorted((r.settings['nx'], r.settings['ny'], r.settings['nz']))
        full = dims[-1]
        assert full in (32, 64, 96, 128, 192)
        assert dims in ([full] * 3, [full // 2, full, full])
        halved += dims[0] != full
        assert 100 <= r.steps <= 2000
        re = 1.0 / r.settings['nu']
        assert 100.0 <= re <= 1e4
        assert r.settings['nx'] % 2 == 0 and r.settings['ny'] % 2 == 0 and r.settings['fdm_rtol'] == 1e-10
    assert 0.4 < halved / len(reqs) < 0.6
    again = poisson_requests(30.0, rate=20.0, seed=4)
    assert [(r.arrival, r.steps, r.settings['nx']) for r in again] == [(r.arrival, r.steps, r.settings['nx']) for r in reqs]


def test_summary():
    st = ServeStats()
    for k in range(10):
        r = Request(k, {'nx': 8, 'ny': 8, 'nz': 8}, 5, float(k), k)
        r.admitted, r.finished = k + 0.5, k + 2.0
        st.completed.append(r)
    st.step_ms = [1.0] * 50
    st.step_cells = [512] * 50
    st.wall = 2.0
    s = st.summary()
    assert s['latency_s_p50'] == 2.0 and s['queue_wait_s_p99'] == 0.5
    assert abs(s['gpu_mcell_steps_per_s'] - 512 * 50 / 0.05 / 1e6) < 1e-9
    assert abs(s['mcell_steps_per_s'] - 512 * 50 / 2.0 / 1e6) < 1e-9


def test_summary_censors_unfinished_requests_at_the_end():
    st = ServeStats()
    for k in range(4):
        r = Request(k, {'nx': 8, 'ny': 8, 'nz': 8}, 5, float(k), k)
        r.admitted, r.finished = k + 0.5, k + 1.0
        st.completed.append(r)
    for k in range(4, 8):                              # still waiting or running when the run ends at t = 10
        st.unfinished.append(Request(k, {'nx': 8, 'ny': 8, 'nz': 8}, 5, float(k), k))
    st.unfinished.append(Request(99, {'nx': 8, 'ny': 8, 'nz': 8}, 5, 12.0, 99))   # never arrived
    st.wall = 10.0
    s = st.summary()
    assert s['completed'] == 4 and s['offered'] == 8 and s['censored'] == 4
    lat = sorted([1.0] * 4 + [10.0 - k for k in range(4, 8)])
    assert s['latency_all_offered_s_p50_lower_bound'] == np.percentile(lat, 50)
    assert s['latency_s_p50'] == 1.0
