"""The statistics bench.py reports ZPEs and times to accuracy with (CPU, synthetic series).

* `extrapolate_dt0`: the quadratic through three time steps returns the intercept of any
  quadratic exactly, with the Lagrange weights (1/3, -2, 8/3) at dt = 5, 2.5, 1.25, and the
  propagated error sqrt(sum (w_i se_i)^2); the linear check through the two smallest steps
  returns the intercept of a straight line.
* `blocking_se`: the largest estimate over block sizes, so conservative: on white noise between
  sigma / sqrt(N) and 1.5 times it (the estimates from 16 - 32 blocks scatter by about 20 %); on
  an AR(1) series it recovers the inflation sqrt((1 + rho) / (1 - rho)) that the naive standard
  error misses, within the same bounds.
* `zpe_of`: the reference's estimator, the mean of vref over steps T/4 .. T, in cm^-1.
* `time_to_accuracy`: the first checkpoint from which the running estimate stays inside the
  band to the end, not the first one that touches it; None for a run that ends outside.
"""
import numpy as np
import pytest

import bench

WN = bench.WN


def test_extrapolation_weights_and_intercept():
    dts = (5.0, 2.5, 1.25)
    a, b, c = 4677.6, -0.3, 0.05
    zs = [a + b * d + c * d * d for d in dts]
    ses = [0.17, 0.17, 0.16]
    r = bench.extrapolate_dt0(dts, zs, ses)
    np.testing.assert_allclose(r['lagrange_weights'], [1 / 3, -2.0, 8 / 3], rtol=1e-12)
    assert r['zpe0_cm'] == pytest.approx(a, abs=1e-9)
    assert r['se_cm'] == pytest.approx(np.sqrt((0.17 / 3) ** 2 + (2 * 0.17) ** 2 + (8 / 3 * 0.16) ** 2), rel=1e-12)
    # a straight line: both the quadratic and the two-smallest-step line return its intercept
    zl = [a + b * d for d in dts]
    r2 = bench.extrapolate_dt0(dts, zl, ses)
    assert r2['zpe0_cm'] == pytest.approx(a, abs=1e-9)
    assert r2['linear_two_smallest_cm'] == pytest.approx(a, abs=1e-9)
    # a planted mistake: a straight line through all three points misses a quadratic's intercept
    slope, icpt = np.polyfit(dts, zs, 1)
    assert abs(icpt - a) > 0.05


def test_blocking_error_on_white_and_correlated_noise():
    rs = np.random.RandomState(3)
    n = 1 << 16
    x = rs.normal(0.0, 1.0, n)
    assert 0.9 / np.sqrt(n) < bench.blocking_se(x) < 1.5 / np.sqrt(n)
    rho = 0.9
    y = np.empty(n)
    y[0] = 0.0
    e = rs.normal(0.0, np.sqrt(1 - rho * rho), n)
    for i in range(1, n):
        y[i] = rho * y[i - 1] + e[i]
    true_se = np.sqrt((1 + rho) / (1 - rho)) / np.sqrt(n)
    naive = y.std(ddof=1) / np.sqrt(n)
    assert naive < 0.35 * true_se
    assert 0.8 * true_se < bench.blocking_se(y) < 1.5 * true_se


def test_zpe_estimator_uses_the_last_three_quarters():
    v = np.concatenate([np.full(250, 100.0), np.full(750, 10.0)]) * WN
    z, se = bench.zpe_of(v)
    assert z == pytest.approx(10.0, abs=1e-9) and se == pytest.approx(0.0, abs=1e-9)
    z2, _ = bench.zpe_of(v, 400)  # steps 100..400: 150 at 100 and 150 at 10
    assert z2 == pytest.approx(55.0, abs=1e-9)


def test_time_to_accuracy_requires_staying_in_the_band():
    target, every = 1000.0, 250
    rs = np.random.RandomState(5)
    noise = rs.normal(0.0, 0.5, 20000)
    # settles: far off for the first 2 000 steps, on target after
    v = np.where(np.arange(20000) < 2000, target + 40.0, target) + noise
    r = bench.time_to_accuracy(v * WN, target, ms_per_step=0.01, tol=1.0, sigma=0.3, every=every)
    assert r is not None and r['steps'] % every == 0
    assert abs(r['zpe_cm'] - target) <= 1.0 and r['se_cm'] <= 0.3
    assert r['seconds'] == pytest.approx(r['steps'] * 0.01e-3)
    # every later checkpoint also lies inside the band
    for t in range(r['steps'], 20001, every):
        assert abs(bench.zpe_of(v * WN, t)[0] - target) <= 1.0
    # a run that passes through the band and ends 3 cm^-1 off never settles
    w = target - 20.0 + np.linspace(0.0, 23.0, 20000) + noise
    assert bench.time_to_accuracy(w * WN, target, 0.01, 1.0, 0.3, every) is None
    # a run that touches the band early, leaves it, and comes back (the excursion falls out of
    # the estimator's window, steps T/4 .. T, later): the later entry counts
    u = np.full(20000, target) + noise
    u[1000:3000] += 12.0
    ru = bench.time_to_accuracy(u * WN, target, 0.01, 1.0, 0.3, every)
    first_touch = next(t for t in range(every, 20001, every) if abs(bench.zpe_of(u * WN, t)[0] - target) <= 1.0)
    assert ru is not None and ru['steps'] > first_touch
