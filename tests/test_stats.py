"""ESS / tau_int estimator pinned on AR(1) series, whose tau_int = (1 + phi) / (2 (1 - phi))."""
import numpy as np
import pytest

from chemhmc_amd import stats


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.8, 0.95])
def test_tau_int_ar1(phi):
    rng = np.random.default_rng(int(phi * 100) + 1)
    n = 200000
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi ** 2)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    tau, _ = stats.tau_int(x, c=6.0)
    want = (1 + phi) / (2 * (1 - phi))
    assert tau == pytest.approx(want, rel=0.06)
    assert stats.ess(x, c=6.0) == pytest.approx(n / (2 * tau))


def test_mean_err_and_pressure():
    rng = np.random.default_rng(0)
    x = rng.normal(3.0, 2.0, size=40000)
    m, err, tau = stats.mean_err(x)
    assert abs(m - 3.0) < 4 * err and err == pytest.approx(2.0 / np.sqrt(40000), rel=0.1)
    assert stats.zscore((1.0, 0.3), (1.0, 0.4)) == 0.0
    assert stats.pressure(np.array([0.0]), 100, 125.0, 0.5)[0] == pytest.approx(100 / 125.0 / 0.5)


def test_no_information_series():
    """A chain that never moved (constant U/N, as with zero acceptance) has no tau_int: NaN, not n / 2."""
    x = np.full(20, -5.3600555704884645)
    tau, w = stats.tau_int(x)
    assert np.isnan(tau) and w == 0
    assert np.isnan(stats.ess(x)) and not stats.reliable(x)
    # rounding-level spread around a large mean is no information either
    y = 1e6 + np.array([0.0, 1e-10, 0.0, -1e-10] * 5)
    assert np.isnan(stats.tau_int(y)[0])
    # a chain with a handful of accepted moves is informative but not reliable
    z = np.repeat([1.0, 2.0, 1.5, 3.0], 5)
    assert np.isfinite(stats.tau_int(z)[0]) and not stats.reliable(z)
    m, err, t = stats.mean_err(x)
    assert m == pytest.approx(x[0], rel=1e-15) and np.isnan(err) and np.isnan(t)
