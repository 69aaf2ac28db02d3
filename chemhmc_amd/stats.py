"""Time-series statistics of HMC observables: integrated autocorrelation time with Sokal's automatic
windowing, effective sample size and tau-corrected errors.  (The reference prints V and K per
trajectory, code/modules/HMC.cpp:79,94-95, and has no autocorrelation analysis.)

A series that carries no information about its autocorrelation -- one that never changed (an HMC chain
that accepted nothing), or whose spread is at the rounding level of its mean -- has no tau_int: the
estimators return NaN for it instead of the n / 2 that the windowing would produce from the rounding
residue of the mean.
"""
from __future__ import annotations

import numpy as np

# spread below this fraction of |mean| is rounding noise, not fluctuation
_REL_NOISE = 1e-13
# Sokal's guidance: tau_int is trustworthy when the window closed and the series is many tau long
MIN_TAUS = 50.0


def informative(x: np.ndarray) -> bool:
    """True if the series fluctuates above the rounding level of its mean."""
    x = np.asarray(x, dtype=np.float64)
    if x.size < 4 or not np.all(np.isfinite(x)):
        return False
    spread = float(x.max() - x.min())
    return spread > _REL_NOISE * max(float(np.abs(x).max()), np.finfo(np.float64).tiny)


def autocorr(x: np.ndarray) -> np.ndarray:
    """Normalised autocorrelation rho(t), t = 0..n-1 (FFT, zero-padded, biased estimator)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    d = x - x.mean()
    m = 1 << (2 * n - 1).bit_length()
    f = np.fft.rfft(d, m)
    c = np.fft.irfft(f * np.conj(f), m)[:n]
    if c[0] <= 0:
        return np.concatenate([[1.0], np.zeros(n - 1)])
    return c / c[0]


def tau_window(x: np.ndarray, c: float = 5.0) -> tuple[float, int, bool]:
    """(tau_int, window W, closed): tau = 1/2 + sum_{t=1}^{W} rho(t), W the smallest window with
    W >= c * tau(W) (Sokal, 'Monte Carlo methods in statistical mechanics', 1997); closed is False when
    no such window exists within the series.  tau is NaN for a series without information."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if not informative(x):
        return float("nan"), 0, False
    rho = autocorr(x)
    taus = 0.5 + np.cumsum(rho[1:])
    for w in range(1, n - 1):
        if w >= c * taus[w - 1]:
            return float(max(taus[w - 1], 0.5)), w, True
    return float(max(taus[-1], 0.5)), n - 1, False


def tau_int(x: np.ndarray, c: float = 5.0) -> tuple[float, int]:
    """Integrated autocorrelation time and its window (see tau_window); NaN without information."""
    t, w, _ = tau_window(x, c)
    return t, w


def reliable(x: np.ndarray, c: float = 5.0) -> bool:
    """tau_int is usable: the series is informative, the window closed and n >= MIN_TAUS * tau."""
    t, _, closed = tau_window(x, c)
    return bool(np.isfinite(t) and closed and np.asarray(x).size >= MIN_TAUS * t)


def ess(x: np.ndarray, c: float = 5.0) -> float:
    """Effective sample size n / (2 tau_int); NaN without information."""
    t, _ = tau_int(x, c)
    return x.size / (2.0 * t)


def mean_err(x: np.ndarray, c: float = 5.0) -> tuple[float, float, float]:
    """Mean, tau-corrected standard error sqrt(2 tau var / n), and tau_int (NaN error and tau for a
    series without information)."""
    x = np.asarray(x, dtype=np.float64)
    t, _ = tau_int(x, c)
    var = x.var(ddof=1) if x.size > 1 else 0.0
    return float(x.mean()), float(np.sqrt(2.0 * t * var / max(x.size, 1))), t


def zscore(a: tuple[float, float], b: tuple[float, float]) -> float:
    """|mean_a - mean_b| / sqrt(err_a^2 + err_b^2)."""
    den = np.hypot(a[1], b[1])
    return float(abs(a[0] - b[0]) / den) if den > 0 else (0.0 if a[0] == b[0] else np.inf)


def pressure(W: np.ndarray, N: int, volume: float, beta: float) -> np.ndarray:
    """Virial pressure P = rho T + W / (3 V), T = 1/beta (reduced units)."""
    return N / (volume * beta) + np.asarray(W) / (3.0 * volume)
