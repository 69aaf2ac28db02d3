"""Time-series statistics of HMC observables: integrated autocorrelation time with Sokal's automatic
windowing, effective sample size and tau-corrected errors.  (The reference prints V and K per
trajectory, code/modules/HMC.cpp:79,94-95, and has no autocorrelation analysis.)
"""
from __future__ import annotations

import numpy as np


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


def tau_int(x: np.ndarray, c: float = 5.0) -> tuple[float, int]:
    """Integrated autocorrelation time tau = 1/2 + sum_{t=1}^{W} rho(t), W the smallest window with
    W >= c * tau(W) (Sokal, 'Monte Carlo methods in statistical mechanics', 1997)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n < 4:
        return 0.5, 0
    rho = autocorr(x)
    taus = 0.5 + np.cumsum(rho[1:])
    for w in range(1, n - 1):
        if w >= c * taus[w - 1]:
            return float(max(taus[w - 1], 0.5)), w
    return float(max(taus[-1], 0.5)), n - 1


def ess(x: np.ndarray, c: float = 5.0) -> float:
    """Effective sample size n / (2 tau_int)."""
    t, _ = tau_int(x, c)
    return x.size / (2.0 * t)


def mean_err(x: np.ndarray, c: float = 5.0) -> tuple[float, float, float]:
    """Mean, tau-corrected standard error sqrt(2 tau var / n), and tau_int."""
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
