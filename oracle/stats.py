"""Oracle: energy statistics, integrated autocorrelation time and ESS.
Test infrastructure only (see oracle/__init__.py).

The reference computes only naive statistics (GPU/VMC.py:513-534,785-789):
mean, var = <E^2> - <E>^2, error = sqrt(var/Ns).  tau_int and ESS are new:
Sokal's automatic windowing on the chain-averaged autocorrelation.
"""
import numpy as np


def naive_stats(E):
    E = np.asarray(E, dtype=np.float64)
    mean = E.mean()
    var = (E ** 2).mean() - mean ** 2
    return mean, var, np.sqrt(max(var, 0.0) / E.size)


def tau_int(series, c=5.0):
    """series: (n_chains, T).  rho(t) = mean over chains of the chain
    autocovariance at lag t over the pooled variance; tau(W) = 1 + 2 sum_{t<=W}
    rho(t); W = smallest t with t >= c tau(t)."""
    x = np.asarray(series, dtype=np.float64)
    if x.ndim == 1:
        x = x[None, :]
    n, T = x.shape
    x = x - x.mean(axis=1, keepdims=True)
    var = (x * x).mean()
    if var == 0 or T < 2:
        return 1.0
    tau = 1.0
    for t in range(1, T):
        rho = (x[:, :-t] * x[:, t:]).sum() / (n * (T - t)) / var
        tau += 2.0 * rho
        if t >= c * tau:
            break
    return max(tau, 1.0)


def ess(series, c=5.0):
    x = np.asarray(series)
    return x.size / tau_int(x, c)
