"""Energy statistics on the device: naive mean / variance / error (as the
reference, GPU/VMC.py:513-534,785-789) plus the integrated autocorrelation
time and effective sample size, which the reference does not compute.

tau_int = 1 + 2 sum_{t=1}^{W} rho(t), rho averaged over chains, with Sokal's
automatic window W = min{t : t >= c tau(t)} (c = 5)."""
import torch


def naive_stats(E):
    E = E.to(torch.float64)
    m = E.mean()
    var = (E * E).mean() - m * m
    return float(m), float(var), float(torch.sqrt(torch.clamp(var, min=0.0) / E.numel()))


def tau_int(series, c=5.0):
    """series: (n_chains, T) tensor of per-chain observables (e.g. E_loc
    after each sweep)."""
    x = series.to(torch.float64)
    if x.dim() == 1:
        x = x.unsqueeze(0)
    n, T = x.shape
    x = x - x.mean(dim=1, keepdim=True)
    var = (x * x).mean()
    if T < 2 or float(var) == 0.0:
        return 1.0
    # all lags at once: rho(t) for t < T
    rho = torch.stack([(x[:, :-t] * x[:, t:]).mean() for t in range(1, T)]) / var
    rho = rho.cpu()
    tau = 1.0
    for t in range(1, T):
        tau += 2.0 * float(rho[t - 1])
        if t >= c * tau:
            break
    return max(tau, 1.0)


def ess(series, c=5.0):
    return series.numel() / tau_int(series, c)
