"""Oracle: stochastic reconfiguration and the SGD update.
Test infrastructure only (see oracle/__init__.py).
"""
import math

import numpy as np


def energy_gradient(O, E, energy_mean):
    """g = <E O> - <E><O>, no factor 2 (GPU/optimizer.py:414-434)."""
    Ns = O.shape[0]
    return (E @ O) / Ns - energy_mean * O.mean(axis=0)


def minsr(O, E, energy_mean, diag_shift):
    """Sample-space SR (GPU/optimizer.py:546-563):
    T = Ot Ot^T + shift I, Ot = (O - mean O)/sqrt(Ns), eps = (E - Ebar)/sqrt(Ns),
    dp = Ot^T T^-1 eps."""
    Ns = O.shape[0]
    Ot = (O - O.mean(axis=0)) / math.sqrt(Ns)
    eps = (E - energy_mean) / math.sqrt(Ns)
    T = Ot @ Ot.T + diag_shift * np.eye(Ns)
    return Ot.T @ np.linalg.solve(T, eps)


def sr_dense(O, E, energy_mean, diag_shift):
    """Parameter-space SR (optimizer.py:158-187 with a dense solve):
    (S + shift I) dp = g, S = O^T O / Ns - Obar Obar^T."""
    Ns = O.shape[0]
    Ob = O.mean(axis=0)
    S = O.T @ O / Ns - np.outer(Ob, Ob)
    g = energy_gradient(O, E, energy_mean)
    return np.linalg.solve(S + diag_shift * np.eye(O.shape[1]), g)


def minres(matvec, b, rtol=1e-5, maxiter=100):
    """Paige-Saunders MINRES as restated in GPU/torch_utils.py:747-857
    (convergence ||r|| < rtol ||A|| ||x||)."""
    b = np.asarray(b, dtype=np.float64)
    bnorm = float(np.linalg.norm(b))
    if bnorm == 0:
        return np.zeros_like(b), 0
    x = np.zeros_like(b)
    r1, r2 = b.copy(), b.copy()
    beta = bnorm
    cs, sn, oldb, dbar, epsln, phibar = -1.0, 0.0, 0.0, 0.0, 0.0, bnorm
    Anorm2 = 0.0
    w = np.zeros_like(b)
    w2 = np.zeros_like(b)
    info = maxiter
    for itn in range(1, maxiter + 1):
        v = r2 / beta
        y = matvec(v)
        if itn >= 2:
            y = y - (beta / oldb) * r1
        alfa = float(v @ y)
        y = y - (alfa / beta) * r2
        r1, r2 = r2, y
        oldb, beta = beta, float(np.linalg.norm(r2))
        oldeps = epsln
        delta = cs * dbar + sn * alfa
        gbar = sn * dbar - cs * alfa
        epsln = sn * beta
        dbar = -cs * beta
        gamma = max(math.sqrt(gbar ** 2 + beta ** 2), 1e-300)
        cs, sn = gbar / gamma, beta / gamma
        phi = cs * phibar
        phibar = sn * phibar
        w1, w2 = w2, w
        w = (v - oldeps * w1 - delta * w2) / gamma
        x = x + phi * w
        Anorm2 += alfa ** 2 + oldb ** 2 + beta ** 2
        if abs(phibar) < rtol * math.sqrt(Anorm2) * float(np.linalg.norm(x)):
            info = 0
            break
        if beta == 0.0:
            info = 0
            break
    return x, info


def sr_minres(O, E, energy_mean, diag_shift, rtol=1e-4, maxiter=100):
    """GPU/optimizer.py:414-462: MINRES on (S + shift I) with the matvec
    O^T(O x)/Ns - (Obar.x) Obar + shift x."""
    Ns = O.shape[0]
    Ob = O.mean(axis=0)
    g = energy_gradient(O, E, energy_mean)

    def mv(x):
        return O.T @ (O @ x) / Ns - (Ob @ x) * Ob + diag_shift * x
    return minres(mv, g, rtol=rtol, maxiter=maxiter)


def sgd(params, dp, lr):
    """p <- p - lr dp (GPU/optimizer.py:178-186; optimizer.py:339-340)."""
    return params - lr * dp


def outlier_mask(E, O, clip_factor):
    """GPU/VMC.py:476-498: drop the samples whose |O_s| exceeds clip_factor
    times the median norm; torch's median of an even count is the lower of
    the two middle values.  Dropped rows of O become 0, their energies the
    mean of the kept ones.  Returns (E, O, keep mask)."""
    E = np.array(E, dtype=np.float64)
    O = np.array(O, dtype=np.float64)
    norms = np.sqrt((O * O).sum(axis=1))
    med = np.sort(norms)[(norms.size - 1) // 2]
    keep = norms <= clip_factor * med
    if not keep.all():
        E[~keep] = E[keep].mean()
        O[~keep] = 0.0
    return E, O, keep
