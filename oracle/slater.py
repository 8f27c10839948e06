"""Oracle: Slater-determinant amplitudes, local energy, log-derivatives, exact
enumeration.  Test infrastructure only (see oracle/__init__.py).

psi(x) = det(M[occ, :]) with occ the netket-ordered occupied orbitals in
ascending order (CPU semantics, model.py:7437-7457; GPU slater.py:36-61
uses argsort(desc)[:Nf], whose tie order the build fixes to ascending).
"""
import numpy as np

from .hubbard import quimb_to_netket


def occupied_rows(fxs):
    """(B, Nf) ascending netket indices of occupied orbitals."""
    net = quimb_to_netket(np.atleast_2d(fxs))
    return np.array([np.nonzero(r)[0] for r in net])


def slogdet(M, fxs):
    """(sign, ln|det|) per config; _base.py:101-109 clamps |psi| at 1e-45."""
    occ = occupied_rows(fxs)
    A = M[occ]                                   # (B, Nf, Nf)
    s, l = np.linalg.slogdet(A)
    l = np.maximum(l, np.log(1e-45))
    return s, l


def amplitude(M, fxs):
    s, l = slogdet(M, fxs)
    return s * np.exp(l)


def local_energy(H, M, fxs, logamp=None):
    """E_loc(x) = sum_eta H[eta, x] psi(eta) / psi(x) with the GPU term-major
    get_conn (GPU/vmc_utils.py:322-476, log form at :450-462)."""
    fxs = np.atleast_2d(fxs)
    B = fxs.shape[0]
    s0, l0 = slogdet(M, fxs) if logamp is None else logamp
    etas, coeffs, bids = H.get_conn_batch(fxs)
    if etas.shape[0] == 0:
        return np.zeros(B)
    s1, l1 = slogdet(M, etas)
    terms = coeffs * s1 * s0[bids] * np.exp(l1 - l0[bids])
    E = np.zeros(B)
    np.add.at(E, bids, terms)
    return E


def log_derivative(M, fxs):
    """O[s, i*Nf + j] = d ln|psi(x_s)| / d M[i, j] = (A^-1)[j, r(i)] for the
    occupied orbital i in row r(i), else 0 (closed form of the per-sample
    vmap(grad) in GPU/vmc_utils.py:1389-1634)."""
    fxs = np.atleast_2d(fxs)
    occ = occupied_rows(fxs)
    A = M[occ]
    Ainv = np.linalg.inv(A)                      # (B, Nf, Nf)
    B, n_orb, nf = fxs.shape[0], M.shape[0], M.shape[1]
    O = np.zeros((B, n_orb, nf))
    for b in range(B):
        O[b, occ[b], :] = Ainv[b].T              # row r(i) <- column r of A^-1
    return O.reshape(B, n_orb * nf)


def exact_energy_and_grad(H, M, states=None):
    """Full-Hilbert expectation and gradient (variational_state.py:232-281).

    Returns (E, grad) with grad = 2 (<E_loc O> - <E><O>), as the exact path
    has the factor 2 (variational_state.py:271); sampled paths use no 2.
    """
    states = H.all_states() if states is None else states
    s, l = slogdet(M, states)
    lmax = l.max()
    w = np.exp(2.0 * (l - lmax))
    w /= w.sum()
    E_loc = local_energy(H, M, states, logamp=(s, l))
    E = float(np.dot(w, E_loc))
    O = log_derivative(M, states)
    g = 2.0 * (np.einsum('s,s,sp->p', w, E_loc, O) - E * (w @ O))
    return E, g


def exact_energy_dense(H, M, states=None):
    """<psi|H|psi>/<psi|psi> through the dense matrix (independent of the
    local-energy code path)."""
    states = H.all_states() if states is None else states
    psi = amplitude(M, states)
    Hd = H.to_dense(states)
    return float(psi @ Hd @ psi / (psi @ psi))
