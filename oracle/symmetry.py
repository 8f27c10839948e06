"""Oracle: point-group symmetry projection of a fermionic wavefunction.
Test infrastructure only (see oracle/__init__.py).

Restates GPU/models/symmetry.py:
  * site permutations, row-major sites s = x Ly + y (symmetry.py:80-160):
      C4v (Lx == Ly = L), order [e, C4, C4^2, C4^3, s_v, s_v', s_d, s_d']:
        (x, y) -> (x, y), (L-1-y, x), (L-1-x, L-1-y), (y, L-1-x),
                  (x, L-1-y), (L-1-x, y), (y, x), (L-1-y, L-1-x)
      C2v (Lx != Ly), order [e, C2, s_v, s_v']:
        (x, y) -> (x, y), (Lx-1-x, Ly-1-y), (x, Ly-1-y), (Lx-1-x, y)
    perm[g][s] = flat index of the image; x_g = x[perm[g]];
  * characters of the one-dimensional irreps A1, A2, B1, B2
    (symmetry.py:28-52);
  * fermionic sign (symmetry.py:660-770): modes interleaved [d0, u0, d1, u1,
    ...]; the mode permutation of a site permutation p maps 2i -> 2p(i),
    2i+1 -> 2p(i)+1; sign(g, n) = (-1)^(number of pairs m1 < m2 of occupied
    modes with q(m1) > q(m2)), where q is the mode permutation of the
    INVERSE site permutation of g (symmetry.py:789-806);
  * psi^r(n) = (1/|G|) sum_g chi_r(g) sign(g, n) psi(n[perm[g]])
    (symmetry.py:893-951), in log form with a max-shifted sum and the 1e-45
    clamp.
"""
import numpy as np

C4V_CHARACTERS = {"A1": [1, 1, 1, 1, 1, 1, 1, 1], "A2": [1, 1, 1, 1, -1, -1, -1, -1],
                  "B1": [1, -1, 1, -1, 1, 1, -1, -1], "B2": [1, -1, 1, -1, -1, -1, 1, 1]}
C2V_CHARACTERS = {"A1": [1, 1, 1, 1], "A2": [1, 1, -1, -1], "B1": [1, -1, 1, -1],
                  "B2": [1, -1, -1, 1]}


def permutations(Lx, Ly):
    s = np.arange(Lx * Ly)
    x, y = s // Ly, s % Ly
    if Lx == Ly:
        L = Lx
        imgs = [(x, y), (L - 1 - y, x), (L - 1 - x, L - 1 - y), (y, L - 1 - x),
                (x, L - 1 - y), (L - 1 - x, y), (y, x), (L - 1 - y, L - 1 - x)]
    else:
        imgs = [(x, y), (Lx - 1 - x, Ly - 1 - y), (x, Ly - 1 - y), (Lx - 1 - x, y)]
    return np.array([a * Ly + b for a, b in imgs], dtype=np.int64)


def characters(Lx, Ly, irrep):
    table = C4V_CHARACTERS if Lx == Ly else C2V_CHARACTERS
    return np.array(table[irrep], dtype=np.float64)


def inverse(perm):
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    return inv


def mode_permutation(site_perm):
    N = site_perm.size
    q = np.empty(2 * N, dtype=np.int64)
    q[0::2] = 2 * site_perm
    q[1::2] = 2 * site_perm + 1
    return q


def fermion_sign(perm, x):
    """sign(g, x) for one site permutation perm (= perm[g]) and one
    configuration x (quimb codes): inversions of q = mode permutation of
    perm^-1 among the occupied interleaved modes, counted pair by pair."""
    q = mode_permutation(inverse(np.asarray(perm)))
    n = np.stack([(np.asarray(x) & 1), (np.asarray(x) >> 1) & 1], axis=-1).reshape(-1)
    occ = np.nonzero(n)[0]
    inv = 0
    for a in range(len(occ)):
        for b in range(a + 1, len(occ)):
            inv += int(q[occ[a]] > q[occ[b]])
    return 1.0 if inv % 2 == 0 else -1.0


def project_log(logamp, Lx, Ly, irrep, xs):
    """(sign, ln|psi^r|) of configurations xs (B, N) from logamp(configs) ->
    (sign, ln|psi|) of the base wavefunction."""
    P = permutations(Lx, Ly)
    chi = characters(Lx, Ly, irrep)
    xs = np.atleast_2d(np.asarray(xs))
    G = P.shape[0]
    out_s, out_l = np.zeros(len(xs)), np.zeros(len(xs))
    for b, x in enumerate(xs):
        xg = x[P]                                  # (G, N)
        s, l = logamp(xg)
        f = np.array([fermion_sign(P[g], x) for g in range(G)])
        m = l.max()
        w = (chi * f * s * np.exp(l - m)).sum()
        out_s[b] = np.sign(w)
        out_l[b] = m + np.log(max(abs(w), 1e-45)) - np.log(G)
    return out_s, out_l


def operator_matrix(H, perm):
    """Dense matrix of (U psi)(n) = sign(g, n) psi(n[perm]) on H's basis:
    U[k, k'] = sign(g, n_k) when n_{k'} = n_k[perm]."""
    states = H.all_states()
    index = {tuple(int(v) for v in s): k for k, s in enumerate(states)}
    U = np.zeros((len(states), len(states)))
    for k, s in enumerate(states):
        U[k, index[tuple(int(v) for v in s[perm])]] = fermion_sign(perm, s)
    return U
