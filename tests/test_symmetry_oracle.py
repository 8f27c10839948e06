"""The symmetry-projection oracle pinned on the CPU.  The fermionic
operators (U_g psi)(n) = sign(g, n) psi(n[perm_g]) built from the oracle's
permutations and signs must (i) form a representation of the point group
(U_a U_b = U_ab with perm_ab = perm_a[perm_b]), (ii) commute with the dense
symmray-order Hubbard Hamiltonian, and (iii) map every non-degenerate ground
state into exactly one irrep.  Two planted sign mistakes (the permutation of
g instead of g^-1; the up-block-then-down-block mode order instead of the
interleaved one) each break (i) or (ii).  The group tables follow the
reference's own checks (symmetry.py:163-201)."""
import numpy as np
import pytest

from oracle import hubbard as ohub
from oracle import symmetry as osy

CASES = [(2, 2, 1, 1), (3, 3, 2, 2), (4, 2, 2, 2), (2, 3, 2, 1)]


def _ops(H, sign_fn):
    states = H.all_states()
    index = {tuple(int(v) for v in s): k for k, s in enumerate(states)}
    P = osy.permutations(H.Lx, H.Ly)
    U = []
    for p in P:
        u = np.zeros((len(states), len(states)))
        for k, s in enumerate(states):
            u[k, index[tuple(int(v) for v in s[p])]] = sign_fn(p, s)
        U.append(u)
    return P, U


def _checks(Lx, Ly, nu, nd, sign_fn):
    H = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    Hm = H.to_dense()
    P, U = _ops(H, sign_fn)
    comm = max(np.abs(Hm @ u - u @ Hm).max() for u in U)
    hom = 0.0
    for a in range(len(P)):
        for b in range(len(P)):
            c = [k for k in range(len(P)) if np.array_equal(P[k], P[a][P[b]])]
            assert len(c) == 1                          # closure
            hom = max(hom, np.abs(U[a] @ U[b] - U[c[0]]).max())
    return Hm, U, comm, hom


def test_group_tables():
    for L in (2, 3, 4, 8):
        P = osy.permutations(L, L)
        assert P.shape == (8, L * L)
        for p in P:
            assert np.array_equal(np.sort(p), np.arange(L * L))
        assert np.array_equal(P[1][P[1]], P[2])               # C4 C4 = C4^2
        assert np.array_equal(P[3][P[1]], P[0])               # C4^3 C4 = e
        assert np.array_equal(P[6][P[6]], P[0])               # sigma_d^2 = e
    P = osy.permutations(4, 2)
    assert P.shape == (4, 8)
    for a in range(4):
        assert np.array_equal(P[a][P[a]], P[0])               # C2v: every element an involution
    for table in (osy.C4V_CHARACTERS, osy.C2V_CHARACTERS):
        X = np.array(list(table.values()), dtype=float)
        assert np.allclose(X @ X.T, X.shape[1] * np.eye(4))   # orthogonal characters


@pytest.mark.parametrize("case", CASES)
def test_operators_are_a_symmetry_representation(case):
    Hm, U, comm, hom = _checks(*case, osy.fermion_sign)
    assert comm == 0.0 and hom == 0.0
    w, v = np.linalg.eigh(Hm)
    if w[1] - w[0] > 1e-8:                                     # non-degenerate ground state
        Lx, Ly = case[:2]
        table = osy.C4V_CHARACTERS if Lx == Ly else osy.C2V_CHARACTERS
        norms = sorted(np.linalg.norm(sum(c * u for c, u in zip(table[r], U)) @ v[:, 0])
                       / len(U) for r in table)
        assert np.allclose(norms, [0, 0, 0, 1], atol=1e-10)


def _sign_of_g(perm, x):
    q = osy.mode_permutation(np.asarray(perm))             # planted: g, not g^-1
    n = np.stack([np.asarray(x) & 1, (np.asarray(x) >> 1) & 1], axis=-1).reshape(-1)
    occ = np.nonzero(n)[0]
    inv = sum(int(q[occ[a]] > q[occ[b]]) for a in range(len(occ)) for b in range(a + 1, len(occ)))
    return -1.0 if inv % 2 else 1.0


def _sign_block_order(perm, x):
    N = len(x)
    pinv = osy.inverse(np.asarray(perm))
    q = np.concatenate([pinv, pinv + N])                   # planted: up block then down block
    n = np.concatenate([(np.asarray(x) >> 1) & 1, np.asarray(x) & 1])
    occ = np.nonzero(n)[0]
    inv = sum(int(q[occ[a]] > q[occ[b]]) for a in range(len(occ)) for b in range(a + 1, len(occ)))
    return -1.0 if inv % 2 else 1.0


@pytest.mark.parametrize("fn", [_sign_of_g, _sign_block_order])
def test_planted_sign_mistakes_are_caught(fn):
    _, _, comm, hom = _checks(3, 3, 2, 2, fn)
    assert comm > 0.5 or hom > 0.5


def test_projection_of_an_eigenstate():
    """project_log applied to the exact ground state (as a table lookup)
    returns it on its irrep and (numerically) zero on the others."""
    H = ohub.Hubbard(4, 2, 1.0, 8.0, 2, 2)
    w, v = np.linalg.eigh(H.to_dense())
    psi = v[:, 0]
    index = {tuple(int(a) for a in s): k for k, s in enumerate(H.all_states())}

    def logamp(xs):
        a = np.array([psi[index[tuple(int(t) for t in x)]] for x in xs])
        with np.errstate(divide="ignore"):
            return np.sign(a), np.log(np.maximum(np.abs(a), 1e-300))
    xs = H.all_states()[::37]
    s, l = osy.project_log(logamp, 4, 2, "A1", xs)
    ref_s, ref_l = logamp(xs)
    big = ref_l > np.log(1e-8)
    assert np.array_equal(s[big], ref_s[big]) and np.allclose(l[big], ref_l[big], atol=1e-10)
    for r in ("A2", "B1", "B2"):
        _, l2 = osy.project_log(logamp, 4, 2, r, xs)
        assert np.all(l2 < ref_l.max() - 20)
