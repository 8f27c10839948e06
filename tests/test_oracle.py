"""CPU tests that pin the FP64 oracle before it is trusted as the checker.

Each test fixes a part of the oracle against something independent of its
own formula: the reference's shipped run data, closed forms, textbook
identities, scipy, or brute force on tiny systems.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import hubbard, philox, sampler, slater, sr, stats

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_4x2_slater.json")


@pytest.fixture(scope="module")
def ref4x2():
    return json.load(open(GOLD))


@pytest.fixture(scope="module")
def H4x2():
    return hubbard.Hubbard(4, 2, 1.0, 8.0, 4, 4)


# ------------------------------------------------------------------ lattice
def test_edge_counts_and_order():
    for (Lx, Ly), ne in {(4, 2): 10, (4, 4): 24, (6, 6): 60, (8, 8): 112}.items():
        g = hubbard.SquareLattice(Lx, Ly)
        assert len(g.edges()) == ne
        assert sorted(g.sweep_edges()) == sorted(g.edges())
    g = hubbard.SquareLattice(4, 2)
    assert g.edges()[:3] == [(0, 1), (0, 2), (1, 3)]
    # sweep = horizontal edges row by row, then vertical edges column by column
    assert g.sweep_edges() == [(0, 1), (2, 3), (4, 5), (6, 7),
                               (0, 2), (2, 4), (4, 6), (1, 3), (3, 5), (5, 7)]


def test_encodings_roundtrip():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 4, size=(50, 9))
    assert np.array_equal(hubbard.netket_to_quimb(hubbard.quimb_to_netket(x)), x)
    s = hubbard.quimb_to_symmray(x)
    assert np.array_equal(s[:, 1::2], (x >> 1) & 1) and np.array_equal(s[:, 0::2], x & 1)


# -------------------------------------------------------------- Hamiltonian
def test_two_site_closed_form():
    """1x2 Hubbard, 1 up + 1 down: E0 = (U - sqrt(U^2 + 16 t^2)) / 2."""
    for t, U in [(1.0, 8.0), (0.7, 3.0)]:
        H = hubbard.Hubbard(1, 2, t, U, 1, 1)
        e0 = np.linalg.eigvalsh(H.to_dense())[0]
        assert e0 == pytest.approx((U - math.sqrt(U * U + 16 * t * t)) / 2, abs=1e-12)


@pytest.mark.parametrize("shape", [(2, 2, 2, 2), (3, 2, 2, 1)])
def test_dense_h_symmetric_and_basis_change(shape):
    """The symmray-phase H is symmetric and has the spectrum of the
    netket-phase H (the two differ by a diagonal sign change of basis)."""
    H = hubbard.Hubbard(shape[0], shape[1], 1.0, 8.0, shape[2], shape[3])
    st = H.all_states()
    Hs = H.to_dense(st)
    Hn = H.to_dense(st, conn=H.get_conn_netket)
    assert np.abs(Hs - Hs.T).max() == 0.0
    assert np.allclose(np.linalg.eigvalsh(Hs), np.linalg.eigvalsh(Hn), atol=1e-10)
    D = np.diag(np.sign(np.diag(Hs @ Hn.T) + 0.5))  # any diagonal +-1 works
    assert D.shape == Hs.shape


def test_get_conn_batch_matches_single(H4x2):
    rng = np.random.default_rng(1)
    st = H4x2.all_states()
    x = st[rng.choice(len(st), 40, replace=False)]
    etas, coeffs, bids = H4x2.get_conn_batch(x)
    for b in range(len(x)):
        e1, c1 = H4x2.get_conn(x[b])
        sel = bids == b
        d = {}
        for e, c in zip(etas[sel], coeffs[sel]):
            d[tuple(e)] = d.get(tuple(e), 0.0) + c
        assert d == {tuple(e): c for e, c in zip(e1, c1)}


def test_hilbert_size_and_order(H4x2):
    st = H4x2.all_states()
    assert len(st) == 4900 == H4x2.hilbert_size
    # first state: ups on sites 0..3, downs on sites 0..3 -> code 3 there
    assert list(st[0]) == [3, 3, 3, 3, 0, 0, 0, 0]
    assert list(st[1]) == [3, 3, 3, 2, 1, 0, 0, 0]


# ---------------------------------------------------- reference fixtures
@pytest.mark.parametrize("step,k", [(1, 1), (26, 26)])
def test_checkpoint_energy_matches_logged_run(ref4x2, H4x2, step, k):
    """Exact <H> of the reference's checkpoint M (rows 0..15) equals the
    reference's own VMC estimate for that step within 4 sigma.  This pins
    the symmray-H x netket-Slater sign pairing end to end."""
    M = np.array(ref4x2["checkpoints"][str(step)]["M16"])
    E, _ = slater.exact_energy_and_grad(H4x2, M)
    z = (E - ref4x2["mean"][k]) / ref4x2["error"][k]
    assert abs(z) < 4.0, (E, ref4x2["mean"][k], z)


def test_wrong_phase_convention_is_rejected(ref4x2, H4x2):
    """The netket-phase Hamiltonian gives an energy many sigma away from the
    logged run: the fixture discriminates the sign convention."""
    M = np.array(ref4x2["checkpoints"]["26"]["M16"])
    st = H4x2.all_states()
    psi = slater.amplitude(M, st)
    Hn = H4x2.to_dense(st, conn=H4x2.get_conn_netket)
    En = psi @ Hn @ psi / (psi @ psi)
    assert abs(En - ref4x2["mean"][26]) / ref4x2["error"][26] > 10


def test_inert_rows_constant(ref4x2):
    norms = [ref4x2["checkpoints"][s]["inert_rows_norm"] for s in ("1", "26", "50")]
    assert max(norms) - min(norms) < 1e-9


def test_exact_energy_two_routes(H4x2):
    rng = np.random.default_rng(2)
    H = hubbard.Hubbard(2, 2, 1.0, 8.0, 2, 2)
    M = rng.normal(0, 0.1, (8, 4))
    E1, _ = slater.exact_energy_and_grad(H, M)
    assert E1 == pytest.approx(slater.exact_energy_dense(H, M), rel=1e-12)


# ---------------------------------------------------------------- Slater
def test_log_derivative_finite_difference():
    rng = np.random.default_rng(3)
    H = hubbard.Hubbard(2, 2, 1.0, 8.0, 2, 2)
    M = rng.normal(0, 1.0, (8, 4))
    x = H.all_states()[[0, 7, 20, 35]]
    O = slater.log_derivative(M, x)
    h = 1e-6
    for p in range(M.size):
        Mp, Mm = M.copy(), M.copy()
        Mp.flat[p] += h
        Mm.flat[p] -= h
        fd = (slater.slogdet(Mp, x)[1] - slater.slogdet(Mm, x)[1]) / (2 * h)
        assert np.allclose(O[:, p], fd, atol=1e-6)


def test_exact_gradient_finite_difference():
    rng = np.random.default_rng(4)
    H = hubbard.Hubbard(2, 2, 1.0, 4.0, 2, 2)
    M = rng.normal(0, 1.0, (8, 4))
    _, g = slater.exact_energy_and_grad(H, M)
    h = 1e-6
    for p in range(0, M.size, 3):
        Mp, Mm = M.copy(), M.copy()
        Mp.flat[p] += h
        Mm.flat[p] -= h
        fd = (slater.exact_energy_dense(H, Mp) - slater.exact_energy_dense(H, Mm)) / (2 * h)
        assert g[p] == pytest.approx(fd, abs=1e-5)


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32_10."""
    kat = [
        ((0, 0, 0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 6, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for args, want in kat:
        assert tuple(int(v) for v in philox.philox4x32_10(*args)) == want


def test_philox_draws_uniform():
    u, b0, b2, ua = philox.proposal_draws(123, 7, 3, np.arange(200000))
    for v in (u, ua):
        assert 0 <= v.min() and v.max() < 1
        assert abs(v.mean() - 0.5) < 0.005
    assert abs(b0.mean() - 0.5) < 0.005 and abs(b2.mean() - 0.5) < 0.005
    assert abs(np.corrcoef(u, ua)[0, 1]) < 0.01


# ---------------------------------------------------------------- sampler
def test_proposal_table():
    """Every (s_i, s_j) pair and branch of GPU/sampler.py:47-137."""
    cases = []
    for si in range(4):
        for sj in range(4):
            for u in (0.1, 0.9):           # exchange / hop
                for b in (0, 1):
                    cases.append((si, sj, u, b))
    nmap = [0, 1, 1, 2]
    for si, sj, u, b in cases:
        ni, nj, ch = sampler.propose(np.array([si]), np.array([sj]), np.array([u]),
                                     np.array([b]), np.array([b]))
        if si == sj:
            assert not ch[0] and (ni[0], nj[0]) == (si, sj)
            continue
        assert ch[0]
        dn = abs(nmap[si] - nmap[sj])
        if u < 0.75 or dn == 1:
            want = (sj, si)
        elif dn == 0:
            want = (3, 0) if b else (0, 3)
        else:
            want = (2, 1) if b else (1, 2)
        assert (ni[0], nj[0]) == want
        # particle numbers per spin are conserved
        assert ((si >> 1) + (sj >> 1)) == ((ni[0] >> 1) + (nj[0] >> 1))
        assert ((si & 1) + (sj & 1)) == ((ni[0] & 1) + (nj[0] & 1))


def test_sampler_reproduces_exact_distribution():
    """Long Metropolis chains on 2x2 (36 states) sample |psi|^2."""
    rng = np.random.default_rng(5)
    H = hubbard.Hubbard(2, 2, 1.0, 8.0, 2, 2)
    M = rng.normal(0, 1.0, (8, 4))
    states = H.all_states()
    p = slater.amplitude(M, states) ** 2
    p /= p.sum()
    idx = {tuple(s): k for k, s in enumerate(states)}
    B = 256
    x = states[rng.choice(len(states), B, p=p)]
    counts = np.zeros(len(states))
    edges = H.graph.sweep_edges()
    for sw in range(60):
        x, _, _, _ = sampler.sweep(lambda c: slater.slogdet(M, c), x, edges,
                                   sampler.PhiloxDraws(9, sw), 1)
        if sw >= 10:
            for s in x:
                counts[idx[tuple(s)]] += 1
    emp = counts / counts.sum()
    tv = 0.5 * np.abs(emp - p).sum()
    assert tv < 0.05, tv


# --------------------------------------------------------------------- SR
def test_minsr_equals_param_space_sr():
    rng = np.random.default_rng(6)
    Ns, Np = 60, 25
    O = rng.normal(size=(Ns, Np))
    E = rng.normal(size=Ns)
    for shift in (1e-4, 1e-2):
        a = sr.minsr(O, E, E.mean(), shift)
        b = sr.sr_dense(O, E, E.mean(), shift)
        assert np.allclose(a, b, rtol=1e-7, atol=1e-9)


def test_minres_matches_scipy_and_dense():
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(7)
    Ns, Np = 80, 40
    O = rng.normal(size=(Ns, Np))
    E = rng.normal(size=Ns)
    x, info = sr.sr_minres(O, E, E.mean(), 1e-3, rtol=1e-10, maxiter=500)
    assert info == 0
    assert np.allclose(x, sr.sr_dense(O, E, E.mean(), 1e-3), rtol=1e-6, atol=1e-8)
    A = O.T @ O / Ns - np.outer(O.mean(0), O.mean(0)) + 1e-3 * np.eye(Np)
    g = sr.energy_gradient(O, E, E.mean())
    xs, _ = spla.minres(A, g, rtol=1e-10, maxiter=500)
    assert np.allclose(x, xs, rtol=1e-6, atol=1e-8)


# ------------------------------------------------------------------ stats
def test_tau_int_ar1():
    rng = np.random.default_rng(8)
    phi, n, T = 0.6, 400, 2000
    x = np.zeros((n, T))
    e = rng.normal(size=(n, T))
    for t in range(1, T):
        x[:, t] = phi * x[:, t - 1] + e[:, t]
    tau = stats.tau_int(x[:, 100:])
    assert tau == pytest.approx((1 + phi) / (1 - phi), rel=0.05)
    assert stats.tau_int(rng.normal(size=(100, 500))) == pytest.approx(1.0, abs=0.1)


# ------------------------------------------------------------------ NNBF
def _nnbf_small(seed=9, scale=0.3):
    from oracle import nnbf
    rng = np.random.default_rng(seed)
    H = hubbard.Hubbard(2, 2, 1.0, 8.0, 2, 2)
    params = nnbf.init_params(rng, 4, 4, hidden=6, n_layers=2, m_std=1.0)
    params = [p + 0.1 * rng.normal(size=p.shape) for p in params]   # nonzero biases
    return nnbf, H, params, scale


def test_nnbf_zero_scale_is_slater():
    nnbf, H, params, _ = _nnbf_small()
    x = H.all_states()
    s, l, _ = nnbf.forward(params, x, 2, 0.0)
    s0, l0 = slater.slogdet(params[0], x)
    assert np.array_equal(s, s0) and np.allclose(l, l0, atol=1e-13)


def test_nnbf_constant_output_reduces_to_slater():
    """With W_out = 0 the backflow is the constant b_out, so psi is the Slater
    determinant of M + s reshape(b_out, (2N, n_f)) (row-major); with W_1 = 0
    as well, h_1 = act(b_1) and the backflow is W_out act(W_2 act(b_1) + b_2)
    + b_out.  Both are compared with slater.slogdet at every basis state:
    a transposed reshape, a dropped bias or swapped layers fail them."""
    from oracle import nnbf
    nnbf_mod, H, params, scale = _nnbf_small()
    M, W1, b1, W2, b2, Wo, bo = params
    x = H.all_states()
    n_orb, nf = M.shape
    rng = np.random.default_rng(4)
    for act in ("relu", "tanh"):
        f = (lambda z: np.maximum(z, 0.0)) if act == "relu" else np.tanh
        for p, D in (([M, W1, b1, W2, b2, np.zeros_like(Wo), bo], bo),
                     ([M, np.zeros_like(W1), b1, W2, b2, Wo, bo],
                      Wo @ f(W2 @ f(b1) + b2) + bo)):
            s, l, _ = nnbf.forward(p, x, 2, scale, act)
            s0, l0 = slater.slogdet(M + scale * D.reshape(n_orb, nf), x)
            assert np.array_equal(s, s0) and np.allclose(l, l0, rtol=1e-13, atol=1e-13)
        # a transposed reshape of the constant backflow is caught
        C = rng.normal(size=n_orb * nf)
        p = [M, W1, b1, W2, b2, np.zeros_like(Wo), C]
        _, l, _ = nnbf.forward(p, x, 2, 1.0, act)
        _, lt = slater.slogdet(M + C.reshape(nf, n_orb).T, x)
        assert not np.allclose(l, lt, atol=1e-6)


def test_nnbf_log_derivative_finite_difference():
    nnbf, H, params, scale = _nnbf_small()
    x = H.all_states()[[1, 9, 30]]
    O = nnbf.log_derivative(params, x, 2, scale)
    flat = np.concatenate([p.ravel() for p in params])
    sizes = [p.size for p in params]

    def unflat(v):
        out, off = [], 0
        for p, n in zip(params, sizes):
            out.append(v[off:off + n].reshape(p.shape))
            off += n
        return out
    h = 1e-6
    rng = np.random.default_rng(0)
    for k in rng.choice(flat.size, 60, replace=False):
        vp, vm = flat.copy(), flat.copy()
        vp[k] += h
        vm[k] -= h
        fd = (nnbf.forward(unflat(vp), x, 2, scale)[1] - nnbf.forward(unflat(vm), x, 2, scale)[1]) / (2 * h)
        assert np.allclose(O[:, k], fd, atol=2e-6), k


def test_nnbf_param_count_config3():
    from oracle import nnbf
    p = nnbf.init_params(np.random.default_rng(0), 36, 32, hidden=256, n_layers=2)
    assert sum(a.size for a in p) == 678912
