"""GPU parity of the chi-truncated boundary-MPS fPEPS / NN-fTNS amplitudes
(fPEPS_Model_GPU(max_bond=chi): bmps.hip) against the oracle's boundary MPS
(oracle.fpeps.bmps_*, QR canonical form and SVD truncation; the device
truncates through right Grams and a Jacobi eigensolver per Z2 block).

Tolerances: FP64 both sides, ln|psi| within 1e-8 and signs equal for every
configuration whose truncations all have a relative singular-value gap
(s[chi-1] - s[chi]) / s[0] of at least 1e-6 (below that the truncated state
is ill-conditioned and both sides are right to differ; the tests assert such
configurations are rare); local energies 1e-8 relative to the batch's
largest; sweeps with injected draws reproduce the oracle's configurations and
signs exactly."""
import numpy as np
import pytest
import torch

from helpers import random_configs, tie_free_draws
from oracle import fpeps as fp
from oracle import hubbard as ohub
from vmc_torch_amd import (MetropolisExchangeSpinfulSamplerGPU,
                           MetropolisExchangeSpinfulSamplerReuse_GPU, NNfTNS_Model_GPU,
                           evaluate_energy, evaluate_energy_reuse, fPEPS_Model_GPU,
                           fPEPS_Model_reuse_GPU)
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch

pytestmark = pytest.mark.gpu

GAP = 1e-6


def _plain(Lx, Ly, D, chi, seed, device):
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, np.random.default_rng(seed))
    return fPEPS_Model_GPU(Lx, Ly, D, theta=th, max_bond=chi).to(device), sh, th


def _oracle(sh, th, xs, chi, nn=None, eta=0.0, mode="add"):
    out = []
    for x in np.atleast_2d(xs):
        t = th
        if nn is not None:
            y, _ = fp.mlp_forward(nn, x)
            t = fp.theta_eff(th, y[0], eta, mode)
        g = []
        s, l = fp.bmps_log_amplitude(sh, t, x, chi, g)
        out.append((s, l, min(g) if g else np.inf))
    a = np.array(out)
    return a[:, 0], a[:, 1], a[:, 2]


def _check(s, l, so, lo, go, min_frac=0.9):
    s, l = s.cpu().numpy(), l.cpu().numpy()
    ok = go >= GAP
    assert ok.mean() >= min_frac
    assert np.array_equal(s[ok], so[ok])
    nz = ok & (so != 0)
    assert nz.sum() >= ok.sum() // 2
    assert np.abs(l[nz] - lo[nz]).max() < 1e-8


@pytest.mark.parametrize("Lx,Ly,D,chi", [(4, 4, 2, 3), (4, 5, 2, 3), (6, 4, 2, 3), (4, 4, 4, 8),
                                         (5, 4, 3, 5), (6, 6, 4, 16), (3, 6, 3, 4),
                                         (8, 8, 2, 6)])
def test_bmps_forward_matches_oracle(device, Lx, Ly, D, chi):
    m, sh, th = _plain(Lx, Ly, D, chi, 31 + Lx + Ly + D + chi, device)
    n = 24 if sh.N > 30 else 48
    x = random_configs(np.random.default_rng(3), n, sh.N, sh.N // 2, sh.N // 2)
    s, l = m.forward_log(torch.tensor(x, device=device))
    so, lo, go = _oracle(sh, th, x, chi)
    assert np.isfinite(go).any()          # the case truncates
    _check(s, l, so, lo, go)


@pytest.mark.parametrize("Lx,Ly,D,chi", [(6, 3, 2, 8), (5, 2, 3, 9), (6, 4, 1, 1), (6, 4, 1, 32)])
def test_bmps_untruncated_equals_exact_contraction(device, Lx, Ly, D, chi):
    """Lattices beyond the exact kernels with chi at or above every bond:
    the boundary MPS equals the exact zipper contraction (D = 1, chi = 32:
    buffers laid out for bonds of 32, bonds staying 1)."""
    m, sh, th = _plain(Lx, Ly, D, chi, 5 + Lx + Ly, device)
    rng = np.random.default_rng(4)
    if D == 1:
        # even bonds only: singly occupied sites vanish, so empty / doubly occupied ones
        x = np.array([3 * (rng.permutation(sh.N) < sh.N // 2) for _ in range(32)])
    else:
        x = random_configs(rng, 32, sh.N, sh.N // 2, sh.N // 2)
    s, l = m.forward_log(torch.tensor(x, device=device))
    want = fp.amplitudes(sh, th, x)
    s, l = s.cpu().numpy(), l.cpu().numpy()
    nz = np.abs(want) > 1e-12 * np.abs(want).max()
    assert nz.sum() >= 8
    assert np.array_equal(s[nz], np.sign(want[nz]))
    assert np.abs(l[nz] - np.log(np.abs(want[nz]))).max() < 1e-9


@pytest.mark.parametrize("correction", ["additive", "multiplicative"])
def test_bmps_nnftn_forward_matches_oracle(device, correction):
    Lx, Ly, D, chi, H = 4, 5, 2, 3, 8
    sh = fp.PEPSShape(Lx, Ly, D)
    rng = np.random.default_rng(12)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, H, sh.n_params, 0.2)
    m = NNfTNS_Model_GPU(Lx, Ly, D, hidden_dim=H, nn_eta=0.7, theta=th, max_bond=chi,
                         correction=correction)
    with torch.no_grad():
        for p, v in zip(m.params[1:], nn):
            p.copy_(torch.as_tensor(v))
    m = m.to(device)
    x = random_configs(np.random.default_rng(5), 32, sh.N, sh.N // 2, sh.N // 2)
    s, l = m.forward_log(torch.tensor(x, device=device))
    mode = "add" if correction == "additive" else "mul"
    so, lo, go = _oracle(sh, th, x, chi, nn, m.nn_eta, mode)
    _check(s, l, so, lo, go)


def test_bmps_passes_agree(device):
    """More configurations than one pass of the kernels' workspace (2048):
    every pass gives the values of a small batch, and a sample of them the
    oracle's."""
    m, sh, th = _plain(4, 4, 2, 3, 77, device)
    x = random_configs(np.random.default_rng(6), 2100, sh.N, 8, 8)
    xd = torch.tensor(x, device=device)
    s, l = m.forward_log(xd)
    s2, l2 = m.forward_log(xd[2040:].clone())
    assert torch.equal(s[2040:], s2) and torch.equal(l[2040:], l2)
    so, lo, go = _oracle(sh, th, x[2040:], 3)
    _check(s2, l2, so, lo, go)


def _oracle_eloc(sh, th, x, chi):
    """E_loc from the oracle's boundary-MPS amplitudes, and which walkers
    have every truncation (theirs and their connected configurations')
    well conditioned."""
    Ho = ohub.Hubbard(sh.Lx, sh.Ly, 1.0, 8.0, sh.N // 2, sh.N // 2)
    s0, l0, g0 = _oracle(sh, th, x, chi)
    etas, coeffs, bids = Ho.get_conn_batch(x)
    se, le, ge = _oracle(sh, th, etas, chi)
    Eo = np.zeros(x.shape[0])
    good = g0 >= GAP
    for e_s, e_l, e_g, c, b in zip(se, le, ge, coeffs, bids):
        Eo[b] += c * e_s * s0[b] * np.exp(e_l - l0[b]) if e_s != 0 else 0.0
        good[b] &= e_g >= GAP
    return Eo, good


def test_bmps_local_energy_matches_oracle(device):
    m, sh, th = _plain(4, 4, 2, 3, 41, device)
    H = spinful_Fermi_Hubbard_square_lattice_torch(4, 4, 1.0, 8.0, 16, n_fermions_per_spin=(8, 8))
    x = random_configs(np.random.default_rng(8), 12, sh.N, 8, 8)
    E = m.local_energy(torch.tensor(x, device=device), H).cpu().numpy()
    Eo, good = _oracle_eloc(sh, th, x, 3)
    assert good.sum() >= 8
    assert np.allclose(E[good], Eo[good], rtol=1e-8, atol=1e-8 * np.abs(Eo).max())


# (6, 8, 2, 16): every row environment of both sides, bonds up to 16 = the
# largest boundary tensor and projection of the 32 x 32 working matrices
@pytest.mark.parametrize("Lx,Ly,D,chi", [(4, 5, 2, 3), (6, 4, 2, 3), (5, 4, 3, 5), (6, 6, 4, 16),
                                         (6, 8, 2, 16)])
def test_bmps_log_derivative_matches_oracle(device, Lx, Ly, D, chi):
    """The reference's cheap gradient (environments held fixed) against the
    oracle's, entry by entry (1e-8 of the row's largest), dense and as
    factors."""
    m, sh, th = _plain(Lx, Ly, D, chi, 61 + Lx + D, device)
    x = random_configs(np.random.default_rng(9), 6 if sh.N > 30 else 10, sh.N, sh.N // 2,
                       sh.N // 2)
    xd = torch.tensor(x, device=device)
    O, (s, l) = m.log_derivative(xd, dtype=torch.float64)
    FO, (sf, lf) = m.log_derivative_factored(xd)
    assert torch.equal(s, sf) and torch.equal(l, lf)
    Of = FO.materialize().cpu().numpy()
    O = O.cpu().numpy()
    n_ok = 0
    for k in range(len(x)):
        gaps = []
        g, so, lo = fp.bmps_cheap_log_derivative(sh, th, x[k], chi, gaps)
        if gaps and min(gaps) < GAP:
            continue
        n_ok += 1
        sc = np.abs(g).max()
        assert np.abs(O[k] - g).max() < 1e-8 * sc
        assert np.abs(Of[k] - g).max() < 1e-8 * sc
        assert float(s[k]) == so and abs(float(l[k]) - lo) < 1e-8
    assert n_ok >= len(x) * 2 // 3


def test_bmps_log_derivative_exact_limit(device):
    """A 6-row lattice with chi at every bond: the exact log-derivative."""
    m, sh, th = _plain(6, 3, 2, 8, 71, device)
    x = random_configs(np.random.default_rng(11), 8, sh.N, 9, 9)
    O, _ = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float64)
    O = O.cpu().numpy()
    for k in range(len(x)):
        g = fp.log_derivative(sh, th, x[k])
        assert np.abs(O[k] - g).max() < 1e-9 * np.abs(g).max()


def test_bmps_nnftn_log_derivative_matches_oracle(device):
    Lx, Ly, D, chi, H = 4, 5, 2, 3, 8
    sh = fp.PEPSShape(Lx, Ly, D)
    rng = np.random.default_rng(13)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, H, sh.n_params, 0.2)
    m = NNfTNS_Model_GPU(Lx, Ly, D, hidden_dim=H, nn_eta=0.7, theta=th, max_bond=chi)
    with torch.no_grad():
        for p, v in zip(m.params[1:], nn):
            p.copy_(torch.as_tensor(v))
    m = m.to(device)
    x = random_configs(np.random.default_rng(14), 8, sh.N, 10, 10)
    O, _ = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float64)
    O = O.cpu().numpy()
    _, _, go = _oracle(sh, th, x, chi, nn, m.nn_eta)
    n_ok = 0
    for k in range(len(x)):
        if go[k] < GAP:
            continue
        n_ok += 1
        g = fp.nnftn_log_derivative(sh, th, nn, m.nn_eta, x[k], chi=chi)
        assert np.abs(O[k] - g).max() < 1e-8 * np.abs(g).max()
    assert n_ok >= 5


def test_bmps_vmc_step_matches_oracle(device):
    """One SR step of the reference-shaped driver with a chi-truncated fPEPS
    equals the oracle's step on the same configurations (boundary-MPS local
    energies, cheap gradients, minSR) within 1e-7 of the largest update."""
    from oracle import sr as osr
    from vmc_torch_amd import MetropolisExchangeSpinfulSamplerGPU
    from vmc_torch_amd.optimizer import MinSRGPU, SGDGPU
    from vmc_torch_amd.VMC import VMC_GPU, VMCLoopConfig
    m, sh, th = _plain(4, 4, 2, 3, 81, device)
    B, lr = 48, 0.05
    Hg = spinful_Fermi_Hubbard_square_lattice_torch(4, 4, 1.0, 8.0, 16, n_fermions_per_spin=(8, 8))
    Hg.precompute_hops_gpu(device)
    fxs = torch.tensor(random_configs(np.random.default_rng(12), B, 16, 8, 8), device=device)
    vmc = VMC_GPU(MetropolisExchangeSpinfulSamplerGPU(seed=4), MinSRGPU("fp64"), SGDGPU(lr))
    cfg = VMCLoopConfig(vmc_steps=1, ns_per_rank=B, n_params=m.n_params, nsites=16,
                        learning_rate=lr, diag_shift=1e-3, burn_in_steps=2,
                        outlier_clip_factor=0.0, use_log_amp=True)
    log = []
    vmc.run_vmc_loop(fxs, m, Hg, Hg.graph, config=cfg,
                     on_step_end=lambda d: log.append(d["fxs"].cpu().numpy().copy()))
    x = log[0]
    E, good = _oracle_eloc(sh, th, x, 3)
    assert good.all()
    O = np.array([fp.bmps_cheap_log_derivative(sh, th, c, 3)[0] for c in x])
    dp = osr.minsr(O, E, E.mean(), 1e-3)
    got = m.params[0].detach().cpu().numpy()
    want = th - lr * dp
    assert np.abs(got - want).max() <= 1e-7 * np.abs(lr * dp).max()


def test_bmps_sweep_matches_oracle(device):
    m, sh, th = _plain(4, 4, 2, 3, 52, device)
    gaps = []

    def la(cs):
        s, l, g = _oracle(sh, th, cs, 3)
        gaps.extend(g.tolist())
        return s, np.maximum(l, np.log(1e-45))
    H = spinful_Fermi_Hubbard_square_lattice_torch(4, 4, 1.0, 8.0, 16, n_fermions_per_spin=(8, 8))
    edges = H.graph.sweep_edges()
    rng = np.random.default_rng(10)
    x0 = random_configs(rng, 12, sh.N, 8, 8)
    draws, xo, so, lo = tie_free_draws(la, x0, edges, 1, rng, margin=1e-6)
    assert min(gaps) >= GAP           # no ill-conditioned truncation on the oracle's path
    xd = torch.tensor(x0, device=device)
    dd = tuple(torch.tensor(a, device=device) for a in draws)
    sign, logabs, nacc = m.mcmc_sweep(xd, H.graph.edge_tensor(device, "sweep"), 1, 0, 0,
                                      draws=dd)
    assert np.array_equal(xd.cpu().numpy(), xo)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-8


def test_bmps_options(device):
    with pytest.raises(ValueError):
        fPEPS_Model_GPU(4, 4, 2, max_bond=3, amp="fp32")
    assert fPEPS_Model_reuse_GPU(4, 4, 2, max_bond=3)._bmps   # boundary-cache reuse
    with pytest.raises(NotImplementedError):
        fPEPS_Model_GPU(6, 4, 4, max_bond=20)          # max(chi, D) * D > 64
    # chi at or above D^2 on Lx <= 4 keeps the exact kernels
    m = fPEPS_Model_GPU(4, 4, 2, max_bond=4)
    assert not m._bmps
    assert fPEPS_Model_GPU(4, 4, 2, max_bond=3)._bmps


@pytest.mark.parametrize("Lx,Ly,D,chi", [(6, 4, 2, 3), (4, 4, 4, 8), (5, 4, 3, 5), (3, 4, 2, 2)])
def test_bmps_sweep_cache_bit_identical(device, Lx, Ly, D, chi):
    """Sweeps that resume every proposal from the walkers' cached boundaries
    give the full recontraction's walkers, signs, ln|psi| and acceptance
    counts bit for bit (same Philox draws)."""
    N = Lx * Ly
    m, sh, th = _plain(Lx, Ly, D, chi, 91 + Lx + D, device)
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, 2 * (N // 2),
                                                   n_fermions_per_spin=(N // 2, N // 2))
    x0 = torch.tensor(random_configs(np.random.default_rng(21), 64, N, N // 2, N // 2),
                      device=device)
    out = []
    for cache in (True, False):
        m.bmps_cache = cache
        m._struct_obj = None
        m._state = None          # fresh walker state (acceptance counts from zero)
        xd = x0.clone()
        sign, logabs, nacc = m.mcmc_sweep(xd, H.graph.edge_tensor(device, "sweep"), 2, 7, 0)
        out.append((xd, sign.clone(), logabs.clone(), nacc.clone()))
    m.bmps_cache = True
    m._struct_obj = None
    assert int(out[0][3].sum()) > 0
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("Lx,Ly,D,chi", [(6, 4, 2, 3), (5, 4, 3, 5)])
def test_bmps_local_energy_cache_bit_identical(device, Lx, Ly, D, chi):
    """Local energies with every connected configuration resumed from its
    walker's cached boundaries equal the full recontraction's bit for bit."""
    N = Lx * Ly
    m, sh, th = _plain(Lx, Ly, D, chi, 95 + Lx + D, device)
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, 2 * (N // 2),
                                                   n_fermions_per_spin=(N // 2, N // 2))
    x = torch.tensor(random_configs(np.random.default_rng(22), 40, N, N // 2, N // 2),
                     device=device)
    out = []
    for cache in (True, False):
        m.bmps_cache = cache
        m._struct_obj = None
        m._state = None
        out.append(m.local_energy(x, H).clone())
    m.bmps_cache = True
    m._struct_obj = None
    assert torch.isfinite(out[0]).all()
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("Lx,Ly,D,chi", [(6, 4, 2, 3), (5, 4, 3, 5)])
def test_bmps_reuse_model_equals_full_recontraction(device, Lx, Ly, D, chi):
    """fPEPS_Model_reuse_GPU with a truncating max_bond: the reference-named
    reuse sampler and evaluate_energy_reuse run from the walkers' cached
    boundaries and give the plain model's full recontraction bit for bit
    (walkers, signs, ln|psi|, acceptance counts, E_loc); the reuse model's
    _full methods recontract every row and agree too."""
    N = Lx * Ly
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, np.random.default_rng(131 + Lx + D))
    full = fPEPS_Model_GPU(Lx, Ly, D, theta=th, max_bond=chi).to(device)
    full.bmps_cache = False
    reuse = fPEPS_Model_reuse_GPU(Lx, Ly, D, theta=th, max_bond=chi).to(device)
    assert reuse._bmps and reuse.bmps_cache
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, 2 * (N // 2),
                                                   n_fermions_per_spin=(N // 2, N // 2))
    x0 = torch.tensor(random_configs(np.random.default_rng(23), 48, N, N // 2, N // 2),
                      device=device)
    xa, xb = x0.clone(), x0.clone()
    sa = MetropolisExchangeSpinfulSamplerGPU(seed=77)
    sb = MetropolisExchangeSpinfulSamplerReuse_GPU(seed=77)
    for _ in range(2):
        _, (sga, la) = sa.step(xa, full, H.graph, use_log_amp=True)
        _, (sgb, lb) = sb.step(xb, reuse, H.graph, use_log_amp=True)
    assert torch.equal(xa, xb)
    assert torch.equal(sga, sgb)
    assert torch.equal(la, lb)
    assert torch.equal(sa.last_acceptance[0], sb.last_acceptance[0])
    assert int(sa.last_acceptance[0].sum()) > 0
    Ea = evaluate_energy(xa, full, H)[1]
    Eb = evaluate_energy_reuse(xb, reuse, H)[1]
    assert torch.isfinite(Ea).all()
    assert torch.equal(Ea, Eb)
    assert torch.equal(reuse.local_energy_full(xb, H), Ea)
    assert reuse.bmps_cache
