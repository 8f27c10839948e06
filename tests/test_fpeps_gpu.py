"""GPU parity of the fPEPS / NN-fTNS kernels (exact boundary contraction,
environments for O_k, FP64 MLP backflow) against the FP64 oracle.

Tolerances: amplitudes and FP64 log-derivatives 1e-10 relative (the same
exact contraction in a different order); FP32 O 1e-6 relative to the row
scale; local energies 1e-9 relative; sweeps bit-exact in configurations and
signs with tie-free injected draws."""
import numpy as np
import pytest
import torch

from helpers import random_configs, tie_free_draws
from oracle import fpeps as fp
from oracle import hubbard as ohub
from vmc_torch_amd import NNfTNS_Model_GPU, fPEPS_Model_GPU
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch

pytestmark = pytest.mark.gpu


def _plain(Lx, Ly, D, seed, device):
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, np.random.default_rng(seed))
    m = fPEPS_Model_GPU(Lx, Ly, D, theta=th).to(device)
    return m, sh, th


def _nn(Lx, Ly, D, H, seed, device, eta=0.8, act="relu", correction="additive"):
    sh = fp.PEPSShape(Lx, Ly, D)
    rng = np.random.default_rng(seed)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, H, sh.n_params, 0.2)
    if correction == "multiplicative":
        nn[5] = rng.normal(0.0, 0.3, size=sh.n_params)     # nonzero output bias
    m = NNfTNS_Model_GPU(Lx, Ly, D, hidden_dim=H, nn_eta=eta, activation=act, theta=th,
                         correction=correction)
    with torch.no_grad():
        for p, v in zip(m.params[1:], nn):
            p.copy_(torch.as_tensor(v))
    return m.to(device), sh, th, nn


def _oracle_logamp(f):
    def la(cs):
        a = np.array([f(c) for c in np.atleast_2d(cs)])
        s = np.sign(a)
        with np.errstate(divide="ignore"):
            l = np.maximum(np.log(np.abs(a)), np.log(1e-45))
        return s, l
    return la


@pytest.mark.parametrize("Lx,Ly,D", [(4, 4, 4), (3, 4, 2), (2, 6, 4), (4, 3, 3), (4, 8, 2)])
def test_forward_matches_oracle(device, Lx, Ly, D):
    m, sh, th = _plain(Lx, Ly, D, 11 + Lx + Ly + D, device)
    rng = np.random.default_rng(1)
    x = rng.integers(0, 4, size=(40, sh.N))
    s, l = m.forward_log(torch.tensor(x, device=device))
    want = fp.amplitudes(sh, th, x)
    s, l = s.cpu().numpy(), l.cpu().numpy()
    for k in range(len(x)):
        if want[k] == 0.0:
            assert s[k] == 0.0
        else:
            assert s[k] == np.sign(want[k])
            assert l[k] == pytest.approx(np.log(abs(want[k])), abs=1e-10)


def test_nnftn_forward_matches_oracle(device):
    for (Lx, Ly, D, H) in [(2, 3, 2, 8), (4, 4, 4, 16)]:
        m, sh, th, nn = _nn(Lx, Ly, D, H, 3, device)
        x = random_configs(np.random.default_rng(2), 24, sh.N, sh.N // 2, sh.N // 2)
        s, l = m.forward_log(torch.tensor(x, device=device))
        want = np.array([fp.nnftn_amplitude(sh, th, nn, m.nn_eta, c) for c in x])
        assert np.array_equal(s.cpu().numpy(), np.sign(want))
        assert np.abs(l.cpu().numpy() - np.log(np.abs(want))).max() < 1e-10


@pytest.mark.parametrize("Lx,Ly,D", [(4, 4, 4), (3, 4, 2), (2, 3, 3)])
def test_logderiv_matches_oracle(device, Lx, Ly, D):
    m, sh, th = _plain(Lx, Ly, D, 5, device)
    x = random_configs(np.random.default_rng(3), 12, sh.N, sh.N // 2, sh.N // 2)
    xd = torch.tensor(x, device=device)
    O, (s, l) = m.log_derivative(xd, dtype=torch.float64)
    O32, _ = m.log_derivative(xd, dtype=torch.float32)
    O, O32 = O.cpu().numpy(), O32.cpu().numpy()
    for k in range(len(x)):
        want = fp.log_derivative(sh, th, x[k])
        sc = np.abs(want).max()
        assert np.abs(O[k] - want).max() < 1e-10 * sc
        assert np.abs(O32[k] - want).max() < 1e-6 * sc


@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_nnftn_logderiv_matches_oracle(device, act):
    m, sh, th, nn = _nn(2, 3, 2, 8, 4, device, act=act)
    x = random_configs(np.random.default_rng(4), 10, sh.N, 3, 3)
    O, _ = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float64)
    O = O.cpu().numpy()
    assert O.shape[1] == m.n_params
    for k in range(len(x)):
        want = fp.nnftn_log_derivative(sh, th, nn, m.nn_eta, x[k], act)
        assert np.abs(O[k] - want).max() < 1e-10 * np.abs(want).max()


def test_local_energy_matches_oracle(device):
    for kind in ("plain", "nn"):
        if kind == "plain":
            m, sh, th = _plain(4, 4, 4, 6, device)
            nn = None
        else:
            m, sh, th, nn = _nn(3, 2, 2, 8, 6, device)
        N = sh.N
        nu = nd = N // 2
        H = spinful_Fermi_Hubbard_square_lattice_torch(sh.Lx, sh.Ly, 1.0, 8.0, nu + nd,
                                                       n_fermions_per_spin=(nu, nd))
        x = random_configs(np.random.default_rng(7), 16, N, nu, nd)
        E = m.local_energy(torch.tensor(x, device=device), H).cpu().numpy()
        Ho = ohub.Hubbard(sh.Lx, sh.Ly, 1.0, 8.0, nu, nd)
        Eo = fp.local_energy(Ho, sh, th, x, nn, m.eta)
        assert np.allclose(E, Eo, rtol=1e-9, atol=1e-9 * np.abs(Eo).max())


def test_sweep_matches_oracle(device):
    for kind in ("plain", "nn"):
        if kind == "plain":
            m, sh, th = _plain(4, 4, 4, 8, device)
            f = lambda c: fp.amplitude(sh, th, c)          # noqa: E731
            nu = nd = 8
        else:
            m, sh, th, nn = _nn(2, 3, 2, 8, 8, device)
            f = lambda c: fp.nnftn_amplitude(sh, th, nn, m.nn_eta, c)   # noqa: E731
            nu = nd = 3
        H = spinful_Fermi_Hubbard_square_lattice_torch(sh.Lx, sh.Ly, 1.0, 8.0, nu + nd,
                                                       n_fermions_per_spin=(nu, nd))
        edges = H.graph.sweep_edges()
        rng = np.random.default_rng(9)
        x0 = random_configs(rng, 24, sh.N, nu, nd)
        draws, xo, so, lo = tie_free_draws(_oracle_logamp(f), x0, edges, 2, rng)
        xd = torch.tensor(x0, device=device)
        dd = tuple(torch.tensor(a, device=device) for a in draws)
        sign, logabs, nacc = m.mcmc_sweep(xd, H.graph.edge_tensor(device, "sweep"), 2, 0, 0,
                                          draws=dd)
        assert np.array_equal(xd.cpu().numpy(), xo)
        assert np.array_equal(sign.cpu().numpy(), so)
        assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-10


def test_config4_shape(device):
    """4x4, D = 4, MLP 32 -> 128 -> 128 -> 3200: 436,736 parameters."""
    m = NNfTNS_Model_GPU(4, 4, 4, hidden_dim=128, nn_eta=1.0,
                         generator=torch.Generator().manual_seed(0)).to(device)
    assert m.n_params == 436736
    x = random_configs(np.random.default_rng(10), 32, 16, 8, 8)
    O, (s, l) = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float32)
    assert O.shape == (32, 436736)
    assert bool(torch.isfinite(O).all())
    sh = fp.PEPSShape(4, 4, 4)
    th = m.params[0].detach().cpu().numpy()
    nn = [p.detach().cpu().numpy() for p in m.params[1:]]
    want = np.array([fp.nnftn_amplitude(sh, th, nn, 1.0, c) for c in x[:4]])
    assert np.abs(l.cpu().numpy()[:4] - np.log(np.abs(want))).max() < 1e-10


MODE = {"additive": "add", "multiplicative": "mul"}


@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_nnftn_multiplicative_matches_oracle(device, act):
    """theta(x) = theta (1 + eta MLP(n(x))): amplitudes, FP64 O (dense and as
    factors multiplied out), local energies and a tie-free sweep against the
    oracle's multiplicative form (1e-10 / 1e-9, bit-exact configurations)."""
    m, sh, th, nn = _nn(2, 3, 2, 8, 14, device, act=act, correction="multiplicative")
    x = random_configs(np.random.default_rng(15), 12, sh.N, 3, 3)
    xd = torch.tensor(x, device=device)
    s, l = m.forward_log(xd)
    want = np.array([fp.nnftn_amplitude(sh, th, nn, m.nn_eta, c, act, "mul") for c in x])
    assert np.array_equal(s.cpu().numpy(), np.sign(want))
    assert np.abs(l.cpu().numpy() - np.log(np.abs(want))).max() < 1e-10
    O, _ = m.log_derivative(xd, dtype=torch.float64)
    FO, _ = m.log_derivative_factored(xd)
    Of = FO.materialize().cpu().numpy()
    O = O.cpu().numpy()
    for k in range(len(x)):
        g = fp.nnftn_log_derivative(sh, th, nn, m.nn_eta, x[k], act, "mul")
        sc = np.abs(g).max()
        assert np.abs(O[k] - g).max() < 1e-10 * sc
        assert np.abs(Of[k] - g).max() < 1e-10 * sc
    Hm = spinful_Fermi_Hubbard_square_lattice_torch(2, 3, 1.0, 8.0, 6, n_fermions_per_spin=(3, 3))
    E = m.local_energy(xd, Hm).cpu().numpy()
    Eo = fp.local_energy(ohub.Hubbard(2, 3, 1.0, 8.0, 3, 3), sh, th, x, nn, m.nn_eta, act, "mul")
    assert np.allclose(E, Eo, rtol=1e-9, atol=1e-9 * np.abs(Eo).max())
    f = lambda c: fp.nnftn_amplitude(sh, th, nn, m.nn_eta, c, act, "mul")   # noqa: E731
    edges = Hm.graph.sweep_edges()
    rng = np.random.default_rng(16)
    x0 = random_configs(rng, 24, sh.N, 3, 3)
    draws, xo, so, lo = tie_free_draws(_oracle_logamp(f), x0, edges, 2, rng)
    xd0 = torch.tensor(x0, device=device)
    dd = tuple(torch.tensor(a, device=device) for a in draws)
    sign, logabs, _ = m.mcmc_sweep(xd0, Hm.graph.edge_tensor(device, "sweep"), 2, 0, 0,
                                   draws=dd)
    assert np.array_equal(xd0.cpu().numpy(), xo)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-10


def test_nnftn_multiplicative_config4_factored_sr(device):
    """The 4x4 D = 4 configuration with the multiplicative correction: the
    factored SR step equals the dense FP64 step on the same O (1e-7, as the
    additive form's test) and the parameter count is that of the additive
    form (the correction changes the map, not the parameters)."""
    from vmc_torch_amd import kernels
    m = NNfTNS_Model_GPU(4, 4, 4, hidden_dim=32, nn_eta=0.5, correction="multiplicative",
                         generator=torch.Generator().manual_seed(5)).to(device)
    x = torch.tensor(random_configs(np.random.default_rng(17), 200, 16, 8, 8), device=device)
    O, (s, l) = m.log_derivative(x, dtype=torch.float64)
    FO, (sf, lf) = m.log_derivative_factored(x)
    assert torch.equal(s, sf) and torch.equal(l, lf)
    Of = FO.materialize()
    assert float((Of - O).abs().max()) <= 1e-13 * float(O.abs().max())
    E = torch.tensor(np.random.default_rng(18).normal(size=200) + 1.0, device=device)
    dp, _, info = kernels.sr_minsr_factored(FO, E, None, 1e-3)
    dpd, _, _ = kernels.sr_minsr(Of.contiguous(), E, None, 1e-3, "fp64")
    assert int(info.item()) == 0
    rel = float((dp - dpd).norm() / dpd.norm())
    assert rel < 1e-7
