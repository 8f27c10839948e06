"""GPU parity of the FP32 matrix-core fPEPS / NN-fTNS amplitudes
(amp="fp32": the column contraction on v_mfma_f32_16x16x4f32) against the
FP64 oracle.

Tolerances (FP32 slices and products, FP64 log scale): for every
configuration whose amplitude is at least 1e-3 of the batch's median
nonzero amplitude (amplitudes of random configurations spread over many
decades; far smaller ones are sums with deep cancellation), the sign equals
the oracle's and ln|psi| is within 1e-3; the median ln|psi| error is below
1e-5; amplitudes that vanish by parity vanish exactly.  Local
energies within 1e-4 relative to the batch's largest; sweeps with injected
draws whose accept uniforms sit at least 1e-2 (relative) away from their
ratios reproduce the oracle's configurations and signs."""
import numpy as np
import pytest
import torch

from helpers import random_configs, tie_free_draws
from oracle import fpeps as fp
from oracle import hubbard as ohub
from vmc_torch_amd import NNfTNS_Model_GPU, fPEPS_Model_GPU, fPEPS_Model_reuse_GPU
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch

pytestmark = pytest.mark.gpu


def _plain(Lx, Ly, D, seed, device, amp="fp32"):
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, np.random.default_rng(seed))
    return fPEPS_Model_GPU(Lx, Ly, D, theta=th, amp=amp).to(device), sh, th


def _nn(Lx, Ly, D, H, seed, device, eta=0.8, correction="additive"):
    sh = fp.PEPSShape(Lx, Ly, D)
    rng = np.random.default_rng(seed)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, H, sh.n_params, 0.2)
    m = NNfTNS_Model_GPU(Lx, Ly, D, hidden_dim=H, nn_eta=eta, activation="relu", theta=th,
                         correction=correction, amp="fp32")
    with torch.no_grad():
        for p, v in zip(m.params[1:], nn):
            p.copy_(torch.as_tensor(v))
    return m.to(device), sh, th, nn


def _check_amplitudes(s, l, want):
    s, l = s.cpu().numpy(), l.cpu().numpy()
    zero = want == 0.0
    assert np.all(s[zero] == 0.0)
    a = np.abs(want)
    nz = a > 0.0
    big = a >= 1e-3 * np.median(a[nz])
    assert big.sum() >= nz.sum() // 2
    assert np.array_equal(s[big], np.sign(want[big]))
    err = np.abs(l[big] - np.log(a[big]))
    assert err.max() < 1e-3
    assert np.median(err) < 1e-5


@pytest.mark.parametrize("Lx,Ly,D", [(4, 4, 4), (3, 4, 2), (2, 6, 4), (4, 3, 3), (4, 8, 2),
                                     (2, 2, 1)])
def test_fp32_forward_matches_oracle(device, Lx, Ly, D):
    m, sh, th = _plain(Lx, Ly, D, 11 + Lx + Ly + D, device)
    rng = np.random.default_rng(1)
    x = rng.integers(0, 4, size=(200, sh.N))
    s, l = m.forward_log(torch.tensor(x, device=device))
    _check_amplitudes(s, l, fp.amplitudes(sh, th, x))


def test_fp32_forward_half_filled_matches_fp64(device):
    """4x4 D=4 at half filling, 2048 walkers: FP32 against the FP64 kernels."""
    m32, sh, th = _plain(4, 4, 4, 21, device)
    m64, _, _ = _plain(4, 4, 4, 21, device, amp="fp64")
    x = torch.tensor(random_configs(np.random.default_rng(2), 2048, sh.N, 8, 8), device=device)
    s32, l32 = m32.forward_log(x)
    s64, l64 = m64.forward_log(x)
    want = (s64 * torch.exp(l64 - l64.max())).cpu().numpy()
    _check_amplitudes(s32, l32 - l64.max(), want)


@pytest.mark.parametrize("correction", ["additive", "multiplicative"])
def test_fp32_nnftn_forward_matches_oracle(device, correction):
    for (Lx, Ly, D, H) in [(2, 3, 2, 8), (4, 4, 4, 16)]:
        m, sh, th, nn = _nn(Lx, Ly, D, H, 3, device, correction=correction)
        x = random_configs(np.random.default_rng(2), 64, sh.N, sh.N // 2, sh.N // 2)
        s, l = m.forward_log(torch.tensor(x, device=device))
        mode = "add" if correction == "additive" else "mul"
        want = np.array([fp.nnftn_amplitude(sh, th, nn, m.nn_eta, c, mode=mode) for c in x])
        _check_amplitudes(s, l, want)


def test_fp32_local_energy_matches_oracle(device):
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
        x = random_configs(np.random.default_rng(7), 32, N, nu, nd)
        E = m.local_energy(torch.tensor(x, device=device), H).cpu().numpy()
        Ho = ohub.Hubbard(sh.Lx, sh.Ly, 1.0, 8.0, nu, nd)
        Eo = fp.local_energy(Ho, sh, th, x, nn, m.eta)
        assert np.abs(E - Eo).max() < 1e-4 * np.abs(Eo).max()


def test_fp32_sweep_matches_oracle(device):
    m, sh, th = _plain(4, 4, 4, 8, device)

    def la(cs):
        a = np.array([fp.amplitude(sh, th, c) for c in np.atleast_2d(cs)])
        with np.errstate(divide="ignore"):
            return np.sign(a), np.maximum(np.log(np.abs(a)), np.log(1e-45))
    H = spinful_Fermi_Hubbard_square_lattice_torch(4, 4, 1.0, 8.0, 16, n_fermions_per_spin=(8, 8))
    edges = H.graph.sweep_edges()
    rng = np.random.default_rng(9)
    x0 = random_configs(rng, 24, sh.N, 8, 8)
    draws, xo, so, lo = tie_free_draws(la, x0, edges, 2, rng, margin=1e-2)
    xd = torch.tensor(x0, device=device)
    dd = tuple(torch.tensor(a, device=device) for a in draws)
    sign, logabs, nacc = m.mcmc_sweep(xd, H.graph.edge_tensor(device, "sweep"), 2, 0, 0,
                                      draws=dd)
    assert np.array_equal(xd.cpu().numpy(), xo)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-3


def test_fp32_options():
    with pytest.raises(ValueError):
        fPEPS_Model_GPU(2, 2, 2, amp="bf16")
    with pytest.raises(ValueError):
        fPEPS_Model_reuse_GPU(2, 2, 2, amp="fp32")
