"""GPU parity of the NNBF path (libvmcx through vmc_torch_amd.NNBF_GPU)
against the FP64 oracle (oracle/nnbf.py).

The MLP runs in FP32 (gemm="fp32") or BF16 with FP32 accumulation
(gemm="bf16"); the determinant is FP64.  Tolerances, stated per test:
  FP32 MLP : ln|psi| within 2e-4 absolute, O within 2e-3 of max|O| per
             parameter block, E_loc within 1e-3 relative;
  BF16 MLP : against the oracle with the same BF16 operand rounding,
             ln|psi| within 2e-4 absolute.
Signs, configurations and accept decisions (away from ties) are identical.
"""
import numpy as np
import pytest
import torch

from helpers import random_configs, tie_free_draws
from oracle import hubbard as ohub
from oracle import nnbf as onn
from oracle import sampler as osamp
from vmc_torch_amd import NNBF_GPU
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch

pytestmark = pytest.mark.gpu

# Lx, Ly, nu, nd, hidden; the determinant layouts NFP = 8, 32, 8, 64, 16, 64
# (nf = 35: 64 with identity padding)
CASES = [(4, 2, 4, 4, 32), (6, 6, 16, 16, 64), (3, 3, 4, 3, 16), (8, 8, 32, 32, 32),
         (4, 4, 5, 4, 32), (6, 6, 18, 17, 32)]


def _model(Lx, Ly, nu, nd, hidden, gemm, device, scale=0.3, act="relu", seed=0):
    g = torch.Generator().manual_seed(seed)
    m = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=hidden, activation=act, bf_scale=scale,
                 gemm=gemm, init="xavier", m_std=0.3, generator=g, chunk=4096)
    with torch.no_grad():
        for p in m.params:
            if p.dim() == 1:    # nonzero biases exercise the bias paths
                p.copy_(0.05 * torch.randn(p.shape, generator=g, dtype=p.dtype))
    m = m.to(device)
    # the oracle sees the FP32-rounded weights the kernels use
    params = [p.detach().to(torch.float32).double().cpu().numpy() for p in m.params]
    return m, params


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("act", ["relu", "tanh"])
def test_nnbf_forward_fp32(device, case, act):
    Lx, Ly, nu, nd, hid = case
    m, params = _model(Lx, Ly, nu, nd, hid, "fp32", device, act=act)
    x = random_configs(np.random.default_rng(1), 64, Lx * Ly, nu, nd)
    s, l = m.forward_log(torch.tensor(x, device=device))
    so, lo, _ = onn.forward(params, x, 2, m.bf_scale, act)
    assert np.array_equal(s.cpu().numpy(), so)
    assert np.abs(l.cpu().numpy() - lo).max() < 2e-4


@pytest.mark.parametrize("scale", [0.3, 0.01])
def test_nnbf_forward_bf16(device, scale):
    """BF16 GEMM mode against the oracle with the same BF16 operand rounding
    (round to nearest even of activations and weights of the dense layers):
    ln|psi| within 2e-4, identical signs."""
    Lx, Ly, nu, nd, hid = CASES[1]
    m, params = _model(Lx, Ly, nu, nd, hid, "bf16", device, scale=scale)
    x = random_configs(np.random.default_rng(2), 128, Lx * Ly, nu, nd)
    s, l = m.forward_log(torch.tensor(x, device=device))
    so, lo, _ = onn.forward(params, x, 2, m.bf_scale, bf16_gemm=True)
    assert np.array_equal(s.cpu().numpy(), so)
    assert np.abs(l.cpu().numpy() - lo).max() < 2e-4


@pytest.mark.parametrize("case", CASES)
def test_nnbf_logderiv_fp32(device, case):
    Lx, Ly, nu, nd, hid = case
    m, params = _model(Lx, Ly, nu, nd, hid, "fp32", device)
    x = random_configs(np.random.default_rng(3), 24, Lx * Ly, nu, nd)
    O, (s, l) = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float64)
    Oo = onn.log_derivative(params, x, 2, m.bf_scale)
    O = O.cpu().numpy()
    assert O.shape == Oo.shape
    off = 0
    for p in params:
        blk = slice(off, off + p.size)
        scale = max(np.abs(Oo[:, blk]).max(), 1e-12)
        assert np.abs(O[:, blk] - Oo[:, blk]).max() / scale < 2e-3, p.shape
        off += p.size
    O32, _ = m.log_derivative(torch.tensor(x, device=device), dtype=torch.float32)
    assert np.allclose(O32.cpu().numpy(), O, rtol=1e-5, atol=1e-6 * np.abs(Oo).max())


@pytest.mark.parametrize("case", CASES[:2] + CASES[3:])
def test_nnbf_local_energy_fp32(device, case):
    Lx, Ly, nu, nd, hid = case
    m, params = _model(Lx, Ly, nu, nd, hid, "fp32", device)
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    x = random_configs(np.random.default_rng(4), 40, Lx * Ly, nu, nd)
    xd = torch.tensor(x, device=device)
    E = m.local_energy(xd, H).cpu().numpy()
    Eo = onn.local_energy(ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd), params, x, 2, m.bf_scale)
    assert np.allclose(E, Eo, rtol=1e-3, atol=1e-3 * np.abs(Eo).max())


@pytest.mark.parametrize("case", CASES)
def test_nnbf_sweep_matches_oracle(device, case):
    Lx, Ly, nu, nd, hid = case
    m, params = _model(Lx, Ly, nu, nd, hid, "fp32", device)
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    edges = H.graph.sweep_edges()
    rng = np.random.default_rng(5)
    x0 = random_configs(rng, 48, Lx * Ly, nu, nd)
    draws, xo, so, lo = tie_free_draws(
        lambda c: onn.forward(params, c, 2, m.bf_scale)[:2], x0, edges, 2, rng)
    xd = torch.tensor(x0, device=device)
    dd = tuple(torch.tensor(a, device=device) for a in draws)
    sign, logabs, nacc = m.mcmc_sweep(xd, H.graph.edge_tensor(device, "sweep"), 2, 0, 0,
                                      draws=dd)
    assert np.array_equal(xd.cpu().numpy(), xo)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 2e-4
    # the carried first-layer state equals a fresh forward pass
    z1 = m._state["z1"].clone()
    m.invalidate()
    st = m.walker_state(xd)
    assert torch.allclose(z1, st["z1"], atol=1e-5)


def test_nnbf_config3_shape(device):
    """6x6, 72 -> 256 -> 256 -> 2304, 678,912 parameters, BF16 GEMMs."""
    m, params = _model(6, 6, 16, 16, 256, "bf16", device, scale=0.01)
    assert m.n_params == 678912
    x = random_configs(np.random.default_rng(6), 16, 36, 16, 16)
    xd = torch.tensor(x, device=device)
    O, (s, l) = m.log_derivative(xd, dtype=torch.float32)
    assert O.shape == (16, 678912)
    so, lo, _ = onn.forward(params, x, 2, m.bf_scale, bf16_gemm=True)
    assert np.abs(l.cpu().numpy() - lo).max() < 2e-4
    assert bool(torch.isfinite(O).all())
