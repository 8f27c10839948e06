"""GPU parity of the NNBF path (libvmcx through vmc_torch_amd.NNBF_GPU)
against the FP64 oracle (oracle/nnbf.py).

The MLP runs in FP64 (gemm="fp64", the reference's float64 model), FP32
(gemm="fp32") or BF16 with FP32 accumulation (gemm="bf16"); the determinant
is FP64.  Tolerances, stated per test:
  FP64 MLP : ln|psi| 1e-10 absolute, O and E_loc 1e-9 relative (the same
             arithmetic in another order), sweeps bit-exact in
             configurations and signs;
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


def test_nnbf_8x8_bf16_target_shape(device):
    """The target's shape: 8x8 U=8, 32 up + 32 down, NNBF 128 -> 256 -> 256
    -> 8192 (ReLU, bf_scale 0.01, Xavier, M ~ N(0, 0.1^2)), BF16 GEMMs with
    FP32 accumulation.  Against the oracle with the same BF16 operand
    rounding (64 walkers): identical signs, ln|psi| within 2e-4, E_loc within
    1e-3 relative.  Against the FP64 oracle (the FP64 weights, no rounding)
    on 1024 walkers sampled from |psi|^2 (30 sweeps of the plain sampler
    from random configurations): the energy estimate per site within 1e-4
    (the mixed-precision form's accuracy bound); the per-sample differences are printed.
    (On unsampled random configurations near a node of psi a single E_loc can
    move by O(1): the bound is on the estimate over sampled walkers.)"""
    Lx, Ly, nu, nd = 8, 8, 32, 32
    N = 64
    m = NNBF_GPU(N, nu + nd, hidden_dim=256, activation="relu", bf_scale=0.01, gemm="bf16",
                 init="xavier", m_std=0.1, generator=torch.Generator().manual_seed(3)).to(device)
    p64 = [p.detach().cpu().numpy().copy() for p in m.params]
    p32 = [p.detach().to(torch.float32).double().cpu().numpy() for p in m.params]
    Hd = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                    n_fermions_per_spin=(nu, nd))
    Ho = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    xd = torch.tensor(random_configs(np.random.default_rng(30), 1024, N, nu, nd), device=device)
    m.mcmc_sweep(xd, Hd.graph.edge_tensor(device, "sweep"), 30, 7, 0)
    x = xd.cpu().numpy()
    s, l = m.forward_log(xd[:64])
    so, lo, _ = onn.forward(p32, x[:64], 2, 0.01, bf16_gemm=True)
    assert np.array_equal(s.cpu().numpy(), so)
    assert np.abs(l.cpu().numpy() - lo).max() < 2e-4
    E = m.local_energy(xd, Hd).cpu().numpy()
    Eb = onn.local_energy(Ho, p32, x[:64], 2, 0.01, bf16_gemm=True, chunk=16)
    assert np.allclose(E[:64], Eb, rtol=1e-3, atol=1e-3 * np.abs(Eb).max())
    E64 = onn.local_energy(Ho, p64, x, 2, 0.01, chunk=64)
    d = np.abs(E - E64) / N
    dE = abs(E.mean() - E64.mean()) / N
    print(f"  |dE/site| vs FP64 oracle over {len(x)} sampled walkers: {dE:.2e} "
          f"(per sample: mean {d.mean():.2e}, median {np.median(d):.2e}, max {d.max():.2e})",
          flush=True)
    assert dE < 1e-4



@pytest.mark.parametrize("case", CASES)
def test_nnbf_fp64_matches_oracle(device, case):
    """gemm="fp64": every layer in FP64, so the device equals the FP64 oracle
    to rounding: ln|psi| within 1e-10, O (dense and as factors multiplied
    out) within 1e-9 of its largest entry per parameter block, E_loc within
    1e-9 relative, and a tie-free sweep bit-exact."""
    Lx, Ly, nu, nd, hid = case
    m, _ = _model(Lx, Ly, nu, nd, hid, "fp64", device)
    assert m.gemm == "fp64" and m.z_dtype == torch.float64
    params = [p.detach().cpu().numpy() for p in m.params]
    N = Lx * Ly
    x = random_configs(np.random.default_rng(11), 32, N, nu, nd)
    xd = torch.tensor(x, device=device)
    s, l = m.forward_log(xd)
    so, lo, _ = onn.forward(params, x, 2, m.bf_scale)
    assert np.array_equal(s.cpu().numpy(), so)
    assert np.abs(l.cpu().numpy() - lo).max() < 1e-10
    O, _ = m.log_derivative(xd, dtype=torch.float64)
    FO, _ = m.log_derivative_factored(xd)
    Oo = onn.log_derivative(params, x, 2, m.bf_scale)
    for got in (O.cpu().numpy(), FO.materialize().cpu().numpy()):
        off = 0
        for p in params:
            blk = slice(off, off + p.size)
            sc = max(np.abs(Oo[:, blk]).max(), 1e-300)
            assert np.abs(got[:, blk] - Oo[:, blk]).max() <= 1e-9 * sc, p.shape
            off += p.size
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    E = m.local_energy(xd, H).cpu().numpy()
    Eo = onn.local_energy(ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd), params, x, 2, m.bf_scale)
    assert np.allclose(E, Eo, rtol=1e-9, atol=1e-9 * np.abs(Eo).max())
    edges = H.graph.sweep_edges()
    rng = np.random.default_rng(12)
    x0 = random_configs(rng, 32, N, nu, nd)
    draws, xo, so, lo = tie_free_draws(
        lambda c: onn.forward(params, c, 2, m.bf_scale)[:2], x0, edges, 2, rng)
    xd0 = torch.tensor(x0, device=device)
    dd = tuple(torch.tensor(a, device=device) for a in draws)
    sign, logabs, _ = m.mcmc_sweep(xd0, H.graph.edge_tensor(device, "sweep"), 2, 0, 0,
                                   draws=dd)
    assert np.array_equal(xd0.cpu().numpy(), xo)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-10
    z1 = m._state["z1"].clone()
    m.invalidate()
    assert torch.allclose(z1, m.walker_state(xd0)["z1"], rtol=0, atol=1e-12)


def test_nnbf_default_is_fp64():
    """The reference's NNBF_GPU defaults to float64 (NNBF.py:36-162): so does
    the drop-in, with FP64 arithmetic unless a reduced-precision mode is
    asked for."""
    assert NNBF_GPU(4, 4, hidden_dim=8).gemm == "fp64"
    assert NNBF_GPU(4, 4, hidden_dim=8, gemm="bf16").gemm == "bf16"
    with pytest.raises(ValueError):
        NNBF_GPU(4, 4, hidden_dim=8, dtype=torch.float32, gemm="fp64")



@pytest.mark.parametrize("case", [(8, 8, 32, 32, 256), (6, 6, 16, 16, 256), (4, 4, 8, 8, 256),
                                  (6, 6, 18, 17, 64), (4, 2, 4, 4, 32), (3, 3, 4, 3, 48)])
def test_nnbf_bf16_matrix_core_layers(device, case):
    """The library's own BF16 matrix-core layers (only the occupied orbital
    rows of the output layer formed) against the rocBLAS GEMMs on the same
    BF16 operands: identical signs and ln|psi| within 1e-4 (FP32 accumulation
    in another order); against the BF16-rounding oracle, identical signs,
    ln|psi| within 2e-5 on average and 2e-3 at most.  The oracle rounds
    FP64 pre-activations to BF16 where the GPU rounds FP32 ones: an
    activation that lies within FP32 rounding of a BF16 rounding boundary
    rounds the other way (one BF16 ulp, 2^-8 relative), which moves one
    configuration's ln|psi| by up to a few 1e-4 at bf_scale 0.3 on 8x8.
    Shapes cover n_f a multiple of 4 and not (35), hidden widths 256 / 64 /
    48 / 32 and a last partial block of 64 configurations."""
    Lx, Ly, nu, nd, hid = case
    m, params = _model(Lx, Ly, nu, nd, hid, "bf16", device, scale=0.3)
    m_lib = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=hid, activation="relu", bf_scale=0.3,
                     gemm="bf16", bf16_layers="rocblas").to(device)
    with torch.no_grad():
        for a, b in zip(m_lib.params, m.params):
            a.copy_(b)
    x = random_configs(np.random.default_rng(40), 100, Lx * Ly, nu, nd)
    xd = torch.tensor(x, device=device)
    s1, l1 = m.forward_log(xd)
    s2, l2 = m_lib.forward_log(xd)
    assert torch.equal(s1, s2)
    assert (l1 - l2).abs().max().item() < 1e-4
    so, lo, _ = onn.forward(params, x, 2, m.bf_scale, bf16_gemm=True)
    assert np.array_equal(s1.cpu().numpy(), so)
    d = np.abs(l1.cpu().numpy() - lo)
    assert d.mean() < 2e-5 and d.max() < 2e-3, (float(d.mean()), float(d.max()))


@pytest.mark.parametrize("case", [(8, 8, 32, 32, 256), (6, 6, 16, 16, 256), (4, 4, 8, 8, 64),
                                  (6, 6, 18, 17, 64)])
def test_nnbf_bf16_fp32_determinant(device, case):
    """det="fp32" (the mixed-precision amplitude: FP32 elimination of entries
    formed in FP64) against the FP64 elimination of the same BF16 network on
    100 random configurations: identical signs, ln|psi| within 1e-4 on
    average and 1e-2 at most (FP32 LU: ~cond(A) x 6e-8 relative; random
    configurations at bf_scale 0.3 include badly conditioned ones); against
    the BF16-rounding oracle, where the BF16 rounding flips add theirs,
    within 5e-4 on average and 5e-2 at most.  The bound that matters for the
    mixed form, the energy over sampled walkers, is the next test's."""
    Lx, Ly, nu, nd, hid = case
    m64, params = _model(Lx, Ly, nu, nd, hid, "bf16", device, scale=0.3)
    m32 = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=hid, activation="relu", bf_scale=0.3,
                   gemm="bf16", det="fp32").to(device)
    with torch.no_grad():
        for a, b in zip(m32.params, m64.params):
            a.copy_(b)
    xd = torch.tensor(random_configs(np.random.default_rng(41), 100, Lx * Ly, nu, nd),
                      device=device)
    s1, l1 = m32.forward_log(xd)
    s2, l2 = m64.forward_log(xd)
    d = (l1 - l2).abs()
    assert torch.equal(s1, s2)
    assert d.mean().item() < 1e-4 and d.max().item() < 1e-2, (d.mean().item(), d.max().item())
    so, lo, _ = onn.forward(params, xd.cpu().numpy(), 2, m64.bf_scale, bf16_gemm=True)
    assert np.array_equal(s1.cpu().numpy(), so)
    do = np.abs(l1.cpu().numpy() - lo)
    assert do.mean() < 5e-4 and do.max() < 5e-2, (do.mean(), do.max())


def test_nnbf_8x8_mixed_form_energy(device):
    """The target's mixed form (BF16 network on the matrix-core layers, FP32
    determinants) on 1024 walkers sampled from |psi|^2: the energy estimate per
    site within 1e-4 of the FP64 oracle's (FP64 weights, FP64 everything)."""
    Lx, Ly, nu, nd = 8, 8, 32, 32
    N = 64
    m = NNBF_GPU(N, nu + nd, hidden_dim=256, activation="relu", bf_scale=0.01, gemm="bf16",
                 det="fp32", init="xavier", m_std=0.1,
                 generator=torch.Generator().manual_seed(3)).to(device)
    p64 = [p.detach().cpu().numpy().copy() for p in m.params]
    Hd = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                    n_fermions_per_spin=(nu, nd))
    Ho = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    xd = torch.tensor(random_configs(np.random.default_rng(30), 1024, N, nu, nd), device=device)
    m.mcmc_sweep(xd, Hd.graph.edge_tensor(device, "sweep"), 30, 7, 0)
    m.invalidate()
    E = m.local_energy(xd, Hd).cpu().numpy()
    E64 = onn.local_energy(Ho, p64, xd.cpu().numpy(), 2, 0.01, chunk=64)
    d = np.abs(E - E64) / N
    dE = abs(E.mean() - E64.mean()) / N
    print(f"  mixed form |dE/site| vs FP64 oracle over 1024 sampled walkers: {dE:.2e} "
          f"(per sample: mean {d.mean():.2e}, max {d.max():.2e})", flush=True)
    assert dE < 1e-4


def _neighbour_swaps(x, Lx, Ly, rng):
    """Each row with the contents of two neighbouring sites exchanged (a hop,
    an up/down exchange, or no change when the two sites hold the same),
    as the sampler's moves do."""
    xp = x.copy()
    for b in range(x.shape[0]):
        s = int(rng.integers(Lx * Ly))
        i, j = divmod(s, Ly)
        nb = [(i + di) * Ly + (j + dj) for di, dj in ((1, 0), (0, 1), (-1, 0), (0, -1))
              if 0 <= i + di < Lx and 0 <= j + dj < Ly]
        t = nb[int(rng.integers(len(nb)))]
        xp[b, s], xp[b, t] = x[b, t], x[b, s]
    return xp


MOVED_CASES = [("fp64", (4, 4, 5, 4, 32)), ("fp32", (6, 6, 16, 16, 64)),
               ("bf16", (8, 8, 32, 32, 256)), ("bf16_fp32det", (8, 8, 32, 32, 256)),
               ("bf16", (6, 6, 18, 17, 32))]


@pytest.mark.parametrize("kind,case", MOVED_CASES)
def test_nnbf_forward_moved(device, kind, case):
    """Amplitudes of proposed configurations from the walkers' first-layer
    pre-activations (forward_log_moved) against the full forward of the
    proposed configurations.  FP64: ln|psi| within 1e-10, z1 within 1e-12;
    FP32 network: z1 within 1e-5, ln|psi| within 1e-4; BF16 layers: the
    first layer in FP32 differs in its last bits from the full sum, which
    can move a BF16 rounding of h1 -- ln|psi| mean difference below 2e-5,
    maximum below 2e-3 (the tolerance of the matrix-core layers against the
    library GEMMs).  Signs identical; z1p = first_layer(proposed) as above;
    rows without a change reproduce the walker's own amplitude."""
    Lx, Ly, nu, nd, hid = case
    gemm = "bf16" if kind.startswith("bf16") else kind
    g = torch.Generator().manual_seed(3)
    kw = dict(det="fp32") if kind == "bf16_fp32det" else {}
    m = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=hid, activation="relu", bf_scale=0.05,
                 gemm=gemm, init="xavier", m_std=0.3, generator=g, chunk=4096, **kw).to(device)
    rng = np.random.default_rng(11)
    x = random_configs(rng, 700, Lx * Ly, nu, nd)
    xp = _neighbour_swaps(x, Lx, Ly, rng)
    xd, xpd = torch.tensor(x, device=device), torch.tensor(xp, device=device)
    s0, l0, z1 = m.first_layer(xd)
    s_ref, l_ref = m.forward_log(xd)
    assert torch.equal(s0, s_ref) and torch.equal(l0, l_ref)
    sp, lp, z1p = m.forward_log_moved(xd, xpd, z1)
    sf, lf, z1f = m.first_layer(xpd)
    assert torch.equal(sp, sf)
    dl = (lp - lf).abs().cpu().numpy()
    dz = (z1p - z1f).abs().max().item()
    same = np.all(x == xp, axis=1)
    assert same.any() and (~same).any()
    if kind == "fp64":
        assert dl.max() < 1e-10 and dz < 1e-12
    elif kind == "fp32":
        assert dl.max() < 1e-4 and dz < 1e-5
    else:
        assert dz < 1e-5
        assert dl.mean() < 2e-5 and dl.max() < 2e-3
    # unchanged rows: the walker's own pre-activations and amplitude
    sm = torch.tensor(same, device=device)
    assert torch.equal(z1p[sm], z1[sm])
    assert torch.equal(lp[sm], l0[sm]) and torch.equal(sp[sm], s0[sm])


def test_masked_rows(device):
    from vmc_torch_amd import kernels
    src = torch.randn(37, 19, dtype=torch.float64, device=device)
    dst = torch.randn(37, 19, dtype=torch.float64, device=device)
    mask = (torch.arange(37, device=device) % 3 == 0).to(torch.uint8)
    ref = torch.where(mask.bool()[:, None], src, dst)
    kernels.masked_rows(mask, src, dst)
    assert torch.equal(dst, ref)
