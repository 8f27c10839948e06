"""GPU parity of minSR on a factored O (vmcx_sr_minsr_factored) and of the
NNBF factor output (vmcx_nnbf_logderiv_factors).

Tolerances:
  * factored solve vs the oracle's dense minSR and vs vmcx_sr_minsr (FP64
    Gram) on the same O: 1e-9 relative (FP64 throughout, different
    summation order);
  * NNBF factors vs the dense O of vmcx_nnbf_logderiv (FP64 output of FP32
    arithmetic): the dense kernel rounds each outer product (and bf_scale gM)
    to FP32, the factored form multiplies the FP32 factors in FP64, so
    entries agree within 4 FP32 roundings (5e-7 relative); sign and ln|psi|
    are bit-identical (same forward);
  * VMC loop with the factored SR step vs the dense FP64 one fed the same O
    (the factors multiplied out; the driver's own dense O rounds each outer
    product to FP32, a ~6e-8 change that the solve amplifies by cond(T)):
    configurations identical, energies 1e-9 and parameters 1e-8 relative
    after three steps."""
import numpy as np
import pytest
import torch

from helpers import random_configs
from oracle import sr as osr
from vmc_torch_amd import NNBF_GPU, MetropolisExchangeSpinfulSamplerGPU, kernels
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
from vmc_torch_amd.optimizer import MinSRGPU, SGDGPU
from vmc_torch_amd.VMC import VMC_GPU, VMCLoopConfig

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


def _random_factored(Ns, k0, seed, device):
    """Factors of widths (k0, 5, 3, 16) side by side in one (Ns, width)
    tensor with a row stride wider than the sum; blocks: plain, outer with
    bias, outer without bias, scaled outer with bias sharing factor 0."""
    rng = np.random.default_rng(seed)
    k = [k0, 5, 3, 16]
    cols = []
    c = 0
    for kk in k:
        cols.append((c, kk))
        c += kk
    F = torch.zeros(Ns, c + 3, dtype=torch.float64)
    for (c0, kk) in cols:
        F[:, c0:c0 + kk] = torch.tensor(rng.normal(size=(Ns, kk)) * 0.5 + rng.normal(size=kk))
    blocks, off = [], 0
    for a, b, s, bias in [(0, -1, 1.0, False), (2, 1, 1.0, True), (3, 2, 1.0, False),
                          (0, 3, 0.3, True)]:
        w = off
        off += k[a] * (k[b] if b >= 0 else 1)
        bo = -1
        if bias:
            bo = off
            off += k[a]
        blocks.append((a, b, s, w, bo))
    FO = kernels.FactoredO(F.to(device)[:, :c], cols, blocks, off)
    E = rng.normal(size=Ns) * 2.0 + 3.0
    return FO, E


@pytest.mark.parametrize("Ns,k0", [(1, 7), (129, 300), (300, 4096), (1000, 64)])
def test_factored_minsr_matches_oracle_and_dense(device, Ns, k0):
    print(f"factored minSR Ns={Ns} k0={k0}", flush=True)
    FO, E = _random_factored(Ns, k0, Ns + k0, device)
    O = FO.materialize()
    Ed = torch.tensor(E, device=device)
    want = osr.minsr(O.cpu().numpy(), E, E.mean(), 1e-3)
    dp, stats, info = kernels.sr_minsr_factored(FO, Ed, None, 1e-3)
    assert int(info.item()) == 0
    assert _rel(dp.cpu().numpy(), want) < 1e-9
    dpd, statsd, _ = kernels.sr_minsr(O.contiguous(), Ed, None, 1e-3, "fp64")
    assert _rel(dp.cpu().numpy(), dpd.cpu().numpy()) < 1e-9
    s, sd = stats.cpu().numpy(), statsd.cpu().numpy()
    assert s[0] == pytest.approx(sd[0], rel=1e-14)
    assert s[1] == pytest.approx(sd[1], rel=1e-12, abs=1e-14)
    assert s[2] < 1e-10
    assert s[3] == pytest.approx(sd[3], rel=1e-8, abs=1e-12)
    assert s[4] == 0


def test_factored_minsr_given_energy_mean(device):
    FO, E = _random_factored(200, 40, 5, device)
    O = FO.materialize().cpu().numpy()
    want = osr.minsr(O, E, 2.5, 1e-2)
    dp, _, _ = kernels.sr_minsr_factored(FO, torch.tensor(E, device=device), 2.5, 1e-2)
    assert _rel(dp.cpu().numpy(), want) < 1e-9


CASES = [(4, 2, 4, 4, 32), (4, 4, 5, 4, 32), (8, 8, 32, 32, 32)]


def _model(Lx, Ly, nu, nd, hidden, gemm, device, seed=0):
    g = torch.Generator().manual_seed(seed)
    m = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=hidden, activation="relu", bf_scale=0.3,
                 gemm=gemm, init="xavier", m_std=0.3, generator=g, chunk=4096)
    with torch.no_grad():
        for p in m.params:
            if p.dim() == 1:
                p.copy_(0.05 * torch.randn(p.shape, generator=g, dtype=p.dtype))
    return m.to(device)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("gemm", ["fp32", "bf16"])
def test_nnbf_factors_match_dense_O(device, case, gemm):
    Lx, Ly, nu, nd, hid = case
    print(f"nnbf factors {case} {gemm}", flush=True)
    m = _model(Lx, Ly, nu, nd, hid, gemm, device)
    x = torch.tensor(random_configs(np.random.default_rng(4), 40, Lx * Ly, nu, nd),
                     device=device)
    O, (s, l) = m.log_derivative(x, dtype=torch.float64)
    FO, (sf, lf) = m.log_derivative_factored(x)
    assert torch.equal(s, sf) and torch.equal(l, lf)
    assert FO.shape == tuple(O.shape)
    Of = FO.materialize()
    err = (Of - O).abs()
    assert bool((err <= 5e-7 * O.abs() + 1e-300).all()), float(err.max())


def test_nnbf_factored_sr_step(device):
    Lx, Ly, nu, nd, hid = 4, 4, 7, 7, 32
    m = _model(Lx, Ly, nu, nd, hid, "bf16", device, seed=3)
    x = torch.tensor(random_configs(np.random.default_rng(5), 512, Lx * Ly, nu, nd),
                     device=device)
    FO, _ = m.log_derivative_factored(x)
    E = torch.tensor(np.random.default_rng(6).normal(size=512) + 1.0, device=device)
    ph = []
    dp, stats, info = kernels.sr_minsr_factored(FO, E, None, 1e-3, phase_ms=ph)
    assert int(info.item()) == 0 and len(ph) == 6 and all(v >= 0 for v in ph)
    dpd, statsd, _ = kernels.sr_minsr(FO.materialize(), E, None, 1e-3, "fp64")
    assert _rel(dp.cpu().numpy(), dpd.cpu().numpy()) < 1e-9
    assert stats[3].item() == pytest.approx(statsd[3].item(), rel=1e-8)


@pytest.mark.parametrize("kind,clip", [("nnbf", 0.0), ("nnftns", 0.0), ("nnbf", 1.3),
                                       ("nnftns", 1.3)])
def test_vmc_loop_factored_matches_dense(device, kind, clip):
    Lx, Ly, nu, nd, hid, B = 4, 2, 3, 3, 16, 256
    N, nf = Lx * Ly, nu + nd
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    x0 = torch.tensor(random_configs(np.random.default_rng(8), B, N, nu, nd), device=device)
    runs = []

    def dense_of_factors(fxs, model, **kw):
        FO, aux = model.log_derivative_factored(fxs)
        return FO.materialize(), aux
    for gram in ("fp64", "factored"):
        if kind == "nnbf":
            model = _model(Lx, Ly, nu, nd, hid, "fp32", device, seed=9)
        else:
            model = _peps(Lx, Ly, 2, hid, 9, device)
        sampler = MetropolisExchangeSpinfulSamplerGPU(seed=11, refresh_every=64)
        vmc = VMC_GPU(sampler, MinSRGPU(gram), SGDGPU(0.05),
                      compute_grads_fn=dense_of_factors)
        cfg = VMCLoopConfig(vmc_steps=3, ns_per_rank=B, n_params=model.n_params, nsites=N,
                            learning_rate=0.05, diag_shift=1e-3, burn_in_steps=4,
                            outlier_clip_factor=clip, use_log_amp=True)
        log = []
        vmc.run_vmc_loop(x0.clone(), model, H, H.graph, config=cfg,
                         on_step_end=lambda d: log.append((d["fxs"].cpu().numpy().copy(),
                                                           d["energy_mean"],
                                                           d["n_outliers"])))
        runs.append((log, [p.detach().cpu().numpy().copy() for p in model.params]))
    (ld, pd), (lf, pf) = runs
    for k in range(3):
        assert np.array_equal(ld[k][0], lf[k][0]), f"step {k}: configurations differ"
        assert lf[k][1] == pytest.approx(ld[k][1], rel=1e-9)
        # the outlier mask drops the same samples from the dense and the
        # factored O (GPU/VMC.py:476-498)
        assert lf[k][2] == ld[k][2]
        if clip > 0:
            assert lf[k][2] > 0
    for a, b in zip(pf, pd):
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()


def _peps(Lx, Ly, D, H, seed, device):
    from oracle import fpeps as fp
    from vmc_torch_amd import NNfTNS_Model_GPU, fPEPS_Model_GPU
    sh = fp.PEPSShape(Lx, Ly, D)
    rng = np.random.default_rng(seed)
    th = fp.random_theta(sh, rng)
    if H == 0:
        return fPEPS_Model_GPU(Lx, Ly, D, theta=th).to(device)
    nn = fp.mlp_init(rng, 2 * sh.N, H, sh.n_params, 0.2)
    m = NNfTNS_Model_GPU(Lx, Ly, D, hidden_dim=H, nn_eta=0.8, activation="relu", theta=th)
    with torch.no_grad():
        for p, v in zip(m.params[1:], nn):
            p.copy_(torch.as_tensor(v))
    return m.to(device)


@pytest.mark.parametrize("H", [0, 16])
def test_peps_factors_match_dense_O(device, H):
    """fPEPS / NN-fTNS (FP64 throughout): the dense O and the factors
    multiplied out agree to FP64 rounding; the factored SR step equals the
    dense one on the same O within 1e-7 relative.  The factored form centres
    the Gram after forming it, so its FP64 rounding is relative to the
    uncentred |O O^T| (large column means here), and with Np = 96 < Ns = 300
    (plain fPEPS) T is rank-deficient and lifted by the shift alone, whose
    condition number carries that rounding to ~1e-8 in dp."""
    print(f"peps factors H={H}", flush=True)
    m = _peps(4, 2, 2, H, 12, device)
    x = torch.tensor(random_configs(np.random.default_rng(13), 300, 8, 3, 3), device=device)
    O, (s, l) = m.log_derivative(x, dtype=torch.float64)
    FO, (sf, lf) = m.log_derivative_factored(x)
    assert torch.equal(s, sf) and torch.equal(l, lf)
    assert FO.shape == tuple(O.shape)
    Of = FO.materialize()
    assert float((Of - O).abs().max()) <= 1e-13 * float(O.abs().max())
    E = torch.tensor(np.random.default_rng(14).normal(size=300) + 1.0, device=device)
    dp, _, info = kernels.sr_minsr_factored(FO, E, None, 1e-3)
    dpd, _, _ = kernels.sr_minsr(Of, E, None, 1e-3, "fp64")
    assert int(info.item()) == 0
    assert _rel(dp.cpu().numpy(), dpd.cpu().numpy()) < 1e-7
