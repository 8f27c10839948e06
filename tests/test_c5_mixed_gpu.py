"""The batched learned-proposal configuration (C5) on the GPU: an uneven
batch of 4x4 (n = 1), 6x6 (n = 0.89) and 8x8 (n = 1) U = 8 Hubbard lattices,
each with an NNBF of the family 2N -> 256 -> 256 -> 2N n_f (ReLU, BF16 GEMMs
with FP32 accumulation, bf_scale 0.01), and ONE BF16 proposal net (3 x 3x3
conv, 32 channels) shared by the three.

Parity (32 walkers per lattice, three proposals, injected uniforms): the
device chain's moves, proposed configurations and accept decisions equal
those computed on the host from the net's logits (the device's own logits:
their BF16 rounding is pinned against the oracle's bfloat16 restatement, one
bfloat16 step of the logit scale, here and in test_proposal_gpu.py) and the
|psi|^2 ratios of the BF16-rounding NNBF oracle (oracle/nnbf.py,
bf16_gemm=True), with every choice uniform at least 1e-9 from a CDF boundary
and every accept uniform at least 0.5 % (relative) from its acceptance
ratio; ln|psi| within 2e-4 and E_loc within 1e-3 relative of that oracle.
Full size (8192 + 4096 + 4096 walkers): particle numbers conserved, the
E/site of the learned and the plain chains equal within 5 standard errors
of the chain means (MH with the reverse proposal keeps |psi|^2 exact)."""
import numpy as np
import pytest
import torch

from helpers import random_configs
from oracle import hubbard as ohub
from oracle import nnbf as onn
from oracle import proposal as op
from vmc_torch_amd import NNBF_GPU, kernels
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
from vmc_torch_amd.sampler import LearnedProposalSamplerGPU, MetropolisExchangeSpinfulSamplerGPU

pytestmark = pytest.mark.gpu

LATTICES = [(4, 4, 8, 8, 8192), (6, 6, 16, 16, 4096), (8, 8, 32, 32, 4096)]


def _nnbf(Lx, Ly, nu, nd, seed, device):
    m = NNBF_GPU(Lx * Ly, nu + nd, hidden_dim=256, activation="relu", bf_scale=0.01,
                 gemm="bf16", init="xavier", m_std=0.1,
                 generator=torch.Generator().manual_seed(seed)).to(device)
    params = [p.detach().to(torch.float32).double().cpu().numpy() for p in m.params]
    return m, params


def _logsoftmax(L):
    valid = np.isfinite(L)
    return op.log_q(np.where(valid, L, 0.0), valid)


def _away(u, bounds, margin, rng):
    while np.min(np.abs(bounds - u)) < margin:
        u = rng.random()
    return u


def test_c5_learned_chain_matches_oracle(device):
    net = LearnedProposalSamplerGPU.init_params(32, torch.Generator().manual_seed(0)).to(device)
    th32 = net.double().cpu().numpy()
    rng = np.random.default_rng(3)
    B, steps = 32, 3
    for n_i, (Lx, Ly, nu, nd, _) in enumerate(LATTICES):
        N = Lx * Ly
        print(f"C5 parity {Lx}x{Ly}", flush=True)
        m, params = _nnbf(Lx, Ly, nu, nd, n_i, device)
        x = random_configs(rng, B, N, nu, nd)
        xd = torch.tensor(x, device=device)
        sampler = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=net, seed=9)
        assert sampler.precision == "bf16"

        def logamp(c):
            return onn.forward(params, np.atleast_2d(c), 2, 0.01, bf16_gemm=True)[1]
        draws = []
        n_acc = np.zeros(B, np.int64)
        for step in range(steps):
            Lx_ = kernels.lp_logits(net, 32, Lx, Ly, torch.tensor(x, device=device),
                                    precision="bf16").cpu().numpy()
            # the device logits against the oracle's bfloat16 restatement
            for b in range(0, B, 8):
                want, valid, _ = op.forward(th32, 32, Lx, Ly, x[b], bf16=True)
                assert np.array_equal(np.isfinite(Lx_[b]), valid)
                sc = max(1.0, np.abs(want[valid]).max())
                assert np.abs(Lx_[b][valid] - want[valid]).max() < 2.0 ** -8 * sc
            uc, ua = np.empty(B), np.empty(B)
            xp = x.copy()
            lqf = np.empty(B)
            mv = np.empty(B, np.int64)
            for b in range(B):
                lq = _logsoftmax(Lx_[b])
                cdf = np.cumsum(np.exp(lq)[np.isfinite(lq)])
                uc[b] = _away(rng.random(), cdf, 1e-9, rng)
                mv[b] = op.sample_move(lq, uc[b])
                xp[b] = op.apply_move(x[b], Lx, Ly, mv[b])
                lqf[b] = lq[mv[b]]
            Lp = kernels.lp_logits(net, 32, Lx, Ly, torch.tensor(xp, device=device),
                                   precision="bf16").cpu().numpy()
            lqr = np.array([_logsoftmax(Lp[b])[mv[b]] for b in range(B)])
            r = np.exp(2.0 * (logamp(xp) - logamp(x)) + lqr - lqf)
            for b in range(B):
                ua[b] = _away(rng.random(), np.array([r[b]]), 5e-3 * r[b], rng)
            acc = ua < r
            x = np.where(acc[:, None], xp, x)
            n_acc += acc
            draws.append((torch.tensor(uc, device=device), torch.tensor(ua, device=device)))
        sign, logabs = sampler.sweeps(xd, m, steps, draws=draws)
        assert np.array_equal(xd.cpu().numpy(), x)
        assert np.array_equal(sampler.last_acceptance[0].cpu().numpy(), n_acc)
        assert n_acc.sum() > 0
        so, lo, _ = onn.forward(params, x, 2, 0.01, bf16_gemm=True)
        assert np.array_equal(sign.cpu().numpy(), so)
        assert np.abs(logabs.cpu().numpy() - lo).max() < 2e-4
        Hd = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                        n_fermions_per_spin=(nu, nd))
        E = m.local_energy(xd, Hd).cpu().numpy()
        Eo = onn.local_energy(ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd), params, x, 2, 0.01,
                              bf16_gemm=True, chunk=8)
        assert np.allclose(E, Eo, rtol=1e-3, atol=1e-3 * np.abs(Eo).max())


def test_c5_full_size_learned_vs_plain(device):
    net = LearnedProposalSamplerGPU.init_params(32, torch.Generator().manual_seed(1)).to(device)
    burn, T = 30, 24
    for n_i, (Lx, Ly, nu, nd, B) in enumerate(LATTICES):
        N = Lx * Ly
        print(f"C5 full size {Lx}x{Ly} B={B}", flush=True)
        m, _ = _nnbf(Lx, Ly, nu, nd, n_i, device)
        H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                       n_fermions_per_spin=(nu, nd))
        H.precompute_hops_gpu(device)
        x0 = torch.tensor(random_configs(np.random.default_rng(40 + n_i), B, N, nu, nd),
                          device=device)
        k = max(1, H.graph.edge_tensor(device, "sweep").shape[0] // 4)
        res = {}
        for form in ("plain", "learned"):
            x = x0.clone()
            if form == "plain":
                s = MetropolisExchangeSpinfulSamplerGPU(seed=11 + n_i)

                def adv(n):
                    s.sweeps(x, m, H.graph, n)
            else:
                s = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=net, seed=21 + n_i)

                def adv(n):
                    s.sweeps(x, m, n * k)
            adv(burn)
            series = []
            for _ in range(T):
                adv(1)
                series.append(m.local_energy(x, H) / N)
            xc = x.cpu().numpy()
            assert np.all(((xc >> 1) & 1).sum(1) == nu) and np.all((xc & 1).sum(1) == nd)
            S = torch.stack(series, 1)
            cm = S.mean(1)
            res[form] = (float(cm.mean()), float(cm.std()) / np.sqrt(B))
        (ep, sp), (el, sl) = res["plain"], res["learned"]
        print(f"  {Lx}x{Ly}: plain {ep:.5f} +- {sp:.5f}, learned {el:.5f} +- {sl:.5f}",
              flush=True)
        assert abs(ep - el) < 5.0 * np.hypot(sp, sl)
