"""End-to-end parity of the reference-shaped drop-in path (VMC_GPU with
MetropolisExchangeSpinfulSamplerGPU, evaluate_energy, compute_grads_gpu,
MinSRGPU, SGDGPU) against the FP64 oracle's VMC loop on the same Philox
draws: identical walker configurations each step, parameters within 1e-9."""
import numpy as np
import pytest
import torch

from oracle import hubbard as ohub
from oracle import sampler as osamp
from oracle import slater as osl
from oracle import sr as osr
from vmc_torch_amd import (MetropolisExchangeSpinfulSamplerGPU, SlaterDeterminant_GPU,
                           spinful_Fermi_Hubbard_square_lattice_torch)
from vmc_torch_amd.optimizer import MinSRGPU, SGDGPU
from vmc_torch_amd.VMC import VMC_GPU, VMCLoopConfig

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Lx,Ly,nu,nd,B", [(4, 2, 4, 4, 256), (4, 4, 7, 7, 128)])
def test_vmc_loop_matches_oracle(device, Lx, Ly, nu, nd, B):
    N, nf = Lx * Ly, nu + nd
    steps, burn, seed, lr, shift = 3, 5, 99, 0.05, 1e-3
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf, n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    model = SlaterDeterminant_GPU(2 * N, nf, init_std=0.1,
                                  generator=torch.Generator().manual_seed(1)).to(device)
    M0 = model.M.detach().cpu().numpy().copy()
    rng = np.random.default_rng(2)
    x0 = np.zeros((B, N), np.int64)
    for b in range(B):
        x0[b, rng.choice(N, nu, replace=False)] += 2
        x0[b, rng.choice(N, nd, replace=False)] += 1
    fxs = torch.tensor(x0, device=device)
    sampler = MetropolisExchangeSpinfulSamplerGPU(seed=seed, refresh_every=64)
    vmc = VMC_GPU(sampler, MinSRGPU("fp64"), SGDGPU(lr))
    cfg = VMCLoopConfig(vmc_steps=steps, ns_per_rank=B, n_params=2 * N * nf, nsites=N,
                        learning_rate=lr, diag_shift=shift, burn_in_steps=burn,
                        outlier_clip_factor=0.0, use_log_amp=True)
    log = []
    vmc.run_vmc_loop(fxs, model, H, H.graph, config=cfg,
                     on_step_end=lambda d: log.append((d["fxs"].cpu().numpy().copy(),
                                                       d["energy_mean"])))
    # oracle loop
    Ho = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    edges = Ho.graph.sweep_edges()
    M = M0.copy()
    x = x0.copy()
    x, _, _, _ = osamp.sweep(lambda c: osl.slogdet(M, c), x, edges, osamp.PhiloxDraws(seed, 0), burn)
    for k in range(steps):
        x, _, _, _ = osamp.sweep(lambda c: osl.slogdet(M, c), x, edges,
                                 osamp.PhiloxDraws(seed, burn + k), 1)
        assert np.array_equal(log[k][0], x), f"step {k}: configurations differ"
        E = osl.local_energy(Ho, M, x)
        assert log[k][1] == pytest.approx(E.mean(), rel=1e-10)
        O = osl.log_derivative(M, x)
        dp = osr.minsr(O, E, E.mean(), shift)
        M = M - lr * dp.reshape(M.shape)
    assert np.allclose(model.M.detach().cpu().numpy(), M, rtol=1e-9, atol=1e-10)
