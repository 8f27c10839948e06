"""Captured sampling graphs (vmc_torch_amd.graphs) against the eager calls
they record, on the GPU.  Every replay is compared bit for bit with the eager
sequence on copies of the same walkers (same kernels, same inputs, draws from
the same Philox counters); replays are interleaved with eager work and host
synchronisations, and several replays are checked, so a graph node that
misbehaves only on later replays is caught."""
import numpy as np
import pytest
import torch

from helpers import random_configs
from vmc_torch_amd import NNBF_GPU, SlaterDeterminant_GPU, kernels
from vmc_torch_amd.graphs import LearnedProposalGraph, SlaterSampleGraph
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch
from vmc_torch_amd.sampler import LearnedProposalSamplerGPU

pytestmark = pytest.mark.gpu


def _H(Lx, Ly, nu, nd, device):
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    return H


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_slater_sample_graph_matches_eager(device, dtype):
    Lx, Ly, nu, nd, B, n_sw = 6, 6, 16, 16, 512, 2
    N, nf = 36, 32
    H = _H(Lx, Ly, nu, nd, device)
    m = SlaterDeterminant_GPU(2 * N, nf, dtype=dtype, init_std=0.2,
                              generator=torch.Generator().manual_seed(1)).to(device)
    x0 = torch.tensor(random_configs(np.random.default_rng(2), B, N, nu, nd), device=device)
    fx = x0.clone()
    g = SlaterSampleGraph(m, fx, H, n_sweeps=n_sw, seed=77, sweep_offset=5)
    # eager twin on its own buffers
    xe = x0.clone()
    st = kernels.SlaterWalkerState.empty(B, N, nf, device, dtype=dtype)
    M64 = m.M.detach().to(torch.float64)
    M32 = m.M.detach() if dtype == torch.float32 else None
    kernels.slater_refresh(M64, xe, st)
    es, eg = H.graph.edge_tensor(device, "sweep"), H.graph.edge_tensor(device, "graph")
    for k in range(4):
        E, O = g.replay()
        kernels.slater_sweep(M64, es, xe, st, n_sweeps=n_sw, seed=77, sweep_offset=5 + k * n_sw,
                             refresh_every=64, M32=M32)
        Ee = kernels.slater_local_energy(M64, eg, 1.0, 8.0, xe, st, M32=M32)
        Oe = kernels.slater_logderiv(st, N, nf)
        torch.cuda.synchronize(device)
        assert torch.equal(fx, xe), f"replay {k}: configurations"
        assert torch.equal(E, Ee) and torch.equal(O, Oe), f"replay {k}: E_loc / O"
        assert torch.equal(g.state.logabs, st.logabs) and torch.equal(g.state.sign, st.sign)
        _ = (xe.float() * 2).sum().item()        # unrelated eager work between replays
    assert int(g.counter.item()) == 5 + 4 * n_sw


def test_learned_proposal_graph_two_lattices(device):
    """One graph over two lattice instances (4x4 Slater, 6x6 NNBF with the
    BF16 matrix-core layers) sharing one BF16 proposal net, 3 and 5
    proposals per replay, against the eager sampler."""
    net = LearnedProposalSamplerGPU.init_params(32, torch.Generator().manual_seed(3)).to(device)
    specs = [(4, 4, 8, 8, 300, 3), (6, 6, 16, 16, 200, 5)]
    inst, twins = [], []
    for i, (Lx, Ly, nu, nd, B, k) in enumerate(specs):
        N = Lx * Ly
        if i == 0:
            mk = lambda: SlaterDeterminant_GPU(2 * N, nu + nd, init_std=0.3,   # noqa: E731
                                               generator=torch.Generator().manual_seed(4))
        else:
            mk = lambda: NNBF_GPU(N, nu + nd, hidden_dim=256, activation="relu",  # noqa: E731
                                  bf_scale=0.05, gemm="bf16", init="xavier", m_std=0.1,
                                  generator=torch.Generator().manual_seed(5))
        x0 = torch.tensor(random_configs(np.random.default_rng(10 + i), B, N, nu, nd),
                          device=device)
        s1 = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=net, seed=30 + i)
        s2 = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=net, seed=30 + i)
        inst.append((s1, mk().to(device), x0.clone()))
        twins.append((s2, mk().to(device), x0.clone(), k))
    g = LearnedProposalGraph(inst, [k for *_, k in specs])
    for rep in range(3):
        out = g.replay()
        for i, ((s2, m2, x2, k), (s1, m1, x1)) in enumerate(zip(twins, inst)):
            sg, la = s2.sweeps(x2, m2, k)
            torch.cuda.synchronize(device)
            assert torch.equal(x1, x2), f"replay {rep}, instance {i}: configurations"
            assert torch.equal(out[i][0], sg) and torch.equal(out[i][1], la)
        _ = sum(float(x.sum()) for *_, x in inst)     # eager work between replays
    n_acc = [int(I["n_acc"].sum()) for I in g.inst]
    assert all(n > 0 for n in n_acc)
    g.sync_counters()
    assert inst[0][0].step_counter == 3 * 3 and inst[1][0].step_counter == 3 * 5
