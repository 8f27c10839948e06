"""The autotuner on the GPU: every launch width gives bit-identical walker
results (each walker's arithmetic does not depend on how walkers are packed
into workgroups), tune_slater installs the fastest width, and its cache
file reinstalls it."""
import numpy as np
import pytest
import torch

from helpers import random_configs
from vmc_torch_amd import SlaterDeterminant_GPU, autotune, kernels
from vmc_torch_amd.hamiltonian import spinful_Fermi_Hubbard_square_lattice_torch

pytestmark = pytest.mark.gpu


def _case(device, dtype, Lx=6, Ly=6, nu=16, nd=16, B=777):
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    m = SlaterDeterminant_GPU(2 * Lx * Ly, nu + nd, dtype=dtype, init_std=0.2,
                              generator=torch.Generator().manual_seed(5)).to(device)
    x = torch.tensor(random_configs(np.random.default_rng(6), B, Lx * Ly, nu, nd),
                     device=device)
    return H, m, x


def _run(m, H, x0, device):
    """refresh, 2 sweeps, local energies: everything the three kinds touch"""
    nf = m.n_fermions
    M64 = m.M.detach().to(torch.float64)
    M32 = m.M.detach() if m.M.dtype == torch.float32 else None
    x = x0.clone()
    st = kernels.SlaterWalkerState.empty(x.shape[0], x.shape[1], nf, device, dtype=m.M.dtype)
    kernels.slater_refresh(M64, x, st)
    kernels.slater_sweep(M64, H.graph.edge_tensor(device, "sweep"), x, st, n_sweeps=2, seed=9,
                         refresh_every=0, M32=M32)
    E = kernels.slater_local_energy(M64, H.graph.edge_tensor(device, "graph"), H.t, H.U, x, st,
                                    M32=M32)
    torch.cuda.synchronize(device)
    return x, st, E


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_every_width_is_bit_identical(device, dtype):
    H, m, x0 = _case(device, dtype)
    nfp = kernels.slater_nfp(m.n_fermions)
    fp32 = dtype == torch.float32
    ref = _run(m, H, x0, device)
    try:
        for w in (1, 2, 8):
            for kind in autotune.KINDS:
                autotune.set_waves(kind, nfp, fp32, w)
            x, st, E = _run(m, H, x0, device)
            assert torch.equal(x, ref[0]), w
            assert torch.equal(st.ainv, ref[1].ainv) and torch.equal(st.logabs, ref[1].logabs)
            assert torch.equal(st.n_accept, ref[1].n_accept)
            assert torch.equal(E, ref[2]), w
    finally:
        for kind in autotune.KINDS:
            autotune.set_waves(kind, nfp, fp32, 0)


def test_tune_installs_fastest_and_cache_reinstalls(device, tmp_path):
    H, m, x0 = _case(device, torch.float64, B=2048)
    nfp = kernels.slater_nfp(m.n_fermions)
    cache = str(tmp_path / "widths.json")
    x_before = x0.clone()
    try:
        res = autotune.tune_slater(m, x0, H, repeats=3, cache=cache)
        assert torch.equal(x0, x_before)             # the caller's walkers untouched
        for kind, r in res.items():
            assert r["waves"] == min(r["ms"], key=r["ms"].get)
            assert autotune.get_waves(kind, nfp, False) == r["waves"]
            print(f"  {kind}: " + ", ".join(f"{w} waves {t:.3f} ms" for w, t in r["ms"].items()),
                  flush=True)
        for kind in autotune.KINDS:
            autotune.set_waves(kind, nfp, False, 0)
        assert autotune.load_cache(cache, device) == 3
        for kind, r in res.items():
            assert autotune.get_waves(kind, nfp, False) == r["waves"]
    finally:
        for kind in autotune.KINDS:
            autotune.set_waves(kind, nfp, False, 0)
