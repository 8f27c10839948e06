"""GPU parity of the Slater walker kernels against the FP64 oracle, called
through the C-ABI (vmc_torch_amd.kernels -> libvmcx.so).

Tolerances: integer outputs (configurations, slots, signs, connected
configurations, batch ids, accept decisions) must be identical; floating
outputs agree with the FP64 oracle to the relative tolerance stated in each
test (FP64 kernels: 1e-9 or tighter)."""
import numpy as np
import pytest
import torch

from helpers import edges_tensor, setup_case
from oracle import sampler as osamp
from oracle import slater as osl
from vmc_torch_amd import kernels

pytestmark = pytest.mark.gpu

SHAPES = [  # Lx, Ly, n_up, n_dn -> NFP 8, 16, 16, 32, 64, 8
    (4, 2, 4, 4), (4, 4, 8, 8), (4, 4, 7, 7), (6, 6, 16, 16), (8, 8, 32, 32), (3, 3, 4, 3)]


def _state(xd, N, nf):
    return kernels.SlaterWalkerState.empty(xd.shape[0], N, nf, xd.device)


@pytest.mark.parametrize("shape", SHAPES)
def test_refresh_matches_slogdet_and_inverse(device, shape):
    Lx, Ly, nu, nd = shape
    H, M, x, Md, xd, _ = setup_case(Lx, Ly, nu, nd, 96, seed=11, std=0.1)
    N, nf = Lx * Ly, nu + nd
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    torch.cuda.synchronize()
    s, l = osl.slogdet(M, x)
    assert np.array_equal(st.sign.cpu().numpy(), s)
    assert np.allclose(st.logabs.cpu().numpy(), l, rtol=1e-12, atol=1e-11)
    A = M[osl.occupied_rows(x)]
    inv = np.linalg.inv(A)                                   # (B, nf, nf)
    got = st.ainv.cpu().numpy()[:, :nf, :nf].transpose(0, 2, 1)  # [b][j][k]
    scale = np.abs(inv).max(axis=(1, 2), keepdims=True)
    assert np.abs(got - inv).max() / scale.max() < 1e-10
    # slot map: sorted position of every occupied orbital, 0xFF when empty
    occ = osl.occupied_rows(x)
    want = np.full((x.shape[0], 2 * N), 255)
    for b in range(x.shape[0]):
        want[b, occ[b]] = np.arange(nf)
    assert np.array_equal(st.slot.cpu().numpy(), want)
    # log-amplitude only (no walker state): LU instead of Gauss-Jordan, the
    # same signs and ln|det| within 1e-12 relative
    s2, l2 = kernels.slater_refresh(Md, xd)
    assert torch.equal(s2, st.sign)
    assert torch.allclose(l2, st.logabs, rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("shape", SHAPES)
def test_local_energy_matches_oracle(device, shape):
    Lx, Ly, nu, nd = shape
    H, M, x, Md, xd, _ = setup_case(Lx, Ly, nu, nd, 64, seed=12, std=0.1)
    N, nf = Lx * Ly, nu + nd
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    e = edges_tensor(H.graph.edges())
    E = kernels.slater_local_energy(Md, e, H.t, H.U, xd, st).cpu().numpy()
    Eo = osl.local_energy(H, M, x)
    assert np.allclose(E, Eo, rtol=1e-9, atol=1e-9 * np.abs(Eo).max())


@pytest.mark.parametrize("shape", SHAPES[:5])
def test_get_conn_matches_oracle(device, shape):
    Lx, Ly, nu, nd = shape
    H, M, x, Md, xd, rng = setup_case(Lx, Ly, nu, nd, 300, seed=13)
    e = edges_tensor(H.graph.edges())
    etas, coeffs, bids = kernels.hubbard_get_conn(xd, e, H.t, H.U)
    oe, oc, ob = H.get_conn_batch(x)
    assert np.array_equal(etas.cpu().numpy(), oe)
    assert np.array_equal(coeffs.cpu().numpy(), oc)
    assert np.array_equal(bids.cpu().numpy(), ob)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("refresh_every", [0, 3])
def test_sweep_injected_draws_bit_exact(device, shape, refresh_every):
    Lx, Ly, nu, nd = shape
    B, n_sw = 48, 3
    H, M, x, Md, xd, rng = setup_case(Lx, Ly, nu, nd, B, seed=14, std=0.3)
    N, nf = Lx * Ly, nu + nd
    edges = H.graph.sweep_edges()
    E = len(edges)
    uc = rng.random((n_sw, E, B))
    ua = rng.random((n_sw, E, B))
    bits = rng.integers(0, 4, (n_sw, E, B)).astype(np.uint8)
    xo, so, lo, nacc = osamp.sweep(lambda c: osl.slogdet(M, c), x, edges,
                                   osamp.InjectedDraws(uc, bits, ua), n_sw)
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    draws = tuple(torch.tensor(a, device=device) for a in (uc, bits, ua))
    kernels.slater_sweep(Md, edges_tensor(edges), xd, st, n_sweeps=n_sw,
                         refresh_every=refresh_every, draws=draws)
    assert np.array_equal(xd.cpu().numpy(), xo)
    assert np.array_equal(st.n_accept.cpu().numpy(), nacc)
    assert np.array_equal(st.sign.cpu().numpy(), so)
    assert np.allclose(st.logabs.cpu().numpy(), lo, rtol=1e-10, atol=1e-9)
    assert nacc.sum() > 0


@pytest.mark.parametrize("shape", [SHAPES[0], SHAPES[3], SHAPES[4]])
def test_sweep_philox_matches_oracle(device, shape):
    Lx, Ly, nu, nd = shape
    B, n_sw, seed, off = 40, 2, 0x1234_5678_9ABC, 17
    H, M, x, Md, xd, rng = setup_case(Lx, Ly, nu, nd, B, seed=15, std=0.3)
    N, nf = Lx * Ly, nu + nd
    edges = H.graph.sweep_edges()
    xo, so, lo, nacc = osamp.sweep(lambda c: osl.slogdet(M, c), x, edges,
                                   osamp.PhiloxDraws(seed, off), n_sw)
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    kernels.slater_sweep(Md, edges_tensor(edges), xd, st, n_sweeps=n_sw, seed=seed,
                         sweep_offset=off, refresh_every=64)
    assert np.array_equal(xd.cpu().numpy(), xo)
    assert np.array_equal(st.sign.cpu().numpy(), so)
    assert np.allclose(st.logabs.cpu().numpy(), lo, rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("shape", SHAPES)
def test_logderiv_matches_oracle(device, shape):
    Lx, Ly, nu, nd = shape
    H, M, x, Md, xd, _ = setup_case(Lx, Ly, nu, nd, 32, seed=16)
    N, nf = Lx * Ly, nu + nd
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    Oo = osl.log_derivative(M, x)
    O = kernels.slater_logderiv(st, N, nf).cpu().numpy()
    assert np.allclose(O, Oo, rtol=1e-10, atol=1e-10 * np.abs(Oo).max())
    O32 = kernels.slater_logderiv(st, N, nf, dtype=torch.float32).cpu().numpy()
    assert np.allclose(O32, Oo, rtol=1e-6, atol=1e-6 * np.abs(Oo).max())


def test_full_size_state_consistency(device):
    """8x8 half filling, 8192 walkers: after 30 sweeps of Sherman-Morrison
    updates the carried state equals an exact rebuild (drift bound), the
    electron numbers are conserved, and E_loc agrees."""
    Lx, Ly, nu, nd = 8, 8, 32, 32
    B = 8192
    H, M, x, Md, xd, _ = setup_case(Lx, Ly, nu, nd, B, seed=17, std=0.1)
    N, nf = 64, 64
    st = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st)
    e_sweep = edges_tensor(H.graph.sweep_edges())
    kernels.slater_sweep(Md, e_sweep, xd, st, n_sweeps=30, seed=5, refresh_every=64)
    fx = xd.cpu().numpy()
    assert ((fx >> 1) & 1).sum(1).tolist() == [nu] * B
    assert (fx & 1).sum(1).tolist() == [nd] * B
    e_graph = edges_tensor(H.graph.edges())
    E1 = kernels.slater_local_energy(Md, e_graph, 1.0, 8.0, xd, st)
    st2 = _state(xd, N, nf)
    kernels.slater_refresh(Md, xd, st2)
    E2 = kernels.slater_local_energy(Md, e_graph, 1.0, 8.0, xd, st2)
    assert torch.equal(st.sign, st2.sign)
    assert torch.allclose(st.logabs, st2.logabs, rtol=1e-10, atol=1e-9)
    assert torch.allclose(E1, E2, rtol=1e-8, atol=1e-8)
    acc = st.n_accept.double().mean().item() / (30 * e_sweep.shape[0])
    assert 0.01 < acc < 0.99
    # spot-check E_loc against the oracle on a subset
    sub = slice(0, 64)
    Eo = osl.local_energy(H, M, fx[sub])
    assert np.allclose(E2[sub].cpu().numpy(), Eo, rtol=1e-9, atol=1e-8)
