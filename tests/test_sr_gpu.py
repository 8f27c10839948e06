"""GPU parity of the SR solver (centring, Gram, blocked FP64 Cholesky,
triangular solves, preconditioned CG) and the SGD update against the FP64
oracle.

Tolerances: FP64 Gram 1e-9 relative; reduced-precision Grams with PCG on the
exact operator 1e-6 (FP64 O) / 2e-5 (FP32 O) relative."""
import numpy as np
import pytest
import torch

from oracle import sr as osr
from vmc_torch_amd import kernels

pytestmark = pytest.mark.gpu


def _problem(Ns, Np, seed):
    rng = np.random.default_rng(seed)
    O = rng.normal(size=(Ns, Np)) * 0.4 + rng.normal(size=Np)   # nonzero column means
    E = rng.normal(size=Ns) * 2.0 + 3.0
    return O, E


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a - b)


@pytest.mark.parametrize("Ns,Np", [(1, 5), (100, 60), (129, 300), (257, 200), (1000, 700),
                                   (300, 5000), (64, 40000), (640, 300), (1537, 400),
                                   (2100, 300)])
def test_minsr_fp64_matches_oracle(device, Ns, Np):
    O, E = _problem(Ns, Np, Ns)
    want = osr.minsr(O, E, E.mean(), 1e-3)
    dp, stats, info = kernels.sr_minsr(torch.tensor(O, device=device),
                                       torch.tensor(E, device=device), None, 1e-3, "fp64")
    assert int(info.item()) == 0
    assert _rel(dp.cpu().numpy(), want) < 1e-9
    s = stats.cpu().numpy()
    assert s[0] == pytest.approx(E.mean(), rel=1e-13)
    assert s[1] == pytest.approx(E.var(), rel=1e-10)
    g = osr.energy_gradient(O, E, E.mean())
    assert s[3] == pytest.approx(np.linalg.norm(g), rel=1e-9)


@pytest.mark.parametrize("Ns,Np", [(40, 6), (64, 12), (300, 20), (400, 140), (700, 300)])
def test_minsr_reports_first_nonpositive_minor(device, Ns, Np):
    """Ot Ot^T has rank Np, so with shift -1e-9 the leading minors of
    T = Ot Ot^T + shift I are positive definite up to order Np (smallest
    eigenvalue >= 5e-8 here) and not at Np + 1: info must be Np + 1, in the
    first or second 8 x 8 half of a leaf, a later leaf or a later diagonal
    block."""
    shift = -1e-9
    O, E = _problem(Ns, Np, Ns + 7)
    Ot = (O - O.mean(0)) / np.sqrt(Ns)
    T = Ot @ Ot.T + shift * np.eye(Ns)
    assert np.linalg.eigvalsh(T[:Np, :Np])[0] > 4e-8
    assert np.linalg.eigvalsh(T[:Np + 1, :Np + 1])[0] < 0
    _, _, info = kernels.sr_minsr(torch.tensor(O, device=device), torch.tensor(E, device=device),
                                  None, shift, "fp64", max_iter=0)
    assert int(info.item()) == Np + 1


@pytest.mark.parametrize("odt", [torch.float64, torch.float32])
@pytest.mark.parametrize("gram", ["fp32", "bf16", "bf16x3"])
@pytest.mark.parametrize("shift", [1e-2, 1e-4])
def test_minsr_mixed_precision_pcg(device, odt, gram, shift):
    """Reduced-precision Gram + FP64 Cholesky as preconditioner + PCG on the
    exact operator converges to the FP64 solution even when the Gram rounding
    exceeds the shift.  Bound: 1e-6 relative (FP64 O), 2e-5 (FP32 O, whose
    own rounding is 6e-8 relative)."""
    O, E = _problem(600, 400, 7)
    want = osr.minsr(O, E, E.mean(), shift)
    Od = torch.tensor(O, device=device, dtype=odt)
    dp, stats, info = kernels.sr_minsr(Od, torch.tensor(E, device=device), None, shift, gram,
                                       max_iter=60, tol=1e-10)
    assert int(info.item()) == 0
    tol = 1e-6 if odt == torch.float64 else 2e-5
    assert _rel(dp.cpu().numpy(), want) < tol
    assert stats[4].item() <= 60


def test_minsr_split_bf16_gram_needs_fewer_iterations(device):
    """The three-product BF16 Gram (hi hi^T + hi lo^T + lo hi^T) is a closer
    preconditioner than the one-product BF16 Gram: fewer PCG iterations to
    the same tolerance, same solution."""
    O, E = _problem(600, 400, 9)
    want = osr.minsr(O, E, E.mean(), 1e-4)
    Ed = torch.tensor(E, device=device)
    it = {}
    for gram in ("bf16", "bf16x3"):
        dp, stats, info = kernels.sr_minsr(torch.tensor(O, device=device, dtype=torch.float32),
                                           Ed, None, 1e-4, gram, max_iter=100, tol=1e-10)
        assert int(info.item()) == 0
        assert _rel(dp.cpu().numpy(), want) < 2e-5
        it[gram] = stats[4].item()
    print(f"  PCG iterations: bf16 {it['bf16']:.0f}, bf16x3 {it['bf16x3']:.0f}")
    assert it["bf16x3"] < it["bf16"]


@pytest.mark.parametrize("gram", ["bf16", "bf16x3"])
def test_minsr_indefinite_low_precision_gram(device, gram):
    """More samples than parameters and a tiny shift: the BF16 Gram's
    rounding leaves it indefinite at the shift.  The preconditioner's shift
    is raised until its Cholesky succeeds (info 0); PCG still solves the
    exact system (operator shift unchanged) to the FP64 solution."""
    O, E = _problem(400, 150, 10)
    shift = 1e-7
    want = osr.minsr(O, E, E.mean(), shift)
    dp, stats, info = kernels.sr_minsr(torch.tensor(O, device=device, dtype=torch.float64),
                                       torch.tensor(E, device=device), None, shift, gram,
                                       max_iter=400, tol=1e-10)
    print(f"  {gram}: info {int(info.item())}, PCG iterations {stats[4].item():.0f}, "
          f"residual {stats[2].item():.2e}, dp error {_rel(dp.cpu().numpy(), want):.2e}")
    assert int(info.item()) == 0
    assert _rel(dp.cpu().numpy(), want) < 1e-5


def test_minsr_inplace_fp32(device):
    O, E = _problem(300, 500, 8)
    want = osr.minsr(O, E, E.mean(), 1e-3)
    Od = torch.tensor(O, device=device, dtype=torch.float32)
    dp, stats, info = kernels.sr_minsr(Od, torch.tensor(E, device=device), None, 1e-3, "bf16",
                                       max_iter=60, tol=1e-10, inplace=True)
    assert _rel(dp.cpu().numpy(), want) < 2e-5
    Ot = (O - O.mean(0)) / np.sqrt(300)
    assert np.allclose(Od.cpu().numpy(), Ot, atol=1e-5)


def test_minsr_given_energy_mean(device):
    O, E = _problem(200, 90, 3)
    em = E.mean() + 0.37
    want = osr.minsr(O, E, em, 1e-4)
    dp, _, info = kernels.sr_minsr(torch.tensor(O, device=device),
                                   torch.tensor(E, device=device), em, 1e-4, "fp64")
    assert _rel(dp.cpu().numpy(), want) < 1e-8


def test_minsr_equals_param_space(device):
    """Woodbury: the sample-space solve equals the parameter-space SR."""
    O, E = _problem(150, 80, 4)
    want = osr.sr_dense(O, E, E.mean(), 1e-3)
    dp, _, _ = kernels.sr_minsr(torch.tensor(O, device=device),
                                torch.tensor(E, device=device), None, 1e-3, "fp64")
    assert _rel(dp.cpu().numpy(), want) < 1e-8


def test_minsr_full_size_residual(device):
    """Ns = Np = 8192 (the 8x8 Slater shape): size-independent check that
    (Ot Ot^T + shift) alpha = eps, through dp = Ot^T alpha and the gradient."""
    g = torch.Generator(device=device).manual_seed(0)
    Ns = Np = 8192
    O = torch.randn(Ns, Np, generator=g, device=device, dtype=torch.float64) * 0.3
    E = torch.randn(Ns, generator=g, device=device, dtype=torch.float64)
    dp, stats, info = kernels.sr_minsr(O, E, None, 1e-4, "fp64")
    assert int(info.item()) == 0
    s = stats.cpu().numpy()
    eps = (E - E.mean()) / Ns ** 0.5
    assert s[2] < 1e-9
    # dp solves the parameter-space system too: (S + shift) dp = g
    Ot = (O - O.mean(0)) / Ns ** 0.5
    lhs = Ot.T @ (Ot @ dp) + 1e-4 * dp
    rhs = Ot.T @ eps
    assert float((lhs - rhs).norm() / rhs.norm()) < 1e-7


def test_sgd_update(device):
    p = torch.randn(1000, dtype=torch.float64, device=device)
    d = torch.randn(1000, dtype=torch.float64, device=device)
    want = p - 0.05 * d
    kernels.sgd_update(p, d, 0.05)
    # one fused multiply-add per entry vs torch's multiply then subtract
    assert torch.allclose(p, want, rtol=1e-15, atol=1e-15)


# ------------------------------------------------------------ MINRES (parameter space)
@pytest.mark.parametrize("Ns,Np", [(1, 5), (37, 53), (200, 129), (513, 1100)])
def test_minres_matches_oracle(device, Ns, Np):
    """Same Lanczos recurrence in FP64: the converged solutions agree far
    below the solver tolerance (bound 1e-8 relative at rtol 1e-10)."""
    O, E = _problem(Ns, Np, 100 + Ns)
    want, winfo = osr.sr_minres(O, E, E.mean(), 1e-3, rtol=1e-10, maxiter=600)
    dp, stats, info = kernels.sr_minres(torch.tensor(O, device=device),
                                        torch.tensor(E, device=device), None, 1e-3,
                                        rtol=1e-10, maxiter=600)
    assert int(info.item()) == winfo == 0
    assert _rel(dp.cpu().numpy(), want) < 1e-8
    s = stats.cpu().numpy()
    assert s[0] == pytest.approx(E.mean(), rel=1e-13)
    g = osr.energy_gradient(O, E, E.mean())
    assert s[3] == pytest.approx(np.linalg.norm(g), rel=1e-10)
    # the converged MINRES solution solves the dense SR system
    dense = osr.sr_dense(O, E, E.mean(), 1e-3)
    assert _rel(dp.cpu().numpy(), dense) < 1e-6


def test_minres_truncated_iterates_match(device):
    """Stopped after 4 iterations (not converged): info = maxiter and the
    4th iterate equals the oracle's 4th iterate."""
    O, E = _problem(300, 250, 5)
    want, winfo = osr.sr_minres(O, E, 2.5, 1e-4, rtol=1e-14, maxiter=4)
    dp, stats, info = kernels.sr_minres(torch.tensor(O, device=device),
                                        torch.tensor(E, device=device), 2.5, 1e-4,
                                        rtol=1e-14, maxiter=4)
    assert winfo == 4 and int(info.item()) == 4
    assert stats[4].item() == 4
    assert _rel(dp.cpu().numpy(), want) < 1e-11


def test_minres_fp32_o_and_gradient_only(device):
    O, E = _problem(400, 300, 11)
    O32 = O.astype(np.float32)
    want, _ = osr.sr_minres(O32.astype(np.float64), E, E.mean(), 1e-2, rtol=1e-10,
                            maxiter=400)
    dp, _, info = kernels.sr_minres(torch.tensor(O32, device=device),
                                    torch.tensor(E, device=device), None, 1e-2, rtol=1e-10,
                                    maxiter=400)
    assert int(info.item()) == 0
    assert _rel(dp.cpu().numpy(), want) < 1e-8
    g, _, _ = kernels.sr_minres(torch.tensor(O, device=device), torch.tensor(E, device=device),
                                None, 1e-2, maxiter=0)
    assert _rel(g.cpu().numpy(), osr.energy_gradient(O, E, E.mean())) < 1e-13


def test_minres_zero_rhs(device):
    O, _ = _problem(50, 40, 3)
    E = np.full(50, 1.25)
    dp, stats, info = kernels.sr_minres(torch.tensor(O, device=device),
                                        torch.tensor(E, device=device), None, 1e-3)
    assert int(info.item()) == 0
    assert float(dp.abs().max()) == 0.0
    assert stats[4].item() == 0


def test_minres_preconditioner_interface(device):
    from vmc_torch_amd.optimizer import DistributedSRMinresGPU
    O, E = _problem(256, 180, 9)
    want, _ = osr.sr_minres(O, E, E.mean(), 1e-3, rtol=5e-5, maxiter=100)
    pre = DistributedSRMinresGPU(rtol=5e-5, maxiter=100)
    dp, secs, info = pre.solve(local_o=torch.tensor(O, device=device),
                               local_energies=torch.tensor(E, device=device),
                               energy_mean=float(E.mean()), total_samples=256, n_params=180,
                               diag_shift=1e-3, device=device, run_sr=True)
    assert info == 0 and secs >= 0
    assert _rel(dp.cpu().numpy(), want) < 1e-8
