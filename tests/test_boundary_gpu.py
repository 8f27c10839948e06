"""GPU checks of the drop-in boundary's error and corner behaviour against
the FP64 oracle:
  * vmcx_energy_stats (compensated) against exactly rounded sums (math.fsum);
  * the energy gradient on a factored O (vmcx_sr_factored_gradient) and
    MinSRGPU.solve(run_sr=False) (GPU/optimizer.py:581-614), dense and
    factored, against oracle.sr.energy_gradient, 1e-12 relative;
  * the Cholesky fallback of MinSRGPU (GPU/optimizer.py:557-561): a failing
    factorisation is solved again at a raised shift, info 1;
  * the NaN/Inf guard of compute_grads_gpu (GPU/vmc_utils.py:1330-1386);
  * two SR solves enqueued on two streams at once equal the same solves
    run one after the other (per-stream side streams and rocBLAS handles)."""
import math

import numpy as np
import pytest
import torch

from oracle import sr as osr
from test_sr_factored_gpu import _random_factored
from test_sr_gpu import _problem, _rel
from vmc_torch_amd import SlaterDeterminant_GPU, kernels
from vmc_torch_amd.optimizer import MinSRGPU
from vmc_torch_amd.vmc_utils import compute_grads_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 7, 1000, 16384])
def test_energy_stats_compensated(device, n):
    rng = np.random.default_rng(n)
    E = 1e8 + rng.normal(size=n) * 1e-3          # naive FP64 sums lose digits here
    mean = math.fsum(E) / n
    var = math.fsum((e - mean) ** 2 for e in E) / n
    st = kernels.energy_stats(torch.tensor(E, device=device)).cpu().numpy()
    assert st[0] == pytest.approx(mean, rel=1e-15)
    assert st[1] == pytest.approx(var, rel=1e-6, abs=1e-12)


@pytest.mark.parametrize("Ns,k0,given", [(1, 7, None), (300, 500, None), (257, 64, 2.5)])
def test_factored_gradient_matches_oracle(device, Ns, k0, given):
    FO, E = _random_factored(Ns, k0, Ns + 3 * k0, device)
    O = FO.materialize().cpu().numpy()
    em = E.mean() if given is None else given
    want = osr.energy_gradient(O, E, em)
    g, stats = kernels.sr_factored_gradient(FO, torch.tensor(E, device=device), given)
    assert _rel(g.cpu().numpy(), want) < 1e-12
    s = stats.cpu().numpy()
    assert s[0] == pytest.approx(E.mean(), rel=1e-13)
    assert s[3] == pytest.approx(np.linalg.norm(want), rel=1e-12)


def test_minsr_run_sr_false_returns_gradient(device):
    O, E = _problem(300, 120, 5)
    Od, Ed = torch.tensor(O, device=device), torch.tensor(E, device=device)
    for gram in ("fp64", "fp32"):
        g, _, info = MinSRGPU(gram).solve(local_o=Od, local_energies=Ed, energy_mean=E.mean(),
                                          total_samples=300, n_params=120, diag_shift=1e-3,
                                          device=device, run_sr=False)
        assert info is None
        assert _rel(g.cpu().numpy(), osr.energy_gradient(O, E, E.mean())) < 1e-12
    FO, E2 = _random_factored(200, 40, 6, device)
    O2 = FO.materialize().cpu().numpy()
    g, _, info = MinSRGPU("factored").solve(
        local_o=FO, local_energies=torch.tensor(E2, device=device), energy_mean=1.5,
        total_samples=200, n_params=FO.n_params, diag_shift=1e-3, device=device, run_sr=False)
    assert info is None and _rel(g.cpu().numpy(), osr.energy_gradient(O2, E2, 1.5)) < 1e-12


@pytest.mark.parametrize("factored", [False, True])
def test_minsr_cholesky_fallback(device, factored):
    """Ot Ot^T has rank Np < Ns, so at shift -1e-9 the factorisation fails at
    minor Np + 1; the solver retries at shift 1e-8 and returns that step."""
    Ns = 300
    if factored:
        FO, E = _random_factored(Ns, 20, 21, device)
        Np, lo = FO.n_params, FO
        O = FO.materialize().cpu().numpy()
        sv = np.linalg.svd(O - O.mean(0), compute_uv=False)
        rank = int((sv > 1e-9 * sv[0]).sum())
    else:
        Np = 40
        O, E = _problem(Ns, Np, 31)
        lo, rank = torch.tensor(O, device=device), Np
    assert rank < Ns
    solver = MinSRGPU("factored" if factored else "fp64")
    dp, _, info = solver.solve(local_o=lo, local_energies=torch.tensor(E, device=device),
                               energy_mean=E.mean(), total_samples=Ns, n_params=Np,
                               diag_shift=-1e-9, device=device)
    assert info == 1
    assert solver.last_fallback["diag_shift"] == pytest.approx(1e-8)
    assert solver.last_fallback["cholesky_info"] >= 1
    want = osr.minsr(O, E, E.mean(), 1e-8)
    assert _rel(dp.cpu().numpy(), want) < 1e-6


def test_compute_grads_rejects_nan(device):
    m = SlaterDeterminant_GPU(16, 8, init_std=0.1,
                              generator=torch.Generator().manual_seed(0)).to(device)
    x = torch.tensor([[3, 3, 3, 3, 0, 0, 0, 0], [0, 0, 0, 0, 3, 3, 3, 3]], device=device)
    compute_grads_gpu(x, m, use_log_amp=True)
    with torch.no_grad():
        m.M[0, 0] = float("nan")
    m.invalidate()
    with pytest.raises(ValueError, match="NaN/Inf"):
        compute_grads_gpu(x, m, use_log_amp=True)


def test_two_streams_do_not_share_lookahead(device):
    """Two factorisations enqueued on two streams before either is waited on
    give the results of the same solves run alone (the look-ahead side
    streams and events are per caller stream)."""
    probs = [_problem(1100, 700, 41), _problem(1300, 500, 42)]
    alone = []
    for O, E in probs:
        dp, _, info = kernels.sr_minsr(torch.tensor(O, device=device),
                                       torch.tensor(E, device=device), None, 1e-3, "fp64")
        assert int(info.item()) == 0
        alone.append(dp.cpu().numpy())
    streams = [torch.cuda.Stream(device), torch.cuda.Stream(device)]
    inputs = [(torch.tensor(O, device=device), torch.tensor(E, device=device)) for O, E in probs]
    torch.cuda.synchronize(device)
    outs = []
    for (Od, Ed), st in zip(inputs, streams):
        with torch.cuda.stream(st):
            outs.append(kernels.sr_minsr(Od, Ed, None, 1e-3, "fp64",
                                         workspace=kernels.SRWorkspace()))
    torch.cuda.synchronize(device)
    for (dp, _, info), want in zip(outs, alone):
        assert int(info.item()) == 0
        assert _rel(dp.cpu().numpy(), want) < 1e-12
