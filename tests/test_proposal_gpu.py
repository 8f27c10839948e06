"""GPU parity of the learned proposal (conv net logits, inverse-CDF move
choice, reverse probability at x', Metropolis-Hastings acceptance, backward
pass) against the FP64 oracle, and the end-to-end property that the chain
samples |psi|^2.

Tolerances: logits and ln q 1e-5 (FP32 convolutions); gradients 1e-4
relative; moves, configurations and accept decisions identical with
injected uniforms kept 1e-4 away from every CDF boundary and acceptance
ratio; the sampled distribution of a 2x2 system within total variation
0.02 of |psi|^2 (4096 chains x 200 steps)."""
import itertools

import numpy as np
import pytest
import torch

from helpers import random_configs
from oracle import proposal as op
from oracle import slater as osl
from vmc_torch_amd import (MetropolisExchangeSpinfulSamplerGPU, SlaterDeterminant_GPU, kernels,
                           spinful_Fermi_Hubbard_square_lattice_torch)
from oracle import hubbard as ohub
from vmc_torch_amd.optimizer import MinSRGPU, SGDGPU
from vmc_torch_amd.sampler import LearnedProposalSamplerGPU
from vmc_torch_amd.stats import tau_int
from vmc_torch_amd.VMC import VMC_GPU, VMCLoopConfig

pytestmark = pytest.mark.gpu


def _net(C, seed, scale=1.0, head_bias=True):
    th = op.init_params(np.random.default_rng(seed), C, scale)
    if head_bias:
        th[-3:] = np.random.default_rng(seed + 1).normal(size=3)
    return th, torch.tensor(th, dtype=torch.float32)


@pytest.mark.parametrize("C,Lx,Ly", [(32, 4, 4), (8, 3, 5), (32, 8, 8), (16, 6, 6)])
def test_logits_match_oracle(device, C, Lx, Ly):
    th, P = _net(C, Lx * Ly)
    N = Lx * Ly
    x = random_configs(np.random.default_rng(3), 6, N, N // 3, N // 4)
    L = kernels.lp_logits(P.to(device), C, Lx, Ly, torch.tensor(x, device=device)).cpu().numpy()
    th32 = P.double().numpy()
    for b in range(len(x)):
        want, valid, _ = op.forward(th32, C, Lx, Ly, x[b])
        assert np.array_equal(np.isfinite(L[b]), valid)
        assert np.abs(L[b][valid] - want[valid]).max() < 1e-5 * max(1.0, np.abs(want).max())


def _slater(N, nf, seed, device):
    M = np.random.default_rng(seed).normal(0.0, 1.0, (2 * N, nf))
    m = SlaterDeterminant_GPU(2 * N, nf).to(device)
    with torch.no_grad():
        m.params[0].copy_(torch.tensor(M))
    return m, M


def test_step_matches_oracle(device):
    C, Lx, Ly = 16, 4, 4
    N, nu, nd = 16, 5, 4
    th, P = _net(C, 5, scale=2.0)
    th32 = P.double().numpy()
    model, M = _slater(N, nu + nd, 6, device)
    rng = np.random.default_rng(7)
    x0 = random_configs(rng, 32, N, nu, nd)

    def logamp(c):
        return osl.slogdet(M, np.asarray(c)[None])[1][0]
    uc = rng.random(32)
    ua = rng.random(32)
    want = []
    for b in range(32):
        lf, vf, _ = op.forward(th32, C, Lx, Ly, x0[b])
        cdf = np.cumsum(np.exp(op.log_q(lf, vf))[vf])
        while np.min(np.abs(cdf - uc[b])) < 1e-4:       # keep away from CDF boundaries
            uc[b] = rng.random()
        m = op.sample_move(op.log_q(lf, vf), uc[b])
        xp = op.apply_move(x0[b], Lx, Ly, m)
        lr, vr, _ = op.forward(th32, C, Lx, Ly, xp)
        r = np.exp(2 * (logamp(xp) - logamp(x0[b])) + op.log_q(lr, vr)[m] - op.log_q(lf, vf)[m])
        if abs(ua[b] - r) < 1e-4 * max(r, 1e-300):
            ua[b] = 0.5 * min(r, 1.0)
        want.append(op.mh_step(th32, C, Lx, Ly, x0[b], logamp, uc[b], ua[b]))
    xd = torch.tensor(x0, device=device)
    s = LearnedProposalSamplerGPU(Lx, Ly, channels=C, params=P.to(device))
    sign, logabs = s.sweeps(xd, model, 1, draws=[(torch.tensor(uc, device=device),
                                                  torch.tensor(ua, device=device))])
    got = xd.cpu().numpy()
    for b in range(32):
        assert np.array_equal(got[b], want[b][0])
    so, lo = osl.slogdet(M, got)
    assert np.array_equal(sign.cpu().numpy(), so)
    assert np.abs(logabs.cpu().numpy() - lo).max() < 1e-10
    n_acc, _ = s.last_acceptance
    assert np.array_equal(n_acc.cpu().numpy(), np.array([w[1] for w in want], dtype=np.int64))


def test_backward_matches_oracle(device):
    C, Lx, Ly = 8, 3, 4
    th, P = _net(C, 9)
    th32 = P.double().numpy()
    x = random_configs(np.random.default_rng(8), 5, 12, 4, 4)
    rng = np.random.default_rng(9)
    dl = rng.normal(size=(5, op.n_moves(Lx, Ly)))
    g = kernels.lp_backward(P.to(device), C, Lx, Ly, torch.tensor(x, device=device),
                            torch.tensor(dl, device=device)).cpu().numpy()
    want = sum(op.backward(th32, C, Lx, Ly, x[b], dl[b]) for b in range(5))
    assert np.abs(g - want).max() < 1e-4 * np.abs(want).max()


@pytest.mark.parametrize("Lx,Ly", [(4, 4), (8, 8), (6, 6), (3, 5), (2, 2)])
def test_bf16_logits_match_rounding_oracle(device, Lx, Ly):
    """The BF16 matrix-core net against the oracle's bfloat16 restatement:
    within one bfloat16 step (2^-8) of the logit scale -- the device sums in
    FP32 and in another order, so an activation can round to the
    neighbouring bfloat16 value and carry that step to the logits -- and, in
    total, at most half as far from it as from the FP32 net.  40
    configurations: several workgroups, the last one partly filled."""
    C, N = 32, Lx * Ly
    th, P = _net(C, N + 100)
    x = random_configs(np.random.default_rng(5), 40, N, max(1, N // 3), max(1, N // 4))
    L = kernels.lp_logits(P.to(device), C, Lx, Ly, torch.tensor(x, device=device),
                          precision="bf16").cpu().numpy()
    th32 = P.double().numpy()
    e16 = e32 = 0.0
    for b in range(len(x)):
        want, valid, _ = op.forward(th32, C, Lx, Ly, x[b], bf16=True)
        w32, _, _ = op.forward(th32, C, Lx, Ly, x[b])
        assert np.array_equal(np.isfinite(L[b]), valid)
        if not valid.any():
            continue
        d = np.abs(L[b][valid] - want[valid])
        assert d.max() < 2.0 ** -8 * max(1.0, np.abs(want[valid]).max())
        e16 += d.sum()
        e32 += np.abs(L[b][valid] - w32[valid]).sum()
    print(f"  summed |error| vs bf16 oracle {e16:.3e}, vs fp32 oracle {e32:.3e}")
    assert e16 < 0.5 * e32


@pytest.mark.parametrize("C,precision", [(8, "fp32"), (32, "bf16")])
def test_chain_samples_psi_squared(device, C, precision):
    """2x2, one up and one down electron: the learned-proposal chain's
    histogram equals |psi|^2 (an untrained net with small weights, so the
    chain mixes within the run; stationarity for any net is pinned exactly on
    the CPU), with the FP32 and with the BF16 net."""
    N = 4
    model, M = _slater(N, 2, 11, device)
    th, P = _net(C, 12, scale=0.3, head_bias=False)
    s = LearnedProposalSamplerGPU(2, 2, channels=C, params=P.to(device), seed=5,
                                  precision=precision)
    B = 4096
    x0 = np.tile(np.array([[2, 1, 0, 0]], np.int64), (B, 1))
    xd = torch.tensor(x0, device=device)
    s.sweeps(xd, model, 200)                                 # burn-in
    code = torch.tensor([64, 16, 4, 1], device=device)
    hist = torch.zeros(256, dtype=torch.int64, device=device)
    for _ in range(200):
        s.sweeps(xd, model, 1)
        hist += torch.bincount((xd * code).sum(1), minlength=256)
    states = [np.array(c) for c in itertools.product(range(4), repeat=4)
              if sum((v >> 1) & 1 for v in c) == 1 and sum(v & 1 for v in c) == 1]
    p = np.array([np.exp(2 * osl.slogdet(M, c[None])[1][0]) for c in states])
    p /= p.sum()
    h = hist.cpu().numpy()
    emp = np.array([h[int(c @ np.array([64, 16, 4, 1]))] for c in states], dtype=np.float64)
    assert emp.sum() == 200 * B
    emp /= emp.sum()
    assert 0.5 * np.abs(emp - p).sum() < 0.02


@pytest.mark.parametrize("C,precision", [(8, "fp32"), (32, "bf16")])
def test_two_move_chain_samples_psi_squared(device, C, precision):
    """moves_per_proposal = 2 (two chained moves, path acceptance): the 2x2
    histogram equals |psi|^2 as with single moves; on 2x3 with two electrons
    of each spin the composite proposals move more electrons at once."""
    N = 4
    model, M = _slater(N, 2, 11, device)
    th, P = _net(C, 12, scale=0.3, head_bias=False)
    s = LearnedProposalSamplerGPU(2, 2, channels=C, params=P.to(device), seed=6,
                                  precision=precision, moves_per_proposal=2)
    B = 4096
    x0 = np.tile(np.array([[2, 1, 0, 0]], np.int64), (B, 1))
    xd = torch.tensor(x0, device=device)
    s.sweeps(xd, model, 200)                                 # burn-in
    code = torch.tensor([64, 16, 4, 1], device=device)
    hist = torch.zeros(256, dtype=torch.int64, device=device)
    for _ in range(200):
        s.sweeps(xd, model, 1)
        hist += torch.bincount((xd * code).sum(1), minlength=256)
    states = [np.array(c) for c in itertools.product(range(4), repeat=4)
              if sum((v >> 1) & 1 for v in c) == 1 and sum(v & 1 for v in c) == 1]
    p = np.array([np.exp(2 * osl.slogdet(M, c[None])[1][0]) for c in states])
    p /= p.sum()
    h = hist.cpu().numpy()
    emp = np.array([h[int(c @ np.array([64, 16, 4, 1]))] for c in states], dtype=np.float64)
    assert emp.sum() == 200 * B
    emp /= emp.sum()
    assert 0.5 * np.abs(emp - p).sum() < 0.02
    assert (xd >= 0).all() and ((xd >> 1) & 1).sum(1).eq(1).all() and (xd & 1).sum(1).eq(1).all()


def test_training_lowers_heat_bath_loss(device):
    Lx, Ly, N = 4, 4, 16
    model, _ = _slater(N, 8, 13, device)
    s = LearnedProposalSamplerGPU(Lx, Ly, channels=16, generator=torch.Generator().manual_seed(0),
                                  device=device)
    x = torch.tensor(random_configs(np.random.default_rng(14), 128, N, 4, 4), device=device)
    l0 = s.train_step(x, model, lr=0.0)
    for _ in range(30):
        s.train_step(x, model, lr=0.05)
    l1 = s.train_step(x, model, lr=0.0)
    assert l1 < l0


def test_shared_net_across_lattices(device):
    """One net serves the 4x4 (n = 1), 6x6 (n = 0.89) and 8x8 (n = 1)
    instances; particle numbers are conserved and moves get accepted."""
    P = LearnedProposalSamplerGPU.init_params(32, torch.Generator().manual_seed(1)).to(device)
    for (Lx, Ly, nu, nd) in [(4, 4, 8, 8), (6, 6, 16, 16), (8, 8, 32, 32)]:
        N = Lx * Ly
        model, _ = _slater(N, nu + nd, N, device)
        s = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=P, seed=N)
        x = torch.tensor(random_configs(np.random.default_rng(N), 256, N, nu, nd), device=device)
        s.sweeps(x, model, 5)
        xc = x.cpu().numpy()
        assert np.all(((xc >> 1) & 1).sum(1) == nu) and np.all((xc & 1).sum(1) == nd)
        n_acc, n = s.last_acceptance
        assert 0 < float(n_acc.double().mean()) / n <= 1.0


def _afm_slater(Lx, Ly, nu, nd, delta, device):
    """Mean-field antiferromagnet: per spin sigma the lowest orbitals of
    -t (nearest-neighbour hops) + sigma * delta * (-1)^(x+y); rows are the
    modes [u_0..u_{N-1} | d_0..d_{N-1}], columns nu up then nd down orbitals.
    A structured |psi|^2 on which most local moves are rejected."""
    N = Lx * Ly
    h = np.zeros((N, N))
    for s in range(N):
        x, y = divmod(s, Ly)
        if y + 1 < Ly:
            h[s, s + 1] = h[s + 1, s] = -1.0
        if x + 1 < Lx:
            h[s, s + Ly] = h[s + Ly, s] = -1.0
    stag = np.diag([(-1.0) ** (s // Ly + s % Ly) for s in range(N)])
    M = np.zeros((2 * N, nu + nd))
    M[:N, :nu] = np.linalg.eigh(h + delta * stag)[1][:, :nu]
    M[N:, nu:] = np.linalg.eigh(h - delta * stag)[1][:, :nd]
    m = SlaterDeterminant_GPU(2 * N, nu + nd).to(device)
    with torch.no_grad():
        m.params[0].copy_(torch.tensor(M))
    return m


def test_trained_net_gains_over_untrained(device, capsys):
    """What the trained network gains.  4x4 U=8 Hubbard, 8 up + 8 down, on a
    mean-field antiferromagnet (staggered field 3): the same learned-proposal
    chain with the untrained net and with the net after 300 heat-bath
    training steps on the chain's own walkers, 24 proposals per sample (as
    many as one plain sweep has edges), 2048 chains x 150 samples each.
    Measured on MI355X: acceptance per proposal 3.3 % untrained -> 74 %
    trained, tau_int of E_loc 3.6 -> 1.5 samples; the plain sweep sampler has
    tau_int 6.0 per 24-proposal sweep.  All three chains sample |psi|^2: the
    mean energies agree within 5 standard errors."""
    Lx, Ly, nu, nd, B, T, k = 4, 4, 8, 8, 2048, 150, 24
    N = Lx * Ly
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nu + nd,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    model = _afm_slater(Lx, Ly, nu, nd, 3.0, device)
    x0 = torch.tensor(random_configs(np.random.default_rng(2), B, N, nu, nd), device=device)
    plain = MetropolisExchangeSpinfulSamplerGPU(seed=11)
    plain.sweeps(x0, model, H.graph, 100)

    def run(advance):
        series, acc = [], 0.0
        for _ in range(T):
            acc += advance()
            series.append(model.local_energy(x, H) / N)
        s = torch.stack(series, 1)
        return (acc / T, tau_int(s), float(s.mean()),
                float(s.mean(1).std()) / np.sqrt(B))

    x = x0.clone()

    def plain_step():
        plain.sweeps(x, model, H.graph, 1)
        n, m = plain.last_acceptance
        return float(n.double().mean()) / m
    res = {"plain": run(plain_step)}
    P0 = LearnedProposalSamplerGPU.init_params(32, torch.Generator().manual_seed(5)).to(device)
    trainer = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=P0.clone(), seed=9)
    xt = x0.clone()
    for _ in range(300):
        trainer.sweeps(xt, model, 2)
        trainer.train_step(xt[:1024], model, lr=0.05)
    for tag, P, seed in [("untrained", P0, 7), ("trained", trainer.params, 13)]:
        lp = LearnedProposalSamplerGPU(Lx, Ly, channels=32, params=P.clone(), seed=seed)
        x = x0.clone()
        lp.sweeps(x, model, 2 * k)

        def lp_step():
            lp.sweeps(x, model, k)
            n, m = lp.last_acceptance
            return float(n.double().mean()) / m
        res[tag] = run(lp_step)
    with capsys.disabled():
        for tag, (acc, tau, e, se) in res.items():
            print(f"\n  {tag:9s}: acceptance {acc:.4f}, tau_int {tau:.2f} per {k} proposals, "
                  f"E/site {e:.5f} +- {se:.5f}")
    acc_u, tau_u, e_u, se_u = res["untrained"]
    acc_t, tau_t, e_t, se_t = res["trained"]
    _, tau_p, e_p, se_p = res["plain"]
    assert acc_t > 5.0 * acc_u
    assert tau_t < 0.7 * tau_u and tau_t < tau_p
    for e, se in [(e_u, se_u), (e_t, se_t)]:
        assert abs(e - e_p) < 5.0 * np.hypot(se, se_p)


def test_vmc_loop_with_learned_proposal(device):
    """The learned proposal as the sampler of VMC_GPU (4x2, 4 up + 4 down,
    Slater, 1024 walkers, 6 proposals per sample), its net trained on the
    chain's walkers after every step (on_step_end) while the SR step runs with
    learning rate 0, so every sample is drawn from one known |psi|^2: each
    step's energy agrees with the exact <H> (enumeration of all 4900 states,
    FP64 oracle) within 5 standard errors, the particle numbers are
    conserved and the parameters stay put.  (With a moving wavefunction a few
    proposals per step lag behind it, for any sampler.)"""
    Lx, Ly, nu, nd, B, steps = 4, 2, 4, 4, 1024, 6
    N, nf = Lx * Ly, nu + nd
    H = spinful_Fermi_Hubbard_square_lattice_torch(Lx, Ly, 1.0, 8.0, nf,
                                                   n_fermions_per_spin=(nu, nd))
    H.precompute_hops_gpu(device)
    model = SlaterDeterminant_GPU(2 * N, nf, init_std=0.1,
                                  generator=torch.Generator().manual_seed(1)).to(device)
    fxs = torch.tensor(random_configs(np.random.default_rng(2), B, N, nu, nd), device=device)
    lp = LearnedProposalSamplerGPU(Lx, Ly, channels=32, seed=3, device=device,
                                   proposals_per_step=6,
                                   generator=torch.Generator().manual_seed(4))
    vmc = VMC_GPU(lp, MinSRGPU("fp64"), SGDGPU(0.0))
    cfg = VMCLoopConfig(vmc_steps=steps, ns_per_rank=B, n_params=2 * N * nf, nsites=N,
                        learning_rate=0.0, diag_shift=1e-3, burn_in_steps=60,
                        outlier_clip_factor=0.0, use_log_amp=True)
    Ho = ohub.Hubbard(Lx, Ly, 1.0, 8.0, nu, nd)
    M0 = model.M.detach().cpu().numpy().copy()
    e_exact = osl.exact_energy_and_grad(Ho, M0)[0]
    log = []

    def on_end(d):
        log.append((d["energy_mean"], d["energy_var"]))
        x = d["fxs"].cpu().numpy()
        assert np.all(((x >> 1) & 1).sum(1) == nu) and np.all((x & 1).sum(1) == nd)
        lp.train_step(d["fxs"], model, lr=0.05)
    vmc.run_vmc_loop(fxs, model, H, H.graph, config=cfg, on_step_end=on_end)
    assert len(log) == steps
    assert np.array_equal(model.M.detach().cpu().numpy(), M0)
    for e, var in log:
        assert abs(e - e_exact) < 5.0 * np.sqrt(var / B), (log, e_exact)


@pytest.mark.parametrize("C,precision", [(16, "fp32"), (32, "bf16")])
def test_cached_logits_chain_equals_full_evaluation(device, C, precision):
    """The sampler evaluates the net once per proposal (logits at x carried
    over from the accepted proposals); the chain equals the one that scores
    x and x' afresh every step (vmcx_lp_propose), configuration for
    configuration, with the same Philox draws."""
    N, nu, nd, B, steps = 16, 6, 5, 128, 25
    model, _ = _slater(N, nu + nd, 21, device)
    P = LearnedProposalSamplerGPU.init_params(C, torch.Generator().manual_seed(22)).to(device)
    x0 = torch.tensor(random_configs(np.random.default_rng(23), B, N, nu, nd), device=device)
    a = x0.clone()
    s = LearnedProposalSamplerGPU(4, 4, channels=C, params=P, seed=24, precision=precision)
    s.sweeps(a, model, steps)
    b = x0.clone()
    sign, logabs = model.forward_log(b)
    sign, logabs = sign.clone(), logabs.clone()
    for step in range(steps):
        move, lqf, prop, lqr = kernels.lp_propose(P, C, 4, 4, b, 24, step, precision=precision)
        ps, pl = model.forward_log(prop)
        kernels.lp_accept(b, sign, logabs, move, prop, ps, pl, lqf, lqr, 24, step)
    assert torch.equal(a, b)


def test_accept_commit_fused_matches_separate(device):
    """vmcx_lp_accept_commit equals vmcx_lp_accept + vmcx_lp_commit_logits +
    vmcx_masked_rows bit for bit (Philox draws and injected uniforms)."""
    from vmc_torch_amd import kernels
    rng = np.random.default_rng(40)
    B, N, M, H = 1000, 36, 180, 72
    x = torch.tensor(rng.integers(0, 4, size=(B, N)), device=device)
    prop = torch.tensor(rng.integers(0, 4, size=(B, N)), device=device)
    move = torch.tensor(rng.integers(-1, M, size=B), dtype=torch.int32, device=device)
    f64 = lambda *s: torch.tensor(rng.normal(size=s), device=device)   # noqa: E731
    sign, logabs, ps, pl, lqf, lqr = (f64(B) for _ in range(6))
    lx, lp = f64(B, M), f64(B, M)
    z1, z1p = f64(B, H).float(), f64(B, H).float()
    for u_acc in (None, torch.tensor(rng.random(B), device=device)):
        a = [t.clone() for t in (x, sign, logabs, lx, z1)]
        na = torch.zeros(B, dtype=torch.int64, device=device)
        acc_a = kernels.lp_accept(a[0], a[1], a[2], move, prop, ps, pl, lqf, lqr, 5, 17,
                                  u_acc=u_acc, n_accept=na)
        kernels.lp_commit_logits(acc_a, lp, a[3])
        kernels.masked_rows(acc_a, z1p, a[4])
        b = [t.clone() for t in (x, sign, logabs, lx, z1)]
        nb = torch.zeros(B, dtype=torch.int64, device=device)
        acc_b = torch.empty(B, dtype=torch.uint8, device=device)
        kernels.lp_accept_commit(b[0], b[1], b[2], move, prop, ps, pl, lqf, lqr, 5, 17, lp, b[3],
                                 acc_b, n_accept=nb, u_acc=u_acc, z1p=z1p, z1=b[4])
        assert 0 < int(acc_a.sum()) < B
        assert torch.equal(acc_a, acc_b) and torch.equal(na, nb)
        for ta, tb in zip(a, b):
            assert torch.equal(ta, tb)
