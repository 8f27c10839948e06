"""Metropolis samplers with the reference's interface (GPU/sampler.py).

MetropolisExchangeSpinfulSamplerGPU keeps the proposal semantics of
propose_exchange_or_hopping_vec (GPU/sampler.py:47-137) and the sweep of
step (GPU/sampler.py:216-388): edges in row_edges-then-col_edges order, one
proposal per edge per walker, acceptance u < |psi'|^2/|psi|^2.  The whole
sweep runs in one HIP launch per sweep with Sherman-Morrison updates; the
draws come from Philox4x32-10 keyed by (seed, sweep, edge, walker).
"""
import contextlib
from typing import Tuple

import torch



class SamplerGPU:
    """GPU/sampler.py:140-197: MCMC only (no energies or gradients)."""

    def step(self, fxs, model, graph, **kwargs) -> Tuple[torch.Tensor, torch.Tensor]:
        raise NotImplementedError

    def burn_in(self, fxs, model, graph, n_steps, **kwargs):
        for _ in range(n_steps):
            fxs, _ = self.step(fxs, model, graph, **kwargs)
        return fxs


class MetropolisExchangeSpinfulSamplerGPU(SamplerGPU):
    """Exchange / hopping sampler for spinful fermions.

    Args:
        hopping_rate: fraction of proposals that are hoppings (0.25).
        seed: Philox key (64-bit).
        refresh_every: exact rebuild of A^-1 after this many accepted moves
            (checked at sweep boundaries); 0 disables it.
    """

    def __init__(self, hopping_rate: float = 0.25, seed: int = 0, refresh_every: int = 64):
        self.hopping_rate = hopping_rate
        self.seed = int(seed)
        self.refresh_every = int(refresh_every)
        self.sweep_counter = 0
        self.last_acceptance = None

    @torch.inference_mode()
    def sweeps(self, fxs, model, graph, n_sweeps=1, draws=None):
        """n_sweeps sweeps in place (model.mcmc_sweep runs them in libvmcx);
        returns (sign, log|psi|) of the walkers afterwards."""
        edges = graph.edge_tensor(fxs.device, "sweep")
        sign, logabs, n_acc = model.mcmc_sweep(fxs, edges, 0, self.seed, self.sweep_counter,
                                               self.hopping_rate, self.refresh_every)
        acc0 = n_acc.clone()
        sign, logabs, n_acc = model.mcmc_sweep(fxs, edges, n_sweeps, self.seed,
                                               self.sweep_counter, self.hopping_rate,
                                               self.refresh_every, draws=draws)
        self.sweep_counter += n_sweeps
        if n_sweeps:
            self.last_acceptance = (n_acc - acc0, n_sweeps * edges.shape[0])
        return sign, logabs

    @torch.inference_mode()
    def step(self, fxs, model, graph, compile=False, verbose=False, use_log_amp=False,
             draws=None, **kwargs):
        """One sweep over all edges.  Returns (fxs, amps) or, with
        use_log_amp, (fxs, (signs, log_abs)); fxs is updated in place."""
        sign, logabs = self.sweeps(fxs, model, graph, 1, draws=draws)
        if use_log_amp:
            return fxs, (sign.clone(), logabs.clone())
        return fxs, sign * torch.exp(logabs)

    def burn_in(self, fxs, model, graph, n_steps, **kwargs):
        self.sweeps(fxs, model, graph, int(n_steps))
        return fxs


class MetropolisExchangeSpinfulSamplerReuse_GPU(MetropolisExchangeSpinfulSamplerGPU):
    """Exchange / hopping sampler with environment reuse
    (GPU/sampler.py:391-700): the same proposals, Philox draws and
    Metropolis tests as MetropolisExchangeSpinfulSamplerGPU, in the same edge
    order (row edges row by row, then column edges column by column); every
    proposal is evaluated from the walker's cached column environments
    (vmcx_peps_sweep_reuse), so the chain equals the plain sampler's up to
    rounding.  The model must be a models.fPEPS_Model_reuse_GPU."""

    def _check(self, model):
        from .models import fPEPS_Model_reuse_GPU
        if not isinstance(model, fPEPS_Model_reuse_GPU):
            raise TypeError("MetropolisExchangeSpinfulSamplerReuse_GPU needs an "
                            f"fPEPS_Model_reuse_GPU, got {type(model).__name__}")

    def sweeps(self, fxs, model, graph, n_sweeps=1, draws=None):
        self._check(model)
        return super().sweeps(fxs, model, graph, n_sweeps, draws=draws)


class LearnedProposalSamplerGPU(SamplerGPU):
    """Metropolis-Hastings with a learned local proposal (no counterpart in the
    reference; its external-proposal sampler MetropolisMPSSamplerSpinful,
    sampler.py:2120-2238, uses the same acceptance min(1, P'/P q/q'),
    :2214-2218).

    A 3-layer 3x3 convolutional net on the (n_up, n_dn) grid scores every
    local move (hop of an up or a down electron across an edge, exchange of a
    down and an up electron); q(m|x) is the softmax over the valid moves and
    the acceptance uses q(m|x') from the net at the proposed configuration,
    so the chain samples |psi|^2 exactly for any net.  `train_step` fits the
    net to the heat-bath distribution over the move set,
    q*(m|x) ~ |psi(m(x))|^2, by cross entropy.  The net (FP32, `params`) is
    size-independent: one instance can serve lattices of different shapes
    (pass the same tensor).  precision: "bf16" (layers 2 and 3 on the BF16
    matrix cores, C = 32) or "fp32"; default bf16 when C = 32.  The chain is
    exact in either precision: x and x' are scored by the same net.
    `step` (one sample, the sampler interface of the reference) makes
    `proposals_per_step` proposals; `burn_in(n)` makes n of them.
    moves_per_proposal = 2: each proposal chains two moves, m1 from q(.|x)
    and m2 from q(.|x1), so up to four electrons move at once; the
    acceptance is that of the path, |psi(x2)|^2 q(m2'|x2) q(m1'|x1) /
    (|psi(x)|^2 q(m1|x) q(m2|x1)) with m' the reverse moves (the reverse
    path is the same two-move mechanism run from x2), which keeps |psi|^2
    exact; the net runs at x1 and x2, the wavefunction only at x2.
    """

    def __init__(self, Lx, Ly, channels=32, seed=0, params=None, generator=None,
                 device=None, precision=None, proposals_per_step=1, moves_per_proposal=1):
        from . import kernels
        self.k = kernels
        self.Lx, self.Ly, self.C = int(Lx), int(Ly), int(channels)
        if precision is None:
            precision = "bf16" if self.C == 32 else "fp32"
        if precision not in ("bf16", "fp32") or (precision == "bf16" and self.C != 32):
            raise ValueError("precision must be 'fp32' or 'bf16' (bf16 needs channels = 32)")
        self.precision = precision
        self.proposals_per_step = int(proposals_per_step)
        if self.proposals_per_step < 1:
            raise ValueError("proposals_per_step must be >= 1")
        self.moves_per_proposal = int(moves_per_proposal)
        if self.moves_per_proposal not in (1, 2):
            raise ValueError("moves_per_proposal must be 1 or 2")
        self.seed = int(seed)
        self.step_counter = 0
        self.last_acceptance = None
        if params is None:
            params = self.init_params(self.C, generator)
        self.params = params if device is None else params.to(device)

    @staticmethod
    def init_params(C, generator=None):
        """Kaiming-uniform convolutions, head N(0, 0.1^2), zero biases."""
        shapes = [(C, 2, 3, 3), (C,), (C, C, 3, 3), (C,), (C, C, 3, 3), (C,), (3, 2 * C), (3,)]
        out = []
        for sh in shapes:
            if len(sh) == 4:
                a = (6.0 / (sh[1] * 9)) ** 0.5
                out.append(torch.empty(sh).uniform_(-a, a, generator=generator).reshape(-1))
            elif len(sh) == 2:
                out.append((torch.randn(sh, generator=generator) * 0.1).reshape(-1))
            else:
                out.append(torch.zeros(sh).reshape(-1))
        return torch.cat(out).to(torch.float32)

    @torch.inference_mode()
    def sweeps(self, fxs, model, n_steps=1, draws=None):
        """n_steps proposals per walker; returns (sign, log|psi|).  draws: list
        of (u_choice, u_acc) device arrays per step (tests)."""
        if self.params.device != fxs.device:
            self.params = self.params.to(fxs.device)
        # the model's parameters are constant over the proposals: one check
        # on entry instead of a value comparison (a host synchronisation) per
        # proposal
        frozen = model.frozen() if hasattr(model, "frozen") else contextlib.nullcontext()
        # models with a first-layer state (NNBF) evaluate a proposed
        # configuration from its walker's pre-activations and the moved
        # electrons' weight columns
        moved = hasattr(model, "forward_log_moved")
        with frozen:
            if moved:
                sign, logabs, z1 = model.first_layer(fxs)
            else:
                sign, logabs = model.forward_log(fxs)
                sign, logabs = sign.clone(), logabs.clone()
            n_acc = torch.zeros(fxs.shape[0], dtype=torch.int64, device=fxs.device)
            # logits at the current walkers: evaluated once, then carried over
            # from the proposed configurations of the accepted moves
            lx = self.k.lp_logits(self.params, self.C, self.Lx, self.Ly, fxs,
                                  precision=self.precision) if n_steps else None
            acc = None
            z1p = None
            for s in range(n_steps):
                uc = ua = None
                if self.moves_per_proposal == 2:
                    # two chained moves: draws 2c (m1) and 2c + 1 (m2, and the
                    # acceptance's other Philox words)
                    uc2 = None
                    if draws is not None:
                        uc, uc2, ua = draws[s]
                    c2 = 2 * self.step_counter
                    m1, lqf1, x1, lqr1, lp1 = self.k.lp_propose_given(
                        self.params, self.C, self.Lx, self.Ly, fxs, lx, self.seed, c2, uc,
                        precision=self.precision)
                    m2, lqf2, prop, lqr2, lp = self.k.lp_propose_given(
                        self.params, self.C, self.Lx, self.Ly, x1, lp1, self.seed, c2 + 1, uc2,
                        precision=self.precision)
                    move = torch.where(m1 < 0, m1, m2)
                    lqf, lqr = lqf1 + lqf2, lqr1 + lqr2
                    if moved:
                        ps, pl, z1p = model.forward_log_moved(fxs, prop, z1)
                    else:
                        ps, pl = model.forward_log(prop)
                    if acc is None:
                        acc = torch.empty(fxs.shape[0], dtype=torch.uint8, device=fxs.device)
                    self.k.lp_accept_commit(fxs, sign, logabs, move, prop, ps, pl, lqf, lqr,
                                            self.seed, c2 + 1, lp, lx, acc, n_accept=n_acc,
                                            u_acc=ua, z1p=z1p if moved else None,
                                            z1=z1 if moved else None)
                    self.step_counter += 1
                    continue
                if draws is not None:
                    uc, ua = draws[s]
                move, lqf, prop, lqr, lp = self.k.lp_propose_given(
                    self.params, self.C, self.Lx, self.Ly, fxs, lx, self.seed, self.step_counter,
                    uc, precision=self.precision)
                if moved:
                    ps, pl, z1p = model.forward_log_moved(fxs, prop, z1)
                else:
                    ps, pl = model.forward_log(prop)
                if acc is None:
                    acc = torch.empty(fxs.shape[0], dtype=torch.uint8, device=fxs.device)
                self.k.lp_accept_commit(fxs, sign, logabs, move, prop, ps, pl, lqf, lqr,
                                        self.seed, self.step_counter, lp, lx, acc,
                                        n_accept=n_acc, u_acc=ua, z1p=z1p if moved else None,
                                        z1=z1 if moved else None)
                self.step_counter += 1
        if hasattr(model, "invalidate"):
            model.invalidate()      # the model's cached walker state predates the moves
        self.last_acceptance = (n_acc, n_steps)
        return sign, logabs

    @torch.inference_mode()
    def step(self, fxs, model, graph=None, compile=False, verbose=False, use_log_amp=False,
             **kwargs):
        sign, logabs = self.sweeps(fxs, model, self.proposals_per_step)
        if use_log_amp:
            return fxs, (sign, logabs)
        return fxs, sign * torch.exp(logabs)

    def burn_in(self, fxs, model, graph=None, n_steps=1, **kwargs):
        self.sweeps(fxs, model, int(n_steps))
        return fxs

    @torch.inference_mode()
    def train_step(self, fxs, model, lr=1e-2):
        """One gradient step of the heat-bath cross entropy on the walkers
        `fxs`; returns the loss before the step."""
        if self.params.device != fxs.device:
            self.params = self.params.to(fxs.device)
        B = fxs.shape[0]
        logits = self.k.lp_logits(self.params, self.C, self.Lx, self.Ly, fxs,
                                  precision=self.precision)
        valid = torch.isfinite(logits)
        cand = self.k.lp_candidates(self.Lx, self.Ly, fxs)
        _, lc = model.forward_log(cand.reshape(-1, fxs.shape[1]))
        target = torch.where(valid, 2.0 * lc.reshape(B, -1), torch.full_like(logits, -torch.inf))
        qstar = torch.softmax(target, dim=1)
        logq = torch.log_softmax(logits, dim=1)
        loss = -(qstar * torch.where(valid, logq, torch.zeros_like(logq))).sum() / B
        dl = torch.where(valid, (torch.exp(logq) - qstar) / B, torch.zeros_like(logits))
        grad = self.k.lp_backward(self.params, self.C, self.Lx, self.Ly, fxs, dl.contiguous())
        self.k.lp_axpy(-lr, grad, self.params)
        return float(loss)
