"""Metropolis samplers with the reference's interface (GPU/sampler.py).

MetropolisExchangeSpinfulSamplerGPU keeps the proposal semantics of
propose_exchange_or_hopping_vec (GPU/sampler.py:47-137) and the sweep of
step (GPU/sampler.py:216-388): edges in row_edges-then-col_edges order, one
proposal per edge per walker, acceptance u < |psi'|^2/|psi|^2.  The whole
sweep runs in one HIP launch per sweep with Sherman-Morrison updates; the
draws come from Philox4x32-10 keyed by (seed, sweep, edge, walker).
"""
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
