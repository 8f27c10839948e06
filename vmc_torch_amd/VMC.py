"""VMC driver with the reference's interface (GPU/VMC.py:151-847), on one
GPU: sampler.step -> evaluate_energy_fn -> compute_grads_fn ->
preconditioner.solve -> optimizer.step, with the walker batch on the device.
"""
import time
from dataclasses import dataclass, fields
from typing import Any, Callable, Dict, Optional

import torch

from . import kernels
from .vmc_utils import compute_grads_gpu, evaluate_energy


@dataclass
class VMCLoopConfig:
    """GPU/VMC.py:88-148."""
    vmc_steps: int = 1
    ns_per_rank: int = 1
    grad_batch_size: int = 0
    n_params: int = 0
    nsites: int = 1
    learning_rate: float = 0.1
    diag_shift: float = 1e-4
    burn_in_steps: int = 0
    run_sr: bool = True
    use_log_amp: bool = True
    outlier_clip_factor: float = 100.0
    lr_scheduler: object = None
    show_progress: bool = False
    step_offset: int = 0

    def __init__(self, **kwargs):
        for f in fields(self):
            setattr(self, f.name, kwargs.pop(f.name, f.default))
        for k, v in kwargs.items():
            setattr(self, k, v)


class VMC_GPU:
    """GPU/VMC.py:151-172: dependency-injected driver."""

    def __init__(self, sampler, preconditioner=None, optimizer=None,
                 evaluate_energy_fn: Callable = evaluate_energy,
                 compute_grads_fn: Callable = compute_grads_gpu):
        self.sampler = sampler
        self.preconditioner = preconditioner
        self.optimizer = optimizer
        self.evaluate_energy_fn = evaluate_energy_fn
        self.compute_grads_fn = compute_grads_fn

    def _run_sampling_phase(self, fxs, model, hamiltonian, graph, ns_per_rank,
                            grad_batch_size=0, burn_in=False, burn_in_steps=0,
                            outlier_clip_factor=0.0, use_log_amp=True, **kw):
        """GPU/VMC.py:316-511: sweep, E_loc, O per chunk of B samples."""
        B = fxs.shape[0]
        t_samp = t_locE = t_grad = 0.0
        factored = (getattr(self.preconditioner, "factored", False)
                    and hasattr(model, "log_derivative_factored"))
        if getattr(self.preconditioner, "factored", False) and not factored:
            raise TypeError("the factored SR step needs a model with log_derivative_factored")
        if burn_in and burn_in_steps:
            fxs = self.sampler.burn_in(fxs, model, graph, burn_in_steps)
        E_list, O_list = [], []
        count = 0
        while count < ns_per_rank:
            needed = min(B, ns_per_rank - count)
            t0 = time.time()
            fxs, amps = self.sampler.step(fxs, model, graph, use_log_amp=use_log_amp)
            t1 = time.time()
            _, E = self.evaluate_energy_fn(fxs, model, hamiltonian, amps, use_log_amp=use_log_amp)
            t2 = time.time()
            if factored:
                O, aux = model.log_derivative_factored(fxs)
            else:
                O, aux = self.compute_grads_fn(fxs, model, batch_size=grad_batch_size,
                                               use_log_amp=use_log_amp)
                if not use_log_amp:
                    O = O / aux.unsqueeze(1)
            t3 = time.time()
            t_samp += t1 - t0
            t_locE += t2 - t1
            t_grad += t3 - t2
            E_list.append(E[:needed])
            O_list.append(O[:needed])
            count += needed
        E = torch.cat(E_list) if len(E_list) > 1 else E_list[0]
        if factored:
            O = kernels.FactoredO.cat(O_list) if len(O_list) > 1 else O_list[0]
            if outlier_clip_factor > 0:
                raise NotImplementedError("outlier clipping needs the dense O (gram != 'factored')")
        else:
            O = torch.cat(O_list) if len(O_list) > 1 else O_list[0]
        if outlier_clip_factor > 0 and not factored:
            # GPU/VMC.py:476-498
            norms = O.norm(dim=1)
            mask = norms <= outlier_clip_factor * norms.median()
            if not bool(mask.all()):
                O = O.clone()
                E = E.clone()
                O[~mask] = 0.0
                E[~mask] = E[mask].mean()
        times = {"t_samp": t_samp, "t_locE": t_locE, "t_grad": t_grad}
        return (E, O), fxs, t_samp + t_locE + t_grad, times

    def compute_global_energy_stats(self, local_energies, world_size=1):
        """GPU/VMC.py:513-534 (single device)."""
        n = local_energies.numel() * world_size
        s = local_energies.sum()
        s2 = (local_energies * local_energies).sum()
        if world_size > 1 and torch.distributed.is_initialized():
            torch.distributed.all_reduce(s)
            torch.distributed.all_reduce(s2)
        mean = float(s) / n
        return n, mean, float(s2) / n - mean * mean

    def solve_sr_step(self, local_o, local_energies, energy_mean, total_samples, n_params,
                      diag_shift, device, run_sr):
        return self.preconditioner.solve(
            local_o=local_o, local_energies=local_energies, energy_mean=energy_mean,
            total_samples=total_samples, n_params=n_params, diag_shift=diag_shift,
            device=device, run_sr=run_sr)

    def apply_parameter_update(self, model, dp, learning_rate, device):
        self.optimizer.step(model=model, direction=dp, device=device, learning_rate=learning_rate)

    def run_vmc_loop(self, fxs, model, hamiltonian, graph, rank=0, world_size=1,
                     config: VMCLoopConfig = None,
                     on_step_end: Optional[Callable[[Dict[str, Any]], None]] = None):
        """GPU/VMC.py:645-847."""
        device = next(model.parameters()).device
        history = []
        for step in range(config.vmc_steps):
            t0 = time.time()
            (E, O), fxs, sample_time, times = self._run_sampling_phase(
                fxs, model, hamiltonian, graph, config.ns_per_rank, config.grad_batch_size,
                burn_in=(step == 0), burn_in_steps=config.burn_in_steps,
                outlier_clip_factor=config.outlier_clip_factor, use_log_amp=config.use_log_amp)
            n, emean, evar = self.compute_global_energy_stats(E, world_size)
            dp, t_sr, info = self.solve_sr_step(O, E, emean, n, config.n_params,
                                                config.diag_shift, device, config.run_sr)
            lr = config.learning_rate
            if config.lr_scheduler is not None:
                lr = config.learning_rate = config.lr_scheduler(step + config.step_offset)
            self.apply_parameter_update(model, dp, lr, device)
            e_site = emean / config.nsites
            err = (max(evar, 0.0) / n) ** 0.5 / config.nsites
            history.append(e_site)
            if on_step_end is not None:
                on_step_end({"step": step + config.step_offset, "energy_mean": emean,
                             "energy_var": evar, "energy_per_site": e_site,
                             "error_per_site": err, "total_samples": n,
                             "sample_time": sample_time, "phase_times": times,
                             "sr_time": t_sr, "total_time": time.time() - t0,
                             "solver_info": info, "fxs": fxs, "local_energies": E})
        return history, fxs
