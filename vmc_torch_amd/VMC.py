"""VMC driver with the reference's interface (GPU/VMC.py:151-847), on one
GPU: sampler.step -> evaluate_energy_fn -> compute_grads_fn ->
preconditioner.solve -> optimizer.step, with the walker batch on the device.
"""
import time
import warnings
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

    @classmethod
    def from_vmc_config(cls, cfg, *, n_params, nsites):
        """GPU/VMC.py:120-148: every field of a run-script config dataclass
        (plus its extra attributes), with resume_step -> step_offset and
        sr_diag_shift -> diag_shift; n_params and nsites are given."""
        import dataclasses
        names = {f.name for f in dataclasses.fields(cfg)}
        kwargs = {name: getattr(cfg, name) for name in names}
        for k, v in vars(cfg).items():
            if k not in names and k not in kwargs:
                kwargs[k] = v
        if hasattr(cfg, "resume_step"):
            kwargs["step_offset"] = cfg.resume_step
        if "sr_diag_shift" in kwargs:
            kwargs["diag_shift"] = kwargs.pop("sr_diag_shift")
        kwargs["n_params"] = n_params
        kwargs["nsites"] = nsites
        return cls(**kwargs)


def factored_row_norms(F, factors, blocks):
    """|O_s| of a factored O (kernels.FactoredO layout) from its factors: a
    plain block contributes scale^2 |F_a,s|^2, an outer-product block
    scale^2 |F_a,s|^2 |F_b,s|^2 and its bias block scale^2 |F_a,s|^2."""
    sq = [(F[:, c:c + k] ** 2).sum(dim=1) for c, k in factors]
    n2 = torch.zeros(F.shape[0], dtype=F.dtype, device=F.device)
    for a, b, sc, _wo, bo in blocks:
        term = sq[a] if b < 0 else sq[a] * (sq[b] + (1.0 if bo >= 0 else 0.0))
        n2 += (sc * sc) * term
    return n2.sqrt()


def outlier_mask(E, O, clip_factor):
    """GPU/VMC.py:476-498: keep the samples whose per-sample log-derivative
    norm is at most clip_factor times the median norm (torch's median: the
    lower middle value); the dropped rows of O become 0 and their energies
    the mean of the kept ones.  O is a dense (Ns, Np) tensor or a
    kernels.FactoredO, whose row norms come from its factors (an
    outer-product block contributes |a|^2 |b|^2, its bias block |a|^2) and
    whose dropped rows are zeroed in every factor.  Returns (E, O, n_dropped);
    E and O are copies when rows are dropped, the inputs otherwise."""
    if clip_factor <= 0:
        return E, O, 0
    if isinstance(O, kernels.FactoredO):
        norms = factored_row_norms(O.F, O.factors, O.blocks)
    else:
        norms = O.norm(dim=1)
    mask = norms <= clip_factor * norms.median()
    n_drop = int((~mask).sum())
    if n_drop == 0:
        return E, O, 0
    E = E.clone()
    E[~mask] = E[mask].mean()
    if isinstance(O, kernels.FactoredO):
        F = O.F.clone()
        F[~mask] = 0.0
        O = kernels.FactoredO(F, O.factors, O.blocks, O.n_params)
    else:
        O = O.clone()
        O[~mask] = 0.0
    return E, O, n_drop


def _all_finite(*tensors):
    """One device-side reduction per tensor, one host read in total."""
    flags = [torch.isfinite(t).all() for t in tensors if t is not None and t.numel()]
    return bool(torch.stack(flags).all()) if flags else True


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
        self.last_n_dropped = 0

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
            if not _all_finite(O.F):
                # GPU/vmc_utils.py:1330-1386 (_check_grads_amps)
                raise ValueError("NaN/Inf in gradients (factors of the log-derivative)")
        else:
            O = torch.cat(O_list) if len(O_list) > 1 else O_list[0]
        E, O, self.last_n_dropped = outlier_mask(E, O, outlier_clip_factor)
        times = {"t_samp": t_samp, "t_locE": t_locE, "t_grad": t_grad}
        return (E, O), fxs, t_samp + t_locE + t_grad, times

    def compute_global_energy_stats(self, local_energies, world_size=1):
        """GPU/VMC.py:513-534 on one device: (n, mean, variance), the sums
        compensated (vmcx_energy_stats) where the reference's are naive."""
        if world_size != 1:
            raise NotImplementedError("one GPU: the walker batch stands in for the ranks")
        E = local_energies.to(torch.float64).contiguous()
        st = kernels.energy_stats(E).cpu()
        return E.numel(), float(st[0]), float(st[1])

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
            if emean != emean:
                # GPU/VMC.py:705-723 reports NaN inputs of the SR step
                warnings.warn(f"[WARNING] SR inputs: local_E has NaN at step "
                              f"{step + config.step_offset}")
            dp, t_sr, info = self.solve_sr_step(O, E, emean, n, config.n_params,
                                                config.diag_shift, device, config.run_sr)
            lr = config.learning_rate
            if config.lr_scheduler is not None:
                lr = config.learning_rate = config.lr_scheduler(step + config.step_offset)
            dp_t = torch.as_tensor(dp, device=device)
            skipped = not _all_finite(dp_t)
            if skipped:
                # GPU/VMC.py:757-769 warns about a NaN/Inf dp (and the CPU driver,
                # VMC.py:173-175, raises); here the update is skipped so the
                # parameters stay finite
                warnings.warn(f"[WARNING] SR dp has NaN/Inf at step {step + config.step_offset}:"
                              " parameter update skipped")
            else:
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
                             "solver_info": info, "update_skipped": skipped,
                             "n_outliers": self.last_n_dropped, "fxs": fxs,
                             "local_energies": E})
        return history, fxs
