"""SR preconditioners and parameter updates with the reference's interface
(GPU/optimizer.py: PreconditionerGPU.solve, OptimizerGPU.step)."""
import time
from typing import Any, Optional, Tuple

import torch

from . import kernels


class Scheduler:
    """GPU/optimizer.py:26-127 (callable step -> learning rate)."""

    def __call__(self, step):
        raise NotImplementedError


class TrivialScheduler(Scheduler):
    def __init__(self, init_lr):
        self.init_lr = init_lr

    def __call__(self, step):
        return self.init_lr


class DecayScheduler(Scheduler):
    """lr = init_lr * decay_rate ** (step // patience), floored at min_lr."""

    def __init__(self, init_lr, decay_rate=0.9, patience=50, min_lr=1e-4):
        self.init_lr, self.decay_rate = init_lr, decay_rate
        self.patience, self.min_lr = patience, min_lr

    def __call__(self, step):
        return max(self.init_lr * self.decay_rate ** (step // self.patience), self.min_lr)


class OptimizerGPU:
    """GPU/optimizer.py:130-175: params updated in place, in parameters() order."""

    def __init__(self, learning_rate: float = 1e-3):
        self.lr = learning_rate

    def step(self, model, direction, device=None, learning_rate: Optional[float] = None):
        raise NotImplementedError

    def reset(self):
        pass


class SGDGPU(OptimizerGPU):
    """p <- p - lr * dp (GPU/optimizer.py:178-186) as one HIP axpy per
    parameter tensor."""

    @torch.no_grad()
    def step(self, model, direction, device=None, learning_rate=None):
        lr = self.lr if learning_rate is None else learning_rate
        params = list(model.parameters())
        dev = params[0].device
        d = torch.as_tensor(direction, device=dev, dtype=torch.float64).contiguous()
        total = sum(p.numel() for p in params)
        if d.numel() != total:
            # checked before any parameter is touched
            raise ValueError(f"direction has {d.numel()} entries, parameters {total}")
        off = 0
        for p in params:
            n = p.numel()
            kernels.sgd_update(p.data.view(-1), d[off:off + n], lr)
            off += n
        if hasattr(model, "invalidate"):
            model.invalidate()


class PreconditionerGPU:
    """GPU/optimizer.py:236-251."""

    def solve(self, *, local_o, local_energies, energy_mean: float, total_samples: int,
              n_params: int, diag_shift: float, device, run_sr: bool) -> Tuple[Any, float, Any]:
        raise NotImplementedError


class DistributedSRMinresGPU(PreconditionerGPU):
    """Parameter-space SR by MINRES (GPU/optimizer.py:846-882, solver
    :257-465, torch_minres GPU/torch_utils.py:747-857), on one device.

    The matvec O^T(O x)/Ns - (Obar.x) Obar + shift x runs in libvmcx with
    FP64 accumulation; the Lanczos scalars stay on the device, so the solve
    does not synchronise the host per iteration.  run_sr=False returns the
    energy gradient g = <E O> - <E><O>.  use_scipy is accepted and ignored
    (there is no host path).  Returns (dp, seconds, info) with info 0 when
    converged and maxiter otherwise.
    """

    def __init__(self, rtol: float = 5e-5, maxiter: int = 100, use_scipy: bool = False,
                 sync_info: bool = True):
        self.rtol = rtol
        self.maxiter = maxiter
        self.use_scipy = use_scipy
        self.sync_info = sync_info
        self.workspace = kernels.MinresWorkspace()
        self.last_stats = None

    def solve(self, *, local_o, local_energies, energy_mean, total_samples, n_params,
              diag_shift, device, run_sr=True):
        t0 = time.time()
        dt = local_o.dtype if local_o.dtype in (torch.float32, torch.float64) else torch.float64
        O = local_o.to(device=device, dtype=dt).contiguous()
        E = local_energies.to(device=device, dtype=torch.float64).contiguous()
        if O.shape[0] != total_samples or O.shape[1] != n_params:
            raise ValueError("local_o must be (total_samples, n_params) on one GPU")
        dp, stats, info = kernels.sr_minres(O, E, energy_mean, diag_shift, self.rtol,
                                            self.maxiter if run_sr else 0, self.workspace)
        self.last_stats = stats
        if not run_sr:
            return dp, time.time() - t0, None
        info_v = int(info.item()) if self.sync_info else info
        return dp, time.time() - t0, info_v


class MinSRGPU(PreconditionerGPU):
    """Sample-space SR (GPU/optimizer.py:471-614, MinSRGPU at :873-895).

    gram: "fp64" (plain form), "fp32", "bf16" or "bf16x3" (Gram operands in
    reduced precision with FP32 accumulation; the FP64 Cholesky factor of
    that Gram then preconditions conjugate gradients on the exact operator,
    at most `max_iter` iterations to relative residual `tol`), or "factored"
    (the same FP64 step as "fp64" for models made of dense layers: local_o
    is a kernels.FactoredO from model.log_derivative_factored and the Gram
    is formed from the Grams of its factors, no Ns x Np matrix).
    inplace: centre O in place (O is overwritten; saves an Ns x Np copy).
    Returns (dp device tensor, seconds, info) with info = 0 on success.
    When the Cholesky factorisation fails (a non-positive leading minor, the
    reference's cue to fall back to lstsq, GPU/optimizer.py:557-561), the
    step is solved again with the shift raised tenfold (a shift <= 0 is
    first replaced by max(10 |shift|, 1e-8)), at most
    `max_retries` times, and info = 1 as for the reference's fallback;
    last_fallback records the failing minor and the shift that succeeded
    (None if every retry failed: then dp is NaN, and the driver skips the
    update).  run_sr=False returns the energy gradient
    g = <E O> - <E><O> (GPU/optimizer.py:581-614).
    """

    def __init__(self, gram="fp64", max_iter=None, tol=1e-8, sync_info=True, inplace=False,
                 max_retries=3):
        if gram not in kernels.GRAM_CODES and gram != "factored":
            raise ValueError(f"unknown gram {gram!r}")
        self.gram = gram
        # "factored": the driver hands over kernels.FactoredO (models with
        # log_derivative_factored); the FP64 Gram is formed from the factors
        self.factored = gram == "factored"
        self.factored_workspace = kernels.SRFactoredWorkspace()
        self.max_iter = (0 if gram in ("fp64", "factored") else 20) if max_iter is None else max_iter
        self.tol = tol
        self.sync_info = sync_info
        self.inplace = inplace
        self.workspace = kernels.SRWorkspace()
        self.minres_workspace = None
        self.max_retries = max_retries
        self.last_stats = None
        self.last_fallback = None

    def solve(self, *, local_o, local_energies, energy_mean, total_samples, n_params,
              diag_shift, device, run_sr=True):
        t0 = time.time()
        E = local_energies.to(device=device, dtype=torch.float64).contiguous()
        if isinstance(local_o, kernels.FactoredO):
            if local_o.shape != (total_samples, n_params):
                raise ValueError("local_o must be (total_samples, n_params) on one GPU")
            if not run_sr:
                g, stats = kernels.sr_factored_gradient(local_o, E, energy_mean,
                                                        self.factored_workspace)
                self.last_stats = stats
                return g, time.time() - t0, None

            def run(shift):
                return kernels.sr_minsr_factored(local_o, E, energy_mean, shift,
                                                 self.factored_workspace)
        else:
            if self.factored:
                raise TypeError("gram='factored' takes a kernels.FactoredO "
                                "(model.log_derivative_factored)")
            dt = local_o.dtype if local_o.dtype in (torch.float32, torch.float64) \
                else torch.float64
            if self.gram == "fp64" or not run_sr:
                dt = torch.float64
            O = local_o.to(device=device, dtype=dt).contiguous()
            if O.shape[0] != total_samples or O.shape[1] != n_params:
                raise ValueError("local_o must be (total_samples, n_params) on one GPU")
            if not run_sr:
                # the MINRES entry point with no iterations returns g
                if self.minres_workspace is None:
                    self.minres_workspace = kernels.MinresWorkspace()
                g, stats, _ = kernels.sr_minres(O, E, energy_mean, 0.0, maxiter=0,
                                                workspace=self.minres_workspace)
                self.last_stats = stats
                return g, time.time() - t0, None

            def run(shift):
                return kernels.sr_minsr(O, E, energy_mean, shift, self.gram, self.max_iter,
                                        self.tol, self.workspace, inplace=self.inplace)
        dp, stats, info = run(diag_shift)
        self.last_stats = stats
        self.last_fallback = None
        if not self.sync_info:
            return dp, time.time() - t0, info
        info_v = int(info.item())
        if info_v == 0:
            return dp, time.time() - t0, 0
        # non-positive leading minor: raise the shift (the reference falls back
        # to lstsq here); info 1 as for the reference's fallback
        failed, shift = info_v, diag_shift
        for _ in range(self.max_retries):
            shift = shift * 10.0 if shift > 0 else max(-10.0 * shift, 1e-8)
            dp, stats, info = run(shift)
            if int(info.item()) == 0:
                self.last_stats = stats
                self.last_fallback = {"cholesky_info": failed, "diag_shift": shift}
                return dp, time.time() - t0, 1
        dp = torch.full_like(dp, float("nan"))
        return dp, time.time() - t0, 1
