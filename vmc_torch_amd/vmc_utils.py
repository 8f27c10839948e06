"""Walker initialisation, local energy and per-sample log-derivatives with
the reference's interface (GPU/vmc_utils.py, GPU/vmc_setup.py)."""
import torch

from . import kernels



def random_initial_config(N_f, N_sites, seed=None):
    """GPU/vmc_utils.py:22-39: pattern [1, 2] * (N/2), the first N - N_f
    sites emptied, then a random permutation (torch CPU generator)."""
    if seed is not None:
        torch.manual_seed(seed)
    cfg = torch.tensor([1, 2] * (N_sites // 2), dtype=torch.int64)
    cfg[: N_sites - N_f] = 0
    cfg = cfg[torch.randperm(N_sites)]
    n1, n2 = int((cfg == 1).sum()), int((cfg == 2).sum())
    if n1 != N_f // 2 or n2 != N_f // 2:
        raise AssertionError(f"Number of spin up and spin down fermions should be {N_f // 2}, "
                             f"but got {n1} and {n2}")
    return cfg


def initialize_walkers(init_fn, batch_size, seed=42, rank=0, device=None):
    """GPU/vmc_setup.py:139-159: walker i uses seed + rank*B + i."""
    rows = [torch.as_tensor(init_fn(seed=seed + rank * batch_size + i), dtype=torch.int64)
            for i in range(batch_size)]
    return torch.stack(rows).to(device)


def _check_model(model, what):
    if not hasattr(model, what):
        raise TypeError(f"{type(model).__name__} has no HIP {what} path")


@torch.inference_mode()
def evaluate_energy(fxs, model, H, current_amps=None, verbose=False, use_log_amp=False, **kwargs):
    """GPU/vmc_utils.py:479-489: returns (E_mean, E_loc[B]).

    E_loc(x) = sum_eta H[eta, x] psi(eta)/psi(x) is formed in libvmcx from the
    model's walker state; current_amps is accepted for interface
    compatibility (the state already carries sign and ln|psi|).
    """
    _check_model(model, "local_energy")
    eloc = model.local_energy(fxs, H)
    return kernels.energy_stats(eloc)[0], eloc


def evaluate_energy_reuse(fxs, model, H, current_amps=None, use_log_amp=False, verbose=False,
                          return_bMPS=False, **kwargs):
    """GPU/vmc_utils.py:584-1320 (evaluate_energy_reuse): (E_mean, E_loc[B])
    with every connected configuration evaluated from its walker's cached
    column environments (vmcx_peps_local_energy_reuse).  model: a
    models.fPEPS_Model_reuse_GPU.  The environments live inside the call, so
    return_bMPS adds None as the third element (the reference's cache
    object)."""
    from .models import fPEPS_Model_reuse_GPU
    if not isinstance(model, fPEPS_Model_reuse_GPU):
        raise TypeError(f"evaluate_energy_reuse needs an fPEPS_Model_reuse_GPU, got "
                        f"{type(model).__name__}")
    eloc = model.local_energy(fxs, H)
    mean = kernels.energy_stats(eloc)[0]
    return (mean, eloc, None) if return_bMPS else (mean, eloc)


def check_grads_amps(O, sign, logabs, configs=None):
    """GPU/vmc_utils.py:1330-1386 (_check_grads_amps): raise ValueError when
    the amplitudes or the per-sample gradients hold NaN or Inf (one device
    reduction per tensor and one host read).  An amplitude is its sign and
    ln|psi|; ln|psi| = -inf is a zero amplitude, which the log form clamps
    to ln(1e-45), so any non-finite value is an error."""
    bad_amp = ~(torch.isfinite(sign) & torch.isfinite(logabs))
    # a NaN or Inf anywhere in a row makes its sum non-finite; the row sums
    # avoid a full-size boolean temporary of O
    bad_o = ~torch.isfinite(O.sum(dim=1)) if O.numel() else torch.zeros_like(bad_amp)
    counts = torch.stack([bad_amp.sum(), bad_o.sum()]).cpu()
    if int(counts[0]):
        idx = bad_amp.nonzero()[:20, 0].tolist()
        where = "" if configs is None else "; first configs: " + "; ".join(
            str(configs[i].tolist()) for i in idx)
        raise ValueError(f"NaN/Inf in amplitudes: {int(counts[0])} out of "
                         f"{sign.numel()} samples{where}")
    if int(counts[1]):
        raise ValueError(f"NaN/Inf in gradients: {int(counts[1])}/{O.shape[0]} samples "
                         "affected")


def compute_grads_gpu(fxs, model, batch_size=None, verbose=False, offload_to_cpu=False,
                      use_log_amp=False, out_dtype=None, **kwargs):
    """GPU/vmc_utils.py:1389-1634: per-sample gradients in the flat
    parameter order.  With use_log_amp, returns (O = d ln|psi|/d theta,
    (signs, log_abs)); otherwise (d psi/d theta, psi) as the reference, whose
    caller divides by psi."""
    _check_model(model, "log_derivative")
    with torch.inference_mode():
        if out_dtype is None:
            O, (sign, logabs) = model.log_derivative(fxs)
        else:
            O, (sign, logabs) = model.log_derivative(fxs, dtype=out_dtype)
        check_grads_amps(O, sign, logabs, fxs)
        if use_log_amp:
            out = O, (sign, logabs)
        else:
            amps = sign * torch.exp(logabs)
            out = O * amps.unsqueeze(1).to(O.dtype), amps
    if offload_to_cpu:
        return (out[0].cpu(), out[1])
    return out
