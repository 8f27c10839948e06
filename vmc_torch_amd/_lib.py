"""ctypes binding of libvmcx.so (the HIP kernels) and tensor-argument checks.

The library is loaded from this package directory only.  There is no CPU
fallback: if the library is missing, or a tensor is not on a HIP device,
every entry point raises.
"""
import ctypes
import os
import re

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvmcx.so")
HEADER = os.path.join(_HERE, "..", "include", "vmcx.h")

_c = ctypes
_P = _c.c_void_p
_I64 = _c.c_int64
_I = _c.c_int
_D = _c.c_double
_U64 = _c.c_uint64

_SIGNATURES = {
    "vmcx_abi_version": (_I, []),
    "vmcx_last_error": (_c.c_char_p, []),
    "vmcx_slater_nfp": (_I, [_I]),
    "vmcx_set_launch_waves": (_I, [_I, _I, _I]),
    "vmcx_get_launch_waves": (_I, [_I, _I]),
    "vmcx_slater_refresh": (_I, [_I64, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
    "vmcx_slater_sweep": (_I, [_I64, _I, _I, _P, _I, _P, _D, _I, _U64, _I64,
                               _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vmcx_counter_add": (_I, [_P, _I64, _P]),
    "vmcx_slater_local_energy": (_I, [_I64, _I, _I, _P, _I, _P, _D, _D, _P, _P,
                                      _P, _P, _P]),
    "vmcx_slater_logderiv": (_I, [_I64, _I, _I, _P, _P, _P, _I, _I64, _P]),
    "vmcx_slater_refresh_f32": (_I, [_I64, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P]),
    "vmcx_slater_sweep_f32": (_I, [_I64, _I, _I, _P, _P, _I, _P, _D, _I, _U64, _I64,
                                   _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "vmcx_slater_local_energy_f32": (_I, [_I64, _I, _I, _P, _I, _P, _D, _D, _P, _P,
                                          _P, _P, _P]),
    "vmcx_slater_logderiv_f32": (_I, [_I64, _I, _I, _P, _P, _P, _I, _I64, _P]),
    "vmcx_sgd_update_f32": (_I, [_I64, _P, _P, _D, _P]),
    "vmcx_hubbard_conn_count": (_I, [_I64, _I, _I, _P, _P, _P, _P]),
    "vmcx_hubbard_conn_fill": (_I, [_I64, _I, _I, _P, _D, _D, _P, _P, _P, _P,
                                    _P, _P]),
    "vmcx_sr_workspace_bytes": (_I64, [_I64, _I64, _I, _I, _I]),
    "vmcx_sr_minsr": (_I, [_I64, _I64, _P, _I, _I64, _I, _P, _D, _D, _I, _I, _D, _P, _I64,
                           _P, _P, _P, _P, _P]),
    "vmcx_sr_factored_workspace_bytes": (_I64, [_I64, _I, _P, _I, _P]),
    "vmcx_sr_minsr_factored": (_I, [_I64, _I, _P, _P, _P, _I, _P, _I64, _P, _D, _D, _P, _I64,
                                    _P, _P, _P, _P, _P]),
    "vmcx_sr_factored_gradient": (_I, [_I64, _I, _P, _P, _P, _I, _P, _I64, _P, _D, _P, _I64,
                                       _P, _P, _P]),
    "vmcx_energy_stats": (_I, [_I64, _P, _P, _P]),
    "vmcx_sr_minres_workspace_bytes": (_I64, [_I64, _I64]),
    "vmcx_sr_minres": (_I, [_I64, _I64, _P, _I, _I64, _P, _D, _D, _D, _I, _P, _I64, _P, _P,
                            _P, _P]),
    "vmcx_sr_minres_noalloc_max_bytes": (_I64, []),
    "vmcx_sr_minres_noalloc": (_I, [_I64, _I64, _P, _I, _I64, _P, _D, _D, _D, _I, _P, _P, _P,
                                    _P]),
    "vmcx_sgd_update": (_I, [_I64, _P, _P, _D, _P]),
    "vmcx_syrk_f64": (_I, [_I64, _I64, _P, _I64, _I, _D, _D, _P, _I64, _I, _I64, _P]),
    "vmcx_nnbf_workspace_bytes": (_I64, [_P, _I64, _I64, _I]),
    "vmcx_nnbf_forward": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _P, _P, _P]),
    "vmcx_nnbf_forward_moved": (_I, [_P, _I64, _P, _P, _P, _P, _I64, _I64, _P, _P, _P, _P]),
    "vmcx_masked_rows": (_I, [_I64, _I64, _P, _P, _P, _P]),
    "vmcx_nnbf_sweep": (_I, [_P, _I64, _I, _P, _D, _I, _U64, _I64, _P, _P, _P, _P, _P,
                             _P, _P, _P, _P, _I64, _P]),
    "vmcx_nnbf_local_energy": (_I, [_P, _I64, _I, _P, _D, _D, _P, _P, _P, _P, _I64, _I64,
                                    _P, _P]),
    "vmcx_nnbf_logderiv": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _I, _I64, _P, _P, _P]),
    "vmcx_nnbf_logderiv_factors": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _I64, _P, _P, _P]),
    "vmcx_to_bf16": (_I, [_I64, _P, _P, _P]),
    "vmcx_lp_n_params": (_I, [_I]),
    "vmcx_lp_n_moves": (_I, [_I, _I]),
    "vmcx_lp_logits": (_I, [_P, _I, _I, _I, _I64, _P, _P, _I, _P]),
    "vmcx_lp_propose": (_I, [_P, _I, _I, _I, _I64, _P, _U64, _I64, _P, _P, _P, _P, _P, _P, _I,
                             _P]),
    "vmcx_lp_accept": (_I, [_I64, _I, _U64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                            _P, _P, _P]),
    "vmcx_lp_accept_commit": (_I, [_I64, _I, _U64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                   _P, _P, _P, _I, _P, _P, _I64, _P, _P, _P]),
    "vmcx_lp_backward_workspace_bytes": (_I64, [_I, _I64]),
    "vmcx_lp_backward": (_I, [_P, _I, _I, _I, _I64, _P, _P, _P, _I64, _P, _P]),
    "vmcx_lp_candidates": (_I, [_I64, _I, _I, _P, _P, _P]),
    "vmcx_lp_axpy": (_I, [_I64, _c.c_float, _P, _P, _P]),
    "vmcx_lp_propose_given": (_I, [_P, _I, _I, _I, _I64, _P, _P, _U64, _I64, _P, _P, _P, _P,
                                   _P, _P, _I, _P, _P]),
    "vmcx_lp_commit_logits": (_I, [_I64, _I, _P, _P, _P, _P]),
    "vmcx_sym_images": (_I, [_I64, _I, _I, _P, _P, _P, _P, _P, _P]),
    "vmcx_peps_dense_size": (_I64, [_P]),
    "vmcx_peps_workspace_bytes": (_I64, [_P, _I64, _I64, _I]),
    "vmcx_peps_forward": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _P, _P]),
    "vmcx_peps_sweep": (_I, [_P, _I64, _I, _P, _D, _I, _U64, _I64, _P, _P, _P, _P, _P, _P,
                             _P, _P, _I64, _P]),
    "vmcx_peps_local_energy": (_I, [_P, _I64, _I, _P, _D, _D, _P, _P, _P, _P, _I64, _I64,
                                    _P, _P]),
    "vmcx_peps_reuse_workspace_bytes": (_I64, [_P, _I64, _I64, _I]),
    "vmcx_peps_sweep_reuse": (_I, [_P, _I64, _I, _P, _D, _I, _U64, _I64, _P, _P, _P, _P, _P,
                                   _P, _P, _P, _I64, _P]),
    "vmcx_peps_local_energy_reuse": (_I, [_P, _I64, _I, _P, _D, _D, _P, _P, _P, _P, _I64,
                                          _I64, _P, _P]),
    "vmcx_peps_logderiv": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _I, _I64, _P, _P, _P]),
    "vmcx_peps_logderiv_factors": (_I, [_P, _I64, _P, _P, _I64, _I64, _P, _I64, _P, _P, _P]),
}


class NnbfModelStruct(ctypes.Structure):
    """vmcx_nnbf_model (include/vmcx.h)."""
    _fields_ = [("n_sites", ctypes.c_int32), ("n_f", ctypes.c_int32),
                ("hidden", ctypes.c_int32), ("act", ctypes.c_int32),
                ("gemm", ctypes.c_int32), ("path", ctypes.c_int32),
                ("bf_scale", ctypes.c_double),
                ("M", _P), ("W1", _P), ("b1", _P), ("W2", _P), ("b2", _P),
                ("Wout", _P), ("bout", _P), ("W2_bf16", _P), ("Wout_bf16", _P),
                ("M64", _P), ("W1_64", _P), ("b1_64", _P), ("W2_64", _P), ("b2_64", _P),
                ("Wout_64", _P), ("bout_64", _P), ("det_fp32", ctypes.c_int32),
                ("reserved2", ctypes.c_int32)]


class SrBlockStruct(ctypes.Structure):
    """vmcx_sr_block (include/vmcx.h)."""
    _fields_ = [("a_factor", ctypes.c_int32), ("b_factor", ctypes.c_int32),
                ("scale", ctypes.c_double), ("w_offset", ctypes.c_int64),
                ("bias_offset", ctypes.c_int64)]


class PepsModelStruct(ctypes.Structure):
    """vmcx_peps_model (include/vmcx.h)."""
    _fields_ = [("Lx", ctypes.c_int32), ("Ly", ctypes.c_int32), ("D", ctypes.c_int32),
                ("n_params", ctypes.c_int32), ("dense_map", _P), ("theta", _P),
                ("hidden", ctypes.c_int32), ("act", ctypes.c_int32), ("eta", ctypes.c_double),
                ("W1", _P), ("b1", _P), ("W2", _P), ("b2", _P), ("Wo", _P), ("bo", _P),
                ("mode", ctypes.c_int32), ("amp_fp32", ctypes.c_int32),
                ("chi", ctypes.c_int32), ("bmps_nocache", ctypes.c_int32)]


_LIB = None


class VmcxError(RuntimeError):
    """An entry point of libvmcx returned an error status."""


def header_functions():
    """Names of the functions declared in include/vmcx.h."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(vmcx_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libvmcx.so once (raises ImportError when it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python -m vmc_torch_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().vmcx_last_error().decode(errors="replace")
        raise VmcxError(f"libvmcx status {rc}: {msg}")


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def require(t, dtype, name, shape=None, device=None):
    """Fail loudly unless t is a contiguous HIP tensor of the given dtype."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on a HIP device (got {t.device}); "
                         "libvmcx has no CPU path")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    return t
