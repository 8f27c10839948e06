"""The allocation-free MINRES entry point (vmcx_sr_minres_noalloc) against
the existing one (vmcx_sr_minres, caller-supplied workspace), both called
through the C-ABI: two calls in a row, the second on a smaller problem, give
bit-identical dp, stats and info (same kernels in the same order, only the
workspace address differs); a problem beyond the fixed scratch is refused
before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from vmc_torch_amd import _lib

pytestmark = pytest.mark.gpu


def _call(lib, noalloc, O, E, shift, maxiter, ws=None):
    Ns, Np = O.shape
    dp = torch.empty(Np, dtype=torch.float64, device=O.device)
    stats = torch.empty(5, dtype=torch.float64, device=O.device)
    info = torch.empty(1, dtype=torch.int32, device=O.device)
    st = _lib.stream_ptr(O.device)
    args = (Ns, Np, _lib.ptr(O), 0 if O.dtype == torch.float64 else 1, O.stride(0),
            _lib.ptr(E), float("nan"), shift, 1e-10, maxiter)
    if noalloc:
        rc = lib.vmcx_sr_minres_noalloc(*args, _lib.ptr(dp), _lib.ptr(stats), _lib.ptr(info), st)
    else:
        rc = lib.vmcx_sr_minres(*args, _lib.ptr(ws), ws.numel(), _lib.ptr(dp), _lib.ptr(stats),
                                _lib.ptr(info), st)
    _lib.check(rc)
    return dp, stats, info


def test_noalloc_minres_matches_workspace_entry_point(device):
    lib = _lib.lib()
    rng = np.random.default_rng(5)
    runs = []
    for Ns, Np, dt in ((700, 3000, torch.float64), (130, 257, torch.float32)):
        print(f"noalloc minres Ns={Ns} Np={Np} {dt}", flush=True)
        O = torch.tensor(rng.normal(size=(Ns, Np)) * 0.3 + rng.normal(size=Np), dtype=dt,
                         device=device)
        E = torch.tensor(rng.normal(size=Ns) + 2.0, device=device)
        need = lib.vmcx_sr_minres_workspace_bytes(Ns, Np)
        assert 0 < need <= lib.vmcx_sr_minres_noalloc_max_bytes()
        ws = torch.empty(need, dtype=torch.uint8, device=device)
        want = _call(lib, False, O, E, 1e-3, 80, ws)
        got = _call(lib, True, O, E, 1e-3, 80)
        torch.cuda.synchronize(device)
        runs.append((want, got))
    for want, got in runs:
        for a, b in zip(want, got):
            assert torch.equal(a, b)
        assert int(got[2].item()) == 0          # converged


def test_noalloc_minres_refuses_oversize(device):
    lib = _lib.lib()
    Ns, Np = 64, 1 << 15
    assert lib.vmcx_sr_minres_workspace_bytes(Ns, Np) > lib.vmcx_sr_minres_noalloc_max_bytes()
    p = ctypes.c_void_p(16)   # never dereferenced: the size check fails first
    rc = lib.vmcx_sr_minres_noalloc(Ns, Np, p, 0, Np, p, float("nan"), 1e-3, 1e-8, 10, p, None,
                                    None, None)
    assert rc == 1 and "20 MiB" in lib.vmcx_last_error().decode()
