"""Environment-reuse entry points on the CPU: argument checks that fail
before any device access, and the host classes' type checks."""
import ctypes

import pytest
import torch

from vmc_torch_amd import (MetropolisExchangeSpinfulSamplerReuse_GPU, _lib, fPEPS_Model_GPU,
                           evaluate_energy_reuse, kernels)


def test_reuse_refuses_configuration_dependent_tensors():
    lib = kernels.lib()
    p = ctypes.c_void_p(16)   # never dereferenced: the model check fails first
    # an NN-fTNS model struct (hidden > 0): theta(x) changes with x
    m = _lib.PepsModelStruct(Lx=2, Ly=2, D=2, n_params=32, dense_map=p, theta=p, hidden=4,
                             act=0, eta=0.5, W1=p, b1=p, W2=p, b2=p, Wo=p, bo=p, mode=0,
                             reserved=0)
    rc = lib.vmcx_peps_local_energy_reuse(ctypes.byref(m), 1, 1, p, 1.0, 8.0, p, p, p, p,
                                          1 << 20, 1, p, None)
    assert rc == 1 and "hidden = 0" in lib.vmcx_last_error().decode()
    edges = (ctypes.c_int32 * 2)(0, 1)
    rc = lib.vmcx_peps_sweep_reuse(ctypes.byref(m), 1, 1, ctypes.cast(edges, ctypes.c_void_p),
                                   0.25, 1, 0, 0, None, None, None, p, p, p, p, p, 1 << 20,
                                   None)
    assert rc == 1 and "hidden = 0" in lib.vmcx_last_error().decode()
    # the environment cache: (Ly + 1) left and right PB x PB blocks per walker
    plain = _lib.PepsModelStruct(Lx=2, Ly=2, D=2, n_params=32, dense_map=p, theta=p)
    extra = lib.vmcx_peps_reuse_workspace_bytes(ctypes.byref(plain), 4, 4, 4) - \
        lib.vmcx_peps_workspace_bytes(ctypes.byref(plain), 4, 4, 4)
    assert extra >= 4 * 2 * 3 * 16 * 16 * 8


def test_reuse_host_classes_check_the_model():
    m = fPEPS_Model_GPU(2, 2, 2)
    x = torch.zeros(1, 4, dtype=torch.int64)
    with pytest.raises(TypeError, match="fPEPS_Model_reuse_GPU"):
        MetropolisExchangeSpinfulSamplerReuse_GPU().sweeps(x, m, None)
    with pytest.raises(TypeError, match="fPEPS_Model_reuse_GPU"):
        evaluate_energy_reuse(x, m, None)
