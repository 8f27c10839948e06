"""FP64 matrix-core SYRK (vmcx_syrk_f64) against float64 products on the
host.  Tolerance: rtol 1e-12 relative to sum_k |M_ik M_jk| (FP64 FMA
accumulation in a different order)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _bound(Mn):
    A = np.abs(Mn)
    return A @ A.T


@pytest.mark.parametrize("n,K,rowmajor,splits", [
    (128, 16, True, 1), (1, 1, True, 1), (130, 37, True, 1), (300, 500, True, 1),
    (300, 500, False, 1), (257, 128, False, 1), (200, 1000, True, 3), (64, 5, True, 4),
    (384, 4096, True, 2)])
def test_syrk_matches_host(n, K, rowmajor, splits):
    from vmc_torch_amd import kernels as K_
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(n * 7919 + K)
    Mh = torch.randn(n, K, generator=g, dtype=torch.float64)
    C0 = torch.randn(n, n, generator=g, dtype=torch.float64)
    alpha, beta = {(True, 1): (0.75, 0.0), (True, 3): (-1.0, 0.5), (False, 1): (-1.0, 1.0)}.get((rowmajor, splits), (0.5, 2.0) if rowmajor else (-1.0, 1.0))
    if rowmajor:
        Md = Mh.to(dev)
    else:
        Md = Mh.t().contiguous().to(dev)          # column-major n x K
    shape = (n, n) if splits == 1 else (splits, n, n)
    Cd = C0.expand(*shape).contiguous().to(dev)   # row-major [n, n] = C^T
    K_.syrk_f64(Md, Cd, alpha=alpha, beta=beta, splits=splits, rowmajor=rowmajor)
    torch.cuda.synchronize()
    Ct = Cd.cpu().numpy()
    Mn = Mh.numpy()
    kc = -(-K // splits)
    lower = np.tril(np.ones((n, n), dtype=bool))
    for s in range(splits):
        Ms = Mn[:, s * kc:(s + 1) * kc]
        want = beta * C0.numpy().T + alpha * (Ms @ Ms.T)   # column-major C, lower part
        got = (Ct if splits == 1 else Ct[s]).T
        err = np.abs(got - want)[lower]
        tol = 1e-12 * (abs(alpha) * _bound(Ms) + abs(beta) * np.abs(C0.numpy().T) + 1e-300)[lower]
        assert np.all(err <= tol), (s, err.max())
        # strictly upper triangle untouched
        assert np.array_equal(got[~lower], C0.numpy().T[~lower])


def test_syrk_strided_rows():
    """Row-major M as a column slice of a wider array (ldx > K)."""
    from vmc_torch_amd import kernels as K_
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(5)
    W = torch.randn(200, 96, generator=g, dtype=torch.float64)
    M = W[:, 10:60]
    Cd = torch.zeros(200, 200, dtype=torch.float64, device=dev)
    K_.syrk_f64(W.to(dev)[:, 10:60], Cd)
    torch.cuda.synchronize()
    Mn = M.numpy()
    lower = np.tril(np.ones((200, 200), dtype=bool))
    assert np.allclose(Cd.cpu().numpy().T[lower], (Mn @ Mn.T)[lower], rtol=0, atol=1e-12 * 50)


def test_syrk_rejects_bad_args():
    from vmc_torch_amd import _lib
    lib = _lib.lib()
    # ldc < n and too many splits: errors before any launch
    assert lib.vmcx_syrk_f64(10, 4, 0, 4, 1, 1.0, 0.0, 1, 5, 1, 0, 0) != 0
    assert lib.vmcx_syrk_f64(10, 4, 1, 4, 1, 1.0, 0.0, 1, 10, 70000, 100, 0) != 0
