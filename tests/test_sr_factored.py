"""CPU checks of minSR on a factored O (vmcx_sr_minsr_factored).

The device path forms the Gram of O from the Grams of its factors
(Kronecker identity, bias rows as a + 1), double-centres it and maps the
solution back one block at a time.  Here that algebra is restated in numpy
and held against the oracle's dense minSR (oracle/sr.py, which forms O and
centres its columns); a planted mistake in the restatement (the bias term
or the centring dropped) must be caught.  The C-ABI's block validation runs
without a device (it fails before any HIP call)."""
import ctypes
import math

import numpy as np
import pytest

from oracle import sr as osr
from vmc_torch_amd import _lib


def _blocks_problem(rng, Ns):
    """Factors of widths (7, 5, 3, 4) and blocks: plain, outer with bias,
    outer without bias, scaled outer with bias sharing factor 0."""
    k = [7, 5, 3, 4]
    F = [rng.normal(size=(Ns, kk)) + rng.normal(size=kk) for kk in k]
    blocks = []
    off = 0

    def add(a, b, s, bias):
        nonlocal off
        w = off
        off += k[a] * (k[b] if b >= 0 else 1)
        bo = -1
        if bias:
            bo = off
            off += k[a]
        blocks.append((a, b, s, w, bo))
    add(0, -1, 1.0, False)
    add(2, 1, 1.0, True)
    add(3, 2, 1.0, False)
    add(0, 3, 0.3, True)
    return F, blocks, off


def _dense(F, blocks, Np):
    Ns = F[0].shape[0]
    O = np.zeros((Ns, Np))
    for a, b, s, w, bo in blocks:
        if b < 0:
            O[:, w:w + F[a].shape[1]] = s * F[a]
            continue
        ka, kb = F[a].shape[1], F[b].shape[1]
        O[:, w:w + ka * kb] = s * (F[a][:, :, None] * F[b][:, None, :]).reshape(Ns, -1)
        if bo >= 0:
            O[:, bo:bo + ka] = s * F[a]
    return O


def _factored_minsr(F, blocks, Np, E, shift, bias_term=True, centre=True):
    """The device algorithm restated: K = sum s^2 G_a o (G_b + 1), T = P K P /
    Ns + shift, dp = O^T P alpha / sqrt(Ns) block by block."""
    Ns = F[0].shape[0]
    G = [f @ f.T for f in F]
    K = np.zeros((Ns, Ns))
    for a, b, s, w, bo in blocks:
        if b < 0:
            K += s * s * G[a]
        else:
            K += s * s * G[a] * (G[b] + (1.0 if (bo >= 0 and bias_term) else 0.0))
    if centre:
        P = np.eye(Ns) - np.ones((Ns, Ns)) / Ns
        K = P @ K @ P
    T = K / Ns + shift * np.eye(Ns)
    eps = (E - E.mean()) / math.sqrt(Ns)
    alpha = np.linalg.solve(T, eps)
    at = (alpha - alpha.mean()) / math.sqrt(Ns)
    dp = np.zeros(Np)
    for a, b, s, w, bo in blocks:
        if b < 0:
            dp[w:w + F[a].shape[1]] = s * F[a].T @ at
            continue
        dp[w:w + F[a].shape[1] * F[b].shape[1]] = (s * F[a].T @ (at[:, None] * F[b])).ravel()
        if bo >= 0:
            dp[bo:bo + F[a].shape[1]] = s * F[a].T @ at
    return dp


@pytest.mark.parametrize("Ns", [1, 9, 40])
def test_factored_algebra_matches_dense_oracle(Ns):
    rng = np.random.default_rng(Ns)
    F, blocks, Np = _blocks_problem(rng, Ns)
    E = rng.normal(size=Ns) + 2.0
    want = osr.minsr(_dense(F, blocks, Np), E, E.mean(), 1e-3)
    got = _factored_minsr(F, blocks, Np, E, 1e-3)
    assert np.linalg.norm(got - want) <= 1e-9 * max(np.linalg.norm(want), 1e-300) + 1e-14


def test_planted_mistakes_are_caught():
    rng = np.random.default_rng(7)
    F, blocks, Np = _blocks_problem(rng, 40)
    E = rng.normal(size=40) + 2.0
    want = osr.minsr(_dense(F, blocks, Np), E, E.mean(), 1e-3)
    for kw in ({"bias_term": False}, {"centre": False}):
        bad = _factored_minsr(F, blocks, Np, E, 1e-3, **kw)
        assert np.linalg.norm(bad - want) > 1e-3 * np.linalg.norm(want), kw


def _c_blocks(blocks):
    arr = (_lib.SrBlockStruct * len(blocks))()
    for i, (a, b, s, w, bo) in enumerate(blocks):
        arr[i].a_factor, arr[i].b_factor, arr[i].scale = a, b, s
        arr[i].w_offset, arr[i].bias_offset = w, bo
    return arr


@pytest.mark.parametrize("blocks,Np,msg", [
    ([(0, -1, 1.0, 0, -1), (1, 0, 1.0, 8, -1)], 24, "tile"),      # gap at [4, 8)
    ([(0, -1, 1.0, 0, -1), (1, 0, 1.0, 2, -1)], 18, "tile"),      # overlap
    ([(0, -1, 1.0, 0, -1), (1, 0, 1.0, 4, -1)], 30, "tile"),      # total != Np
    ([(0, -1, 1.0, 0, 3)], 4, "plain block"),
    ([(2, -1, 1.0, 0, -1)], 4, "a_factor"),
])
def test_block_validation_without_device(blocks, Np, msg):
    lib = _lib.lib()
    k = (ctypes.c_int32 * 2)(4, 4)
    p = ctypes.c_void_p(16)   # never dereferenced: the checks fail first
    ptrs = (ctypes.c_void_p * 2)(16, 16)
    ld = (ctypes.c_int64 * 2)(4, 4)
    rc = lib.vmcx_sr_minsr_factored(8, 2, ptrs, ld, k, len(blocks), _c_blocks(blocks), Np, p,
                                    float("nan"), 1e-3, p, 1 << 30, p, None, None, None, None)
    assert rc == 1
    assert msg in lib.vmcx_last_error().decode()


def test_workspace_bytes():
    lib = _lib.lib()
    k = (ctypes.c_int32 * 2)(8192, 256)
    blocks = _c_blocks([(0, -1, 1.0, 0, -1), (0, 1, 0.1, 8192, -1)])
    need = lib.vmcx_sr_factored_workspace_bytes(4096, 2, k, 2, blocks)
    nn = 4096 * 4096 * 8
    # two factor Grams, partial Grams of the split reduction, T and its copy
    assert need >= 4 * nn
    assert need < 12 * nn
    assert lib.vmcx_sr_factored_workspace_bytes(0, 2, k, 2, blocks) == -1


def test_factor_entry_points_reject_bad_arguments():
    """The factor outputs of the NNBF and fPEPS log-derivatives refuse a
    missing model before any device call."""
    lib = _lib.lib()
    p = ctypes.c_void_p(16)
    rc = lib.vmcx_nnbf_logderiv_factors(None, 4, p, p, 1 << 20, 4, p, 100, p, p, None)
    assert rc == 1 and lib.vmcx_last_error().decode()
    rc = lib.vmcx_peps_logderiv_factors(None, 4, p, p, 1 << 20, 4, p, 100, p, p, None)
    assert rc == 1 and lib.vmcx_last_error().decode()
