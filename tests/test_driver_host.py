"""Host-side logic of the drop-in driver, on the CPU: the outlier mask
(GPU/VMC.py:476-498) against the oracle's restatement, the row norms of a
factored O against the materialised O, VMCLoopConfig.from_vmc_config
(GPU/VMC.py:120-148) and the SGD length check (GPU/optimizer.py:144-186)."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import sr as osr
from vmc_torch_amd.optimizer import SGDGPU
from vmc_torch_amd.VMC import VMCLoopConfig, factored_row_norms, outlier_mask


def test_torch_median_is_the_lower_middle():
    assert float(torch.tensor([1.0, 2.0, 4.0, 30.0]).median()) == 2.0


@pytest.mark.parametrize("Ns", [15, 16])
def test_outlier_mask_matches_oracle(Ns):
    rng = np.random.default_rng(Ns)
    O = rng.normal(size=(Ns, 7))
    O[[2, 9]] *= 50.0                       # two outlier samples
    E = rng.normal(size=Ns)
    for factor in (3.0, 100.0):
        Eo, Oo, keep = osr.outlier_mask(E, O, factor)
        Eg, Og, nd = outlier_mask(torch.tensor(E), torch.tensor(O), factor)
        assert nd == int((~keep).sum())
        assert np.allclose(Eg.numpy(), Eo, rtol=1e-15, atol=0) and np.array_equal(Og.numpy(), Oo)
    assert nd == 0 and Eg.data_ptr() != 0


def test_outlier_mask_median_convention():
    """Norms 1, 2, 4, 30 with factor 10: the lower middle (2) drops the
    30-norm sample, the midpoint (3) would keep it."""
    O = np.diag([1.0, 2.0, 4.0, 30.0])
    E = np.array([1.0, 2.0, 3.0, 4.0])
    Eo, Oo, keep = osr.outlier_mask(E, O, 10.0)
    assert keep.tolist() == [True, True, True, False] and Eo[3] == 2.0
    Eg, Og, nd = outlier_mask(torch.tensor(E), torch.tensor(O), 10.0)
    assert nd == 1 and np.array_equal(Eg.numpy(), Eo) and np.array_equal(Og.numpy(), Oo)


def _nnbf_like_factors(rng, Ns, n_orb, nf, H, s):
    """A random F in the NNBF factor layout [gM | n | dz1 | h1 | dz2 | h2]
    and its blocks (models.NNBF_GPU.log_derivative_factored)."""
    nD = n_orb * nf
    widths = [nD, n_orb, H, H, H, H]
    offs = np.concatenate([[0], np.cumsum(widths)[:-1]])
    factors = list(zip(offs.tolist(), widths))
    o_w1 = nD
    o_b1 = o_w1 + H * n_orb
    o_w2 = o_b1 + H
    o_b2 = o_w2 + H * H
    o_wo = o_b2 + H
    o_bo = o_wo + nD * H
    blocks = [(0, -1, 1.0, 0, -1), (2, 1, 1.0, o_w1, o_b1), (4, 3, 1.0, o_w2, o_b2),
              (0, 5, s, o_wo, o_bo)]
    F = rng.normal(size=(Ns, sum(widths)))
    return F, factors, blocks, o_bo + nD


def _materialize(F, factors, blocks, Np):
    Ns = F.shape[0]
    O = np.zeros((Ns, Np))
    for a, b, sc, wo, bo in blocks:
        ca, ka = factors[a]
        Fa = F[:, ca:ca + ka]
        if b < 0:
            O[:, wo:wo + ka] = sc * Fa
            continue
        cb, kb = factors[b]
        Fb = F[:, cb:cb + kb]
        O[:, wo:wo + ka * kb] = sc * (Fa[:, :, None] * Fb[:, None, :]).reshape(Ns, -1)
        if bo >= 0:
            O[:, bo:bo + ka] = sc * Fa
    return O


def test_factored_row_norms_match_dense():
    rng = np.random.default_rng(3)
    F, factors, blocks, Np = _nnbf_like_factors(rng, 9, 8, 4, 5, 0.37)
    O = _materialize(F, factors, blocks, Np)
    got = factored_row_norms(torch.tensor(F), factors, blocks).numpy()
    assert np.allclose(got, np.linalg.norm(O, axis=1), rtol=1e-13)
    # dropping the bias term (a plausible mistake) is caught
    nob = [(a, b, sc, wo, -1) for a, b, sc, wo, _ in blocks]
    bad = factored_row_norms(torch.tensor(F), factors, nob).numpy()
    assert not np.allclose(bad, np.linalg.norm(O, axis=1), rtol=1e-6)


@dataclasses.dataclass
class _RunConfig:
    vmc_steps: int = 5
    learning_rate: float = 0.02
    sr_diag_shift: float = 5e-4
    resume_step: int = 40
    outlier_clip_factor: float = 7.0


def test_from_vmc_config_maps_reference_names():
    cfg = _RunConfig()
    cfg.extra_flag = "kept"
    c = VMCLoopConfig.from_vmc_config(cfg, n_params=123, nsites=16)
    assert c.diag_shift == 5e-4 and c.step_offset == 40
    assert c.vmc_steps == 5 and c.learning_rate == 0.02 and c.outlier_clip_factor == 7.0
    assert c.n_params == 123 and c.nsites == 16 and c.extra_flag == "kept"
    assert not hasattr(c, "sr_diag_shift")


def test_sgd_rejects_short_direction_before_any_update():
    m = torch.nn.Module()
    m.params = torch.nn.ParameterList([torch.nn.Parameter(torch.ones(3, dtype=torch.float64)),
                                       torch.nn.Parameter(torch.ones(2, dtype=torch.float64))])
    before = [p.detach().clone() for p in m.parameters()]
    with pytest.raises(ValueError, match="direction has 4 entries, parameters 5"):
        SGDGPU(0.1).step(m, torch.zeros(4, dtype=torch.float64))
    assert all(torch.equal(a, b) for a, b in zip(before, m.parameters()))
