"""Walker initialisation (GPU/vmc_utils.py:22-39, GPU/vmc_setup.py:139-159)
pinned against tests/golden/walker_init.json, written by the reference's
torch recipe (tests/golden/make_walker_init.py): the product's
random_initial_config / initialize_walkers and the oracle's restatement give
the fixture's configurations for every case and seed."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import hubbard as ohub
from vmc_torch_amd.vmc_utils import initialize_walkers, random_initial_config

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walker_init.json")
FIX = json.load(open(GOLD))


@pytest.mark.parametrize("case", FIX["cases"], ids=lambda c: f"nf{c['n_f']}_N{c['n_sites']}")
def test_random_initial_config_matches_fixture(case):
    n_f, n = case["n_f"], case["n_sites"]
    for seed, want in case["configs"].items():
        got = random_initial_config(n_f, n, seed=int(seed))
        assert got.dtype == torch.int64 and got.tolist() == want, seed
        assert np.array_equal(ohub.random_initial_config(n_f, n, int(seed)), want), seed
        w = np.asarray(want)
        # invariants of the recipe: N_f / 2 electrons of each spin, no double
        # occupation, N - N_f empty sites
        assert (w == 1).sum() == n_f // 2 and (w == 2).sum() == n_f // 2
        assert (w == 0).sum() == n - n_f and (w == 3).sum() == 0


@pytest.mark.parametrize("rank", [0, 1])
def test_initialize_walkers_seed_layout(rank):
    """Walker i of rank r uses seed + r * B + i."""
    case = next(c for c in FIX["cases"] if c["n_sites"] == 16)
    B, seed = 6, 42
    W = initialize_walkers(lambda seed: random_initial_config(case["n_f"], 16, seed=seed), B,
                           seed=seed, rank=rank)
    assert W.shape == (B, 16) and W.dtype == torch.int64
    for i in range(B):
        assert W[i].tolist() == case["configs"][str(seed + rank * B + i)]


def test_fixture_catches_a_wrong_recipe():
    """Plausible mistakes in a restatement fail the fixture: emptying the
    last N - N_f sites instead of the first, or drawing the permutation
    from a generator that was not seeded by the walker's seed."""
    case = next(c for c in FIX["cases"] if c["n_sites"] == 36)
    n_f, n = case["n_f"], case["n_sites"]
    wrong_end, wrong_seed = 0, 0
    for seed, want in case["configs"].items():
        g = torch.Generator().manual_seed(int(seed))
        cfg = torch.tensor([1, 2] * (n // 2), dtype=torch.int64)
        cfg[n_f:] = 0
        wrong_end += cfg[torch.randperm(n, generator=g)].tolist() != want
        g2 = torch.Generator().manual_seed(int(seed) + 1)
        cfg = torch.tensor([1, 2] * (n // 2), dtype=torch.int64)
        cfg[: n - n_f] = 0
        wrong_seed += cfg[torch.randperm(n, generator=g2)].tolist() != want
    assert wrong_end == len(case["configs"]) and wrong_seed == len(case["configs"])
