"""Write tests/golden/walker_init.json: initial walker configurations by the
reference's recipe (GPU/vmc_utils.py:22-39 random_initial_config,
GPU/vmc_setup.py:139-159 initialize_walkers), computed here with the
literal torch calls of that recipe (torch.manual_seed(seed), the pattern
[1, 2] * (N/2) with its first N - N_f entries emptied, then indexing by
torch.randperm(N)) so that the fixture does not depend on the code it pins.

    python tests/golden/make_walker_init.py
"""
import json
import os

import torch

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "walker_init.json")

# (N_f, N_sites): the 4x2, 4x4 (7 up + 7 down), 6x6 (n = 0.89) and 8x8
# lattices of the benchmark configurations
CASES = [(8, 8), (14, 16), (32, 36), (64, 64)]
SEEDS = list(range(42, 42 + 12)) + [0, 7, 12345]


def reference_recipe(n_f, n_sites, seed):
    torch.manual_seed(seed)
    half = torch.tensor([1, 2] * (n_sites // 2))
    doped = half.clone()
    doped[list(range(n_sites - n_f))] = 0
    perm = torch.randperm(n_sites)
    return doped[perm].tolist()


def main():
    out = {"recipe": "GPU/vmc_utils.py:22-39; seeds of initialize_walkers: "
                     "seed + rank * batch_size + i (GPU/vmc_setup.py:139-159)",
           "torch": torch.__version__, "cases": []}
    for n_f, n in CASES:
        out["cases"].append({"n_f": n_f, "n_sites": n,
                             "configs": {str(s): reference_recipe(n_f, n, s) for s in SEEDS}})
    with open(OUT, "w") as f:
        json.dump(out, f)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
