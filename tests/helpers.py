"""Shared helpers for the parity tests (seeded inputs, oracle wrappers)."""
import numpy as np
import torch

from oracle import hubbard as ohub


def random_configs(rng, B, N, n_up, n_dn):
    x = np.zeros((B, N), np.int64)
    for b in range(B):
        x[b, rng.choice(N, n_up, replace=False)] += 2
        x[b, rng.choice(N, n_dn, replace=False)] += 1
    return x


def setup_case(Lx, Ly, n_up, n_dn, B, seed, std=0.1, device="cuda"):
    """Oracle Hamiltonian, M (numpy + device) and random walkers."""
    rng = np.random.default_rng(seed)
    N = Lx * Ly
    M = rng.normal(0.0, std, (2 * N, n_up + n_dn))
    x = random_configs(rng, B, N, n_up, n_dn)
    H = ohub.Hubbard(Lx, Ly, 1.0, 8.0, n_up, n_dn)
    Md = torch.tensor(M, dtype=torch.float64, device=device)
    xd = torch.tensor(x, dtype=torch.int64, device=device)
    return H, M, x, Md, xd, rng


def edges_tensor(edges, device="cuda"):
    return torch.tensor(edges, dtype=torch.int32, device=device).reshape(-1, 2)
