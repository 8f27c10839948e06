"""Shared helpers for the parity tests (seeded inputs, oracle wrappers)."""
import numpy as np
import torch

from oracle import hubbard as ohub
from oracle import sampler as osamp


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


def tie_free_draws(logamp, x, edges, n_sw, rng, margin=1e-3):
    """Oracle trajectory and injected draws with every accept uniform at
    least `margin` (relative) away from its acceptance ratio."""
    B, E = x.shape[0], len(edges)
    uc = rng.random((n_sw, E, B))
    bits = rng.integers(0, 4, (n_sw, E, B)).astype(np.uint8)
    ua = rng.random((n_sw, E, B))
    x = x.copy()
    s, l = logamp(x)
    for sw in range(n_sw):
        for e, (i, j) in enumerate(edges):
            ni, nj, flag = osamp.propose(x[:, i], x[:, j], uc[sw, e], bits[sw, e] & 1,
                                         (bits[sw, e] >> 1) & 1)
            if not flag.any():
                continue
            sel = np.nonzero(flag)[0]
            prop = x[sel].copy()
            prop[:, i], prop[:, j] = ni[sel], nj[sel]
            ps, pl = logamp(prop)
            r = np.exp(2.0 * (pl - l[sel]))
            u = ua[sw, e, sel]
            close = np.abs(u - r) < margin * np.maximum(r, 1e-300)
            u[close] = 0.5 * np.minimum(r[close], 1.0)
            ua[sw, e, sel] = u
            acc = u < r
            a = sel[acc]
            x[a, i], x[a, j] = ni[a], nj[a]
            s[a], l[a] = ps[acc], pl[acc]
    return (uc, bits, ua), x, s, l
