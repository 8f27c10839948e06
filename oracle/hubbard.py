"""Oracle: square lattice, spinful Fermi-Hubbard Hamiltonian, encodings, get_conn.

Test infrastructure only (see oracle/__init__.py).

Encodings (reference vmc_torch/fermion_utils.py:1448-1498):
  quimb code per site  {0: empty, 1: down, 2: up, 3: up+down}
  netket binary        [u_0 .. u_{N-1} | d_0 .. d_{N-1}]
  symmray interleaved  [d_0, u_0, d_1, u_1, ...]
"""
import itertools
from math import comb

import numpy as np


# ---------------------------------------------------------------- lattice
def edges_2d_square(Lx, Ly):
    """Open-boundary square-lattice edges in quimb's sorted order.

    Follows hamiltonian_torch.py:351 (qtn.edges_2d_square).  quimb is absent,
    so the order is the documented assumption: for each cell (x, y) in
    row-major order, ((x, y), (x, y+1)) then ((x, y), (x+1, y)), which is the
    lexicographically sorted edge list.
    """
    edges = []
    for x in range(Lx):
        for y in range(Ly):
            if y + 1 < Ly:
                edges.append(((x, y), (x, y + 1)))
            if x + 1 < Lx:
                edges.append(((x, y), (x + 1, y)))
    return sorted(edges)


class SquareLattice:
    """hamiltonian_torch.py:344-372: site index x*Ly + y, row/col edge dicts."""

    def __init__(self, Lx, Ly):
        self.Lx, self.Ly = Lx, Ly
        self.n_nodes = Lx * Ly
        self._edges = [(a[0] * Ly + a[1], b[0] * Ly + b[1])
                       for a, b in edges_2d_square(Lx, Ly)]
        self.row_edges, self.col_edges = {}, {}
        for i, j in self._edges:
            ci, cj = divmod(i, Ly), divmod(j, Ly)
            if ci[0] == cj[0]:
                self.row_edges.setdefault(ci[0], []).append((i, j))
            elif ci[1] == cj[1]:
                self.col_edges.setdefault(ci[1], []).append((i, j))

    def edges(self):
        return list(self._edges)

    def sweep_edges(self):
        """Edge order of one sampler sweep: row_edges then col_edges
        (GPU/sampler.py:263-269)."""
        out = []
        for v in self.row_edges.values():
            out.extend(v)
        for v in self.col_edges.values():
            out.extend(v)
        return out


# -------------------------------------------------------------- encodings
def quimb_to_netket(x):
    """fermion_utils.py:1472-1498 (batched over leading dims)."""
    x = np.asarray(x)
    up = ((x == 2) | (x == 3)).astype(np.int64)
    dn = ((x == 1) | (x == 3)).astype(np.int64)
    return np.concatenate([up, dn], axis=-1)


def netket_to_quimb(n):
    """fermion_utils.py:1448-1470: up*2 + down."""
    n = np.asarray(n)
    N = n.shape[-1] // 2
    return (n[..., :N] * 2 + n[..., N:]).astype(np.int64)


def quimb_to_symmray(x):
    """Interleaved [d0, u0, d1, u1, ...] (fermion_utils.py:1575-1579)."""
    x = np.asarray(x)
    up = ((x == 2) | (x == 3)).astype(np.int64)
    dn = ((x == 1) | (x == 3)).astype(np.int64)
    return np.stack([dn, up], axis=-1).reshape(x.shape[:-1] + (2 * x.shape[-1],))


def calc_phase_symmray(xi, xj):
    """fermion_utils.py:1564-1585: (-1)^(# occupied symmray modes of xi
    strictly between the two modes that differ)."""
    xi, xj = np.asarray(xi), np.asarray(xj)
    if len(np.nonzero(xi - xj)[0]) != 2:
        return 1
    si, sj = quimb_to_symmray(xi), quimb_to_symmray(xj)
    p, q = np.nonzero(si - sj)[0]
    return (-1) ** int(np.sum(si[p + 1:q]) % 2)


def calc_phase_netket(xi, xj):
    """fermion_utils.py:1550-1562 (netket-ordered JW phase)."""
    xi, xj = np.asarray(xi), np.asarray(xj)
    if len(np.nonzero(xi - xj)[0]) != 2:
        return 1
    ni, nj = quimb_to_netket(xi), quimb_to_netket(xj)
    p, q = np.nonzero(ni - nj)[0]
    return (-1) ** int(np.sum(ni[p + 1:q]) % 2)


# ------------------------------------------------------------ Hamiltonian
class Hubbard:
    """Spinful Fermi-Hubbard on an OBC square lattice.

    Term dictionary as hamiltonian_torch.py:995-1015: for each edge in
    graph order, (i, j, +1) = -t then (i, j, -1) = -t; then (i,) = U.
    """

    def __init__(self, Lx, Ly, t, U, n_up, n_dn):
        self.Lx, self.Ly, self.t, self.U = Lx, Ly, float(t), float(U)
        self.N = Lx * Ly
        self.n_up, self.n_dn = n_up, n_dn
        self.n_f = n_up + n_dn
        self.graph = SquareLattice(Lx, Ly)
        self.H = {}
        for i, j in self.graph.edges():
            for spin in (1, -1):
                self.H[(i, j, spin)] = -self.t
        for i in range(self.N):
            self.H[(i,)] = self.U
        self._precompute_hops()

    def _precompute_hops(self):
        """hamiltonian_torch.py:562-584 (precompute_hops_gpu)."""
        N = self.N
        self.hop_list, self.diag_list = [], []
        for key, value in self.H.items():
            if len(key) == 3:
                i0, j0, spin = key
                i_net = i0 if spin == 1 else i0 + N
                j_net = j0 if spin == 1 else j0 + N
                p = 2 * i0 + (1 if spin == 1 else 0)
                q = 2 * j0 + (1 if spin == 1 else 0)
                if p > q:
                    p, q = q, p
                self.hop_list.append((i_net, j_net, value, p, q))
            else:
                self.diag_list.append((key[0], value))

    # -- CPU semantics: one config, duplicates merged in a dict
    def get_conn(self, x):
        """hamiltonian_torch.py:1037-1076."""
        x = np.asarray(x)
        sigma = quimb_to_netket(x)
        out = {}
        N = self.N
        for key, value in self.H.items():
            if len(key) == 3:
                i0, j0, spin = key
                i = i0 if spin == 1 else i0 + N
                j = j0 if spin == 1 else j0 + N
                if sigma[i] != sigma[j]:
                    eta = sigma.copy()
                    eta[i], eta[j] = sigma[j], sigma[i]
                    eq = netket_to_quimb(eta)
                    ph = calc_phase_symmray(x, eq)
                    k = tuple(int(v) for v in eq)
                    out[k] = out.get(k, 0.0) + value * ph
            else:
                if x[key[0]] == 3:
                    k = tuple(int(v) for v in x)
                    out[k] = out.get(k, 0.0) + value
        etas = np.array(list(out.keys()), dtype=np.int64).reshape(-1, N)
        return etas, np.array(list(out.values()), dtype=np.float64)

    # -- GPU semantics: batched, term-major, U terms unmerged
    def get_conn_batch(self, fxs):
        """hamiltonian_torch.py:586-655 (get_conn_batch_gpu)."""
        fxs = np.asarray(fxs, dtype=np.int64)
        B, N = fxs.shape
        net = quimb_to_netket(fxs)
        sym = quimb_to_symmray(fxs)
        etas, coeffs, bids = [], [], []
        for i_net, j_net, c, p, q in self.hop_list:
            idx = np.nonzero(net[:, i_net] != net[:, j_net])[0]
            if idx.size == 0:
                continue
            nn = net[idx].copy()
            nn[:, [i_net, j_net]] = nn[:, [j_net, i_net]]
            between = sym[idx, p + 1:q].sum(axis=-1) % 2
            etas.append(netket_to_quimb(nn))
            coeffs.append((1.0 - 2.0 * between) * c)
            bids.append(idx)
        for site, c in self.diag_list:
            idx = np.nonzero(fxs[:, site] == 3)[0]
            if idx.size == 0:
                continue
            etas.append(fxs[idx])
            coeffs.append(np.full(idx.size, c))
            bids.append(idx)
        if not etas:
            return (np.zeros((0, N), np.int64), np.zeros(0), np.zeros(0, np.int64))
        return (np.concatenate(etas), np.concatenate(coeffs),
                np.concatenate(bids).astype(np.int64))

    # -- exact enumeration
    def all_states(self):
        """hamiltonian_torch.py:151-173: product(up combos, down combos) in
        itertools.combinations order, netket -> quimb."""
        N = self.N

        def vecs(m):
            out = []
            for pos in itertools.combinations(range(N), m):
                v = np.zeros(N, np.int64)
                v[list(pos)] = 1
                out.append(v)
            return out

        ups, dns = vecs(self.n_up), vecs(self.n_dn)
        net = np.array([np.concatenate([u, d]) for u, d in itertools.product(ups, dns)])
        return netket_to_quimb(net)

    @property
    def hilbert_size(self):
        return comb(self.N, self.n_up) * comb(self.N, self.n_dn)

    def to_dense(self, states=None, conn=None):
        """hamiltonian_torch.py:424-442: H[eta, sigma] += <eta|H|sigma>."""
        states = self.all_states() if states is None else states
        index = {tuple(int(v) for v in s): k for k, s in enumerate(states)}
        Hm = np.zeros((len(states), len(states)))
        conn = conn or self.get_conn
        for k, s in enumerate(states):
            etas, cs = conn(s)
            for e, c in zip(etas, cs):
                Hm[index[tuple(int(v) for v in e)], k] += c
        return Hm

    def get_conn_netket(self, x):
        """The netket-phase variant (hamiltonian_torch.py:1078-1119)."""
        x = np.asarray(x)
        sigma = quimb_to_netket(x)
        out = {}
        N = self.N
        for key, value in self.H.items():
            if len(key) == 3:
                i0, j0, spin = key
                i = i0 if spin == 1 else i0 + N
                j = j0 if spin == 1 else j0 + N
                if sigma[i] != sigma[j]:
                    eta = sigma.copy()
                    eta[i], eta[j] = sigma[j], sigma[i]
                    eq = netket_to_quimb(eta)
                    k = tuple(int(v) for v in eq)
                    out[k] = out.get(k, 0.0) + value * calc_phase_netket(x, eq)
            elif x[key[0]] == 3:
                k = tuple(int(v) for v in x)
                out[k] = out.get(k, 0.0) + value
        etas = np.array(list(out.keys()), dtype=np.int64).reshape(-1, N)
        return etas, np.array(list(out.values()), dtype=np.float64)


def random_initial_config(n_f, n_sites, seed):
    """GPU/vmc_utils.py:22-39: pattern [1,2]*(N/2), first N - N_f sites
    emptied, then torch.randperm under torch.manual_seed(seed).

    Uses the container's torch CPU generator, which is what the reference
    calls; returns a NumPy int64 array.
    """
    import torch
    g = torch.Generator().manual_seed(int(seed))
    cfg = torch.tensor([1, 2] * (n_sites // 2), dtype=torch.int64)
    cfg[: n_sites - n_f] = 0
    perm = torch.randperm(n_sites, generator=g)
    return cfg[perm].numpy().copy()
