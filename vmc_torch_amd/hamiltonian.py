"""Lattice, Hilbert space and spinful Fermi-Hubbard Hamiltonian with the
reference's interface (vmc_torch/hamiltonian_torch.py).

Only the Hubbard / square-lattice parts on the VMC hot path are provided.
The batched connected-configuration enumeration runs in HIP
(vmcx_hubbard_conn_count / _fill)."""
import itertools
from math import comb

import torch

from . import kernels


def edges_2d_square(Lx, Ly, pbc=False):
    """OBC square-lattice edges in sorted order: for each cell (x, y) in
    row-major order, ((x, y), (x, y+1)) before ((x, y), (x+1, y)) — the
    order of quimb's edges_2d_square used at hamiltonian_torch.py:351."""
    if pbc:
        raise NotImplementedError("PBC not implemented (as in the reference)")
    out = []
    for x in range(Lx):
        for y in range(Ly):
            if y + 1 < Ly:
                out.append(((x, y), (x, y + 1)))
            if x + 1 < Lx:
                out.append(((x, y), (x + 1, y)))
    return sorted(out)


class SquareLattice:
    """hamiltonian_torch.py:344-372 (site index x*Ly + y, row_edges /
    col_edges in first-seen order)."""

    def __init__(self, Lx, Ly, pbc=False):
        self.Lx, self.Ly, self.pbc = Lx, Ly, pbc
        self.n_nodes = Lx * Ly
        self._edges = [(a[0] * Ly + a[1], b[0] * Ly + b[1])
                       for a, b in edges_2d_square(Lx, Ly, pbc)]
        self.row_edges, self.col_edges = {}, {}
        for i, j in self._edges:
            ci, cj = self.from_index_to_2dcoo(i), self.from_index_to_2dcoo(j)
            if ci[0] == cj[0]:
                self.row_edges.setdefault(ci[0], []).append((i, j))
            elif ci[1] == cj[1]:
                self.col_edges.setdefault(ci[1], []).append((i, j))
        self._edge_tensors = {}

    def edges(self):
        return list(self._edges)

    def from_index_to_2dcoo(self, index):
        return index // self.Ly, index % self.Ly

    def sweep_edges(self):
        """The sampler's sweep order: row_edges then col_edges
        (GPU/sampler.py:263-269)."""
        out = []
        for v in self.row_edges.values():
            out.extend(v)
        for v in self.col_edges.values():
            out.extend(v)
        return out

    def edge_tensor(self, device, order="sweep"):
        """int32 [E, 2] device tensor of the edges, cached per device."""
        key = (str(device), order)
        if key not in self._edge_tensors:
            lst = self.sweep_edges() if order == "sweep" else self.edges()
            self._edge_tensors[key] = torch.tensor(lst, dtype=torch.int32, device=device).reshape(-1, 2)
        return self._edge_tensors[key]


class SpinfulFermion:
    """hamiltonian_torch.py:109-222 (U(1)-conserving sector only)."""

    def __init__(self, n_orbitals, n_fermions=None, n_fermions_per_spin=None):
        self.n_orbitals = 2 * n_orbitals
        self.n_fermions = (n_fermions if n_fermions is not None
                           else n_fermions_per_spin[0] + n_fermions_per_spin[1])
        self.n_fermions_per_spin = n_fermions_per_spin
        self.n_fermions_spin_up = n_fermions_per_spin[0] if n_fermions_per_spin else None
        self.n_fermions_spin_down = n_fermions_per_spin[1] if n_fermions_per_spin else None

    @property
    def size(self):
        half = self.n_orbitals // 2
        if self.n_fermions_per_spin is not None:
            return comb(half, self.n_fermions_spin_up) * comb(half, self.n_fermions_spin_down)
        return comb(self.n_orbitals, self.n_fermions)

    def all_states(self):
        """Full sector in the reference order (itertools.product of the up
        and down combinations), quimb codes, int64 (CPU tensor)."""
        if self.n_fermions_per_spin is None:
            raise NotImplementedError("all_states needs n_fermions_per_spin")
        half = self.n_orbitals // 2
        rows = []
        for up, dn in itertools.product(
                itertools.combinations(range(half), self.n_fermions_spin_up),
                itertools.combinations(range(half), self.n_fermions_spin_down)):
            x = [0] * half
            for i in up:
                x[i] += 2
            for i in dn:
                x[i] += 1
            rows.append(x)
        return torch.tensor(rows, dtype=torch.int64)


class spinful_Fermi_Hubbard_square_lattice_torch:
    """hamiltonian_torch.py:995-1076 + SpinfulFermionGPUMixin (553-655)."""

    def __init__(self, Lx, Ly, t, U, N_f, pbc=False, n_fermions_per_spin=None,
                 no_u1_symmetry=False, gpu=True):
        if no_u1_symmetry:
            raise NotImplementedError("only the U(1) sector is on the hot path")
        self.Lx, self.Ly, self.t, self.U = Lx, Ly, float(t), float(U)
        N = Lx * Ly
        if n_fermions_per_spin is None:
            self._hilbert = SpinfulFermion(N, n_fermions=N_f)
        else:
            self._hilbert = SpinfulFermion(N, n_fermions_per_spin=n_fermions_per_spin)
        self._graph = SquareLattice(Lx, Ly, pbc)
        self._H = {}
        for i, j in self._graph.edges():
            for spin in (1, -1):
                self._H[(i, j, spin)] = -self.t
        for i in range(N):
            self._H[(i,)] = self.U
        self.gpu = gpu

    @property
    def graph(self):
        return self._graph

    @property
    def hilbert(self):
        return self._hilbert

    @property
    def H(self):
        return self._H

    def precompute_hops_gpu(self, device):
        """hamiltonian_torch.py:562-584: keeps the hop list (presence of
        `_hop_list` selects the batched path in evaluate_energy) and uploads
        the edge list once."""
        N = self._hilbert.n_orbitals // 2
        hop_list, diag_list = [], []
        for key, value in self._H.items():
            if len(key) == 3:
                i0, j0, spin = key
                p = 2 * i0 + (1 if spin == 1 else 0)
                q = 2 * j0 + (1 if spin == 1 else 0)
                hop_list.append((i0 if spin == 1 else i0 + N, j0 if spin == 1 else j0 + N,
                                 value, min(p, q), max(p, q)))
            else:
                diag_list.append((key[0], value))
        self._hop_list, self._diag_list = hop_list, diag_list
        self._gpu_device = torch.device(device)
        self._graph.edge_tensor(self._gpu_device, "graph")
        self._graph.edge_tensor(self._gpu_device, "sweep")

    def get_conn_batch_gpu(self, fxs):
        """(conn_etas [T, N] int64, conn_coeffs [T] float64, batch_ids [T]
        int64) in term-major order, computed by the HIP kernels."""
        edges = self._graph.edge_tensor(fxs.device, "graph")
        return kernels.hubbard_get_conn(fxs.contiguous(), edges, self.t, self.U)

    def get_conn(self, sigma_quimb):
        """Single-configuration form (hamiltonian_torch.py:1037-1076): the
        batched kernel on one walker; Hubbard hops never produce duplicate
        connected configurations, so no merging is needed beyond the
        diagonal U terms, which are summed."""
        x = torch.as_tensor(sigma_quimb, dtype=torch.int64)
        dev = getattr(self, "_gpu_device", None) or torch.device("cuda")
        etas, coeffs, _ = self.get_conn_batch_gpu(x.reshape(1, -1).to(dev))
        n_hop = int((etas != x.to(dev)).any(dim=1).sum().item())
        if coeffs.shape[0] > n_hop:
            diag = coeffs[n_hop:].sum().reshape(1)
            etas = torch.cat([etas[:n_hop], etas[n_hop:n_hop + 1]])
            coeffs = torch.cat([coeffs[:n_hop], diag])
        return etas, coeffs
