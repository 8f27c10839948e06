"""Symmetry-projected fermionic wavefunction (GPU/models/symmetry.py).

    psi^r(n) = (1/|G|) sum_g chi_r(g) sign(g, n) psi(n[pi_g])

over the point group of the lattice: C4v (8 elements) when Lx == Ly, C2v
(4) otherwise, one-dimensional irreps A1, A2, B1, B2 (symmetry.py:28-52).
sign(g, n) is the fermionic reordering sign in the interleaved mode order
[d0, u0, d1, u1, ...] of the inverse permutation (symmetry.py:789-806).
The images n[pi_g] and the signs come from one HIP kernel
(`vmcx_sym_images`); the base model evaluates all |G| images of every
walker in one batched call.  The wrapper has the model interface the
drivers use: forward / forward_log, params, local_energy (connected
configurations of the Hamiltonian, amplitude ratios of the projected
state), log_derivative (sum over the images of the base model's O weighted
by each image's share of the projected amplitude).  It has no walker-state
sweep; sample it with a sampler that calls forward_log
(`sampler.LearnedProposalSamplerGPU`).
"""
import math

import torch

from . import kernels

# order [e, C4, C4^2, C4^3, sigma_v, sigma_v', sigma_d, sigma_d'] (symmetry.py:28-38)
C4V_CHARACTERS = {"A1": [1, 1, 1, 1, 1, 1, 1, 1], "A2": [1, 1, 1, 1, -1, -1, -1, -1],
                  "B1": [1, -1, 1, -1, 1, 1, -1, -1], "B2": [1, -1, 1, -1, -1, -1, 1, 1]}
C4V_ELEMENT_NAMES = ["e", "C4", "C4^2", "C4^3", "sigma_v", "sigma_v'", "sigma_d", "sigma_d'"]
# order [e, C2, sigma_v, sigma_v'] (symmetry.py:44-52)
C2V_CHARACTERS = {"A1": [1, 1, 1, 1], "A2": [1, 1, -1, -1], "B1": [1, -1, 1, -1],
                  "B2": [1, -1, -1, 1]}
C2V_ELEMENT_NAMES = ["e", "C2", "sigma_v", "sigma_v'"]


def build_c4v_permutations(L):
    """symmetry.py:118-160: perms[g][s] = flat index (row-major x L + y) of
    the image of site s; x[perms[g]] is the transformed configuration."""
    s = torch.arange(L * L, dtype=torch.long)
    x, y = s // L, s % L
    imgs = [(x, y), (L - 1 - y, x), (L - 1 - x, L - 1 - y), (y, L - 1 - x),
            (x, L - 1 - y), (L - 1 - x, y), (y, x), (L - 1 - y, L - 1 - x)]
    return torch.stack([a * L + b for a, b in imgs])


def build_c2v_permutations(Lx, Ly):
    """symmetry.py:56-82."""
    s = torch.arange(Lx * Ly, dtype=torch.long)
    x, y = s // Ly, s % Ly
    imgs = [(x, y), (Lx - 1 - x, Ly - 1 - y), (x, Ly - 1 - y), (Lx - 1 - x, y)]
    return torch.stack([a * Ly + b for a, b in imgs])


def build_symmetry_group(Lx, Ly):
    """symmetry.py:85-100: (perms, character table, element names)."""
    if Lx == Ly:
        return build_c4v_permutations(Lx), C4V_CHARACTERS, C4V_ELEMENT_NAMES
    return build_c2v_permutations(Lx, Ly), C2V_CHARACTERS, C2V_ELEMENT_NAMES


class FermionSymmetryProjectedModel(torch.nn.Module):
    """symmetry.py:770-951: projection of `base_model` onto the irrep
    `irrep` of the lattice's point group, with the fermionic sign."""

    def __init__(self, base_model, Lx, Ly, irrep="A1"):
        super().__init__()
        perms, table, names = build_symmetry_group(Lx, Ly)
        if irrep not in table:
            raise ValueError(f"unknown irrep {irrep!r}; choose from {sorted(table)}")
        self.base_model = base_model
        self.Lx, self.Ly, self.irrep = int(Lx), int(Ly), irrep
        self.n_group = perms.shape[0]
        self.element_names = names
        inv = torch.empty_like(perms)
        for g in range(self.n_group):
            inv[g][perms[g]] = torch.arange(perms.shape[1])
        self.register_buffer("perms", perms)
        self.register_buffer("characters", torch.tensor(table[irrep], dtype=torch.float64))
        # the reference's mode buffers, kept so that its state dicts load
        # (symmetry.py:680-700, 803-806): mode permutations of the inverse
        # site permutations and their inversion masks
        mode = torch.empty(self.n_group, 2 * perms.shape[1], dtype=torch.long)
        mode[:, 0::2] = 2 * inv
        mode[:, 1::2] = 2 * inv + 1
        upper = torch.triu(torch.ones(mode.shape[1], mode.shape[1], dtype=torch.bool), 1)
        self.register_buffer("mode_perms", mode)
        self.register_buffer("mode_inv_masks", (mode.unsqueeze(2) > mode.unsqueeze(1)) & upper)
        self.register_buffer("_perm32", perms.to(torch.int32), persistent=False)
        self.register_buffer("_inv32", inv.to(torch.int32), persistent=False)

    # parameters live in the base model (symmetry.py:280-288)
    @property
    def params(self):
        return self.base_model.params

    @property
    def n_params(self):
        return self.base_model.n_params

    def invalidate(self):
        if hasattr(self.base_model, "invalidate"):
            self.base_model.invalidate()

    def images(self, x):
        """(images [B, G, N], fermionic signs [B, G])."""
        return kernels.sym_images(x.contiguous(), self._perm32, self._inv32)

    def _combine(self, x):
        B, N = x.shape
        G = self.n_group
        img, fs = self.images(x)
        s, l = self.base_model.forward_log(img.reshape(B * G, N))
        eff = s.reshape(B, G) * self.characters * fs
        lg = l.reshape(B, G)
        m = lg.max(dim=1, keepdim=True).values
        w = (eff * torch.exp(lg - m)).sum(dim=1)
        sign = torch.sign(w)
        logabs = m.squeeze(1) + torch.log(w.abs().clamp(min=1e-45)) - math.log(G)
        return sign, logabs, img, eff, lg

    @torch.inference_mode()
    def forward_log(self, x):
        """symmetry.py:924-951: (sign, ln|psi^r|)."""
        sign, logabs, _, _, _ = self._combine(x)
        return sign, logabs

    @torch.inference_mode()
    def forward(self, x):
        sign, logabs = self.forward_log(x)
        return sign * torch.exp(logabs)

    @torch.inference_mode()
    def local_energy(self, fxs, H, chunk=65536):
        """E_loc(x) = sum_eta H[eta, x] psi^r(eta) / psi^r(x) over the
        connected configurations of H (GPU/vmc_utils.py:322-476 in log form);
        the diagonal entries contribute their coefficient."""
        s0, l0 = self.forward_log(fxs)
        etas, coeffs, bids = H.get_conn_batch_gpu(fxs)
        ratio = torch.ones_like(coeffs)
        hop = (etas != fxs[bids]).any(dim=1).nonzero().squeeze(1)
        for a in range(0, hop.numel(), chunk):
            idx = hop[a:a + chunk]
            s1, l1 = self.forward_log(etas[idx])
            ratio[idx] = s1 * s0[bids[idx]] * torch.exp(l1 - l0[bids[idx]])
        eloc = torch.zeros(fxs.shape[0], dtype=torch.float64, device=fxs.device)
        eloc.index_add_(0, bids, coeffs * ratio)
        return eloc

    @torch.inference_mode()
    def log_derivative(self, fxs, dtype=torch.float64, max_bytes=2 << 30):
        """d ln|psi^r| / d theta = sum_g w_g O(x_g), w_g = chi_g sign_g
        s_g s^r exp(l_g - l^r) / |G|, with O the base model's log-derivative
        of each image; returns (O^r [B, Np], (sign, ln|psi^r|))."""
        B, N = fxs.shape
        G = self.n_group
        Np = self.n_params
        sign, logabs, img, eff, lg = self._combine(fxs)
        wts = eff * sign.unsqueeze(1) * torch.exp(lg - logabs.unsqueeze(1)) / G
        out = torch.empty(B, Np, dtype=dtype, device=fxs.device)
        esz = torch.empty((), dtype=dtype).element_size()
        per = max(1, int(max_bytes // (G * Np * esz)))
        for b0 in range(0, B, per):
            b1 = min(B, b0 + per)
            O, _ = self.base_model.log_derivative(img[b0:b1].reshape(-1, N), dtype=dtype)
            out[b0:b1] = torch.bmm(wts[b0:b1].to(dtype).unsqueeze(1),
                                   O.reshape(b1 - b0, G, Np)).squeeze(1)
        self.invalidate()
        return out, (sign, logabs)

    def mcmc_sweep(self, *args, **kwargs):
        raise NotImplementedError(
            "the projected model has no walker-state sweep; sample it with "
            "sampler.LearnedProposalSamplerGPU (which calls forward_log)")
