"""Wavefunction models with the reference's GPU interface
(vmc_torch/GPU/models/_base.py, slater.py).

A model keeps the parameters as an nn.ParameterList (the order defines the
flat O_k / dp layout, _base.py:41-44) and exposes forward / forward_log.
Evaluation runs in libvmcx; the model additionally caches the device-side
walker state (A^-1, slot map, sign, ln|psi|) that the sampler, the local
energy and the log-derivative kernels share.
"""
import torch
import torch.nn as nn

from . import kernels


class WavefunctionModel_GPU(nn.Module):
    """_base.py:21-130 interface: params, forward(x) -> amp[B],
    forward_log(x) -> (sign[B], log|psi|[B])."""

    def __init__(self, params_list):
        super().__init__()
        self.params = nn.ParameterList([
            p if isinstance(p, nn.Parameter) else nn.Parameter(p) for p in params_list])

    def forward_log(self, x):
        raise NotImplementedError

    def forward(self, x):
        s, l = self.forward_log(x)
        return s * torch.exp(l)

    @property
    def n_params(self):
        return sum(p.numel() for p in self.params)


class SlaterDeterminant_GPU(WavefunctionModel_GPU):
    """psi(x) = det(M[occupied, :]), occupied = netket-ordered occupied
    orbitals ascending (GPU/models/slater.py:15-61; CPU model.py:7437-7457).

    M: (n_orbitals = 2*N_sites, n_fermions) float64.  `init_std` scales the
    N(0, 1) draw (the reference uses std 1; the benchmark configurations use
    0.1).
    """

    def __init__(self, n_orbitals, n_fermions, dtype=torch.float64, init_std=1.0,
                 generator=None):
        if dtype != torch.float64:
            raise ValueError("the Slater kernels are FP64")
        M = torch.randn(n_orbitals, n_fermions, dtype=dtype, generator=generator) * init_std
        super().__init__([M])
        self.n_orbitals = n_orbitals
        self.n_fermions = n_fermions
        self.n_sites = n_orbitals // 2
        self._state = None
        self._state_key = None
        self._m_snapshot = None

    @property
    def M(self):
        return self.params[0]

    def forward_log(self, x):
        return kernels.slater_refresh(self.M.detach(), x.contiguous())

    # ---------------------------------------------------- walker state cache
    def _key(self, fxs):
        return (fxs.data_ptr(), tuple(fxs.shape), fxs._version,
                self.M.data_ptr(), self.M._version)

    def _params_unchanged(self):
        snap = self._m_snapshot
        return snap is not None and torch.equal(snap, self.M.detach())

    def walker_state(self, fxs, refresh=False):
        """State for the walker batch `fxs`, rebuilt exactly when fxs or M
        changed since it was last built (M is compared by value, since
        optimizers may write it through .data)."""
        fxs_ok = self._state is not None and self._state_key == self._key(fxs)
        if not refresh and fxs_ok and self._params_unchanged():
            return self._state
        B = fxs.shape[0]
        st = self._state
        if st is None or st.ainv.shape[0] != B or st.ainv.device != fxs.device:
            st = kernels.SlaterWalkerState.empty(B, self.n_sites, self.n_fermions, fxs.device)
        kernels.slater_refresh(self.M.detach(), fxs, st)
        self._state = st
        self._state_key = self._key(fxs)
        self._m_snapshot = self.M.detach().clone()
        return st

    def invalidate(self):
        self._state_key = None
