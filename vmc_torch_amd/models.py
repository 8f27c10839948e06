"""Wavefunction models with the reference's GPU interface
(vmc_torch/GPU/models/_base.py, slater.py, NNBF.py).

A model keeps its parameters as an nn.ParameterList whose order defines the
flat O_k / dp layout (_base.py:41-44) and exposes forward / forward_log.
Beyond the reference interface every model implements the three hot-path
operations on a walker batch, all in libvmcx:
    mcmc_sweep(fxs, edges, n_sweeps, seed, sweep_offset, hopping_rate, draws)
    local_energy(fxs, H) -> E_loc [B]
    log_derivative(fxs, dtype) -> (O [B, Np], (sign, log|psi|))
and caches the device-side walker state they share.
"""
import contextlib
import ctypes

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
        self._snapshot = None
        self._state = None
        self._state_key = None

    def forward_log(self, x):
        raise NotImplementedError

    def forward(self, x):
        s, l = self.forward_log(x)
        return s * torch.exp(l)

    @property
    def n_params(self):
        return sum(p.numel() for p in self.params)

    # ----------------------------------------------- parameter change tracking
    def _flat_params(self):
        return torch.cat([p.detach().reshape(-1) for p in self.params])

    def params_changed(self):
        """True when the parameters differ (by value) from the last call that
        built derived device data; optimizers may write through .data.  While
        the model is frozen (graph capture) the parameters are taken as
        unchanged without looking (the comparison would wait for the device)."""
        if getattr(self, "_frozen", False):
            return False
        flat = self._flat_params()
        if self._snapshot is None or self._snapshot.shape != flat.shape or \
                self._snapshot.device != flat.device or not torch.equal(self._snapshot, flat):
            self._snapshot = flat.clone()
            return True
        return False

    def invalidate(self):
        self._state_key = None
        self._snapshot = None

    @contextlib.contextmanager
    def frozen(self):
        """Parameters promised constant inside the block: no value checks
        (each would synchronise the host), so the model's kernels can be
        captured into a HIP graph.  Derived device data is brought up to date
        on entry."""
        if getattr(self, "_frozen", False):     # nested: the outer block owns it
            yield self
            return
        self._frozen = False
        self._prepare_device_data()
        self._frozen = True
        try:
            yield self
        finally:
            self._frozen = False

    def _prepare_device_data(self):
        """Build the device copies the kernels read (weights, structs)."""

    # True when forward_log launches libvmcx kernels only (no library call that
    # may allocate or record memset nodes), so it can run under graph capture
    graph_safe = False

    def _walker_key(self, fxs):
        """Identity of the walker tensor (storage, shape, in-place version).
        Inference-mode tensors carry no version counter: their key never
        matches, so their state is rebuilt on every call."""
        if fxs.is_inference():
            return object()
        return (fxs.data_ptr(), tuple(fxs.shape), fxs._version)


class SlaterDeterminant_GPU(WavefunctionModel_GPU):
    """psi(x) = det(M[occupied, :]), occupied = netket-ordered occupied
    orbitals ascending (GPU/models/slater.py:15-61; CPU model.py:7437-7457).

    M: (n_orbitals = 2*N_sites, n_fermions), of `dtype` (slater.py:29-34).
    float64: the plain form, FP64 throughout.  float32: the FP32 form — M and
    the walkers' A^-1 in FP32, ratios and Sherman-Morrison updates in FP32,
    the exact refresh, sign and ln|psi| in FP64 (from M widened to FP64,
    which is exact).  `init_std` scales the N(0, 1) draw (the reference uses
    std 1; the benchmark shapes use 0.1).
    """

    def __init__(self, n_orbitals, n_fermions, dtype=torch.float64, init_std=1.0,
                 generator=None):
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("the Slater kernels take float64 or float32")
        M = (torch.randn(n_orbitals, n_fermions, dtype=torch.float64, generator=generator)
             * init_std).to(dtype)
        super().__init__([M])
        self.n_orbitals = n_orbitals
        self.n_fermions = n_fermions
        self.n_sites = n_orbitals // 2
        self._m64 = None

    @property
    def M(self):
        return self.params[0]

    @property
    def fp32(self):
        return self.M.dtype == torch.float32

    graph_safe = True

    def _prepare_device_data(self):
        if self.params_changed():
            self._state_key = None      # the walker state predates the change
        self._m_f64()

    def _m_f64(self):
        """M in FP64 for the exact refresh (M itself in the plain form)."""
        M = self.M.detach()
        if not self.fp32:
            return M
        if self._m64 is None or self._m64[0] is not self._snapshot:
            self._m64 = (self._snapshot, M.to(torch.float64).contiguous())
        return self._m64[1]

    def forward_log(self, x):
        self.params_changed()
        return kernels.slater_refresh(self._m_f64(), x.contiguous())

    def walker_state(self, fxs, refresh=False):
        """State for the walker batch `fxs`, rebuilt exactly when fxs or M
        changed since it was last built."""
        changed = self.params_changed()
        if not refresh and not changed and self._state is not None and \
                self._state_key == self._walker_key(fxs):
            return self._state
        B = fxs.shape[0]
        st = self._state
        if st is None or st.ainv.shape[0] != B or st.ainv.device != fxs.device or \
                st.ainv.dtype != self.M.dtype:
            st = kernels.SlaterWalkerState.empty(B, self.n_sites, self.n_fermions, fxs.device,
                                                 dtype=self.M.dtype)
        kernels.slater_refresh(self._m_f64(), fxs, st)
        self._state, self._state_key = st, self._walker_key(fxs)
        return st

    def mcmc_sweep(self, fxs, edges, n_sweeps, seed, sweep_offset, hopping_rate=0.25,
                   refresh_every=64, draws=None):
        st = self.walker_state(fxs)
        M32 = self.M.detach() if self.fp32 else None
        kernels.slater_sweep(self._m_f64(), edges, fxs, st, n_sweeps=n_sweeps, seed=seed,
                             sweep_offset=sweep_offset, hopping_rate=hopping_rate,
                             refresh_every=refresh_every, draws=draws, M32=M32)
        return st.sign, st.logabs, st.n_accept

    def local_energy(self, fxs, H):
        st = self.walker_state(fxs)
        edges = H.graph.edge_tensor(fxs.device, "graph")
        M32 = self.M.detach() if self.fp32 else None
        return kernels.slater_local_energy(self._m_f64(), edges, H.t, H.U, fxs, st, M32=M32)

    def log_derivative(self, fxs, dtype=torch.float64):
        st = self.walker_state(fxs)
        O = kernels.slater_logderiv(st, self.n_sites, self.n_fermions, dtype=dtype)
        return O, (st.sign.clone(), st.logabs.clone())


class NNBF_GPU(WavefunctionModel_GPU):
    """Neural-network backflow (GPU/models/NNBF.py:36-162):
    psi(x) = det((M + bf_scale * reshape(W_out h + b_out))[occupied]), with
    h the output of `n_layers` = 2 hidden layers of width `hidden_dim`.

    Parameter order [M, W1, b1, W2, b2, W_out, b_out] (NNBF.py:44-48).
    init="reference": M ~ N(0,1), kaiming-uniform W, zero b, W_out = b_out = 0
    (NNBF.py:86-111); init="xavier": M ~ N(0, m_std^2), Xavier-uniform for
    every W including W_out, zero biases (the benchmark configuration).
    gemm: the arithmetic of the network: "fp64" (the weights, layers and
    backflow in FP64, as the reference's float64 model; the default for
    dtype=float64), "fp32" (FP32 copies of the weights, FP32 GEMMs) or
    "bf16" (the two dense layers on BF16 operands with FP32 accumulation).
    bf16_layers (BF16 mode): "matrix_cores" (the library's own GEMMs for the
    forward, only the occupied orbital rows of the output layer formed) or
    "rocblas".  det: the determinant's arithmetic, "fp64" (default), or
    "fp32" in the BF16 mode on the matrix-core layers (the mixed-precision
    form: the amplitude-only passes -- forward, local energy, learned
    proposals -- eliminate in FP32 from entries formed in FP64; A^-1 for the
    plain sweep and O_k stays FP64; pair it with the learned sampler, whose
    acceptance then compares amplitudes of one precision).
    """

    ACTS = {"relu": 0, "tanh": 1}

    GEMMS = {"fp32": 0, "bf16": 1, "fp64": 2}

    def __init__(self, n_sites, n_fermions, hidden_dim=64, n_layers=2, activation="tanh",
                 bf_scale=0.01, dtype=torch.float64, gemm=None, init="reference",
                 m_std=1.0, generator=None, chunk=65536, bf16_layers="matrix_cores",
                 det="fp64"):
        if n_layers != 2:
            raise NotImplementedError("the NNBF kernels take exactly two hidden layers")
        if activation not in self.ACTS:
            raise ValueError(f"activation must be one of {list(self.ACTS)}")
        if gemm is None:
            gemm = "fp64" if dtype == torch.float64 else "fp32"
        if gemm not in self.GEMMS:
            raise ValueError(f"gemm must be one of {list(self.GEMMS)}")
        if gemm == "fp64" and dtype != torch.float64:
            raise ValueError("gemm='fp64' needs dtype=torch.float64")
        n_orb = 2 * n_sites
        g = generator
        if init == "reference":
            M = torch.randn(n_orb, n_fermions, dtype=dtype, generator=g)
            W1 = torch.empty(hidden_dim, n_orb, dtype=dtype)
            W2 = torch.empty(hidden_dim, hidden_dim, dtype=dtype)
            for w in (W1, W2):
                bound = (3.0 / w.shape[1]) ** 0.5 * (5.0 / 3.0)   # kaiming_uniform, tanh gain
                w.uniform_(-bound, bound, generator=g)
            Wo = torch.zeros(n_orb * n_fermions, hidden_dim, dtype=dtype)
        else:
            M = torch.randn(n_orb, n_fermions, dtype=dtype, generator=g) * m_std

            def xav(fo, fi):
                a = (6.0 / (fi + fo)) ** 0.5
                return torch.empty(fo, fi, dtype=dtype).uniform_(-a, a, generator=g)
            W1, W2, Wo = xav(hidden_dim, n_orb), xav(hidden_dim, hidden_dim), \
                xav(n_orb * n_fermions, hidden_dim)
        params = [M, W1, torch.zeros(hidden_dim, dtype=dtype), W2,
                  torch.zeros(hidden_dim, dtype=dtype), Wo,
                  torch.zeros(n_orb * n_fermions, dtype=dtype)]
        super().__init__(params)
        self.n_sites, self.n_fermions, self.hidden = n_sites, n_fermions, hidden_dim
        self.n_layers, self.activation, self.bf_scale = n_layers, activation, bf_scale
        self.gemm = gemm
        if bf16_layers not in ("matrix_cores", "rocblas"):
            raise ValueError("bf16_layers must be 'matrix_cores' or 'rocblas'")
        self.bf16_layers = bf16_layers
        if det not in ("fp64", "fp32"):
            raise ValueError("det must be 'fp64' or 'fp32'")
        if det == "fp32" and not (gemm == "bf16" and bf16_layers == "matrix_cores"):
            raise ValueError("det='fp32' needs gemm='bf16' on the matrix-core layers")
        self.det = det
        self.chunk = chunk
        self._w32 = None
        self._wbf = None
        self._ws = {}
        self._struct_obj = None

    # ------------------------------------------------------ device weights
    @property
    def graph_safe(self):
        """The BF16 forward on the library's own matrix-core layers launches
        libvmcx kernels only."""
        return self.gemm == "bf16" and self.bf16_layers == "matrix_cores"

    def _prepare_device_data(self):
        self._sync_weights()

    @property
    def z_dtype(self):
        """dtype of the first-layer walker state z1."""
        return torch.float64 if self.gemm == "fp64" else torch.float32

    def _sync_weights(self):
        if self.params_changed() or self._struct_obj is None:
            names = ["M", "W1", "b1", "W2", "b2", "Wout", "bout"]
            w64 = None
            if self.gemm == "fp64":
                w64 = {k: p.detach().contiguous() for k, p in zip(names, self.params)}
                self._w32 = w64          # keeps the tensors the struct points to alive
            else:
                self._w32 = {k: p.detach().to(torch.float32).contiguous()
                             for k, p in zip(names, self.params)}
            if self.gemm == "bf16":
                self._wbf = {k: kernels.to_bf16(self._w32[k], torch.empty_like(
                    self._w32[k], dtype=torch.bfloat16)) for k in ("W2", "Wout")}
            else:
                self._wbf = None
            self._struct_obj = kernels.nnbf_struct(
                self.n_sites, self.n_fermions, self.hidden, self.ACTS[self.activation],
                self.GEMMS[self.gemm], self.bf_scale, self._w32, self._wbf, w64,
                path=0 if self.bf16_layers == "matrix_cores" else 1,
                det_fp32=self.det == "fp32")
            self._state_key = None
            return True
        return False

    def _ws_for(self, B, chunk, n_edges, device):
        return kernels.nnbf_workspace(self._struct_obj, B, chunk, n_edges, device, self._ws)

    def forward_log(self, x):
        x = x.contiguous()
        kernels.require(x, torch.int64, "configs")
        self._sync_weights()
        B = x.shape[0]
        sign = torch.empty(B, dtype=torch.float64, device=x.device)
        logabs = torch.empty(B, dtype=torch.float64, device=x.device)
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(sign), kernels.ptr(logabs), None, kernels.stream_ptr(x.device)))
        return sign, logabs

    def first_layer(self, x, out=None):
        """(sign, ln|psi|, z1) of configurations x, z1 = W1 n(x) + b1 ([B][hidden],
        z_dtype): the state forward_log_moved starts from.  out: optional
        (sign, logabs, z1) buffers."""
        x = x.contiguous()
        kernels.require(x, torch.int64, "configs")
        self._sync_weights()
        B = x.shape[0]
        if out is None:
            out = (torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, self.hidden, dtype=self.z_dtype, device=x.device))
        sign, logabs, z1 = out
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(sign), kernels.ptr(logabs), kernels.ptr(z1),
            kernels.stream_ptr(x.device)))
        return sign, logabs, z1

    def forward_log_moved(self, x, xp, z1, out=None):
        """(sign, ln|psi|, z1p) of proposed configurations xp, row b differing
        from x[b] by a few moved electrons: the first layer is z1[b] (x[b]'s
        pre-activations, first_layer) updated by the W1 columns of the
        orbitals whose occupation changed, the rest of the forward as
        forward_log.  z1p: xp's pre-activations (the walker state after an
        accepted move).  out: optional (sign, logabs, z1p) buffers."""
        x, xp = x.contiguous(), xp.contiguous()
        kernels.require(x, torch.int64, "configs")
        kernels.require(xp, torch.int64, "proposed", tuple(x.shape), x.device)
        kernels.require(z1, self.z_dtype, "z1", (x.shape[0], self.hidden), x.device)
        self._sync_weights()
        B = x.shape[0]
        if out is None:
            out = (torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, self.hidden, dtype=self.z_dtype, device=x.device))
        sign, logabs, z1p = out
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward_moved(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(xp), kernels.ptr(z1),
            kernels.ptr(ws), ws.numel(), chunk, kernels.ptr(sign), kernels.ptr(logabs),
            kernels.ptr(z1p), kernels.stream_ptr(x.device)))
        return sign, logabs, z1p

    # --------------------------------------------------------- walker state
    def walker_state(self, fxs, refresh=False):
        changed = self._sync_weights()
        if not refresh and not changed and self._state is not None and \
                self._state_key == self._walker_key(fxs):
            return self._state
        B = fxs.shape[0]
        st = {"z1": torch.empty(B, self.hidden, dtype=self.z_dtype, device=fxs.device),
              "sign": torch.empty(B, dtype=torch.float64, device=fxs.device),
              "logabs": torch.empty(B, dtype=torch.float64, device=fxs.device),
              "n_accept": (self._state["n_accept"] if self._state is not None and
                           self._state["n_accept"].shape[0] == B
                           else torch.zeros(B, dtype=torch.int64, device=fxs.device))}
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward(
            ctypes.byref(self._struct_obj), B, kernels.ptr(fxs), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(st["sign"]), kernels.ptr(st["logabs"]), kernels.ptr(st["z1"]),
            kernels.stream_ptr(fxs.device)))
        self._state, self._state_key = st, self._walker_key(fxs)
        return st

    def mcmc_sweep(self, fxs, edges, n_sweeps, seed, sweep_offset, hopping_rate=0.25,
                   refresh_every=64, draws=None):
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges_host = edges.detach().to("cpu", torch.int32).contiguous()
        uc = bits = ua = None
        if draws is not None:
            uc, bits, ua = draws
        ws = self._ws_for(B, B, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_nnbf_sweep(
            ctypes.byref(self._struct_obj), B, edges_host.shape[0],
            ctypes.c_void_p(edges_host.data_ptr()), float(hopping_rate), int(n_sweeps),
            int(seed) & (2 ** 64 - 1), int(sweep_offset), kernels.ptr(uc), kernels.ptr(bits),
            kernels.ptr(ua), kernels.ptr(fxs), kernels.ptr(st["z1"]), kernels.ptr(st["sign"]),
            kernels.ptr(st["logabs"]), kernels.ptr(st["n_accept"]), kernels.ptr(ws), ws.numel(),
            kernels.stream_ptr(fxs.device)))
        return st["sign"], st["logabs"], st["n_accept"]

    def local_energy(self, fxs, H):
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges = H.graph.edge_tensor(fxs.device, "graph")
        chunk = self.chunk
        ws = self._ws_for(B, chunk, edges.shape[0], fxs.device)
        eloc = torch.empty(B, dtype=torch.float64, device=fxs.device)
        kernels.check(kernels.lib().vmcx_nnbf_local_energy(
            ctypes.byref(self._struct_obj), B, edges.shape[0], kernels.ptr(edges), float(H.t),
            float(H.U), kernels.ptr(fxs), kernels.ptr(st["sign"]), kernels.ptr(st["logabs"]),
            kernels.ptr(ws), ws.numel(), chunk, kernels.ptr(eloc),
            kernels.stream_ptr(fxs.device)))
        return eloc

    def log_derivative(self, fxs, dtype=torch.float32, out=None, chunk=4096):
        self._sync_weights()
        B = fxs.shape[0]
        Np = self.n_params
        O = torch.empty(B, Np, dtype=dtype, device=fxs.device) if out is None else out
        sign = torch.empty(B, dtype=torch.float64, device=fxs.device)
        logabs = torch.empty(B, dtype=torch.float64, device=fxs.device)
        chunk = min(chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, fxs.device)
        code = {torch.float64: 0, torch.float32: 1}[dtype]
        kernels.check(kernels.lib().vmcx_nnbf_logderiv(
            ctypes.byref(self._struct_obj), B, kernels.ptr(fxs), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(O), code, O.stride(0), kernels.ptr(sign), kernels.ptr(logabs),
            kernels.stream_ptr(fxs.device)))
        return O, (sign, logabs)

    def log_derivative_factored(self, fxs, chunk=4096, out=None):
        """The per-sample log-derivative as factors (kernels.FactoredO): row s
        holds [gM | n | dz1 | h1 | dz2 | h2] (2N nf + 2N + 4 hidden numbers in
        FP64) instead of the n_params-wide O row; O = [gM, dz1 n^T, dz1,
        dz2 h1^T, dz2, s gM h2^T, s gM] (s = bf_scale), NNBF.py:44-48 order.
        Returns (FactoredO, (sign, logabs))."""
        self._sync_weights()
        B = fxs.shape[0]
        N, H, nf = self.n_sites, self.hidden, self.n_fermions
        nD, n_orb = 2 * N * nf, 2 * N
        width = nD + n_orb + 4 * H
        F = torch.empty(B, width, dtype=torch.float64, device=fxs.device) if out is None else out
        sign = torch.empty(B, dtype=torch.float64, device=fxs.device)
        logabs = torch.empty(B, dtype=torch.float64, device=fxs.device)
        chunk = min(chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_nnbf_logderiv_factors(
            ctypes.byref(self._struct_obj), B, kernels.ptr(fxs), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(F), F.stride(0), kernels.ptr(sign), kernels.ptr(logabs),
            kernels.stream_ptr(fxs.device)))
        factors = [(0, nD), (nD, n_orb), (nD + n_orb, H), (nD + n_orb + H, H),
                   (nD + n_orb + 2 * H, H), (nD + n_orb + 3 * H, H)]
        o_w1 = nD
        o_b1 = o_w1 + H * n_orb
        o_w2 = o_b1 + H
        o_b2 = o_w2 + H * H
        o_wo = o_b2 + H
        o_bo = o_wo + nD * H
        blocks = [(0, -1, 1.0, 0, -1), (2, 1, 1.0, o_w1, o_b1), (4, 3, 1.0, o_w2, o_b2),
                  (0, 5, float(self.bf_scale), o_wo, o_bo)]
        return kernels.FactoredO(F, factors, blocks, self.n_params), (sign, logabs)


def uses_bmps(Lx, Ly, D, max_bond):
    """True when fPEPS amplitudes go through the chi-truncated boundary MPS:
    max_bond > 0 and either the lattice is beyond the exact column
    contraction (Lx <= 4, Ly <= 8) or max_bond is below the largest bond of
    a side, D^(Lx - Lx // 2)."""
    if max_bond <= 0:
        return False
    if Lx > 4 or Ly > 8:
        return True
    return max_bond < D ** (Lx - Lx // 2)


class fPEPS_Model_GPU(WavefunctionModel_GPU):
    """Z2-symmetric fermionic PEPS (GPU/models/pureTNS.py:24-156).
    max_bond <= 0 (or >= D^2 on Lx <= 4): exact contraction, column by
    column.  Otherwise the reference's boundary MPS with max_bond = chi
    (pureTNS.py:81-146): rows 0 .. Lx/2-1 from the top, Lx-1 .. Lx/2 from the
    bottom, each truncated left to right onto its chi leading Schmidt
    vectors after every absorbed row, the two boundaries then contracted
    exactly (chi <= 32, max(chi, D) * D <= 64, Lx * Ly <= 64, Ly <= 16;
    FP64 amplitudes: forward, sweep, local energy).

    Parameters: one flat FP64 vector of the Z2-even entries of the site
    tensors T_s[p, l, u, d, r] (site-major, row-major), the layout of
    kernels.peps_layout.  Fermionic convention: modes in the Hamiltonian's
    Jordan-Wigner order, psi(x) = (-1)^{C(P,2)} contract_s S_s with
    S_s = T_s[x_s] (-1)^{q(d)(q(r) + P_right(s))} (see include/vmcx.h).
    theta: initial vector (numpy or tensor); default: entries N(0, 1) from
    `generator`, each site tensor scaled to unit norm.
    amp: "fp64" (default) or "fp32": the amplitudes of forward_log, of the
    sweep's proposals and of the local energy's connected configurations
    contracted in FP32 on the matrix cores (the log-derivative stays FP64).
    """

    AMPS = ("fp64", "fp32")

    def __init__(self, Lx, Ly, D, theta=None, generator=None, dtype=torch.float64,
                 max_bond=-1, chunk=65536, _nn=None, act="relu", eta=0.0, _mode=0,
                 amp="fp64"):
        if dtype != torch.float64:
            raise ValueError("the fPEPS parameters are FP64 (amp='fp32' selects FP32 "
                             "contraction of the amplitudes)")
        if amp not in self.AMPS:
            raise ValueError(f"amp must be one of {self.AMPS}")
        self._bmps = uses_bmps(Lx, Ly, D, max_bond)
        if self._bmps:
            if amp != "fp64":
                raise ValueError("the boundary-MPS contraction (max_bond > 0) is FP64")
            cb = max(max_bond, D)
            if not (2 <= Lx <= 16 and 1 <= Ly <= 16 and Lx * Ly <= 64 and 1 <= D <= 4
                    and max_bond <= 32 and cb * D <= 64):
                raise NotImplementedError(
                    "boundary MPS needs 2 <= Lx <= 16, Ly <= 16, Lx * Ly <= 64, D <= 4, "
                    "max_bond <= 32, max(max_bond, D) * D <= 64")
        elif not (2 <= Lx <= 4 and 1 <= Ly <= 8 and 1 <= D <= 4):
            raise NotImplementedError("exact contraction needs 2 <= Lx <= 4, Ly <= 8, D <= 4 "
                                      "(max_bond > 0 selects the boundary MPS)")
        dmap, sizes = kernels.peps_layout(Lx, Ly, D)
        n = int(sum(sizes))
        if theta is None:
            th = torch.randn(n, dtype=torch.float64, generator=generator)
            off = 0
            for k in sizes:
                th[off:off + k] /= th[off:off + k].norm()
                off += k
        else:
            th = torch.as_tensor(theta, dtype=torch.float64).clone().reshape(-1)
            if th.numel() != n:
                raise ValueError(f"theta must have {n} entries")
        super().__init__([th] + ([] if _nn is None else list(_nn)))
        self.register_buffer("dense_map", torch.as_tensor(dmap, dtype=torch.int32))
        self.Lx, self.Ly, self.D, self.n_sites = Lx, Ly, D, Lx * Ly
        self.n_tn_params = n
        self.site_sizes = sizes
        self.chi = max_bond
        self.chunk = chunk
        self.act_name = act
        self.eta = float(eta)
        self.mode = int(_mode)
        self.amp = amp
        # sweeps of a truncated plain fPEPS resume each proposal from the
        # walker's cached boundaries (same amplitudes, bit for bit)
        self.bmps_cache = True
        self._struct_obj = None
        self._ws = {}

    # ------------------------------------------------------ device struct
    def _nn_dict(self):
        return None

    def _sync(self):
        changed = self.params_changed()
        if changed or self._struct_obj is None:
            for p in self.params:
                if not p.data.is_contiguous():
                    p.data = p.data.contiguous()
            self._struct_obj = kernels.peps_struct(
                self.Lx, self.Ly, self.D, self.dense_map, self.params[0].data,
                self._nn_dict(), 0 if self.act_name == "relu" else 1, self.eta, self.mode,
                amp_fp32=self.amp == "fp32", chi=self.chi if self._bmps else 0,
                bmps_cache=self.bmps_cache)
            self._state_key = None
            return True
        return False

    def _ws_for(self, B, chunk, n_edges, device):
        return kernels.peps_workspace(self._struct_obj, B, chunk, n_edges, device, self._ws)

    def forward_log(self, x):
        x = x.contiguous()
        kernels.require(x, torch.int64, "configs")
        self._sync()
        B = x.shape[0]
        sign = torch.empty(B, dtype=torch.float64, device=x.device)
        logabs = torch.empty(B, dtype=torch.float64, device=x.device)
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_peps_forward(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(ws), ws.numel(), chunk,
            kernels.ptr(sign), kernels.ptr(logabs), kernels.stream_ptr(x.device)))
        return sign, logabs

    def first_layer(self, x, out=None):
        """(sign, ln|psi|, z1) of configurations x, z1 = W1 n(x) + b1 ([B][hidden],
        z_dtype): the state forward_log_moved starts from.  out: optional
        (sign, logabs, z1) buffers."""
        x = x.contiguous()
        kernels.require(x, torch.int64, "configs")
        self._sync_weights()
        B = x.shape[0]
        if out is None:
            out = (torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, self.hidden, dtype=self.z_dtype, device=x.device))
        sign, logabs, z1 = out
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(sign), kernels.ptr(logabs), kernels.ptr(z1),
            kernels.stream_ptr(x.device)))
        return sign, logabs, z1

    def forward_log_moved(self, x, xp, z1, out=None):
        """(sign, ln|psi|, z1p) of proposed configurations xp, row b differing
        from x[b] by a few moved electrons: the first layer is z1[b] (x[b]'s
        pre-activations, first_layer) updated by the W1 columns of the
        orbitals whose occupation changed, the rest of the forward as
        forward_log.  z1p: xp's pre-activations (the walker state after an
        accepted move).  out: optional (sign, logabs, z1p) buffers."""
        x, xp = x.contiguous(), xp.contiguous()
        kernels.require(x, torch.int64, "configs")
        kernels.require(xp, torch.int64, "proposed", tuple(x.shape), x.device)
        kernels.require(z1, self.z_dtype, "z1", (x.shape[0], self.hidden), x.device)
        self._sync_weights()
        B = x.shape[0]
        if out is None:
            out = (torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, dtype=torch.float64, device=x.device),
                   torch.empty(B, self.hidden, dtype=self.z_dtype, device=x.device))
        sign, logabs, z1p = out
        chunk = min(self.chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, x.device)
        kernels.check(kernels.lib().vmcx_nnbf_forward_moved(
            ctypes.byref(self._struct_obj), B, kernels.ptr(x), kernels.ptr(xp), kernels.ptr(z1),
            kernels.ptr(ws), ws.numel(), chunk, kernels.ptr(sign), kernels.ptr(logabs),
            kernels.ptr(z1p), kernels.stream_ptr(x.device)))
        return sign, logabs, z1p

    # --------------------------------------------------------- walker state
    def walker_state(self, fxs, refresh=False):
        changed = self._sync()
        if not refresh and not changed and self._state is not None and \
                self._state_key == self._walker_key(fxs):
            return self._state
        B = fxs.shape[0]
        sign, logabs = self.forward_log(fxs)
        st = {"sign": sign, "logabs": logabs,
              "n_accept": (self._state["n_accept"] if self._state is not None and
                           self._state["n_accept"].shape[0] == B
                           else torch.zeros(B, dtype=torch.int64, device=fxs.device))}
        self._state, self._state_key = st, self._walker_key(fxs)
        return st

    def mcmc_sweep(self, fxs, edges, n_sweeps, seed, sweep_offset, hopping_rate=0.25,
                   refresh_every=64, draws=None):
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges_host = edges.detach().to("cpu", torch.int32).contiguous()
        uc = bits = ua = None
        if draws is not None:
            uc, bits, ua = draws
        ws = self._ws_for(B, B, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_peps_sweep(
            ctypes.byref(self._struct_obj), B, edges_host.shape[0],
            ctypes.c_void_p(edges_host.data_ptr()), float(hopping_rate), int(n_sweeps),
            int(seed) & (2 ** 64 - 1), int(sweep_offset), kernels.ptr(uc), kernels.ptr(bits),
            kernels.ptr(ua), kernels.ptr(fxs), kernels.ptr(st["sign"]),
            kernels.ptr(st["logabs"]), kernels.ptr(st["n_accept"]), kernels.ptr(ws), ws.numel(),
            kernels.stream_ptr(fxs.device)))
        return st["sign"], st["logabs"], st["n_accept"]

    def local_energy(self, fxs, H):
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges = H.graph.edge_tensor(fxs.device, "graph")
        chunk = self.chunk
        ws = self._ws_for(B, chunk, edges.shape[0], fxs.device)
        eloc = torch.empty(B, dtype=torch.float64, device=fxs.device)
        kernels.check(kernels.lib().vmcx_peps_local_energy(
            ctypes.byref(self._struct_obj), B, edges.shape[0], kernels.ptr(edges), float(H.t),
            float(H.U), kernels.ptr(fxs), kernels.ptr(st["sign"]), kernels.ptr(st["logabs"]),
            kernels.ptr(ws), ws.numel(), chunk, kernels.ptr(eloc),
            kernels.stream_ptr(fxs.device)))
        return eloc

    def log_derivative(self, fxs, dtype=torch.float64, out=None, chunk=4096):
        self._sync()
        B = fxs.shape[0]
        Np = self.n_params
        O = torch.empty(B, Np, dtype=dtype, device=fxs.device) if out is None else out
        sign = torch.empty(B, dtype=torch.float64, device=fxs.device)
        logabs = torch.empty(B, dtype=torch.float64, device=fxs.device)
        chunk = min(chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, fxs.device)
        code = {torch.float64: 0, torch.float32: 1}[dtype]
        kernels.check(kernels.lib().vmcx_peps_logderiv(
            ctypes.byref(self._struct_obj), B, kernels.ptr(fxs), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(O), code, O.stride(0), kernels.ptr(sign), kernels.ptr(logabs),
            kernels.stream_ptr(fxs.device)))
        return O, (sign, logabs)

    def log_derivative_factored(self, fxs, chunk=4096, out=None):
        """The per-sample log-derivative as factors (kernels.FactoredO): with
        the backflow MLP, row s holds [g | n | dz1 | h1 | dz2 | h2] in FP64
        (g = d ln|psi| / d theta_eff) and O = [g, dz1 n^T, dz1, dz2 h1^T, dz2,
        eta g h2^T, eta g]; plain fPEPS: O = g.  Returns (FactoredO,
        (sign, logabs))."""
        self._sync()
        B = fxs.shape[0]
        N = self.Lx * self.Ly
        P = self.params[0].numel()
        H = int(getattr(self, "hidden", 0) or 0)
        mult = H > 0 and self.mode == 1
        width = (P + 2 * N + 4 * H + (P if mult else 0)) if H > 0 else P
        F = torch.empty(B, width, dtype=torch.float64, device=fxs.device) if out is None else out
        sign = torch.empty(B, dtype=torch.float64, device=fxs.device)
        logabs = torch.empty(B, dtype=torch.float64, device=fxs.device)
        chunk = min(chunk, max(B, 1))
        ws = self._ws_for(B, chunk, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_peps_logderiv_factors(
            ctypes.byref(self._struct_obj), B, kernels.ptr(fxs), kernels.ptr(ws), ws.numel(),
            chunk, kernels.ptr(F), F.stride(0), kernels.ptr(sign), kernels.ptr(logabs),
            kernels.stream_ptr(fxs.device)))
        if H == 0:
            return kernels.FactoredO(F, [(0, P)], [(0, -1, 1.0, 0, -1)], P), (sign, logabs)
        n_orb = 2 * N
        factors = [(0, P), (P, n_orb), (P + n_orb, H), (P + n_orb + H, H),
                   (P + n_orb + 2 * H, H), (P + n_orb + 3 * H, H)]
        # the network's upstream: g itself (additive) or g theta, a seventh
        # factor after h2 (multiplicative)
        up = 0
        if mult:
            factors.append((P + n_orb + 4 * H, P))
            up = 6
        o_w1 = P
        o_b1 = o_w1 + H * n_orb
        o_w2 = o_b1 + H
        o_b2 = o_w2 + H * H
        o_wo = o_b2 + H
        o_bo = o_wo + P * H
        blocks = [(0, -1, 1.0, 0, -1), (2, 1, 1.0, o_w1, o_b1), (4, 3, 1.0, o_w2, o_b2),
                  (up, 5, float(self.nn_eta), o_wo, o_bo)]
        return kernels.FactoredO(F, factors, blocks, self.n_params), (sign, logabs)



class fPEPS_Model_reuse_GPU(fPEPS_Model_GPU):
    """fPEPS with environment reuse (GPU/models/pureTNS.py:163-2375,
    fPEPS_Model_reuse_GPU): the same amplitude as fPEPS_Model_GPU, with
    sweeps and local energies that evaluate each proposal or connected
    configuration from its walker's cached column environments, contracting
    only the one or two columns the move touches (vmcx_peps_sweep_reuse,
    vmcx_peps_local_energy_reuse).  In the exact regime (Lx <= 4, D <= 4,
    max_bond <= 0 or >= D^2) the environments are the exact ones of the
    column-by-column contraction, rebuilt incrementally as accepted moves
    change columns.  With a truncating max_bond = chi the environments are
    the walkers' cached boundary MPSs of both sides, row by row: a proposal
    or connected configuration resumes each side from the boundary of the
    rows before its first changed row (vmcx_peps_sweep / vmcx_peps_local_
    energy with the boundary cache), which gives the full recontraction's
    amplitude bit for bit; the reference's amplitude_reuse
    (pureTNS.py:329-563) contracts the changed rows between a cached top and
    bottom environment instead, a different truncation order.
    `mcmc_sweep` and `local_energy` use the reuse path, so the plain sampler
    and evaluate_energy get it too; the `_full` methods keep the full
    contraction for comparison."""

    def __init__(self, Lx, Ly, D, theta=None, generator=None, dtype=torch.float64,
                 max_bond=-1, chunk=65536, **kwargs):
        if kwargs.get("amp", "fp64") != "fp64":
            raise ValueError("the environment-reuse entry points contract in FP64")
        super().__init__(Lx, Ly, D, theta=theta, generator=generator, dtype=dtype,
                         max_bond=max_bond, chunk=chunk)
        self.radius = 0

    def _ws_reuse(self, B, chunk, n_edges, device):
        return kernels.peps_workspace(self._struct_obj, B, chunk, n_edges, device, self._ws,
                                      reuse=True)

    def mcmc_sweep(self, fxs, edges, n_sweeps, seed, sweep_offset, hopping_rate=0.25,
                   refresh_every=64, draws=None):
        if self._bmps:   # the walkers' cached boundaries
            return fPEPS_Model_GPU.mcmc_sweep(self, fxs, edges, n_sweeps, seed, sweep_offset,
                                              hopping_rate, refresh_every, draws)
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges_host = edges.detach().to("cpu", torch.int32).contiguous()
        uc = bits = ua = None
        if draws is not None:
            uc, bits, ua = draws
        ws = self._ws_reuse(B, B, 0, fxs.device)
        kernels.check(kernels.lib().vmcx_peps_sweep_reuse(
            ctypes.byref(self._struct_obj), B, edges_host.shape[0],
            ctypes.c_void_p(edges_host.data_ptr()), float(hopping_rate), int(n_sweeps),
            int(seed) & (2 ** 64 - 1), int(sweep_offset), kernels.ptr(uc), kernels.ptr(bits),
            kernels.ptr(ua), kernels.ptr(fxs), kernels.ptr(st["sign"]),
            kernels.ptr(st["logabs"]), kernels.ptr(st["n_accept"]), kernels.ptr(ws), ws.numel(),
            kernels.stream_ptr(fxs.device)))
        return st["sign"], st["logabs"], st["n_accept"]

    def local_energy(self, fxs, H):
        if self._bmps:
            return fPEPS_Model_GPU.local_energy(self, fxs, H)
        st = self.walker_state(fxs)
        B = fxs.shape[0]
        edges = H.graph.edge_tensor(fxs.device, "graph")
        chunk = self.chunk
        ws = self._ws_reuse(B, chunk, edges.shape[0], fxs.device)
        eloc = torch.empty(B, dtype=torch.float64, device=fxs.device)
        kernels.check(kernels.lib().vmcx_peps_local_energy_reuse(
            ctypes.byref(self._struct_obj), B, edges.shape[0], kernels.ptr(edges), float(H.t),
            float(H.U), kernels.ptr(fxs), kernels.ptr(st["sign"]), kernels.ptr(st["logabs"]),
            kernels.ptr(ws), ws.numel(), chunk, kernels.ptr(eloc),
            kernels.stream_ptr(fxs.device)))
        return eloc

    def _full(self, fn, *args, **kwargs):
        if not self._bmps:
            return fn(self, *args, **kwargs)
        keep = self.bmps_cache
        self.bmps_cache = False   # every row recontracted
        self._struct_obj = None
        try:
            return fn(self, *args, **kwargs)
        finally:
            self.bmps_cache = keep
            self._struct_obj = None

    def mcmc_sweep_full(self, *args, **kwargs):
        return self._full(fPEPS_Model_GPU.mcmc_sweep, *args, **kwargs)

    def local_energy_full(self, fxs, H):
        return self._full(fPEPS_Model_GPU.local_energy, fxs, H)

    # the reference's names for the reuse path
    def forward_reuse_log(self, x, **kwargs):
        """(sign, ln|psi|) of x; the cached environments are internal to the
        sweep and the local energy, so this is the full contraction."""
        return self.forward_log(x)

class NNfTNS_Model_GPU(fPEPS_Model_GPU):
    """NN-fTNS (GPU/models/NNfTNS.py:280-955): the fPEPS above with a
    configuration-dependent correction of every tensor entry by an MLP.
    correction="additive": theta(x) = theta + nn_eta * MLP(n(x)), the
    reference's _apply_backflow (NNfTNS.py:390-437); "multiplicative":
    theta(x) = theta * (1 + nn_eta * MLP(n(x))), the multiplicative site-tensor
    correction of the NN-fTNS benchmark configuration (no reference form).
    MLP: n = [up | down] occupations (2N) -> hidden_dim -> hidden_dim ->
    n_tn_params, ReLU or tanh, FP64.  Parameter order [theta, W1, b1, W2, b2,
    Wo, bo].  init: Xavier-uniform hidden layers, output layer
    N(0, init_scale^2), zero biases (the model starts near the plain fPEPS)."""

    CORRECTIONS = {"additive": 0, "multiplicative": 1}

    def __init__(self, Lx, Ly, D, hidden_dim=128, nn_eta=1.0, activation="relu",
                 init_scale=1e-2, theta=None, generator=None, dtype=torch.float64, max_bond=-1,
                 chunk=65536, correction="additive", amp="fp64"):
        if activation not in ("relu", "tanh"):
            raise ValueError("activation must be relu or tanh")
        if correction not in self.CORRECTIONS:
            raise ValueError(f"correction must be one of {list(self.CORRECTIONS)}")
        _, sizes = kernels.peps_layout(Lx, Ly, D)
        n = int(sum(sizes))
        N = Lx * Ly
        g = generator

        def xav(fo, fi):
            a = (6.0 / (fi + fo)) ** 0.5
            return torch.empty(fo, fi, dtype=torch.float64).uniform_(-a, a, generator=g)
        th = None
        if theta is None:
            th = torch.randn(n, dtype=torch.float64, generator=g)
            off = 0
            for k in sizes:
                th[off:off + k] /= th[off:off + k].norm()
                off += k
        else:
            th = theta
        W1, W2 = xav(hidden_dim, 2 * N), xav(hidden_dim, hidden_dim)
        Wo = torch.randn(n, hidden_dim, dtype=torch.float64, generator=g) * init_scale
        nnp = [W1, torch.zeros(hidden_dim, dtype=torch.float64), W2,
               torch.zeros(hidden_dim, dtype=torch.float64), Wo,
               torch.zeros(n, dtype=torch.float64)]
        super().__init__(Lx, Ly, D, theta=th, generator=g, dtype=dtype, max_bond=max_bond,
                         chunk=chunk, _nn=nnp, act=activation, eta=nn_eta,
                         _mode=self.CORRECTIONS[correction], amp=amp)
        self.hidden = hidden_dim
        self.nn_eta = float(nn_eta)
        self.correction = correction

    def _nn_dict(self):
        names = ["W1", "b1", "W2", "b2", "Wo", "bo"]
        return {k: p.data for k, p in zip(names, self.params[1:])}
