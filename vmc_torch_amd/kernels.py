"""Tensor-level wrappers of the libvmcx entry points (one call = one C-ABI
call, plus argument checks).  The reference-shaped classes in sampler.py,
models.py, vmc_utils.py and hamiltonian.py are built on these."""
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib, ptr, require, stream_ptr

F64, I64, I32, U8 = torch.float64, torch.int64, torch.int32, torch.uint8


def slater_nfp(n_f):
    nfp = lib().vmcx_slater_nfp(int(n_f))
    if nfp < 0:
        raise ValueError(f"n_f = {n_f} exceeds the 64-fermion limit of the Slater kernels")
    return nfp


@dataclass
class SlaterWalkerState:
    """Device-resident Slater state of a batch of walkers (see vmcx.h).  The
    dtype of ainv selects the form: float64 (plain) or float32 (the FP32
    form: FP32 ratios and updates, FP64 refresh)."""
    ainv: torch.Tensor        # [B, NFP, NFP] float64 / float32, column k of A^-1 contiguous
    slot: torch.Tensor        # [B, 2N] uint8
    sign: torch.Tensor        # [B] float64
    logabs: torch.Tensor      # [B] float64
    since_refresh: torch.Tensor  # [B] int32
    n_accept: torch.Tensor    # [B] int64

    @classmethod
    def empty(cls, B, n_sites, n_f, device, dtype=F64):
        if dtype not in (torch.float64, torch.float32):
            raise TypeError("the Slater walker state is float64 or float32")
        nfp = slater_nfp(n_f)
        return cls(
            ainv=torch.empty(B, nfp, nfp, dtype=dtype, device=device),
            slot=torch.empty(B, 2 * n_sites, dtype=U8, device=device),
            sign=torch.empty(B, dtype=F64, device=device),
            logabs=torch.empty(B, dtype=F64, device=device),
            since_refresh=torch.zeros(B, dtype=I32, device=device),
            n_accept=torch.zeros(B, dtype=I64, device=device),
        )


def _check_m(M, n_sites):
    require(M, F64, "M")
    if M.dim() != 2 or M.shape[0] != 2 * n_sites:
        raise ValueError(f"M must be (2*n_sites, n_f) = ({2 * n_sites}, n_f), got {tuple(M.shape)}")
    return M.shape[1]


def _check_configs(configs, device):
    require(configs, I64, "configs", device=device)
    if configs.dim() != 2:
        raise ValueError("configs must be (B, n_sites)")
    return configs.shape


def slater_refresh(M, configs, state=None, only_if_since=-1, sign=None, logabs=None):
    """Exact rebuild of a walker state (or, with state=None, just the
    batched (sign, ln|psi|) of arbitrary configurations)."""
    B, N = _check_configs(configs, M.device)
    n_f = _check_m(M, N)
    if state is None:
        sign = torch.empty(B, dtype=F64, device=M.device) if sign is None else sign
        logabs = torch.empty(B, dtype=F64, device=M.device) if logabs is None else logabs
        check(lib().vmcx_slater_refresh(B, N, n_f, ptr(M), ptr(configs), None, None,
                                        ptr(sign), ptr(logabs), None, -1,
                                        stream_ptr(M.device)))
        return sign, logabs
    nfp = slater_nfp(n_f)
    f32 = state.ainv.dtype == torch.float32
    require(state.ainv, state.ainv.dtype, "ainv", (B, nfp, nfp), M.device)
    require(state.slot, U8, "slot", (B, 2 * N), M.device)
    require(state.since_refresh, I32, "since_refresh", (B,), M.device)
    fn = lib().vmcx_slater_refresh_f32 if f32 else lib().vmcx_slater_refresh
    check(fn(B, N, n_f, ptr(M), ptr(configs), ptr(state.ainv), ptr(state.slot),
             ptr(state.sign), ptr(state.logabs), ptr(state.since_refresh), int(only_if_since),
             stream_ptr(M.device)))
    return state.sign, state.logabs


def slater_sweep(M, edges, configs, state, n_sweeps=1, seed=0, sweep_offset=0,
                 hopping_rate=0.25, refresh_every=64, draws=None, M32=None, offset_dev=None):
    """n_sweeps Metropolis sweeps in place.  draws = (u_choice, bits, u_acc),
    each (n_sweeps, n_edges, B), replaces the Philox stream.  With a float32
    state the sweep runs in FP32 from M32 (M rounded to FP32; formed here
    when not given) and the exact refresh in FP64 from M.  offset_dev: a
    one-element int64 device tensor added to sweep_offset when the kernel
    runs (graph replays advance it with counter_add)."""
    B, N = _check_configs(configs, M.device)
    n_f = _check_m(M, N)
    nfp = slater_nfp(n_f)
    require(edges, I32, "edges", device=M.device)
    E = edges.shape[0]
    f32 = state.ainv.dtype == torch.float32
    if f32:
        M32 = M.to(torch.float32).contiguous() if M32 is None else M32
        require(M32, torch.float32, "M32", tuple(M.shape), M.device)
    for name, t, dt, shp in (("ainv", state.ainv, state.ainv.dtype, (B, nfp, nfp)),
                             ("slot", state.slot, U8, (B, 2 * N)),
                             ("sign", state.sign, F64, (B,)),
                             ("logabs", state.logabs, F64, (B,)),
                             ("since_refresh", state.since_refresh, I32, (B,)),
                             ("n_accept", state.n_accept, I64, (B,))):
        require(t, dt, name, shp, M.device)
    uc = bits = ua = None
    if draws is not None:
        uc, bits, ua = draws
        require(uc, F64, "u_choice", (n_sweeps, E, B), M.device)
        require(bits, U8, "bits", (n_sweeps, E, B), M.device)
        require(ua, F64, "u_acc", (n_sweeps, E, B), M.device)
    head = (B, N, n_f, ptr(M)) + ((ptr(M32),) if f32 else ())
    fn = lib().vmcx_slater_sweep_f32 if f32 else lib().vmcx_slater_sweep
    check(fn(*head, E, ptr(edges), float(hopping_rate), int(n_sweeps),
             int(seed) & (2 ** 64 - 1), int(sweep_offset), ptr(uc), ptr(bits), ptr(ua),
             int(refresh_every), ptr(configs), ptr(state.ainv), ptr(state.slot),
             ptr(state.sign), ptr(state.logabs), ptr(state.since_refresh),
             ptr(state.n_accept), ptr(offset_dev), stream_ptr(M.device)))


def counter_add(counter, inc):
    """counter[0] += inc on the current stream (graph-capturable)."""
    require(counter, I64, "counter", (1,))
    check(lib().vmcx_counter_add(ptr(counter), int(inc), stream_ptr(counter.device)))


def slater_local_energy(M, edges, t, U, configs, state, out=None, M32=None):
    """E_loc of every walker from its state (a float32 state: ratios from
    M32, M rounded to FP32, formed here when not given)."""
    B, N = _check_configs(configs, M.device)
    n_f = _check_m(M, N)
    require(edges, I32, "edges", device=M.device)
    out = torch.empty(B, dtype=F64, device=M.device) if out is None else require(out, F64, "eloc", (B,))
    if state.ainv.dtype == torch.float32:
        M32 = M.to(torch.float32).contiguous() if M32 is None else M32
        require(M32, torch.float32, "M32", tuple(M.shape), M.device)
        check(lib().vmcx_slater_local_energy_f32(B, N, n_f, ptr(M32), edges.shape[0],
                                                 ptr(edges), float(t), float(U), ptr(configs),
                                                 ptr(state.ainv), ptr(state.slot), ptr(out),
                                                 stream_ptr(M.device)))
        return out
    check(lib().vmcx_slater_local_energy(B, N, n_f, ptr(M), edges.shape[0], ptr(edges),
                                         float(t), float(U), ptr(configs), ptr(state.ainv),
                                         ptr(state.slot), ptr(out), stream_ptr(M.device)))
    return out


def slater_logderiv(state, n_sites, n_f, dtype=F64, out=None):
    B = state.ainv.shape[0]
    np_ = 2 * n_sites * n_f
    if out is None:
        out = torch.empty(B, np_, dtype=dtype, device=state.ainv.device)
    require(out, dtype, "O")
    if out.dim() != 2 or out.shape[0] != B or out.shape[1] < np_:
        raise ValueError("O must be (B, >= 2*n_sites*n_f)")
    code = {torch.float64: 0, torch.float32: 1}[dtype]
    fn = lib().vmcx_slater_logderiv_f32 if state.ainv.dtype == torch.float32 \
        else lib().vmcx_slater_logderiv
    check(fn(B, n_sites, n_f, ptr(state.ainv), ptr(state.slot), ptr(out), code, out.stride(0),
             stream_ptr(state.ainv.device)))
    return out


def hubbard_get_conn(configs, edges, t, U):
    """(conn_etas [T, N] int64, coeffs [T] float64, batch_ids [T] int64),
    term-major as the reference's get_conn_batch_gpu."""
    B, N = _check_configs(configs, configs.device)
    require(edges, I32, "edges", device=configs.device)
    E = edges.shape[0]
    n_terms = 2 * E + N
    dev = configs.device
    counts = torch.empty(n_terms, dtype=I64, device=dev)
    s = stream_ptr(dev)
    check(lib().vmcx_hubbard_conn_count(B, N, E, ptr(edges), ptr(configs), ptr(counts), s))
    offsets = torch.cumsum(counts, 0) - counts
    total = int(counts.sum().item())
    etas = torch.empty(total, N, dtype=I64, device=dev)
    coeffs = torch.empty(total, dtype=F64, device=dev)
    bids = torch.empty(total, dtype=I64, device=dev)
    if total:
        check(lib().vmcx_hubbard_conn_fill(B, N, E, ptr(edges), float(t), float(U),
                                           ptr(configs), ptr(offsets.contiguous()),
                                           ptr(etas), ptr(coeffs), ptr(bids), s))
    return etas, coeffs, bids


__all__ = ["SlaterWalkerState", "slater_nfp", "slater_refresh", "slater_sweep",
           "slater_local_energy", "slater_logderiv", "hubbard_get_conn", "_lib"]


GRAM_CODES = {"fp64": 0, "fp32": 1, "bf16": 2, "bf16x3": 3}


class SRWorkspace:
    """Device workspace of vmcx_sr_minsr, grown on demand and reused."""

    def __init__(self):
        self.buf = None

    def get(self, Ns, Np, o_dtype, gram, flags, device):
        need = int(lib().vmcx_sr_workspace_bytes(int(Ns), int(Np), int(o_dtype),
                                                  GRAM_CODES[gram], int(flags)))
        if need < 0:
            raise ValueError("bad SR sizes")
        if self.buf is None or self.buf.numel() < need or self.buf.device != device:
            self.buf = None
            self.buf = torch.empty(need, dtype=torch.uint8, device=device)
        return self.buf


def sr_minsr(O, E, energy_mean, diag_shift, gram="fp64", max_iter=0, tol=1e-10,
             workspace=None, dp=None, with_stats=True, phase_ms=None, inplace=False):
    """Sample-space SR on the device.  O: (Ns, Np) float64 or float32,
    row-major (overwritten by the centred matrix when inplace=True).
    Returns (dp [Np] float64, stats [5] float64 device tensor or None, info
    [1] int32 device tensor).  phase_ms: optional list, filled with the 5
    phase durations and the mean diagonal-block launch of the factorisation
    (ms; the call then synchronises the stream)."""
    import ctypes
    if O.dtype not in (torch.float64, torch.float32):
        raise TypeError("O must be float64 or float32")
    require(O, O.dtype, "O")
    if O.dim() != 2 or O.stride(1) != 1:
        raise ValueError("O must be a row-major (Ns, Np) matrix")
    Ns, Np = O.shape
    od = 0 if O.dtype == torch.float64 else 1
    require(E, F64, "local_energies", (Ns,), O.device)
    flags = 1 if inplace else 0
    ws = (workspace or SRWorkspace()).get(Ns, Np, od, gram, flags, O.device)
    dp = torch.empty(Np, dtype=F64, device=O.device) if dp is None else require(dp, F64, "dp", (Np,))
    stats = torch.empty(5, dtype=F64, device=O.device) if with_stats else None
    info = torch.empty(1, dtype=I32, device=O.device)
    em = float("nan") if energy_mean is None else float(energy_mean)
    ph = (ctypes.c_float * 6)() if phase_ms is not None else None
    check(lib().vmcx_sr_minsr(Ns, Np, ptr(O), od, O.stride(0), flags, ptr(E), em,
                              float(diag_shift), GRAM_CODES[gram], int(max_iter), float(tol),
                              ptr(ws), ws.numel(), ptr(dp), ptr(stats), ptr(info),
                              ctypes.cast(ph, ctypes.c_void_p) if ph is not None else None,
                              stream_ptr(O.device)))
    if ph is not None:
        phase_ms[:] = list(ph)
    return dp, stats, info


class MinresWorkspace:
    """Device workspace of vmcx_sr_minres, grown on demand and reused."""

    def __init__(self):
        self.buf = None

    def get(self, Ns, Np, device):
        need = int(lib().vmcx_sr_minres_workspace_bytes(int(Ns), int(Np)))
        if need < 0:
            raise ValueError("bad SR sizes")
        if self.buf is None or self.buf.numel() < need or self.buf.device != device:
            self.buf = None
            self.buf = torch.empty(need, dtype=torch.uint8, device=device)
        return self.buf


def sr_minres(O, E, energy_mean, diag_shift, rtol=5e-5, maxiter=100, workspace=None,
              dp=None, with_stats=True):
    """Parameter-space SR by MINRES on the device (S never formed).
    O: (Ns, Np) float64 or float32, row-major, not centred.  Returns
    (dp [Np] float64, stats [5] float64 device tensor or None, info [1]
    int32 device tensor: 0 converged, maxiter not)."""
    if O.dtype not in (torch.float64, torch.float32):
        raise TypeError("O must be float64 or float32")
    require(O, O.dtype, "O")
    if O.dim() != 2 or O.stride(1) != 1:
        raise ValueError("O must be a row-major (Ns, Np) matrix")
    Ns, Np = O.shape
    od = 0 if O.dtype == torch.float64 else 1
    require(E, F64, "local_energies", (Ns,), O.device)
    ws = (workspace or MinresWorkspace()).get(Ns, Np, O.device)
    dp = torch.empty(Np, dtype=F64, device=O.device) if dp is None else require(dp, F64, "dp", (Np,))
    stats = torch.empty(5, dtype=F64, device=O.device) if with_stats else None
    info = torch.empty(1, dtype=I32, device=O.device)
    em = float("nan") if energy_mean is None else float(energy_mean)
    check(lib().vmcx_sr_minres(Ns, Np, ptr(O), od, O.stride(0), ptr(E), em, float(diag_shift),
                               float(rtol), int(maxiter), ptr(ws), ws.numel(), ptr(dp),
                               ptr(stats), ptr(info), stream_ptr(O.device)))
    return dp, stats, info


class FactoredO:
    """Per-sample log-derivatives in factored form, the input of
    vmcx_sr_minsr_factored.  F: (Ns, width) float64 row-major, the factor
    matrices side by side; factors: [(column offset, width)]; blocks:
    [(a, b, scale, w_offset, bias_offset)] with b = -1 for a plain block
    (O[:, w_offset:+k_a] = scale F_a) and otherwise O[:, w_offset:+k_a k_b] =
    scale F_a (x) F_b row by row, plus O[:, bias_offset:+k_a] = scale F_a when
    bias_offset >= 0.  n_params: the width of O."""

    def __init__(self, F, factors, blocks, n_params):
        if not isinstance(F, torch.Tensor) or F.device.type != "cuda" or F.dtype != F64:
            raise TypeError("F must be a float64 HIP tensor")
        if F.dim() != 2 or F.stride(1) != 1 or F.stride(0) < F.shape[1]:
            raise ValueError("F must be a 2-D tensor with unit column stride (rows of a "
                             "wider tensor are allowed)")
        self.F, self.factors, self.blocks = F, list(factors), list(blocks)
        self.n_params = int(n_params)

    @property
    def shape(self):
        return (self.F.shape[0], self.n_params)

    @property
    def dtype(self):
        return F64

    @property
    def device(self):
        return self.F.device

    def factor(self, f):
        c, k = self.factors[f]
        return self.F[:, c:c + k]

    def __getitem__(self, rows):
        return FactoredO(self.F[rows], self.factors, self.blocks, self.n_params)

    @staticmethod
    def cat(parts):
        p0 = parts[0]
        return FactoredO(torch.cat([p.F for p in parts]), p0.factors, p0.blocks, p0.n_params)

    def materialize(self):
        """The dense (Ns, n_params) O (for checks on small cases)."""
        Ns = self.F.shape[0]
        O = torch.zeros(Ns, self.n_params, dtype=F64, device=self.F.device)
        for a, b, sc, wo, bo in self.blocks:
            Fa = self.factor(a)
            if b < 0:
                O[:, wo:wo + Fa.shape[1]] = sc * Fa
                continue
            Fb = self.factor(b)
            ka, kb = Fa.shape[1], Fb.shape[1]
            O[:, wo:wo + ka * kb] = sc * (Fa.unsqueeze(2) * Fb.unsqueeze(1)).reshape(Ns, ka * kb)
            if bo >= 0:
                O[:, bo:bo + ka] = sc * Fa
        return O

    def _c_args(self):
        import ctypes
        nf, nb = len(self.factors), len(self.blocks)
        ptrs = (ctypes.c_void_p * nf)(*[self.F.data_ptr() + 8 * c for c, _ in self.factors])
        lds = (ctypes.c_int64 * nf)(*[self.F.stride(0)] * nf)
        ks = (ctypes.c_int32 * nf)(*[k for _, k in self.factors])
        blk = (_lib.SrBlockStruct * nb)()
        for i, (a, b, sc, wo, bo) in enumerate(self.blocks):
            blk[i].a_factor, blk[i].b_factor, blk[i].scale = a, b, float(sc)
            blk[i].w_offset, blk[i].bias_offset = wo, bo
        return nf, ptrs, lds, ks, nb, blk


class SRFactoredWorkspace:
    """Device workspace of vmcx_sr_minsr_factored, grown on demand."""

    def __init__(self):
        self.buf = None

    def get(self, FO):
        import ctypes
        nf, _, _, ks, nb, blk = FO._c_args()
        need = int(lib().vmcx_sr_factored_workspace_bytes(FO.F.shape[0], nf, ks, nb, blk))
        if need < 0:
            raise ValueError("bad factored SR sizes")
        if self.buf is None or self.buf.numel() < need or self.buf.device != FO.device:
            self.buf = None
            self.buf = torch.empty(need, dtype=torch.uint8, device=FO.device)
        return self.buf


def sr_minsr_factored(FO, E, energy_mean, diag_shift, workspace=None, dp=None,
                      with_stats=True, phase_ms=None):
    """minSR (FP64 Gram, Cholesky) on a FactoredO without forming O.
    Returns (dp [Np] float64, stats [5] or None, info [1] int32); phase_ms:
    optional list filled with [factor Grams, combination + centring,
    factorisation, solve, dp + statistics, diagonal-block launch] in ms."""
    import ctypes
    Ns, Np = FO.shape
    require(E, F64, "local_energies", (Ns,), FO.device)
    ws = (workspace or SRFactoredWorkspace()).get(FO)
    dp = torch.empty(Np, dtype=F64, device=FO.device) if dp is None else require(dp, F64, "dp", (Np,))
    stats = torch.empty(5, dtype=F64, device=FO.device) if with_stats else None
    info = torch.empty(1, dtype=I32, device=FO.device)
    em = float("nan") if energy_mean is None else float(energy_mean)
    ph = (ctypes.c_float * 6)() if phase_ms is not None else None
    nf, ptrs, lds, ks, nb, blk = FO._c_args()
    check(lib().vmcx_sr_minsr_factored(Ns, nf, ptrs, lds, ks, nb, blk, Np, ptr(E), em,
                                       float(diag_shift), ptr(ws), ws.numel(), ptr(dp),
                                       ptr(stats), ptr(info),
                                       ctypes.cast(ph, ctypes.c_void_p) if ph is not None else None,
                                       stream_ptr(FO.device)))
    if ph is not None:
        phase_ms[:] = list(ph)
    return dp, stats, info


def sr_factored_gradient(FO, E, energy_mean, workspace=None, g=None, with_stats=True):
    """The energy gradient g = O^T (E - energy_mean) / Ns of a FactoredO
    (no Gram).  Returns (g [Np] float64, stats [5] or None)."""
    Ns, Np = FO.shape
    require(E, F64, "local_energies", (Ns,), FO.device)
    ws = (workspace or SRFactoredWorkspace()).get(FO)
    g = torch.empty(Np, dtype=F64, device=FO.device) if g is None else require(g, F64, "g", (Np,))
    stats = torch.empty(5, dtype=F64, device=FO.device) if with_stats else None
    em = float("nan") if energy_mean is None else float(energy_mean)
    nf, ptrs, lds, ks, nb, blk = FO._c_args()
    check(lib().vmcx_sr_factored_gradient(Ns, nf, ptrs, lds, ks, nb, blk, Np, ptr(E), em,
                                          ptr(ws), ws.numel(), ptr(g), ptr(stats),
                                          stream_ptr(FO.device)))
    return g, stats


def energy_stats(E, out=None):
    """[mean, variance] of E (compensated sums) as a [2] float64 device
    tensor; nothing is copied to the host."""
    require(E, F64, "local_energies")
    if E.dim() != 1 or E.numel() < 1:
        raise ValueError("local_energies must be a non-empty vector")
    out = torch.empty(2, dtype=F64, device=E.device) if out is None else require(out, F64, "out", (2,))
    check(lib().vmcx_energy_stats(E.numel(), ptr(E), ptr(out), stream_ptr(E.device)))
    return out


def sgd_update(param, dp, lr):
    """param -= lr * dp (in place, flat view; float64 or float32 param, the
    update computed in FP64)."""
    if param.dtype not in (torch.float64, torch.float32):
        raise TypeError("param must be float64 or float32")
    require(param, param.dtype, "param")
    require(dp, F64, "dp", (param.numel(),), param.device)
    fn = lib().vmcx_sgd_update if param.dtype == torch.float64 else lib().vmcx_sgd_update_f32
    check(fn(param.numel(), ptr(param), ptr(dp), float(lr), stream_ptr(param.device)))


def syrk_f64(M, C, alpha=1.0, beta=0.0, splits=1, rowmajor=True):
    """C = beta C + alpha M M^T on the lower triangle of C (matrix cores).
    rowmajor=True: M is a row-major [n, K] tensor (row stride >= K);
    rowmajor=False: M is the column-major n x K matrix stored as a row-major
    [K, n] tensor (column k of M is row k).  C: column-major n x n, i.e. a
    row-major [n, n] tensor holding C^T; with splits > 1, C is [splits, n, n]
    and split s reduces the s-th chunk of k into C[s]."""
    if not isinstance(M, torch.Tensor) or M.device.type != "cuda" or M.dtype != F64:
        raise TypeError("M must be a float64 HIP tensor")
    if M.dim() != 2 or M.stride(1) != 1 or M.stride(0) < M.shape[1]:
        raise ValueError("M must be a 2-D tensor with unit column stride (a row slice "
                         "of a wider tensor is allowed)")
    n, K = (M.shape[0], M.shape[1]) if rowmajor else (M.shape[1], M.shape[0])
    shape = (n, n) if splits == 1 else (splits, n, n)
    require(C, F64, "C", shape, M.device)
    if not C.is_contiguous():
        raise ValueError("C must be contiguous")
    check(lib().vmcx_syrk_f64(n, K, ptr(M), M.stride(0), 1 if rowmajor else 0, float(alpha),
                              float(beta), ptr(C), n, int(splits), n * n,
                              stream_ptr(M.device)))
    return C


# ------------------------------------------------------------------ NNBF
def nnbf_struct(n_sites, n_f, hidden, act, gemm, bf_scale, w32, wbf, w64=None, path=0,
                det_fp32=False):
    """Fill the vmcx_nnbf_model struct.  gemm 0 / 1: w32 = dict M, W1, b1,
    W2, b2, Wout, bout of FP32 device copies (wbf: dict W2, Wout of BF16
    copies for gemm 1, else None); gemm 2: w64 = the same dict of FP64
    device tensors (w32 unused).  path (BF16 mode): 0 = the library's own
    matrix-core layers for the forward, 1 = rocBLAS.  det_fp32 (BF16 mode,
    path 0): the amplitude-only determinants eliminated in FP32."""
    s = _lib.NnbfModelStruct()
    s.n_sites, s.n_f, s.hidden = n_sites, n_f, hidden
    s.act, s.gemm, s.path = act, gemm, int(path)
    s.det_fp32 = 1 if det_fp32 else 0
    s.bf_scale = float(bf_scale)
    names = ("M", "W1", "b1", "W2", "b2", "Wout", "bout")
    for k in names:
        if gemm == 2:
            setattr(s, k, None)
            t = w64[k]
            require(t, torch.float64, k)
            setattr(s, k + "64" if k == "M" else k + "_64", t.data_ptr())
        else:
            t = w32[k]
            require(t, torch.float32, k)
            setattr(s, k, t.data_ptr())
            setattr(s, k + "64" if k == "M" else k + "_64", None)
    s.W2_bf16 = wbf["W2"].data_ptr() if wbf else None
    s.Wout_bf16 = wbf["Wout"].data_ptr() if wbf else None
    return s


def nnbf_workspace(struct, B, chunk, n_edges, device, holder):
    import ctypes
    need = int(lib().vmcx_nnbf_workspace_bytes(ctypes.byref(struct), int(B), int(chunk),
                                                int(n_edges)))
    if need < 0:
        raise ValueError("bad NNBF workspace sizes")
    buf = holder.get("ws")
    if buf is None or buf.numel() < need or buf.device != device:
        buf = torch.empty(need, dtype=torch.uint8, device=device)
        holder["ws"] = buf
    return buf


def to_bf16(x, out):
    check(lib().vmcx_to_bf16(x.numel(), ptr(x), ptr(out), stream_ptr(x.device)))
    return out


# ------------------------------------------------------------------ fPEPS
def peps_layout(Lx, Ly, D):
    """(dense_map int32 numpy [n_dense], per-site allowed sizes): the flat
    theta layout lists the Z2-even entries of each site tensor
    T_s[p, l, u, d, r] (site-major, row-major) -- virtual index k of a bond
    of dimension D has parity k >= D // 2, boundary legs have dimension 1."""
    import numpy as np
    pp = np.array([0, 1, 1, 0])

    def bpar(d):
        return np.zeros(1, dtype=np.int64) if d == 1 else (np.arange(d) >= d // 2).astype(np.int64)
    maps, sizes, k = [], [], 0
    for i in range(Lx * Ly):
        x, y = divmod(i, Ly)
        dl, du = (1 if y == 0 else D), (1 if x == 0 else D)
        dd, dr = (1 if x == Lx - 1 else D), (1 if y == Ly - 1 else D)
        par = (pp[:, None, None, None, None] + bpar(dl)[None, :, None, None, None]
               + bpar(du)[None, None, :, None, None] + bpar(dd)[None, None, None, :, None]
               + bpar(dr)[None, None, None, None, :]) % 2
        ok = (par == 0).ravel()
        m = np.full(ok.size, -1, dtype=np.int32)
        m[ok] = np.arange(k, k + int(ok.sum()), dtype=np.int32)
        k += int(ok.sum())
        maps.append(m)
        sizes.append(int(ok.sum()))
    return np.concatenate(maps), sizes


def peps_struct(Lx, Ly, D, dense_map, theta, nn=None, act=0, eta=0.0, mode=0, amp_fp32=False,
                chi=0, bmps_cache=True):
    """vmcx_peps_model from device tensors: dense_map int32, theta float64,
    nn = dict W1, b1, W2, b2, Wo, bo (float64) or None; mode 0 = additive,
    1 = multiplicative network correction; amp_fp32: amplitudes contracted
    in FP32 on the matrix cores; chi > 0: max_bond of the boundary MPS;
    bmps_cache: sweeps of a truncated plain fPEPS resume proposals from the
    walkers' cached boundaries (False: every proposal from scratch)."""
    s = _lib.PepsModelStruct()
    s.mode, s.amp_fp32 = int(mode), 1 if amp_fp32 else 0
    s.chi, s.bmps_nocache = max(int(chi), 0), 0 if bmps_cache else 1
    require(dense_map, I32, "dense_map")
    require(theta, F64, "theta")
    s.Lx, s.Ly, s.D, s.n_params = int(Lx), int(Ly), int(D), int(theta.numel())
    s.dense_map, s.theta = dense_map.data_ptr(), theta.data_ptr()
    s.hidden = 0 if nn is None else int(nn["b1"].numel())
    s.act, s.eta = int(act), float(eta)
    for k in ("W1", "b1", "W2", "b2", "Wo", "bo"):
        if nn is None:
            setattr(s, k, None)
        else:
            require(nn[k], F64, k)
            setattr(s, k, nn[k].data_ptr())
    return s


def peps_workspace(struct, B, chunk, n_edges, device, holder, reuse=False):
    """Device workspace of the fPEPS entry points (reuse: the environment-
    reuse ones), grown on demand and kept in `holder`."""
    import ctypes
    fn = lib().vmcx_peps_reuse_workspace_bytes if reuse else lib().vmcx_peps_workspace_bytes
    need = int(fn(ctypes.byref(struct), int(B), int(chunk), int(n_edges)))
    if need < 0:
        raise ValueError("bad fPEPS workspace sizes / model")
    buf = holder.get("ws")
    if buf is None or buf.numel() < need or buf.device != device:
        buf = None
        holder.pop("ws", None)
        buf = torch.empty(need, dtype=torch.uint8, device=device)
        holder["ws"] = buf
    return buf


# ------------------------------------------------------- learned proposal
def lp_n_params(C):
    return int(lib().vmcx_lp_n_params(int(C)))


def lp_n_moves(Lx, Ly):
    return int(lib().vmcx_lp_n_moves(int(Lx), int(Ly)))


def _lp_check(params, C, configs, Lx, Ly):
    require(params, torch.float32, "proposal params", (lp_n_params(C),))
    B, N = _check_configs(configs, params.device)
    if N != Lx * Ly:
        raise ValueError("configs must have Lx * Ly sites")
    return B, N


LP_PRECISION = {"fp32": 0, "bf16": 1}


def lp_logits(params, C, Lx, Ly, configs, precision="fp32"):
    """Logits of every move [B, n_moves] (FP64, -inf for invalid moves).
    precision "fp32": FP32 convolutions on the vector ALUs (any C <= 32);
    "bf16": layers 2 and 3 on the BF16 matrix cores (C = 32)."""
    B, _ = _lp_check(params, C, configs, Lx, Ly)
    out = torch.empty(B, lp_n_moves(Lx, Ly), dtype=F64, device=configs.device)
    check(lib().vmcx_lp_logits(ptr(params), int(C), int(Lx), int(Ly), B, ptr(configs), ptr(out),
                               LP_PRECISION[precision], stream_ptr(configs.device)))
    return out


def lp_propose(params, C, Lx, Ly, configs, seed, step, u_choice=None, precision="fp32"):
    """(move int32 [B], ln q(m|x) [B], proposed configs [B, N], ln q(m|x') [B])."""
    B, N = _lp_check(params, C, configs, Lx, Ly)
    dev = configs.device
    logits = torch.empty(B, lp_n_moves(Lx, Ly), dtype=F64, device=dev)
    move = torch.empty(B, dtype=I32, device=dev)
    lqf = torch.empty(B, dtype=F64, device=dev)
    lqr = torch.empty(B, dtype=F64, device=dev)
    prop = torch.empty(B, N, dtype=I64, device=dev)
    if u_choice is not None:
        require(u_choice, F64, "u_choice", (B,), dev)
    check(lib().vmcx_lp_propose(ptr(params), int(C), int(Lx), int(Ly), B, ptr(configs),
                                int(seed) & (2 ** 64 - 1), int(step), ptr(u_choice), ptr(logits),
                                ptr(move), ptr(lqf), ptr(prop), ptr(lqr), LP_PRECISION[precision],
                                stream_ptr(dev)))
    return move, lqf, prop, lqr


def lp_propose_given(params, C, Lx, Ly, configs, logits_x, seed, step, u_choice=None,
                     precision="fp32", step_dev=None, out=None):
    """lp_propose with the logits at `configs` given: (move, ln q(m|x),
    proposed configs, ln q(m|x'), logits at the proposed configs)."""
    B, N = _lp_check(params, C, configs, Lx, Ly)
    dev = configs.device
    M = lp_n_moves(Lx, Ly)
    require(logits_x, F64, "logits_x", (B, M), dev)
    if out is None:
        logits_p = torch.empty(B, M, dtype=F64, device=dev)
        move = torch.empty(B, dtype=I32, device=dev)
        lqf = torch.empty(B, dtype=F64, device=dev)
        lqr = torch.empty(B, dtype=F64, device=dev)
        prop = torch.empty(B, N, dtype=I64, device=dev)
    else:
        move, lqf, prop, lqr, logits_p = out
    if u_choice is not None:
        require(u_choice, F64, "u_choice", (B,), dev)
    check(lib().vmcx_lp_propose_given(ptr(params), int(C), int(Lx), int(Ly), B, ptr(configs),
                                      ptr(logits_x), int(seed) & (2 ** 64 - 1), int(step),
                                      ptr(u_choice), ptr(logits_p), ptr(move), ptr(lqf),
                                      ptr(prop), ptr(lqr), LP_PRECISION[precision],
                                      ptr(step_dev), stream_ptr(dev)))
    return move, lqf, prop, lqr, logits_p


def lp_commit_logits(accepted, logits_p, logits_x):
    B, M = logits_x.shape
    require(logits_p, F64, "logits_p", (B, M), logits_x.device)
    require(accepted, U8, "accepted", (B,), logits_x.device)
    check(lib().vmcx_lp_commit_logits(B, M, ptr(accepted), ptr(logits_p), ptr(logits_x),
                                      stream_ptr(logits_x.device)))


def masked_rows(mask, src, dst):
    """dst[b] = src[b] for the rows b with mask[b] != 0 (uint8 mask [B];
    src and dst of one shape and dtype, rows contiguous)."""
    B = dst.shape[0]
    require(mask, U8, "mask", (B,), dst.device)
    if src.shape != dst.shape or src.dtype != dst.dtype or src.device != dst.device:
        raise ValueError("masked_rows: src and dst differ in shape, dtype or device")
    if not (src.is_contiguous() and dst.is_contiguous()):
        raise ValueError("masked_rows: src and dst must be contiguous")
    row_bytes = (dst.numel() // max(B, 1)) * dst.element_size()
    check(lib().vmcx_masked_rows(B, row_bytes, ptr(mask), ptr(src), ptr(dst),
                                 stream_ptr(dst.device)))


def lp_accept(configs, sign, logabs, move, proposed, psign, plogabs, logq_f, logq_r, seed, step,
              u_acc=None, n_accept=None, step_dev=None, accepted=None):
    B, N = _check_configs(configs, configs.device)
    dev = configs.device
    acc = torch.empty(B, dtype=U8, device=dev) if accepted is None else accepted
    if u_acc is not None:
        require(u_acc, F64, "u_acc", (B,), dev)
    check(lib().vmcx_lp_accept(B, N, int(seed) & (2 ** 64 - 1), int(step), ptr(u_acc), ptr(move),
                               ptr(proposed), ptr(psign), ptr(plogabs), ptr(logq_f), ptr(logq_r),
                               ptr(configs), ptr(sign), ptr(logabs), ptr(n_accept), ptr(acc),
                               ptr(step_dev), stream_ptr(dev)))
    return acc


def lp_accept_commit(configs, sign, logabs, move, proposed, psign, plogabs, logq_f, logq_r,
                     seed, step, logits_p, logits_x, accepted, n_accept=None, step_dev=None,
                     u_acc=None, z1p=None, z1=None):
    """lp_accept, lp_commit_logits(accepted, logits_p, logits_x) and, with z1
    given, masked_rows(accepted, z1p, z1) in one launch (same results)."""
    B, N = _check_configs(configs, configs.device)
    dev = configs.device
    M = logits_x.shape[1]
    require(logits_p, F64, "logits_p", (B, M), dev)
    require(accepted, U8, "accepted", (B,), dev)
    if u_acc is not None:
        require(u_acc, F64, "u_acc", (B,), dev)
    row_bytes = 0
    if z1 is not None:
        if z1p is None or z1p.shape != z1.shape or z1p.dtype != z1.dtype or \
                not (z1.is_contiguous() and z1p.is_contiguous()):
            raise ValueError("lp_accept_commit: z1p and z1 must be contiguous and alike")
        row_bytes = (z1.numel() // max(B, 1)) * z1.element_size()
    check(lib().vmcx_lp_accept_commit(B, N, int(seed) & (2 ** 64 - 1), int(step), ptr(u_acc),
                                      ptr(move), ptr(proposed), ptr(psign), ptr(plogabs),
                                      ptr(logq_f), ptr(logq_r), ptr(configs), ptr(sign),
                                      ptr(logabs), ptr(n_accept), ptr(accepted), ptr(step_dev),
                                      int(M), ptr(logits_p), ptr(logits_x), int(row_bytes),
                                      ptr(z1p), ptr(z1), stream_ptr(dev)))
    return accepted


def lp_backward(params, C, Lx, Ly, configs, dlogits):
    B, _ = _lp_check(params, C, configs, Lx, Ly)
    require(dlogits, F64, "dlogits", (B, lp_n_moves(Lx, Ly)), configs.device)
    need = int(lib().vmcx_lp_backward_workspace_bytes(int(C), B))
    ws = torch.empty(max(need, 1), dtype=U8, device=configs.device)
    grad = torch.empty_like(params)
    check(lib().vmcx_lp_backward(ptr(params), int(C), int(Lx), int(Ly), B, ptr(configs),
                                 ptr(dlogits), ptr(ws), need, ptr(grad),
                                 stream_ptr(configs.device)))
    return grad


def lp_candidates(Lx, Ly, configs):
    B, N = _check_configs(configs, configs.device)
    out = torch.empty(B, lp_n_moves(Lx, Ly), N, dtype=I64, device=configs.device)
    check(lib().vmcx_lp_candidates(B, int(Lx), int(Ly), ptr(configs), ptr(out),
                                   stream_ptr(configs.device)))
    return out


def lp_axpy(a, x, y):
    require(x, torch.float32, "x")
    require(y, torch.float32, "y", (x.numel(),), x.device)
    check(lib().vmcx_lp_axpy(x.numel(), float(a), ptr(x), ptr(y), stream_ptr(x.device)))


def sym_images(configs, perms, inv_perms):
    """(images [B, G, N] int64, fermionic signs [B, G] float64) of the walkers
    under the G site permutations (int32 [G, N], with their inverses)."""
    B, N = _check_configs(configs, configs.device)
    G = perms.shape[0]
    require(perms, torch.int32, "perms", (G, N), configs.device)
    require(inv_perms, torch.int32, "inv_perms", (G, N), configs.device)
    images = torch.empty(B, G, N, dtype=I64, device=configs.device)
    signs = torch.empty(B, G, dtype=F64, device=configs.device)
    check(lib().vmcx_sym_images(B, N, G, ptr(configs), ptr(perms), ptr(inv_perms), ptr(images),
                                ptr(signs), stream_ptr(configs.device)))
    return images, signs
