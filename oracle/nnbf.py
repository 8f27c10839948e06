"""Oracle: neural-network backflow (NNBF) amplitudes and per-sample
log-derivatives.  Test infrastructure only (see oracle/__init__.py).

GPU/models/NNBF.py:36-162 (NNBF_GPU.amplitude):
  n   = [up | down] occupation (2N)
  h_k = act(W_k h_{k-1} + b_k)              (act = ReLU for the configs here)
  dM  = reshape(W_out h_L + b_out, (2N, Nf))
  psi = det((M + bf_scale * dM)[occ ascending, :])
Parameter order (flat O / dp layout): [M, W_1, b_1, ..., W_L, b_L, W_out, b_out]
(NNBF.py:44-48).
"""
import numpy as np

from .hubbard import quimb_to_netket


def unpack(params, n_layers):
    M = params[0]
    layers = [(params[1 + 2 * k], params[2 + 2 * k]) for k in range(n_layers)]
    Wo, bo = params[1 + 2 * n_layers], params[2 + 2 * n_layers]
    return M, layers, Wo, bo


def bf16_round(a):
    """Round to bfloat16 (round to nearest, ties to even), returned as
    float64: the operand rounding of the BF16 GEMM mode."""
    b = np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32)
    b = (b + (((b >> 16) & 1) + 0x7FFF)) & np.uint32(0xFFFF0000)
    return b.view(np.float32).astype(np.float64)


def forward(params, x, n_layers, bf_scale, act="relu", bf16_gemm=False):
    """Returns (sign, ln|psi|, cache) for configs x (B, N).  bf16_gemm: the
    operands of the dense layers after the first (activations and weights)
    are rounded to bfloat16, as in the BF16 GEMM mode."""
    M, layers, Wo, bo = unpack(params, n_layers)
    x = np.atleast_2d(x)
    n = quimb_to_netket(x).astype(np.float64)              # (B, 2N)
    rnd = bf16_round if bf16_gemm else (lambda v: v)
    hs, zs = [n], []
    h = n
    for k, (W, b) in enumerate(layers):
        z = (h @ W.T + b) if k == 0 else (rnd(h) @ rnd(W).T + b)
        zs.append(z)
        h = np.maximum(z, 0.0) if act == "relu" else np.tanh(z)
        hs.append(h)
    D = rnd(h) @ rnd(Wo).T + bo                             # (B, 2N*Nf)
    n_orb, nf = M.shape
    Mx = M[None] + bf_scale * D.reshape(-1, n_orb, nf)
    occ = np.array([np.nonzero(r)[0] for r in quimb_to_netket(x)])
    A = np.take_along_axis(Mx, occ[:, :, None], axis=1)    # (B, Nf, Nf)
    s, l = np.linalg.slogdet(A)
    return s, np.maximum(l, np.log(1e-45)), {"hs": hs, "zs": zs, "A": A, "occ": occ}


def log_derivative(params, x, n_layers, bf_scale, act="relu"):
    """Per-sample d ln|psi| / d theta in the flat parameter order."""
    M, layers, Wo, bo = unpack(params, n_layers)
    s, l, c = forward(params, x, n_layers, bf_scale, act)
    B = c["A"].shape[0]
    n_orb, nf = M.shape
    Ainv = np.linalg.inv(c["A"])
    gM = np.zeros((B, n_orb, nf))
    for b in range(B):
        gM[b, c["occ"][b], :] = Ainv[b].T                  # d ln|det A| / d A = A^-T
    gD = bf_scale * gM.reshape(B, -1)                       # (B, 2N*Nf)
    out = [gM.reshape(B, -1)]
    grads_layers = []
    g = gD @ Wo                                             # d/d h_L
    gWo = np.einsum("bi,bk->bik", gD, c["hs"][-1]).reshape(B, -1)
    for k in range(n_layers - 1, -1, -1):
        z = c["zs"][k]
        dz = g * (z > 0) if act == "relu" else g * (1 - np.tanh(z) ** 2)
        W, _ = layers[k]
        gW = np.einsum("bi,bj->bij", dz, c["hs"][k]).reshape(B, -1)
        grads_layers.append((gW, dz))
        g = dz @ W
    for gW, gb in reversed(grads_layers):
        out += [gW, gb]
    out += [gWo, gD]
    return np.concatenate(out, axis=1)


def init_params(rng, n_sites, n_f, hidden, n_layers, m_std=0.1, xavier_out=True):
    """Seeded parameters for the benchmark shape: M ~ N(0, m_std^2), Xavier
    uniform weights, zero biases."""
    n_orb = 2 * n_sites

    def xavier(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, size=(fo, fi))
    params = [rng.normal(0.0, m_std, (n_orb, n_f))]
    fi = n_orb
    for _ in range(n_layers):
        params += [xavier(hidden, fi), np.zeros(hidden)]
        fi = hidden
    params += [xavier(n_orb * n_f, hidden) if xavier_out else np.zeros((n_orb * n_f, hidden)),
               np.zeros(n_orb * n_f)]
    return params


def local_energy(H, params, x, n_layers, bf_scale, act="relu", bf16_gemm=False,
                 chunk=None):
    """E_loc with the GPU term-major get_conn and NNBF amplitudes
    (GPU/vmc_utils.py:322-476); bf16_gemm as in forward.  chunk: walkers
    per pass (bounds the memory of the connected amplitudes)."""
    x = np.atleast_2d(x)
    if chunk is not None and x.shape[0] > chunk:
        return np.concatenate([local_energy(H, params, x[i:i + chunk], n_layers, bf_scale, act,
                                            bf16_gemm) for i in range(0, x.shape[0], chunk)])
    s0, l0, _ = forward(params, x, n_layers, bf_scale, act, bf16_gemm)
    etas, coeffs, bids = H.get_conn_batch(x)
    E = np.zeros(x.shape[0])
    if etas.shape[0]:
        s1, l1, _ = forward(params, etas, n_layers, bf_scale, act, bf16_gemm)
        np.add.at(E, bids, coeffs * s1 * s0[bids] * np.exp(l1 - l0[bids]))
    return E
