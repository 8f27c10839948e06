"""Oracle: learned Metropolis proposal (conv net -> move logits) with the
exact Metropolis-Hastings correction.  Test infrastructure only (see
oracle/__init__.py).

The reference has no learned proposal; its closest form is the external
(MPS) proposal of MetropolisMPSSamplerSpinful (sampler.py:2120-2238), whose
acceptance is min(1, (P'/P) * (q/q')) (:2214-2218).  This module fixes the
build's design:

  * input: the occupation grid of an Lx x Ly lattice, channels (n_up, n_dn);
  * three 3x3 convolutions with zero padding, C channels, ReLU;
  * moves m = 3 e + t over the graph edges e = (i, j), i < j, in sorted order
    ((s, s+1) then (s, s+Ly) for s ascending), types t = 0 hop of an up
    electron, 1 hop of a down electron, 2 exchange of a down and an up
    electron; valid when t = 0: up bits of i and j differ, t = 1: down bits
    differ, t = 2: {x_i, x_j} = {1, 2};
  * logits[e, t] = bA[t] + A[t, :C] . h3[:, i] + A[t, C:] . h3[:, j];
  * q(m | x) = softmax over the valid moves; every move is its own reverse;
  * step: m ~ q(.|x) by inverse CDF (uniform u_c, moves in index order),
    x' = m(x), accept when u_a < exp(2 (ln|psi(x')| - ln|psi(x)|)) *
    q(m | x') / q(m | x).
Training target (heat bath over the move set): q*(m | x) proportional to
|psi(m(x))|^2 over the valid moves; loss = -sum_m q*(m) ln q(m | x), whose
gradient w.r.t. the logits is q - q*.
Parameter order of the flat vector: W1 (C, 2, 3, 3), b1, W2 (C, C, 3, 3),
b2, W3 (C, C, 3, 3), b3, A (3, 2C), bA (3).

BF16 form (forward(..., bf16=True)): the rounding of the matrix-core net
restated -- every activation h1, h2, h3 rounded to bfloat16 after the ReLU,
W2 and W3 rounded to bfloat16, sums in FP64 here (FP32 on the device), the
head unchanged.  bf16_round is pinned against torch's float32 -> bfloat16
conversion (round to nearest even).
"""
import numpy as np


def edges(Lx, Ly):
    out = []
    for s in range(Lx * Ly):
        x, y = divmod(s, Ly)
        if y + 1 < Ly:
            out.append((s, s + 1))
        if x + 1 < Lx:
            out.append((s, s + Ly))
    return out


def n_moves(Lx, Ly):
    return 3 * len(edges(Lx, Ly))


def param_shapes(C):
    return [(C, 2, 3, 3), (C,), (C, C, 3, 3), (C,), (C, C, 3, 3), (C,), (3, 2 * C), (3,)]


def unpack(theta, C):
    out, off = [], 0
    for sh in param_shapes(C):
        n = int(np.prod(sh))
        out.append(np.asarray(theta[off:off + n]).reshape(sh))
        off += n
    return out


def n_params(C):
    return int(sum(np.prod(s) for s in param_shapes(C)))


def init_params(rng, C, scale=1.0):
    """Kaiming-uniform convolutions, small head, zero biases."""
    ps = []
    for sh in param_shapes(C):
        if len(sh) == 4:
            fan = sh[1] * 9
            a = np.sqrt(6.0 / fan) * scale
            ps.append(rng.uniform(-a, a, size=sh))
        elif sh[0] == 3 and len(sh) == 2:
            ps.append(rng.normal(0.0, 0.1 * scale, size=sh))
        else:
            ps.append(np.zeros(sh))
    return np.concatenate([p.ravel() for p in ps])


def bf16_round(a):
    """Round to the nearest bfloat16 (ties to even), via float32; returns float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _conv(h, W, b):
    """h (Cin, Lx, Ly), W (Cout, Cin, 3, 3) -> (Cout, Lx, Ly), zero padding."""
    Cin, Lx, Ly = h.shape
    hp = np.zeros((Cin, Lx + 2, Ly + 2))
    hp[:, 1:-1, 1:-1] = h
    out = np.zeros((W.shape[0], Lx, Ly)) + b[:, None, None]
    for dx in range(3):
        for dy in range(3):
            out += np.einsum("oc,cxy->oxy", W[:, :, dx, dy], hp[:, dx:dx + Lx, dy:dy + Ly])
    return out


def forward(theta, C, Lx, Ly, x, bf16=False):
    """logits (3E,), valid mask (3E,), cache.  bf16: the matrix-core net's
    rounding (activations and W2, W3 in bfloat16)."""
    W1, b1, W2, b2, W3, b3, A, bA = unpack(theta, C)
    rnd = bf16_round if bf16 else (lambda v: v)
    if bf16:
        W2, W3 = bf16_round(W2), bf16_round(W3)
    x = np.asarray(x).reshape(Lx, Ly)
    h0 = np.stack([(x >> 1) & 1, x & 1]).astype(np.float64)
    z1 = _conv(h0, W1, b1)
    h1 = rnd(np.maximum(z1, 0))
    z2 = _conv(h1, W2, b2)
    h2 = rnd(np.maximum(z2, 0))
    z3 = _conv(h2, W3, b3)
    h3 = rnd(np.maximum(z3, 0)).reshape(C, -1)
    E = edges(Lx, Ly)
    xf = x.reshape(-1)
    logits = np.zeros(3 * len(E))
    valid = np.zeros(3 * len(E), dtype=bool)
    for e, (i, j) in enumerate(E):
        for t in range(3):
            logits[3 * e + t] = bA[t] + A[t, :C] @ h3[:, i] + A[t, C:] @ h3[:, j]
        valid[3 * e] = ((xf[i] >> 1) & 1) != ((xf[j] >> 1) & 1)
        valid[3 * e + 1] = (xf[i] & 1) != (xf[j] & 1)
        valid[3 * e + 2] = {int(xf[i]), int(xf[j])} == {1, 2}
    return logits, valid, (h0, z1, h1, z2, h2, z3, h3)


def log_q(logits, valid):
    """ln q over the moves (-inf for invalid)."""
    lv = np.where(valid, logits, -np.inf)
    m = lv.max()
    return lv - (m + np.log(np.exp(lv - m).sum()))


def apply_move(x, Lx, Ly, m):
    e, t = divmod(int(m), 3)
    i, j = edges(Lx, Ly)[e]
    y = np.array(x, copy=True)
    if t == 0:
        y[i] ^= 2
        y[j] ^= 2
    elif t == 1:
        y[i] ^= 1
        y[j] ^= 1
    else:
        y[i], y[j] = y[j], y[i]
    return y


def sample_move(logq, u):
    """Inverse CDF over the moves in index order (FP64 running sum)."""
    p = np.exp(logq)
    c = 0.0
    last = -1
    for m in range(len(p)):
        if p[m] == 0.0:
            continue
        c += p[m]
        last = m
        if u < c:
            return m
    return last


def mh_step(theta, C, Lx, Ly, x, logamp, u_c, u_a):
    """One proposal for one walker.  logamp(config) -> ln|psi|.
    Returns (new x, accepted, ln q forward, ln q reverse)."""
    lf, vf, _ = forward(theta, C, Lx, Ly, x)
    lq = log_q(lf, vf)
    if not vf.any():
        return np.array(x), False, 0.0, 0.0
    m = sample_move(lq, u_c)
    xp = apply_move(x, Lx, Ly, m)
    lr, vr, _ = forward(theta, C, Lx, Ly, xp)
    lqr = log_q(lr, vr)
    ratio = np.exp(2.0 * (logamp(xp) - logamp(x)) + lqr[m] - lq[m])
    if u_a < ratio:
        return xp, True, lq[m], lqr[m]
    return np.array(x), False, lq[m], lqr[m]


def heat_bath_target(Lx, Ly, x, valid, logamp):
    """q*(m) proportional to |psi(m(x))|^2 over the valid moves."""
    l = np.full(valid.shape, -np.inf)
    for m in np.nonzero(valid)[0]:
        l[m] = 2.0 * logamp(apply_move(x, Lx, Ly, m))
    mx = l[valid].max()
    p = np.where(valid, np.exp(l - mx), 0.0)
    return p / p.sum()


def backward(theta, C, Lx, Ly, x, dlogits):
    """Gradient of sum_m dlogits[m] * logits[m] w.r.t. theta."""
    W1, b1, W2, b2, W3, b3, A, bA = unpack(theta, C)
    _, _, (h0, z1, h1, z2, h2, z3, h3) = forward(theta, C, Lx, Ly, x)
    E = edges(Lx, Ly)
    N = Lx * Ly
    dA = np.zeros_like(A)
    dbA = np.zeros(3)
    dh3 = np.zeros((C, N))
    for e, (i, j) in enumerate(E):
        for t in range(3):
            g = dlogits[3 * e + t]
            dbA[t] += g
            dA[t, :C] += g * h3[:, i]
            dA[t, C:] += g * h3[:, j]
            dh3[:, i] += g * A[t, :C]
            dh3[:, j] += g * A[t, C:]
    dz3 = dh3.reshape(C, Lx, Ly) * (z3 > 0)

    def conv_back(hin, W, dz):
        Cin = hin.shape[0]
        hp = np.zeros((Cin, Lx + 2, Ly + 2))
        hp[:, 1:-1, 1:-1] = hin
        dW = np.zeros_like(W)
        dhp = np.zeros_like(hp)
        for dx in range(3):
            for dy in range(3):
                dW[:, :, dx, dy] = np.einsum("oxy,cxy->oc", dz, hp[:, dx:dx + Lx, dy:dy + Ly])
                dhp[:, dx:dx + Lx, dy:dy + Ly] += np.einsum("oc,oxy->cxy", W[:, :, dx, dy], dz)
        return dW, dz.sum(axis=(1, 2)), dhp[:, 1:-1, 1:-1]
    dW3, db3, dh2 = conv_back(h2, W3, dz3)
    dz2 = dh2 * (z2 > 0)
    dW2, db2, dh1 = conv_back(h1, W2, dz2)
    dz1 = dh1 * (z1 > 0)
    dW1, db1, _ = conv_back(h0, W1, dz1)
    return np.concatenate([g.ravel() for g in (dW1, db1, dW2, db2, dW3, db3, dA, dbA)])


def lp_draws(seed, step, walkers):
    """(u_choice, u_acc) of one learned-proposal step: Philox4x32-10 with
    counter (step lo, step hi, walker, 0x4C50), key (seed lo, seed hi);
    u = top 53 bits of (o0:o1) resp. (o2:o3) times 2^-53."""
    from .philox import TWO_M53, philox4x32_10
    w = np.asarray(walkers, dtype=np.uint64)
    seed, step = int(seed), int(step)
    o0, o1, o2, o3 = philox4x32_10(np.uint64(step & 0xFFFFFFFF),
                                   np.uint64((step >> 32) & 0xFFFFFFFF), w, 0x4C50,
                                   np.uint64(seed & 0xFFFFFFFF),
                                   np.uint64((seed >> 32) & 0xFFFFFFFF))
    x = (o0 << np.uint64(32)) | o1
    y = (o2 << np.uint64(32)) | o3
    return ((x >> np.uint64(11)).astype(np.float64) * TWO_M53,
            (y >> np.uint64(11)).astype(np.float64) * TWO_M53)
