"""Oracle: Z2-symmetric fermionic PEPS (fPEPS) and NN-fTNS amplitudes.
Test infrastructure only (see oracle/__init__.py).

The reference builds these with quimb + symmray (GPU/models/pureTNS.py:24-156,
GPU/models/NNfTNS.py:280-955), neither of which is vendored or installed, so
the fermionic sign convention at that boundary is unpinned.  This module fixes
a convention of its own and pins it against the definition (a Grassmann tensor
network, brute force below):

  * modes in the Jordan-Wigner order of the Hamiltonian: sites row-major
    (site i = x*Ly + y), within a site (down, up); |x> is the product of the
    creation operators in that order; local states p = 0, 1 (down), 2 (up),
    3 (both = c_down^+ c_up^+), parity(p) = 1 for p in {1, 2};
  * site tensor T_s[p, l, u, d, r], Z2-even (entries whose index parities
    sum to odd vanish); virtual index k of a bond of dimension D has parity
    k >= D/2 (D = 1: even); T_s = sum T (c^+)^p th_l th_u th_d th_r with the
    Grassmann variables in that order; a bond pairs th (at the earlier site in
    the JW order) with its partner at the later site, contracted when adjacent
    in that order.

Bringing every (c^+)^p to the front in JW order costs (-1)^{C(P,2)}, P the
number of singly occupied sites; bringing the bond partners together costs
(-1) per crossing of odd bonds drawn as arcs over the JW line.  With the leg
order (l, u, d, r) no two arcs sharing a site cross; the remaining crossings
(vertical bonds of one row interface with each other, and with the vertical
bonds of the interface above at larger column) reduce, with the Z2 constraint
of each tensor, to one local factor per sliced tensor:

  psi(x) = (-1)^{C(P,2)} * contract_s S_s,
  S_s[l,u,d,r] = T_s[x_s,l,u,d,r] * (-1)^{q(d) (q(r) + P_right(s))},

P_right(s) = number of singly occupied sites to the right of s in its row.
`amplitude_grassmann` evaluates the definition directly; the tests check the
two agree on small lattices.

NN-fTNS (NNfTNS.py:390-437): theta(x) = theta + eta * MLP(n(x)), additive on
the flat vector of allowed tensor entries; MLP input n = [up occupations |
down occupations] (2N), two hidden layers.
"""
import itertools

import numpy as np

from .hubbard import quimb_to_netket

PARITY_P = np.array([0, 1, 1, 0])


def bond_parity(D):
    """Parity of each virtual index of a bond of dimension D."""
    if D == 1:
        return np.zeros(1, dtype=np.int64)
    return (np.arange(D) >= D // 2).astype(np.int64)


def site_dims(Lx, Ly, D, x, y):
    """(l, u, d, r) dimensions of site (x, y): 1 on the open boundary."""
    return (1 if y == 0 else D, 1 if x == 0 else D, 1 if x == Lx - 1 else D,
            1 if y == Ly - 1 else D)


def allowed_mask(Lx, Ly, D, x, y):
    """Boolean (4, l, u, d, r) mask of the Z2-even entries."""
    dl, du, dd, dr = site_dims(Lx, Ly, D, x, y)
    par = (PARITY_P[:, None, None, None, None] + bond_parity(dl)[None, :, None, None, None]
           + bond_parity(du)[None, None, :, None, None]
           + bond_parity(dd)[None, None, None, :, None]
           + bond_parity(dr)[None, None, None, None, :]) % 2
    return par == 0


class PEPSShape:
    """Site dimensions, masks and the flat parameter layout (site-major,
    allowed entries of T_s in row-major (p, l, u, d, r) order)."""

    def __init__(self, Lx, Ly, D):
        self.Lx, self.Ly, self.D = Lx, Ly, D
        self.N = Lx * Ly
        self.masks = [allowed_mask(Lx, Ly, D, *divmod(i, Ly)) for i in range(self.N)]
        self.sizes = [int(m.sum()) for m in self.masks]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n_params = int(self.offsets[-1])

    def unpack(self, theta):
        """Flat allowed entries -> list of dense site tensors."""
        out = []
        for i, m in enumerate(self.masks):
            t = np.zeros(m.shape)
            t[m] = theta[self.offsets[i]:self.offsets[i + 1]]
            out.append(t)
        return out

    def pack(self, tensors):
        return np.concatenate([t[m] for t, m in zip(tensors, self.masks)])


def random_theta(shape, rng, normalise=True):
    """Allowed entries ~ N(0, 1), each site tensor then scaled to unit
    Frobenius norm when `normalise` (the benchmark initialisation)."""
    th = rng.normal(size=shape.n_params)
    if normalise:
        for i in range(shape.N):
            a, b = shape.offsets[i], shape.offsets[i + 1]
            th[a:b] /= np.linalg.norm(th[a:b])
    return th


def global_sign(x):
    """(-1)^{C(P,2)}, P = number of singly occupied sites."""
    P = int(np.sum(PARITY_P[np.asarray(x)]))
    return -1.0 if (P * (P - 1) // 2) % 2 else 1.0


def signed_slices(shape, tensors, x):
    """S_s[l,u,d,r] = T_s[x_s] * (-1)^{q(d) (q(r) + P_right(s))}."""
    Lx, Ly, D = shape.Lx, shape.Ly, shape.D
    x = np.asarray(x)
    out = []
    for i in range(shape.N):
        r_, c_ = divmod(i, Ly)
        dl, du, dd, dr = site_dims(Lx, Ly, D, r_, c_)
        pr = int(np.sum(PARITY_P[x[r_ * Ly + c_ + 1:(r_ + 1) * Ly]])) % 2
        qd = bond_parity(dd)[None, None, :, None]
        qr = bond_parity(dr)[None, None, None, :]
        sgn = np.where((qd * (qr + pr)) % 2 == 1, -1.0, 1.0)
        out.append(tensors[i][x[i]] * sgn)
    return out


def _bond_labels(Lx, Ly):
    """einsum index letters: per site (l, u, d, r) bond ids, -1 = open
    boundary (dimension 1)."""
    hid = {}
    vid = {}
    k = 0
    for xx in range(Lx):
        for yy in range(Ly - 1):
            hid[(xx, yy)] = k
            k += 1
    for xx in range(Lx - 1):
        for yy in range(Ly):
            vid[(xx, yy)] = k
            k += 1
    legs = []
    for xx in range(Lx):
        for yy in range(Ly):
            legs.append((hid.get((xx, yy - 1), -1), vid.get((xx - 1, yy), -1),
                         vid.get((xx, yy), -1), hid.get((xx, yy), -1)))
    return legs, k


def contract(shape, slices, open_site=None):
    """Exact contraction of the sliced network, site by site in row-major
    order (a zipper over the row interface): the state carries the down legs
    of the sites already absorbed in the current row, the horizontal bond,
    the up legs of the sites still to come, and (with `open_site`) the four
    legs of that site left open, which yields its environment
    d contract / d S_open [l, u, d, r] instead of a number.  Independent of
    the device's column-pair order."""
    Lx, Ly = shape.Lx, shape.Ly
    state = np.ones([1] * (Ly + 1))        # [h, u_0 .. u_{Ly-1}]
    nextra = 0
    for xx in range(Lx):
        for yy in range(Ly):
            s_idx = xx * Ly + yy
            S = slices[s_idx]                                 # [l, u, d, r]
            if s_idx == open_site:
                dims = S.shape
                S = np.einsum("ia,jb,kc,le->ijklabce", *[np.eye(d) for d in dims])
            state = np.tensordot(state, S, axes=([yy, yy + 1], [0, 1]))
            nd = state.ndim
            p0 = yy + (Ly - yy - 1) + nextra                  # position of the new d leg
            order = (list(range(yy)) + [p0, p0 + 1] + list(range(yy, yy + Ly - yy - 1))
                     + list(range(yy + Ly - yy - 1, p0)) + list(range(p0 + 2, nd)))
            state = np.transpose(state, order)
            if s_idx == open_site:
                nextra = 4
        state = np.moveaxis(state, Ly, 0)   # h (dimension 1) -> front: next row's [h, u..]
    if open_site is None:
        return float(state.reshape(-1)[0])
    return state.reshape(state.shape[Ly + 1:])


def amplitude(shape, theta, x):
    tensors = shape.unpack(theta)
    return global_sign(x) * contract(shape, signed_slices(shape, tensors, x))


def amplitudes(shape, theta, xs):
    return np.array([amplitude(shape, theta, x) for x in np.atleast_2d(xs)])


def log_derivative(shape, theta, x):
    """d ln|psi| / d theta (flat allowed entries): psi is linear in each
    site tensor, d psi / d S_s = environment of s (the network with s cut
    out), mapped back through the slice and its sign."""
    tensors = shape.unpack(theta)
    x = np.asarray(x)
    S = signed_slices(shape, tensors, x)
    g0 = global_sign(x)
    psi = g0 * contract(shape, S)
    out = np.zeros(shape.n_params)
    for i in range(shape.N):
        env = contract(shape, S, open_site=i)
        # dS/dT: S = T[x] * sign  ->  d psi / d T[x] = g0 * env * sign
        unit = np.ones_like(S[i])
        sgn = signed_slices_single(shape, unit, x, i)
        dT = np.zeros(shape.masks[i].shape)
        dT[x[i]] = g0 * env * sgn
        out[shape.offsets[i]:shape.offsets[i + 1]] = dT[shape.masks[i]]
    return out / psi


def signed_slices_single(shape, sliced, x, i):
    """The sign pattern of site i applied to an already sliced tensor."""
    Lx, Ly, D = shape.Lx, shape.Ly, shape.D
    r_, c_ = divmod(i, Ly)
    dl, du, dd, dr = site_dims(Lx, Ly, D, r_, c_)
    pr = int(np.sum(PARITY_P[np.asarray(x)[r_ * Ly + c_ + 1:(r_ + 1) * Ly]])) % 2
    qd = bond_parity(dd)[None, None, :, None]
    qr = bond_parity(dr)[None, None, None, :]
    return sliced * np.where((qd * (qr + pr)) % 2 == 1, -1.0, 1.0)


# --------------------------------------------------------- the definition
def amplitude_grassmann(shape, theta, x):
    """<x|psi> from the Grassmann tensor network by brute force over the
    virtual indices: for each assignment, the sign of the permutation of
    the odd symbols from [p_s, l_s, u_s, d_s, r_s for s in JW order] to
    [p_1 .. p_N, bond pairs (earlier, later) adjacent].  Small lattices."""
    Lx, Ly, D = shape.Lx, shape.Ly, shape.D
    tensors = shape.unpack(theta)
    x = np.asarray(x)
    legs, nb = _bond_labels(Lx, Ly)
    qb = bond_parity(D)
    total = 0.0
    for assign in itertools.product(range(D), repeat=nb):
        val = 1.0
        seq = []   # odd symbols in initial order: ("p", s) or ("b", bond, end)
        for s in range(shape.N):
            lg = legs[s]
            idx = tuple(assign[b] if b >= 0 else 0 for b in lg)
            val *= tensors[s][(x[s],) + idx]
            if val == 0.0:
                break
            if PARITY_P[x[s]]:
                seq.append(("p", s))
            for pos, b in enumerate(lg):      # order l, u, d, r
                if b >= 0 and qb[assign[b]]:
                    # l and u legs are the later end of their bond
                    seq.append(("b", b, 1 if pos < 2 else 0))
        if val == 0.0:
            continue
        # target: p symbols in site order, then each odd bond as (end 0, end 1)
        target = [sym for sym in seq if sym[0] == "p"]
        bonds = sorted({sym[1] for sym in seq if sym[0] == "b"})
        for b in bonds:
            target += [("b", b, 0), ("b", b, 1)]
        pos = {sym: k for k, sym in enumerate(target)}
        perm = [pos[sym] for sym in seq]
        inv = sum(1 for i in range(len(perm)) for j in range(i + 1, len(perm))
                  if perm[i] > perm[j])
        total += -val if inv % 2 else val
    return total


# --------------------------------------------------------------- NN-fTNS
def mlp_init(rng, n_in, hidden, n_out, out_scale):
    """Xavier-uniform hidden layers, N(0, out_scale^2) output layer, zero
    biases: the network starts near the plain fPEPS."""
    def xav(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, size=(fo, fi))
    return [xav(hidden, n_in), np.zeros(hidden), xav(hidden, hidden), np.zeros(hidden),
            rng.normal(0.0, out_scale, size=(n_out, hidden)), np.zeros(n_out)]


def mlp_forward(nn, xs, act="relu"):
    """(B, n_out) output and the cache for the backward pass."""
    W1, b1, W2, b2, Wo, bo = nn
    n = quimb_to_netket(np.atleast_2d(xs)).astype(np.float64)
    f = (lambda z: np.maximum(z, 0.0)) if act == "relu" else np.tanh
    z1 = n @ W1.T + b1
    h1 = f(z1)
    z2 = h1 @ W2.T + b2
    h2 = f(z2)
    return h2 @ Wo.T + bo, (n, z1, h1, z2, h2)


def theta_eff(theta, out, eta, mode="add"):
    """The configuration-dependent tensor entries.  mode "add": theta +
    eta out, the reference's additive backflow (GPU/models/NNfTNS.py:390-437);
    mode "mul": theta * (1 + eta out), the multiplicative site-tensor
    correction of the NN-fTNS benchmark configuration (no reference form)."""
    return theta + eta * out if mode == "add" else theta * (1.0 + eta * out)


def nnftn_amplitude(shape, theta, nn, eta, x, act="relu", mode="add"):
    out, _ = mlp_forward(nn, x, act)
    return amplitude(shape, theta_eff(theta, out[0], eta, mode), x)


def nnftn_log_derivative(shape, theta, nn, eta, x, act="relu", mode="add", chi=0):
    """[d/d theta, d/d (W1, b1, W2, b2, Wo, bo)] of ln|psi| for one config
    (chi > 0: with bmps_cheap_log_derivative's gradient of theta(x)).
    With g = d ln|psi| / d theta_eff: mode "add": d/d theta = g and
    d/d out = eta g; mode "mul": d/d theta = g (1 + eta out) and
    d/d out = eta g theta."""
    out, (n, z1, h1, z2, h2) = mlp_forward(nn, x, act)
    te = theta_eff(theta, out[0], eta, mode)
    g = bmps_cheap_log_derivative(shape, te, x, chi)[0] if chi > 0 else log_derivative(shape, te, x)
    W1, b1, W2, b2, Wo, bo = nn
    if mode == "add":
        dout = eta * g
    else:
        dout = eta * g * theta
        g = g * (1.0 + eta * out[0])
    dWo = np.outer(dout, h2[0])
    dh2 = Wo.T @ dout
    dz2 = dh2 * ((z2[0] > 0) if act == "relu" else (1 - np.tanh(z2[0]) ** 2))
    dW2 = np.outer(dz2, h1[0])
    dh1 = W2.T @ dz2
    dz1 = dh1 * ((z1[0] > 0) if act == "relu" else (1 - np.tanh(z1[0]) ** 2))
    dW1 = np.outer(dz1, n[0])
    return np.concatenate([g, dW1.ravel(), dz1, dW2.ravel(), dz2, dWo.ravel(), dout])


def local_energy(H, shape, theta, xs, nn=None, eta=0.0, act="relu", mode="add", chi=0):
    """E_loc via the term-major get_conn and full amplitudes (chi > 0: the
    chi-truncated boundary-MPS amplitudes)."""
    xs = np.atleast_2d(xs)

    def amp(c):
        if chi > 0:
            t = theta
            if nn is not None:
                t = theta_eff(theta, mlp_forward(nn, c, act)[0][0], eta, mode)
            s, l = bmps_log_amplitude(shape, t, c, chi)
            return s * np.exp(l)
        if nn is None:
            return amplitude(shape, theta, c)
        return nnftn_amplitude(shape, theta, nn, eta, c, act, mode)
    psi = np.array([amp(c) for c in xs])
    etas, coeffs, bids = H.get_conn_batch(xs)
    E = np.zeros(xs.shape[0])
    for e, c, b in zip(etas, coeffs, bids):
        E[b] += c * amp(e) / psi[b]
    return E


# ------------------------------------------- chi-truncated boundary MPS
# The reference contracts an fPEPS amplitude with quimb's boundary MPS
# (GPU/models/pureTNS.py:81-111): rows 0 .. Lx/2-1 absorbed from xmin, rows
# Lx-1 .. Lx/2 from xmax, each with max_bond=chi, cutoff=0, mode="mps",
# canonize=True (GPU/run_scripts/vmc_run_fpeps.py:126-136), then the two
# boundaries contracted exactly.  quimb is absent, so its sweep direction is
# unpinned; this restatement fixes it: after each absorbed row whose bonds
# exceed chi, canonize from the last column to the first (QR), then truncate
# from the first column to the last by SVD, keeping the chi largest singular
# values.  That is the sequential left-to-right Schmidt truncation of the
# boundary state (a TT-SVD of the state with maximal rank chi), which the
# tests check against that dense definition.

def _bmps_init(row, down):
    """The first row of a side as an MPS [a, v, b]: its open vertical legs
    (down legs from the top, up legs from the bottom)."""
    return [S[:, 0, :, :] if down else S[:, :, 0, :] for S in row]


def _bmps_absorb(mps, row, down):
    out = []
    for B, S in zip(mps, row):
        T = np.einsum("avb,lvdr->aldbr", B, S) if down else np.einsum("avb,luvr->alubr", B, S)
        a, l, v, b, r = T.shape
        out.append(T.reshape(a * l, v, b * r))
    return out


def _bmps_compress(mps, chi, gaps, canonize=True):
    """QR canonicalisation right to left, SVD truncation left to right; the
    relative gap (s[chi-1] - s[chi]) / s[0] of every truncation goes to
    `gaps` (a small gap makes the truncated state ill-conditioned).
    canonize=False exists for the tests only (the truncation without the
    canonical form, which they must reject)."""
    mps = list(mps)
    for y in range(len(mps) - 1, 0, -1) if canonize else ():
        a, v, b = mps[y].shape
        Q, R = np.linalg.qr(mps[y].reshape(a, v * b).T)
        mps[y] = Q.T.reshape(Q.shape[1], v, b)
        mps[y - 1] = np.einsum("avb,kb->avk", mps[y - 1], R)
    for y in range(len(mps) - 1):
        a, v, b = mps[y].shape
        U, s, Vt = np.linalg.svd(mps[y].reshape(a * v, b), full_matrices=False)
        k = min(chi, s.size)
        if k < s.size and s[0] > 0:
            gaps.append((s[k - 1] - s[k]) / s[0])
        mps[y] = U[:, :k].reshape(a, v, k)
        mps[y + 1] = np.einsum("k,kb,bvc->kvc", s[:k], Vt[:k], mps[y + 1])
    return mps


def _bmps_normalise(mps):
    lg = 0.0
    out = []
    for T in mps:
        m = np.abs(T).max()
        if m == 0.0:
            return [T for T in mps], -np.inf
        out.append(T / m)
        lg += np.log(m)
    return out, lg


def _bmps_boundaries(shape, slices, chi, down, rows, gaps):
    """(MPS, log scale) after each of `rows` absorbed in order, compressed to
    chi after every absorbed row whose bonds exceed chi (chi <= 0: exact)."""
    Ly = shape.Ly
    mps, lg = _bmps_normalise(_bmps_init([slices[rows[0] * Ly + y] for y in range(Ly)], down))
    out = [(mps, lg)]
    for xx in rows[1:]:
        mps = _bmps_absorb(mps, [slices[xx * Ly + y] for y in range(Ly)], down)
        if chi > 0 and max(T.shape[2] for T in mps) > chi:
            mps = _bmps_compress(mps, chi, gaps)
        mps, l2 = _bmps_normalise(mps)
        lg += l2
        out.append((mps, lg))
    return out


def bmps_side(shape, slices, chi, down, gaps=None):
    """(MPS, log scale) of one side: rows 0 .. Lx/2-1 (down=True) or
    Lx-1 .. Lx/2 (down=False)."""
    Lx = shape.Lx
    h = Lx // 2
    rows = list(range(h)) if down else list(range(Lx - 1, h - 1, -1))
    return _bmps_boundaries(shape, slices, chi, down, rows, [] if gaps is None else gaps)[-1]


def bmps_contract(shape, slices, chi, gaps=None):
    """(sign, ln|value|) of the sliced network by the two boundaries."""
    top, lt = bmps_side(shape, slices, chi, True, gaps)
    bot, lb = bmps_side(shape, slices, chi, False, gaps)
    E = np.ones((1, 1))
    for T, Bo in zip(top, bot):
        E = np.einsum("ab,avc,bvd->cd", E, T, Bo)
    v = float(E[0, 0])
    if v == 0.0 or not np.isfinite(lt + lb):
        return 0.0, -np.inf
    return float(np.sign(v)), float(np.log(abs(v)) + lt + lb)


def bmps_log_amplitude(shape, theta, x, chi, gaps=None):
    """(sign, ln|psi|) of fPEPS_Model_GPU.log_amplitude with max_bond=chi
    (pureTNS.py:114-146) in this module's sign convention."""
    tensors = shape.unpack(theta)
    s, l = bmps_contract(shape, signed_slices(shape, tensors, x), chi, gaps)
    return s * global_sign(x), l


def bmps_log_amplitudes(shape, theta, xs, chi):
    """Batch form; also returns each configuration's smallest truncation gap
    (inf when nothing was truncated)."""
    out = []
    for x in np.atleast_2d(xs):
        g = []
        s, l = bmps_log_amplitude(shape, theta, x, chi, g)
        out.append((s, l, min(g) if g else np.inf))
    a = np.array(out)
    return a[:, 0], a[:, 1], a[:, 2]


def bmps_strip_derivatives(top, row, bot):
    """Z = <top| row |bot> of one row of slices [l, u, d, r] between the
    boundary MPSs top [a, u, b] and bot [a', d, b'], and dZ/dS_y for every
    site y, from left and right strip environments E[a, l, a']."""
    Ly = len(row)
    left = [np.ones((1, 1, 1))]
    for y in range(Ly):
        left.append(np.einsum("xly,xub,ludr,ydc->brc", left[-1], top[y], row[y], bot[y],
                              optimize=True))
    right = [np.ones((1, 1, 1))]
    for y in range(Ly - 1, -1, -1):
        right.append(np.einsum("xub,ludr,ydc,brc->xly", top[y], row[y], bot[y], right[-1],
                               optimize=True))
    right = right[::-1]
    dS = [np.einsum("xly,xub,ydc,brc->ludr", left[y], top[y], bot[y], right[y + 1],
                    optimize=True) for y in range(Ly)]
    return float(left[-1].reshape(-1)[0]), dS


def bmps_cheap_log_derivative(shape, theta, x, chi, gaps=None):
    """d ln|psi| / d theta with the boundary environments held fixed, the
    reference's cheap gradient (GPU/models/pureTNS.py:2230-2334): for every
    row r the derivative of <env above r| row r |env below r> with respect to
    the row's tensors (environments: the chi-truncated boundaries of the rows
    above and below, quimb's compute_x_environments), summed over rows and
    divided by psi of the two half boundaries (the amplitude).  Exact
    (chi <= 0 or no truncation): the exact log-derivative.  `gaps` collects
    the relative gaps of every truncation (bmps_contract's)."""
    Lx, Ly, D = shape.Lx, shape.Ly, shape.D
    x = np.asarray(x)
    tensors = shape.unpack(theta)
    S = signed_slices(shape, tensors, x)
    gaps = [] if gaps is None else gaps
    s_psi, l_psi = bmps_contract(shape, S, chi, gaps)
    g = np.zeros(shape.n_params)
    if s_psi == 0.0:
        return g, 0.0, -np.inf
    tops = [None] + _bmps_boundaries(shape, S, chi, True, list(range(Lx - 1)), gaps)
    bots = _bmps_boundaries(shape, S, chi, False, list(range(Lx - 1, 0, -1)), gaps)[::-1] + [None]
    triv = ([np.ones((1, 1, 1))] * Ly, 0.0)
    g0 = global_sign(x)
    for r in range(Lx):
        (tm, lt), (bm, lb) = tops[r] or triv, bots[r] or triv
        _, dS = bmps_strip_derivatives(tm, [S[r * Ly + y] for y in range(Ly)], bm)
        for y in range(Ly):
            i = r * Ly + y
            unit = np.ones_like(S[i])
            sgn = signed_slices_single(shape, unit, x, i)
            dT = np.zeros(shape.masks[i].shape)
            # d psi / psi: the global sign cancels, 1 / sign(Z) = sign(Z)
            dT[x[i]] = dS[y] * sgn * np.exp(lt + lb - l_psi) * s_psi
            g[shape.offsets[i]:shape.offsets[i + 1]] = dT[shape.masks[i]]
    return g, s_psi * g0, l_psi
