"""The chi-truncated boundary-MPS oracle (oracle.fpeps.bmps_*) pinned without
quimb (absent): the exact limit against the row-major zipper contraction,
the truncation against its dense definition (the boundary state as a full
vector over the open vertical legs, absorbed row by row and projected
left to right onto its chi leading Schmidt vectors at every cut), and the
invariance under gauge transformations of the horizontal bonds, which a
truncation without the canonical form violates."""
import numpy as np
import pytest

from oracle import fpeps as fp


def _slices(sh, th, x):
    return fp.signed_slices(sh, sh.unpack(th), x)


def _even_cfgs(rng, n, N):
    """Configurations with an even number of singly occupied sites (the
    others vanish by Z2 parity)."""
    out = []
    while len(out) < n:
        x = rng.integers(0, 4, size=N)
        if fp.PARITY_P[x].sum() % 2 == 0:
            out.append(x)
    return out


def _dense_row(psi, row, down):
    """Apply one row (slices [l, u, d, r]) to the dense boundary state."""
    L = len(row)
    T = psi[..., None]
    for y, S in enumerate(row):
        T = np.tensordot(T, S, axes=([y, L], [1 if down else 2, 0]))
        T = np.moveaxis(T, -2, y)
    return T[..., 0]


def _dense_truncate(psi, chi):
    L = psi.ndim
    shape = psi.shape
    for c in range(1, L):
        M = psi.reshape(int(np.prod(shape[:c])), -1)
        U, s, _ = np.linalg.svd(M, full_matrices=False)
        k = min(chi, s.size)
        M = U[:, :k] @ (U[:, :k].T @ M)
        psi = M.reshape(shape)
    return psi


def dense_bmps(sh, slices, chi):
    Lx, Ly = sh.Lx, sh.Ly
    h = Lx // 2
    out = []
    for down, rows in ((True, range(h)), (False, range(Lx - 1, h - 1, -1))):
        psi = np.ones([1] * Ly)
        for k, xx in enumerate(rows):
            psi = _dense_row(psi, [slices[xx * Ly + y] for y in range(Ly)], down)
            if k > 0 and chi > 0:
                psi = _dense_truncate(psi, chi)
        out.append(psi)
    return float(np.sum(out[0] * out[1]))


@pytest.mark.parametrize("Lx,Ly,D", [(3, 3, 2), (4, 4, 2), (5, 3, 2), (6, 3, 2), (4, 3, 3),
                                     (2, 4, 4), (6, 2, 3)])
def test_exact_limit_matches_zipper_contraction(Lx, Ly, D):
    rng = np.random.default_rng(Lx * 100 + Ly * 10 + D)
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, rng)
    xs = rng.integers(0, 4, size=(12, sh.N))
    want = np.array([fp.amplitude(sh, th, x) for x in xs])
    scale = np.abs(want).max()
    for chi in (0, D ** Ly):
        s, l = np.array([fp.bmps_log_amplitude(sh, th, x, chi) for x in xs]).T
        got = s * np.exp(l)
        # exact zeros (parity) come back as rounding residue of the QR / SVD
        assert np.allclose(got, want, rtol=1e-10, atol=1e-12 * scale)
    assert np.count_nonzero(want) >= 3


@pytest.mark.parametrize("Lx,Ly,D,chi", [(4, 4, 2, 2), (4, 5, 2, 3), (6, 4, 2, 3), (4, 4, 3, 4),
                                         (5, 4, 2, 2), (6, 5, 2, 3)])
def test_truncation_matches_dense_definition(Lx, Ly, D, chi):
    rng = np.random.default_rng(Lx * 1000 + Ly * 100 + D * 10 + chi)
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, rng)
    n_trunc = 0
    for x in _even_cfgs(rng, 8, sh.N):
        S = _slices(sh, th, x)
        g = []
        s, l = fp.bmps_contract(sh, S, chi, g)
        want = dense_bmps(sh, S, chi)
        assert s * np.exp(l) == pytest.approx(want, rel=1e-9, abs=1e-300)
        exact = dense_bmps(sh, S, 0)
        n_trunc += bool(g) and abs(want - exact) > 1e-6 * abs(exact)
    assert n_trunc >= 4          # the truncation changes the value


def test_dense_check_rejects_truncation_without_canonical_form(monkeypatch):
    """A plausible mistake (SVD truncation of the raw, non-canonical MPS)
    fails the dense definition."""
    rng = np.random.default_rng(5)
    sh = fp.PEPSShape(4, 5, 2)
    th = fp.random_theta(sh, rng)
    xs = _even_cfgs(rng, 6, sh.N)
    want = [dense_bmps(sh, _slices(sh, th, x), 3) for x in xs]
    good = fp._bmps_compress
    monkeypatch.setattr(fp, "_bmps_compress", lambda m, c, g: good(m, c, g, canonize=False))
    bad = 0
    for x, w in zip(xs, want):
        s, l = fp.bmps_contract(sh, _slices(sh, th, x), 3)
        bad += abs(s * np.exp(l) - w) > 1e-6 * abs(w)
    assert bad >= 3


def test_truncated_value_invariant_under_horizontal_gauge():
    """G, G^-1 inserted on every horizontal bond leave the boundary states,
    and so the truncated value, unchanged."""
    rng = np.random.default_rng(8)
    sh = fp.PEPSShape(6, 4, 2)
    th = fp.random_theta(sh, rng)
    for x in _even_cfgs(rng, 4, sh.N):
        S = _slices(sh, th, x)
        Sg = [t.copy() for t in S]
        for xx in range(sh.Lx):
            for y in range(sh.Ly - 1):
                G = rng.normal(size=(sh.D, sh.D)) + 2.0 * np.eye(sh.D)
                i = xx * sh.Ly + y
                Sg[i] = np.einsum("ludr,rk->ludk", Sg[i], G)
                Sg[i + 1] = np.einsum("kl,ludr->kudr", np.linalg.inv(G), Sg[i + 1])
        a = fp.bmps_contract(sh, S, 3)
        b = fp.bmps_contract(sh, Sg, 3)
        assert a[0] == b[0]
        assert a[1] == pytest.approx(b[1], abs=1e-9)


def test_batch_gap_report():
    rng = np.random.default_rng(2)
    sh = fp.PEPSShape(4, 4, 2)
    th = fp.random_theta(sh, rng)
    xs = rng.integers(0, 4, size=(5, sh.N))
    s, l, g = fp.bmps_log_amplitudes(sh, th, xs, 2)
    assert s.shape == l.shape == g.shape == (5,)
    s0, l0, g0 = fp.bmps_log_amplitudes(sh, th, xs, 0)
    assert np.all(np.isinf(g0))


def _strip_einsum(top, row, bot, skip=None):
    """<top| row |bot> (skip=None) or its derivative with respect to row[skip]
    by one einsum over the whole strip."""
    Ly = len(row)
    L = iter("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz")
    ht, hs, hb = ([next(L) for _ in range(Ly + 1)] for _ in range(3))
    vu, vd = ([next(L) for _ in range(Ly)] for _ in range(2))
    terms, ops = [], []
    for y in range(Ly):
        terms += [ht[y] + vu[y] + ht[y + 1], hs[y] + vu[y] + vd[y] + hs[y + 1],
                  hb[y] + vd[y] + hb[y + 1]]
        ops += [top[y], row[y], bot[y]]
    for e in (ht[0], ht[Ly], hs[0], hs[Ly], hb[0], hb[Ly]):
        terms.append(e)
        ops.append(np.ones(1))
    out = ""
    if skip is not None:
        k = 3 * skip + 1
        out = terms[k]
        terms, ops = terms[:k] + terms[k + 1:], ops[:k] + ops[k + 1:]
    return np.einsum(",".join(terms) + "->" + out, *ops)


@pytest.mark.parametrize("Ly,D,chi", [(3, 2, 2), (2, 3, 3), (3, 2, 1)])
def test_strip_derivatives_match_whole_strip_contraction(Ly, D, chi):
    """The strip environments against one einsum per site, and Euler's
    identity for a function linear in each site: sum_y <dZ/dS_y, S_y> = Ly Z."""
    rng = np.random.default_rng(Ly * 10 + D)
    top = [rng.normal(size=(1 if y == 0 else chi, D, 1 if y == Ly - 1 else chi))
           for y in range(Ly)]
    bot = [rng.normal(size=(1 if y == 0 else chi, D, 1 if y == Ly - 1 else chi))
           for y in range(Ly)]
    row = [rng.normal(size=(1 if y == 0 else D, D, D, 1 if y == Ly - 1 else D))
           for y in range(Ly)]
    Z, dS = fp.bmps_strip_derivatives(top, row, bot)
    assert Z == pytest.approx(float(_strip_einsum(top, row, bot)), rel=1e-12)
    for y in range(Ly):
        assert np.allclose(dS[y], _strip_einsum(top, row, bot, y), rtol=1e-12, atol=1e-14)
    assert sum(np.sum(d * s) for d, s in zip(dS, row)) == pytest.approx(Ly * Z, rel=1e-12)


@pytest.mark.parametrize("Lx,Ly,D,chi", [(4, 3, 2, 0), (5, 3, 2, 0), (3, 3, 3, 0), (6, 3, 2, 8),
                                         (2, 3, 2, 0)])
def test_cheap_gradient_exact_limit(Lx, Ly, D, chi):
    """Without truncation the environments are exact, so the reference's
    cheap gradient is the exact log-derivative (zipper environments)."""
    rng = np.random.default_rng(Lx + 10 * Ly + D)
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, rng)
    for x in _even_cfgs(rng, 3, sh.N):
        g, s, l = fp.bmps_cheap_log_derivative(sh, th, x, chi)
        ge = fp.log_derivative(sh, th, x)
        assert np.abs(g - ge).max() < 1e-12 * np.abs(ge).max()
        assert s * np.exp(l) == pytest.approx(fp.amplitude(sh, th, x), rel=1e-12)
