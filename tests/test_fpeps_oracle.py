"""The fPEPS / NN-fTNS oracle pinned without the reference's quimb/symmray
stack (absent here, versions unpinned): the local sign rule against the
Grassmann tensor network evaluated from its definition, Z2 parity
selection, the D = 1 product-state closed form, gradients against finite
differences, eta = 0 reducing NN-fTNS to fPEPS, and the layout shared with
the device library."""
import numpy as np
import pytest

from oracle import fpeps as fp
from vmc_torch_amd import kernels


def _cfgs(rng, n, N):
    return [rng.integers(0, 4, size=N) for _ in range(n)]


@pytest.mark.parametrize("Lx,Ly,D,n", [(2, 2, 4, 30), (2, 3, 2, 30), (3, 2, 2, 30),
                                       (3, 3, 2, 8), (2, 3, 3, 8)])
def test_sign_rule_matches_grassmann_definition(Lx, Ly, D, n):
    rng = np.random.default_rng(10 * Lx + Ly + D)
    sh = fp.PEPSShape(Lx, Ly, D)
    th = fp.random_theta(sh, rng)
    nonzero = 0
    for x in _cfgs(rng, n, sh.N):
        a = fp.amplitude(sh, th, x)
        b = fp.amplitude_grassmann(sh, th, x)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-14)
        nonzero += abs(b) > 1e-12
    assert nonzero >= n // 4


def test_grassmann_check_catches_a_dropped_sign(monkeypatch):
    """A plausible mistake (dropping the q(d) q(r) crossing factor) fails the
    comparison with the definition."""
    rng = np.random.default_rng(3)
    sh = fp.PEPSShape(2, 3, 2)
    th = fp.random_theta(sh, rng)
    xs = _cfgs(rng, 40, sh.N)
    ref = [fp.amplitude_grassmann(sh, th, x) for x in xs]
    orig = fp.signed_slices

    def wrong(shape, tensors, x):
        out = orig(shape, tensors, x)
        # undo the sign for the q(r) part only: use (-1)^{q(d) P_right}
        res = []
        for i, s in enumerate(out):
            r_, c_ = divmod(i, shape.Ly)
            dl, du, dd, dr = fp.site_dims(shape.Lx, shape.Ly, shape.D, r_, c_)
            qd = fp.bond_parity(dd)[None, None, :, None]
            qr = fp.bond_parity(dr)[None, None, None, :]
            res.append(s * np.where((qd * qr) % 2 == 1, -1.0, 1.0))
        return res
    monkeypatch.setattr(fp, "signed_slices", wrong)
    bad = [fp.amplitude(sh, th, x) for x in xs]
    assert max(abs(a - b) for a, b in zip(bad, ref)) > 1e-3


def test_z2_parity_selection():
    rng = np.random.default_rng(4)
    sh = fp.PEPSShape(3, 3, 2)
    th = fp.random_theta(sh, rng)
    for x in _cfgs(rng, 30, sh.N):
        if int(np.sum(fp.PARITY_P[x])) % 2:
            assert fp.amplitude(sh, th, x) == 0.0


def test_product_state_closed_form():
    """D = 1: psi(x) = (-1)^{C(P,2)} prod_s T_s[x_s]."""
    rng = np.random.default_rng(5)
    sh = fp.PEPSShape(3, 4, 1)
    th = fp.random_theta(sh, rng, normalise=False)
    T = sh.unpack(th)
    for x in _cfgs(rng, 20, sh.N):
        want = fp.global_sign(x) * np.prod([T[s].reshape(4)[x[s]] for s in range(sh.N)])
        assert fp.amplitude(sh, th, x) == pytest.approx(want, rel=1e-13, abs=1e-300)


def test_log_derivative_finite_differences():
    rng = np.random.default_rng(6)
    sh = fp.PEPSShape(3, 3, 2)
    th = fp.random_theta(sh, rng)
    x = np.array([1, 2, 3, 0, 1, 2, 0, 3, 0])
    g = fp.log_derivative(sh, th, x)
    psi = fp.amplitude(sh, th, x)
    h = 1e-6
    for k in rng.choice(sh.n_params, 25, replace=False):
        e = np.zeros(sh.n_params)
        e[k] = h
        fd = (fp.amplitude(sh, th + e, x) - fp.amplitude(sh, th - e, x)) / (2 * h) / psi
        assert g[k] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_nnftn_eta_zero_and_gradient():
    rng = np.random.default_rng(7)
    sh = fp.PEPSShape(2, 3, 2)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, 8, sh.n_params, 0.3)
    x = np.array([1, 2, 0, 3, 2, 1])
    assert fp.nnftn_amplitude(sh, th, nn, 0.0, x) == pytest.approx(fp.amplitude(sh, th, x))
    for act in ("relu", "tanh"):
        g = fp.nnftn_log_derivative(sh, th, nn, 0.7, x, act)
        flat = np.concatenate([th] + [p.ravel() for p in nn])
        sizes = [th.size] + [p.size for p in nn]

        def amp(v):
            parts = np.split(v, np.cumsum(sizes)[:-1])
            nnp = [p.reshape(q.shape) for p, q in zip(parts[1:], nn)]
            return fp.nnftn_amplitude(sh, parts[0], nnp, 0.7, x, act)
        psi = amp(flat)
        h = 1e-6
        for k in rng.choice(flat.size, 30, replace=False):
            e = np.zeros(flat.size)
            e[k] = h
            fd = (amp(flat + e) - amp(flat - e)) / (2 * h) / psi
            assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-8)


@pytest.mark.parametrize("mode", ["add", "mul"])
def test_nnftn_constant_output_reduces_to_fpeps(mode):
    """With Wo = 0 the network output is the constant bo, so NN-fTNS is the
    plain fPEPS at theta + eta bo ("add") or theta * (1 + eta bo) ("mul");
    with W1 = 0 as well the output is Wo act(W2 act(b1) + b2) + bo, fixed by
    the layer order.  Each amplitude is compared with the plain contraction
    (pinned against the Grassmann definition) at every basis state of 2x2."""
    rng = np.random.default_rng(12)
    sh = fp.PEPSShape(2, 2, 2)
    th = fp.random_theta(sh, rng)
    W1, b1, W2, b2, Wo, bo = fp.mlp_init(rng, 2 * sh.N, 6, sh.n_params, 0.3)
    b1, b2, bo = rng.normal(size=6), rng.normal(size=6), rng.normal(size=sh.n_params)
    eta = 0.4
    cases = [([W1, b1, W2, b2, np.zeros_like(Wo), bo], bo),
             ([np.zeros_like(W1), b1, W2, b2, Wo, bo],
              Wo @ np.maximum(W2 @ np.maximum(b1, 0.0) + b2, 0.0) + bo)]
    xs = [np.array(c) for c in np.ndindex(4, 4, 4, 4)]
    for nn, out in cases:
        eff = th + eta * out if mode == "add" else th * (1.0 + eta * out)
        for x in xs:
            assert fp.nnftn_amplitude(sh, th, nn, eta, x, "relu", mode) == pytest.approx(
                fp.amplitude(sh, eff, x), rel=1e-12, abs=1e-15)


def test_nnftn_multiplicative_gradient():
    """The multiplicative form's log-derivative against finite differences of
    its amplitude, for both activations."""
    rng = np.random.default_rng(8)
    sh = fp.PEPSShape(2, 3, 2)
    th = fp.random_theta(sh, rng)
    nn = fp.mlp_init(rng, 2 * sh.N, 8, sh.n_params, 0.3)
    nn[1], nn[3], nn[5] = rng.normal(size=8), rng.normal(size=8), rng.normal(size=sh.n_params)
    x = np.array([1, 2, 0, 3, 2, 1])
    for act in ("relu", "tanh"):
        g = fp.nnftn_log_derivative(sh, th, nn, 0.7, x, act, "mul")
        flat = np.concatenate([th] + [p.ravel() for p in nn])
        sizes = [th.size] + [p.size for p in nn]

        def amp(v):
            parts = np.split(v, np.cumsum(sizes)[:-1])
            nnp = [p.reshape(q.shape) for p, q in zip(parts[1:], nn)]
            return fp.nnftn_amplitude(sh, parts[0], nnp, 0.7, x, act, "mul")
        psi = amp(flat)
        h = 1e-6
        for k in rng.choice(flat.size, 40, replace=False):
            e = np.zeros(flat.size)
            e[k] = h
            fd = (amp(flat + e) - amp(flat - e)) / (2 * h) / psi
            assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-8)


@pytest.mark.parametrize("Lx,Ly,D", [(4, 4, 4), (2, 5, 3), (3, 2, 1)])
def test_device_layout_matches_oracle(Lx, Ly, D):
    dmap, sizes = kernels.peps_layout(Lx, Ly, D)
    sh = fp.PEPSShape(Lx, Ly, D)
    assert sizes == sh.sizes
    dense = np.concatenate([m.ravel() for m in sh.masks])
    assert np.array_equal(dmap >= 0, dense)
    assert np.array_equal(dmap[dmap >= 0], np.arange(sh.n_params))


def test_config4_parameter_count():
    """4x4, D = 4: 4 corners x 32 + 8 edges x 128 + 4 bulk x 512 = 3200."""
    assert fp.PEPSShape(4, 4, 4).n_params == 3200
