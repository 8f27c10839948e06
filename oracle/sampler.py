"""Oracle: batched exchange/hop Metropolis sweep with full-determinant
amplitudes.  Test infrastructure only (see oracle/__init__.py).

Follows GPU/sampler.py:47-137 (propose_exchange_or_hopping_vec) and
GPU/sampler.py:216-388 (MetropolisExchangeSpinfulSamplerGPU.step, log-amp
form), with the uniforms supplied explicitly so that the same draws can be
fed to the HIP sampler.
"""
import numpy as np

from . import philox

NMAP = np.array([0, 1, 1, 2])


def propose(col_i, col_j, u_choice, bit_d0, bit_d2, hopping_rate=0.25):
    """GPU/sampler.py:47-137 for a batch of walkers at one edge.

    Returns (new_i, new_j, changed)."""
    col_i = np.asarray(col_i)
    col_j = np.asarray(col_j)
    diff = col_i != col_j                                   # :75
    is_ex = (u_choice < (1 - hopping_rate)) & diff          # :81
    is_hop = (~is_ex) & diff                                # :82
    dn = np.abs(NMAP[col_i] - NMAP[col_j])                  # :89-91
    new_i, new_j = col_i.copy(), col_j.copy()
    swap = is_ex | (is_hop & (dn == 1))                     # :94
    new_i[swap], new_j[swap] = col_j[swap], col_i[swap]
    d0 = is_hop & (dn == 0)                                 # :103
    b0 = np.asarray(bit_d0).astype(bool)
    new_i[d0] = np.where(b0, 3, 0)[d0]                      # :111-115
    new_j[d0] = np.where(b0, 0, 3)[d0]
    d2 = is_hop & (dn == 2)                                 # :119
    b2 = np.asarray(bit_d2).astype(bool)
    new_i[d2] = np.where(b2, 2, 1)[d2]                      # :127-131
    new_j[d2] = np.where(b2, 1, 2)[d2]
    return new_i, new_j, diff


class InjectedDraws:
    """Draw arrays of shape (n_sweeps, n_edges, B)."""

    def __init__(self, u_choice, bits, u_acc):
        self.u_choice, self.bits, self.u_acc = u_choice, bits, u_acc

    def __call__(self, sweep, edge, walkers):
        b = self.bits[sweep, edge, walkers]
        return (self.u_choice[sweep, edge, walkers], b & 1, (b >> 1) & 1,
                self.u_acc[sweep, edge, walkers])


class PhiloxDraws:
    """Philox draws keyed by (seed, global sweep index, edge, walker)."""

    def __init__(self, seed, sweep_offset=0):
        self.seed, self.sweep_offset = seed, sweep_offset

    def __call__(self, sweep, edge, walkers):
        return philox.proposal_draws(self.seed, self.sweep_offset + sweep, edge, walkers)


def sweep(logamp_fn, fxs, edges, draws, n_sweeps=1, hopping_rate=0.25, walkers=None):
    """n_sweeps Metropolis sweeps over `edges` for all walkers.

    logamp_fn(configs) -> (sign, log|psi|) evaluates full amplitudes (the
    reference's model.forward_log).  Acceptance: u_acc < exp(2 (l' - l))
    (GPU/sampler.py:343-353).  Returns (fxs, sign, logabs, n_accept).
    """
    fxs = np.array(fxs, dtype=np.int64, copy=True)
    B = fxs.shape[0]
    walkers = np.arange(B) if walkers is None else np.asarray(walkers)
    s, l = logamp_fn(fxs)
    s, l = np.array(s, dtype=np.float64), np.array(l, dtype=np.float64)
    n_acc = np.zeros(B, dtype=np.int64)
    for sw in range(n_sweeps):
        for e, (i, j) in enumerate(edges):
            u, b0, b2, ua = draws(sw, e, walkers)
            ni, nj, flag = propose(fxs[:, i], fxs[:, j], u, b0, b2, hopping_rate)
            if not flag.any():                              # :284-286
                continue
            prop = fxs[flag].copy()
            prop[:, i], prop[:, j] = ni[flag], nj[flag]
            ps, pl = logamp_fn(prop)
            ratio = np.ones(B)
            ratio[flag] = np.exp(2.0 * (pl - l[flag]))
            acc = flag & (ua < ratio)
            if acc.any():
                sel = np.nonzero(flag)[0]
                fa = acc[flag]
                fxs[sel[fa], i], fxs[sel[fa], j] = ni[sel[fa]], nj[sel[fa]]
                s[sel[fa]], l[sel[fa]] = ps[fa], pl[fa]
                n_acc[sel[fa]] += 1
    return fxs, s, l, n_acc


def uniform_ratio_margin(logamp_fn, fxs, edges, draws, hopping_rate=0.25, walkers=None):
    """Smallest |u_acc - ratio| / ratio over the proposals of one sweep: a
    decision closer than ~1e-12 to its threshold is a floating-point tie
    between two exact evaluations of the same ratio."""
    fxs = np.array(fxs, dtype=np.int64, copy=True)
    B = fxs.shape[0]
    walkers = np.arange(B) if walkers is None else np.asarray(walkers)
    s, l = logamp_fn(fxs)
    margin = np.inf
    for e, (i, j) in enumerate(edges):
        u, b0, b2, ua = draws(0, e, walkers)
        ni, nj, flag = propose(fxs[:, i], fxs[:, j], u, b0, b2, hopping_rate)
        if not flag.any():
            continue
        prop = fxs[flag].copy()
        prop[:, i], prop[:, j] = ni[flag], nj[flag]
        ps, pl = logamp_fn(prop)
        r = np.exp(2.0 * (pl - l[flag]))
        margin = min(margin, float(np.min(np.abs(ua[flag] - r) / np.maximum(r, 1e-300))))
        acc = ua[flag] < r
        sel = np.nonzero(flag)[0][acc]
        fxs[sel, i], fxs[sel, j] = ni[sel], nj[sel]
        s[sel], l[sel] = ps[acc], pl[acc]
    return margin
