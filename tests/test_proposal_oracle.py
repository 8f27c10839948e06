"""The learned-proposal oracle pinned on the CPU: the move list is the
Hamiltonian's edge order, the analytic backward pass matches finite
differences, and the Metropolis-Hastings step leaves |psi|^2 exactly
stationary (transition matrix of a 2x2 system, any net), which a dropped
q(m|x')/q(m|x) factor breaks."""
import itertools

import numpy as np
import pytest

from oracle import hubbard as ohub
from oracle import proposal as op
from oracle import slater as osl


def test_edges_are_the_hamiltonian_order():
    for Lx, Ly in [(4, 4), (3, 5), (8, 8), (2, 1)]:
        want = [(a[0] * Ly + a[1], b[0] * Ly + b[1]) for a, b in ohub.edges_2d_square(Lx, Ly)]
        assert op.edges(Lx, Ly) == want


def test_backward_finite_differences():
    rng = np.random.default_rng(0)
    C, Lx, Ly = 4, 3, 3
    th = op.init_params(rng, C)
    th[op.n_params(C) - 3:] = rng.normal(size=3)            # nonzero head bias
    x = np.array([1, 2, 0, 3, 0, 2, 1, 0, 2])
    logits, valid, _ = op.forward(th, C, Lx, Ly, x)
    dl = rng.normal(size=logits.size)
    g = op.backward(th, C, Lx, Ly, x, dl)
    h = 1e-6
    for k in rng.choice(th.size, 40, replace=False):
        e = np.zeros(th.size)
        e[k] = h
        fp = op.forward(th + e, C, Lx, Ly, x)[0] @ dl
        fm = op.forward(th - e, C, Lx, Ly, x)[0] @ dl
        assert g[k] == pytest.approx((fp - fm) / (2 * h), rel=1e-5, abs=1e-7)


def _states_2x2():
    out = []
    for u in range(4):
        for d in range(4):
            x = np.zeros(4, np.int64)
            x[u] += 2
            x[d] += 1
            out.append(x)
    return out


def _transition_matrix(th, C, logamp, drop_q_ratio=False):
    S = _states_2x2()
    idx = {tuple(x): k for k, x in enumerate(S)}
    P = np.zeros((len(S), len(S)))
    for a, x in enumerate(S):
        lf, vf, _ = op.forward(th, C, 2, 2, x)
        lq = op.log_q(lf, vf)
        for m in np.nonzero(vf)[0]:
            xp = op.apply_move(x, 2, 2, m)
            lr, vr, _ = op.forward(th, C, 2, 2, xp)
            lqr = op.log_q(lr, vr)
            corr = 0.0 if drop_q_ratio else lqr[m] - lq[m]
            acc = min(1.0, np.exp(2 * (logamp(xp) - logamp(x)) + corr))
            P[a, idx[tuple(xp)]] += np.exp(lq[m]) * acc
        P[a, a] += 1.0 - P[a].sum()
    return S, P


def test_mh_step_leaves_psi_squared_stationary():
    rng = np.random.default_rng(1)
    M = rng.normal(size=(8, 2))
    C = 4
    th = op.init_params(rng, C, scale=3.0)

    def logamp(x):
        return osl.slogdet(M, np.asarray(x)[None])[1][0]
    S, P = _transition_matrix(th, C, logamp)
    pi = np.array([np.exp(2 * logamp(x)) for x in S])
    pi /= pi.sum()
    assert np.abs(pi @ P - pi).max() < 1e-12
    # the planted mistake: without q(m|x')/q(m|x) the chain is biased
    _, Pw = _transition_matrix(th, C, logamp, drop_q_ratio=True)
    assert np.abs(pi @ Pw - pi).max() > 1e-3


def test_mh_step_matches_transition_rule():
    rng = np.random.default_rng(2)
    M = rng.normal(size=(8, 2))
    C = 4
    th = op.init_params(rng, C)

    def logamp(x):
        return osl.slogdet(M, np.asarray(x)[None])[1][0]
    x = np.array([2, 1, 0, 0])
    lf, vf, _ = op.forward(th, C, 2, 2, x)
    lq = op.log_q(lf, vf)
    # u_c just above the first valid move's mass picks the second valid move
    valid = np.nonzero(vf)[0]
    u = np.exp(lq[valid[0]]) + 1e-9
    xp, acc, lqf, lqr = op.mh_step(th, C, 2, 2, x, logamp, u, 0.0)
    assert lqf == pytest.approx(lq[valid[1]])
    assert acc and np.array_equal(xp, op.apply_move(x, 2, 2, valid[1]))


def test_draws_are_uniform_and_keyed():
    uc, ua = op.lp_draws(7, 3, np.arange(20000))
    assert 0.0 <= uc.min() and uc.max() < 1.0 and abs(uc.mean() - 0.5) < 0.01
    uc2, _ = op.lp_draws(7, 4, np.arange(5))
    assert not np.array_equal(uc[:5], uc2)


def test_bf16_round_matches_torch():
    """The oracle's bfloat16 rounding equals torch's float32 -> bfloat16
    conversion (round to nearest, ties to even), including the tie cases."""
    import torch
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(0, 10, 20000), rng.normal(0, 1e-3, 2000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, -3.5, 65504.0]]).astype(np.float32)
    # exact ties: float32 values whose low 16 bits are 0x8000
    ties = (rng.integers(0x3F800000 >> 16, 0x47000000 >> 16, 500, dtype=np.uint32) << 16) | 0x8000
    v = np.concatenate([v, ties.astype(np.uint32).view(np.float32)])
    want = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(op.bf16_round(v), want)


def test_bf16_forward_is_the_rounded_net():
    """bf16=True changes the logits only by the bfloat16 rounding of the
    activations and of W2, W3 (relative size ~2^-8), and every activation it
    keeps is a bfloat16 value."""
    rng = np.random.default_rng(4)
    C, Lx, Ly = 32, 4, 4
    th = op.init_params(rng, C)
    x = np.array([1, 2, 0, 3, 0, 2, 1, 0, 2, 1, 2, 1, 0, 3, 1, 2])
    l32, v32, _ = op.forward(th, C, Lx, Ly, x)
    l16, v16, c16 = op.forward(th, C, Lx, Ly, x, bf16=True)
    assert np.array_equal(v32, v16)
    d = np.abs(l16[v16] - l32[v16]).max()
    assert 0.0 < d < 2e-2 * max(1.0, np.abs(l32[v32]).max())
    for h in c16[2], c16[4], c16[6]:            # h1, h2, h3 are bfloat16 values
        assert np.array_equal(op.bf16_round(h), h)


def _transition_matrix_two_moves(th, C, logamp, mistake=None):
    """Two chained moves per proposal with the path acceptance
    |psi(x2)|^2 q(m2|x2) q(m1|x1) / (|psi(x)|^2 q(m1|x) q(m2|x1)) (the moves
    are their own reverses).  mistake="reverse_at_x2": both reverse terms
    from the net at x2; "no_q": the bare |psi|^2 ratio."""
    S = _states_2x2()
    idx = {tuple(x): k for k, x in enumerate(S)}
    P = np.zeros((len(S), len(S)))

    def lq_at(x):
        lf, vf, _ = op.forward(th, C, 2, 2, x)
        return op.log_q(lf, vf), vf
    for a, x in enumerate(S):
        lq0, v0 = lq_at(x)
        for m1 in np.nonzero(v0)[0]:
            x1 = op.apply_move(x, 2, 2, m1)
            lq1, v1 = lq_at(x1)
            for m2 in np.nonzero(v1)[0]:
                x2 = op.apply_move(x1, 2, 2, m2)
                lq2, _ = lq_at(x2)
                if mistake == "no_q":
                    corr = 0.0
                elif mistake == "reverse_at_x2":
                    corr = lq2[m2] + lq2[m1] - lq0[m1] - lq1[m2]
                else:
                    corr = lq2[m2] + lq1[m1] - lq0[m1] - lq1[m2]
                acc = min(1.0, np.exp(2 * (logamp(x2) - logamp(x)) + corr))
                P[a, idx[tuple(x2)]] += np.exp(lq0[m1] + lq1[m2]) * acc
        P[a, a] += 1.0 - P[a].sum()
    return S, P


def test_two_move_path_acceptance_leaves_psi_squared_stationary():
    """The acceptance of the GPU sampler's moves_per_proposal = 2 leaves
    |psi|^2 exactly stationary; the planted mistakes (reverse probabilities
    at the wrong configuration, or none) do not."""
    rng = np.random.default_rng(3)
    M = rng.normal(size=(8, 2))
    C = 4
    th = op.init_params(rng, C, scale=3.0)

    def logamp(x):
        return osl.slogdet(M, np.asarray(x)[None])[1][0]
    S, P = _transition_matrix_two_moves(th, C, logamp)
    pi = np.array([np.exp(2 * logamp(x)) for x in S])
    pi /= pi.sum()
    assert np.abs(P.sum(1) - 1).max() < 1e-12
    assert np.abs(pi @ P - pi).max() < 1e-12
    for mistake in ("reverse_at_x2", "no_q"):
        _, Pw = _transition_matrix_two_moves(th, C, logamp, mistake)
        assert np.abs(pi @ Pw - pi).max() > 1e-4
