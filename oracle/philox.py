"""Oracle: Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw,
"Parallel random numbers: as easy as 1, 2, 3", SC'11) and the mapping from
one Philox block to the four draws a proposal consumes.

Test infrastructure only (see oracle/__init__.py).  Pinned by the
Random123 known-answer vectors in tests/test_oracle.py.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds; all arguments uint32-valued
    arrays (broadcastable).  Returns four uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3, k0, k1 = (np.asarray(v, dtype=np.uint64) & MASK
                              for v in (c0, c1, c2, c3, k0, k1))
    for r in range(10):
        if r:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


TWO_M53 = 2.0 ** -53


def proposal_draws(seed, sweep, edge, walkers):
    """The four draws of one (walker, sweep, edge) proposal.

    counter = (edge, sweep, walker, 0), key = (seed lo, seed hi).
      u_choice = top 53 bits of (o0:o1) * 2^-53
      bit_d0   = o1 & 1,  bit_d2 = (o1 >> 1) & 1
      u_acc    = top 53 bits of (o2:o3) * 2^-53
    """
    w = np.asarray(walkers, dtype=np.uint64)
    seed = int(seed)
    o0, o1, o2, o3 = philox4x32_10(np.uint64(edge), np.uint64(sweep), w, 0,
                                   np.uint64(seed & 0xFFFFFFFF),
                                   np.uint64((seed >> 32) & 0xFFFFFFFF))
    x = (o0 << np.uint64(32)) | o1
    y = (o2 << np.uint64(32)) | o3
    u_choice = (x >> np.uint64(11)).astype(np.float64) * TWO_M53
    u_acc = (y >> np.uint64(11)).astype(np.float64) * TWO_M53
    bit_d0 = (o1 & np.uint64(1)).astype(np.uint8)
    bit_d2 = ((o1 >> np.uint64(1)) & np.uint64(1)).astype(np.uint8)
    return u_choice, bit_d0, bit_d2, u_acc
