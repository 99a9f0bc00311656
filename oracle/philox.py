"""Philox4x32-10 counter-based generator in NumPy (test infrastructure only).

The reference draws its uniforms with ``tf.random.uniform`` (src/vegasflow/
monte_carlo.py:264-266), whose stream cannot be reproduced without
TensorFlow. The HIP engine therefore uses its own generator, Philox4x32-10
(Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11), and this module is the bit-exact CPU restatement the kernels are
checked against. It is pinned by the published Random123 known-answer
vectors and by vectors from an independent implementation (see
tests/golden/philox_vectors.json).

Counter layout used by every kernel (a pure function of the draw, never of
the launch geometry):
    key     = (seed & 0xffffffff, seed >> 32)
    counter = (event & 0xffffffff, event >> 32, draw_id, block)
where ``block`` selects which group of four 32-bit words of the event is
produced.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32 with ``rounds`` rounds.

    All arguments broadcast; they are uint32 arrays (or scalars). Returns a
    tuple of four uint32 arrays.
    """
    c0 = np.asarray(c0, dtype=np.uint32)
    c1 = np.asarray(c1, dtype=np.uint32)
    c2 = np.asarray(c2, dtype=np.uint32)
    c3 = np.asarray(c3, dtype=np.uint32)
    k0 = np.asarray(k0, dtype=np.uint32)
    k1 = np.asarray(k1, dtype=np.uint32)
    c0, c1, c2, c3, k0, k1 = np.broadcast_arrays(c0, c1, c2, c3, k0, k1)
    c0 = c0.copy(); c1 = c1.copy(); c2 = c2.copy(); c3 = c3.copy()
    k0 = k0.copy(); k1 = k1.copy()
    with np.errstate(over="ignore"):
        for r in range(rounds):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            n0 = hi1 ^ c1 ^ k0
            n1 = lo1
            n2 = hi0 ^ c3 ^ k1
            n3 = lo0
            c0, c1, c2, c3 = n0, n1, n2, n3
            if r != rounds - 1:
                k0 = k0 + W0
                k1 = k1 + W1
    return c0, c1, c2, c3


def event_words(seed, draw_id, events, n_blocks):
    """Words of ``n_blocks`` Philox blocks for each event index in ``events``.

    Returns uint32 array (len(events), 4 * n_blocks); word 4*b + i is output i
    of block b.
    """
    ev = np.asarray(events, dtype=np.uint64)
    k0 = np.uint32(int(seed) & 0xFFFFFFFF)
    k1 = np.uint32((int(seed) >> 32) & 0xFFFFFFFF)
    lo = (ev & _MASK).astype(np.uint32)
    hi = (ev >> np.uint64(32)).astype(np.uint32)
    out = np.empty((ev.shape[0], 4 * n_blocks), dtype=np.uint32)
    for b in range(n_blocks):
        r = philox4x32_10(lo, hi, np.uint32(draw_id), np.uint32(b), k0, k1)
        for i in range(4):
            out[:, 4 * b + i] = r[i]
    return out
