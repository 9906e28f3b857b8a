"""Host twin of the kernels' counter-based draws (oracle; test infrastructure).

Philox4x32-10 (Salmon et al., SC'11) in vectorised NumPy uint32/uint64
arithmetic, with the counter layout, uniform conversion, Box-Muller pairing
and Poisson inversion that sdepy_amd/csrc/philox.hpp uses:

    counter = (block, path, stream | sub << 8, epoch), key = (k0, k1)
    normal g of a path's stream -> block g // 4, lane g % 4
    u = (x + 0.5) * 2**-32;  z0, z1 = r cos(2 pi u_b), r sin(2 pi u_b),
    r = sqrt(-2 ln u_a) on the lane pairs (0, 1) and (2, 3).

These replace the NumPy draws of sdepy's sources (``rng.normal``
infrastructure.py:1515, ``rng.multivariate_normal`` :1532, ``rng.poisson``
:1631, ``scipy.stats.norm.rvs`` :1718); value-level parity with NumPy's own
streams is impossible by design, so the draws are pinned on their own:

Pinning
-------
* ``philox4x32_10``: the three Random123 known-answer vectors (counter/key
  all-zero, all-ones, and the pi-digits vector) and, where the script in
  tests/golden has been run, rocRAND's and ATen's Philox source headers
  compiled on the host.
* ``box_muller``: closed-form values at chosen uniforms (u_a = e^-1/2 gives
  r = 1, u_b = 0, 1/4, 1/2 give the axes) and the first two moments.
* ``poisson_inv``: the Poisson pmf (chi-square goodness of fit) and the
  inversion identity P(n <= k) = F(k) at the CDF breakpoints.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

STREAM_NORMAL = 0
STREAM_POISSON = 1
STREAM_JUMP = 2
STREAM_BACKWARD = 16

TWO_PI = 6.283185307179586
POISSON_CAP = 4096


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 of broadcastable uint32 counter words and a scalar key.

    Returns the four uint32 output words.
    """
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint32) for c in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over='ignore'):
        for r in range(10):
            if r:
                k0 = np.uint32((int(k0) + int(W0)) & 0xFFFFFFFF)
                k1 = np.uint32((int(k1) + int(W1)) & 0xFFFFFFFF)
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n1 = (p1 & MASK).astype(np.uint32)
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            n3 = (p0 & MASK).astype(np.uint32)
            c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def u01(x):
    """(x + 0.5) * 2**-32 in float64, exact, in (0, 1)."""
    return (np.asarray(x, dtype=np.float64) + 0.5) * 2.0 ** -32


def box_muller(a, b):
    """Normal pair from two uint32 words (float64)."""
    r = np.sqrt(-2.0 * np.log(u01(a)))
    ang = TWO_PI * u01(b)
    return r * np.cos(ang), r * np.sin(ang)


def block_normals(blocks, paths, c2, epoch, key):
    """Four normals per (block, path): array shaped blocks.shape + (4,)."""
    x, y, z, w = philox4x32_10(blocks, paths, c2, epoch, key[0], key[1])
    n0, n1 = box_muller(x, y)
    n2, n3 = box_muller(z, w)
    return np.stack((n0, n1, n2, n3), axis=-1)


def stream_normals(key, epoch, c2, nslots, nvars, paths):
    """Normals of slots 0..nslots-1, shaped (nslots, nvars, len(paths)).

    Normal (s, v) of a path is number g = s*nvars + v of its stream.
    """
    paths = np.asarray(paths, dtype=np.uint32)
    n = nslots * nvars
    nb = (n + 3) // 4
    blocks = np.arange(nb, dtype=np.uint32)[:, None]
    z = block_normals(blocks, paths[None, :], c2, epoch, key)   # (nb, P, 4)
    z = z.transpose(0, 2, 1).reshape(nb * 4, paths.size)[:n]
    return z.reshape(nslots, nvars, paths.size)


def stream_words(key, epoch, c2, nslots, nvars, paths):
    """Raw uint32 draws of slots 0..nslots-1 (Poisson uniforms)."""
    paths = np.asarray(paths, dtype=np.uint32)
    n = nslots * nvars
    nb = (n + 3) // 4
    blocks = np.arange(nb, dtype=np.uint32)[:, None]
    w = np.stack(philox4x32_10(blocks, paths[None, :], c2, epoch,
                               key[0], key[1]), axis=-1)       # (nb, P, 4)
    w = w.transpose(0, 2, 1).reshape(nb * 4, paths.size)[:n]
    return w.reshape(nslots, nvars, paths.size)


def _slot_blocks(s, nvars):
    g0 = s * nvars
    b0, b1 = g0 // 4, (g0 + nvars - 1) // 4
    return g0, b0, np.arange(b0, b1 + 1, dtype=np.uint32)


def slot_normals(key, epoch, c2, s, nvars, paths):
    """Normals (s, 0..nvars-1) of each path: shape (nvars, len(paths))."""
    paths = np.asarray(paths, dtype=np.uint32)
    g0, b0, blocks = _slot_blocks(s, nvars)
    z = block_normals(blocks[:, None], paths[None, :], c2, epoch, key)
    z = z.transpose(0, 2, 1).reshape(blocks.size * 4, paths.size)
    return z[g0 - 4 * b0:g0 - 4 * b0 + nvars]


def slot_words(key, epoch, c2, s, nvars, paths):
    """Raw uint32 draws (s, 0..nvars-1) of each path: (nvars, len(paths))."""
    paths = np.asarray(paths, dtype=np.uint32)
    g0, b0, blocks = _slot_blocks(s, nvars)
    w = np.stack(philox4x32_10(blocks[:, None], paths[None, :], c2, epoch,
                               key[0], key[1]), axis=-1)
    w = w.transpose(0, 2, 1).reshape(blocks.size * 4, paths.size)
    return w[g0 - 4 * b0:g0 - 4 * b0 + nvars]


def poisson_inv(u, lam, p0, cap=POISSON_CAP):
    """Sequential CDF inversion, vectorised; float64 IEEE operations in the
    same order as the kernel, so counts agree bit for bit."""
    u, lam, p0 = np.broadcast_arrays(np.asarray(u, float),
                                     np.asarray(lam, float),
                                     np.asarray(p0, float))
    n = np.zeros(u.shape, dtype=np.int64)
    p = p0.copy()
    f = p0.copy()
    active = (u > f) & (n < cap)
    while active.any():
        n = np.where(active, n + 1, n)
        p = np.where(active, p * (lam / np.maximum(n, 1).astype(float)), p)
        f = np.where(active, f + p, f)
        active = active & (u > f) & (n < cap)
    return n


# Poisson means above POISSON_SPLIT (poisson_draw in sdepy_amd/csrc/philox.hpp):
# exp(-lam) underflows in FP64 from about 745, so such a count is the sum of
# k = ceil(lam / POISSON_SPLIT) counts of mean lam / k; the first takes the
# slot's own word, part j >= 1 word (j-1) % 4 of block (slot, path,
# c2p | (1 + (j-1) // 4) << 8, epoch).
POISSON_SPLIT = 700.


def poisson_p0(lam):
    """exp(-lam), or exp(-lam / k) for a mean split into k parts."""
    lam = np.asarray(lam, float)
    k = np.maximum(np.ceil(lam / POISSON_SPLIT), 1.)
    return np.where(lam <= POISSON_SPLIT, np.exp(-lam), np.exp(-(lam / k)))


def poisson_draw(words, lam, key, epoch, c2p, slots, paths, p0=None):
    """Poisson counts of means lam from the slots' words (uint32), splitting
    means above POISSON_SPLIT; slots and paths are the counters' c0 and c1
    of each element (broadcast against words)."""
    words, lam, slots, paths = np.broadcast_arrays(
        np.asarray(words, np.uint32), np.asarray(lam, float),
        np.asarray(slots, np.uint64), np.asarray(paths, np.uint64))
    p0 = poisson_p0(lam) if p0 is None else np.broadcast_to(np.asarray(p0, float), lam.shape)
    k = np.maximum(np.ceil(lam / POISSON_SPLIT), 1.)
    part = np.where(lam <= POISSON_SPLIT, lam, lam / k)
    n = poisson_inv(u01(words), part, p0)
    for j in range(1, int(k.max()) if k.size else 1):
        m = k > j
        if not m.any():
            break
        w = philox4x32_10(slots[m].astype(np.uint32), paths[m].astype(np.uint32),
                          c2p | ((1 + (j - 1) // 4) << 8), epoch, key[0], key[1])
        n[m] += poisson_inv(u01(w[(j - 1) & 3]), part[m], p0[m])
    return n


# jump-size laws of the compound Poisson draws (SDEGPU_JUMP_* in sdegpu.h)
JUMP_NORMAL = 0       # p0 + p1 z                  norm_rv(a, b) = (a, b, -)
JUMP_UNIFORM = 1      # p0 + p1 u                  uniform_rv(a, b) = (a, b - a, -)
JUMP_DOUBLE_EXP = 2   # (v <= p2 ? p0 : -p1) E     double_exp_rv(a, b, pa) = (a, b, pa),
#                                                  exp_rv(a) = (a, -, 1)


def jump_total(key, epoch, tag, slot, path, n, law, p0, p1, p2=0.):
    """Sum of the n jump sizes of one (slot, path), in jump order, as the
    kernels draw them (sdepy_amd/csrc/jumps.hpp): jump j reads the block
    (slot, path, tag|2|sub<<8, epoch) of the jump stream; normal and uniform
    laws: sub = j // 4, value j % 4 of the block (Box-Muller pairs for the
    normal law, u01 of the word for the uniform law); double exponential:
    sub = j // 2, E = -ln u01 of word 2 (j % 2), the side from u01 of word
    2 (j % 2) + 1 compared with p2."""
    acc = 0.0
    c2 = tag | STREAM_JUMP
    n = int(n)
    if law == JUMP_DOUBLE_EXP:
        for j in range(0, n, 2):
            w = philox4x32_10(np.uint32(slot), np.uint32(path), c2 | ((j >> 1) << 8),
                              epoch, key[0], key[1])
            for q in range(min(2, n - j)):
                e = -np.log(u01(w[2 * q]))
                side = p0 if u01(w[2 * q + 1]) <= p2 else -p1
                acc = acc + side * e
        return acc
    for j in range(0, n, 4):
        if law == JUMP_NORMAL:
            y = block_normals(np.uint32(slot), np.uint32(path), c2 | ((j >> 2) << 8),
                              epoch, key)
        else:
            y = u01(np.stack(philox4x32_10(np.uint32(slot), np.uint32(path),
                                           c2 | ((j >> 2) << 8), epoch, key[0], key[1])))
        for q in range(min(4, n - j)):
            acc = acc + (p0 + p1 * y[q])
    return acc

