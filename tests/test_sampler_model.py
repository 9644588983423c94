"""The exact parallel evaluation of qfvm's sampling sum, modelled on the CPU.

measure_samples (statevector.hpp:1380-1412) draws the first index k whose running sum
S_k = fl(S_{k-1} + p_k) exceeds u. The GPU sampler (pyquafu_amd/csrc/sample.hip) evaluates S at
every 32-element block end exactly, in parallel, without replaying the additions one by one:

* while S stays in one binade [2^e, 2^(e+1)), S = m 2^(e-52) for an integer m, and
  fl(S + p) = (m + rne(p / 2^(e-52))) 2^(e-52): an integer step. Its size is floor(p / ulp) plus
  one when the fraction is above 1/2, and, at exactly 1/2 (a tie), plus one only when that makes
  the result even, i.e. depending on the parity of the incoming m;
* so a block's effect is a pair (w0, w1): the increment of m for an even / odd incoming m. Pairs
  compose associatively, (f then g)_p = f_p + g_{(p + f_p) mod 2}, which is what makes a parallel
  scan possible;
* a block whose m could reach 2^53 - 1 (the sum could leave the binade) is replayed sequentially
  from its exact start, and the next chunk starts in the new binade.

This module pins that algorithm (chunking, tiles, saturation, crossings, ties) against the plain
sequential loop on crafted inputs; the GPU tests then check the kernels against the oracle.
"""

import math

import numpy as np
import pytest

W = 32
SAT = 1 << 60


def xelem(p: float, e: int):
    lim = math.ldexp(1.0, e + 1)
    if not p < lim:
        return (SAT, SAT)
    r = math.ldexp(p, 52 - e)
    fl = math.floor(r)
    f = r - fl
    q = int(fl)
    if f > 0.5:
        return (q + 1, q + 1)
    if f < 0.5:
        return (q, q)
    return (q + (q & 1), q + ((1 + q) & 1))


def compose(f, g):
    return (min(f[0] + g[f[0] & 1], SAT), min(f[1] + g[(1 + f[1]) & 1], SAT))


def block_sequential(p, j, s):
    for x in p[j * W:(j + 1) * W]:
        s = s + x
    return s


def boundaries_model(p, tile=4, min_chunk=8, max_chunk=64):
    """Exact S at every block end, computed the way the GPU does (chunks of blocks, tiles of
    blocks, one scan of tile transfers per chunk, crossing blocks replayed sequentially)."""
    nblk = len(p) // W
    B = [0.0] * nblk
    nz = [i for i, x in enumerate(p) if x > 0]
    if not nz:
        return B
    b0 = nz[0] // W
    B[b0] = block_sequential(p, b0, 0.0)
    nxt = b0 + 1
    while nxt < nblk:
        s = B[nxt - 1]
        if s < 2.0 ** -1022:
            B[nxt] = block_sequential(p, nxt, s)
            nxt += 1
            continue
        e = math.frexp(s)[1] - 1
        m0 = int(math.ldexp(s, 52 - e))
        length = min(max(min_chunk, nxt - b0), max_chunk)
        end = min(nblk, nxt + length)
        blocks = []
        for j in range(nxt, end):
            x = (0, 0)
            for v in p[j * W:(j + 1) * W]:
                x = compose(x, xelem(v, e))
            blocks.append(x)
        tiles = []
        for t in range(0, len(blocks), tile):
            x = (0, 0)
            for b in blocks[t:t + tile]:
                x = compose(x, b)
            tiles.append(x)
        prefix, acc = [], (0, 0)
        for t in tiles:
            prefix.append(acc)
            acc = compose(acc, t)
        crossing = None
        for ti, pre in enumerate(prefix):
            m = m0 + pre[m0 & 1]
            for k, b in enumerate(blocks[ti * tile:(ti + 1) * tile]):
                m = m + b[m & 1]
                j = nxt + ti * tile + k
                if m >= (1 << 53) - 1:
                    crossing = j if crossing is None else min(crossing, j)
                    break
                B[j] = math.ldexp(float(m), e - 52)
        if crossing is None:
            nxt = end
        else:
            start = B[crossing - 1] if crossing > nxt else s
            B[crossing] = block_sequential(p, crossing, start)
            nxt = crossing + 1
    return B


def boundaries_sequential(p):
    out, s = [], 0.0
    for k, x in enumerate(p):
        s = s + x
        if k % W == W - 1:
            out.append(s)
    return out


def case(name, rng):
    n = 1 << 12
    if name == "random":
        v = rng.normal(size=n) ** 2 + rng.normal(size=n) ** 2
        return list(v / v.sum())
    if name == "ties":
        # S in [0.5, 1): ulp 2^-53; p = odd * 2^-54 lands exactly half-way
        p = [0.5] + [float(rng.integers(1, 9) * 2 + 1) * 2.0**-54 for _ in range(n - 1)]
        return p
    if name == "ties_mixed":
        p = []
        s = 2.0**-20
        for k in range(n):
            e = math.frexp(s)[1] - 1
            if rng.random() < 0.5:
                x = float(2 * rng.integers(0, 5) + 1) * 2.0 ** (e - 53)
            else:
                x = float(rng.random()) * 2.0 ** (e - 8)
            p.append(x)
            s = s + x
        return p
    if name == "growing":
        return [2.0 ** (-60 + k // 64) * float(rng.random()) for k in range(n)]
    if name == "zeros":
        v = rng.random(n)
        v[: n // 3] = 0.0
        v[rng.random(n) < 0.5] = 0.0
        return list(v)
    if name == "jumps":
        v = [0.0] * n
        for k, x in zip([5, 40, 41, 700, 2000, 4095], [1e-300, 1e-200, 3e-100, 1e-3, 0.5, 7.0]):
            v[k] = x
        for k in range(2100, 2400):
            v[k] = 1e-17 * (k % 7)
        return v
    if name == "large":
        return list(rng.random(n) * 1e10)
    if name == "subnormal_start":
        v = list(rng.random(n) * 1e-300)
        v[0] = 5e-324
        return v
    raise KeyError(name)


@pytest.mark.parametrize("name", ["random", "ties", "ties_mixed", "growing", "zeros", "jumps", "large",
                                  "subnormal_start"])
def test_model_reproduces_sequential_sum(name):
    rng = np.random.Generator(np.random.PCG64(sum(map(ord, name))))
    p = case(name, rng)
    want = boundaries_sequential(p)
    got = boundaries_model(p)
    assert got == want  # bit for bit


def test_ties_depend_on_parity():
    # the same block applied to an even and an odd m gives different increments: a tie rounds to even
    f = xelem(2.0**-54, -1)  # S in [0.5, 1): ulp 2^-53, p = half an ulp
    assert f == (0, 1)
    g = compose(f, f)
    assert g == (0, 1)  # even m stays; odd m: m + 1/2 -> m + 1 (even), then + 1/2 -> stays


M_CROSS = (1 << 53) - 1


def tree(xs):
    """Composition in the order of the kernels' shuffle trees: neighbours first, then pairs of
    pairs (len(xs) a power of two)."""
    while len(xs) > 1:
        xs = [compose(xs[i], xs[i + 1]) for i in range(0, len(xs), 2)]
    return xs[0]


def scan(xs):
    """Inclusive scan in the order of the kernels' wavefront scans (Hillis-Steele: step d composes
    the value d lanes back in front)."""
    xs = list(xs)
    d = 1
    while d < len(xs):
        xs = [compose(xs[i - d], xs[i]) if i >= d else xs[i] for i in range(len(xs))]
        d *= 2
    return xs


def serial(xs):
    acc = (0, 0)
    for x in xs:
        acc = compose(acc, x)
    return acc


def agree(m0, f, g):
    a, b = m0 + f[m0 & 1], m0 + g[m0 & 1]
    return a == b or (a >= M_CROSS and b >= M_CROSS)


@pytest.mark.parametrize("kind", ["small", "ties", "near_cross", "saturated"])
def test_parallel_orders_match_the_serial_fold(kind):
    """The GPU composes element pairs in runs of 8, trees of 4 runs (a block) and of 64 blocks (a
    tile), and scans tiles and blocks with wavefront scans. Saturation at 2^60 makes composition
    associative only up to values past the binade; for an incoming m in [2^52, 2^53) every order
    must give the serial fold's m exactly, or both must reach 2^53 - 1 (the block is replayed)."""
    rng = np.random.Generator(np.random.PCG64(sum(map(ord, kind))))
    e = -1  # S in [0.5, 1), ulp 2^-53
    for _ in range(20):
        n = 2048
        if kind == "small":
            p = rng.random(n) * 2.0**-40
        elif kind == "ties":
            p = (2 * rng.integers(0, 6, n) + 1) * 2.0**-54
        elif kind == "near_cross":
            p = rng.random(n) * 2.0**-11 / n * 8  # sums of a few blocks approach 2^53 ulps
        else:
            p = rng.random(n) * 2.0**-45
            p[rng.integers(0, n, 3)] = 1.5  # beyond the binade: saturated elements
        el = [xelem(float(v), e) for v in p]
        runs = [serial(el[i:i + 8]) for i in range(0, n, 8)]
        blocks = [tree(runs[i:i + 4]) for i in range(0, len(runs), 4)]
        tile = tree(blocks)
        whole = serial(el)
        prefix = scan(blocks)
        ends, acc = [], (0, 0)
        for k, x in enumerate(el):  # the serial fold at every block end
            acc = compose(acc, x)
            if k % W == W - 1:
                ends.append(acc)
        for m0 in (1 << 52, (1 << 52) + 1, (1 << 53) - 2 - int(rng.integers(0, 1 << 40))):
            assert agree(m0, tile, whole)
            for b in range(len(blocks)):
                assert agree(m0, prefix[b], ends[b])
