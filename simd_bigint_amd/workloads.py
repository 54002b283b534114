"""Synthetic operands for tests and benchmarks.

xorshift64 with shifts (13, 7, 17) -- the reference's generator
(bench/nt_sweep.c:15-16, tests/p30_test.c:14-15). The state is a
one-element list so A then B are drawn from one stream as the reference does
(bench/nt_sweep.c:46). Large fills are vectorised exactly: xorshift64 is a
linear map T on GF(2)^64, so 2^16 lanes seeded with T^1..T^65536 of the
state are advanced together by the bit-matrix of T^65536.
"""
from __future__ import annotations

import numpy as np

_M = (1 << 64) - 1
_LANES = 1 << 16
_JUMP = None


def _xs(x: int) -> int:
    x ^= (x << 13) & _M
    x ^= x >> 7
    x ^= (x << 17) & _M
    return x


def _xs_vec(v: np.ndarray) -> np.ndarray:
    v = v ^ (v << np.uint64(13))
    v = v ^ (v >> np.uint64(7))
    v = v ^ (v << np.uint64(17))
    return v


def _jump_matrix() -> np.ndarray:
    """column b = T^_LANES (1 << b)."""
    global _JUMP
    if _JUMP is None:
        cols = np.array([1 << b for b in range(64)], dtype=np.uint64)
        for _ in range(_LANES):
            cols = _xs_vec(cols)
        _JUMP = cols
    return _JUMP


def _apply(mat: np.ndarray, v: np.ndarray) -> np.ndarray:
    out = np.zeros_like(v)
    one = np.uint64(1)
    for b in range(64):
        bit = (v >> np.uint64(b)) & one
        out ^= (np.uint64(0) - bit) & mat[b]
    return out


def xorshift_fill(n: int, st: list) -> np.ndarray:
    """n limbs from state st[0] (updated), identical to n calls of xr()."""
    out = np.empty(n, dtype=np.uint64)
    x = int(st[0])
    if n <= 2 * _LANES:
        for i in range(n):
            x = _xs(x)
            out[i] = x
        st[0] = x
        return out
    head = np.empty(_LANES, dtype=np.uint64)
    for j in range(_LANES):
        x = _xs(x)
        head[j] = x
    mat = _jump_matrix()
    cur = head
    pos = 0
    while pos < n:
        take = min(_LANES, n - pos)
        out[pos:pos + take] = cur[:take]
        pos += take
        if pos < n:
            cur = _apply(mat, cur)
    st[0] = int(out[-1])
    return out


def pattern_fill(n: int, pat: int, st: list) -> np.ndarray:
    """pat 0: xorshift limbs; 1: all ones; 2: all ones every 32nd limb, zero
    elsewhere (tests/p30_test.c:18-22)."""
    if pat == 1:
        return np.full(n, _M, dtype=np.uint64)
    if pat == 2:
        a = np.zeros(n, dtype=np.uint64)
        a[::32] = np.uint64(_M)
        return a
    return xorshift_fill(n, st)


def rlimb64_fill(n: int, st: list) -> np.ndarray:
    """biased limbs of tests/mpn_mul_u64_test.cpp:19-27: xr() & 7 picks 0,
    ~0, 1, 2^63 or a fresh xr() limb."""
    out = np.empty(n, dtype=np.uint64)
    x = int(st[0])
    for i in range(n):
        x = _xs(x)
        k = x & 7
        if k == 0:
            v = 0
        elif k == 1:
            v = _M
        elif k == 2:
            v = 1
        elif k == 3:
            v = 1 << 63
        else:
            x = _xs(x)
            v = x
        out[i] = v
    st[0] = x
    return out


def to_int(a) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint64).tobytes(), "little")


def from_int(v: int, n: int) -> np.ndarray:
    return np.frombuffer(v.to_bytes(8 * n, "little"), dtype=np.uint64).copy()


def all_ones_product(an: int, bn: int) -> np.ndarray:
    """(B^a - 1)(B^b - 1) = B^(a+b) - B^a - B^b + 1 with B = 2^64, as limbs
    (a >= b, either order taken): limb 0 is 1, limbs 1 .. b-1 are 0, limbs
    b .. a-1 are B - 1, limb a is B - 2 and limbs a+1 .. a+b-1 are B - 1."""
    a, b = max(an, bn), min(an, bn)
    r = np.full(a + b, np.uint64(2**64 - 1), dtype=np.uint64)
    r[0] = 1
    r[1:b] = 0
    r[a] = np.uint64(2**64 - 2)
    return r


def batch_workload(n: int = 4096, seed: int = 4, lo: int = 64, hi: int = 65536,
                   sparse: float = 0.10, ones: float = 0.05):
    """The batched mixed-size workload: n independent pairs, an and bn drawn
    independently log-uniform in [lo, hi] limbs; a `sparse` fraction of pairs
    with 90% zero limbs, a `ones` fraction all-ones, the rest uniform random
    u64. Returns a list of (a, b) uint64 arrays."""
    rng = np.random.default_rng(seed)
    span = np.log2(hi / lo)
    out = []
    for _ in range(n):
        an = int(lo * 2 ** (rng.random() * span))
        bn = int(lo * 2 ** (rng.random() * span))
        u = rng.random()
        if u < ones:
            a = np.full(an, _M, dtype=np.uint64)
            b = np.full(bn, _M, dtype=np.uint64)
        else:
            a = rng.integers(0, 2**64, an, dtype=np.uint64)
            b = rng.integers(0, 2**64, bn, dtype=np.uint64)
            if u < ones + sparse:
                a[rng.random(an) < 0.9] = 0
                b[rng.random(bn) < 0.9] = 0
        out.append((a, b))
    return out


_FP_BASE = {}


def fingerprint32(limbs, q: int) -> int:
    """(the integer of the u64 limbs) mod q for a prime q < 2^31, in O(n)
    numpy work: the 32-bit words h_j times 2^(32 j) mod q, summed mod q."""
    assert 2 < q < 2**31
    h = np.ascontiguousarray(limbs, dtype=np.uint64).view(np.uint32).astype(np.int64)
    n = len(h)
    if n == 0:
        return 0
    B = 4096
    if q not in _FP_BASE:
        base = np.empty(B, dtype=np.int64)
        x, w = 1, (1 << 32) % q
        for i in range(B):
            base[i] = x
            x = x * w % q
        _FP_BASE[q] = (base, x)                  # x = 2^(32 B) mod q
    base, stepB = _FP_BASE[q]
    nb = (n + B - 1) // B
    heads = np.empty(nb, dtype=np.int64)
    y = 1
    for k in range(nb):
        heads[k] = y
        y = y * stepB % q
    pw = (heads[:, None] * base[None, :] % q).reshape(-1)[:n]
    return int(((h % q) * pw % q).sum() % q)
