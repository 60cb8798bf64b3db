"""Synthetic body sets for the 3-D Laplace FMM workloads.

The uniform-cube sets reproduce the reference drivers' initialisation
(3d/fmm.cxx:22-31: ``srand48(seed)``, then per body X[0], X[1], X[2], q from
consecutive ``drand48()`` draws), so that a body set built here is the same
array the reference would build.  ``drand48`` is the POSIX 48-bit linear
congruential generator x <- (0x5DEECE66D*x + 0xB) mod 2^48, returned as
x/2^48; it is evaluated here with a vectorised jump-ahead.
"""
from __future__ import annotations

import numpy as np

BODY_DTYPE = np.dtype([("X", "<f8", (3,)), ("q", "<f8"), ("p", "<f8"), ("F", "<f8", (3,))])
assert BODY_DTYPE.itemsize == 64  # bit layout of the reference Body (3d/exafmm.h:14-19)

_A = 0x5DEECE66D
_C = 0xB
_MASK = (1 << 48) - 1


def _mulmod48(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """(a*b) mod 2^48 for uint64 arrays holding 48-bit values."""
    m24 = np.uint64((1 << 24) - 1)
    ah, al = a >> np.uint64(24), a & m24
    bh, bl = b >> np.uint64(24), b & m24
    mid = (ah * bl + al * bh) & m24
    return ((mid << np.uint64(24)) + al * bl) & np.uint64(_MASK)


def drand48(seed: int, count: int) -> np.ndarray:
    """The first ``count`` values of drand48() after srand48(seed)."""
    x0 = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    block = min(count, 4096)
    states = np.empty(block, dtype=np.uint64)
    x = x0
    for i in range(block):
        x = (_A * x + _C) & _MASK
        states[i] = x
    out = np.empty(count, dtype=np.float64)
    out[:block] = states.astype(np.float64) / float(1 << 48)
    if count > block:
        # jump of `block` steps: x_{k+block} = Aj*x_k + Cj (mod 2^48)
        aj, cj = 1, 0
        for _ in range(block):
            aj, cj = (_A * aj) & _MASK, (_A * cj + _C) & _MASK
        Aj = np.uint64(aj)
        Cj = np.uint64(cj)
        pos = block
        cur = states
        while pos < count:
            cur = (_mulmod48(cur, np.full_like(cur, Aj)) + Cj) & np.uint64(_MASK)
            take = min(block, count - pos)
            out[pos:pos + take] = cur[:take].astype(np.float64) / float(1 << 48)
            pos += take
    return out


def uniform_cube(n: int, seed: int = 0, charge: str = "sym") -> np.ndarray:
    """N bodies uniform in [0,1)^3 in the reference's draw order.

    charge: "sym" -> q = 2*u-1 (U[-1,1)), "unit_over_n" -> q = u/N (U[0,1)/N),
    "neutral" -> q = u-0.5 minus the mean (3d/fmm.cxx:25-33).
    """
    r = drand48(seed, 4 * n).reshape(n, 4)
    b = np.zeros(n, dtype=BODY_DTYPE)
    b["X"] = r[:, :3]
    u = r[:, 3]
    if charge == "sym":
        b["q"] = 2.0 * u - 1.0
    elif charge == "unit_over_n":
        b["q"] = u / n
    elif charge == "neutral":
        q = u - 0.5
        b["q"] = q - q.sum() / n
    else:
        raise ValueError(charge)
    return b


def reference_driver_bodies(n: int, seed: int = 0) -> np.ndarray:
    """Bodies of 3d/fmm.cxx and test/test_fmm.h: X in [-pi,pi)^3, neutral q."""
    r = drand48(seed, 4 * n).reshape(n, 4)
    b = np.zeros(n, dtype=BODY_DTYPE)
    b["X"] = r[:, :3] * 2 * np.pi - np.pi
    q = r[:, 3] - 0.5
    avg = 0.0
    for v in q:  # sequential sum, as the driver accumulates it
        avg += v
    avg /= n
    b["q"] = q - avg
    return b


def plummer(n: int, a: float, rmax: float, rng: np.random.Generator, center=(0.5, 0.5, 0.5)) -> np.ndarray:
    """Plummer sphere positions: r = a/sqrt(u^(-2/3)-1), rejected to r <= rmax,
    isotropic direction, centred at `center`.  Returns an (n,3) array."""
    out = np.empty((n, 3))
    got = 0
    while got < n:
        m = int((n - got) * 1.3) + 16
        u = rng.random(m)
        u = u[u > 0]
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r <= rmax]
        k = min(r.size, n - got)
        r = r[:k]
        v = rng.normal(size=(k, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        out[got:got + k] = np.asarray(center) + v * r[:, None]
        got += k
    return out


def sample_targets(n: int, count: int) -> np.ndarray:
    """Target indices used for the error check: stride sampling of the
    tree-ordered bodies (3d/fmm.cxx:57-61)."""
    count = min(count, n)
    stride = n // count
    return np.arange(count, dtype=np.int64) * stride


def rel_l2(test: np.ndarray, ref: np.ndarray) -> float:
    """sqrt(sum |test-ref|^2 / sum |ref|^2), normalised by the FP64 direct values."""
    d = np.asarray(test, dtype=np.float64) - np.asarray(ref, dtype=np.float64)
    return float(np.sqrt(np.sum(d * d) / np.sum(np.asarray(ref, dtype=np.float64) ** 2)))


def plummer_background(n: int, seed: int) -> np.ndarray:
    """The learned-order workload: 90% of the bodies in a Plummer sphere of
    scale radius 0.05 centred in the unit cube and truncated at r = 0.5, 10%
    uniform in [0,1)^3, charges uniform in [-1,1)."""
    rng = np.random.default_rng(seed)
    b = np.zeros(n, dtype=BODY_DTYPE)
    nb = n // 10
    b["X"][: n - nb] = plummer(n - nb, 0.05, 0.5, rng)
    b["X"][n - nb:] = rng.random((nb, 3))
    b["q"] = rng.uniform(-1, 1, n)
    return b


def batched_problems(count: int, seed: int) -> list:
    """The batched workload: ``count`` independent problems, N log-uniform in
    [10^3, 10^5]; even-indexed problems uniform in [0,1)^3, odd-indexed ones
    Plummer spheres (scale radius 0.1, truncated at 10 scale radii, centred
    at 0.5); charges U[0,1)/N; expansion order drawn from {6, 8, 12}.
    Returns a list of (bodies, P)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(10.0 ** rng.uniform(3.0, 5.0)))
        b = np.zeros(n, dtype=BODY_DTYPE)
        if i % 2 == 0:
            b["X"] = rng.random((n, 3))
        else:
            b["X"] = plummer(n, 0.1, 1.0, rng)
        b["q"] = rng.random(n) / n
        out.append((b, int(rng.choice([6, 8, 12]))))
    return out
