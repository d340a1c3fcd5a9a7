"""Synthetic particle sets of the benchmark workloads (host, numpy).

All sets lie in the root cell [-1/2, 1/2)^3 of side 1 centred at 0."""
from __future__ import annotations

import numpy as np


def uniform(N: int, rng) -> np.ndarray:
    """i.i.d. uniform in [0,1)^3, shifted into the root cell."""
    return rng.random((N, 3)) - 0.5


def _directions(n: int, rng) -> np.ndarray:
    d = rng.standard_normal((n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def plummer(N: int, rng, a: float = 0.1, rmax: float = 1.0) -> np.ndarray:
    """Plummer sphere of scale a (radii by inverting the cumulative mass
    M(<r) = r^3 / (r^2 + a^2)^(3/2)), truncated to r <= rmax, then mapped
    from [-rmax, rmax]^3 into the root cell."""
    out = np.empty((0, 3))
    while out.shape[0] < N:
        m = 2 * (N - out.shape[0]) + 16
        u = rng.random(m)
        u = u[(u > 0) & (u < 1)]
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = r[r <= rmax]
        out = np.concatenate([out, r[:, None] * _directions(r.size, rng)])
    return 0.5 / rmax * out[:N]


def sphere_surface(N: int, rng, radius: float = 0.4) -> np.ndarray:
    """Uniform on the sphere of the given radius about the centre."""
    return radius * _directions(N, rng)


def uneven_instances(count: int = 64, nmin: int = 4096, nmax: int = 262144, seed: int = 0):
    """count instances, N log-uniform in [nmin, nmax]; the first half Plummer
    (a = 0.1, r <= 1), the second half on a sphere of radius 0.4; charges
    uniform [0, 1].  Returns (list of positions, list of charges)."""
    rng = np.random.default_rng(seed)
    Ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(int)
    pos, q = [], []
    for i, n in enumerate(Ns):
        p = plummer(int(n), rng) if i < count // 2 else sphere_surface(int(n), rng)
        pos.append(p)
        q.append(rng.uniform(0.0, 1.0, size=int(n)))
    return pos, q
