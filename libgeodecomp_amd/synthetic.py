"""Seeded synthetic inputs shared by the benchmark, the tests and the C++
initializers (libgeodecomp/io/synthetic.h implements the same hash).

Every value is a pure function of (seed, linear cell index) through
SplitMix64, so CPU and GPU runs see identical inputs without transfers and
without depending on a sequential RNG's call order.  The reference seeds its
MT19937 per coordinate for the same reason
(src/libgeodecomp/io/initializer.h seedRNG); bit-compatibility with that
generator is not needed because both sides of every comparison here share
this initializer.
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20261202

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, index: np.ndarray) -> np.ndarray:
    """SplitMix64 output for stream `seed` at positions `index` (uint64)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (index.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, offset: int = 0, chunk: int = 1 << 24) -> np.ndarray:
    """Doubles in [0,1) with 53 random bits, for positions offset..offset+n-1."""
    out = np.empty(n, dtype=np.float64)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        idx = np.arange(offset + s, offset + e, dtype=np.uint64)
        out[s:e] = (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return out


def jacobi_initial(nx: int, ny: int, nz: int, seed: int = DEFAULT_SEED) -> np.ndarray:
    """Interior values uniform[0,1), shape (nz, ny, nx), float64."""
    return uniform01(seed, nx * ny * nz).reshape(nz, ny, nx)


def life_initial(nx: int, ny: int, density: float = 0.3, seed: int = DEFAULT_SEED,
                 chunk: int = 1 << 24) -> np.ndarray:
    """Bernoulli(density) alive cells, shape (ny, nx), uint8 0/1."""
    n = nx * ny
    out = np.empty(n, dtype=np.uint8)
    thr = np.uint64(int(density * 2.0 ** 53))
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        idx = np.arange(s, e, dtype=np.uint64)
        out[s:e] = ((splitmix64(seed, idx) >> np.uint64(11)) < thr).astype(np.uint8)
    return out.reshape(ny, nx)


def lbm_initial_velocity(nx: int, ny: int, nz: int, amplitude: float = 1e-3,
                         seed: int = DEFAULT_SEED) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Velocity noise uniform[-a, a] per component, three (nz, ny, nx) arrays."""
    n = nx * ny * nz
    u = [(2.0 * uniform01(seed + k, n) - 1.0) * amplitude for k in (1, 2, 3)]
    return tuple(x.reshape(nz, ny, nx) for x in u)


def poisson_ensemble(count: int, sizes=tuple(range(32, 193, 16)), seed: int = DEFAULT_SEED):
    """Ensemble of Poisson problems on n^3 cubes (n drawn from `sizes`): the
    source is a sum of 1-8 Gaussians (centres uniform in the cube, widths
    uniform[2,12] cells, amplitudes N(0,1)) evaluated in FP64 at the cell
    centres and stored FP32; each of the six faces gets a Dirichlet value
    uniform[-1,1].  Returns a list of (n, faces[6], f) with f the padded
    (n+2)^3 float32 block (halo zero), x fastest."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.choice(np.asarray(sizes)))
        k = int(rng.integers(1, 9))
        centres = rng.uniform(0.0, n, size=(k, 3))
        widths = rng.uniform(2.0, 12.0, size=k)
        amps = rng.standard_normal(k)
        faces = rng.uniform(-1.0, 1.0, size=6)
        ax = np.arange(n, dtype=np.float64)
        f = np.zeros((n, n, n), dtype=np.float64)
        for (cx, cy, cz), w, a in zip(centres, widths, amps):
            gx = np.exp(-(ax - cx) ** 2 / (2 * w * w))
            gy = np.exp(-(ax - cy) ** 2 / (2 * w * w))
            gz = np.exp(-(ax - cz) ** 2 / (2 * w * w))
            f += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        fp = np.zeros((n + 2, n + 2, n + 2), dtype=np.float32)
        fp[1:-1, 1:-1, 1:-1] = f.astype(np.float32)
        out.append((n, faces, fp))
    return out
