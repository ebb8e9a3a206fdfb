"""Synthetic cases of the benchmark configurations (grid nodes only; the
divergence right-hand side is generated on the device by
``Solver.case_rhs``).

Grids: x_i = i Lx/nx, z_k = k Lz/nz, eta_j = j/ny,
s(eta) = (1 + tanh(gamma (2 eta - 1)) / tanh(gamma)) / 2,
y = zb + (Ly - zb) s(eta) with the terrain zb = amp sin(2 pi x/(Lx/4)) cos(2 pi z/(Lz/4)).
Domain 2 pi x 2 x pi, periodic in x and z, walls in y.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

LX, LY, LZ = 2 * np.pi, 2.0, np.pi


def make_nodes(nx, ny, nz, gamma=2.0, amp=0.0, Lx=LX, Ly=LY, Lz=LZ):
    """Node coordinates x, y, z, each (nz+1, ny+1, nx+1) (Fortran (0:nx,0:ny,0:nz))."""
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    xi = i * Lx / nx
    zk = k * Lz / nz
    eta = j / ny
    s = 0.5 * (1.0 + np.tanh(gamma * (2.0 * eta - 1.0)) / np.tanh(gamma)) if gamma > 0 else eta
    zb = amp * np.sin(2 * np.pi * xi[None, :] / (Lx / 4.0)) * np.cos(2 * np.pi * zk[:, None] / (Lz / 4.0))
    shape = (nz + 1, ny + 1, nx + 1)
    x = np.broadcast_to(xi[None, None, :], shape).copy()
    z = np.broadcast_to(zk[:, None, None], shape).copy()
    y = zb[:, None, :] + (Ly - zb[:, None, :]) * s[None, :, None]
    return x, np.ascontiguousarray(y), z


@dataclass
class Config:
    name: str
    nx: int
    ny: int
    nz: int
    gamma: float
    amp: float
    nlevel: int
    tol: float
    seed: int
    text: str

    def nodes(self):
        return make_nodes(self.nx, self.ny, self.nz, self.gamma, self.amp)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


H = LY
CONFIGS = {
    "c1": Config("c1", 64, 32, 64, 2.0, 0.0, 4, 1e-8, 1001,
                 "channel 64x32x64, tanh 2.0, FP64, L=4, 1e-8"),
    "c2": Config("c2", 256, 128, 256, 1.5, 0.1 * H, 6, 1e-10, 1002,
                 "terrain 256x128x256 (zb=0.1H sin cos), tanh 1.5, L=6, 1e-10, FP64 MG vs FP32 V + FP64 refinement"),
    "c3": Config("c3", 512, 256, 512, 2.0, 0.1 * H, 7, 1e-10, 1003,
                 "warped 512x256x512, tanh 2.0, L=7, warm start p_prev + 1e-3 std N(0,1), 1e-10"),
}


# Batched workload: 64 independent problems, sizes drawn uniformly from five
# families, terrain amplitude U[0, 0.15] H, tanh 1.5 stretching, each with its
# own seeded divergence rhs.  Levels leave at least two wall-normal cells on
# the coarsest level (32^3 therefore uses 5 levels).
C4_SIZES = [(32, 32, 32, 5), (64, 64, 64, 6), (96, 48, 96, 5), (128, 64, 128, 6), (192, 96, 192, 5)]


def c4_problems(n=64, seed0=2000, sizes=C4_SIZES):
    probs = []
    for b in range(n):
        rng = np.random.default_rng(seed0 + b)
        nx, ny, nz, nlevel = sizes[int(rng.integers(len(sizes)))]
        amp = float(rng.uniform(0.0, 0.15)) * H
        x, y, z = make_nodes(nx, ny, nz, 1.5, amp)
        probs.append(dict(x=x, y=y, z=z, nlevel=nlevel, seed=seed0 + b, shape=(nx, ny, nz), amp=amp))
    return probs


# Serving workload: cases with sizes log-uniform between 32^3 and
# 256x128x256 cells, aspect 2:1:2, dims multiples of 2^(L-1) (at least two
# wall-normal cells on the coarsest level), terrain amplitude 0.1 H,
# tanh 1.5 stretching.  Requests arrive as a Poisson process; each picks a
# case and a tolerance of 1e-6 or 1e-10 with equal probability.
def c5_cases(n_cases=24, seed=3000):
    rng = np.random.default_rng(seed)
    out = []
    lo, hi = np.log(32.0 ** 3), np.log(256.0 * 128 * 256)
    for c in range(n_cases):
        cells = float(np.exp(rng.uniform(lo, hi)))
        a = (cells / 4.0) ** (1.0 / 3.0)
        nlevel = int(max(2, min(6, np.floor(np.log2(a)) - 1)))
        m = 2 ** (nlevel - 1)
        a = int(max(2 * m, round(a / m) * m))
        nx, ny, nz = 2 * a, a, 2 * a
        x, y, z = make_nodes(nx, ny, nz, 1.5, 0.1 * H)
        out.append(dict(case_id=c, x=x, y=y, z=z, nlevel=nlevel, shape=(nx, ny, nz), rhs_seed=3500 + c))
    return out


def c5_requests(n_req, n_cases, rate=200.0, seed=3000):
    """(arrival time s, case, tolerance, perturbation seed) per request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    reqs = []
    for r in range(n_req):
        q = np.random.default_rng(3001 + r)
        reqs.append((float(t[r]), int(q.integers(n_cases)), 1e-6 if q.random() < 0.5 else 1e-10, 3001 + r))
    return reqs
