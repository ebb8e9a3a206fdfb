"""Shared test inputs: grids and seeded right-hand sides built with the
oracle's case generator, and conversions between the oracle's ghosted
arrays and the reference-shaped arrays the C ABI takes."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

DOMAIN = dict(Lx=2 * np.pi, Ly=2.0, Lz=np.pi)


def grid(nx, ny, nz, gamma=2.0, amp=0.0, ip=0, jp=1, kp=0):
    return O.make_grid(nx, ny, nz, gamma=gamma, amp=amp, ip=ip, jp=jp, kp=kp, **DOMAIN)


def ref_metrics(g, gm):
    """oracle ghosted level-1 metrics -> reference shapes (g11 (0:nx,1:ny,1:nz) ...)"""
    nx, ny, nz = g.nx, g.ny, g.nz
    out = []
    for q in range(9):
        d = q // 3
        a = gm[q]
        if d == 0:
            out.append(np.ascontiguousarray(a[1:nz + 1, 1:ny + 1, 0:nx + 1]))
        elif d == 1:
            out.append(np.ascontiguousarray(a[1:nz + 1, 0:ny + 1, 1:nx + 1]))
        else:
            out.append(np.ascontiguousarray(a[0:nz + 1, 1:ny + 1, 1:nx + 1]))
    return out


def interior(a):
    return np.ascontiguousarray(a[1:-1, 1:-1, 1:-1])


def ghosted(g, a_int):
    out = g.zeros()
    out[1:-1, 1:-1, 1:-1] = a_int
    return out


def rel_l2_meanfree(p, q):
    """||(p - mean p) - (q - mean q)|| / ||q - mean q|| over the interior."""
    a = interior(p) if p.ndim == 3 and p.shape[0] > 2 else p
    b = interior(q) if q.ndim == 3 and q.shape[0] > 2 else q
    a = a - a.mean()
    b = b - b.mean()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def random_cs(g, seed, scale=1e-2):
    """Boundary fluxes cs1..cs6 (reference shapes) on wall faces, zero on
    periodic ones."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = g.nx, g.ny, g.nz
    cs = [np.zeros((nz, ny)), np.zeros((nz, ny)), np.zeros((nz, nx)), np.zeros((nz, nx)),
          np.zeros((ny, nx)), np.zeros((ny, nx))]
    if g.ip == 1:
        cs[0][:] = scale * rng.standard_normal(cs[0].shape)
        cs[1][:] = scale * rng.standard_normal(cs[1].shape)
    if g.jp == 1:
        cs[2][:] = scale * rng.standard_normal(cs[2].shape)
        cs[3][:] = scale * rng.standard_normal(cs[3].shape)
    if g.kp == 1:
        cs[4][:] = scale * rng.standard_normal(cs[4].shape)
        cs[5][:] = scale * rng.standard_normal(cs[5].shape)
    return cs


def driver_grid():
    """Nodes of the Fortran driver's grid (ples_amd/fortran/ples_mg_driver.f90):
    32x16x32, tanh 1.5 stretching, terrain amplitude 0.2."""
    g = grid(32, 16, 32, gamma=1.5, amp=0.2)
    return g.x, g.y, g.z
