"""Pins of the CPU oracle (no GPU).

The reference ships no test vectors for this path (its test directory does
not exist), so each oracle function is pinned by what the mathematics fixes:
  * metrics: on affine grids x = M xi the face metrics J^-1 g^{ij} equal
    det(M) (M^-1 M^-T)_{ij} exactly, for central and one-sided formulas and
    on every multigrid level (numpy.linalg.inv as the independent check);
  * operator: the 19-point flux form is exact for quadratic pressures on
    affine grids (A p = det(M) * lap(p)) and annihilates affine ones;
  * wall ghosts: the ghost fill reproduces an affine pressure when the
    boundary flux equals its exact flux;
  * transfers: restriction of a constant, the sum rule, prolongation of
    constants and of linear fields;
  * solver: the V-cycle solution equals an exact FFT + tridiagonal solve of
    the channel operator (closed-form coefficients) and a dense solve of the
    assembled operator on a small warped grid.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases


def affine_grid(nx, ny, nz, M, ip, jp, kp):
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    X = np.stack([i, j, k], axis=-1).astype(float) @ np.asarray(M, float).T
    return O.Grid(nx, ny, nz, ip, jp, kp, np.ascontiguousarray(X[..., 0]),
                  np.ascontiguousarray(X[..., 1]), np.ascontiguousarray(X[..., 2]))


# periodic directions must not shear (cellep shifts only along the direction)
M_PER = [[0.3, 0.05, 0.0], [0.0, 0.2, 0.0], [0.0, -0.07, 0.25]]
M_WALL = [[0.3, 0.05, 0.02], [0.04, 0.2, -0.03], [0.01, -0.07, 0.25]]


def exact_metrics(M):
    M = np.asarray(M, float)
    Mi = np.linalg.inv(M)
    det = np.linalg.det(M)
    return det * (Mi @ Mi.T), det, det * Mi


@pytest.mark.parametrize("M,flags", [(M_PER, (0, 1, 0)), (M_WALL, (1, 1, 1)), (M_PER, (0, 1, 1))])
def test_metrics_affine_exact_all_levels(M, flags):
    g = affine_grid(16, 8, 16, M, *flags)
    G, det, cof = exact_metrics(M)
    for lev in range(3):
        st = 1 << lev
        d = g.dims(lev)
        gm, csv, giac = O.metrics(g, lev, want_csv=(lev == 0), want_giac=(lev == 0))
        nx, ny, nz = d.nx, d.ny, d.nz
        faces = [np.s_[1:nz + 1, 1:ny + 1, 0:nx + 1], np.s_[1:nz + 1, 0:ny + 1, 1:nx + 1],
                 np.s_[0:nz + 1, 1:ny + 1, 1:nx + 1]]
        for q in range(9):
            a, b = q // 3, q % 3
            # coarse metrics are built from spacing st and divided by pot = st
            assert np.allclose(gm[q][faces[a]], G[a, b], rtol=1e-12, atol=1e-13), (lev, q)
        if lev == 0:
            for q in range(9):
                assert np.allclose(csv[q][faces[q // 3]], cof[q // 3, q % 3], rtol=1e-12, atol=1e-13)
            assert np.allclose(giac[1:-1, 1:-1, 1:-1], det, rtol=1e-12)
        del st


def _cell_centres(g):
    # centroid of the 8 nodes of every ghosted cell (affine grid: exact)
    M = np.array([g.x[0, 0, 1] - g.x[0, 0, 0], g.y[0, 0, 1] - g.y[0, 0, 0], g.z[0, 0, 1] - g.z[0, 0, 0]])
    Mj = np.array([g.x[0, 1, 0] - g.x[0, 0, 0], g.y[0, 1, 0] - g.y[0, 0, 0], g.z[0, 1, 0] - g.z[0, 0, 0]])
    Mk = np.array([g.x[1, 0, 0] - g.x[0, 0, 0], g.y[1, 0, 0] - g.y[0, 0, 0], g.z[1, 0, 0] - g.z[0, 0, 0]])
    k, j, i = np.meshgrid(np.arange(g.nz + 2), np.arange(g.ny + 2), np.arange(g.nx + 2), indexing="ij")
    xi = np.stack([i - 0.5, j - 0.5, k - 0.5], axis=-1)
    return xi @ np.stack([M, Mj, Mk], axis=1).T


@pytest.mark.parametrize("M,flags", [(M_PER, (0, 1, 0)), (M_WALL, (1, 1, 1))])
def test_operator_exact_on_quadratics(M, flags):
    g = affine_grid(12, 8, 10, M, *flags)
    gm, _, _ = O.metrics(g, 0)
    X = _cell_centres(g)
    Q = np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.1], [-0.2, 0.1, 0.7]])
    lin = np.array([0.4, -1.1, 0.9])
    p_quad = np.ascontiguousarray(np.einsum("...a,ab,...b->...", X, Q, X) + X @ lin)
    p_lin = np.ascontiguousarray(X @ lin + 3.0)
    det = np.linalg.det(np.asarray(M))
    zero = g.zeros()
    r = O.resid(g, 0, p_quad, zero, gm)
    assert np.allclose(r[1:-1, 1:-1, 1:-1], det * 2 * np.trace(Q), rtol=1e-9)
    r = O.resid(g, 0, p_lin, zero, gm)
    assert np.abs(r[1:-1, 1:-1, 1:-1]).max() < 1e-12


def test_wall_ghosts_reproduce_affine_pressure():
    """eta walls with a sheared eta direction (cross metrics g21, g23 != 0),
    periodic xi and zeta: p = c * y is periodic, and with the boundary flux
    set to its exact flux the wall ghosts must be the affine values."""
    M = M_PER
    g = affine_grid(12, 8, 10, M, 0, 1, 0)
    gm, _, _ = O.metrics(g, 0)
    assert np.abs(gm[3]).max() > 1e-3 and np.abs(gm[5]).max() > 1e-3
    X = _cell_centres(g)
    grad = np.array([0.0, -1.1, 0.0])
    p_ex = np.ascontiguousarray(X @ grad)
    G, det, _ = exact_metrics(M)
    flux = G @ (np.asarray(M, float).T @ grad)
    dt = 0.7
    nx, ny, nz = g.nx, g.ny, g.nz
    cs = [np.zeros((nz, ny)), np.zeros((nz, ny)), np.full((nz, nx), flux[1] * dt),
          np.full((nz, nx), flux[1] * dt), np.zeros((ny, nx)), np.zeros((ny, nx))]
    p = p_ex.copy()
    p[:, 0, :] = 0.0
    p[:, -1, :] = 0.0
    O.ghost_fill(g, 0, p, gm, cs, 1.0, dt, 40)
    assert np.abs(p - p_ex).max() < 1e-12


def test_restriction_and_prolongation_invariants():
    g = cases.grid(16, 8, 16)
    rng = np.random.default_rng(0)
    rf = g.zeros()
    rf[1:-1, 1:-1, 1:-1] = rng.standard_normal((16, 8, 16))
    rc = O.restrict(g, 1, rf)
    assert np.isclose(8 * rc[1:-1, 1:-1, 1:-1].sum(), -rf[1:-1, 1:-1, 1:-1].sum())
    const = g.zeros()
    const[1:-1, 1:-1, 1:-1] = 2.5
    assert np.allclose(O.restrict(g, 1, const)[1:-1, 1:-1, 1:-1], -2.5)
    # prolongation of a coarse constant (ghosts included) adds it everywhere
    ec = np.full(g.gshape(1), 1.25)
    pf = g.zeros()
    O.prolong(g, 1, ec, pf)
    assert np.allclose(pf, 1.25)
    # linear in the cell-centre coordinate (ghosts included): reproduced exactly
    shape_c = g.gshape(1)
    kc, jc, ic = np.meshgrid(np.arange(shape_c[0]) - 0.5, np.arange(shape_c[1]) - 0.5,
                             np.arange(shape_c[2]) - 0.5, indexing="ij")
    ec = np.ascontiguousarray(0.3 * ic - 0.7 * jc + 1.1 * kc)
    pf = g.zeros()
    O.prolong(g, 1, ec, pf)
    shape_f = g.gshape(0)
    kf, jf, i_f = np.meshgrid(np.arange(shape_f[0]) - 0.5, np.arange(shape_f[1]) - 0.5,
                              np.arange(shape_f[2]) - 0.5, indexing="ij")
    # fine centre (f - 1/2) maps to coarse coordinate (f + 1/2)/2 - 1/2 in cell units
    expect = 0.3 * ((i_f + 1) / 2 - 0.5) - 0.7 * ((jf + 1) / 2 - 0.5) + 1.1 * ((kf + 1) / 2 - 0.5)
    assert np.allclose(pf, expect, atol=1e-12)


def _channel_exact(nx, ny, nz, gamma, rhs_int):
    """Exact discrete solution of the channel operator: FFT in x,z, Thomas in y.
    Coefficients from the channel geometry: dy_j cell heights, face distance
    (y_{j+1} - y_{j-1})/2 between the centres of neighbouring eta faces."""
    Lx, Ly, Lz = 2 * np.pi, 2.0, np.pi
    dx, dz = Lx / nx, Lz / nz
    eta = np.arange(ny + 1) / ny
    y = Ly * 0.5 * (1 + np.tanh(gamma * (2 * eta - 1)) / np.tanh(gamma))
    dy = np.diff(y)
    a11 = dy * dz / dx                                   # per cell j (x-faces)
    a33 = dx * dy / dz
    a22 = dx * dz / (0.5 * (y[2:] - y[:-2]))             # interior eta faces 1..ny-1
    B = np.fft.fft2(rhs_int, axes=(0, 2))                # rhs_int[k, j, i]
    mx = np.fft.fftfreq(nx) * nx
    mz = np.fft.fftfreq(nz) * nz
    lx = -4 * np.sin(np.pi * mx / nx) ** 2
    lz = -4 * np.sin(np.pi * mz / nz) ** 2
    P = np.zeros_like(B)
    for kk in range(nz):
        for ii in range(nx):
            diag = a11 * lx[ii] + a33 * lz[kk]
            Amat = np.zeros((ny, ny))
            for j in range(ny):
                Amat[j, j] = diag[j]
                if j > 0:
                    Amat[j, j] -= a22[j - 1]
                    Amat[j, j - 1] += a22[j - 1]
                if j < ny - 1:
                    Amat[j, j] -= a22[j]
                    Amat[j, j + 1] += a22[j]
            rhs = B[kk, :, ii]
            if ii == 0 and kk == 0:
                P[kk, :, ii] = np.linalg.lstsq(Amat, rhs, rcond=None)[0]
            else:
                P[kk, :, ii] = np.linalg.solve(Amat, rhs)
    return np.real(np.fft.ifft2(P, axes=(0, 2)))


def test_multi_matches_exact_channel_solution():
    nx, ny, nz, gamma = 16, 16, 16, 1.0
    g = cases.grid(nx, ny, nz, gamma=gamma, amp=0.0)
    rhs_g = O.case_rhs(g, 1001)
    fi = g.zeros()
    n, rm, _ = O.multi(g, 3, fi, rhs_g, eps=1e-13, ficycle=3000)
    assert rm < 1e-12
    ex = _channel_exact(nx, ny, nz, gamma, rhs_g[1:-1, 1:-1, 1:-1])
    assert cases.rel_l2_meanfree(fi[1:-1, 1:-1, 1:-1], ex) < 1e-10


def test_multi_matches_dense_solve_warped():
    g = cases.grid(8, 4, 8, gamma=1.0, amp=0.2)
    gm, _, _ = O.metrics(g, 0)
    nint = 8 * 4 * 8
    Amat = np.zeros((nint, nint))
    zero = g.zeros()
    for c in range(nint):
        p = g.zeros()
        p[1:-1, 1:-1, 1:-1].flat[c] = 1.0
        O.ghost_fill(g, 0, p, gm, None, 0.0, 1.0, 80)
        Amat[:, c] = O.resid(g, 0, p, zero, gm)[1:-1, 1:-1, 1:-1].ravel()
    rhs_g = O.case_rhs(g, 5)
    sol = np.linalg.lstsq(Amat, rhs_g[1:-1, 1:-1, 1:-1].ravel(), rcond=None)[0]
    fi = g.zeros()
    n, rm, _ = O.multi(g, 2, fi, rhs_g, eps=1e-13, ficycle=5000)
    assert rm < 1e-12
    assert cases.rel_l2_meanfree(fi[1:-1, 1:-1, 1:-1].ravel(), sol) < 1e-9


def test_reference_multi_diverges_with_inverted_coarse_wall_metric():
    """The reference schedule (SOR 1.6, 5/10 sweeps) on config-1 stretching
    with 4 coarse cells across the channel: the coarse wall g22 is negative
    and the V-cycle diverges; with one level less it converges."""
    g = cases.grid(32, 16, 32, gamma=2.0)
    gm3, _, _ = O.metrics(g, 2)
    assert (gm3[4][1:-1, 0, 1:-1] < 0).all()
    rhs_g = O.case_rhs(g, 1001)
    fi = g.zeros()
    _, rm_bad, _ = O.multi(g, 3, fi, rhs_g, eps=1e-8, ficycle=40)
    fi = g.zeros()
    _, rm_ok, _ = O.multi(g, 2, fi, rhs_g, eps=1e-8, ficycle=400)
    assert not (rm_bad < 1e-8)
    assert rm_ok < 1e-8


def test_case_rhs_is_divergence_of_fluxes_and_zero_mean():
    g = cases.grid(16, 8, 16, gamma=2.0, amp=0.2)
    rhs = O.case_rhs(g, 3)
    assert abs(rhs[1:-1, 1:-1, 1:-1].mean()) < 1e-14
    # telescoping: with periodic x,z and zero wall flux the raw divergence sums to 0
    _, csv, _ = O.metrics(g, 0, want_csv=True)
    u, v, w = O.velocity(g, 3)
    uc, vc, wc = O.contra(g, csv, u, v, w)
    raw = O.diver(g, uc, vc, wc, 1.0)
    assert abs(raw[1:-1, 1:-1, 1:-1].sum()) < 1e-10 * np.abs(raw).max()
    # the generator is N(0,1)
    assert abs(u[1:-1, 1:-1, 1:-1].std() - 1.0) < 0.05


def test_gradie_of_affine_pressure():
    g = affine_grid(8, 8, 8, M_WALL, 1, 1, 1)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    X = _cell_centres(g)
    grad = np.array([0.4, -1.1, 0.9])
    p = np.ascontiguousarray(X @ grad)
    dt = 0.5
    gra1, gra2, gra3, cg1, cg2, cg3 = O.gradie(g, p, csv, giac, gm, dt)
    # gra = -dt * grad p in the interior (fsrc/gradie.f03:31-66)
    for q, a in enumerate((gra1, gra2, gra3)):
        assert np.allclose(a[1:-1, 1:-1, 1:-1], -dt * grad[q], rtol=1e-10)
    G, _, _ = exact_metrics(M_WALL)
    flux = G @ (np.asarray(M_WALL).T @ grad)
    assert np.allclose(cg1[1:-1, 1:-1, 1:-1], flux[0], rtol=1e-10)
