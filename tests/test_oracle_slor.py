"""Pins of the oracle's eta-line relaxation (islor = 1): tridag, sett and
solut_sndrcv_slor (no GPU).

  * tridag (fsrc/tridag.f03:11-30) against scipy's banded solver;
  * sett's line coefficients (fsrc/sett.f03:119-143) against the operator:
    on a grid without cross metrics the tridiagonal rows of a column are
    exactly the operator applied to that column (its wall ghosts at the
    homogeneous flux relation);
  * one SLOR sweep (fsrc/multigrid.f03:2050-2199): the first column relaxed
    (all its neighbours still at their old values) with omega = 1 satisfies
    its cells' equations and its wall rows' flux relations exactly, and
    omega scales that update linearly;
  * the exact discrete solution is a fixed point of the sweep on warped
    grids with all nine metrics, walls, and wall fluxes, for any omega:
    every explicit term of the line system must be the operator's;
  * the SLOR V-cycle converges to the same pressure as the SOR V-cycle
    (itself pinned by the exact channel solution and a dense solve).
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import oracle as O
from tests import cases


def test_tridag_matches_banded_solve():
    rng = np.random.default_rng(2)
    for n in (2, 3, 7, 40):
        a = rng.standard_normal(n)
        c = rng.standard_normal(n)
        b = 3.0 + np.abs(a) + np.abs(c) + rng.random(n)
        b *= np.where(rng.random(n) < 0.5, -1.0, 1.0)
        a[0] = 0.0
        c[-1] = 0.0
        f = rng.standard_normal(n)
        ab = np.zeros((3, n))
        ab[0, 1:] = c[:-1]
        ab[1] = b
        ab[2, :-1] = a[1:]
        ref = solve_banded((1, 1), ab, f)
        got = O.tridag(a, b, c, f)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-14), n


def _orthogonal_grid(ip, kp):
    # channel nodes (no terrain): every cross metric vanishes
    return cases.grid(8, 12, 6, gamma=1.5, amp=0.0, ip=ip, kp=kp)


@pytest.mark.parametrize("ip,kp", [(0, 0), (1, 1)])
@pytest.mark.parametrize("level", [0, 1])
def test_sett_rows_are_the_operator_on_a_column(ip, kp, level):
    g = _orthogonal_grid(ip, kp)
    gm, _, _ = O.metrics(g, level)
    for q in (1, 2, 3, 5, 6, 7):
        assert np.abs(gm[q]).max() < 1e-12 * np.abs(gm[4]).max()
    aa, bb, cc = O.sett(g, level, gm)
    d = g.dims(level)
    rng = np.random.default_rng(7)
    zero = g.zeros(level)
    for (i, k) in ((1, 1), (d.nx, d.nz), (2, 3)):
        p = g.zeros(level)
        p[k, 1:-1, i] = rng.standard_normal(d.ny)
        O.ghost_fill(g, level, p, gm, None, 0.0, 1.0, 4)
        Ap = O.resid(g, level, p, zero, gm)[k, :, i]
        v = p[k, :, i]
        T = aa[k, :, i] * np.roll(v, 1) + bb[k, :, i] * v + cc[k, :, i] * np.roll(v, -1)
        assert np.allclose(T[1:-1], Ap[1:-1], rtol=1e-12, atol=1e-13 * np.abs(Ap).max()), (i, k)
        # wall rows: the homogeneous flux relation F_y = 0 holds for the filled ghosts
        assert abs(T[0]) < 1e-13 * np.abs(Ap).max() and abs(T[-1]) < 1e-13 * np.abs(Ap).max()


@pytest.mark.parametrize("ip,kp", [(0, 0), (1, 1), (1, 0)])
def test_first_slor_column_solves_its_equations(ip, kp):
    """omega = 1: after the sweep, column (1, 1) -- relaxed first, so every
    value it reads is the old one -- satisfies A p = rh on its cells (its
    x/z wall ghosts refilled from it) and its wall rows carry the fluxes."""
    g = _orthogonal_grid(ip, kp)
    gm, _, _ = O.metrics(g, 0)
    rng = np.random.default_rng(4)
    p0 = rng.standard_normal(g.gshape())
    rh = cases.ghosted(g, rng.standard_normal((g.nz, g.ny, g.nx)))
    cs = cases.random_cs(g, 9, scale=0.3)
    dt = 0.8
    p = p0.copy()
    O.slor_sweep(g, 0, p, rh, gm, cs, an=1.0, dt=dt, omega=1.0, nbc=5)
    col = p[1, :, 1].copy()
    # the state the first column saw: the ghost fill and the periodic-k planes
    # of the sweep's start; replace the column by its relaxed values, refill
    # the ghosts (orthogonal grid: the ghosts next to column (1, 1) depend on
    # that column only) and take the residual of its cells
    q = p0.copy()
    O.ghost_fill(g, 0, q, gm, cs, 1.0, dt, 5)
    if kp == 0:
        q[0, 1:-1, 1:-1], q[-1, 1:-1, 1:-1] = q[-2, 1:-1, 1:-1], q[1, 1:-1, 1:-1].copy()
    q[1, :, 1] = col
    O.ghost_fill(g, 0, q, gm, cs, 1.0, dt, 3)
    assert np.allclose(q[1, :, 1], col, rtol=1e-14, atol=1e-14)
    r = O.resid(g, 0, q, rh, gm)[1, 1:-1, 1]
    assert np.abs(r).max() < 1e-12 * np.abs(rh).max(), np.abs(r).max()
    # eta-wall rows: g22 (p1 - p0) = cs3 / dt, g22 (p_ny+1 - p_ny) = cs4 / dt
    g22 = gm[4]
    assert np.isclose(g22[1, 0, 1] * (col[1] - col[0]), cs[2][0, 0] / dt, rtol=1e-12)
    assert np.isclose(g22[1, g.ny, 1] * (col[-1] - col[-2]), cs[3][0, 0] / dt, rtol=1e-12)
    # omega scales the update of the first column
    ph = p0.copy()
    O.slor_sweep(g, 0, ph, rh, gm, cs, an=1.0, dt=dt, omega=0.5, nbc=5)
    pf = p0.copy()
    O.ghost_fill(g, 0, pf, gm, cs, 1.0, dt, 5)
    assert np.allclose(ph[1, :, 1] - pf[1, :, 1], 0.5 * (col - pf[1, :, 1]), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("ip,kp", [(0, 0), (1, 1), (0, 1), (1, 0)])
def test_exact_solution_is_a_fixed_point_of_the_slor_sweep(ip, kp):
    g = cases.grid(16, 8, 16, gamma=1.5, amp=0.15, ip=ip, kp=kp)
    gm, _, _ = O.metrics(g, 0)
    # terrain grid: the cross metrics g12, g21, g23, g32 do not vanish (g13 and
    # g31 do: the xi and zeta grid lines stay orthogonal)
    for q in (1, 3, 5, 7):
        assert np.abs(gm[q]).max() > 1e-3 * np.abs(gm[4]).max(), q
    cs = cases.random_cs(g, 21, scale=0.05)
    dt = 0.5
    rhs = O.case_rhs(g, 8)
    O.project(g, rhs, cs, dt)
    fi = g.zeros()
    n, rm, _ = O.multi(g, 2, fi, rhs, cs=cs, dt=dt, eps=1e-14, ficycle=4000)
    assert rm < 1e-13
    O.ghost_fill(g, 0, fi, gm, cs, 1.0, dt, 200)
    scale = np.abs(fi[1:-1, 1:-1, 1:-1]).max()
    for omega in (1.0, 1.5):
        p = fi.copy()
        O.slor_sweep(g, 0, p, rhs, gm, cs, an=1.0, dt=dt, omega=omega, nbc=5)
        err = np.abs(p[1:-1, :, 1:-1] - fi[1:-1, :, 1:-1]).max() / scale
        assert err < 1e-10, (omega, err)
    # a perturbed state is not: the sweep moves it back towards the solution
    p = fi.copy()
    p[5, 3, 7] += 1e-3 * scale
    before = np.abs(p - fi).max()
    O.slor_sweep(g, 0, p, rhs, gm, cs, an=1.0, dt=dt, omega=1.0, nbc=5)
    assert np.abs(p[1:-1, 1:-1, 1:-1] - fi[1:-1, 1:-1, 1:-1]).max() < 0.5 * before


@pytest.mark.parametrize("shape,nlevel,flags", [((16, 16, 16), 3, (0, 0)), ((16, 8, 16), 2, (1, 1))])
def test_slor_vcycle_converges_to_the_sor_solution(shape, nlevel, flags):
    ip, kp = flags
    g = cases.grid(*shape, gamma=1.0, amp=0.1, ip=ip, kp=kp)
    rhs = O.case_rhs(g, 3)
    fi_sor, fi_slor = g.zeros(), g.zeros()
    n0, r0, _ = O.multi(g, nlevel, fi_sor, rhs, eps=1e-13, ficycle=3000, islor=0)
    n1, r1, h1 = O.multi(g, nlevel, fi_slor, rhs, eps=1e-13, ficycle=3000, islor=1)
    assert r0 < 1e-13 and r1 < 1e-13
    assert n1 < n0  # the line relaxation is the stronger smoother on stretched cells
    assert cases.rel_l2_meanfree(fi_slor, fi_sor) < 1e-10


@pytest.mark.parametrize("islor", [0, 1])
def test_k_slab_sweeps_reach_the_same_solution(islor):
    """orc_hier_multi on k-slabs (the reference's MPI slab semantics, used by
    the CPU baseline) converges to the global-domain solution for any slab
    count, and one hierarchy serves solves of different slab counts."""
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.2)
    rhs = O.case_rhs(g, 3)
    H = O.Hier(g, 3)
    out = {}
    for ns in (1, 4, 2, 8):
        fi = g.zeros()
        n, rm, _ = H.multi(fi, rhs, eps=1e-13, ficycle=3000, islor=islor, omega=1.0 if islor else 1.6, nslab=ns)
        assert rm < 1e-13, (ns, rm)
        out[ns] = fi
    with pytest.raises(ValueError):
        H.multi(g.zeros(), rhs, eps=1e-8, ficycle=3, nslab=16)   # 32 planes / (16 slabs * 4) < 1
    H.close()
    for ns in (2, 4, 8):
        assert cases.rel_l2_meanfree(out[ns], out[1]) < 1e-9, ns
