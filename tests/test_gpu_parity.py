"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Tolerances: FP64 kernels agree with the oracle to 1e-12 relative (the only
differences are fused multiply-adds and summation order on the device);
converged pressures agree to 1e-9 relative L2 after mean removal (the
pressure parity bar of the solver), mixed precision included.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ples_amd import mg
from tests import cases

pytestmark = pytest.mark.gpu


def _solver(g, nlevel, precision="fp64", metrics=None, **kw):
    return mg.Solver(g.x, g.y, g.z, g.ip, g.jp, g.kp, nlevel, metrics=metrics, precision=precision, **kw)


def _refine(a, axis):
    """Node array (nz+1, ny+1, nx+1) refined by two along `axis` (2: xi, 1: eta):
    the nodes at the even indices, midpoints between them."""
    a = np.moveaxis(a, axis, -1)
    out = np.empty(a.shape[:-1] + (2 * a.shape[-1] - 1,))
    out[..., ::2] = a
    out[..., 1::2] = 0.5 * (a[..., 1:] + a[..., :-1])
    return np.ascontiguousarray(np.moveaxis(out, -1, axis))


def _oracle_level_metrics(g, lev, nx, ny):
    """Oracle metrics of level `lev` (1-based) for a device level of nx, ny
    cells.  A level keeping twice the full-coarsening xi (eta) count has node
    stride 2^(lev-2) there and 2^(lev-1) in zeta: the oracle's level `lev` on
    the nodes refined by two in that direction reads exactly those nodes (the
    midpoints never), and divides by 2^(lev-1) where the device divides by
    sx sy / sz, a factor 2 per kept direction."""
    st = 1 << (lev - 1)
    fx, fy = nx * st // g.nx, ny * st // g.ny
    if fx == 1 and fy == 1:
        return O.metrics(g, lev - 1)[0]
    x, y, z = g.x, g.y, g.z
    for f, axis in ((fx, 2), (fy, 1)):
        if f == 2:
            x, y, z = _refine(x, axis), _refine(y, axis), _refine(z, axis)
    gu = O.Grid(fx * g.nx, fy * g.ny, g.nz, g.ip, g.jp, g.kp, x, y, z)
    return [fx * fy * r for r in O.metrics(gu, lev - 1)[0]]


@pytest.mark.parametrize("semi", [1, 2, 0])
@pytest.mark.parametrize("shape,gamma,amp,ip,kp,nlevel", [
    ((32, 16, 32), 1.5, 0.2, 0, 0, 3),
    ((24, 16, 16), 1.0, 0.15, 1, 1, 3),
    ((16, 8, 32), 1.5, 0.0, 0, 1, 2),
    ((16, 16, 32), 1.5, 0.05, 1, 0, 3),
])
def test_level_metrics_match_oracle(shape, gamma, amp, ip, kp, nlevel, semi):
    """Every level against the oracle's mul_met.  The grids with dz <= dx/1.41
    (the first, third and fourth) get a first coarse level that keeps the xi
    and eta counts (semicoarsen=1) or the xi count (2); semicoarsen=0 is the
    reference's hierarchy."""
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    with _solver(g, nlevel, semicoarsen=semi) as s:
        anis = shape != (24, 16, 16)
        kx = semi in (1, 2) and anis
        ky = semi == 1 and anis
        for lev in range(1, nlevel + 1):
            mine = s.level_metrics(lev)
            nx, ny, nz = s.level_shape(lev)
            up = 1 if lev > 1 else 0
            assert nz == g.nz >> (lev - 1)
            assert nx == g.nx >> (lev - 1 - (up if kx else 0))
            assert ny == g.ny >> (lev - 1 - (up if ky else 0))
            ref = _oracle_level_metrics(g, lev, nx, ny)
            for q in range(9):
                scale = np.abs(ref[q]).max() + 1e-300
                err = np.abs(mine[q] - ref[q]).max() / scale
                assert err < 1e-12, (lev, q, err)


def test_coarse_wall_metric_fallback():
    """gamma = 2 with 4 coarse cells across the channel: the reference's
    second-order one-sided wall derivative (fsrc/multigrid.f03:2695-2754)
    inverts the coarse wall-face Jacobian (negative g22); the device uses the
    first-order one-sided derivative there and agrees everywhere else."""
    g = cases.grid(32, 16, 32, gamma=2.0, amp=0.2)
    with _solver(g, 3, semicoarsen=0) as s:
        mine = s.level_metrics(3)
    ref, _, _ = O.metrics(g, 2)
    nx, ny, nz = 8, 4, 8
    assert (ref[4][1:nz + 1, 0, 1:nx + 1] < 0).all()       # the reference's value
    assert (mine[4][1:nz + 1, 0, 1:nx + 1] > 0).all()      # the device fallback
    assert (mine[4][1:nz + 1, ny, 1:nx + 1] > 0).all()
    inner = np.s_[1:nz + 1, 1:ny, 1:nx + 1]
    assert np.allclose(mine[4][inner], ref[4][inner], rtol=1e-12)
    for q in (0, 1, 2, 6, 7, 8):
        assert np.allclose(mine[q], ref[q], rtol=1e-12, atol=1e-14 * np.abs(ref[q]).max())


def test_metrica_vectors_and_giac():
    g = cases.grid(32, 16, 24, gamma=2.0, amp=0.2)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    with _solver(g, 2) as s:
        G, V, J = s.metrica()
    ref_g = cases.ref_metrics(g, gm)
    ref_v = cases.ref_metrics(g, csv)
    for q in range(9):
        assert np.allclose(G[q], ref_g[q], rtol=1e-12, atol=1e-14 * np.abs(ref_g[q]).max())
        assert np.allclose(V[q], ref_v[q], rtol=1e-12, atol=1e-14 * np.abs(ref_v[q]).max())
    assert np.allclose(J, cases.interior(giac), rtol=1e-12)


@pytest.mark.parametrize("ip,kp,gamma,amp", [(0, 0, 2.0, 0.2), (1, 1, 1.0, 0.1), (1, 0, 1.5, 0.1)])
def test_residual_matches_oracle(ip, kp, gamma, amp):
    # xi walls on a strongly warped and stretched grid make the reference's
    # ghost iteration diverge (|g12/g11| >> 1 on thin wall cells); the cases
    # here have a ghost fixed point, checked first.
    g = cases.grid(32, 16, 16, gamma=gamma, amp=amp, ip=ip, kp=kp)
    rng = np.random.default_rng(5)
    fi = rng.standard_normal(g.gshape())
    rhs = rng.standard_normal((g.nz, g.ny, g.nx))
    cs = cases.random_cs(g, 11)
    dt = 0.37
    gm, _, _ = O.metrics(g, 0)
    p = fi.copy()
    O.ghost_fill(g, 0, p, gm, cs, 1.0, dt, 80)
    p2 = p.copy()
    O.ghost_fill(g, 0, p2, gm, cs, 1.0, dt, 20)
    assert np.abs(p2 - p).max() < 1e-13  # converged ghost fixed point
    r_ref = O.resid(g, 0, p, cases.ghosted(g, rhs), gm)
    with _solver(g, 2, halo_passes=60) as s:
        r = s.residual(fi, rhs, cs, dt)
    ref = cases.interior(r_ref)
    err = np.abs(r - ref).max() / np.abs(ref).max()
    assert err < 1e-11, err


def _oracle_solution(g, nlevel, rhs_g, cs=None, dt=1.0):
    fi = g.zeros()
    n, rm, _ = O.multi(g, nlevel, fi, rhs_g, cs=cs, dt=dt, eps=1e-13, ficycle=3000)
    assert rm < 1e-12, (n, rm)
    return fi


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_matches_oracle_warped(precision):
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.2)
    rhs_g = O.case_rhs(g, 1002)
    ref = _oracle_solution(g, 3, rhs_g)
    with _solver(g, 3, precision) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=100)
    assert res.ok, res
    assert res.rel_res < 1e-12
    err = cases.rel_l2_meanfree(fi, ref)
    assert err < 1e-9, err


def test_solve_with_wall_fluxes_and_walls_everywhere():
    g = cases.grid(16, 16, 16, gamma=1.0, amp=0.1, ip=1, kp=1)
    rhs_g = O.case_rhs(g, 77)
    cs = cases.random_cs(g, 3)
    dt = 0.5
    # make the rhs compatible with the boundary fluxes
    O.project(g, rhs_g, cs, dt)
    ref = _oracle_solution(g, 2, rhs_g, cs=cs, dt=dt)
    with _solver(g, 3) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), cs=cs, dt=dt, eps=1e-12, ficycle=200)
    assert res.ok, res
    err = cases.rel_l2_meanfree(fi, ref)
    assert err < 1e-9, err


@pytest.mark.parametrize("ip,kp", [(0, 0), (1, 1)])
def test_exact_coarse_solve(ip, kp):
    """The dense coarsest-level solve is exact: on one level (the coarse solve
    is then the whole V-cycle) the solve with wall fluxes converges in at
    most two refinement steps and matches the oracle; on three levels the FP64
    and mixed solves with it match the oracle's solution (walls everywhere on
    the moderately warped grid: on strongly warped walled grids the default
    wall-ghost passes do not converge, a limitation of the red/black ghost
    fill, not of the coarse solve)."""
    g = cases.grid(8, 8, 8, gamma=1.0, amp=0.1, ip=ip, kp=kp)
    rhs_g = O.case_rhs(g, 91)
    cs = cases.random_cs(g, 5)
    dt = 0.5
    O.project(g, rhs_g, cs, dt)
    ref = _oracle_solution(g, 1, rhs_g, cs=cs, dt=dt)
    with _solver(g, 1, coarse_exact=1) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), cs=cs, dt=dt, eps=1e-12, ficycle=50)
    assert res.ok and res.cycles <= 2, res
    assert cases.rel_l2_meanfree(fi, ref) < 1e-9
    g = cases.grid(32, 16, 32, gamma=1.5 if ip == 0 else 1.0, amp=0.2 if ip == 0 else 0.1, ip=ip, kp=kp)
    rhs_g = O.case_rhs(g, 92)
    ref = _oracle_solution(g, 3, rhs_g)
    for prec in ("fp64", "mixed"):
        with _solver(g, 3, prec, coarse_exact=1) as s:
            fi = np.zeros(s.fi_shape)
            res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=300)
            assert res.ok, (prec, res)
        assert cases.rel_l2_meanfree(fi, ref) < 1e-9, prec


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_is_deterministic(precision):
    """The same solve gives bitwise the same pressure, cycles and residual when
    repeated on one handle and on a second handle (deterministic reductions,
    no kernel reading what another block of its launch writes)."""
    g = cases.grid(64, 32, 64, gamma=1.5, amp=0.2)
    rhs = cases.interior(O.case_rhs(g, 1004))
    out = []
    for rep in range(2):
        with _solver(g, 4, precision) as s:
            for _ in range(2):
                fi = np.zeros(s.fi_shape)
                res = s.multi(fi, rhs, eps=1e-11, ficycle=200)
                assert res.ok, res
                out.append((fi, res.cycles, res.rel_res))
    for fi, cyc, rr in out[1:]:
        assert cyc == out[0][1] and rr == out[0][2]
        assert np.array_equal(fi, out[0][0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_walls_everywhere_strongly_warped(precision):
    """Walls in every direction on a strongly warped grid (gamma 1.5, terrain
    amplitude 0.2): the wall-ghost passes contract slowly there, and the
    automatic pass count measured at create must carry the outer residual to
    its fixed point (a fixed 24 passes stalls near 1e-7)."""
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.2, ip=1, kp=1)
    rhs_g = O.case_rhs(g, 92)
    ref = _oracle_solution(g, 3, rhs_g)
    with _solver(g, 3, precision) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=300)
    assert res.ok, res
    assert cases.rel_l2_meanfree(fi, ref) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_matches_oracle_partial_blocks(precision):
    """A fine level whose xi count (48) fills 48 of the 64 lanes of its
    smoother and residual blocks, so the last lanes of every block row are
    idle: the terrain grid's packed 16-bit smoother metrics and the column
    kernels leave them out and the solve still matches the oracle."""
    g = cases.grid(48, 16, 24, gamma=1.5, amp=0.2)
    rhs_g = O.case_rhs(g, 1006)
    ref = _oracle_solution(g, 3, rhs_g)
    with _solver(g, 3, precision) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
    assert res.ok, res
    assert cases.rel_l2_meanfree(fi, ref) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_solve_matches_oracle_sheared(precision):
    """Grid lines skewed in the xi-zeta plane as well (x shifted by a function
    of z, periodic over the zeta period): every cross metric is nonzero, so the
    kernels read all nine metric arrays (on the terrain grids g13 and g31 are
    identically zero and are not read)."""
    g0 = cases.grid(32, 16, 32, gamma=1.5, amp=0.2)
    g = O.Grid(g0.nx, g0.ny, g0.nz, 0, 1, 0, g0.x + 0.15 * np.sin(2.0 * g0.z), g0.y, g0.z)
    gm, _, _ = O.metrics(g, 0)
    assert np.abs(gm[2]).max() > 1e-3 and np.abs(gm[6]).max() > 1e-3
    rhs_g = O.case_rhs(g, 1003)
    ref = _oracle_solution(g, 3, rhs_g)
    with _solver(g, 3, precision) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
    assert res.ok, res
    assert cases.rel_l2_meanfree(fi, ref) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sweep_schedule_same_solution(precision):
    """The Chebyshev sweep schedule changes the damping of each sweep, not the
    solution: both schedules converge to the oracle's pressure, the schedule
    in no more cycles than constant damping."""
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.2)
    rhs_g = O.case_rhs(g, 1002)
    ref = _oracle_solution(g, 3, rhs_g)
    cyc = {}
    for sched in (0, 1):
        with _solver(g, 3, precision, sweep_schedule=sched) as s:
            fi = np.zeros(s.fi_shape)
            res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
            assert res.ok, (sched, res)
            cyc[sched] = res.cycles
        assert cases.rel_l2_meanfree(fi, ref) < 1e-9, sched
    assert cyc[1] <= cyc[0], cyc


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape,gamma,amp,ip,kp,nlevel", [
    ((32, 16, 32), 0.0, 0.0, 0, 0, 3),
    ((32, 16, 32), 1.5, 0.2, 0, 0, 3),
    ((32, 16, 32), 1.0, 0.0, 1, 1, 3),
])
def test_semicoarsening_same_solution(precision, shape, gamma, amp, ip, kp, nlevel):
    """A first coarse level coarsened in zeta only (semicoarsen=1) or in eta
    and zeta (2) (zeta coupling four times the xi coupling here) changes the
    hierarchy, not the solution: all three hierarchies converge to the
    oracle's pressure.  On the uniform grid the eta-zeta one needs fewer
    cycles (measured 9 against 11 FP64 with GCR; zeta only: 11)."""
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    rhs_g = O.case_rhs(g, 1004)
    if ip or kp:
        O.project(g, rhs_g, None, 1.0)
    ref = _oracle_solution(g, nlevel, rhs_g)
    cyc = {}
    for semi in (0, 1, 2):
        with _solver(g, nlevel, precision, semicoarsen=semi) as s:
            assert s.level_shape(2) == (g.nx >> (0 if semi else 1), g.ny >> (0 if semi == 1 else 1), g.nz >> 1)
            fi = np.zeros(s.fi_shape)
            res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
            assert res.ok, (semi, res)
            cyc[semi] = res.cycles
        assert cases.rel_l2_meanfree(fi, ref) < 1e-9, semi
    if gamma == 0.0 and amp == 0.0:
        assert cyc[2] < cyc[0] and cyc[1] <= cyc[0], cyc


def test_diverging_wall_ghosts_refused_at_create():
    """xi and zeta walls on a steep terrain (slope 1.6 along zeta over thin
    zeta cells): the oracle's ghost iteration grows without bound, and the
    library refuses the grid at create (unsupported) instead of returning
    non-finite pressures; the milder terrain is accepted and solved."""
    g = cases.grid(16, 16, 32, gamma=1.5, amp=0.2, ip=1, kp=1)
    gm, _, _ = O.metrics(g, 0)
    p = np.random.default_rng(1).standard_normal(g.gshape())
    O.ghost_fill(g, 0, p, gm, None, 0.0, 1.0, 20)
    assert np.abs(p).max() > 1e6
    with pytest.raises(mg.PlesMGError) as e:
        _solver(g, 3)
    assert e.value.status == 7
    g2 = cases.grid(16, 16, 32, gamma=1.5, amp=0.05, ip=1, kp=1)
    rhs_g = O.case_rhs(g2, 1005)
    O.project(g2, rhs_g, None, 1.0)
    ref = _oracle_solution(g2, 3, rhs_g)
    with _solver(g2, 3) as s:
        fi = np.zeros(s.fi_shape)
        res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
    assert res.ok, res
    assert cases.rel_l2_meanfree(fi, ref) < 1e-9


def test_mixed_equals_fp64_c1_shape():
    g = cases.grid(64, 32, 64, gamma=2.0, amp=0.0)
    rhs = cases.interior(O.case_rhs(g, 1001))
    sols = {}
    for prec in ("fp64", "mixed"):
        with _solver(g, 4, prec) as s:
            fi = np.zeros(s.fi_shape)
            res = s.multi(fi, rhs, eps=1e-11, ficycle=200)
            assert res.ok and res.rel_res < 1e-11, (prec, res)
            sols[prec] = fi
    assert cases.rel_l2_meanfree(sols["mixed"], sols["fp64"]) < 1e-9


def test_case_rhs_matches_oracle():
    g = cases.grid(32, 16, 32, gamma=2.0, amp=0.2)
    ref = cases.interior(O.case_rhs(g, 4242))
    with _solver(g, 2) as s:
        s.case_rhs(4242)
        rhs = s.get_rhs()
    assert np.abs(rhs - ref).max() < 1e-10 * np.abs(ref).max()


def test_diver_and_gradie_match_oracle():
    g = cases.grid(24, 16, 16, gamma=2.0, amp=0.2)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    u, v, w = O.velocity(g, 9)
    uc, vc, wc = O.contra(g, csv, u, v, w)
    dt = 0.25
    rhs_ref = O.diver(g, uc, vc, wc, dt)
    fi = np.random.default_rng(1).standard_normal(g.gshape())
    outs_ref = O.gradie(g, fi, csv, giac, gm, dt)
    nx, ny, nz = g.nx, g.ny, g.nz
    with _solver(g, 2) as s:
        rhs = s.diver(uc[1:nz + 1, 1:ny + 1, 0:nx + 1], vc[1:nz + 1, 0:ny + 1, 1:nx + 1],
                      wc[0:nz + 1, 1:ny + 1, 1:nx + 1], dt)
        outs = s.gradie(fi, dt)
    assert np.allclose(rhs, cases.interior(rhs_ref), rtol=1e-12, atol=1e-12)
    for q in range(3):
        ref = cases.interior(outs_ref[q])
        assert np.abs(outs[q] - ref).max() < 1e-12 * np.abs(ref).max()
    sl = [np.s_[1:nz + 1, 1:ny + 1, 0:nx + 1], np.s_[1:nz + 1, 0:ny + 1, 1:nx + 1],
          np.s_[0:nz + 1, 1:ny + 1, 1:nx + 1]]
    for q in range(3):
        ref = outs_ref[3 + q][sl[q]]
        assert np.abs(outs[3 + q] - ref).max() < 1e-12 * np.abs(ref).max()


def test_kernel_timing_leaves_results_unchanged():
    """Kernel timing, on every refinement step or sampled (one step in four
    eager and timed, the rest replayed from their graphs), gives bitwise the
    untimed solve; sampling records about a quarter of the launches."""
    g = cases.grid(64, 32, 64, gamma=2.0, amp=0.2)
    out = {}
    with _solver(g, 4, "mixed") as s:
        s.case_rhs(1003)
        for every in (None, 1, 4):
            s.set_timing(["smooth_fine"] if every else None, every=every or 1)
            s.zero_fi()
            r = s.solve_resident(1e-10, 200)
            assert r.ok, (every, r)
            n = s.kernel_stats("smooth_fine")[0] if every else 0
            out[every] = (r.cycles, s.get_fi(), n)
        s.set_timing(None)
    for every in (1, 4):
        assert out[every][0] == out[None][0]
        assert np.array_equal(out[every][1], out[None][1]), every
    n1, n4 = out[1][2], out[4][2]
    assert n1 > 0 and 0 < n4 < n1 and abs(4 * n4 - n1) <= 4 * 6 + 6, (n1, n4)


def test_warm_start_resident_and_determinism():
    g = cases.grid(64, 32, 64, gamma=2.0, amp=0.2)
    with _solver(g, 4, "mixed") as s:
        s.case_rhs(1003)
        s.zero_fi()
        r1 = s.solve_resident(1e-10, 200)
        assert r1.ok, r1
        p1 = s.get_fi()
        s.perturb_fi(1004, 1e-3)
        r2 = s.solve_resident(1e-10, 200)
        assert r2.ok and r2.cycles < r1.cycles, (r1, r2)
        s.zero_fi()
        r3 = s.solve_resident(1e-10, 200)
        p3 = s.get_fi()
    assert r3.cycles == r1.cycles
    assert np.array_equal(p1, p3)  # bitwise deterministic


def test_tuned_launch_shapes_same_solution():
    """Launch shapes from the tuning table change the smoother's segments and
    the residual's blocks, not the solution: fixed-rule (autotune=-1) and
    table solves agree to the parity bar, and handles created with the
    default options (autotune=0) or with timing (1) report the table's shapes
    once it is loaded."""
    assert mg.tuning_load()
    g = cases.grid(64, 32, 64, gamma=1.5, amp=0.2)
    # nx ny nz ip kp f32 fine half -> P rP rbx use16
    table = {}
    for line in open(mg.TUNING_TABLE):
        if not line.startswith("#"):
            v = list(map(int, line.split()))
            table[tuple(v[:8])] = tuple(v[8:11])
    out = []
    for tune in (-1, 0, 1):
        with _solver(g, 4, "mixed", autotune=tune) as s:
            s.case_rhs(77)
            s.zero_fi()
            r = s.solve_resident(1e-11, 200)
            assert r.ok
            out.append(s.get_fi())
            if tune >= 0:
                assert s.tiles(1) == table[(64, 32, 64, 0, 0, 1, 1, 1)]
    assert cases.rel_l2_meanfree(out[1], out[0]) < 1e-9
    assert np.array_equal(out[1], out[2])


_SEG_REF = {}


def _seg_reference():
    """oracle solution of the 32x48x32 case (computed once per session)"""
    if not _SEG_REF:
        g = cases.grid(32, 48, 32, gamma=1.5, amp=0.2)
        rhs = O.case_rhs(g, 321)
        fiv = g.zeros()
        _, rmax, _ = O.multi(g, 2, fiv, rhs, eps=1e-13, ficycle=3000)
        _SEG_REF.update(g=g, rhs=rhs, fiv=fiv, rmax=rmax)
    return _SEG_REF["g"], _SEG_REF["rhs"], _SEG_REF["fiv"], _SEG_REF["rmax"]


@pytest.mark.parametrize("P", [1, 2, 3, 4, 6, 12])
def test_segment_counts_same_solution(P, tmp_path):
    """Every segment count of the line smoother (powers of two and the
    three-way splits) solves the same problem to the same solution: the
    fine level's P is forced through a tuning-table entry (a shape no other
    test uses) and compared with the oracle."""
    g, rhs, fiv, rmax = _seg_reference()
    table = tmp_path / "t.txt"
    rows = []
    for f32 in (0, 1):
        rows.append(f"32 48 32 0 0 {f32} 1 {P} {P} 32")
    table.write_text("\n".join(rows) + "\n")
    assert mg.tuning_load(str(table))
    with _solver(g, 3, "fp64", autotune=1, stop_mode="abs_max") as s:
        assert s.tiles(1)[0] == P
        fi = np.zeros(s.fi_shape)
        r = s.multi(fi, cases.interior(rhs), eps=1e-13, ficycle=300)
        assert r.ok, r
    assert rmax < 1e-13
    assert cases.rel_l2_meanfree(fi, fiv) < 1e-9


def test_vcycle_stops_at_first_dense_level():
    """With coarse_exact the V-cycle descends to the first coarse level of at
    most 2048 cells on grids periodic in xi and zeta (512 with xi/zeta walls),
    solves it exactly and still reaches the oracle's pressure; the deeper
    levels stay built (their shapes are reported) but are not visited."""
    g = cases.grid(64, 32, 64, gamma=1.5, amp=0.2)
    rhs_g = O.case_rhs(g, 93)
    ref = _oracle_solution(g, 3, rhs_g)  # the discrete solution does not depend on the level count
    for prec in ("fp64", "mixed"):
        with _solver(g, 5, prec) as s:
            # 64x32x64, 32x16x32, 16x8x16 (2048 cells: exact), then 8x4x8, 4x2x4 unvisited
            assert s.cycle_levels() == 3
            assert s.level_shape(5) == (4, 2, 4)
            fi = np.zeros(s.fi_shape)
            res = s.multi(fi, cases.interior(rhs_g), eps=1e-12, ficycle=200)
            assert res.ok, (prec, res)
        assert cases.rel_l2_meanfree(fi, ref) < 1e-9, prec
        with _solver(g, 5, prec, coarse_exact=0) as s:
            assert s.cycle_levels() == 5
    gw = cases.grid(32, 16, 32, gamma=1.0, amp=0.1, ip=1, kp=1)
    with _solver(gw, 3, "fp64") as s:
        assert s.cycle_levels() == 3  # 16x8x16 is above the walled limit: down to 8x4x8
