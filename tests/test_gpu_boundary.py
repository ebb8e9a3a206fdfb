"""The drop-in boundary on the GPU:

  * the reference's own level-1 metric arrays g11..g33
    (fsrc/myarrays_metri3.f03:11-20, the shapes of fsrc/iniz_metrica.f03:87-109)
    passed through ples_mg_grid.g are the metrics the handle solves with: on a
    warped grid with walls in every direction they reproduce the node-built
    handle's metrics and pressure and the oracle's solution, and metrics that
    differ from the nodes' (scaled) are used as given;
  * ples_mg_project_rhs: the compatibility projection on host arrays makes the
    rhs integral equal the wall-flux integral (the oracle's orc_project);
  * the resident solve projects its rhs at every call;
  * the device's metric construction on node maps quadratic in each index
    with walls everywhere equals the closed form on every level (the pin of
    the oracle's one-sided wall derivatives, here for the device's k_metrics).
"""
import numpy as np
import pytest

from oracle import oracle as O
from ples_amd import mg
from tests import cases
from tests.test_oracle_pins import FACES, M_Q, V_Q, closed_form_metrics, quad_grid

pytestmark = pytest.mark.gpu


def _walled_grid():
    return cases.grid(32, 16, 32, gamma=1.5, amp=0.1, ip=1, kp=1)


def test_reference_metric_arrays_through_the_boundary():
    g = _walled_grid()
    L = 3
    gm, _, _ = O.metrics(g, 0)
    ref_shapes = cases.ref_metrics(g, gm)   # g11 (0:jx,1:jy,1:jz) ... in C order
    rhs_g = O.case_rhs(g, 41)
    cs = cases.random_cs(g, 42, scale=0.02)
    dt = 0.6
    O.project(g, rhs_g, cs, dt)
    ref = g.zeros()
    _, rm, _ = O.multi(g, L, ref, rhs_g, cs=cs, dt=dt, eps=1e-13, ficycle=3000)
    assert rm < 1e-13
    rhs = cases.interior(rhs_g)
    out = {}
    for name, met in (("nodes", None), ("arrays", ref_shapes)):
        with mg.Solver(g.x, g.y, g.z, 1, 1, 1, L, metrics=met, precision="fp64") as s:
            got = s.level_metrics(1)
            for q in range(9):
                scale = np.abs(gm[q]).max() + 1e-300
                assert np.abs(got[q] - gm[q]).max() / scale < 1e-12, (name, q)
            fi = np.zeros(s.fi_shape)
            r = s.multi(fi, rhs, cs=cs, dt=dt, eps=1e-12, ficycle=400)
            assert r.ok, (name, r)
            out[name] = fi
    assert cases.rel_l2_meanfree(out["arrays"], out["nodes"]) < 1e-11
    assert cases.rel_l2_meanfree(out["arrays"], ref) < 1e-9
    # metrics that are not the nodes' are used as given
    scaled = [m * (1.0 + 0.01 * q) for q, m in enumerate(ref_shapes)]
    with mg.Solver(g.x, g.y, g.z, 1, 1, 1, L, metrics=scaled, precision="fp64") as s:
        got = s.level_metrics(1)
        for q in range(9):
            want = gm[q] * (1.0 + 0.01 * q)
            assert np.abs(got[q] - want).max() <= 1e-14 * np.abs(want).max(), q
    # shapes are checked before anything reaches the library
    bad = list(ref_shapes)
    bad[4] = bad[4][:, :-1, :]
    with pytest.raises(ValueError):
        mg.Solver(g.x, g.y, g.z, 1, 1, 1, L, metrics=bad)


def test_project_rhs_export_matches_oracle():
    g = _walled_grid()
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((g.nz, g.ny, g.nx)) + 0.3
    cs = cases.random_cs(g, 6, scale=0.05)
    dt = 0.4
    ref = cases.ghosted(g, rhs.copy())
    removed_ref = O.project(g, ref, cs, dt)
    with mg.Solver(g.x, g.y, g.z, 1, 1, 1, 3) as s:
        mine = rhs.copy()
        removed = s.project_rhs(mine, cs, dt)
        assert np.allclose(s.get_rhs(), mine, rtol=0, atol=0)
    assert abs(removed - removed_ref) < 1e-13 * abs(removed_ref)
    assert np.abs(mine - cases.interior(ref)).max() < 1e-13
    flux = sum(c.sum() * sgn for c, sgn in zip(cs, (-1, 1, -1, 1, -1, 1))) / dt
    assert abs(mine.sum() - flux) < 1e-11 * np.abs(rhs).sum()


def test_resident_solve_projects_its_rhs():
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.1)
    rhs = cases.interior(O.case_rhs(g, 9)) + 0.25    # inconsistent: nonzero mean
    with mg.Solver(g.x, g.y, g.z, 0, 1, 0, 3, precision="mixed") as s:
        s.set_rhs(rhs)
        s.zero_fi()
        r = s.solve_resident(1e-10, 200)
        assert r.ok and abs(r.rhs_mean_removed - 0.25) < 1e-12, r
        assert abs(s.get_rhs().mean()) < 1e-14


def test_device_metrics_exact_on_quadratic_node_maps():
    nx, ny, nz, L = 16, 8, 16, 3
    g = quad_grid(nx, ny, nz, M_Q, V_Q, 1, 1, 1)
    with mg.Solver(g.x, g.y, g.z, 1, 1, 1, L, precision="fp64", semicoarsen=0) as s:
        for lev in range(1, L + 1):
            st = 1 << (lev - 1)
            shape = s.level_shape(lev)
            got = s.level_metrics(lev)
            ref, _, _ = closed_form_metrics(M_Q, V_Q, shape, st)
            for q in range(9):
                sl = FACES[q // 3](*shape)
                scale = np.abs(ref[q][sl]).max() + np.abs(ref[4 * (q // 3)][sl]).max()
                err = np.abs(got[q][sl] - ref[q][sl]).max() / scale
                assert err < 1e-12, (lev, q, err)
