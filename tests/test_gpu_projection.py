"""GPU parity of the time-loop steps around the solve (update, contra,
vel_up, check_divergence) and of the one-call projection step against the
CPU oracle, through the C ABI.

Tolerances: the elementwise FP64 kernels agree with the oracle to 1e-13
relative (fused multiply-adds only); the projection step's pressure agrees to
1e-9 relative L2 after mean removal and its velocities to 1e-8 relative (both
sides solved to a max-norm residual of 1e-13), and the divergence after the
correction equals dt times the device residual inside the field.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ples_amd import mg
from tests import cases
from tests.test_oracle_projection import oracle_projection_step

pytestmark = pytest.mark.gpu

LAYOUTS = [((16, 12, 20), 1.5, 0.2, 0, 0), ((12, 8, 10), 1.0, 0.1, 1, 1)]


def _solver(g, nlevel=2, **kw):
    return mg.Solver(g.x, g.y, g.z, g.ip, g.jp, g.kp, nlevel, **kw)


def _faces(g, fc):
    nx, ny, nz = g.nx, g.ny, g.nz
    return [np.ascontiguousarray(fc[0][1:nz + 1, 1:ny + 1, 0:nx + 1]),
            np.ascontiguousarray(fc[1][1:nz + 1, 0:ny + 1, 1:nx + 1]),
            np.ascontiguousarray(fc[2][0:nz + 1, 1:ny + 1, 1:nx + 1])]


def _close(a, b, rtol):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale


@pytest.mark.parametrize("shape,gamma,amp,ip,kp", LAYOUTS)
def test_update_matches_oracle(shape, gamma, amp, ip, kp):
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    rng = np.random.default_rng(31)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    dvel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    ref = [a.copy() for a in vel]
    O.update(g, ref, dvel)
    with _solver(g) as s:
        s.update(vel, dvel)
    for c in range(3):
        assert _close(vel[c], ref[c], 1e-15)  # fused multiply-adds in the extrapolation only


@pytest.mark.parametrize("shape,gamma,amp,ip,kp", LAYOUTS)
@pytest.mark.parametrize("potenziale", [0, 1])
def test_contra_matches_oracle(shape, gamma, amp, ip, kp, potenziale):
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    rng = np.random.default_rng(32)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    old = [rng.standard_normal(g.gshape()) for _ in range(3)]
    ref = [a.copy() for a in old]
    cs_ref = O.contra_step(g, csv, vel, ref, potenziale)
    fc = _faces(g, old)
    with _solver(g) as s:
        cs = s.contra(vel, fc, potenziale)
    for d, want in enumerate(_faces(g, ref)):
        assert _close(fc[d], want, 1e-13), d
    walls = (ip, 1, kp)
    for q in range(6):
        if walls[q // 2]:
            assert _close(cs[q], cs_ref[q], 1e-13), q


def test_vel_up_and_check_divergence_match_oracle():
    g = cases.grid(16, 12, 20, gamma=1.5, amp=0.2, ip=0, kp=0)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    rng = np.random.default_rng(33)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    fc0 = [rng.standard_normal(g.gshape()) for _ in range(3)]
    gra = [rng.standard_normal(g.gshape()) for _ in range(3)]
    cgra = [rng.standard_normal(g.gshape()) for _ in range(3)]
    dt = 0.37
    vref = [a.copy() for a in vel]
    fref = [a.copy() for a in fc0]
    O.vel_up(g, vref, fref, gra, cgra, dt, 0)
    fc = _faces(g, fc0)
    with _solver(g) as s:
        s.vel_up(vel, fc, [cases.interior(a) for a in gra], _faces(g, cgra), dt, 0)
        div = s.check_divergence(fc)
    for c in range(3):
        assert _close(vel[c], vref[c], 1e-15)
    for d, want in enumerate(_faces(g, fref)):
        assert _close(fc[d], want, 1e-15)
    div_ref = O.check_divergence(g, fref, giac)
    assert np.allclose(div, div_ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_projection_step_matches_oracle_chain(precision):
    g = cases.grid(32, 16, 32, gamma=1.5, amp=0.2)
    rng = np.random.default_rng(41)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    dvel = [0.1 * rng.standard_normal(g.gshape()) for _ in range(3)]
    fc = [np.zeros(g.gshape()) for _ in range(3)]
    dt = 0.5
    vref = [a.copy() for a in vel]
    fref = [a.copy() for a in fc]
    fiv = g.zeros()
    cs_ref, rhs_ref, _, rmax_ref, _, giac = oracle_projection_step(g, 3, vref, dvel, fref, fiv, dt, 1e-13)
    assert rmax_ref < 1e-13

    fcd = _faces(g, fc)
    fi = np.zeros(g.gshape())
    with _solver(g, 3, precision=precision, stop_mode="abs_max") as s:
        res, cs, div = s.projection_step(vel, dvel, fcd, fi, dt=dt, eps=1e-13, ficycle=300)
        assert res.ok, res
        rhs_dev = s.get_rhs()
        r = s.residual(fi, rhs_dev, cs=cs, dt=dt)
    assert cases.rel_l2_meanfree(fi, fiv) < 1e-9
    for q in (2, 3):
        assert _close(cs[q], cs_ref[q], 1e-13)
    for c in range(3):
        assert _close(cases.interior(vel[c]), cases.interior(vref[c]), 1e-8)
    for d, want in enumerate(_faces(g, fref)):
        assert _close(fcd[d], want, 1e-8)
    # divergence inside the field = dt * residual there (device quantities only)
    assert div[0] <= dt * np.abs(r[1:-1, 1:-1, 1:-1]).max() * (1 + 1e-6) + 1e-14
    assert div[0] < 1e-10
