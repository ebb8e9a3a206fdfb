"""Pins of the oracle's time-loop steps around the solve (no GPU):
update, contra, vel_up and check_divergence, and the chain
update -> contra -> diver -> multi -> gradie -> vel_up.

  * update: the wall ghost is the parabola through the first three cells
    evaluated at the wall face (half a cell out), so any field quadratic in
    the cell index is reproduced exactly there; periodic ghosts are copies;
  * contra: a constant velocity gives face fluxes U . (face vector) on every
    face, and the discrete divergence of those fluxes vanishes to round-off
    because the face vectors of every cell close (geometric conservation), a
    property that any shifted face index breaks on a warped grid;
  * vel_up: closed form (increments, periodic images, untouched wall fluxes);
  * check_divergence: brute force in numpy;
  * chain: after the correction the divergence of every cell not touching a
    wall equals -dt times the pressure residual of that cell, because gradie's
    face gradients are the operator's face fluxes.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import cases


def _quadratic(g, shape):
    k, j, i = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    return 0.3 * i * i - 0.2 * j * j + 0.1 * k * k + 0.05 * i * j - 0.07 * j * k + 0.4 * i - j + 2.0, (i, j, k)


@pytest.mark.parametrize("flags", [(0, 1, 0), (1, 1, 1)])
def test_update_wall_extrapolation_exact_on_quadratics(flags):
    ip, jp, kp = flags
    g = cases.grid(8, 6, 10, gamma=1.5, amp=0.1, ip=ip, kp=kp)
    q, (i, j, k) = _quadratic(g, g.gshape())
    rng = np.random.default_rng(3)
    vel = [q.copy(), 2.0 * q, -q + 1.0]
    dvel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    base = [v.copy() for v in vel]
    O.update(g, vel, dvel)
    nx, ny, nz = g.nx, g.ny, g.nz
    I = np.s_[1:-1, 1:-1, 1:-1]
    for c in range(3):
        u = vel[c]
        assert np.array_equal(u[I], base[c][I] + dvel[c][I])
        new = base[c] + dvel[c]  # interior after the increment
        # wall extrapolation of the incremented interior; exact for the quadratic part
        def par(a1, a2, a3):
            return 1.875 * a1 - 1.25 * a2 + .375 * a3
        if ip:
            assert np.allclose(u[1:-1, 1:-1, 0], par(new[1:-1, 1:-1, 1], new[1:-1, 1:-1, 2], new[1:-1, 1:-1, 3]))
        else:
            assert np.array_equal(u[1:-1, 1:-1, 0], u[1:-1, 1:-1, nx])
            assert np.array_equal(u[1:-1, 1:-1, nx + 1], u[1:-1, 1:-1, 1])
        assert np.allclose(u[1:-1, 0, 1:-1], par(new[1:-1, 1, 1:-1], new[1:-1, 2, 1:-1], new[1:-1, 3, 1:-1]))
        if not kp:
            assert np.array_equal(u[0, 1:-1, 1:-1], u[nz, 1:-1, 1:-1])
    # the quadratic itself (zero increment): ghost = value at the wall face
    vel = [q.copy() for _ in range(3)]
    O.update(g, vel, [np.zeros(g.gshape())] * 3)
    qf = lambda ii, jj, kk: (0.3 * ii * ii - 0.2 * jj * jj + 0.1 * kk * kk + 0.05 * ii * jj - 0.07 * jj * kk
                             + 0.4 * ii - jj + 2.0)
    J = np.arange(1, ny + 1)[None, :, None]
    K = np.arange(1, nz + 1)[:, None, None]
    Ii = np.arange(1, nx + 1)[None, None, :]
    assert np.allclose(vel[0][1:-1, 0:1, 1:-1], qf(Ii, 0.5, K), rtol=0, atol=1e-12)
    assert np.allclose(vel[0][1:-1, ny + 1:ny + 2, 1:-1], qf(Ii, ny + 0.5, K), rtol=0, atol=1e-12)
    if ip:
        assert np.allclose(vel[0][1:-1, 1:-1, 0:1], qf(0.5, J, K), rtol=0, atol=1e-12)
    if kp:
        assert np.allclose(vel[0][0:1, 1:-1, 1:-1], qf(Ii, J, 0.5), rtol=0, atol=1e-12)
    # edges are not touched
    assert vel[0][0, 0, 0] == q[0, 0, 0]


@pytest.mark.parametrize("flags", [(0, 1, 0), (1, 1, 1)])
@pytest.mark.parametrize("potenziale", [0, 1])
def test_contra_constant_velocity_and_closure(flags, potenziale):
    ip, jp, kp = flags
    g = cases.grid(12, 8, 10, gamma=1.5, amp=0.2, ip=ip, kp=kp)
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    U = (0.3, -1.1, 0.7)
    vel = [np.full(g.gshape(), U[c]) for c in range(3)]
    rng = np.random.default_rng(5)
    old = [rng.standard_normal(g.gshape()) for _ in range(3)]
    fc = [o.copy() for o in old]
    cs = O.contra_step(g, csv, vel, fc, potenziale)
    nx, ny, nz = g.nx, g.ny, g.nz
    boxes = [np.s_[1:nz + 1, 1:ny + 1, 0:nx + 1], np.s_[1:nz + 1, 0:ny + 1, 1:nx + 1],
             np.s_[0:nz + 1, 1:ny + 1, 1:nx + 1]]
    for d in range(3):
        exact = sum(U[c] * csv[3 * d + c] for c in range(3))
        assert np.allclose(fc[d][boxes[d]], exact[boxes[d]], rtol=0, atol=1e-14)
    # wall flux changes
    e0 = sum(U[c] * csv[3 + c] for c in range(3))
    want3 = -old[1][1:nz + 1, 0, 1:nx + 1] if potenziale else e0[1:nz + 1, 0, 1:nx + 1] - old[1][1:nz + 1, 0, 1:nx + 1]
    assert np.allclose(cs[2], want3, rtol=0, atol=1e-14)
    want4 = -old[1][1:nz + 1, ny, 1:nx + 1] if potenziale else e0[1:nz + 1, ny, 1:nx + 1] - old[1][1:nz + 1, ny, 1:nx + 1]
    assert np.allclose(cs[3], want4, rtol=0, atol=1e-14)
    if ip:
        ex = sum(U[c] * csv[c] for c in range(3))
        want1 = -old[0][1:nz + 1, 1:ny + 1, 0] if potenziale else ex[1:nz + 1, 1:ny + 1, 0] - old[0][1:nz + 1, 1:ny + 1, 0]
        assert np.allclose(cs[0], want1, rtol=0, atol=1e-14)
    else:
        assert not cs[0].any() and not cs[1].any()
    # closure: the divergence of a constant velocity vanishes
    div = O.check_divergence(g, fc, giac)
    scale = np.abs(fc[0]).max()
    assert div[2] < 1e-13 * scale, div


def test_vel_up_closed_form():
    g = cases.grid(8, 6, 10, gamma=1.5, amp=0.1, ip=0, kp=0)
    rng = np.random.default_rng(11)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    fc = [rng.standard_normal(g.gshape()) for _ in range(3)]
    gra = [rng.standard_normal(g.gshape()) for _ in range(3)]
    cgra = [rng.standard_normal(g.gshape()) for _ in range(3)]
    v0 = [a.copy() for a in vel]
    f0 = [a.copy() for a in fc]
    dt = 0.3
    O.vel_up(g, vel, fc, gra, cgra, dt, 0)
    nx, ny, nz = g.nx, g.ny, g.nz
    I = np.s_[1:-1, 1:-1, 1:-1]
    for c in range(3):
        assert np.array_equal(vel[c][I], v0[c][I] + gra[c][I])
        assert np.array_equal(vel[c][1:nz + 1, :, 0], vel[c][1:nz + 1, :, nx])    # xi images, j = 0..ny+1
        assert np.array_equal(vel[c][0, :, 1:nx + 1], vel[c][nz, :, 1:nx + 1])    # zeta images
    # uc on faces 0..nx (periodic), vc on 1..ny-1 (eta walls untouched), wc on 0..nz
    assert np.allclose(fc[0][1:nz + 1, 1:ny + 1, 0:nx + 1], (f0[0] - dt * cgra[0])[1:nz + 1, 1:ny + 1, 0:nx + 1])
    assert np.allclose(fc[1][1:nz + 1, 1:ny, 1:nx + 1], (f0[1] - dt * cgra[1])[1:nz + 1, 1:ny, 1:nx + 1])
    assert np.array_equal(fc[1][1:nz + 1, 0, 1:nx + 1], f0[1][1:nz + 1, 0, 1:nx + 1])
    assert np.array_equal(fc[1][1:nz + 1, ny, 1:nx + 1], f0[1][1:nz + 1, ny, 1:nx + 1])
    assert np.allclose(fc[2][0:nz + 1, 1:ny + 1, 1:nx + 1], (f0[2] - dt * cgra[2])[0:nz + 1, 1:ny + 1, 1:nx + 1])


def test_check_divergence_brute_force():
    g = cases.grid(9, 7, 6, gamma=1.0, amp=0.1, ip=1, kp=1)
    rng = np.random.default_rng(2)
    fc = [rng.standard_normal(g.gshape()) for _ in range(3)]
    giac = 1.0 + rng.random(g.gshape())
    out = O.check_divergence(g, fc, giac)
    d = (fc[0][1:-1, 1:-1, 1:-1] - fc[0][1:-1, 1:-1, :-2] + fc[1][1:-1, 1:-1, 1:-1] - fc[1][1:-1, :-2, 1:-1]
         + fc[2][1:-1, 1:-1, 1:-1] - fc[2][:-2, 1:-1, 1:-1])
    dg = d / giac[1:-1, 1:-1, 1:-1]
    ins = np.s_[1:-1, 1:-1, 1:-1]
    assert out[0] == np.abs(d[ins]).max() and out[1] == np.abs(dg[ins]).max()
    assert out[2] == np.abs(d).max() and out[3] == np.abs(dg).max()


def oracle_projection_step(g, nlevel, vel, dvel, fc, fi, dt, eps, potenziale=0):
    """The reference sequence update -> contra -> diver -> multi -> gradie ->
    vel_up with the oracle (fsrc/stratiParticle.f03:2678-2937); in place."""
    gm, csv, giac = O.metrics(g, 0, want_csv=True, want_giac=True)
    O.update(g, vel, dvel)
    cs = O.contra_step(g, csv, vel, fc, potenziale)
    rhs = O.diver(g, fc[0], fc[1], fc[2], dt)
    ncyc, rmax, _ = O.multi(g, nlevel, fi, rhs, cs=cs, dt=dt, eps=eps, ficycle=400)
    outs = O.gradie(g, fi, csv, giac, gm, dt)
    O.vel_up(g, vel, fc, outs[:3], outs[3:], dt, potenziale)
    return cs, rhs, ncyc, rmax, gm, giac


def test_projection_chain_divergence_is_minus_dt_residual():
    g = cases.grid(16, 8, 16, gamma=1.5, amp=0.2)
    rng = np.random.default_rng(21)
    vel = [rng.standard_normal(g.gshape()) for _ in range(3)]
    dvel = [0.1 * rng.standard_normal(g.gshape()) for _ in range(3)]
    fc = [np.zeros(g.gshape()) for _ in range(3)]
    fi = g.zeros()
    dt = 0.5
    cs, rhs, ncyc, rmax, gm, giac = oracle_projection_step(g, 2, vel, dvel, fc, fi, dt, 1e-9)
    assert rmax < 1e-9
    # residual of the returned pressure (ghosts as the solve left them)
    r = O.resid(g, 0, fi, rhs, gm)
    ins = np.s_[2:-2, 2:-2, 2:-2]
    nx, ny, nz = g.nx, g.ny, g.nz
    d = (fc[0][1:-1, 1:-1, 1:-1] - fc[0][1:-1, 1:-1, :-2] + fc[1][1:-1, 1:-1, 1:-1] - fc[1][1:-1, :-2, 1:-1]
         + fc[2][1:-1, 1:-1, 1:-1] - fc[2][:-2, 1:-1, 1:-1])
    # cells away from the eta walls: div = -dt * r exactly up to round-off
    inner = np.s_[:, 1:-1, :]
    assert np.allclose(d[inner], -dt * r[1:-1, 1:-1, 1:-1][inner], rtol=0, atol=1e-12)
    div = O.check_divergence(g, fc, giac)
    assert div[0] <= dt * rmax * (1 + 1e-9) + 1e-13
    assert div[0] < 1e-8 * np.abs(rhs).max() * dt
