"""Kernel-level GPU parity: one restriction, one prolongation, one residual
and one smoother sweep of a multigrid level (ples_mg_level_op) against the
oracle on the same inputs.

  * restriction (k_restrict, fsrc/restrict_sndrcv.f03:32-53), prolongation
    (k_prolong_blk, fsrc/prolong.f03:64-261), the residual (k_resid,
    fsrc/resid_sndrcv.f03:80-151) and the residual fused with the
    restriction (k_resid_restrict) against orc_restrict, orc_prolong and
    orc_resid: FP64 to 1e-12 relative (summation order only);
  * the eta-line smoother (k_yline, every segment count P the level's row
    count allows, register and staged variants): one sweep with omega = 1
    is the line-Jacobi step p + M^-1 (A p - b), M the column tridiagonal of
    -s^2 g22 couplings and the non-wall normal metrics, solved here with
    scipy's banded solver per column.  The smoother stores its line factors
    in FP32 (it is a preconditioner; the operator and residuals stay FP64),
    so an FP64 sweep matches the exact step to 1e-5 of the update's size
    (measured: 1e-6);
    the eta-wall ghosts it writes satisfy the homogeneous flux relation with
    the input's neighbours (1e-12), and it writes the periodic images.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded

from oracle import oracle as O
from ples_amd import mg
from tests import cases

pytestmark = pytest.mark.gpu


def _solver(g, nlevel, precision="fp64", **kw):
    return mg.Solver(g.x, g.y, g.z, g.ip, g.jp, g.kp, nlevel, precision=precision, autotune=-1, **kw)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


# walls in xi or zeta with a mild terrain (steep terrain against xi/zeta walls
# has no wall-ghost fixed point: create refuses it)
GRIDS = [((32, 16, 32), 0, 0, 3, 1.5, 0.15), ((24, 16, 16), 1, 1, 3, 1.0, 0.1), ((16, 16, 32), 1, 0, 3, 1.5, 0.05)]


@pytest.mark.parametrize("shape,ip,kp,L,gamma,amp", GRIDS)
def test_transfers_and_residual_match_oracle(shape, ip, kp, L, gamma, amp):
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    rng = np.random.default_rng(17)
    with _solver(g, L) as s:
        for lev in range(1, L):
            gm, _, _ = O.metrics(g, lev - 1)
            pot = float(1 << (lev - 1))
            fine = g.gshape(lev - 1)
            coarse = g.gshape(lev)
            assert s.level_shape(lev) == (fine[2] - 2, fine[1] - 2, fine[0] - 2)
            # restriction of a residual (interior children only are read)
            rf = g.zeros(lev - 1)
            rf[1:-1, 1:-1, 1:-1] = rng.standard_normal((fine[0] - 2, fine[1] - 2, fine[2] - 2))
            rc = s.level_op(lev, mg.OP_RESTRICT, rf)
            ref = O.restrict(g, lev, rf)
            assert _rel(rc[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) < 1e-14
            # prolongation-and-add over the ghost-inclusive range
            ec = rng.standard_normal(coarse)
            pf = rng.standard_normal(fine)
            out = s.level_op(lev, mg.OP_PROLONG, ec, pf)
            ref = pf.copy()
            O.prolong(g, lev, ec, ref)
            assert _rel(out, ref) < 1e-14
            # residual and residual fused with the restriction
            p = rng.standard_normal(fine)
            b = g.zeros(lev - 1)
            b[1:-1, 1:-1, 1:-1] = rng.standard_normal((fine[0] - 2, fine[1] - 2, fine[2] - 2))
            r = s.level_op(lev, mg.OP_RESID, p, b)
            r_ref = O.resid(g, lev - 1, p, b, gm)
            assert _rel(r[1:-1, 1:-1, 1:-1], r_ref[1:-1, 1:-1, 1:-1]) < 1e-12
            rc = s.level_op(lev, mg.OP_RESID_RESTRICT, p, b)
            ref = O.restrict(g, lev, r_ref)
            assert _rel(rc[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) < 1e-12
            del pot


def _line_step(g, lev, gm, p, b, omega):
    """p + omega M^-1 (A p - b) in the interior, column by column."""
    d = g.dims(lev)
    nx, ny, nz = d.nx, d.ny, d.nz
    s2 = 1.0 / float(1 << lev) ** 2
    r = O.resid(g, lev, p, b, gm)
    out = p.copy()
    g11, g22, g33 = gm[0], gm[4], gm[8]
    for k in range(1, nz + 1):
        for i in range(1, nx + 1):
            j = np.arange(1, ny + 1)
            g11p = np.zeros(ny) if (d.ip == 1 and i == nx) else g11[k, 1:-1, i]
            g11m = np.zeros(ny) if (d.ip == 1 and i == 1) else g11[k, 1:-1, i - 1]
            g33p = np.zeros(ny) if (d.kp == 1 and k == nz) else g33[k, 1:-1, i]
            g33m = np.zeros(ny) if (d.kp == 1 and k == 1) else g33[k - 1, 1:-1, i]
            g22p = np.where(j == ny, 0.0, g22[k, 1:-1, i])
            g22m = np.where(j == 1, 0.0, g22[k, 0:-2, i])
            den = s2 * (g11p + g11m + g22p + g22m + g33p + g33m)
            ab = np.zeros((3, ny))
            ab[0, 1:] = -s2 * g22[k, 1:ny, i]      # super-diagonal c_j = -s^2 g22(j)
            ab[1] = den
            ab[2, :-1] = -s2 * g22[k, 1:ny, i]     # sub-diagonal a_{j+1} = -s^2 g22(j)
            out[k, 1:-1, i] = p[k, 1:-1, i] + omega * solve_banded((1, 1), ab, r[k, 1:-1, i])
    return out


def _seg_counts(ny):
    return [P for P in (1, 2, 3, 4, 6, 8, 12, 16) if ny % P == 0 and ny // P >= 4]


@pytest.mark.parametrize("shape,ip,kp,lev,gamma,amp", [
    ((16, 48, 8), 0, 0, 1, 1.5, 0.15), ((16, 64, 8), 0, 0, 1, 1.5, 0.15),
    ((16, 48, 16), 1, 1, 1, 1.0, 0.02), ((16, 96, 16), 0, 0, 2, 1.5, 0.15)])
def test_line_smoother_sweep_is_the_line_jacobi_step(shape, ip, kp, lev, gamma, amp):
    g = cases.grid(*shape, gamma=gamma, amp=amp, ip=ip, kp=kp)
    gm, _, _ = O.metrics(g, lev - 1)
    rng = np.random.default_rng(23)
    sh = g.gshape(lev - 1)
    p = rng.standard_normal(sh)
    if ip == 0:  # periodic images consistent with the interior, as a sweep leaves them
        p[:, :, 0], p[:, :, -1] = p[:, :, -2], p[:, :, 1]
    if kp == 0:
        p[0], p[-1] = p[-2], p[1]
    b = g.zeros(lev - 1)
    b[1:-1, 1:-1, 1:-1] = rng.standard_normal((sh[0] - 2, sh[1] - 2, sh[2] - 2))
    ref = _line_step(g, lev - 1, gm, p, b, 1.0)
    upd = np.abs(ref - p)[1:-1, 1:-1, 1:-1].max()
    ny = sh[1] - 2
    nx, nz = sh[2] - 2, sh[0] - 2
    with _solver(g, lev + 1) as s:
        for P in _seg_counts(ny):
            out = s.level_op(lev, mg.OP_SMOOTH, p, b, P=P, omega=1.0)
            err = np.abs(out - ref)[1:-1, 1:-1, 1:-1].max() / upd
            assert err < 1e-5, (P, err)
            # eta-wall ghosts: homogeneous flux relation with the input's neighbours
            q = 0.25
            lo = gm[4][1:-1, 0, 1:-1] * (out[1:-1, 1, 1:-1] - out[1:-1, 0, 1:-1]) + q * (
                gm[3][1:-1, 0, 1:-1] * (p[1:-1, 1, 2:] + p[1:-1, 0, 2:] - p[1:-1, 1, :-2] - p[1:-1, 0, :-2]) +
                gm[5][1:-1, 0, 1:-1] * (p[2:, 1, 1:-1] + p[2:, 0, 1:-1] - p[:-2, 1, 1:-1] - p[:-2, 0, 1:-1]))
            hi = gm[4][1:-1, ny, 1:-1] * (out[1:-1, ny + 1, 1:-1] - out[1:-1, ny, 1:-1]) + q * (
                gm[3][1:-1, ny, 1:-1] * (p[1:-1, ny + 1, 2:] + p[1:-1, ny, 2:] - p[1:-1, ny + 1, :-2] -
                                         p[1:-1, ny, :-2]) +
                gm[5][1:-1, ny, 1:-1] * (p[2:, ny + 1, 1:-1] + p[2:, ny, 1:-1] - p[:-2, ny + 1, 1:-1] -
                                         p[:-2, ny, 1:-1]))
            scale = np.abs(gm[4][1:-1, 0, 1:-1]).max() * np.abs(out).max()
            if ip == 0 and kp == 0:
                assert np.abs(lo).max() < 1e-12 * scale and np.abs(hi).max() < 1e-12 * scale, P
            # periodic images written by the sweep
            if ip == 0:
                assert np.array_equal(out[1:-1, :, 0], out[1:-1, :, nx]), P
                assert np.array_equal(out[1:-1, :, nx + 1], out[1:-1, :, 1]), P
            if kp == 0:
                assert np.array_equal(out[0, :, 1:-1], out[nz, :, 1:-1]), P
    # omega scales the update (same segments)
    with _solver(g, lev + 1) as s:
        P = _seg_counts(ny)[0]
        full = s.level_op(lev, mg.OP_SMOOTH, p, b, P=P, omega=1.0)
        half = s.level_op(lev, mg.OP_SMOOTH, p, b, P=P, omega=0.5)
        assert np.abs((half - p) - 0.5 * (full - p))[1:-1, 1:-1, 1:-1].max() < 1e-12 * upd


def test_fp32_level_ops_match_oracle():
    """The mixed solver's FP32 levels (FP32 metrics in the smoother): the same
    operations to FP32 accuracy."""
    g = cases.grid(16, 48, 16, gamma=1.5, amp=0.15)
    gm, _, _ = O.metrics(g, 0)
    rng = np.random.default_rng(31)
    sh = g.gshape(0)
    p = rng.standard_normal(sh)
    p[:, :, 0], p[:, :, -1] = p[:, :, -2], p[:, :, 1]
    p[0], p[-1] = p[-2], p[1]
    b = g.zeros()
    b[1:-1, 1:-1, 1:-1] = rng.standard_normal((sh[0] - 2, sh[1] - 2, sh[2] - 2))
    with _solver(g, 3, precision="mixed", smoother_fp16=0) as s:
        ref = _line_step(g, 0, gm, p, b, 1.0)
        upd = np.abs(ref - p)[1:-1, 1:-1, 1:-1].max()
        for P in _seg_counts(48):
            out = s.level_op(1, mg.OP_SMOOTH, p, b, P=P, omega=1.0)
            assert np.abs(out - ref)[1:-1, 1:-1, 1:-1].max() / upd < 1e-4, P
        r_ref = O.resid(g, 0, p, b, gm)
        rc = s.level_op(1, mg.OP_RESID_RESTRICT, p, b)
        ref = O.restrict(g, 1, r_ref)
        assert _rel(rc[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]) < 1e-5
        ec = rng.standard_normal(g.gshape(1))
        pf = rng.standard_normal(sh)
        out = s.level_op(1, mg.OP_PROLONG, ec, pf)
        ref = pf.copy()
        O.prolong(g, 1, ec, ref)
        assert _rel(out, ref) < 1e-6


def test_bandwidth_probes():
    """The roofline's measured ceilings: copy, read-only and twelve-stream
    probes return rates below the HBM peak, the twelve-stream one (4-byte
    accesses) at most the read-only one (16-byte accesses) plus noise."""
    nbytes = 256 << 20
    copy = mg.copy_bandwidth(nbytes, 5)
    read = mg.read_bandwidth(nbytes, 5)
    streams = mg.streams_bandwidth(nbytes, 5)
    for v in (copy, read, streams):
        assert 100.0 < v < 8000.0 * 1.05, (copy, read, streams)
    assert streams < 1.1 * read, (streams, read)
