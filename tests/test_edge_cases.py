"""Edge cases of the C ABI: argument and dimension errors (checked before any
device work, so they run without a GPU), and on the GPU: non-finite input,
zero rhs, a zero cycle cap, a single level, uneven and minimal grids, and the
reference's deeper k ghost layers of fi (deepl = deepr = 2)."""
import ctypes as C

import numpy as np
import pytest

from ples_amd import cases as PC
from ples_amd import mg
from tests import cases


def _nodes(nx, ny, nz, gamma=1.5, amp=0.1):
    return PC.make_nodes(nx, ny, nz, gamma, amp)


@pytest.mark.parametrize("dims,nlevel,jp,status", [
    ((30, 16, 32), 3, 1, 4),               # 30 not divisible by 4
    ((32, 4, 32), 3, 1, 4),                # one wall-normal cell on the coarsest level
    ((32, 16, 32), 17, 1, 4),              # too many levels
    ((32, 16, 32), 2, 0, 7),               # eta periodic: unsupported (as in the reference)
])
def test_create_rejects_bad_grids(dims, nlevel, jp, status):
    x, y, z = _nodes(*dims)
    with pytest.raises(mg.PlesMGError) as e:
        mg.Solver(x, y, z, 0, jp, 0, nlevel)
    assert e.value.status == status


def test_null_arguments_are_errors():
    L = mg.lib()
    assert L.ples_mg_create(None, None, None) == 6
    assert L.ples_mg_solve(None, None, 1, 1, None, None, None, None, None, None, None, 1.0, 1e-8, 10, None) == 6
    assert L.ples_mg_destroy(None) == 0
    g = mg._grid(*_nodes(32, 16, 32), 0, 1, 0, 3)
    o = mg._opts("fp64")
    o.omega = 2.5  # outside (0, 2)
    h = C.c_void_p()
    assert L.ples_mg_create(C.byref(g), C.byref(o), C.byref(h)) == 6
    assert mg.estimate_bytes(*_nodes(30, 16, 32), nlevel=3) == 0
    assert mg.estimate_bytes(*_nodes(32, 16, 32), nlevel=3) > 0


@pytest.mark.gpu
def test_nonfinite_rhs_reported():
    x, y, z = _nodes(32, 16, 32)
    with mg.Solver(x, y, z, 0, 1, 0, 3, precision="mixed") as s:
        rhs = np.zeros(s.rhs_shape)
        rhs[3, 4, 5] = np.nan
        fi = np.zeros(s.fi_shape)
        with pytest.raises(mg.PlesMGError) as e:
            s.multi(fi, rhs, eps=1e-8, ficycle=50)
        assert e.value.status == 2


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zero_rhs_and_zero_cycle_cap(precision):
    x, y, z = _nodes(32, 16, 32)
    with mg.Solver(x, y, z, 0, 1, 0, 3, precision=precision, stop_mode="abs_max") as s:
        fi = np.zeros(s.fi_shape)
        r = s.multi(fi, np.zeros(s.rhs_shape), eps=1e-12, ficycle=50)
        assert r.ok and r.cycles == 0 and not fi.any()
        rng = np.random.default_rng(3)
        rhs = rng.standard_normal(s.rhs_shape)
        r = s.multi(fi, rhs, eps=1e-12, ficycle=0)
        assert r.status == mg.NOT_CONVERGED and r.cycles == 0


@pytest.mark.gpu
@pytest.mark.parametrize("shape,nlevel,ip,kp", [
    ((8, 4, 8), 1, 0, 0),        # single level: the coarse solver is the whole solve
    ((48, 24, 16), 3, 0, 0),     # uneven dims
    ((12, 6, 20), 2, 1, 1),      # walls everywhere, smallest wall-normal extent
    ((96, 8, 16), 3, 0, 1),      # long and flat
])
def test_uneven_and_small_grids_match_oracle(shape, nlevel, ip, kp):
    from oracle import oracle as O
    g = cases.grid(*shape, gamma=1.0, amp=0.1, ip=ip, kp=kp)
    rhs = O.case_rhs(g, 11)
    fiv = g.zeros()
    _, rmax, _ = O.multi(g, max(1, min(nlevel, 2)), fiv, rhs, eps=1e-13, ficycle=3000)
    with mg.Solver(g.x, g.y, g.z, ip, 1, kp, nlevel, precision="fp64", stop_mode="abs_max") as s:
        fi = np.zeros(s.fi_shape)
        r = s.multi(fi, cases.interior(rhs), eps=1e-13, ficycle=400)
        assert r.ok, r
    assert rmax < 1e-13
    assert cases.rel_l2_meanfree(fi, fiv) < 1e-9


@pytest.mark.gpu
def test_deeper_k_ghost_layers():
    x, y, z = _nodes(32, 16, 32)
    rng = np.random.default_rng(5)
    with mg.Solver(x, y, z, 0, 1, 0, 3, precision="fp64") as s:
        rhs = rng.standard_normal(s.rhs_shape)
        rhs -= rhs.mean()
        fi1 = np.zeros(s.fi_shape)
        s.multi(fi1, rhs, eps=1e-10, ficycle=100)
        fi2 = np.full((s.jz + 4, s.jy + 2, s.jx + 2), 7.0)
        fi2[1:-1] = 0.0
        s.multi(fi2, rhs, eps=1e-10, ficycle=100, deepl=2, deepr=2)
    assert np.array_equal(fi2[1:-1], fi1)
    assert (fi2[0] == 7.0).all() and (fi2[-1] == 7.0).all()  # planes outside 0..jz+1 untouched
