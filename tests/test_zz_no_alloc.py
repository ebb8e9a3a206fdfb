"""The no-allocation entry point: ples_mg_project_rhs_noalloc on device
arrays gives the results of ples_mg_project_rhs (host arrays through a
handle) on two calls in a row, the second on a smaller grid (the scratch
array keeps the first call's partial sums), to FP64 summation-order
accuracy: 1e-14 of the largest term of the sums."""
import ctypes

import numpy as np
import pytest

from ples_amd import mg
from tests import cases

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime (test plumbing)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p.value)
        assert self.L.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, p, shape):
        out = np.zeros(shape)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _case(shape, seed):
    g = cases.grid(*shape, gamma=1.5, amp=0.1)
    rng = np.random.default_rng(seed)
    rhs = rng.standard_normal((g.nz, g.ny, g.nx)) + 0.2
    cs = cases.random_cs(g, seed + 1, scale=0.05)   # eta walls: cs3, cs4; zero elsewhere
    return g, rhs, cs


def test_noalloc_projection_equals_the_handle_path_twice():
    hip = _Hip()
    try:
        for shape, seed in (((48, 24, 48), 3), ((32, 16, 16), 4)):   # the second call is smaller
            g, rhs, cs = _case(shape, seed)
            dt = 0.7
            ref = rhs.copy()
            with mg.Solver(g.x, g.y, g.z, 0, 1, 0, 2) as s:
                removed_ref = s.project_rhs(ref, cs, dt)
            d_rhs = hip.upload(rhs)
            d_cs = [hip.upload(c) if (q in (2, 3)) else None for q, c in enumerate(cs)]
            d_rem = hip.upload(np.zeros(1))
            mg.project_rhs_noalloc(d_rhs, g.nx, g.ny, g.nz, d_cs, dt, d_rem)
            got = hip.download(d_rhs, rhs.shape)
            removed = hip.download(d_rem, (1,))[0]
            scale = (np.abs(rhs).sum() + sum(np.abs(c).sum() for c in cs) / dt) / rhs.size
            assert abs(removed - removed_ref) <= 1e-14 * scale, (shape, removed, removed_ref)
            assert np.abs(got - ref).max() <= 1e-14 * scale + 1e-15 * np.abs(rhs).max(), shape
            # compatibility: the rhs integral equals the boundary-flux integral
            flux = (cs[3].sum() - cs[2].sum()) / dt
            assert abs(got.sum() - flux) <= 1e-12 * np.abs(rhs).sum()
    finally:
        hip.close()


def test_noalloc_rejects_bad_arguments():
    with pytest.raises(mg.PlesMGError):
        mg.project_rhs_noalloc(None, 8, 8, 8)
    with pytest.raises(mg.PlesMGError):
        mg.project_rhs_noalloc(1, 1 << 12, 1 << 12, 1 << 8)   # 2^32 cells: beyond the 32-bit indices
