"""ctypes binding of the CPU oracle (oracle/ples_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared with.
Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this module.  Arrays use the ghosted layout described in
ples_oracle.h, held as numpy arrays of shape (nz+2, ny+2, nx+2) (C order,
i fastest), i.e. ``a[k, j, i]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libples_oracle.so")
_lib = None


class Dims(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ip", C.c_int), ("jp", C.c_int), ("kp", C.c_int)]


_dp = C.POINTER(C.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_make_nodes.argtypes = [C.c_int] * 3 + [C.c_double] * 5 + [_dp] * 3
        L.orc_gauss.argtypes = [C.c_uint64, C.c_uint32, C.c_int64]
        L.orc_gauss.restype = C.c_double
        L.orc_make_velocity.argtypes = [C.POINTER(Dims), C.c_uint64, _dp, _dp, _dp]
        L.orc_metrics.argtypes = [C.POINTER(Dims), _dp, _dp, _dp, C.c_int,
                                  C.POINTER(_dp), C.POINTER(_dp), _dp]
        L.orc_ghost_fill.argtypes = [C.POINTER(Dims), _dp, C.POINTER(_dp), C.POINTER(_dp),
                                     C.c_double, C.c_double, C.c_int]
        L.orc_sor_sweep.argtypes = [C.POINTER(Dims), _dp, _dp, C.POINTER(_dp), C.c_double, C.c_double]
        L.orc_resid.argtypes = [C.POINTER(Dims), _dp, _dp, C.POINTER(_dp), C.c_double, _dp]
        L.orc_restrict.argtypes = [C.POINTER(Dims), _dp, _dp]
        L.orc_prolong.argtypes = [C.POINTER(Dims), _dp, _dp]
        L.orc_multi.argtypes = [C.POINTER(Dims), C.c_int, _dp, _dp, _dp, C.POINTER(_dp), _dp, _dp,
                                C.POINTER(_dp), C.c_double, C.c_double, C.c_double, C.c_int,
                                C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_multi.restype = C.c_int
        L.orc_multi_ex.argtypes = [C.POINTER(Dims), C.c_int, _dp, _dp, _dp, C.POINTER(_dp), _dp, _dp,
                                   C.POINTER(_dp), C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_multi_ex.restype = C.c_int
        L.orc_tridag.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_sett.argtypes = [C.POINTER(Dims), C.POINTER(_dp), C.c_double, _dp, _dp, _dp]
        L.orc_slor_sweep.argtypes = [C.POINTER(Dims), _dp, _dp, C.POINTER(_dp), C.POINTER(_dp),
                                     C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_edge_fill.argtypes = [C.POINTER(Dims), _dp]
        _bind_hier(L)
        L.orc_contra.argtypes = [C.POINTER(Dims), C.POINTER(_dp), _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_diver.argtypes = [C.POINTER(Dims), _dp, _dp, _dp, C.c_double, _dp]
        L.orc_project.argtypes = [C.POINTER(Dims), _dp, C.POINTER(_dp), C.c_double]
        L.orc_project.restype = C.c_double
        L.orc_gradie.argtypes = [C.POINTER(Dims), _dp, C.POINTER(_dp), _dp, C.POINTER(_dp),
                                 C.c_double] + [_dp] * 6
        L.orc_update.argtypes = [C.POINTER(Dims), C.POINTER(_dp), C.POINTER(_dp)]
        L.orc_contra_step.argtypes = [C.POINTER(Dims), C.POINTER(_dp), C.POINTER(_dp), C.POINTER(_dp),
                                      C.POINTER(_dp), C.c_int]
        L.orc_vel_up.argtypes = [C.POINTER(Dims)] + [C.POINTER(_dp)] * 4 + [C.c_double, C.c_int]
        L.orc_check_divergence.argtypes = [C.POINTER(Dims), C.POINTER(_dp), _dp, _dp]
        for f in ("orc_absmax", "orc_l2", "orc_mean"):
            getattr(L, f).argtypes = [C.POINTER(Dims), _dp]
            getattr(L, f).restype = C.c_double
        _lib = L
    return _lib


def _bind_hier(L):
    L.orc_hier_create.argtypes = [C.POINTER(Dims), C.c_int, _dp, _dp, _dp, C.POINTER(_dp)]
    L.orc_hier_create.restype = C.c_void_p
    L.orc_hier_destroy.argtypes = [C.c_void_p]
    L.orc_hier_multi.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(_dp), C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.orc_hier_multi.restype = C.c_int


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _parr(arrs):
    if arrs is None:
        return None
    return (_dp * len(arrs))(*[_p(a) for a in arrs])


class Grid:
    """Finest-level dims + node coordinates (numpy (nz+1, ny+1, nx+1))."""

    def __init__(self, nx, ny, nz, ip=0, jp=1, kp=0, x=None, y=None, z=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.ip, self.jp, self.kp = ip, jp, kp
        self.x, self.y, self.z = x, y, z

    def dims(self, level: int = 0) -> Dims:
        st = 1 << level
        return Dims(self.nx // st, self.ny // st, self.nz // st, self.ip, self.jp, self.kp)

    def gshape(self, level: int = 0):
        d = self.dims(level)
        return (d.nz + 2, d.ny + 2, d.nx + 2)

    def zeros(self, level: int = 0):
        return np.zeros(self.gshape(level))


def make_grid(nx, ny, nz, Lx=2 * np.pi, Ly=2.0, Lz=np.pi, gamma=2.0, amp=0.0,
              ip=0, jp=1, kp=0) -> Grid:
    L = lib()
    sh = (nz + 1, ny + 1, nx + 1)
    x, y, z = np.zeros(sh), np.zeros(sh), np.zeros(sh)
    L.orc_make_nodes(nx, ny, nz, Lx, Ly, Lz, gamma, amp, _p(x), _p(y), _p(z))
    return Grid(nx, ny, nz, ip, jp, kp, x, y, z)


def metrics(g: Grid, level: int = 0, want_csv=False, want_giac=False):
    L = lib()
    d = g.dims(0)
    gm = [g.zeros(level) for _ in range(9)]
    csv = [g.zeros(level) for _ in range(9)] if want_csv else None
    giac = g.zeros(level) if want_giac else None
    L.orc_metrics(C.byref(d), _p(g.x), _p(g.y), _p(g.z), 1 << level, _parr(gm), _parr(csv),
                  _p(giac) if giac is not None else None)
    return gm, csv, giac


def velocity(g: Grid, seed: int):
    L = lib()
    u, v, w = g.zeros(), g.zeros(), g.zeros()
    d = g.dims(0)
    L.orc_make_velocity(C.byref(d), seed, _p(u), _p(v), _p(w))
    return u, v, w


def contra(g: Grid, csv, u, v, w):
    L = lib()
    uc, vc, wc = g.zeros(), g.zeros(), g.zeros()
    d = g.dims(0)
    L.orc_contra(C.byref(d), _parr(csv), _p(u), _p(v), _p(w), _p(uc), _p(vc), _p(wc))
    return uc, vc, wc


def diver(g: Grid, uc, vc, wc, dt=1.0):
    L = lib()
    rhs = g.zeros()
    d = g.dims(0)
    L.orc_diver(C.byref(d), _p(uc), _p(vc), _p(wc), dt, _p(rhs))
    return rhs


def project(g: Grid, rhs, cs=None, dt=1.0):
    d = g.dims(0)
    return lib().orc_project(C.byref(d), _p(rhs), _parr(cs), dt)


def case_rhs(g: Grid, seed: int, dt: float = 1.0):
    """Seeded divergence RHS: N(0,1) cell velocity -> contra fluxes (wall
    fluxes zero) -> diver -> zero-mean projection."""
    _, csv, _ = metrics(g, 0, want_csv=True)
    u, v, w = velocity(g, seed)
    uc, vc, wc = contra(g, csv, u, v, w)
    rhs = diver(g, uc, vc, wc, dt)
    project(g, rhs, None, dt)
    return rhs


def ghost_fill(g: Grid, level, p, gm, cs=None, an=0.0, dt=1.0, npass=5):
    d = g.dims(level)
    lib().orc_ghost_fill(C.byref(d), _p(p), _parr(gm), _parr(cs), an, dt, npass)


def sor_sweep(g: Grid, level, p, rh, gm, omega=1.6):
    d = g.dims(level)
    lib().orc_sor_sweep(C.byref(d), _p(p), _p(rh), _parr(gm), float(1 << level), omega)


def resid(g: Grid, level, p, rh, gm):
    d = g.dims(level)
    r = g.zeros(level)
    lib().orc_resid(C.byref(d), _p(p), _p(rh), _parr(gm), float(1 << level), _p(r))
    return r


def restrict(g: Grid, level_c, rf):
    d = g.dims(level_c)
    rc = g.zeros(level_c)
    lib().orc_restrict(C.byref(d), _p(rf), _p(rc))
    return rc


def prolong(g: Grid, level_c, ec, pf):
    d = g.dims(level_c)
    lib().orc_prolong(C.byref(d), _p(ec), _p(pf))


def multi(g: Grid, nlevel, fi, rhs, cs=None, dt=1.0, omega=1.6, eps=1e-8, ficycle=1000,
          kss1=5, kssc=10, nbc=5, g1=None, islor=0):
    d = g.dims(0)
    hist = np.zeros(max(ficycle, 1))
    rm = C.c_double(0.0)
    n = lib().orc_multi_ex(C.byref(d), nlevel, _p(g.x), _p(g.y), _p(g.z), _parr(g1), _p(fi), _p(rhs),
                           _parr(cs), dt, omega, eps, ficycle, kss1, kssc, nbc, islor, C.byref(rm), _p(hist))
    return n, rm.value, hist[:n]


def tridag(aa, bb, cc, ff):
    """Solve with the reference's Thomas algorithm; returns the solution
    (the inputs are copied, the reference overwrites its own)."""
    a, b, c, f = (np.array(v, dtype=np.float64) for v in (aa, bb, cc, ff))
    lib().orc_tridag(len(f), _p(a), _p(b), _p(c), _p(f))
    return f


def sett(g: Grid, level, gm):
    """Line coefficients aa, bb, cc (ghosted layout of the level)."""
    d = g.dims(level)
    aa, bb, cc = g.zeros(level), g.zeros(level), g.zeros(level)
    lib().orc_sett(C.byref(d), _parr(gm), float(1 << level), _p(aa), _p(bb), _p(cc))
    return aa, bb, cc


def slor_sweep(g: Grid, level, p, rh, gm, cs=None, an=0.0, dt=1.0, omega=1.6, nbc=5):
    d = g.dims(level)
    lib().orc_slor_sweep(C.byref(d), _p(p), _p(rh), _parr(gm), _parr(cs), an, dt, float(1 << level), omega, nbc)


def edge_fill(g: Grid, level, p):
    lib().orc_edge_fill(C.byref(g.dims(level)), _p(p))


class Hier:
    """orc_hier: the level hierarchy built once, then timed solves.  `path`
    selects the library (the OpenMP baseline build for timing)."""

    def __init__(self, g: Grid, nlevel, path=None):
        self.L = lib() if path is None else _load_baseline(path)
        self.g = g
        d = g.dims(0)
        self.h = self.L.orc_hier_create(C.byref(d), nlevel, _p(g.x), _p(g.y), _p(g.z), None)

    def multi(self, fi, rhs, cs=None, dt=1.0, omega=1.6, eps=1e-8, ficycle=1000, kss1=5, kssc=10,
              nbc=5, islor=0, nslab=1):
        hist = np.zeros(max(ficycle, 1))
        rm = C.c_double(0.0)
        n = self.L.orc_hier_multi(self.h, _p(fi), _p(rhs), _parr(cs), dt, omega, eps, ficycle, kss1, kssc,
                                  nbc, islor, nslab, C.byref(rm), _p(hist))
        if n < 0:
            raise ValueError(f"nslab={nslab} does not divide the planes of every level")
        return n, rm.value, hist[:n]

    def close(self):
        if self.h:
            self.L.orc_hier_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_baselines = {}


def build_baseline(out_dir=None) -> str:
    """Compile the OpenMP baseline build (-O3 -march=native) on this host."""
    out = out_dir or os.path.join(_HERE, "_build")
    subprocess.run(["make", "-s", "-C", _HERE, "baseline", f"OUT={out}"], check=True)
    return os.path.join(out, "libples_oracle_omp.so")


def _load_baseline(path):
    if path not in _baselines:
        L = C.CDLL(path)
        _bind_hier(L)
        _baselines[path] = L
    return _baselines[path]


def gradie(g: Grid, fi, csv, giac, gm, dt=1.0):
    d = g.dims(0)
    outs = [g.zeros() for _ in range(6)]
    lib().orc_gradie(C.byref(d), _p(fi), _parr(csv), _p(giac), _parr(gm), dt, *[_p(a) for a in outs])
    return outs


def absmax(g: Grid, r, level=0):
    d = g.dims(level)
    return lib().orc_absmax(C.byref(d), _p(r))


def l2(g: Grid, r, level=0):
    d = g.dims(level)
    return lib().orc_l2(C.byref(d), _p(r))


def mean(g: Grid, p, level=0):
    d = g.dims(level)
    return lib().orc_mean(C.byref(d), _p(p))


def update(g: Grid, vel, dvel):
    """In place on vel (3 ghosted arrays)."""
    lib().orc_update(C.byref(g.dims(0)), _parr(vel), _parr(dvel))


def contra_step(g: Grid, csv, vel, fc, potenziale=0):
    """In place on fc = [uc, vc, wc] (ghosted face layout); returns cs1..cs6
    in the reference shapes as numpy (C order: cs1 (nz, ny) ...)."""
    cs = [np.zeros((g.nz, g.ny)), np.zeros((g.nz, g.ny)), np.zeros((g.nz, g.nx)), np.zeros((g.nz, g.nx)),
          np.zeros((g.ny, g.nx)), np.zeros((g.ny, g.nx))]
    lib().orc_contra_step(C.byref(g.dims(0)), _parr(csv), _parr(vel), _parr(fc), _parr(cs), potenziale)
    return cs


def vel_up(g: Grid, vel, fc, gra, cgra, dt=1.0, potenziale=0):
    lib().orc_vel_up(C.byref(g.dims(0)), _parr(vel), _parr(fc), _parr(gra), _parr(cgra), dt, potenziale)


def check_divergence(g: Grid, fc, giac):
    out = np.zeros(4)
    lib().orc_check_divergence(C.byref(g.dims(0)), _parr(fc), _p(giac), _p(out))
    return out
