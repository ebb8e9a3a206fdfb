"""Python host binding of libples_mg.so (the MI355X pressure multigrid).

Mirrors the reference's solver interface: ``Solver.multi(fi, rhs, cs, dt,
eps, ficycle)`` plays the role of ``call multi(eps, ficycle, nlevel, ...)``
(fsrc/multigrid.f03:42-48, called at fsrc/stratiParticle.f03:2881-2887):
the pressure ``fi`` is the warm start and is overwritten with the solution
including its ghost layers.

Arrays are numpy float64 in the reference's Fortran memory order, seen from
numpy as C-order arrays with reversed axes:
  fi   (jz+2, jy+2, jx+2)    == fi(0:jx+1, 0:jy+1, 0:jz+1)
  rhs  (jz, jy, jx)          == rhs(1:jx, 1:jy, 1:jz)
  g11..g13 (jz, jy, jx+1), g21..g23 (jz, jy+1, jx), g31..g33 (jz+1, jy, jx)
  cs1, cs2 (jz, jy); cs3, cs4 (jz, jx); cs5, cs6 (jy, jx)
  nodes x, y, z (jz+1, jy+1, jx+1)

There is no CPU fallback: if the HIP library is missing or no GPU is
present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libples_mg.so")

OK = 0
NOT_CONVERGED = 1
BUSY = 8
FP64 = 0
MIXED = 1
STOP_REL_L2 = 0
STOP_ABS_MAX = 1

OP_SMOOTH, OP_RESID, OP_RESTRICT, OP_RESID_RESTRICT, OP_PROLONG = range(5)

K_SMOOTH, K_SMOOTH_FINE, K_RESID, K_RESID_FINE, K_OUTER_APPLY, K_OUTER_UPDATE, K_TRANSFER, K_HALO = range(8)
KERNEL_FAMILIES = {
    "smooth": K_SMOOTH, "smooth_fine": K_SMOOTH_FINE, "resid": K_RESID, "resid_fine": K_RESID_FINE,
    "outer_apply": K_OUTER_APPLY, "outer_update": K_OUTER_UPDATE, "transfer": K_TRANSFER, "halo": K_HALO,
}

_dp = C.POINTER(C.c_double)


class _Grid(C.Structure):
    _fields_ = [("jx", C.c_int), ("jy", C.c_int), ("jz", C.c_int),
                ("ip", C.c_int), ("jp", C.c_int), ("kp", C.c_int), ("nlevel", C.c_int),
                ("x", _dp), ("y", _dp), ("z", _dp), ("g", _dp * 9)]


class _Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("nu_pre", C.c_int), ("nu_post", C.c_int),
                ("nu_coarse", C.c_int), ("omega", C.c_double), ("krylov", C.c_int),
                ("stop_mode", C.c_int), ("project_rhs", C.c_int), ("halo_passes", C.c_int),
                ("device", C.c_int), ("graphs", C.c_int), ("inner_cycles", C.c_int),
                ("autotune", C.c_int), ("fp32_krylov", C.c_int), ("smoother_fp16", C.c_int),
                ("coarse_exact", C.c_int), ("sweep_schedule", C.c_int), ("semicoarsen", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int), ("cycles", C.c_int), ("rel_res", C.c_double),
                ("abs_res", C.c_double), ("rhs_mean_removed", C.c_double), ("seconds", C.c_double)]


class _ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("warm_cases", C.c_int),
                ("device", C.c_int)]


@dataclass
class Result:
    status: int
    cycles: int
    rel_res: float
    abs_res: float
    rhs_mean_removed: float
    seconds: float

    @property
    def ok(self) -> bool:
        return self.status == OK


class PlesMGError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} ({status})")


_lib = None


def lib():
    """Load libples_mg.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ples_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.ples_mg_default_opts.argtypes = [C.POINTER(_Opts)]
        L.ples_mg_status_string.argtypes = [C.c_int]
        L.ples_mg_status_string.restype = C.c_char_p
        L.ples_mg_create.argtypes = [C.POINTER(_Grid), C.POINTER(_Opts), C.POINTER(C.c_void_p)]
        L.ples_mg_destroy.argtypes = [C.c_void_p]
        L.ples_mg_solve.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp] + [_dp] * 6 + \
            [C.c_double, C.c_double, C.c_int, C.POINTER(_Result)]
        L.ples_mg_residual.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp] + [_dp] * 6 + \
            [C.c_double, _dp]
        L.ples_mg_metrica.argtypes = [C.c_void_p, C.POINTER(_dp), C.POINTER(_dp), _dp]
        L.ples_mg_level_metrics.argtypes = [C.c_void_p, C.c_int, C.POINTER(_dp)]
        L.ples_mg_level_shape.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 3
        L.ples_mg_diver.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, _dp]
        L.ples_mg_gradie.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double] + [_dp] * 6
        L.ples_mg_update.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ples_mg_contra.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(_dp), C.c_int]
        L.ples_mg_vel_up.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(_dp), C.POINTER(_dp), C.c_double, C.c_int]
        L.ples_mg_check_divergence.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.ples_mg_projection_step.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp,
                                              _dp, _dp, _dp, C.POINTER(_dp), _dp, C.c_double, C.c_double,
                                              C.c_int, C.c_int, C.POINTER(_Result), _dp]
        L.ples_mg_batch_create.argtypes = [C.c_int, C.POINTER(_Grid), C.POINTER(_Opts), C.POINTER(C.c_void_p)]
        L.ples_mg_batch_destroy.argtypes = [C.c_void_p]
        L.ples_mg_batch_size.argtypes = [C.c_void_p]
        L.ples_mg_batch_handle.argtypes = [C.c_void_p, C.c_int]
        L.ples_mg_batch_handle.restype = C.c_void_p
        L.ples_mg_batch_device_bytes.argtypes = [C.c_void_p]
        L.ples_mg_batch_device_bytes.restype = C.c_size_t
        L.ples_mg_batch_solve.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(_Result)]
        L.ples_mg_batch_set_grouped.argtypes = [C.c_void_p, C.c_int]
        L.ples_mg_batch_families.argtypes = [C.c_void_p]
        L.ples_mg_tiles.argtypes = [C.c_void_p, C.c_int] + [C.POINTER(C.c_int)] * 3
        L.ples_mg_cycle_levels.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.ples_mg_copy_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.ples_mg_read_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.ples_mg_streams_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.ples_mg_tuning_load.argtypes = [C.c_char_p]
        L.ples_mg_restart_write.argtypes = [C.c_char_p] + [C.c_int] * 4 + [C.c_double] * 3 + [_dp] * 8
        L.ples_mg_restart_read.argtypes = [C.c_char_p] + [C.c_int] * 4 + [_dp] * 3 + [_dp] * 8
        L.ples_mg_tuning_save.argtypes = [C.c_char_p]
        L.ples_mg_tuning_clear.argtypes = []
        L.ples_mg_project_rhs.argtypes = [C.c_void_p, _dp] + [_dp] * 6 + [C.c_double, _dp]
        L.ples_mg_project_rhs_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 6 + \
            [C.c_double, C.c_void_p, C.c_void_p]
        L.ples_mg_level_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp]
        L.ples_mg_server_default_opts.argtypes = [C.POINTER(_ServerOpts)]
        L.ples_mg_server_create.argtypes = [C.POINTER(_ServerOpts), C.POINTER(_Opts), C.POINTER(C.c_void_p)]
        L.ples_mg_server_destroy.argtypes = [C.c_void_p]
        L.ples_mg_server_add_case.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(_Grid), C.c_uint64, C.c_double]
        L.ples_mg_server_submit.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                            C.POINTER(C.c_int64)]
        L.ples_mg_server_step.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.ples_mg_server_wait.argtypes = [C.c_void_p]
        L.ples_mg_server_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(_Result), C.POINTER(C.c_int)]
        L.ples_mg_server_pending.argtypes = [C.c_void_p]
        L.ples_mg_server_get_fi.argtypes = [C.c_void_p, C.c_uint64, _dp]
        L.ples_mg_server_set_fi.argtypes = [C.c_void_p, C.c_uint64, _dp]
        L.ples_mg_server_stats.argtypes = [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                           C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.ples_mg_case_rhs.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
        L.ples_mg_zero_fi.argtypes = [C.c_void_p]
        L.ples_mg_perturb_fi.argtypes = [C.c_void_p, C.c_uint64, C.c_double]
        L.ples_mg_solve_resident.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(_Result)]
        for f in ("ples_mg_get_fi", "ples_mg_set_fi", "ples_mg_get_rhs", "ples_mg_set_rhs"):
            getattr(L, f).argtypes = [C.c_void_p, _dp]
        L.ples_mg_stream.argtypes = [C.c_void_p]
        L.ples_mg_stream.restype = C.c_void_p
        L.ples_mg_device_bytes.argtypes = [C.c_void_p]
        L.ples_mg_device_bytes.restype = C.c_size_t
        L.ples_mg_estimate_bytes.argtypes = [C.POINTER(_Grid), C.POINTER(_Opts)]
        L.ples_mg_estimate_bytes.restype = C.c_size_t
        L.ples_mg_set_timing.argtypes = [C.c_void_p, C.c_int]
        L.ples_mg_set_timing_every.argtypes = [C.c_void_p, C.c_int]
        L.ples_mg_kernel_stats.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_long),
                                           C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def status_string(s: int) -> str:
    return lib().ples_mg_status_string(s).decode()


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(_dp)


def _need(a, shape, what):
    """Raise unless a is a float64 array of exactly this shape: the C ABI
    reads and writes caller memory by the reference shapes."""
    if a is None:
        return
    if tuple(np.shape(a)) != tuple(shape):
        raise ValueError(f"{what}: shape {tuple(np.shape(a))}, expected {tuple(shape)}")


def _check(st: int, what: str):
    if st != OK:
        raise PlesMGError(st, what)


def default_opts() -> dict:
    o = _Opts()
    lib().ples_mg_default_opts(C.byref(o))
    return {f: getattr(o, f) for f, _ in _Opts._fields_}


def _opts(precision="fp64", **kw) -> _Opts:
    o = _Opts()
    lib().ples_mg_default_opts(C.byref(o))
    o.precision = {"fp64": FP64, "mixed": MIXED}[precision] if isinstance(precision, str) else int(precision)
    for k, v in kw.items():
        if k == "stop_mode" and isinstance(v, str):
            v = {"rel_l2": STOP_REL_L2, "abs_max": STOP_ABS_MAX}[v]
        setattr(o, k, v)
    return o


def _grid(x, y, z, ip, jp, kp, nlevel, metrics=None):
    jz, jy, jx = (s - 1 for s in x.shape)
    g = _Grid(jx, jy, jz, ip, jp, kp, nlevel, _p(x), _p(y), _p(z))
    if metrics is not None:
        for q in range(9):
            g.g[q] = _p(metrics[q])
    return g


def copy_bandwidth(nbytes=1 << 30, reps=20, device=0) -> float:
    """Measured device copy bandwidth in GB/s (read + write)."""
    g = C.c_double()
    _check(lib().ples_mg_copy_bandwidth(device, nbytes, reps, C.byref(g)), "ples_mg_copy_bandwidth")
    return g.value


def project_rhs_noalloc(rhs_dev, jx, jy, jz, cs_dev=None, dt=1.0, removed_dev=None, stream=None):
    """ples_mg_project_rhs_noalloc on device pointers (ints): the projection of
    a device rhs with no device allocation (enqueued on `stream`)."""
    cs = list(cs_dev) if cs_dev is not None else [None] * 6
    _check(lib().ples_mg_project_rhs_noalloc(rhs_dev, jx, jy, jz, *cs, dt, removed_dev, stream),
           "ples_mg_project_rhs_noalloc")


def read_bandwidth(nbytes=1 << 30, reps=20, device=0) -> float:
    """Measured device read bandwidth in GB/s (read-only sweep)."""
    g = C.c_double()
    _check(lib().ples_mg_read_bandwidth(device, nbytes, reps, C.byref(g)), "ples_mg_read_bandwidth")
    return g.value


def streams_bandwidth(nbytes=1 << 30, reps=20, device=0) -> float:
    """Measured GB/s of eleven 4-byte FP32 read streams and one written stream."""
    g = C.c_double()
    _check(lib().ples_mg_streams_bandwidth(device, nbytes, reps, C.byref(g)), "ples_mg_streams_bandwidth")
    return g.value


def restart_write(path, ti, bbx, dbbx, u, v, w, fi, rhov, uc, vc, wc):
    """Write the reference's restart file (old_res_form layout).  Arrays in C
    order: u, v, w, fi (jz+2, jy+2, jx+2); rhov (jz+2, jy+2, jx+2, nscal);
    uc (jz, jy, jx+1), vc (jz, jy+1, jx), wc (jz+1, jy, jx)."""
    jz, jy, jx = (d - 2 for d in u.shape)
    nscal = rhov.shape[-1] if rhov is not None else 0
    arrs = [np.ascontiguousarray(a, np.float64) for a in (u, v, w, fi)]
    rh = np.ascontiguousarray(rhov, np.float64) if nscal else None
    fl = [np.ascontiguousarray(a, np.float64) for a in (uc, vc, wc)]
    _check(lib().ples_mg_restart_write(path.encode(), jx, jy, jz, nscal, ti, bbx, dbbx, *[_p(a) for a in arrs],
                                       _p(rh) if rh is not None else None, *[_p(a) for a in fl]),
           "ples_mg_restart_write")


def restart_read(path, jx, jy, jz, nscal):
    """Read a restart file; returns (ti, bbx, dbbx, u, v, w, fi, rhov, uc, vc, wc)."""
    cells = (jz + 2, jy + 2, jx + 2)
    u, v, w, fi = (np.zeros(cells) for _ in range(4))
    rhov = np.zeros(cells + (nscal,)) if nscal else None
    uc, vc, wc = np.zeros((jz, jy, jx + 1)), np.zeros((jz, jy + 1, jx)), np.zeros((jz + 1, jy, jx))
    t, b1, b2 = C.c_double(), C.c_double(), C.c_double()
    _check(lib().ples_mg_restart_read(path.encode(), jx, jy, jz, nscal, C.byref(t), C.byref(b1), C.byref(b2),
                                      *[_p(a) for a in (u, v, w, fi)], _p(rhov) if rhov is not None else None,
                                      *[_p(a) for a in (uc, vc, wc)]), "ples_mg_restart_read")
    return t.value, b1.value, b2.value, u, v, w, fi, rhov, uc, vc, wc


TUNING_TABLE = os.path.join(_HERE, "tuning", "mi355x.txt")


def tuning_load(path=TUNING_TABLE) -> bool:
    """Load a launch-shape table for autotune=1 handles; False if absent."""
    if not os.path.exists(path):
        return False
    return lib().ples_mg_tuning_load(path.encode()) == OK


def tuning_clear():
    """Empty the process-wide launch-shape table (fixed rule from now on)."""
    _check(lib().ples_mg_tuning_clear(), "ples_mg_tuning_clear")


def tuning_save(path):
    _check(lib().ples_mg_tuning_save(path.encode()), "ples_mg_tuning_save")


def estimate_bytes(x, y, z, ip=0, jp=1, kp=0, nlevel=4, precision="fp64", **kw) -> int:
    g = _grid(x, y, z, ip, jp, kp, nlevel)
    return int(lib().ples_mg_estimate_bytes(C.byref(g), C.byref(_opts(precision, **kw))))


class Solver:
    """Device-resident multigrid for one grid (one handle, one HIP stream)."""

    @classmethod
    def _borrow(cls, handle, x, y, z, ip, jp, kp, nlevel):
        """A view of a handle owned elsewhere (a batch): never destroyed here."""
        s = cls.__new__(cls)
        s._x, s._y, s._z = x, y, z
        s.jz, s.jy, s.jx = (d - 1 for d in x.shape)
        s.ip, s.jp, s.kp, s.nlevel = ip, jp, kp, nlevel
        s._metrics = None
        s._h = C.c_void_p(handle)
        s._owned = False
        return s

    def __init__(self, x, y, z, ip=0, jp=1, kp=0, nlevel=4, metrics=None, precision="fp64", **opts):
        self._owned = True
        self._x, self._y, self._z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self.jz, self.jy, self.jx = (s - 1 for s in self._x.shape)
        self.ip, self.jp, self.kp, self.nlevel = ip, jp, kp, nlevel
        self._metrics = None if metrics is None else [np.ascontiguousarray(m, np.float64) for m in metrics]
        if self._metrics is not None:
            if len(self._metrics) != 9:
                raise ValueError("metrics: nine arrays g11..g33")
            for q, m in enumerate(self._metrics):
                _need(m, self.face_shapes()[q // 3], f"metrics[{q}]")
        g = _grid(self._x, self._y, self._z, ip, jp, kp, nlevel, self._metrics)
        o = _opts(precision, **opts)
        h = C.c_void_p()
        _check(lib().ples_mg_create(C.byref(g), C.byref(o), C.byref(h)), "ples_mg_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and getattr(self, "_owned", True):
            lib().ples_mg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- shapes
    @property
    def fi_shape(self):
        return (self.jz + 2, self.jy + 2, self.jx + 2)

    @property
    def rhs_shape(self):
        return (self.jz, self.jy, self.jx)

    def cs_shapes(self):
        return [(self.jz, self.jy)] * 2 + [(self.jz, self.jx)] * 2 + [(self.jy, self.jx)] * 2

    def _cs(self, cs):
        if cs is None:
            return [None] * 6
        if len(cs) != 6:
            raise ValueError("cs: six arrays cs1..cs6")
        out = [None if c is None else np.ascontiguousarray(c, np.float64) for c in cs]
        for q, c in enumerate(out):
            _need(c, self.cs_shapes()[q], f"cs{q + 1}")
        return out

    # ---- drop-in
    def multi(self, fi, rhs, cs=None, dt=1.0, eps=1e-8, ficycle=1000, deepl=1, deepr=1) -> Result:
        """Solve in place (fi = warm start -> solution with ghosts).  fi holds
        k planes 1-deepl .. jz+deepr as the reference allocates it."""
        assert fi.shape == (self.jz + deepl + deepr, self.jy + 2, self.jx + 2) and rhs.shape == self.rhs_shape
        assert fi.flags["C_CONTIGUOUS"] and fi.dtype == np.float64
        cs = self._cs(cs)
        r = _Result()
        rhs = np.ascontiguousarray(rhs, np.float64)
        st = lib().ples_mg_solve(self._h, _p(fi), deepl, deepr, _p(rhs), *[_p(c) for c in cs], dt, eps,
                                 ficycle, C.byref(r))
        if st not in (OK, NOT_CONVERGED):
            raise PlesMGError(st, "ples_mg_solve")
        return Result(r.status, r.cycles, r.rel_res, r.abs_res, r.rhs_mean_removed, r.seconds)

    def residual(self, fi, rhs, cs=None, dt=1.0):
        _need(fi, self.fi_shape, "fi")
        _need(rhs, self.rhs_shape, "rhs")
        cs = self._cs(cs)
        out = np.zeros(self.rhs_shape)
        _check(lib().ples_mg_residual(self._h, _p(np.ascontiguousarray(fi)), 1, 1,
                                      _p(np.ascontiguousarray(rhs, np.float64)),
                                      *[_p(c) for c in cs], dt, _p(out)), "ples_mg_residual")
        return out

    def metrica(self):
        """Level-1 metrics g[9], face vectors csv[9], giac in the reference shapes."""
        shapes = [(self.jz, self.jy, self.jx + 1)] * 3 + [(self.jz, self.jy + 1, self.jx)] * 3 + \
                 [(self.jz + 1, self.jy, self.jx)] * 3
        g = [np.zeros(s) for s in shapes]
        v = [np.zeros(s) for s in shapes]
        giac = np.zeros(self.rhs_shape)
        _check(lib().ples_mg_metrica(self._h, (_dp * 9)(*[_p(a) for a in g]),
                                     (_dp * 9)(*[_p(a) for a in v]), _p(giac)), "ples_mg_metrica")
        return g, v, giac

    def level_shape(self, level: int):
        """(nx, ny, nz) cell counts of level `level` (1-based)."""
        v = [C.c_int(), C.c_int(), C.c_int()]
        _check(lib().ples_mg_level_shape(self._h, level, *[C.byref(a) for a in v]), "ples_mg_level_shape")
        return tuple(a.value for a in v)

    def level_metrics(self, level: int):
        """FP64 metrics of level `level` (1-based) in the ghosted layout."""
        nx, ny, nz = self.level_shape(level)
        sh = (nz + 2, ny + 2, nx + 2)
        g = [np.zeros(sh) for _ in range(9)]
        _check(lib().ples_mg_level_metrics(self._h, level, (_dp * 9)(*[_p(a) for a in g])),
               "ples_mg_level_metrics")
        return g

    def project_rhs(self, rhs, cs=None, dt=1.0):
        """Zero-mean compatibility projection in place; returns the value removed."""
        _need(rhs, self.rhs_shape, "rhs")
        if not (rhs.flags["C_CONTIGUOUS"] and rhs.dtype == np.float64):
            raise TypeError("rhs: C-contiguous float64, projected in place")
        cs = self._cs(cs)
        rem = C.c_double()
        _check(lib().ples_mg_project_rhs(self._h, _p(rhs), *[_p(c) for c in cs], dt, C.byref(rem)),
               "ples_mg_project_rhs")
        return rem.value

    def level_op(self, level, op, in0, in1=None, P=0, omega=1.0):
        """One kernel-level operation of level `level` (1-based) on ghosted
        level arrays (see ples_mg_level_op); returns the output array."""
        nx, ny, nz = self.level_shape(level)
        sh = (nz + 2, ny + 2, nx + 2)
        if op in (OP_RESTRICT, OP_RESID_RESTRICT, OP_PROLONG):
            cx, cy, cz = self.level_shape(level + 1)
            csh = (cz + 2, cy + 2, cx + 2)
        else:
            csh = sh
        in_sh = csh if op == OP_PROLONG else sh
        out_sh = csh if op in (OP_RESTRICT, OP_RESID_RESTRICT) else sh
        _need(in0, in_sh, "in0")
        if in1 is not None:
            _need(in1, sh, "in1")
        a0 = np.ascontiguousarray(in0, np.float64)
        a1 = None if in1 is None else np.ascontiguousarray(in1, np.float64)
        out = np.zeros(out_sh)
        _check(lib().ples_mg_level_op(self._h, level, op, P, omega, _p(a0), _p(a1), _p(out)), "ples_mg_level_op")
        return out

    def diver(self, uc, vc, wc, dt=1.0):
        for a, s, nm in zip((uc, vc, wc), self.face_shapes(), ("uc", "vc", "wc")):
            _need(a, s, nm)
        out = np.zeros(self.rhs_shape)
        _check(lib().ples_mg_diver(self._h, _p(np.ascontiguousarray(uc)), _p(np.ascontiguousarray(vc)),
                                   _p(np.ascontiguousarray(wc)), dt, _p(out)), "ples_mg_diver")
        return out

    def gradie(self, fi, dt=1.0):
        _need(fi, self.fi_shape, "fi")
        jx, jy, jz = self.jx, self.jy, self.jz
        outs = [np.zeros((jz, jy, jx)) for _ in range(3)] + \
               [np.zeros((jz, jy, jx + 1)), np.zeros((jz, jy + 1, jx)), np.zeros((jz + 1, jy, jx))]
        _check(lib().ples_mg_gradie(self._h, _p(np.ascontiguousarray(fi)), 1, 1, dt,
                                    *[_p(a) for a in outs]), "ples_mg_gradie")
        return outs

    # ---- time-loop steps around the solve (arrays in C order: reversed reference shapes)
    def face_shapes(self):
        jx, jy, jz = self.jx, self.jy, self.jz
        return [(jz, jy, jx + 1), (jz, jy + 1, jx), (jz + 1, jy, jx)]

    def _cs_out(self):
        return [np.zeros(s) for s in self.cs_shapes()]

    def _check_cells_faces(self, vel, fc):
        for a in vel:
            _need(a, self.fi_shape, "vel")
        for a, s in zip(fc, self.face_shapes()):
            _need(a, s, "fc")
            if not (a.flags["C_CONTIGUOUS"] and a.dtype == np.float64):
                raise TypeError("fc: C-contiguous float64 arrays, updated in place")

    def update(self, vel, dvel):
        """In place on vel = [u, v, w] (fi_shape); dvel shaped fi_shape too."""
        for a in vel:
            _need(a, self.fi_shape, "vel")
            if not (a.flags["C_CONTIGUOUS"] and a.dtype == np.float64):
                raise TypeError("vel: C-contiguous float64 arrays, updated in place")
        for a in dvel:
            _need(a, self.fi_shape, "dvel")
        dv = [np.ascontiguousarray(a, np.float64) for a in dvel]
        _check(lib().ples_mg_update(self._h, *[_p(a) for a in vel], 1, 1, *[_p(a) for a in dv]),
               "ples_mg_update")

    def contra(self, vel, fc, potenziale=0):
        """In place on fc = [uc, vc, wc]; returns cs1..cs6 (zeros for periodic directions)."""
        self._check_cells_faces(vel, fc)
        cs = self._cs_out()
        _check(lib().ples_mg_contra(self._h, *[_p(np.ascontiguousarray(a)) for a in vel], 1, 1,
                                    *[_p(a) for a in fc], (_dp * 6)(*[_p(a) for a in cs]), potenziale),
               "ples_mg_contra")
        return cs

    def vel_up(self, vel, fc, gra, cgra, dt=1.0, potenziale=0):
        self._check_cells_faces(vel, fc)
        for a in gra:
            _need(a, self.rhs_shape, "gra")
        for a, s in zip(cgra, self.face_shapes()):
            _need(a, s, "cgra")
        gra = [np.ascontiguousarray(a, np.float64) for a in gra]
        cgra = [np.ascontiguousarray(a, np.float64) for a in cgra]
        _check(lib().ples_mg_vel_up(self._h, *[_p(a) for a in vel], 1, 1, *[_p(a) for a in fc],
                                    (_dp * 3)(*[_p(a) for a in gra]), (_dp * 3)(*[_p(a) for a in cgra]),
                                    dt, potenziale), "ples_mg_vel_up")

    def check_divergence(self, fc):
        for a, s in zip(fc, self.face_shapes()):
            _need(a, s, "fc")
        out = np.zeros(4)
        _check(lib().ples_mg_check_divergence(self._h, *[_p(np.ascontiguousarray(a)) for a in fc], _p(out)),
               "ples_mg_check_divergence")
        return out

    def projection_step(self, vel, dvel, fc, fi, dt=1.0, eps=1e-10, ficycle=200, potenziale=0):
        """update -> contra -> diver -> multi -> gradie -> vel_up -> check_divergence,
        in place on vel, fc, fi.  Returns (Result, cs1..cs6, div maxima)."""
        self._check_cells_faces(vel, fc)
        for a in dvel:
            _need(a, self.fi_shape, "dvel")
        _need(fi, self.fi_shape, "fi")
        cs = self._cs_out()
        dv = [np.ascontiguousarray(a, np.float64) for a in dvel]
        div = np.zeros(4)
        r = _Result()
        st = lib().ples_mg_projection_step(self._h, *[_p(a) for a in vel], 1, 1, *[_p(a) for a in dv],
                                           *[_p(a) for a in fc], (_dp * 6)(*[_p(a) for a in cs]), _p(fi),
                                           dt, eps, ficycle, potenziale, C.byref(r), _p(div))
        if st not in (OK, NOT_CONVERGED):
            raise PlesMGError(st, "ples_mg_projection_step")
        return Result(r.status, r.cycles, r.rel_res, r.abs_res, r.rhs_mean_removed, r.seconds), cs, div

    # ---- device-resident
    def case_rhs(self, seed: int, dt: float = 1.0):
        _check(lib().ples_mg_case_rhs(self._h, C.c_uint64(seed), dt), "ples_mg_case_rhs")

    def zero_fi(self):
        _check(lib().ples_mg_zero_fi(self._h), "ples_mg_zero_fi")

    def perturb_fi(self, seed: int, rel: float):
        _check(lib().ples_mg_perturb_fi(self._h, C.c_uint64(seed), rel), "ples_mg_perturb_fi")

    def solve_resident(self, eps=1e-10, maxcycles=200) -> Result:
        r = _Result()
        st = lib().ples_mg_solve_resident(self._h, eps, maxcycles, C.byref(r))
        if st not in (OK, NOT_CONVERGED):
            raise PlesMGError(st, "ples_mg_solve_resident")
        return Result(r.status, r.cycles, r.rel_res, r.abs_res, r.rhs_mean_removed, r.seconds)

    def get_fi(self):
        a = np.zeros(self.fi_shape)
        _check(lib().ples_mg_get_fi(self._h, _p(a)), "ples_mg_get_fi")
        return a

    def set_fi(self, fi):
        _need(fi, self.fi_shape, "fi")
        _check(lib().ples_mg_set_fi(self._h, _p(np.ascontiguousarray(fi, np.float64))), "ples_mg_set_fi")

    def get_rhs(self):
        a = np.zeros(self.rhs_shape)
        _check(lib().ples_mg_get_rhs(self._h, _p(a)), "ples_mg_get_rhs")
        return a

    def set_rhs(self, rhs):
        _need(rhs, self.rhs_shape, "rhs")
        _check(lib().ples_mg_set_rhs(self._h, _p(np.ascontiguousarray(rhs, np.float64))), "ples_mg_set_rhs")

    @property
    def stream(self) -> int:
        return lib().ples_mg_stream(self._h)

    @property
    def device_bytes(self) -> int:
        return int(lib().ples_mg_device_bytes(self._h))

    def cycle_levels(self) -> int:
        """Levels the V-cycle visits (nlevel, or fewer with coarse_exact)."""
        v = C.c_int()
        _check(lib().ples_mg_cycle_levels(self._h, C.byref(v)), "ples_mg_cycle_levels")
        return v.value

    def tiles(self, level: int):
        """(smoother segments P, residual segments rP, residual lanes rbx) of a level."""
        v = [C.c_int(), C.c_int(), C.c_int()]
        _check(lib().ples_mg_tiles(self._h, level, *[C.byref(a) for a in v]), "ples_mg_tiles")
        return tuple(a.value for a in v)

    def set_timing(self, families=None, every=1):
        """Time the given kernel families (names or ids); None/False = off,
        True = all.  every: time one refinement step in `every` (the others
        replay their graphs)."""
        if families is True:
            mask = 0xFF
        elif not families:
            mask = 0
        else:
            mask = 0
            for f in families:
                mask |= 1 << (KERNEL_FAMILIES[f] if isinstance(f, str) else int(f))
        _check(lib().ples_mg_set_timing(self._h, mask), "ples_mg_set_timing")
        _check(lib().ples_mg_set_timing_every(self._h, int(every)), "ples_mg_set_timing_every")

    def kernel_stats(self, family) -> tuple[int, float, float]:
        fam = KERNEL_FAMILIES[family] if isinstance(family, str) else int(family)
        n, ms, by = C.c_long(0), C.c_double(0), C.c_double(0)
        _check(lib().ples_mg_kernel_stats(self._h, fam, C.byref(n), C.byref(ms), C.byref(by)),
               "ples_mg_kernel_stats")
        return n.value, ms.value, by.value


class Batch:
    """n independent problems solved together (ples_mg_batch_*).  problems:
    list of dicts with x, y, z (node arrays), nlevel and optional ip, kp.
    grouped: one launch per kernel over each family of same-shape problems
    (the default), else every problem's kernels on its own stream."""

    def __init__(self, problems, precision="mixed", grouped=True, **opts):
        self._probs = []
        grids = (_Grid * len(problems))()
        for q, pr in enumerate(problems):
            x, y, z = (np.ascontiguousarray(pr[c], dtype=np.float64) for c in ("x", "y", "z"))
            ip, kp = pr.get("ip", 0), pr.get("kp", 0)
            self._probs.append((x, y, z, ip, 1, kp, pr["nlevel"]))
            grids[q] = _grid(x, y, z, ip, 1, kp, pr["nlevel"])
        o = _opts(precision, **opts)
        b = C.c_void_p()
        _check(lib().ples_mg_batch_create(len(problems), grids, C.byref(o), C.byref(b)), "ples_mg_batch_create")
        self._b = b
        self.solvers = [Solver._borrow(lib().ples_mg_batch_handle(b, q), *self._probs[q])
                        for q in range(len(problems))]
        self.set_grouped(grouped)

    def set_grouped(self, on: bool):
        _check(lib().ples_mg_batch_set_grouped(self._b, 1 if on else 0), "ples_mg_batch_set_grouped")

    def families(self) -> int:
        """Families of the last grouped graph (0 before a grouped solve)."""
        return int(lib().ples_mg_batch_families(self._b))

    def __len__(self):
        return len(self._probs)

    def device_bytes(self) -> int:
        return int(lib().ples_mg_batch_device_bytes(self._b))

    def solve(self, eps=1e-8, maxcycles=200, steps_per_launch=0):
        res = (_Result * len(self))()
        st = lib().ples_mg_batch_solve(self._b, eps, maxcycles, steps_per_launch, res)
        if st not in (OK, NOT_CONVERGED):
            raise PlesMGError(st, "ples_mg_batch_solve")
        return [Result(r.status, r.cycles, r.rel_res, r.abs_res, r.rhs_mean_removed, r.seconds) for r in res]

    def close(self):
        if getattr(self, "_b", None):
            self.solvers = []
            lib().ples_mg_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Serving loop (ples_mg_server_*): cases with resident warm starts,
    HBM-budget admission, continuous batching of running requests."""

    def __init__(self, hbm_budget=0, max_active=8, warm_cases=16, precision="mixed", **opts):
        so = _ServerOpts()
        lib().ples_mg_server_default_opts(C.byref(so))
        so.hbm_budget, so.max_active, so.warm_cases = int(hbm_budget), max_active, warm_cases
        o = _opts(precision, **opts)
        sv = C.c_void_p()
        _check(lib().ples_mg_server_create(C.byref(so), C.byref(o), C.byref(sv)), "ples_mg_server_create")
        self._sv = sv
        self._cases = {}

    def add_case(self, case_id, x, y, z, nlevel, rhs_seed, dt=1.0, ip=0, kp=0):
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        self._cases[case_id] = (x.shape, (x, y, z))
        g = _grid(x, y, z, ip, 1, kp, nlevel)
        _check(lib().ples_mg_server_add_case(self._sv, case_id, C.byref(g), rhs_seed, dt), "ples_mg_server_add_case")

    def submit(self, case_id, perturb_seed, perturb_rel=1e-3, eps=1e-10, maxcycles=200) -> int:
        t = C.c_int64()
        _check(lib().ples_mg_server_submit(self._sv, case_id, perturb_seed, perturb_rel, eps, maxcycles,
                                           C.byref(t)), "ples_mg_server_submit")
        return t.value

    def step(self) -> int:
        n = C.c_int()
        _check(lib().ples_mg_server_step(self._sv, C.byref(n)), "ples_mg_server_step")
        return n.value

    def wait(self):
        _check(lib().ples_mg_server_wait(self._sv), "ples_mg_server_wait")

    def poll(self, ticket):
        """(Result, warm) once done, else None."""
        r = _Result()
        w = C.c_int()
        st = lib().ples_mg_server_poll(self._sv, ticket, C.byref(r), C.byref(w))
        if st == BUSY:
            return None
        _check(st, "ples_mg_server_poll")
        return Result(r.status, r.cycles, r.rel_res, r.abs_res, r.rhs_mean_removed, r.seconds), bool(w.value)

    def pending(self) -> int:
        return int(lib().ples_mg_server_pending(self._sv))

    def drain(self):
        while self.pending():
            if self.step() == 0:
                self.wait()

    def get_fi(self, case_id):
        (nzp, nyp, nxp), _ = self._cases[case_id]
        a = np.zeros((nzp + 1, nyp + 1, nxp + 1))
        _check(lib().ples_mg_server_get_fi(self._sv, case_id, _p(a)), "ples_mg_server_get_fi")
        return a

    def set_fi(self, case_id, fi):
        """Start the case's next request from this pressure (shape of get_fi),
        e.g. the fi of a restart file."""
        (nzp, nyp, nxp), _ = self._cases[case_id]
        a = np.ascontiguousarray(fi, np.float64)
        if a.shape != (nzp + 1, nyp + 1, nxp + 1):
            raise ValueError(f"fi shape {a.shape} != {(nzp + 1, nyp + 1, nxp + 1)}")
        _check(lib().ples_mg_server_set_fi(self._sv, case_id, _p(a)), "ples_mg_server_set_fi")

    def stats(self) -> dict:
        u, r = C.c_size_t(), C.c_int()
        c, e, k = C.c_long(), C.c_long(), C.c_long()
        _check(lib().ples_mg_server_stats(self._sv, C.byref(u), C.byref(r), C.byref(c), C.byref(e), C.byref(k)),
               "ples_mg_server_stats")
        return dict(used_bytes=u.value, resident=r.value, completed=c.value, evictions=e.value, cold_starts=k.value)

    def close(self):
        if getattr(self, "_sv", None):
            lib().ples_mg_server_destroy(self._sv)
            self._sv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
