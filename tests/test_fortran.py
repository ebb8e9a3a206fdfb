"""Fortran binding (ples_amd/fortran/ples_mg_f.f90): the module compiles with
the image's Fortran compiler and its bind(C) interfaces link against every
entry point they name; on a GPU the driver program runs a complete
`ples_multi` solve (the drop-in for the reference's `call multi(...)`,
fsrc/stratiParticle.f03:2881-2887) and its result matches the same solve made
through the Python binding."""
import os
import re
import subprocess

import numpy as np
import pytest

from ples_amd import build as B

pytestmark = pytest.mark.skipif(not os.path.exists(B.FC), reason="no Fortran compiler")


def _parse(out):
    m = re.search(r"status=(\d+) cycles=(\d+) resmax=\s*(\S+) check=\s*(\S+) fi_mean=\s*(\S+) fi_rms=\s*(\S+)", out)
    assert m, out
    st, cyc = int(m.group(1)), int(m.group(2))
    return st, cyc, [float(m.group(i)) for i in range(3, 7)]


def test_fortran_module_links(tmp_path):
    """Compile the module and the driver from the sources into a scratch
    folder: every bind(C) name must resolve in libples_mg.so."""
    if not os.path.exists(B.LIB):
        pytest.skip("library not built")
    mod = tmp_path / "m"
    mod.mkdir()
    subprocess.run([B.FC, "-c", os.path.join(B.FDIR, "ples_mg_f.f90"), "-J", str(mod),
                    "-o", str(mod / "f.o")], check=True)
    subprocess.run([B.FC, os.path.join(B.FDIR, "ples_mg_driver.f90"), str(mod / "f.o"), "-J", str(mod),
                    "-I", str(mod), "-L", B.LIBDIR, "-lples_mg", "-Wl,--no-undefined",
                    "-o", str(tmp_path / "drv")], check=True)
    assert (tmp_path / "drv").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fortran_driver_solve(prec):
    """The Fortran driver's ples_multi solve converges to its max-norm
    tolerance (1e-11) and gives the same pressure (mean, rms) as the Python
    binding's solve of the same grid and rhs."""
    assert os.path.exists(B.FDRIVER), "Fortran driver not built"
    out = subprocess.run([B.FDRIVER, prec], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    st, cyc, (resmax, check, fmean, frms) = _parse(out.stdout)
    assert st == 0 and cyc > 0
    assert resmax < 1e-11 and check < 2e-11

    from ples_amd import mg
    from tests.cases import driver_grid

    x, y, z = driver_grid()
    with mg.Solver(x, y, z, 0, 1, 0, 3, precision=prec, stop_mode="abs_max") as s:
        s.case_rhs(1002)
        rhs = s.get_rhs()
        fi = np.zeros((34, 18, 34))
        res = s.multi(fi, rhs, eps=1e-11, ficycle=200)
        assert res.ok
    inner = fi[1:-1, 1:-1, 1:-1]
    assert abs(inner.mean() - fmean) <= 1e-9 * frms + 1e-14
    assert abs(inner.std() - frms) <= 1e-9 * frms
