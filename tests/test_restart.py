"""Restart file of the flow field (ples_mg_restart_write / _read) in the
layout PLES reads at restart (fsrc/restart.f03:33-98, format 10e18.10):
round trip to the format's ten significant digits, and agreement with
Fortran's own formatted I/O (a Fortran program using the same record
statements, built with the image's Fortran compiler): our data records are
character-identical to Fortran's, Fortran reads our file, we read Fortran's."""
import os
import subprocess

import numpy as np
import pytest

from ples_amd import build as B
from ples_amd import mg

JX, JY, JZ, NSCAL = 5, 3, 4, 2
HERE = os.path.dirname(os.path.abspath(__file__))


def _fields(seed):
    rng = np.random.default_rng(seed)
    cells = (JZ + 2, JY + 2, JX + 2)
    scale = 10.0 ** rng.integers(-12, 12, size=cells)
    u, v, w, fi = (rng.standard_normal(cells) * scale for _ in range(4))
    u[0, 0, 0] = 0.0
    v[1, 1, 1] = -0.0
    rhov = rng.standard_normal(cells + (NSCAL,))
    uc, vc, wc = rng.standard_normal((JZ, JY, JX + 1)), rng.standard_normal((JZ, JY + 1, JX)), \
        rng.standard_normal((JZ + 1, JY, JX))
    return 0.375, 1.25, -0.5, u, v, w, fi, rhov, uc, vc, wc


def _close(a, b):
    # 0.dddddddddd x 10^e: half a unit in the tenth digit is at most 5e-10 of the value
    return np.allclose(a, b, rtol=5.0001e-10, atol=0)


def test_round_trip(tmp_path):
    f = _fields(1)
    path = str(tmp_path / "new_res_form")
    mg.restart_write(path, *f)
    g = mg.restart_read(path, JX, JY, JZ, NSCAL)
    assert g[:3] == f[:3]
    for a, b in zip(f[3:], g[3:]):
        assert _close(b, a)


def _fortran_helper(tmp_path):
    if not os.path.exists(B.FC):
        pytest.skip("no Fortran compiler")
    exe = str(tmp_path / "restart_io")
    subprocess.run([B.FC, "-O0", os.path.join(HERE, "fortran", "restart_io.f90"), "-o", exe], check=True,
                   cwd=str(tmp_path))
    return exe


def _pack(f):
    ti, bbx, dbbx, u, v, w, fi, rhov, uc, vc, wc = f
    cell = np.concatenate([np.stack([u, v, w, fi], axis=-1), rhov], axis=-1)  # (k, j, i, 4+nscal)
    return np.concatenate([[ti, bbx, dbbx], cell.ravel(), uc.ravel(), vc.ravel(), wc.ravel()])


def test_agrees_with_fortran_formatted_io(tmp_path):
    exe = _fortran_helper(tmp_path)
    f = _fields(2)
    binp = tmp_path / "in.bin"
    _pack(f).tofile(binp)
    ffile = str(tmp_path / "fortran_res")
    subprocess.run([exe, "write", str(binp), ffile, str(JX), str(JY), str(JZ), str(NSCAL)], check=True)
    ours = str(tmp_path / "our_res")
    mg.restart_write(ours, *f)
    # data records (after the three header records) are character-identical
    fl = open(ffile).read().splitlines()[3:]
    ol = open(ours).read().splitlines()[3:]
    assert len(fl) == len(ol)
    assert fl == ol
    # we read Fortran's file
    g = mg.restart_read(ffile, JX, JY, JZ, NSCAL)
    for a, b in zip(f[3:], g[3:]):
        assert _close(b, a)
    # Fortran reads ours
    bout = tmp_path / "out.bin"
    subprocess.run([exe, "read", ours, str(bout), str(JX), str(JY), str(JZ), str(NSCAL)], check=True)
    back = np.fromfile(bout)
    assert back.shape == _pack(f).shape
    assert _close(back, _pack(f))
