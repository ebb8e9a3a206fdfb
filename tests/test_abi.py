"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from ples_amd import mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ples_\w+)\s*\(", text):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(mg.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_status_strings_and_defaults():
    assert mg.status_string(0) == "ok"
    assert "dimensions" in mg.status_string(4)
    o = mg.default_opts()
    assert o["precision"] == mg.FP64 and o["krylov"] == 1 and o["nu_pre"] == 3 and o["nu_post"] == 3
    assert o["graphs"] == 1 and o["fp32_krylov"] == 1 and o["autotune"] == 0 and o["smoother_fp16"] == 1
    assert o["coarse_exact"] == 1 and o["sweep_schedule"] == 1 and o["semicoarsen"] == 0


def test_tuning_table_parses():
    """The committed launch-shape table loads (no GPU needed) and every
    entry is a valid shape: P, rP powers of two dividing ny, rbx*rP <= 512."""
    assert mg.tuning_load()
    n = 0
    for line in open(mg.TUNING_TABLE):
        if line.startswith("#"):
            continue
        f = list(map(int, line.split()))
        if len(f) == 10:  # older layout: no 16-bit smoother metrics
            f = f[:7] + [0] + f[7:] + [0]
        nx, ny, nz, ip, kp, f32, fine, half, P, rP, rbx, use16 = f
        assert ny % P == 0 and ny % rP == 0 and rbx * rP <= 512
        assert use16 in (0, 1) and half in (0, 1) and use16 <= half and half <= f32
        n += 1
    assert n > 20


def test_tuning_loader_rejects_segments_below_four_rows(tmp_path):
    """A table entry whose smoother or residual segments would be shorter
    than the four rows the autotuner's candidates keep is not loaded."""
    bad = tmp_path / "bad.txt"
    bad.write_text("32 16 32 0 0 1 1 1 16 8 32 1\n32 16 32 0 0 1 1 1 4 16 32 1\n")
    assert not mg.tuning_load(str(bad))
    good = tmp_path / "good.txt"
    good.write_text("32 16 32 0 0 1 1 1 4 2 64 1\n")
    assert mg.tuning_load(str(good))
    mg.tuning_clear()
