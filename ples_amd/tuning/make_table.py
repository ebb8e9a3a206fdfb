"""Build the autotuner's launch-shape table for the benchmark workloads on
this GPU: creates every level shape of the default workloads with
autotune=1 (each level's candidate shapes are timed on the device) and
writes the table that `ples_mg_tuning_load` reads.

    python -m ples_amd.tuning.make_table OUT.txt
"""
from __future__ import annotations

import sys

from ples_amd import cases as PC
from ples_amd import mg


def main(out: str):
    shapes = []
    for name in ("c2", "c3"):
        cfg = PC.CONFIGS[name]
        shapes.append((cfg.nodes(), cfg.nlevel))
    seen = set()
    for p in PC.c4_problems() + PC.c5_cases():
        key = (p["shape"], p["nlevel"])
        if key not in seen:
            seen.add(key)
            shapes.append(((p["x"], p["y"], p["z"]), p["nlevel"]))
    for (x, y, z), nlevel in shapes:
        for prec in ("mixed", "fp64"):
            with mg.Solver(x, y, z, 0, 1, 0, nlevel, precision=prec, autotune=1) as s:
                print(prec, x.shape, [s.tiles(l) for l in range(1, nlevel + 1)], flush=True)
    mg.tuning_save(out)


if __name__ == "__main__":
    main(sys.argv[1])
