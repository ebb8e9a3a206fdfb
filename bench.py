"""Benchmark: Poisson solves/s to 1e-10 relative residual on one MI355X.

Workload (default): the terrain-following 256x128x256 grid (tanh 1.5
stretching, terrain amplitude 0.1 H), 6 multigrid levels, seeded Gaussian
divergence right-hand side, solved from zero to ||b - A p||/||b|| < 1e-10.
One step = one complete solve.  `value` is the mixed-precision solver (FP64
iterative refinement with FP64 residuals around an FP32 GCR preconditioned by
FP32 V-cycles); the FP64 multigrid is timed on the same steps and reported
beside it.  Both use 3/3 sweeps and autotuned launch shapes.  The
512x256x512 warm-start solve is reported under "c3", the 64-problem batch
under "c4_batch" and the Poisson-arrival serving stream under "c5_serving".

The roofline entry is the fine-level eta-line smoother of the timed mixed
solves: algorithmic bytes per launch (the library's per-launch byte count)
over the launch duration measured with HIP events on the solver's stream,
inside the timed region.

The cpu_baseline entry times the CPU oracle (the reference algorithm: SOR
1.6, 5/10 sweeps per level visit, ghost fill 5x per sweep) on a bounded
sample: three V-cycles of the same 256x128x256 solve on one core; solves/s
extrapolate the measured time per cycle to the cycles needed for 1e-10 at
the contraction measured over the sampled cycles.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X, AMD published


def solver_for(cfg, precision, **kw):
    """Both precisions with the library defaults (3/3 sweeps) and the launch
    shapes chosen by the autotuner."""
    from ples_amd import mg

    x, y, z = cfg.nodes()
    kw.setdefault("autotune", 1)
    return mg.Solver(x, y, z, 0, 1, 0, cfg.nlevel, precision=precision, **kw)


def time_solves(s, steps, warmup, tol, prepare, timing_fams=None):
    """prepare(s) sets up the start of a step (device work, inside the step)."""
    import ctypes

    res = None
    for _ in range(warmup):
        prepare(s)
        res = s.solve_resident(tol, 400)
    if timing_fams:
        s.set_timing(timing_fams)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    cycles = []
    for _ in range(steps):
        prepare(s)
        res = s.solve_resident(tol, 400)
        cycles.append(res.cycles)
        if not res.ok:
            raise RuntimeError(f"solve did not converge: {res}")
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    return dt, res, cycles


def rel_l2_meanfree(p, q):
    """||(p - mean p) - (q - mean q)||_2 / ||q - mean q||_2 over the interior of
    two ghosted pressures (the pressure is defined up to a constant)."""
    a = p[1:-1, 1:-1, 1:-1]
    b = q[1:-1, 1:-1, 1:-1]
    a = a - a.mean()
    b = b - b.mean()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# tolerance of the FP64 reference pressure every result is compared with
REF_TOL = 1e-12


def traffic_per_launch():
    """HBM bytes per launch of the fine-level mixed smoother at the default
    workload, from the committed rocprofv3 PMC summary (FETCH_SIZE scaled by
    the copy-kernel factor, plus WRITE_SIZE), or None without one."""
    path = os.path.join(ROOT, "profiles", "pmc_smooth_fine_c2_256x128x256.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def bench_c4(steps):
    """64 uneven problems in one batch (graph launches of the whole batch)
    against the same 64 solved one after another on their own handles."""
    import ctypes

    from ples_amd import cases as PC
    from ples_amd import mg

    hip = ctypes.CDLL("libamdhip64.so")
    probs = PC.c4_problems()
    tol = 1e-8
    out = {"problems": len(probs), "cells": int(sum(int(np.prod(p["shape"])) for p in probs)), "tolerance": tol}
    with mg.Batch(probs, precision="mixed") as b:
        for p, s in zip(probs, b.solvers):
            s.case_rhs(p["seed"])

        def run():
            for s in b.solvers:
                s.zero_fi()
            return b.solve(tol, 200)

        run()  # warm-up (graph capture)
        hip.hipDeviceSynchronize()
        reps = max(1, steps // 3)
        t = time.perf_counter()
        for _ in range(reps):
            res = run()
        hip.hipDeviceSynchronize()
        dt = (time.perf_counter() - t) / reps
        assert all(r.ok for r in res), res
        out.update(batch_solves_per_s=len(probs) / dt, batch_ms=1e3 * dt,
                   cycles_max=max(r.cycles for r in res), cycles_mean=float(np.mean([r.cycles for r in res])))
        # the same problems one at a time (each handle's own graph-replayed loop)
        hip.hipDeviceSynchronize()
        t = time.perf_counter()
        for s in b.solvers:
            s.zero_fi()
            r = s.solve_resident(tol, 200)
            assert r.ok
        hip.hipDeviceSynchronize()
        dts = time.perf_counter() - t
        out.update(sequential_solves_per_s=len(probs) / dts, sequential_ms=1e3 * dts,
                   speedup_batch_vs_sequential=dts / dt)
    return out


def bench_c5(n_req=400, n_cases=16, rate=200.0, max_active=16):
    """Serving stream: Poisson arrivals in real time into the serving loop."""
    from ples_amd import cases as PC
    from ples_amd import mg

    cases = PC.c5_cases(n_cases)
    reqs = PC.c5_requests(n_req, n_cases, rate)
    lat, warm, cyc = [], [], []
    with mg.Server(max_active=max_active, warm_cases=16, precision="mixed") as sv:
        for c in cases:
            sv.add_case(c["case_id"], c["x"], c["y"], c["z"], c["nlevel"], c["rhs_seed"])
        # warm-up: one request per case (kernel loading, graph capture), untimed
        for c in cases[:16]:
            sv.submit(c["case_id"], 1, 1e-3, 1e-6, 200)
        sv.drain()
        t0 = time.perf_counter()
        nxt, open_t = 0, {}
        while nxt < len(reqs) or open_t:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt][0] <= now:
                at, c, eps, seed = reqs[nxt]
                open_t[sv.submit(c, seed, 1e-3, eps, 200)] = at
                nxt += 1
            progressed = sv.step()
            for tk in list(open_t):
                r = sv.poll(tk)
                if r is not None:
                    res, w = r
                    assert res.ok, res
                    lat.append(time.perf_counter() - t0 - open_t.pop(tk))
                    warm.append(w)
                    cyc.append(res.cycles)
            if not progressed and open_t:
                if sv.pending():
                    sv.wait()
            elif not open_t and nxt < len(reqs):
                time.sleep(max(0.0, min(1e-3, reqs[nxt][0] - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        st = sv.stats()
    lat = np.array(lat)
    return {"requests": n_req, "rate_per_s": rate, "cases": n_cases, "warm_cases": 16, "max_active": max_active,
            "resident_bytes": st["used_bytes"],
            "solves_per_s": n_req / wall, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "mean_ms": 1e3 * float(lat.mean()),
            "warm_fraction": float(np.mean(warm)), "cycles_mean": float(np.mean(cyc)),
            "evictions": st["evictions"], "cold_starts": st["cold_starts"], "wall_s": wall}


def cpu_baseline(cfg):
    import numpy as np

    from oracle import oracle as O

    O.build()
    x, y, z = cfg.nodes()
    g = O.Grid(cfg.nx, cfg.ny, cfg.nz, 0, 1, 0, x, y, z)
    rhs = O.case_rhs(g, cfg.seed)
    b0 = float(np.abs(rhs[1:-1, 1:-1, 1:-1]).max())
    t = time.perf_counter()
    O.multi(g, cfg.nlevel, g.zeros(), rhs, eps=1e-300, ficycle=0)
    t_setup = time.perf_counter() - t
    ncyc = 3
    t = time.perf_counter()
    _, _, hist = O.multi(g, cfg.nlevel, g.zeros(), rhs, eps=1e-300, ficycle=ncyc)
    t_cyc = (time.perf_counter() - t - t_setup) / ncyc
    rho = float(hist[-1] / hist[-2])
    need = math.log(1e-10 * b0 / hist[-1]) / math.log(rho) if rho < 1 else float("inf")
    cycles = ncyc + max(0.0, need)
    return {
        "value": 1.0 / (t_cyc * cycles),
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"3 V-cycles of the reference algorithm (CPU oracle: SOR 1.6, 5/10 sweeps, "
                   f"5 ghost passes per sweep, L={cfg.nlevel}) on the {cfg.nx}x{cfg.ny}x{cfg.nz} case: "
                   f"{t_cyc:.2f} s/cycle, contraction {rho:.3f}, {cycles:.0f} cycles to 1e-10 extrapolated"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--no-c3", action="store_true", help="skip the 512x256x512 warm-start line")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-c4", action="store_true", help="skip the 64-problem batch")
    ap.add_argument("--no-c5", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-live-timing", action="store_true",
                    help="no per-kernel HIP events in the timed region (roofline fields null)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    from ples_amd import cases as PC

    cfg = PC.CONFIGS[args.config]
    from ples_amd import mg

    tuned = mg.tuning_load()  # committed launch-shape table of this GPU (else tuned live)
    copy_gbs = mg.copy_bandwidth(1 << 30, 20)
    print(f"[bench] measured copy bandwidth {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)

    def from_zero(s):
        s.zero_fi()

    out = {}
    for prec in ("mixed", "fp64"):
        with solver_for(cfg, prec) as s:
            s.case_rhs(cfg.seed)
            fams = ["smooth_fine", "resid_fine"] if prec == "mixed" else ["smooth_fine"]
            if args.no_live_timing:
                fams = []
            dt, res, cyc = time_solves(s, args.steps, args.warmup, cfg.tol, from_zero, fams)
            stats = {f: s.kernel_stats(f) for f in ("smooth_fine", "resid_fine")}
            out[prec] = dict(seconds=dt, res=res, cycles=cyc, stats=stats, fi=s.get_fi())
            if prec == "fp64":  # reference pressure: the FP64 solve continued to REF_TOL
                s.set_timing(None)
                s.solve_resident(REF_TOL, 400)
                out["ref_fi"] = s.get_fi()
            print(f"[bench] {cfg.name} {prec}: {args.steps / dt:.2f} solves/s, {1e3 * dt / args.steps:.2f} ms/solve, "
                  f"cycles {cyc[-1]}, rel_res {res.rel_res:.2e}", file=sys.stderr, flush=True)

    mixed, fp64 = out["mixed"], out["fp64"]
    n, ms, by = mixed["stats"]["smooth_fine"]
    avg_ms = ms / max(n, 1)
    achieved = (by / max(n, 1)) / (avg_ms * 1e-3) / 1e9 if n else None
    rn, rms, rby = mixed["stats"]["resid_fine"]
    resid_gbs = (rby / max(rn, 1)) / ((rms / max(rn, 1)) * 1e-3) / 1e9 if rn else None

    c3 = None
    if not args.no_c3 and args.config != "c3":
        c3cfg = PC.CONFIGS["c3"]
        c3, c3_fi = {}, {}
        for prec in ("mixed", "fp64"):
            with solver_for(c3cfg, prec) as s:
                s.case_rhs(c3cfg.seed)
                s.zero_fi()
                base = s.solve_resident(c3cfg.tol, 400)  # p_prev

                def warm(sv):
                    sv.perturb_fi(1004, 1e-3)

                steps3 = max(2, args.steps // 2)
                dt, res, cyc = time_solves(s, steps3, 1, c3cfg.tol, warm,
                                           [] if args.no_live_timing else ["smooth_fine"])
                n3, ms3, by3 = s.kernel_stats("smooth_fine")
                c3[prec] = {"solves_per_s": steps3 / dt, "ms_per_solve": 1e3 * dt / steps3,
                            "cycles": cyc[-1], "cycles_from_zero": base.cycles, "rel_res": res.rel_res,
                            "smoother_fine_GBs": (by3 / n3) / ((ms3 / n3) * 1e-3) / 1e9 if n3 else None}
                c3_fi[prec] = s.get_fi()
                if prec == "fp64":
                    s.set_timing(None)
                    s.solve_resident(REF_TOL, 400)
                    c3_fi["ref"] = s.get_fi()
                print(f"[bench] c3 {prec}: {c3[prec]}", file=sys.stderr, flush=True)
        c3["speedup_mixed_vs_fp64"] = c3["mixed"]["solves_per_s"] / c3["fp64"]["solves_per_s"]
        c3["pressure_rel_l2_vs_fp64_1e-12"] = {p: rel_l2_meanfree(c3_fi[p], c3_fi["ref"]) for p in ("mixed", "fp64")}
        c3["pressure_rel_l2_mixed_vs_fp64"] = rel_l2_meanfree(c3_fi["mixed"], c3_fi["fp64"])
        del c3_fi

    c4 = None
    if not args.no_c4:
        c4 = bench_c4(args.steps)
        print(f"[bench] c4: {c4}", file=sys.stderr, flush=True)
    c5 = None
    if not args.no_c5:
        c5 = bench_c5()
        print(f"[bench] c5: {c5}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu and int(os.environ.get("RANK", "0")) == 0:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(cfg)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    value = args.steps / mixed["seconds"]
    line = {
        "metric": "Poisson solves/s to 1e-10 rel-res (mixed vs FP64 MG, 1 GPU)",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded Gaussian divergence rhs generated on device)",
        "config": {"workload": cfg.text, "cells": cfg.cells, "levels": cfg.nlevel, "tolerance": cfg.tol,
                   "launch_shapes": "tuning table" if tuned else "tuned at create",
                   "precision": "FP32 V-cycles, FP64 GCR refinement and outer residual"},
        "fp64_solves_per_s": args.steps / fp64["seconds"],
        "speedup_mixed_vs_fp64": fp64["seconds"] / mixed["seconds"],
        "cycles": {"mixed": mixed["cycles"][-1], "fp64": fp64["cycles"][-1]},
        "rel_res": {"mixed": mixed["res"].rel_res, "fp64": fp64["res"].rel_res},
        "pressure_rel_l2_vs_fp64_1e-12": {p: rel_l2_meanfree(out[p]["fi"], out["ref_fi"]) for p in ("mixed", "fp64")},
        "pressure_rel_l2_mixed_vs_fp64": rel_l2_meanfree(out["mixed"]["fi"], out["fp64"]["fi"]),
        "roofline": {
            "kernel": "k_yline<float, packed 16-bit metrics> fine level (eta-line smoother)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic_per_launch(),
            "copy_measured_GBs": copy_gbs,
            "frac_of_copy": achieved / copy_gbs if achieved else None,
            "launches": n,
            "avg_us": 1e3 * avg_ms,
            "bytes_per_launch": by / max(n, 1),
            "resid_fine_GBs": resid_gbs,
        },
        "cpu_baseline": cpu,
        "c3": c3,
        "c4_batch": c4,
        "c5_serving": c5,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
