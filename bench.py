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

Every solve projects its rhs (FP64 compatibility projection) inside the
timed step.  "pressure_bar" reports, per precision, the loosest tolerance of
1e-10, 1e-11 and 1e-12 at which the pressure is within 1e-9 (relative L2,
mean-free) of the FP64 solve continued to 1e-13, with its solves/s.
"dropin" times the host-array entry points the Fortran loop calls: the solve
on host arrays (PCIe transfers included) and the one-call projection step.

The cpu_baseline entry times the CPU oracle (the reference algorithm: point
SOR 1.6, 5/10 sweeps per level visit, ghost fill 5x per sweep, the
reference's max-norm stop) built -O3 -march=native with OpenMP on this host:
complete solves, on one core and on k-slabs over the host cores with the
reference's MPI slab semantics; the 256x128x256 solve at the GPU's level
count (the value) and at the reference's cap of 4 levels, the 64x32x64
channel as the median of repeated solves (3 levels: with 4 the reference's
coarse wall metric inverts and its V-cycle diverges).
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
        s.set_timing(timing_fams, every=TIMING_EVERY)
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
REF_TOL = 1e-13
# live kernel timing: one refinement step in TIMING_EVERY launched eagerly with
# start/end events, the others replayed from their graphs (the timed run keeps
# the graphs' launch rate)
TIMING_EVERY = 4


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
    """64 uneven problems in one batch (graph launches of the whole batch,
    grouped: one launch per kernel over each shape family) against the same
    batch with every problem on its own stream and against the 64 solved
    one after another on their own handles."""
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
        out.update(batch_solves_per_s=len(probs) / dt, batch_ms=1e3 * dt, families=b.families(),
                   cycles_max=max(r.cycles for r in res), cycles_mean=float(np.mean([r.cycles for r in res])))
        # the same batch with every problem's kernels on its own stream
        b.set_grouped(False)
        run()
        hip.hipDeviceSynchronize()
        t = time.perf_counter()
        for _ in range(reps):
            res_s = run()
        hip.hipDeviceSynchronize()
        dt_s = (time.perf_counter() - t) / reps
        assert all(r.ok for r in res_s), res_s
        out.update(per_stream_solves_per_s=len(probs) / dt_s, per_stream_ms=1e3 * dt_s)
        b.set_grouped(True)
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


def _drive_stream(sv, reqs, t_scale=1.0):
    """Submit reqs (arrival s, case, eps, seed) at their arrival times
    (t_scale 0: all at once) while stepping the server; returns latencies,
    warm flags, cycles and the wall time."""
    lat, warm, cyc = [], [], []
    t0 = time.perf_counter()
    nxt, open_t = 0, {}
    while nxt < len(reqs) or open_t:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt][0] * t_scale <= now:
            at, c, eps, seed = reqs[nxt]
            open_t[sv.submit(c, seed, 1e-3, eps, 200)] = at * t_scale
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
            time.sleep(max(0.0, min(1e-3, reqs[nxt][0] * t_scale - (time.perf_counter() - t0))))
    return np.array(lat), warm, cyc, time.perf_counter() - t0


def bench_c5(n_req=400, n_cases=16, rate=200.0, max_active=16, warm_cases=16):
    """Serving: Poisson arrivals in real time into the serving loop; then the
    capacity: the same number of requests all queued at once (closed loop at
    saturation).  With more cases than warm slots, HBM admission evicts idle
    cases and cold starts build handles."""
    from ples_amd import cases as PC
    from ples_amd import mg

    cases = PC.c5_cases(n_cases)
    reqs = PC.c5_requests(n_req, n_cases, rate)
    with mg.Server(max_active=max_active, warm_cases=warm_cases, precision="mixed") as sv:
        for c in cases:
            sv.add_case(c["case_id"], c["x"], c["y"], c["z"], c["nlevel"], c["rhs_seed"])
        # warm-up: one request per case (kernel loading, graph capture), untimed
        for c in cases:
            sv.submit(c["case_id"], 1, 1e-3, 1e-6, 200)
        sv.drain()
        st0 = sv.stats()
        lat, warm, cyc, wall = _drive_stream(sv, reqs, 1.0)
        st = sv.stats()
        # capacity: every request queued at t = 0
        lat_c, _, _, wall_c = _drive_stream(sv, [(0.0, c, e, s + 100000) for _, c, e, s in reqs], 0.0)
    return {"requests": n_req, "rate_per_s": rate, "cases": n_cases, "warm_cases": warm_cases,
            "max_active": max_active, "resident_bytes": st["used_bytes"],
            "solves_per_s": n_req / wall, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "mean_ms": 1e3 * float(lat.mean()),
            "warm_fraction": float(np.mean(warm)), "cycles_mean": float(np.mean(cyc)),
            "evictions": st["evictions"] - st0["evictions"], "cold_starts": st["cold_starts"] - st0["cold_starts"],
            "wall_s": wall,
            "capacity_solves_per_s": n_req / wall_c, "capacity_p50_ms": 1e3 * float(np.percentile(lat_c, 50)),
            "capacity_p99_ms": 1e3 * float(np.percentile(lat_c, 99))}


def _lscpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _timed_solves(H, g, rhs, L, eps, nslab, min_solves, min_s, max_solves=20):
    """Complete solves from zero until min_solves ran and min_s elapsed (at
    most max_solves); returns the median seconds, cycles, the final
    relative L2 residual under the oracle's operator."""
    from oracle import oracle as O

    times, n = [], 0
    fi = g.zeros()
    t_all = time.perf_counter()
    while True:
        fi = g.zeros()
        t = time.perf_counter()
        n, rm, _ = H.multi(fi, rhs, eps=eps, ficycle=2000, nslab=nslab)
        times.append(time.perf_counter() - t)
        if (len(times) >= min_solves and time.perf_counter() - t_all >= min_s) or len(times) >= max_solves:
            break
    gm, _, _ = O.metrics(g, 0)
    r = O.resid(g, 0, fi, rhs, gm)
    rel = float(np.linalg.norm(r[1:-1, 1:-1, 1:-1]) / np.linalg.norm(rhs[1:-1, 1:-1, 1:-1]))
    return float(np.median(times)), n, rel, len(times)


def cpu_baseline(cfg):
    """The reference algorithm (CPU oracle) timed on this host's cores."""
    from ples_amd import cases as PC
    from oracle import oracle as O

    path = O.build_baseline(os.path.join(ROOT, "oracle", "_build", "host"))
    cores = len(os.sched_getaffinity(0))
    out = {"cores_available": cores, "cpu_model": _lscpu_model(), "build": "gcc -O3 -march=native -fopenmp"}

    def slabs(nz, L):
        # the reference keeps at least two planes per rank on its coarsest
        # level (its level cap by planes per rank, fsrc/iniz_metrica.f03:22-47)
        ns = cores
        while ns > 1 and nz % (ns << L):
            ns -= 1
        return ns

    def case(c):
        x, y, z = c.nodes()
        g = O.Grid(c.nx, c.ny, c.nz, 0, 1, 0, x, y, z)
        rhs = O.case_rhs(g, c.seed)
        # the reference's max-norm stop, set so the relative L2 residual meets the tolerance
        return g, rhs, 0.4 * c.tol * float(np.abs(rhs).max())

    # 64x32x64 channel, 3 levels: median of repeated complete solves
    c1 = PC.CONFIGS["c1"]
    g, rhs, eps = case(c1)
    H = O.Hier(g, 3, path=path)
    rows = {}
    for label, ns in (("1core", 1), ("all_cores", slabs(c1.nz, 3))):
        t, n, rel, k = _timed_solves(H, g, rhs, 3, eps, ns, 3 if ns == 1 else 20, 2.0)
        rows[label] = {"solves_per_s": 1.0 / t, "s_per_solve": t, "cycles": n, "rel_res": rel, "threads": ns,
                       "solves_timed": k}
    H.close()
    out["c1_channel_64x32x64_L3"] = rows
    print(f"[bench] cpu c1: {rows}", file=sys.stderr, flush=True)
    # the benchmark workload: complete solves at the GPU's level count and at
    # the reference's cap of 4 levels (all cores), one core timed per cycle
    g, rhs, eps = case(cfg)
    rows = {}
    for L in (cfg.nlevel, min(cfg.nlevel, 4)):
        H = O.Hier(g, L, path=path)
        ns = slabs(cfg.nz, L)
        t, n, rel, k = _timed_solves(H, g, rhs, L, eps, ns, 1, 0.0, 1)
        rows[f"L{L}_all_cores"] = {"solves_per_s": 1.0 / t, "s_per_solve": t, "cycles": n, "rel_res": rel,
                                   "threads": ns}
        if L == cfg.nlevel:
            fi = g.zeros()
            t1 = time.perf_counter()
            H.multi(fi, rhs, eps=1e-300, ficycle=2, nslab=1)
            per_cycle = (time.perf_counter() - t1) / 2
            rows[f"L{L}_1core"] = {"solves_per_s": 1.0 / (per_cycle * n), "s_per_cycle": per_cycle,
                                   "cycles": n, "threads": 1,
                                   "note": "2 cycles timed on one core x the cycle count of the complete solve"}
        H.close()
        print(f"[bench] cpu {cfg.name} L{L}: {rows}", file=sys.stderr, flush=True)
    out[f"{cfg.name}_L"] = rows
    main = rows[f"L{cfg.nlevel}_all_cores"]
    out.update({
        "value": main["solves_per_s"],
        "unit": "solves/s",
        "cores": main["threads"],
        "kind": "port",
        "sample": (f"complete solves of the reference algorithm (CPU oracle: point SOR 1.6, 5/10 sweeps, 5 ghost "
                   f"passes per sweep, max-norm stop at 0.4 tol max|b|) on the {cfg.nx}x{cfg.ny}x{cfg.nz} case, "
                   f"L={cfg.nlevel}, {main['threads']} k-slab threads: {main['cycles']} cycles, "
                   f"{main['s_per_solve']:.1f} s, final rel. L2 residual {main['rel_res']:.1e}"),
    })
    return out


def bench_pressure_bar(cfg, ref_fi, steps):
    """Per precision: solves/s at 1e-11 and 1e-12 and the pressure error
    against the FP64 solve continued to 1e-13; the loosest tolerance that
    meets the 1e-9 pressure bar."""
    res = {}
    for prec in ("mixed", "fp64"):
        rows = {}
        with solver_for(cfg, prec) as s:
            s.case_rhs(cfg.seed)
            for tol in (1e-11, 1e-12):
                dt, r, cyc = time_solves(s, steps, 1, tol, lambda sv: sv.zero_fi())
                rows[f"{tol:.0e}"] = {"solves_per_s": steps / dt, "cycles": cyc[-1], "rel_res": r.rel_res,
                                      "pressure_rel_l2": rel_l2_meanfree(s.get_fi(), ref_fi)}
        res[prec] = rows
    return res


def bench_dropin(cfg, steps):
    """The host-array entry points at the benchmark size: ples_mg_solve
    (fi, rhs and cs from host arrays over PCIe, projection, solve, fi back
    with its ghosts) and ples_mg_projection_step (update, contra, diver,
    projection, solve, gradie, vel_up, check_divergence, all fields in and
    out)."""
    from ples_amd import mg

    out = {}
    x, y, z = cfg.nodes()
    with solver_for(cfg, "mixed") as s:
        s.case_rhs(cfg.seed)
        rhs = s.get_rhs()
        fi = np.zeros(s.fi_shape)
        s.multi(fi.copy(), rhs, eps=cfg.tol, ficycle=400)  # warm-up
        t = time.perf_counter()
        for _ in range(steps):
            f = np.zeros(s.fi_shape)
            r = s.multi(f, rhs, eps=cfg.tol, ficycle=400)
            assert r.ok, r
        dt = (time.perf_counter() - t) / steps
        out["solve_host_arrays"] = {"solves_per_s": 1.0 / dt, "ms_per_solve": 1e3 * dt, "cycles": r.cycles,
                                    "rel_res": r.rel_res}
        rng = np.random.default_rng(5)
        sh = s.fi_shape
        vel = [rng.standard_normal(sh) for _ in range(3)]
        dvel = [1e-2 * rng.standard_normal(sh) for _ in range(3)]
        fc = [np.zeros(a) for a in s.face_shapes()]
        s.contra(vel, fc)
        f = np.zeros(sh)
        s.projection_step([v.copy() for v in vel], dvel, [a.copy() for a in fc], f, dt=1.0, eps=cfg.tol)
        n2 = max(2, steps // 2)
        t = time.perf_counter()
        for _ in range(n2):
            v2 = [v.copy() for v in vel]
            f2 = [a.copy() for a in fc]
            f = np.zeros(sh)
            r, _, div = s.projection_step(v2, dvel, f2, f, dt=1.0, eps=cfg.tol)
            assert r.ok, r
        dt = (time.perf_counter() - t) / n2
        out["projection_step_host_arrays"] = {"steps_per_s": 1.0 / dt, "ms_per_step": 1e3 * dt,
                                              "cycles": r.cycles, "divmax_inside": float(div[0])}
    return out


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
    ap.add_argument("--no-extra", action="store_true", help="skip c1, the pressure bar and the drop-in lines")
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
    read_gbs = mg.read_bandwidth(1 << 30, 20)
    streams_gbs = mg.streams_bandwidth(1 << 30, 20)
    print(f"[bench] measured copy bandwidth {copy_gbs:.0f} GB/s, read {read_gbs:.0f} GB/s, "
          f"12 streams {streams_gbs:.0f} GB/s", file=sys.stderr, flush=True)

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

    extra = {}
    if not args.no_extra:
        bar = bench_pressure_bar(cfg, out["ref_fi"], max(3, args.steps // 2))
        for prec in ("mixed", "fp64"):
            bar[prec][f"{cfg.tol:.0e}"] = {"solves_per_s": args.steps / out[prec]["seconds"],
                                           "cycles": out[prec]["cycles"][-1], "rel_res": out[prec]["res"].rel_res,
                                           "pressure_rel_l2": rel_l2_meanfree(out[prec]["fi"], out["ref_fi"])}
            ok = [t for t in ("1e-10", "1e-11", "1e-12") if t in bar[prec] and bar[prec][t]["pressure_rel_l2"] <= 1e-9]
            bar[prec]["loosest_tol_meeting_1e-9"] = ok[0] if ok else None
        sp = {}
        if bar["mixed"]["loosest_tol_meeting_1e-9"] and bar["fp64"]["loosest_tol_meeting_1e-9"]:
            sp = bar["mixed"][bar["mixed"]["loosest_tol_meeting_1e-9"]]["solves_per_s"] / \
                bar["fp64"][bar["fp64"]["loosest_tol_meeting_1e-9"]]["solves_per_s"]
        bar["speedup_mixed_vs_fp64_at_the_bar"] = sp or None
        extra["pressure_bar"] = bar
        print(f"[bench] pressure bar: {bar}", file=sys.stderr, flush=True)
        extra["dropin"] = bench_dropin(cfg, max(3, args.steps // 2))
        print(f"[bench] dropin: {extra['dropin']}", file=sys.stderr, flush=True)
        c1cfg = PC.CONFIGS["c1"]
        c1 = {}
        for prec in ("fp64", "mixed"):
            with solver_for(c1cfg, prec) as s:
                s.case_rhs(c1cfg.seed)
                dt, r, cyc = time_solves(s, args.steps, args.warmup, c1cfg.tol, from_zero)
                c1[prec] = {"solves_per_s": args.steps / dt, "ms_per_solve": 1e3 * dt / args.steps,
                            "cycles": cyc[-1], "rel_res": r.rel_res}
        extra["c1"] = c1
        print(f"[bench] c1: {c1}", file=sys.stderr, flush=True)

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
        c3["pressure_rel_l2_vs_fp64_1e-13"] = {p: rel_l2_meanfree(c3_fi[p], c3_fi["ref"]) for p in ("mixed", "fp64")}
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
        # 24 cases on the 16 warm slots: admission evicts idle cases
        c5["admission_bound_24_cases"] = bench_c5(n_req=200, n_cases=24)
        print(f"[bench] c5 (24 cases): {c5['admission_bound_24_cases']}", file=sys.stderr, flush=True)

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
        "data": "synthetic (seeded Gaussian divergence rhs generated on device, projected in every step)",
        "config": {"workload": cfg.text, "cells": cfg.cells, "levels": cfg.nlevel, "tolerance": cfg.tol,
                   "launch_shapes": "tuning table" if tuned else "tuned at create",
                   "precision": ("mixed: FP32 V-cycles (the smoothers and the fused V-cycle residual read "
                                 "16-bit copies of the metrics, the smoothers 16-bit line spikes; the FP32 "
                                 "Krylov operator FP32 metrics), exact coarse solve with FP64 vectors and an "
                                 "FP32 dense map, FP64 refinement, outer residual and rhs projection; fp64: "
                                 "FP64 throughout (FP32 line factors in the smoother)")},
        "fp64_solves_per_s": args.steps / fp64["seconds"],
        "speedup_mixed_vs_fp64": fp64["seconds"] / mixed["seconds"],
        "cycles": {"mixed": mixed["cycles"][-1], "fp64": fp64["cycles"][-1]},
        "rel_res": {"mixed": mixed["res"].rel_res, "fp64": fp64["res"].rel_res},
        "pressure_rel_l2_vs_fp64_1e-13": {p: rel_l2_meanfree(out[p]["fi"], out["ref_fi"]) for p in ("mixed", "fp64")},
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
            "read_measured_GBs": read_gbs,
            "frac_of_copy": achieved / copy_gbs if achieved else None,
            "streams12_measured_GBs": streams_gbs,
            "frac_of_streams12": achieved / streams_gbs if achieved else None,
            "launches": n,
            "timed_steps": f"1 in {TIMING_EVERY} refinement steps (eager, per-dispatch events); the rest graph-replayed",
            "avg_us": 1e3 * avg_ms,
            "bytes_per_launch": by / max(n, 1),
            "resid_fine_GBs": resid_gbs,
        },
        "cpu_baseline": cpu,
        **extra,
        "c3": c3,
        "c4_batch": c4,
        "c5_serving": c5,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
