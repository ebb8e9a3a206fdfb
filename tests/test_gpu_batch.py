"""Batched solves (ples_mg_batch_*): every problem of a batch of uneven
sizes converges and its pressure is bitwise the one the same handle options
give when the problem is solved alone, whatever the number of refinement
steps per graph launch, with grouped launches (one launch per kernel over the
problems of one shape) and with every problem on its own stream.  (The single
solve itself is checked against the oracle in test_gpu_parity.)"""
import numpy as np
import pytest

from ples_amd import cases as PC
from ples_amd import mg

pytestmark = pytest.mark.gpu

SIZES = [(32, 16, 32, 3), (48, 24, 48, 4), (64, 32, 64, 4), (32, 32, 32, 5), (16, 8, 16, 2), (64, 16, 32, 3)]


def _problems(n=6):
    return PC.c4_problems(n=n, seed0=2100, sizes=SIZES)


def _alone(pr, precision, tol, fp32_krylov):
    with mg.Solver(pr["x"], pr["y"], pr["z"], 0, 1, 0, pr["nlevel"], precision=precision,
                   fp32_krylov=fp32_krylov) as s:
        s.case_rhs(pr["seed"])
        s.zero_fi()
        r = s.solve_resident(tol, 200)
        return r, s.get_fi()


# mixed with fp32_krylov 1 (default): the batch runs the single solve's
# iterative refinement (inner FP32 GCR) driven on the device; with 0, FP64 GCR
# steps around the FP32 V-cycle, as the single solve with that option
@pytest.mark.parametrize("precision,grouped,fp32_krylov",
                         [("mixed", True, 1), ("mixed", False, 1), ("mixed", True, 0), ("fp64", True, 1)])
def test_batch_equals_alone(precision, grouped, fp32_krylov):
    probs = _problems(10)   # 10 problems over 6 shapes: some families hold several
    shapes = {(p["shape"], p["nlevel"]) for p in probs}
    assert len(shapes) < len(probs)
    tol = 1e-8
    ref = [_alone(pr, precision, tol, fp32_krylov) for pr in probs]
    with mg.Batch(probs, precision=precision, grouped=grouped, fp32_krylov=fp32_krylov) as b:
        for steps in (2, 1, 4):
            for pr, s in zip(probs, b.solvers):
                s.case_rhs(pr["seed"])
                s.zero_fi()
            res = b.solve(tol, 200, steps_per_launch=steps)
            for q, (r, s) in enumerate(zip(res, b.solvers)):
                assert r.ok and r.rel_res < tol, (steps, q, r)
                assert r.cycles == ref[q][0].cycles, (steps, q, r.cycles, ref[q][0].cycles)
                assert np.array_equal(s.get_fi(), ref[q][1]), (steps, q)
            assert b.families() == (len(shapes) if grouped else 0), (steps, b.families())


def test_batch_warm_restart_and_mixed_tolerances():
    probs = _problems()
    with mg.Batch(probs, precision="mixed") as b:
        for pr, s in zip(probs, b.solvers):
            s.case_rhs(pr["seed"])
            s.zero_fi()
        first = b.solve(1e-10, 200)
        assert all(r.ok for r in first)
        # warm start: perturbed solutions converge in fewer cycles
        for q, s in enumerate(b.solvers):
            s.perturb_fi(7000 + q, 1e-3)
        second = b.solve(1e-10, 200)
        for r1, r2 in zip(first, second):
            assert r2.ok and r2.cycles < r1.cycles, (r1, r2)
