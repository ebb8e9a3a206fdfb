"""Serving loop (ples_mg_server_*): requests that run concurrently with
others give bitwise the results they give alone (same case history, same
tolerance), every request meets its own tolerance, warm starts are used for
resident cases, and the HBM-budget admission evicts idle cases (cold starts)
without affecting correctness."""
import numpy as np
import pytest

from ples_amd import cases as PC
from ples_amd import mg

pytestmark = pytest.mark.gpu

SHAPES = [(32, 16, 32, 3), (48, 24, 48, 4), (64, 32, 64, 4), (40, 20, 40, 3), (16, 8, 16, 2), (32, 32, 32, 4)]


def _cases():
    out = []
    for c, (nx, ny, nz, L) in enumerate(SHAPES):
        x, y, z = PC.make_nodes(nx, ny, nz, 1.5, 0.2)
        out.append(dict(case_id=c, x=x, y=y, z=z, nlevel=L, rhs_seed=4000 + c))
    return out


def _requests():
    # three time steps per case, interleaved, mixed tolerances
    reqs = []
    for step in range(3):
        for c in range(len(SHAPES)):
            reqs.append((c, 5000 + 10 * c + step, 1e-6 if (c + step) % 2 else 1e-10))
    return reqs


def _server(cases, **kw):
    sv = mg.Server(**kw)
    for c in cases:
        sv.add_case(c["case_id"], c["x"], c["y"], c["z"], c["nlevel"], c["rhs_seed"])
    return sv


def test_results_same_alone_and_in_the_mix():
    cases = _cases()
    reqs = _requests()
    # in the mix: everything submitted at once, up to 4 running together
    with _server(cases, max_active=4, warm_cases=16) as sv:
        tickets = [sv.submit(c, seed, 1e-3, eps, 200) for c, seed, eps in reqs]
        sv.drain()
        mix = [sv.poll(t) for t in tickets]
        fi_mix = {c["case_id"]: sv.get_fi(c["case_id"]) for c in cases}
        st = sv.stats()
    assert st["completed"] == len(reqs) and st["evictions"] == 0
    for (c, seed, eps), (r, warm) in zip(reqs, mix):
        assert r.ok and r.rel_res < eps, (c, r)
    # warm starts from the second step of each case on
    assert [w for _, w in mix] == [False] * len(SHAPES) + [True] * (2 * len(SHAPES))
    # alone: one request at a time, each drained before the next
    with _server(cases, max_active=1, warm_cases=16) as sv:
        alone = []
        for c, seed, eps in reqs:
            t = sv.submit(c, seed, 1e-3, eps, 200)
            sv.drain()
            alone.append(sv.poll(t))
        fi_alone = {c["case_id"]: sv.get_fi(c["case_id"]) for c in cases}
    for (r1, w1), (r2, w2) in zip(mix, alone):
        assert r1.cycles == r2.cycles and w1 == w2
        assert r1.rel_res == r2.rel_res
    for c in fi_mix:
        assert np.array_equal(fi_mix[c], fi_alone[c]), c


def test_hbm_budget_admission_evicts_idle_cases():
    cases = _cases()[:4]
    with mg.Solver(cases[2]["x"], cases[2]["y"], cases[2]["z"], 0, 1, 0, cases[2]["nlevel"], precision="mixed") as s:
        per_case = s.device_bytes
    # room for about two of the largest cases at a time
    with _server(cases, max_active=2, warm_cases=16, hbm_budget=int(2.6 * per_case + (128 << 20))) as sv:
        tickets = []
        for step in range(2):
            for c in range(4):
                tickets.append(sv.submit(c, 6000 + 10 * c + step, 1e-3, 1e-8, 200))
        sv.drain()
        out = [sv.poll(t) for t in tickets]
        st = sv.stats()
    assert all(r.ok and r.rel_res < 1e-8 for r, _ in out)
    assert st["evictions"] > 0 and st["cold_starts"] > 4
    assert st["used_bytes"] <= int(2.6 * per_case + (128 << 20))


def test_served_case_resumes_from_its_pressure_and_restart_file(tmp_path):
    """A case resumed on a fresh server from its pressure (ples_mg_server_set_fi)
    gives bitwise the result of continuing on the server that solved it; resumed
    from the flow-field restart file (ten significant digits) it starts warm,
    meets its tolerance and needs fewer cycles than a cold start."""
    c = _cases()[2]
    cid = c["case_id"]
    steps = [(7000, 1e-10), (7001, 1e-10), (7002, 1e-10)]

    def run(sv, seed, eps):
        t = sv.submit(cid, seed, 1e-3, eps, 200)
        sv.drain()
        return sv.poll(t)

    with _server([c], max_active=1) as sv:
        for seed, eps in steps[:2]:
            assert run(sv, seed, eps)[0].ok
        fi = sv.get_fi(cid)
        cont, w = run(sv, *steps[2])
        assert w and cont.ok
        fi_cont = sv.get_fi(cid)
    with _server([c], max_active=1) as sv:
        sv.set_fi(cid, fi)  # not resident yet: kept until admission
        r, w = run(sv, *steps[2])
        assert w and r.cycles == cont.cycles and r.rel_res == cont.rel_res
        assert np.array_equal(sv.get_fi(cid), fi_cont)
        sv.set_fi(cid, fi)  # resident: copied at once
        r, w = run(sv, *steps[2])
        assert w and r.cycles == cont.cycles and r.rel_res == cont.rel_res

    zc = np.zeros(fi.shape)
    jz, jy, jx = (d - 2 for d in fi.shape)
    path = str(tmp_path / "new_res_form")
    mg.restart_write(path, 1.5, 0.0, 0.0, zc, zc, zc, fi, None, np.zeros((jz, jy, jx + 1)),
                     np.zeros((jz, jy + 1, jx)), np.zeros((jz + 1, jy, jx)))
    fi_file = mg.restart_read(path, jx, jy, jz, 0)[6]
    assert np.allclose(fi_file, fi, rtol=5.0001e-10, atol=0)
    with _server([c], max_active=1) as sv:
        cold, w = run(sv, *steps[2])
        assert not w and cold.ok
    with _server([c], max_active=1) as sv:
        sv.set_fi(cid, fi_file)
        r, w = run(sv, *steps[2])
        assert w and r.ok and r.rel_res < 1e-10
        assert abs(r.cycles - cont.cycles) <= 1 and r.cycles < cold.cycles
