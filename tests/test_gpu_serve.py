"""Continuous-batching server: results are the same alone and in the mix,
requests overlap on the device, failures are reported per request."""
import numpy as np
import pytest

from conftest import HALDANE, HMCASE0, make_param, rel_l2

pytestmark = pytest.mark.gpu

SAME_TOL = 0.0  # a solve's arithmetic does not depend on what runs beside it: bitwise equal


def _requests(tmp_path):
    from hhgamd import tbgen
    files = {}
    for nb in (2, 4, 8):
        files[nb] = str(tmp_path / f"syn{nb}_tb.dat")
        tbgen.write_tb(files[nb], *tbgen.synthetic(nb, seed=40 + nb))
    out = [
        make_param(HALDANE, nr=(48, 60), tmax=12.0, tmax_field=10.0, pol=(0.6, 0.8, 0.0)),
        make_param(HMCASE0, nr=(30, 30), tmax=10.0, tmax_field=9.0),
        dict(make_param(files[4], nr=(12, 16), tmax=8.0, tmax_field=7.0, pol=(0.6, 0.8, 0.0)), precision="fp32"),
        make_param(files[8], nr=(8, 10), tmax=6.0, tmax_field=5.0, pol=(0.8, 0.6, 0.0)),
        make_param(files[2], nr=(20, 24), tmax=10.0, tmax_field=9.0, pol=(0.6, 0.8, 0.0)),
        dict(make_param(HALDANE, nr=(40, 40), tmax=15.0, tmax_field=12.0), precision="fp32"),
    ]
    return out * 2  # every request twice, so the mix holds duplicates


def served_alone(params, precision=None):
    """Each request through the server with nothing else in flight."""
    from hhgamd import serve
    out = []
    with serve.Server(hbm_budget_gb=64.0, max_inflight=1, host_threads=1) as one:
        for p in params:
            out.append(one.result(one.submit(p))[1])
    return out


def test_server_results_equal_alone(tmp_path):
    from hhgamd import lib, serve
    params = _requests(tmp_path)
    alone = served_alone(params)
    # a served solve takes the launch rules of batch members; against the same
    # solve run by itself through the C ABI it can differ only in the
    # summation order of its current
    for p, Ja in zip(params, alone):
        with lib.Solve(p) as s:
            s.run()
            assert rel_l2(Ja, s.current()[1]) <= 1e-13
    with serve.Server(hbm_budget_gb=8.0, max_inflight=4, host_threads=4) as srv:
        ids = [srv.submit(p) for p in params]
        for rid, Ja in zip(ids, alone):
            _, J = srv.result(rid)
            err = rel_l2(J, Ja)
            print(f"request {rid}: mix vs alone rel L2 = {err:.1e}")
            assert err <= SAME_TOL
        st = srv.stats()
        print(st)
        assert st["completed"] == len(params) and st["failed"] == 0
        assert st["peak_inflight"] > 1  # solves overlapped on the device


def test_server_vs_oracle(tmp_path, oracle_lib):
    from hhgamd import serve
    p = make_param(HMCASE0, nr=(12, 14), tmax=20.0, tmax_field=15.0, pol=(0.96, 0.25, 0.0))
    with serve.Server(hbm_budget_gb=2.0, max_inflight=2) as srv:
        _, J = srv.result(srv.submit(p))
    _, Jref = oracle_lib.run(HMCASE0, p, mode="perk")
    assert rel_l2(J, Jref) <= 1e-10


def test_server_failures_are_per_request(tmp_path):
    from hhgamd import serve
    good = make_param(HALDANE, nr=(8, 8), tmax=2.0, tmax_field=1.5)
    with serve.Server(hbm_budget_gb=0.01, max_inflight=2) as srv:
        bad_file = srv.submit(dict(good, path_tb=str(tmp_path / "missing_tb.dat")))
        too_big = srv.submit(make_param(HALDANE, nr=(240, 240)))
        ok = srv.submit(good)
        assert srv.wait(bad_file) == "failed" and "cannot open" in srv.error(bad_file)
        assert srv.wait(too_big) == "failed" and "budget" in srv.error(too_big)
        assert srv.wait(ok) == "done"
        t, J = srv.result(ok)
        assert len(t) == int(2.0 / 0.02197) and np.isfinite(J).all()


def test_admission_policies_give_the_same_results(tmp_path):
    """Arrival order and least-work-first admission change only when a request
    runs, never its result."""
    from hhgamd import serve
    params = _requests(tmp_path)[:6]
    out = {}
    for policy in ("fifo", "least-work"):
        with serve.Server(hbm_budget_gb=8.0, max_inflight=3, host_threads=4, policy=policy) as srv:
            ids = [srv.submit(p) for p in params]
            out[policy] = [srv.result(i)[1] for i in ids]
    for a, b in zip(out["fifo"], out["least-work"]):
        assert np.array_equal(a, b)
