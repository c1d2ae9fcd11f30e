"""Serving loop: requests of different grids and body counts arriving while
others are on the GPU are admitted by HBM footprint, and each result is
bitwise the same as the same request solved alone on its own handle."""
import numpy as np
import pytest

import timstochrigidbodiessim_amd as fa
from timstochrigidbodiessim_amd import serve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _requests(count, precision_mix=True):
    reqs = serve.poisson_requests(rate=2000.0, count=count, seed=5, grids=(48, 64, 96), rods=(20, 300))
    for r in reqs:
        if precision_mix and r.rid % 2:
            r.precision = fa.FCM_PREC_FP64
        r.max_iter = 300
        r.make_geometry()
    return reqs


def _alone(srv, req):
    key, cfg = srv.config_for(req)
    ref, x, q, F6 = req.geometry
    Minv, N = srv.refs[(req.n, req.nbpb, req.a)]
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(F6)
        s.mobility_prepare(req.seed)
        sol, rep = s.solve(None, tol=req.tol, max_iter=req.max_iter, precision=req.precision)
    return sol, rep


def test_served_results_equal_alone():
    reqs = _requests(10)
    with serve.MobilityServer(max_jobs=6, worker_streams=4, steps_per_round=2, max_krylov=300,
                              keep_solutions=True) as srv:
        srv.prepare_refs(reqs)
        st = srv.run(reqs)
        say("serve stats", st, "max active", srv.max_active, "handles created", srv.created)
        assert st["requests"] == len(reqs)
        assert srv.max_active > 1, "requests never overlapped on the device"
        for r in reqs:
            sol, rep = _alone(srv, r)
            say(r.rid, r.n, r.NB, r.precision, r.report, "alone", rep["iterations"])
            assert r.report["status"] == "converged"
            assert r.report["iterations"] == rep["iterations"]
            assert np.array_equal(r.solution, sol), f"request {r.rid} differs from its solve alone"


def test_admission_respects_hbm_budget():
    reqs = _requests(6, precision_mix=False)
    biggest = 0
    with serve.MobilityServer(max_jobs=8, worker_streams=2, steps_per_round=2, max_krylov=300) as probe:
        for r in reqs:
            biggest = max(biggest, fa.footprint(probe.config_for(r)[1]))
    budget = int(2.2 * biggest)  # room for two of the largest handles at a time
    with serve.MobilityServer(hbm_budget=budget, max_jobs=8, worker_streams=2, steps_per_round=2,
                              max_krylov=300) as srv:
        srv.prepare_refs(reqs)
        st = srv.run(reqs)
        say("budgeted serve", st, "max active", srv.max_active, "used", srv.used, "budget", budget)
        assert st["requests"] == len(reqs)
        assert srv.max_used <= budget
