"""The batched, serving and time-step workloads at their full sizes.

* 64 uneven instances (grids 32-128^3, 50-2000 rods of 6-12 spheres,
  log-uniform; FP32+refine, tol 1e-8): batched equals alone bitwise, and the
  residual of every converged instance is computed with the oracle's
  saddle-point product (gmres_solver.cpp:72-83).
* A Poisson stream of 40 requests (64-256^3, 100-20,000 rods of 10 spheres,
  kT = 1): every served result equals the same request solved alone, bitwise,
  and converged results have an oracle residual below the tolerance.
* 512^3 with 150,000 rods of 12 spheres (1.8 M blobs, kT = 1, sterics cut off
  at 2.2a): the device cell list's barrier forces against the oracle's brute
  force on a sub-box, then two Euler-Maruyama steps with property checks
  (converged, true FP64 residual below tol, finite unit quaternions, bodies
  moved).

Instances whose rod counts cannot fit their box (volume fraction above 1:
2,000 rods of 12 in 32^3) stop at the iteration cap in FP32+refine and FP64
alike; they are still checked bitwise and by status."""
import math

import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from timstochrigidbodiessim_amd import serve

pytestmark = pytest.mark.gpu

DT = np.pi * 1.86 ** 3 / 100.0
CAP = 800


def say(*a):
    print(*a, flush=True)


def _oracle_residual(n, ref, x, q, Minv, N, rhs, sol, kT=0.0):
    oc = po.cfg(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=kT, dt=DT)
    o = po.Fcm(oc, len(x) * len(ref))
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    b = po.Bodies(ref, x, q, Minv, N)
    r = np.linalg.norm(rhs - po.saddle_matvec(o, b, sol)) / np.linalg.norm(rhs)
    o.close()
    return r


def test_config3_64_uneven_instances_batched():
    rng = np.random.default_rng(3)
    insts = []
    say("config 3: set-up")
    for i in range(64):
        n = int(rng.choice([32, 48, 64, 96, 128]))
        NB = int(round(math.exp(rng.uniform(math.log(50), math.log(2000)))))
        nbpb = int(rng.integers(6, 13))
        cfg = fa.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=0.0, dt=DT, max_blobs=NB * nbpb, max_bodies=NB,
                             max_blobs_per_body=nbpb, max_krylov=CAP, fp32=True, random_field=False)
        ref, x, q, F6, _ = fa.synth_rods(100 + i, NB, nbpb, 1.86, [float(n)] * 3, max_attempts=5)
        s = fa.FcmSolver(cfg)
        s.set_geometry(ref)
        _, Minv, N = s.ref_matrices()
        s.set_bodies(x, q)
        s.set_body_forces(F6)
        s.assemble_rhs(use_random=False)
        insts.append((s, n, ref, x, q, F6, Minv, N))
    try:
        alone = []
        for k, (s, *_) in enumerate(insts):
            say("config 3: alone", k)
            sol, rep = s.solve(None, tol=1e-8, max_iter=CAP, precision=fa.FCM_PREC_FP32)
            alone.append((sol, rep))
        say("config 3: batch")
        with fa.FcmBatch(max_jobs=64, worker_streams=4, steps_per_round=4) as b:
            for s, *_ in insts:
                b.submit(s, tol=1e-8, max_iter=CAP, precision=fa.FCM_PREC_FP32)
            b.run()
            done = {id(s): rep for s, rep in b.take()}
        assert len(done) == 64
        conv = 0
        for (s, n, ref, x, q, F6, Minv, N), (sol, rep) in zip(insts, alone):
            brep = done[id(s)]
            assert brep["iterations"] == rep["iterations"] and brep["status"] == rep["status"]
            assert np.array_equal(s.solution(), sol), "batched solution differs from the solve alone"
            phi = len(x) * len(ref) * 4 / 3 * np.pi * 1.86 ** 3 / n ** 3
            if rep["status"] == "converged":
                conv += 1
                rhs = np.concatenate([np.zeros(3 * len(x) * len(ref)), -F6.ravel()])
                r = _oracle_residual(n, ref, x, q, Minv, N, rhs, sol)
                say(f"  {n}^3 {len(x)} x {len(ref)} phi {phi:.2f}: {rep['iterations']} it, oracle residual {r:.3e}")
                assert r < 1.01e-8, r
            else:
                say(f"  {n}^3 {len(x)} x {len(ref)} phi {phi:.2f}: {rep['status']} at {rep['iterations']}")
                assert rep["status"] == "max_iter"
        say("config 3: converged", conv, "of 64")
        assert conv >= 48
    finally:
        for s, *_ in insts:
            s.close()


def test_config4_poisson_stream_served():
    reqs = serve.poisson_requests(rate=4.0, count=40, seed=11, grids=(64, 96, 128, 192, 256), rods=(100, 20000),
                                  nbpb=10, precision=fa.FCM_PREC_FP32, tol=1e-8, max_iter=CAP)
    for r in reqs:
        r.make_geometry()
    say("config 4: serving")
    with serve.MobilityServer(max_jobs=16, worker_streams=4, steps_per_round=4, max_krylov=CAP,
                              keep_solutions=True) as srv:
        srv.prepare_refs(reqs)
        st = srv.run(reqs)
        say("config 4:", st, "max active", srv.max_active)
        assert st["requests"] == 40 and srv.max_active > 1
        conv = 0
        for r in reqs:
            key, cfg = srv.config_for(r)
            ref, x, q, F6 = r.geometry
            Minv, N = srv.refs[(r.n, r.nbpb, r.a)]
            say("config 4: alone", r.rid, r.n, r.NB)
            with fa.FcmSolver(cfg) as s:
                s.set_geometry(ref)
                s.set_bodies(x, q)
                s.set_ref_matrices(Minv, N)
                s.set_body_forces(F6)
                s.mobility_prepare(r.seed)
                sol, rep = s.solve(None, tol=r.tol, max_iter=r.max_iter, precision=r.precision)
                Vr = s.random_interp()
            assert r.report["iterations"] == rep["iterations"] and r.report["status"] == rep["status"]
            assert np.array_equal(r.solution, sol), f"request {r.rid} differs from its solve alone"
            if rep["status"] == "converged":
                conv += 1
                rhs = np.concatenate([-Vr, -np.asarray(F6).ravel()])
                res = _oracle_residual(r.n, ref, x, q, Minv, N, rhs, sol, kT=1.0)
                say(f"  {r.n}^3 {r.NB} rods: {rep['iterations']} it, oracle residual {res:.3e}")
                assert res < 1.01e-8, res
            else:
                assert rep["status"] == "max_iter"
        say("config 4: converged", conv, "of 40")
        assert conv >= 30


def test_config5_512_cubed_time_steps():
    n, NB, nbpb, a = 512, 150000, 12, 1.86
    cutoff = 2.2 * a
    say("config 5: geometry")
    ref, x, q, _, nov = fa.synth_rods(5, NB, nbpb, a, [float(n)] * 3, max_attempts=20)
    nb = NB * nbpb
    cfg = fa.make_config(n, ng=7, dx=1.0, a=a, mu=1.0, kT=1.0, dt=DT, max_blobs=nb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=600, fp32=True, random_field=True)
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        s.set_body_forces(np.zeros((NB, 6)))
        say("config 5: barrier forces")
        F = s.barrier_forces(cutoff)
        Y = s.blob_positions().reshape(-1, 3)
        # the oracle's brute force on every body with a blob in a 80^3 corner
        # sub-box: the forces on blobs farther than the cutoff inside it are
        # complete there (all their neighbours lie in the sub-box's bodies)
        inbox = np.all(Y < 80.0, axis=1).reshape(NB, nbpb).any(axis=1)
        bodies = np.nonzero(inbox)[0]
        sel = (bodies[:, None] * nbpb + np.arange(nbpb)[None, :]).ravel()
        oc = po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=1.0, dt=DT)
        Fo = po.barrier_forces(oc, Y[sel], nbpb, cutoff).reshape(-1, 3)
        inner = np.all((Y[sel] > cutoff + 0.5) & (Y[sel] < 80.0 - cutoff - 0.5), axis=1)
        say(f"  {len(sel)} blobs in the sub-box, {inner.sum()} checked, "
            f"{np.count_nonzero(np.abs(Fo[inner]).sum(1))} with a force")
        assert inner.sum() > 1000 and np.abs(Fo[inner]).max() > 0
        e = np.linalg.norm(F[sel][inner] - Fo[inner]) / np.linalg.norm(Fo[inner])
        say("  barrier forces vs oracle", e)
        assert e < 1e-12, e
        x0, q0 = s.bodies()
        for k in range(2):
            say("config 5: step", k)
            rep = s.em_step(1 + k, cutoff, tol=1e-8, max_iter=600, precision=fa.FCM_PREC_FP32)
            say("  ", rep)
            assert rep["status"] == "converged"
            assert 0 < rep["true_rel_residual"] < 1e-8
        x1, q1 = s.bodies()
    assert np.all(np.isfinite(x1)) and np.all(np.isfinite(q1))
    assert np.allclose(np.linalg.norm(q1, axis=1), 1.0, atol=1e-12)
    d = np.asarray(x1) - np.asarray(x0)
    d -= n * np.round(d / n)
    step = np.linalg.norm(d, axis=1)
    say("  body displacement over two steps: mean", step.mean(), "max", step.max())
    assert np.all(step > 0) and step.max() < 5.0
