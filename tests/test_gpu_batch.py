"""Batched solves (fcm_batch_*): many independent, uneven instances advanced
together give bitwise the same solutions, iteration counts and refinement
counts as the same solves run one at a time, including jobs submitted while
others are in flight (continuous batching).  The single solve is itself
checked against the CPU oracle in test_gpu_gmres.py."""
import numpy as np
import pytest

import timstochrigidbodiessim_amd as fa

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_instance(seed, n, NB, nbpb, kT=1.0, max_krylov=200):
    cfg = fa.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=kT, dt=0.20216, max_blobs=NB * nbpb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=max_krylov, fp32=True, random_field=True)
    ref, x, q, F6, _ = fa.synth_rods(seed, NB, nbpb, 1.86, [float(n)] * 3, max_attempts=50)
    s = fa.FcmSolver(cfg)
    s.set_geometry(ref)
    _, Minv, N = s.ref_matrices()
    s.set_bodies(x, q)
    s.set_ref_matrices(Minv, N)
    s.set_body_forces(F6)
    return s


# (grid, rods, spheres per rod, precision); every rod is shorter than its box
# (a rod of 12 spheres spans 44.6 > 32, which makes the system nearly singular)
CASES = [(32, 20, 6, fa.FCM_PREC_FP64), (48, 40, 8, fa.FCM_PREC_FP32), (64, 60, 10, fa.FCM_PREC_FP64),
         (32, 30, 7, fa.FCM_PREC_FP32), (64, 100, 12, fa.FCM_PREC_FP32), (48, 15, 9, fa.FCM_PREC_FP64),
         (96, 80, 10, fa.FCM_PREC_FP32), (32, 10, 6, fa.FCM_PREC_FP32)]


def _alone(insts):
    out = []
    for i, (s, prec) in enumerate(insts):
        s.mobility_prepare(100 + i)
        sol, rep = s.solve(None, tol=1e-8, max_iter=200, precision=prec)
        out.append((sol, rep))
    return out


@pytest.mark.parametrize("workers", [4, 0])
def test_batch_equals_alone(workers):
    say("batch vs alone, worker streams", workers)
    insts = [(make_instance(7 + i, n, NB, nbpb), prec) for i, (n, NB, nbpb, prec) in enumerate(CASES)]
    alone = _alone(insts)
    for (sol, rep), (n, NB, nbpb, prec) in zip(alone, CASES):
        say(n, NB, nbpb, prec, rep)
        assert rep["status"] == "converged"
    with fa.FcmBatch(max_jobs=16, worker_streams=workers, steps_per_round=3) as b:
        for i, (s, prec) in enumerate(insts):
            s.mobility_prepare(100 + i)
            b.submit(s, tol=1e-8, max_iter=200, precision=prec)
        b.run()
        done = b.take()
    assert len(done) == len(insts)
    by_solver = {id(s): rep for s, rep in done}
    for (s, prec), (sol, rep) in zip(insts, alone):
        brep = by_solver[id(s)]
        assert brep["status"] == rep["status"]
        assert brep["iterations"] == rep["iterations"], (brep, rep)
        assert brep["refinements"] == rep["refinements"]
        assert np.array_equal(s.solution(), sol), "batched solution differs from the solve alone"
    for s, _ in insts:
        s.close()


def test_continuous_batching_submit_in_flight():
    say("continuous batching")
    insts = [(make_instance(31 + i, n, NB, nbpb), prec) for i, (n, NB, nbpb, prec) in enumerate(CASES[:6])]
    alone = _alone(insts)
    finished = {}
    with fa.FcmBatch(max_jobs=4, worker_streams=2, steps_per_round=2) as b:
        pending = list(range(len(insts)))
        active = 0
        steps = 0
        while pending or active:
            # admit while slots are free (fewer slots than jobs: later ones wait)
            while pending and active < 4:
                i = pending.pop(0)
                s, prec = insts[i]
                s.mobility_prepare(100 + i)
                b.submit(s, tol=1e-8, max_iter=200, precision=prec)
                active += 1
            b.step()
            steps += 1
            for s, rep in b.take():
                finished[id(s)] = rep
                active -= 1
            assert steps < 5000
    for (s, prec), (sol, rep) in zip(insts, alone):
        brep = finished[id(s)]
        assert brep["iterations"] == rep["iterations"]
        assert np.array_equal(s.solution(), sol)
    for s, _ in insts:
        s.close()


def test_batch_zero_rhs_and_iteration_cap():
    say("batch edge cases")
    s1 = make_instance(3, 32, 10, 6, kT=0.0)
    s2 = make_instance(4, 32, 12, 6)
    s1.set_body_forces(np.zeros((10, 6)))
    s1.assemble_rhs(use_random=False)
    s2.mobility_prepare(9)
    with fa.FcmBatch(max_jobs=4, worker_streams=2, steps_per_round=4) as b:
        b.submit(s1, tol=1e-8, max_iter=50, precision=fa.FCM_PREC_FP32)
        b.submit(s2, tol=1e-14, max_iter=7, precision=fa.FCM_PREC_FP64)
        b.run()
        reps = {id(s): r for s, r in b.take()}
    assert reps[id(s1)]["status"] == "zero_rhs"
    assert not np.any(s1.solution())
    assert reps[id(s2)]["status"] == "max_iter" and reps[id(s2)]["iterations"] == 7
    s1.close()
    s2.close()


def make_channel_instance(seed, n, NB, nbpb, max_krylov=200):
    """Rods with every sphere between slip walls at y = 0 and y = n/2."""
    cfg = fa.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=1.0, dt=0.20216, max_blobs=NB * nbpb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=max_krylov, fp32=True, random_field=True,
                         slip_channel=True)
    margin = 1.86 * (nbpb - 1) + 0.5
    ref, x, q, F6, _ = fa.synth_rods(seed, NB, nbpb, 1.86, [float(n), n / 2 - 2 * margin, float(n)], max_attempts=50)
    x[:, 1] += margin
    s = fa.FcmSolver(cfg)
    s.set_geometry(ref)
    _, Minv, N = s.ref_matrices()
    s.set_bodies(x, q)
    s.set_ref_matrices(Minv, N)
    s.set_body_forces(F6)
    return s


def test_batch_mixes_channel_and_periodic_instances():
    """A batch advances each handle with its own geometry: slip-channel and
    periodic instances together give bitwise their solves alone."""
    say("batch, channel and periodic")
    insts = [(make_channel_instance(51, 64, 30, 6), fa.FCM_PREC_FP32), (make_instance(52, 32, 20, 6), fa.FCM_PREC_FP64),
             (make_channel_instance(53, 48, 20, 4), fa.FCM_PREC_FP64), (make_instance(54, 48, 40, 8), fa.FCM_PREC_FP32)]
    alone = _alone(insts)
    for sol, rep in alone:
        assert rep["status"] == "converged", rep
    with fa.FcmBatch(max_jobs=8, worker_streams=2, steps_per_round=3) as b:
        for i, (s, prec) in enumerate(insts):
            s.mobility_prepare(100 + i)
            b.submit(s, tol=1e-8, max_iter=200, precision=prec)
        b.run()
        done = b.take()
    by_solver = {id(s): rep for s, rep in done}
    for (s, prec), (sol, rep) in zip(insts, alone):
        assert by_solver[id(s)]["iterations"] == rep["iterations"]
        assert np.array_equal(s.solution(), sol)
    for s, _ in insts:
        s.close()


@pytest.mark.parametrize("steps_per_round", [1, 2, 5])
def test_round_length_does_not_change_results(steps_per_round):
    """The step that finalises a converging inner solve is a step of its own
    (its mat-vec gated off), so it can fall anywhere within a round of graph
    steps, or open one: every round length gives bitwise the solve alone,
    including an iteration cap reached by a finalising step."""
    say("round length", steps_per_round)
    cases = [(32, 20, 6, fa.FCM_PREC_FP64, 200), (48, 40, 8, fa.FCM_PREC_FP32, 200), (32, 30, 7, fa.FCM_PREC_FP64, 9)]
    insts = [(make_instance(51 + i, n, NB, nbpb), prec, cap) for i, (n, NB, nbpb, prec, cap) in enumerate(cases)]
    alone = []
    for i, (s, prec, cap) in enumerate(insts):
        s.mobility_prepare(200 + i)
        alone.append(s.solve(None, tol=1e-8, max_iter=cap, precision=prec))
    assert alone[2][1]["status"] == "max_iter" and alone[2][1]["iterations"] == 9
    with fa.FcmBatch(max_jobs=4, worker_streams=0, steps_per_round=steps_per_round) as b:
        for i, (s, prec, cap) in enumerate(insts):
            s.mobility_prepare(200 + i)
            b.submit(s, tol=1e-8, max_iter=cap, precision=prec)
        b.run()
        reps = {id(s): r for s, r in b.take()}
    for (s, prec, cap), (sol, rep) in zip(insts, alone):
        brep = reps[id(s)]
        assert (brep["status"], brep["iterations"], brep["refinements"]) == (
            rep["status"], rep["iterations"], rep["refinements"]), (brep, rep)
        assert np.array_equal(s.solution(), sol)
    for s, _, _ in insts:
        s.close()
