"""GPU saddle-point solves against the oracle (gmres_solver.cpp:148-306).

FP64 plain solve: relative L2 error in blob forces and body velocities
<= 1e-10 against the oracle solve at the same tolerance.  FP32 bulk solve with
FP64 iterative refinement: true FP64 residual below the tolerance, and its
FP64-oracle error is stated (asserted <= 1e-6)."""
import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, rel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def setup_pair(nx, NB, nbpb, seed, kT=0.0, max_attempts=2000):
    gc, oc = pair_cfg(nx, kT=kT, dt=0.20216, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      max_krylov=150)
    L = [float(nx)] * 3
    ref, x, q, F6, nov = fa.synth_rods(seed, NB, nbpb, 1.86, L, max_attempts=max_attempts)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    b = po.Bodies(ref, x, q, Minv, N)
    s = fa.FcmSolver(gc)
    s.set_geometry(ref)
    s.set_bodies(x, q)
    s.set_ref_matrices(Minv, N)
    s.set_body_forces(F6)
    return s, o, b, F6, oc


def test_config1_fp64_solve_matches_oracle():
    """64^3, 100 rods x 10 spheres, a = 1.86, kT = 0, tol 1e-8 (config 1)."""
    say("config1 fp64")
    s, o, b, F6, oc = setup_pair(64, 100, 10, seed=1)
    rhs = np.concatenate([np.zeros(3000), -F6.ravel()])
    st, ref, it_o, hist_o = po.gmres(o, b, rhs, tol=1e-8, max_iter=150)
    assert st == 0
    sol, rep = s.solve(rhs, tol=1e-8, max_iter=150, hist=True)
    say("oracle iters", it_o, "gpu", rep)
    assert rep["status"] == "converged"
    assert abs(rep["iterations"] - it_o) <= 1
    assert rel(sol[3000:], ref[3000:]) < 1e-10, rel(sol[3000:], ref[3000:])
    assert rel(sol[:3000], ref[:3000]) < 1e-10, rel(sol[:3000], ref[:3000])
    n = min(len(hist_o), len(rep["hist"]))
    assert np.allclose(rep["hist"][:n], hist_o[:n], rtol=1e-6, atol=1e-14)
    # bitwise repeatable
    sol2, rep2 = s.solve(rhs, tol=1e-8, max_iter=150)
    assert np.array_equal(sol, sol2)
    s.close()


def test_config1_fp32_refined_solve():
    say("config1 fp32+refine")
    s, o, b, F6, oc = setup_pair(64, 100, 10, seed=1)
    rhs = np.concatenate([np.zeros(3000), -F6.ravel()])
    st, ref, it_o, _ = po.gmres(o, b, rhs, tol=1e-8, max_iter=150)
    sol, rep = s.solve(rhs, tol=1e-8, max_iter=150, precision=fa.FCM_PREC_FP32)
    say("fp32 refine", rep, "oracle err", rel(sol, ref))
    assert rep["status"] == "converged"
    assert rep["true_rel_residual"] < 1e-8
    true = np.linalg.norm(rhs - po.saddle_matvec(o, b, sol)) / np.linalg.norm(rhs)
    assert true < 2e-8
    assert rel(sol[3000:], ref[3000:]) < 1e-6
    s.close()


def test_fluctuating_solve_matches_oracle():
    """kT = 1 with the seeded stress stream fed to both sides."""
    say("fluctuating solve")
    s, o, b, F6, oc = setup_pair(32, 12, 6, seed=2, kT=1.0)
    o.fill_stress(77)
    o.random_field()
    Vr = o.random_interp()
    rhs = np.concatenate([-Vr, -F6.ravel()])
    st, ref, it_o, _ = po.gmres(o, b, rhs, tol=1e-8, max_iter=150)
    rep = s.mobility_solve(77, tol=1e-8, max_iter=150)
    sol = s.solution()
    say("oracle iters", it_o, "gpu", rep)
    assert rep["status"] == "converged"
    assert rel(sol, ref) < 1e-10, rel(sol, ref)


def test_zero_rhs_and_iteration_cap():
    say("zero rhs / cap")
    s, o, b, F6, oc = setup_pair(32, 6, 5, seed=3)
    sol, rep = s.solve(np.zeros(b.n), tol=1e-8, max_iter=20)
    assert rep["status"] == "zero_rhs" and np.all(sol == 0)
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal(b.n)
    sol, rep = s.solve(rhs, tol=1e-15, max_iter=3)
    assert rep["status"] == "max_iter" and rep["iterations"] == 3
    st, ref, it, _ = po.gmres(o, b, rhs, tol=1e-15, max_iter=3)
    assert st == 1 and rel(sol, ref) < 1e-10
    s.close()


def test_one_body_reference_pose_one_iteration():
    say("one body")
    gc, oc = pair_cfg(32, max_blobs=10, max_bodies=1, max_blobs_per_body=10)
    ref, _, _, _, _ = fa.synth_rods(0, 1, 10, 1.86, [32.0] * 3)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(np.array([[16.0, 16.0, 16.0]]), np.array([[1.0, 0, 0, 0]]))
        s.ref_matrices()
        F = np.array([0.3, -1.0, 0.5, 0.0, 0.7, -0.2])
        rhs = np.concatenate([np.zeros(30), -F])
        sol, rep = s.solve(rhs, tol=1e-8, max_iter=10)
        assert rep["iterations"] == 1 and rep["status"] in ("converged", "breakdown")
        res = s.saddle_matvec(sol) - rhs
        assert np.linalg.norm(res) < 1e-10 * np.linalg.norm(rhs)


def test_long_vector_solve_residual():
    """A system of more than 2^18 unknowns, where the basis passes move two
    rows per access: FP64 and FP32+refine solves reach the tolerance, checked
    with the oracle's own saddle product, and agree with each other."""
    say("long vectors")
    NB, nbpb, a = 7000, 12, 0.8
    gc, oc = pair_cfg(96, a=a, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, max_krylov=600)
    ref, x, q, F6, nov = fa.synth_rods(5, NB, nbpb, a, [96.0] * 3, max_attempts=200)
    n = 3 * NB * nbpb + 6 * NB
    assert n >= 1 << 18
    rhs = np.concatenate([np.zeros(3 * NB * nbpb), -F6.ravel()])
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        Mg, Minv, N = s.ref_matrices()
        sol64, rep64 = s.solve(rhs, tol=1e-8, max_iter=600)
        sol32, rep32 = s.solve(rhs, tol=1e-8, max_iter=600, precision=fa.FCM_PREC_FP32)
        say("fp64", rep64, "fp32+refine", rep32)
    assert rep64["status"] == "converged" and rep32["status"] == "converged"
    assert rep32["true_rel_residual"] < 1e-8
    o = po.Fcm(oc, NB * nbpb)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    b = po.Bodies(ref, x, q, Minv, N)
    for sol in (sol64, sol32):
        true = np.linalg.norm(rhs - po.saddle_matvec(o, b, sol)) / np.linalg.norm(rhs)
        assert true < 2e-8, true
    # two solutions within the same residual bound differ by what the conditioning
    # allows: 1.1e-5 over the whole vector (the blob forces of these small spheres),
    # 4e-7 in the body velocities, as measured
    nl = 3 * NB * nbpb
    say("fp32 vs fp64: all", rel(sol32, sol64), "body velocities", rel(sol32[nl:], sol64[nl:]))
    assert rel(sol32, sol64) < 1e-4
    assert rel(sol32[nl:], sol64[nl:]) < 1e-5


def test_fp32_history_longer_than_krylov_capacity():
    """An FP32+refine solve may take more iterations in all than the Krylov
    capacity (each inner solve restarts the basis); the residual history keeps
    the first max_krylov + 2 entries and the solve is unaffected."""
    say("fp32 history beyond capacity")
    NB, nbpb, cap = 40, 8, 12
    gc, oc = pair_cfg(48, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, max_krylov=cap)
    ref, x, q, F6, _ = fa.synth_rods(8, NB, nbpb, 1.86, [48.0] * 3, max_attempts=200)
    rhs = np.concatenate([np.zeros(3 * NB * nbpb), -F6.ravel()])
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        sol, rep = s.solve(rhs, tol=1e-8, max_iter=400, precision=fa.FCM_PREC_FP32, max_refine=60, hist=True)
        say(rep["iterations"], rep["refinements"], rep["status"], len(rep["hist"]))
        assert rep["status"] == "converged" and rep["iterations"] > cap + 2
        h = rep["hist"]
        assert len(h) == cap + 2
        assert np.all(np.isfinite(h)) and np.all(h > 0) and np.all(h <= 1.0 + 1e-12)
        sol2, rep2 = s.solve(rhs, tol=1e-8, max_iter=400, precision=fa.FCM_PREC_FP32, max_refine=60)
        assert np.array_equal(sol, sol2) and rep2["iterations"] == rep["iterations"]
