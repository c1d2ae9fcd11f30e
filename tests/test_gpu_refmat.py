"""The single-body reference matrices computed on the GPU against the oracle's
(produce_reference_matrices, fcm_fluid_solver.cpp:842-881): the mobility M of
3 n_b unit-force products, ref_Minv = inv_sympd(M) and ref_N = inv(K^T Minv K)
(a pseudo-inverse where K^T Minv K is singular: collinear blobs and single
spheres have no resistance to spin about their axis).  Each matrix is compared
directly, none is injected, on the periodic box and the slip channel, for
collinear rods, one sphere, and the reference's own 22-sphere rods of turned
pairs (collection_of_rigid_bodies.cpp:23,56-70).  Then P^-1 and a solve with
the GPU's own matrices match the oracle's P^-1 and solve with its matrices."""
import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, random_quats, rel, rod
from _gdc import reference_rods

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = [(32, "collinear10", False), (48, "reference22", False), (48, "sphere", False), (48, "reference22", True),
         (48, "collinear6", True), (32, "reference12", False)]


def _geometry(kind):
    if kind.startswith("collinear"):
        return rod(int(kind[9:]), 1.86)
    if kind.startswith("reference"):
        return reference_rods(int(kind[9:]), 1.86)
    return rod(1, 1.86)


@pytest.mark.parametrize("n,kind,channel", CASES)
def test_reference_matrices_match_oracle(n, kind, channel):
    say("reference matrices", n, kind, "channel" if channel else "periodic")
    ref = _geometry(kind)
    nbpb = len(ref)
    gc, oc = pair_cfg(n, max_blobs=nbpb, max_bodies=1, max_blobs_per_body=nbpb, slip_channel=channel)
    o = po.Fcm(oc, nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.close()
    assert st == 0
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        Mg, Minvg, Ng = s.ref_matrices()
    eM, eMi, eN = rel(Mg, M), rel(Minvg, Minv), rel(Ng, N)
    say("  M", eM, "Minv", eMi, "N", eN)
    assert eM < 1e-12 and eMi < 1e-12 and eN < 1e-12, (eM, eMi, eN)
    # the pseudo-inverse drops the same directions on both sides: N is
    # singular exactly where the oracle's is
    w, wg = np.linalg.eigvalsh(0.5 * (N + N.T)), np.linalg.eigvalsh(0.5 * (Ng + Ng.T))
    assert np.sum(np.abs(w) < 1e-12 * np.abs(w).max()) == np.sum(np.abs(wg) < 1e-12 * np.abs(wg).max())


@pytest.mark.parametrize("kind,channel", [("reference22", False), ("collinear10", False), ("reference22", True)])
def test_Pinv_and_solve_with_gpu_matrices_match_oracle(kind, channel):
    """With the matrices each side computed itself, P^-1 and the saddle-point
    solve agree (the solve to 1e-10, the reference's FP64 bar)."""
    say("P^-1 and solve, own matrices", kind, "channel" if channel else "periodic")
    n = 64
    ref = _geometry(kind)
    nbpb, NB = len(ref), 12
    gc, oc = pair_cfg(n, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, slip_channel=channel,
                      max_krylov=200)
    rng = np.random.default_rng(41)
    x = rng.uniform(0, n, (NB, 3))
    if channel:
        half = 0.5 * np.abs(ref).max() + 1.0
        x[:, 1] = rng.uniform(half, n / 2 - half, NB)
        # keep every sphere between the walls: bodies rotated only about y
        th = rng.uniform(0, 2 * np.pi, NB)
        q = np.stack([np.cos(th / 2), np.zeros(NB), np.sin(th / 2), np.zeros(NB)], axis=1)
    else:
        q = random_quats(rng, NB)
    F6 = rng.standard_normal((NB, 6))
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    b = po.Bodies(ref, x, q, Minv, N)
    # collinear rods: no axial torque, as the synthetic workloads do
    if kind.startswith("collinear"):
        for k in range(NB):
            axis = po.quat_matrix(q[k])[:, 0]
            F6[k, 3:] -= axis * (axis @ F6[k, 3:])
    rhs = np.concatenate([np.zeros(3 * NB * nbpb), -F6.ravel()])
    st, sol_ref, it, _ = po.gmres(o, b, rhs, tol=1e-10, max_iter=200)
    assert st == 0
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        v = rng.standard_normal(b.n)
        e = rel(s.apply_Pinv(v), po.apply_Pinv(b, v))
        say("  P^-1", e)
        assert e < 1e-11, e
        sol, rep = s.solve(rhs, tol=1e-10, max_iter=200)
        say("  solve", rep, "oracle iterations", it, "error", rel(sol, sol_ref))
        assert rep["status"] == "converged"
        assert rel(sol, sol_ref) < 1e-9
    o.close()
