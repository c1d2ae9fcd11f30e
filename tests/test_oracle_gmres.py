"""Pins the oracle's saddle-point solve (gmres_solver.cpp:58-306) and the
reference matrices (fcm_fluid_solver.cpp:842-881).

Checks: dense LU on small systems, the one-body exact-preconditioner
known answer (one iteration), the Givens residual estimate against the
explicit residual, iteration cap and zero right-hand side status codes.
"""
import numpy as np
import pytest

import pyoracle as po


def rod(nbpb, a):
    return np.array([[2 * a * (m - (nbpb - 1) / 2), 0.0, 0.0] for m in range(nbpb)])


def random_quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]


def setup(c, blob_ref, x, q):
    nbpb = len(blob_ref)
    f = po.Fcm(c, max(nbpb, len(x) * nbpb))
    st, M, Minv, N = f.ref_matrices(blob_ref)
    assert st == 0
    b = po.Bodies(blob_ref, x, q, Minv, N)
    f.set_up_gaussians(po.fill_blob_pos(c, blob_ref, x, q))
    return f, b, M, Minv, N


def project_axial(F, q, axis_body=np.array([1.0, 0, 0])):
    F = F.copy().reshape(-1, 6)
    for n in range(len(F)):
        e = po.quat_matrix(q[n]) @ axis_body
        F[n, 3:] -= np.dot(F[n, 3:], e) * e
    return F.ravel()


def test_ref_matrices_symmetric_spd_and_singular_axis_for_collinear_rod():
    c = po.cfg(32, ng=7, a=1.86)
    f = po.Fcm(c, 10)
    st, M, Minv, N = f.ref_matrices(rod(10, 1.86))
    assert st == 0
    assert np.max(np.abs(M - M.T)) < 1e-13 * np.max(np.abs(M))
    assert np.min(np.linalg.eigvalsh(M)) > 0
    assert np.allclose(Minv @ M, np.eye(30), atol=1e-9)
    # axial spin (body x axis) is in ker K, so N has a zero row/column there
    assert np.max(np.abs(N[3, :])) < 1e-12 * np.max(np.abs(N))
    assert np.max(np.abs(N[:, 3])) < 1e-12 * np.max(np.abs(N))


def test_one_body_at_reference_pose_converges_in_one_iteration():
    c = po.cfg(32, ng=7, a=1.86)
    ref = rod(6, 1.86)
    x = np.array([[16.0, 16.0, 16.0]])
    q = np.array([[1.0, 0, 0, 0]])
    f, b, *_ = setup(c, ref, x, q)
    rng = np.random.default_rng(0)
    F = project_axial(rng.standard_normal(6), q)
    rhs = np.concatenate([np.zeros(3 * 6), -F])
    st, sol, iters, hist = po.gmres(f, b, rhs, tol=1e-8, max_iter=20)
    assert st in (0, 3) and iters == 1
    res = po.saddle_matvec(f, b, sol) - rhs
    assert np.linalg.norm(res) < 1e-10 * np.linalg.norm(rhs)


def _dense(f, b):
    n = b.n
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = po.saddle_matvec(f, b, e)
    return A


def test_gmres_matches_dense_lu_non_collinear_bodies():
    c = po.cfg(16, ng=7, a=1.0, dx=1.0)
    rng = np.random.default_rng(1)
    ref = rng.standard_normal((4, 3)) * 1.5
    x = rng.uniform(0, 16, (3, 3))
    q = random_quats(rng, 3)
    f, b, *_ = setup(c, ref, x, q)
    A = _dense(f, b)
    rhs = rng.standard_normal(b.n)
    x_lu = np.linalg.solve(A, rhs)
    st, sol, iters, hist = po.gmres(f, b, rhs, tol=1e-13, max_iter=b.n)
    assert st in (0, 3)
    assert np.linalg.norm(sol - x_lu) < 1e-9 * np.linalg.norm(x_lu)
    x2 = np.zeros(b.n)
    assert po.lib().or_solve_dense(b.n, np.ascontiguousarray(A.ravel()), rhs, x2) == 0
    assert np.allclose(x2, x_lu, rtol=1e-9, atol=1e-12)


def test_givens_estimate_equals_explicit_residual():
    c = po.cfg(16, ng=7, a=1.0)
    rng = np.random.default_rng(2)
    ref = rng.standard_normal((5, 3)) * 1.5
    x = rng.uniform(0, 16, (4, 3))
    q = random_quats(rng, 4)
    f, b, *_ = setup(c, ref, x, q)
    rhs = rng.standard_normal(b.n)
    for tol in (1e-3, 1e-6, 1e-9):
        st, sol, iters, hist = po.gmres(f, b, rhs, tol=tol, max_iter=200)
        assert st == 0
        assert hist[-1] < tol and (iters == 1 or hist[-2] >= tol)
        true = np.linalg.norm(rhs - po.saddle_matvec(f, b, sol)) / np.linalg.norm(rhs)
        assert abs(true - hist[-1]) < 1e-10 + 1e-6 * hist[-1]


def test_iteration_cap_and_zero_rhs():
    c = po.cfg(16, ng=7, a=1.0)
    rng = np.random.default_rng(3)
    ref = rng.standard_normal((4, 3))
    x = rng.uniform(0, 16, (3, 3))
    q = random_quats(rng, 3)
    f, b, *_ = setup(c, ref, x, q)
    rhs = rng.standard_normal(b.n)
    st, sol, iters, hist = po.gmres(f, b, rhs, tol=1e-14, max_iter=3)
    assert st == 1 and iters == 3
    st, sol, iters, hist = po.gmres(f, b, np.zeros(b.n), tol=1e-8, max_iter=10)
    assert st == 2 and np.all(sol == 0)


def test_Pinv_is_exact_inverse_for_one_body_at_reference_pose():
    c = po.cfg(24, ng=7, a=1.2)
    rng = np.random.default_rng(4)
    ref = rng.standard_normal((5, 3))
    x = np.array([[12.0, 12.0, 12.0]])
    q = np.array([[1.0, 0, 0, 0]])
    f, b, *_ = setup(c, ref, x, q)
    v = rng.standard_normal(b.n)
    w = po.saddle_matvec(f, b, po.apply_Pinv(b, v))
    assert np.linalg.norm(w - v) < 1e-10 * np.linalg.norm(v)


def test_Pinv_rotation_equivariance():
    """P^{-1} for a rotated body equals the reference-frame P^{-1} conjugated
    by the rotation (blockwise), independent of how it is computed."""
    c = po.cfg(24, ng=7, a=1.2)
    rng = np.random.default_rng(5)
    ref = rng.standard_normal((3, 3))
    f = po.Fcm(c, 3)
    st, M, Minv, N = f.ref_matrices(ref)
    q = random_quats(rng, 1)
    R = po.quat_matrix(q[0])
    b_rot = po.Bodies(ref, np.zeros(3), q, Minv, N)
    b_id = po.Bodies(ref, np.zeros(3), np.array([[1.0, 0, 0, 0]]), Minv, N)
    v = rng.standard_normal(b_rot.n)
    Rb = np.kron(np.eye(5), R)       # 3 blobs + 2 body 3-vectors
    out_rot = po.apply_Pinv(b_rot, v)
    out_id = po.apply_Pinv(b_id, Rb.T @ v)
    assert np.allclose(out_rot, Rb @ out_id, atol=1e-12 * np.max(np.abs(out_rot)))
