"""Oracle composition of one generalised drifter-corrector step (main.cpp:
222-312 with GEN_DRIFTER_CORRECTOR; collection_of_rigid_bodies.cpp:566-621,
869-953) from the oracle's operators, for the GPU tests."""
import numpy as np

import pyoracle as po


def reference_rods(nbpb, a):
    """The reference's rods (collection_of_rigid_bodies.cpp:55-67): pairs of
    spheres at +-a across the axis, each pair turned by 90 degrees."""
    pos = []
    for pair in range(nbpb // 2):
        th = pair * 0.5 * np.pi
        x = np.sqrt(2.0) * a * (pair - 0.25 * nbpb + 0.5)
        pos.append([x, a * np.cos(th), a * np.sin(th)])
        pos.append([x, -a * np.cos(th), -a * np.sin(th)])
    return np.array(pos)


def ktk_inv(ref):
    """(K^T K)^-1 with K = [I, rcross(r_m)], rcross(r) v = v x r (:117-128)."""
    rows = []
    for r in ref:
        rc = np.array([[0.0, r[2], -r[1]], [-r[2], 0.0, r[0]], [r[1], -r[0], 0.0]])
        rows.append(np.hstack([np.eye(3), rc]))
    K = np.vstack(rows)
    return np.linalg.inv(K.T @ K)


def rigid_fit(ref, q, V, KTKi):
    """U_n = diag(R,R) KTKi diag(R^T,R^T) [sum v; sum (R r_m) x v] (:583-599)."""
    b = po.mul_KT(ref, q, V).reshape(-1, 6)
    q = np.asarray(q).reshape(-1, 4)
    U = np.zeros_like(b)
    for n in range(len(b)):
        R = po.quat_matrix(q[n])
        u = KTKi @ np.concatenate([R.T @ b[n, :3], R.T @ b[n, 3:]])
        U[n] = np.concatenate([R @ u[:3], R @ u[3:]])
    return U


def gdc_step(o, oc, ref, x, q, Minv, N, seed, cutoff, delta, tol, F6=None, shear=False):
    """Returns (x1, q1, divU, sol, midstep blob positions) for one step from
    bodies (x, q); o is an oracle Fcm of the same grid (with its shear field
    set when shear is true: the mid-step RHS then subtracts its blob
    velocities, main.cpp:270-274 and gmres_solver.cpp:62-64)."""
    x = np.asarray(x, dtype=float).ravel()
    q = np.asarray(q, dtype=float).ravel()
    NB = x.size // 3
    KTKi = ktk_inv(ref)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    o.fill_stress(seed)
    o.random_field()
    V = o.random_interp()
    U0 = rigid_fit(ref, q, V, KTKi)
    xm, qm = po.end_of_step_update(oc, U0.ravel(), 0.5 * oc.dt, x, q)
    divU = 0.0
    for i in range(6):
        e = delta * oc.a if i < 3 else delta
        E = np.zeros((NB, 6))
        E[:, i] = e
        xd, qd = po.end_of_step_update(oc, E.ravel(), 1.0, x, q)
        o.set_up_gaussians(po.fill_blob_pos(oc, ref, xd, qd))
        Ud = rigid_fit(ref, qd, o.random_interp(), KTKi)
        divU += np.sum((Ud[:, i] - U0[:, i]) / e)
    nu = 1.0 + 0.5 * oc.dt * divU
    Ym = po.fill_blob_pos(oc, ref, xm, qm)
    o.set_up_gaussians(Ym)
    Vm = o.random_interp()
    if shear:
        Vm = Vm + o.shear_interp()
    Fb = po.mul_KT(ref, qm, po.barrier_forces(oc, Ym, len(ref), cutoff)) if cutoff > 0 else np.zeros(6 * NB)
    if F6 is not None:
        Fb = Fb + np.asarray(F6).ravel()
    rhs = np.concatenate([-Vm, -Fb])
    st, sol, it, _ = po.gmres(o, po.Bodies(ref, xm, qm, Minv, N), rhs, tol=tol, max_iter=300)
    assert st == 0
    x1, q1 = po.end_of_step_update(oc, sol[3 * NB * len(ref):], nu * oc.dt, x, q)
    return x1, q1, divU, sol, Ym.reshape(-1, 3)
