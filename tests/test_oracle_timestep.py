"""Oracle pins for the time-step pieces (config 5): soft-sphere barrier forces
(linked_list.cpp:192-281,362-382), the exponential map (quaternion.cpp:128-145),
the quaternion product and the end-of-step body update
(collection_of_rigid_bodies.cpp:759-790, particle.cpp:45-56).  Checked against
closed-form numbers that follow from the reference's constants (U0 = 20 kT,
b = a/2), Newton's third law, and scipy's independent rotation code."""
import numpy as np
from scipy.spatial.transform import Rotation

import pyoracle as po


def _cfg(n=32, a=1.0, kT=1.0):
    return po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=0.1)


def _pair(c, y0, y1, cutoff, nbpb=1):
    Y = np.array([y0, y1], dtype=float)
    return po.barrier_forces(c, Y, nbpb, cutoff).reshape(2, 3)


def test_barrier_force_known_answers():
    c = _cfg(a=1.0, kT=1.0)  # U0 = 20, b = 0.5: |F| = 40 inside 2a, 40 exp((2 - r)/0.5) outside
    F = _pair(c, [10.0, 10.0, 10.0], [12.5, 10.0, 10.0], cutoff=3.0)
    assert np.allclose(F[0], [-14.715177646857693, 0, 0], rtol=1e-14, atol=0)
    assert np.array_equal(F[1], -F[0])
    F = _pair(c, [10.0, 10.0, 10.0], [10.0, 11.0, 10.0], cutoff=3.0)  # overlapping: constant 40
    assert np.allclose(F[0], [0, -40.0, 0], rtol=1e-15)
    F = _pair(c, [10.0, 10.0, 10.0], [10.0, 10.0, 13.5], cutoff=3.0)  # beyond the cutoff
    assert not F.any()
    c2 = _cfg(a=1.5, kT=2.0)  # U0 = 40, b = 0.75: at r = 3.3, 40/0.75 exp(-0.4)
    F = _pair(c2, [5.0, 5.0, 5.0], [5.0, 5.0, 8.3], cutoff=4.0)
    assert np.allclose(F[0], [0, 0, -(40 / 0.75) * np.exp(-0.3 / 0.75)], rtol=1e-13)


def test_barrier_force_uses_nearest_image_and_skips_own_body():
    c = _cfg(n=32, a=1.0)
    F = _pair(c, [0.5, 7.0, 7.0], [31.0, 7.0, 7.0], cutoff=3.0)  # separation 1.5 through the x face
    assert np.allclose(F[0], [40.0, 0, 0], rtol=1e-15)
    F = _pair(c, [10.0, 10.0, 10.0], [11.0, 10.0, 10.0], cutoff=3.0, nbpb=2)  # one body of two blobs
    assert not F.any()


def test_barrier_forces_sum_to_zero_and_are_translation_invariant():
    c = _cfg(n=24, a=1.2)
    rng = np.random.default_rng(0)
    Y = rng.uniform(0, 24, (60, 3))
    F = po.barrier_forces(c, Y, 3, 2.64).reshape(-1, 3)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).max()
    assert np.abs(F).max() > 0
    G = po.barrier_forces(c, np.mod(Y + np.array([3.0, 5.0, 7.0]), 24.0), 3, 2.64).reshape(-1, 3)
    assert np.allclose(F, G, rtol=1e-10, atol=1e-12)


def test_lie_exp_matches_scipy_rotvec():
    rng = np.random.default_rng(1)
    for _ in range(5):
        u = rng.standard_normal(3)
        q = po.lie_exp(u)
        assert abs(np.linalg.norm(q) - 1) < 1e-15
        # scipy quaternions are (x, y, z, w)
        ref = Rotation.from_rotvec(u).as_quat()
        assert np.allclose(np.r_[ref[3], ref[:3]] * np.sign(ref[3] * q[0]), q, atol=1e-15)
    assert np.array_equal(po.lie_exp([1e-11, 0, 0]), [1.0, 0, 0, 0])


def test_quaternion_product_is_hamilton():
    i, j, k = np.array([0, 1.0, 0, 0]), np.array([0, 0, 1.0, 0]), np.array([0, 0, 0, 1.0])
    assert np.array_equal(po.quat_mul(i, j), k)
    assert np.array_equal(po.quat_mul(j, i), -k)
    rng = np.random.default_rng(2)
    p, q = rng.standard_normal(4), rng.standard_normal(4)
    assert np.isclose(np.linalg.norm(po.quat_mul(p, q)), np.linalg.norm(p) * np.linalg.norm(q), rtol=1e-14)


def test_end_of_step_update_matches_scipy_composition():
    c = _cfg(n=16)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 16, (4, 3))
    q = rng.standard_normal((4, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    V6 = rng.standard_normal((4, 6)) * 20
    s = 0.37
    x1, q1 = po.end_of_step_update(c, V6, s, x, q)
    assert np.allclose(x1.reshape(-1, 3), np.mod(x + s * V6[:, :3], 16.0), atol=1e-13)
    for n in range(4):
        R0 = Rotation.from_quat(np.r_[q[n, 1:], q[n, 0]]).as_matrix()
        R1 = Rotation.from_quat(np.r_[q1[4 * n + 1:4 * n + 4], q1[4 * n]]).as_matrix()
        assert np.allclose(R1, Rotation.from_rotvec(s * V6[n, 3:]).as_matrix() @ R0, atol=1e-13)


def _adjacent(ci, cj, n):
    """Cells of an n^3 periodic grid that are the same or neighbours."""
    d = np.abs(ci - cj)
    d = np.minimum(d, n - d)
    return np.all(d <= 1, axis=-1)


def test_reference_cell_traversal_covers_every_pair_once_on_three_cells():
    """On a 3^3 cell grid every cell neighbours every other one, so the
    reference's traversal (each cell against itself and its 13 half-shell
    cells, linked_list.cpp:192-281) must give every pair exactly once: the
    all-pairs brute force with no cutoff."""
    c = _cfg(n=24, a=1.0)
    rng = np.random.default_rng(3)
    Y = rng.uniform(0, 24, (90, 3))
    F = po.barrier_forces_cells(c, Y, 3, 3).reshape(-1, 3)
    G = po.barrier_forces(c, Y, 3, 1e9).reshape(-1, 3)
    assert np.abs(G).max() > 0
    assert np.allclose(F, G, rtol=1e-12, atol=1e-14 * np.abs(G).max())


def test_reference_cell_traversal_is_the_sum_over_neighbouring_cell_pairs():
    """For n >= 4 the interaction set is the pairs of blobs of different bodies
    whose cells floor(n Y/L) are the same or neighbours (periodically), at any
    distance: the result is the superposition of the two-blob forces over
    exactly those pairs.  Channel wall springs come on top."""
    for n_cells, channel, seed in ((4, False, 5), (6, False, 6), (5, True, 7)):
        c = po.cfg(20, ng=7, dx=1.0, a=1.0, mu=1.0, kT=1.3, dt=0.1, slip_channel=channel)
        rng = np.random.default_rng(seed)
        nbpb = 2
        Y = rng.uniform(0, 20, (48, 3))
        if channel:
            Y[:, 1] = rng.uniform(0.5, 9.5, 48)
        F = po.barrier_forces_cells(c, Y, nbpb, n_cells).reshape(-1, 3)
        cells = np.floor(n_cells * (Y / 20.0)).astype(int)
        # wall springs alone: each blob as its own body, far from every other
        W = np.stack([po.barrier_forces(c, Y[i:i + 1], 1, 1e-9).reshape(3) for i in range(len(Y))])
        S = W.copy()
        c0 = po.cfg(20, ng=7, dx=1.0, a=1.0, mu=1.0, kT=1.3, dt=0.1)
        far = 0
        for i in range(len(Y)):
            for j in range(len(Y)):
                if i // nbpb == j // nbpb or not _adjacent(cells[i], cells[j], n_cells):
                    continue
                S[i] += po.barrier_forces(c0, Y[[i, j]], 1, 1e9).reshape(2, 3)[0]
                d = Y[i] - Y[j]
                d -= 20.0 * np.round(d / 20.0)
                far += np.linalg.norm(d) > 20.0 / n_cells
        assert far > 0, "some interacting pairs must lie farther apart than one cell width"
        assert np.allclose(F, S, rtol=1e-12, atol=1e-13 * np.abs(S).max()), (n_cells, channel)
        if not channel:
            assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).max()


def _interp_and_gradient(o, oc, Y, names=("ux_rand", "uy_rand", "uz_rand")):
    """Blob velocities V_m = dx^3 sum g_x g_y g_z u of the oracle's field and
    their analytic derivatives dV_m/dY_m (3 x 3 per blob): the weight
    g(Y) = exp(-(xg dx - Y)^2 / (2 a^2/pi)) / (sqrt 2 a) of node xg = int(Y/dx)
    - ng/2 + i + 1 (fcm_fluid_solver.cpp:131-165) has dg/dY = g 2 (xg dx - Y) /
    (2 a^2/pi); the stencil does not move under a displacement that keeps
    int(Y/dx)."""
    ng, dx, a = oc.ng, oc.dx, oc.a
    n = (oc.nx, oc.ny, oc.nz)
    pad = po.pad_of(oc)
    u = [o.field(c).reshape(oc.nx, oc.ny, pad)[:, :, :oc.nz] for c in names]
    Y = np.asarray(Y).reshape(-1, 3)
    V = np.zeros((len(Y), 3))
    G = np.zeros((len(Y), 3, 3))  # [blob, component, derivative axis]
    s2 = 2 * a * a / np.pi
    for m, y in enumerate(Y):
        w, dw, idx = [], [], []
        for d in range(3):
            xg = int(y[d] / dx) - ng // 2 + np.arange(1, ng + 1)
            g = np.exp(-(xg * dx - y[d]) ** 2 / s2) / (np.sqrt(2.0) * a)
            w.append(g)
            dw.append(g * 2 * (xg * dx - y[d]) / s2)
            idx.append(np.mod(xg, n[d]))
        for c in range(3):
            box = u[c][np.ix_(idx[0], idx[1], idx[2])]
            V[m, c] = dx ** 3 * np.einsum("i,j,k,ijk", w[0], w[1], w[2], box)
            G[m, c, 0] = dx ** 3 * np.einsum("i,j,k,ijk", dw[0], w[1], w[2], box)
            G[m, c, 1] = dx ** 3 * np.einsum("i,j,k,ijk", w[0], dw[1], w[2], box)
            G[m, c, 2] = dx ** 3 * np.einsum("i,j,k,ijk", w[0], w[1], dw[2], box)
    return V, G


def _skew(e):
    return np.array([[0.0, -e[2], e[1]], [e[2], 0.0, -e[0]], [-e[1], e[0], 0.0]])


def test_gdc_divU_checker_matches_the_analytic_derivative():
    """tests/_gdc.py's divU (calculate_divU, collection_of_rigid_bodies.cpp:
    869-953: forward differences of the rigid fit of the interpolated random
    field under six body displacements) against the derivative taken
    analytically: d/de of U_n = diag(R,R) (K^T K)^-1 diag(R^T,R^T) [sum v;
    sum (R r_m) x v] with v = V(Y(e)) from the analytic gradient of the
    Gaussian interpolation and, for rotations, dR/de = [e_i]x R.  The
    finite-difference error must be first order in the step: 100 times
    smaller at step 1e-5 than at 1e-3, and below 1e-4 relative.  A dropped
    term (the fit's dependence on orientation, a blob's own displacement)
    or a wrong sign leaves an error that does not shrink with the step."""
    import _gdc
    n, NB, nbpb, a = 24, 3, 6, 1.86
    oc = po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=1.0, dt=0.2)
    ref = _gdc.reference_rods(nbpb, a)
    rng = np.random.default_rng(12)
    x = rng.uniform(0, n, (NB, 3))
    q = rng.standard_normal((NB, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    fd = {}
    for delta in (1e-3, 1e-5):
        fd[delta] = _gdc.gdc_step(o, oc, ref, x, q, Minv, N, seed=31, cutoff=0.0, delta=delta, tol=1e-10)[2]
    # the field the checker differentiated (same seed), and its analytic derivative
    Y = po.fill_blob_pos(oc, ref, x, q).reshape(-1, 3)
    o.set_up_gaussians(Y.ravel())
    o.fill_stress(31)
    o.random_field()
    V, G = _interp_and_gradient(o, oc, Y)
    assert np.allclose(V.ravel(), o.random_interp(), rtol=1e-12, atol=1e-14 * np.abs(V).max())
    KTKi = _gdc.ktk_inv(ref)
    an = 0.0
    for b in range(NB):
        R = po.quat_matrix(q[b]).reshape(3, 3)
        d = (R @ ref.T).T  # R r_m
        vb, Gb = V[b * nbpb:(b + 1) * nbpb], G[b * nbpb:(b + 1) * nbpb]

        def fit(Rm, v):
            dd = (Rm @ ref.T).T
            w = np.concatenate([v.sum(0), np.cross(dd, v).sum(0)])
            B = np.kron(np.eye(2), Rm)
            return B @ KTKi @ B.T @ w

        for i in range(6):
            e = np.eye(3)[i % 3]
            if i < 3:
                dv = Gb @ e  # every blob moves by e
                dU = fit(R, dv)
            else:
                dY = np.cross(e, d)
                dv = np.einsum("mcd,md->mc", Gb, dY)
                dR = _skew(e) @ R
                w = np.concatenate([vb.sum(0), np.cross(d, vb).sum(0)])
                dw = np.concatenate([dv.sum(0), (np.cross(dY, vb) + np.cross(d, dv)).sum(0)])
                B, dB = np.kron(np.eye(2), R), np.kron(np.eye(2), dR)
                dU = dB @ KTKi @ B.T @ w + B @ KTKi @ dB.T @ w + B @ KTKi @ B.T @ dw
            an += dU[i]
    e3, e5 = abs(fd[1e-3] - an), abs(fd[1e-5] - an)
    print("divU analytic", an, "fd 1e-3", fd[1e-3], "fd 1e-5", fd[1e-5])
    assert abs(an) > 1e-3
    # forward differences: the error falls with the step, 100x for 100x
    assert e5 < 1e-4 * abs(an) and 0.005 < e5 / e3 < 0.02, (e3, e5)
    o.close()
