"""Config-5 pieces on the GPU against the oracle: barrier forces on the device
cell list (linked_list.cpp:192-281), the end-of-step body update
(collection_of_rigid_bodies.cpp:759-790) and whole Euler-Maruyama steps
(main.cpp:222-312): body positions and quaternions after each step agree with
the oracle's step to 1e-9."""
import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, random_quats, rel, rod

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n,NB,nbpb,a,cutoff", [(32, 60, 6, 1.86, 2.2 * 1.86), (16, 20, 3, 1.0, 6.0),
                                                (48, 200, 8, 1.86, 2.2 * 1.86), (24, 40, 4, 1.5, 5.0)])
def test_barrier_forces_match_oracle(n, NB, nbpb, a, cutoff):
    say("barrier", n, NB, nbpb, a, cutoff)
    gc, oc = pair_cfg(n, a=a, kT=1.0, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb)
    rng = np.random.default_rng(n + NB)
    ref = rod(nbpb, a)
    x = rng.uniform(0, n, (NB, 3))
    q = random_quats(rng, NB)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        F = s.barrier_forces(cutoff)
        Fo = po.barrier_forces(oc, po.fill_blob_pos(oc, ref, x, q), nbpb, cutoff).reshape(-1, 3)
        assert np.abs(Fo).max() > 0
        assert rel(F, Fo) < 1e-12, rel(F, Fo)
        assert np.array_equal(F, s.barrier_forces(cutoff))


def test_end_of_step_update_matches_oracle():
    gc, oc = pair_cfg(32, dt=0.2, max_blobs=50, max_bodies=10, max_blobs_per_body=5)
    rng = np.random.default_rng(9)
    ref = rod(5, 1.86)
    x = rng.uniform(0, 32, (10, 3))
    q = random_quats(rng, 10)
    V6 = rng.standard_normal((10, 6)) * 30
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.end_of_step_update(V6, nu=0.8)
        xg, qg = s.bodies()
        xo, qo = po.end_of_step_update(oc, V6, 0.8 * 0.2, x, q)
        assert np.allclose(xg.ravel(), xo, rtol=0, atol=1e-13)
        assert np.allclose(qg.ravel(), qo, rtol=0, atol=1e-14)
        assert rel(s.blob_positions(), po.fill_blob_pos(oc, ref, xo, qo)) < 1e-14


def test_euler_maruyama_steps_match_oracle():
    n, NB, nbpb, dt, cutoff, tol = 32, 12, 6, 0.20216, 2.2 * 1.86, 1e-10
    gc, oc = pair_cfg(n, kT=1.0, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, max_krylov=150)
    ref, x, q, _, _ = fa.synth_rods(21, NB, nbpb, 1.86, [float(n)] * 3, max_attempts=20)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    xo, qo = x.ravel().copy(), q.ravel().copy()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(np.zeros((NB, 6)))
        for step in range(3):
            seed = 500 + step
            Y = po.fill_blob_pos(oc, ref, xo, qo)
            o.set_up_gaussians(Y)
            o.fill_stress(seed)
            o.random_field()
            Vr = o.random_interp()
            f = po.barrier_forces(oc, Y, nbpb, cutoff)
            F6 = po.mul_KT(ref, qo, f)
            rhs = np.concatenate([-Vr, -F6])
            b = po.Bodies(ref, xo, qo, Minv, N)
            st, sol, it, _ = po.gmres(o, b, rhs, tol=tol, max_iter=150)
            assert st == 0
            xo, qo = po.end_of_step_update(oc, sol[3 * NB * nbpb:], dt, xo, qo)
            rep = s.em_step(seed, cutoff, tol=tol, max_iter=150)
            xg, qg = s.bodies()
            dx = np.abs(xg.ravel() - xo)
            dx = np.minimum(dx, n - dx)
            say("step", step, rep, "oracle iterations", it, "max |dx|", dx.max(), "max |dq|",
                np.abs(qg.ravel() - qo).max())
            assert rep["status"] == "converged"
            assert dx.max() < 1e-9
            assert np.abs(qg.ravel() - qo).max() < 1e-9
    o.close()


def _gdc_pair(n, NB, nbpb, channel, rng):
    import _gdc
    a = 1.86
    ref = _gdc.reference_rods(nbpb, a)
    gc, oc = pair_cfg(n, kT=1.0, dt=0.20216, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      max_krylov=300, slip_channel=channel)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    return _gdc, ref, gc, oc, o, Minv, N


def test_gdc_steps_match_oracle():
    """Generalised drifter-corrector steps (main.cpp:235-274, 308-312):
    divU from the six finite differences and the positions after each step."""
    n, NB, nbpb, cutoff, tol = 32, 8, 6, 2.2 * 1.86, 1e-10
    rng = np.random.default_rng(3)
    _gdc, ref, gc, oc, o, Minv, N = _gdc_pair(n, NB, nbpb, False, rng)
    x = rng.uniform(0, n, (NB, 3))
    q = random_quats(rng, NB)
    xo, qo = x.ravel().copy(), q.ravel().copy()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(np.zeros((NB, 6)))
        for step in range(2):
            seed = 11 + step
            xo, qo, divU, sol, _ = _gdc.gdc_step(o, oc, ref, xo, qo, Minv, N, seed, cutoff, 1e-5, tol)
            rep = s.gdc_step(seed, cutoff, fd_delta=1e-5, tol=tol, max_iter=300)
            xg, qg = s.bodies()
            d = np.abs(xg.ravel() - xo)
            d = np.minimum(d, n - d)
            say("gdc step", step, rep, "oracle divU", divU, "max |dx|", d.max())
            assert rep["status"] == "converged" and not rep["redo"]
            assert abs(rep["divU"] - divU) < 1e-6 * abs(divU)
            assert d.max() < 1e-9
            assert np.abs(qg.ravel() - qo).max() < 1e-9
    o.close()


def test_gdc_channel_midstep_wall_overlap_matches_oracle():
    """In a slip channel the step is taken again when the mid-step puts a blob
    on or beyond a wall (main.cpp:237-246): the GPU's verdict matches the
    oracle's mid-step for each seed, a refused step leaves the bodies as they
    were, an accepted one matches the oracle's step."""
    n, NB, nbpb, cutoff, tol = 32, 8, 6, 2.2 * 1.86, 1e-10
    rng = np.random.default_rng(3)
    _gdc, ref, gc, oc, o, Minv, N = _gdc_pair(n, NB, nbpb, True, rng)
    x = rng.uniform(0, n, (NB, 3))
    x[:, 1] = rng.uniform(5, 11, NB)
    x[0, 1] = 1.86 + 1e-3  # two spheres of body 0 sit 1e-3 above the lower wall
    q = np.tile([1.0, 0.0, 0.0, 0.0], (NB, 1))
    seen = set()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        for seed in (20, 21, 22, 23):
            s.set_bodies(x, q)
            s.set_ref_matrices(Minv, N)
            s.set_body_forces(np.zeros((NB, 6)))
            x1, q1, divU, sol, Ym = _gdc.gdc_step(o, oc, ref, x, q, Minv, N, seed, cutoff, 1e-5, tol)
            # the oracle's mid-step configuration decides
            o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
            o.fill_stress(seed)
            o.random_field()
            U0 = _gdc.rigid_fit(ref, q, o.random_interp(), _gdc.ktk_inv(ref))
            xm, qm = po.end_of_step_update(oc, U0.ravel(), 0.5 * oc.dt, x.ravel(), q.ravel())
            ym = po.fill_blob_pos(oc, ref, xm, qm).reshape(-1, 3)[:, 1]
            expect = bool(((ym == 0.0) | (ym >= n / 2)).any())
            rep = s.gdc_step(seed, cutoff, fd_delta=1e-5, tol=tol, max_iter=300)
            say("seed", seed, "redo", rep["redo"], "oracle", expect)
            assert rep["redo"] == expect
            seen.add(expect)
            xg, qg = s.bodies()
            if expect:
                assert np.array_equal(xg, x) and np.array_equal(qg, q)
            else:
                d = np.abs(xg.ravel() - x1)
                d = np.minimum(d, n - d)
                assert d.max() < 1e-9 and np.abs(qg.ravel() - q1).max() < 1e-9
    assert seen == {True, False}
    o.close()


def test_blob_mode_steps_match_oracle():
    """SIMULATING_BLOBS (RIGID_BODY_TYPE 0, main.cpp:290-304): free spheres, no
    saddle-point solve.  Each step composes the entry points: the fluctuating
    field's blob velocities, the barrier forces, the deterministic mobility
    product of those forces, and every sphere moved by DT (V_rand + M F)
    (a one-sphere body per blob, rotation zero)."""
    n, NB, dt, cutoff = 32, 60, 0.20216, 2.2 * 1.86
    gc, oc = pair_cfg(n, kT=1.0, dt=dt, max_blobs=NB, max_bodies=NB, max_blobs_per_body=1)
    rng = np.random.default_rng(17)
    x = rng.uniform(0, n, (NB, 3))
    q = np.tile([1.0, 0.0, 0.0, 0.0], (NB, 1))
    ref = np.zeros((1, 3))
    o = po.Fcm(oc, NB)
    xo, qo = x.ravel().copy(), q.ravel().copy()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        for step in range(3):
            seed = 40 + step
            Y = po.fill_blob_pos(oc, ref, xo, qo)
            o.set_up_gaussians(Y)
            o.fill_stress(seed)
            o.random_field()
            Vr = o.random_interp()
            F = po.barrier_forces(oc, Y, 1, cutoff)
            V = Vr + o.mobility(F)
            V6 = np.zeros((NB, 6))
            V6[:, :3] = V.reshape(-1, 3)
            xo, qo = po.end_of_step_update(oc, V6.ravel(), dt, xo, qo)

            s.random_field(None, seed)
            Vg = s.random_interp()
            Fg = s.barrier_forces(cutoff)
            Vg = Vg + s.mobility(Fg.ravel())
            V6g = np.zeros((NB, 6))
            V6g[:, :3] = Vg.reshape(-1, 3)
            s.end_of_step_update(V6g, nu=1.0)
            xg, qg = s.bodies()
            d = np.abs(xg.ravel() - xo)
            d = np.minimum(d, n - d)
            say("blob step", step, "max |dx|", d.max())
            assert d.max() < 1e-11
            assert np.array_equal(qg.ravel(), qo)  # no rotation
    o.close()


def test_gdc_channel_shear_steps_match_oracle():
    """A drifter-corrector step in a slip channel with shear forcing: the
    mid-step RHS carries the shear blob velocities at the mid-step positions."""
    n, NB, nbpb, cutoff, tol = 32, 6, 6, 2.2 * 1.86, 1e-10
    rng = np.random.default_rng(5)
    _gdc, ref, gc, oc, o, Minv, N = _gdc_pair(n, NB, nbpb, True, rng)
    o.shear_field(1.5, 1.0)
    x = rng.uniform(0, n, (NB, 3))
    x[:, 1] = rng.uniform(6.0, 10.0, NB)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (NB, 1))
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(np.zeros((NB, 6)))
        s.shear_field(1.5, 1.0)
        x1, q1, divU, sol, _ = _gdc.gdc_step(o, oc, ref, x, q, Minv, N, 31, cutoff, 1e-5, tol, shear=True)
        rep = s.gdc_step(31, cutoff, fd_delta=1e-5, tol=tol, max_iter=300)
        assert not rep["redo"] and rep["status"] == "converged"
        assert abs(rep["divU"] - divU) < 1e-6 * abs(divU)
        assert rel(s.solution(), sol) < 1e-10, rel(s.solution(), sol)
        xg, qg = s.bodies()
        d = np.abs(xg.ravel() - x1)
        d = np.minimum(d, n - d)
        assert d.max() < 1e-9 and np.abs(qg.ravel() - q1).max() < 1e-9
    o.close()


@pytest.mark.parametrize("channel", [False, True])
def test_fp32_refined_steps_track_fp64_steps(channel):
    """Time steps with the FP32+refine solve (the benchmark's form) stay with the
    FP64 steps over three Euler-Maruyama steps, periodic and in a slip channel.
    Both solves stop at a relative residual of 1e-8, so their solutions differ
    by up to the conditioning times that (about 1e-6 relative here): the body
    displacements must agree to 1e-5 of their size."""
    n, NB, nbpb, dt, cutoff = 32, 10, 6, 0.20216, 2.2 * 1.86
    import _gdc
    ref = _gdc.reference_rods(nbpb, 1.86)
    rng = np.random.default_rng(23)
    x = rng.uniform(0, n, (NB, 3))
    if channel:
        x[:, 1] = rng.uniform(6.0, 10.0, NB)
    q = random_quats(rng, NB)
    out = {}
    for prec in (fa.FCM_PREC_FP64, fa.FCM_PREC_FP32):
        gc, _ = pair_cfg(n, kT=1.0, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                         max_krylov=300, slip_channel=channel)
        with fa.FcmSolver(gc) as s:
            s.set_geometry(ref)
            s.ref_matrices()
            s.set_bodies(x, q)
            s.set_body_forces(np.zeros((NB, 6)))
            for k in range(3):
                rep = s.em_step(60 + k, cutoff, tol=1e-8, max_iter=300, precision=prec)
                assert rep["status"] == "converged", rep
            out[prec] = s.bodies()
    d64 = out[fa.FCM_PREC_FP64][0] - x
    d32 = out[fa.FCM_PREC_FP32][0] - x
    d64 -= n * np.round(d64 / n)
    d32 -= n * np.round(d32 / n)
    say("fp32 vs fp64 steps", channel, np.abs(d32 - d64).max(), np.abs(d64).max())
    assert np.abs(d32 - d64).max() < 1e-5 * np.abs(d64).max()
    assert np.abs(out[fa.FCM_PREC_FP32][1] - out[fa.FCM_PREC_FP64][1]).max() < 1e-5


@pytest.mark.parametrize("n,NB,nbpb,cells,channel", [(32, 60, 6, 6, False), (24, 40, 4, 3, False),
                                                     (48, 200, 8, 6, False), (32, 30, 4, 6, True),
                                                     (48, 80, 5, 7, False)])
def test_reference_cell_interaction_set_matches_oracle(n, NB, nbpb, cells, channel):
    """ll_set_cells: the reference's interaction set on a cells^3 grid, every
    pair in the same or neighbouring cells at any distance (linked_list.cpp:
    23-102,192-281), against the oracle's restatement of that traversal."""
    say("reference cells", n, NB, nbpb, cells, channel)
    a = 1.86
    gc, oc = pair_cfg(n, a=a, kT=1.0, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      slip_channel=channel)
    rng = np.random.default_rng(n + NB + cells)
    ref = rod(nbpb, a)
    x = rng.uniform(0, n, (NB, 3))
    if channel:
        x[:, 1] = rng.uniform(a + 0.5, n / 2 - a - 0.5, NB)
        th = rng.uniform(0, 2 * np.pi, NB)  # rods turned about y keep their spheres at the centre's y
        q = np.stack([np.cos(th / 2), np.zeros(NB), np.sin(th / 2), np.zeros(NB)], axis=1)
    else:
        q = random_quats(rng, NB)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        Y = po.fill_blob_pos(oc, ref, x, q)
        Fo = po.barrier_forces_cells(oc, Y, nbpb, cells).reshape(-1, 3)
        s.set_steric_cells(cells)
        F = s.barrier_forces(1.0)  # the cutoff only switches the forces on
        say("  rel", rel(F, Fo))
        assert np.abs(Fo).max() > 0
        assert rel(F, Fo) < 1e-12, rel(F, Fo)
        assert np.array_equal(F, s.barrier_forces(1.0))
        s.set_steric_cells(0)
        Fc = s.barrier_forces(2.2 * a)
        assert rel(Fc, po.barrier_forces(oc, Y, nbpb, 2.2 * a).reshape(-1, 3)) < 1e-12


def test_euler_maruyama_step_with_reference_cells_matches_oracle():
    n, NB, nbpb, dt, tol, cells = 32, 12, 6, 0.20216, 1e-10, 6
    gc, oc = pair_cfg(n, kT=1.0, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, max_krylov=150)
    ref, x, q, _, _ = fa.synth_rods(23, NB, nbpb, 1.86, [float(n)] * 3, max_attempts=20)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    xo, qo = x.ravel().copy(), q.ravel().copy()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(np.zeros((NB, 6)))
        s.set_steric_cells(cells)
        for step in range(2):
            seed = 700 + step
            Y = po.fill_blob_pos(oc, ref, xo, qo)
            o.set_up_gaussians(Y)
            o.fill_stress(seed)
            o.random_field()
            rhs = np.concatenate([-o.random_interp(),
                                  -po.mul_KT(ref, qo, po.barrier_forces_cells(oc, Y, nbpb, cells))])
            st, sol, it, _ = po.gmres(o, po.Bodies(ref, xo, qo, Minv, N), rhs, tol=tol, max_iter=150)
            assert st == 0
            xo, qo = po.end_of_step_update(oc, sol[3 * NB * nbpb:], dt, xo, qo)
            rep = s.em_step(seed, 1.0, tol=tol, max_iter=150)
            xg, qg = s.bodies()
            dx = np.abs(xg.ravel() - xo)
            dx = np.minimum(dx, n - dx)
            say("step", step, rep, "max |dx|", dx.max())
            assert rep["status"] == "converged"
            assert dx.max() < 1e-9 and np.abs(qg.ravel() - qo).max() < 1e-9
    o.close()
