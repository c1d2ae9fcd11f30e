"""GPU parity of the slip-channel mode (SLIP_CHANNEL, config.hpp:33) against the
oracle's channel mode, through the C ABI: force images after spreading on
every spreading kernel and on the fused column path
(fcm_fluid_solver.cpp:602-630), truncated interpolation on every
interpolation kernel (:387-413, :450-474), the image-symmetric seeded stress
(:186-217), the mid-channel reference body (collection_of_rigid_bodies.cpp:
139-147), the wall springs (:640-662), full FP64 and FP32+refine solves and
Euler-Maruyama steps.  Tolerances are those of the periodic tests: 1e-12
relative for FP64 operators, 2e-6 for the FP32 bulk path, 1e-10 for solves."""
import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, random_quats, rel, rod

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-6


def say(*a):
    print(*a, flush=True)


def _channel_bodies(rng, NB, nbpb, a, n, margin=1.0):
    """Rods placed with every blob in margin <= y <= n/2 - margin."""
    ref = rod(nbpb, a)
    xs, qs = [], []
    while len(xs) < NB:
        x = rng.uniform(0, n, 3)
        x[1] = rng.uniform(0, n / 2)
        q = random_quats(rng, 1)[0]
        y = po.fill_blob_pos(po.cfg(n), ref, x[None, :], q[None, :]).reshape(-1, 3)[:, 1]
        if y.min() >= margin and y.max() <= n / 2 - margin:
            xs.append(x)
            qs.append(q)
    return ref, np.array(xs), np.array(qs)


@pytest.mark.parametrize("dims,ng,nb", [((32, 32, 32), 7, 300), ((16, 24, 16), 5, 80), ((48, 32, 96), 7, 300),
                                        ((32, 48, 32), 8, 400)])
def test_channel_spread_interp_variants_match_oracle(dims, ng, nb):
    """Blobs anywhere in y, so stencils wrap through y = 0 and cross y = ny/2."""
    say("channel variants", dims, ng, nb)
    dx = 0.9
    gc, oc = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb, slip_channel=True)
    rng = np.random.default_rng(31)
    Y = rng.uniform(0, 1, (nb, 3)) * np.array(dims)[None, :] * dx
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    o.set_up_gaussians(Y)
    o.spread(F)
    ref_f = np.stack([po.from_padded(oc, o.field(n)) for n in ("fx", "fy", "fz")])
    u = rng.standard_normal((3,) + dims)
    for c, name in enumerate(("ux", "uy", "uz")):
        o.field(name)[:] = po.to_padded(oc, u[c])
    ref_V = o.interp()
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        for sv, iz in ((0, 1), (1, 2), (2, 3), (0, 3)):
            s.set_tuning((sv, sv), (iz, iz))
            for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
                s.spread(F, prec)
                got = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
                assert rel(got, ref_f) < tol, ("spread", sv, prec, rel(got, ref_f))
                for c in range(3):
                    s.set_field(c, po.to_padded(oc, u[c]), prec)
                V = s.interp(prec)
                assert rel(V, ref_V) < tol, ("interp", iz, prec, rel(V, ref_V))


@pytest.mark.parametrize("dims,ng,nb", [((32, 32, 32), 7, 300), ((16, 16, 64), 7, 600), ((32, 16, 48), 8, 200),
                                        ((16, 16, 256), 7, 200), ((16, 16, 512), 8, 200), ((32, 24, 32), 7, 200),
                                        ((32, 12, 32), 5, 100)])
def test_channel_mobility_column_and_separate_match_oracle(dims, ng, nb):
    say("channel mobility", dims, ng, nb)
    dx = 0.9
    gc, oc = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb, slip_channel=True)
    rng = np.random.default_rng(32)
    Y = rng.uniform(0, 1, (nb, 3)) * np.array(dims)[None, :] * dx
    Y[:, 1] *= 0.5  # inside the channel, stencils still cross both walls
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    o.set_up_gaussians(Y)
    ref = o.mobility(F)
    gp, op = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb)
    po_ = po.Fcm(op, nb)
    po_.set_up_gaussians(Y)
    assert rel(po_.mobility(F), ref) > 1e-3  # the channel operator is not the periodic one
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        t = s.tuning()
        # (32, 12, 32) is too short in y for the column path: separate passes only
        assert t["column"] == ([0, 0] if dims[1] < 16 else [1, 1]), t
        for column in (((1, 1), (0, 0)) if t["column"] == [1, 1] else ((0, 0),)):
            s.set_tuning(t["spread_variant"], t["interp_izb"], column=column)
            for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
                V = s.mobility(F, prec)
                assert rel(V, ref) < tol, (column, prec, rel(V, ref))
                assert np.array_equal(V, s.mobility(F, prec))


def test_channel_random_field_seeded_stream_matches_oracle():
    say("channel random field")
    dims = (16, 16, 32)
    gc, oc = pair_cfg(*dims, kT=1.0, dt=0.20216, max_blobs=40, slip_channel=True)
    o = po.Fcm(oc, 40)
    o.fill_stress(4321)
    o.random_field()
    ref = np.stack([po.from_padded(oc, o.field(n)) for n in ("ux_rand", "uy_rand", "uz_rand")])
    rng = np.random.default_rng(33)
    Y = rng.uniform(0, 16, (40, 3))
    Y[:, 1] *= 0.5
    o.set_up_gaussians(Y)
    refV = o.random_interp()
    with fa.FcmSolver(gc) as s:
        s.random_field(None, 4321)
        got = np.stack([po.from_padded(oc, s.get_field(3 + c)) for c in range(3)])
        assert rel(got, ref) < FP64_TOL, rel(got, ref)
        s.set_up_gaussians(Y)
        assert rel(s.random_interp(), refV) < FP64_TOL


def test_channel_ref_matrices_Pinv_and_saddle_product_match_oracle():
    say("channel ref matrices")
    n, NB, nbpb = 48, 10, 6
    gc, oc = pair_cfg(n, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, slip_channel=True)
    rng = np.random.default_rng(34)
    ref, x, q = _channel_bodies(rng, NB, nbpb, 1.86, n)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    assert st == 0
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        Mg, Minvg, Ng = s.ref_matrices()
        assert rel(Mg, M) < 1e-12, rel(Mg, M)
        s.set_ref_matrices(Minv, N)
        b = po.Bodies(ref, x, q, Minv, N)
        o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
        v = rng.standard_normal(b.n)
        assert rel(s.saddle_matvec(v), po.saddle_matvec(o, b, v)) < FP64_TOL


def _solve_pair(n, NB, nbpb, seed, kT=0.0):
    gc, oc = pair_cfg(n, kT=kT, dt=0.20216, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      max_krylov=200, slip_channel=True)
    rng = np.random.default_rng(seed)
    ref, x, q = _channel_bodies(rng, NB, nbpb, 1.86, n)
    F6 = rng.standard_normal((NB, 6))
    for i in range(NB):  # a rod takes no torque about its own axis
        axis = po.quat_matrix(q[i]).reshape(3, 3)[:, 0]
        F6[i, 3:] -= (F6[i, 3:] @ axis) * axis
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    b = po.Bodies(ref, x, q, Minv, N)
    s = fa.FcmSolver(gc)
    s.set_geometry(ref)
    s.set_bodies(x, q)
    s.set_ref_matrices(Minv, N)
    s.set_body_forces(F6)
    return s, o, b, F6


def test_channel_fp64_and_fp32_refined_solves_match_oracle():
    say("channel solves")
    NB, nbpb = 40, 6
    s, o, b, F6 = _solve_pair(64, NB, nbpb, seed=35)
    n3 = 3 * NB * nbpb
    rhs = np.concatenate([np.zeros(n3), -F6.ravel()])
    st, ref, it_o, _ = po.gmres(o, b, rhs, tol=1e-8, max_iter=200)
    assert st == 0
    sol, rep = s.solve(rhs, tol=1e-8, max_iter=200)
    say("oracle iters", it_o, "gpu", rep)
    assert rep["status"] == "converged"
    assert abs(rep["iterations"] - it_o) <= 1
    assert rel(sol, ref) < 1e-10, rel(sol, ref)
    sol32, rep32 = s.solve(rhs, tol=1e-8, max_iter=200, precision=fa.FCM_PREC_FP32)
    say("fp32 refine", rep32)
    assert rep32["status"] == "converged"
    true = np.linalg.norm(rhs - po.saddle_matvec(o, b, sol32)) / np.linalg.norm(rhs)
    assert true < 2e-8
    assert rel(sol32[n3:], ref[n3:]) < 1e-6
    s.close()


def test_channel_fluctuating_solve_matches_oracle():
    say("channel fluctuating solve")
    s, o, b, F6 = _solve_pair(32, 8, 3, seed=36, kT=1.0)
    o.fill_stress(91)
    o.random_field()
    Vr = o.random_interp()
    rhs = np.concatenate([-Vr, -F6.ravel()])
    st, ref, it_o, _ = po.gmres(o, b, rhs, tol=1e-8, max_iter=200)
    rep = s.mobility_solve(91, tol=1e-8, max_iter=200)
    say("oracle iters", it_o, "gpu", rep)
    assert rep["status"] == "converged"
    assert rel(s.solution(), ref) < 1e-10, rel(s.solution(), ref)
    s.close()


def test_channel_wall_springs_and_steps_match_oracle():
    """Barrier forces with the wall springs, then Euler-Maruyama steps."""
    n, NB, nbpb, dt, cutoff, tol = 32, 10, 3, 0.20216, 2.2 * 1.86, 1e-10
    gc, oc = pair_cfg(n, kT=1.0, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      max_krylov=200, slip_channel=True)
    rng = np.random.default_rng(37)
    ref, x, q = _channel_bodies(rng, NB, nbpb, 1.86, n, margin=0.5)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    xo, qo = x.ravel().copy(), q.ravel().copy()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        F = s.barrier_forces(cutoff)
        Fo = po.barrier_forces(oc, po.fill_blob_pos(oc, ref, x, q), nbpb, cutoff).reshape(-1, 3)
        Fp = po.barrier_forces(po.cfg(n, kT=1.0), po.fill_blob_pos(oc, ref, x, q), nbpb, cutoff).reshape(-1, 3)
        assert np.abs(Fo - Fp).max() > 0  # some blob is within reach of a wall
        assert rel(F, Fo) < 1e-12, rel(F, Fo)
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(np.zeros((NB, 6)))
        for step in range(2):
            seed = 700 + step
            Y = po.fill_blob_pos(oc, ref, xo, qo)
            o.set_up_gaussians(Y)
            o.fill_stress(seed)
            o.random_field()
            Vr = o.random_interp()
            f = po.barrier_forces(oc, Y, nbpb, cutoff)
            F6 = po.mul_KT(ref, qo, f)
            rhs = np.concatenate([-Vr, -F6])
            b = po.Bodies(ref, xo, qo, Minv, N)
            st, sol, it, _ = po.gmres(o, b, rhs, tol=tol, max_iter=200)
            assert st == 0
            xo, qo = po.end_of_step_update(oc, sol[3 * NB * nbpb:], dt, xo, qo)
            rep = s.em_step(seed, cutoff, tol=tol, max_iter=200)
            xg, qg = s.bodies()
            d = np.abs(xg.ravel() - xo)
            d = np.minimum(d, n - d)
            say("step", step, rep, "oracle iterations", it, "max |dx|", d.max())
            assert rep["status"] == "converged"
            assert d.max() < 1e-9
            assert np.abs(qg.ravel() - qo).max() < 1e-9
    o.close()


def test_channel_stokes_entry_takes_the_fields_as_given():
    """fcm_stokes (produce_deterministic_velocity_field) solves on the force
    fields as stored: only spreading makes the images, so a field without the
    mirror symmetry gives the oracle's periodic-style answer."""
    say("channel stokes entry")
    dims = (16, 24, 16)
    gc, oc = pair_cfg(*dims, dx=0.8, mu=1.3, max_blobs=4, slip_channel=True)
    rng = np.random.default_rng(38)
    f = rng.standard_normal((3,) + dims)
    o = po.Fcm(oc)
    for c, name in enumerate(("fx", "fy", "fz")):
        o.field(name)[:] = po.to_padded(oc, f[c])
    o.stokes()
    ref = np.stack([po.from_padded(oc, o.field(n)) for n in ("ux", "uy", "uz")])
    with fa.FcmSolver(gc) as s:
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            for c in range(3):
                s.set_field(c, po.to_padded(oc, f[c]), prec)
            s.stokes(prec)
            got = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
            assert rel(got, ref) < tol, (prec, rel(got, ref))


def test_channel_wall_overlap_and_reset_to_start_of_step():
    """check_for_channel_wall_overlap after a step that carries a blob past the
    upper wall, then reset_to_start_of_step restores bodies and blobs exactly
    (collection_of_rigid_bodies.cpp:207-239,759-796; main.cpp:304-326)."""
    say("channel wall overlap / reset")
    n, NB, nbpb, dt = 32, 4, 3, 0.2
    gc, oc = pair_cfg(n, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb, slip_channel=True)
    ref = rod(nbpb, 1.86)
    x = np.array([[8.0, 8.0, 8.0], [20.0, 12.0, 20.0], [10.0, 5.0, 5.0], [25.0, 9.0, 25.0]])
    q = np.tile([1.0, 0.0, 0.0, 0.0], (NB, 1))
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        with pytest.raises(fa.FcmError):
            s.reset_to_start_of_step()  # no step to undo
        assert not s.channel_wall_overlap()
        Y0 = s.blob_positions()
        V6 = np.zeros((NB, 6))
        V6[1, 1] = 4.5 / dt  # body 1 (all blobs at y = 12) to y = 16.5 > LY/2
        s.end_of_step_update(V6, nu=1.0)
        xg, _ = s.bodies()
        assert abs(xg[1, 1] - 16.5) < 1e-12
        assert s.channel_wall_overlap()
        s.reset_to_start_of_step()
        xr, qr = s.bodies()
        assert np.array_equal(xr, x) and np.array_equal(qr, q)
        assert np.array_equal(s.blob_positions(), Y0)
        assert not s.channel_wall_overlap()
        V6[1, 1] = 3.9 / dt  # to y = 15.9: inside
        s.end_of_step_update(V6, nu=1.0)
        assert not s.channel_wall_overlap()
        V6[:] = 0.0
        V6[2, 1] = -5.0 / dt  # body 2 to y = 0 exactly: on the lower wall
        s.end_of_step_update(V6, nu=1.0)
        assert s.channel_wall_overlap()
    gp, _ = pair_cfg(n, dt=dt, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb)
    with fa.FcmSolver(gp) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        V6 = np.zeros((NB, 6))
        V6[1, 1] = 4.5 / dt
        s.end_of_step_update(V6, nu=1.0)
        assert not s.channel_wall_overlap()  # no walls in a periodic box
