"""The reference's shipped configuration through the boundary (config.hpp:
33,47-50,62-69,88): slip channel, 3 rods of 22 spheres in turned pairs
(collection_of_rigid_bodies.cpp:23,56-70), BLOB_RADIUS 1, DX = a / (1.86
sqrt(pi)), NUM_FCM_GAUSSIAN_POINTS 20.  Stencils this wide run the separate
spreading / FFT / interpolation passes (the column kernels serve ng <= 8).
Every operator is compared with the oracle (FP64 <= 1e-12, FP32 <= 2e-6),
the reference matrices directly, and a fluctuating solve to 1e-10."""
import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, rel
from _gdc import reference_rods

pytestmark = pytest.mark.gpu

A = 1.0
DX = A * 0.3033277330901916
DT = np.pi * A ** 3 / 100.0
NG = 20


def say(*a):
    print(*a, flush=True)


def _bodies(rng, NB, n, channel):
    """Rod centres in the box (between the walls in a channel, the rods
    turned about y only so that every sphere stays within a of its centre's y)."""
    L = n * DX
    x = rng.uniform(0, L, (NB, 3))
    if channel:
        x[:, 1] = rng.uniform(0.15 * L, 0.35 * L, NB)
        th = rng.uniform(0, 2 * np.pi, NB)
        q = np.stack([np.cos(th / 2), np.zeros(NB), np.sin(th / 2), np.zeros(NB)], axis=1)
    else:
        q = rng.standard_normal((NB, 4))
        q /= np.linalg.norm(q, axis=1)[:, None]
    return x, q


def _pair(n, NB, channel, kT=1.0, max_krylov=200):
    ref = reference_rods(22, A)
    nb = NB * len(ref)
    gc, oc = pair_cfg(n, ng=NG, dx=DX, a=A, mu=1.0, kT=kT, dt=DT, max_blobs=nb, max_bodies=NB,
                      max_blobs_per_body=len(ref), max_krylov=max_krylov, slip_channel=channel)
    return ref, gc, oc


@pytest.mark.parametrize("n,channel", [(64, True), (48, False)])
def test_operators_at_ng20_match_oracle(n, channel):
    say("ng 20 operators", n, "channel" if channel else "periodic")
    NB = 3
    ref, gc, oc = _pair(n, NB, channel)
    rng = np.random.default_rng(60 + n)
    x, q = _bodies(rng, NB, n, channel)
    Y = po.fill_blob_pos(oc, ref, x, q)
    nb = Y.size // 3
    if channel:
        assert np.all((Y[1::3] > 0) & (Y[1::3] < 0.5 * n * DX))
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    st, M, Minv, N = o.ref_matrices(ref)
    assert st == 0
    o.set_up_gaussians(Y)
    o.spread(F)
    f_ref = np.stack([po.from_padded(oc, o.field(c)) for c in ("fx", "fy", "fz")])
    u = rng.standard_normal((3, n, n, n))
    for c, name in enumerate(("ux", "uy", "uz")):
        o.field(name)[:] = po.to_padded(oc, u[c])
    V_int = o.interp()
    V_mob = o.mobility(F)
    b = po.Bodies(ref, x, q, Minv, N)
    v = rng.standard_normal(b.n)
    Av = po.saddle_matvec(o, b, v)
    o.fill_stress(9)
    o.random_field()
    Vr = o.random_interp()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        Mg, Minvg, Ng = s.ref_matrices()
        e = (rel(Mg, M), rel(Minvg, Minv), rel(Ng, N))
        say("  reference matrices", e)
        assert max(e) < 1e-12, e
        s.set_bodies(x, q)
        assert rel(s.blob_positions().ravel(), Y) < 1e-15
        for prec, tol in ((fa.FCM_PREC_FP64, 1e-12), (fa.FCM_PREC_FP32, 2e-6)):
            s.spread(F, prec)
            got = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
            es = rel(got, f_ref)
            for c in range(3):
                s.set_field(c, po.to_padded(oc, u[c]), prec)
            ei = rel(s.interp(prec), V_int)
            em = rel(s.mobility(F, prec), V_mob)
            ea = rel(s.saddle_matvec(v, prec), Av)
            say("  ", prec, "spread", es, "interp", ei, "mobility", em, "saddle", ea)
            assert max(es, ei, em, ea) < tol, (prec, es, ei, em, ea)
        s.random_field(None, 9)
        er = rel(s.random_interp(), Vr)
        say("  random blob velocities", er)
        assert er < 1e-12, er
    o.close()


def test_fluctuating_solve_at_ng20_matches_oracle():
    """The reference's default mode (slip channel, kT = 1) solved to 1e-10
    on both sides: body velocities and blob forces within 1e-10."""
    n, NB = 64, 3
    say("ng 20 fluctuating solve", n)
    ref, gc, oc = _pair(n, NB, True)
    rng = np.random.default_rng(71)
    x, q = _bodies(rng, NB, n, True)
    F6 = rng.standard_normal((NB, 6))
    nb = NB * len(ref)
    o = po.Fcm(oc, nb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    o.fill_stress(5)
    o.random_field()
    rhs = np.concatenate([-o.random_interp(), -F6.ravel()])
    st, sol_ref, it, _ = po.gmres(o, po.Bodies(ref, x, q, Minv, N), rhs, tol=1e-10, max_iter=200)
    assert st == 0
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        s.set_body_forces(F6)
        for prec in (fa.FCM_PREC_FP64, fa.FCM_PREC_FP32):
            rep = s.mobility_solve(5, tol=1e-10, max_iter=200, precision=prec)
            sol = s.solution()
            say("  ", prec, rep, "oracle iterations", it, "error", rel(sol, sol_ref))
            assert rep["status"] == "converged"
            assert rel(sol, sol_ref) < (1e-10 if prec == fa.FCM_PREC_FP64 else 1e-7)
    o.close()


def test_reference_default_grid_256_saddle_product():
    """config.hpp as shipped: 256^3, ng = 20, the slip channel and 3 rods of
    22 spheres; one saddle-point product against the oracle's."""
    n, NB = 256, 3
    say("ng 20 at 256^3")
    ref, gc, oc = _pair(n, NB, True, max_krylov=10)
    rng = np.random.default_rng(81)
    x, q = _bodies(rng, NB, n, True)
    o = po.Fcm(oc, NB * len(ref))
    o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
    Minv, N = np.eye(3 * len(ref)), np.eye(6)
    b = po.Bodies(ref, x, q, Minv, N)
    v = rng.standard_normal(b.n)
    Av = po.saddle_matvec(o, b, v)
    o.close()
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        assert s.tuning()["column"] == [0, 0], "ng = 20 must run the separate passes"
        for prec, tol in ((fa.FCM_PREC_FP64, 1e-12), (fa.FCM_PREC_FP32, 2e-6)):
            e = rel(s.saddle_matvec(v, prec), Av)
            say("  ", prec, e)
            assert e < tol, (prec, e)
