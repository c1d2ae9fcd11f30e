"""Pins the oracle's slip-channel mode (SLIP_CHANNEL, config.hpp:33): the image
system in y of fcm_fluid_solver.cpp:186-217 (stress), :387-413 and :450-474
(truncated interpolation), :602-630 (force images), the mid-channel reference
body of collection_of_rigid_bodies.cpp:139-147 and the wall springs of
:640-662.

The checks compare the channel operators with the periodic ones on the same
inputs and test what the image system must produce: zero normal velocity on
both walls, even/odd symmetry of the fields about the walls, a symmetric
positive mobility for blobs away from the walls, the slip-wall image
asymptotics, and the doubled stress variance on the wall planes.
"""
import numpy as np
import pytest

import pyoracle as po


def _pair(n=32, **kw):
    cp = po.cfg(n, **kw)
    cc = po.cfg(n, slip_channel=True, **kw)
    return cp, cc


def _grid(f, name):
    c = f.c
    return f.field(name).reshape(c.nx, c.ny, f.pad).copy()


def _interior_blobs(rng, nb, n, margin=6.0):
    """Blobs whose 7-point y stencil stays inside rows 1 .. n/2 - 1."""
    Y = rng.uniform(0, n, (nb, 3))
    Y[:, 1] = rng.uniform(margin, n / 2 - margin, nb)
    return Y


def test_channel_spread_is_periodic_spread_with_images():
    """Rows 1 .. ny/2 - 1 hold the periodic spread; fy vanishes on the wall rows;
    rows above ny/2 are the mirror image (fy odd), so deposits that wrap through
    y = 0 (blob at y = 1.3) are discarded."""
    rng = np.random.default_rng(3)
    n = 32
    cp, cc = _pair(n)
    Y = _interior_blobs(rng, 6, n)
    Y = np.vstack([Y, [[7.2, 1.3, 20.1], [15.0, 14.2, 3.3]]])
    F = rng.standard_normal(Y.shape)
    out = {}
    for key, c in (("p", cp), ("c", cc)):
        f = po.Fcm(c, len(Y))
        f.set_up_gaussians(Y)
        f.spread(F)
        out[key] = [_grid(f, nm) for nm in ("fx", "fy", "fz")]
    h = n // 2
    for d in range(3):
        assert np.array_equal(out["c"][d][:, 1:h], out["p"][d][:, 1:h])
    # the blob at y = 1.3 reaches row ny - 1 by wrapping: discarded
    assert np.abs(out["p"][0][:, n - 1]).max() > 1e-3
    for d, sign in ((0, 1.0), (1, -1.0), (2, 1.0)):
        for j in range(h + 1, n):
            assert np.array_equal(out["c"][d][:, j], sign * out["c"][d][:, n - j])
    assert np.array_equal(out["c"][0][:, 0], out["p"][0][:, 0])
    assert not out["c"][1][:, 0].any() and not out["c"][1][:, h].any()
    assert np.abs(out["p"][1][:, 0]).max() > 1e-3  # the periodic fy there was not zero


def test_channel_velocity_slip_walls_and_symmetry():
    """u_y = 0 on both walls; u_x, u_z even and u_y odd about them."""
    rng = np.random.default_rng(5)
    n = 32
    _, cc = _pair(n)
    Y = rng.uniform(0, n, (12, 3))
    Y[:, 1] = rng.uniform(0.5, n / 2 - 0.5, 12)  # blobs anywhere in the channel
    F = rng.standard_normal(Y.shape)
    f = po.Fcm(cc, len(Y))
    f.set_up_gaussians(Y)
    f.spread(F)
    f.stokes()
    ux, uy, uz = (_grid(f, nm)[:, :, :n] for nm in ("ux", "uy", "uz"))
    scale = max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    h = n // 2
    mirror = (n - np.arange(n)) % n
    # f_y is odd, so it has no ky = ny/2 content and u_x, u_z are exactly even
    assert np.abs(ux - ux[:, mirror]).max() < 1e-13 * scale
    assert np.abs(uz - uz[:, mirror]).max() < 1e-13 * scale
    # u_y picks up a small even part from the Nyquist plane only
    vy = _drop_y_nyquist(uy)
    assert np.abs(uy - vy).max() < 1e-4 * scale
    assert np.abs(vy[:, 0]).max() < 1e-13 * scale
    assert np.abs(vy[:, h]).max() < 1e-13 * scale
    assert np.abs(vy + vy[:, mirror]).max() < 1e-13 * scale
    assert np.abs(uy[:, 3]).max() > 1e-3 * scale


def _drop_y_nyquist(u):
    """The field without its ky = ny/2 plane.  The reference's projection uses
    +pi/dx for qy on that plane whichever way the mode is read, so the Stokes
    solve commutes with the reflection y -> -y on every other plane only: the
    image symmetry is exact after this plane is removed."""
    U = np.fft.fft(u, axis=1)
    U[:, u.shape[1] // 2] = 0
    return np.fft.ifft(U, axis=1).real


def _mobility_matrix(c, Y):
    f = po.Fcm(c, len(Y))
    f.set_up_gaussians(Y)
    n3 = 3 * len(Y)
    M = np.zeros((n3, n3))
    for k in range(n3):
        e = np.zeros(n3)
        e[k] = 1.0
        M[:, k] = f.mobility(e)
    return M


def test_channel_mobility_symmetric_positive_and_not_periodic():
    """Away from the walls the image system is a Galerkin form of a symmetric
    positive operator (the reflection commutes with the Stokes projection only
    when fy is odd), so M is symmetric and positive definite up to the
    ky = ny/2 plane (a few 1e-7 here; an even fy leaves an asymmetry of the
    size of the image terms, ~1e-2)."""
    rng = np.random.default_rng(11)
    n = 32
    cp, cc = _pair(n)
    Y = _interior_blobs(rng, 5, n)
    Mc = _mobility_matrix(cc, Y)
    Mp = _mobility_matrix(cp, Y)
    assert np.abs(Mc - Mc.T).max() < 2e-6 * np.abs(Mc).max()
    assert np.linalg.eigvalsh(0.5 * (Mc + Mc.T)).min() > 0
    assert np.abs(Mc - Mp).max() > 1e-3 * np.abs(Mp).max()


def test_channel_self_mobility_matches_image_stokeslet():
    """A blob at height h above a slip wall feels the periodic mobility plus
    the cross mobility of its mirror image (x, LY - y, z) under the mirrored
    force (F_y negated).  The two differ only where the image's 7-point
    stencil is offset by one node from the mirror of the blob's, a weight of
    a few 1e-3 on an already small cross term."""
    n = 64
    cp, cc = _pair(n)
    for h in (7.3, 11.6):
        y = np.array([[31.4, h, 30.2]])
        img = np.array([[31.4, n - h, 30.2]])
        Mc = _mobility_matrix(cc, y)
        Mp2 = _mobility_matrix(cp, np.vstack([y, img]))
        pred = Mp2[:3, :3] @ np.eye(3) + Mp2[:3, 3:] @ np.diag([1.0, -1.0, 1.0])
        assert np.abs(Mc - pred).max() < 2e-3 * np.abs(Mc).max()
        # no penetration: the wall lowers the normal mobility
        assert Mc[1, 1] < 0.97 * Mp2[1, 1]


def test_channel_interp_truncates_rows_at_and_above_half():
    """Interpolation skips stencil rows j >= ny/2: a field living only there
    gives zero, a field living only below gives the periodic result."""
    rng = np.random.default_rng(7)
    n = 32
    cp, cc = _pair(n)
    h = n // 2
    Y = rng.uniform(0, n, (8, 3))
    Y[:4, 1] = rng.uniform(h - 3.0, h - 0.2, 4)   # stencils reach rows >= h
    Y[4:, 1] = rng.uniform(0.2, 3.0, 4)           # stencils wrap to rows near ny
    fp, fc = po.Fcm(cp, len(Y)), po.Fcm(cc, len(Y))
    for f in (fp, fc):
        f.set_up_gaussians(Y)
    u = rng.standard_normal((3, n, n, fp.pad))
    for part, rows in (("high", slice(h, n)), ("low", slice(0, h))):
        for f in (fp, fc):
            for d, nm in enumerate(("ux", "uy", "uz")):
                g = np.zeros((n, n, fp.pad))
                g[:, rows] = u[d][:, rows]
                f.field(nm)[:] = g.ravel()
        Vp, Vc = fp.interp(), fc.interp()
        if part == "high":
            assert not Vc.any()
            assert np.abs(Vp).min() > 0
        else:
            assert np.array_equal(Vc, Vp)


def test_channel_stress_images_and_wall_variance():
    """sxy, syz vanish on the wall rows and are odd; the others are even; the
    wall rows carry twice the variance of the interior rows."""
    n = 32
    cp, cc = _pair(n, kT=1.0, dt=0.2)
    fp, fc = po.Fcm(cp), po.Fcm(cc)
    fp.fill_stress(17)
    fc.fill_stress(17)
    h = n // 2
    mirror = (n - np.arange(n)) % n
    var1 = np.sqrt(2.0 * 1.0 * 1.0 / (1.0 * 0.2))
    for s, nm in enumerate(("sxx", "syy", "szz", "sxy", "sxz", "syz")):
        sc = _grid(fc, nm)[:, :, :n]
        sp = _grid(fp, nm)[:, :, :n]
        odd = nm in ("sxy", "syz")
        assert np.array_equal(sc, (-1.0 if odd else 1.0) * sc[:, mirror])
        assert np.array_equal(sc[:, 1:h], sp[:, 1:h])  # same stream below the walls
        wall = np.concatenate([sc[:, 0].ravel(), sc[:, h].ravel()])
        if odd:
            assert not wall.any()
        else:
            v = (np.sqrt(2.0) if s < 3 else 1.0) * var1
            ratio = wall.var() / (v * v)
            assert abs(ratio - 2.0) < 0.25, (nm, ratio)
            assert abs(sc[:, 1:h].var() / (v * v) - 1.0) < 0.1


def test_channel_random_velocity_slip_walls():
    n = 32
    _, cc = _pair(n, kT=1.0, dt=0.2)
    f = po.Fcm(cc)
    f.fill_stress(23)
    f.random_field()
    ux, uy, uz = (_grid(f, nm)[:, :, :n] for nm in ("ux_rand", "uy_rand", "uz_rand"))
    scale = np.abs(uy).max()
    mirror = (n - np.arange(n)) % n
    vy, vx = _drop_y_nyquist(uy), _drop_y_nyquist(ux)
    assert np.abs(vy[:, 0]).max() < 1e-12 * scale and np.abs(vy[:, n // 2]).max() < 1e-12 * scale
    assert np.abs(vy + vy[:, mirror]).max() < 1e-12 * scale
    assert np.abs(vx - vx[:, mirror]).max() < 1e-12 * np.abs(ux).max()


def test_channel_reference_body_sits_mid_channel():
    """The single-body matrices are formed at y = LY/4; there the rod's
    mobility is symmetric and differs from the periodic box-centre one."""
    n = 32
    cp, cc = _pair(n)
    ref = np.array([[2 * 1.86 * (m - 1.5), 0.0, 0.0] for m in range(4)])
    st_c, Mc, Minv_c, Nc = po.Fcm(cc, 4).ref_matrices(ref)
    st_p, Mp, _, _ = po.Fcm(cp, 4).ref_matrices(ref)
    assert st_c == 0 and st_p == 0
    Y = ref + np.array([n / 2, n / 4, n / 2])
    assert np.allclose(Mc, _mobility_matrix(cc, Y), rtol=0, atol=1e-15)
    assert np.abs(Mc - Mc.T).max() < 2e-6 * np.abs(Mc).max()
    assert np.abs(Mc - Mp).max() > 1e-3 * np.abs(Mp).max()


@pytest.mark.parametrize("y,expect", [(1.0, 1), (4.09, 1), (4.2, 0), (11.0, 0), (12.0, -1), (15.5, -1),
                                      (16.5, 0)])
def test_channel_wall_springs(y, expect):
    """Linear springs of range 2.2a push blobs off both walls, zero elsewhere
    (the blob's partner body is far away, so only the wall acts)."""
    n = 32
    c = po.cfg(n, kT=1.0, slip_channel=True)
    Y = np.array([[5.0, y, 5.0], [20.0, 8.0, 20.0]])
    F = po.barrier_forces(c, Y, 1, 3.0).reshape(-1, 3)
    a, L = 1.86, 2.2 * 1.86
    K = 40.0 / a / L
    want = {1: K * (L - y), 0: 0.0, -1: -K * (L - n / 2 + y)}[expect]
    assert F[0, 1] == pytest.approx(want, rel=1e-14, abs=0)
    assert F[0, 0] == 0 and F[0, 2] == 0
    Fp = po.barrier_forces(po.cfg(n, kT=1.0), Y, 1, 3.0).reshape(-1, 3)
    assert not Fp[0].any()
