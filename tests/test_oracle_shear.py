"""Pins the oracle's shear forcing (SHEARING_FORCES, config.hpp:32):
produce_shear_velocity_field (fcm_fluid_solver.cpp:703-840),
calculate_shear_blob_velocities (:488-550) and the x-velocity profile of
main.cpp:371-433.

The forcing fx = scale kT sin(2 pi kappa y/ny)/a depends on y only and points
along x, so q.f = 0 on every mode it has and the Stokes solve reduces to
u_x = f_x/(mu q_y^2) mode by mode: in a periodic box that is the single-mode
closed form, in a slip channel it is the 1-D spectral solve of the
image-extended forcing, computed here with numpy."""
import numpy as np
import pytest

import pyoracle as po

SCALE, KT, A, MU = 1.7, 1.3, 1.86, 1.2


def _field(f, name):
    c = f.c
    return f.field(name).reshape(c.nx, c.ny, f.pad)[:, :, :c.nz].copy()


@pytest.mark.parametrize("kappa", [1.0, 3.0])
def test_periodic_shear_flow_is_the_single_mode_closed_form(kappa):
    n, ny, dx = 16, 24, 0.8
    c = po.cfg(n, ny, n, dx=dx, a=A, mu=MU, kT=KT)
    f = po.Fcm(c)
    f.shear_field(SCALE, kappa)
    ux, uy, uz = (_field(f, nm) for nm in ("ux_shear", "uy_shear", "uz_shear"))
    q = 2 * np.pi * kappa / (ny * dx)
    y = np.arange(ny)
    want = SCALE * KT / A * np.sin(2 * np.pi * kappa * y / ny) / (MU * (q * q + 1e-16))
    scale = np.abs(want).max()
    assert np.abs(ux - want[None, :, None]).max() < 1e-12 * scale
    assert np.abs(uy).max() < 1e-14 * scale and np.abs(uz).max() < 1e-14 * scale


@pytest.mark.parametrize("kappa", [1.0, 2.0, 1.5])
def test_channel_shear_flow_is_the_image_extended_1d_solve(kappa):
    n, ny, dx = 16, 32, 1.0
    c = po.cfg(n, ny, n, dx=dx, a=A, mu=MU, kT=KT, slip_channel=True)
    f = po.Fcm(c)
    f.shear_field(SCALE, kappa)
    ux = _field(f, "ux_shear")
    y = np.arange(ny)
    fy = SCALE * KT / A * np.sin(2 * np.pi * kappa * y / ny)
    fy[ny // 2 + 1:] = fy[1:ny // 2][::-1]  # rows above ny/2 mirror rows below (fx even)
    F = np.fft.fft(fy)
    k = np.fft.fftfreq(ny, d=1.0 / ny)
    k[ny // 2] = ny // 2  # the reference's positive Nyquist wavenumber
    q2 = (2 * np.pi * k / (ny * dx)) ** 2 + 1e-16
    U = F / (MU * q2)
    U[0] = 0.0
    want = np.fft.ifft(U).real
    scale = np.abs(want).max()
    assert np.abs(ux - want[None, :, None]).max() < 1e-12 * scale
    assert np.abs(ux - ux[:, (ny - y) % ny]).max() < 1e-13 * scale  # even about both walls


def test_shear_blob_velocities_and_profile():
    """The blob velocity is the Gaussian-weighted profile (x, z weights sum
    out); the x-velocity profile adds the shear flow to the flow of the blob
    forces."""
    n = 24
    c = po.cfg(n, a=A, mu=MU, kT=KT, slip_channel=True)
    rng = np.random.default_rng(4)
    Y = rng.uniform(0, n, (9, 3))
    Y[:, 1] = rng.uniform(0.5, n / 2 - 0.5, 9)
    f = po.Fcm(c, 9)
    f.shear_field(SCALE, 1.0)
    f.set_up_gaussians(Y)
    V = f.shear_interp().reshape(-1, 3)
    prof = _field(f, "ux_shear")[0, :, 0]
    (gx, gy, gz), (ix, iy, iz) = f.gaussians()
    keep = iy < n // 2  # truncated Gaussians in a channel
    want = gx.sum(1) * gz.sum(1) * (gy * prof[iy] * keep).sum(1)
    assert np.abs(V[:, 0] - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(V[:, 1:]).max() < 1e-13 * np.abs(want).max()
    lam = rng.standard_normal(27)
    f.spread(lam)
    f.stokes()
    ux = _field(f, "ux")
    p = f.x_velocity_profile()
    assert np.abs(p - (ux.mean(axis=(0, 2)) + prof)).max() < 1e-13 * np.abs(p).max()
