"""Pins the oracle's transforms (restating FFTW 2 rfftwnd_mpi, fft_solver.cpp:21-42).

Checks: brute-force O(N^2) DFT, numpy.fft, plane waves and deltas, the
round trip = N*I, Parseval, and FFTW's documented halfcomplex inverse
semantics for the kz = 0 and kz = nz/2 planes (no imaginary part stored).
"""
import numpy as np
import pytest

import pyoracle as po


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 24, 25, 30, 32, 48, 49, 64, 96])
def test_fft1d_matches_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        a = po.fft1d(x, sign)
        b = po.dft1d(x, sign)
        assert np.max(np.abs(a - b)) <= 1e-13 * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("n", [8, 12, 64, 256, 512])
def test_fft1d_matches_numpy(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(po.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * np.sqrt(n))
    assert np.allclose(po.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * np.sqrt(n))


def test_fft1d_plane_wave_and_delta():
    n = 24
    k0 = 5
    x = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    X = po.fft1d(x, -1)
    expect = np.zeros(n, complex)
    expect[k0] = n
    assert np.max(np.abs(X - expect)) < 1e-12
    d = np.zeros(n, complex)
    d[3] = 1.0
    D = po.fft1d(d, -1)
    assert np.max(np.abs(D - np.exp(-2j * np.pi * 3 * np.arange(n) / n))) < 1e-14


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 12, 12), (16, 12, 10), (6, 10, 8), (4, 6, 9)])
def test_r2c_matches_numpy_rfftn_in_transposed_order(dims):
    nx, ny, nz = dims
    c = po.cfg(nx, ny, nz)
    rng = np.random.default_rng(sum(dims))
    r = rng.standard_normal((nx, ny, nz))
    out = po.fft3d_r2c(c, po.to_padded(c, r))
    Z = po.kspace_to_complex(c, out)
    ref = np.fft.rfftn(r)
    assert np.max(np.abs(Z - ref)) < 1e-12 * np.sqrt(nx * ny * nz)


@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 12, 12), (16, 12, 10), (6, 10, 8)])
def test_c2r_round_trip_is_N_times_identity(dims):
    nx, ny, nz = dims
    c = po.cfg(nx, ny, nz)
    rng = np.random.default_rng(7 + sum(dims))
    r = rng.standard_normal((nx, ny, nz))
    k = po.fft3d_r2c(c, po.to_padded(c, r))
    back = po.from_padded(c, po.fft3d_c2r(c, k))
    assert np.max(np.abs(back - nx * ny * nz * r)) < 1e-11 * nx * ny * nz


def test_parseval():
    nx, ny, nz = 10, 8, 12
    c = po.cfg(nx, ny, nz)
    rng = np.random.default_rng(3)
    r = rng.standard_normal((nx, ny, nz))
    Z = po.kspace_to_complex(c, po.fft3d_r2c(c, po.to_padded(c, r)))
    w = np.full(nz // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    energy = np.sum(w[None, None, :] * np.abs(Z) ** 2) / (nx * ny * nz)
    assert abs(energy - np.sum(r * r)) < 1e-10 * np.sum(r * r)


def _halfcomplex_reference(c, Z):
    """FFTW's documented C2R: complex inverse along x and y, then for each
    (x,y) the halfcomplex->real transform
      y[n] = Re X0 + (-1)^n Re X_{N/2} + 2 sum_{k=1}^{N/2-1} (Re X_k cos - Im X_k sin)(2 pi k n/N)
    evaluated here as a direct sum (not an FFT)."""
    nx, ny, nz = c.nx, c.ny, c.nz
    W = np.fft.ifft(np.fft.ifft(Z, axis=0), axis=1) * nx * ny
    n = np.arange(nz)
    out = np.zeros((nx, ny, nz))
    out += W[:, :, 0].real[:, :, None]
    out += ((-1.0) ** n)[None, None, :] * W[:, :, nz // 2].real[:, :, None]
    for k in range(1, nz // 2):
        ang = 2 * np.pi * k * n / nz
        out += 2 * (W[:, :, k].real[:, :, None] * np.cos(ang)[None, None, :]
                    - W[:, :, k].imag[:, :, None] * np.sin(ang)[None, None, :])
    return out


@pytest.mark.parametrize("dims", [(8, 8, 8), (6, 10, 12)])
def test_c2r_non_hermitian_input_follows_halfcomplex_semantics(dims):
    """Arbitrary (non-Hermitian) complex input, including complex DC and
    Nyquist planes: the oracle must match FFTW's halfcomplex definition."""
    nx, ny, nz = dims
    c = po.cfg(nx, ny, nz)
    rng = np.random.default_rng(11)
    Z = rng.standard_normal((nx, ny, nz // 2 + 1)) + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))
    got = po.from_padded(c, po.fft3d_c2r(c, po.complex_to_kspace(c, Z)))
    ref = _halfcomplex_reference(c, Z)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))


def test_transposed_order_index_of_a_single_mode():
    """A single real plane wave cos(2 pi (kx x/NX + ky y/NY + kz z/NZ)) lands at
    index (ky*NX + kx)*pad + 2kz with value N/2 (transposed order)."""
    nx, ny, nz = 8, 6, 10
    c = po.cfg(nx, ny, nz)
    kx, ky, kz = 3, 2, 4
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    r = np.cos(2 * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
    out = po.fft3d_r2c(c, po.to_padded(c, r))
    pad = po.pad_of(c)
    idx = (ky * nx + kx) * pad + 2 * kz
    assert abs(out[idx] - nx * ny * nz / 2) < 1e-10
    out[idx] = 0.0
    assert np.max(np.abs(out)) < 1e-10
