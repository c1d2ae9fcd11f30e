"""Pins the oracle's FCM operators against closed forms, brute force and
invariants (the reference ships no tests or golden vectors).

Covers fcm_fluid_solver.cpp:71-165 (tables, Gaussians), :552-600 (spread),
:636-701 (Stokes solve), :363-424 (interpolation), :167-361 (fluctuating
stress), the seeded stress stream, and the quaternion/K/K^T support math.
"""
import numpy as np
import pytest

import pyoracle as po

PI = np.pi


def _rand_positions(rng, nb, L):
    return rng.uniform(0, 1, (nb, 3)) * np.asarray(L)[None, :]


# --------------------------------------------------------------------------
# tables and Gaussians
# --------------------------------------------------------------------------

def test_wavenumber_tables_nyquist_positive():
    c = po.cfg(8, 6, 10, dx=0.5)
    f = po.Fcm(c)
    qx, qy, qp = f.tables()
    LX, LY, LZ = 8 * 0.5, 6 * 0.5, 10 * 0.5
    assert np.allclose(qx, 2 * PI / LX * np.array([0, 1, 2, 3, 4, -3, -2, -1]))
    assert np.allclose(qy, 2 * PI / LY * np.array([0, 1, 2, 3, -2, -1]))
    assert np.allclose(qp, 2 * PI / LZ * np.repeat(np.arange(6), 2))


def test_gaussian_anchor_formula_and_closed_form():
    """xc = int(Y/dx); node = xc - ng/2 + i + 1, wrapped; weight is the
    normalised 1-D Gaussian of sigma = a/sqrt(pi) at the unwrapped node."""
    c = po.cfg(16, ng=7, dx=1.0, a=1.86)
    f = po.Fcm(c, 4)
    Y = np.array([[5.3, 0.2, 15.9], [0.0, 8.5, 1.0]])
    f.set_up_gaussians(Y)
    g, ix = f.gaussians()
    assert list(ix[0][0]) == [3, 4, 5, 6, 7, 8, 9]
    assert list(ix[1][0]) == [14, 15, 0, 1, 2, 3, 4]       # yc = 0 -> -2..4 wrapped
    assert list(ix[2][0]) == [13, 14, 15, 0, 1, 2, 3]      # zc = 15 -> 13..19 wrapped
    sigma = 1.86 / np.sqrt(PI)
    for d in range(3):
        xc = int(Y[0, d])
        nodes = xc - 3 + np.arange(1, 8)
        expect = np.exp(-(nodes - Y[0, d]) ** 2 / (2 * sigma ** 2)) / np.sqrt(2 * PI * sigma ** 2)
        assert np.allclose(g[d][0], expect, rtol=1e-13, atol=0)


def test_gaussian_weights_sum_to_one_for_wide_stencil():
    c = po.cfg(32, ng=16, dx=1.0, a=1.86)
    f = po.Fcm(c, 8)
    rng = np.random.default_rng(1)
    Y = _rand_positions(rng, 8, [32, 32, 32])
    f.set_up_gaussians(Y)
    g, _ = f.gaussians()
    for d in range(3):
        assert np.allclose(g[d].sum(axis=1), 1.0, atol=1e-8)


def test_gaussian_whole_cell_translation_is_exact():
    c = po.cfg(16, ng=7, dx=1.0)
    f = po.Fcm(c, 8)
    rng = np.random.default_rng(2)
    Y = np.round(_rand_positions(rng, 5, [13, 13, 13]) * 1024) / 1024   # exact after +3
    f.set_up_gaussians(Y)
    g1, i1 = f.gaussians()
    f.set_up_gaussians(Y + 3.0)
    g2, i2 = f.gaussians()
    for d in range(3):
        assert np.array_equal(g1[d], g2[d])
        assert np.array_equal((i1[d] + 3) % 16, i2[d])


# --------------------------------------------------------------------------
# spread / interpolation
# --------------------------------------------------------------------------

def _brute_spread(c, Y, F):
    """Node-driven restatement: for each node and blob, the stencil offset o
    is the unique value in [1-ng/2, 1-ng/2+n) congruent to node-anchor; the
    blob contributes only if o <= ng - ng/2."""
    n = np.array([c.nx, c.ny, c.nz])
    ng = c.ng
    omin, omax = 1 - ng // 2, ng - ng // 2
    sig2 = c.a ** 2 / PI
    out = np.zeros((3, c.nx, c.ny, c.nz))
    for b in range(len(Y)):
        w = []
        for d in range(3):
            anc = int(Y[b, d] / c.dx)
            nodes = np.arange(n[d])
            o = (nodes - anc - omin) % n[d] + omin
            val = np.exp(-((anc + o) * c.dx - Y[b, d]) ** 2 / (2 * sig2)) / np.sqrt(2 * PI * sig2)
            w.append(np.where(o <= omax, val, 0.0))
        wt = w[0][:, None, None] * w[1][None, :, None] * w[2][None, None, :]
        for comp in range(3):
            out[comp] += F[3 * b + comp] * wt
    return out


@pytest.mark.parametrize("dims,ng", [((12, 12, 12), 7), ((16, 10, 8), 7), ((10, 10, 10), 10)])
def test_spread_matches_bruteforce(dims, ng):
    c = po.cfg(*dims, ng=ng)
    rng = np.random.default_rng(4)
    nb = 9
    Y = _rand_positions(rng, nb, dims)
    F = rng.standard_normal(3 * nb)
    f = po.Fcm(c, nb)
    f.set_up_gaussians(Y)
    f.spread(F)
    got = np.stack([po.from_padded(c, f.field(n)) for n in ("fx", "fy", "fz")])
    ref = _brute_spread(c, Y, F)
    assert np.max(np.abs(got - ref)) < 1e-14 * np.max(np.abs(ref))


def test_interp_is_adjoint_of_spread():
    c = po.cfg(16, 12, 14, ng=7, dx=0.7)
    rng = np.random.default_rng(5)
    nb = 20
    Y = _rand_positions(rng, nb, [16 * 0.7, 12 * 0.7, 14 * 0.7])
    F = rng.standard_normal(3 * nb)
    f = po.Fcm(c, nb)
    f.set_up_gaussians(Y)
    f.spread(F)
    spread = np.stack([po.from_padded(c, f.field(n)) for n in ("fx", "fy", "fz")])
    u = rng.standard_normal((3, 16, 12, 14))
    for comp, name in enumerate(("ux", "uy", "uz")):
        f.field(name)[:] = po.to_padded(c, u[comp])
    V = f.interp()
    lhs = np.sum(spread * u) * c.dx ** 3
    rhs = np.dot(F, V)
    assert abs(lhs - rhs) < 1e-12 * (abs(lhs) + 1)


def test_interp_of_constant_field():
    c = po.cfg(12, ng=7, dx=0.9)
    rng = np.random.default_rng(6)
    Y = _rand_positions(rng, 4, [12 * 0.9] * 3)
    f = po.Fcm(c, 4)
    f.set_up_gaussians(Y)
    for comp, name in enumerate(("ux", "uy", "uz")):
        f.field(name)[:] = po.to_padded(c, np.full((12, 12, 12), comp + 1.0))
    V = f.interp().reshape(-1, 3)
    g, _ = f.gaussians()
    s = g[0].sum(1) * g[1].sum(1) * g[2].sum(1) * c.dx ** 3
    assert np.allclose(V, s[:, None] * np.array([1.0, 2.0, 3.0])[None, :], rtol=1e-14)


# --------------------------------------------------------------------------
# Stokes solve (fcm_fluid_solver.cpp:636-701)
# --------------------------------------------------------------------------

def _grid(c):
    return np.meshgrid(np.arange(c.nx) * c.dx, np.arange(c.ny) * c.dx, np.arange(c.nz) * c.dx, indexing="ij")


def _stokes(c, fxyz):
    f = po.Fcm(c)
    for comp, name in enumerate(("fx", "fy", "fz")):
        f.field(name)[:] = po.to_padded(c, fxyz[comp])
    f.stokes()
    return np.stack([po.from_padded(c, f.field(n)) for n in ("ux", "uy", "uz")])


def test_stokes_shear_wave_exact():
    c = po.cfg(16, 12, 8, dx=0.5, mu=1.7)
    X, Yg, Z = _grid(c)
    LY = c.ny * c.dx
    k = 2 * PI / LY
    fx = np.cos(k * Yg)
    u = _stokes(c, [fx, 0 * fx, 0 * fx])
    assert np.max(np.abs(u[0] - fx / (c.mu * k * k))) < 1e-13
    assert np.max(np.abs(u[1:])) < 1e-13


def test_stokes_compressive_wave_projected_out_and_mean_removed():
    c = po.cfg(12, 12, 12)
    X, Yg, Z = _grid(c)
    k = 2 * PI / 12
    u = _stokes(c, [np.cos(k * X) + 3.0, 0 * X + 1.0, 0 * X])
    assert np.max(np.abs(u)) < 1e-13


def test_stokes_nyquist_plane_keeps_only_the_real_part():
    """f_x = cos(2 pi x/L) (-1)^z lives on the kz = nz/2 plane where the
    reference projects with q3 = +pi/dx (not zeroed), which is not Hermitian
    in (kx, ky).  FFTW's C2R keeps only Re there, so u_z = 0 and
    u_x = pi^2 f_x / (mu (k^2 + pi^2)^2)."""
    c = po.cfg(8, 8, 8, dx=1.0, mu=1.0)
    X, Yg, Z = _grid(c)
    k = 2 * PI / 8
    fx = np.cos(k * X) * (-1.0) ** Z
    u = _stokes(c, [fx, 0 * fx, 0 * fx])
    q2 = k * k + PI * PI
    assert np.max(np.abs(u[0] - PI * PI * fx / q2 ** 2)) < 1e-13
    assert np.max(np.abs(u[1])) < 1e-13
    assert np.max(np.abs(u[2])) < 1e-13


def test_stokes_divergence_free_off_nyquist():
    c = po.cfg(12, 10, 8, dx=0.8)
    rng = np.random.default_rng(8)
    u = _stokes(c, rng.standard_normal((3, 12, 10, 8)))
    U = [np.fft.fftn(u[i]) for i in range(3)]
    kx = 2 * PI * np.fft.fftfreq(12, 0.8)
    ky = 2 * PI * np.fft.fftfreq(10, 0.8)
    kz = 2 * PI * np.fft.fftfreq(8, 0.8)
    KX, KY, KZ = np.meshgrid(kx, ky, kz, indexing="ij")
    div = KX * U[0] + KY * U[1] + KZ * U[2]
    mask = (np.abs(KX) < PI / 0.8 - 1e-9) & (np.abs(KY) < PI / 0.8 - 1e-9) & (np.abs(KZ) < PI / 0.8 - 1e-9)
    assert np.max(np.abs(div[mask])) < 1e-11 * np.max(np.abs(U[0]))


# --------------------------------------------------------------------------
# full mobility
# --------------------------------------------------------------------------

def _mobility_matrix(c, Y):
    f = po.Fcm(c, len(Y))
    f.set_up_gaussians(Y)
    n = 3 * len(Y)
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = f.mobility(e)
    return M


def test_mobility_symmetric_positive_definite():
    c = po.cfg(16, 16, 16, ng=7)
    rng = np.random.default_rng(9)
    Y = _rand_positions(rng, 5, [16] * 3)
    M = _mobility_matrix(c, Y)
    assert np.max(np.abs(M - M.T)) < 1e-15 * 1e3 * np.max(np.abs(M))
    assert np.min(np.linalg.eigvalsh(0.5 * (M + M.T))) > 0


def test_self_mobility_matches_hasimoto():
    """Single blob, wide stencil: M_self = (1 - 2.837297 a/L + 4.19 (a/L)^3)/(6 pi mu a)."""
    L = 48
    c = po.cfg(L, ng=14, dx=1.0, a=1.86, mu=1.0)
    M = _mobility_matrix(c, np.array([[20.3, 17.6, 31.1]]))
    r = 1.86 / L
    expect = (1 - 2.837297 * r + 4.19 * r ** 3) / (6 * PI * 1.86)
    assert np.allclose(np.diag(M), expect, rtol=1e-3)
    assert np.max(np.abs(M - np.diag(np.diag(M)))) < 1e-6 * expect


# --------------------------------------------------------------------------
# seeded stress stream and fluctuating-stress field (fcm_fluid_solver.cpp:167-361)
# --------------------------------------------------------------------------

def test_splitmix64_known_answers():
    assert po.splitmix64_seq(0, 1)[0] == 0xE220A8397B1DCDAF
    mask = (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return z ^ (z >> 31)

    s = 1234567
    expect = [mix((s + (i + 1) * 0x9E3779B97F4A7C15) & mask) for i in range(5)]
    assert po.splitmix64_seq(s, 5) == expect


def test_stress_normal_statistics():
    nn = 200000
    x = np.array([po.stress_normal(42, 3, i, nn) for i in range(nn)])
    assert abs(x.mean()) < 5 / np.sqrt(nn)
    assert abs(x.var() - 1) < 6 * np.sqrt(2 / nn)
    assert abs(np.mean(x ** 4) - 3) < 0.05
    y = np.array([po.stress_normal(42, 4, i, nn) for i in range(2000)])
    assert abs(np.corrcoef(x[:2000], y)[0, 1]) < 0.1
    assert po.stress_normal(42, 3, 17, nn) == x[17]


def test_fill_stress_variances():
    c = po.cfg(16, dx=0.5, kT=2.0, mu=1.5, dt=0.3)
    f = po.Fcm(c)
    f.fill_stress(7)
    var1 = np.sqrt(2 * 2.0 * 1.5 / (0.5 ** 3 * 0.3))
    var2 = np.sqrt(2) * var1
    for name, v in (("sxx", var2), ("syy", var2), ("szz", var2), ("sxy", var1), ("sxz", var1), ("syz", var1)):
        s = po.from_padded(c, f.field(name))
        assert abs(s.std() / v - 1) < 0.05
        assert np.all(f.field(name).reshape(16, 16, -1)[:, :, 16:] == 0)
    nn = 16 ** 3
    assert f.field("syz")[0] == var1 * po.stress_normal(7, 5, 0, nn)


def _random_field_from(c, stress):
    f = po.Fcm(c)
    for name, s in stress.items():
        f.field(name)[:] = po.to_padded(c, s)
    f.random_field()
    return np.stack([po.from_padded(c, f.field(n)) for n in ("ux_rand", "uy_rand", "uz_rand")])


def test_random_field_shear_stress_kat():
    """sxy = cos(k y): f = div S has f_x = -k sin(k y), so u_x = -sin(k y)/(mu k)."""
    c = po.cfg(10, 12, 8, dx=0.5, mu=1.3)
    X, Yg, Z = _grid(c)
    k = 2 * PI / (12 * 0.5)
    zero = 0 * X
    u = _random_field_from(c, {"sxy": np.cos(k * Yg), "sxx": zero, "syy": zero, "szz": zero, "sxz": zero, "syz": zero})
    assert np.max(np.abs(u[0] + np.sin(k * Yg) / (c.mu * k))) < 1e-12
    assert np.max(np.abs(u[1:])) < 1e-12


def test_random_field_nyquist_wavenumbers_zeroed_in_divergence():
    c = po.cfg(8, 8, 8)
    X, Yg, Z = _grid(c)
    zero = 0 * X
    u = _random_field_from(c, {"sxx": (-1.0) ** X, "syy": (-1.0) ** Yg, "szz": (-1.0) ** Z,
                               "sxy": (-1.0) ** X, "sxz": (-1.0) ** Z, "syz": (-1.0) ** Yg})
    assert np.max(np.abs(u)) < 1e-13


def test_random_field_compressive_stress_projected_out():
    c = po.cfg(8, 8, 12)
    X, Yg, Z = _grid(c)
    zero = 0 * X
    u = _random_field_from(c, {"szz": np.cos(2 * PI * Z / 12), "sxx": zero, "syy": zero, "sxy": zero, "sxz": zero, "syz": zero})
    assert np.max(np.abs(u)) < 1e-13


# --------------------------------------------------------------------------
# quaternion / K / K^T
# --------------------------------------------------------------------------

def test_quaternion_matrix_is_rodrigues_rotation():
    rng = np.random.default_rng(10)
    for _ in range(5):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        th = rng.uniform(-3, 3)
        q = np.concatenate([[np.cos(th / 2)], np.sin(th / 2) * n])
        K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
        Rr = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
        assert np.allclose(po.quat_matrix(q), Rr, atol=1e-14)


def test_K_and_KT_are_adjoint_and_rigid():
    rng = np.random.default_rng(11)
    NB, nbpb = 3, 4
    ref = rng.standard_normal((nbpb, 3))
    q = rng.standard_normal((NB, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    v = rng.standard_normal(6 * NB)
    lam = rng.standard_normal(3 * NB * nbpb)
    assert abs(np.dot(po.mul_K(ref, q, v), lam) - np.dot(v, po.mul_KT(ref, q, lam))) < 1e-12
    # pure rotation about z of body 0: blob velocity = omega x (R r)
    v = np.zeros(6 * NB)
    v[5] = 2.0
    out = po.mul_K(ref, q, v).reshape(NB, nbpb, 3)
    R = po.quat_matrix(q[0])
    for m in range(nbpb):
        assert np.allclose(out[0, m], np.cross([0, 0, 2.0], R @ ref[m]), atol=1e-14)
    assert np.allclose(out[1:], 0)


def test_fill_blob_pos_wraps_into_box():
    c = po.cfg(10, dx=1.0)
    ref = np.array([[0, 0, 0], [3.0, 0, 0], [-3.0, 0, 0]])
    x = np.array([[9.5, 0.5, 5.0]])
    q = np.array([[1.0, 0, 0, 0]])
    Y = po.fill_blob_pos(c, ref, x, q).reshape(-1, 3)
    assert np.allclose(Y, [[9.5, 0.5, 5.0], [2.5, 0.5, 5.0], [6.5, 0.5, 5.0]])


def test_dense_inverses():
    rng = np.random.default_rng(12)
    A = rng.standard_normal((9, 9))
    A = A @ A.T + 9 * np.eye(9)
    st, Ai = po.inv_sympd(A)
    assert st == 0 and np.allclose(Ai, np.linalg.inv(A), atol=1e-13)
    B = rng.standard_normal((6, 4))
    S = B @ B.T                       # rank 4
    dropped, Sp = po.pinv_sym(S)
    assert dropped == 2
    assert np.allclose(Sp, np.linalg.pinv(S, rcond=1e-10), atol=1e-10)
