"""GPU parity of each operator on the path against the CPU oracle, through
the C ABI (libfcm_hip.so).  Tolerances: FP64 relative L2 <= 1e-12 (operators)
and the FP32 bulk path <= 2e-6 relative L2.  Also checks bitwise
repeatability of spreading and the full mobility product."""
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


GRIDS = [(16, 16, 16), (12, 12, 12), (32, 16, 24), (64, 64, 64), (24, 48, 32)]


@pytest.mark.parametrize("dims", GRIDS)
def test_fft_r2c_matches_oracle(dims):
    say("fft r2c", dims)
    gc, oc = pair_cfg(*dims)
    rng = np.random.default_rng(1)
    r = rng.standard_normal(dims)
    with fa.FcmSolver(gc) as s:
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            s.set_field(0, po.to_padded(oc, r), prec)
            s.fft_transform(0, prec)
            got = s.get_kspace(0, prec)
            ref = po.fft3d_r2c(oc, po.to_padded(oc, r))
            assert rel(got, ref) < tol, (prec, rel(got, ref))


@pytest.mark.parametrize("dims", GRIDS)
def test_fft_c2r_matches_oracle_non_hermitian(dims):
    say("fft c2r", dims)
    gc, oc = pair_cfg(*dims)
    nx, ny, nz = dims
    rng = np.random.default_rng(2)
    Z = rng.standard_normal((nx, ny, nz // 2 + 1)) + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))
    k = po.complex_to_kspace(oc, Z)
    ref = po.from_padded(oc, po.fft3d_c2r(oc, k))
    with fa.FcmSolver(gc) as s:
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            s.set_kspace(0, k, prec)
            s.fft_inverse_transform(0, prec)
            got = po.from_padded(oc, s.get_field(0, prec))
            assert rel(got, ref) < tol, (prec, rel(got, ref))


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 24, 16), (64, 64, 64)])
def test_stokes_matches_oracle(dims):
    say("stokes", dims)
    gc, oc = pair_cfg(*dims, dx=0.8, mu=1.3)
    rng = np.random.default_rng(3)
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


def _blobs(rng, nb, dims, dx):
    return rng.uniform(0, 1, (nb, 3)) * np.array(dims)[None, :] * dx


@pytest.mark.parametrize("dims,ng,nb", [((16, 16, 16), 7, 40), ((32, 24, 16), 7, 200), ((64, 64, 64), 7, 1000),
                                        ((12, 12, 12), 5, 30), ((32, 32, 32), 12, 100), ((16, 8, 8), 7, 30),
                                        ((32, 16, 16), 8, 80), ((24, 24, 48), 5, 100), ((48, 32, 96), 7, 300)])
def test_spread_and_interp_match_oracle(dims, ng, nb):
    say("spread/interp", dims, ng, nb)
    dx = 0.9
    gc, oc = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb)
    rng = np.random.default_rng(4)
    Y = _blobs(rng, nb, dims, dx)
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
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            s.spread(F, prec)
            got = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
            assert rel(got, ref_f) < tol, ("spread", prec, rel(got, ref_f))
            s.spread(F, prec)
            again = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
            assert np.array_equal(got, again), "spreading is not bitwise repeatable"
            for c in range(3):
                s.set_field(c, po.to_padded(oc, u[c]), prec)
            V = s.interp(prec)
            assert rel(V, ref_V) < tol, ("interp", prec, rel(V, ref_V))


@pytest.mark.parametrize("dims,nb", [((32, 32, 32), 200), ((64, 64, 64), 1000)])
def test_mobility_matches_oracle_and_is_repeatable(dims, nb):
    say("mobility", dims, nb)
    gc, oc = pair_cfg(*dims, max_blobs=nb)
    rng = np.random.default_rng(5)
    Y = _blobs(rng, nb, dims, 1.0)
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    o.set_up_gaussians(Y)
    ref = o.mobility(F)
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        V = s.mobility(F)
        assert rel(V, ref) < FP64_TOL, rel(V, ref)
        assert np.array_equal(V, s.mobility(F))
        V32 = s.mobility(F, fa.FCM_PREC_FP32)
        assert rel(V32, ref) < FP32_TOL, rel(V32, ref)


@pytest.mark.parametrize("dims", [(16, 16, 16), (32, 24, 16)])
def test_random_field_injected_stress_matches_oracle(dims):
    say("random field injected", dims)
    gc, oc = pair_cfg(*dims, dx=0.7, mu=1.2, kT=1.0, dt=0.2, max_blobs=50)
    rng = np.random.default_rng(6)
    S = rng.standard_normal((6,) + dims)
    o = po.Fcm(oc, 50)
    for i, name in enumerate(("sxx", "syy", "szz", "sxy", "sxz", "syz")):
        o.field(name)[:] = po.to_padded(oc, S[i])
    o.random_field()
    ref = np.stack([po.from_padded(oc, o.field(n)) for n in ("ux_rand", "uy_rand", "uz_rand")])
    Y = _blobs(rng, 50, dims, 0.7)
    o.set_up_gaussians(Y)
    refV = o.random_interp()
    with fa.FcmSolver(gc) as s:
        s.random_field(S.reshape(6, -1))
        got = np.stack([po.from_padded(oc, s.get_field(3 + c)) for c in range(3)])
        assert rel(got, ref) < FP64_TOL, rel(got, ref)
        s.set_up_gaussians(Y)
        assert rel(s.random_interp(), refV) < FP64_TOL


def test_random_field_seeded_stream_matches_oracle():
    say("random field seeded")
    dims = (16, 16, 32)
    gc, oc = pair_cfg(*dims, kT=1.0, dt=0.20216, max_blobs=20)
    o = po.Fcm(oc, 20)
    o.fill_stress(1234)
    o.random_field()
    ref = np.stack([po.from_padded(oc, o.field(n)) for n in ("ux_rand", "uy_rand", "uz_rand")])
    with fa.FcmSolver(gc) as s:
        s.random_field(None, 1234)
        got = np.stack([po.from_padded(oc, s.get_field(3 + c)) for c in range(3)])
        assert rel(got, ref) < FP64_TOL, rel(got, ref)
        s.random_field(None, 1235)
        other = np.stack([po.from_padded(oc, s.get_field(3 + c)) for c in range(3)])
        assert rel(other, ref) > 0.5


def test_bodies_K_KT_and_positions_match_oracle():
    say("bodies")
    gc, oc = pair_cfg(32, max_blobs=6 * 5, max_bodies=6, max_blobs_per_body=5)
    rng = np.random.default_rng(7)
    ref = rng.standard_normal((5, 3)) * 2
    x = rng.uniform(0, 32, (6, 3))
    q = random_quats(rng, 6)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        assert rel(s.blob_positions(), po.fill_blob_pos(oc, ref, x, q)) < 1e-15
        v6 = rng.standard_normal(36)
        assert rel(s.mul_K(v6), po.mul_K(ref, q, v6)) < 1e-14
        v3 = rng.standard_normal(90)
        assert rel(s.mul_KT(v3), po.mul_KT(ref, q, v3)) < 1e-14


def test_ref_matrices_and_Pinv_match_oracle():
    say("ref matrices / Pinv")
    gc, oc = pair_cfg(32, max_blobs=8 * 10, max_bodies=8, max_blobs_per_body=10)
    ref = rod(10, 1.86)
    o = po.Fcm(oc, 80)
    st, M, Minv, N = o.ref_matrices(ref)
    rng = np.random.default_rng(8)
    x = rng.uniform(0, 32, (8, 3))
    q = random_quats(rng, 8)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        Mg, Minvg, Ng = s.ref_matrices()
        assert rel(Mg, M) < 1e-12, rel(Mg, M)
        s.set_ref_matrices(Minv, N)
        b = po.Bodies(ref, x, q, Minv, N)
        v = rng.standard_normal(b.n)
        assert rel(s.apply_Pinv(v), po.apply_Pinv(b, v)) < 1e-12
        o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
        assert rel(s.saddle_matvec(v), po.saddle_matvec(o, b, v)) < FP64_TOL


@pytest.mark.parametrize("nbpb,NB", [(1, 40), (2, 30), (3, 25), (5, 12), (6, 50), (10, 30), (12, 300), (13, 20),
                                     (16, 9), (21, 4), (22, 6), (24, 5), (30, 3)])
def test_Pinv_every_body_size_matches_oracle(nbpb, NB):
    """P^-1 on bodies of 1 to 30 spheres: the packed kernel (2 to 24 spheres,
    whose row threads also form the six body rows; 300 bodies of 12 take more
    than one pass of a workgroup; 6, 10 and 12 spheres have the row count
    compiled in) and the one-workgroup-per-body kernel (1 sphere, whose three
    rows could not, and more than 24)."""
    say("Pinv nbpb", nbpb, NB)
    n = 48
    gc, oc = pair_cfg(n, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb)
    # spheres 2a apart on a 3 x 2 x k lattice: compact and well conditioned
    i = np.arange(nbpb)
    ref = 2 * 1.86 * np.stack([i % 3, (i // 3) % 2, i // 6], axis=1).astype(float)
    ref -= ref.mean(axis=0)
    o = po.Fcm(oc, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    rng = np.random.default_rng(nbpb)
    x = rng.uniform(0, n, (NB, 3))
    q = random_quats(rng, NB)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        s.set_ref_matrices(Minv, N)
        b = po.Bodies(ref, x, q, Minv, N)
        v = rng.standard_normal(b.n)
        assert rel(s.apply_Pinv(v), po.apply_Pinv(b, v)) < 1e-12, rel(s.apply_Pinv(v), po.apply_Pinv(b, v))
        o.set_up_gaussians(po.fill_blob_pos(oc, ref, x, q))
        assert rel(s.saddle_matvec(v), po.saddle_matvec(o, b, v)) < FP64_TOL


@pytest.mark.parametrize("dims,ng,nb", [((32, 24, 16), 7, 200), ((64, 64, 64), 7, 1000), ((16, 8, 8), 7, 30),
                                        ((48, 32, 96), 7, 300), ((32, 16, 16), 8, 80), ((24, 24, 48), 5, 100),
                                        ((16, 16, 16), 7, 60), ((32, 48, 32), 8, 400)])
def test_spread_interp_launch_variants_match_oracle(dims, ng, nb):
    """Every tunable launch shape (spreading kernel, interpolation tile depth)
    keeps the parity bar and bitwise repeatability."""
    say("variants", dims, ng, nb)
    dx = 0.9
    gc, oc = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb)
    rng = np.random.default_rng(14)
    Y = _blobs(rng, nb, dims, dx)
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
        for sv, iz in ((0, 1), (1, 1), (1, 2), (0, 2), (0, 3), (2, 3), (2, 1)):
            s.set_tuning((sv, sv), (iz, iz))
            for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
                s.spread(F, prec)
                got = np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)])
                assert rel(got, ref_f) < tol, ("spread", sv, prec, rel(got, ref_f))
                s.spread(F, prec)
                assert np.array_equal(got, np.stack([po.from_padded(oc, s.get_field(c, prec)) for c in range(3)]))
                for c in range(3):
                    s.set_field(c, po.to_padded(oc, u[c]), prec)
                V = s.interp(prec)
                assert rel(V, ref_V) < tol, ("interp", iz, prec, rel(V, ref_V))
        t = s.autotune(fa.FCM_PREC_FP32, reps=2)
        say("autotuned", t)
        assert t["spread_variant"][1] in (0, 1, 2) and t["interp_izb"][1] in (1, 2, 3)


# the column path of the mobility product on every z length it is built for,
# dense lists (several staging rounds per patch), ng from 2 to 8, wrap-around
# stencils on small boxes and non-cubic grids
COLUMN_CASES = [((32, 32, 32), 7, 200), ((16, 16, 32), 7, 2000), ((32, 16, 48), 8, 300), ((16, 24, 96), 5, 150),
                ((24, 16, 64), 2, 50), ((32, 32, 128), 6, 400), ((16, 16, 192), 7, 200), ((16, 16, 256), 7, 300),
                ((16, 16, 512), 8, 300)]


@pytest.mark.parametrize("dims,ng,nb", COLUMN_CASES)
def test_column_mobility_matches_oracle(dims, ng, nb):
    say("column mobility", dims, ng, nb)
    dx = 0.9
    gc, oc = pair_cfg(*dims, ng=ng, dx=dx, a=1.5, max_blobs=nb)
    rng = np.random.default_rng(15)
    Y = _blobs(rng, nb, dims, dx)
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    o.set_up_gaussians(Y)
    ref = o.mobility(F)
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        t = s.tuning()
        assert t["column"] == [1, 1], t
        col = {}
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            V = s.mobility(F, prec)
            assert rel(V, ref) < tol, ("column", prec, rel(V, ref))
            assert np.array_equal(V, s.mobility(F, prec)), "column path is not bitwise repeatable"
            col[prec] = V
        s.set_tuning(t["spread_variant"], t["interp_izb"], column=(0, 0))
        for prec, tol in ((fa.FCM_PREC_FP64, FP64_TOL), (fa.FCM_PREC_FP32, FP32_TOL)):
            V = s.mobility(F, prec)
            assert rel(V, ref) < tol, ("separate", prec, rel(V, ref))
            assert rel(V, col[prec]) < tol


def test_column_path_needs_a_qualifying_grid():
    gc, _ = pair_cfg(12, 12, 12, max_blobs=20)
    with fa.FcmSolver(gc) as s:
        t = s.tuning()
        assert t["column"] == [0, 0]
        with pytest.raises(fa.FcmError):
            s.set_tuning(t["spread_variant"], t["interp_izb"], column=(1, 1))
