"""ctypes binding of libfcm_hip.so (the C ABI declared in include/fcm_hip.h).

The library is built in-tree by ``python -m timstochrigidbodiessim_amd.build``.
Nothing here falls back to another implementation: if the library is missing
or a call fails, an exception is raised.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfcm_hip.so")

FCM_PREC_FP64 = 0
FCM_PREC_FP32 = 1
FCM_ERR_ARG = 1
FCM_ERR_CAPACITY = 6
FCM_NOALLOC_SCRATCH_BYTES = 48 << 20  # include/fcm_hip.h

GMRES_STATUS = {0: "converged", 1: "max_iter", 2: "zero_rhs", 3: "breakdown"}

# every symbol include/fcm_hip.h declares
EXPORTS = [
    "fcm_create", "fcm_destroy", "fcm_last_error", "fcm_sync", "fcm_query_footprint", "fcm_device_memory",
    "fcm_set_up_gaussians", "fcm_spread", "fcm_stokes", "fcm_interp", "fcm_mobility",
    "fcm_random_field", "fcm_random_interp", "fcm_get_field", "fcm_set_field",
    "fcm_ref_matrices", "fcm_set_ref_matrices",
    "fft_transform", "fft_inverse_transform", "fcm_get_kspace", "fcm_set_kspace",
    "rb_set_geometry", "rb_set_bodies", "rb_set_body_forces", "rb_get_blob_positions",
    "rb_mul_K", "rb_mul_KT", "ll_barrier_forces", "ll_set_cells", "rb_end_of_step_update", "rb_get_bodies",
    "fcm_shear_field", "fcm_shear_interp", "fcm_x_velocity_profile",
    "rb_channel_wall_overlap", "rb_reset_to_start_of_step", "fcm_em_step", "fcm_gdc_step",
    "rb_advance_to_midstep", "rb_calculate_divU",
    "gmres_assemble_rhs", "gmres_saddle_matvec", "gmres_apply_Pinv", "gmres_solve",
    "gmres_get_solution", "fcm_mobility_solve", "fcm_mobility_prepare",
    "fcm_batch_create", "fcm_batch_destroy", "fcm_batch_submit", "fcm_batch_step", "fcm_batch_take",
    "fcm_batch_run",
    "fcm_get_tuning", "fcm_set_tuning", "fcm_autotune", "fcm_autotune_cached", "fcm_clear_tuning_cache",
    "fcm_profile_matvec", "fcm_profile_copy", "ll_barrier_forces_noalloc", "fcm_synth_rods",
]

PASS_NAMES = ["spread", "z_forward", "y_forward", "x_project", "y_inverse", "z_inverse", "interp", "KT",
              "interp_reduce"]


TUNING_PAIRS = ("spread_variant", "interp_izb", "column", "col_points", "y_points", "x_points")


class FcmTuning(C.Structure):
    _fields_ = [(k, C.c_int * 2) for k in TUNING_PAIRS] + [("col_cx", C.c_int)]

    def as_dict(self):
        d = {k: list(getattr(self, k)) for k in TUNING_PAIRS}
        d["col_cx"] = int(self.col_cx)
        return d

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k in TUNING_PAIRS:
            setattr(t, k, (C.c_int * 2)(*d[k]))
        t.col_cx = int(d["col_cx"])
        return t


class FcmConfig(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("dx", C.c_double), ("a", C.c_double), ("mu", C.c_double), ("kT", C.c_double),
                ("dt", C.c_double), ("max_blobs", C.c_int), ("max_bodies", C.c_int),
                ("max_blobs_per_body", C.c_int), ("max_krylov", C.c_int), ("fp32", C.c_int),
                ("random_field", C.c_int), ("slip_channel", C.c_int)]


class GmresOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int),
                ("inner_tol", C.c_double), ("max_refine", C.c_int)]


class GmresReport(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("refinements", C.c_int),
                ("rel_residual", C.c_double), ("true_rel_residual", C.c_double),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {"status": GMRES_STATUS.get(self.status, self.status), "iterations": self.iterations,
                "refinements": self.refinements, "rel_residual": self.rel_residual,
                "true_rel_residual": self.true_rel_residual, "seconds": self.seconds}


_dp = C.POINTER(C.c_double)
_lib = None


class FcmError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FcmError(f"{LIB_PATH} is not built; run python -m timstochrigidbodiessim_amd.build")
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        L.fcm_last_error.restype = C.c_char_p
        L.fcm_create.argtypes = [C.POINTER(FcmConfig), C.POINTER(H)]
        L.fcm_destroy.argtypes = [H]
        L.fcm_sync.argtypes = [H]
        L.fcm_query_footprint.argtypes = [C.POINTER(FcmConfig), C.POINTER(C.c_size_t)]
        L.fcm_device_memory.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.fcm_set_up_gaussians.argtypes = [H, _dp, C.c_int]
        L.fcm_spread.argtypes = [H, _dp, C.c_int]
        L.fcm_stokes.argtypes = [H, C.c_int]
        L.fcm_interp.argtypes = [H, _dp, C.c_int]
        L.fcm_mobility.argtypes = [H, _dp, _dp, C.c_int]
        L.fcm_random_field.argtypes = [H, _dp, C.c_uint64]
        L.fcm_random_interp.argtypes = [H, _dp]
        L.fcm_get_field.argtypes = [H, C.c_int, C.c_int, _dp]
        L.fcm_set_field.argtypes = [H, C.c_int, C.c_int, _dp]
        L.fcm_ref_matrices.argtypes = [H, _dp, _dp, _dp]
        L.fcm_set_ref_matrices.argtypes = [H, _dp, _dp]
        L.fft_transform.argtypes = [H, C.c_int, C.c_int]
        L.fft_inverse_transform.argtypes = [H, C.c_int, C.c_int]
        L.fcm_get_kspace.argtypes = [H, C.c_int, C.c_int, _dp]
        L.fcm_set_kspace.argtypes = [H, C.c_int, C.c_int, _dp]
        L.rb_set_geometry.argtypes = [H, C.c_int, _dp]
        L.rb_set_bodies.argtypes = [H, C.c_int, _dp, _dp]
        L.rb_set_body_forces.argtypes = [H, _dp]
        L.rb_get_blob_positions.argtypes = [H, _dp]
        L.rb_mul_K.argtypes = [H, _dp, _dp]
        L.rb_mul_KT.argtypes = [H, _dp, _dp]
        L.ll_barrier_forces.argtypes = [H, C.c_double, _dp]
        L.ll_set_cells.argtypes = [H, C.c_int]
        L.ll_barrier_forces_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, C.c_double, C.c_double,
                                                C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.rb_end_of_step_update.argtypes = [H, _dp, C.c_double]
        L.rb_get_bodies.argtypes = [H, _dp, _dp]
        L.rb_channel_wall_overlap.argtypes = [H, C.POINTER(C.c_int)]
        L.fcm_shear_field.argtypes = [H, C.c_double, C.c_double]
        L.fcm_shear_interp.argtypes = [H, _dp]
        L.fcm_x_velocity_profile.argtypes = [H, _dp, _dp]
        L.rb_reset_to_start_of_step.argtypes = [H]
        L.fcm_em_step.argtypes = [H, C.c_uint64, C.c_double, C.POINTER(GmresOpts), C.POINTER(GmresReport)]
        L.fcm_gdc_step.argtypes = [H, C.c_uint64, C.c_double, C.c_double, C.POINTER(GmresOpts),
                                   C.POINTER(GmresReport), C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.rb_advance_to_midstep.argtypes = [H, _dp, C.POINTER(C.c_int)]
        L.rb_calculate_divU.argtypes = [H, C.c_double, C.POINTER(C.c_double)]
        L.gmres_assemble_rhs.argtypes = [H, C.c_int]
        L.gmres_saddle_matvec.argtypes = [H, _dp, _dp, C.c_int]
        L.gmres_apply_Pinv.argtypes = [H, _dp, _dp]
        L.gmres_solve.argtypes = [H, _dp, _dp, C.POINTER(GmresOpts), C.POINTER(GmresReport), _dp]
        L.gmres_get_solution.argtypes = [H, _dp]
        L.fcm_mobility_solve.argtypes = [H, C.c_uint64, C.POINTER(GmresOpts), C.POINTER(GmresReport)]
        L.fcm_mobility_prepare.argtypes = [H, C.c_uint64]
        L.fcm_batch_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(H)]
        L.fcm_batch_destroy.argtypes = [H]
        L.fcm_batch_submit.argtypes = [H, H, C.POINTER(GmresOpts)]
        L.fcm_batch_step.argtypes = [H, C.POINTER(C.c_int)]
        L.fcm_batch_take.argtypes = [H, C.POINTER(H), C.POINTER(GmresReport), C.c_int, C.POINTER(C.c_int)]
        L.fcm_batch_run.argtypes = [H]
        L.fcm_get_tuning.argtypes = [H, C.POINTER(FcmTuning)]
        L.fcm_set_tuning.argtypes = [H, C.POINTER(FcmTuning)]
        L.fcm_autotune.argtypes = [H, C.c_int, C.c_int, C.POINTER(FcmTuning)]
        L.fcm_clear_tuning_cache.argtypes = []
        L.fcm_autotune_cached.argtypes = [H, C.c_int, C.c_int, C.POINTER(FcmTuning), C.POINTER(C.c_int)]
        L.fcm_profile_matvec.argtypes = [H, C.c_int, C.c_int, _dp]
        L.fcm_profile_copy.argtypes = [H, C.c_size_t, C.c_int, _dp]
        L.fcm_synth_rods.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_double, _dp, C.c_int, _dp, _dp,
                                     _dp, _dp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise FcmError(f"libfcm_hip error {rc}: {lib().fcm_last_error().decode()}")


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def synth_rods(seed, NB, nbpb, a, L, max_attempts=50, forces=True):
    """Seeded rods (see fcm_synth_rods in include/fcm_hip.h).  Returns
    blob_ref (nbpb x 3), x (NB x 3), q (NB x 4), F6 (NB x 6) and the number of
    rods that could not be placed without overlap."""
    ref = np.zeros(3 * nbpb)
    x = np.zeros(3 * NB)
    q = np.zeros(4 * NB)
    F6 = np.zeros(6 * NB) if forces else None
    nov = C.c_int(0)
    Lh = _f64(L)
    _chk(lib().fcm_synth_rods(seed, NB, nbpb, a, _ptr(Lh), max_attempts, _ptr(ref), _ptr(x), _ptr(q),
                              _ptr(F6), C.byref(nov)))
    return ref.reshape(nbpb, 3), x.reshape(NB, 3), q.reshape(NB, 4), (F6.reshape(NB, 6) if forces else None), nov.value


def clear_tuning_cache():
    """Empty the library's shape-class cache of tuned launch shapes."""
    _chk(lib().fcm_clear_tuning_cache())


def device_memory():
    """(free, total) bytes of device memory."""
    f, t = C.c_size_t(), C.c_size_t()
    _chk(lib().fcm_device_memory(C.byref(f), C.byref(t)))
    return f.value, t.value


def make_config(nx, ny=None, nz=None, ng=7, dx=1.0, a=1.86, mu=1.0, kT=0.0, dt=1.0, max_blobs=1,
                max_bodies=1, max_blobs_per_body=1, max_krylov=200, fp32=False, random_field=False,
                slip_channel=False):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return FcmConfig(nx, ny, nz, ng, dx, a, mu, kT, dt, max_blobs, max_bodies, max_blobs_per_body,
                     max_krylov, int(fp32), int(random_field), int(bool(slip_channel)))


def footprint(cfg):
    b = C.c_size_t(0)
    _chk(lib().fcm_query_footprint(C.byref(cfg), C.byref(b)))
    return b.value


class FcmSolver:
    """One grid + blobs + bodies + GMRES basis on the device (an fcm_handle)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        h = C.c_void_p()
        _chk(self.L.fcm_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.nb = 0
        self.NB = 0
        self.nbpb = 0
        self.pad = 2 * (cfg.nz // 2 + 1)

    def close(self):
        if self.h:
            self.L.fcm_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self):
        return 3 * self.NB * self.nbpb + 6 * self.NB

    # fcm_fluid_solver
    def set_up_gaussians(self, Y):
        Y = _f64(Y).ravel()
        self.nb = len(Y) // 3
        _chk(self.L.fcm_set_up_gaussians(self.h, _ptr(Y), self.nb))

    def spread(self, F, precision=FCM_PREC_FP64):
        _chk(self.L.fcm_spread(self.h, _ptr(_f64(F).ravel()), precision))

    def stokes(self, precision=FCM_PREC_FP64):
        _chk(self.L.fcm_stokes(self.h, precision))

    def interp(self, precision=FCM_PREC_FP64):
        V = np.zeros(3 * self.nb)
        _chk(self.L.fcm_interp(self.h, _ptr(V), precision))
        return V

    def mobility(self, F, precision=FCM_PREC_FP64):
        V = np.zeros(3 * self.nb)
        _chk(self.L.fcm_mobility(self.h, _ptr(_f64(F).ravel()), _ptr(V), precision))
        return V

    def random_field(self, stress6=None, seed=0):
        s = None if stress6 is None else _f64(stress6).ravel()
        _chk(self.L.fcm_random_field(self.h, _ptr(s), seed))

    def random_interp(self):
        V = np.zeros(3 * self.nb)
        _chk(self.L.fcm_random_interp(self.h, _ptr(V)))
        return V

    def get_field(self, which, precision=FCM_PREC_FP64):
        c = self.cfg
        out = np.zeros(c.nx * c.ny * self.pad)
        _chk(self.L.fcm_get_field(self.h, which, precision, _ptr(out)))
        return out

    def set_field(self, which, arr, precision=FCM_PREC_FP64):
        _chk(self.L.fcm_set_field(self.h, which, precision, _ptr(_f64(arr).ravel())))

    def get_kspace(self, which, precision=FCM_PREC_FP64):
        c = self.cfg
        out = np.zeros(c.nx * c.ny * self.pad)
        _chk(self.L.fcm_get_kspace(self.h, which, precision, _ptr(out)))
        return out

    def set_kspace(self, which, arr, precision=FCM_PREC_FP64):
        _chk(self.L.fcm_set_kspace(self.h, which, precision, _ptr(_f64(arr).ravel())))

    def fft_transform(self, which, precision=FCM_PREC_FP64):
        _chk(self.L.fft_transform(self.h, which, precision))

    def fft_inverse_transform(self, which, precision=FCM_PREC_FP64):
        _chk(self.L.fft_inverse_transform(self.h, which, precision))

    def ref_matrices(self):
        n3 = 3 * self.nbpb
        M = np.zeros((n3, n3)); Minv = np.zeros((n3, n3)); N = np.zeros((6, 6))
        _chk(self.L.fcm_ref_matrices(self.h, _ptr(M), _ptr(Minv), _ptr(N)))
        return M, Minv, N

    def set_ref_matrices(self, Minv, N):
        _chk(self.L.fcm_set_ref_matrices(self.h, _ptr(_f64(Minv).ravel().copy()), _ptr(_f64(N).ravel().copy())))

    # collection_of_rigid_bodies
    def set_geometry(self, blob_ref):
        r = _f64(blob_ref).ravel()
        self.nbpb = len(r) // 3
        _chk(self.L.rb_set_geometry(self.h, self.nbpb, _ptr(r)))

    def set_bodies(self, x, q):
        x = _f64(x).ravel(); q = _f64(q).ravel()
        self.NB = len(q) // 4
        _chk(self.L.rb_set_bodies(self.h, self.NB, _ptr(x), _ptr(q)))
        self.nb = self.NB * self.nbpb

    def set_body_forces(self, F6):
        _chk(self.L.rb_set_body_forces(self.h, _ptr(_f64(F6).ravel())))

    def blob_positions(self):
        Y = np.zeros(3 * self.nb)
        _chk(self.L.rb_get_blob_positions(self.h, _ptr(Y)))
        return Y

    def mul_K(self, v6):
        out = np.zeros(3 * self.nb)
        _chk(self.L.rb_mul_K(self.h, _ptr(_f64(v6).ravel()), _ptr(out)))
        return out

    def mul_KT(self, v3):
        out = np.zeros(6 * self.NB)
        _chk(self.L.rb_mul_KT(self.h, _ptr(_f64(v3).ravel()), _ptr(out)))
        return out

    # gmres_solver
    def barrier_forces(self, cutoff):
        """Soft-sphere barrier forces (3 Nb), kept on the device for the next RHS."""
        F = np.zeros(3 * self.NB * self.nbpb)
        _chk(self.L.ll_barrier_forces(self.h, cutoff, _ptr(F)))
        return F.reshape(-1, 3)

    def set_steric_cells(self, cells_per_dim):
        """The reference's interaction set on a cells_per_dim^3 grid (no cutoff), or 0 for the cutoff rule."""
        _chk(self.L.ll_set_cells(self.h, int(cells_per_dim)))

    def end_of_step_update(self, V6=None, nu=1.0):
        """Move the bodies by nu*dt*V6 (host [U; Omega] per body, or the device solution)."""
        _chk(self.L.rb_end_of_step_update(self.h, None if V6 is None else _ptr(_f64(V6).ravel()), nu))

    def shear_field(self, scale, wavenumber):
        """Set the steady shear flow (scale 0: off).  The RHS subtracts the blob
        velocities stored by shear_interp() (or a drifter-corrector mid-step)."""
        _chk(self.L.fcm_shear_field(self.h, scale, wavenumber))

    def shear_interp(self):
        """calculate_shear_blob_velocities: computes, stores (for the RHS) and returns Vshear."""
        V = np.zeros(3 * self.nb)
        _chk(self.L.fcm_shear_interp(self.h, _ptr(V)))
        return V

    def x_velocity_profile(self, lam=None):
        """Mean x velocity per y row of the flow of blob forces lam (None: the last solution) plus the shear flow."""
        p = np.zeros(self.cfg.ny)
        _chk(self.L.fcm_x_velocity_profile(self.h, None if lam is None else _ptr(_f64(lam).ravel()), _ptr(p)))
        return p

    def channel_wall_overlap(self):
        """True if a blob is on or beyond a slip-channel wall (always False in a periodic box)."""
        v = C.c_int(0)
        _chk(self.L.rb_channel_wall_overlap(self.h, C.byref(v)))
        return bool(v.value)

    def reset_to_start_of_step(self):
        """Undo the last end-of-step update (bodies, blob positions, Gaussians)."""
        _chk(self.L.rb_reset_to_start_of_step(self.h))

    def bodies(self):
        x = np.zeros(3 * self.NB)
        q = np.zeros(4 * self.NB)
        _chk(self.L.rb_get_bodies(self.h, _ptr(x), _ptr(q)))
        return x.reshape(-1, 3), q.reshape(-1, 4)

    def em_step(self, seed, cutoff, tol=1e-8, max_iter=200, precision=FCM_PREC_FP64, inner_tol=0.0, max_refine=0):
        """One Euler-Maruyama step on the device; returns the solve report."""
        o = GmresOpts(tol, max_iter, precision, inner_tol, max_refine)
        r = GmresReport()
        _chk(self.L.fcm_em_step(self.h, seed, cutoff, C.byref(o), C.byref(r)))
        return r.as_dict()

    def gdc_step(self, seed, cutoff, fd_delta=1e-5, tol=1e-8, max_iter=200, precision=FCM_PREC_FP64, inner_tol=0.0,
                 max_refine=0):
        """One generalised drifter-corrector step on the device; returns the solve
        report with divU and redo (a mid-step wall overlap: nothing moved)."""
        o = GmresOpts(tol, max_iter, precision, inner_tol, max_refine)
        r = GmresReport()
        d = C.c_double(0.0)
        redo = C.c_int(0)
        _chk(self.L.fcm_gdc_step(self.h, seed, cutoff, fd_delta, C.byref(o), C.byref(r), C.byref(d), C.byref(redo)))
        out = r.as_dict()
        out["divU"] = d.value
        out["redo"] = bool(redo.value)
        return out

    def assemble_rhs(self, use_random=True):
        _chk(self.L.gmres_assemble_rhs(self.h, int(use_random)))

    def saddle_matvec(self, v, precision=FCM_PREC_FP64):
        out = np.zeros(self.n)
        _chk(self.L.gmres_saddle_matvec(self.h, _ptr(_f64(v).ravel()), _ptr(out), precision))
        return out

    def apply_Pinv(self, v):
        out = np.zeros(self.n)
        _chk(self.L.gmres_apply_Pinv(self.h, _ptr(_f64(v).ravel()), _ptr(out)))
        return out

    def solve(self, rhs=None, tol=1e-8, max_iter=200, precision=FCM_PREC_FP64, inner_tol=0.0, max_refine=0,
              want_solution=True, hist=False):
        o = GmresOpts(tol, max_iter, precision, inner_tol, max_refine)
        r = GmresReport()
        soln = np.zeros(self.n) if want_solution else None
        hh = np.zeros(max(1, max_iter)) if hist else None
        rr = None if rhs is None else _f64(rhs).ravel()
        _chk(self.L.gmres_solve(self.h, _ptr(rr), _ptr(soln), C.byref(o), C.byref(r), _ptr(hh)))
        d = r.as_dict()
        if hist:
            d["hist"] = hh[:min(r.iterations, self.cfg.max_krylov + 2)].copy()
        return soln, d

    def mobility_solve(self, seed, tol=1e-8, max_iter=200, precision=FCM_PREC_FP64, inner_tol=0.0, max_refine=0):
        o = GmresOpts(tol, max_iter, precision, inner_tol, max_refine)
        r = GmresReport()
        _chk(self.L.fcm_mobility_solve(self.h, seed, C.byref(o), C.byref(r)))
        return r.as_dict()

    def tuning(self):
        """The handle's launch shapes: per precision [FP64, FP32] spread_variant,
        interp_izb, column, col_points, y_points, x_points, and col_cx."""
        t = FcmTuning()
        _chk(self.L.fcm_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, spread_variant=None, interp_izb=None, column=None, **shapes):
        """Set launch shapes; fields not given keep their current values."""
        d = self.tuning()
        for k, v in (("spread_variant", spread_variant), ("interp_izb", interp_izb), ("column", column)):
            if v is not None:
                d[k] = list(v)
        for k, v in shapes.items():
            if k not in d:
                raise KeyError(k)
            d[k] = int(v) if k == "col_cx" else list(v)
        _chk(self.L.fcm_set_tuning(self.h, C.byref(FcmTuning.from_dict(d))))

    def autotune(self, precision=FCM_PREC_FP32, reps=5):
        t = FcmTuning()
        _chk(self.L.fcm_autotune(self.h, precision, reps, C.byref(t)))
        return t.as_dict()

    def autotune_cached(self, precision=FCM_PREC_FP32, reps=5):
        """(tuning, from_cache): the shape class's cached tuning, or a fresh autotune."""
        t = FcmTuning()
        hit = C.c_int(0)
        _chk(self.L.fcm_autotune_cached(self.h, precision, reps, C.byref(t), C.byref(hit)))
        return t.as_dict(), bool(hit.value)

    def mobility_prepare(self, seed):
        """Seeded fluctuating field + RHS on the device, not waited for."""
        _chk(self.L.fcm_mobility_prepare(self.h, seed))

    def profile_matvec(self, precision=FCM_PREC_FP64, reps=10):
        ms = np.zeros(len(PASS_NAMES))
        _chk(self.L.fcm_profile_matvec(self.h, precision, reps, _ptr(ms)))
        return dict(zip(PASS_NAMES, ms.tolist()))

    def profile_copy(self, nbytes, reps=5):
        """Mean device ms of a plain copy of nbytes (the byte-counter reference)."""
        ms = np.zeros(1)
        _chk(self.L.fcm_profile_copy(self.h, nbytes, reps, _ptr(ms)))
        return float(ms[0])

    def solution(self):
        s = np.zeros(self.n)
        _chk(self.L.gmres_get_solution(self.h, _ptr(s)))
        return s


class FcmBatch:
    """Many independent solves advanced together (fcm_batch_* in the C ABI).

    submit() queues the solve of a solver's assembled RHS; step() launches one
    round and returns the number of unfinished jobs; take() returns finished
    (solver, report) pairs in finishing order."""

    def __init__(self, max_jobs=64, worker_streams=4, steps_per_round=4):
        self.L = lib()
        self.b = C.c_void_p()
        _chk(self.L.fcm_batch_create(max_jobs, worker_streams, steps_per_round, C.byref(self.b)))
        self.max_jobs = max_jobs
        self._owner = {}

    def close(self):
        if self.b:
            self.L.fcm_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, solver, tol=1e-8, max_iter=200, precision=FCM_PREC_FP64, inner_tol=0.0, max_refine=0):
        o = GmresOpts(tol, max_iter, precision, inner_tol, max_refine)
        _chk(self.L.fcm_batch_submit(self.b, solver.h, C.byref(o)))
        self._owner[solver.h.value] = solver

    def step(self):
        n = C.c_int()
        _chk(self.L.fcm_batch_step(self.b, C.byref(n)))
        return n.value

    def take(self):
        hs = (C.c_void_p * self.max_jobs)()
        rs = (GmresReport * self.max_jobs)()
        n = C.c_int()
        _chk(self.L.fcm_batch_take(self.b, hs, rs, self.max_jobs, C.byref(n)))
        out = []
        for i in range(n.value):
            out.append((self._owner.pop(hs[i], None), rs[i].as_dict()))
        return out

    def run(self):
        _chk(self.L.fcm_batch_run(self.b))
