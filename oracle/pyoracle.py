"""ctypes binding for the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It loads oracle/liboracle.so (built by oracle/Makefile) and
wraps the FP64 restatement of the reference's FCM mobility path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("dx", C.c_double), ("a", C.c_double), ("mu", C.c_double),
                ("kT", C.c_double), ("dt", C.c_double), ("slip_channel", C.c_int)]


class OrBodies(C.Structure):
    _fields_ = [("NB", C.c_int), ("nbpb", C.c_int),
                ("blob_ref", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
                ("q", C.POINTER(C.c_double)), ("Minv", C.POINTER(C.c_double)),
                ("N", C.POINTER(C.c_double))]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_splitmix64_next.restype = C.c_uint64
        L.or_splitmix64_next.argtypes = [C.POINTER(C.c_uint64)]
        L.or_stress_normal.restype = C.c_double
        L.or_stress_normal.argtypes = [C.c_uint64, C.c_int, C.c_uint64, C.c_uint64]
        for fn in ("or_dft1d", "or_fft1d"):
            getattr(L, fn).argtypes = [C.c_int, _dp, _dp, C.c_int]
            getattr(L, fn).restype = None
        for fn in ("or_fft3d_r2c", "or_fft3d_c2r"):
            getattr(L, fn).argtypes = [C.POINTER(OrCfg), _dp]
            getattr(L, fn).restype = None
        L.or_fcm_create.restype = C.c_void_p
        L.or_fcm_create.argtypes = [C.POINTER(OrCfg), C.c_int]
        L.or_fcm_destroy.argtypes = [C.c_void_p]
        L.or_fcm_pad.argtypes = [C.c_void_p]
        L.or_fcm_total.argtypes = [C.c_void_p]
        L.or_fcm_total.restype = C.c_size_t
        L.or_fcm_field.argtypes = [C.c_void_p, C.c_int]
        L.or_fcm_field.restype = C.POINTER(C.c_double)
        L.or_fcm_tables.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_fcm_set_up_gaussians.argtypes = [C.c_void_p, _dp, C.c_int]
        L.or_fcm_get_gaussians.argtypes = [C.c_void_p, _dp, _dp, _dp, _ip, _ip, _ip]
        L.or_fcm_spread.argtypes = [C.c_void_p, _dp]
        L.or_fcm_stokes.argtypes = [C.c_void_p]
        L.or_fcm_interp.argtypes = [C.c_void_p, _dp]
        L.or_fcm_fill_stress.argtypes = [C.c_void_p, C.c_uint64]
        L.or_fcm_random_field.argtypes = [C.c_void_p]
        L.or_fcm_random_interp.argtypes = [C.c_void_p, _dp]
        L.or_fcm_shear_field.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.or_fcm_shear_interp.argtypes = [C.c_void_p, _dp]
        L.or_fcm_x_velocity_profile.argtypes = [C.c_void_p, _dp]
        L.or_fcm_mobility.argtypes = [C.c_void_p, _dp, _dp]
        L.or_quat_matrix.argtypes = [_dp, _dp]
        L.or_fill_blob_pos.argtypes = [C.POINTER(OrCfg), C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.or_mul_K.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.or_mul_KT.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.or_barrier_forces.argtypes = [C.POINTER(OrCfg), C.c_int, C.c_int, _dp, C.c_double, _dp]
        L.or_barrier_forces_cells.argtypes = [C.POINTER(OrCfg), C.c_int, C.c_int, _dp, C.c_int, _dp]
        L.or_lie_exp.argtypes = [_dp, _dp]
        L.or_quat_mul.argtypes = [_dp, _dp, _dp]
        L.or_end_of_step_update.argtypes = [C.POINTER(OrCfg), C.c_int, _dp, C.c_double, _dp, _dp]
        L.or_inv_sympd.argtypes = [C.c_int, _dp, _dp]
        L.or_pinv_sym.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.or_solve_dense.argtypes = [C.c_int, _dp, _dp, _dp]
        L.or_ref_matrices.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp]
        L.or_saddle_matvec.argtypes = [C.c_void_p, C.POINTER(OrBodies), _dp, _dp]
        L.or_apply_Pinv.argtypes = [C.POINTER(OrBodies), _dp, _dp]
        L.or_gmres_orth_seconds.argtypes = [C.c_size_t, C.c_int]
        L.or_gmres_orth_seconds.restype = C.c_double
        L.or_gmres.argtypes = [C.c_void_p, C.POINTER(OrBodies), _dp, _dp, C.c_double, C.c_int,
                               C.POINTER(C.c_int), _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def cfg(nx, ny=None, nz=None, ng=7, dx=1.0, a=1.86, mu=1.0, kT=0.0, dt=1.0, slip_channel=False):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return OrCfg(nx, ny, nz, ng, dx, a, mu, kT, dt, int(bool(slip_channel)))


def pad_of(c):
    return 2 * (c.nz // 2 + 1)


def stress_normal(seed, comp, node, nnodes):
    return lib().or_stress_normal(seed, comp, node, nnodes)


def splitmix64_seq(state, count):
    s = C.c_uint64(state)
    return [lib().or_splitmix64_next(C.byref(s)) for _ in range(count)]


def dft1d(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    inp = _f64(np.stack([x.real, x.imag], -1).ravel())
    out = np.zeros_like(inp)
    lib().or_dft1d(len(x), inp, out, sign)
    return out[0::2] + 1j * out[1::2]


def fft1d(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    inp = _f64(np.stack([x.real, x.imag], -1).ravel())
    out = np.zeros_like(inp)
    lib().or_fft1d(len(x), inp, out, sign)
    return out[0::2] + 1j * out[1::2]


def fft3d_r2c(c, real_padded):
    f = _f64(real_padded).copy()
    lib().or_fft3d_r2c(C.byref(c), f)
    return f


def fft3d_c2r(c, kspace_padded):
    f = _f64(kspace_padded).copy()
    lib().or_fft3d_c2r(C.byref(c), f)
    return f


def to_padded(c, real):
    """[nx][ny][nz] -> reference layout [nx][ny][pad]."""
    out = np.zeros((c.nx, c.ny, pad_of(c)))
    out[:, :, :c.nz] = real
    return out.ravel()


def from_padded(c, flat):
    return flat.reshape(c.nx, c.ny, pad_of(c))[:, :, :c.nz].copy()


def kspace_to_complex(c, flat):
    """transposed-order k-space -> complex array indexed [kx][ky][kz]."""
    a = flat.reshape(c.ny, c.nx, pad_of(c) // 2, 2)
    z = a[..., 0] + 1j * a[..., 1]
    return np.transpose(z, (1, 0, 2)).copy()


def complex_to_kspace(c, z):
    """complex [kx][ky][kz] -> transposed-order flat array."""
    zt = np.transpose(z, (1, 0, 2))
    out = np.zeros((c.ny, c.nx, pad_of(c) // 2, 2))
    out[..., 0] = zt.real
    out[..., 1] = zt.imag
    return out.ravel()


def quat_matrix(q):
    R = np.zeros(9)
    lib().or_quat_matrix(_f64(q), R)
    return R.reshape(3, 3)


def fill_blob_pos(c, blob_ref, x, q):
    NB = len(x) // 3 if np.ndim(x) == 1 else len(x)
    nbpb = np.size(blob_ref) // 3
    Y = np.zeros(3 * NB * nbpb)
    lib().or_fill_blob_pos(C.byref(c), NB, nbpb, _f64(blob_ref).ravel(), _f64(x).ravel(), _f64(q).ravel(), Y)
    return Y


def mul_K(blob_ref, q, in6):
    NB = np.size(q) // 4
    nbpb = np.size(blob_ref) // 3
    out = np.zeros(3 * NB * nbpb)
    lib().or_mul_K(NB, nbpb, _f64(blob_ref).ravel(), _f64(q).ravel(), _f64(in6).ravel(), out)
    return out


def mul_KT(blob_ref, q, in3):
    NB = np.size(q) // 4
    nbpb = np.size(blob_ref) // 3
    out = np.zeros(6 * NB)
    lib().or_mul_KT(NB, nbpb, _f64(blob_ref).ravel(), _f64(q).ravel(), _f64(in3).ravel(), out)
    return out


def barrier_forces(c, Y, nbpb, cutoff):
    """Soft-sphere barrier forces (linked_list.cpp:192-281) by brute force."""
    Y = _f64(Y).ravel()
    F = np.zeros_like(Y)
    lib().or_barrier_forces(C.byref(c), Y.size // 3, nbpb, Y, cutoff, F)
    return F


def barrier_forces_cells(c, Y, nbpb, ncell):
    """The reference's cell-list interaction set (linked_list.cpp:23-102,192-281)
    on an ncell^3 grid: neighbouring-cell pairs at any distance."""
    Y = _f64(Y).ravel()
    F = np.zeros_like(Y)
    lib().or_barrier_forces_cells(C.byref(c), Y.size // 3, nbpb, Y, ncell, F)
    return F


def lie_exp(u):
    q = np.zeros(4)
    lib().or_lie_exp(_f64(u).ravel(), q)
    return q


def quat_mul(p, q):
    out = np.zeros(4)
    lib().or_quat_mul(_f64(p).ravel(), _f64(q).ravel(), out)
    return out


def end_of_step_update(c, V6, s, x, q):
    """Returns the updated (x, q) copies (collection_of_rigid_bodies.cpp:759-790)."""
    x = _f64(x).ravel().copy()
    q = _f64(q).ravel().copy()
    lib().or_end_of_step_update(C.byref(c), x.size // 3, _f64(V6).ravel(), s, x, q)
    return x, q


def inv_sympd(A):
    A = _f64(A)
    out = np.zeros_like(A)
    st = lib().or_inv_sympd(A.shape[0], A.ravel().copy(), out.ravel())
    return st, out


def pinv_sym(A, rtol=1e-9):
    A = _f64(A)
    out = np.zeros_like(A)
    dropped = lib().or_pinv_sym(A.shape[0], A.ravel().copy(), out.ravel(), rtol)
    return dropped, out


class Fcm:
    """Wrapper of the oracle's fluid solver object."""

    FIELDS = ["ux", "uy", "uz", "fx", "fy", "fz", "sxx", "syy", "szz", "sxy", "sxz", "syz",
              "ux_rand", "uy_rand", "uz_rand", "ux_shear", "uy_shear", "uz_shear"]

    def __init__(self, c, max_blobs=1):
        self.c = c
        self.L = lib()
        self.h = self.L.or_fcm_create(C.byref(c), max_blobs)
        self.pad = pad_of(c)
        self.total = self.L.or_fcm_total(self.h)
        self.nb = 0

    def close(self):
        if self.h:
            self.L.or_fcm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        i = self.FIELDS.index(name) if isinstance(name, str) else name
        p = self.L.or_fcm_field(self.h, i)
        return np.ctypeslib.as_array(p, shape=(self.total,))

    def tables(self):
        qx = np.zeros(self.c.nx); qy = np.zeros(self.c.ny); qp = np.zeros(self.pad)
        self.L.or_fcm_tables(self.h, qx, qy, qp)
        return qx, qy, qp

    def set_up_gaussians(self, Y):
        Y = _f64(Y).ravel()
        self.nb = len(Y) // 3
        self.L.or_fcm_set_up_gaussians(self.h, Y, self.nb)

    def gaussians(self):
        n = self.nb * self.c.ng
        g = [np.zeros(n) for _ in range(3)]
        ix = [np.zeros(n, dtype=np.int32) for _ in range(3)]
        self.L.or_fcm_get_gaussians(self.h, g[0], g[1], g[2], ix[0], ix[1], ix[2])
        sh = (self.nb, self.c.ng)
        return [x.reshape(sh) for x in g], [x.reshape(sh) for x in ix]

    def spread(self, F):
        self.L.or_fcm_spread(self.h, _f64(F).ravel())

    def stokes(self):
        self.L.or_fcm_stokes(self.h)

    def interp(self):
        V = np.zeros(3 * self.nb)
        self.L.or_fcm_interp(self.h, V)
        return V

    def fill_stress(self, seed):
        self.L.or_fcm_fill_stress(self.h, seed)

    def random_field(self):
        self.L.or_fcm_random_field(self.h)

    def random_interp(self):
        V = np.zeros(3 * self.nb)
        self.L.or_fcm_random_interp(self.h, V)
        return V

    def shear_field(self, scale, kappa):
        self.L.or_fcm_shear_field(self.h, scale, kappa)

    def shear_interp(self):
        V = np.zeros(3 * self.nb)
        self.L.or_fcm_shear_interp(self.h, V)
        return V

    def x_velocity_profile(self):
        p = np.zeros(self.c.ny)
        self.L.or_fcm_x_velocity_profile(self.h, p)
        return p

    def mobility(self, F):
        V = np.zeros(3 * self.nb)
        self.L.or_fcm_mobility(self.h, _f64(F).ravel(), V)
        return V

    def ref_matrices(self, blob_ref):
        nbpb = np.size(blob_ref) // 3
        n3 = 3 * nbpb
        M = np.zeros((n3, n3)); Minv = np.zeros((n3, n3)); N = np.zeros((6, 6))
        st = self.L.or_ref_matrices(self.h, nbpb, _f64(blob_ref).ravel(), M.ravel(), Minv.ravel(), N.ravel())
        self.nb = nbpb
        return st, M, Minv, N


class Bodies:
    def __init__(self, blob_ref, x, q, Minv, N):
        self.blob_ref = _f64(blob_ref).ravel().copy()
        self.x = _f64(x).ravel().copy()
        self.q = _f64(q).ravel().copy()
        self.Minv = _f64(Minv).ravel().copy()
        self.N = _f64(N).ravel().copy()
        self.NB = len(self.q) // 4
        self.nbpb = len(self.blob_ref) // 3
        P = C.POINTER(C.c_double)
        self.s = OrBodies(self.NB, self.nbpb,
                          self.blob_ref.ctypes.data_as(P), self.x.ctypes.data_as(P),
                          self.q.ctypes.data_as(P), self.Minv.ctypes.data_as(P),
                          self.N.ctypes.data_as(P))

    @property
    def n(self):
        return 3 * self.NB * self.nbpb + 6 * self.NB


def saddle_matvec(fcm, bodies, v):
    out = np.zeros(bodies.n)
    lib().or_saddle_matvec(fcm.h, C.byref(bodies.s), _f64(v), out)
    return out


def apply_Pinv(bodies, v):
    out = np.zeros(bodies.n)
    lib().or_apply_Pinv(C.byref(bodies.s), _f64(v), out)
    return out


def gmres_orth_seconds(n, iters):
    """Seconds of one MGS step of the oracle's GMRES against `iters` vectors of length n."""
    return lib().or_gmres_orth_seconds(n, iters)


def gmres(fcm, bodies, rhs, tol=1e-8, max_iter=200):
    soln = np.zeros(bodies.n)
    iters = C.c_int(0)
    hist = np.zeros(max(1, max_iter))
    st = lib().or_gmres(fcm.h, C.byref(bodies.s), _f64(rhs), soln, tol, max_iter, C.byref(iters), hist)
    return st, soln, iters.value, hist[:iters.value].copy()
