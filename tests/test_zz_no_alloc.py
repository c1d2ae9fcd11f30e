"""ll_barrier_forces_noalloc (the entry point that allocates nothing) against
ll_barrier_forces on the same blobs, through the C ABI: bitwise equal forces,
for the cutoff rule and for the reference's cell interaction set, on two calls
in a row (the second with fewer blobs, in the scratch the first one left),
plus the capacity and argument errors, which are decided on the host."""
import ctypes as C

import numpy as np
import pytest

import timstochrigidbodiessim_amd as fa
from timstochrigidbodiessim_amd import _lib
from _common import pair_cfg, random_quats, rod

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _forces_by_handle(n, NB, nbpb, a, cutoff, cells, channel, seed):
    gc, oc = pair_cfg(n, a=a, kT=1.3, max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                      slip_channel=channel)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, n, (NB, 3))
    if channel:
        # rods turned about y only keep every sphere at its centre's y, between the walls
        x[:, 1] = rng.uniform(a + 0.5, n / 2 - a - 0.5, NB)
        th = rng.uniform(0, 2 * np.pi, NB)
        q = np.stack([np.cos(th / 2), np.zeros(NB), np.sin(th / 2), np.zeros(NB)], axis=1)
    else:
        q = random_quats(rng, NB)
    with fa.FcmSolver(gc) as s:
        s.set_geometry(rod(nbpb, a))
        s.set_bodies(x, q)
        s.set_steric_cells(cells)
        F = s.barrier_forces(cutoff)
        Y = s.blob_positions().reshape(-1, 3)
    return Y, F


def _hip():
    """The HIP runtime libfcm_hip.so is linked against (already loaded by
    fa.lib(); torch bundles a runtime of its own, so device buffers for the
    ABI come from this one)."""
    fa.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


H2D, D2H = 1, 2


def _noalloc(Y, nbpb, n, a, cutoff, cells, channel):
    hip = _hip()
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    G = np.full_like(Y, np.nan)
    nbytes = Y.nbytes
    Yd, Fd, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(Yd), nbytes) == 0 and hip.hipMalloc(C.byref(Fd), nbytes) == 0
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        assert hip.hipMemcpy(Yd, Y.ctypes.data, nbytes, H2D) == 0
        assert hip.hipMemcpy(Fd, G.ctypes.data, nbytes, H2D) == 0  # NaN: every entry must be written
        L = np.array([float(n)] * 3)
        rc = fa.lib().ll_barrier_forces_noalloc(Yd, Y.shape[0], nbpb, _lib._ptr(L), a, 1.3, cutoff, cells,
                                                int(channel), Fd, st)
        assert rc == 0, fa.lib().fcm_last_error()
        assert hip.hipStreamSynchronize(st) == 0
        assert hip.hipMemcpy(G.ctypes.data, Fd, nbytes, D2H) == 0
    finally:
        hip.hipStreamDestroy(st)
        hip.hipFree(Yd)
        hip.hipFree(Fd)
    return G


@pytest.mark.parametrize("cells,channel", [(0, False), (6, False), (0, True)])
def test_noalloc_barrier_forces_equal_the_handle_path(cells, channel):
    a = 1.86
    cutoff = 2.2 * a if cells == 0 else 1.0
    runs = [(48, 400, 8, 11), (48, 90, 8, 12)]  # the second call has fewer blobs
    for n, NB, nbpb, seed in runs:
        say("noalloc barrier forces", n, NB, nbpb, "cells", cells, "channel", channel)
        Y, F = _forces_by_handle(n, NB, nbpb, a, cutoff, cells, channel, seed)
        G = _noalloc(Y, nbpb, n, a, cutoff, cells, channel)
        assert np.abs(F).max() > 0
        assert np.array_equal(F, G), np.abs(F - G).max()


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    say("noalloc capacity")
    L = _lib._ptr(np.array([512.0, 512.0, 512.0]))
    lib = fa.lib()
    dummy = C.c_void_p(16)  # never dereferenced: the checks fail on the host first
    # 2 nb + 3 cells + 1 ints beyond the static array
    big = _lib.FCM_NOALLOC_SCRATCH_BYTES // 8 + 1
    rc = lib.ll_barrier_forces_noalloc(dummy, big, 12, L, 1.86, 1.0, 4.092, 0, 0, dummy, None)
    assert rc == _lib.FCM_ERR_CAPACITY and b"static scratch" in lib.fcm_last_error()
    # 1024^3 cells alone exceed it
    rc = lib.ll_barrier_forces_noalloc(dummy, 10, 1, L, 1.86, 1.0, 0.1, 0, 0, dummy, None)
    assert rc == _lib.FCM_ERR_CAPACITY
    rc = lib.ll_barrier_forces_noalloc(dummy, 10, 1, L, 1.86, 1.0, 1.0, 2, 0, dummy, None)
    assert rc == _lib.FCM_ERR_ARG
    rc = lib.ll_barrier_forces_noalloc(None, 10, 1, L, 1.86, 1.0, 1.0, 0, 0, dummy, None)
    assert rc == _lib.FCM_ERR_ARG
