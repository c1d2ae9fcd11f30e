"""The launch-shape autotuner (fcm_autotune / fcm_autotune_cached): every
shape it can pick keeps the mobility product's parity with the oracle (FP64
<= 1e-12, FP32 <= 2e-6), the choice is cached by shape class and taken by new
handles of that class, and shapes a handle cannot run are refused.

The candidates are the ones fcm_autotune times: column-path patch width 4 or 2
and z-pass points per thread 8 or 16, y-pass points 16 or 8 and (FP32) x-pass
points 16 or 8 (with the columns per workgroup each implies), and the column
path against the separate passes."""
import itertools

import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import pair_cfg, rel

pytestmark = pytest.mark.gpu

TOL = {fa.FCM_PREC_FP64: 1e-12, fa.FCM_PREC_FP32: 2e-6}


def say(*a):
    print(*a, flush=True)


def _setup(dims, nb, channel=False, seed=31):
    gc, oc = pair_cfg(*dims, ng=7, dx=1.0, a=1.86, max_blobs=nb, slip_channel=channel)
    rng = np.random.default_rng(seed)
    Y = rng.uniform(0, 1, (nb, 3)) * np.array(dims, dtype=float)[None, :]
    if channel:
        Y[:, 1] = rng.uniform(2.5, dims[1] / 2 - 2.5, nb)
    F = rng.standard_normal(3 * nb)
    o = po.Fcm(oc, nb)
    o.set_up_gaussians(Y)
    ref = o.mobility(F)
    o.close()
    return gc, Y, F, ref


@pytest.mark.parametrize("dims,nb,channel", [((64, 64, 64), 1500, False), ((128, 64, 96), 3000, False),
                                             ((64, 64, 64), 800, True), ((32, 64, 256), 1200, False)])
def test_every_tunable_shape_keeps_parity(dims, nb, channel):
    say("tuning shapes", dims, nb, "channel" if channel else "periodic")
    gc, Y, F, ref = _setup(dims, nb, channel)
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        base = s.tuning()
        assert base["col_cx"] in (2, 4)
        n_ok = 0
        for cx, cp, yp, xp in itertools.product((4, 2), (8, 16), (16, 8), (16, 8)):
            if yp == 8 and dims[1] < 64 or xp == 8 and dims[0] < 64:
                continue
            s.set_tuning(column=(1, 1), col_cx=cx, col_points=(cp, cp), y_points=(yp, yp), x_points=(0, xp))
            for prec in (fa.FCM_PREC_FP64, fa.FCM_PREC_FP32):
                V = s.mobility(F, prec)
                e = rel(V, ref)
                assert e < TOL[prec], (cx, cp, yp, xp, prec, e)
            n_ok += 1
        say("  column shapes checked:", n_ok)
        assert n_ok >= 8
        # the separate passes with the tuned FFT shapes
        for yp, xp in ((16, 16), (8, 8)):
            if yp == 8 and (dims[1] < 64 or dims[0] < 64):
                continue
            s.set_tuning(column=(0, 0), y_points=(yp, yp), x_points=(0, xp))
            for prec in (fa.FCM_PREC_FP64, fa.FCM_PREC_FP32):
                e = rel(s.mobility(F, prec), ref)
                assert e < TOL[prec], ("separate", yp, xp, prec, e)
        # whatever the autotuner picks
        for prec in (fa.FCM_PREC_FP32, fa.FCM_PREC_FP64):
            t = s.autotune(prec, reps=2)
            say("  autotuned", prec, t)
            e = rel(s.mobility(F, prec), ref)
            assert e < TOL[prec], ("autotuned", t, prec, e)


def test_tuning_cache_by_shape_class():
    say("tuning cache")
    dims, nb = (64, 64, 64), 1500
    gc, Y, F, ref = _setup(dims, nb)
    with fa.FcmSolver(gc) as a:
        a.set_up_gaussians(Y)
        t, hit = a.autotune_cached(fa.FCM_PREC_FP32, reps=2)
        assert not hit
        # a new handle of the same class takes the cached choice at blob set-up
        with fa.FcmSolver(gc) as b:
            b.set_up_gaussians(Y)
            tb = b.tuning()
            for k in ("spread_variant", "interp_izb", "column", "col_points", "y_points", "x_points"):
                assert tb[k][1] == t[k][1], (k, tb, t)
            assert tb["col_cx"] == t["col_cx"]
            t2, hit = b.autotune_cached(fa.FCM_PREC_FP32, reps=2)
            assert hit and t2 == b.tuning()
            # the same shapes on the same blobs: bitwise the same product
            assert np.array_equal(b.mobility(F, fa.FCM_PREC_FP32), a.mobility(F, fa.FCM_PREC_FP32))
        # another blob count class is not taken from the cache
        gc2, Y2, F2, ref2 = _setup(dims, 300, seed=5)
        with fa.FcmSolver(gc2) as c:
            c.set_up_gaussians(Y2)
            _, hit = c.autotune_cached(fa.FCM_PREC_FP32, reps=2)
            assert not hit
            assert rel(c.mobility(F2, fa.FCM_PREC_FP32), ref2) < TOL[fa.FCM_PREC_FP32]
    fa.clear_tuning_cache()
    with fa.FcmSolver(gc) as d:
        d.set_up_gaussians(Y)
        _, hit = d.autotune_cached(fa.FCM_PREC_FP32, reps=1)
        assert not hit


def test_shapes_a_handle_cannot_run_are_refused():
    gc, Y, F, ref = _setup((32, 32, 48), 200)
    with fa.FcmSolver(gc) as s:
        s.set_up_gaussians(Y)
        for bad in ({"col_cx": 3}, {"x_points": (16, 16)}, {"y_points": (8, 8)}, {"col_points": (12, 12)},
                    {"x_points": (0, 8)}):
            with pytest.raises(fa.FcmError):
                s.set_tuning(**bad)
        t = s.tuning()
        assert rel(s.mobility(F, fa.FCM_PREC_FP64), ref) < 1e-12 and t == s.tuning()
    gc, _ = pair_cfg(12, 12, 12, max_blobs=20)
    with fa.FcmSolver(gc) as s:
        assert s.tuning()["col_cx"] == 0
        with pytest.raises(fa.FcmError):
            s.set_tuning(col_cx=4)
