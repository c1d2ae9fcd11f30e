"""CPU-only checks of the HIP library: it loads, exports every entry point
include/fcm_hip.h declares, and its host-only functions work.  No call here
touches a GPU."""
import os
import re

import numpy as np
import pytest

import timstochrigidbodiessim_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "fcm_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    L = fa.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(fa.EXPORTS) == syms


def test_query_footprint_and_rejects_bad_configs():
    cfg = fa.make_config(256, max_blobs=240000, max_bodies=20000, max_blobs_per_body=12, max_krylov=120,
                         fp32=True, random_field=True)
    b = fa.footprint(cfg)
    assert 1.5e9 < b < 1e10
    small = fa.footprint(fa.make_config(64, max_blobs=1000, max_bodies=100, max_blobs_per_body=10))
    assert small < b
    with pytest.raises(fa.FcmError):
        fa.footprint(fa.make_config(14))          # 7 = nz/2 is not 2^a 3^b
    with pytest.raises(fa.FcmError):
        fa.footprint(fa.make_config(16, ng=20))   # stencil wider than the grid
    with pytest.raises(fa.FcmError):
        fa.footprint(fa.make_config(64, ng=21))   # wider than the reference's 20 points
    assert fa.footprint(fa.make_config(64, ng=20)) > 0  # NUM_FCM_GAUSSIAN_POINTS 20 (config.hpp:69)


def test_synth_rods_geometry_and_forces():
    L = [64.0, 64.0, 64.0]
    ref, x, q, F6, nov = fa.synth_rods(3, 100, 10, 1.86, L, max_attempts=5000)
    assert ref.shape == (10, 3) and np.allclose(ref[:, 1:], 0)
    assert np.allclose(np.diff(ref[:, 0]), 2 * 1.86)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    assert np.all((x >= 0) & (x < 64))
    assert nov == 0
    # axial torque removed
    for n in range(100):
        s, v = q[n, 0], q[n, 1:]
        e = np.array([1 - 2 * (v[1] ** 2 + v[2] ** 2), 2 * (v[0] * v[1] + s * v[2]), 2 * (v[0] * v[2] - s * v[1])])
        assert abs(np.dot(F6[n, 3:], e)) < 1e-12
    # no two blobs of different rods closer than 2a (minimum image)
    P = []
    for n in range(100):
        s, v = q[n, 0], q[n, 1:]
        R = np.array([[1 - 2 * (v[1] ** 2 + v[2] ** 2), 2 * (v[0] * v[1] - s * v[2]), 2 * (v[0] * v[2] + s * v[1])],
                      [2 * (v[0] * v[1] + s * v[2]), 1 - 2 * (v[0] ** 2 + v[2] ** 2), 2 * (v[1] * v[2] - s * v[0])],
                      [2 * (v[0] * v[2] - s * v[1]), 2 * (v[1] * v[2] + s * v[0]), 1 - 2 * (v[0] ** 2 + v[1] ** 2)]])
        P.append((x[n][None, :] + ref @ R.T) % 64.0)
    P = np.array(P)
    flat = P.reshape(-1, 3)
    body = np.repeat(np.arange(100), 10)
    d = flat[:, None, :] - flat[None, :, :]
    d -= 64.0 * np.round(d / 64.0)
    r = np.sqrt((d ** 2).sum(-1))
    mask = body[:, None] != body[None, :]
    assert r[mask].min() >= 2 * 1.86
    # same seed, same output
    ref2, x2, q2, F62, nov2 = fa.synth_rods(3, 100, 10, 1.86, L, max_attempts=5000)
    assert np.array_equal(x, x2) and np.array_equal(q, q2) and np.array_equal(F6, F62)


def test_synth_rods_overcrowded_reports_overlaps():
    ref, x, q, F6, nov = fa.synth_rods(5, 400, 12, 1.86, [32.0, 32.0, 32.0], max_attempts=5)
    assert nov > 0
