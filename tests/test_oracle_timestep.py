"""Oracle pins for the time-step pieces (config 5): soft-sphere barrier forces
(linked_list.cpp:192-281,362-382), the exponential map (quaternion.cpp:128-145),
the quaternion product and the end-of-step body update
(collection_of_rigid_bodies.cpp:759-790, particle.cpp:45-56).  Checked against
closed-form numbers that follow from the reference's constants (U0 = 20 kT,
b = a/2), Newton's third law, and scipy's independent rotation code."""
import numpy as np
from scipy.spatial.transform import Rotation

import pyoracle as po


def _cfg(n=32, a=1.0, kT=1.0):
    return po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=0.1)


def _pair(c, y0, y1, cutoff, nbpb=1):
    Y = np.array([y0, y1], dtype=float)
    return po.barrier_forces(c, Y, nbpb, cutoff).reshape(2, 3)


def test_barrier_force_known_answers():
    c = _cfg(a=1.0, kT=1.0)  # U0 = 20, b = 0.5: |F| = 40 inside 2a, 40 exp((2 - r)/0.5) outside
    F = _pair(c, [10.0, 10.0, 10.0], [12.5, 10.0, 10.0], cutoff=3.0)
    assert np.allclose(F[0], [-14.715177646857693, 0, 0], rtol=1e-14, atol=0)
    assert np.array_equal(F[1], -F[0])
    F = _pair(c, [10.0, 10.0, 10.0], [10.0, 11.0, 10.0], cutoff=3.0)  # overlapping: constant 40
    assert np.allclose(F[0], [0, -40.0, 0], rtol=1e-15)
    F = _pair(c, [10.0, 10.0, 10.0], [10.0, 10.0, 13.5], cutoff=3.0)  # beyond the cutoff
    assert not F.any()
    c2 = _cfg(a=1.5, kT=2.0)  # U0 = 40, b = 0.75: at r = 3.3, 40/0.75 exp(-0.4)
    F = _pair(c2, [5.0, 5.0, 5.0], [5.0, 5.0, 8.3], cutoff=4.0)
    assert np.allclose(F[0], [0, 0, -(40 / 0.75) * np.exp(-0.3 / 0.75)], rtol=1e-13)


def test_barrier_force_uses_nearest_image_and_skips_own_body():
    c = _cfg(n=32, a=1.0)
    F = _pair(c, [0.5, 7.0, 7.0], [31.0, 7.0, 7.0], cutoff=3.0)  # separation 1.5 through the x face
    assert np.allclose(F[0], [40.0, 0, 0], rtol=1e-15)
    F = _pair(c, [10.0, 10.0, 10.0], [11.0, 10.0, 10.0], cutoff=3.0, nbpb=2)  # one body of two blobs
    assert not F.any()


def test_barrier_forces_sum_to_zero_and_are_translation_invariant():
    c = _cfg(n=24, a=1.2)
    rng = np.random.default_rng(0)
    Y = rng.uniform(0, 24, (60, 3))
    F = po.barrier_forces(c, Y, 3, 2.64).reshape(-1, 3)
    assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).max()
    assert np.abs(F).max() > 0
    G = po.barrier_forces(c, np.mod(Y + np.array([3.0, 5.0, 7.0]), 24.0), 3, 2.64).reshape(-1, 3)
    assert np.allclose(F, G, rtol=1e-10, atol=1e-12)


def test_lie_exp_matches_scipy_rotvec():
    rng = np.random.default_rng(1)
    for _ in range(5):
        u = rng.standard_normal(3)
        q = po.lie_exp(u)
        assert abs(np.linalg.norm(q) - 1) < 1e-15
        # scipy quaternions are (x, y, z, w)
        ref = Rotation.from_rotvec(u).as_quat()
        assert np.allclose(np.r_[ref[3], ref[:3]] * np.sign(ref[3] * q[0]), q, atol=1e-15)
    assert np.array_equal(po.lie_exp([1e-11, 0, 0]), [1.0, 0, 0, 0])


def test_quaternion_product_is_hamilton():
    i, j, k = np.array([0, 1.0, 0, 0]), np.array([0, 0, 1.0, 0]), np.array([0, 0, 0, 1.0])
    assert np.array_equal(po.quat_mul(i, j), k)
    assert np.array_equal(po.quat_mul(j, i), -k)
    rng = np.random.default_rng(2)
    p, q = rng.standard_normal(4), rng.standard_normal(4)
    assert np.isclose(np.linalg.norm(po.quat_mul(p, q)), np.linalg.norm(p) * np.linalg.norm(q), rtol=1e-14)


def test_end_of_step_update_matches_scipy_composition():
    c = _cfg(n=16)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 16, (4, 3))
    q = rng.standard_normal((4, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    V6 = rng.standard_normal((4, 6)) * 20
    s = 0.37
    x1, q1 = po.end_of_step_update(c, V6, s, x, q)
    assert np.allclose(x1.reshape(-1, 3), np.mod(x + s * V6[:, :3], 16.0), atol=1e-13)
    for n in range(4):
        R0 = Rotation.from_quat(np.r_[q[n, 1:], q[n, 0]]).as_matrix()
        R1 = Rotation.from_quat(np.r_[q1[4 * n + 1:4 * n + 4], q1[4 * n]]).as_matrix()
        assert np.allclose(R1, Rotation.from_rotvec(s * V6[n, 3:]).as_matrix() @ R0, atol=1e-13)


def _adjacent(ci, cj, n):
    """Cells of an n^3 periodic grid that are the same or neighbours."""
    d = np.abs(ci - cj)
    d = np.minimum(d, n - d)
    return np.all(d <= 1, axis=-1)


def test_reference_cell_traversal_covers_every_pair_once_on_three_cells():
    """On a 3^3 cell grid every cell neighbours every other one, so the
    reference's traversal (each cell against itself and its 13 half-shell
    cells, linked_list.cpp:192-281) must give every pair exactly once: the
    all-pairs brute force with no cutoff."""
    c = _cfg(n=24, a=1.0)
    rng = np.random.default_rng(3)
    Y = rng.uniform(0, 24, (90, 3))
    F = po.barrier_forces_cells(c, Y, 3, 3).reshape(-1, 3)
    G = po.barrier_forces(c, Y, 3, 1e9).reshape(-1, 3)
    assert np.abs(G).max() > 0
    assert np.allclose(F, G, rtol=1e-12, atol=1e-14 * np.abs(G).max())


def test_reference_cell_traversal_is_the_sum_over_neighbouring_cell_pairs():
    """For n >= 4 the interaction set is the pairs of blobs of different bodies
    whose cells floor(n Y/L) are the same or neighbours (periodically), at any
    distance: the result is the superposition of the two-blob forces over
    exactly those pairs.  Channel wall springs come on top."""
    for n_cells, channel, seed in ((4, False, 5), (6, False, 6), (5, True, 7)):
        c = po.cfg(20, ng=7, dx=1.0, a=1.0, mu=1.0, kT=1.3, dt=0.1, slip_channel=channel)
        rng = np.random.default_rng(seed)
        nbpb = 2
        Y = rng.uniform(0, 20, (48, 3))
        if channel:
            Y[:, 1] = rng.uniform(0.5, 9.5, 48)
        F = po.barrier_forces_cells(c, Y, nbpb, n_cells).reshape(-1, 3)
        cells = np.floor(n_cells * (Y / 20.0)).astype(int)
        # wall springs alone: each blob as its own body, far from every other
        W = np.stack([po.barrier_forces(c, Y[i:i + 1], 1, 1e-9).reshape(3) for i in range(len(Y))])
        S = W.copy()
        c0 = po.cfg(20, ng=7, dx=1.0, a=1.0, mu=1.0, kT=1.3, dt=0.1)
        far = 0
        for i in range(len(Y)):
            for j in range(len(Y)):
                if i // nbpb == j // nbpb or not _adjacent(cells[i], cells[j], n_cells):
                    continue
                S[i] += po.barrier_forces(c0, Y[[i, j]], 1, 1e9).reshape(2, 3)[0]
                d = Y[i] - Y[j]
                d -= 20.0 * np.round(d / 20.0)
                far += np.linalg.norm(d) > 20.0 / n_cells
        assert far > 0, "some interacting pairs must lie farther apart than one cell width"
        assert np.allclose(F, S, rtol=1e-12, atol=1e-13 * np.abs(S).max()), (n_cells, channel)
        if not channel:
            assert np.abs(F.sum(axis=0)).max() < 1e-12 * np.abs(F).max()
