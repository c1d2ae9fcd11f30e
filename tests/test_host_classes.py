"""The reference-named C++ host classes (timstochrigidbodiessim_amd/host):
host-only math on the CPU; on the GPU, time steps driven through
fcm_fluid_solver / collection_of_rigid_bodies / linked_list / gmres_solver in
main.cpp's order, checked against the oracle (first solve <= 1e-10) and
bitwise against the device-resident fcm_em_step path."""
import os
import struct
import subprocess

import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa

EXE = os.path.join(os.path.dirname(fa.__file__), "fcm_host_check")


def _need_exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (python -m timstochrigidbodiessim_amd.build)")


def test_host_math_classes():
    _need_exe()
    r = subprocess.run([EXE, "math"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def _write_problem(path, n, NB, nbpb, steps, prec, seed, a, kT, dt, tol, cutoff, ref, x, q, F6, channel=None):
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", n, NB, nbpb, steps, prec))
        f.write(struct.pack("<Q", seed))
        f.write(struct.pack("<5d", a, kT, dt, tol, cutoff))
        for arr in (ref, x, q, F6):
            f.write(np.ascontiguousarray(arr, dtype="<f8").tobytes())
        if channel is not None:  # (slip_channel, shear_force_scale, shear_wavenumber[, algorithm])
            f.write(struct.pack("<i2d", *channel[:3]))
            if len(channel) > 3:
                f.write(struct.pack("<i", channel[3]))


def _read_result(path, nsys, NB, steps):
    raw = open(path, "rb").read()
    off, sols, its = 0, [], []
    for _ in range(steps):
        sols.append(np.frombuffer(raw, "<f8", nsys, off).copy())
        off += 8 * nsys
        its.append(struct.unpack_from("<i", raw, off)[0])
        off += 4
    x = np.frombuffer(raw, "<f8", 3 * NB, off).reshape(NB, 3)
    off += 24 * NB
    q = np.frombuffer(raw, "<f8", 4 * NB, off).reshape(NB, 4)
    return sols, its, x, q


@pytest.mark.gpu
def test_host_classes_time_steps(tmp_path):
    _need_exe()
    n, NB, nbpb, steps, seed, a, kT, dt, tol, cutoff = 32, 10, 6, 3, 77, 1.86, 1.0, 0.20216, 1e-10, 2.2 * 1.86
    ref, x, q, F6, _ = fa.synth_rods(5, NB, nbpb, a, [float(n)] * 3, max_attempts=20)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_problem(inp, n, NB, nbpb, steps, fa.FCM_PREC_FP64, seed, a, kT, dt, tol, cutoff, ref, x, q, F6)
    r = subprocess.run([EXE, "steps", inp, out], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    nb = NB * nbpb
    sols, its, xh, qh = _read_result(out, 3 * nb + 6 * NB, NB, steps)
    # first step against the oracle
    oc = po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=dt)
    o = po.Fcm(oc, nb)
    st, M, Minv, N = o.ref_matrices(ref)
    Y = po.fill_blob_pos(oc, ref, x, q)
    o.set_up_gaussians(Y)
    o.fill_stress(seed)
    o.random_field()
    Vr = o.random_interp()
    f = po.barrier_forces(oc, Y, nbpb, cutoff)
    rhs = np.concatenate([-Vr, -(po.mul_KT(ref, q, f) + F6.ravel())])
    st, sol, it, _ = po.gmres(o, po.Bodies(ref, x, q, Minv, N), rhs, tol=tol, max_iter=200)
    o.close()
    err = np.linalg.norm(sols[0] - sol) / np.linalg.norm(sol)
    print("host-class first solve vs oracle", err, "iterations", its[0], it)
    assert err < 1e-10
    # every step bitwise equal to the device-resident Euler-Maruyama step
    cfg = fa.make_config(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=dt, max_blobs=nb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=200, fp32=False, random_field=True)
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        s.set_body_forces(F6)
        for k in range(steps):
            rep = s.em_step(seed + k, cutoff, tol=tol, max_iter=200)
            assert rep["iterations"] == its[k]
        xd, qd = s.bodies()
    assert np.array_equal(xd, xh) and np.array_equal(qd, qh)


@pytest.mark.gpu
def test_host_classes_channel_shear_time_steps(tmp_path):
    """The host-class Euler-Maruyama loop in a slip channel with a shear field
    set (main.cpp:98-102, 304-326): as in main.cpp, the EM branch never
    computes Vshear, so the RHS carries no shear term (gmres_solver.cpp:62-64
    subtracts an unfilled Vshear); first solve against the oracle, every step
    bitwise equal to the device-resident step with the same shear field."""
    _need_exe()
    n, NB, nbpb, steps, seed, a, kT, dt, tol, cutoff = 32, 8, 3, 3, 91, 1.86, 1.0, 0.20216, 1e-10, 2.2 * 1.86
    scale, kappa = 2.0, 1.0
    margin = a * (nbpb - 1) + 0.5
    ref, x, q, F6, _ = fa.synth_rods(6, NB, nbpb, a, [float(n), n / 2 - 2 * margin, float(n)], max_attempts=20)
    x[:, 1] += margin
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_problem(inp, n, NB, nbpb, steps, fa.FCM_PREC_FP64, seed, a, kT, dt, tol, cutoff, ref, x, q, F6,
                   channel=(1, scale, kappa))
    r = subprocess.run([EXE, "steps", inp, out], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "redone" not in r.stdout  # the rods stay clear of the walls over these steps
    nb = NB * nbpb
    sols, its, xh, qh = _read_result(out, 3 * nb + 6 * NB, NB, steps)
    oc = po.cfg(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=dt, slip_channel=True)
    o = po.Fcm(oc, nb)
    st, M, Minv, N = o.ref_matrices(ref)
    o.shear_field(scale, kappa)
    Y = po.fill_blob_pos(oc, ref, x, q)
    o.set_up_gaussians(Y)
    o.fill_stress(seed)
    o.random_field()
    Vr = o.random_interp()
    f = po.barrier_forces(oc, Y, nbpb, cutoff)
    rhs = np.concatenate([-Vr, -(po.mul_KT(ref, q, f) + F6.ravel())])
    st, sol, it, _ = po.gmres(o, po.Bodies(ref, x, q, Minv, N), rhs, tol=tol, max_iter=200)
    o.close()
    err = np.linalg.norm(sols[0] - sol) / np.linalg.norm(sol)
    print("host-class first channel+shear solve vs oracle", err, "iterations", its[0], it)
    assert err < 1e-10
    cfg = fa.make_config(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=dt, max_blobs=nb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=200, fp32=False, random_field=True, slip_channel=True)
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.shear_field(scale, kappa)
        s.set_bodies(x, q)
        s.set_body_forces(F6)
        for k in range(steps):
            rep = s.em_step(seed + k, cutoff, tol=tol, max_iter=200)
            assert not s.channel_wall_overlap()
            assert rep["iterations"] == its[k]
        xd, qd = s.bodies()
    assert np.array_equal(xd, xh) and np.array_equal(qd, qh)


@pytest.mark.gpu
def test_host_classes_drifter_corrector_steps(tmp_path):
    """The host-class loop's GEN_DRIFTER_CORRECTOR branch (main.cpp:235-268:
    advance_to_midstep, calculate_divU, the mid-step blob velocities) gives
    bitwise the device-resident drifter-corrector steps."""
    _need_exe()
    import _gdc
    n, NB, nbpb, steps, seed, a, kT, dt, tol, cutoff = 32, 8, 6, 2, 57, 1.86, 1.0, 0.20216, 1e-10, 2.2 * 1.86
    ref = _gdc.reference_rods(nbpb, a)
    rng = np.random.default_rng(8)
    x = rng.uniform(0, n, (NB, 3))
    q = rng.standard_normal((NB, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    F6 = np.zeros((NB, 6))
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_problem(inp, n, NB, nbpb, steps, fa.FCM_PREC_FP64, seed, a, kT, dt, tol, cutoff, ref, x, q, F6,
                   channel=(0, 0.0, 0.0, 1))
    r = subprocess.run([EXE, "steps", inp, out], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "redone" not in r.stdout
    nb = NB * nbpb
    sols, its, xh, qh = _read_result(out, 3 * nb + 6 * NB, NB, steps)
    cfg = fa.make_config(n, ng=7, dx=1.0, a=a, mu=1.0, kT=kT, dt=dt, max_blobs=nb, max_bodies=NB,
                         max_blobs_per_body=nbpb, max_krylov=200, fp32=False, random_field=True)
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.ref_matrices()
        s.set_bodies(x, q)
        s.set_body_forces(F6)
        for k in range(steps):
            rep = s.gdc_step(seed + k, cutoff, fd_delta=1e-5, tol=tol, max_iter=200)
            assert not rep["redo"] and rep["iterations"] == its[k]
            assert np.array_equal(s.solution(), sols[k])
        xd, qd = s.bodies()
    assert np.array_equal(xd, xh) and np.array_equal(qd, qh)
