"""Writes tests/golden/fcm_small.npz: small oracle inputs and outputs kept as
fixtures, so that a change in the oracle or the GPU path shows up against
fixed numbers.  Inputs: a 16^3 grid (dx 0.9, a 1.2, mu 1.1, kT 1, dt 0.2),
3 rods of 5 spheres from a fixed seed, N(0,1) body forces.  Outputs: the
seeded random blob velocities, the mobility of a fixed force vector, the
single-body reference matrices, the saddle-point solve (tol 1e-12) and its
iteration count.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as po  # noqa: E402


def make():
    n, nbpb, NB = 16, 5, 3
    c = po.cfg(n, ng=7, dx=0.9, a=1.2, mu=1.1, kT=1.0, dt=0.2)
    rng = np.random.default_rng(2024)
    ref = np.array([[2.4 * (m - 2), 0.0, 0.0] for m in range(nbpb)])
    x = rng.uniform(0, n * 0.9, (NB, 3))
    q = rng.standard_normal((NB, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    F6 = rng.standard_normal((NB, 6))
    for b in range(NB):  # no torque about the rod axis (collinear rods)
        R = np.zeros(9)
        po.lib().or_quat_matrix(q[b], R)
        axis = R.reshape(3, 3)[:, 0]
        F6[b, 3:] -= axis * (F6[b, 3:] @ axis)
    Fmob = rng.standard_normal(3 * NB * nbpb)
    o = po.Fcm(c, NB * nbpb)
    st, M, Minv, N = o.ref_matrices(ref)
    Y = po.fill_blob_pos(c, ref, x, q)
    o.set_up_gaussians(Y)
    Vmob = o.mobility(Fmob)
    o.fill_stress(99)
    o.random_field()
    Vr = o.random_interp()
    rhs = np.concatenate([-Vr, -F6.ravel()])
    st, sol, it, hist = po.gmres(o, po.Bodies(ref, x, q, Minv, N), rhs, tol=1e-12, max_iter=60)
    o.close()
    return dict(n=n, nbpb=nbpb, NB=NB, dx=0.9, a=1.2, mu=1.1, kT=1.0, dt=0.2, stress_seed=99, ref=ref, x=x, q=q,
                F6=F6, Fmob=Fmob, Y=Y, Vmob=Vmob, Vrand=Vr, M=M, Minv=Minv, N=N, rhs=rhs, sol=sol, iterations=it,
                status=st)


if __name__ == "__main__":
    np.savez(os.path.join(HERE, "fcm_small.npz"), **make())
    print("wrote", os.path.join(HERE, "fcm_small.npz"))
