"""Shared helpers for the parity tests: matching configurations for the HIP
library and the CPU oracle."""
import numpy as np

import pyoracle as po
import timstochrigidbodiessim_amd as fa


def pair_cfg(nx, ny=None, nz=None, ng=7, dx=1.0, a=1.86, mu=1.0, kT=0.0, dt=1.0, max_blobs=64, max_bodies=8,
             max_blobs_per_body=16, max_krylov=200, fp32=True, random_field=True, slip_channel=False):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    gc = fa.make_config(nx, ny, nz, ng=ng, dx=dx, a=a, mu=mu, kT=kT, dt=dt, max_blobs=max_blobs,
                        max_bodies=max_bodies, max_blobs_per_body=max_blobs_per_body, max_krylov=max_krylov,
                        fp32=fp32, random_field=random_field, slip_channel=slip_channel)
    oc = po.cfg(nx, ny, nz, ng=ng, dx=dx, a=a, mu=mu, kT=kT, dt=dt, slip_channel=slip_channel)
    return gc, oc


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def rod(nbpb, a):
    return np.array([[2 * a * (m - (nbpb - 1) / 2), 0.0, 0.0] for m in range(nbpb)])


def random_quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1)[:, None]
