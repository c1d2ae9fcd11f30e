"""Fixed numbers: tests/golden/fcm_small.npz (written by
tests/golden/make_golden.py from the oracle).  The oracle must keep
reproducing them (CPU), and the GPU path must match them to the parity bar."""
import os

import numpy as np
import pytest

import pyoracle as po
import timstochrigidbodiessim_amd as fa
from _common import rel

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "fcm_small.npz"))


def test_oracle_reproduces_golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    d = mg.make()
    for k in ("Y", "Vmob", "Vrand", "M", "Minv", "N", "rhs"):
        assert rel(d[k], G[k]) < 1e-13, k
    assert d["iterations"] == int(G["iterations"]) and d["status"] == int(G["status"]) == 0
    assert rel(d["sol"], G["sol"]) < 1e-12


@pytest.mark.gpu
def test_gpu_matches_golden():
    n, nbpb, NB = int(G["n"]), int(G["nbpb"]), int(G["NB"])
    cfg = fa.make_config(n, ng=7, dx=float(G["dx"]), a=float(G["a"]), mu=float(G["mu"]), kT=float(G["kT"]),
                         dt=float(G["dt"]), max_blobs=NB * nbpb, max_bodies=NB, max_blobs_per_body=nbpb,
                         max_krylov=60, fp32=True, random_field=True)
    with fa.FcmSolver(cfg) as s:
        s.set_geometry(G["ref"])
        M, Minv, N = s.ref_matrices()
        assert rel(M, G["M"]) < 1e-12
        s.set_bodies(G["x"], G["q"])
        assert rel(s.blob_positions(), G["Y"]) < 1e-15
        assert rel(s.mobility(G["Fmob"]), G["Vmob"]) < 1e-12
        s.set_ref_matrices(G["Minv"], G["N"])
        s.set_body_forces(G["F6"])
        rep = s.mobility_solve(int(G["stress_seed"]), tol=1e-12, max_iter=60)
        assert rel(s.random_interp(), G["Vrand"]) < 1e-12
        assert rep["status"] == "converged" and abs(rep["iterations"] - int(G["iterations"])) <= 1
        assert rel(s.solution(), G["sol"]) < 1e-10
