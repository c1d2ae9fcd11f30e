"""Run the saddle-point mat-vec of the 256^3 workload a few times per
precision, plus a plain device copy of known size, as a target for
rocprofv3 (kernel statistics and FETCH_SIZE / WRITE_SIZE counters):

    python -m timstochrigidbodiessim_amd.profile_matvec [--reps 3] [--grid 256]
    python3 timstochrigidbodiessim_amd/profile_matvec.py ...   (the same, as a file)

The copy (fcm_profile_copy) reads and writes `--copy-mb` MiB with a plain
streaming kernel; its counter values calibrate the byte counts of the FCM
kernels."""
import argparse
import os
import sys
import time

if __package__ in (None, ""):  # run as a file (rocprofv3 -- python3 .../profile_matvec.py)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from timstochrigidbodiessim_amd import _lib as L
else:
    from . import _lib as L


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--rods", type=int, default=20000)
    ap.add_argument("--spheres", type=int, default=12)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-mb", type=int, default=512)
    ap.add_argument("--spread-variant", type=int, default=None, help="0 or 1 (both precisions)")
    ap.add_argument("--interp-izb", type=int, default=None, help="1, 2 or 3 (both precisions)")
    ap.add_argument("--autotune", action="store_true")
    a = ap.parse_args()
    n = a.grid
    t0 = time.perf_counter()
    ref, x, q, F6, _ = L.synth_rods(2, a.rods, a.spheres, 1.86, [float(n)] * 3, max_attempts=20)
    nb = a.rods * a.spheres
    cfg = L.make_config(n, ng=7, dx=1.0, a=1.86, mu=1.0, kT=1.0, dt=0.20216, max_blobs=nb, max_bodies=a.rods,
                        max_blobs_per_body=a.spheres, max_krylov=4, fp32=True, random_field=True)
    with L.FcmSolver(cfg) as s:
        s.set_geometry(ref)
        s.set_bodies(x, q)
        print(f"[profile] set-up {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        if a.autotune:
            s.autotune(L.FCM_PREC_FP32, 3)
            s.autotune(L.FCM_PREC_FP64, 3)
        if a.spread_variant is not None or a.interp_izb is not None:
            t = s.tuning()
            sv = [a.spread_variant] * 2 if a.spread_variant is not None else t["spread_variant"]
            iz = [a.interp_izb] * 2 if a.interp_izb is not None else t["interp_izb"]
            s.set_tuning(sv, iz)
        print(f"[profile] launch shapes {s.tuning()}", flush=True)
        for name, prec in (("fp32", L.FCM_PREC_FP32), ("fp64", L.FCM_PREC_FP64)):
            ms = s.profile_matvec(prec, a.reps)
            print(f"[profile] {name}: " + ", ".join(f"{k} {v:.3f} ms" for k, v in ms.items()), flush=True)
        nbytes = a.copy_mb << 20
        ms = s.profile_copy(nbytes, a.reps)
        print(f"[profile] copy: {a.copy_mb} MiB read + {a.copy_mb} MiB written per launch, {ms:.3f} ms, "
              f"{2 * nbytes / (ms * 1e-3) / 1e9:.0f} GB/s", flush=True)


if __name__ == "__main__":
    main()
