"""Serving loop for fluctuating mobility solves: continuous batching with
admission by HBM footprint.

The reference runs one simulation per MPI job (main.cpp); a service that
answers many independent mobility solves of different sizes is new here.
Requests arrive on a clock (a Poisson stream in the benchmark).  Each one is
admitted as soon as a batch slot is free and its handle fits in the HBM
budget: an idle handle of the same shape class is reused (its device buffers
stay allocated), otherwise a new one is created if fcm_query_footprint says
it fits, after idle handles of other shapes are released to make room.  The
admitted request's set-up (body state, seeded fluctuating field, RHS) is
queued on its handle's own stream, so it overlaps with the solves already on
the device; the solve itself joins the running batch (fcm_batch_submit) and
is advanced with everyone else's, round by round.  Latency is measured from
arrival to the host seeing the finished result.
"""
import collections
import math
import time

import numpy as np

from . import _lib as L


class Request:
    """One mobility solve: grid n^3 (dx = 1), NB rods of nbpb spheres of
    radius a, seeded geometry, forces and fluctuating field."""

    def __init__(self, rid, arrival, n, NB, nbpb=10, seed=0, precision=L.FCM_PREC_FP32, tol=1e-8, max_iter=400,
                 a=1.86, kT=1.0, dt=None):
        self.rid = rid
        self.arrival = arrival
        self.n = n
        self.NB = NB
        self.nbpb = nbpb
        self.seed = seed
        self.precision = precision
        self.tol = tol
        self.max_iter = max_iter
        self.a = a
        self.kT = kT
        self.dt = dt if dt is not None else math.pi * a ** 3 / 100.0
        self.geometry = None  # (blob_ref, x, q, F6)
        self.t_admit = None
        self.t_done = None
        self.report = None
        self.solution = None

    def make_geometry(self, max_attempts=5):
        ref, x, q, F6, _ = L.synth_rods(self.seed, self.NB, self.nbpb, self.a, [float(self.n)] * 3,
                                        max_attempts=max_attempts)
        self.geometry = (ref, x, q, F6)

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival


def poisson_requests(rate, count, seed=0, grids=(64, 96, 128, 192, 256), rods=(100, 20000), nbpb=10,
                     precision=L.FCM_PREC_FP32, tol=1e-8, max_iter=400):
    """A Poisson stream: exponential inter-arrival times at `rate` per second,
    grid drawn uniformly from `grids`, rod count log-uniform in `rods`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice(grids))
        NB = int(round(math.exp(rng.uniform(math.log(rods[0]), math.log(rods[1])))))
        out.append(Request(i, t, n, NB, nbpb, seed=1000 + i, precision=precision, tol=tol, max_iter=max_iter))
    return out


def _pow2(v, lo):
    p = lo
    while p < v:
        p *= 2
    return p


class MobilityServer:
    """Continuous-batching request loop over one FcmBatch."""

    def __init__(self, hbm_budget=None, max_jobs=32, worker_streams=4, steps_per_round=4, max_krylov=400,
                 keep_solutions=False, autotune=True, tune_reps=3):
        self.batch = L.FcmBatch(max_jobs=max_jobs, worker_streams=worker_streams, steps_per_round=steps_per_round)
        free, _ = L.device_memory()
        self.budget = int(hbm_budget) if hbm_budget else int(0.85 * free)
        self.max_jobs = max_jobs
        self.max_krylov = max_krylov
        self.keep_solutions = keep_solutions
        # launch shapes: each shape class is tuned once, by its first request, in
        # both precisions (an FP32 solve refines with FP64 products); later
        # requests, and new handles of the class, take the library's cached choice,
        # so a request runs the same shapes as the same request solved alone
        self.autotune = autotune
        self.tune_reps = tune_reps
        self.tunings_run = 0  # (shape class, precision) pairs timed
        self.used = 0
        self.idle = collections.defaultdict(list)  # shape key -> idle solvers
        self.bytes_of = {}
        self.refs = {}  # (n, nbpb, a) -> (Minv, N)
        self.active = {}  # id(solver) -> (request, key, solver)
        self.created = 0
        self.max_active = 0
        self.max_used = 0

    # ---- handles ---------------------------------------------------------
    def config_for(self, req):
        """Shape class of a request: the rod capacity is rounded up to a power
        of two so that handles are reused across requests of similar size."""
        cap = _pow2(req.NB, 64)
        cfg = L.make_config(req.n, ng=7, dx=1.0, a=req.a, mu=1.0, kT=req.kT, dt=req.dt, max_blobs=cap * req.nbpb,
                            max_bodies=cap, max_blobs_per_body=req.nbpb, max_krylov=self.max_krylov, fp32=True,
                            random_field=True)
        key = (req.n, cap, req.nbpb, req.a, req.kT, req.dt)
        return key, cfg

    def _acquire(self, req):
        key, cfg = self.config_for(req)
        if self.idle[key]:
            return key, self.idle[key].pop()
        need = L.footprint(cfg)
        if self.used + need > self.budget:
            # release idle handles of other shapes, largest first
            idle = sorted(((self.bytes_of[id(s)], k, s) for k, v in self.idle.items() for s in v),
                          key=lambda t: -t[0])
            for b, k, s in idle:
                if self.used + need <= self.budget:
                    break
                self.idle[k].remove(s)
                self._release(s)
        if self.used + need > self.budget:
            return None
        s = L.FcmSolver(cfg)
        self.bytes_of[id(s)] = need
        self.used += need
        self.max_used = max(self.max_used, self.used)
        self.created += 1
        return key, s

    def _release(self, s):
        self.used -= self.bytes_of.pop(id(s))
        s.close()

    def prepare_refs(self, requests):
        """Compute the single-body reference matrices (fcm_fluid_solver.cpp:842-881)
        of every (grid, rod) shape in `requests` ahead of time."""
        for req in requests:
            rkey = (req.n, req.nbpb, req.a)
            if rkey in self.refs:
                continue
            key, cfg = self.config_for(req)
            ref = np.array([[2 * req.a * (m - (req.nbpb - 1) / 2), 0.0, 0.0] for m in range(req.nbpb)])
            with L.FcmSolver(L.make_config(req.n, ng=7, dx=1.0, a=req.a, mu=1.0, max_blobs=req.nbpb, max_bodies=1,
                                           max_blobs_per_body=req.nbpb, max_krylov=1)) as s:
                s.set_geometry(ref)
                _, Minv, N = s.ref_matrices()
            self.refs[rkey] = (Minv, N)

    def _setup(self, req, s):
        ref, x, q, F6 = req.geometry
        s.set_geometry(ref)
        rkey = (req.n, req.nbpb, req.a)
        if rkey not in self.refs:
            _, Minv, N = s.ref_matrices()
            self.refs[rkey] = (Minv, N)
        Minv, N = self.refs[rkey]
        s.set_bodies(x, q)
        if self.autotune:
            for prec in (L.FCM_PREC_FP32, L.FCM_PREC_FP64):
                _, hit = s.autotune_cached(prec, reps=self.tune_reps)
                self.tunings_run += 0 if hit else 1
        s.set_ref_matrices(Minv, N)
        s.set_body_forces(F6)
        s.mobility_prepare(req.seed)

    def try_admit(self, req, now):
        if len(self.active) >= self.max_jobs:
            return False
        got = self._acquire(req)
        if got is None:
            return False
        key, s = got
        if req.geometry is None:
            req.make_geometry()
        self._setup(req, s)
        self.batch.submit(s, tol=req.tol, max_iter=req.max_iter, precision=req.precision)
        req.t_admit = now
        self.active[id(s)] = (req, key, s)
        self.max_active = max(self.max_active, len(self.active))
        return True

    def collect(self, now):
        done = []
        for s, rep in self.batch.take():
            req, key, s2 = self.active.pop(id(s))
            req.t_done = now
            req.report = rep
            if self.keep_solutions:
                req.solution = s2.solution()
            self.idle[key].append(s2)
            done.append(req)
        return done

    # ---- the loop --------------------------------------------------------
    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (their .arrival in seconds from the start) and
        return the statistics; each request gets t_admit, t_done, report."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        t0 = clock()
        finished = []
        while pending or queue or self.active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while queue and self.try_admit(queue[0], now):
                queue.popleft()
            if self.active:
                self.batch.step()
                finished += self.collect(clock() - t0)
            elif queue:
                raise RuntimeError(f"request {queue[0].rid} can never be admitted (HBM budget {self.budget} bytes)")
            elif pending:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - (clock() - t0))))
        wall = clock() - t0
        return self.stats(finished, wall)

    @staticmethod
    def stats(finished, wall):
        lat = np.array([r.latency for r in finished]) if finished else np.zeros(1)
        conv = sum(1 for r in finished if r.report and r.report["status"] == "converged")
        return {"requests": len(finished), "converged": conv, "wall_s": wall,
                "solves_per_s": len(finished) / wall if wall > 0 else 0.0,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "mean_latency_s": float(lat.mean()),
                "iterations": int(sum(r.report["iterations"] for r in finished if r.report))}

    def close(self):
        for v in self.idle.values():
            for s in v:
                self._release(s)
        self.idle.clear()
        for req, key, s in list(self.active.values()):
            self._release(s)
        self.active.clear()
        self.batch.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
