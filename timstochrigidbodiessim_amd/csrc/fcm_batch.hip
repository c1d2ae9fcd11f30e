// fcm_batch.hip -- solve jobs and the batch driver.
//
// A job is one saddle-point solve (gmres_solver.cpp:148-306) of one handle:
// plain FP64 GMRES, or FP32-bulk inner GMRES solves refined against the FP64
// residual.  A batch advances many jobs together: each step launches a round
// of Arnoldi steps for every job that is iterating (captured once into a
// hipGraph whose branches run on worker streams), then copies the 16-double
// control block of every job into one pinned snapshot.  The host reads the
// snapshot of the PREVIOUS round while the device runs the current one, and
// only then enqueues the per-job transitions (finish, refinement residual,
// next inner solve), so the device is never left waiting for the host.
//
// The single-handle solve (gmres_solve, fcm_mobility_solve) is a batch of one
// on the handle's own stream, so a job gives bitwise the same result alone
// and in a batch: its kernels, their arguments and their order are the same.
#include "fcm_handle.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>

using namespace fcm_host;

namespace {

constexpr int CTL = 16;  // doubles per control block: dctl[0..8) then int ictl[16]

enum class Phase { INNER, WAIT_REFINE, DRAIN, TAKEN };

struct Job {
  fcm_handle h = nullptr;
  int prec = FCM_PREC_FP64;
  double tol = 1e-8, inner_tol = 0.0;
  int max_iter = 1, max_ref = 12;
  Phase phase = Phase::INNER;
  gmres_report rep{};
  long long valid_after = 0;  // first snapshot that reflects the last enqueued transition
  int slot = -1;
  int stream = 0;
  double cost = 0.0;
};

bool graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("FCM_HIP_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

}  // namespace

struct fcm_batch_s {
  hipStream_t master = nullptr;
  bool own_master = false;
  std::vector<hipStream_t> workers;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  int chunk = 4;
  bool graphs = true;

  std::vector<Job*> jobs;        // submitted, not yet taken
  std::deque<Job*> finished;     // done on the device, not yet taken
  int cap = 0;
  std::vector<Job*> slot_job;
  const double** d_src = nullptr;  // device table: control block of each slot (null if free)
  double* d_snap = nullptr;        // cap x CTL gather buffer
  double* h_snap[2] = {nullptr, nullptr};
  hipEvent_t snap_ev[2] = {nullptr, nullptr};
  long long snap_seq[2] = {0, 0};
  bool snap_pending[2] = {false, false};
  int cur = 0;
  long long seq = 0;

  hipGraphExec_t exec = nullptr;
  struct Member {
    fcm_handle h;
    Job* job;
    unsigned long long epoch;
    int prec;
  };
  std::vector<Member> members;  // jobs captured in exec
};

namespace fcm_host {

namespace {

int batch_init(fcm_batch_s* B, hipStream_t master, int nworkers, int chunk, int cap) {
  B->chunk = std::max(1, chunk);
  B->cap = std::max(1, cap);
  B->graphs = graphs_enabled();
  if (master) {
    B->master = master;
  } else {
    HIPCHK(hipStreamCreateWithFlags(&B->master, hipStreamNonBlocking));
    B->own_master = true;
  }
  for (int i = 0; i < nworkers; ++i) {
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    B->workers.push_back(s);
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    B->join.push_back(e);
  }
  HIPCHK(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
  for (int i = 0; i < 2; ++i) {
    HIPCHK(hipEventCreateWithFlags(&B->snap_ev[i], hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&B->h_snap[i], (size_t)B->cap * CTL * sizeof(double), hipHostMallocDefault));
  }
  HIPCHK(hipMalloc((void**)&B->d_src, (size_t)B->cap * sizeof(double*)));
  HIPCHK(hipMalloc((void**)&B->d_snap, (size_t)B->cap * CTL * sizeof(double)));
  HIPCHK(hipMemsetAsync((void*)B->d_src, 0, (size_t)B->cap * sizeof(double*), B->master));
  HIPCHK(hipStreamSynchronize(B->master));
  B->slot_job.assign(B->cap, nullptr);
  return FCM_OK;
}

// h->batches lists the batches holding a job of h or a captured graph that
// references h, so that destroying h can detach it first.
void ref_handle(fcm_batch_s* B, fcm_handle h) {
  if (std::find(h->batches.begin(), h->batches.end(), B) == h->batches.end()) h->batches.push_back(B);
}

void unref_handle(fcm_batch_s* B, fcm_handle h) {
  for (const Job* j : B->jobs)
    if (j->h == h) return;
  for (const auto& m : B->members)
    if (m.h == h) return;
  h->batches.erase(std::remove(h->batches.begin(), h->batches.end(), B), h->batches.end());
}

void drop_graph(fcm_batch_s* B) {
  if (B->exec) {
    (void)hipStreamSynchronize(B->master);
    (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
  }
  std::vector<fcm_handle> hs;
  for (const auto& m : B->members) hs.push_back(m.h);
  B->members.clear();
  for (fcm_handle h : hs) unref_handle(B, h);
}

void batch_free(fcm_batch_s* B) {
  if (!B) return;
  if (B->master) (void)hipStreamSynchronize(B->master);
  drop_graph(B);
  for (Job* j : B->jobs) {
    auto& v = j->h->batches;
    v.erase(std::remove(v.begin(), v.end(), B), v.end());
    delete j;
  }
  for (Job* j : B->finished) delete j;
  for (auto s : B->workers) (void)hipStreamDestroy(s);
  for (auto e : B->join) (void)hipEventDestroy(e);
  if (B->fork) (void)hipEventDestroy(B->fork);
  for (int i = 0; i < 2; ++i) {
    if (B->snap_ev[i]) (void)hipEventDestroy(B->snap_ev[i]);
    if (B->h_snap[i]) (void)hipHostFree(B->h_snap[i]);
  }
  if (B->d_src) (void)hipFree((void*)B->d_src);
  if (B->d_snap) (void)hipFree(B->d_snap);
  if (B->own_master && B->master) (void)hipStreamDestroy(B->master);
  delete B;
}

// entries of the residual history a handle stores (its hist_all buffer)
int hist_capacity(const fcm_handle h) { return h->cfg.max_krylov + 2; }

int inner_max_iter(const Job* j) {
  const fcm_handle h = j->h;
  const long long left = (long long)j->max_iter - j->rep.iterations;
  return (int)std::max<long long>(1, std::min<long long>({left, (long long)h->cfg.max_krylov, (long long)h->nsys()}));
}

// Each inner FP32 solve reduces the current residual by 1e-2 (or just enough to
// reach the tolerance): restarting the Krylov space every ~75 iterations keeps
// the orthogonalisation, whose cost grows with the square of the basis size,
// well below that of one long solve, at about the same iteration count.  The
// inner basis is stored in FP32 (half the bytes per pass over it); the FP64
// refinement of the true residual removes its rounding from the solution.
double inner_tol(const Job* j, double bn, double rn) {
  double itol = j->inner_tol > 0 ? j->inner_tol : std::max(0.5 * j->tol * bn / rn, 1e-2);
  return std::min(itol, 0.5);
}

// enqueue the start of a job (on the master stream)
int job_begin(fcm_batch_s* B, Job* j) {
  fcm_handle h = j->h;
  hipStream_t st = B->master;
  const size_t n = h->nsys();
  HIPCHK(hipEventRecord(h->ev0, st));
  if (j->prec == FCM_PREC_FP32) {
    HIPCHK(axpby(0.0, h->rhs, 0.0, h->xacc, n, st));
    HIPCHK(axpby(1.0, h->rhs, 0.0, h->res, n, st));
    HIPCHK(dots(h->rhs, 0, 1, h->rhs, n, h->kw, h->gm_ctl + 7, st));
    HIPCHK(dots(h->res, 0, 1, h->res, n, h->kw, h->gm.red, st));
    HIPCHK(gm_launch_start(h->gm, inner_tol(j, 1.0, 1.0), inner_max_iter(j), h->res, h->Qf(), n, st));
  } else {
    HIPCHK(dots(h->rhs, 0, 1, h->rhs, n, h->kw, h->gm.red, st));
    HIPCHK(gm_launch_start(h->gm, j->tol, inner_max_iter(j), h->rhs, h->Q, n, st));
  }
  j->phase = Phase::INNER;
  j->valid_after = B->seq + 1;
  return FCM_OK;
}

int job_final(fcm_batch_s* B, Job* j) {
  fcm_handle h = j->h;
  if (j->prec == FCM_PREC_FP32) HIPCHK(axpby(1.0, h->xacc, 0.0, h->sol, h->nsys(), B->master));
  HIPCHK(hipEventRecord(h->ev1, B->master));
  j->phase = Phase::DRAIN;
  j->valid_after = B->seq + 1;
  return FCM_OK;
}

// the inner GMRES of a job has stopped (status word set in the snapshot)
int job_inner_done(fcm_batch_s* B, Job* j, const double* ctl) {
  fcm_handle h = j->h;
  hipStream_t st = B->master;
  const size_t n = h->nsys();
  const int* ic = reinterpret_cast<const int*>(ctl + 8);
  const int status = ic[GM_STATUS] - 1;
  int nc = ic[GM_NCOLS];
  double rel = ctl[4];
  if (status == FCM_GMRES_ZERO_RHS) {
    nc = 0;
    rel = 0.0;
  }
  if (j->prec == FCM_PREC_FP32) HIPCHK(gm_launch_finish(h->gm, h->Qf(), h->ldq, h->tmp, n, st));
  else HIPCHK(gm_launch_finish(h->gm, h->Q, h->ldq, h->tmp, n, st));
  BodyData bdy = h->body();
  HIPCHK(apply_pinv(bdy, h->tmp, h->sol, nullptr, st));
  // hist_all keeps the first hist_capacity(h) entries of the whole solve's history
  const int keep = std::min(nc, hist_capacity(h) - j->rep.iterations);
  if (keep > 0)
    HIPCHK(hipMemcpyAsync(h->hist_all + j->rep.iterations, h->gm.hist, (size_t)keep * sizeof(double),
                          hipMemcpyDeviceToDevice, st));
  j->rep.iterations += nc;
  j->rep.rel_residual = rel;
  if (j->prec != FCM_PREC_FP32) {
    j->rep.status = status;
    j->rep.true_rel_residual = -1.0;
    return job_final(B, j);
  }
  j->rep.refinements += 1;
  HIPCHK(axpby(1.0, h->sol, 1.0, h->xacc, n, st));
  if (int rc = saddle_dev(h, h->xacc, h->tmp, FCM_PREC_FP64, st)) return rc;
  HIPCHK(axpby(1.0, h->rhs, 0.0, h->res, n, st));
  HIPCHK(axpby(-1.0, h->tmp, 1.0, h->res, n, st));
  HIPCHK(dots(h->res, 0, 1, h->res, n, h->kw, h->gm_ctl + 6, st));
  j->phase = Phase::WAIT_REFINE;
  j->valid_after = B->seq + 1;
  return FCM_OK;
}

// the FP64 residual of the refined iterate is in the snapshot
int job_refine_decide(fcm_batch_s* B, Job* j, const double* ctl) {
  fcm_handle h = j->h;
  const double rn = std::sqrt(ctl[6]), bn = std::sqrt(ctl[7]);
  if (bn == 0.0) {
    j->rep.status = FCM_GMRES_ZERO_RHS;
    j->rep.iterations = 0;
    j->rep.refinements = 0;
    j->rep.rel_residual = 0.0;
    j->rep.true_rel_residual = 0.0;
    return job_final(B, j);
  }
  j->rep.true_rel_residual = rn / bn;
  if (rn / bn < j->tol) {
    j->rep.status = FCM_GMRES_CONVERGED;
    return job_final(B, j);
  }
  if (j->rep.iterations >= j->max_iter || j->rep.refinements >= j->max_ref) {
    j->rep.status = FCM_GMRES_MAX_ITER;
    return job_final(B, j);
  }
  const size_t n = h->nsys();
  HIPCHK(dots(h->res, 0, 1, h->res, n, h->kw, h->gm.red, B->master));
  HIPCHK(gm_launch_start(h->gm, inner_tol(j, bn, rn), inner_max_iter(j), h->res, h->Qf(), n, B->master));
  j->phase = Phase::INNER;
  j->valid_after = B->seq + 1;
  return FCM_OK;
}

bool iterating(const Job* j) { return j->phase == Phase::INNER || j->phase == Phase::WAIT_REFINE; }

// launch (or capture) one round: chunk Arnoldi steps of every iterating job
int launch_round(fcm_batch_s* B, const std::vector<Job*>& js) {
  const int P = (int)B->workers.size();
  if (P > 0) {
    HIPCHK(hipEventRecord(B->fork, B->master));
    for (int p = 0; p < P; ++p) HIPCHK(hipStreamWaitEvent(B->workers[p], B->fork, 0));
  }
  for (int c = 0; c < B->chunk; ++c)
    for (Job* j : js) {
      hipStream_t s = P > 0 ? B->workers[j->stream] : B->master;
      if (int rc = launch_arnoldi_step(j->h, j->prec, s)) return rc;
    }
  if (P > 0) {
    for (int p = 0; p < P; ++p) {
      HIPCHK(hipEventRecord(B->join[p], B->workers[p]));
      HIPCHK(hipStreamWaitEvent(B->master, B->join[p], 0));
    }
  }
  return FCM_OK;
}

// balance the iterating jobs over the worker streams by their cost per step
void assign_streams(fcm_batch_s* B, std::vector<Job*>& js) {
  const int P = (int)B->workers.size();
  if (P == 0) return;
  std::vector<Job*> order(js);
  std::stable_sort(order.begin(), order.end(), [](const Job* a, const Job* b) { return a->cost > b->cost; });
  std::vector<double> load(P, 0.0);
  for (Job* j : order) {
    const int p = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    j->stream = p;
    load[p] += j->cost;
  }
}

bool graph_current(const fcm_batch_s* B, const std::vector<Job*>& js) {
  if (!B->exec) return false;
  size_t live = 0;
  for (Job* j : js) {
    bool found = false;
    for (const auto& m : B->members)
      if (m.job == j && m.h == j->h) {
        if (m.epoch != j->h->epoch || m.prec != j->prec) return false;
        found = true;
        break;
      }
    if (!found) return false;
    ++live;
  }
  // drop captured steps of jobs that have left once they are a quarter of the graph
  return 4 * (B->members.size() - live) < B->members.size() || B->members.size() - live == 0;
}

int ensure_graph(fcm_batch_s* B, std::vector<Job*>& js) {
  if (graph_current(B, js)) return FCM_OK;
  drop_graph(B);
  assign_streams(B, js);
  HIPCHK(hipStreamBeginCapture(B->master, hipStreamCaptureModeThreadLocal));
  const int rc = launch_round(B, js);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(B->master, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(ec);
  const hipError_t ei = hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ei);
  for (Job* j : js) {
    B->members.push_back({j->h, j, j->h->epoch, j->prec});
    ref_handle(B, j->h);
  }
  return FCM_OK;
}

int process_snapshot(fcm_batch_s* B, const double* snap, long long sid) {
  for (size_t i = 0; i < B->jobs.size(); ++i) {
    Job* j = B->jobs[i];
    if (sid < j->valid_after) continue;
    const double* ctl = snap + (size_t)j->slot * CTL;
    const int* ic = reinterpret_cast<const int*>(ctl + 8);
    int rc = FCM_OK;
    if (j->phase == Phase::INNER) {
      if (ic[GM_STATUS] != 0) rc = job_inner_done(B, j, ctl);
    } else if (j->phase == Phase::WAIT_REFINE) {
      rc = job_refine_decide(B, j, ctl);
    } else if (j->phase == Phase::DRAIN) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, j->h->ev0, j->h->ev1));
      j->rep.seconds = ms * 1e-3;
      j->h->have_sol = true;
      if (j->h->st != B->master) HIPCHK(hipStreamWaitEvent(j->h->st, j->h->ev1, 0));
      HIPCHK(gm_set_ptr(B->d_src, j->slot, nullptr, B->master));
      B->slot_job[j->slot] = nullptr;
      j->phase = Phase::TAKEN;
      B->finished.push_back(j);
    }
    if (rc) return rc;
  }
  std::vector<fcm_handle> left;
  for (const Job* j : B->jobs)
    if (j->phase == Phase::TAKEN) left.push_back(j->h);
  B->jobs.erase(std::remove_if(B->jobs.begin(), B->jobs.end(), [](const Job* j) { return j->phase == Phase::TAKEN; }),
                B->jobs.end());
  for (fcm_handle h : left) unref_handle(B, h);
  return FCM_OK;
}

int batch_submit(fcm_batch_s* B, fcm_handle h, const gmres_opts* o) {
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  if (!h->have_rhs) return fail(FCM_ERR_STATE, "no right-hand side");
  const int prec = o ? o->precision : FCM_PREC_FP64;
  if (prec == FCM_PREC_FP32 && !h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
  for (Job* j : B->jobs)
    if (j->h == h) return fail(FCM_ERR_STATE, "handle already has a job in this batch");
  if (!h->batches.empty())
    for (fcm_batch_s* other : h->batches)
      if (other != B)
        for (Job* j : other->jobs)
          if (j->h == h) return fail(FCM_ERR_STATE, "handle has a job in another batch");
  int slot = -1;
  for (int i = 0; i < B->cap; ++i)
    if (!B->slot_job[i]) {
      slot = i;
      break;
    }
  if (slot < 0) return fail(FCM_ERR_CAPACITY, "batch is full");
  const int max_iter = o ? o->max_iter : h->cfg.max_krylov;
  if (max_iter < 1) return fail(FCM_ERR_ARG, "max_iter must be >= 1");
  Job* j = new Job();
  j->h = h;
  j->prec = prec;
  j->tol = o ? o->tol : 1e-8;
  j->inner_tol = o ? o->inner_tol : 0.0;
  j->max_iter = max_iter;
  j->max_ref = (o && o->max_refine > 0) ? o->max_refine : 12;
  j->rep.status = FCM_GMRES_MAX_ITER;
  j->rep.true_rel_residual = -1.0;
  j->slot = slot;
  j->cost = (double)h->g.field_words() * (prec == FCM_PREC_FP32 ? 8.0 : 16.0) + 100.0 * h->bd.nb;
  // the job's set-up work on the handle's stream precedes it on the master stream
  if (h->st != B->master) {
    HIPCHK(hipEventRecord(h->ev1, h->st));
    HIPCHK(hipStreamWaitEvent(B->master, h->ev1, 0));
  }
  HIPCHK(gm_set_ptr(B->d_src, slot, h->gm_ctl, B->master));
  B->slot_job[slot] = j;
  B->jobs.push_back(j);
  ref_handle(B, h);
  if (int rc = job_begin(B, j)) {
    B->jobs.pop_back();
    B->slot_job[slot] = nullptr;
    delete j;
    return rc;
  }
  return FCM_OK;
}

int batch_step(fcm_batch_s* B, int* unfinished) {
  std::vector<Job*> js;
  for (Job* j : B->jobs)
    if (iterating(j)) js.push_back(j);
  bool any_inner = false;
  for (Job* j : js) any_inner = any_inner || j->phase == Phase::INNER;
  if (any_inner) {
    if (B->graphs) {
      if (int rc = ensure_graph(B, js)) return rc;
      HIPCHK(hipGraphLaunch(B->exec, B->master));
    } else {
      std::vector<Job*> inner;
      for (Job* j : js)
        if (j->phase == Phase::INNER) inner.push_back(j);
      assign_streams(B, inner);
      if (int rc = launch_round(B, inner)) return rc;
    }
  }
  const int cur = B->cur, prev = cur ^ 1;
  if (!B->jobs.empty()) {
    HIPCHK(gm_gather_ctl(B->d_src, B->d_snap, B->cap, B->master));
    HIPCHK(hipMemcpyAsync(B->h_snap[cur], B->d_snap, (size_t)B->cap * CTL * sizeof(double), hipMemcpyDeviceToHost,
                          B->master));
    HIPCHK(hipEventRecord(B->snap_ev[cur], B->master));
    B->snap_seq[cur] = ++B->seq;
    B->snap_pending[cur] = true;
  }
  if (B->snap_pending[prev]) {
    HIPCHK(hipEventSynchronize(B->snap_ev[prev]));
    B->snap_pending[prev] = false;
    if (int rc = process_snapshot(B, B->h_snap[prev], B->snap_seq[prev])) return rc;
  }
  B->cur = prev;
  if (unfinished) *unfinished = (int)B->jobs.size();
  return FCM_OK;
}

int batch_run(fcm_batch_s* B) {
  int left = 1;
  while (left > 0)
    if (int rc = batch_step(B, &left)) return rc;
  return FCM_OK;
}

}  // namespace

int solve_single(fcm_handle h, const gmres_opts* o, gmres_report* rep, double* hist) {
  if (!h->solo) {
    fcm_batch_s* B = new fcm_batch_s();
    if (int rc = batch_init(B, h->st, 0, 4, 1)) {
      batch_free(B);
      return rc;
    }
    h->solo = B;
  }
  fcm_batch_s* B = h->solo;
  if (int rc = batch_submit(B, h, o)) return rc;
  if (int rc = batch_run(B)) return rc;
  Job* j = B->finished.front();
  B->finished.pop_front();
  const gmres_report r = j->rep;
  delete j;
  HIPCHK(hipStreamSynchronize(h->st));
  const int nh = std::min(r.iterations, hist_capacity(h));
  if (hist && nh > 0) HIPCHK(hipMemcpy(hist, h->hist_all, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost));
  if (rep) *rep = r;
  return FCM_OK;
}

void batch_forget(fcm_handle h) {
  std::vector<fcm_batch_s*> bs = h->batches;
  for (fcm_batch_s* B : bs) {
    drop_graph(B);
    for (Job* j : B->jobs)
      if (j->h == h) {
        // a handle destroyed with its job in flight: wait for the device and drop the job
        (void)hipStreamSynchronize(B->master);
        (void)gm_set_ptr(B->d_src, j->slot, nullptr, B->master);
        (void)hipStreamSynchronize(B->master);
        B->slot_job[j->slot] = nullptr;
        j->phase = Phase::TAKEN;
      }
    B->jobs.erase(std::remove_if(B->jobs.begin(), B->jobs.end(),
                                 [h](const Job* j) { return j->h == h; }),
                  B->jobs.end());
    for (Job*& j : B->finished)
      if (j->h == h) j->h = nullptr;
  }
  h->batches.clear();
}

void batch_destroy_solo(fcm_handle h) {
  if (h->solo) {
    batch_free(h->solo);
    h->solo = nullptr;
  }
}

}  // namespace fcm_host

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int fcm_batch_create(int max_jobs, int worker_streams, int steps_per_round, fcm_batch* out) {
  if (!out || max_jobs < 1 || worker_streams < 0 || worker_streams > 8 || steps_per_round < 1)
    return fail(FCM_ERR_ARG, "bad batch parameters");
  fcm_batch_s* B = new fcm_batch_s();
  if (int rc = batch_init(B, nullptr, worker_streams, steps_per_round, max_jobs)) {
    batch_free(B);
    return rc;
  }
  *out = B;
  return FCM_OK;
}

int fcm_batch_destroy(fcm_batch b) {
  batch_free(b);
  return FCM_OK;
}

int fcm_batch_submit(fcm_batch b, fcm_handle h, const gmres_opts* opts) {
  if (!b || !h) return fail(FCM_ERR_ARG, "null argument");
  return batch_submit(b, h, opts);
}

int fcm_batch_step(fcm_batch b, int* unfinished) {
  if (!b) return fail(FCM_ERR_ARG, "null batch");
  return batch_step(b, unfinished);
}

int fcm_batch_take(fcm_batch b, fcm_handle* handles, gmres_report* reps, int max, int* count) {
  if (!b || !count || max < 0) return fail(FCM_ERR_ARG, "null argument");
  int c = 0;
  while (c < max && !b->finished.empty()) {
    Job* j = b->finished.front();
    b->finished.pop_front();
    if (handles) handles[c] = j->h;
    if (reps) reps[c] = j->rep;
    delete j;
    ++c;
  }
  *count = c;
  return FCM_OK;
}

int fcm_batch_run(fcm_batch b) {
  if (!b) return fail(FCM_ERR_ARG, "null batch");
  return batch_run(b);
}

}  // extern "C"
