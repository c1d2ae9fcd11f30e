// fcm_capi.hip -- host side of libfcm_hip.so: the C ABI of include/fcm_hip.h.
//
// A handle owns one periodic grid, its blobs and bodies, the GMRES basis and
// a HIP stream (default priority).  Host code here only sequences kernels,
// moves the small per-iteration scalars of GMRES and runs the Givens/least-
// squares algebra of gmres_solver.cpp:212-284 on those scalars.
#include "fcm_internal.hpp"
#include "../../include/fcm_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fcm_dense.hpp"

using namespace fcm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return fail(FCM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr double PI = 3.14159265358979323846264338327950288;

int largest_divisor_le(int n, int cap) {
  for (int d = std::min(n, cap); d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

struct fcm_handle_s {
  fcm_config cfg{};
  GridGeom g;
  BlobBins bb;
  ProjParams pp{};
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  c2<double>* f64[3] = {nullptr, nullptr, nullptr};
  c2<float>* f32[3] = {nullptr, nullptr, nullptr};
  c2<double>* stress[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool have_urand = false;
  void* twbuf64 = nullptr;
  void* twbuf32 = nullptr;
  TwiddleSet<double> tw64;
  TwiddleSet<float> tw32;

  // blobs
  BlobData bd;
  double* dY = nullptr;
  bool have_blobs = false;

  // bodies
  int nbpb = 0, NB = 0;
  double *d_ref = nullptr, *d_x = nullptr, *d_q = nullptr, *d_R = nullptr, *d_dpos = nullptr;
  double *d_Minv = nullptr, *d_N = nullptr, *d_F6 = nullptr;
  bool have_geometry = false, have_bodies = false, have_ref = false, have_forces = false;
  std::vector<double> h_ref;

  // vectors
  size_t nmax = 0, ldq = 0;
  double* Q = nullptr;        // (max_krylov + 1) columns
  double* z = nullptr;        // preconditioned vector
  double* rhs = nullptr;
  double* sol = nullptr;
  double* xacc = nullptr;
  double* res = nullptr;
  double* tmp = nullptr;
  double* Vrand = nullptr;
  double* Vtmp = nullptr;
  double* red = nullptr;      // reductions, 2*(max_krylov+2)+8
  double* coef = nullptr;     // coefficients of the reorthogonalisation pass
  KrylovWork kw;
  // device-controlled GMRES
  GmDev gm;
  double* gm_ctl = nullptr;   // 16 doubles: dctl[0..8) then ictl (int[16])
  double* w = nullptr;        // A P^-1 q_k of the current step (fixed address for graph capture)
  double* hctl = nullptr;     // pinned host snapshots, 3 x 16 doubles
  hipEvent_t pev[2] = {nullptr, nullptr};
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    long long key[3] = {-1, -1, -1};
  } graph[2];
  bool have_rhs = false, have_sol = false;
  std::vector<void*> allocs;

  BodyData body() const {
    BodyData b;
    b.NB = NB;
    b.nbpb = nbpb;
    b.ref = d_ref;
    b.x = d_x;
    b.q = d_q;
    b.R = d_R;
    b.dpos = d_dpos;
    b.Y = dY;
    b.Minv = d_Minv;
    b.Nref = d_N;
    return b;
  }
  size_t n3() const { return 3 * (size_t)bd.nb; }
  size_t nsys() const { return 3 * (size_t)NB * nbpb + 6 * (size_t)NB; }
  FieldPtrs<double> fp64() const {
    FieldPtrs<double> p{};
    for (int i = 0; i < 3; ++i) p.p[i] = f64[i];
    return p;
  }
  FieldPtrs<float> fp32() const {
    FieldPtrs<float> p{};
    for (int i = 0; i < 3; ++i) p.p[i] = f32[i];
    return p;
  }
  FieldPtrs<double> fpstress() const {
    FieldPtrs<double> p{};
    for (int i = 0; i < 6; ++i) p.p[i] = stress[i];
    return p;
  }
  FieldPtrs<double> fprand() const {
    FieldPtrs<double> p{};
    for (int i = 0; i < 3; ++i) p.p[i] = stress[i];
    return p;
  }
};

namespace {

struct Sizes {
  size_t field64, field32, twbytes64, twbytes32, total;
  size_t nmax, ldq;
};

bool valid_cfg(const fcm_config* c, std::string& why) {
  const int n[3] = {c->nx, c->ny, c->nz};
  for (int d = 0; d < 3; ++d) {
    if (n[d] < 8 || n[d] % 2 != 0) { why = "grid sizes must be even and >= 8"; return false; }
    if (d < 2 && !fft_len_supported(n[d])) { why = "unsupported FFT length " + std::to_string(n[d]); return false; }
  }
  if (!fft_len_supported(c->nz / 2)) { why = "unsupported z length " + std::to_string(c->nz); return false; }
  if (c->ng < 1 || c->ng > 16) { why = "ng must be in 1..16"; return false; }
  if (c->ng > std::min(c->nx, std::min(c->ny, c->nz))) { why = "ng exceeds the grid"; return false; }
  if (!(c->dx > 0) || !(c->a > 0) || !(c->mu > 0)) { why = "dx, a and mu must be positive"; return false; }
  if (c->max_blobs < 1 || c->max_bodies < 1 || c->max_blobs_per_body < 1 || c->max_krylov < 1) {
    why = "capacities must be positive";
    return false;
  }
  if (c->max_krylov > 2048) {
    why = "max_krylov must be <= 2048";
    return false;
  }
  return true;
}

GridGeom make_geom(const fcm_config* c) {
  GridGeom g;
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
  g.nzc = c->nz / 2 + 1;
  g.cpitch = round_up(g.nzc, 8);  // 128-byte rows for FP64
  g.dx = c->dx;
  return g;
}

Sizes sizes_of(const fcm_config* c) {
  Sizes s{};
  GridGeom g = make_geom(c);
  s.field64 = g.field_words() * sizeof(double2);
  s.field32 = g.field_words() * sizeof(float2);
  const size_t ntw = (size_t)c->nx + c->ny + c->nz / 2 + c->nz / 2 + 1;
  s.twbytes64 = ntw * sizeof(double2);
  s.twbytes32 = ntw * sizeof(float2);
  s.nmax = 3 * (size_t)c->max_blobs + 6 * (size_t)c->max_bodies;
  s.ldq = (s.nmax + 255) / 256 * 256;
  const int nb = c->max_blobs;
  BlobBins bb;
  bb.nbx = c->nx / largest_divisor_le(c->nx, 4);
  bb.nby = c->ny / largest_divisor_le(c->ny, 8);
  bb.nbz = c->nz / largest_divisor_le(c->nz, 8);
  const size_t nbins = (size_t)bb.nbx * bb.nby * bb.nbz;
  size_t t = 0;
  t += 3 * s.field64 + s.twbytes64;
  if (c->fp32) t += 3 * s.field32 + s.twbytes32;
  if (c->random_field) t += 6 * s.field64;
  t += (size_t)nb * (3 * 8 + 4 + 4 + 16 + 3 * c->ng * 12 + 8 * 4) + (3 * nbins + 2) * 4;
  {
    GridGeom g2 = make_geom(c);
    t += 3 * (size_t)(spread_tiles(g2) + 1) * 4;
  }
  const int n3b = 3 * c->max_blobs_per_body;
  t += (size_t)(3 * n3b + n3b * n3b + 36) * 8 + (size_t)c->max_bodies * (3 + 4 + 9 + 6) * 8 + (size_t)nb * 3 * 8;
  t += (size_t)(c->max_krylov + 1) * s.ldq * 8 + 6 * s.ldq * 8 + 2 * (size_t)nb * 3 * 8;
  const size_t nblk = (s.nmax + 1023) / 1024;
  t += nblk * (2 * (size_t)c->max_krylov + 6) * 8 + 2 * (2 * (size_t)c->max_krylov + 16) * 8;
  const size_t m2 = (size_t)(c->max_krylov + 2) * (c->max_krylov + 2);
  t += (2 * m2 + 8 * (size_t)(c->max_krylov + 2) + 16) * 8 + s.ldq * 8;
  s.total = t;
  return s;
}

template <typename T> T* dalloc(fcm_handle h, size_t count, hipError_t& err) {
  void* p = nullptr;
  if (count == 0) count = 1;
  err = hipMalloc(&p, count * sizeof(T));
  if (err == hipSuccess) h->allocs.push_back(p);
  return static_cast<T*>(p);
}

template <typename T>
void twiddles(int n, int sign_unused, std::vector<c2<T>>& out) {
  (void)sign_unused;
  for (int m = 0; m < n; ++m) {
    const long double ang = -2.0L * (long double)PI * (long double)m / (long double)n;
    out.push_back(mk<T>((T)cosl(ang), (T)sinl(ang)));
  }
}

template <typename T>
int upload_twiddles(fcm_handle h, void** buf, TwiddleSet<T>& ts) {
  const GridGeom& g = h->g;
  std::vector<c2<T>> v;
  twiddles<T>(g.nx, -1, v);
  const size_t oy = v.size();
  twiddles<T>(g.ny, -1, v);
  const size_t oz = v.size();
  twiddles<T>(g.nz / 2, -1, v);
  const size_t ou = v.size();
  for (int k = 0; k <= g.nz / 2; ++k) {
    const long double ang = -2.0L * (long double)PI * (long double)k / (long double)g.nz;
    v.push_back(mk<T>((T)cosl(ang), (T)sinl(ang)));
  }
  hipError_t e;
  c2<T>* d = dalloc<c2<T>>(h, v.size(), e);
  HIPCHK(e);
  HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(c2<T>), hipMemcpyHostToDevice));
  *buf = d;
  ts.x = d;
  ts.y = d + oy;
  ts.z = d + oz;
  ts.zu = d + ou;
  return FCM_OK;
}

int check(fcm_handle h) {
  if (!h) return fail(FCM_ERR_ARG, "null handle");
  return FCM_OK;
}

// zero-fill with a kernel-free path: hipMemsetAsync is not used in sequences
// a caller could capture into a graph; these fills run at set-up only.
int zero_field64(fcm_handle h, c2<double>* p) {
  HIPCHK(hipMemsetAsync(p, 0, h->g.field_words() * sizeof(double2), h->st));
  return FCM_OK;
}

int sync_stream(fcm_handle h) {
  HIPCHK(hipStreamSynchronize(h->st));
  return FCM_OK;
}

// ---------------------------------------------------------------------------
// mobility product on device vectors: Vout = dx^3 J S Fin (- K zV if given)
// ---------------------------------------------------------------------------
int mobility_dev(fcm_handle h, const double* Fin, double* Vout, int precision, const KTerm& kt,
                 const int* gate = nullptr) {
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated (fcm_config.fp32 = 0)");
    FieldPtrs<float> fp = h->fp32();
    fp.gate = gate;
    HIPCHK(spread<float>(h->g, h->bb, h->bd, Fin, fp, h->st));
    HIPCHK(stokes_solve<float>(h->g, fp, h->pp, h->tw32, h->st));
    HIPCHK(interp<float>(h->g, h->bb, h->bd, fp, Vout, kt, h->st));
  } else {
    FieldPtrs<double> fp = h->fp64();
    fp.gate = gate;
    HIPCHK(spread<double>(h->g, h->bb, h->bd, Fin, fp, h->st));
    HIPCHK(stokes_solve<double>(h->g, fp, h->pp, h->tw64, h->st));
    HIPCHK(interp<double>(h->g, h->bb, h->bd, fp, Vout, kt, h->st));
  }
  return FCM_OK;
}

// saddle-point product (gmres_solver.cpp:72-83) on device vectors
int saddle_dev(fcm_handle h, const double* in, double* out, int precision) {
  const size_t n3 = h->n3();
  KTerm kt;
  kt.zV = in + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  int rc = mobility_dev(h, in, out, precision, kt);
  if (rc) return rc;
  BodyData b = h->body();
  HIPCHK(mul_KT(b, in, out + n3, h->st));
  HIPCHK(axpby(0.0, out + n3, -1.0, out + n3, 6 * (size_t)h->NB, h->st));
  return FCM_OK;
}

// w = A P^{-1} q  (one Arnoldi product; the K^T term comes out of the P^{-1} kernel)
int arnoldi_product(fcm_handle h, const double* q, double* w, int precision) {
  const size_t n3 = h->n3();
  BodyData b = h->body();
  HIPCHK(apply_pinv(b, q, h->z, w + n3, h->st));
  KTerm kt;
  kt.zV = h->z + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  return mobility_dev(h, h->z, w, precision, kt);
}

double dev_norm(fcm_handle h, const double* v, size_t n, int& rc) {
  rc = FCM_OK;
  hipError_t e = dots(v, 0, 1, v, n, h->kw, h->red, h->st);
  if (e != hipSuccess) { rc = fail(FCM_ERR_HIP, hipGetErrorString(e)); return 0; }
  double s = 0;
  e = hipMemcpyAsync(&s, h->red, sizeof(double), hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  if (e != hipSuccess) { rc = fail(FCM_ERR_HIP, hipGetErrorString(e)); return 0; }
  return std::sqrt(s);
}

struct InnerResult {
  int status = FCM_GMRES_MAX_ITER;
  int iters = 0;
  double rel = 0;
};

// One Arnoldi step under device control: w = A P^-1 q_k with k read on the
// device, then the DCGS2 basis update and the Givens step.  Every kernel is
// gated on the status word, so steps queued after convergence do nothing.
int launch_arnoldi_step(fcm_handle h, int precision) {
  const size_t n = h->nsys(), n3 = h->n3();
  const int* gate = h->gm.ictl + GM_STATUS;
  BodyData b = h->body();
  b.gate = gate;
  b.kcol = h->gm.ictl + GM_K;
  b.ldq = h->ldq;
  HIPCHK(apply_pinv(b, h->Q, h->z, h->w + n3, h->st));
  KTerm kt;
  kt.zV = h->z + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  if (int rc = mobility_dev(h, h->z, h->w, precision, kt, gate)) return rc;
  HIPCHK(gm_launch_basis(h->gm, h->Q, h->ldq, h->w, n, h->kw.partial, gm_blocks(n), h->st));
  return FCM_OK;
}

bool graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("FCM_HIP_GRAPHS");
    return !(e && e[0] == '0');
  }();
  return on;
}

// The Arnoldi step captured once into a hipGraph per precision; recaptured
// when the blob or body counts baked into its kernel arguments change.
int step_graph(fcm_handle h, int precision, hipGraphExec_t* out) {
  auto& gc = h->graph[precision == FCM_PREC_FP32 ? 1 : 0];
  const long long key[3] = {h->bd.nb, h->NB, h->nbpb};
  if (gc.exec && std::memcmp(gc.key, key, sizeof key) == 0) {
    *out = gc.exec;
    return FCM_OK;
  }
  if (gc.exec) {
    (void)hipGraphExecDestroy(gc.exec);
    gc.exec = nullptr;
  }
  HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  const int rc = launch_arnoldi_step(h, precision);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(h->st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  HIPCHK(ec);
  const hipError_t ei = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIPCHK(ei);
  std::memcpy(gc.key, key, sizeof key);
  *out = gc.exec;
  return FCM_OK;
}

// Unrestarted right-preconditioned GMRES, x0 = 0 (gmres_solver.cpp:148-306)
// on device vectors b -> x.  The Arnoldi loop, the Givens least squares of
// the reference and the stopping test run on the device (fcm_gmres.hip); the
// host queues steps in chunks and reads the status word of the chunk before
// the one it has just queued, so the device never waits for the host.
int gmres_device(fcm_handle h, const double* b, double* x, double tol, int max_iter_in, int precision,
                 InnerResult& out, double* hist) {
  const size_t n = h->nsys();
  const int max_iter = (int)std::min<size_t>((size_t)std::min(max_iter_in, h->cfg.max_krylov), n);
  if (max_iter < 1) return fail(FCM_ERR_ARG, "max_iter must be >= 1");
  if (precision == FCM_PREC_FP32 && !h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
  out = InnerResult{};
  HIPCHK(dots(b, 0, 1, b, n, h->kw, h->gm.red, h->st));
  HIPCHK(gm_launch_start(h->gm, tol, max_iter, b, h->Q, n, h->st));
  hipGraphExec_t ge = nullptr;
  if (graphs_enabled())
    if (int rc = step_graph(h, precision, &ge)) return rc;
  constexpr int CHUNK = 4;
  constexpr size_t CTL = 16;  // doubles per control snapshot
  int launched = 0, slot = 0, pending = -1;
  while (launched < max_iter) {
    const int c = std::min(CHUNK, max_iter - launched);
    for (int i = 0; i < c; ++i) {
      if (ge) {
        HIPCHK(hipGraphLaunch(ge, h->st));
      } else if (int rc = launch_arnoldi_step(h, precision)) {
        return rc;
      }
    }
    launched += c;
    HIPCHK(hipMemcpyAsync(h->hctl + slot * CTL, h->gm_ctl, CTL * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipEventRecord(h->pev[slot], h->st));
    if (pending >= 0) {
      HIPCHK(hipEventSynchronize(h->pev[pending]));
      const int* ic = reinterpret_cast<const int*>(h->hctl + pending * CTL + 8);
      if (ic[GM_STATUS] != 0) break;
    }
    pending = slot;
    slot ^= 1;
  }
  HIPCHK(gm_launch_finish(h->gm, h->Q, h->ldq, h->tmp, n, h->st));
  BodyData bdy = h->body();
  HIPCHK(apply_pinv(bdy, h->tmp, x, nullptr, h->st));
  double* snap = h->hctl + 2 * CTL;
  HIPCHK(hipMemcpyAsync(snap, h->gm_ctl, CTL * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const int* ic = reinterpret_cast<const int*>(snap + 8);
  if (ic[GM_STATUS] == 0) return fail(FCM_ERR_STATE, "GMRES control did not terminate");
  out.status = ic[GM_STATUS] - 1;
  out.iters = out.status == FCM_GMRES_ZERO_RHS ? 0 : ic[GM_NCOLS];
  out.rel = out.status == FCM_GMRES_ZERO_RHS ? 0.0 : snap[4];
  if (hist && out.iters > 0)
    HIPCHK(hipMemcpy(hist, h->gm.hist, (size_t)out.iters * sizeof(double), hipMemcpyDeviceToHost));
  return FCM_OK;
}

int solve_dev(fcm_handle h, const gmres_opts* o, gmres_report* rep, double* hist) {
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  const size_t n = h->nsys();
  gmres_report r{};
  HIPCHK(hipEventRecord(h->ev0, h->st));
  const double tol = o ? o->tol : 1e-8;
  const int max_iter = o ? o->max_iter : h->cfg.max_krylov;
  const int prec = o ? o->precision : FCM_PREC_FP64;
  InnerResult ir;
  int rc;
  if (prec != FCM_PREC_FP32) {
    rc = gmres_device(h, h->rhs, h->sol, tol, max_iter, FCM_PREC_FP64, ir, hist);
    if (rc) return rc;
    r.status = ir.status;
    r.iterations = ir.iters;
    r.rel_residual = ir.rel;
    r.true_rel_residual = -1.0;
  } else {
    // FP32-bulk inner solves, FP64 outer residual (iterative refinement)
    const double bn = dev_norm(h, h->rhs, n, rc);
    if (rc) return rc;
    if (bn == 0.0) {
      HIPCHK(axpby(0.0, h->rhs, 0.0, h->sol, n, h->st));
      r.status = FCM_GMRES_ZERO_RHS;
    } else {
      const int max_ref = (o && o->max_refine > 0) ? o->max_refine : 8;
      HIPCHK(axpby(0.0, h->rhs, 0.0, h->xacc, n, h->st));
      HIPCHK(axpby(1.0, h->rhs, 0.0, h->res, n, h->st));
      double rn = bn;
      r.status = FCM_GMRES_MAX_ITER;
      for (int k = 0; k < max_ref; ++k) {
        double itol = (o && o->inner_tol > 0) ? o->inner_tol : std::max(0.5 * tol * bn / rn, 1e-5);
        itol = std::min(itol, 0.5);
        rc = gmres_device(h, h->res, h->sol, itol, max_iter - r.iterations, FCM_PREC_FP32, ir,
                        hist ? hist + r.iterations : nullptr);
        if (rc) return rc;
        r.iterations += ir.iters;
        r.refinements = k + 1;
        r.rel_residual = ir.rel;
        HIPCHK(axpby(1.0, h->sol, 1.0, h->xacc, n, h->st));
        rc = saddle_dev(h, h->xacc, h->tmp, FCM_PREC_FP64);
        if (rc) return rc;
        HIPCHK(axpby(1.0, h->rhs, 0.0, h->res, n, h->st));
        HIPCHK(axpby(-1.0, h->tmp, 1.0, h->res, n, h->st));
        rn = dev_norm(h, h->res, n, rc);
        if (rc) return rc;
        r.true_rel_residual = rn / bn;
        if (rn / bn < tol) { r.status = FCM_GMRES_CONVERGED; break; }
        if (r.iterations >= max_iter) break;
      }
      HIPCHK(axpby(1.0, h->xacc, 0.0, h->sol, n, h->st));
    }
  }
  HIPCHK(hipEventRecord(h->ev1, h->st));
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  r.seconds = ms * 1e-3;
  h->have_sol = true;
  if (rep) *rep = r;
  return FCM_OK;
}

int setup_blobs_device(fcm_handle h, int nb) {
  h->bd.nb = nb;
  h->bd.Y = h->dY;
  HIPCHK(setup_gaussians(h->g, h->bb, h->bd, h->cfg.a, h->st));
  h->have_blobs = true;
  return FCM_OK;
}

// copy a real-space field between device layout (row pitch 2*cpitch reals)
// and the reference layout (row pitch pad = 2*(nz/2+1))
template <typename T>
int field_to_host(fcm_handle h, const c2<T>* d, double* out) {
  const GridGeom& g = h->g;
  const size_t rows = (size_t)g.nx * g.ny;
  const int pad = 2 * g.nzc;
  std::vector<T> buf(rows * 2 * g.cpitch);
  HIPCHK(hipMemcpyAsync(buf.data(), d, buf.size() * sizeof(T), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (size_t r = 0; r < rows; ++r)
    for (int k = 0; k < pad; ++k) out[r * pad + k] = (k < g.nz) ? (double)buf[r * 2 * g.cpitch + k] : 0.0;
  return FCM_OK;
}

template <typename T>
int field_from_host(fcm_handle h, c2<T>* d, const double* in, bool ref_layout) {
  const GridGeom& g = h->g;
  const size_t rows = (size_t)g.nx * g.ny;
  const int pad = 2 * g.nzc;
  std::vector<T> buf(rows * 2 * g.cpitch, T(0));
  for (size_t r = 0; r < rows; ++r)
    for (int k = 0; k < g.nz; ++k) buf[r * 2 * g.cpitch + k] = (T)in[ref_layout ? r * pad + k : r * g.nz + k];
  HIPCHK(hipMemcpyAsync(d, buf.data(), buf.size() * sizeof(T), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return FCM_OK;
}

template <typename T>
int kspace_to_host(fcm_handle h, const c2<T>* d, double* out) {
  const GridGeom& g = h->g;
  std::vector<c2<T>> buf(g.field_words());
  HIPCHK(hipMemcpyAsync(buf.data(), d, buf.size() * sizeof(c2<T>), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  const int pad = 2 * g.nzc;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i)
      for (int k = 0; k < g.nzc; ++k) {
        const c2<T> v = buf[((size_t)i * g.ny + j) * g.cpitch + k];
        out[((size_t)j * g.nx + i) * pad + 2 * k] = (double)v.x;
        out[((size_t)j * g.nx + i) * pad + 2 * k + 1] = (double)v.y;
      }
  return FCM_OK;
}

template <typename T>
int kspace_from_host(fcm_handle h, c2<T>* d, const double* in) {
  const GridGeom& g = h->g;
  std::vector<c2<T>> buf(g.field_words(), mk<T>(T(0), T(0)));
  const int pad = 2 * g.nzc;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i)
      for (int k = 0; k < g.nzc; ++k)
        buf[((size_t)i * g.ny + j) * g.cpitch + k] =
            mk<T>((T)in[((size_t)j * g.nx + i) * pad + 2 * k], (T)in[((size_t)j * g.nx + i) * pad + 2 * k + 1]);
  HIPCHK(hipMemcpyAsync(d, buf.data(), buf.size() * sizeof(c2<T>), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return FCM_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* fcm_last_error(void) { return g_err.c_str(); }

int fcm_query_footprint(const fcm_config* cfg, size_t* bytes) {
  std::string why;
  if (!cfg || !bytes) return fail(FCM_ERR_ARG, "null argument");
  if (!valid_cfg(cfg, why)) return fail(FCM_ERR_ARG, why);
  *bytes = sizes_of(cfg).total;
  return FCM_OK;
}

int fcm_create(const fcm_config* cfg, fcm_handle* out) {
  std::string why;
  if (!cfg || !out) return fail(FCM_ERR_ARG, "null argument");
  if (!valid_cfg(cfg, why)) return fail(FCM_ERR_ARG, why);
  fcm_handle h = new fcm_handle_s();
  h->cfg = *cfg;
  h->g = make_geom(cfg);
  BlobBins& bb = h->bb;
  bb.bx = largest_divisor_le(cfg->nx, 4);
  bb.by = largest_divisor_le(cfg->ny, 8);
  bb.bz = largest_divisor_le(cfg->nz, 8);
  bb.nbx = cfg->nx / bb.bx;
  bb.nby = cfg->ny / bb.by;
  bb.nbz = cfg->nz / bb.bz;
  bb.sx = (bb.nbx % 2 == 0 && 2 * bb.bx <= 8) ? 2 : 1;
  bb.sy = (bb.nby % 2 == 0 && 2 * bb.by <= 16) ? 2 : 1;
  bb.sz = (bb.nbz % 2 == 0 && 2 * bb.bz <= 16) ? 2 : 1;
  while (bb.sy * bb.by * bb.sz * bb.bz > 256) {
    if (bb.sz > 1) bb.sz = 1; else bb.sy = 1;
  }
  const double LX = cfg->dx * cfg->nx, LY = cfg->dx * cfg->ny, LZ = cfg->dx * cfg->nz;
  h->pp.facx = 2.0 * PI / LX;
  h->pp.facy = 2.0 * PI / LY;
  h->pp.facz = 2.0 * PI / LZ;
  h->pp.inv_muN = 1.0 / (cfg->mu * (double)cfg->nx * cfg->ny * cfg->nz);
  h->pp.nx = cfg->nx; h->pp.ny = cfg->ny; h->pp.nz = cfg->nz;

  auto cleanup_fail = [&](int code, const std::string& m) {
    fcm_destroy(h);
    return fail(code, m);
  };
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) return cleanup_fail(FCM_ERR_HIP, "stream");
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
    return cleanup_fail(FCM_ERR_HIP, "event");
  const Sizes s = sizes_of(cfg);
  hipError_t e = hipSuccess;
  const size_t fw = h->g.field_words();
  for (int i = 0; i < 3 && e == hipSuccess; ++i) h->f64[i] = dalloc<c2<double>>(h, fw, e);
  if (cfg->fp32)
    for (int i = 0; i < 3 && e == hipSuccess; ++i) h->f32[i] = dalloc<c2<float>>(h, fw, e);
  if (cfg->random_field)
    for (int i = 0; i < 6 && e == hipSuccess; ++i) h->stress[i] = dalloc<c2<double>>(h, fw, e);
  if (e != hipSuccess) return cleanup_fail(FCM_ERR_NOMEM, "grid allocation failed");
  if (upload_twiddles<double>(h, &h->twbuf64, h->tw64)) return cleanup_fail(FCM_ERR_NOMEM, "twiddles");
  if (cfg->fp32 && upload_twiddles<float>(h, &h->twbuf32, h->tw32)) return cleanup_fail(FCM_ERR_NOMEM, "twiddles");

  const int nb = cfg->max_blobs, nbins = bb.nbins();
  h->dY = dalloc<double>(h, 3 * (size_t)nb, e);
  if (e == hipSuccess) h->bd.bin_of = dalloc<int>(h, nb, e);
  if (e == hipSuccess) h->bd.bin_count = dalloc<int>(h, nbins, e);
  if (e == hipSuccess) h->bd.bin_start = dalloc<int>(h, nbins + 1, e);
  if (e == hipSuccess) h->bd.bin_cursor = dalloc<int>(h, nbins, e);
  if (e == hipSuccess) h->bd.sorted = dalloc<int>(h, nb, e);
  if (e == hipSuccess) h->bd.rec = dalloc<int4>(h, nb, e);
  if (e == hipSuccess) h->bd.w64 = dalloc<double>(h, 3 * (size_t)nb * cfg->ng, e);
  if (e == hipSuccess) h->bd.w32 = dalloc<float>(h, 3 * (size_t)nb * cfg->ng, e);
  {
    const int nt = spread_tiles(h->g);
    if (e == hipSuccess) h->bd.tile_count = dalloc<int>(h, nt, e);
    if (e == hipSuccess) h->bd.tile_start = dalloc<int>(h, nt + 1, e);
    if (e == hipSuccess) h->bd.tile_cursor = dalloc<int>(h, nt, e);
    if (e == hipSuccess) h->bd.tile_list = dalloc<int>(h, 8 * (size_t)nb, e);
  }
  h->bd.ng = cfg->ng;
  const int n3b = 3 * cfg->max_blobs_per_body;
  if (e == hipSuccess) h->d_ref = dalloc<double>(h, n3b, e);
  if (e == hipSuccess) h->d_Minv = dalloc<double>(h, (size_t)n3b * n3b, e);
  if (e == hipSuccess) h->d_N = dalloc<double>(h, 36, e);
  if (e == hipSuccess) h->d_x = dalloc<double>(h, 3 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_q = dalloc<double>(h, 4 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_R = dalloc<double>(h, 9 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_F6 = dalloc<double>(h, 6 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_dpos = dalloc<double>(h, 3 * (size_t)nb, e);
  h->nmax = s.nmax;
  h->ldq = s.ldq;
  if (e == hipSuccess) h->Q = dalloc<double>(h, (size_t)(cfg->max_krylov + 1) * s.ldq, e);
  if (e == hipSuccess) h->z = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->rhs = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->sol = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->xacc = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->res = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->tmp = dalloc<double>(h, s.ldq, e);
  if (e == hipSuccess) h->Vrand = dalloc<double>(h, 3 * (size_t)nb, e);
  if (e == hipSuccess) h->Vtmp = dalloc<double>(h, 3 * (size_t)nb, e);
  h->kw.nblk = krylov_blocks(s.nmax);
  if (e == hipSuccess) h->kw.partial = dalloc<double>(h, (size_t)h->kw.nblk * (2 * (size_t)cfg->max_krylov + 6), e);
  if (e == hipSuccess) h->red = dalloc<double>(h, 2 * (size_t)cfg->max_krylov + 16, e);
  if (e == hipSuccess) h->coef = dalloc<double>(h, 2 * (size_t)cfg->max_krylov + 16, e);
  {
    const int M = cfg->max_krylov;
    const size_t m2 = (size_t)(M + 2) * (M + 2);
    GmDev& G = h->gm;
    G.M = M;
    if (e == hipSuccess) h->gm_ctl = dalloc<double>(h, 16, e);
    G.dctl = h->gm_ctl;
    G.ictl = reinterpret_cast<int*>(h->gm_ctl + 8);
    if (e == hipSuccess) G.Hraw = dalloc<double>(h, m2, e);
    if (e == hipSuccess) G.R = dalloc<double>(h, m2, e);
    if (e == hipSuccess) G.CS = dalloc<double>(h, 8 * (size_t)(M + 2), e);
    if (e == hipSuccess) {
      G.SN = G.CS + (M + 2);
      G.beta = G.SN + (M + 2);
      G.hprov = G.beta + (M + 2);
      G.hnew = G.hprov + (M + 2);
      G.hist = G.hnew + (M + 2);
      G.y = G.hist + (M + 2);
    }
    G.coef = h->coef;
    G.red = h->red;
    if (e == hipSuccess) h->w = dalloc<double>(h, s.ldq, e);
  }
  if (e != hipSuccess) return cleanup_fail(FCM_ERR_NOMEM, "device allocation failed");
  if (hipHostMalloc((void**)&h->hctl, 3 * 16 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return cleanup_fail(FCM_ERR_NOMEM, "pinned host allocation failed");
  if (hipEventCreateWithFlags(&h->pev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&h->pev[1], hipEventDisableTiming) != hipSuccess)
    return cleanup_fail(FCM_ERR_HIP, "event");
  for (int i = 0; i < 3; ++i)
    if (zero_field64(h, h->f64[i])) return cleanup_fail(FCM_ERR_HIP, "memset");
  if (hipMemsetAsync(h->Vrand, 0, 3 * (size_t)nb * sizeof(double), h->st) != hipSuccess ||
      hipMemsetAsync(h->d_F6, 0, 6 * (size_t)cfg->max_bodies * sizeof(double), h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return cleanup_fail(FCM_ERR_HIP, "memset");
  *out = h;
  return FCM_OK;
}

int fcm_destroy(fcm_handle h) {
  if (!h) return FCM_OK;
  if (h->st) (void)hipStreamSynchronize(h->st);
  for (auto& gc : h->graph)
    if (gc.exec) (void)hipGraphExecDestroy(gc.exec);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->hctl) (void)hipHostFree(h->hctl);
  for (auto& ev : h->pev)
    if (ev) (void)hipEventDestroy(ev);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return FCM_OK;
}

int fcm_sync(fcm_handle h) {
  if (int rc = check(h)) return rc;
  return sync_stream(h);
}

int fcm_set_up_gaussians(fcm_handle h, const double* Y, int nb) {
  if (int rc = check(h)) return rc;
  if (!Y || nb < 0) return fail(FCM_ERR_ARG, "bad positions");
  if (nb > h->cfg.max_blobs) return fail(FCM_ERR_CAPACITY, "more blobs than max_blobs");
  HIPCHK(hipMemcpyAsync(h->dY, Y, 3 * (size_t)nb * sizeof(double), hipMemcpyHostToDevice, h->st));
  int rc = setup_blobs_device(h, nb);
  if (rc) return rc;
  return sync_stream(h);
}

int fcm_spread(fcm_handle h, const double* F, int precision) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  HIPCHK(hipMemcpyAsync(h->Vtmp, F, h->n3() * sizeof(double), hipMemcpyHostToDevice, h->st));
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    HIPCHK(spread<float>(h->g, h->bb, h->bd, h->Vtmp, h->fp32(), h->st));
  } else {
    HIPCHK(spread<double>(h->g, h->bb, h->bd, h->Vtmp, h->fp64(), h->st));
  }
  return sync_stream(h);
}

int fcm_stokes(fcm_handle h, int precision) {
  if (int rc = check(h)) return rc;
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    HIPCHK(stokes_solve<float>(h->g, h->fp32(), h->pp, h->tw32, h->st));
  } else {
    HIPCHK(stokes_solve<double>(h->g, h->fp64(), h->pp, h->tw64, h->st));
  }
  return sync_stream(h);
}

int fcm_interp(fcm_handle h, double* V, int precision) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  KTerm kt;
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    HIPCHK(interp<float>(h->g, h->bb, h->bd, h->fp32(), h->Vtmp, kt, h->st));
  } else {
    HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fp64(), h->Vtmp, kt, h->st));
  }
  HIPCHK(hipMemcpyAsync(V, h->Vtmp, h->n3() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_mobility(fcm_handle h, const double* F, double* V, int precision) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  const size_t n3 = h->n3();
  HIPCHK(hipMemcpyAsync(h->tmp, F, n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
  int rc = mobility_dev(h, h->tmp, h->Vtmp, precision, KTerm{});
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(V, h->Vtmp, n3 * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_random_field(fcm_handle h, const double* stress6, uint64_t seed) {
  if (int rc = check(h)) return rc;
  if (!h->stress[0]) return fail(FCM_ERR_STATE, "random field not allocated (fcm_config.random_field = 0)");
  const fcm_config& c = h->cfg;
  // fcm_fluid_solver.cpp:169-170
  const double var1 = std::sqrt(2.0 * c.kT * c.mu / (c.dx * c.dx * c.dx * c.dt));
  const double var2 = std::sqrt(2.0) * var1;
  const double scale[6] = {var2, var2, var2, var1, var1, var1};
  if (stress6) {
    const size_t nn = (size_t)c.nx * c.ny * c.nz;
    for (int i = 0; i < 6; ++i) {
      int rc = field_from_host<double>(h, h->stress[i], stress6 + i * nn, false);
      if (rc) return rc;
    }
  }
  HIPCHK(random_velocity<double>(h->g, h->fpstress(), stress6 == nullptr, (unsigned long long)seed, scale, h->pp,
                                 h->tw64, h->st));
  h->have_urand = true;
  return sync_stream(h);
}

int fcm_random_interp(fcm_handle h, double* V) {
  if (int rc = check(h)) return rc;
  if (!h->have_urand) return fail(FCM_ERR_STATE, "no fluctuating velocity field");
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fprand(), h->Vrand, KTerm{}, h->st));
  if (V) HIPCHK(hipMemcpyAsync(V, h->Vrand, h->n3() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_get_field(fcm_handle h, int which, int precision, double* out) {
  if (int rc = check(h)) return rc;
  if (!out || which < 0 || which > 5) return fail(FCM_ERR_ARG, "bad field");
  if (which >= 3) {
    if (!h->have_urand) return fail(FCM_ERR_STATE, "no fluctuating velocity field");
    return field_to_host<double>(h, h->stress[which - 3], out);
  }
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    return field_to_host<float>(h, h->f32[which], out);
  }
  return field_to_host<double>(h, h->f64[which], out);
}

int fcm_set_field(fcm_handle h, int which, int precision, const double* in) {
  if (int rc = check(h)) return rc;
  if (!in || which < 0 || which > 2) return fail(FCM_ERR_ARG, "bad field");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    return field_from_host<float>(h, h->f32[which], in, true);
  }
  return field_from_host<double>(h, h->f64[which], in, true);
}

int fft_transform(fcm_handle h, int which, int precision) {
  if (int rc = check(h)) return rc;
  if (which < 0 || which > 2) return fail(FCM_ERR_ARG, "bad field");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    FieldPtrs<float> p{};
    p.p[0] = h->f32[which];
    HIPCHK(fft3d_forward<float>(h->g, p, 1, h->tw32, h->st));
  } else {
    FieldPtrs<double> p{};
    p.p[0] = h->f64[which];
    HIPCHK(fft3d_forward<double>(h->g, p, 1, h->tw64, h->st));
  }
  return sync_stream(h);
}

int fft_inverse_transform(fcm_handle h, int which, int precision) {
  if (int rc = check(h)) return rc;
  if (which < 0 || which > 2) return fail(FCM_ERR_ARG, "bad field");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    FieldPtrs<float> p{};
    p.p[0] = h->f32[which];
    HIPCHK(fft3d_inverse<float>(h->g, p, 1, h->tw32, h->st));
  } else {
    FieldPtrs<double> p{};
    p.p[0] = h->f64[which];
    HIPCHK(fft3d_inverse<double>(h->g, p, 1, h->tw64, h->st));
  }
  return sync_stream(h);
}

int fcm_get_kspace(fcm_handle h, int which, int precision, double* out) {
  if (int rc = check(h)) return rc;
  if (!out || which < 0 || which > 2) return fail(FCM_ERR_ARG, "bad field");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    return kspace_to_host<float>(h, h->f32[which], out);
  }
  return kspace_to_host<double>(h, h->f64[which], out);
}

int fcm_set_kspace(fcm_handle h, int which, int precision, const double* in) {
  if (int rc = check(h)) return rc;
  if (!in || which < 0 || which > 2) return fail(FCM_ERR_ARG, "bad field");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
    return kspace_from_host<float>(h, h->f32[which], in);
  }
  return kspace_from_host<double>(h, h->f64[which], in);
}

int rb_set_geometry(fcm_handle h, int nbpb, const double* blob_ref) {
  if (int rc = check(h)) return rc;
  if (!blob_ref || nbpb < 1) return fail(FCM_ERR_ARG, "bad geometry");
  if (nbpb > h->cfg.max_blobs_per_body || nbpb > h->cfg.max_blobs)
    return fail(FCM_ERR_CAPACITY, "more blobs per body than max_blobs_per_body");
  h->nbpb = nbpb;
  h->h_ref.assign(blob_ref, blob_ref + 3 * (size_t)nbpb);
  HIPCHK(hipMemcpyAsync(h->d_ref, blob_ref, 3 * (size_t)nbpb * sizeof(double), hipMemcpyHostToDevice, h->st));
  h->have_geometry = true;
  h->have_bodies = false;
  h->have_ref = false;
  return sync_stream(h);
}

int rb_set_bodies(fcm_handle h, int NB, const double* x, const double* q) {
  if (int rc = check(h)) return rc;
  if (!h->have_geometry) return fail(FCM_ERR_STATE, "rb_set_geometry first");
  if (!x || !q || NB < 1) return fail(FCM_ERR_ARG, "bad bodies");
  if (NB > h->cfg.max_bodies || (size_t)NB * h->nbpb > (size_t)h->cfg.max_blobs)
    return fail(FCM_ERR_CAPACITY, "more bodies than the capacities allow");
  h->NB = NB;
  HIPCHK(hipMemcpyAsync(h->d_x, x, 3 * (size_t)NB * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_q, q, 4 * (size_t)NB * sizeof(double), hipMemcpyHostToDevice, h->st));
  BodyData b = h->body();
  HIPCHK(bodies_update(h->g, b, h->st));
  int rc = setup_blobs_device(h, NB * h->nbpb);
  if (rc) return rc;
  h->have_bodies = true;
  h->have_rhs = false;
  return sync_stream(h);
}

int rb_set_body_forces(fcm_handle h, const double* F6) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  if (!F6) return fail(FCM_ERR_ARG, "null forces");
  HIPCHK(hipMemcpyAsync(h->d_F6, F6, 6 * (size_t)h->NB * sizeof(double), hipMemcpyHostToDevice, h->st));
  h->have_forces = true;
  return sync_stream(h);
}

int rb_get_blob_positions(fcm_handle h, double* Y) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs || !Y) return fail(FCM_ERR_STATE, "no blob positions");
  HIPCHK(hipMemcpyAsync(Y, h->dY, h->n3() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int rb_mul_K(fcm_handle h, const double* in6, double* out3) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  HIPCHK(hipMemcpyAsync(h->tmp, in6, 6 * (size_t)h->NB * sizeof(double), hipMemcpyHostToDevice, h->st));
  BodyData b = h->body();
  HIPCHK(mul_K(b, h->tmp, h->Vtmp, h->st));
  HIPCHK(hipMemcpyAsync(out3, h->Vtmp, h->n3() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int rb_mul_KT(fcm_handle h, const double* in3, double* out6) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  HIPCHK(hipMemcpyAsync(h->Vtmp, in3, h->n3() * sizeof(double), hipMemcpyHostToDevice, h->st));
  BodyData b = h->body();
  HIPCHK(mul_KT(b, h->Vtmp, h->tmp, h->st));
  HIPCHK(hipMemcpyAsync(out6, h->tmp, 6 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_set_ref_matrices(fcm_handle h, const double* Minv, const double* N) {
  if (int rc = check(h)) return rc;
  if (!h->have_geometry) return fail(FCM_ERR_STATE, "rb_set_geometry first");
  if (!Minv || !N) return fail(FCM_ERR_ARG, "null matrices");
  const size_t n3 = 3 * (size_t)h->nbpb;
  HIPCHK(hipMemcpyAsync(h->d_Minv, Minv, n3 * n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_N, N, 36 * sizeof(double), hipMemcpyHostToDevice, h->st));
  h->have_ref = true;
  return sync_stream(h);
}

int fcm_ref_matrices(fcm_handle h, double* Msingle, double* Minv, double* N) {
  if (int rc = check(h)) return rc;
  if (!h->have_geometry) return fail(FCM_ERR_STATE, "rb_set_geometry first");
  const fcm_config& c = h->cfg;
  const int nbpb = h->nbpb, n3 = 3 * nbpb;
  // body 0 at the box centre, identity orientation (collection_of_rigid_bodies.cpp:141-149)
  const double L[3] = {c.dx * c.nx, c.dx * c.ny, c.dx * c.nz};
  std::vector<double> Y(n3);
  for (int m = 0; m < nbpb; ++m)
    for (int d = 0; d < 3; ++d) {
      const double p = 0.5 * L[d] + h->h_ref[3 * m + d];
      Y[3 * m + d] = p - L[d] * std::floor(p / L[d]);
    }
  HIPCHK(hipMemcpyAsync(h->dY, Y.data(), n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
  int rc = setup_blobs_device(h, nbpb);
  if (rc) return rc;
  // 3 nbpb unit-force solves (fcm_fluid_solver.cpp:853-864)
  std::vector<double> M((size_t)n3 * n3), e(n3, 0.0), col(n3);
  for (int n = 0; n < n3; ++n) {
    std::fill(e.begin(), e.end(), 0.0);
    e[n] = 1.0;
    HIPCHK(hipMemcpyAsync(h->tmp, e.data(), n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
    rc = mobility_dev(h, h->tmp, h->Vtmp, FCM_PREC_FP64, KTerm{});
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(col.data(), h->Vtmp, n3 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int i = 0; i < n3; ++i) M[(size_t)i * n3 + n] = col[i];
  }
  std::vector<double> Mi((size_t)n3 * n3), Nr(36);
  if (dense::inv_spd(n3, M.data(), Mi.data()) != 0) return fail(FCM_ERR_SINGULAR, "single-body mobility not SPD");
  dense::ref_N(nbpb, h->h_ref.data(), Mi.data(), Nr.data());
  if (Msingle) std::memcpy(Msingle, M.data(), M.size() * sizeof(double));
  if (Minv) std::memcpy(Minv, Mi.data(), Mi.size() * sizeof(double));
  if (N) std::memcpy(N, Nr.data(), 36 * sizeof(double));
  rc = fcm_set_ref_matrices(h, Mi.data(), Nr.data());
  if (rc) return rc;
  // restore the blob state of the current bodies
  h->have_blobs = false;
  if (h->have_bodies) {
    BodyData b = h->body();
    HIPCHK(bodies_update(h->g, b, h->st));
    rc = setup_blobs_device(h, h->NB * h->nbpb);
    if (rc) return rc;
  }
  return sync_stream(h);
}

int gmres_assemble_rhs(fcm_handle h, int use_random) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  HIPCHK(rhs_assemble(use_random ? h->Vrand : nullptr, h->have_forces ? h->d_F6 : nullptr, h->rhs, h->n3(),
                      6 * (size_t)h->NB, h->st));
  h->have_rhs = true;
  return sync_stream(h);
}

int gmres_saddle_matvec(fcm_handle h, const double* in, double* out, int precision) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  const size_t n = h->nsys();
  HIPCHK(hipMemcpyAsync(h->xacc, in, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  int rc = saddle_dev(h, h->xacc, h->tmp, precision);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, h->tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int gmres_apply_Pinv(fcm_handle h, const double* in, double* out) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  const size_t n = h->nsys();
  HIPCHK(hipMemcpyAsync(h->xacc, in, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  BodyData b = h->body();
  HIPCHK(apply_pinv(b, h->xacc, h->tmp, nullptr, h->st));
  HIPCHK(hipMemcpyAsync(out, h->tmp, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int gmres_solve(fcm_handle h, const double* rhs, double* soln, const gmres_opts* opts, gmres_report* rep,
                double* resid_hist) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  const size_t n = h->nsys();
  if (rhs) {
    HIPCHK(hipMemcpyAsync(h->rhs, rhs, n * sizeof(double), hipMemcpyHostToDevice, h->st));
    h->have_rhs = true;
  }
  if (!h->have_rhs) return fail(FCM_ERR_STATE, "no right-hand side");
  int rc = solve_dev(h, opts, rep, resid_hist);
  if (rc) return rc;
  if (soln) {
    HIPCHK(hipMemcpyAsync(soln, h->sol, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    return sync_stream(h);
  }
  return FCM_OK;
}

int gmres_get_solution(fcm_handle h, double* soln) {
  if (int rc = check(h)) return rc;
  if (!h->have_sol || !soln) return fail(FCM_ERR_STATE, "no solution");
  HIPCHK(hipMemcpyAsync(soln, h->sol, h->nsys() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_profile_matvec(fcm_handle h, int precision, int reps, double* ms_per_pass) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies || !ms_per_pass || reps < 1) return fail(FCM_ERR_STATE, "bodies must be set; reps >= 1");
  if (precision == FCM_PREC_FP32 && !h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
  const size_t n = h->nsys(), n3 = h->n3();
  std::vector<double> v(n);
  uint64_t st = 12345;
  for (size_t i = 0; i < n; ++i) {  // a fixed pseudo-random input vector
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    v[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  HIPCHK(hipMemcpyAsync(h->xacc, v.data(), n * sizeof(double), hipMemcpyHostToDevice, h->st));
  hipEvent_t ev[9];
  for (int i = 0; i < 9; ++i) HIPCHK(hipEventCreate(&ev[i]));
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  KTerm kt;
  kt.zV = h->xacc + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  BodyData b = h->body();
  int rc = FCM_OK;
  for (int r = 0; r < reps && rc == FCM_OK; ++r) {
    HIPCHK(hipEventRecord(ev[0], h->st));
    if (precision == FCM_PREC_FP32) {
      HIPCHK(spread<float>(h->g, h->bb, h->bd, h->xacc, h->fp32(), h->st));
      HIPCHK(hipEventRecord(ev[1], h->st));
      for (int p = 0; p < 5; ++p) {
        HIPCHK(stokes_pass<float>(h->g, h->fp32(), p, h->pp, h->tw32, h->st));
        HIPCHK(hipEventRecord(ev[2 + p], h->st));
      }
      HIPCHK(interp<float>(h->g, h->bb, h->bd, h->fp32(), h->tmp, kt, h->st));
    } else {
      HIPCHK(spread<double>(h->g, h->bb, h->bd, h->xacc, h->fp64(), h->st));
      HIPCHK(hipEventRecord(ev[1], h->st));
      for (int p = 0; p < 5; ++p) {
        HIPCHK(stokes_pass<double>(h->g, h->fp64(), p, h->pp, h->tw64, h->st));
        HIPCHK(hipEventRecord(ev[2 + p], h->st));
      }
      HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fp64(), h->tmp, kt, h->st));
    }
    HIPCHK(hipEventRecord(ev[7], h->st));
    HIPCHK(mul_KT(b, h->xacc, h->tmp + n3, h->st));
    HIPCHK(hipEventRecord(ev[8], h->st));
    HIPCHK(hipEventSynchronize(ev[8]));
    for (int i = 0; i < 8; ++i) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      acc[i] += ms;
    }
  }
  for (int i = 0; i < 9; ++i) (void)hipEventDestroy(ev[i]);
  for (int i = 0; i < 8; ++i) ms_per_pass[i] = acc[i] / reps;
  return rc;
}

int fcm_mobility_solve(fcm_handle h, uint64_t seed, const gmres_opts* opts, gmres_report* rep) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  const bool fluct = h->cfg.kT > 0.0 && h->stress[0] != nullptr;
  if (fluct) {
    const fcm_config& c = h->cfg;
    const double var1 = std::sqrt(2.0 * c.kT * c.mu / (c.dx * c.dx * c.dx * c.dt));
    const double var2 = std::sqrt(2.0) * var1;
    const double scale[6] = {var2, var2, var2, var1, var1, var1};
    HIPCHK(random_velocity<double>(h->g, h->fpstress(), true, (unsigned long long)seed, scale, h->pp, h->tw64, h->st));
    h->have_urand = true;
    HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fprand(), h->Vrand, KTerm{}, h->st));
  }
  HIPCHK(rhs_assemble(fluct ? h->Vrand : nullptr, h->have_forces ? h->d_F6 : nullptr, h->rhs, h->n3(),
                      6 * (size_t)h->NB, h->st));
  h->have_rhs = true;
  return solve_dev(h, opts, rep, nullptr);
}

}  // extern "C"
