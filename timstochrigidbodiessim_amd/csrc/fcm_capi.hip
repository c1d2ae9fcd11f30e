// fcm_capi.hip -- host side of libfcm_hip.so: the C ABI of include/fcm_hip.h.
//
// A handle owns one periodic grid, its blobs and bodies, the GMRES basis and
// a HIP stream (default priority).  Host code here only sequences kernels,
// moves the small per-iteration scalars of GMRES and runs the Givens/least-
// squares algebra of gmres_solver.cpp:212-284 on those scalars.
#include "fcm_handle.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "fcm_dense.hpp"

using namespace fcm;

namespace fcm_host {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace fcm_host

using namespace fcm_host;

namespace {

constexpr double PI = 3.14159265358979323846264338327950288;

int largest_divisor_le(int n, int cap) {
  for (int d = std::min(n, cap); d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

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
  if (c->ng < 1 || c->ng > 20) { why = "ng must be in 1..20"; return false; }
  if (c->slip_channel != 0 && c->slip_channel != 1) { why = "slip_channel must be 0 or 1"; return false; }
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
  g.cpitch = round_up(g.nzc, 16);  // rows start on 128-byte lines in FP32 and FP64
  g.dx = c->dx;
  g.channel = c->slip_channel ? 1 : 0;
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
  t += (size_t)nb * (3 * 8 + 4 + 4 + 16 + 3 * c->ng * 12 + 8 * 4 + 16) + (3 * nbins + 2) * 4;
  if (c->slip_channel) t += (size_t)nb * 3 * c->ng * 12;  // truncated interpolation weights
  {
    GridGeom g2 = make_geom(c);
    t += 3 * (size_t)(spread_tiles(g2) + 1) * 4;
    BlobBins bb2;
    bb2.bx = largest_divisor_le(c->nx, 4);
    bb2.by = largest_divisor_le(c->ny, 8);
    bb2.bz = largest_divisor_le(c->nz, 8);
    // the column buffers are sized for the narrower of the two patch widths
    if (col_patch_cx(g2, bb2, c->ng) > 0)
      t += (3 * (size_t)col_groups(g2, 2) + 1) * 4 + col_list_capacity(nb) * 20 + col_zst_words(g2, 2) * 4 +
           (size_t)nb * COL_SLOTS_MAX * 3 * 8 + (size_t)nb * COL_WIN * 8 * (c->slip_channel ? 2 : 1) +
           col_list_capacity(nb) * COL_EWIN * 8;
  }
  const int n3b = 3 * c->max_blobs_per_body;
  t += (size_t)(3 * n3b + 2 * n3b * n3b + 36) * 8 + (size_t)c->max_bodies * (3 + 4 + 9 + 6) * 8 + (size_t)nb * 3 * 8;
  t += (size_t)(c->max_krylov + 1) * s.ldq * 8 + 6 * s.ldq * 8 + 2 * (size_t)nb * 3 * 8;
  const size_t nblk = (s.nmax + 1023) / 1024;
  t += nblk * (2 * (size_t)c->max_krylov + 6) * 8 + 2 * (2 * (size_t)c->max_krylov + 16) * 8;
  t += (size_t)c->max_bodies * 12 * 8 + (size_t)nb * (3 * 8 + 2 * 4);  // time-step buffers
  t += (size_t)c->max_bodies * 7 * 8 + 4;                                 // bodies before a step, flag
  t += (size_t)c->max_bodies * 32 * 8 + 42 * 8;                           // drifter-corrector buffers
  const size_t m2 = (size_t)(c->max_krylov + 2) * (c->max_krylov + 2);
  t += (2 * m2 + 9 * (size_t)(c->max_krylov + 2) + 16) * 8 + s.ldq * 8;
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

}  // namespace

namespace fcm_host {

// spread -> Stokes solve -> interpolation on one precision's fields; on the
// column path the z passes are done inside spreading and interpolation
template <typename T>
int mobility_passes(fcm_handle h, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw, bool column, const double* Fin,
                    double* Vout, const KTerm& kt, hipStream_t st) {
  // a slip channel's force images are applied by the forward y pass as it reads
  if (column && h->bd.col_cx > 0) {
    HIPCHK(spread_zfwd<T>(h->g, h->bd, Fin, fp, tw, st, false));
    for (int p = 1; p <= 3; ++p) HIPCHK(stokes_pass<T>(h->g, fp, p, h->pp, tw, st, p == 1 && h->g.channel));
    HIPCHK(zinv_interp<T>(h->g, h->bd, fp, tw, Vout, kt, st));
    return FCM_OK;
  }
  HIPCHK(spread<T>(h->g, h->bb, h->bd, Fin, fp, st, false));
  HIPCHK(stokes_solve<T>(h->g, fp, h->pp, tw, st, h->g.channel != 0));
  HIPCHK(interp<T>(h->g, h->bb, h->bd, fp, Vout, kt, st));
  return FCM_OK;
}

// the passes of the column path with events between them (the z slots stay empty)
constexpr int PROFILE_COL_REPS = 4;

template <typename T>
int profile_column(fcm_handle h, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw, const KTerm& kt, hipEvent_t* ev) {
  // the fused column kernels are timed over PROFILE_COL_REPS back-to-back
  // launches each (both are idempotent), as a single launch between two events
  // also counts its dispatch
  for (int r = 0; r < PROFILE_COL_REPS; ++r) HIPCHK(spread_zfwd<T>(h->g, h->bd, h->xacc, fp, tw, h->st, false));
  HIPCHK(hipEventRecord(ev[1], h->st));
  HIPCHK(hipEventRecord(ev[2], h->st));
  for (int p = 1; p <= 3; ++p) {
    HIPCHK(stokes_pass<T>(h->g, fp, p, h->pp, tw, h->st, p == 1 && h->g.channel));
    HIPCHK(hipEventRecord(ev[2 + p], h->st));
  }
  HIPCHK(hipEventRecord(ev[6], h->st));
  HIPCHK(zinv_interp<T>(h->g, h->bd, fp, tw, h->tmp, kt, h->st, ev[7], PROFILE_COL_REPS));
  return FCM_OK;
}

int mobility_dev(fcm_handle h, const double* Fin, double* Vout, int precision, const KTerm& kt, hipStream_t st,
                 const int* gate) {
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  if (precision == FCM_PREC_FP32) {
    if (!h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated (fcm_config.fp32 = 0)");
    FieldPtrs<float> fp = h->fp32();
    fp.gate = gate;
    return mobility_passes<float>(h, fp, h->tw32, h->bd.col_fused[1] != 0, Fin, Vout, kt, st);
  }
  FieldPtrs<double> fp = h->fp64();
  fp.gate = gate;
  return mobility_passes<double>(h, fp, h->tw64, h->bd.col_fused[0] != 0, Fin, Vout, kt, st);
}

int saddle_dev(fcm_handle h, const double* in, double* out, int precision, hipStream_t st) {
  const size_t n3 = h->n3();
  KTerm kt;
  kt.zV = in + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  int rc = mobility_dev(h, in, out, precision, kt, st);
  if (rc) return rc;
  BodyData b = h->body();
  HIPCHK(mul_KT(b, in, out + n3, st));
  HIPCHK(axpby(0.0, out + n3, -1.0, out + n3, 6 * (size_t)h->NB, st));
  return FCM_OK;
}

// One Arnoldi step under device control: w = A P^-1 q_k with k read on the
// device (the K^T term comes out of the P^-1 kernel), then the DCGS2 basis
// update and the Givens step.  Every kernel is gated on the status word, so
// steps queued after convergence do nothing.
int launch_arnoldi_step(fcm_handle h, int precision, hipStream_t st) {
  const size_t n = h->nsys(), n3 = h->n3();
  const int* gate = h->gm.ictl + GM_GATE;
  BodyData b = h->body();
  b.gate = gate;
  b.kcol = h->gm.ictl + GM_K;
  b.ldq = h->ldq;
  // the inner solves of FP32+refine keep their basis in FP32
  const bool fp32 = precision == FCM_PREC_FP32;
  if (fp32) HIPCHK(apply_pinv(b, h->Qf(), h->z, h->w + n3, st));
  else HIPCHK(apply_pinv(b, h->Q, h->z, h->w + n3, st));
  KTerm kt;
  kt.zV = h->z + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  if (int rc = mobility_dev(h, h->z, h->w, precision, kt, st, gate)) return rc;
  if (fp32) HIPCHK(gm_launch_basis(h->gm, h->Qf(), h->ldq, h->w, n, h->kw.partial, gm_blocks(n), st));
  else HIPCHK(gm_launch_basis(h->gm, h->Q, h->ldq, h->w, n, h->kw.partial, gm_blocks(n), st));
  return FCM_OK;
}

}  // namespace fcm_host

namespace fcm_host {
int apply_cached_tuning(fcm_handle h);
}  // namespace fcm_host

namespace {

// body forces for the RHS: F6 (rb_set_body_forces) plus K^T of the blob forces
// (ll_barrier_forces), as assemble_RHS forms -K^T blob_forces (gmres_solver.cpp:58-70)
int body_forces_dev(fcm_handle h, const double** out) {
  *out = h->have_forces ? h->d_F6 : nullptr;
  if (!h->have_blob_forces) return FCM_OK;
  BodyData b = h->body();
  HIPCHK(mul_KT(b, h->d_fblob, h->d_F6tot, h->st));
  if (h->have_forces) HIPCHK(axpby(1.0, h->d_F6, 1.0, h->d_F6tot, 6 * (size_t)h->NB, h->st));
  *out = h->d_F6tot;
  return FCM_OK;
}

// calculate_shear_blob_velocities (fcm_fluid_solver.cpp:488-550): the shear
// field interpolated at the current blobs, stored as the handle's Vshear
int compute_shear_velocities(fcm_handle h) {
  if (!h->have_shear) return FCM_OK;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fpshear(), h->Vshear, KTerm{}, h->st));
  h->have_vshear = true;
  h->vshear_nb = h->bd.nb;
  return FCM_OK;
}

// the Vshear that assemble_RHS subtracts (gmres_solver.cpp:62-64): the stored
// vector of the last calculate_shear_blob_velocities, which main.cpp fills only
// at the drifter-corrector mid-step (main.cpp:270-274); null if none is stored
int shear_velocities_dev(fcm_handle h, const double** out) {
  *out = (h->have_shear && h->have_vshear && h->vshear_nb == h->bd.nb) ? h->Vshear : nullptr;
  return FCM_OK;
}

int setup_blobs_device(fcm_handle h, int nb) {
  const int nb_before = h->bd.nb;
  h->bd.nb = nb;
  h->bd.Y = h->dY;
  ++h->epoch;
  HIPCHK(setup_gaussians(h->g, h->bb, h->bd, h->cfg.a, h->st));
  h->have_blobs = true;
  // a handle's first blobs (or a new blob count) select its shape class
  if (!h->tuning_set && (nb != nb_before || !h->cache_checked)) {
    h->cache_checked = true;
    return fcm_host::apply_cached_tuning(h);
  }
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

int fcm_device_memory(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return fail(FCM_ERR_ARG, "null argument");
  HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
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
  if (!spread_fast_ok(h->g, bb, cfg->ng) && !spread_generic_ok(h->g, bb, cfg->ng)) {
    delete h;
    return fail(FCM_ERR_ARG, "stencil too wide for the spreading tiles of this grid");
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
  if (e == hipSuccess) h->bd.woff = dalloc<float4>(h, (size_t)nb, e);
  if (cfg->slip_channel) {
    if (e == hipSuccess) h->bd.iw64 = dalloc<double>(h, 3 * (size_t)nb * cfg->ng, e);
    if (e == hipSuccess) h->bd.iw32 = dalloc<float>(h, 3 * (size_t)nb * cfg->ng, e);
  } else {
    h->bd.iw64 = h->bd.w64;
    h->bd.iw32 = h->bd.w32;
  }
  {
    const int nt = spread_tiles(h->g);
    if (e == hipSuccess) h->bd.tile_count = dalloc<int>(h, nt, e);
    if (e == hipSuccess) h->bd.tile_start = dalloc<int>(h, nt + 1, e);
    if (e == hipSuccess) h->bd.tile_cursor = dalloc<int>(h, nt, e);
    if (e == hipSuccess) h->bd.tile_list = dalloc<int>(h, 8 * (size_t)nb, e);
  }
  h->bd.ng = cfg->ng;
  h->bd.col_cx = col_patch_cx(h->g, bb, cfg->ng);
  if (h->bd.col_cx > 0) {
    const int ng_ = col_groups(h->g, 2);  // sized for either patch width
    if (e == hipSuccess) h->bd.col_count = dalloc<int>(h, ng_, e);
    if (e == hipSuccess) h->bd.col_start = dalloc<int>(h, (size_t)ng_ + 1, e);
    if (e == hipSuccess) h->bd.col_cursor = dalloc<int>(h, ng_, e);
    if (e == hipSuccess) h->bd.col_list = dalloc<int>(h, col_list_capacity(nb), e);
    if (e == hipSuccess) h->bd.col_ent = dalloc<int4>(h, col_list_capacity(nb), e);
    if (e == hipSuccess) h->bd.col_zst = dalloc<int>(h, col_zst_words(h->g, 2), e);
    if (e == hipSuccess) h->bd.cwin64 = dalloc<double>(h, (size_t)nb * COL_WIN, e);
    if (cfg->slip_channel) {
      if (e == hipSuccess) h->bd.icwin64 = dalloc<double>(h, (size_t)nb * COL_WIN, e);
    } else {
      h->bd.icwin64 = h->bd.cwin64;
    }
    if (e == hipSuccess) h->bd.ewin64 = dalloc<double>(h, col_list_capacity(nb) * COL_EWIN, e);
    if (e == hipSuccess) h->bd.col_part = dalloc<double>(h, (size_t)nb * COL_SLOTS_MAX * 3, e);
  } else {
    h->bd.col_fused[0] = h->bd.col_fused[1] = 0;
  }
  h->bd.ng = cfg->ng;
  const int n3b = 3 * cfg->max_blobs_per_body;
  if (e == hipSuccess) h->d_ref = dalloc<double>(h, n3b, e);
  if (e == hipSuccess) h->d_Minv = dalloc<double>(h, (size_t)n3b * n3b, e);
  if (e == hipSuccess) h->d_MinvT = dalloc<double>(h, (size_t)n3b * n3b, e);
  if (e == hipSuccess) h->d_N = dalloc<double>(h, 36, e);
  if (e == hipSuccess) h->d_x = dalloc<double>(h, 3 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_q = dalloc<double>(h, 4 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_R = dalloc<double>(h, 9 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_F6 = dalloc<double>(h, 6 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_F6tot = dalloc<double>(h, 6 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_V6 = dalloc<double>(h, 6 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_x0 = dalloc<double>(h, 3 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_q0 = dalloc<double>(h, 4 * (size_t)cfg->max_bodies, e);
  if (e == hipSuccess) h->d_flag = dalloc<int>(h, 1, e);
  {
    const size_t nbd = (size_t)cfg->max_bodies;
    if (e == hipSuccess) h->d_xs = dalloc<double>(h, 3 * nbd, e);
    if (e == hipSuccess) h->d_qs = dalloc<double>(h, 4 * nbd, e);
    if (e == hipSuccess) h->d_xm = dalloc<double>(h, 3 * nbd, e);
    if (e == hipSuccess) h->d_qm = dalloc<double>(h, 4 * nbd, e);
    if (e == hipSuccess) h->d_U0 = dalloc<double>(h, 6 * nbd, e);
    if (e == hipSuccess) h->d_Ud = dalloc<double>(h, 6 * nbd, e);
    if (e == hipSuccess) h->d_b6 = dalloc<double>(h, 6 * nbd, e);
    if (e == hipSuccess) h->d_KTKinv = dalloc<double>(h, 36, e);
    if (e == hipSuccess) h->d_divu = dalloc<double>(h, 6, e);
  }
  if (e == hipSuccess) h->d_fblob = dalloc<double>(h, 3 * (size_t)nb, e);
  if (e == hipSuccess) h->sw.cell_of = dalloc<int>(h, nb, e);
  if (e == hipSuccess) h->sw.sorted = dalloc<int>(h, nb, e);
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
    if (e == hipSuccess) h->hist_all = dalloc<double>(h, (size_t)M + 2, e);
  }
  if (e != hipSuccess) return cleanup_fail(FCM_ERR_NOMEM, "device allocation failed");

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
  batch_forget(h);
  batch_destroy_solo(h);
  for (void* p : h->allocs) (void)hipFree(p);
  for (void* p : h->lazy) (void)hipFree(p);
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
  int rc = mobility_dev(h, h->tmp, h->Vtmp, precision, KTerm{}, h->st);
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
  {
    // ref_KTKinv = inv(K^T K), K = [I, rcross(r_m)] over the blobs
    // (collection_of_rigid_bodies.cpp:117-128); a body whose blobs are collinear
    // (K^T K singular about the axis) gets the pseudo-inverse
    double KTK[36] = {0.0};
    for (int m = 0; m < nbpb; ++m) {
      const double* r = blob_ref + 3 * m;
      // rows of K for this blob: K(3m + c, .) = [e_c, (rcross(r))_c.], rcross(r) v = v x r
      const double Kb[3][6] = {{1, 0, 0, 0, r[2], -r[1]}, {0, 1, 0, -r[2], 0, r[0]}, {0, 0, 1, r[1], -r[0], 0}};
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j)
          for (int c = 0; c < 3; ++c) KTK[6 * i + j] += Kb[c][i] * Kb[c][j];
    }
    double inv[36];
    if (dense::inv_spd(6, KTK, inv) != 0) dense::pinv_sym(6, KTK, inv, 1e-9);
    HIPCHK(hipMemcpyAsync(h->d_KTKinv, inv, 36 * sizeof(double), hipMemcpyHostToDevice, h->st));
  }
  h->have_geometry = true;
  ++h->epoch;
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
  h->have_prev = false;
  h->gdc_mid = false;
  h->have_rhs = false;
  h->have_blob_forces = false;
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
  // the transposed copy lets the preconditioner stage ref_Minv with contiguous loads
  std::vector<double> MT(n3 * n3);
  for (size_t i = 0; i < n3; ++i)
    for (size_t k = 0; k < n3; ++k) MT[k * n3 + i] = Minv[i * n3 + k];
  HIPCHK(hipMemcpyAsync(h->d_Minv, Minv, n3 * n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_MinvT, MT.data(), n3 * n3 * sizeof(double), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_N, N, 36 * sizeof(double), hipMemcpyHostToDevice, h->st));
  h->have_ref = true;
  return sync_stream(h);
}

int fcm_ref_matrices(fcm_handle h, double* Msingle, double* Minv, double* N) {
  if (int rc = check(h)) return rc;
  if (!h->have_geometry) return fail(FCM_ERR_STATE, "rb_set_geometry first");
  const fcm_config& c = h->cfg;
  const int nbpb = h->nbpb, n3 = 3 * nbpb;
  // body 0 at the box centre, or at y = LY/4 in a slip channel, identity
  // orientation (collection_of_rigid_bodies.cpp:139-149)
  const double L[3] = {c.dx * c.nx, c.dx * c.ny, c.dx * c.nz};
  const double at[3] = {0.5, c.slip_channel ? 0.25 : 0.5, 0.5};
  std::vector<double> Y(n3);
  for (int m = 0; m < nbpb; ++m)
    for (int d = 0; d < 3; ++d) {
      const double p = at[d] * L[d] + h->h_ref[3 * m + d];
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
    rc = mobility_dev(h, h->tmp, h->Vtmp, FCM_PREC_FP64, KTerm{}, h->st);
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
  const double* F6 = nullptr;
  if (int rc = body_forces_dev(h, &F6)) return rc;
  const double* Vs = nullptr;
  if (int rc = shear_velocities_dev(h, &Vs)) return rc;
  HIPCHK(rhs_assemble(use_random ? h->Vrand : nullptr, Vs, F6, h->rhs, h->n3(), 6 * (size_t)h->NB, h->st));
  h->have_rhs = true;
  return sync_stream(h);
}

int gmres_saddle_matvec(fcm_handle h, const double* in, double* out, int precision) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  const size_t n = h->nsys();
  HIPCHK(hipMemcpyAsync(h->xacc, in, n * sizeof(double), hipMemcpyHostToDevice, h->st));
  int rc = saddle_dev(h, h->xacc, h->tmp, precision, h->st);
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
  int rc = solve_single(h, opts, rep, resid_hist);
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
  // events: 0 start, 1 spread, 2..6 the five FFT slots, 7 interpolation, 8 reduction, 9 K^T
  hipEvent_t ev[10];
  for (int i = 0; i < 10; ++i) HIPCHK(hipEventCreate(&ev[i]));
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  KTerm kt;
  kt.zV = h->xacc + n3;
  kt.dpos = h->d_dpos;
  kt.nbpb = h->nbpb;
  BodyData b = h->body();
  int rc = FCM_OK;
  const bool column = h->bd.col_cx > 0 && h->bd.col_fused[precision == FCM_PREC_FP32 ? 1 : 0] != 0;
  // rep -1 is an untimed warm-up (first launches load code objects)
  for (int r = -1; r < reps && rc == FCM_OK; ++r) {
    HIPCHK(hipEventRecord(ev[0], h->st));
    if (column) {
      rc = precision == FCM_PREC_FP32 ? profile_column<float>(h, h->fp32(), h->tw32, kt, ev)
                                      : profile_column<double>(h, h->fp64(), h->tw64, kt, ev);
      if (rc) break;
    } else if (precision == FCM_PREC_FP32) {
      HIPCHK(spread<float>(h->g, h->bb, h->bd, h->xacc, h->fp32(), h->st, false));
      HIPCHK(hipEventRecord(ev[1], h->st));
      for (int p = 0; p < 5; ++p) {
        HIPCHK(stokes_pass<float>(h->g, h->fp32(), p, h->pp, h->tw32, h->st, p == 1 && h->g.channel));
        HIPCHK(hipEventRecord(ev[2 + p], h->st));
      }
      HIPCHK(interp<float>(h->g, h->bb, h->bd, h->fp32(), h->tmp, kt, h->st));
    } else {
      HIPCHK(spread<double>(h->g, h->bb, h->bd, h->xacc, h->fp64(), h->st, false));
      HIPCHK(hipEventRecord(ev[1], h->st));
      for (int p = 0; p < 5; ++p) {
        HIPCHK(stokes_pass<double>(h->g, h->fp64(), p, h->pp, h->tw64, h->st, p == 1 && h->g.channel));
        HIPCHK(hipEventRecord(ev[2 + p], h->st));
      }
      HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fp64(), h->tmp, kt, h->st));
    }
    if (!column) HIPCHK(hipEventRecord(ev[7], h->st));
    HIPCHK(hipEventRecord(ev[8], h->st));
    HIPCHK(mul_KT(b, h->xacc, h->tmp + n3, h->st));
    HIPCHK(hipEventRecord(ev[9], h->st));
    HIPCHK(hipEventSynchronize(ev[9]));
    for (int i = 0; i < 9 && r >= 0; ++i) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      acc[i] += ms;
    }
  }
  for (int i = 0; i < 10; ++i) (void)hipEventDestroy(ev[i]);
  // slots: spread, z fwd, y fwd, x, y inv, z inv, interpolation, K^T, reduction
  const double mean[9] = {acc[0] / reps, acc[1] / reps, acc[2] / reps, acc[3] / reps, acc[4] / reps,
                          acc[5] / reps, acc[6] / reps, acc[8] / reps, acc[7] / reps};
  for (int i = 0; i < 9; ++i) ms_per_pass[i] = mean[i];
  if (column) {
    ms_per_pass[1] = ms_per_pass[5] = 0.0;  // back-to-back events only: the z passes ran inside
    ms_per_pass[0] /= PROFILE_COL_REPS;
    ms_per_pass[6] /= PROFILE_COL_REPS;
  }
  else ms_per_pass[8] = 0.0;
  return rc;
}

int fcm_get_tuning(fcm_handle h, fcm_tuning* t) {
  if (int rc = check(h)) return rc;
  if (!t) return fail(FCM_ERR_ARG, "null tuning");
  // the shapes the passes run with: a default (0) is reported as what it resolves to
  for (int p = 0; p < 2; ++p) {
    t->spread_variant[p] = h->bd.spread_variant[p];
    t->interp_izb[p] = h->bd.interp_izb[p];
    t->column[p] = h->bd.col_fused[p];
    t->col_points[p] =
        h->bd.col_cx > 0 ? col_shape_pm(p == FCM_PREC_FP32 ? 4 : 8, h->g.nz, h->bd.col_cx, h->bd.col_pm[p]) : 0;
    t->y_points[p] = h->g.ypm[p] ? h->g.ypm[p] : 16;
    t->x_points[p] = h->g.xpm[p] ? h->g.xpm[p] : (p == FCM_PREC_FP32 ? 16 : 8);
  }
  t->col_cx = h->bd.col_cx;
  return FCM_OK;
}

}  // extern "C"

namespace {

bool pass_points_ok(int n, int pm) { return pm == 0 || pm == 16 || (pm == 8 && n >= 64); }

// why `t` is not a tuning this handle can run, or null
const char* tuning_problem(const fcm_handle h, const fcm_tuning* t) {
  const bool colgrid = h->bd.col_cx > 0;
  if (colgrid) {
    if (t->col_cx != 2 && t->col_cx != 4) return "col_cx must be 2 or 4 on this grid";
    if (h->g.nx % t->col_cx != 0) return "col_cx must divide nx";
  } else if (t->col_cx != 0) {
    return "the grid has no column path (col_cx must be 0)";
  }
  for (int p = 0; p < 2; ++p) {
    if (t->spread_variant[p] < 0 || t->spread_variant[p] > 2 || t->interp_izb[p] < 0 || t->interp_izb[p] > 3)
      return "spread_variant must be 0..2 and interp_izb 0..3";
    if (t->column[p] < 0 || t->column[p] > 1 || (t->column[p] == 1 && !colgrid)) return "column must be 0 or 1";
    if (t->col_points[p] != 0 && t->col_points[p] != 8 && t->col_points[p] != 16) return "col_points must be 0, 8 or 16";
    if (colgrid && !col_shape_valid(p == FCM_PREC_FP32 ? 4 : 8, h->g.nz, t->col_cx, t->col_points[p]))
      return "no column kernels for this patch width and z-pass points per thread";
    if (!pass_points_ok(h->g.ny, t->y_points[p]) || t->y_points[p] < 0) return "y_points must be 0, 16 or 8 (ny >= 64)";
    if (p == FCM_PREC_FP64 ? (t->x_points[p] != 0 && t->x_points[p] != 8)
                           : (t->x_points[p] < 0 || !pass_points_ok(h->g.nx, t->x_points[p])))
      return "x_points must be 0, 16 or 8 (nx >= 64); FP64 keeps 8 (0 or 8)";
  }
  return nullptr;
}

int apply_tuning(fcm_handle h, const fcm_tuning* t) {
  HIPCHK(hipStreamSynchronize(h->st));
  for (int p = 0; p < 2; ++p) {
    h->bd.spread_variant[p] = t->spread_variant[p];
    h->bd.interp_izb[p] = t->interp_izb[p];
    h->bd.col_fused[p] = t->column[p];
    h->bd.col_pm[p] = t->col_points[p];
    h->g.ypm[p] = t->y_points[p];
    h->g.xpm[p] = t->x_points[p];
  }
  if (t->col_cx != h->bd.col_cx) {
    h->bd.col_cx = t->col_cx;
    // the patch lists follow the patch width
    if (h->have_blobs) HIPCHK(build_col_lists(h->g, h->bb, h->bd, h->st));
  }
  ++h->epoch;
  return sync_stream(h);
}

// Tuned launch shapes cached by shape class: grid, stencil, channel and the
// power of two at or above the blob count.  fcm_autotune stores its choice;
// fcm_autotune_cached and the blob set-up of a handle whose tuning was never
// set explicitly take it.  One cache per process, guarded by a mutex.
struct ShapeClass {
  int nx, ny, nz, ng, channel, nb2;
  bool operator<(const ShapeClass& o) const {
    return std::tie(nx, ny, nz, ng, channel, nb2) < std::tie(o.nx, o.ny, o.nz, o.ng, o.channel, o.nb2);
  }
};
struct CachedTuning {
  fcm_tuning t{};
  bool have[2] = {false, false};
};
std::mutex g_tune_mu;
std::map<ShapeClass, CachedTuning> g_tune_cache;

ShapeClass shape_class(const fcm_handle h) {
  int nb2 = 1;
  while (nb2 < h->bd.nb) nb2 *= 2;
  return ShapeClass{h->cfg.nx, h->cfg.ny, h->cfg.nz, h->cfg.ng, h->cfg.slip_channel, nb2};
}

// the cached choice for precision p of h's class into t (the rest of t kept)
bool cache_lookup(const fcm_handle h, int p, fcm_tuning* t) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const auto it = g_tune_cache.find(shape_class(h));
  if (it == g_tune_cache.end() || !it->second.have[p]) return false;
  const fcm_tuning& c = it->second.t;
  t->spread_variant[p] = c.spread_variant[p];
  t->interp_izb[p] = c.interp_izb[p];
  t->column[p] = c.column[p];
  t->col_points[p] = c.col_points[p];
  t->y_points[p] = c.y_points[p];
  t->x_points[p] = c.x_points[p];
  // the patch width is shared by both precisions; the FP32 choice wins
  if (p == FCM_PREC_FP32 || !it->second.have[FCM_PREC_FP32]) t->col_cx = c.col_cx;
  return true;
}

void cache_store(const fcm_handle h, int p, const fcm_tuning* t) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  CachedTuning& c = g_tune_cache[shape_class(h)];
  c.t.spread_variant[p] = t->spread_variant[p];
  c.t.interp_izb[p] = t->interp_izb[p];
  c.t.column[p] = t->column[p];
  c.t.col_points[p] = t->col_points[p];
  c.t.y_points[p] = t->y_points[p];
  c.t.x_points[p] = t->x_points[p];
  if (p == FCM_PREC_FP32 || !c.have[FCM_PREC_FP32]) c.t.col_cx = t->col_cx;
  c.have[p] = true;
}

}  // namespace

int fcm_set_tuning(fcm_handle h, const fcm_tuning* t) {
  if (int rc = check(h)) return rc;
  if (!t) return fail(FCM_ERR_ARG, "null tuning");
  if (const char* why = tuning_problem(h, t)) return fail(FCM_ERR_ARG, std::string("bad tuning: ") + why);
  h->tuning_set = true;
  return apply_tuning(h, t);
}

namespace fcm_host {

// at a handle's blob set-up: the cached tuning of its shape class, unless the
// handle's tuning was set explicitly
int apply_cached_tuning(fcm_handle h) {
  if (h->tuning_set) return FCM_OK;
  fcm_tuning t;
  if (int rc = fcm_get_tuning(h, &t)) return rc;
  bool any = false;
  for (int p = 0; p < 2; ++p) any = cache_lookup(h, p, &t) || any;
  if (!any || tuning_problem(h, &t)) return FCM_OK;
  return apply_tuning(h, &t);
}

}  // namespace fcm_host

extern "C" {

int fcm_autotune(fcm_handle h, int precision, int reps, fcm_tuning* chosen) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  if (precision != FCM_PREC_FP32 && precision != FCM_PREC_FP64) return fail(FCM_ERR_ARG, "bad precision");
  if (precision == FCM_PREC_FP32 && !h->f32[0]) return fail(FCM_ERR_STATE, "FP32 path not allocated");
  reps = std::max(1, reps);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const size_t n3 = h->n3();
  HIPCHK(axpby(0.0, h->tmp, 0.0, h->tmp, n3, h->st));  // spread input: zeros (same work)
  auto time_it = [&](auto&& launch, float& ms) -> int {
    if (int rc = launch()) return rc;  // warm-up
    HIPCHK(hipEventRecord(e0, h->st));
    for (int r = 0; r < reps; ++r)
      if (int rc = launch()) return rc;
    HIPCHK(hipEventRecord(e1, h->st));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    return FCM_OK;
  };
  const KTerm kt{};
  const int pi = precision == FCM_PREC_FP32 ? 1 : 0;
  auto spread_once = [&]() -> int {
    if (precision == FCM_PREC_FP32) HIPCHK(spread<float>(h->g, h->bb, h->bd, h->tmp, h->fp32(), h->st, false));
    else HIPCHK(spread<double>(h->g, h->bb, h->bd, h->tmp, h->fp64(), h->st, false));
    return FCM_OK;
  };
  auto interp_once = [&]() -> int {
    if (precision == FCM_PREC_FP32) HIPCHK(interp<float>(h->g, h->bb, h->bd, h->fp32(), h->Vtmp, kt, h->st));
    else HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fp64(), h->Vtmp, kt, h->st));
    return FCM_OK;
  };
  auto product_once = [&]() -> int {
    if (precision == FCM_PREC_FP32)
      return mobility_passes<float>(h, h->fp32(), h->tw32, h->bd.col_fused[1] != 0, h->tmp, h->Vtmp, kt, h->st);
    return mobility_passes<double>(h, h->fp64(), h->tw64, h->bd.col_fused[0] != 0, h->tmp, h->Vtmp, kt, h->st);
  };
  // the fastest value of one choice, by timing `once` for each candidate
  int rc = FCM_OK;
  auto pick = [&](int* knob, std::initializer_list<int> cands, auto&& once, auto&& valid) {
    float best = 0.f;
    int best_v = *knob;
    bool first = true;
    for (int v : cands) {
      if (rc || !valid(v)) continue;
      *knob = v;
      float ms = 0.f;
      rc = time_it(once, ms);
      if (!rc && (first || ms < best)) {
        best = ms;
        best_v = v;
        first = false;
      }
    }
    *knob = best_v;
  };
  auto any = [](int) { return true; };
  // separate-path spreading and interpolation tiles
  pick(&h->bd.spread_variant[pi], {0, 1, 2}, spread_once, any);
  pick(&h->bd.interp_izb[pi], {1, 2, 3}, interp_once, any);
  // FFT pass shapes (points per thread, and with them the columns per workgroup)
  pick(&h->g.ypm[pi], {16, 8}, product_once, [&](int v) { return pass_points_ok(h->g.ny, v); });
  if (precision == FCM_PREC_FP32)
    pick(&h->g.xpm[pi], {16, 8}, product_once, [&](int v) { return pass_points_ok(h->g.nx, v); });
  if (!rc && h->bd.col_cx > 0) {
    // the column path: patch width (shared by both precisions, so only the FP32
    // tuning moves it when the handle has an FP32 path) and z-pass points per
    // thread, then the column path against the separate passes
    const int cx0 = h->bd.col_cx;
    h->bd.col_fused[pi] = 1;
    const bool tune_cx = precision == FCM_PREC_FP32 || !h->f32[0];
    int best_cx = cx0, best_pm = h->bd.col_pm[pi];
    float best = 0.f;
    bool first = true;
    for (int cx : {4, 2}) {
      if (rc || (!tune_cx && cx != cx0) || h->g.nx % cx != 0) continue;
      if (cx != h->bd.col_cx) {
        h->bd.col_cx = cx;
        if (hipError_t e = build_col_lists(h->g, h->bb, h->bd, h->st)) rc = fail(FCM_ERR_HIP, hipGetErrorString(e));
      }
      for (int pm : {8, 16}) {
        if (rc || !col_shape_valid(precision == FCM_PREC_FP32 ? 4 : 8, h->g.nz, cx, pm)) continue;
        h->bd.col_pm[pi] = pm;
        float ms = 0.f;
        rc = time_it(product_once, ms);
        if (!rc && (first || ms < best)) {
          best = ms;
          best_cx = cx;
          best_pm = pm;
          first = false;
        }
      }
    }
    if (!rc && best_cx != h->bd.col_cx) {
      h->bd.col_cx = best_cx;
      if (hipError_t e = build_col_lists(h->g, h->bb, h->bd, h->st)) rc = fail(FCM_ERR_HIP, hipGetErrorString(e));
    }
    h->bd.col_pm[pi] = best_pm;
    pick(&h->bd.col_fused[pi], {1, 0}, product_once, any);
  }
  ++h->epoch;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  fcm_tuning t;
  if (int r2 = fcm_get_tuning(h, &t)) return r2;
  cache_store(h, precision, &t);
  h->tuning_set = true;
  if (chosen) *chosen = t;
  return sync_stream(h);
}

int fcm_clear_tuning_cache(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_cache.clear();
  return FCM_OK;
}

int fcm_autotune_cached(fcm_handle h, int precision, int reps, fcm_tuning* chosen, int* from_cache) {
  if (int rc = check(h)) return rc;
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  if (precision != FCM_PREC_FP32 && precision != FCM_PREC_FP64) return fail(FCM_ERR_ARG, "bad precision");
  fcm_tuning t;
  if (int rc = fcm_get_tuning(h, &t)) return rc;
  if (cache_lookup(h, precision, &t) && !tuning_problem(h, &t)) {
    if (from_cache) *from_cache = 1;
    h->tuning_set = true;
    if (int rc = apply_tuning(h, &t)) return rc;
    if (chosen) *chosen = t;
    return FCM_OK;
  }
  if (from_cache) *from_cache = 0;
  return fcm_autotune(h, precision, reps, chosen);
}

int fcm_profile_copy(fcm_handle h, size_t bytes, int reps, double* ms_per_copy) {
  if (int rc = check(h)) return rc;
  if (!ms_per_copy || bytes < 16 || reps < 1) return fail(FCM_ERR_ARG, "bad copy arguments");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(FCM_ERR_NOMEM, "copy buffers");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMemsetAsync(a, 0, bytes, h->st);
  if (e == hipSuccess) e = copy16(a, b, bytes, h->st);  // warm-up
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, h->st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = copy16(a, b, bytes, h->st);
  if (e == hipSuccess) e = hipEventRecord(e1, h->st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  HIPCHK(e);
  *ms_per_copy = ms / reps;
  return FCM_OK;
}

int fcm_mobility_solve(fcm_handle h, uint64_t seed, const gmres_opts* opts, gmres_report* rep) {
  if (int rc = fcm_mobility_prepare(h, seed)) return rc;
  return solve_single(h, opts, rep, nullptr);
}

int fcm_mobility_prepare(fcm_handle h, uint64_t seed) {
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
  const double* F6 = nullptr;
  if (int rc = body_forces_dev(h, &F6)) return rc;
  const double* Vs = nullptr;
  if (int rc = shear_velocities_dev(h, &Vs)) return rc;
  HIPCHK(rhs_assemble(fluct ? h->Vrand : nullptr, Vs, F6, h->rhs, h->n3(), 6 * (size_t)h->NB, h->st));
  h->have_rhs = true;
  return FCM_OK;
}

int ll_set_cells(fcm_handle h, int cells_per_dim) {
  if (int rc = check(h)) return rc;
  if (cells_per_dim != 0 && (cells_per_dim < 3 || cells_per_dim > 1024))
    return fail(FCM_ERR_ARG, "cells_per_dim must be 0 or in 3..1024");
  h->ll_cells = cells_per_dim;
  h->have_blob_forces = false;
  h->have_rhs = false;
  return FCM_OK;
}

int ll_barrier_forces(fcm_handle h, double cutoff, double* F3) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  if (!(cutoff > 0.0)) return fail(FCM_ERR_ARG, "cutoff must be positive");
  const fcm_config& c = h->cfg;
  const double L[3] = {c.dx * c.nx, c.dx * c.ny, c.dx * c.nz};
  const StericCells cg = steric_cells(L, cutoff, h->ll_cells);
  const size_t ncells = (size_t)cg.nc[0] * cg.nc[1] * cg.nc[2];
  if ((int)ncells > h->sw.cap_cells) {
    HIPCHK(hipStreamSynchronize(h->st));
    for (void* p : h->lazy) (void)hipFree(p);
    h->lazy.clear();
    h->sw.cap_cells = 0;
    void *a = nullptr, *b = nullptr, *d = nullptr;
    HIPCHK(hipMalloc(&a, ncells * sizeof(int)));
    h->lazy.push_back(a);
    HIPCHK(hipMalloc(&b, (ncells + 1) * sizeof(int)));
    h->lazy.push_back(b);
    HIPCHK(hipMalloc(&d, ncells * sizeof(int)));
    h->lazy.push_back(d);
    h->sw.count = static_cast<int*>(a);
    h->sw.start = static_cast<int*>(b);
    h->sw.cursor = static_cast<int*>(d);
    h->sw.cap_cells = (int)ncells;
  }
  const int nb = h->NB * h->nbpb;
  HIPCHK(barrier_forces(h->dY, nb, h->nbpb, cg, cutoff, h->ll_cells > 0, c.kT, c.a, c.slip_channel != 0, h->sw,
                        h->d_fblob, h->st));
  h->have_blob_forces = true;
  h->have_rhs = false;
  if (F3) HIPCHK(hipMemcpyAsync(F3, h->d_fblob, 3 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

namespace {

hipError_t copy_bodies(fcm_handle h, double* xd, double* qd, const double* xsrc, const double* qsrc) {
  hipError_t e = hipMemcpyAsync(xd, xsrc, 3 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st);
  if (e == hipSuccess) e = hipMemcpyAsync(qd, qsrc, 4 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st);
  return e;
}

// blob positions, rotations and Gaussians of the bodies in d_x, d_q
int place_bodies(fcm_handle h) {
  BodyData b = h->body();
  HIPCHK(bodies_update(h->g, b, h->st));
  return setup_blobs_device(h, h->NB * h->nbpb);
}

}  // namespace

int rb_end_of_step_update(fcm_handle h, const double* V6, double nu) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  const double* v = nullptr;
  if (V6) {
    HIPCHK(hipMemcpyAsync(h->d_V6, V6, 6 * (size_t)h->NB * sizeof(double), hipMemcpyHostToDevice, h->st));
    v = h->d_V6;
  } else {
    if (!h->have_sol) return fail(FCM_ERR_STATE, "no solution to take the body velocities from");
    v = h->sol + h->n3();
  }
  // after a drifter-corrector mid-step the update applies to the start-of-step
  // bodies (bodies vs midstep_bodies, collection_of_rigid_bodies.cpp:566-621, 759-790)
  if (h->gdc_mid) {
    HIPCHK(copy_bodies(h, h->d_x, h->d_q, h->d_xs, h->d_qs));
    h->gdc_mid = false;
  }
  // temp_bodies = bodies (collection_of_rigid_bodies.cpp:761), for reset_to_start_of_step
  HIPCHK(hipMemcpyAsync(h->d_x0, h->d_x, 3 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_q0, h->d_q, 4 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st));
  HIPCHK(bodies_step(h->g, h->d_x, h->d_q, h->NB, v, nu * h->cfg.dt, h->st));
  BodyData b = h->body();
  HIPCHK(bodies_update(h->g, b, h->st));
  if (int rc = setup_blobs_device(h, h->NB * h->nbpb)) return rc;
  h->have_prev = true;
  h->have_blob_forces = false;
  h->have_rhs = false;
  return sync_stream(h);
}

int fcm_shear_field(fcm_handle h, double shear_force_scale, double shear_wavenumber) {
  if (int rc = check(h)) return rc;
  h->have_vshear = false;  // the stored blob velocities belonged to the previous field
  h->have_rhs = false;
  if (shear_force_scale == 0.0) {
    h->have_shear = false;
    return FCM_OK;
  }
  hipError_t e = hipSuccess;
  for (int i = 0; i < 3 && e == hipSuccess; ++i)
    if (!h->ushear[i]) h->ushear[i] = dalloc<c2<double>>(h, h->g.field_words(), e);
  if (e == hipSuccess && !h->Vshear) h->Vshear = dalloc<double>(h, 3 * (size_t)h->cfg.max_blobs, e);
  if (e != hipSuccess) return fail(FCM_ERR_NOMEM, "shear field allocation failed");
  const FieldPtrs<double> fp = h->fp64();
  HIPCHK(shear_force(h->g, fp, shear_force_scale * h->cfg.kT, h->cfg.a, shear_wavenumber, h->st));
  HIPCHK(channel_images<double>(h->g, fp, h->st));
  HIPCHK(stokes_solve<double>(h->g, fp, h->pp, h->tw64, h->st));
  for (int i = 0; i < 3; ++i)
    HIPCHK(hipMemcpyAsync(h->ushear[i], h->f64[i], h->g.field_words() * sizeof(double2), hipMemcpyDeviceToDevice,
                          h->st));
  h->have_shear = true;
  h->have_rhs = false;
  return sync_stream(h);
}

int fcm_shear_interp(fcm_handle h, double* V) {
  if (int rc = check(h)) return rc;
  if (!h->have_shear) return fail(FCM_ERR_STATE, "no shear field (fcm_shear_field)");
  if (int rc = compute_shear_velocities(h)) return rc;
  h->have_rhs = false;
  if (V) HIPCHK(hipMemcpyAsync(V, h->Vshear, h->n3() * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_x_velocity_profile(fcm_handle h, const double* lambda, double* prof) {
  if (int rc = check(h)) return rc;
  if (!prof) return fail(FCM_ERR_ARG, "null argument");
  if (!h->have_blobs) return fail(FCM_ERR_STATE, "no blob positions set");
  const double* F = nullptr;
  if (lambda) {
    HIPCHK(hipMemcpyAsync(h->tmp, lambda, h->n3() * sizeof(double), hipMemcpyHostToDevice, h->st));
    F = h->tmp;
  } else {
    if (!h->have_sol) return fail(FCM_ERR_STATE, "no solution to take the blob forces from");
    F = h->sol;
  }
  if (!h->d_prof) {
    hipError_t e = hipSuccess;
    h->d_prof = dalloc<double>(h, h->g.ny, e);
    if (e != hipSuccess) return fail(FCM_ERR_NOMEM, "profile buffer");
  }
  const FieldPtrs<double> fp = h->fp64();
  HIPCHK(spread<double>(h->g, h->bb, h->bd, F, fp, h->st));
  HIPCHK(stokes_solve<double>(h->g, fp, h->pp, h->tw64, h->st));
  HIPCHK(x_velocity_profile(h->g, fp, h->have_shear ? h->fpshear() : FieldPtrs<double>{}, h->d_prof, h->st));
  HIPCHK(hipMemcpyAsync(prof, h->d_prof, h->g.ny * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int rb_advance_to_midstep(fcm_handle h, const double* V, int* overlap) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  const double* v = h->Vrand;
  if (V) {
    HIPCHK(hipMemcpyAsync(h->Vtmp, V, h->n3() * sizeof(double), hipMemcpyHostToDevice, h->st));
    v = h->Vtmp;
  } else if (!h->have_urand) {
    return fail(FCM_ERR_STATE, "no fluctuating blob velocities (fcm_random_field, fcm_random_interp)");
  }
  if (h->gdc_mid) HIPCHK(copy_bodies(h, h->d_x, h->d_q, h->d_xs, h->d_qs));  // a step begun again
  else HIPCHK(copy_bodies(h, h->d_xs, h->d_qs, h->d_x, h->d_q));
  BodyData b = h->body();
  HIPCHK(mul_KT(b, v, h->d_b6, h->st));
  HIPCHK(body_fit(b, h->d_b6, h->d_KTKinv, h->d_U0, h->st));
  HIPCHK(bodies_step(h->g, h->d_x, h->d_q, h->NB, h->d_U0, 0.5 * h->cfg.dt, h->st));
  HIPCHK(copy_bodies(h, h->d_xm, h->d_qm, h->d_x, h->d_q));
  if (int rc = place_bodies(h)) return rc;
  h->gdc_mid = true;
  h->have_blob_forces = false;
  h->have_rhs = false;
  int ov = 0;
  if (h->g.channel) {
    HIPCHK(wall_overlap(h->dY, h->NB * h->nbpb, 0.5 * h->g.dx * h->g.ny, h->d_flag, h->st));
    HIPCHK(hipMemcpyAsync(&ov, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  if (ov) {  // main.cpp:237-246: the step is begun again, the bodies stay where they were
    HIPCHK(copy_bodies(h, h->d_x, h->d_q, h->d_xs, h->d_qs));
    if (int rc = place_bodies(h)) return rc;
    h->gdc_mid = false;
  }
  if (overlap) *overlap = ov;
  return sync_stream(h);
}

int rb_calculate_divU(fcm_handle h, double fd_delta, double* divU) {
  if (int rc = check(h)) return rc;
  if (!h->gdc_mid) return fail(FCM_ERR_STATE, "rb_advance_to_midstep first");
  if (!(fd_delta > 0.0)) return fail(FCM_ERR_ARG, "fd_delta must be positive");
  const int NB = h->NB;
  for (int i = 0; i < 6; ++i) {
    const double e = i < 3 ? fd_delta * h->cfg.a : fd_delta;
    HIPCHK(copy_bodies(h, h->d_x, h->d_q, h->d_xs, h->d_qs));
    HIPCHK(fill_body6(h->d_b6, NB, i, e, h->st));
    HIPCHK(bodies_step(h->g, h->d_x, h->d_q, NB, h->d_b6, 1.0, h->st));
    if (int rc = place_bodies(h)) return rc;
    HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fprand(), h->Vtmp, KTerm{}, h->st));
    BodyData bd = h->body();
    HIPCHK(mul_KT(bd, h->Vtmp, h->d_b6, h->st));
    HIPCHK(body_fit(bd, h->d_b6, h->d_KTKinv, h->d_Ud, h->st));
    HIPCHK(divu_sum(h->d_Ud, h->d_U0, NB, i, e, h->d_divu, h->st));
  }
  double acc[6];
  HIPCHK(hipMemcpyAsync(acc, h->d_divu, 6 * sizeof(double), hipMemcpyDeviceToHost, h->st));
  // back to the mid-step configuration (fill_all_blob_pos_at_midstep, main.cpp:263)
  HIPCHK(copy_bodies(h, h->d_x, h->d_q, h->d_xm, h->d_qm));
  if (int rc = place_bodies(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  double d = 0.0;
  for (int i = 0; i < 6; ++i) d += acc[i];
  if (divU) *divU = d;
  return FCM_OK;
}

int fcm_gdc_step(fcm_handle h, uint64_t seed, double cutoff, double fd_delta, const gmres_opts* opts,
                 gmres_report* rep, double* divU_out, int* redo) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies || !h->have_ref) return fail(FCM_ERR_STATE, "bodies and reference matrices must be set");
  if (!h->stress[0] || !(h->cfg.kT > 0.0)) return fail(FCM_ERR_STATE, "the drifter-corrector step needs kT > 0 and the random field");
  if (!(fd_delta > 0.0)) return fail(FCM_ERR_ARG, "fd_delta must be positive");
  const fcm_config& c = h->cfg;
  if (redo) *redo = 0;
  // main.cpp:230-233: the fluctuating field and its blob velocities at the start
  const double var1 = std::sqrt(2.0 * c.kT * c.mu / (c.dx * c.dx * c.dx * c.dt));
  const double var2 = std::sqrt(2.0) * var1;
  const double scale[6] = {var2, var2, var2, var1, var1, var1};
  HIPCHK(random_velocity<double>(h->g, h->fpstress(), true, (unsigned long long)seed, scale, h->pp, h->tw64, h->st));
  h->have_urand = true;
  HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fprand(), h->Vrand, KTerm{}, h->st));
  int overlap = 0;
  if (int rc = rb_advance_to_midstep(h, nullptr, &overlap)) return rc;
  if (overlap) {
    if (redo) *redo = 1;
    return FCM_OK;
  }
  double divU = 0.0;
  if (int rc = rb_calculate_divU(h, fd_delta, &divU)) return rc;
  const double nu = 1.0 + 0.5 * c.dt * divU;  // main.cpp:254
  if (divU_out) *divU_out = divU;
  // main.cpp:262-266, 278, 308-312 at the mid-step: blob velocities of the same
  // field, forces, the RHS and the solve; then the start bodies move by nu DT V
  HIPCHK(interp<double>(h->g, h->bb, h->bd, h->fprand(), h->Vrand, KTerm{}, h->st));
  h->have_blob_forces = false;
  if (cutoff > 0.0)
    if (int rc = ll_barrier_forces(h, cutoff, nullptr)) return rc;
  const double* F6 = nullptr;
  if (int rc = body_forces_dev(h, &F6)) return rc;
  if (int rc = compute_shear_velocities(h)) return rc;  // main.cpp:270-274
  const double* Vs = nullptr;
  if (int rc = shear_velocities_dev(h, &Vs)) return rc;
  HIPCHK(rhs_assemble(h->Vrand, Vs, F6, h->rhs, h->n3(), 6 * (size_t)h->NB, h->st));
  h->have_rhs = true;
  if (int rc = solve_single(h, opts, rep, nullptr)) return rc;
  return rb_end_of_step_update(h, nullptr, nu);
}

int rb_channel_wall_overlap(fcm_handle h, int* overlap) {
  if (int rc = check(h)) return rc;
  if (!overlap) return fail(FCM_ERR_ARG, "null argument");
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  *overlap = 0;
  if (!h->g.channel) return FCM_OK;
  HIPCHK(wall_overlap(h->dY, h->NB * h->nbpb, 0.5 * h->g.dx * h->g.ny, h->d_flag, h->st));
  HIPCHK(hipMemcpyAsync(overlap, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int rb_reset_to_start_of_step(fcm_handle h) {
  if (int rc = check(h)) return rc;
  if (!h->have_prev) return fail(FCM_ERR_STATE, "no end-of-step update to undo");
  HIPCHK(hipMemcpyAsync(h->d_x, h->d_x0, 3 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->d_q, h->d_q0, 4 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToDevice, h->st));
  BodyData b = h->body();
  HIPCHK(bodies_update(h->g, b, h->st));
  if (int rc = setup_blobs_device(h, h->NB * h->nbpb)) return rc;
  h->have_prev = false;
  h->have_blob_forces = false;
  h->have_rhs = false;
  return sync_stream(h);
}

int rb_get_bodies(fcm_handle h, double* x, double* q) {
  if (int rc = check(h)) return rc;
  if (!h->have_bodies) return fail(FCM_ERR_STATE, "rb_set_bodies first");
  if (x) HIPCHK(hipMemcpyAsync(x, h->d_x, 3 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToHost, h->st));
  if (q) HIPCHK(hipMemcpyAsync(q, h->d_q, 4 * (size_t)h->NB * sizeof(double), hipMemcpyDeviceToHost, h->st));
  return sync_stream(h);
}

int fcm_em_step(fcm_handle h, uint64_t seed, double cutoff, const gmres_opts* opts, gmres_report* rep) {
  if (int rc = check(h)) return rc;
  if (cutoff > 0.0)
    if (int rc = ll_barrier_forces(h, cutoff, nullptr)) return rc;
  if (int rc = fcm_mobility_prepare(h, seed)) return rc;
  if (int rc = solve_single(h, opts, rep, nullptr)) return rc;
  return rb_end_of_step_update(h, nullptr, 1.0);
}

}  // extern "C"
