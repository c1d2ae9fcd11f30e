// fcm_handle.hpp -- the state behind an fcm_handle and the host helpers
// shared by the C ABI (fcm_capi.hip) and the solve/batch driver
// (fcm_batch.hip).  Internal to libfcm_hip.so.
#pragma once
#include "fcm_internal.hpp"
#include "../../include/fcm_hip.h"

#include <string>
#include <vector>

struct fcm_batch_s;

namespace fcm_host {
using namespace fcm;

int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) return ::fcm_host::fail(FCM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace fcm_host

using namespace fcm;

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
  // steady shear velocity field (produce_shear_velocity_field), allocated by
  // the first fcm_shear_field; its blob velocities join every RHS while set
  c2<double>* ushear[3] = {nullptr, nullptr, nullptr};
  double* Vshear = nullptr;
  double* d_prof = nullptr;
  bool have_shear = false;
  bool have_vshear = false;     // Vshear holds calculate_shear_blob_velocities' result
  int vshear_nb = 0;            // for this many blobs
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
  double *d_Minv = nullptr, *d_MinvT = nullptr, *d_N = nullptr, *d_F6 = nullptr;
  double* d_fblob = nullptr;    // 3 Nb blob forces (barrier forces), blob order
  double* d_F6tot = nullptr;    // 6 NB body forces + K^T blob forces, for the RHS
  double* d_V6 = nullptr;       // 6 NB body velocities for the end-of-step update
  double *d_x0 = nullptr, *d_q0 = nullptr;  // bodies before the last end-of-step update
  // generalised drifter-corrector step: start-of-step and mid-step bodies, the
  // rigid fits U0 and U (6 NB each), a 6 NB scratch, (K^T K)^-1 of the
  // reference body and the six finite-difference sums
  double *d_xs = nullptr, *d_qs = nullptr, *d_xm = nullptr, *d_qm = nullptr;
  double *d_U0 = nullptr, *d_Ud = nullptr, *d_b6 = nullptr, *d_KTKinv = nullptr, *d_divu = nullptr;
  bool gdc_mid = false;  // the device bodies are a drifter-corrector mid-step (start in d_xs)
  int* d_flag = nullptr;        // one-word result of the wall-overlap check
  bool have_prev = false;
  StericWork sw;                // cell list of the barrier forces
  int ll_cells = 0;             // > 0: the reference's cells^3 interaction set (ll_set_cells)
  std::vector<void*> lazy;      // cell arrays, grown on demand
  bool have_blob_forces = false;
  bool have_geometry = false, have_bodies = false, have_ref = false, have_forces = false;
  std::vector<double> h_ref;

  // vectors
  size_t nmax = 0, ldq = 0;
  double* Q = nullptr;        // (max_krylov + 1) columns of ldq words: FP64, or FP32 in the inner solves
  float* Qf() const { return reinterpret_cast<float*>(Q); }  // of FP32+refine (the same columns, ldq floats apart)
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
  double* hist_all = nullptr; // residual history of the whole (refined) solve, max_krylov + 2
  unsigned long long epoch = 0; // bumped whenever kernel arguments of an Arnoldi step change
  bool tuning_set = false;      // launch shapes set explicitly (fcm_set_tuning, fcm_autotune*)
  bool cache_checked = false;   // the shape-class cache has been consulted
  fcm_batch_s* solo = nullptr;  // single-solve driver on this handle's stream
  std::vector<fcm_batch_s*> batches;  // batches whose captured graphs may reference this handle
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
    b.MinvT = d_MinvT;
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
  FieldPtrs<double> fpshear() const {
    FieldPtrs<double> p{};
    for (int i = 0; i < 3; ++i) p.p[i] = ushear[i];
    return p;
  }
};


namespace fcm_host {

// mobility product on device vectors: Vout = dx^3 J S Fin (- K zV if given)
int mobility_dev(fcm_handle h, const double* Fin, double* Vout, int precision, const KTerm& kt, hipStream_t st,
                 const int* gate = nullptr);
// saddle-point product (gmres_solver.cpp:72-83) on device vectors
int saddle_dev(fcm_handle h, const double* in, double* out, int precision, hipStream_t st);
// one device-controlled Arnoldi step of the handle's GMRES state
int launch_arnoldi_step(fcm_handle h, int precision, hipStream_t st);
// solve the assembled right-hand side h->rhs into h->sol on the handle's stream
int solve_single(fcm_handle h, const gmres_opts* o, gmres_report* rep, double* hist);
// detach a handle from every batch that may still reference it (before it is freed)
void batch_forget(fcm_handle h);
void batch_destroy_solo(fcm_handle h);

}  // namespace fcm_host
