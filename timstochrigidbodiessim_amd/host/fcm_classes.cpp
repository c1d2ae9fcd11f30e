// fcm_classes.cpp -- the reference's host classes over libfcm_hip.so.
#include "fcm_classes.hpp"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>

namespace tsrb {

void check(int rc) {
  if (rc != FCM_OK) throw fcm_error(rc, std::string("libfcm_hip: ") + fcm_last_error());
}

// ---- quaternion -------------------------------------------------------------------
quaternion::quaternion(double q0, double q1, double q2, double q3) : scalar_part(q0), vector_part{q1, q2, q3} {}

double quaternion::norm() const {
  return std::sqrt(scalar_part * scalar_part + vector_part[0] * vector_part[0] + vector_part[1] * vector_part[1] +
                   vector_part[2] * vector_part[2]);
}

void quaternion::normalise_in_place() {
  const double n = norm();
  scalar_part /= n;
  for (double& v : vector_part) v /= n;
}

quaternion quaternion::conj() const { return quaternion(scalar_part, -vector_part[0], -vector_part[1], -vector_part[2]); }

// R = I + 2 s [v]x + 2 [v]x^2 for the unit quaternion (s, v)
void quaternion::matrix(double R[9]) const {
  const double s = scalar_part, a = vector_part[0], b = vector_part[1], c = vector_part[2];
  R[0] = 1.0 - 2.0 * (b * b + c * c);
  R[1] = 2.0 * (a * b - s * c);
  R[2] = 2.0 * (a * c + s * b);
  R[3] = 2.0 * (a * b + s * c);
  R[4] = 1.0 - 2.0 * (a * a + c * c);
  R[5] = 2.0 * (b * c - s * a);
  R[6] = 2.0 * (a * c - s * b);
  R[7] = 2.0 * (b * c + s * a);
  R[8] = 1.0 - 2.0 * (a * a + b * b);
}

void quaternion::rotate(const double v[3], double out[3]) const {
  double R[9];
  matrix(R);
  for (int i = 0; i < 3; ++i) out[i] = R[3 * i] * v[0] + R[3 * i + 1] * v[1] + R[3 * i + 2] * v[2];
}

quaternion quaternion::operator*(const quaternion& p) const {
  const double* u = vector_part;
  const double* w = p.vector_part;
  return quaternion(scalar_part * p.scalar_part - (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]),
                    scalar_part * w[0] + p.scalar_part * u[0] + (u[1] * w[2] - u[2] * w[1]),
                    scalar_part * w[1] + p.scalar_part * u[1] + (u[2] * w[0] - u[0] * w[2]),
                    scalar_part * w[2] + p.scalar_part * u[2] + (u[0] * w[1] - u[1] * w[0]));
}

quaternion quaternion::operator*(double s) const {
  return quaternion(s * scalar_part, s * vector_part[0], s * vector_part[1], s * vector_part[2]);
}

quaternion quaternion::operator+(const quaternion& p) const {
  return quaternion(scalar_part + p.scalar_part, vector_part[0] + p.vector_part[0], vector_part[1] + p.vector_part[1],
                    vector_part[2] + p.vector_part[2]);
}

quaternion lie_exp(const double u[3]) {
  const double theta = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (theta < 1e-10) return quaternion(1.0, 0.0, 0.0, 0.0);
  const double cs = std::cos(0.5 * theta), sn = std::sin(0.5 * theta);
  return quaternion(cs, sn * u[0] / theta, sn * u[1] / theta, sn * u[2] / theta);
}

void particle::update(const double g[6], const double L[3]) {
  for (int k = 0; k < 3; ++k) {
    x[k] += g[k];
    x[k] -= L[k] * std::floor(x[k] / L[k]);
  }
  q = lie_exp(g + 3) * q;
}

// ---- gaussian_random ---------------------------------------------------------------
namespace {
uint64_t g_state = 0x853c49e6748fea9bULL;

uint64_t splitmix64() {
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double uniform01() { return (double)(splitmix64() >> 11) * (1.0 / 9007199254740992.0); }
}  // namespace

void seed_gaussian_random(uint64_t seed) { g_state = seed; }

double gaussian_random() {
  for (;;) {
    const double x1 = 2.0 * uniform01() - 1.0, x2 = 2.0 * uniform01() - 1.0;
    const double w = x1 * x1 + x2 * x2;
    if (w < 1.0 && w > 0.0) return x1 * std::sqrt(-2.0 * std::log(w) / w);
  }
}

// ---- fft_solver ------------------------------------------------------------------------
void fft_solver::initialise(fcm_handle h, const fcm_config& cfg) {
  h_ = h;
  pad = 2 * (cfg.nz / 2 + 1);
  local_nx = cfg.nx;
  local_x_start = 0;
  local_ny_after_transpose = cfg.ny;
  local_y_start_after_transpose = 0;
  total_local_size = cfg.nx * cfg.ny * pad;
}

void fft_solver::transform(double* in) {
  check(fcm_set_field(h_, 0, FCM_PREC_FP64, in));
  check(fft_transform(h_, 0, FCM_PREC_FP64));
  check(fcm_get_kspace(h_, 0, FCM_PREC_FP64, in));
}

void fft_solver::inverse_transform(double* in) {
  check(fcm_set_kspace(h_, 0, FCM_PREC_FP64, in));
  check(fft_inverse_transform(h_, 0, FCM_PREC_FP64));
  check(fcm_get_field(h_, 0, FCM_PREC_FP64, in));
}

// ---- fcm_fluid_solver -------------------------------------------------------------------
fcm_fluid_solver::~fcm_fluid_solver() {
  if (h_) fcm_destroy(h_);
}

void fcm_fluid_solver::initialise(const fcm_config& cfg, int total_num_blobs) {
  if (h_) fcm_destroy(h_);
  h_ = nullptr;
  cfg_ = cfg;
  if (total_num_blobs > cfg.max_blobs) throw fcm_error(FCM_ERR_CAPACITY, "total_num_blobs > max_blobs");
  check(fcm_create(&cfg_, &h_));
  num_blobs_ = total_num_blobs;
  V.assign(3 * (size_t)total_num_blobs, 0.0);
  fft.initialise(h_, cfg_);
}

void fcm_fluid_solver::set_up_gaussians(const std::vector<double>& Y) {
  num_blobs_ = (int)(Y.size() / 3);
  V.assign(Y.size(), 0.0);
  check(fcm_set_up_gaussians(h_, Y.data(), num_blobs_));
}

void fcm_fluid_solver::set_up_force_distribution(const std::vector<double>& F) {
  check(fcm_spread(h_, F.data(), FCM_PREC_FP64));
}

void fcm_fluid_solver::produce_deterministic_velocity_field() { check(fcm_stokes(h_, FCM_PREC_FP64)); }

void fcm_fluid_solver::calculate_deterministic_blob_velocities() {
  V.resize(3 * (size_t)num_blobs_);
  check(fcm_interp(h_, V.data(), FCM_PREC_FP64));
}

void fcm_fluid_solver::produce_random_velocity_field(uint64_t seed) { check(fcm_random_field(h_, nullptr, seed)); }

void fcm_fluid_solver::calculate_random_blob_velocities() {
  V.resize(3 * (size_t)num_blobs_);
  check(fcm_random_interp(h_, V.data()));
}

void fcm_fluid_solver::produce_shear_velocity_field(double shear_force_scale, double shear_wavenumber) {
  check(fcm_shear_field(h_, shear_force_scale, shear_wavenumber));
}

void fcm_fluid_solver::calculate_shear_blob_velocities() {
  Vshear.resize(3 * (size_t)num_blobs_);
  check(fcm_shear_interp(h_, Vshear.data()));
}

std::vector<double> fcm_fluid_solver::x_velocity_profile(const std::vector<double>& lambda) {
  std::vector<double> p(cfg_.ny);
  check(fcm_x_velocity_profile(h_, lambda.data(), p.data()));
  return p;
}

void fcm_fluid_solver::produce_reference_matrices(collection_of_rigid_bodies& all_bodies) {
  const size_t n3 = 3 * (size_t)all_bodies.blobs_per_body;
  ref_Minv.assign(n3 * n3, 0.0);
  ref_N.assign(36, 0.0);
  check(fcm_ref_matrices(h_, nullptr, ref_Minv.data(), ref_N.data()));
}

std::vector<double> fcm_fluid_solver::field(int which) const {
  std::vector<double> out((size_t)cfg_.nx * cfg_.ny * 2 * (cfg_.nz / 2 + 1));
  check(fcm_get_field(h_, which, FCM_PREC_FP64, out.data()));
  return out;
}

// ---- linked_list ----------------------------------------------------------------------------
void linked_list::initialise(fcm_handle h, double cutoff_radius) {
  h_ = h;
  cutoff = cutoff_radius;
}

void linked_list::apply_barrier_forces(std::vector<double>& blob_forces) const {
  if (!(cutoff > 0.0)) return;
  std::vector<double> F(blob_forces.size());
  check(ll_barrier_forces(h_, cutoff, F.data()));
  for (size_t i = 0; i < F.size(); ++i) blob_forces[i] += F[i];
}

// ---- collection_of_rigid_bodies ----------------------------------------------------------------
void collection_of_rigid_bodies::initialise(fcm_fluid_solver& fcm, const std::vector<double>& ref_pos,
                                            const std::vector<particle>& init, double steric_cutoff) {
  h_ = fcm.handle();
  const fcm_config& c = fcm.config();
  L_[0] = c.dx * c.nx;
  L_[1] = c.dx * c.ny;
  L_[2] = c.dx * c.nz;
  blob_ref_pos = ref_pos;
  blobs_per_body = (int)(ref_pos.size() / 3);
  bodies = init;
  total_num_blobs = blobs_per_body * (int)bodies.size();
  blob_forces.assign(3 * (size_t)total_num_blobs, 0.0);
  body_forces.assign(6 * bodies.size(), 0.0);
  check(rb_set_geometry(h_, blobs_per_body, blob_ref_pos.data()));
  list.initialise(h_, steric_cutoff);
  fill_all_blob_pos();
}

void collection_of_rigid_bodies::upload() const {
  std::vector<double> x(3 * bodies.size()), q(4 * bodies.size());
  for (size_t n = 0; n < bodies.size(); ++n) {
    for (int k = 0; k < 3; ++k) x[3 * n + k] = bodies[n].x[k];
    q[4 * n] = bodies[n].q.scalar_part;
    for (int k = 0; k < 3; ++k) q[4 * n + 1 + k] = bodies[n].q.vector_part[k];
  }
  check(rb_set_bodies(h_, (int)bodies.size(), x.data(), q.data()));
}

void collection_of_rigid_bodies::fill_all_blob_pos() {
  upload();
  all_blob_pos.assign(3 * (size_t)total_num_blobs, 0.0);
  check(rb_get_blob_positions(h_, all_blob_pos.data()));
}

void collection_of_rigid_bodies::calculate_forces() {
  std::fill(blob_forces.begin(), blob_forces.end(), 0.0);
  list.apply_barrier_forces(blob_forces);
}

std::vector<double> collection_of_rigid_bodies::multiply_by_K_at_midstep(const std::vector<double>& in) const {
  std::vector<double> out(3 * (size_t)total_num_blobs);
  check(rb_mul_K(h_, in.data(), out.data()));
  return out;
}

std::vector<double> collection_of_rigid_bodies::multiply_by_KT_at_midstep(const std::vector<double>& in) const {
  std::vector<double> out(6 * bodies.size());
  check(rb_mul_KT(h_, in.data(), out.data()));
  return out;
}

bool collection_of_rigid_bodies::end_of_step_update(const std::vector<double>& V, double nu) {
  check(rb_end_of_step_update(h_, V.data(), nu));
  std::vector<double> x(3 * bodies.size()), q(4 * bodies.size());
  check(rb_get_bodies(h_, x.data(), q.data()));
  for (size_t n = 0; n < bodies.size(); ++n) {
    for (int k = 0; k < 3; ++k) bodies[n].x[k] = x[3 * n + k];
    bodies[n].q = quaternion(q[4 * n], q[4 * n + 1], q[4 * n + 2], q[4 * n + 3]);
  }
  all_blob_pos.assign(3 * (size_t)total_num_blobs, 0.0);
  check(rb_get_blob_positions(h_, all_blob_pos.data()));
  return check_for_channel_wall_overlap();
}

bool collection_of_rigid_bodies::advance_to_midstep(const std::vector<double>& V) {
  int overlap = 0;
  check(rb_advance_to_midstep(h_, V.data(), &overlap));
  return overlap != 0;
}

double collection_of_rigid_bodies::calculate_divU(fcm_fluid_solver&, double delta) {
  double d = 0.0;
  check(rb_calculate_divU(h_, delta, &d));
  return d;
}

void collection_of_rigid_bodies::fill_all_blob_pos_at_midstep() {
  all_blob_pos.assign(3 * (size_t)total_num_blobs, 0.0);
  check(rb_get_blob_positions(h_, all_blob_pos.data()));
}

bool collection_of_rigid_bodies::check_for_channel_wall_overlap() const {
  int overlap = 0;
  check(rb_channel_wall_overlap(h_, &overlap));
  return overlap != 0;
}

void collection_of_rigid_bodies::reset_to_start_of_step() {
  check(rb_reset_to_start_of_step(h_));
  std::vector<double> x(3 * bodies.size()), q(4 * bodies.size());
  check(rb_get_bodies(h_, x.data(), q.data()));
  for (size_t n = 0; n < bodies.size(); ++n) {
    for (int k = 0; k < 3; ++k) bodies[n].x[k] = x[3 * n + k];
    bodies[n].q = quaternion(q[4 * n], q[4 * n + 1], q[4 * n + 2], q[4 * n + 3]);
  }
  all_blob_pos.assign(3 * (size_t)total_num_blobs, 0.0);
  check(rb_get_blob_positions(h_, all_blob_pos.data()));
}

namespace {
void put_sci(std::ofstream& f, double v) { f << v << ' '; }

// the reference's data files are appended to, std::scientific, precision 6
std::ofstream open_append(const std::string& path) {
  std::ofstream f(path, std::ios::app);
  if (!f) throw fcm_error(FCM_ERR_ARG, "cannot open " + path);
  f << std::scientific << std::setprecision(6);
  return f;
}
}  // namespace

void collection_of_rigid_bodies::write_data_to_file(double t, const std::string& name) const {
  {
    std::ofstream f = open_append(name + ".dat");
    put_sci(f, t);
    for (const particle& b : bodies) {
      for (int k = 0; k < 3; ++k) put_sci(f, b.x[k]);
      put_sci(f, b.q.scalar_part);
      for (int k = 0; k < 3; ++k) put_sci(f, b.q.vector_part[k]);
    }
    f << '\n';
  }
  std::ofstream f = open_append(name + "_blob_forces.dat");
  put_sci(f, t);
  for (double v : blob_forces) put_sci(f, v);
  f << '\n';
}

void collection_of_rigid_bodies::write_blob_references(const std::string& name) const {
  std::ofstream f(name + "_blob_references.dat");
  if (!f) throw fcm_error(FCM_ERR_ARG, "cannot open " + name + "_blob_references.dat");
  f << std::scientific << std::setprecision(6);
  for (int d = 0; d < 3; ++d) {
    for (int m = 0; m < blobs_per_body; ++m) put_sci(f, blob_ref_pos[3 * (size_t)m + d]);
    f << '\n';
  }
}

int collection_of_rigid_bodies::read_initial_conditions(const std::string& file, int sample) {
  const size_t nbody = bodies.size();
  std::vector<double> state(7 * nbody), chosen;
  std::ifstream f(file);
  if (!f) throw fcm_error(FCM_ERR_ARG, "cannot open " + file);
  // one state: a time, then x y z q0 q1 q2 q3 per body (whitespace separated,
  // lines do not matter), counted as the reference counts them (:272-292)
  int n = 0;
  double t;
  while (f >> t) {
    for (size_t i = 0; i < 7 * nbody; ++i) f >> state[i];
    if (!f) break;
    ++n;
    if (sample == 0 || n == sample) chosen = state;
  }
  if (n == 0) throw fcm_error(FCM_ERR_ARG, file + " holds no complete state");
  if (sample < 0 || sample > n) throw fcm_error(FCM_ERR_ARG, "state " + std::to_string(sample) + " of " +
                                                              std::to_string(n) + " requested from " + file);
  for (size_t i = 0; i < nbody; ++i) {
    const double* s = chosen.data() + 7 * i;
    for (int k = 0; k < 3; ++k) bodies[i].x[k] = s[k];
    bodies[i].q = quaternion(s[3], s[4], s[5], s[6]);
  }
  if (h_) fill_all_blob_pos();  // not yet initialised: initialise uploads them
  return n;
}

void write_par_file(const std::string& name, const run_parameters& p) {
  std::ofstream f(name + ".par");
  if (!f) throw fcm_error(FCM_ERR_ARG, "cannot open " + name + ".par");
  f << p.num_bodies << " %% NUM_BODIES\n";
  f << p.blob_radius << " %% BLOB_RADIUS\n";
  f << p.fluid_viscosity << " %% FLUID_VISCOSITY\n";
  f << p.energy_scale << " %% ENERGY_SCALE\n";
  f << p.LX << " %% LX\n";
  f << p.LY << " %% LY\n";
  f << p.LZ << " %% LZ\n";
  f << p.slip_channel << " %% SLIP_CHANNEL\n";
  f << p.steps_per_diffusive_time << " %% STEPS_PER_DIFFUSIVE_TIME\n";
  f << p.num_diffusive_times << " %% NUM_DIFFUSIVE_TIMES\n";
  f << p.data_saves_per_diffusive_time << " %% DATA_SAVES_PER_DIFFUSIVE_TIME\n";
  if (p.shearing) {
    f << p.shear_force_scale << " %% SHEAR_FORCE_SCALE\n";
    f << p.shear_wavenumber << " %% SHEAR_WAVENUMBER\n";
  }
}

void append_fluid_x_velocity(const std::string& name, double t, const std::vector<double>& profile) {
  std::ofstream f(name + "_fluid_x_velocity_in_y.dat", std::ios::app);
  if (!f) throw fcm_error(FCM_ERR_ARG, "cannot open " + name + "_fluid_x_velocity_in_y.dat");
  f << t << ' ';
  for (double u : profile) f << u << ' ';
  f << '\n';
}

void append_velocities(const std::string& name, double t, const std::vector<double>& soln, int total_num_blobs,
                       int num_bodies) {
  const size_t off = 3 * (size_t)total_num_blobs, nv = 6 * (size_t)num_bodies;
  if (soln.size() < off + nv) throw fcm_error(FCM_ERR_ARG, "solution shorter than [lambda; V]");
  std::ofstream f = open_append(name + "_velocities.dat");
  put_sci(f, t);
  for (size_t i = 0; i < nv; ++i) put_sci(f, soln[off + i]);
  f << '\n';
}

// ---- gmres_solver -----------------------------------------------------------------------------
void gmres_solver::initialise(int total_num_blobs, int num_bodies, const gmres_opts& o) {
  num_blobs = total_num_blobs;
  system_size = 3 * total_num_blobs + 6 * num_bodies;
  RHS.assign(system_size, 0.0);
  opts = o;
}

void gmres_solver::assemble_RHS(const collection_of_rigid_bodies& all_bodies, const fcm_fluid_solver& fcm) {
  const std::vector<double> KTf = all_bodies.multiply_by_KT_at_midstep(all_bodies.blob_forces);
  const size_t n3 = 3 * (size_t)num_blobs;
  for (size_t i = 0; i < n3; ++i) RHS[i] = -fcm.V[i];
  if (fcm.Vshear.size() == n3)  // SHEARING_FORCES (gmres_solver.cpp:62-64)
    for (size_t i = 0; i < n3; ++i) RHS[i] -= fcm.Vshear[i];
  for (size_t i = 0; i < KTf.size(); ++i) RHS[n3 + i] = -(KTf[i] + all_bodies.body_forces[i]);
}

std::vector<double> gmres_solver::multiply_by_Pinv(const std::vector<double>& in, const collection_of_rigid_bodies&,
                                                   const fcm_fluid_solver& fcm) {
  std::vector<double> out(in.size());
  check(gmres_apply_Pinv(fcm.handle(), in.data(), out.data()));
  return out;
}

std::vector<double> gmres_solver::multiply_by_saddle_point_matrix(const std::vector<double>& in,
                                                                  const collection_of_rigid_bodies&,
                                                                  fcm_fluid_solver& fcm) {
  std::vector<double> out(in.size());
  check(gmres_saddle_matvec(fcm.handle(), in.data(), out.data(), FCM_PREC_FP64));
  return out;
}

std::vector<double> gmres_solver::invert_saddle_point_system(const collection_of_rigid_bodies&, fcm_fluid_solver& fcm,
                                                             int, double) {
  std::vector<double> soln(system_size);
  residual_history.assign(opts.max_iter > 0 ? opts.max_iter : 1, 0.0);
  check(gmres_solve(fcm.handle(), RHS.data(), soln.data(), &opts, &report, residual_history.data()));
  residual_history.resize(report.iterations);
  return soln;
}

}  // namespace tsrb
