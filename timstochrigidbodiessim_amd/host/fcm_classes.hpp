// fcm_classes.hpp -- the reference's host classes, kept by name and meaning,
// over the C ABI of libfcm_hip.so (include/fcm_hip.h).
//
// A program written against the reference (fcm_fluid_solver.hpp,
// fft_solver.hpp, gmres_solver.hpp, collection_of_rigid_bodies.hpp,
// linked_list.hpp, particle.hpp, quaternion.hpp, parallel_random_numbers.hpp)
// keeps its call sequence: the same classes, methods and public members, with
// std::vector<double> in place of Armadillo vectors (blob positions Nb x 3
// row-major, blob vectors interleaved [x0,y0,z0,...], body vectors [U; Omega]
// per body, GMRES vectors [lambda; V]).  What differs:
//   * compile-time config.hpp becomes a runtime fcm_config passed to
//     fcm_fluid_solver::initialise;
//   * one process drives the whole grid on one GPU (no MPI ranks), and all grid
//     state lives on the device: the solver owns an fcm_handle that the other
//     classes share;
//   * produce_random_velocity_field takes the seed of the stress stream (the
//     reference draws from an unseeded SPRNG stream);
//   * errors are exceptions (fcm_error) carrying the library's status code.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fcm_hip.h"

namespace tsrb {

struct fcm_error : std::runtime_error {
  int code;
  fcm_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// throws fcm_error unless rc == FCM_OK
void check(int rc);

// ---- quaternion.hpp / quaternion.cpp ------------------------------------------
class quaternion {
 public:
  double scalar_part = 1.0;
  double vector_part[3] = {0.0, 0.0, 0.0};

  quaternion() = default;
  quaternion(double q0, double q1, double q2, double q3);
  double norm() const;
  void normalise_in_place();
  quaternion conj() const;
  // quaternion::matrix (quaternion.cpp:61-103), row-major
  void matrix(double R[9]) const;
  void rotate(const double v[3], double out[3]) const;
  quaternion operator*(const quaternion& p) const;
  quaternion operator*(double s) const;
  quaternion operator+(const quaternion& p) const;
};

// quaternion.cpp:128-145
quaternion lie_exp(const double u[3]);

// ---- particle.hpp ---------------------------------------------------------------
class particle {
 public:
  double x[3] = {0.0, 0.0, 0.0};
  quaternion q;
  // particle::update(const vec& grand_update) (particle.cpp:45-56) in a box L
  void update(const double grand_update[6], const double L[3]);
};

// ---- parallel_random_numbers.hpp -------------------------------------------------
// gaussian_random (parallel_random_numbers.cpp:8-19): Marsaglia's polar method,
// here over a seeded SplitMix64 stream instead of SPRNG.
void seed_gaussian_random(uint64_t seed);
double gaussian_random();

// ---- fft_solver.hpp ---------------------------------------------------------------
// In-place unnormalised R2C / C2R of a host field in the reference layout
// ([nx][ny][pad] reals in, transposed-order complex out), on the device.
class fft_solver {
 public:
  int local_nx = 0, local_x_start = 0, local_ny_after_transpose = 0, local_y_start_after_transpose = 0;
  int total_local_size = 0, pad = 0;
  void initialise(fcm_handle h, const fcm_config& cfg);
  void transform(double* in);
  void inverse_transform(double* in);

 private:
  fcm_handle h_ = nullptr;
};

class collection_of_rigid_bodies;

// ---- fcm_fluid_solver.hpp ---------------------------------------------------------
class fcm_fluid_solver {
 public:
  std::vector<double> V;         // blob velocities of the last interpolation (3 Nb)
  std::vector<double> Vshear;    // shear-flow blob velocities (SHEARING_FORCES), 3 Nb
  std::vector<double> ref_Minv;  // (3 n_b)^2, row-major
  std::vector<double> ref_N;     // 6 x 6
  fft_solver fft;

  fcm_fluid_solver() = default;
  ~fcm_fluid_solver();
  fcm_fluid_solver(const fcm_fluid_solver&) = delete;
  fcm_fluid_solver& operator=(const fcm_fluid_solver&) = delete;

  // initialise (fcm_fluid_solver.cpp:29-129); total_num_blobs <= cfg.max_blobs
  void initialise(const fcm_config& cfg, int total_num_blobs);
  void set_up_gaussians(const std::vector<double>& Y);
  void set_up_force_distribution(const std::vector<double>& F);
  void produce_deterministic_velocity_field();
  void calculate_deterministic_blob_velocities();
  void produce_random_velocity_field(uint64_t seed);
  void calculate_random_blob_velocities();
  void produce_reference_matrices(collection_of_rigid_bodies& all_bodies);
  // SHEARING_FORCES (fcm_fluid_solver.cpp:488-550, 703-840): the steady shear
  // flow once, then its blob velocities at the current positions each step
  void produce_shear_velocity_field(double shear_force_scale, double shear_wavenumber);
  void calculate_shear_blob_velocities();
  // main.cpp:371-433: mean x velocity per y row of the flow of the blob forces
  // lambda plus the shear flow
  std::vector<double> x_velocity_profile(const std::vector<double>& lambda);

  // real-space field `which` (0..2 ux,uy,uz / fx,fy,fz after a spread; 3..5 the
  // random velocity) in the reference layout [nx][ny][pad]
  std::vector<double> field(int which) const;
  fcm_handle handle() const { return h_; }
  const fcm_config& config() const { return cfg_; }

 private:
  fcm_handle h_ = nullptr;
  fcm_config cfg_{};
  int num_blobs_ = 0;
};

// ---- linked_list.hpp ---------------------------------------------------------------
// The cell list lives on the device; apply_barrier_forces adds the soft-sphere
// forces (linked_list.cpp:192-281) of the current blob positions to blob_forces.
class linked_list {
 public:
  double cutoff = 0.0;
  void initialise(fcm_handle h, double cutoff_radius);
  void apply_barrier_forces(std::vector<double>& blob_forces) const;

 private:
  fcm_handle h_ = nullptr;
};

// ---- collection_of_rigid_bodies.hpp -------------------------------------------------
class collection_of_rigid_bodies {
 public:
  int blobs_per_body = 0;
  int total_num_blobs = 0;
  std::vector<double> blob_ref_pos;  // n_b x 3 body-frame offsets
  std::vector<double> all_blob_pos;  // Nb x 3
  std::vector<particle> bodies;
  std::vector<double> blob_forces;   // 3 Nb
  std::vector<double> body_forces;   // 6 NB external forces and torques
  linked_list list;

  // geometry and initial state; shares fcm's device state
  void initialise(fcm_fluid_solver& fcm, const std::vector<double>& ref_pos, const std::vector<particle>& init,
                  double steric_cutoff);
  void fill_all_blob_pos();
  void calculate_forces();
  std::vector<double> multiply_by_K_at_midstep(const std::vector<double>& in) const;
  std::vector<double> multiply_by_KT_at_midstep(const std::vector<double>& in) const;
  // collection_of_rigid_bodies.cpp:759-790; returns whether the step must be
  // redone: a blob on or beyond a slip-channel wall (SKIP_ON_OVERLAP is false,
  // so blob-blob overlaps never force a redo)
  bool end_of_step_update(const std::vector<double>& V, double nu);
  // :207-239 (false in a periodic box)
  bool check_for_channel_wall_overlap() const;
  // GEN_DRIFTER_CORRECTOR (:566-621, 869-953): the mid-step from the blob
  // velocities V (true: a wall overlap, nothing moved), divU by finite
  // differences of step delta (FINITE_DIFFERENCE_DELTA_FACTOR, config.hpp:73),
  // and the mid-step blob positions (:fill_all_blob_pos_at_midstep)
  bool advance_to_midstep(const std::vector<double>& V);
  double calculate_divU(fcm_fluid_solver& fcm, double delta = 1e-5);
  void fill_all_blob_pos_at_midstep();
  // :792-796: the bodies before the last end_of_step_update
  void reset_to_start_of_step();

  // ---- on-disk formats (collection_of_rigid_bodies.cpp:245-339, 798-834) ----
  // Appends one line "t x y z q0 q1 q2 q3 ..." (every body) to <name>.dat and one
  // line "t f_0 ... f_{3Nb-1}" to <name>_blob_forces.dat, "%.6e" numbers each
  // followed by a space (write_data_to_file, :798-834).
  void write_data_to_file(double t, const std::string& name) const;
  // <name>_blob_references.dat: three lines, the x, y and z body-frame offsets
  // of the blobs (seed_the_bodies, :245-263)
  void write_blob_references(const std::string& name) const;
  // Reads a state file in the .dat format (READ_IN_INITIAL_CONDITIONS, :265-339)
  // into bodies: sample 0 takes the last state (IC_READ_IN_LAST), sample k >= 1
  // the k-th (IC_READ_IN_RANDOM draws k uniformly from 1..n; the caller draws
  // it).  Returns the number of states n in the file; throws if it has none or
  // sample > n.  Blob positions are refreshed on the device.
  int read_initial_conditions(const std::string& file, int sample);

 private:
  fcm_handle h_ = nullptr;
  double L_[3] = {0.0, 0.0, 0.0};
  void upload() const;
};

// ---- main.cpp output files --------------------------------------------------------
// The run parameters main.cpp:36-66 writes to <SIMULATION_NAME>.par.
struct run_parameters {
  int num_bodies = 0;
  double blob_radius = 0.0, fluid_viscosity = 0.0, energy_scale = 0.0;
  double LX = 0.0, LY = 0.0, LZ = 0.0;
  int slip_channel = 0;
  double steps_per_diffusive_time = 0.0, num_diffusive_times = 0.0, data_saves_per_diffusive_time = 0.0;
  bool shearing = false;
  double shear_force_scale = 0.0, shear_wavenumber = 0.0;
};
// "<value> %% NAME" lines, values in the stream's default format (main.cpp:36-66)
void write_par_file(const std::string& name, const run_parameters& p);
// appends "t U_0 Omega_0 ... " (the body part of the saddle solution, "%.6e"
// each followed by a space) to <name>_velocities.dat (main.cpp:343-367)
void append_velocities(const std::string& name, double t, const std::vector<double>& soln, int total_num_blobs,
                       int num_bodies);
// appends "t u_0 ... u_{ny-1} " in the stream's default format to
// <name>_fluid_x_velocity_in_y.dat (main.cpp:371-433)
void append_fluid_x_velocity(const std::string& name, double t, const std::vector<double>& profile);

// ---- gmres_solver.hpp ----------------------------------------------------------------
class gmres_solver {
 public:
  int num_blobs = 0;
  int system_size = 0;
  std::vector<double> RHS;
  gmres_opts opts{1e-8, 200, FCM_PREC_FP64, 0.0, 0};
  gmres_report report{};
  std::vector<double> residual_history;

  void initialise(int total_num_blobs, int num_bodies, const gmres_opts& o);
  // RHS = [-fcm.V; -(K^T blob_forces + body_forces)] (gmres_solver.cpp:58-70)
  void assemble_RHS(const collection_of_rigid_bodies& all_bodies, const fcm_fluid_solver& fcm);
  std::vector<double> multiply_by_Pinv(const std::vector<double>& in, const collection_of_rigid_bodies& all_bodies,
                                       const fcm_fluid_solver& fcm);
  std::vector<double> multiply_by_saddle_point_matrix(const std::vector<double>& in,
                                                      const collection_of_rigid_bodies& all_bodies,
                                                      fcm_fluid_solver& fcm);
  std::vector<double> invert_saddle_point_system(const collection_of_rigid_bodies& all_bodies, fcm_fluid_solver& fcm,
                                                 int myrank, double t);
};

}  // namespace tsrb
