// fcm_host_check.cpp -- drives the reference-named host classes the way the
// reference's main.cpp time loop does (main.cpp:222-312, EULER_MARUYAMA).
//
//   fcm_host_check math
//       host-only checks of quaternion, lie_exp, particle::update and
//       gaussian_random (no GPU needed)
//   fcm_host_check steps <in.bin> <out.bin>
//       reads a problem, runs its time steps through fcm_fluid_solver,
//       collection_of_rigid_bodies, linked_list and gmres_solver, and writes
//       every step's GMRES solution and the final body state
//   fcm_host_check io <name> <state file> <sample>
//       writes <name>.dat (two saves), <name>_blob_forces.dat,
//       <name>_blob_references.dat, <name>.par, <name>_velocities.dat and
//       <name>_fluid_x_velocity_in_y.dat for
//       three two-sphere bodies with fixed values (no GPU needed), then reads
//       state <sample> (0: the last) of <state file> into the three bodies and
//       prints them, one body per line, "%.17g"
//
// in.bin (little-endian): int32 n, NB, nbpb, steps, precision; uint64 seed;
// float64 a, kT, dt, tol, cutoff; float64 ref[3 nbpb], x[3 NB], q[4 NB], F6[6 NB];
// optionally int32 slip_channel, float64 shear_force_scale, shear_wavenumber
// [, int32 algorithm: 0 Euler-Maruyama, 1 generalised drifter-corrector].
// out.bin: per step float64 soln[3 Nb + 6 NB] and int32 iterations; then
// float64 x[3 NB], q[4 NB].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fcm_classes.hpp"

using namespace tsrb;

namespace {

int fails = 0;

void expect(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++fails;
}

int run_math() {
  // lie_exp of a rotation about z by theta is that rotation
  const double th = 0.731;
  const double u[3] = {0.0, 0.0, th};
  double R[9];
  lie_exp(u).matrix(R);
  const double Rz[9] = {std::cos(th), -std::sin(th), 0, std::sin(th), std::cos(th), 0, 0, 0, 1};
  double err = 0;
  for (int i = 0; i < 9; ++i) err = std::fmax(err, std::fabs(R[i] - Rz[i]));
  expect(err < 1e-15, "lie_exp(theta e_z) is the rotation about z");
  // Hamilton product: i j = k
  const quaternion i(0, 1, 0, 0), j(0, 0, 1, 0);
  const quaternion k = i * j;
  expect(k.scalar_part == 0 && k.vector_part[0] == 0 && k.vector_part[1] == 0 && k.vector_part[2] == 1, "i j = k");
  // matrix(p q) = matrix(p) matrix(q)
  quaternion p(0.3, -0.2, 0.9, 0.1), q(-0.5, 0.4, 0.1, 0.7);
  p.normalise_in_place();
  q.normalise_in_place();
  double Rp[9], Rq[9], Rpq[9];
  p.matrix(Rp);
  q.matrix(Rq);
  (p * q).matrix(Rpq);
  err = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int m = 0; m < 3; ++m) s += Rp[3 * r + m] * Rq[3 * m + c];
      err = std::fmax(err, std::fabs(s - Rpq[3 * r + c]));
    }
  expect(err < 1e-14, "matrix(p q) = matrix(p) matrix(q)");
  // particle::update wraps into the box and composes the rotation on the left
  particle b;
  b.x[0] = 9.5;
  b.x[1] = 0.25;
  b.x[2] = 3.0;
  const double L[3] = {10.0, 10.0, 10.0};
  const double g[6] = {1.0, -0.5, 0.0, 0.0, 0.0, th};
  b.update(g, L);
  expect(std::fabs(b.x[0] - 0.5) < 1e-15 && std::fabs(b.x[1] - 9.75) < 1e-15 && b.x[2] == 3.0, "particle wraps");
  b.q.matrix(R);
  err = 0;
  for (int m = 0; m < 9; ++m) err = std::fmax(err, std::fabs(R[m] - Rz[m]));
  expect(err < 1e-15, "particle rotation is lie_exp(omega) q");
  // gaussian_random moments
  seed_gaussian_random(12345);
  double s1 = 0, s2 = 0, s4 = 0;
  const int N = 400000;
  for (int m = 0; m < N; ++m) {
    const double z = gaussian_random();
    s1 += z;
    s2 += z * z;
    s4 += z * z * z * z;
  }
  s1 /= N;
  s2 /= N;
  s4 /= N;
  expect(std::fabs(s1) < 0.01 && std::fabs(s2 - 1) < 0.01 && std::fabs(s4 - 3) < 0.05, "gaussian_random moments");
  return fails ? 1 : 0;
}

template <typename T> bool rd(std::FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }
template <typename T> void wr(std::FILE* f, const T* p, size_t n) { std::fwrite(p, sizeof(T), n, f); }

int run_steps(const char* in_path, const char* out_path) {
  std::FILE* f = std::fopen(in_path, "rb");
  if (!f) return 2;
  int hdr[5];
  uint64_t seed;
  double par[5];
  if (!rd(f, hdr, 5) || !rd(f, &seed, 1) || !rd(f, par, 5)) return 2;
  const int n = hdr[0], NB = hdr[1], nbpb = hdr[2], steps = hdr[3], precision = hdr[4];
  const double a = par[0], kT = par[1], dt = par[2], tol = par[3], cutoff = par[4];
  std::vector<double> ref(3 * nbpb), x(3 * NB), q(4 * NB), F6(6 * NB);
  if (!rd(f, ref.data(), ref.size()) || !rd(f, x.data(), x.size()) || !rd(f, q.data(), q.size()) ||
      !rd(f, F6.data(), F6.size()))
    return 2;
  // optional trailer: int32 slip_channel; float64 shear_force_scale, shear_wavenumber;
  // int32 algorithm (0 Euler-Maruyama, 1 generalised drifter-corrector)
  int chan = 0, gdc = 0;
  double shear[2] = {0.0, 0.0};
  if (rd(f, &chan, 1) && !rd(f, shear, 2)) return 2;
  if (!rd(f, &gdc, 1)) gdc = 0;
  std::fclose(f);
  const int nb = NB * nbpb;

  fcm_config cfg{n, n, n, 7, 1.0, a, 1.0, kT, dt, nb, NB, nbpb, 200, precision == FCM_PREC_FP32, 1, chan};
  fcm_fluid_solver fcm;
  fcm.initialise(cfg, nb);
  std::vector<particle> init(NB);
  for (int b = 0; b < NB; ++b) {
    for (int k = 0; k < 3; ++k) init[b].x[k] = x[3 * b + k];
    init[b].q = quaternion(q[4 * b], q[4 * b + 1], q[4 * b + 2], q[4 * b + 3]);
  }
  collection_of_rigid_bodies all_bodies;
  all_bodies.initialise(fcm, ref, init, cutoff);
  all_bodies.body_forces = F6;
  fcm.produce_reference_matrices(all_bodies);
  if (shear[0] != 0.0) fcm.produce_shear_velocity_field(shear[0], shear[1]);  // main.cpp:98-102
  gmres_solver gmres;
  gmres.initialise(nb, NB, gmres_opts{tol, 200, precision, 0.0, 0});

  std::FILE* o = std::fopen(out_path, "wb");
  if (!o) return 2;
  double t = 0;
  int redone = 0;
  for (int step = 0; step < steps;) {
    // main.cpp:230-233; a redone step draws from a seed of its own
    const uint64_t s = seed + step + 1000 * (uint64_t)redone;
    all_bodies.fill_all_blob_pos();
    fcm.produce_random_velocity_field(s);
    fcm.set_up_gaussians(all_bodies.all_blob_pos);
    fcm.calculate_random_blob_velocities();
    double nu = 1.0;
    if (gdc) {  // main.cpp:235-268
      if (all_bodies.advance_to_midstep(fcm.V)) {
        std::printf("step %d: redone (the mid-step reached a channel wall)\n", step);
        ++redone;
        continue;
      }
      const double divU = all_bodies.calculate_divU(fcm);
      nu = 1.0 + 0.5 * dt * divU;
      std::printf("step %d: divU %.17g\n", step, divU);
      all_bodies.fill_all_blob_pos_at_midstep();
      fcm.set_up_gaussians(all_bodies.all_blob_pos);
      fcm.calculate_random_blob_velocities();
      // main.cpp:270-274; the Euler-Maruyama branch never computes Vshear
      if (shear[0] != 0.0) fcm.calculate_shear_blob_velocities();
    }
    // main.cpp:285, 305-310
    all_bodies.calculate_forces();
    gmres.assemble_RHS(all_bodies, fcm);
    std::vector<double> soln = gmres.invert_saddle_point_system(all_bodies, fcm, 0, t);
    std::vector<double> V(soln.begin() + 3 * nb, soln.end());
    if (all_bodies.end_of_step_update(V, nu)) {  // main.cpp:314-326
      std::printf("step %d: redone (a blob reached a channel wall)\n", step);
      all_bodies.reset_to_start_of_step();
      ++redone;
      continue;
    }
    wr(o, soln.data(), soln.size());
    wr(o, &gmres.report.iterations, 1);
    std::printf("step %d: %d iterations, status %d\n", step, gmres.report.iterations, gmres.report.status);
    t += dt;
    ++step;
  }
  for (const particle& b : all_bodies.bodies) wr(o, b.x, 3);
  for (const particle& b : all_bodies.bodies) {
    const double qq[4] = {b.q.scalar_part, b.q.vector_part[0], b.q.vector_part[1], b.q.vector_part[2]};
    wr(o, qq, 4);
  }
  std::fclose(o);
  return 0;
}

int run_io(const std::string& name, const std::string& state_file, int sample) {
  collection_of_rigid_bodies c;
  c.blobs_per_body = 2;
  c.blob_ref_pos = {-1.86, 0.0, 0.0, 1.86, 0.0, 0.0};
  c.bodies.resize(3);
  const double xs[3][3] = {{1.5, 2.25, 3.125}, {-0.0, 1e-300, 123456789.0}, {63.999999951, 0.1, 7.0}};
  for (int n = 0; n < 3; ++n) {
    for (int k = 0; k < 3; ++k) c.bodies[n].x[k] = xs[n][k];
    c.bodies[n].q = quaternion(0.5 + 0.1 * n, -0.5, 0.5, std::sqrt(0.75 - (0.5 + 0.1 * n) * (0.5 + 0.1 * n)));
  }
  c.total_num_blobs = 6;
  c.blob_forces.resize(18);
  for (int i = 0; i < 18; ++i) c.blob_forces[i] = (i - 9) * 0.3125e-3;
  c.write_blob_references(name);
  c.write_data_to_file(0.0, name);
  c.write_data_to_file(0.20216, name);
  run_parameters p;
  p.num_bodies = 3;
  p.blob_radius = 1.86;
  p.fluid_viscosity = 1.0;
  p.energy_scale = 1.0;
  p.LX = 64.0;
  p.LY = 128.0;
  p.LZ = 64.0;
  p.slip_channel = 1;
  p.steps_per_diffusive_time = 100;
  p.num_diffusive_times = 100;
  p.data_saves_per_diffusive_time = 10;
  write_par_file(name, p);
  std::vector<double> soln(3 * 6 + 6 * 3);
  for (size_t i = 0; i < soln.size(); ++i) soln[i] = 1.0 / (1.0 + (double)i);
  append_velocities(name, 0.20216, soln, 6, 3);
  append_fluid_x_velocity(name, 0.20216, {0.5, -1.25e-7, 3.0, 123456.789, -0.0});
  const int n = c.read_initial_conditions(state_file, sample);
  std::printf("states %d\n", n);
  for (const particle& b : c.bodies)
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", b.x[0], b.x[1], b.x[2], b.q.scalar_part,
                b.q.vector_part[0], b.q.vector_part[1], b.q.vector_part[2]);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "math") == 0) return run_math();
    if (argc >= 4 && std::strcmp(argv[1], "steps") == 0) return run_steps(argv[2], argv[3]);
    if (argc >= 5 && std::strcmp(argv[1], "io") == 0) return run_io(argv[2], argv[3], std::atoi(argv[4]));
  } catch (const fcm_error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 3;
  }
  std::fprintf(stderr, "usage: %s math | steps <in.bin> <out.bin> | io <name> <state file> <sample>\n", argv[0]);
  return 2;
}
