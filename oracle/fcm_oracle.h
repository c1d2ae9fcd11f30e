/*
 * fcm_oracle.h -- CPU oracle for the FCM mobility solve (TEST INFRASTRUCTURE).
 *
 * This is a plain-C FP64 restatement of the reference's hot path, written to
 * check the HIP product against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never calls it.
 *
 * Layouts follow the reference exactly:
 *   real field      [nx][ny][pad], pad = 2*(nz/2+1)   index (i*ny + j)*pad + k
 *                   (fcm_fluid_solver.cpp:574-586, fft_solver.cpp:26)
 *   k-space field   "transposed order" [ny][nx][nz/2+1] complex,
 *                   index (j*nx + i)*pad + 2*kz (+1 for Im)
 *                   (fcm_fluid_solver.cpp:655-670, fft_solver.cpp:34)
 *   blob vectors    interleaved [x0,y0,z0,x1,...]      (fcm_fluid_solver.cpp:590-594)
 *   body vectors    [U(3), Omega(3)] per body           (collection_of_rigid_bodies.cpp:715-723)
 *   GMRES vector    [lambda(3*Nb); V(6*NB)]             (gmres_solver.cpp:18)
 */
#ifndef FCM_ORACLE_H
#define FCM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;   /* NUM_FFT_GRID_POINTS_{X,Y,Z}  config.hpp:65-67 */
  int ng;           /* NUM_FCM_GAUSSIAN_POINTS     config.hpp:69    */
  double dx;        /* DX                          config.hpp:88    */
  double a;         /* BLOB_RADIUS                 config.hpp:48    */
  double mu;        /* FLUID_VISCOSITY             config.hpp:49    */
  double kT;        /* ENERGY_SCALE                config.hpp:50    */
  double dt;        /* DT                          config.hpp:84    */
  int slip_channel; /* SLIP_CHANNEL                config.hpp:33: image system in y,
                       fluid in 0 < y < LY/2 between slip walls at y = 0, LY/2 */
} or_cfg;

typedef struct or_fcm or_fcm;

/* ---- random streams (seeded, counter based) ---------------------------- */
uint64_t or_splitmix64_next(uint64_t* state);
double   or_stress_normal(uint64_t seed, int comp, uint64_t node, uint64_t nnodes);

/* ---- 1-D / 3-D transforms (FFTW 2 rfftwnd_mpi semantics) -------------- */
void or_dft1d(int n, const double* in, double* out, int sign);      /* O(n^2) brute force */
void or_fft1d(int n, const double* in, double* out, int sign);      /* mixed radix       */
void or_fft3d_r2c(const or_cfg* c, double* field);                  /* in place           */
void or_fft3d_c2r(const or_cfg* c, double* field);                  /* in place           */

/* ---- fluid solver object (fcm_fluid_solver.hpp:18-76) ------------------ */
or_fcm* or_fcm_create(const or_cfg* c, int max_blobs);
void    or_fcm_destroy(or_fcm* f);
int     or_fcm_pad(const or_fcm* f);
size_t  or_fcm_total(const or_fcm* f);
double* or_fcm_field(or_fcm* f, int which); /* 0..11 ux uy uz fx fy fz sxx syy szz sxy sxz syz; 12..14 u*_rand;
                                               15..17 u*_shear */
void    or_fcm_tables(or_fcm* f, double* qx, double* qy, double* qpad);
void    or_fcm_set_up_gaussians(or_fcm* f, const double* Y, int nb);
void    or_fcm_get_gaussians(const or_fcm* f, double* gx, double* gy, double* gz, int* ix, int* iy, int* iz);
void    or_fcm_spread(or_fcm* f, const double* F);
void    or_fcm_stokes(or_fcm* f);
void    or_fcm_interp(or_fcm* f, double* V);
void    or_fcm_fill_stress(or_fcm* f, uint64_t seed);
void    or_fcm_random_field(or_fcm* f);
void    or_fcm_random_interp(or_fcm* f, double* V);
void    or_fcm_shear_field(or_fcm* f, double scale, double kappa);
void    or_fcm_shear_interp(or_fcm* f, double* V);
void    or_fcm_x_velocity_profile(or_fcm* f, double* prof);
void    or_fcm_mobility(or_fcm* f, const double* F, double* V);

/* ---- rigid bodies (collection_of_rigid_bodies.cpp) ---------------------- */
void or_quat_matrix(const double q[4], double R[9]);
void or_fill_blob_pos(const or_cfg* c, int NB, int nbpb, const double* blob_ref,
                      const double* x, const double* q, double* Y);
void or_mul_K(int NB, int nbpb, const double* blob_ref, const double* q,
              const double* in6, double* out3);
void or_mul_KT(int NB, int nbpb, const double* blob_ref, const double* q,
               const double* in3, double* out6);

/* ---- time-step pieces (config 5) --------------------------------------- */
/* Soft-sphere barrier forces between blobs of different bodies
 * (linked_list.cpp:192-281, magnitude interaction_magnitude :362-382 with
 * SOFT_SPHERE: U0 = 20 kT, b = a/2, |F| = U0/b for r < 2a, else
 * U0 exp((2a - r)/b)/b), acting along the minimum-image separation
 * x1 - x2 - L trunc(2 (x1 - x2)/L) (:255-257), for pairs closer than
 * `cutoff` (the reference's cell traversal bounds the range instead).
 * Brute force over all pairs, in blob order; coincident pairs are skipped. */
void or_barrier_forces(const or_cfg* c, int Nb, int nbpb, const double* Y, double cutoff, double* F);
/* The reference's own interaction set (linked_list.cpp:23-102,192-281):
 * every pair of blobs of different bodies in one cell of an ncell^3 grid or
 * in neighbouring cells (its 13-cell half shell), at any distance, in the
 * reference's head/list traversal order. */
void or_barrier_forces_cells(const or_cfg* c, int Nb, int nbpb, const double* Y, int ncell, double* F);
/* quaternion.cpp:128-145: exp map of a rotation vector (identity below 1e-10) */
void or_lie_exp(const double u[3], double q[4]);
/* p * q (quaternion.hpp operator*) */
void or_quat_mul(const double p[4], const double q[4], double out[4]);
/* end_of_step_update (collection_of_rigid_bodies.cpp:759-790) with
 * particle::update (particle.cpp:45-56): x += s U, q = lie_exp(s Omega) q,
 * x wrapped into [0, L); V6 holds [U; Omega] per body, s = nu DT. */
void or_end_of_step_update(const or_cfg* c, int NB, const double* V6, double s, double* x, double* q);

/* ---- small dense algebra ------------------------------------------------ */
int  or_inv_sympd(int n, const double* A, double* Ainv);
int  or_pinv_sym(int n, const double* A, double* Ainv, double rtol);
int  or_solve_dense(int n, const double* A, const double* b, double* x);

/* ---- preconditioner reference matrices (fcm_fluid_solver.cpp:842-881) --- */
int  or_ref_matrices(or_fcm* f, int nbpb, const double* blob_ref,
                     double* Msingle, double* Minv, double* N);

/* ---- GMRES saddle-point solve (gmres_solver.cpp) ------------------------ */
typedef struct {
  int NB, nbpb;
  const double* blob_ref;   /* 3*nbpb, column m = (x,y,z) of blob m   */
  const double* x;          /* 3*NB  body centres                       */
  const double* q;          /* 4*NB  body quaternions (q0,q1,q2,q3)     */
  const double* Minv;       /* (3nbpb)^2 row-major                      */
  const double* N;          /* 36 row-major                             */
} or_bodies;

/* Wall seconds of one modified Gram-Schmidt step of or_gmres (gmres_solver.cpp:
 * 198-210) against `iter` basis vectors of length n, on synthetic vectors: the
 * orthogonalisation share of a CPU solve for the bench's baseline. */
double or_gmres_orth_seconds(size_t n, int iter);
void or_saddle_matvec(or_fcm* f, const or_bodies* b, const double* in, double* out);
void or_apply_Pinv(const or_bodies* b, const double* in, double* out);
int  or_gmres(or_fcm* f, const or_bodies* b, const double* rhs, double* soln,
              double tol, int max_iter, int* iters, double* resid_hist);

#ifdef __cplusplus
}
#endif

#endif
