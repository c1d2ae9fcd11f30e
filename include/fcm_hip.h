/*
 * fcm_hip.h -- C ABI of the MI355X-native FCM mobility solver (libfcm_hip.so).
 *
 * Drop-in boundary for the reference's hot path: the public methods of
 * fcm_fluid_solver, fft_solver, gmres_solver and collection_of_rigid_bodies
 * (timwestwood/TimStochRigidBodiesSim).  Each entry point cites the
 * reference interface it replaces.  All functions return an int status
 * (FCM_OK = 0); no C++ exception crosses this boundary.  A handle is used by
 * one host thread at a time; work is ordered on the handle's HIP stream and
 * every entry point that returns host data synchronises that stream.
 *
 * Units and layouts follow the reference:
 *   blob vectors    interleaved [x0,y0,z0,x1,...]      (fcm_fluid_solver.cpp:590-594)
 *   body vectors    [U(3), Omega(3)] per body           (collection_of_rigid_bodies.cpp:715-723)
 *   saddle vectors  [lambda(3*Nb); V(6*NB)]             (gmres_solver.cpp:18)
 *   real fields     [nx][ny][pad], pad = 2*(nz/2+1)     (fcm_fluid_solver.cpp:574-586)
 *   k-space fields  transposed order [ny][nx][nz/2+1] complex (fft_solver.cpp:34)
 */
#ifndef FCM_HIP_H
#define FCM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCM_OK 0
#define FCM_ERR_ARG 1        /* invalid argument or unsupported size        */
#define FCM_ERR_HIP 2        /* HIP runtime error (message in fcm_last_error) */
#define FCM_ERR_NOMEM 3      /* device allocation failed                    */
#define FCM_ERR_STATE 4      /* call made before the state it needs is set  */
#define FCM_ERR_SINGULAR 5   /* single-body mobility matrix not SPD         */
#define FCM_ERR_CAPACITY 6   /* exceeds the capacities given at creation     */

/* GMRES outcomes (gmres_solver.cpp:256-300 prints these; here they are codes) */
#define FCM_GMRES_CONVERGED 0
#define FCM_GMRES_MAX_ITER 1   /* iteration cap reached, iterate returned     */
#define FCM_GMRES_ZERO_RHS 2   /* ||RHS|| = 0, solution 0 (the reference divides by 0) */
#define FCM_GMRES_BREAKDOWN 3  /* exact Krylov breakdown, solution exact      */

#define FCM_PREC_FP64 0        /* every grid pass in FP64                    */
#define FCM_PREC_FP32 1        /* spread/FFT/projection/interp in FP32        */

typedef struct fcm_handle_s* fcm_handle;

/* Runtime replacement of the compile-time parameters of config.hpp. */
typedef struct {
  int nx, ny, nz;      /* NUM_FFT_GRID_POINTS_{X,Y,Z} (config.hpp:65-67): nx, ny and nz/2
                          each 2^a or 3 2^a between 4 and 512 (nx, ny >= 8)            */
  int ng;              /* NUM_FCM_GAUSSIAN_POINTS (config.hpp:69), <= 20 (the reference's
                          default); the column kernels serve ng <= 8, wider stencils
                          run the separate passes                                      */
  double dx;           /* DX (config.hpp:88), decoupled from the blob radius             */
  double a;            /* BLOB_RADIUS (config.hpp:48)                                    */
  double mu;           /* FLUID_VISCOSITY (config.hpp:49)                                */
  double kT;           /* ENERGY_SCALE (config.hpp:50)                                   */
  double dt;           /* DT (config.hpp:84), sets the fluctuating-stress variance       */
  int max_blobs;       /* capacity: blobs                                                */
  int max_bodies;      /* capacity: bodies                                               */
  int max_blobs_per_body;
  int max_krylov;      /* capacity: GMRES iterations (basis columns - 1), GMRES_MAX_ITER */
  int fp32;            /* also allocate the FP32 bulk path (FCM_PREC_FP32 solves)       */
  int random_field;    /* allocate the six stress fields for the fluctuating velocity   */
  int slip_channel;    /* SLIP_CHANNEL (config.hpp:33), 0 or 1: image system in y, slip
                          walls at y = 0 and y = LY/2, blobs in between                */
} fcm_config;

/* ---- lifetime ------------------------------------------------------------ */
/* fcm_fluid_solver::initialise, fcm_fluid_solver.cpp:29-129 (+ fft_solver::initialise,
 * fft_solver.cpp:17-30; gmres_solver::initialise, gmres_solver.cpp:14-34) */
int fcm_create(const fcm_config* cfg, fcm_handle* out);
/* fcm_fluid_solver::~fcm_fluid_solver, fcm_fluid_solver.cpp:10-25 */
int fcm_destroy(fcm_handle h);
/* message of the last failing call on this thread */
const char* fcm_last_error(void);
int fcm_sync(fcm_handle h);
/* HBM bytes fcm_create would allocate for cfg (serving admission) */
int fcm_query_footprint(const fcm_config* cfg, size_t* bytes);
/* free and total device memory of the current device (admission by footprint) */
int fcm_device_memory(size_t* free_bytes, size_t* total_bytes);

/* ---- fcm_fluid_solver ----------------------------------------------------- */
/* set_up_gaussians(const mat& Y), fcm_fluid_solver.cpp:131-165; Y is nb x 3 (host) */
int fcm_set_up_gaussians(fcm_handle h, const double* Y, int nb);
/* set_up_force_distribution(const vec& F), fcm_fluid_solver.cpp:552-634; F host 3*nb */
int fcm_spread(fcm_handle h, const double* F, int precision);
/* produce_deterministic_velocity_field(), fcm_fluid_solver.cpp:636-701 */
int fcm_stokes(fcm_handle h, int precision);
/* calculate_deterministic_blob_velocities(), fcm_fluid_solver.cpp:363-424; V host 3*nb */
int fcm_interp(fcm_handle h, double* V, int precision);
/* spread + Stokes + interpolation: the FCM mobility product V = M F */
int fcm_mobility(fcm_handle h, const double* F, double* V, int precision);
/* produce_random_velocity_field(), fcm_fluid_solver.cpp:167-361.  stress6 is
 * NULL (draw from the seeded stream `seed`) or six host fields sxx syy szz sxy
 * sxz syz, each nx*ny*nz reals (C order x,y,z), already scaled by the variances */
int fcm_random_field(fcm_handle h, const double* stress6, uint64_t seed);
/* calculate_random_blob_velocities(), fcm_fluid_solver.cpp:426-487; V host 3*nb */
int fcm_random_interp(fcm_handle h, double* V);
/* the public grid arrays ux..fz (fcm_fluid_solver.hpp:23-28) in the reference
 * layout [nx][ny][pad]: which = 0,1,2 -> velocity/force grid component x,y,z of
 * the given precision; 3,4,5 -> fluctuating velocity ux_rand..uz_rand */
int fcm_get_field(fcm_handle h, int which, int precision, double* out);
int fcm_set_field(fcm_handle h, int which, int precision, const double* in);
/* single-body reference matrices, produce_reference_matrices,
 * fcm_fluid_solver.cpp:842-881 (body at the box centre, identity
 * orientation; N is a pseudo-inverse so collinear rods are allowed).
 * Needs rb_set_geometry.  Any output may be NULL.  The results are kept
 * and used by gmres_apply_Pinv / gmres_solve. */
int fcm_ref_matrices(fcm_handle h, double* Msingle, double* Minv, double* N);
/* install ref_Minv / ref_N computed elsewhere (fcm_fluid_solver.hpp:59-60) */
int fcm_set_ref_matrices(fcm_handle h, const double* Minv, const double* N);

/* ---- fft_solver ------------------------------------------------------------ */
/* transform(double*), fft_solver.cpp:32-36: in-place R2C of grid component
 * `which` (0..2); read the spectrum with fcm_get_kspace */
int fft_transform(fcm_handle h, int which, int precision);
/* inverse_transform(double*), fft_solver.cpp:38-42 */
int fft_inverse_transform(fcm_handle h, int which, int precision);
int fcm_get_kspace(fcm_handle h, int which, int precision, double* out_transposed);
int fcm_set_kspace(fcm_handle h, int which, int precision, const double* in_transposed);

/* ---- collection_of_rigid_bodies ------------------------------------------ */
/* blob_ref_pos (collection_of_rigid_bodies.hpp:27): nbpb x 3 body-frame offsets */
int rb_set_geometry(fcm_handle h, int nbpb, const double* blob_ref);
/* body state (x, q) of all bodies; computes all_blob_pos (fill_all_blob_pos,
 * collection_of_rigid_bodies.cpp:155-173) and set_up_gaussians on the device */
int rb_set_bodies(fcm_handle h, int NB, const double* x, const double* q);
/* body forces/torques (6 per body) used by the right-hand side */
int rb_set_body_forces(fcm_handle h, const double* F6);
int rb_get_blob_positions(fcm_handle h, double* Y);
/* multiply_by_K_at_midstep, collection_of_rigid_bodies.cpp:707-731 */
int rb_mul_K(fcm_handle h, const double* in6, double* out3);
/* multiply_by_KT_at_midstep, collection_of_rigid_bodies.cpp:733-757 */
int rb_mul_KT(fcm_handle h, const double* in3, double* out6);

/* ---- linked_list and the time step (config 5) ------------------------------ */
/* linked_list::populate + apply_barrier_forces (linked_list.cpp:80-102,192-281):
 * soft-sphere repulsion between blobs of different bodies (interaction_magnitude,
 * linked_list.cpp:362-382, SOFT_SPHERE: U0 = 20 kT, b = a/2) along the nearest-
 * image separation, for pairs closer than `cutoff`, on a device cell list with
 * cells at least `cutoff` wide.  The forces stay on the device as the blob
 * forces of the next RHS (assemble_RHS adds -K^T of them, gmres_solver.cpp:58-70);
 * F3 (3 Nb, blob order) receives a copy unless NULL.  Coincident pairs are
 * skipped (the reference divides by r = 0). */
int ll_barrier_forces(fcm_handle h, double cutoff, double* F3);
/* The reference's own interaction set for every later barrier-force
 * evaluation of the handle (ll_barrier_forces, fcm_em_step, fcm_gdc_step):
 * linked_list::initialise/populate (linked_list.cpp:23-102) on a
 * cells_per_dim^3 grid (NUM_LINKED_LIST_CELLS_PER_DIM, config.hpp:71, is 6),
 * cell floor(n Y/L) per axis, and every pair of blobs of different bodies in
 * the same or neighbouring cells interacting at any distance (the cell and
 * its 13-cell half shell, apply_barrier_forces :192-281); `cutoff` then only
 * switches the forces on (> 0).  0 (the default) restores pairs closer than
 * the cutoff on cells at least the cutoff wide.  cells_per_dim is 0 or 3..1024
 * (with fewer than 3 cells the reference's half shell names one cell twice). */
int ll_set_cells(fcm_handle h, int cells_per_dim);
/* end_of_step_update (collection_of_rigid_bodies.cpp:759-790): every body moves
 * by nu*DT*[U; Omega] with particle::update (particle.cpp:45-56): x += ..., q =
 * lie_exp(...) * q (quaternion.cpp:128-145), x wrapped into the box; then the
 * blob positions and Gaussians are recomputed.  V6: host [U; Omega] per body, or
 * NULL for the body part of the last solution on the device.  DT = fcm_config.dt. */
int rb_end_of_step_update(fcm_handle h, const double* V6, double nu);
/* current body positions (3 NB) and quaternions (4 NB); either may be NULL */
int rb_get_bodies(fcm_handle h, double* x, double* q);
/* ---- shear forcing (SHEARING_FORCES, config.hpp:32) ------------------------ */
/* produce_shear_velocity_field (fcm_fluid_solver.cpp:703-840): the steady
 * flow of the force density fx = scale kT sin(2 pi kappa y / ny) / a (images
 * in a slip channel), solved like produce_deterministic_velocity_field and kept
 * on the device (3 FP64 fields, allocated by the first call; the working
 * fields are overwritten, as the reference's fx..uz are).  As in the
 * reference, assemble_RHS (gmres_assemble_rhs, fcm_mobility_solve/_prepare,
 * fcm_em_step) subtracts the STORED shear blob velocities Vshear
 * (gmres_solver.cpp:62-64), which only fcm_shear_interp
 * (calculate_shear_blob_velocities) and the drifter-corrector mid-step of
 * fcm_gdc_step (main.cpp:270-274) fill; main.cpp's Euler-Maruyama loop never
 * fills them, so its steps carry no shear term.  A new field, or scale = 0
 * (off), drops the stored velocities. */
int fcm_shear_field(fcm_handle h, double shear_force_scale, double shear_wavenumber);
/* calculate_shear_blob_velocities (fcm_fluid_solver.cpp:488-550): the shear
 * field interpolated at the current blobs, stored as the Vshear every later
 * RHS subtracts, and copied to V (3 Nb, host; NULL: device only) */
int fcm_shear_interp(fcm_handle h, double* V);
/* main.cpp:371-433: the x velocity averaged over x and z on each of the ny rows
 * (prof[ny], host), of the flow of the blob forces lambda (3 Nb host, or NULL
 * for the lambda part of the last solution) plus the shear field if set; the
 * blob forces are spread at the current positions (images in a channel). */
int fcm_x_velocity_profile(fcm_handle h, const double* lambda, double* prof);

/* check_for_channel_wall_overlap (collection_of_rigid_bodies.cpp:207-239): in a
 * slip channel *overlap = 1 if any blob is at y == 0 or y >= LY/2 (wrapped
 * positions), else 0; always 0 in a periodic box.  end_of_step_update returns
 * this flag (:784-790) and main.cpp:304-326 then redoes the step. */
int rb_channel_wall_overlap(fcm_handle h, int* overlap);
/* reset_to_start_of_step (collection_of_rigid_bodies.cpp:792-796): the bodies,
 * blob positions and Gaussians return to their state before the last
 * rb_end_of_step_update (or fcm_em_step); the next step can then be redone
 * with a fresh stress seed.  FCM_ERR_STATE if there is no step to undo. */
int rb_reset_to_start_of_step(fcm_handle h);

/* ---- gmres_solver ----------------------------------------------------------- */
typedef struct {
  double tol;          /* GMRES_TOL (config.hpp:62): stop when |beta_j|/||RHS|| < tol */
  int max_iter;        /* GMRES_MAX_ITER (config.hpp:63)                             */
  int precision;       /* FCM_PREC_FP64: plain solve; FCM_PREC_FP32: inner solves with
                          the grid passes and the Krylov basis in FP32 (coefficients,
                          least squares and P^-1 in FP64), and FP64 iterative
                          refinement on the true residual */
  double inner_tol;    /* FP32 form: relative tolerance of each inner solve
                          (0: max(1e-2, tol ||b|| / (2 ||r||)))                  */
  int max_refine;      /* FP32 form: maximum outer refinements (0: 12)              */
} gmres_opts;

typedef struct {
  int status;               /* FCM_GMRES_* */
  int iterations;           /* total inner GMRES iterations (mat-vecs in the Krylov loop) */
  int refinements;          /* outer refinement steps (FP32 form), 0 for FP64 */
  double rel_residual;      /* Givens estimate |beta|/||b|| of the last inner solve (FP64: of the solve) */
  double true_rel_residual; /* ||b - A x||/||b||, FP64 operator (FP32 form; FP64 form if check_residual) */
  double seconds;           /* device time of the solve */
} gmres_report;

/* assemble_RHS, gmres_solver.cpp:58-70: RHS = [-V_rand; -(F + K^T f)] with V_rand
 * from the last fcm_random_interp (zero if none), F from rb_set_body_forces and
 * f the blob forces of ll_barrier_forces (each zero if not set) */
int gmres_assemble_rhs(fcm_handle h, int use_random);
/* multiply_by_saddle_point_matrix, gmres_solver.cpp:72-83 (host vectors) */
int gmres_saddle_matvec(fcm_handle h, const double* in, double* out, int precision);
/* multiply_by_Pinv, gmres_solver.cpp:85-146 (host vectors) */
int gmres_apply_Pinv(fcm_handle h, const double* in, double* out);
/* invert_saddle_point_system, gmres_solver.cpp:148-306.  rhs: host vector or
 * NULL (use the device RHS from gmres_assemble_rhs).  soln: host output or
 * NULL (keep it on the device; read with gmres_get_solution).  resid_hist may
 * be NULL or hold max_iter entries (|beta_j|/||RHS||, the inner solves' histories
 * one after another in the FP32 form); it receives the first
 * min(iterations, max_krylov + 2) of them. */
int gmres_solve(fcm_handle h, const double* rhs, double* soln, const gmres_opts* opts, gmres_report* rep,
                double* resid_hist);
int gmres_get_solution(fcm_handle h, double* soln);

/* ---- one whole fluctuating mobility solve, state resident on the device ---- */
/* produce_random_velocity_field + calculate_random_blob_velocities +
 * assemble_RHS + invert_saddle_point_system (main.cpp:231-233,308-310) */
int fcm_mobility_solve(fcm_handle h, uint64_t seed, const gmres_opts* opts, gmres_report* rep);
/* The right-hand side half of fcm_mobility_solve (main.cpp:231-233 then
 * gmres_solver.cpp:58-70): the seeded fluctuating field, its interpolation
 * and the assembled RHS, left on the device and NOT waited for, so that it
 * overlaps with other work; a following fcm_batch_submit is ordered after it. */
int fcm_mobility_prepare(fcm_handle h, uint64_t seed);
/* one Euler-Maruyama step (main.cpp:222-312 with EULER_MARUYAMA, nu = 1):
 * barrier forces (cutoff > 0), the seeded fluctuating field and its
 * interpolation, the RHS, the saddle-point solve and the body update, all on
 * the device.  In a slip channel the caller then asks rb_channel_wall_overlap
 * and, on 1, calls rb_reset_to_start_of_step and steps again with a new seed
 * (main.cpp:304-326). */
int fcm_em_step(fcm_handle h, uint64_t seed, double cutoff, const gmres_opts* opts, gmres_report* rep);
/* one generalised drifter-corrector step (main.cpp:222-312 with
 * GEN_DRIFTER_CORRECTOR), on the device: the seeded fluctuating field and its
 * blob velocities V; advance_to_midstep (collection_of_rigid_bodies.cpp:
 * 566-621): U0 = the rigid fit of V per body, (K_ref^T K_ref)^-1 in the body
 * frame, and the mid-step bodies = start bodies moved by DT/2 U0;
 * calculate_divU (:869-953): for each of the six directions every body is
 * displaced by e (fd_delta * a for translations, fd_delta for rotations;
 * FINITE_DIFFERENCE_DELTA_FACTOR, config.hpp:73, is 1e-5), the same field is
 * interpolated and divU += sum_n (U_n - U0_n)_i / e; then at the mid-step the
 * blob velocities of the same field, the barrier forces (cutoff > 0), the RHS
 * and the saddle-point solve, and the start bodies move by nu DT V with
 * nu = 1 + DT divU / 2.  divU (may be NULL) receives divU.  In a slip channel,
 * *redo = 1 (may be NULL) means the mid-step put a blob on or beyond a wall:
 * nothing moved and the step is to be taken again with a new seed
 * (main.cpp:237-246); after the step, rb_channel_wall_overlap and
 * rb_reset_to_start_of_step apply as for fcm_em_step.  A body whose K^T K is
 * singular (collinear blobs) uses its pseudo-inverse where the reference
 * inverts. */
int fcm_gdc_step(fcm_handle h, uint64_t seed, double cutoff, double fd_delta, const gmres_opts* opts,
                 gmres_report* rep, double* divU, int* redo);
/* The pieces of that step for the reference's own loop.
 * advance_to_midstep (collection_of_rigid_bodies.cpp:566-621): from the blob
 * velocities V (3 Nb host, or NULL for those of the last fcm_random_interp),
 * U0 per body and the mid-step bodies = bodies moved by DT/2 U0; the device's
 * blobs (positions, Gaussians, K) become the mid-step ones, the bodies stay
 * the start-of-step ones (the next rb_end_of_step_update moves them).  In a
 * slip channel *overlap = 1 (may be NULL) when the mid-step touches a wall:
 * then nothing changes and the step is begun again. */
int rb_advance_to_midstep(fcm_handle h, const double* V, int* overlap);
/* calculate_divU (:869-953) after rb_advance_to_midstep, with the random
 * field in place: six displaced copies of the start bodies, the field's blob
 * velocities and rigid fits; the blobs are the mid-step ones again afterwards
 * (fill_all_blob_pos_at_midstep, main.cpp:263). */
int rb_calculate_divU(fcm_handle h, double fd_delta, double* divU);

/* ---- batches of independent solves (no counterpart in the reference) ------ */
/* The reference solves one system per process (main.cpp:308-310).  A batch
 * advances many handles' solves together: every fcm_batch_step launches one
 * round of `steps_per_round` Arnoldi steps of every iterating job (one hipGraph
 * launch whose branches run on `worker_streams` streams; 0 runs them on the
 * batch's own stream), then reads the control words of the round before and
 * enqueues the finished jobs' next stage.  A job's result is bitwise the same
 * as gmres_solve on its own.  Jobs may be submitted at any time (continuous
 * batching); a handle may hold one job at a time and must not be changed while
 * its job runs.  Not thread-safe: one host thread per batch. */
typedef struct fcm_batch_s* fcm_batch;
int fcm_batch_create(int max_jobs, int worker_streams, int steps_per_round, fcm_batch* out);
int fcm_batch_destroy(fcm_batch b);
/* queue a solve of the handle's assembled RHS (gmres_assemble_rhs or
 * fcm_mobility_prepare); the solution stays on the device (gmres_get_solution) */
int fcm_batch_submit(fcm_batch b, fcm_handle h, const gmres_opts* opts);
/* one round; *unfinished = jobs submitted and not yet finished */
int fcm_batch_step(fcm_batch b, int* unfinished);
/* move up to max finished jobs out: their handles and reports, in finishing order */
int fcm_batch_take(fcm_batch b, fcm_handle* handles, gmres_report* reps, int max, int* count);
/* step until every submitted job has finished */
int fcm_batch_run(fcm_batch b);

/* ---- launch shapes (no counterpart in the reference) ---------------------------- */
/* Tunable launch shapes of a handle's mat-vec, per precision ([FCM_PREC_FP64],
 * [FCM_PREC_FP32]): the spreading kernel (0: stencil-start specialised x loop,
 * 1: branch-free tile-shifted weight tables, 2: as 1 with a ballot over the
 * blobs touching each wavefront's patch) and the interpolation tile (1 or 2:
 * 8x8x8 or 8x8x16 nodes with all three components staged; 3: 8x16x16 (FP32) or
 * 8x16x8 (FP64) staged one component at a time; 0 picks 2 when it divides).  Results do not depend
 * on the choice beyond rounding; FP64 parity holds for all.  `column` selects
 * the mobility product's path: 1 (the default where the grid qualifies: x and y
 * multiples of 4 and >= 16, nz in {32,48,64,96,128,192,256,512}, ng <= 8)
 * spreads onto x-y patches of whole z rows and does the forward z pass in the
 * same kernel, and does the inverse z pass inside interpolation; 0 runs the
 * spreading, five FFT passes and interpolation as separate kernels.  The
 * stand-alone entry points (fcm_spread, fcm_stokes, fcm_interp) always use the
 * separate kernels. */
typedef struct {
  int spread_variant[2];
  int interp_izb[2];
  int column[2];
  /* z-pass points per thread of the column kernels: 8 or 16 (0: the plan that
     pads the patch rows least) */
  int col_points[2];
  /* y FFT passes of the mobility product: 16 or 8 points per thread (0: 16),
     with as many columns per workgroup as fit 256 threads and 64 KiB of LDS
     (8 only where ny >= 64) */
  int y_points[2];
  /* fused x pass (forward FFT, projection, inverse): 16 or 8 in FP32 (0: 16;
     8 only where nx >= 64); FP64 keeps 8 (0 or 8) */
  int x_points[2];
  /* patch x extent of the column path, 4 or 2 (0 where the grid has none);
     one per handle, as the patch lists serve both precisions */
  int col_cx;
} fcm_tuning;
/* the shapes the passes run with: a default (0) is reported as the shape it
   resolves to, so the result can be passed back to fcm_set_tuning */
int fcm_get_tuning(fcm_handle h, fcm_tuning* t);
/* FCM_ERR_ARG for a tuning the handle cannot run (the message says which field) */
int fcm_set_tuning(fcm_handle h, const fcm_tuning* t);
/* Times the candidates of one precision on the handle's current blobs (`reps`
 * launches each, HIP events on the handle's stream) and keeps the fastest:
 * spreading and interpolation tiles of the separate passes, the y and x pass
 * shapes, then for the column path each patch width (only when tuning FP32,
 * or FP64 on a handle without the FP32 path) and z-pass points per thread, and
 * the column path against the separate passes.  Returns the handle's tuning
 * and stores it in a process-wide cache under the handle's shape class (grid,
 * ng, slip channel, the power of two at or above the blob count); a handle
 * whose tuning was never set takes its class's cached tuning when its blobs
 * are set up. */
int fcm_autotune(fcm_handle h, int precision, int reps, fcm_tuning* chosen);
/* As fcm_autotune, but a shape class already in the cache is applied without
 * timing anything (*from_cache = 1; may be NULL).  For batches and serving,
 * where many handles share a class. */
int fcm_autotune_cached(fcm_handle h, int precision, int reps, fcm_tuning* chosen, int* from_cache);
/* empty the shape-class cache (handles keep the tunings they have) */
int fcm_clear_tuning_cache(void);

/* ---- measurement hook (no counterpart in the reference) ----------------------- */
/* Runs one saddle-point mat-vec `reps` times (after one untimed warm-up) on a
 * fixed input with HIP events between its passes, on the handle's stream
 * (`reps` + 1 launches of each kernel in all); ms_per_pass[0..8] receives the
 * mean device time of: spread, z forward, y forward, x forward+projection+x
 * inverse, y inverse, z inverse, interpolation(+K), K^T, and the reduction of
 * the interpolation partial sums.  On the column path the spread slot holds
 * the fused spread + z forward kernel, the interpolation slot the fused z
 * inverse + interpolation kernel (each timed over 4 back-to-back launches and
 * divided by 4, so that one launch's dispatch is not counted), and the two z
 * slots are 0; on the separate path the last slot is 0. */
int fcm_profile_matvec(fcm_handle h, int precision, int reps, double* ms_per_pass);
/* A plain streaming copy of `bytes` (16-byte words) between two scratch buffers,
 * `reps` times, on the handle's stream; the reference point for HBM byte counters
 * and for the achievable copy bandwidth.  *ms_per_copy = mean device time. */
int fcm_profile_copy(fcm_handle h, size_t bytes, int reps, double* ms_per_copy);

/* ---- an entry point that allocates nothing -------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the blob forces of ll_barrier_forces, bitwise the same,
 * with the cell list in a static device array of FCM_NOALLOC_SCRATCH_BYTES on
 * every device the library runs on (held until the process ends) instead of
 * memory the library allocates.  Y (nb x 3, wrapped into [0, L)) and F (3 nb)
 * are DEVICE pointers; body of blob i = i / nbpb; L the box (dx n per axis);
 * a, kT as fcm_config; the interaction set is ll_set_cells's (cells_per_dim 0:
 * pairs closer than `cutoff`; 3..1024: the reference's cells, `cutoff` > 0 is
 * then only a switch); slip_channel adds the wall springs.  Work is enqueued
 * on `stream` (a hipStream_t; NULL: the null stream) and not waited for.
 * FCM_ERR_CAPACITY if the cell list (2 nb + 3 cells + 1 ints) exceeds the
 * array.  Calls that overlap on one device share the array and race.  The
 * first call on a device resolves the array's address there (make it before
 * any graph capture). */
#define FCM_NOALLOC_SCRATCH_BYTES (48u << 20)
int ll_barrier_forces_noalloc(const double* Y, int nb, int nbpb, const double L[3], double a, double kT,
                              double cutoff, int cells_per_dim, int slip_channel, double* F, void* stream);

/* ---- synthetic inputs (host only; no GPU needed) ---------------------------- */
/* Rods of nbpb spheres spaced 2a on the body x axis, centres uniform in the box
 * with rejection of inter-body blob pairs closer than 2a (minimum image), at
 * most max_attempts per rod (then accepted with overlap), uniform random
 * quaternions (normalised 4-Gaussian, quaternion.cpp:10-17), N(0,1) forces and
 * torques with the axial torque removed.  Outputs may not be NULL except F6. */
int fcm_synth_rods(uint64_t seed, int NB, int nbpb, double a, const double L[3], int max_attempts, double* blob_ref,
                   double* x, double* q, double* F6, int* n_overlapping);

#ifdef __cplusplus
}
#endif

#endif /* FCM_HIP_H */
