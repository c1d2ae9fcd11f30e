// fcm_internal.hpp -- declarations shared by the kernel translation units and
// the host orchestration behind the C ABI.
#pragma once

#include "fcm_common.hpp"

namespace fcm {

// Grid geometry.  Fields are complex arrays [nx][ny][cpitch]; real-space rows
// hold nz reals in their first nz real slots.
struct GridGeom {
  int nx = 0, ny = 0, nz = 0, nzc = 0, cpitch = 0;
  int channel = 0;  // slip-channel image system in y (walls at rows 0 and ny/2)
  double dx = 1.0;
  // launch shapes of the y and x FFT passes of the mobility product per
  // precision (0 FP64, 1 FP32): points per thread, 8 or 16 (0: the default,
  // 16; the x pass keeps 8 in FP64), with as many columns per workgroup as fit
  // 256 threads and 64 KiB of LDS; lengths below 64 always use the default
  int ypm[2] = {0, 0}, xpm[2] = {0, 0};
  size_t field_words() const { return (size_t)nx * ny * cpitch; }  // complex words per field
};

// grid fields handed to a kernel; if gate is set and *gate != 0 the kernel
// returns at once (device-controlled iterations after convergence)
template <typename T> struct FieldPtrs {
  c2<T>* p[6];
  const int* gate;
};

// device twiddle tables: x (nx), y (ny), z (nz/2) and the R2C untangling
// factors zu[k] = exp(-2 pi i k / nz), k = 0..nz/2
template <typename T> struct TwiddleSet {
  const c2<T>* x = nullptr;
  const c2<T>* y = nullptr;
  const c2<T>* z = nullptr;
  const c2<T>* zu = nullptr;
};

struct ProjParams {
  double facx, facy, facz;  // 2 pi / L per axis (fcm_fluid_solver.cpp:75,97,118)
  double inv_muN;           // 1 / (mu nx ny nz)
  int nx, ny, nz;
};

template <typename T>
hipError_t fft3d_forward(const GridGeom& g, const FieldPtrs<T>& fp, int nf, const TwiddleSet<T>& tws, hipStream_t st);
template <typename T>
hipError_t fft3d_inverse(const GridGeom& g, const FieldPtrs<T>& fp, int nf, const TwiddleSet<T>& tws, hipStream_t st);
// images: the forward y pass applies the slip-channel image rule as it reads
// (a slip channel's mobility product, whose spreading leaves the images out)
template <typename T>
hipError_t stokes_solve(const GridGeom& g, const FieldPtrs<T>& fp, const ProjParams& pp, const TwiddleSet<T>& tws,
                        hipStream_t st, bool images = false);
// one pass of stokes_solve: 0 z forward, 1 y forward, 2 x forward+projection+
// x inverse, 3 y inverse, 4 z inverse (all three fields)
template <typename T>
hipError_t stokes_pass(const GridGeom& g, const FieldPtrs<T>& fp, int pass, const ProjParams& pp,
                       const TwiddleSet<T>& tws, hipStream_t st, bool images = false);
template <typename T>
hipError_t random_velocity(const GridGeom& g, const FieldPtrs<T>& stress6, bool generate, unsigned long long seed,
                           const double scale[6], const ProjParams& pp, const TwiddleSet<T>& tws, hipStream_t st);

// ---------------------------------------------------------------------------
// blobs: Gaussian set-up, cell binning, spreading and interpolation
// ---------------------------------------------------------------------------
struct BlobBins {
  int bx = 4, by = 8, bz = 8;       // bin extent (nodes) per axis, divides the grid
  int nbx = 0, nby = 0, nbz = 0;    // bins per axis
  int sx = 2, sy = 2, sz = 2;       // spread tile = (sx*bx, sy*by, sz*bz) nodes
  int nbins() const { return nbx * nby * nbz; }
};

struct BlobData {
  int nb = 0;                 // number of blobs
  int ng = 7;                 // Gaussian points per axis
  const double* Y = nullptr;  // nb x 3 positions
  int* bin_of = nullptr;      // nb
  int* bin_count = nullptr;   // nbins
  int* bin_start = nullptr;   // nbins + 1
  int* bin_cursor = nullptr;  // nbins
  int* sorted = nullptr;      // nb, blob ids in bin order (ascending id within a bin)
  int4* rec = nullptr;        // nb, sorted order: (anchor x, y, z wrapped, id)
  int4* anc = nullptr;        // nb, sorted order: unwrapped anchors
  double* w64 = nullptr;      // nb x 3 x ng, sorted order
  float* w32 = nullptr;       // nb x 3 x ng, sorted order
  // FP32 column kernels: the weights are evaluated from the offset of stencil
  // node 0 (node position - blob position, x y z; sorted order) as
  // wa32 exp2(wc32 xx^2), xx = offset + k dx
  float4* woff = nullptr;
  float wa32 = 0.f, wc32 = 0.f;
  // interpolation weights: w64/w32 themselves, or in a slip channel a copy whose
  // y weights are zero on rows >= ny/2 (truncated Gaussians)
  double* iw64 = nullptr;
  float* iw32 = nullptr;
  int* tile_count = nullptr;  // spreading tiles: relevant blobs per tile
  int* tile_start = nullptr;  // ntiles + 1
  int* tile_cursor = nullptr; // ntiles (scan scratch)
  int* tile_list = nullptr;   // sorted slots of the blobs relevant to each tile (<= 8 per blob)
  // launch-shape choices (tunable per grid): spreading kernel 0 = stencil-start
  // specialised, 1 = branch-free shifted tables
  // indexed by precision (0 FP64, 1 FP32)
  int spread_variant[2] = {0, 0};
  int interp_izb[2] = {0, 0};  // interpolation tile depth in z bins: 0 = 2 when it divides, 1, 2
  // column path of the mobility product (fcm_column.hip): x-y patches of
  // col_cx x COL_CY nodes with all of z; per patch the blobs whose x-y stencil
  // touches it, ordered by stencil-start z, and a z prefix table
  int col_cx = 0;              // patch x extent, 4 or 2 (0: the grid does not qualify)
  int col_pm[2] = {0, 0};      // z-pass points per thread per precision: 8, 16 (0: the shape's default)
  int* col_count = nullptr;    // npatch x nbz: entries per (patch, z group)
  int* col_start = nullptr;    // npatch x nbz + 1 (exclusive scan)
  int* col_cursor = nullptr;   // scan scratch
  int* col_list = nullptr;     // sorted blob slots, COL_SLOTS_MAX per blob at most
  int4* col_ent = nullptr;     // entry records in list order (slot, id, packed offsets, partial slot)
  int* col_zst = nullptr;      // npatch x (nz + 1): first entry with stencil-start z >= k
  // FP64 weight windows of the column kernels, COL_WIN words per sorted blob:
  // for each patch a blob's stencil can touch in x (y), the patch's cx (4) x (y)
  // weights, then the ng z weights, zero-padded (built from w64 / iw64 per
  // patch width; the FP32 kernels evaluate their weights from woff instead)
  double* cwin64 = nullptr;
  double* icwin64 = nullptr;  // interpolation: cwin64 itself, or in a slip channel from iw64
  // FP64 spreading weights of every list entry, COL_EWIN words in list order
  // (gx 4, gy 4, gz 8: the entry's patch cut out of cwin64), so that spreading
  // streams them with the entries instead of gathering them through the blob
  double* ewin64 = nullptr;
  void* col_part = nullptr;    // interpolation partial sums: nb x COL_SLOTS_MAX x 3 (T)
  int col_fused[2] = {1, 1};   // per precision (0 FP64, 1 FP32): use the column path in the mobility product
};


// spreading tiles of the main path (0 if the grid does not qualify)
bool spread_fast_ok(const GridGeom& g, const BlobBins& bb, int ng);
bool spread_generic_ok(const GridGeom& g, const BlobBins& bb, int ng);
int spread_tiles(const GridGeom& g);
// slip-channel image rule on the three force fields after spreading (no-op
// unless g.channel)
template <typename T>
hipError_t channel_images(const GridGeom& g, const FieldPtrs<T>& fp, hipStream_t st);

hipError_t setup_gaussians(const GridGeom& g, const BlobBins& bb, BlobData& bd, double a, hipStream_t st);

// constraint term folded into interpolation: out[id] = dx^3 J u - (U + W x d_id)
struct KTerm {
  const double* zV = nullptr;  // 6 per body (U, Omega), null: no K term
  const double* dpos = nullptr;  // 3 per blob, R_n r_m
  int nbpb = 1;
};

// V[id] = dx^3 (v0, v1, v2) - (U + Omega x d_id) when the K term is present
// (the first row of the saddle product, gmres_solver.cpp:72-83 with
// collection_of_rigid_bodies.cpp:723)
__device__ __forceinline__ void store_blob_velocity(double* V, const KTerm& kt, double dx3, int id, double v0,
                                                    double v1, double v2) {
  v0 *= dx3; v1 *= dx3; v2 *= dx3;
  if (kt.zV) {
    const int body = id / kt.nbpb;
    const double* u = kt.zV + 6 * (size_t)body;
    const double* d = kt.dpos + 3 * (size_t)id;
    v0 -= u[0] + (u[4] * d[2] - u[5] * d[1]);
    v1 -= u[1] + (u[5] * d[0] - u[3] * d[2]);
    v2 -= u[2] + (u[3] * d[1] - u[4] * d[0]);
  }
  V[3 * (size_t)id] = v0;
  V[3 * (size_t)id + 1] = v1;
  V[3 * (size_t)id + 2] = v2;
}

// images: apply the slip-channel image rule to the spread fields (the mobility
// product passes false: its forward y pass applies the rule as it reads)
template <typename T>
hipError_t spread(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                  hipStream_t st, bool images = true);
template <typename T>
hipError_t interp(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const FieldPtrs<T>& fp, double* V,
                  const KTerm& kt, hipStream_t st);

// ---------------------------------------------------------------------------
// column path (fcm_column.hip): spreading fused with the forward z R2C pass,
// and the inverse z C2R pass fused with interpolation
// ---------------------------------------------------------------------------
constexpr int COL_CY = 4;          // patch y extent
constexpr int COL_SLOTS_MAX = 15;  // partial sums per blob: up to 5 x 3 patches
constexpr int COL_WIN = 32;        // window words per blob: x windows 0-11, y windows 12-23, z 24-31
constexpr int COL_EWIN = 16;       // FP64 spreading weights per list entry: x 0-3, y 4-7, z 8-15
// patch x extent usable on this grid for both precisions (0: none)
int col_patch_cx(const GridGeom& g, const BlobBins& bb, int ng);
// the column kernels exist for word size tbytes at this nz, patch width cx and
// z-pass points per thread pm (0: the default plan)
bool col_shape_valid(int tbytes, int nz, int cx, int pm);
// the z-pass points per thread those kernels run with (0: no such kernels)
int col_shape_pm(int tbytes, int nz, int cx, int pm);
size_t col_list_capacity(int nb);
size_t col_zst_words(const GridGeom& g, int cx);
int col_groups(const GridGeom& g, int cx);  // npatch x nbz
hipError_t build_col_lists(const GridGeom& g, const BlobBins& bb, BlobData& bd, hipStream_t st);
template <typename T>
hipError_t spread_zfwd(const GridGeom& g, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                       const TwiddleSet<T>& tw, hipStream_t st, bool images = true);
// inverse z pass + interpolation partial sums, then the fixed-order reduction
// into V (with the K term).  For timing, `between` is recorded between the two
// and the first kernel is launched `interp_reps` times (it is idempotent).
template <typename T>
hipError_t zinv_interp(const GridGeom& g, const BlobData& bd, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw,
                       double* V, const KTerm& kt, hipStream_t st, hipEvent_t between = nullptr, int interp_reps = 1);

// ---------------------------------------------------------------------------
// rigid bodies, preconditioner, Krylov kernels
// ---------------------------------------------------------------------------
struct BodyData {
  int NB = 0, nbpb = 0;
  const double* ref = nullptr;   // 3 x nbpb body-frame blob offsets
  const double* x = nullptr;     // 3 NB
  const double* q = nullptr;     // 4 NB
  double* R = nullptr;           // 9 NB (row-major)
  double* dpos = nullptr;        // 3 nb (world-frame offsets R r_m)
  double* Y = nullptr;           // 3 nb wrapped positions
  const double* Minv = nullptr;  // (3 nbpb)^2
  const double* MinvT = nullptr; // (3 nbpb)^2, transposed
  const double* Nref = nullptr;  // 36
  const int* gate = nullptr;     // P^-1 kernel returns when *gate != 0
  const int* kcol = nullptr;     // if set, the P^-1 input is column *kcol of a basis with stride ldq
  size_t ldq = 0;
};

hipError_t bodies_update(const GridGeom& g, BodyData& b, hipStream_t st);
// *flag = 1 if any blob sits on the lower wall (y == 0) or at or beyond the
// upper one (y >= half_ly), else 0
hipError_t wall_overlap(const double* Y, int nb, double half_ly, int* flag, hipStream_t st);
// U = diag(R,R) KTKinv diag(R^T,R^T) KTv per body (the body-frame rigid fit of
// advance_to_midstep / calculate_divU)
hipError_t body_fit(const BodyData& b, const double* KTv, const double* KTKinv, double* U, hipStream_t st);
// V6[6n + k] = (k == i) ? e : 0 for every body n
hipError_t fill_body6(double* V6, int NB, int i, double e, hipStream_t st);
// acc[i] = sum_n (U[6n+i] - U0[6n+i]) / e (one workgroup, fixed order)
hipError_t divu_sum(const double* U, const double* U0, int NB, int i, double e, double* acc, hipStream_t st);
// the shear forcing fx = (scale kT) sin(2 pi kappa y / ny) / a on every real node,
// fy = fz = 0 (fcm_fluid_solver.cpp:705-729)
hipError_t shear_force(const GridGeom& g, const FieldPtrs<double>& fp, double scale_kT, double a, double kappa,
                       hipStream_t st);
// prof[y] = mean over x and z of u[0] (+ add[0] if set) on every y row of real fields
hipError_t x_velocity_profile(const GridGeom& g, const FieldPtrs<double>& u, const FieldPtrs<double>& add,
                              double* prof, hipStream_t st);
// end_of_step_update (collection_of_rigid_bodies.cpp:759-790): body n moves by
// s [U; Omega] (V6 with stride vstride doubles between bodies' 6-vectors)
hipError_t bodies_step(const GridGeom& g, double* x, double* q, int NB, const double* V6, double s, hipStream_t st);

// counting sort of n items by key (count[] already holds the key histogram);
// items keep ascending index order inside a key
hipError_t sort_by_key(const int* key_of, int n, int nkeys, const int* count, int* start, int* cursor, int* sorted,
                       hipStream_t st);
hipError_t fill_ints(int* p, int n, int v, hipStream_t st);
// start[0..n] = exclusive scan of count[0..n), cursor[0..n) = 0 (one workgroup)
hipError_t exclusive_scan(const int* count, int* start, int* cursor, int n, hipStream_t st);

// soft-sphere barrier forces between blobs of different bodies within a cutoff
// (linked_list.cpp:80-102,192-281,362-382) on a cell list of cells >= cutoff
struct StericCells {
  int nc[3];
  double L[3];
};
// The cell grid of a barrier-force evaluation.  cells_per_dim > 0: the
// reference's NUM_LINKED_LIST_CELLS_PER_DIM^3 grid (config.hpp:71), every pair
// in neighbouring cells interacting at any distance (linked_list.cpp:23-102);
// else cells at least `cutoff` wide (at most 1024 per axis), pairs closer than
// the cutoff.
inline StericCells steric_cells(const double L[3], double cutoff, int cells_per_dim) {
  StericCells cg;
  for (int d = 0; d < 3; ++d) {
    cg.L[d] = L[d];
    if (cells_per_dim > 0) {
      cg.nc[d] = cells_per_dim;
    } else {
      const double w = L[d] / cutoff;
      cg.nc[d] = w >= 1024.0 ? 1024 : (w < 1.0 ? 1 : (int)w);
    }
  }
  return cg;
}
struct StericWork {
  int* cell_of = nullptr;  // nb
  int* sorted = nullptr;   // nb
  int* count = nullptr;    // ncells
  int* start = nullptr;    // ncells + 1
  int* cursor = nullptr;   // ncells
  int cap_cells = 0;
};
// all_pairs: every pair in neighbouring cells (no cutoff), else pairs closer than cutoff
hipError_t barrier_forces(const double* Y, int nb, int nbpb, const StericCells& cg, double cutoff, bool all_pairs,
                          double kT, double a, bool channel, StericWork& w, double* F, hipStream_t st);
// out = P^{-1} in; if wKT != null also writes wKT = -K^T out_lambda (6 per body)
hipError_t apply_pinv(const BodyData& b, const double* in, double* out, double* wKT, hipStream_t st);
// the same with the input in FP32 (a column of an FP32 Krylov basis)
hipError_t apply_pinv(const BodyData& b, const float* in, double* out, double* wKT, hipStream_t st);
hipError_t mul_K(const BodyData& b, const double* in6, double* out3, hipStream_t st);
hipError_t mul_KT(const BodyData& b, const double* in3, double* out6, hipStream_t st);

struct KrylovWork {
  int nblk = 0;              // row blocks for the reductions
  double* partial = nullptr; // nblk x max_cols
  double* red = nullptr;     // reduced values (max_cols + 8)
};

int krylov_blocks(size_t n);
// red[i] = sum_r Q[i][r] w[r], i < ncols   (deterministic two-stage)
hipError_t dots(const double* Q, size_t ldq, int ncols, const double* w, size_t n, KrylovWork& kw, double* red_out,
                hipStream_t st);
// w -= sum_i h[i] Q[i]; if norm_out != null, *norm_out = ||w||^2 after the update
hipError_t update(const double* Q, size_t ldq, int ncols, const double* h, double* w, size_t n, KrylovWork& kw,
                  double* norm_out, hipStream_t st);
// red[0..ncols) = Q[c].v1, red[ncols..2 ncols) = Q[c].v2 in one pass (either vector may be null)
hipError_t dots2(const double* Q, size_t ldq, int ncols, const double* v1, const double* v2, size_t n, KrylovWork& kw,
                 double* red_out, hipStream_t st);
// lagged reorthogonalisation pass, coef = [s(m) | t(mt) | 1/nu | pk]:
//   qt <- (qt - sum_{i<m} s_i Q_i)/nu (if qt), w <- w - sum_{i<mt} t_i Q_i - pk qt (if w),
//   *norm_out = ||w||^2 (or ||qt||^2 when w is null)
hipError_t update2(const double* Q, size_t ldq, int m, int mt, const double* coef, double* qt, double* w, size_t n,
                   KrylovWork& kw, double* norm_out, hipStream_t st);
// y = sum_i c[i] Q[i]
hipError_t combine(const double* Q, size_t ldq, int ncols, const double* c, double* y, size_t n, hipStream_t st);
// y = a*x (a from device pointer, taken as 1/sqrt(*a) if invsqrt)
hipError_t scale_vec(const double* x, double* y, size_t n, const double* a, bool invsqrt, hipStream_t st);
hipError_t copy16(const void* src, void* dst, size_t bytes, hipStream_t st);
hipError_t axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t st);  // y = a x + b y
// b = [-Vrand - Vshear; -F6] (gmres_solver.cpp:58-70); null inputs count as zero
hipError_t rhs_assemble(const double* Vrand, const double* Vshear, const double* F6, double* b, size_t n3, size_t n6,
                        hipStream_t st);

// ---------------------------------------------------------------------------
// device-resident GMRES control (fcm_gmres.hip)
// ---------------------------------------------------------------------------
// GM_FIN: 0 while iterating, 1 from the step that predicts convergence until
// the next step (a finalising step) has finalised the new vector, 2 for the
// rest of that step.  GM_GATE: nonzero when the mat-vec kernels of a step must
// not run (the solve has stopped, or the step is a finalising step).
// GM_ARR_FIN: arrival counter of the finalising update, whose last block ends
// the finalising step (zero between launches; the last block resets it).
enum {
  GM_K = 0, GM_PROV = 1, GM_STATUS = 2, GM_NCOLS = 3, GM_FIN = 4, GM_MAXIT = 5, GM_M = 6, GM_MT = 7, GM_HS = 8,
  GM_GATE = 9, GM_ARR_FIN = 10, GM_NINT = 11
};

struct GmDev {
  int* ictl = nullptr;     // GM_* words; status 0 = running, 1 + FCM_GMRES_* when done
  double* dctl = nullptr;  // 0 tol, 1 ||b||, 2 eta, 3 nu, 4 last residual, 5 1/||w||
  double* Hraw = nullptr;  // (M+2)^2 unrotated Hessenberg
  double* R = nullptr;     // (M+2)^2 rotated
  double *CS = nullptr, *SN = nullptr, *beta = nullptr, *hprov = nullptr, *hnew = nullptr;  // M+2
  double* coef = nullptr;  // 2M+8
  double* red = nullptr;   // 2M+8
  double* hist = nullptr;  // M+2
  double* y = nullptr;     // M+2
  int M = 0;
};

int gm_blocks(size_t n);
// the basis in FP64 (FP64 solves) or FP32 (the inner solves of FP32+refine);
// every coefficient, the new vector and the least squares stay in FP64
hipError_t gm_launch_start(const GmDev& G, double tol, int max_iter, const double* b, double* Q, size_t n,
                           hipStream_t st);
hipError_t gm_launch_start(const GmDev& G, double tol, int max_iter, const double* b, float* Q, size_t n,
                           hipStream_t st);
hipError_t gm_launch_basis(const GmDev& G, double* Q, size_t ldq, double* wbuf, size_t n, double* partial,
                           int nblk, hipStream_t st);
hipError_t gm_launch_basis(const GmDev& G, float* Q, size_t ldq, double* wbuf, size_t n, double* partial,
                           int nblk, hipStream_t st);
hipError_t gm_launch_finish(const GmDev& G, const double* Q, size_t ldq, double* x, size_t n, hipStream_t st);
hipError_t gm_launch_finish(const GmDev& G, const float* Q, size_t ldq, double* x, size_t n, hipStream_t st);
// copy the 16-double control blocks src[0..n) into dst (device memory)
hipError_t gm_gather_ctl(const double* const* src, double* dst, int n, hipStream_t st);
// table[idx] = p, stream-ordered
hipError_t gm_set_ptr(const double** table, int idx, const double* p, hipStream_t st);

}  // namespace fcm
