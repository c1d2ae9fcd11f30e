// fcm_gmres.hip -- device-resident control of the GMRES saddle-point solve.
//
// gmres_solver.cpp:148-306 runs the Arnoldi loop, the Givens rotations and
// the stopping test on the host.  Here every scalar of that loop lives in
// device memory (GmDev) and is updated by small single-workgroup kernels, so
// one Arnoldi iteration is a fixed sequence of launches that reads nothing
// back: it can be captured once into a hipGraph and replayed.  Once the
// stopping test fires, the status word is set and every later kernel of the
// sequence returns at once, so the host only polls the status every few
// iterations.
//
// The basis is built by classical Gram-Schmidt with delayed
// reorthogonalisation (DCGS2, Swirydowicz et al., Numer. Linear Algebra
// Appl. 2021): per iteration one pass over the basis computes the
// projections of the new vector and of the previous, once-orthogonalised
// vector; a second pass finalises the previous vector and updates the new
// one.  The Hessenberg column of the previous step is corrected with the
// reorthogonalisation coefficients and only then committed to the Givens
// least squares, whose algebra (including the sign convention of the
// rotations) is the reference's.
#include "fcm_internal.hpp"

namespace fcm {

static constexpr int GT = 256;   // threads of the control kernels
static constexpr int KRT = 256;  // threads of the basis kernels
static constexpr int KRP = 4;    // rows per thread
static constexpr int KRB = KRT * KRP;

// Rows of the basis passes.  Thread t of block x holds KRP rows, element q at
//   MAP 1: x KRB + t + KRT q, one word at a time -- faster on short vectors
//          (the batches of small systems ran 8% faster);
//   MAP 2: x KRB + 2 t + 2 KRT (q / 2) + q % 2, two words at a time (16-byte
//          loads of an FP64 basis: about 15% faster on the update pass at n = 840k);
//   MAP 4: x KRB + 4 t + q, four words at a time (16-byte loads of an FP32 basis).
// The basis columns are 256-word aligned, so every vector access is aligned.
// Rows at or past n read as zero and are not written.  Every vector the passes
// touch (basis columns in TQ, the new vector in FP64) uses the same rows.
template <int MAP> __device__ __forceinline__ size_t gm_row(int q) {
  const size_t b = (size_t)blockIdx.x * KRB;
  if constexpr (MAP == 1) return b + threadIdx.x + (size_t)q * KRT;
  else if constexpr (MAP == 2) return b + 2 * threadIdx.x + (size_t)(q / 2) * 2 * KRT + (q % 2);
  else return b + 4 * threadIdx.x + q;
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <int MAP, typename T>
__device__ __forceinline__ void gm_load(const T* __restrict__ v, size_t n, double (&o)[KRP]) {
  if constexpr (MAP == 1) {
#pragma unroll
    for (int q = 0; q < KRP; ++q) {
      const size_t r = gm_row<1>(q);
      o[q] = r < n ? (double)v[r] : 0.0;
    }
  } else if constexpr (MAP == 2) {
    using V2 = typename Vec2<T>::type;
#pragma unroll
    for (int p = 0; p < KRP / 2; ++p) {
      const size_t r = gm_row<2>(2 * p);
      if (r + 1 < n) {
        const V2 x = *reinterpret_cast<const V2*>(v + r);
        o[2 * p] = x.x;
        o[2 * p + 1] = x.y;
      } else {
        o[2 * p] = r < n ? (double)v[r] : 0.0;
        o[2 * p + 1] = 0.0;
      }
    }
  } else {
    static_assert(KRP == 4, "four rows per thread");
    const size_t r = gm_row<4>(0);
    if (r + 3 < n) {
      if constexpr (sizeof(T) == 4) {
        const float4 x = *reinterpret_cast<const float4*>(v + r);
        o[0] = x.x;
        o[1] = x.y;
        o[2] = x.z;
        o[3] = x.w;
      } else {
        const double2 x = *reinterpret_cast<const double2*>(v + r), y = *reinterpret_cast<const double2*>(v + r + 2);
        o[0] = x.x;
        o[1] = x.y;
        o[2] = y.x;
        o[3] = y.y;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) o[q] = r + q < n ? (double)v[r + q] : 0.0;
    }
  }
}

template <int MAP, typename T>
__device__ __forceinline__ void gm_store(T* __restrict__ v, size_t n, const double (&x)[KRP]) {
  if constexpr (MAP == 1) {
#pragma unroll
    for (int q = 0; q < KRP; ++q) {
      const size_t r = gm_row<1>(q);
      if (r < n) v[r] = (T)x[q];
    }
  } else if constexpr (MAP == 2) {
    using V2 = typename Vec2<T>::type;
#pragma unroll
    for (int p = 0; p < KRP / 2; ++p) {
      const size_t r = gm_row<2>(2 * p);
      if (r + 1 < n) {
        V2 t;
        t.x = (T)x[2 * p];
        t.y = (T)x[2 * p + 1];
        *reinterpret_cast<V2*>(v + r) = t;
      } else if (r < n) {
        v[r] = (T)x[2 * p];
      }
    }
  } else {
    const size_t r = gm_row<4>(0);
    if (r + 3 < n) {
      if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4*>(v + r) = make_float4((float)x[0], (float)x[1], (float)x[2], (float)x[3]);
      } else {
        double2 a, b;
        a.x = x[0];
        a.y = x[1];
        b.x = x[2];
        b.y = x[3];
        *reinterpret_cast<double2*>(v + r) = a;
        *reinterpret_cast<double2*>(v + r + 2) = b;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (r + q < n) v[r + q] = (T)x[q];
    }
  }
}

__device__ __forceinline__ double* Hp(const GmDev& G, int i, int j) { return G.Hraw + (size_t)i * (G.M + 2) + j; }
__device__ __forceinline__ double* Rp(const GmDev& G, int i, int j) { return G.R + (size_t)i * (G.M + 2) + j; }

// block sum over GT threads (fixed order)
__device__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return r;
}

// The Givens algebra of gmres_solver.cpp:214-252 on a column held in LDS.
// All threads call givens_stage (copies rotations 0..k-1 to LDS); thread 0
// alone calls givens_sweep, which applies them to col[0..k+1] in order and
// forms the new rotation; commit stores the rotation and beta, and the rotated
// column is then written to R by all threads (givens_store).  Returns
// |beta_{k+1}| / ||b||.
static constexpr int GCAP = 2048 + 4;

__device__ __forceinline__ void givens_stage(const GmDev& G, int k, double* sCS, double* sSN) {
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    sCS[i] = G.CS[i];
    sSN[i] = G.SN[i];
  }
}

__device__ double givens_sweep(const GmDev& G, int k, double* col, const double* sCS, const double* sSN,
                               bool commit) {
  double a = col[0];
#pragma unroll 8
  for (int i = 1; i <= k; ++i) {
    const double b = col[i], c = sCS[i - 1], sn = sSN[i - 1];
    col[i - 1] = c * a + sn * b;  // temp = CS h(i-1) + SN h(i)
    a = c * b - sn * a;           // h(i) = CS h(i) - SN h(i-1)
  }
  col[k] = a;
  double cs, sn;
  if (col[k + 1] == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (fabs(col[k + 1]) > fabs(col[k])) {
    const double temp = col[k] / col[k + 1];
    sn = (2.0 * double(col[k + 1] > 0) - 1.0) / sqrt(1.0 + temp * temp);
    cs = temp * sn;
  } else {
    const double temp = col[k + 1] / col[k];
    cs = (2.0 * double(col[k] > 0) - 1.0) / sqrt(1.0 + temp * temp);
    sn = temp * cs;
  }
  const double b1 = -sn * G.beta[k];
  if (commit) {
    col[k] = cs * col[k] + sn * col[k + 1];
    G.CS[k] = cs;
    G.SN[k] = sn;
    G.beta[k + 1] = b1;
    G.beta[k] = cs * G.beta[k];
  }
  return fabs(b1) / G.dctl[1];
}

__device__ __forceinline__ void givens_store(const GmDev& G, int k, const double* col) {
  for (int i = threadIdx.x; i <= k; i += blockDim.x) *Rp(G, i, k) = col[i];
}

__device__ __forceinline__ void set_done(const GmDev& G, int status, int ncols, double rel) {
  G.ictl[GM_STATUS] = 1 + status;
  G.ictl[GM_GATE] = 1;
  G.ictl[GM_NCOLS] = ncols;
  G.dctl[4] = rel;
}

// ---- set-up -----------------------------------------------------------------
// red[0] = ||b||^2 from a preceding reduction
__global__ void __launch_bounds__(GT) gm_start_kernel(GmDev G, double tol, int max_iter) {
  // Hraw and R need no clearing: every entry a step reads was written earlier
  // in the same solve (column i holds rows 0..i+1 once it is final).
  if (threadIdx.x == 0) {
    const double nrm = sqrt(G.red[0]);
    G.dctl[0] = tol;
    G.dctl[1] = nrm;
    G.dctl[2] = 0.0;
    G.dctl[3] = 1.0;
    G.dctl[4] = 1.0;
    G.ictl[GM_K] = 0;
    G.ictl[GM_PROV] = 0;
    G.ictl[GM_FIN] = 0;
    G.ictl[GM_MAXIT] = max_iter;
    G.ictl[GM_NCOLS] = 0;
    G.ictl[GM_ARR_FIN] = 0;
    G.ictl[GM_STATUS] = nrm == 0.0 ? 1 + 2 /* zero rhs */ : 0;
    G.ictl[GM_GATE] = nrm == 0.0 ? 1 : 0;
    G.beta[0] = nrm;
  }
}

// Q col 0 = b / ||b||
template <typename TQ>
__global__ void gm_q0_kernel(GmDev G, const double* __restrict__ b, TQ* __restrict__ Q, size_t n) {
  if (G.ictl[GM_STATUS] != 0) return;
  const double s = 1.0 / G.dctl[1];
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (size_t)gridDim.x * blockDim.x)
    Q[r] = (TQ)(b[r] * s);
}

// ---- basis passes ------------------------------------------------------------
// main step: columns 0..k against v1 = Q col k (if provisional) and v2 = w
// finalising step (FIN = 1): columns 0..k+1 against v1 = Q col k+1
// Grid (row blocks, DOT_CG): each block takes every DOT_CG-th group of 8
// columns, so that enough waves stream the basis at once (8 against 4: FP64
// solves 0.9% faster at 256^3, FP32+refine unchanged; 2: FP32+refine 0.8% slower).
static constexpr int DOT_CG = 8;

template <int MAP, typename TQ>
__global__ void __launch_bounds__(KRT) gm_dots_kernel(GmDev G, const TQ* __restrict__ Q, size_t ldq,
                                                      const double* __restrict__ wbuf, size_t n,
                                                      double* __restrict__ partial, int nblk) {
  if (G.ictl[GM_STATUS] != 0) return;
  const int mode = G.ictl[GM_FIN] == 1 ? 1 : 0;
  const int k = G.ictl[GM_K];
  const bool prov = G.ictl[GM_PROV] != 0;
  const int ncols = mode == 0 ? k + 1 : k + 2;
  if ((int)blockIdx.y * 8 >= ncols) return;
  const TQ* v1 = mode == 0 ? (prov ? Q + (size_t)k * ldq : nullptr) : Q + (size_t)(k + 1) * ldq;
  const double* v2 = mode == 0 ? wbuf : nullptr;
  __shared__ double red[16 * 16];  // [row of 16 lanes][value]
  double a[KRP], b[KRP];
  if (v1) {
    gm_load<MAP>(v1, n, a);
  } else {
#pragma unroll
    for (int q = 0; q < KRP; ++q) a[q] = 0.0;
  }
  if (v2) {
    gm_load<MAP>(v2, n, b);
  } else {
#pragma unroll
    for (int q = 0; q < KRP; ++q) b[q] = 0.0;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c0 = blockIdx.y * 8; c0 < ncols; c0 += DOT_CG * 8) {
    double qv[8][KRP];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if (c0 + c < ncols) {
        gm_load<MAP>(Q + (size_t)(c0 + c) * ldq, n, qv[c]);
      } else {
#pragma unroll
        for (int q = 0; q < KRP; ++q) qv[c][q] = 0.0;
      }
    }
    double acc[16];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      acc[c] = 0.0;
      acc[8 + c] = 0.0;
#pragma unroll
      for (int qq = 0; qq < KRP; ++qq) {
        acc[c] += qv[c][qq] * a[qq];
        acc[8 + c] += qv[c][qq] * b[qq];
      }
    }
    // each row of 16 lanes reduce-scatters the 16 sums; the 16 row sums of
    // every value are then added in a fixed order
    red[(wv * 4 + (lane >> 4)) * 16 + rs16_index(lane)] = row_reduce_scatter16(acc, lane);
    __syncthreads();
    if (threadIdx.x < 16) {
      const int c = threadIdx.x & 7, which = threadIdx.x >> 3;
      const bool have = which == 0 ? v1 != nullptr : v2 != nullptr;
      double q[4];
#pragma unroll
      for (int w = 0; w < 4; ++w)
        q[w] = (red[(4 * w) * 16 + threadIdx.x] + red[(4 * w + 1) * 16 + threadIdx.x]) +
               (red[(4 * w + 2) * 16 + threadIdx.x] + red[(4 * w + 3) * 16 + threadIdx.x]);
      const double v = (q[0] + q[1]) + (q[2] + q[3]);
      if (have && c0 + c < ncols) partial[((size_t)which * ncols + c0 + c) * nblk + blockIdx.x] = v;
    }
    __syncthreads();
  }
}

// G.red[c] = sum_blk partial[c][blk] for rows c = blockIdx.x, blockIdx.x +
// gridDim.x, ... < rows; fixed order (thread-strided sums, then a tree in LDS)
__device__ void reduce_rows(const GmDev& G, const double* __restrict__ partial, int nblk, int rows, double* s) {
  for (int c = blockIdx.x; c < rows; c += gridDim.x) {
    const double* p = partial + (size_t)c * nblk;
    double a = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) a += p[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) G.red[c] = s[0];
    __syncthreads();
  }
}

// True in every thread of the block that arrives last at counter `word` of a
// multi-block kernel (every block calls it once, after its global writes); that
// block then sees what the others wrote and resets the counter.
__device__ bool last_block(const GmDev& G, int word) {
  __shared__ int s_last;
  __threadfence();  // this block's writes before its arrival
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&G.ictl[word], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!s_last) return false;
  __threadfence();  // the other blocks' writes after their arrival
  if (threadIdx.x == 0) G.ictl[word] = 0;
  return true;
}

// end of an iteration: the next one works on column k + 1, provisional after
// a main step, final after a finalising step (whose mat-vec gate reopens)
__device__ __forceinline__ void end_iteration(const GmDev& G, bool finalised) {
  G.ictl[GM_PROV] = finalised ? 0 : 1;
  G.ictl[GM_K] += 1;
  if (finalised) {
    G.ictl[GM_FIN] = 2;
    G.ictl[GM_GATE] = 0;
  }
}

// hnew_l = (proj_l - sum_i H(l,i) s_i) / nu: the projection of A P^-1 q_k on
// q_l of the final basis, corrected for the provisional input (one wavefront,
// H rows are contiguous)
__device__ void hs_row(const GmDev& G, int l, int k, int lane) {
  const double* s = G.red;
  const double* t = G.red + (k + 1);
  const double* row = Hp(G, l, 0);
  double c = 0.0;
  for (int i = (l > 0 ? l - 1 : 0) + lane; i < k; i += 64) c += row[i] * s[i];
  c = wave_sum(c);
  if (lane == 0) {
    const double proj = l < k ? t[l] : G.coef[2 * k + 1];
    G.hnew[l] = (proj - c) / G.dctl[3];
  }
}

// main step: q = (Q col k - sum_{i<m} s_i Q_i) / nu written back (if m >= 0
// and provisional), w <- w - sum_{i<mt} t_i Q_i - pk q, partial ||w||^2; after
// a provisional input the wavefronts also form hnew_0..k (one row each, for
// control step 2).
// finalising step: Q col k+1 <- (Q col k+1 - sum_{i<=k} s_i Q_i) / nu, and the
// last block ends the iteration.
template <int MAP, typename TQ>
__global__ void __launch_bounds__(KRT) gm_update_kernel(GmDev G, TQ* __restrict__ Q, size_t ldq,
                                                        double* __restrict__ wbuf, size_t n,
                                                        double* __restrict__ partial) {
  if (G.ictl[GM_STATUS] != 0) return;
  const int mode = G.ictl[GM_FIN] == 1 ? 1 : 0;
  const int k = G.ictl[GM_K];
  const int m = G.ictl[GM_M], mt = G.ictl[GM_MT];
  if (mode == 0 && G.ictl[GM_HS] != 0) {
    constexpr int WPB = KRT / 64;
    for (int l = (int)blockIdx.x * WPB + (int)(threadIdx.x >> 6); l <= k; l += (int)gridDim.x * WPB)
      hs_row(G, l, k, threadIdx.x & 63);
  }
  TQ* qt = mode == 0 ? (m > 0 ? Q + (size_t)k * ldq : nullptr) : Q + (size_t)(k + 1) * ldq;
  double* w = mode == 0 ? wbuf : nullptr;
  __shared__ double cs[2048 + 4];
  __shared__ double red[4];
  const int ncoef = m + mt + 2;
  for (int i = threadIdx.x; i < ncoef && i < 2048 + 4; i += KRT) cs[i] = G.coef[i];
  __syncthreads();
  auto S = [&](int i) { return i < 2048 + 4 ? cs[i] : G.coef[i]; };
  const double inv_nu = S(m + mt), pk = S(m + mt + 1);
  double as[KRP], at[KRP];
#pragma unroll
  for (int q = 0; q < KRP; ++q) as[q] = at[q] = 0.0;
  const int mm = m > mt ? m : mt;
  constexpr int UC = 4;  // columns whose loads are in flight together
  for (int c0 = 0; c0 < mm; c0 += UC) {
    double qv[UC][KRP];
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      if (c0 + u < mm) {
        gm_load<MAP>(Q + (size_t)(c0 + u) * ldq, n, qv[u]);
      } else {
#pragma unroll
        for (int q = 0; q < KRP; ++q) qv[u][q] = 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int c = c0 + u;
      const double sv = (qt && c < m) ? S(c) : 0.0;
      const double tv = (w && c < mt) ? S(m + c) : 0.0;
#pragma unroll
      for (int qq = 0; qq < KRP; ++qq) {
        as[qq] += sv * qv[u][qq];
        at[qq] += tv * qv[u][qq];
      }
    }
  }
  // the finalised previous vector, as the basis stores it (rounded to TQ), and
  // the new vector; rows at or past n are zero throughout
  double qk[KRP], acc = 0.0;
#pragma unroll
  for (int q = 0; q < KRP; ++q) qk[q] = 0.0;
  if (qt) {
    gm_load<MAP>(qt, n, qk);
#pragma unroll
    for (int q = 0; q < KRP; ++q) qk[q] = (double)(TQ)((qk[q] - as[q]) * inv_nu);
    gm_store<MAP>(qt, n, qk);
  }
  if (w) {
    double x[KRP];
    gm_load<MAP>(w, n, x);
#pragma unroll
    for (int q = 0; q < KRP; ++q) {
      x[q] = x[q] - at[q] - pk * qk[q];
      acc += x[q] * x[q];
    }
    gm_store<MAP>(w, n, x);
  }
  if (mode == 0) {
    acc = wave_sum(acc);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  } else if (last_block(G, GM_ARR_FIN)) {
    // the finalised vector is written: the next iteration works on column k + 1
    if (threadIdx.x == 0) end_iteration(G, true);
  }
}

// ---- control steps (one workgroup of GT threads) ----------------------------------
// A step of the Arnoldi sequence is: the mat-vec (gated), projections,
// reduction, control step 1 (or 3), update (which also forms H s), control
// step 2, scaling.
// When control step 2 predicts convergence, the next step is a finalising
// step: its mat-vec kernels return at once (GM_GATE), and its projections,
// reduction, control and update kernels finalise the new vector (control
// step 3 in place of step 1); control step 2 and the scaling return at once.
// So the finalisation, needed about once per solve, costs no launches in the
// other steps.
struct CtlShared {
  double sh[8];
  double col[GCAP], sCS[GCAP], sSN[GCAP];
  int done;  // the solve has stopped
};

// after the main dots pass: finalise column k-1 (when the previous vector was
// provisional), commit it to the least squares, and set up the update
__device__ void ctl1_step(const GmDev& G, CtlShared& L) {
  double* sh = L.sh;
  double *col = L.col, *sCS = L.sCS, *sSN = L.sSN;
  int& s_done = L.done;
  const int k = G.ictl[GM_K];
  const bool prov = G.ictl[GM_PROV] != 0;
  const int ncols = k + 1;
  const double* s = G.red;           // Q.v1, k+1 values (s_0..s_{k-1}, alpha)
  const double* t = G.red + ncols;   // Q.v2, k+1 values
  if (prov) {
    // every load first (one round trip): the projections, the provisional
    // column k-1 of H (finalised here with the reorthogonalisation
    // coefficients s) and the rotations; the update's coefficients are the
    // projections themselves
    const double eta = G.dctl[2];
    const double alpha = s[k], tau = t[k];
    double ss = 0.0, st = 0.0;
    for (int i = threadIdx.x; i < k; i += GT) {
      const double si = s[i], ti = t[i];
      ss += si * si;
      st += si * ti;
      const double h = G.hprov[i] + eta * si;
      *Hp(G, i, k - 1) = h;
      col[i] = h;
      G.coef[i] = si;
      G.coef[k + i] = ti;
    }
    givens_stage(G, k - 1, sCS, sSN);
    ss = block_sum(ss, sh);
    st = block_sum(st, sh);
    const double nu = sqrt(fmax(alpha - ss, 0.0));
    if (threadIdx.x == 0) {
      *Hp(G, k, k - 1) = eta * nu;
      col[k] = eta * nu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const double rel = givens_sweep(G, k - 1, col, sCS, sSN, true);
      G.hist[k - 1] = rel;
      s_done = 0;
      if (rel < G.dctl[0] || k == G.ictl[GM_MAXIT]) {
        set_done(G, rel < G.dctl[0] ? 0 : 1, k, rel);
        s_done = 1;
      } else if (!(nu > 0.0)) {
        set_done(G, 3, k, rel);
        s_done = 1;
      }
      G.dctl[3] = nu;
      // the projections of A P^-1 q_k on q_0..q_k are finished by the update (hs_row)
      G.ictl[GM_HS] = 1;
      G.ictl[GM_M] = k;
      G.ictl[GM_MT] = k;
      G.coef[2 * k] = 1.0 / nu;
      G.coef[2 * k + 1] = (tau - st) / nu;
    }
    __syncthreads();
    givens_store(G, k - 1, col);
  } else {
    for (int l = threadIdx.x; l <= k; l += GT) {
      G.hnew[l] = t[l];
      G.coef[l] = t[l];
    }
    if (threadIdx.x == 0) {
      G.coef[k + 1] = 1.0;
      G.coef[k + 2] = 0.0;
      G.ictl[GM_M] = 0;
      G.ictl[GM_MT] = k + 1;
      G.ictl[GM_HS] = 0;
      G.dctl[3] = 1.0;
    }
  }
}

// rows of the projections: 2 (k+1) in a main step, k + 2 in a finalising one
__global__ void __launch_bounds__(256) gm_reduce_kernel(GmDev G, const double* __restrict__ partial, int nblk) {
  if (G.ictl[GM_STATUS] != 0) return;
  __shared__ double s[256];
  const int k = G.ictl[GM_K];
  reduce_rows(G, partial, nblk, G.ictl[GM_FIN] == 1 ? k + 2 : 2 * (k + 1), s);
}

__device__ void ctl3_step(const GmDev& G, CtlShared& L);

// control step 1, or step 3 in a finalising step (one workgroup); the first
// main step after a finalising one clears its mark
__global__ void __launch_bounds__(GT) gm_ctl1_kernel(GmDev G) {
  if (G.ictl[GM_STATUS] != 0) return;
  __shared__ CtlShared L;
  const int fin = G.ictl[GM_FIN];
  if (fin == 1) {
    ctl3_step(G, L);
    return;
  }
  __syncthreads();  // every thread has read the mark
  if (fin == 2 && threadIdx.x == 0) G.ictl[GM_FIN] = 0;
  ctl1_step(G, L);
}

// after the main update: normalisation, breakdown, predicted convergence
// (one block; ||w||^2 from the update pass's block partials, fixed order)
__global__ void __launch_bounds__(GT) gm_ctl2_kernel(GmDev G, const double* __restrict__ partial, int nblk) {
  if (G.ictl[GM_STATUS] != 0 || G.ictl[GM_FIN] != 0) return;
  __shared__ double col[GCAP], sCS[GCAP], sSN[GCAP], s[GT];
  __shared__ int s_brk;
  const int k = G.ictl[GM_K];
  const double nu = G.dctl[3];
  // every load first (one round trip): the update pass's block partials of
  // ||w||^2 (summed as reduce_rows does), the new column and the rotations
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += GT) a += partial[i];
  for (int l = threadIdx.x; l <= k; l += GT) {
    const double h = G.hnew[l];
    G.hprov[l] = h;
    col[l] = h;
  }
  givens_stage(G, k, sCS, sSN);
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = GT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) G.red[0] = s[0];
  const double etar = sqrt(s[0]);
  const double eta = etar / nu;
  if (threadIdx.x == 0) {
    G.dctl[2] = eta;
    s_brk = !(etar > 0.0);
    if (s_brk) {
      // exact breakdown: the Krylov space is invariant, column k is exact
      for (int i = 0; i <= k; ++i) *Hp(G, i, k) = col[i];
      *Hp(G, k + 1, k) = 0.0;
      col[k + 1] = 0.0;
      const double rel = givens_sweep(G, k, col, sCS, sSN, true);
      G.hist[k] = rel;
      set_done(G, 3, k + 1, rel);
    } else {
      G.dctl[5] = 1.0 / etar;
      col[k + 1] = eta;
      const double est = givens_sweep(G, k, col, sCS, sSN, false);
      if (est < G.dctl[0] || k + 1 == G.ictl[GM_MAXIT]) {
        // the next step finalises the new vector (its mat-vec is gated off)
        G.ictl[GM_FIN] = 1;
        G.ictl[GM_GATE] = 1;
      } else {
        end_iteration(G, false);
      }
    }
  }
  __syncthreads();
  if (s_brk) givens_store(G, k, col);
}

// Q col k+1 = w / ||w|| (control step 2 has already advanced K unless the
// next step finalises)
template <typename TQ>
__global__ void gm_scale_kernel(GmDev G, const double* __restrict__ wbuf, TQ* __restrict__ Q, size_t ldq, size_t n) {
  const int fin = G.ictl[GM_FIN];
  if (G.ictl[GM_STATUS] != 0 || fin == 2) return;
  const int col = fin == 1 ? G.ictl[GM_K] + 1 : G.ictl[GM_K];
  const double s = G.dctl[5];
  TQ* q = Q + (size_t)col * ldq;
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (size_t)gridDim.x * blockDim.x)
    q[r] = (TQ)(wbuf[r] * s);
}

// control step 3, after the projections of a finalising step
__device__ void ctl3_step(const GmDev& G, CtlShared& L) {
  double* sh = L.sh;
  double *col = L.col, *sCS = L.sCS, *sSN = L.sSN;
  int& s_done = L.done;
  const int k = G.ictl[GM_K];
  const double* s = G.red;  // k+2 values: s_0..s_k, alpha
  double ss = 0.0;
  for (int i = threadIdx.x; i <= k; i += GT) ss += s[i] * s[i];
  ss = block_sum(ss, sh);
  const double nu2 = sqrt(fmax(s[k + 1] - ss, 0.0));
  const double eta = G.dctl[2];
  for (int i = threadIdx.x; i <= k; i += GT) {
    const double h = G.hprov[i] + eta * s[i];
    *Hp(G, i, k) = h;
    col[i] = h;
  }
  if (threadIdx.x == 0) {
    *Hp(G, k + 1, k) = eta * nu2;
    col[k + 1] = eta * nu2;
  }
  givens_stage(G, k, sCS, sSN);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rel = givens_sweep(G, k, col, sCS, sSN, true);
    G.hist[k] = rel;
    s_done = 0;
    if (rel < G.dctl[0] || k + 1 == G.ictl[GM_MAXIT]) {
      set_done(G, rel < G.dctl[0] ? 0 : 1, k + 1, rel);
      s_done = 1;
    } else if (!(nu2 > 0.0)) {
      set_done(G, 3, k + 1, rel);
      s_done = 1;
    }
  }
  __syncthreads();
  givens_store(G, k, col);
  if (s_done) return;
  for (int i = threadIdx.x; i <= k; i += GT) G.coef[i] = s[i];
  if (threadIdx.x == 0) {
    G.coef[k + 1] = 1.0 / nu2;
    G.coef[k + 2] = 0.0;
    G.ictl[GM_M] = k + 1;
    G.ictl[GM_MT] = 0;
  }
}

// back substitution (gmres_solver.cpp:260-276) into G.y; one wavefront
__global__ void __launch_bounds__(64) gm_backsub_kernel(GmDev G) {
  const int st = G.ictl[GM_STATUS] - 1;
  const int nc = G.ictl[GM_NCOLS];
  if (st == 2 || nc <= 0) return;  // zero right-hand side: nothing to solve
  __shared__ double ys[2048 + 4];
  const int lane = threadIdx.x;
  for (int i = nc - 1; i >= 0; --i) {
    double a = 0.0;
    for (int j = i + 1 + lane; j < nc; j += 64) a += *Rp(G, i, j) * ys[j];
    a = wave_sum(a);
    if (lane == 0) ys[i] = (G.beta[i] - a) / *Rp(G, i, i);
    __syncthreads();
  }
  for (int j = lane; j < nc; j += 64) G.y[j] = ys[j];
}

// x = sum_{i < ncols} y_i Q_i (zero for a zero right-hand side)
template <typename TQ>
__global__ void gm_combine_kernel(GmDev G, const TQ* __restrict__ Q, size_t ldq, double* __restrict__ x, size_t n) {
  const int st = G.ictl[GM_STATUS] - 1;
  const int nc = st == 2 ? 0 : G.ictl[GM_NCOLS];
  for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nc; ++i) s += G.y[i] * (double)Q[(size_t)i * ldq + r];
    x[r] = s;
  }
}

// ---- control snapshots ----------------------------------------------------------------
// dst[16 j + e] = src[j][e] for the control blocks of a batch's jobs (null: zeros)
__global__ void gm_gather_kernel(const double* const* __restrict__ src, double* __restrict__ dst, int n) {
  const int j = blockIdx.x * 16 + (threadIdx.x >> 4), e = threadIdx.x & 15;
  if (j < n) {
    const double* p = src[j];
    dst[(size_t)j * 16 + e] = p ? p[e] : 0.0;
  }
}

__global__ void gm_set_ptr_kernel(const double** table, int idx, const double* p) {
  if (threadIdx.x == 0) table[idx] = p;
}

hipError_t gm_gather_ctl(const double* const* src, double* dst, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gm_gather_kernel, dim3((n + 15) / 16), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

hipError_t gm_set_ptr(const double** table, int idx, const double* p, hipStream_t st) {
  hipLaunchKernelGGL(gm_set_ptr_kernel, dim3(1), dim3(64), 0, st, table, idx, p);
  return hipGetLastError();
}

// ---- host-side launch helpers -------------------------------------------------------
int gm_blocks(size_t n) { return (int)((n + KRB - 1) / KRB); }

template <typename TQ>
static hipError_t gm_launch_start_t(const GmDev& G, double tol, int max_iter, const double* b, TQ* Q, size_t n,
                                    hipStream_t st) {
  hipLaunchKernelGGL(gm_start_kernel, dim3(1), dim3(GT), 0, st, G, tol, max_iter);
  hipLaunchKernelGGL(gm_q0_kernel<TQ>, dim3(1024), dim3(256), 0, st, G, b, Q, n);
  return hipGetLastError();
}
hipError_t gm_launch_start(const GmDev& G, double tol, int max_iter, const double* b, double* Q, size_t n,
                           hipStream_t st) {
  return gm_launch_start_t(G, tol, max_iter, b, Q, n, st);
}
hipError_t gm_launch_start(const GmDev& G, double tol, int max_iter, const double* b, float* Q, size_t n,
                           hipStream_t st) {
  return gm_launch_start_t(G, tol, max_iter, b, Q, n, st);
}

template <int MAP, typename TQ>
static void gm_launch_basis_t(const GmDev& G, TQ* Q, size_t ldq, double* wbuf, size_t n, double* partial, int nblk,
                              hipStream_t st) {
  // projections, reduce, control 1 (3), update (with H s), norm reduce +
  // control 2, scale; each kernel reads from GM_FIN whether the step is a
  // finalising one
  hipLaunchKernelGGL((gm_dots_kernel<MAP, TQ>), dim3(nblk, DOT_CG), dim3(KRT), 0, st, G, Q, ldq, wbuf, n, partial,
                     nblk);
  hipLaunchKernelGGL(gm_reduce_kernel, dim3(256), dim3(256), 0, st, G, partial, nblk);
  hipLaunchKernelGGL(gm_ctl1_kernel, dim3(1), dim3(GT), 0, st, G);
  hipLaunchKernelGGL((gm_update_kernel<MAP, TQ>), dim3(nblk), dim3(KRT), 0, st, G, Q, ldq, wbuf, n, partial);
  hipLaunchKernelGGL(gm_ctl2_kernel, dim3(1), dim3(GT), 0, st, G, partial, nblk);
  hipLaunchKernelGGL(gm_scale_kernel<TQ>, dim3(1024), dim3(256), 0, st, G, wbuf, Q, ldq, n);
}

// 16-byte accesses on long vectors (two FP64 or four FP32 words), single
// words on short ones
hipError_t gm_launch_basis(const GmDev& G, double* Q, size_t ldq, double* wbuf, size_t n, double* partial,
                           int nblk, hipStream_t st) {
  if (n >= ((size_t)1 << 18))
    gm_launch_basis_t<2>(G, Q, ldq, wbuf, n, partial, nblk, st);
  else
    gm_launch_basis_t<1>(G, Q, ldq, wbuf, n, partial, nblk, st);
  return hipGetLastError();
}
hipError_t gm_launch_basis(const GmDev& G, float* Q, size_t ldq, double* wbuf, size_t n, double* partial,
                           int nblk, hipStream_t st) {
  if (n >= ((size_t)1 << 18))
    gm_launch_basis_t<4>(G, Q, ldq, wbuf, n, partial, nblk, st);
  else
    gm_launch_basis_t<1>(G, Q, ldq, wbuf, n, partial, nblk, st);
  return hipGetLastError();
}

template <typename TQ>
static hipError_t gm_launch_finish_t(const GmDev& G, const TQ* Q, size_t ldq, double* x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(gm_backsub_kernel, dim3(1), dim3(64), 0, st, G);
  hipLaunchKernelGGL(gm_combine_kernel<TQ>, dim3(2048), dim3(256), 0, st, G, Q, ldq, x, n);
  return hipGetLastError();
}
hipError_t gm_launch_finish(const GmDev& G, const double* Q, size_t ldq, double* x, size_t n, hipStream_t st) {
  return gm_launch_finish_t(G, Q, ldq, x, n, st);
}
hipError_t gm_launch_finish(const GmDev& G, const float* Q, size_t ldq, double* x, size_t n, hipStream_t st) {
  return gm_launch_finish_t(G, Q, ldq, x, n, st);
}

}  // namespace fcm
