// fcm_solver.hip -- rigid-body operators, the block-diagonal preconditioner
// and the Krylov vector kernels of the saddle-point solve (FP64 throughout).
#include <algorithm>

#include "fcm_internal.hpp"

namespace fcm {

__device__ __forceinline__ double wrapd(double p, double L) { return p - L * floor(p / L); }

// quaternion.cpp:61-103 (row-major R)
__device__ __forceinline__ void quat_matrix(const double* q, double* R) {
  const double s = q[0], v0 = q[1], v1 = q[2], v2 = q[3];
  double M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double t = 2.0 * v0 * v0; M[4] -= t; M[8] -= t;
  t = 2.0 * v1 * v1; M[0] -= t; M[8] -= t;
  t = 2.0 * v2 * v2; M[0] -= t; M[4] -= t;
  t = 2.0 * v0 * v1; M[3] += t; M[1] += t;
  t = 2.0 * v0 * v2; M[6] += t; M[2] += t;
  t = 2.0 * v1 * v2; M[5] += t; M[7] += t;
  t = 2.0 * s * v2; M[3] += t; M[1] -= t;
  t = 2.0 * s * v1; M[6] -= t; M[2] += t;
  t = 2.0 * s * v0; M[7] += t; M[5] -= t;
  for (int i = 0; i < 9; ++i) R[i] = M[i];
}

// one thread per blob: R of its body, world-frame offset R r_m, wrapped
// position x + R r_m (collection_of_rigid_bodies.cpp:155-173)
__global__ void bodies_kernel(BodyData b, double LX, double LY, double LZ) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = b.NB * b.nbpb;
  if (id >= nb) return;
  const int n = id / b.nbpb, m = id % b.nbpb;
  double R[9];
  quat_matrix(b.q + 4 * (size_t)n, R);
  if (m == 0)
    for (int i = 0; i < 9; ++i) b.R[9 * (size_t)n + i] = R[i];
  const double* r = b.ref + 3 * m;
  const double d0 = R[0] * r[0] + R[1] * r[1] + R[2] * r[2];
  const double d1 = R[3] * r[0] + R[4] * r[1] + R[5] * r[2];
  const double d2 = R[6] * r[0] + R[7] * r[1] + R[8] * r[2];
  b.dpos[3 * (size_t)id] = d0;
  b.dpos[3 * (size_t)id + 1] = d1;
  b.dpos[3 * (size_t)id + 2] = d2;
  const double* x = b.x + 3 * (size_t)n;
  b.Y[3 * (size_t)id] = wrapd(x[0] + d0, LX);
  b.Y[3 * (size_t)id + 1] = wrapd(x[1] + d1, LY);
  b.Y[3 * (size_t)id + 2] = wrapd(x[2] + d2, LZ);
}

hipError_t bodies_update(const GridGeom& g, BodyData& b, hipStream_t st) {
  const int nb = b.NB * b.nbpb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(bodies_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, b, g.dx * g.nx, g.dx * g.ny,
                     g.dx * g.nz);
  return hipGetLastError();
}

// The rigid-body velocity that best fits blob velocities in the body frame
// (collection_of_rigid_bodies.cpp:583-599, 925-944): from b = K^T V (the
// world-frame sums [sum v; sum (R r_m) x v]), U = diag(R, R) KTKinv
// diag(R^T, R^T) b with KTKinv = (K_ref^T K_ref)^-1.  One thread per body.
__global__ void body_fit_kernel(BodyData b, const double* __restrict__ KTv, const double* __restrict__ KTKinv,
                                double* __restrict__ U) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= b.NB) return;
  const double* R = b.R + 9 * (size_t)n;
  const double* v = KTv + 6 * (size_t)n;
  double bb[6], u[6];
  for (int h = 0; h < 2; ++h)
    for (int i = 0; i < 3; ++i)  // (R^T b)_i
      bb[3 * h + i] = R[i] * v[3 * h] + R[3 + i] * v[3 * h + 1] + R[6 + i] * v[3 * h + 2];
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += KTKinv[6 * i + k] * bb[k];
    u[i] = s;
  }
  for (int h = 0; h < 2; ++h)
    for (int i = 0; i < 3; ++i)
      U[6 * (size_t)n + 3 * h + i] = R[3 * i] * u[3 * h] + R[3 * i + 1] * u[3 * h + 1] + R[3 * i + 2] * u[3 * h + 2];
}

hipError_t body_fit(const BodyData& b, const double* KTv, const double* KTKinv, double* U, hipStream_t st) {
  if (b.NB <= 0) return hipSuccess;
  hipLaunchKernelGGL(body_fit_kernel, dim3((b.NB + 255) / 256), dim3(256), 0, st, b, KTv, KTKinv, U);
  return hipGetLastError();
}

// V6 of every body = e (the common displacement of calculate_divU's finite
// difference, collection_of_rigid_bodies.cpp:877-897)
__global__ void fill_body6_kernel(double* __restrict__ V6, int NB, int i, double e) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < 6 * NB) V6[r] = (r % 6 == i) ? e : 0.0;
}

hipError_t fill_body6(double* V6, int NB, int i, double e, hipStream_t st) {
  if (NB <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_body6_kernel, dim3((6 * NB + 255) / 256), dim3(256), 0, st, V6, NB, i, e);
  return hipGetLastError();
}

// acc[i] = sum over bodies of (U_n(i) - U0_n(i)) / e, one workgroup, fixed
// order (strided partial sums, then a tree): one direction of calculate_divU
// (collection_of_rigid_bodies.cpp:946)
__global__ void __launch_bounds__(256) divu_kernel(const double* __restrict__ U, const double* __restrict__ U0, int NB,
                                                   int i, double e, double* __restrict__ acc) {
  __shared__ double red[256];
  double s = 0.0;
  for (int n = threadIdx.x; n < NB; n += 256) s += (U[6 * (size_t)n + i] - U0[6 * (size_t)n + i]) / e;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) acc[i] = red[0];
}

hipError_t divu_sum(const double* U, const double* U0, int NB, int i, double e, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(divu_kernel, dim3(1), dim3(256), 0, st, U, U0, NB, i, e, acc);
  return hipGetLastError();
}

// check_for_channel_wall_overlap (collection_of_rigid_bodies.cpp:207-239): one
// workgroup strides over the wrapped blob positions and ORs y == 0 || y >= LY/2
__global__ void __launch_bounds__(1024) wall_overlap_kernel(const double* __restrict__ Y, int nb, double half_ly,
                                                            int* __restrict__ flag) {
  int hit = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double y = Y[3 * (size_t)i + 1];
    hit |= (y == 0.0 || y >= half_ly) ? 1 : 0;
  }
  hit = __syncthreads_or(hit);
  if (threadIdx.x == 0) *flag = hit ? 1 : 0;
}

hipError_t wall_overlap(const double* Y, int nb, double half_ly, int* flag, hipStream_t st) {
  hipLaunchKernelGGL(wall_overlap_kernel, dim3(1), dim3(1024), 0, st, Y, nb, half_ly, flag);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// P^{-1}, gmres_solver.cpp:85-146, evaluated in the body frame: with R the
// body rotation, Minv_n = (I(x)R) Minv_ref (I(x)R)^T and N_n = diag(R,R) N_ref
// diag(R,R)^T, so every product can act on rotated vectors instead of on
// rotated matrices:
//   b' = R^T b_lambda,  t = Minv_ref b',  s' = -diag(R^T,R^T) b_V - K_ref^T t,
//   V' = N_ref s',  out_V = diag(R,R) V',  out_lambda = R (Minv_ref (K_ref V' + b')).
// One workgroup per body; optionally also -K^T out_lambda for the mat-vec.
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void __launch_bounds__(128) pinv_kernel(BodyData b, const TI* __restrict__ in, double* __restrict__ out,
                                                   double* __restrict__ wKT) {
  extern __shared__ double sm[];
  const int n3 = 3 * b.nbpb;
  double* bp = sm;            // n3
  double* tt = bp + n3;       // n3
  double* gg = tt + n3;       // n3
  double* oo = gg + n3;       // n3
  double* sv = oo + n3;       // 6
  double* vv = sv + 6;        // 6
  double* Rs = vv + 6;        // 9
  double* ref = Rs + 9;       // n3
  if (b.gate && *b.gate) return;
  if (b.kcol) in += (size_t)(*b.kcol) * b.ldq;
  const int n = blockIdx.x;
  const int t = threadIdx.x;
  const size_t nlam = 3 * (size_t)b.NB * b.nbpb;
  const TI* inl = in + (size_t)n3 * n;
  const TI* inv = in + nlam + 6 * (size_t)n;
  if (t < 9) Rs[t] = b.R[9 * (size_t)n + t];
  for (int i = t; i < n3; i += blockDim.x) ref[i] = b.ref[i];
  __syncthreads();
  for (int i = t; i < n3; i += blockDim.x) {
    const int m = i / 3, c = i % 3;
    const TI* v = inl + 3 * m;
    bp[i] = Rs[c] * (double)v[0] + Rs[3 + c] * (double)v[1] + Rs[6 + c] * (double)v[2];  // (R^T v)_c
  }
  __syncthreads();
  for (int i = t; i < n3; i += blockDim.x) {
    const double* row = b.Minv + (size_t)i * n3;
    double s = 0.0;
    for (int k = 0; k < n3; ++k) s += row[k] * bp[k];
    tt[i] = s;
  }
  __syncthreads();
  if (t < 6) {
    const int c = t % 3;
    // (R^T b_V)_c for the translational (t<3) or rotational (t>=3) part
    const TI* v = inv + (t < 3 ? 0 : 3);
    const double rv = Rs[c] * (double)v[0] + Rs[3 + c] * (double)v[1] + Rs[6 + c] * (double)v[2];
    double k = 0.0;
    if (t < 3) {
      for (int m = 0; m < b.nbpb; ++m) k += tt[3 * m + c];
    } else {
      const int c1 = (c + 1) % 3, c2i = (c + 2) % 3;
      for (int m = 0; m < b.nbpb; ++m) k += ref[3 * m + c1] * tt[3 * m + c2i] - ref[3 * m + c2i] * tt[3 * m + c1];
    }
    sv[t] = -rv - k;
  }
  __syncthreads();
  if (t < 6) {
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += b.Nref[t * 6 + k] * sv[k];
    vv[t] = s;
  }
  __syncthreads();
  for (int i = t; i < n3; i += blockDim.x) {
    const int m = i / 3, c = i % 3;
    const int c1 = (c + 1) % 3, c2i = (c + 2) % 3;
    // K_ref V' = U' + Omega' x r_m
    gg[i] = vv[c] + (vv[3 + c1] * ref[3 * m + c2i] - vv[3 + c2i] * ref[3 * m + c1]) + bp[i];
  }
  __syncthreads();
  for (int i = t; i < n3; i += blockDim.x) {
    const double* row = b.Minv + (size_t)i * n3;
    double s = 0.0;
    for (int k = 0; k < n3; ++k) s += row[k] * gg[k];
    oo[i] = s;
  }
  __syncthreads();
  double* outl = out + (size_t)n3 * n;
  for (int i = t; i < n3; i += blockDim.x) {
    const int m = i / 3, c = i % 3;
    const double* o = oo + 3 * m;
    const double v = Rs[3 * c] * o[0] + Rs[3 * c + 1] * o[1] + Rs[3 * c + 2] * o[2];
    outl[i] = v;
    gg[i] = v;  // keep the world-frame result for K^T
  }
  if (t < 6) {
    const int c = t % 3;
    const double* v = vv + (t < 3 ? 0 : 3);
    out[nlam + 6 * (size_t)n + t] = Rs[3 * c] * v[0] + Rs[3 * c + 1] * v[1] + Rs[3 * c + 2] * v[2];
  }
  if (wKT) {
    __syncthreads();
    if (t < 6) {
      // -K^T lambda: -[sum lambda_m ; sum (R r_m) x lambda_m]   (collection_of_rigid_bodies.cpp:749)
      const int c = t % 3;
      double s = 0.0;
      const double* d = b.dpos + 3 * (size_t)n * b.nbpb;
      if (t < 3) {
        for (int m = 0; m < b.nbpb; ++m) s += gg[3 * m + c];
      } else {
        const int c1 = (c + 1) % 3, c2i = (c + 2) % 3;
        for (int m = 0; m < b.nbpb; ++m) s += d[3 * m + c1] * gg[3 * m + c2i] - d[3 * m + c2i] * gg[3 * m + c1];
      }
      wKT[6 * (size_t)n + t] = -s;
    }
  }
}

// The same products with one thread per (body, row): a pass of a workgroup
// takes BP = 256 / 3n_b bodies (7 at n_b = 12, 98% of the lanes), and the
// workgroup loops over passes so that ref_Minv (transposed: the threads of a
// body read consecutive words; copied from the transposed matrix kept in HBM,
// a contiguous read) and ref_N are staged in LDS once.  Every sum
// runs over k in the order of pinv_kernel, so the results are the same bit for
// bit.  Used when 6 <= 3 n_b <= 72 (the threads of a body's rows also compute
// its six body rows, and the LDS stays under 64 KiB).  N3 = 3 n_b when it is
// compiled in (the products then unroll with constant LDS offsets), 0 for any.
// Each thread loads the column and the row of its body's rotation that its
// component c needs, instead of all nine words, so that no register array is
// indexed at run time (that had cost a select per word and use).
template <typename TI, int N3>
__global__ void __launch_bounds__(256) pinv_pack_kernel(BodyData b, const TI* __restrict__ in,
                                                        double* __restrict__ out, double* __restrict__ wKT) {
  extern __shared__ double sm[];
  if (b.gate && *b.gate) return;
  if (b.kcol) in += (size_t)(*b.kcol) * b.ldq;
  const int n3 = N3 > 0 ? N3 : 3 * b.nbpb;
  const int BP = 256 / n3;
  double* MT = sm;               // n3 x n3, MT[k n3 + i] = Minv[i n3 + k]
  double* NR = MT + n3 * n3;     // 36
  double* ref = NR + 36;         // n3
  double* bp = ref + n3;         // BP x n3 each: R^T v, Minv bp, g, Minv g
  double* tt = bp + BP * n3;
  double* gg = tt + BP * n3;
  double* oo = gg + BP * n3;
  double* sv = oo + BP * n3;     // BP x 6
  double* vv = sv + BP * 6;      // BP x 6
  const int t = threadIdx.x;
  for (int e = t; e < n3 * n3; e += 256) MT[e] = b.MinvT[e];
  if (t < 36) NR[t] = b.Nref[t];
  for (int i = t; i < n3; i += 256) ref[i] = b.ref[i];
  const int q = t / n3, i = t - (t / n3) * n3, m = i / 3, c = i - (i / 3) * 3;
  const int c1 = (c + 1) % 3, c2i = (c + 2) % 3;
  const size_t nlam = 3 * (size_t)b.NB * b.nbpb;
  double* bpq = bp + q * n3;
  double* ttq = tt + q * n3;
  double* ggq = gg + q * n3;
  double* ooq = oo + q * n3;
  double* svq = sv + q * 6;
  double* vvq = vv + q * 6;
  for (int base = blockIdx.x * BP; base < b.NB; base += gridDim.x * BP) {
    const int n = base + q;
    const bool live = q < BP && n < b.NB;
    // every global load of the pass first: column c and row c of the body's
    // rotation, this row's blob vector, the body vector (rows 0..5) and
    // components c1, c2 of the blob offset R r_m (for -K^T of the output)
    double Rc[3] = {0.0, 0.0, 0.0}, Rr[3] = {0.0, 0.0, 0.0}, v3[3] = {0.0, 0.0, 0.0}, w3[3] = {0.0, 0.0, 0.0};
    double dc1 = 0.0, dc2 = 0.0;
    if (live) {
      const double* Rn = b.R + 9 * (size_t)n;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        Rc[r] = Rn[3 * r + c];
        Rr[r] = Rn[3 * c + r];
      }
      const TI* v = in + (size_t)n3 * n + 3 * m;
      v3[0] = v[0]; v3[1] = v[1]; v3[2] = v[2];
      if (i < 6) {
        const TI* vb = in + nlam + 6 * (size_t)n + (i < 3 ? 0 : 3);
        w3[0] = vb[0]; w3[1] = vb[1]; w3[2] = vb[2];
      }
      if (wKT) {
        const double* d = b.dpos + 3 * ((size_t)n * b.nbpb + m);
        dc1 = d[c1];
        dc2 = d[c2i];
      }
    }
    if (live) bpq[i] = Rc[0] * v3[0] + Rc[1] * v3[1] + Rc[2] * v3[2];  // (R^T v)_c
    __syncthreads();
    if (live) {
      double s = 0.0;
#pragma unroll 12
      for (int k = 0; k < n3; ++k) s += MT[k * n3 + i] * bpq[k];
      ttq[i] = s;
    }
    __syncthreads();
    if (live && i < 6) {
      const int cc = c;  // i % 3
      const double rv = Rc[0] * w3[0] + Rc[1] * w3[1] + Rc[2] * w3[2];
      double k = 0.0;
      if (i < 3) {
#pragma unroll 4
        for (int mm = 0; mm < b.nbpb; ++mm) k += ttq[3 * mm + cc];
      } else {
        const int d1 = (cc + 1) % 3, d2 = (cc + 2) % 3;
#pragma unroll 4
        for (int mm = 0; mm < b.nbpb; ++mm) k += ref[3 * mm + d1] * ttq[3 * mm + d2] - ref[3 * mm + d2] * ttq[3 * mm + d1];
      }
      svq[i] = -rv - k;
    }
    __syncthreads();
    if (live && i < 6) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += NR[i * 6 + k] * svq[k];
      vvq[i] = s;
    }
    __syncthreads();
    if (live) ggq[i] = vvq[c] + (vvq[3 + c1] * ref[3 * m + c2i] - vvq[3 + c2i] * ref[3 * m + c1]) + bpq[i];
    __syncthreads();
    if (live) {
      double s = 0.0;
#pragma unroll 12
      for (int k = 0; k < n3; ++k) s += MT[k * n3 + i] * ggq[k];
      ooq[i] = s;
    }
    __syncthreads();
    if (live) {
      const double* o = ooq + 3 * m;
      const double v = Rr[0] * o[0] + Rr[1] * o[1] + Rr[2] * o[2];
      out[(size_t)n3 * n + i] = v;
      ggq[i] = v;
    }
    if (live && i < 6) {
      const double* v = vvq + (i < 3 ? 0 : 3);  // row c = i % 3 of R
      out[nlam + 6 * (size_t)n + i] = Rr[0] * v[0] + Rr[1] * v[1] + Rr[2] * v[2];
    }
    if (wKT) {
      __syncthreads();
      // torque term of this row: (d_m x lambda_m)_c, from the blob's three outputs
      if (live) {
        const double* l = ggq + 3 * m;
        ttq[i] = dc1 * l[c2i] - dc2 * l[c1];
      }
      __syncthreads();
      if (live && i < 6) {
        const int cc = i % 3;
        double s = 0.0;
        if (i < 3) {
  #pragma unroll 4
        for (int mm = 0; mm < b.nbpb; ++mm) s += ggq[3 * mm + cc];
        } else {
  #pragma unroll 4
        for (int mm = 0; mm < b.nbpb; ++mm) s += ttq[3 * mm + cc];
        }
        wKT[6 * (size_t)n + i] = -s;
      }
    }
    __syncthreads();  // the next pass overwrites the per-body rows
  }
}

template <typename TI>
static hipError_t apply_pinv_t(const BodyData& b, const TI* in, double* out, double* wKT, hipStream_t st) {
  if (b.NB == 0) return hipSuccess;
  const int n3 = 3 * b.nbpb;
  // the packed kernel's rows double as the six body rows: it needs 3 n_b >= 6
  if (n3 >= 6 && n3 <= 72) {
    const int BP = 256 / n3;
    const size_t lds = sizeof(double) * ((size_t)n3 * n3 + 36 + n3 + (size_t)BP * (4 * n3 + 12));
    const int groups = (b.NB + BP - 1) / BP;
    const dim3 grid(groups < 2048 ? groups : 2048);
    switch (n3) {
      case 18: hipLaunchKernelGGL((pinv_pack_kernel<TI, 18>), grid, dim3(256), lds, st, b, in, out, wKT); break;
      case 30: hipLaunchKernelGGL((pinv_pack_kernel<TI, 30>), grid, dim3(256), lds, st, b, in, out, wKT); break;
      case 36: hipLaunchKernelGGL((pinv_pack_kernel<TI, 36>), grid, dim3(256), lds, st, b, in, out, wKT); break;
      default: hipLaunchKernelGGL((pinv_pack_kernel<TI, 0>), grid, dim3(256), lds, st, b, in, out, wKT); break;
    }
    return hipGetLastError();
  }
  const size_t lds = sizeof(double) * (5 * (size_t)n3 + 21);
  hipLaunchKernelGGL(pinv_kernel<TI>, dim3(b.NB), dim3(128), lds, st, b, in, out, wKT);
  return hipGetLastError();
}
hipError_t apply_pinv(const BodyData& b, const double* in, double* out, double* wKT, hipStream_t st) {
  return apply_pinv_t(b, in, out, wKT, st);
}
hipError_t apply_pinv(const BodyData& b, const float* in, double* out, double* wKT, hipStream_t st) {
  return apply_pinv_t(b, in, out, wKT, st);
}

// K [U;W] per blob: U + W x (R r_m)   (collection_of_rigid_bodies.cpp:707-731)
__global__ void mulK_kernel(BodyData b, const double* __restrict__ in6, double* __restrict__ out3) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= b.NB * b.nbpb) return;
  const int n = id / b.nbpb;
  const double* u = in6 + 6 * (size_t)n;
  const double* d = b.dpos + 3 * (size_t)id;
  out3[3 * (size_t)id] = u[0] + (u[4] * d[2] - u[5] * d[1]);
  out3[3 * (size_t)id + 1] = u[1] + (u[5] * d[0] - u[3] * d[2]);
  out3[3 * (size_t)id + 2] = u[2] + (u[3] * d[1] - u[4] * d[0]);
}

// K^T lambda per body: [sum lambda_m; sum (R r_m) x lambda_m]   (:733-757)
__global__ void mulKT_kernel(BodyData b, const double* __restrict__ in3, double* __restrict__ out6) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= b.NB) return;
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int m = 0; m < b.nbpb; ++m) {
    const size_t id = (size_t)n * b.nbpb + m;
    const double* l = in3 + 3 * id;
    const double* d = b.dpos + 3 * id;
    a[0] += l[0]; a[1] += l[1]; a[2] += l[2];
    a[3] += d[1] * l[2] - d[2] * l[1];
    a[4] += d[2] * l[0] - d[0] * l[2];
    a[5] += d[0] * l[1] - d[1] * l[0];
  }
  for (int i = 0; i < 6; ++i) out6[6 * (size_t)n + i] = a[i];
}

hipError_t mul_K(const BodyData& b, const double* in6, double* out3, hipStream_t st) {
  const int nb = b.NB * b.nbpb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(mulK_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, b, in6, out3);
  return hipGetLastError();
}

hipError_t mul_KT(const BodyData& b, const double* in3, double* out6, hipStream_t st) {
  if (b.NB == 0) return hipSuccess;
  hipLaunchKernelGGL(mulKT_kernel, dim3((b.NB + 255) / 256), dim3(256), 0, st, b, in3, out6);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Krylov kernels.  Rows are split into fixed blocks of KB rows; each block
// writes one partial per column and a second kernel sums the partials of a
// column in a fixed tree order, so every dot product is bitwise repeatable.
// ---------------------------------------------------------------------------
static constexpr int KT = 256;        // threads per block
static constexpr int KRPT = 4;        // rows per thread
static constexpr int KB = KT * KRPT;  // rows per block
static constexpr int KCOLS = 8;       // columns reduced together

int krylov_blocks(size_t n) { return (int)((n + KB - 1) / KB); }

__device__ __forceinline__ void block_reduce8(double (&v)[KCOLS], double* red_lds /* 4 x KCOLS */) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < KCOLS; ++c) v[c] = wave_sum(v[c]);
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < KCOLS; ++c) red_lds[wv * KCOLS + c] = v[c];
  __syncthreads();
  if (threadIdx.x < KCOLS) {
    const int c = threadIdx.x;
    v[c] = ((red_lds[c] + red_lds[KCOLS + c]) + (red_lds[2 * KCOLS + c] + red_lds[3 * KCOLS + c]));
  }
}

__global__ void __launch_bounds__(KT) dots_kernel(const double* __restrict__ Q, size_t ldq, int ncols,
                                                  const double* __restrict__ w, size_t n, double* __restrict__ partial,
                                                  int nblk) {
  __shared__ double red[4 * KCOLS];
  const size_t r0 = (size_t)blockIdx.x * KB + threadIdx.x;
  double wv[KRPT];
#pragma unroll
  for (int k = 0; k < KRPT; ++k) {
    const size_t r = r0 + (size_t)k * KT;
    wv[k] = r < n ? w[r] : 0.0;
  }
  for (int c0 = 0; c0 < ncols; c0 += KCOLS) {
    double acc[KCOLS];
#pragma unroll
    for (int c = 0; c < KCOLS; ++c) {
      acc[c] = 0.0;
      if (c0 + c < ncols) {
        const double* q = Q + (size_t)(c0 + c) * ldq;
#pragma unroll
        for (int k = 0; k < KRPT; ++k) {
          const size_t r = r0 + (size_t)k * KT;
          if (r < n) acc[c] += q[r] * wv[k];
        }
      }
    }
    block_reduce8(acc, red);
    if (threadIdx.x < KCOLS && c0 + (int)threadIdx.x < ncols)
      partial[(size_t)(c0 + threadIdx.x) * nblk + blockIdx.x] = acc[threadIdx.x];
    __syncthreads();
  }
}

// out[c] = sum over blocks of partial[c][blk], fixed tree order; one block per column
__global__ void __launch_bounds__(256) reduce_cols_kernel(const double* __restrict__ partial, int nblk,
                                                          double* __restrict__ out, int sqrt_flag) {
  __shared__ double s[256];
  const double* p = partial + (size_t)blockIdx.x * nblk;
  double a = 0.0;
  for (int i = threadIdx.x; i < nblk; i += 256) a += p[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt_flag ? sqrt(s[0]) : s[0];
}

__global__ void __launch_bounds__(KT) update_kernel(const double* __restrict__ Q, size_t ldq, int ncols,
                                                    const double* __restrict__ h, double* __restrict__ w, size_t n,
                                                    double* __restrict__ partial, int nblk) {
  __shared__ double hs[1024];
  __shared__ double red[4 * KCOLS];
  for (int i = threadIdx.x; i < ncols && i < 1024; i += KT) hs[i] = h[i];
  __syncthreads();
  const size_t r0 = (size_t)blockIdx.x * KB + threadIdx.x;
  double wv[KRPT];
#pragma unroll
  for (int k = 0; k < KRPT; ++k) {
    const size_t r = r0 + (size_t)k * KT;
    wv[k] = r < n ? w[r] : 0.0;
  }
  for (int c = 0; c < ncols; ++c) {
    const double hc = c < 1024 ? hs[c] : h[c];
    const double* q = Q + (size_t)c * ldq;
#pragma unroll
    for (int k = 0; k < KRPT; ++k) {
      const size_t r = r0 + (size_t)k * KT;
      if (r < n) wv[k] -= hc * q[r];
    }
  }
  double acc[KCOLS];
#pragma unroll
  for (int c = 0; c < KCOLS; ++c) acc[c] = 0.0;
#pragma unroll
  for (int k = 0; k < KRPT; ++k) {
    const size_t r = r0 + (size_t)k * KT;
    if (r < n) {
      w[r] = wv[k];
      acc[0] += wv[k] * wv[k];
    }
  }
  if (partial) {
    block_reduce8(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
  }
}

// Two dot products per basis column in one pass over the basis:
// partial[c][blk] = Q[c].v1 (c < ncols, skipped if v1 is null) and
// partial[ncols + c][blk] = Q[c].v2 (skipped if v2 is null).
__global__ void __launch_bounds__(KT) dots2_kernel(const double* __restrict__ Q, size_t ldq, int ncols,
                                                   const double* __restrict__ v1, const double* __restrict__ v2,
                                                   size_t n, double* __restrict__ partial, int nblk) {
  __shared__ double red[4 * KCOLS];
  const size_t r0 = (size_t)blockIdx.x * KB + threadIdx.x;
  double a[KRPT], b[KRPT];
#pragma unroll
  for (int k = 0; k < KRPT; ++k) {
    const size_t r = r0 + (size_t)k * KT;
    a[k] = (v1 && r < n) ? v1[r] : 0.0;
    b[k] = (v2 && r < n) ? v2[r] : 0.0;
  }
  constexpr int H = KCOLS / 2;
  for (int c0 = 0; c0 < ncols; c0 += H) {
    double acc[KCOLS];
#pragma unroll
    for (int c = 0; c < H; ++c) {
      acc[c] = 0.0;
      acc[H + c] = 0.0;
      if (c0 + c < ncols) {
        const double* q = Q + (size_t)(c0 + c) * ldq;
#pragma unroll
        for (int k = 0; k < KRPT; ++k) {
          const size_t r = r0 + (size_t)k * KT;
          if (r < n) {
            const double qv = q[r];
            acc[c] += qv * a[k];
            acc[H + c] += qv * b[k];
          }
        }
      }
    }
    block_reduce8(acc, red);
    if (threadIdx.x < KCOLS) {
      const int c = threadIdx.x % H, which = threadIdx.x / H;
      if (c0 + c < ncols && (which == 0 ? v1 != nullptr : v2 != nullptr))
        partial[((size_t)which * ncols + c0 + c) * nblk + blockIdx.x] = acc[threadIdx.x];
    }
    __syncthreads();
  }
}

// One pass of the lagged (delayed) reorthogonalisation:
//   if qt: q = (qt - sum_{i<m} s_i Q_i) * inv_nu, written back over qt
//   if w:  w' = w - sum_{i<mt} t_i Q_i - pk * q     (q term only if qt)
// and the per-block partial of ||w'||^2 (or of ||q||^2 when w is null).
// coef = [s(m) | t(mt) | inv_nu | pk]
__global__ void __launch_bounds__(KT) update2_kernel(const double* __restrict__ Q, size_t ldq, int m, int mt,
                                                     const double* __restrict__ coef, double* __restrict__ qt,
                                                     double* __restrict__ w, size_t n, double* __restrict__ partial) {
  __shared__ double cs[2048 + 4];
  __shared__ double red[4 * KCOLS];
  const int ncoef = m + mt + 2;
  for (int i = threadIdx.x; i < ncoef && i < 2048 + 4; i += KT) cs[i] = coef[i];
  __syncthreads();
  const double* sc = coef;  // fallbacks for very long bases read global memory
  auto S = [&](int i) { return i < 2048 + 4 ? cs[i] : sc[i]; };
  const double inv_nu = S(m + mt), pk = S(m + mt + 1);
  const size_t r0 = (size_t)blockIdx.x * KB + threadIdx.x;
  double as[KRPT], at[KRPT];
#pragma unroll
  for (int k = 0; k < KRPT; ++k) as[k] = at[k] = 0.0;
  const int mm = m > mt ? m : mt;
  for (int c = 0; c < mm; ++c) {
    const double sv = (qt && c < m) ? S(c) : 0.0;
    const double tv = (w && c < mt) ? S(m + c) : 0.0;
    const double* q = Q + (size_t)c * ldq;
#pragma unroll
    for (int k = 0; k < KRPT; ++k) {
      const size_t r = r0 + (size_t)k * KT;
      if (r < n) {
        const double qv = q[r];
        as[k] += sv * qv;
        at[k] += tv * qv;
      }
    }
  }
  double acc[KCOLS];
#pragma unroll
  for (int c = 0; c < KCOLS; ++c) acc[c] = 0.0;
#pragma unroll
  for (int k = 0; k < KRPT; ++k) {
    const size_t r = r0 + (size_t)k * KT;
    if (r < n) {
      double qk = 0.0;
      if (qt) {
        qk = (qt[r] - as[k]) * inv_nu;
        qt[r] = qk;
      }
      if (w) {
        const double v = w[r] - at[k] - pk * qk;
        w[r] = v;
        acc[0] += v * v;
      } else {
        acc[0] += qk * qk;
      }
    }
  }
  block_reduce8(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(KT) combine_kernel(const double* __restrict__ Q, size_t ldq, int ncols,
                                                     const double* __restrict__ c, double* __restrict__ y, size_t n) {
  const size_t r = (size_t)blockIdx.x * KT + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int i = 0; i < ncols; ++i) s += c[i] * Q[(size_t)i * ldq + r];
  y[r] = s;
}

__global__ void scale_kernel(const double* __restrict__ x, double* __restrict__ y, size_t n, const double* a,
                             int invsqrt) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double s = invsqrt ? 1.0 / sqrt(*a) : *a;
  y[r] = x[r] * s;
}

__global__ void axpby_kernel(double a, const double* x, double b, double* y, size_t n) {  // x may alias y
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // a zero coefficient drops its operand, so stale non-finite data is never read into y
  if (r < n) y[r] = (a == 0.0 ? 0.0 : a * x[r]) + (b == 0.0 ? 0.0 : b * y[r]);
}

__global__ void rhs_kernel(const double* __restrict__ Vrand, const double* __restrict__ Vshear,
                           const double* __restrict__ F6, double* __restrict__ b, size_t n3, size_t n6) {
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n3) {
    double v = Vrand ? -Vrand[r] : 0.0;
    if (Vshear) v -= Vshear[r];  // RHS.subvec(0, 3 Nb - 1) -= fcm.Vshear (gmres_solver.cpp:62-64)
    b[r] = v;
  } else if (r < n3 + n6) {
    b[r] = F6 ? -F6[r - n3] : 0.0;
  }
}

hipError_t dots(const double* Q, size_t ldq, int ncols, const double* w, size_t n, KrylovWork& kw, double* red_out,
                hipStream_t st) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(dots_kernel, dim3(kw.nblk), dim3(KT), 0, st, Q, ldq, ncols, w, n, kw.partial, kw.nblk);
  hipLaunchKernelGGL(reduce_cols_kernel, dim3(ncols), dim3(256), 0, st, kw.partial, kw.nblk, red_out, 0);
  return hipGetLastError();
}

hipError_t update(const double* Q, size_t ldq, int ncols, const double* h, double* w, size_t n, KrylovWork& kw,
                  double* norm_out, hipStream_t st) {
  hipLaunchKernelGGL(update_kernel, dim3(kw.nblk), dim3(KT), 0, st, Q, ldq, ncols, h, w, n,
                     norm_out ? kw.partial : (double*)nullptr, kw.nblk);
  if (norm_out) hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(256), 0, st, kw.partial, kw.nblk, norm_out, 0);
  return hipGetLastError();
}

hipError_t dots2(const double* Q, size_t ldq, int ncols, const double* v1, const double* v2, size_t n, KrylovWork& kw,
                 double* red_out, hipStream_t st) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(dots2_kernel, dim3(kw.nblk), dim3(KT), 0, st, Q, ldq, ncols, v1, v2, n, kw.partial, kw.nblk);
  // red_out[0..ncols) = Q.v1, red_out[ncols..2 ncols) = Q.v2 (absent vectors leave zeros untouched)
  if (v1) hipLaunchKernelGGL(reduce_cols_kernel, dim3(ncols), dim3(256), 0, st, kw.partial, kw.nblk, red_out, 0);
  if (v2)
    hipLaunchKernelGGL(reduce_cols_kernel, dim3(ncols), dim3(256), 0, st, kw.partial + (size_t)ncols * kw.nblk, kw.nblk,
                       red_out + ncols, 0);
  return hipGetLastError();
}

hipError_t update2(const double* Q, size_t ldq, int m, int mt, const double* coef, double* qt, double* w, size_t n,
                   KrylovWork& kw, double* norm_out, hipStream_t st) {
  hipLaunchKernelGGL(update2_kernel, dim3(kw.nblk), dim3(KT), 0, st, Q, ldq, m, mt, coef, qt, w, n, kw.partial);
  hipLaunchKernelGGL(reduce_cols_kernel, dim3(1), dim3(256), 0, st, kw.partial, kw.nblk, norm_out, 0);
  return hipGetLastError();
}

hipError_t combine(const double* Q, size_t ldq, int ncols, const double* c, double* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(combine_kernel, dim3((unsigned)((n + KT - 1) / KT)), dim3(KT), 0, st, Q, ldq, ncols, c, y, n);
  return hipGetLastError();
}

hipError_t scale_vec(const double* x, double* y, size_t n, const double* a, bool invsqrt, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, y, n, a, invsqrt ? 1 : 0);
  return hipGetLastError();
}

hipError_t axpby(double a, const double* x, double b, double* y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, x, b, y, n);
  return hipGetLastError();
}

// plain streaming copy (16-byte words), the calibration kernel of the byte counters
__global__ void __launch_bounds__(256) copy16_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                     size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t copy16(const void* src, void* dst, size_t bytes, hipStream_t st) {
  const size_t n = bytes / 16;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(copy16_kernel, dim3(8192), dim3(256), 0, st, static_cast<const float4*>(src),
                     static_cast<float4*>(dst), n);
  return hipGetLastError();
}

hipError_t rhs_assemble(const double* Vrand, const double* Vshear, const double* F6, double* b, size_t n3, size_t n6,
                        hipStream_t st) {
  const size_t n = n3 + n6;
  hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Vrand, Vshear, F6, b, n3, n6);
  return hipGetLastError();
}

// produce_shear_velocity_field's forcing (fcm_fluid_solver.cpp:705-729):
// fx = amp sin(kappa 2 pi y / ny) on the nz real nodes of each row, the padding
// and fy, fz zero; amp = SHEAR_FORCE_SCALE kT / a is applied as the reference
// orders it, (scale kT) sin(...) / a.  One thread per real word.
__global__ void shear_force_kernel(FieldPtrs<double> fp, GridGeom g, double sk, double a, double kappa) {
  constexpr double PI = 3.14159265358979323846264338327950288;  // config.hpp:86
  const size_t rw = 2 * (size_t)g.cpitch;  // real words per row
  const size_t n = (size_t)g.nx * g.ny * rw;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(w % rw);
    const int y = (int)((w / rw) % g.ny);
    double* fx = reinterpret_cast<double*>(fp.p[0]);
    double* fy = reinterpret_cast<double*>(fp.p[1]);
    double* fz = reinterpret_cast<double*>(fp.p[2]);
    fx[w] = k < g.nz ? sk * sin(kappa * 2.0 * PI * (double)y / (double)g.ny) / a : 0.0;
    fy[w] = 0.0;
    fz[w] = 0.0;
  }
}

hipError_t shear_force(const GridGeom& g, const FieldPtrs<double>& fp, double amp_sk, double a, double kappa,
                       hipStream_t st) {
  const size_t n = (size_t)g.nx * g.ny * 2 * g.cpitch;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(shear_force_kernel, dim3(blocks), dim3(256), 0, st, fp, g, amp_sk, a, kappa);
  return hipGetLastError();
}

// main.cpp:371-433: one workgroup per y row sums u_x (+ the shear u_x) over x
// and the nz real nodes of z, then divides by nx nz
__global__ void __launch_bounds__(256) x_profile_kernel(FieldPtrs<double> u, FieldPtrs<double> add, GridGeom g,
                                                        double* __restrict__ prof) {
  __shared__ double red[256];
  const int y = blockIdx.x;
  const size_t rw = 2 * (size_t)g.cpitch;
  const double* ux = reinterpret_cast<const double*>(u.p[0]);
  const double* us = reinterpret_cast<const double*>(add.p[0]);
  double s = 0.0;
  for (int w = threadIdx.x; w < g.nx * g.nz; w += blockDim.x) {
    const int x = w / g.nz, k = w - (w / g.nz) * g.nz;
    const size_t i = ((size_t)x * g.ny + y) * rw + k;
    s += ux[i] + (us ? us[i] : 0.0);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) prof[y] = red[0] / (double)(g.nx * g.nz);
}

hipError_t x_velocity_profile(const GridGeom& g, const FieldPtrs<double>& u, const FieldPtrs<double>& add,
                              double* prof, hipStream_t st) {
  hipLaunchKernelGGL(x_profile_kernel, dim3(g.ny), dim3(256), 0, st, u, add, g, prof);
  return hipGetLastError();
}

}  // namespace fcm
