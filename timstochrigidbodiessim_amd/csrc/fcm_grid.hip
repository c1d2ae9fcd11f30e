// fcm_grid.hip -- Gaussian set-up, cell binning, spreading and interpolation.
//
// set_up_gaussians (fcm_fluid_solver.cpp:131-165): per blob and axis the
// anchor cell xc = int(Y/dx) and the ng weights
//   g(i) = exp(-(xg*dx - Y)^2 / (2 a^2/pi)) / (sqrt(2) a),  xg = xc - ng/2 + i + 1,
// stored in bin-sorted order.  Blobs are binned by anchor cell into bins of
// (bx, by, bz) nodes; within a bin they keep ascending blob id.
//
// Spreading (fcm_fluid_solver.cpp:552-600) is a deterministic gather: one
// workgroup owns a tile of nodes, stages every blob whose stencil meets the
// tile in LDS (in bin order), and each thread accumulates a column of nodes
// in registers in that fixed order -- no atomics, bitwise repeatable.
//
// Interpolation (fcm_fluid_solver.cpp:363-424) stages one bin's node box
// (bin + stencil halo) in LDS and reduces each blob's ng^3 products with a
// wavefront reduction; the rigid-body term -K U (gmres_solver.cpp:79) can be
// folded into the same write.
#include "fcm_internal.hpp"

namespace fcm {

static constexpr double FCM_PI = 3.14159265358979323846264338327950288;

__global__ void fill_int_kernel(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ int wrapi(int v, int n) {
  int r = v % n;
  return r < 0 ? r + n : r;
}

__global__ void bin_kernel(const double* __restrict__ Y, int nb, GridGeom g, BlobBins bb, int* bin_of, int* count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int ax = wrapi((int)(Y[3 * b] / g.dx), g.nx);
  const int ay = wrapi((int)(Y[3 * b + 1] / g.dx), g.ny);
  const int az = wrapi((int)(Y[3 * b + 2] / g.dx), g.nz);
  const int bin = ((ax / bb.bx) * bb.nby + ay / bb.by) * bb.nbz + az / bb.bz;
  bin_of[b] = bin;
  atomicAdd(&count[bin], 1);
}

// exclusive scan of count[0..n) into start[0..n], cursor reset; one block
__global__ void __launch_bounds__(1024) scan_kernel(const int* count, int* start, int* cursor, int n) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int lo = t * chunk, hi = min(n, lo + chunk);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += count[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = lo; i < hi; ++i) {
    start[i] = run;
    run += count[i];
    cursor[i] = 0;
  }
  if (t == 1023) start[n] = part[1023];
}

__global__ void scatter_kernel(const int* bin_of, int nb, const int* start, int* cursor, int* sorted) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int bin = bin_of[b];
  sorted[start[bin] + atomicAdd(&cursor[bin], 1)] = b;
}

// restore ascending id order inside each bin (bins are small)
__global__ void bin_sort_kernel(const int* start, int nbins, int* sorted) {
  const int bin = blockIdx.x * blockDim.x + threadIdx.x;
  if (bin >= nbins) return;
  const int lo = start[bin], hi = start[bin + 1];
  for (int i = lo + 1; i < hi; ++i) {
    const int v = sorted[i];
    int j = i - 1;
    while (j >= lo && sorted[j] > v) {
      sorted[j + 1] = sorted[j];
      --j;
    }
    sorted[j + 1] = v;
  }
}

__global__ void record_kernel(const double* __restrict__ Y, const int* __restrict__ sorted, int nb, GridGeom g,
                              int ng, double a, int4* rec, double* w64, float* w32) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const int id = sorted[s];
  const int ngdh = ng / 2;
  const double anorm = 1.0 / (sqrt(2.0) * a);
  const double anorm2 = 2.0 * a * a / FCM_PI;
  const int n[3] = {g.nx, g.ny, g.nz};
  int wa[3];
  for (int d = 0; d < 3; ++d) {
    const double y = Y[3 * id + d];
    const int xc = (int)(y / g.dx);
    wa[d] = wrapi(xc, n[d]);
    for (int i = 0; i < ng; ++i) {
      const int xg = xc - ngdh + (i + 1);
      const double xx = xg * g.dx - y;
      const double w = anorm * exp(-xx * xx / anorm2);
      w64[((size_t)s * 3 + d) * ng + i] = w;
      w32[((size_t)s * 3 + d) * ng + i] = (float)w;
    }
  }
  rec[s] = make_int4(wa[0], wa[1], wa[2], id);
}

hipError_t setup_gaussians(const GridGeom& g, const BlobBins& bb, BlobData& bd, double a, hipStream_t st) {
  const int nb = bd.nb, nbins = bb.nbins();
  hipLaunchKernelGGL(fill_int_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, bd.bin_count, nbins, 0);
  if (nb > 0)
    hipLaunchKernelGGL(bin_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.Y, nb, g, bb, bd.bin_of, bd.bin_count);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, bd.bin_count, bd.bin_start, bd.bin_cursor, nbins);
  if (nb > 0) {
    hipLaunchKernelGGL(scatter_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.bin_of, nb, bd.bin_start,
                       bd.bin_cursor, bd.sorted);
    hipLaunchKernelGGL(bin_sort_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, bd.bin_start, nbins, bd.sorted);
    hipLaunchKernelGGL(record_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.Y, bd.sorted, nb, g, bd.ng, a,
                       bd.rec, bd.w64, bd.w32);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// spreading
// ---------------------------------------------------------------------------
static constexpr int SP_TXMAX = 8;    // tile x extent held in registers
static constexpr int SP_NGMAX = 16;   // largest stencil supported
static constexpr int SP_BATCH = 128;  // candidate blobs staged per round
static constexpr int SP_MAXC = 512;   // candidate bins per tile

__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

template <typename T> __device__ __forceinline__ const T* wsel(const double* w64, const float* w32);
template <> __device__ __forceinline__ const double* wsel<double>(const double* w64, const float*) { return w64; }
template <> __device__ __forceinline__ const float* wsel<float>(const double*, const float* w32) { return w32; }

struct SpreadArgs {
  GridGeom g;
  BlobBins bb;
  int ng;
  const int* bin_start;
  const int4* rec;
  const double* w64;
  const float* w32;
  const double* F;
};

template <typename T>
__global__ void __launch_bounds__(256) spread_kernel(SpreadArgs A, FieldPtrs<T> fp) {
  const GridGeom& g = A.g;
  const BlobBins& bb = A.bb;
  const int ng = A.ng, omin = 1 - ng / 2, omax = ng - ng / 2;
  const int TX = bb.sx * bb.bx, TY = bb.sy * bb.by, TZ = bb.sz * bb.bz;
  const int ntx = g.nx / TX, nty = g.ny / TY, ntz = g.nz / TZ;
  int tile = blockIdx.x;
  const int tz = tile % ntz; tile /= ntz;
  const int ty = tile % nty;
  const int tx = tile / nty;
  const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
  (void)ntx;

  __shared__ int c_start[SP_MAXC], c_off[SP_MAXC + 1];
  __shared__ int s_ncand;
  __shared__ int s_wave_cnt[4];
  __shared__ T r_F[SP_BATCH][3];
  __shared__ T r_gx[SP_BATCH][SP_TXMAX];
  __shared__ T r_gy[SP_BATCH][SP_NGMAX];
  __shared__ T r_gz[SP_BATCH][SP_NGMAX];
  __shared__ int r_ay[SP_BATCH], r_az[SP_BATCH];

  const int t = threadIdx.x;
  if (t == 0) {
    int lo[3], cnt[3];
    const int org[3] = {x0, y0, z0}, ext[3] = {TX, TY, TZ}, bsz[3] = {bb.bx, bb.by, bb.bz},
              nbd[3] = {bb.nbx, bb.nby, bb.nbz};
    for (int d = 0; d < 3; ++d) {
      lo[d] = floor_div(org[d] - omax, bsz[d]);
      const int hi = floor_div(org[d] + ext[d] - 1 - omin, bsz[d]);
      cnt[d] = hi - lo[d] + 1;
      if (cnt[d] >= nbd[d]) { lo[d] = 0; cnt[d] = nbd[d]; }
    }
    int nc = 0, off = 0;
    for (int i = 0; i < cnt[0]; ++i)
      for (int j = 0; j < cnt[1]; ++j)
        for (int k = 0; k < cnt[2]; ++k) {
          if (nc >= SP_MAXC) continue;
          const int bin = (wrapi(lo[0] + i, bb.nbx) * bb.nby + wrapi(lo[1] + j, bb.nby)) * bb.nbz + wrapi(lo[2] + k, bb.nbz);
          const int s0 = A.bin_start[bin], s1 = A.bin_start[bin + 1];
          c_start[nc] = s0;
          c_off[nc] = off;
          off += s1 - s0;
          ++nc;
        }
    c_off[nc] = off;
    s_ncand = nc;
  }
  __syncthreads();
  const int ncand = s_ncand;
  const int ntot = c_off[ncand];

  const int ly = t / TZ, lz = t % TZ;
  const bool active = ly < TY;
  const int yn = y0 + (active ? ly : 0), zn = z0 + lz;
  T acc[SP_TXMAX][3];
#pragma unroll
  for (int xi = 0; xi < SP_TXMAX; ++xi) acc[xi][0] = acc[xi][1] = acc[xi][2] = T(0);

  const T* W = wsel<T>(A.w64, A.w32);
  const int lane = t & 63, wv = t >> 6;

  for (int base = 0; base < ntot; base += SP_BATCH) {
    const int c = base + t;
    bool rel = false;
    int s = 0;
    int4 r = make_int4(0, 0, 0, 0);
    if (t < SP_BATCH && c < ntot) {
      int lo = 0, hi = ncand - 1;  // last segment with c_off <= c
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c_off[mid] <= c) lo = mid; else hi = mid - 1;
      }
      s = c_start[lo] + (c - c_off[lo]);
      r = A.rec[s];
      bool rx = false, ry = false, rz = false;
      for (int i = 0; i < TX; ++i) rx |= cyc(x0 + i - r.x, g.nx, omin) <= omax;
      for (int i = 0; i < TY; ++i) ry |= cyc(y0 + i - r.y, g.ny, omin) <= omax;
      for (int i = 0; i < TZ; ++i) rz |= cyc(z0 + i - r.z, g.nz, omin) <= omax;
      rel = rx && ry && rz;
    }
    // order-preserving compaction of the relevant candidates
    const unsigned long long m = __ballot(rel);
    const int below = __popcll(m & ((1ULL << lane) - 1ULL));
    if (lane == 0) s_wave_cnt[wv] = __popcll(m);
    __syncthreads();
    int woff = 0, nrel = 0;
    for (int i = 0; i < 4; ++i) {
      if (i < wv) woff += s_wave_cnt[i];
      nrel += s_wave_cnt[i];
    }
    if (rel) {
      const int p = woff + below;
      const T* ws = W + (size_t)s * 3 * ng;
      r_F[p][0] = T(A.F[3 * r.w]);
      r_F[p][1] = T(A.F[3 * r.w + 1]);
      r_F[p][2] = T(A.F[3 * r.w + 2]);
      for (int i = 0; i < SP_TXMAX; ++i) {
        T v = T(0);
        if (i < TX) {
          const int o = cyc(x0 + i - r.x, g.nx, omin);
          if (o <= omax) v = ws[o - omin];
        }
        r_gx[p][i] = v;
      }
      for (int i = 0; i < ng; ++i) {
        r_gy[p][i] = ws[ng + i];
        r_gz[p][i] = ws[2 * ng + i];
      }
      r_ay[p] = r.y;
      r_az[p] = r.z;
    }
    __syncthreads();
    if (active) {
      for (int p = 0; p < nrel; ++p) {
        const int oy = cyc(yn - r_ay[p], g.ny, omin);
        const int oz = cyc(zn - r_az[p], g.nz, omin);
        if (oy <= omax && oz <= omax) {
          const T w = r_gy[p][oy - omin] * r_gz[p][oz - omin];
          const T f0 = r_F[p][0] * w, f1 = r_F[p][1] * w, f2 = r_F[p][2] * w;
#pragma unroll
          for (int xi = 0; xi < SP_TXMAX; ++xi) {
            const T gx = r_gx[p][xi];
            acc[xi][0] += f0 * gx;
            acc[xi][1] += f1 * gx;
            acc[xi][2] += f2 * gx;
          }
        }
      }
    }
    __syncthreads();
  }
  if (active) {
    const size_t rstride = 2 * (size_t)g.cpitch;  // reals per row
#pragma unroll
    for (int xi = 0; xi < SP_TXMAX; ++xi) {
      if (xi < TX) {
        const size_t idx = ((size_t)(x0 + xi) * g.ny + yn) * rstride + zn;
        reinterpret_cast<T*>(fp.p[0])[idx] = acc[xi][0];
        reinterpret_cast<T*>(fp.p[1])[idx] = acc[xi][1];
        reinterpret_cast<T*>(fp.p[2])[idx] = acc[xi][2];
      }
    }
  }
}

template <typename T>
hipError_t spread(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                  hipStream_t st) {
  if (bd.ng > SP_NGMAX || bb.sx * bb.bx > SP_TXMAX || bb.sy * bb.by * bb.sz * bb.bz > 256) return hipErrorInvalidValue;
  SpreadArgs A{g, bb, bd.ng, bd.bin_start, bd.rec, bd.w64, bd.w32, F};
  const int ntiles = (g.nx / (bb.sx * bb.bx)) * (g.ny / (bb.sy * bb.by)) * (g.nz / (bb.sz * bb.bz));
  hipLaunchKernelGGL((spread_kernel<T>), dim3(ntiles), dim3(256), 0, st, A, fp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// interpolation
// ---------------------------------------------------------------------------
struct InterpArgs {
  GridGeom g;
  BlobBins bb;
  int ng;
  const int* bin_start;
  const int4* rec;
  const double* w64;
  const float* w32;
  double* V;
  KTerm kt;
  double dx3;
};

__device__ __forceinline__ void write_velocity(const InterpArgs& A, int id, double v0, double v1, double v2) {
  v0 *= A.dx3; v1 *= A.dx3; v2 *= A.dx3;
  if (A.kt.zV) {
    const int body = id / A.kt.nbpb;
    const double* u = A.kt.zV + 6 * (size_t)body;
    const double* d = A.kt.dpos + 3 * (size_t)id;
    // U + Omega x d   (collection_of_rigid_bodies.cpp:723)
    v0 -= u[0] + (u[4] * d[2] - u[5] * d[1]);
    v1 -= u[1] + (u[5] * d[0] - u[3] * d[2]);
    v2 -= u[2] + (u[3] * d[1] - u[4] * d[0]);
  }
  A.V[3 * (size_t)id] = v0;
  A.V[3 * (size_t)id + 1] = v1;
  A.V[3 * (size_t)id + 2] = v2;
}

// one bin per workgroup; node box of (bin + ng - 1) per axis staged in LDS
template <typename T>
__global__ void __launch_bounds__(256) interp_box_kernel(InterpArgs A, FieldPtrs<T> fp) {
  extern __shared__ unsigned char smem_raw[];
  T* box = reinterpret_cast<T*>(smem_raw);
  const GridGeom& g = A.g;
  const BlobBins& bb = A.bb;
  const int ng = A.ng, omin = 1 - ng / 2;
  int bin = blockIdx.x;
  const int s0 = A.bin_start[bin], s1 = A.bin_start[bin + 1];
  if (s1 == s0) return;
  const int bzi = bin % bb.nbz; bin /= bb.nbz;
  const int byi = bin % bb.nby;
  const int bxi = bin / bb.nby;
  const int HX = bb.bx + ng - 1, HY = bb.by + ng - 1, HZ = bb.bz + ng - 1;
  const int X0 = bxi * bb.bx + omin, Y0 = byi * bb.by + omin, Z0 = bzi * bb.bz + omin;
  const int nbox = HX * HY * HZ;
  const size_t rstride = 2 * (size_t)g.cpitch;
  const T* f0 = reinterpret_cast<const T*>(fp.p[0]);
  const T* f1 = reinterpret_cast<const T*>(fp.p[1]);
  const T* f2 = reinterpret_cast<const T*>(fp.p[2]);
  for (int i = threadIdx.x; i < nbox; i += blockDim.x) {
    const int hz = i % HZ, hy = (i / HZ) % HY, hx = i / (HZ * HY);
    const size_t idx = ((size_t)wrapi(X0 + hx, g.nx) * g.ny + wrapi(Y0 + hy, g.ny)) * rstride + wrapi(Z0 + hz, g.nz);
    box[i] = f0[idx];
    box[nbox + i] = f1[idx];
    box[2 * nbox + i] = f2[idx];
  }
  __syncthreads();
  const T* W = wsel<T>(A.w64, A.w32);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int s = s0 + wv; s < s1; s += nw) {
    const int4 r = A.rec[s];
    const T* ws = W + (size_t)s * 3 * ng;
    const int lx = r.x - bxi * bb.bx, ly = r.y - byi * bb.by, lz = r.z - bzi * bb.bz;
    T a0 = T(0), a1 = T(0), a2 = T(0);
    for (int p = lane; p < ng * ng; p += 64) {
      const int jy = p / ng, jz = p % ng;
      const int hyz = (ly + jy) * HZ + (lz + jz);
      T s0v = T(0), s1v = T(0), s2v = T(0);
      for (int jx = 0; jx < ng; ++jx) {
        const int i = (lx + jx) * HY * HZ + hyz;
        const T gx = ws[jx];
        s0v += gx * box[i];
        s1v += gx * box[nbox + i];
        s2v += gx * box[2 * nbox + i];
      }
      const T wyz = ws[ng + jy] * ws[2 * ng + jz];
      a0 += wyz * s0v;
      a1 += wyz * s1v;
      a2 += wyz * s2v;
    }
    const double v0 = warp_sum((double)a0), v1 = warp_sum((double)a1), v2 = warp_sum((double)a2);
    if (lane == 0) write_velocity(A, r.w, v0, v1, v2);
  }
}

// fallback for stencils whose node box does not fit in LDS: one wavefront per
// blob gathers straight from the grid
template <typename T>
__global__ void __launch_bounds__(256) interp_gather_kernel(InterpArgs A, FieldPtrs<T> fp, int nb) {
  const GridGeom& g = A.g;
  const int ng = A.ng, omin = 1 - ng / 2;
  const int s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= nb) return;
  const int4 r = A.rec[s];
  const T* W = wsel<T>(A.w64, A.w32);
  const T* ws = W + (size_t)s * 3 * ng;
  const size_t rstride = 2 * (size_t)g.cpitch;
  const T* f0 = reinterpret_cast<const T*>(fp.p[0]);
  const T* f1 = reinterpret_cast<const T*>(fp.p[1]);
  const T* f2 = reinterpret_cast<const T*>(fp.p[2]);
  T a0 = T(0), a1 = T(0), a2 = T(0);
  for (int p = lane; p < ng * ng; p += 64) {
    const int jy = p / ng, jz = p % ng;
    const int y = wrapi(r.y + omin + jy, g.ny), z = wrapi(r.z + omin + jz, g.nz);
    T s0v = T(0), s1v = T(0), s2v = T(0);
    for (int jx = 0; jx < ng; ++jx) {
      const size_t idx = ((size_t)wrapi(r.x + omin + jx, g.nx) * g.ny + y) * rstride + z;
      const T gx = ws[jx];
      s0v += gx * f0[idx];
      s1v += gx * f1[idx];
      s2v += gx * f2[idx];
    }
    const T wyz = ws[ng + jy] * ws[2 * ng + jz];
    a0 += wyz * s0v;
    a1 += wyz * s1v;
    a2 += wyz * s2v;
  }
  const double v0 = warp_sum((double)a0), v1 = warp_sum((double)a1), v2 = warp_sum((double)a2);
  if (lane == 0) write_velocity(A, r.w, v0, v1, v2);
}

template <typename T>
hipError_t interp(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const FieldPtrs<T>& fp, double* V,
                  const KTerm& kt, hipStream_t st) {
  InterpArgs A{g, bb, bd.ng, bd.bin_start, bd.rec, bd.w64, bd.w32, V, kt, g.dx * g.dx * g.dx};
  if (bd.nb == 0) return hipSuccess;
  const size_t lds = 3 * sizeof(T) * (size_t)(bb.bx + bd.ng - 1) * (bb.by + bd.ng - 1) * (bb.bz + bd.ng - 1);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL((interp_box_kernel<T>), dim3(bb.nbins()), dim3(256), lds, st, A, fp);
  } else {
    hipLaunchKernelGGL((interp_gather_kernel<T>), dim3((bd.nb + 3) / 4), dim3(256), 0, st, A, fp, bd.nb);
  }
  return hipGetLastError();
}

template hipError_t spread<double>(const GridGeom&, const BlobBins&, const BlobData&, const double*, const FieldPtrs<double>&, hipStream_t);
template hipError_t spread<float>(const GridGeom&, const BlobBins&, const BlobData&, const double*, const FieldPtrs<float>&, hipStream_t);
template hipError_t interp<double>(const GridGeom&, const BlobBins&, const BlobData&, const FieldPtrs<double>&, double*, const KTerm&, hipStream_t);
template hipError_t interp<float>(const GridGeom&, const BlobBins&, const BlobData&, const FieldPtrs<float>&, double*, const KTerm&, hipStream_t);

}  // namespace fcm
