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
                              int ng, double a, int4* rec, double* w64, float* w32, double* iw64, float* iw32,
                              float4* woff) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const int id = sorted[s];
  const int ngdh = ng / 2;
  const double anorm = 1.0 / (sqrt(2.0) * a);
  const double anorm2 = 2.0 * a * a / FCM_PI;
  const int n[3] = {g.nx, g.ny, g.nz};
  int wa[3];
  float off[3];
  for (int d = 0; d < 3; ++d) {
    const double y = Y[3 * id + d];
    const int xc = (int)(y / g.dx);
    wa[d] = wrapi(xc, n[d]);
    off[d] = (float)((xc - ngdh + 1) * g.dx - y);
    for (int i = 0; i < ng; ++i) {
      const int xg = xc - ngdh + (i + 1);
      const double xx = xg * g.dx - y;
      const double w = anorm * exp(-xx * xx / anorm2);
      w64[((size_t)s * 3 + d) * ng + i] = w;
      w32[((size_t)s * 3 + d) * ng + i] = (float)w;
      if (g.channel) {
        // interpolation drops stencil rows at and above the upper wall
        // (fcm_fluid_solver.cpp:387-413): the wrapped row index decides
        const double wi = (d == 1 && wrapi(xg, g.ny) >= g.ny / 2) ? 0.0 : w;
        iw64[((size_t)s * 3 + d) * ng + i] = wi;
        iw32[((size_t)s * 3 + d) * ng + i] = (float)wi;
      }
    }
  }
  rec[s] = make_int4(wa[0], wa[1], wa[2], id);
  if (woff) woff[s] = make_float4(off[0], off[1], off[2], 0.f);
}

hipError_t build_tile_lists(const GridGeom& g, const BlobBins& bb, BlobData& bd, hipStream_t st);

hipError_t sort_by_key(const int* key_of, int n, int nkeys, const int* count, int* start, int* cursor, int* sorted,
                       hipStream_t st) {
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, count, start, cursor, nkeys);
  if (n > 0) {
    hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, key_of, n, start, cursor, sorted);
    hipLaunchKernelGGL(bin_sort_kernel, dim3((nkeys + 255) / 256), dim3(256), 0, st, start, nkeys, sorted);
  }
  return hipGetLastError();
}

hipError_t exclusive_scan(const int* count, int* start, int* cursor, int n, hipStream_t st) {
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, count, start, cursor, n);
  return hipGetLastError();
}

hipError_t fill_ints(int* p, int n, int v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t setup_gaussians(const GridGeom& g, const BlobBins& bb, BlobData& bd, double a, hipStream_t st) {
  const int nb = bd.nb, nbins = bb.nbins();
  // anorm exp(-xx^2 / anorm2) = wa exp2(wc xx^2), anorm = 1 / (sqrt(2) a), anorm2 = 2 a^2 / pi
  bd.wa32 = (float)(1.0 / (sqrt(2.0) * a));
  bd.wc32 = (float)(-FCM_PI / (2.0 * a * a) / log(2.0));
  hipLaunchKernelGGL(fill_int_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, bd.bin_count, nbins, 0);
  if (nb > 0)
    hipLaunchKernelGGL(bin_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.Y, nb, g, bb, bd.bin_of, bd.bin_count);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, bd.bin_count, bd.bin_start, bd.bin_cursor, nbins);
  if (nb > 0) {
    hipLaunchKernelGGL(scatter_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.bin_of, nb, bd.bin_start,
                       bd.bin_cursor, bd.sorted);
    hipLaunchKernelGGL(bin_sort_kernel, dim3((nbins + 255) / 256), dim3(256), 0, st, bd.bin_start, nbins, bd.sorted);
    hipLaunchKernelGGL(record_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, bd.Y, bd.sorted, nb, g, bd.ng, a,
                       bd.rec, bd.w64, bd.w32, bd.iw64, bd.iw32, bd.woff);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = build_tile_lists(g, bb, bd, st);
  if (e != hipSuccess) return e;
  return build_col_lists(g, bb, bd, st);
}

// ---------------------------------------------------------------------------
// slip-channel images of the force fields (fcm_fluid_solver.cpp:602-630): fy is
// zero on the wall rows y = 0 and y = ny/2, and each row y > ny/2 becomes the
// mirror of row ny - y, fx and fz even, fy odd.  Whole rows are copied, so the
// rule serves real rows and z-transformed rows alike (it acts on y only and
// commutes with the z transform).  One thread per complex word, grid-stride.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) channel_image_kernel(FieldPtrs<T> fp, GridGeom g) {
  if (fp.gate && *fp.gate) return;
  const int half = g.ny / 2, cp = g.cpitch;
  const size_t words = (size_t)g.nx * half * cp;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(w % cp);
    const size_t r = w / cp;
    const int t = (int)(r % half);
    const size_t plane = (r / half) * g.ny;  // first row of this x plane
    if (t == 0) {
      fp.p[1][plane * cp + k] = mk<T>(T(0), T(0));
      fp.p[1][(plane + half) * cp + k] = mk<T>(T(0), T(0));
    } else {
      const size_t dst = (plane + half + t) * cp + k, src = (plane + half - t) * cp + k;
      fp.p[0][dst] = fp.p[0][src];
      const c2<T> v = fp.p[1][src];
      fp.p[1][dst] = mk<T>(-v.x, -v.y);
      fp.p[2][dst] = fp.p[2][src];
    }
  }
}

template <typename T>
hipError_t channel_images(const GridGeom& g, const FieldPtrs<T>& fp, hipStream_t st) {
  if (!g.channel) return hipSuccess;
  const size_t words = (size_t)g.nx * (g.ny / 2) * g.cpitch;
  const int blocks = (int)std::min<size_t>((words + 255) / 256, 8192);
  hipLaunchKernelGGL((channel_image_kernel<T>), dim3(blocks), dim3(256), 0, st, fp, g);
  return hipGetLastError();
}
template hipError_t channel_images<double>(const GridGeom&, const FieldPtrs<double>&, hipStream_t);
template hipError_t channel_images<float>(const GridGeom&, const FieldPtrs<float>&, hipStream_t);

// ---------------------------------------------------------------------------
// spreading
// ---------------------------------------------------------------------------
static constexpr int SP_TXMAX = 8;    // tile x extent held in registers
static constexpr int SP_NGMAX = 20;   // largest stencil supported (NUM_FCM_GAUSSIAN_POINTS, config.hpp:69)
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
  if (fp.gate && *fp.gate) return;
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

// ---------------------------------------------------------------------------
// spreading, main path (grid sides multiples of 8, nx >= 16, ng <= 8).
// Tile = 8 (x, registers) x 8 (y) x 8*NW (z) nodes; wavefront w owns the
// 8 x 8 (y, z) patch z in [8w, 8w+8).  The blobs whose stencil meets a tile
// are listed once per set of positions (tile_list); a spread stages them in
// LDS (force, raw weights, stencil start per axis relative to the tile and a
// mask of the wavefronts they touch).  A wavefront skips blobs that miss its
// patch, and the x loop is specialised per stencil start so that only
// stencil nodes are touched.  The tile leaves through LDS as whole z rows.
// Blocks are mapped to tiles so that each XCD works on a contiguous range.
// ---------------------------------------------------------------------------
static constexpr int SF_BATCH = 128;

__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
  const int per = (ntiles + 7) / 8;
  return (b % 8) * per + b / 8;
}

__device__ __forceinline__ int wrap_half(int v, int n) {
  if (v < -(n / 2)) v += n; else if (v >= n - n / 2) v -= n;
  return v;
}

// relevance of a blob (wrapped anchor r) to the tile at (x0, y0, z0)
__device__ __forceinline__ bool tile_relevant(const int4& r, int x0, int y0, int z0, int TZ, const GridGeom& g,
                                              int ng, int omin) {
  const int sx = wrap_half(r.x + omin - x0, g.nx);
  const int sy = wrap_half(r.y + omin - y0, g.ny);
  const int sz = wrap_half(r.z + omin - z0, g.nz);
  const bool rx = (sx <= 7) && (sx + ng - 1 >= 0);
  const bool ry = g.ny <= 8 || ((sy <= 7) && (sy + ng - 1 >= 0));
  const bool rz = g.nz <= TZ || ((sz <= TZ - 1) && (sz + ng - 1 >= 0));
  return rx && ry && rz;
}

// pass 0: count the relevant blobs of every tile; pass 1: write them, in bin
// order, from tile_start[tile]
__global__ void __launch_bounds__(256) tile_list_kernel(GridGeom g, BlobBins bb, int ng, const int* bin_start,
                                                        const int4* rec, int NW, int pass, int* tile_count,
                                                        const int* tile_start, int* tile_list) {
  const int omin = 1 - ng / 2, omax = ng - ng / 2;
  const int TZ = 8 * NW;
  const int nty = g.ny / 8, ntz = g.nz / TZ;
  int tile = blockIdx.x;
  const int tz = tile % ntz; tile /= ntz;
  const int ty = tile % nty;
  const int tx = tile / nty;
  const int x0 = tx * 8, y0 = ty * 8, z0 = tz * TZ;
  __shared__ int c_start[SP_MAXC], c_off[SP_MAXC + 1];
  __shared__ int s_ncand;
  __shared__ int s_wave_cnt[4];
  const int t = threadIdx.x;
  if (t == 0) {
    int lo[3], cnt[3];
    const int org[3] = {x0, y0, z0}, ext[3] = {8, 8, TZ}, bsz[3] = {bb.bx, bb.by, bb.bz},
              nbd[3] = {bb.nbx, bb.nby, bb.nbz};
    for (int d = 0; d < 3; ++d) {
      lo[d] = floor_div(org[d] - omax, bsz[d]);
      const int hi = floor_div(org[d] + ext[d] - 1 - omin, bsz[d]);
      cnt[d] = hi - lo[d] + 1;
      if (cnt[d] >= nbd[d]) { lo[d] = 0; cnt[d] = nbd[d]; }
    }
    int nc = 0, off = 0;
    for (int i = 0; i < cnt[0]; ++i)
      for (int j = 0; j < cnt[1]; ++j) {
        const int bxy = wrapi(lo[0] + i, bb.nbx) * bb.nby + wrapi(lo[1] + j, bb.nby);
        int k = 0;
        while (k < cnt[2] && nc < SP_MAXC) {
          const int bz0 = wrapi(lo[2] + k, bb.nbz);
          int run = 1;
          while (k + run < cnt[2] && bz0 + run < bb.nbz) ++run;
          const int b0 = bxy * bb.nbz + bz0;
          c_start[nc] = bin_start[b0];
          c_off[nc] = off;
          off += bin_start[b0 + run] - bin_start[b0];
          ++nc;
          k += run;
        }
      }
    c_off[nc] = off;
    s_ncand = nc;
  }
  __syncthreads();
  const int ncand = s_ncand, ntot = c_off[ncand];
  const int lane = t & 63, wv = t >> 6;
  int written = 0;
  for (int base = 0; base < ntot; base += 256) {
    const int c = base + t;
    bool rel = false;
    int s = 0;
    if (c < ntot) {
      int lo = 0, hi = ncand - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (c_off[mid] <= c) lo = mid; else hi = mid - 1;
      }
      s = c_start[lo] + (c - c_off[lo]);
      rel = tile_relevant(rec[s], x0, y0, z0, TZ, g, ng, omin);
    }
    const unsigned long long m = __ballot(rel);
    const int below = __popcll(m & ((1ULL << lane) - 1ULL));
    if (lane == 0) s_wave_cnt[wv] = __popcll(m);
    __syncthreads();
    int woff = 0, nrel = 0;
    for (int i = 0; i < 4; ++i) {
      if (i < wv) woff += s_wave_cnt[i];
      nrel += s_wave_cnt[i];
    }
    if (pass == 1 && rel) tile_list[tile_start[blockIdx.x] + written + woff + below] = s;
    written += nrel;
    __syncthreads();
  }
  if (pass == 0 && t == 0) tile_count[blockIdx.x] = written;
}

template <typename T, int S>
__device__ __forceinline__ void sf_accum(T (&acc)[8][3], T f0, T f1, T f2, const T* gxw) {
#pragma unroll
  for (int xi = 0; xi < 8; ++xi) {
    if (xi - S >= 0 && xi - S < 8) {
      const T gx = gxw[xi - S];
      acc[xi][0] += f0 * gx;
      acc[xi][1] += f1 * gx;
      acc[xi][2] += f2 * gx;
    }
  }
}

struct SpreadListArgs {
  GridGeom g;
  int ng, NW;
  const int* tile_start;
  const int* tile_list;
  const int4* rec;
  const double* w64;
  const float* w32;
  const double* F;
};

template <typename T>
__global__ void __launch_bounds__(256) spread_fast_kernel(SpreadListArgs A, FieldPtrs<T> fp) {
  if (fp.gate && *fp.gate) return;
  const GridGeom& g = A.g;
  const int ng = A.ng, omin = 1 - ng / 2, NW = A.NW;
  const int TZ = 8 * NW;
  const int nty = g.ny / 8, ntz = g.nz / TZ;
  const int ntiles = (g.nx / 8) * nty * ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int tt = tile;
  const int tz = tt % ntz; tt /= ntz;
  const int ty = tt % nty;
  const int tx = tt / nty;
  const int x0 = tx * 8, y0 = ty * 8, z0 = tz * TZ;

  constexpr int REC_WORDS = SF_BATCH * (3 + 8 + 8 + 8);
  constexpr int STAGE_WORDS = 3 * 8 * 8 * 32;
  __shared__ T buf[REC_WORDS > STAGE_WORDS ? REC_WORDS : STAGE_WORDS];
  __shared__ int r_sx[SF_BATCH], r_sy[SF_BATCH], r_sz[SF_BATCH], r_mask[SF_BATCH];
  T (*r_F)[3] = reinterpret_cast<T(*)[3]>(buf);
  T (*r_gx)[8] = reinterpret_cast<T(*)[8]>(buf + SF_BATCH * 3);
  T (*r_gy)[8] = reinterpret_cast<T(*)[8]>(buf + SF_BATCH * 11);
  T (*r_gz)[8] = reinterpret_cast<T(*)[8]>(buf + SF_BATCH * 19);

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int lz = lane & 7, ly = lane >> 3;
  const int zl = 8 * wv + lz;
  const int L0 = A.tile_start[tile], L1 = A.tile_start[tile + 1];
  T acc[8][3];
#pragma unroll
  for (int xi = 0; xi < 8; ++xi) acc[xi][0] = acc[xi][1] = acc[xi][2] = T(0);
  const T* W = wsel<T>(A.w64, A.w32);

  for (int base = L0; base < L1; base += SF_BATCH) {
    const int cnt = min(SF_BATCH, L1 - base);
    if (t < cnt) {
      const int s = A.tile_list[base + t];
      const int4 r = A.rec[s];
      const T* ws = W + (size_t)s * 3 * ng;
      T wx[8], wy[8], wz[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        wx[i] = i < ng ? ws[i] : T(0);
        wy[i] = i < ng ? ws[ng + i] : T(0);
        wz[i] = i < ng ? ws[2 * ng + i] : T(0);
      }
      const double f0 = A.F[3 * (size_t)r.w], f1 = A.F[3 * (size_t)r.w + 1], f2 = A.F[3 * (size_t)r.w + 2];
      const int sx = wrap_half(r.x + omin - x0, g.nx);
      const int sy = wrap_half(r.y + omin - y0, g.ny);
      const int sz = wrap_half(r.z + omin - z0, g.nz);
      int mask = 0;
      for (int i = 0; i < TZ; ++i) {
        int oz = i - sz;
        if (oz < 0) oz += g.nz; else if (oz >= g.nz) oz -= g.nz;
        if (oz < ng) mask |= 1 << (i >> 3);
      }
      r_F[t][0] = T(f0);
      r_F[t][1] = T(f1);
      r_F[t][2] = T(f2);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        r_gx[t][i] = wx[i];
        r_gy[t][i] = wy[i];
        r_gz[t][i] = wz[i];
      }
      r_sx[t] = sx;
      r_sy[t] = sy;
      r_sz[t] = sz;
      r_mask[t] = mask;
    }
    __syncthreads();
    if (wv < NW) {
      for (int p = 0; p < cnt; ++p) {
        if (!((r_mask[p] >> wv) & 1)) continue;
        int oy = ly - r_sy[p];
        if (oy < 0) oy += g.ny; else if (oy >= g.ny) oy -= g.ny;
        int oz = zl - r_sz[p];
        if (oz < 0) oz += g.nz; else if (oz >= g.nz) oz -= g.nz;
        if (oy < ng && oz < ng) {
          const T wgt = r_gy[p][oy] * r_gz[p][oz];
          const T f0 = r_F[p][0] * wgt, f1 = r_F[p][1] * wgt, f2 = r_F[p][2] * wgt;
          const T* gxw = r_gx[p];
          switch (r_sx[p]) {
            case -7: sf_accum<T, -7>(acc, f0, f1, f2, gxw); break;
            case -6: sf_accum<T, -6>(acc, f0, f1, f2, gxw); break;
            case -5: sf_accum<T, -5>(acc, f0, f1, f2, gxw); break;
            case -4: sf_accum<T, -4>(acc, f0, f1, f2, gxw); break;
            case -3: sf_accum<T, -3>(acc, f0, f1, f2, gxw); break;
            case -2: sf_accum<T, -2>(acc, f0, f1, f2, gxw); break;
            case -1: sf_accum<T, -1>(acc, f0, f1, f2, gxw); break;
            case 0: sf_accum<T, 0>(acc, f0, f1, f2, gxw); break;
            case 1: sf_accum<T, 1>(acc, f0, f1, f2, gxw); break;
            case 2: sf_accum<T, 2>(acc, f0, f1, f2, gxw); break;
            case 3: sf_accum<T, 3>(acc, f0, f1, f2, gxw); break;
            case 4: sf_accum<T, 4>(acc, f0, f1, f2, gxw); break;
            case 5: sf_accum<T, 5>(acc, f0, f1, f2, gxw); break;
            case 6: sf_accum<T, 6>(acc, f0, f1, f2, gxw); break;
            default: sf_accum<T, 7>(acc, f0, f1, f2, gxw); break;
          }
        }
      }
    }
    __syncthreads();
  }
  // stage [comp][x][y][z] and write whole z rows
  T* stage = buf;
  if (wv < NW) {
#pragma unroll
    for (int xi = 0; xi < 8; ++xi)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) stage[((cc * 8 + xi) * 8 + ly) * TZ + zl] = acc[xi][cc];
  }
  __syncthreads();
  const size_t rstride = 2 * (size_t)g.cpitch;
  const int total = 3 * 8 * 8 * TZ;
  for (int e = t; e < total; e += 256) {
    const int zi = e % TZ;
    const int row = e / TZ;  // (cc * 8 + xi) * 8 + yi
    const int yi = row & 7, xi = (row >> 3) & 7, cc = row >> 6;
    const size_t idx = ((size_t)(x0 + xi) * g.ny + (y0 + yi)) * rstride + z0 + zi;
    reinterpret_cast<T*>(fp.p[cc])[idx] = stage[e];
  }
}

// Branch-free variant of the same tiling: the staging thread of a blob shifts
// its weights into tile coordinates -- F_c * g_x at each of the 8 tile x
// nodes (24 values), g_y at the 8 tile rows, g_z at the TZ tile columns, zero
// outside the stencil -- so that every lane of a wavefront whose patch the
// blob touches does the same 24 fused multiply-adds, reading the x/force
// table as LDS broadcasts and only its own y and z weights.
static constexpr int SD_BATCH = 64;

template <typename T>
__global__ void __launch_bounds__(256) spread_dense_kernel(SpreadListArgs A, FieldPtrs<T> fp) {
  if (fp.gate && *fp.gate) return;
  const GridGeom& g = A.g;
  const int ng = A.ng, omin = 1 - ng / 2, NW = A.NW;
  const int TZ = 8 * NW;
  const int nty = g.ny / 8, ntz = g.nz / TZ;
  const int ntiles = (g.nx / 8) * nty * ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int tt = tile;
  const int tz = tt % ntz; tt /= ntz;
  const int ty = tt % nty;
  const int tx = tt / nty;
  const int x0 = tx * 8, y0 = ty * 8, z0 = tz * TZ;

  constexpr int TAB = SD_BATCH * (24 + 8 + 32);
  constexpr int STAGE = 3 * 8 * 8 * 32;
  __shared__ T buf[TAB > STAGE ? TAB : STAGE];
  __shared__ int s_mask[SD_BATCH];
  T (*s_fg)[24] = reinterpret_cast<T(*)[24]>(buf);
  T (*s_gy)[8] = reinterpret_cast<T(*)[8]>(buf + SD_BATCH * 24);
  T (*s_gz)[32] = reinterpret_cast<T(*)[32]>(buf + SD_BATCH * 32);

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int lz = lane & 7, ly = lane >> 3;
  const int zl = 8 * wv + lz;
  const int L0 = A.tile_start[tile], L1 = A.tile_start[tile + 1];
  T acc[8][3];
#pragma unroll
  for (int xi = 0; xi < 8; ++xi) acc[xi][0] = acc[xi][1] = acc[xi][2] = T(0);
  const T* W = wsel<T>(A.w64, A.w32);

  for (int base = L0; base < L1; base += SD_BATCH) {
    const int cnt = min(SD_BATCH, L1 - base);
    if (t < cnt) {
      const int s = A.tile_list[base + t];
      const int4 r = A.rec[s];
      const T* ws = W + (size_t)s * 3 * ng;
      const T f[3] = {T(A.F[3 * (size_t)r.w]), T(A.F[3 * (size_t)r.w + 1]), T(A.F[3 * (size_t)r.w + 2])};
      const int sx = wrap_half(r.x + omin - x0, g.nx);
      const int sy = wrap_half(r.y + omin - y0, g.ny);
      const int sz = wrap_half(r.z + omin - z0, g.nz);
#pragma unroll
      for (int xi = 0; xi < 8; ++xi) {
        int o = xi - sx;
        if (o < 0) o += g.nx; else if (o >= g.nx) o -= g.nx;
        const T gx = o < ng ? ws[o] : T(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) s_fg[t][3 * xi + c] = f[c] * gx;
      }
#pragma unroll
      for (int yi = 0; yi < 8; ++yi) {
        int o = yi - sy;
        if (o < 0) o += g.ny; else if (o >= g.ny) o -= g.ny;
        s_gy[t][yi] = o < ng ? ws[ng + o] : T(0);
      }
      int mask = 0;
      for (int zi = 0; zi < TZ; ++zi) {
        int o = zi - sz;
        if (o < 0) o += g.nz; else if (o >= g.nz) o -= g.nz;
        const bool in = o < ng;
        s_gz[t][zi] = in ? ws[2 * ng + o] : T(0);
        if (in) mask |= 1 << (zi >> 3);
      }
      s_mask[t] = mask;
    }
    __syncthreads();
    if (wv < NW) {
      for (int p = 0; p < cnt; ++p) {
        if (!((s_mask[p] >> wv) & 1)) continue;
        const T w = s_gy[p][ly] * s_gz[p][zl];
        const T* fg = s_fg[p];
#pragma unroll
        for (int xi = 0; xi < 8; ++xi)
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[xi][c] += w * fg[3 * xi + c];
      }
    }
    __syncthreads();
  }
  T* stage = buf;
  if (wv < NW) {
#pragma unroll
    for (int xi = 0; xi < 8; ++xi)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) stage[((cc * 8 + xi) * 8 + ly) * TZ + zl] = acc[xi][cc];
  }
  __syncthreads();
  const size_t rstride = 2 * (size_t)g.cpitch;
  const int total = 3 * 8 * 8 * TZ;
  for (int e = t; e < total; e += 256) {
    const int zi = e % TZ;
    const int row = e / TZ;
    const int yi = row & 7, xi = (row >> 3) & 7, cc = row >> 6;
    const size_t idx = ((size_t)(x0 + xi) * g.ny + (y0 + yi)) * rstride + z0 + zi;
    reinterpret_cast<T*>(fp.p[cc])[idx] = stage[e];
  }
}

// Variant 2: as variant 1, with the staging spread over the whole workgroup
// (one blob per thread, SB blobs per batch: 256 in FP32, 128 in FP64), the z
// weights kept raw (a lane picks its own from the blob's stencil start), and a
// wavefront finding the blobs that touch its patch with one ballot per 64
// blobs and walking the set bits in a scalar loop.  Per blob a lane reads its
// y and z weight and three broadcast 16-byte words (8 shifted x weights and
// the force), forms w F_c and does 24 fused multiply-adds.
template <typename T> struct SpreadBatch { static constexpr int value = sizeof(T) == 4 ? 256 : 128; };

template <typename T>
__global__ void __launch_bounds__(256) spread_bits_kernel(SpreadListArgs A, FieldPtrs<T> fp) {
  if (fp.gate && *fp.gate) return;
  const GridGeom& g = A.g;
  const int ng = A.ng, omin = 1 - ng / 2, NW = A.NW;
  const int TZ = 8 * NW;
  const int nty = g.ny / 8, ntz = g.nz / TZ;
  const int ntiles = (g.nx / 8) * nty * ntz;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  int tt = tile;
  const int tz = tt % ntz; tt /= ntz;
  const int ty = tt % nty;
  const int tx = tt / nty;
  const int x0 = tx * 8, y0 = ty * 8, z0 = tz * TZ;

  constexpr int B = SpreadBatch<T>::value;
  constexpr int TAB = B * (8 + 4 + 8 + 8);
  constexpr int STAGE = 3 * 8 * 8 * 32;
  __shared__ __align__(16) T buf[TAB > STAGE ? TAB : STAGE];
  __shared__ int s_mask[B], s_sz[B];
  T (*s_gx)[8] = reinterpret_cast<T(*)[8]>(buf);
  T (*s_f)[4] = reinterpret_cast<T(*)[4]>(buf + B * 8);
  T (*s_gy)[8] = reinterpret_cast<T(*)[8]>(buf + B * 12);
  T (*s_gz)[8] = reinterpret_cast<T(*)[8]>(buf + B * 20);

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int lz = lane & 7, ly = lane >> 3;
  const int zl = 8 * wv + lz;
  const int L0 = A.tile_start[tile], L1 = A.tile_start[tile + 1];
  T acc[8][3];
#pragma unroll
  for (int xi = 0; xi < 8; ++xi) acc[xi][0] = acc[xi][1] = acc[xi][2] = T(0);
  const T* W = wsel<T>(A.w64, A.w32);

  for (int base = L0; base < L1; base += B) {
    const int cnt = min(B, L1 - base);
    if (t < cnt) {
      const int s = A.tile_list[base + t];
      const int4 r = A.rec[s];
      const T* ws = W + (size_t)s * 3 * ng;
      s_f[t][0] = T(A.F[3 * (size_t)r.w]);
      s_f[t][1] = T(A.F[3 * (size_t)r.w + 1]);
      s_f[t][2] = T(A.F[3 * (size_t)r.w + 2]);
      s_f[t][3] = T(0);
      const int sx = wrap_half(r.x + omin - x0, g.nx);
      const int sy = wrap_half(r.y + omin - y0, g.ny);
      const int sz = wrap_half(r.z + omin - z0, g.nz);
#pragma unroll
      for (int xi = 0; xi < 8; ++xi) {
        int o = xi - sx;
        if (o < 0) o += g.nx; else if (o >= g.nx) o -= g.nx;
        s_gx[t][xi] = o < ng ? ws[o] : T(0);
      }
#pragma unroll
      for (int yi = 0; yi < 8; ++yi) {
        int o = yi - sy;
        if (o < 0) o += g.ny; else if (o >= g.ny) o -= g.ny;
        s_gy[t][yi] = o < ng ? ws[ng + o] : T(0);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) s_gz[t][i] = i < ng ? ws[2 * ng + i] : T(0);
      // wavefront w covers tile z in [8w, 8w + 8): the stencil covers z = sz + o, o < ng (mod nz)
      int mask = 0;
      for (int w = 0; w < NW; ++w) {
        int o0 = 8 * w - sz;  // first offset of the slab, wrapped into [0, nz)
        if (o0 < 0) o0 += g.nz; else if (o0 >= g.nz) o0 -= g.nz;
        const int o7 = o0 + 7 < g.nz ? o0 + 7 : o0 + 7 - g.nz;  // last offset
        if (o0 < ng || o7 < ng || (o0 > o7)) mask |= 1 << w;  // o0 > o7: the slab wraps through offset 0
      }
      s_mask[t] = mask;
      s_sz[t] = sz;
    }
    __syncthreads();
    if (wv < NW) {
#pragma unroll
      for (int q = 0; q < B / 64; ++q) {
        const int pl = 64 * q + lane;
        const bool mine = pl < cnt && ((s_mask[pl] >> wv) & 1);
        unsigned long long bits = __ballot(mine);
        while (bits) {
          const int p = 64 * q + __builtin_ctzll(bits);
          bits &= bits - 1;
          int oz = zl - s_sz[p];
          if (oz < 0) oz += g.nz; else if (oz >= g.nz) oz -= g.nz;
          const T gz = s_gz[p][oz < 8 ? oz : 7];
          const T w = s_gy[p][ly] * (oz < ng ? gz : T(0));
          const T wf0 = w * s_f[p][0], wf1 = w * s_f[p][1], wf2 = w * s_f[p][2];
#pragma unroll
          for (int xi = 0; xi < 8; ++xi) {
            const T gx = s_gx[p][xi];
            acc[xi][0] += wf0 * gx;
            acc[xi][1] += wf1 * gx;
            acc[xi][2] += wf2 * gx;
          }
        }
      }
    }
    __syncthreads();
  }
  T* stage = buf;
  if (wv < NW) {
#pragma unroll
    for (int xi = 0; xi < 8; ++xi)
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) stage[((cc * 8 + xi) * 8 + ly) * TZ + zl] = acc[xi][cc];
  }
  __syncthreads();
  const size_t rstride = 2 * (size_t)g.cpitch;
  const int total = 3 * 8 * 8 * TZ;
  for (int e = t; e < total; e += 256) {
    const int zi = e % TZ;
    const int row = e / TZ;
    const int yi = row & 7, xi = (row >> 3) & 7, cc = row >> 6;
    const size_t idx = ((size_t)(x0 + xi) * g.ny + (y0 + yi)) * rstride + z0 + zi;
    reinterpret_cast<T*>(fp.p[cc])[idx] = stage[e];
  }
}

// the generic spreading kernel's limits: stencil, tile shape, and the bins
// whose blobs can reach a tile (a tile of ext nodes meets at most
// (ext + ng - 2) / b + 2 bins of b nodes per axis), which it lists in LDS
bool spread_generic_ok(const GridGeom& g, const BlobBins& bb, int ng) {
  if (ng > SP_NGMAX || bb.sx * bb.bx > SP_TXMAX || bb.sy * bb.by * bb.sz * bb.bz > 256) return false;
  const int ext[3] = {bb.sx * bb.bx, bb.sy * bb.by, bb.sz * bb.bz}, b[3] = {bb.bx, bb.by, bb.bz},
            nb[3] = {bb.nbx, bb.nby, bb.nbz};
  long long cand = 1;
  for (int d = 0; d < 3; ++d) cand *= std::min(nb[d], (ext[d] + ng - 2) / b[d] + 2);
  (void)g;
  return cand <= SP_MAXC;
}

bool spread_fast_ok(const GridGeom& g, const BlobBins& bb, int ng) {
  return ng <= 8 && g.nx >= 16 && g.nx % 8 == 0 && g.ny % 8 == 0 && g.nz % 8 == 0 && bb.bx == 4 && bb.by == 8 &&
         bb.bz == 8;
}

int spread_tiles(const GridGeom& g) {
  const int NW = g.nz % 32 == 0 ? 4 : (g.nz % 16 == 0 ? 2 : 1);
  return (g.nx / 8) * (g.ny / 8) * (g.nz / (8 * NW));
}

hipError_t build_tile_lists(const GridGeom& g, const BlobBins& bb, BlobData& bd, hipStream_t st) {
  if (!spread_fast_ok(g, bb, bd.ng)) return hipSuccess;
  const int NW = g.nz % 32 == 0 ? 4 : (g.nz % 16 == 0 ? 2 : 1);
  const int nt = spread_tiles(g);
  hipLaunchKernelGGL(tile_list_kernel, dim3(nt), dim3(256), 0, st, g, bb, bd.ng, bd.bin_start, bd.rec, NW, 0,
                     bd.tile_count, (const int*)nullptr, (int*)nullptr);
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, bd.tile_count, bd.tile_start, bd.tile_cursor, nt);
  hipLaunchKernelGGL(tile_list_kernel, dim3(nt), dim3(256), 0, st, g, bb, bd.ng, bd.bin_start, bd.rec, NW, 1,
                     bd.tile_count, (const int*)bd.tile_start, bd.tile_list);
  return hipGetLastError();
}

template <typename T>
hipError_t spread(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                  hipStream_t st, bool images) {
  if (spread_fast_ok(g, bb, bd.ng)) {
    const int NW = g.nz % 32 == 0 ? 4 : (g.nz % 16 == 0 ? 2 : 1);
    const int nt = spread_tiles(g);
    SpreadListArgs A{g, bd.ng, NW, bd.tile_start, bd.tile_list, bd.rec, bd.w64, bd.w32, F};
    const int variant = bd.spread_variant[sizeof(T) == 4 ? 1 : 0];
    if (variant == 2)
      hipLaunchKernelGGL((spread_bits_kernel<T>), dim3(8 * ((nt + 7) / 8)), dim3(256), 0, st, A, fp);
    else if (variant == 1)
      hipLaunchKernelGGL((spread_dense_kernel<T>), dim3(8 * ((nt + 7) / 8)), dim3(256), 0, st, A, fp);
    else
      hipLaunchKernelGGL((spread_fast_kernel<T>), dim3(8 * ((nt + 7) / 8)), dim3(256), 0, st, A, fp);
  } else {
    SpreadArgs A{g, bb, bd.ng, bd.bin_start, bd.rec, bd.w64, bd.w32, F};
    if (!spread_generic_ok(g, bb, bd.ng)) return hipErrorInvalidValue;
    const int ntiles = (g.nx / (bb.sx * bb.bx)) * (g.ny / (bb.sy * bb.by)) * (g.nz / (bb.sz * bb.bz));
    hipLaunchKernelGGL((spread_kernel<T>), dim3(ntiles), dim3(256), 0, st, A, fp);
  }
  const hipError_t e = hipGetLastError();
  return e != hipSuccess || !images ? e : channel_images<T>(g, fp, st);
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
  store_blob_velocity(A.V, A.kt, A.dx3, id, v0, v1, v2);
}

// one bin per workgroup; node box of (bin + ng - 1) per axis staged in LDS
template <typename T>
__global__ void __launch_bounds__(256) interp_box_kernel(InterpArgs A, FieldPtrs<T> fp) {
  if (fp.gate && *fp.gate) return;
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
  if (fp.gate && *fp.gate) return;
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

// ---------------------------------------------------------------------------
// interpolation, main path (same grid conditions as the main spreading path).
// One workgroup per tile of 8 x 8 x 8*IZB nodes (2 x 1 x IZB bins): the node
// box of the tile plus the stencil halo is staged in LDS (all three
// components in FP32, one component per pass in FP64), then each wavefront
// reduces one blob at a time over its ng x ng (y, z) lanes.
// ---------------------------------------------------------------------------
template <typename T> struct IfNc { static constexpr int value = sizeof(T) == 4 ? 3 : 1; };

template <typename T, int NG>
__global__ void __launch_bounds__(256) interp_fast_kernel(InterpArgs A, FieldPtrs<T> fp, int IZB) {
  if (fp.gate && *fp.gate) return;
  constexpr int NC = IfNc<T>::value;
  constexpr int OMIN = 1 - NG / 2;
  constexpr int HX = 8 + NG - 1, HY = 8 + NG - 1;
  extern __shared__ unsigned char smem_raw[];
  T* box = reinterpret_cast<T*>(smem_raw);
  __shared__ int s_lo[2], s_hi[2];
  __shared__ double vres[256][3];
  const GridGeom& g = A.g;
  const BlobBins& bb = A.bb;
  const int TZ = 8 * IZB;
  const int nty = g.ny / 8, ntz = g.nz / TZ;
  const int ntiles = (g.nx / 8) * nty * ntz;
  int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int tz = tile % ntz; tile /= ntz;
  const int ty = tile % nty;
  const int tx = tile / nty;
  const int x0 = tx * 8, y0 = ty * 8, z0 = tz * TZ;
  const int t = threadIdx.x;
  if (t < 2) {
    const int b0 = ((2 * tx + t) * bb.nby + ty) * bb.nbz + tz * IZB;
    s_lo[t] = A.bin_start[b0];
    s_hi[t] = A.bin_start[b0 + IZB];
  }
  __syncthreads();
  const int n0 = s_hi[0] - s_lo[0], nblob = n0 + (s_hi[1] - s_lo[1]);
  if (nblob == 0) return;
  const int HZ = TZ + NG - 1;
  const int nbox = HX * HY * HZ;
  const size_t rstride = 2 * (size_t)g.cpitch;
  const T* W = wsel<T>(A.w64, A.w32);
  const int lane = t & 63, wv = t >> 6;
  const int X0 = x0 + OMIN, Y0 = y0 + OMIN, Z0 = z0 + OMIN;
  const int zl = t & 31;
  int zg = Z0 + zl;
  if (zg < 0) zg += g.nz; else if (zg >= g.nz) zg -= g.nz;
  const bool zok = zl < HZ;
  size_t rowoff[2];
  bool yok[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int hy = (t >> 5) + 8 * j;
    int yg = Y0 + hy;
    if (yg < 0) yg += g.ny; else if (yg >= g.ny) yg -= g.ny;
    yok[j] = hy < HY && zok;
    rowoff[j] = (size_t)yg * rstride + zg;
  }
  const int jy = lane / NG, jz = lane - (lane / NG) * NG;
  const bool pair_ok = lane < NG * NG;

  for (int chunk = 0; chunk < nblob; chunk += 256) {
    const int nch = min(256, nblob - chunk);
    for (int c0 = 0; c0 < 3; c0 += NC) {
      if (c0 > 0 || chunk > 0) __syncthreads();
      // stage the node box: all loads of half the x rows are issued before
      // their LDS stores
      constexpr int HALF = (HX + 1) / 2;
#pragma unroll
      for (int h2 = 0; h2 < 2; ++h2) {
        T v[HALF][2][NC];
#pragma unroll
        for (int a = 0; a < HALF; ++a) {
          const int hx = h2 * HALF + a;
          int xg = X0 + hx;
          if (xg < 0) xg += g.nx; else if (xg >= g.nx) xg -= g.nx;
          const size_t xoff = (size_t)xg * g.ny * rstride;
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const bool ok = hx < HX && yok[j];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
              v[a][j][cc] = ok ? reinterpret_cast<const T*>(fp.p[c0 + cc])[xoff + rowoff[j]] : T(0);
          }
        }
#pragma unroll
        for (int a = 0; a < HALF; ++a) {
          const int hx = h2 * HALF + a;
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            const int hy = (t >> 5) + 8 * j;
            if (hx < HX && yok[j])
#pragma unroll
              for (int cc = 0; cc < NC; ++cc) box[cc * nbox + (hx * HY + hy) * HZ + zl] = v[a][j][cc];
          }
        }
      }
      __syncthreads();
      for (int k = wv; k < nch; k += 4) {
        const int kk = chunk + k;
        const int sidx = __builtin_amdgcn_readfirstlane(kk < n0 ? s_lo[0] + kk : s_lo[1] + (kk - n0));
        const int4 r = A.rec[sidx];
        const T* ws = W + (size_t)sidx * 3 * NG;
        const int lx = r.x - x0, ly = r.y - y0, lz = r.z - z0;
        T a[NC];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) a[cc] = T(0);
        if (pair_ok) {
          const int hyz = (ly + jy) * HZ + (lz + jz);
          T sv[NC];
#pragma unroll
          for (int cc = 0; cc < NC; ++cc) sv[cc] = T(0);
#pragma unroll
          for (int jx = 0; jx < NG; ++jx) {
            const int i = (lx + jx) * HY * HZ + hyz;
            const T gx = ws[jx];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) sv[cc] += gx * box[cc * nbox + i];
          }
          const T wyz = ws[NG + jy] * ws[2 * NG + jz];
#pragma unroll
          for (int cc = 0; cc < NC; ++cc) a[cc] = wyz * sv[cc];
        }
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
          const T v = wave_sum(a[cc]);
          if (lane == 0) vres[k][c0 + cc] = (double)v;
        }
      }
    }
    __syncthreads();
    for (int k = t; k < nch; k += 256) {
      const int kk = chunk + k;
      const int sidx = kk < n0 ? s_lo[0] + kk : s_lo[1] + (kk - n0);
      write_velocity(A, A.rec[sidx].w, vres[k][0], vres[k][1], vres[k][2]);
    }
  }
}

// Interpolation on 8 x TY x TZ tiles (TY, TZ = 16 or 8), one velocity
// component per pass: the node box of the tile plus the stencil halo, one
// component at a time, fits in 27-35 KB of LDS, so four workgroups share a CU
// and the halo is re-read 3.3x instead of 4.2x.  Blobs are taken by anchor
// bin (the tile owns 2 x 2^(TY/16) x 2^(TZ/16)... bins, listed as x,y bin pairs
// with contiguous z bins), one wavefront per blob, lanes over (jy, jz), the x
// stencil in a register loop, DPP reduction.
template <typename T, int NG, int TY, int TZ>
__global__ void __launch_bounds__(256) interp_slab_kernel(InterpArgs A, FieldPtrs<T> fp) {
  if (fp.gate && *fp.gate) return;
  constexpr int OMIN = 1 - NG / 2;
  constexpr int HX = 8 + NG - 1, HY = TY + NG - 1, HZ = TZ + NG - 1;
  constexpr int NBOX = HX * HY * HZ;
  constexpr int NYB = TY / 8, NZB = TZ / 8;  // y and z bins per tile
  constexpr int NR = 2 * NYB;                // bin ranges: (x bin, y bin) pairs
  static_assert(HZ <= 32, "z rows are loaded by 32 lanes");
  __shared__ T box[NBOX];
  __shared__ double vres[256][3];
  __shared__ int s_lo[NR], s_cnt[NR + 1];
  const GridGeom& g = A.g;
  const BlobBins& bb = A.bb;
  const int nty = g.ny / TY, ntz = g.nz / TZ;
  const int ntiles = (g.nx / 8) * nty * ntz;
  int tile = xcd_tile(blockIdx.x, ntiles);
  if (tile >= ntiles) return;
  const int tz = tile % ntz; tile /= ntz;
  const int ty = tile % nty;
  const int tx = tile / nty;
  const int x0 = tx * 8, y0 = ty * TY, z0 = tz * TZ;
  const int t = threadIdx.x;
  if (t == 0) {
    int acc = 0;
    for (int r = 0; r < NR; ++r) {
      const int xb = 2 * tx + (r / NYB), yb = ty * NYB + (r % NYB);
      const int b0 = (xb * bb.nby + yb) * bb.nbz + tz * NZB;
      s_lo[r] = A.bin_start[b0];
      s_cnt[r] = acc;
      acc += A.bin_start[b0 + NZB] - A.bin_start[b0];
    }
    s_cnt[NR] = acc;
  }
  __syncthreads();
  const int nblob = s_cnt[NR];
  if (nblob == 0) return;
  const size_t rstride = 2 * (size_t)g.cpitch;
  const T* W = wsel<T>(A.w64, A.w32);
  const int lane = t & 63, wv = t >> 6;
  // the box is loaded plane by plane in x; thread t owns the (y, z) entries
  // e = t + 256 j of every plane, whose offsets are fixed for the tile
  constexpr int HYZ = HY * HZ;
  constexpr int JN = (HYZ + 255) / 256;
  size_t yzoff[JN];
  bool yzok[JN];
#pragma unroll
  for (int j = 0; j < JN; ++j) {
    const int e = t + 256 * j;
    yzok[j] = e < HYZ;
    const int hy = e / HZ, hz = e - (e / HZ) * HZ;
    int yg = y0 + OMIN + hy, zg = z0 + OMIN + hz;
    if (yg < 0) yg += g.ny; else if (yg >= g.ny) yg -= g.ny;
    if (zg < 0) zg += g.nz; else if (zg >= g.nz) zg -= g.nz;
    yzoff[j] = yzok[j] ? (size_t)yg * rstride + zg : 0;
  }
  const int jy = lane / NG, jz = lane - (lane / NG) * NG;
  const bool pair_ok = lane < NG * NG;

  for (int chunk = 0; chunk < nblob; chunk += 256) {
    const int nch = min(256, nblob - chunk);
    for (int c = 0; c < 3; ++c) {
      __syncthreads();
      const T* __restrict__ src = reinterpret_cast<const T*>(fp.p[c]);
      T pv[HX][JN];
#pragma unroll
      for (int hx = 0; hx < HX; ++hx) {
        int xg = x0 + OMIN + hx;
        if (xg < 0) xg += g.nx; else if (xg >= g.nx) xg -= g.nx;
        const T* plane = src + (size_t)xg * g.ny * rstride;
#pragma unroll
        for (int j = 0; j < JN; ++j) pv[hx][j] = yzok[j] ? plane[yzoff[j]] : T(0);
      }
#pragma unroll
      for (int hx = 0; hx < HX; ++hx)
#pragma unroll
        for (int j = 0; j < JN; ++j)
          if (yzok[j]) box[hx * HYZ + t + 256 * j] = pv[hx][j];
      __syncthreads();
      for (int k = wv; k < nch; k += 4) {
        const int kk = chunk + k;
        int r = 0;
#pragma unroll
        for (int q = 1; q < NR; ++q) r += (kk >= s_cnt[q]) ? 1 : 0;
        const int sidx = __builtin_amdgcn_readfirstlane(s_lo[r] + (kk - s_cnt[r]));
        const int4 rc = A.rec[sidx];
        const T* ws = W + (size_t)sidx * 3 * NG;
        const int lx = rc.x - x0, ly = rc.y - y0, lz = rc.z - z0;
        T a = T(0);
        if (pair_ok) {
          const int hyz = (ly + jy) * HZ + (lz + jz);
          T sv = T(0);
#pragma unroll
          for (int jx = 0; jx < NG; ++jx) sv += ws[jx] * box[(lx + jx) * HY * HZ + hyz];
          a = ws[NG + jy] * ws[2 * NG + jz] * sv;
        }
        const T v = wave_sum(a);
        if (lane == 0) vres[k][c] = (double)v;
      }
    }
    __syncthreads();
    for (int k = t; k < nch; k += 256) {
      const int kk = chunk + k;
      int r = 0;
#pragma unroll
      for (int q = 1; q < NR; ++q) r += (kk >= s_cnt[q]) ? 1 : 0;
      const int sidx = s_lo[r] + (kk - s_cnt[r]);
      write_velocity(A, A.rec[sidx].w, vres[k][0], vres[k][1], vres[k][2]);
    }
  }
}

static bool interp_fast_ok(const GridGeom& g, const BlobBins& bb, int ng) {
  return spread_fast_ok(g, bb, ng) && bb.nbx % 2 == 0;
}

template <typename T>
hipError_t interp(const GridGeom& g, const BlobBins& bb, const BlobData& bd, const FieldPtrs<T>& fp, double* V,
                  const KTerm& kt, hipStream_t st) {
  InterpArgs A{g, bb, bd.ng, bd.bin_start, bd.rec, bd.iw64, bd.iw32, V, kt, g.dx * g.dx * g.dx};
  if (bd.nb == 0) return hipSuccess;
  const int izb_choice = bd.interp_izb[sizeof(T) == 4 ? 1 : 0];
  if (izb_choice == 3 && interp_fast_ok(g, bb, bd.ng) && (bd.ng == 7 || bd.ng == 8) && g.nz % 16 == 0 &&
      g.ny % (sizeof(T) == 4 ? 16 : 16) == 0) {
    constexpr int TZ = sizeof(T) == 4 ? 16 : 8;
    const int ntiles = (g.nx / 8) * (g.ny / 16) * (g.nz / TZ);
    if (bd.ng == 7)
      hipLaunchKernelGGL((interp_slab_kernel<T, 7, 16, TZ>), dim3(8 * ((ntiles + 7) / 8)), dim3(256), 0, st, A, fp);
    else
      hipLaunchKernelGGL((interp_slab_kernel<T, 8, 16, TZ>), dim3(8 * ((ntiles + 7) / 8)), dim3(256), 0, st, A, fp);
    return hipGetLastError();
  }
  if (interp_fast_ok(g, bb, bd.ng) && (bd.ng == 7 || bd.ng == 8)) {
    const int IZB = (bd.interp_izb[sizeof(T) == 4 ? 1 : 0] == 1 || bb.nbz % 2 != 0) ? 1 : 2;
    const int TZ = 8 * IZB;
    const size_t lds = IfNc<T>::value * sizeof(T) * (size_t)(8 + bd.ng - 1) * (8 + bd.ng - 1) * (TZ + bd.ng - 1);
    const int ntiles = (g.nx / 8) * (g.ny / 8) * (g.nz / TZ);
    if (bd.ng == 7)
      hipLaunchKernelGGL((interp_fast_kernel<T, 7>), dim3(8 * ((ntiles + 7) / 8)), dim3(256), lds, st, A, fp, IZB);
    else
      hipLaunchKernelGGL((interp_fast_kernel<T, 8>), dim3(8 * ((ntiles + 7) / 8)), dim3(256), lds, st, A, fp, IZB);
    return hipGetLastError();
  }
  const size_t lds = 3 * sizeof(T) * (size_t)(bb.bx + bd.ng - 1) * (bb.by + bd.ng - 1) * (bb.bz + bd.ng - 1);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL((interp_box_kernel<T>), dim3(bb.nbins()), dim3(256), lds, st, A, fp);
  } else {
    hipLaunchKernelGGL((interp_gather_kernel<T>), dim3((bd.nb + 3) / 4), dim3(256), 0, st, A, fp, bd.nb);
  }
  return hipGetLastError();
}

template hipError_t spread<double>(const GridGeom&, const BlobBins&, const BlobData&, const double*, const FieldPtrs<double>&, hipStream_t, bool);
template hipError_t spread<float>(const GridGeom&, const BlobBins&, const BlobData&, const double*, const FieldPtrs<float>&, hipStream_t, bool);
template hipError_t interp<double>(const GridGeom&, const BlobBins&, const BlobData&, const FieldPtrs<double>&, double*, const KTerm&, hipStream_t);
template hipError_t interp<float>(const GridGeom&, const BlobBins&, const BlobData&, const FieldPtrs<float>&, double*, const KTerm&, hipStream_t);

}  // namespace fcm
