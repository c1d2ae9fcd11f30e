// fcm_column.hip -- the patch lists of the column path and its FP64 kernels
// (the kernels and the design are described in fcm_column.hpp).
#include "fcm_column.hpp"

namespace fcm {

// ---------------------------------------------------------------------------
// patch lists
// ---------------------------------------------------------------------------
struct ColListArgs {
  GridGeom g;
  BlobBins bb;
  int ng, cx, npx, npy, nb;
  const int* bin_start;
  const int4* rec;
  int* count;
  const int* start;
  int* list;
  int4* ent;
  int* zst;
};

__device__ __forceinline__ long long col_key(const int4* rec, int s, int omin, int nz) {
  return ((long long)col_wrapi(rec[s].z + omin, nz) << 32) | (unsigned)s;
}

// One 64-thread block per patch; thread b takes the group of blobs with
// stencil start z in [8b, 8b + 8).  Pass 0 counts, pass 1 writes the group in
// (stencil start z, slot) order and its part of the z prefix table.
__global__ void __launch_bounds__(64) col_list_kernel(ColListArgs A, int pass) {
  const GridGeom& g = A.g;
  const BlobBins& bb = A.bb;
  const int ng = A.ng, omin = 1 - ng / 2, cx = A.cx;
  const int p = blockIdx.x;
  const int px = p / A.npy, py = p - (p / A.npy) * A.npy;
  const int x0 = cx * px, y0 = COL_CY * py;
  // anchor bins whose stencils can reach the patch in x and y
  int xlo = col_fdiv(x0 - omin - ng + 1, bb.bx);
  int xn = col_fdiv(x0 + cx - 1 - omin, bb.bx) - xlo + 1;
  if (xn >= bb.nbx) { xlo = 0; xn = bb.nbx; }
  int ylo = col_fdiv(y0 - omin - ng + 1, bb.by);
  int yn = col_fdiv(y0 + COL_CY - 1 - omin, bb.by) - ylo + 1;
  if (yn >= bb.nby) { ylo = 0; yn = bb.nby; }
  const int nbz = bb.nbz;
  for (int b = threadIdx.x; b < nbz; b += 64) {
    // stencil starts [8b, 8b + 8) <=> anchors [8b - omin, 8b + 7 - omin]
    const int zlo = col_fdiv(8 * b - omin, bb.bz), zhi = col_fdiv(8 * b + 7 - omin, bb.bz);
    const int gi = p * nbz + b;
    const int pos0 = pass == 1 ? A.start[gi] : 0;
    int pos = pos0;
    for (int i = 0; i < xn; ++i)
      for (int j = 0; j < yn; ++j) {
        const int bxy = col_wrapi(xlo + i, bb.nbx) * bb.nby + col_wrapi(ylo + j, bb.nby);
        for (int zb = zlo; zb <= zhi; ++zb) {
          const int bin = bxy * nbz + col_wrapi(zb, nbz);
          const int s1 = A.bin_start[bin + 1];
          for (int s = A.bin_start[bin]; s < s1; ++s) {
            const int4 r = A.rec[s];
            if ((col_wrapi(r.z + omin, g.nz) >> 3) != b) continue;
            const int rx = col_wrap_half(col_wrapi(r.x + omin, g.nx) - x0, g.nx);
            const int ry = col_wrap_half(col_wrapi(r.y + omin, g.ny) - y0, g.ny);
            if (rx > cx - 1 || rx + ng - 1 < 0 || ry > COL_CY - 1 || ry + ng - 1 < 0) continue;
            if (pass == 1) A.list[pos] = s;
            ++pos;
          }
        }
      }
    if (pass == 0) {
      A.count[gi] = pos;
      continue;
    }
    for (int i = pos0 + 1; i < pos; ++i) {  // insertion sort of the group
      const int v = A.list[i];
      const long long kv = col_key(A.rec, v, omin, g.nz);
      int j = i - 1;
      while (j >= pos0 && col_key(A.rec, A.list[j], omin, g.nz) > kv) {
        A.list[j + 1] = A.list[j];
        --j;
      }
      A.list[j + 1] = v;
    }
    int* zst = A.zst + (size_t)p * (g.nz + 1);
    int i = pos0;
    for (int k = 0; k < 8; ++k) {
      const int z = 8 * b + k;
      while (i < pos && col_wrapi(A.rec[A.list[i]].z + omin, g.nz) < z) ++i;
      zst[z] = i;
    }
    if (b == nbz - 1) zst[g.nz] = pos;
    for (int e = pos0; e < pos; ++e) {
      const int sl = A.list[e];
      const int4 r = A.rec[sl];
      const int sxw = col_wrapi(r.x + omin, g.nx), syw = col_wrapi(r.y + omin, g.ny);
      const int rx = col_wrap_half(sxw - x0, g.nx), ry = col_wrap_half(syw - y0, g.ny);
      const int dst = col_wrapi(px - sxw / cx, A.npx) * 3 + col_wrapi(py - syw / COL_CY, A.npy);
      const int dxw = col_wrapi(px - sxw / cx, A.npx), dyw = col_wrapi(py - syw / COL_CY, A.npy);
      A.ent[e] = make_int4(sl, r.w,
                           (rx + 8) | ((ry + 8) << 5) | (col_wrapi(r.z + omin, g.nz) << 10) | (dxw << 20) | (dyw << 23),
                           dst * A.nb + sl);
    }
  }
}

// The FP64 weight windows of the column kernels (COL_WIN words per sorted
// blob; the FP32 kernels evaluate their weights from the stencil offsets).
// A blob whose stencil starts at x node sx touches the patches floor(sx / cx)
// + d, d = 0, 1, ...; in patch d its weights for the patch's columns ii are
// g[cx d - sx mod cx + ii] (zero outside 0..ng-1), stored at words cx d + ii.
// The same for y with COL_CY (words 12 + 4 d + jj), and the ng z weights at
// 24..31.  A list entry then reads its weights as three aligned vectors of one
// 256-byte window instead of gathering them word by word.
__device__ void col_window_one(const double* __restrict__ ws, int ng, int cx, int rx0, int ry0,
                               double* __restrict__ out) {
  for (int q = 0; q < 12; ++q) {
    const int d = q / cx, ii = q - (q / cx) * cx;
    const int o = cx * d - rx0 + ii;
    out[q] = (o >= 0 && o < ng) ? ws[o] : 0.0;
  }
  for (int q = 0; q < 12; ++q) {
    const int d = q / COL_CY, jj = q - (q / COL_CY) * COL_CY;
    const int o = COL_CY * d - ry0 + jj;
    out[12 + q] = (o >= 0 && o < ng) ? ws[ng + o] : 0.0;
  }
  for (int q = 0; q < 8; ++q) out[24 + q] = q < ng ? ws[2 * ng + q] : 0.0;
}

__global__ void __launch_bounds__(256) col_window_kernel(GridGeom g, int ng, int cx, int nb, const int4* __restrict__ rec,
                                                          const double* __restrict__ w64, double* __restrict__ o64) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nb) return;
  const int omin = 1 - ng / 2;
  const int4 r = rec[s];
  const int rx0 = col_wrapi(r.x + omin, g.nx) % cx, ry0 = col_wrapi(r.y + omin, g.ny) % COL_CY;
  col_window_one(w64 + (size_t)s * 3 * ng, ng, cx, rx0, ry0, o64 + (size_t)s * COL_WIN);
}

// The FP64 spreading weights of list entry i (COL_EWIN words), cut from its
// blob's window: the x weights of its patch (words cx dx.., zero past cx), the
// y weights (12 + 4 dy..) and the z weights; the values col_win_load reads.
__global__ void __launch_bounds__(256) col_ewin_kernel(int cx, const int* __restrict__ total,
                                                        const int4* __restrict__ ent, const double* __restrict__ w64,
                                                        double* __restrict__ o64) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= *total) return;
  const int4 e = ent[i];
  const double* wn = w64 + (size_t)e.x * COL_WIN;
  double* out = o64 + (size_t)i * COL_EWIN;
  const int dx = col_dx(e.z), dy = col_dy(e.z);
  for (int q = 0; q < 4; ++q) out[q] = q < cx ? wn[cx * dx + q] : 0.0;
  for (int q = 0; q < 4; ++q) out[4 + q] = wn[12 + 4 * dy + q];
  for (int q = 0; q < 8; ++q) out[8 + q] = wn[24 + q];
}

static bool col_nz_supported(int nz) {
  switch (nz) {
#define FCM_COL_CASE(v) case v:
    FCM_COL_NZ(FCM_COL_CASE)
#undef FCM_COL_CASE
    return true;
    default:
      return false;
  }
}


int col_patch_cx(const GridGeom& g, const BlobBins& bb, int ng) {
  if (ng < 2 || ng > 8) return 0;
  if (bb.bx != 4 || bb.by != 8 || bb.bz != 8) return 0;
  if (g.nx < 16 || g.ny < 16 || g.nz < 16 || g.ny % COL_CY != 0 || !col_nz_supported(g.nz)) return 0;
  const int cx = col_cx_for(g.nz);
  return g.nx % cx == 0 ? cx : 0;
}

size_t col_list_capacity(int nb) { return (size_t)COL_SLOTS_MAX * (size_t)(nb > 0 ? nb : 1); }

size_t col_zst_words(const GridGeom& g, int cx) {
  if (cx <= 0) return 1;
  return (size_t)(g.nx / cx) * (g.ny / COL_CY) * (g.nz + 1);
}

int col_groups(const GridGeom& g, int cx) {
  if (cx <= 0) return 1;
  return (g.nx / cx) * (g.ny / COL_CY) * (g.nz / 8);
}

hipError_t build_col_lists(const GridGeom& g, const BlobBins& bb, BlobData& bd, hipStream_t st) {
  const int cx = bd.col_cx;
  if (cx <= 0 || !bd.col_list || !bd.col_ent || !bd.cwin64 || !bd.ewin64) return hipSuccess;
  ColListArgs A{g, bb, bd.ng, cx, g.nx / cx, g.ny / COL_CY, bd.nb, bd.bin_start, bd.rec, bd.col_count, bd.col_start,
                bd.col_list, bd.col_ent, bd.col_zst};
  const int npatch = A.npx * A.npy;
  hipLaunchKernelGGL(col_list_kernel, dim3(npatch), dim3(64), 0, st, A, 0);
  hipError_t e = exclusive_scan(bd.col_count, bd.col_start, bd.col_cursor, col_groups(g, cx), st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(col_list_kernel, dim3(npatch), dim3(64), 0, st, A, 1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (bd.nb > 0) {
    const dim3 grid((bd.nb + 255) / 256);
    hipLaunchKernelGGL(col_window_kernel, grid, dim3(256), 0, st, g, bd.ng, cx, bd.nb, bd.rec, bd.w64, bd.cwin64);
    if (bd.icwin64 != bd.cwin64)
      hipLaunchKernelGGL(col_window_kernel, grid, dim3(256), 0, st, g, bd.ng, cx, bd.nb, bd.rec, bd.iw64,
                         bd.icwin64);
    // every entry a list can hold (the count of this geometry is read on the device)
    const size_t cap = col_list_capacity(bd.nb);
    hipLaunchKernelGGL(col_ewin_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, cx,
                       bd.col_start + col_groups(g, cx), bd.col_ent, bd.cwin64, bd.ewin64);
  }
  return hipGetLastError();
}

// the shapes (patch width, z-pass points per thread) the column kernels are
// built for, per word size and nz
int col_shape_pm(int tbytes, int nz, int cx, int pm) {
  switch (nz) {
#define FCM_COL_PM(v)                                                                              \
  case v:                                                                                          \
    if (cx == 4) return tbytes == 4 ? col_pm_pick<float, v, 4>(pm) : col_pm_pick<double, v, 4>(pm); \
    if (cx == 2) return tbytes == 4 ? col_pm_pick<float, v, 2>(pm) : col_pm_pick<double, v, 2>(pm); \
    return 0;
    FCM_COL_NZ(FCM_COL_PM)
#undef FCM_COL_PM
    default:
      return 0;
  }
}

bool col_shape_valid(int tbytes, int nz, int cx, int pm) { return col_shape_pm(tbytes, nz, cx, pm) != 0; }

template hipError_t spread_zfwd<double>(const GridGeom&, const BlobData&, const double*, const FieldPtrs<double>&,
                                        const TwiddleSet<double>&, hipStream_t, bool);
template hipError_t zinv_interp<double>(const GridGeom&, const BlobData&, const FieldPtrs<double>&,
                                        const TwiddleSet<double>&, double*, const KTerm&, hipStream_t, hipEvent_t, int);

}  // namespace fcm
