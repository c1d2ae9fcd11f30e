// fcm_steric.hip -- steric repulsion between rigid bodies on a device cell list,
// and the end-of-step body update.
//
// linked_list.cpp builds a 6^3 cell list on the host (populate, :80-102) and
// walks each cell against itself and 13 neighbours, adding the soft-sphere
// force of every pair of blobs in different bodies to both blobs
// (apply_barrier_forces, :192-281).  Here blobs are counting-sorted by cell
// (ascending blob index inside a cell), and every blob GATHERS the forces of
// the blobs in its 27 surrounding cells in a fixed order: no atomics on the
// forces, bitwise repeatable, and the same per-pair terms as the reference
// (action and reaction are computed by the two blobs separately and are exact
// negatives).  Two interaction sets: the reference's (an N^3 grid, every pair
// in the same or neighbouring cells at any distance: for N >= 3 the 27 cells
// around a blob are the cell, its 13 half-shell cells and theirs mirrored) or
// pairs closer than a cutoff, on cells at least the cutoff wide.
#include "fcm_internal.hpp"

namespace fcm {

// find_containing_cell (linked_list.cpp:74-78): floor(nc (y / L)), kept in range
__device__ __forceinline__ int cell_coord(double y, double L, int nc) {
  int c = (int)floor(nc * (y / L));
  return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

__global__ void ll_cell_kernel(const double* __restrict__ Y, int nb, StericCells cg, int* __restrict__ cell_of,
                               int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int cx = cell_coord(Y[3 * (size_t)i], cg.L[0], cg.nc[0]);
  const int cy = cell_coord(Y[3 * (size_t)i + 1], cg.L[1], cg.nc[1]);
  const int cz = cell_coord(Y[3 * (size_t)i + 2], cg.L[2], cg.nc[2]);
  const int c = (cx * cg.nc[1] + cy) * cg.nc[2] + cz;
  cell_of[i] = c;
  atomicAdd(&count[c], 1);
}

// neighbour cell offsets along one axis: -1, 0, 1, or every cell when the axis
// has fewer than 3 cells (so that no cell is visited twice)
__device__ __forceinline__ int axis_cells(int c, int nc, int k) {
  if (nc < 3) return k;
  int v = c + k - 1;
  return v < 0 ? v + nc : (v >= nc ? v - nc : v);
}

__global__ void __launch_bounds__(256) ll_force_kernel(const double* __restrict__ Y, int nb, int nbpb, StericCells cg,
                                                       const int* __restrict__ start, const int* __restrict__ sorted,
                                                       double cut2, double U0, double bl, double two_a,
                                                       double wall_L, double wall_K, double* __restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double x1 = Y[3 * (size_t)i], y1 = Y[3 * (size_t)i + 1], z1 = Y[3 * (size_t)i + 2];
  const int body1 = i / nbpb;
  const int cx = cell_coord(x1, cg.L[0], cg.nc[0]);
  const int cy = cell_coord(y1, cg.L[1], cg.nc[1]);
  const int cz = cell_coord(z1, cg.L[2], cg.nc[2]);
  const int kx = cg.nc[0] < 3 ? cg.nc[0] : 3, ky = cg.nc[1] < 3 ? cg.nc[1] : 3, kz = cg.nc[2] < 3 ? cg.nc[2] : 3;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int a = 0; a < kx; ++a) {
    const int ax = axis_cells(cx, cg.nc[0], a);
    for (int b = 0; b < ky; ++b) {
      const int ay = axis_cells(cy, cg.nc[1], b);
      for (int c = 0; c < kz; ++c) {
        const int az = axis_cells(cz, cg.nc[2], c);
        const int cell = (ax * cg.nc[1] + ay) * cg.nc[2] + az;
        for (int s = start[cell]; s < start[cell + 1]; ++s) {
          const int j = sorted[s];
          if (j / nbpb == body1) continue;  // blobs of one body do not repel (linked_list.cpp:215)
          double d0 = x1 - Y[3 * (size_t)j], d1 = y1 - Y[3 * (size_t)j + 1], d2 = z1 - Y[3 * (size_t)j + 2];
          // nearest image, linked_list.cpp:255-257
          d0 -= cg.L[0] * trunc(2.0 * d0 / cg.L[0]);
          d1 -= cg.L[1] * trunc(2.0 * d1 / cg.L[1]);
          d2 -= cg.L[2] * trunc(2.0 * d2 / cg.L[2]);
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (!(r2 < cut2) || r2 == 0.0) continue;
          const double r = sqrt(r2);
          // interaction_magnitude, linked_list.cpp:377 (SOFT_SPHERE, dynamics)
          const double mag = r < two_a ? U0 / bl : U0 * exp((two_a - r) / bl) / bl;
          fx += mag * (d0 / r);
          fy += mag * (d1 / r);
          fz += mag * (d2 / r);
        }
      }
    }
  }
  if (wall_L > 0.0) {
    // slip-channel wall springs (collection_of_rigid_bodies.cpp:640-662)
    const double hy = 0.5 * cg.L[1];
    if (y1 <= wall_L) fy += wall_K * (wall_L - y1);
    else if ((hy - y1 <= wall_L) && (y1 <= hy)) fy -= wall_K * (wall_L - hy + y1);
  }
  F[3 * (size_t)i] = fx;
  F[3 * (size_t)i + 1] = fy;
  F[3 * (size_t)i + 2] = fz;
}

hipError_t barrier_forces(const double* Y, int nb, int nbpb, const StericCells& cg, double cutoff, bool all_pairs,
                          double kT, double a, bool channel, StericWork& w, double* F, hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  const int ncells = cg.nc[0] * cg.nc[1] * cg.nc[2];
  if (ncells > w.cap_cells) return hipErrorInvalidValue;
  hipError_t e = fill_ints(w.count, ncells, 0, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ll_cell_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, Y, nb, cg, w.cell_of, w.count);
  e = sort_by_key(w.cell_of, nb, ncells, w.count, w.start, w.cursor, w.sorted, st);
  if (e != hipSuccess) return e;
  const double U0 = 20.0 * kT, bl = 0.5 * a;  // linked_list.cpp:367-368
  const double wall_L = channel ? 2.2 * a : 0.0, wall_K = channel ? (40.0 * kT / a) / wall_L : 0.0;
  hipLaunchKernelGGL(ll_force_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, Y, nb, nbpb, cg, w.start, w.sorted,
                     all_pairs ? HUGE_VAL : cutoff * cutoff, U0, bl, 2.0 * a, wall_L, wall_K, F);
  return hipGetLastError();
}

// quaternion.cpp:128-145 then the product lie_exp(u) * q (quaternion.hpp)
__device__ void lie_exp_mul(const double* u, double* q) {
  const double theta = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (theta < 1e-10) return;  // identity
  const double cs = cos(0.5 * theta), sn = sin(0.5 * theta);
  const double p0 = cs, p1 = sn * u[0] / theta, p2 = sn * u[1] / theta, p3 = sn * u[2] / theta;
  const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
  const double s = p0 * q0 - (p1 * q1 + p2 * q2 + p3 * q3);
  const double v0 = p0 * q1 + q0 * p1 + (p2 * q3 - p3 * q2);
  const double v1 = p0 * q2 + q0 * p2 + (p3 * q1 - p1 * q3);
  const double v2 = p0 * q3 + q0 * p3 + (p1 * q2 - p2 * q1);
  q[0] = s;
  q[1] = v0;
  q[2] = v1;
  q[3] = v2;
}

// particle::update (particle.cpp:45-56) with the grand update s V6[6n..6n+5]
__global__ void bodies_step_kernel(double* __restrict__ x, double* __restrict__ q, int NB,
                                   const double* __restrict__ V6, double s, double LX, double LY, double LZ) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= NB) return;
  const double L[3] = {LX, LY, LZ};
  double u[3];
  for (int k = 0; k < 3; ++k) {
    const double v = x[3 * (size_t)n + k] + s * V6[6 * (size_t)n + k];
    x[3 * (size_t)n + k] = v - L[k] * floor(v / L[k]);
    u[k] = s * V6[6 * (size_t)n + 3 + k];
  }
  double qq[4];
  for (int k = 0; k < 4; ++k) qq[k] = q[4 * (size_t)n + k];
  lie_exp_mul(u, qq);
  for (int k = 0; k < 4; ++k) q[4 * (size_t)n + k] = qq[k];
}

hipError_t bodies_step(const GridGeom& g, double* x, double* q, int NB, const double* V6, double s, hipStream_t st) {
  if (NB <= 0) return hipSuccess;
  hipLaunchKernelGGL(bodies_step_kernel, dim3((NB + 255) / 256), dim3(256), 0, st, x, q, NB, V6, s, g.dx * g.nx,
                     g.dx * g.ny, g.dx * g.nz);
  return hipGetLastError();
}

}  // namespace fcm
