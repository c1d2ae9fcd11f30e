// fcm_column.hpp -- the column path's kernels and their launchers (internal to
// libfcm_hip.so; fcm_column.hip builds the FP64 and list-building side,
// fcm_column32.hip the FP32 kernels).
#pragma once
// spreading and interpolation of the mobility product, each
// fused with the z pass of the FFT that touches the same grid rows.
//
// A patch is cx x 4 grid rows (x, y) with all of z.  For every patch the blobs
// whose x-y stencil touches it are listed once per set of positions, ordered by
// the z of their first stencil node ("stencil start"), with a prefix table
// over z.  One workgroup owns a patch and one thread owns one z:
//
//   spread + z R2C (col_spread_kernel): the thread sums F gx gy gz over exactly
//   the blobs whose z stencil covers its z -- a contiguous range of the ordered
//   list -- for every node of the patch rows, in registers.  The rows go
//   through LDS into the nz/2-point complex FFT and the R2C untangling, and
//   only k-space rows are written (fcm_fluid_solver.cpp:552-600 followed by the
//   z stage of fft_solver.cpp:32-36).
//
//   z C2R + interpolation (col_interp_kernel): the k-space rows are brought
//   back to real space in LDS (imaginary parts of kz = 0 and nz/2 ignored, as
//   FFTW's halfcomplex format stores none) and left there z-major; here one
//   thread owns one or two entries of the list instead of a z, and forms
//   gz sum_xy gx gy u over the entry's stencil in a fixed order.  The patch's
//   partial sum goes to the blob's slot for this patch; col_reduce_kernel
//   adds the (at most 5 x 3) slots of a blob in a fixed order and writes
//   dx^3 J u - K V (fcm_fluid_solver.cpp:363-424 preceded by the z stage of
//   fft_solver.cpp:38-42).
//
// Nothing is accumulated with atomics, so both are bitwise repeatable.  Per
// product the grid is written once and read once in k space, instead of real
// space being written, read back, transformed and written again on each side.
#include "fcm_internal.hpp"

#include <type_traits>

namespace fcm {

__device__ __forceinline__ int col_wrapi(int v, int n) {
  const int r = v % n;
  return r < 0 ? r + n : r;
}
__device__ __forceinline__ int col_fdiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__device__ __forceinline__ int col_wrap_half(int v, int n) {
  if (v < -(n / 2)) v += n; else if (v >= n - n / 2) v -= n;
  return v;
}
// blocks are spread over the 8 XCDs so that each works on a contiguous range
// of patches (neighbouring patches share blobs in that XCD's L2)
__device__ __forceinline__ int col_xcd(int b, int n) {
  const int per = (n + 7) / 8;
  return (b % 8) * per + b / 8;
}

struct ColArgs {
  GridGeom g;
  int ng, npx, npy;
  const int4* ent;  // entry records of the patch lists (col_entry)
  const int* zst;
  const double* w64;   // FP64 weight windows (COL_WIN words per sorted blob)
  const double* ew64;  // FP64 spreading: weights per list entry (COL_EWIN words)
  const float4* woff;  // FP32: offsets of stencil node 0 per sorted blob
  float wa, wc, wdx;   // FP32 weight wa exp2(wc xx^2), node spacing
};

// One entry of a patch list: (sorted blob slot, blob id, packed offsets, index
// of the blob's partial-sum slot for this patch: slot-major, dst nb + blob, so
// that the reduction reads consecutive blobs).  The packed word holds the
// stencil start relative to the patch in x and y (+8, 5 bits each), the
// wrapped stencil start in z (bits 10-19) and which of the blob's patches this
// is in x (bits 20-22) and y (bits 23-24), the index of its weight windows.
__device__ __forceinline__ int col_rx(int pk) { return (pk & 31) - 8; }
__device__ __forceinline__ int col_ry(int pk) { return ((pk >> 5) & 31) - 8; }
__device__ __forceinline__ int col_sz(int pk) { return (pk >> 10) & 1023; }
__device__ __forceinline__ int col_dx(int pk) { return (pk >> 20) & 7; }
__device__ __forceinline__ int col_dy(int pk) { return (pk >> 23) & 3; }

// N consecutive words from a 16-byte aligned address (8-byte aligned for two
// FP32 words), as 16-byte loads
template <typename T, int N>
__device__ __forceinline__ void col_vload(const T* __restrict__ p, T* o) {
  if constexpr (sizeof(T) == 4 && N % 4 == 0) {
#pragma unroll
    for (int i = 0; i < N / 4; ++i) {
      const float4 v = reinterpret_cast<const float4*>(p)[i];
      o[4 * i] = v.x; o[4 * i + 1] = v.y; o[4 * i + 2] = v.z; o[4 * i + 3] = v.w;
    }
  } else if constexpr (sizeof(T) == 4 && N == 2) {
    const float2 v = *reinterpret_cast<const float2*>(p);
    o[0] = v.x; o[1] = v.y;
  } else {
    static_assert(sizeof(T) == 8 && N % 2 == 0, "vector load shape");
#pragma unroll
    for (int i = 0; i < N / 2; ++i) {
      const double2 v = reinterpret_cast<const double2*>(p)[i];
      o[2 * i] = v.x; o[2 * i + 1] = v.y;
    }
  }
}

template <typename T, int N>
__device__ __forceinline__ void col_vstore(T* p, const T* v) {
  static_assert(sizeof(T) == 4 && N % 4 == 0, "16-byte stores of FP32 words");
#pragma unroll
  for (int i = 0; i < N / 4; ++i)
    reinterpret_cast<float4*>(p)[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
}

// x, y and z weights of a list entry over its patch, from the blob's FP64
// windows (zero outside the stencil)
template <typename T, int CX>
__device__ __forceinline__ void col_win_load(const T* __restrict__ win, int slot, int pk, T* gx, T* gy, T* gz) {
  const T* wn = win + (size_t)slot * COL_WIN;
  col_vload<T, CX>(wn + CX * col_dx(pk), gx);
#pragma unroll
  for (int ii = CX; ii < 4; ++ii) gx[ii] = T(0);
  col_vload<T, 4>(wn + 12 + 4 * col_dy(pk), gy);
  col_vload<T, 8>(wn + 24, gz);
}

// FP32 weights of one list entry for its patch, evaluated from the blob's
// stencil offsets: x for the patch's CX columns (stencil node ii - rx), y for
// its 4 rows (node jj - ry; in a slip channel interpolation drops the rows at
// and above the upper wall, as the FP64 interpolation weights do), z for the
// ng nodes.  The same expression as the FP64 weights, in FP32 from the offset
// rounded to FP32 (relative error a few 1e-7).
template <int CX>
__device__ __forceinline__ void col_win_eval(const ColArgs& A, int slot, int pk, bool wall, int y0, float* gx,
                                             float* gy, float* gz) {
  const float4 o = A.woff[slot];
  const int rx = col_rx(pk), ry = col_ry(pk), ng = A.ng;
  auto w = [&](float x0, int k) {
    const float xx = fmaf((float)k, A.wdx, x0);
    return A.wa * exp2f(A.wc * (xx * xx));
  };
#pragma unroll
  for (int ii = 0; ii < 4; ++ii) {
    const int k = ii - rx;
    gx[ii] = (ii < CX && k >= 0 && k < ng) ? w(o.x, k) : 0.f;
  }
#pragma unroll
  for (int jj = 0; jj < COL_CY; ++jj) {
    const int k = jj - ry;
    gy[jj] = (k >= 0 && k < ng && !(wall && y0 + jj >= A.g.ny / 2)) ? w(o.y, k) : 0.f;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) gz[q] = q < ng ? w(o.z, q) : 0.f;
}

// ---------------------------------------------------------------------------
// launch shapes
// ---------------------------------------------------------------------------
constexpr int col_threads(int nz) { return (nz + 63) / 64 * 64; }  // one thread per z
constexpr int col_max(int a, int b) { return a > b ? a : b; }
constexpr int col_min(int a, int b) { return a < b ? a : b; }
// largest odd number <= v (LDS planes of an odd number of words are not
// aligned with one another in the 64 banks)
constexpr int col_odd_le(int v) { return (v % 2 == 1) ? v : v - 1; }

// points per thread of the z pass: the plan whose rounds pad the 3 cx 4 rows
// least, then the one with the smaller LDS scratch
constexpr int col_pm(int NZ, int CX) {
  const int nt = col_threads(NZ), rows = 3 * CX * COL_CY;
  int best = 16, best_pad = 1 << 30, best_lds = 1 << 30;
  for (int pm : {16, 8}) {
    const FftPlan pl = make_plan(NZ / 2, pm);
    if (!pl.ok || nt % pl.T != 0) continue;
    const int rr = nt / pl.T;
    const int pad = (rows + rr - 1) / rr * rr;
    const int lds = (NZ / 2 + 1) * (rr + 1);
    if (pad < best_pad || (pad == best_pad && lds < best_lds)) {
      best_pad = pad;
      best_lds = lds;
      best = pm;
    }
  }
  return best;
}

// LDS layout of the z pass scratch: element e of FFT column col at
// e ls + (e >> sh) sb + col.
// Chosen per shape with a model of the 64 LDS banks (a b64 access is served
// 32 lanes at a time, a b128 access 16 at a time) over every exchange of the
// transform, within the LDS the kernel has anyway at its occupancy; ls = 0
// keeps the plain layout (ls = rows + 1, no skew).  At nz = 256 in FP32 this
// takes the bank cycles of the exchanges of one round from 228 to 124 per
// wavefront (forward) and from 208 to 80 (inverse).
struct ColLayout {
  int ls, sh, sb;
};
constexpr ColLayout col_layout(int tbytes, int nz) {
  if (tbytes == 4) {
    switch (nz) {
      case 32: return {72, 2, 8};
      case 48: return {48, 2, 8};
      case 64: return {24, 3, 8};
      case 96: return {48, 1, 4};
      case 128: return {20, 3, 4};
      case 192: return {14, 1, 0};
      case 256: return {18, 4, 2};
      case 512: return {17, 5, 3};
      default: return {0, 0, 0};
    }
  }
  switch (nz) {
    case 32: return {40, 3, 8};
    case 48: return {20, 1, 0};
    case 64: return {20, 3, 4};
    case 96: return {18, 1, 0};
    case 128: return {18, 3, 2};
    case 256: return {17, 4, 1};
    default: return {0, 0, 0};  // nz = 512: the skew costs the inverse kernel a wave per SIMD
  }
}

// Lane pairing of the C2R input.  The point k of the inverse z transform needs
// X[k] and X[M - k]; with the first radix of the inverse plan equal to 2, the
// word X[M - k] of thread tcol's slot s is thread (TC - tcol)'s slot P - 1 - s,
// and for tcol = 0 it is the thread's own slot c2r_self_slot(s) (none for k = 0,
// whose partner is the Nyquist bin X[M]).  Checked at compile time over every
// thread; where it does not hold, both words are loaded.
constexpr int c2r_in_index(const FftPlan& pl, int tcol, int s) {
  return tcol + (s / pl.r[0]) * pl.T + (s % pl.r[0]) * (pl.N / pl.r[0]);
}
constexpr int c2r_self_slot(const FftPlan& pl, int s) {
  const int want = (pl.N - c2r_in_index(pl, 0, s)) % pl.N;
  if (want == 0) return -1;
  for (int q = 0; q < pl.P; ++q)
    if (c2r_in_index(pl, 0, q) == want) return q;
  return -2;
}
constexpr bool c2r_lane_pairs(const FftPlan& pl) {
  if (pl.T > 64 || 64 % pl.T != 0) return false;
  for (int tcol = 1; tcol < pl.T; ++tcol)
    for (int s = 0; s < pl.P; ++s)
      if (c2r_in_index(pl, pl.T - tcol, pl.P - 1 - s) != pl.N - c2r_in_index(pl, tcol, s)) return false;
  for (int s = 0; s < pl.P; ++s)
    if (c2r_self_slot(pl, s) == -2 || (c2r_self_slot(pl, s) == -1) != (c2r_in_index(pl, 0, s) == 0)) return false;
  return true;
}

// Lane pairing of the R2C untangle.  The forward z transform leaves bin
// k = r2c_out_index(tcol, s) in slot s of thread tcol; the untangle of k needs
// Z[M - k], which is thread (TC - tcol)'s slot P - 1 - s, or for tcol = 0 the
// thread's own slot r2c_self_slot(s).  Checked at compile time over every
// thread; where it does not hold, the bins go through LDS.
constexpr int r2c_out_index(const FftPlan& pl, int tcol, int s) {
  const int rl = pl.r[pl.ns - 1];
  return tcol + (s / rl) * pl.T + (s % rl) * (pl.N / rl);
}
constexpr int r2c_self_slot(const FftPlan& pl, int s) {
  const int want = (pl.N - r2c_out_index(pl, 0, s)) % pl.N;
  for (int q = 0; q < pl.P; ++q)
    if (r2c_out_index(pl, 0, q) == want) return q;
  return -1;
}
constexpr bool r2c_lane_pairs(const FftPlan& pl) {
  if (!pl.ok || pl.T > 64 || 64 % pl.T != 0) return false;
  for (int tcol = 1; tcol < pl.T; ++tcol)
    for (int s = 0; s < pl.P; ++s)
      if (r2c_out_index(pl, pl.T - tcol, pl.P - 1 - s) != pl.N - r2c_out_index(pl, tcol, s)) return false;
  for (int s = 0; s < pl.P; ++s)
    if (r2c_self_slot(pl, s) < 0) return false;
  return true;
}

// waves per SIMD the interpolation kernel is compiled for: at nz = 512 a
// workgroup is 8 waves, and the 140 registers FP32 takes unconstrained leave
// one workgroup per CU
constexpr int col_interp_waves(int nz, int tbytes) { return (nz >= 512 && tbytes == 4) ? 4 : 1; }

// z-major row stride (words) for rr rows of tbytes-byte words: a 16-byte read
// covers 4 of the 64 LDS banks, and 16 lanes reading rows at 16 consecutive z
// hit 16 distinct groups when the stride in 16-byte units is odd
constexpr int col_ldu(int rr, int tbytes) {
  const int q = tbytes == 4 ? 8 : 4, r = tbytes == 4 ? 4 : 2;
  int v = rr;
  while (v % q != r) ++v;
  return v;
}

// patch x extent per nz by default: FP64 rows of a 4 x 4 patch exceed the LDS at nz = 512
constexpr int col_cx_for(int nz) { return nz <= 256 ? 4 : 2; }

template <typename T, int NZ, int CX, int PM>
struct ColumnShape {
  static constexpr int NT = col_threads(NZ);
  static constexpr int M = NZ / 2;                // complex FFT length of the z pass
  static constexpr FftPlan PL = make_plan(M, PM);
  static constexpr bool PLAN_OK = PL.ok && PL.T > 0 && NT % PL.T == 0 && (PM == 8 || PM == 16);
  static constexpr int P = PL.P, TC = PLAN_OK ? PL.T : 1;
  static constexpr int RR = NT / TC;              // rows transformed at once (one round)
  static constexpr int NROW = 3 * CX * COL_CY;    // rows of a patch (3 components)
  static constexpr int NRND = (NROW + RR - 1) / RR;
  // LDS for staging the blob lists (about 38 KB per 256 threads in FP32, 52 KB
  // in FP64: four or two workgroups per CU at nz = 256)
  static constexpr int STAGE = col_min((sizeof(T) == 4 ? 38912 : 53248) / 256 * NT, 65536);
  static constexpr int plain_words(int n) { return col_max((M - 1) * (n * RR + 1) + n * RR, n * RR * (M + 1)); }
  // all rounds transformed together (one set of barriers) when their scratch
  // fits in the staging LDS (FP32 only), else one round at a time: a larger
  // scratch would cost workgroups per CU (at nz = 256 in FP32, three rounds at
  // once held three workgroups per CU instead of four and took 15% longer)
  static constexpr int NCB = (sizeof(T) == 4 && plain_words(NRND) * 8 <= STAGE) ? NRND : 1;
  // the bank-skewed layouts are fitted to the default patch width and plan;
  // other shapes take the plain layout
  static constexpr bool DEFAULT_SHAPE = CX == col_cx_for(NZ) && PM == col_pm(NZ, CX);
  static constexpr ColLayout LAY = DEFAULT_SHAPE ? col_layout((int)sizeof(T), NZ) : ColLayout{0, 0, 0};
  // FFT scratch: element e of row r at e LS + (e >> SSH) SB + r
  static constexpr int LS = LAY.ls > 0 ? LAY.ls : NCB * RR + 1;
  static constexpr int SSH = LAY.ls > 0 ? LAY.sh : 0, SB = LAY.ls > 0 ? LAY.sb : 0;
  static constexpr int RS = M + 1;                // row-major rows: complex words per row
  static constexpr int FFT_WORDS =
      col_max((M - 1) * LS + ((M - 1) >> SSH) * SB + NCB * RR, NCB * RR * RS);
  static constexpr int FFT_BYTES = FFT_WORDS * 2 * (int)sizeof(T);
  static __device__ __forceinline__ int slot(int e, int col) { return fft_slot<SSH, SB>(e, LS, col); }
  static __device__ __forceinline__ int row0(int r) { return r * RS; }
  // staging of the blob lists: budget per workgroup, then planes of CAP entries
  static constexpr int BUDGET = col_max(STAGE, FFT_BYTES);
  // staged entry: gx 4, gy 4, F 3, stencil-start z, gz 8.  In FP32 one 80-byte
  // record per entry (four 16-byte reads per (entry, z) step, and records 16
  // entries apart are the first to share banks); in FP64 planes of single words
  // (entry i of field f at f S_CAP + i; 160-byte records shared banks every
  // 8 entries and were 3% slower)
  static constexpr bool S_AOS = sizeof(T) == 4;
  static constexpr int S_ENT = 20 * (int)sizeof(T) - (S_AOS ? 0 : (int)sizeof(T) - 4);
  static constexpr int S_CAP = col_odd_le(BUDGET / S_ENT);
  static constexpr int S_LDS = col_max(FFT_BYTES, S_CAP * S_ENT);
  // interpolation: the rows of a round kept z-major, E_LDU words per z, a row
  // stride whose 16-byte reads by 16 lanes at neighbouring z fall on distinct
  // banks; E_KR list entries per thread per pass over the transform
  static constexpr int E_LDU = col_ldu(RR, (int)sizeof(T));
  // row z starts at z E_LDU + ((z >> 4) & 1) E_PAD: without the shift, the
  // transform's writes of rows z and z + 16 fell on the same banks
  static constexpr int E_PAD = E_LDU - RR;
  static __device__ __forceinline__ int urow(int z) { return z * E_LDU + ((z >> 4) & 1) * E_PAD; }
  static constexpr int E_KR = 2;
  static constexpr int E_LDS = col_max(FFT_BYTES, (NZ * E_LDU + E_PAD) * (int)sizeof(T));
  // a shape the kernels can be built for: a z plan that tiles the threads in
  // whole y columns of rows, FFT columns that do not overlap, room to stage 15
  // entries, and the LDS of a CU
  static constexpr bool ok = PLAN_OK && RR % COL_CY == 0 && LS >= NCB * RR && S_CAP >= 15 &&
                             S_LDS + 2 * M * 2 * (int)sizeof(T) <= 163840 &&
                             E_LDS + 2 * M * 2 * (int)sizeof(T) <= 163840;
};

// the range [a1, b1) u [a2, b2) of the patch list holds exactly the blobs
// whose stencil covers z (stencil start in [z - ng + 1, z] modulo nz).  Written
// with selects: as branches the compiler kept the four words in scratch.
template <int NZ>
__device__ __forceinline__ void col_ranges(const int* __restrict__ zs, int z, int ng, int& a1, int& b1, int& a2, int& b2) {
  const int lo = z - ng + 1;
  const bool in = z < NZ, wrap = lo < 0;
  const int za = wrap ? NZ + lo : lo, zb = wrap ? NZ : z + 1;
  a1 = in ? zs[za] : 0;
  b1 = in ? zs[zb] : 0;
  a2 = in && wrap ? zs[0] : 0;
  b2 = in && wrap ? zs[z + 1] : 0;
}

// ---------------------------------------------------------------------------
// spreading + forward z R2C
//
// LDS while accumulating: planes of S_CAP entries (entry i of field f at
// f S_CAP + i), so that the lanes of a wavefront, which read neighbouring
// entries of the list, read neighbouring banks.  Then, per round of RR rows:
// the rows go row-major through LDS into registers, the FFT runs with a
// padded scratch stride, and the untangled k-space rows are stored.  In FP32
// the untangle takes each bin's partner Z[M - k] from the lane that holds it
// (no LDS round trip; spreading 0.125 -> 0.117 ms at 256^3); in FP64 the extra
// registers would spill and the bins go through LDS.
// ---------------------------------------------------------------------------
template <typename T, int NZ, int CX, int PM>
__global__ void __launch_bounds__(col_threads(NZ))
    col_spread_kernel(ColArgs A, const double* __restrict__ F, FieldPtrs<T> fp, const c2<T>* __restrict__ tw,
                      const c2<T>* __restrict__ twz) {
  using SH = ColumnShape<T, NZ, CX, PM>;
  static_assert(SH::ok, "column shape");
  using V = c2<T>;
  constexpr int NT = SH::NT, M = SH::M, P = SH::P, TC = SH::TC, RR = SH::RR, NROW = SH::NROW, NRND = SH::NRND;
  constexpr int CY = COL_CY, CAP = SH::S_CAP, LS = SH::LS, NCB = SH::NCB;
  __shared__ __align__(16) unsigned char smem[SH::S_LDS];
  __shared__ V s_tw[M], s_twz[M];  // untangle twiddles W^k, k < M
  if (fp.gate && *fp.gate) return;
  const GridGeom& g = A.g;
  const int npatch = A.npx * A.npy;
  const int p = col_xcd(blockIdx.x, npatch);
  if (p >= npatch) return;
  const int px = p / A.npy, py = p - (p / A.npy) * A.npy;
  const int x0 = CX * px, y0 = CY * py;
  // slip channel: rows above ny/2 are images, made from their mirror rows when
  // they are read, so a patch wholly above ny/2 has nothing to write
  if (g.channel && y0 > g.ny / 2) return;
  const int t = threadIdx.x;
  const int ng = A.ng;
  for (int i = t; i < M; i += NT) {
    s_tw[i] = tw[i];
    s_twz[i] = twz[i];
  }
  __syncthreads();
  // this patch's z prefix table, read from HBM (a few words per thread)
  const int* zs = A.zst + (size_t)p * (NZ + 1);
  const int L0 = zs[0], L1 = zs[NZ];
  int a1, b1, a2, b2;
  col_ranges<NZ>(zs, t, ng, a1, b1, a2, b2);

  T* sw = reinterpret_cast<T*>(smem);  // records (gx 0-3, gy 4-7, F 8-10, z 11, gz 12-19) or planes (gz 11-18)
  int* ssz = reinterpret_cast<int*>(sw + 19 * CAP);  // planes: stencil-start z
  const double* EW = A.ew64;
  T acc[NROW];  // row (c CX + ii) CY + jj of this thread's z
#pragma unroll
  for (int r = 0; r < NROW; ++r) acc[r] = T(0);

  for (int c0 = L0; c0 < L1; c0 += CAP) {
    const int cnt = min(CAP, L1 - c0);
    __syncthreads();
    // two entries per thread per step, all of their loads in flight together
    for (int i0 = 0; i0 < cnt; i0 += 2 * NT) {
      int4 e[2];
      T gx[2][4], gy[2][4], gz[2][8], f[2][3];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int i = i0 + t + u * NT;
        e[u] = i < cnt ? A.ent[c0 + i] : make_int4(0, 0, 8 | (8 << 5), 0);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        if constexpr (sizeof(T) == 8) {
          // the entry's own copy of its weights, read in list order
          const int i = i0 + t + u * NT;
          const T* ew = EW + (size_t)(c0 + (i < cnt ? i : 0)) * COL_EWIN;
          col_vload<T, 4>(ew, gx[u]);
          col_vload<T, 4>(ew + 4, gy[u]);
          col_vload<T, 8>(ew + 8, gz[u]);
        } else {
          col_win_eval<CX>(A, e[u].x, e[u].z, false, y0, gx[u], gy[u], gz[u]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) f[u][c] = T(F[3 * (size_t)e[u].y + c]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int i = i0 + t + u * NT;
        if (i < cnt) {
          if constexpr (SH::S_AOS) {
            T* r = sw + (size_t)i * 20;
            const T fz[4] = {f[u][0], f[u][1], f[u][2], __int_as_float(col_sz(e[u].z))};
            col_vstore<T, 4>(r, gx[u]);
            col_vstore<T, 4>(r + 4, gy[u]);
            col_vstore<T, 4>(r + 8, fz);
            col_vstore<T, 8>(r + 12, gz[u]);
          } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              sw[k * CAP + i] = gx[u][k];
              sw[(4 + k) * CAP + i] = gy[u][k];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) sw[(8 + c) * CAP + i] = f[u][c];
#pragma unroll
            for (int k = 0; k < 8; ++k) sw[(11 + k) * CAP + i] = gz[u][k];
            ssz[i] = col_sz(e[u].z);
          }
        }
      }
    }
    __syncthreads();
    const int e1a = max(a1, c0), e1b = min(b1, c0 + cnt);
    const int e2a = max(a2, c0), e2b = min(b2, c0 + cnt);
    const int n1 = max(0, e1b - e1a), n2 = max(0, e2b - e2a);
    for (int k = 0; k < n1 + n2; ++k) {
      const int i = (k < n1 ? e1a + k : e2a + (k - n1)) - c0;
      T gx[4], gy[4], fz[4], w;
      if constexpr (SH::S_AOS) {
        const T* r = sw + (size_t)i * 20;
        col_vload<T, 4>(r + 8, fz);
        int o = t - __float_as_int(fz[3]);
        if (o < 0) o += NZ;
        w = r[12 + o];
        col_vload<T, 4>(r, gx);
        col_vload<T, 4>(r + 4, gy);
      } else {
        int o = t - ssz[i];
        if (o < 0) o += NZ;
        w = sw[(11 + o) * CAP + i];
#pragma unroll
        for (int q = 0; q < CX; ++q) gx[q] = sw[q * CAP + i];
#pragma unroll
        for (int q = 0; q < CY; ++q) gy[q] = sw[(4 + q) * CAP + i];
#pragma unroll
        for (int c = 0; c < 3; ++c) fz[c] = sw[(8 + c) * CAP + i];
      }
      const T wf0 = w * fz[0], wf1 = w * fz[1], wf2 = w * fz[2];
#pragma unroll
      for (int jj = 0; jj < CY; ++jj) {
        const T wy0 = wf0 * gy[jj], wy1 = wf1 * gy[jj], wy2 = wf2 * gy[jj];
#pragma unroll
        for (int ii = 0; ii < CX; ++ii) {
          acc[(0 * CX + ii) * CY + jj] += wy0 * gx[ii];
          acc[(1 * CX + ii) * CY + jj] += wy1 * gx[ii];
          acc[(2 * CX + ii) * CY + jj] += wy2 * gx[ii];
        }
      }
    }
  }
  V* area = reinterpret_cast<V*>(smem);
  T* rowsT = reinterpret_cast<T*>(smem);
  const int rloc = t / TC, tcol = t - (t / TC) * TC;
  constexpr FftPlan FPL = make_plan(M, PM);
  constexpr bool OPAIRS = sizeof(T) == 4 && TC > 1 && r2c_lane_pairs(FPL);
  const int partner = (t & 63) - tcol + ((TC - tcol) & (TC - 1));  // lane of thread TC - tcol of this row
#pragma unroll
  for (int qg = 0; qg < NRND; qg += NCB) {
    __syncthreads();  // staging (first group) or the previous group's scratch is free
    if (t < NZ) {
#pragma unroll
      for (int r = 0; r < NCB * RR; ++r)
        if (qg * RR + r < NROW) rowsT[2 * SH::row0(r) + t] = acc[qg * RR + r];
    }
    __syncthreads();
    V v[NCB][P];
#pragma unroll
    for (int c = 0; c < NCB; ++c)
#pragma unroll
      for (int s = 0; s < P; ++s) v[c][s] = area[SH::row0(rloc + c * RR) + fft_in_index<M, false, PM>(tcol, s)];
    __syncthreads();
    fft_columns<T, M, -1, false, PM, NCB, SH::SSH, SH::SB>(v, area, LS, rloc, RR, tcol, s_tw);
    if constexpr (OPAIRS) {
      // untangle straight from registers: bin k's partner Z[M - k] is held by
      // the lane of thread TC - tcol (slot P - 1 - s), or by thread 0 itself
#pragma unroll
      for (int c = 0; c < NCB; ++c) {
        const int row = (qg + c) * RR + rloc;
        const bool valid = row < NROW;
        const int cc = row / (CX * CY), rem = row - (row / (CX * CY)) * (CX * CY);
        const int ii = rem / CY, jj = rem - (rem / CY) * CY;
        V* out = fp.p[valid ? cc : 0] + ((size_t)(x0 + (valid ? ii : 0)) * g.ny + (y0 + (valid ? jj : 0))) * g.cpitch;
#pragma unroll
        for (int s = 0; s < P; ++s) {
          const int k = fft_out_index<M, false, PM>(tcol, s);
          const V sh = mk<T>(__shfl(v[c][P - 1 - s].x, partner), __shfl(v[c][P - 1 - s].y, partner));
          const int self = r2c_self_slot(FPL, s);  // a constant once unrolled
          const V pz = tcol == 0 ? v[c][self >= 0 ? self : 0] : sh;
          // X[k] = e + W^k o with e = (Z[k] + conj Z[M-k]) / 2, o = (Z[k] - conj Z[M-k]) / 2i
          const V zm = cconj(pz);
          const V e = cscale(cadd(v[c][s], zm), T(0.5));
          const V d = csub(v[c][s], zm);
          const V o = mk<T>(T(0.5) * d.y, T(-0.5) * d.x);
          const V tt = cmul(s_twz[k], o);
          if (valid) {
            out[k] = cadd(e, tt);
            if (k == 0) out[M] = cconj(csub(e, tt));
          }
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < NCB; ++c)
#pragma unroll
        for (int s = 0; s < P; ++s) area[SH::slot(fft_out_index<M, false, PM>(tcol, s), rloc + c * RR)] = v[c][s];
      __syncthreads();
#pragma unroll
      for (int c = 0; c < NCB; ++c) {
        const int row = (qg + c) * RR + rloc;
        if (row < NROW) {
          const int cc = row / (CX * CY), rem = row - (row / (CX * CY)) * (CX * CY);
          const int ii = rem / CY, jj = rem - (rem / CY) * CY;
          V* out = fp.p[cc] + ((size_t)(x0 + ii) * g.ny + (y0 + jj)) * g.cpitch;
          // untangle, one bin pair (k, M - k) per step
          const int zc = rloc + c * RR;
          for (int k = tcol; k <= M / 2; k += TC) {
            V xk, xmk;
            r2c_untangle_pair<T>(area[SH::slot(k, zc)], area[SH::slot(k == 0 ? 0 : M - k, zc)], s_twz[k], xk, xmk);
            out[k] = xk;
            if (M - k != k) out[M - k] = xmk;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// inverse z C2R + interpolation partial sums, one thread per list entry
//
// The entries of the patch list are taken E_KR per thread, their x, y and z
// weights held in registers (loaded at the start, so the first transform
// hides their latency).  Per round, the RR rows' k-space words are read
// straight from HBM (each C2R point k needs X[k] and X[M - k]; the second
// word comes from the partner lane where the plan allows), transformed
// through the padded LDS scratch and left in LDS z-major (row r of node z at
// urow(z) + r, rows z and z + 16 offset by the padding).  Each thread then reads, for every stencil z of each of its
// entries, the round's rows at that z (16-byte reads) and adds
// gz sum_x gx sum_y gy u to the entry's three sums, while the next round's
// k-space words are in flight.  Every lane does the same work whatever the z
// density of the patch (one thread per z ran each wavefront as long as its
// busiest lane: 24 entry steps against a mean of 10 at 256^3), each entry's
// sum runs over its stencil in a fixed order, and no per-(entry, z) slots are
// staged (they filled the LDS and split a patch's list into two passes).  A
// patch with more than E_KR NT entries repeats the transform per pass.
// FP32 at 256^3: 0.138 -> 0.113 ms; FP64 0.282 -> 0.240 ms.
// ---------------------------------------------------------------------------
template <typename T, int NZ, int CX, int PM>
__global__ void __launch_bounds__(col_threads(NZ), col_interp_waves(NZ, (int)sizeof(T)))
    col_interp_kernel(ColArgs A, FieldPtrs<T> fp, const c2<T>* __restrict__ tw, const c2<T>* __restrict__ twz,
                      T* __restrict__ part) {
  using SH = ColumnShape<T, NZ, CX, PM>;
  static_assert(SH::ok, "column shape");
  using V = c2<T>;
  constexpr int NT = SH::NT, M = SH::M, P = SH::P, TC = SH::TC, RR = SH::RR, NROW = SH::NROW, NRND = SH::NRND;
  constexpr int CY = COL_CY, LS = SH::LS, KR = SH::E_KR;
  static_assert(RR % CY == 0, "rows of a round come in whole y columns");
  __shared__ __align__(16) unsigned char smem[SH::E_LDS];
  __shared__ V s_tw[M], s_twz[M];  // FFT twiddles, C2R twiddles W^k
  if (fp.gate && *fp.gate) return;
  const GridGeom& g = A.g;
  const int npatch = A.npx * A.npy;
  const int p = col_xcd(blockIdx.x, npatch);
  if (p >= npatch) return;
  const int px = p / A.npy, py = p - (p / A.npy) * A.npy;
  const int x0 = CX * px, y0 = CY * py;
  const int t = threadIdx.x;
  const int ng = A.ng;
  for (int i = t; i < M; i += NT) {
    s_tw[i] = tw[i];
    s_twz[i] = twz[i];
  }
  __syncthreads();  // the C2R twiddles are read before the first transform's barriers
  V* area = reinterpret_cast<V*>(smem);
  T* uc = reinterpret_cast<T*>(smem);
  const int rloc = t / TC, tcol = t - (t / TC) * TC;
  constexpr FftPlan IPL = reverse_plan(make_plan(M, PM));
  constexpr bool PAIRS = c2r_lane_pairs(IPL);
  const int partner = (t & 63) - tcol + ((TC - tcol) & (TC - 1));  // lane of thread TC - tcol of this row
  const int* zs = A.zst + (size_t)p * (NZ + 1);
  const int L0 = zs[0], L1 = zs[NZ];
  const double* W = A.w64;
  // C2R input of round qg's rows, straight from HBM: each point k needs X[k]
  // and X[M - k] (imaginary parts of k = 0 and M ignored)
  V xk[P], xn, xm_nopair[PAIRS ? 1 : P];
  auto load_round = [&](int qg) {
    const int row = qg * RR + rloc;
    const bool valid = row < NROW;
    const int cc = row / (CX * CY), rem = row - (row / (CX * CY)) * (CX * CY);
    const int ii = rem / CY, jj = rem - (rem / CY) * CY;
    const V* in = fp.p[valid ? cc : 0] + ((size_t)(x0 + (valid ? ii : 0)) * g.ny + (y0 + (valid ? jj : 0))) * g.cpitch;
#pragma unroll
    for (int s = 0; s < P; ++s) xk[s] = valid ? in[fft_in_index<M, true, PM>(tcol, s)] : mk<T>(T(0), T(0));
    xn = mk<T>(T(0), T(0));
    if (PAIRS && valid && tcol == 0) xn = in[M];
    if constexpr (!PAIRS) {
#pragma unroll
      for (int s = 0; s < P; ++s) xm_nopair[s] = valid ? in[M - fft_in_index<M, true, PM>(tcol, s)] : mk<T>(T(0), T(0));
    }
  };
  for (int base = L0; base < L1; base += KR * NT) {
    T gx[KR][4], gy[KR][4], gz[KR][8], acc[KR][3];
    int z0[KR], oid[KR];
    bool val[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      const int i = base + k * NT + t;
      val[k] = i < L1;
      const int4 e = val[k] ? A.ent[i] : make_int4(0, 0, 8 | (8 << 5), 0);
      if constexpr (sizeof(T) == 8) col_win_load<double, CX>(W, e.x, e.z, gx[k], gy[k], gz[k]);
      else col_win_eval<CX>(A, e.x, e.z, g.channel != 0, y0, gx[k], gy[k], gz[k]);
      z0[k] = col_sz(e.z);
      oid[k] = e.w;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[k][c] = T(0);
    }
#pragma unroll
    for (int qg = 0; qg < NRND; ++qg) {
      if (qg == 0) load_round(0);
      V v[1][P];
#pragma unroll
      for (int s = 0; s < P; ++s) {
        const int k = fft_in_index<M, true, PM>(tcol, s);
        V xm;
        if constexpr (PAIRS) {
          const V sh = mk<T>(__shfl(xk[P - 1 - s].x, partner), __shfl(xk[P - 1 - s].y, partner));
          const int self = c2r_self_slot(IPL, s);
          const V own = self >= 0 ? xk[self >= 0 ? self : 0] : xn;
          xm = tcol == 0 ? own : sh;
        } else {
          xm = xm_nopair[s];
        }
        V xs = xk[s];
        if (k == 0) { xs.y = T(0); xm.y = T(0); }
        const V sum = cadd(xs, cconj(xm));
        const V dif = csub(xs, cconj(xm));
        const V rot = cmulc(dif, s_twz[k]);
        v[0][s] = cadd(sum, mk<T>(-rot.y, rot.x));
      }
      __syncthreads();  // the previous round's rows have been read
      fft_columns<T, M, +1, true, PM, 1, SH::SSH, SH::SB>(v, area, LS, rloc, RR, tcol, s_tw);
      // fft_columns returns after the barrier that follows its last LDS reads
#pragma unroll
      for (int s = 0; s < P; ++s) {
        const int m = fft_out_index<M, true, PM>(tcol, s);
        uc[SH::urow(2 * m) + rloc] = v[0][s].x;
        uc[SH::urow(2 * m + 1) + rloc] = v[0][s].y;
      }
      __syncthreads();
      // the next round's input is loaded while this round's rows are used
      if (qg + 1 < NRND) load_round(qg + 1);
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        if (!val[k]) continue;
#pragma unroll
        for (int o = 0; o < 8; ++o) {
          if (o >= ng) break;
          int z = z0[k] + o;
          if (z >= NZ) z -= NZ;
          const T* ur = uc + SH::urow(z);
          T a[3] = {T(0), T(0), T(0)};
#pragma unroll
          for (int g4 = 0; g4 < RR / CY; ++g4) {
            const int R = qg * RR + g4 * CY;  // a constant once unrolled
            if (R < NROW) {
              const int c = R / (CX * CY), ii = (R % (CX * CY)) / CY;
              T u4[CY];
#pragma unroll
              for (int j = 0; j < CY; ++j) u4[j] = ur[g4 * CY + j];
              T b = gy[k][0] * u4[0];
#pragma unroll
              for (int j = 1; j < CY; ++j) b += gy[k][j] * u4[j];
              a[c] += gx[k][ii] * b;
            }
          }
          // components with rows in this round (constants once unrolled)
          const int clo = (qg * RR) / (CX * CY), chi = col_min(NROW - 1, qg * RR + RR - 1) / (CX * CY);
#pragma unroll
          for (int c = 0; c < 3; ++c)
            if (c >= clo && c <= chi) acc[k][c] += gz[k][o] * a[c];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KR; ++k)
      if (val[k]) {
        T* out = part + (size_t)oid[k] * 3;
        out[0] = acc[k][0];
        out[1] = acc[k][1];
        out[2] = acc[k][2];
      }
  }
}

// sum of a blob's patch partials in a fixed order, then dx^3 J u - K V.  The
// loads of every slot a blob can have (XM x 3 for patches CX wide, ng <= 8)
// are issued before the record that says which of them are its own arrives;
// the others are read but not added.
template <typename T, int CX>
__global__ void __launch_bounds__(256) col_reduce_kernel(GridGeom g, int ng, int nb, const int4* __restrict__ rec,
                                                         const T* __restrict__ part, double* V, KTerm kt, double dx3,
                                                         const int* gate) {
  constexpr int XM = (CX - 1 + 7) / CX + 1, YM = (COL_CY - 1 + 7) / COL_CY + 1;
  static_assert(XM * YM <= COL_SLOTS_MAX, "partial-sum slots");
  if (gate && *gate) return;
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nb) return;
  T pv[XM][YM][3];
#pragma unroll
  for (int dx = 0; dx < XM; ++dx)
#pragma unroll
    for (int dy = 0; dy < YM; ++dy) {
      const T* pp = part + ((size_t)(dx * 3 + dy) * nb + s) * 3;
      pv[dx][dy][0] = pp[0];
      pv[dx][dy][1] = pp[1];
      pv[dx][dy][2] = pp[2];
    }
  const int omin = 1 - ng / 2;
  const int4 r = rec[s];
  const int sxw = col_wrapi(r.x + omin, g.nx), syw = col_wrapi(r.y + omin, g.ny);
  const int nxp = (sxw % CX + ng - 1) / CX + 1, nyp = (syw % COL_CY + ng - 1) / COL_CY + 1;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
  for (int dx = 0; dx < XM; ++dx)
#pragma unroll
    for (int dy = 0; dy < YM; ++dy)
      if (dx < nxp && dy < nyp) {
        v0 += (double)pv[dx][dy][0];
        v1 += (double)pv[dx][dy][1];
        v2 += (double)pv[dx][dy][2];
      }
  store_blob_velocity(V, kt, dx3, r.w, v0, v1, v2);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define FCM_COL_NZ(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(512)

namespace {

ColArgs col_args(const GridGeom& g, const BlobData& bd, bool interp) {
  return ColArgs{g,          bd.ng,      g.nx / bd.col_cx, g.ny / COL_CY, bd.col_ent, bd.col_zst,
                 interp ? bd.icwin64 : bd.cwin64, interp ? nullptr : bd.ewin64,
                 bd.woff, bd.wa32, bd.wc32, (float)g.dx};
}

// points per thread of the z pass: the tuned request (8 or 16) or the default
// plan of the shape; 0 if the shape cannot be built
template <typename T, int NZ, int CX>
constexpr int col_pm_pick(int req) {
  constexpr int PD = col_pm(NZ, CX);
  if (req == 8 && ColumnShape<T, NZ, CX, 8>::ok) return 8;
  if (req == 16 && ColumnShape<T, NZ, CX, 16>::ok) return 16;
  if (req != 0 && req != PD) return 0;
  return ColumnShape<T, NZ, CX, PD>::ok ? PD : 0;
}

template <typename T, int NZ, int CX, int PM>
hipError_t launch_spread(const ColArgs& A, const double* F, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw,
                         hipStream_t st) {
  if constexpr (!ColumnShape<T, NZ, CX, PM>::ok) {
    return hipErrorInvalidValue;
  } else {
    using SH = ColumnShape<T, NZ, CX, PM>;
    const int npatch = A.npx * A.npy;
    hipLaunchKernelGGL((col_spread_kernel<T, NZ, CX, PM>), dim3(8 * ((npatch + 7) / 8)), dim3(SH::NT), 0, st, A, F,
                       fp, tw.z, tw.zu);
    return hipGetLastError();
  }
}

template <typename T, int NZ, int CX, int PM>
hipError_t launch_interp(const ColArgs& A, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw, T* part, hipStream_t st) {
  if constexpr (!ColumnShape<T, NZ, CX, PM>::ok) {
    return hipErrorInvalidValue;
  } else {
    using SH = ColumnShape<T, NZ, CX, PM>;
    const int npatch = A.npx * A.npy;
    hipLaunchKernelGGL((col_interp_kernel<T, NZ, CX, PM>), dim3(8 * ((npatch + 7) / 8)), dim3(SH::NT), 0, st, A, fp,
                       tw.z, tw.zu, part);
    return hipGetLastError();
  }
}

// the kernel pair of (nz, cx, pm) for one side of the product
template <typename T, int NZ>
hipError_t col_launch(bool interp, const ColArgs& A, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                      const TwiddleSet<T>& tw, T* part, hipStream_t st) {
  const int req = bd.col_pm[sizeof(T) == 4 ? 1 : 0];
  auto go = [&](auto cxc) -> hipError_t {
    constexpr int CX = decltype(cxc)::value;
    switch (col_pm_pick<T, NZ, CX>(req)) {
      case 8:
        return interp ? launch_interp<T, NZ, CX, 8>(A, fp, tw, part, st) : launch_spread<T, NZ, CX, 8>(A, F, fp, tw, st);
      case 16:
        return interp ? launch_interp<T, NZ, CX, 16>(A, fp, tw, part, st)
                      : launch_spread<T, NZ, CX, 16>(A, F, fp, tw, st);
      default:
        return hipErrorInvalidValue;
    }
  };
  if (bd.col_cx == 4) return go(std::integral_constant<int, 4>{});
  if (bd.col_cx == 2) return go(std::integral_constant<int, 2>{});
  return hipErrorInvalidValue;
}

}  // namespace

template <typename T>
hipError_t spread_zfwd(const GridGeom& g, const BlobData& bd, const double* F, const FieldPtrs<T>& fp,
                       const TwiddleSet<T>& tw, hipStream_t st, bool images) {
  if (bd.col_cx <= 0) return hipErrorInvalidValue;
  const ColArgs A = col_args(g, bd, false);
  hipError_t e = hipErrorInvalidValue;
  switch (g.nz) {
#define FCM_COL_SPREAD(v) \
  case v: e = col_launch<T, v>(false, A, bd, F, fp, tw, nullptr, st); break;
    FCM_COL_NZ(FCM_COL_SPREAD)
#undef FCM_COL_SPREAD
    default:
      break;
  }
  return e != hipSuccess || !images ? e : channel_images<T>(g, fp, st);
}

template <typename T>
hipError_t zinv_interp(const GridGeom& g, const BlobData& bd, const FieldPtrs<T>& fp, const TwiddleSet<T>& tw,
                       double* V, const KTerm& kt, hipStream_t st, hipEvent_t between, int interp_reps) {
  if (bd.col_cx <= 0) return hipErrorInvalidValue;
  if (bd.nb == 0) return hipSuccess;
  const ColArgs A = col_args(g, bd, true);
  T* part = static_cast<T*>(bd.col_part);
  hipError_t e = hipErrorInvalidValue;
  for (int r = 0; r < interp_reps; ++r) {
    switch (g.nz) {
#define FCM_COL_INTERP(v) \
  case v: e = col_launch<T, v>(true, A, bd, nullptr, fp, tw, part, st); break;
      FCM_COL_NZ(FCM_COL_INTERP)
#undef FCM_COL_INTERP
      default:
        break;
    }
    if (e != hipSuccess) return e;
  }
  if (between && (e = hipEventRecord(between, st)) != hipSuccess) return e;
  const dim3 grid((bd.nb + 255) / 256);
  if (bd.col_cx == 4)
    hipLaunchKernelGGL((col_reduce_kernel<T, 4>), grid, dim3(256), 0, st, g, bd.ng, bd.nb, bd.rec, (const T*)part, V,
                       kt, g.dx * g.dx * g.dx, fp.gate);
  else
    hipLaunchKernelGGL((col_reduce_kernel<T, 2>), grid, dim3(256), 0, st, g, bd.ng, bd.nb, bd.rec, (const T*)part, V,
                       kt, g.dx * g.dx * g.dx, fp.gate);
  return hipGetLastError();
}

}  // namespace fcm
