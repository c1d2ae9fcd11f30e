// fcm_fft.hip -- 3-D real FFT passes and the Fourier-space Stokes solve.
//
// Grid storage: each field is an array of complex words [nx][ny][cpitch]
// (cpitch >= nz/2+1, rounded up so that every row starts on a 128-byte
// boundary).  In real space row (x,y) holds nz reals in its first nz real
// slots.  A 3-D transform is five passes over HBM per field:
//   z: R2C of every row, as an nz/2 complex FFT plus an untangling step
//   y: C2C down the columns of every x-plane
//   x: C2C down the columns of every y-plane; for the Stokes solve this pass
//      is fused with the projection and the inverse x transform (3 fields
//      in, 3 fields out, one read and one write each)
// followed by inverse y and inverse z (C2R).  The arithmetic follows the
// reference's FFTW usage (fft_solver.cpp:21-42): unnormalised transforms, and
// on the inverse the imaginary parts of the kz = 0 and kz = nz/2 planes are
// ignored (FFTW's halfcomplex format stores none).
#include "fcm_internal.hpp"

namespace fcm {

// ---------------------------------------------------------------------------
// block shapes
// ---------------------------------------------------------------------------
template <typename T, int N, int PM = 16, int TMAX = 256> struct ColShape {
  static constexpr FftPlan PL = make_plan(N, PM);
  static constexpr int TC = PL.T;
  // columns per block: at most 16, at most TMAX threads, at most 64 KiB LDS
  static constexpr int C0 = TMAX / TC < 16 ? TMAX / TC : 16;
  static constexpr int CL = (65536 / (N * (int)sizeof(c2<T>))) < C0 ? (65536 / (N * (int)sizeof(c2<T>))) : C0;
  static constexpr int C = CL < 1 ? 1 : CL;
  static constexpr int THREADS = C * TC;
};

template <typename T, int M> struct RowShape {
  static constexpr FftPlan PL = make_plan(M);
  static constexpr int TC = PL.T;
  // rows per block: at most 256 threads and at most 64 KiB of LDS
  static constexpr int R0 = 256 / TC;
  static constexpr int RL = 65536 / ((M + 1) * (int)sizeof(c2<T>));
  static constexpr int RR = RL < R0 ? RL : R0;
  static constexpr int THREADS = RR * TC;
};

// ---------------------------------------------------------------------------
// z pass, forward R2C.  If GEN, the real input is drawn from the seeded
// stress stream instead of being read (fused fluctuating-stress generation).
// ---------------------------------------------------------------------------
struct StressGen {
  unsigned long long seed;
  double scale[6];
  int comp0;    // stream component of field 0 of this launch
  int ny;       // rows per x plane
  int channel;  // slip-channel image system in y
};

__device__ __forceinline__ unsigned long long sm_mix(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// The N(0,1) draws of (component, node) and (component, node + 1): for each,
// the SplitMix64 stream keyed by the seed is read at position
// (comp*nnodes + node)*64 + t, Marsaglia polar method, first variate, at most
// 32 tries, else 0 (parallel_random_numbers.cpp:8-19).  One rejection loop
// serves both draws (the second starts when the first is accepted), so a wave
// runs as many passes as its busiest lane needs for the pair, and the
// transform x1 sqrt(-2 ln w / w) runs once per draw after the loop instead of
// inside it under every lane's acceptance.  Same values as one draw at a time.
__device__ void stress_normal_pair(unsigned long long seed, int comp, unsigned long long node,
                                   unsigned long long nnodes, double& a, double& b) {
  const unsigned long long G = 0x9E3779B97F4A7C15ULL;
  const unsigned long long key = sm_mix(seed + G);
  const unsigned long long base0 = ((unsigned long long)comp * nnodes + node) * 64ULL;
  double xa = 0.0, wa = 0.0, xb = 0.0, wb = 0.0;  // w = 0: no acceptance (the draw is 0)
  bool second = false;
  int t = 0;
  for (int pass = 0; pass < 64; ++pass) {
    const unsigned long long base = second ? base0 + 64ULL : base0;
    const double u1 = (double)(sm_mix(key + G * (base + 2ULL * t + 1ULL)) >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(sm_mix(key + G * (base + 2ULL * t + 2ULL)) >> 11) * (1.0 / 9007199254740992.0);
    const double x1 = 2.0 * u1 - 1.0, x2 = 2.0 * u2 - 1.0;
    const double w = x1 * x1 + x2 * x2;
    const bool ok = w < 1.0 && w > 0.0;
    if (!second) {
      if (ok) { xa = x1; wa = w; }
      if (ok || t == 31) { second = true; t = 0; continue; }
    } else {
      if (ok) { xb = x1; wb = w; }
      if (ok || t == 31) break;
    }
    ++t;
  }
  a = wa > 0.0 ? xa * sqrt((-2.0 * log(wa)) / wa) : 0.0;
  b = wb > 0.0 ? xb * sqrt((-2.0 * log(wb)) / wb) : 0.0;
}

template <typename T, int M, bool GEN>
__global__ void __launch_bounds__(256) zfwd_kernel(FieldPtrs<T> fp, int rows, int cpitch,
                                                   const c2<T>* __restrict__ tw, const c2<T>* __restrict__ twz,
                                                   StressGen sg, int nz_total_nodes_hi, int nz_total_nodes_lo) {
  if (fp.gate && *fp.gate) return;
  using V = c2<T>;
  using SH = RowShape<T, M>;
  constexpr FftPlan PL = make_plan(M);
  constexpr int P = PL.P, TC = SH::TC, RR = SH::RR;
  constexpr int LS = RR | 1;  // odd element stride: rows of a column fall in different LDS banks
  __shared__ V lds[(M + 1) * LS];
  const int t = threadIdx.x;
  const int tcol = t % TC, rloc = t / TC;
  const int row = blockIdx.x * RR + rloc;
  const bool valid = row < rows;
  V* base = fp.p[blockIdx.y] + (size_t)(valid ? row : 0) * cpitch;
  V v[P];
  if (GEN) {
    const unsigned long long nnodes = ((unsigned long long)nz_total_nodes_hi << 32) | (unsigned)nz_total_nodes_lo;
    const int comp = sg.comp0 + blockIdx.y;
    double sc = sg.scale[comp];
    int srow = row;
    if (sg.channel) {
      // slip channel (fcm_fluid_solver.cpp:186-217): the wall rows y = 0 and
      // y = ny/2 carry sqrt(2) times the scaled draw and no sxy, syz (stream
      // components 3 and 5); a row y > ny/2 repeats row ny - y, sxy and syz negated
      const int y = row % sg.ny, half = sg.ny / 2;
      const bool odd = comp == 3 || comp == 5;
      if (y == 0 || y == half) {
        sc = odd ? 0.0 : sqrt(2.0) * sc;
      } else if (y > half) {
        srow = row - y + (sg.ny - y);
        if (odd) sc = -sc;
      }
    }
    const bool draw = valid && sc != 0.0;
    const int nz = 2 * M;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<M, false>(tcol, s);
      const unsigned long long node = (unsigned long long)srow * nz + 2 * e;
      double na = 0.0, nb = 0.0;
      if (draw) stress_normal_pair(sg.seed, comp, node, nnodes, na, nb);
      const double a = draw ? sc * na : 0.0;
      const double b = draw ? sc * nb : 0.0;
      v[s] = mk<T>(T(a), T(b));
    }
  } else {
    const V* in = base;  // row of nz reals viewed as M complex words
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<M, false>(tcol, s);
      v[s] = valid ? in[e] : mk<T>(T(0), T(0));
    }
  }
  fft_column<T, M, -1, false>(v, lds, LS, rloc, tcol, tw);
#pragma unroll
  for (int s = 0; s < P; ++s) lds[fft_out_index<M, false>(tcol, s) * LS + rloc] = v[s];
  __syncthreads();
  // untangle, one bin pair (k, M - k) per step
  for (int k = tcol; k <= M / 2; k += TC) {
    V xk, xmk;
    r2c_untangle_pair<T>(lds[k * LS + rloc], lds[(k == 0 ? 0 : M - k) * LS + rloc], twz[k], xk, xmk);
    if (valid) {
      base[k] = xk;
      if (M - k != k) base[M - k] = xmk;
    }
  }
}

// z pass, inverse C2R (halfcomplex: Im of X[0] and X[M] ignored)
template <typename T, int M>
__global__ void __launch_bounds__(256) zinv_kernel(FieldPtrs<T> fp, int rows, int cpitch,
                                                   const c2<T>* __restrict__ tw, const c2<T>* __restrict__ twz) {
  if (fp.gate && *fp.gate) return;
  using V = c2<T>;
  using SH = RowShape<T, M>;
  constexpr FftPlan PL = make_plan(M);
  constexpr int P = PL.P, TC = SH::TC, RR = SH::RR;
  constexpr int LS = RR | 1;  // odd element stride: rows of a column fall in different LDS banks
  __shared__ V lds[(M + 1) * LS];
  const int t = threadIdx.x;
  const int tcol = t % TC, rloc = t / TC;
  const int row = blockIdx.x * RR + rloc;
  const bool valid = row < rows;
  V* base = fp.p[blockIdx.y] + (size_t)(valid ? row : 0) * cpitch;
  for (int k = tcol; k <= M; k += TC) {
    V x = valid ? base[k] : mk<T>(T(0), T(0));
    if (k == 0 || k == M) x.y = T(0);
    lds[k * LS + rloc] = x;
  }
  __syncthreads();
  V v[P];
#pragma unroll
  for (int s = 0; s < P; ++s) {
    const int k = fft_in_index<M, true>(tcol, s);
    const V xk = lds[k * LS + rloc];
    const V xm = cconj(lds[(M - k) * LS + rloc]);
    const V sum = cadd(xk, xm);
    const V dif = csub(xk, xm);
    const V rot = cmulc(dif, twz[k]);          // W^{-k} (X[k] - conj X[M-k])
    v[s] = cadd(sum, mk<T>(-rot.y, rot.x));    // + i * rot
  }
  __syncthreads();
  fft_column<T, M, +1, true>(v, lds, LS, rloc, tcol, tw);
  if (valid) {
#pragma unroll
    for (int s = 0; s < P; ++s) base[fft_out_index<M, true>(tcol, s)] = v[s];
  }
}

// ---------------------------------------------------------------------------
// column passes (y or x), one field per blockIdx.z
//   element e of column kz: data[base + e*estride + kz]
//   y pass: blockIdx.y = x index, base = x*ny*cpitch, estride = cpitch
//   x pass: blockIdx.y = y index, base = y*cpitch,    estride = ny*cpitch
// ---------------------------------------------------------------------------
// MIRROR (forward y pass over the three force fields of a slip channel): the
// slip-channel image rule is applied as the columns are read, so spreading
// need not write it: element y > N/2 is read from N - y, field 1 (fy) negated
// there and zero at y = 0 and y = N/2.  On a field that already holds the
// images the read returns what is stored, so the rule may be applied twice.
template <typename T, int N, int S, bool MIRROR = false, int PM = 16>
__global__ void __launch_bounds__(256) cols_kernel(FieldPtrs<T> fp, size_t plane_stride, size_t estride, int nzc,
                                                   const c2<T>* __restrict__ tw) {
  if (fp.gate && *fp.gate) return;
  using V = c2<T>;
  using SH = ColShape<T, N, PM>;
  constexpr FftPlan PL = make_plan(N, PM);
  constexpr int P = PL.P, C = SH::C;
  constexpr bool REV = (S > 0);
  __shared__ V lds[N * C];
  __shared__ V s_tw[N];
  const int t = threadIdx.x;
  for (int i = t; i < N; i += SH::THREADS) s_tw[i] = tw[i];
  __syncthreads();
  const int col = t % C, tcol = t / C;
  const int kz = blockIdx.x * C + col;
  const bool valid = kz < nzc;
  // a uniform plane base plus 32-bit byte offsets (scalar-base addressing);
  // the Nyquist-column remap of xproj_kernel does not pay here, as the Nyquist
  // columns of different planes are a plane apart (about 2% slower at 256^3)
  char* base = reinterpret_cast<char*>(fp.p[blockIdx.z] + (size_t)blockIdx.y * plane_stride);
  const unsigned colb = (unsigned)((valid ? kz : 0) * sizeof(V));
  const unsigned eb = (unsigned)(estride * sizeof(V));
  V v[P];
  if constexpr (MIRROR) {
    static_assert(S < 0, "the image rule is read by the forward pass");
    const bool fy = blockIdx.z == 1;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<N, REV, PM>(tcol, s);
      const bool up = e > N / 2;
      V x = valid ? *reinterpret_cast<const V*>(base + (colb + (unsigned)(up ? N - e : e) * eb)) : mk<T>(T(0), T(0));
      if (fy) {
        if (up) x = mk<T>(-x.x, -x.y);
        else if (e == 0 || e == N / 2) x = mk<T>(T(0), T(0));
      }
      v[s] = x;
    }
  } else {
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<N, REV, PM>(tcol, s);
      v[s] = valid ? *reinterpret_cast<const V*>(base + (colb + (unsigned)e * eb)) : mk<T>(T(0), T(0));
    }
  }
  fft_column<T, N, S, REV, PM>(v, lds, C, col, tcol, s_tw);
  if (valid) {
#pragma unroll
    for (int s = 0; s < P; ++s)
      *reinterpret_cast<V*>(base + (colb + (unsigned)fft_out_index<N, REV, PM>(tcol, s) * eb)) = v[s];
  }
}

// ---------------------------------------------------------------------------
// Fourier-space Stokes projection, fcm_fluid_solver.cpp:653-684.  For each
// stored mode, with q = (qx, qy, qz) (Nyquist wavenumbers positive):
//   |q|^2 = qy^2 + qx^2 + 1e-16 + qz^2,
//   u = (f - q (q.f)/|q|^2) / (|q|^2 mu nx ny nz),
// and the k = 0 mode set to zero (:642-651, :686-695).
// ---------------------------------------------------------------------------
// 1 / x: the hardware reciprocal in FP32 (1 ulp, against the ten-instruction
// IEEE division), a division in FP64
__device__ __forceinline__ float proj_recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double proj_recip(double x) { return 1.0 / x; }

template <typename T>
__device__ __forceinline__ void project_mode(c2<T>& a, c2<T>& b, c2<T>& c, T q1, T q2, T q3, T inv_muN, bool k0) {
  const T qq = q2 * q2 + q1 * q1 + T(1e-16);
  const T norm = proj_recip(qq + q3 * q3);
  const T kr = (q1 * a.x + q2 * b.x + q3 * c.x) * norm;
  const T ki = (q1 * a.y + q2 * b.y + q3 * c.y) * norm;
  const T s = norm * inv_muN;
  if (k0) {
    a = mk<T>(T(0), T(0)); b = a; c = a;
    return;
  }
  a = mk<T>(s * (a.x - q1 * kr), s * (a.y - q1 * ki));
  b = mk<T>(s * (b.x - q2 * kr), s * (b.y - q2 * ki));
  c = mk<T>(s * (c.x - q3 * kr), s * (c.y - q3 * ki));
}

__device__ __forceinline__ int signed_k(int k, int n) { return k <= n / 2 ? k : k - n; }

// forward x FFT of 3 fields -> projection -> inverse x FFT, in place.
// points per thread of the fused x passes: three (or more) fields stay in
// registers between the forward and inverse transforms
// (8 points per thread in FP32 as well, with 512-thread workgroups, halves the
// registers but doubles the LDS exchanges per transform: 0.146 ms against
// 0.102 ms at 256^3, so FP32 keeps 16)
template <typename T> struct XPM { static constexpr int value = sizeof(T) == 8 ? 8 : 16; };
template <typename T> struct XTMAX { static constexpr int value = 256; };

template <typename T, int N, int PM = XPM<T>::value>
__global__ void __launch_bounds__(XTMAX<T>::value) xproj_kernel(FieldPtrs<T> fp, int ny, int cpitch, int nzc,
                                                    ProjParams pp, const c2<T>* __restrict__ tw) {
  if (fp.gate && *fp.gate) return;
  using V = c2<T>;
  using SH = ColShape<T, N, PM, XTMAX<T>::value>;
  constexpr FftPlan PL = make_plan(N, PM);
  constexpr int P = PL.P, C = SH::C;
  __shared__ V lds[N * C];
  __shared__ V s_tw[N];
  const int t = threadIdx.x;
  for (int i = t; i < N; i += SH::THREADS) s_tw[i] = tw[i];
  __syncthreads();
  const int col = t % C, tcol = t / C;
  // When nz/2 is a multiple of C, the last column block of a y row would hold
  // the Nyquist column kz = nz/2 alone; instead, the last blocks each take the
  // Nyquist columns of C consecutive y rows (a row apart), and the surplus ones
  // return (0.095 -> 0.092 ms FP32 at 256^3).
  const int M = nzc - 1;
  const bool tail = (M % C == 0) && (int)blockIdx.x == M / C;
  if (tail && (int)blockIdx.y * C >= ny) return;
  const int ky = tail ? (int)blockIdx.y * C + col : (int)blockIdx.y;
  const int kz = tail ? M : (int)blockIdx.x * C + col;
  const bool valid = tail ? ky < ny : kz < nzc;
  // uniform bases plus 32-bit byte offsets
  const size_t rowb = tail ? (size_t)0 : (size_t)blockIdx.y * cpitch;
  const unsigned colb = valid ? (unsigned)(((tail ? (size_t)ky * cpitch : (size_t)0) + kz) * sizeof(V)) : 0u;
  const unsigned eb = (unsigned)((size_t)ny * cpitch * sizeof(V));
  V v[3][P];
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const char* d = reinterpret_cast<const char*>(fp.p[f] + rowb);
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<N, false, PM>(tcol, s);
      v[f][s] = valid ? *reinterpret_cast<const V*>(d + (colb + (unsigned)e * eb)) : mk<T>(T(0), T(0));
    }
    fft_column<T, N, -1, false, PM>(v[f], lds, C, col, tcol, s_tw);
  }
  const T q2 = T(signed_k(ky, pp.ny) * pp.facy);
  const T q3 = T(kz * pp.facz);
  const T inv_muN = T(pp.inv_muN);
#pragma unroll
  for (int s = 0; s < P; ++s) {
    const int kx = fft_out_index<N, false, PM>(tcol, s);
    const T q1 = T(signed_k(kx, pp.nx) * pp.facx);
    project_mode<T>(v[0][s], v[1][s], v[2][s], q1, q2, q3, inv_muN, (kx | ky | kz) == 0);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    fft_column<T, N, +1, true, PM>(v[f], lds, C, col, tcol, s_tw);
    if (valid) {
      char* d = reinterpret_cast<char*>(fp.p[f] + rowb);
#pragma unroll
      for (int s = 0; s < P; ++s)
        *reinterpret_cast<V*>(d + (colb + (unsigned)fft_out_index<N, true, PM>(tcol, s) * eb)) = v[f][s];
    }
  }
}

// forward x FFT of the 6 stress fields -> f = i q'.S with the Nyquist
// wavenumbers zeroed (fcm_fluid_solver.cpp:243-292) -> projection -> inverse
// x FFT, written to fields 0..2 of fp (in place over sxx, syy, szz).
// Stream/field order: sxx syy szz sxy sxz syz.
template <typename T, int N>
__global__ void __launch_bounds__(XTMAX<T>::value) xstress_kernel(FieldPtrs<T> fp, int ny, int cpitch, int nzc,
                                                      ProjParams pp, const c2<T>* __restrict__ tw) {
  if (fp.gate && *fp.gate) return;
  using V = c2<T>;
  constexpr int PM = XPM<T>::value;
  using SH = ColShape<T, N, PM, XTMAX<T>::value>;
  constexpr FftPlan PL = make_plan(N, PM);
  constexpr int P = PL.P, C = SH::C;
  __shared__ V lds[N * C];
  const int t = threadIdx.x;
  const int col = t % C, tcol = t / C;
  const int kz = blockIdx.x * C + col;
  const int ky = blockIdx.y;
  const bool valid = kz < nzc;
  const size_t estride = (size_t)ny * cpitch;
  const size_t off = (size_t)ky * cpitch + (valid ? kz : 0);
  const T q2n = (ky == pp.ny / 2) ? T(0) : T(signed_k(ky, pp.ny) * pp.facy);
  const T q3n = (kz == pp.nz / 2) ? T(0) : T(kz * pp.facz);
  V acc[3][P];
#pragma unroll
  for (int s = 0; s < P; ++s) acc[0][s] = acc[1][s] = acc[2][s] = mk<T>(T(0), T(0));
  for (int f = 0; f < 6; ++f) {
    V v[P];
    const V* d = fp.p[f] + off;
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int e = fft_in_index<N, false, PM>(tcol, s);
      v[s] = valid ? d[(size_t)e * estride] : mk<T>(T(0), T(0));
    }
    fft_column<T, N, -1, false, PM>(v, lds, C, col, tcol, tw);
#pragma unroll
    for (int s = 0; s < P; ++s) {
      const int kx = fft_out_index<N, false, PM>(tcol, s);
      const T q1n = (kx == pp.nx / 2) ? T(0) : T(signed_k(kx, pp.nx) * pp.facx);
      switch (f) {
        case 0: acc[0][s] = cadd(acc[0][s], cscale(v[s], q1n)); break;
        case 1: acc[1][s] = cadd(acc[1][s], cscale(v[s], q2n)); break;
        case 2: acc[2][s] = cadd(acc[2][s], cscale(v[s], q3n)); break;
        case 3: acc[0][s] = cadd(acc[0][s], cscale(v[s], q2n)); acc[1][s] = cadd(acc[1][s], cscale(v[s], q1n)); break;
        case 4: acc[0][s] = cadd(acc[0][s], cscale(v[s], q3n)); acc[2][s] = cadd(acc[2][s], cscale(v[s], q1n)); break;
        default: acc[1][s] = cadd(acc[1][s], cscale(v[s], q3n)); acc[2][s] = cadd(acc[2][s], cscale(v[s], q2n)); break;
      }
    }
  }
  const T q2 = T(signed_k(ky, pp.ny) * pp.facy);
  const T q3 = T(kz * pp.facz);
  const T inv_muN = T(pp.inv_muN);
#pragma unroll
  for (int s = 0; s < P; ++s) {
    const int kx = fft_out_index<N, false, PM>(tcol, s);
    const T q1 = T(signed_k(kx, pp.nx) * pp.facx);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c][s] = mk<T>(-acc[c][s].y, acc[c][s].x);  // times i
    project_mode<T>(acc[0][s], acc[1][s], acc[2][s], q1, q2, q3, inv_muN, (kx | ky | kz) == 0);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    fft_column<T, N, +1, true, PM>(acc[f], lds, C, col, tcol, tw);
    if (valid) {
      V* d = fp.p[f] + off;
#pragma unroll
      for (int s = 0; s < P; ++s) d[(size_t)fft_out_index<N, true, PM>(tcol, s) * estride] = acc[f][s];
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
static hipError_t launch_z(const GridGeom& g, const FieldPtrs<T>& fp, int nf, bool fwd, const TwiddleSet<T>& tws,
                           hipStream_t st, const StressGen* gen) {
  const int M = g.nz / 2;
  const int rows = g.nx * g.ny;
  const unsigned long long nn = (unsigned long long)g.nx * g.ny * g.nz;
  StressGen sg{};
  if (gen) sg = *gen;
  switch (M) {
#define FCM_CASE(v)                                                                                             \
  case v: {                                                                                                     \
    using SH = RowShape<T, v>;                                                                                  \
    dim3 grid((rows + SH::RR - 1) / SH::RR, nf);                                                                \
    if (!fwd)                                                                                                   \
      hipLaunchKernelGGL((zinv_kernel<T, v>), grid, dim3(SH::THREADS), 0, st, fp, rows, g.cpitch, tws.z, tws.zu); \
    else if (gen)                                                                                               \
      hipLaunchKernelGGL((zfwd_kernel<T, v, true>), grid, dim3(SH::THREADS), 0, st, fp, rows, g.cpitch, tws.z,  \
                         tws.zu, sg, (int)(nn >> 32), (int)(nn & 0xffffffffULL));                               \
    else                                                                                                        \
      hipLaunchKernelGGL((zfwd_kernel<T, v, false>), grid, dim3(SH::THREADS), 0, st, fp, rows, g.cpitch, tws.z, \
                         tws.zu, sg, 0, 0);                                                                     \
    break;                                                                                                      \
  }
    FCM_FOR_EACH_N(FCM_CASE)
#undef FCM_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N, int PM>
static hipError_t cols_launch(const GridGeom& g, const FieldPtrs<T>& fp, int nf, int nplanes, size_t plane_stride,
                              size_t estride, bool fwd, bool mirror, const c2<T>* tw, hipStream_t st) {
  using SH = ColShape<T, N, PM>;
  const dim3 grid((g.nzc + SH::C - 1) / SH::C, nplanes, nf);
  if (mirror)
    hipLaunchKernelGGL((cols_kernel<T, N, -1, true, PM>), grid, dim3(SH::THREADS), 0, st, fp, plane_stride, estride,
                       g.nzc, tw);
  else if (fwd)
    hipLaunchKernelGGL((cols_kernel<T, N, -1, false, PM>), grid, dim3(SH::THREADS), 0, st, fp, plane_stride, estride,
                       g.nzc, tw);
  else
    hipLaunchKernelGGL((cols_kernel<T, N, 1, false, PM>), grid, dim3(SH::THREADS), 0, st, fp, plane_stride, estride,
                       g.nzc, tw);
  return hipGetLastError();
}

template <typename T>
static hipError_t launch_cols(const GridGeom& g, const FieldPtrs<T>& fp, int nf, int axis, bool fwd,
                              const TwiddleSet<T>& tws, hipStream_t st, bool mirror = false) {
  if (mirror && (axis != 1 || !fwd || nf != 3)) return hipErrorInvalidValue;
  const int N = axis == 1 ? g.ny : g.nx;
  const int nplanes = axis == 1 ? g.nx : g.ny;
  const size_t plane_stride = axis == 1 ? (size_t)g.ny * g.cpitch : (size_t)g.cpitch;
  const size_t estride = axis == 1 ? (size_t)g.cpitch : (size_t)g.ny * g.cpitch;
  const c2<T>* tw = axis == 1 ? tws.y : tws.x;
  // the tuned shape applies to the y passes; x passes of the stand-alone
  // transforms keep the default
  const int pm = axis == 1 ? g.ypm[sizeof(T) == 4 ? 1 : 0] : 0;
  switch (N) {
#define FCM_CASE(v)                                                                                              \
  case v: {                                                                                                      \
    if constexpr (v >= 64)                                                                                       \
      if (pm == 8) return cols_launch<T, v, 8>(g, fp, nf, nplanes, plane_stride, estride, fwd, mirror, tw, st);  \
    return cols_launch<T, v, 16>(g, fp, nf, nplanes, plane_stride, estride, fwd, mirror, tw, st);               \
  }
    FCM_FOR_EACH_N(FCM_CASE)
#undef FCM_CASE
    default:
      return hipErrorInvalidValue;
  }
}

template <typename T>
static hipError_t launch_xproj(const GridGeom& g, const FieldPtrs<T>& fp, bool stress, const ProjParams& pp,
                               const TwiddleSet<T>& tws, hipStream_t st) {
  const int pm = g.xpm[sizeof(T) == 4 ? 1 : 0];
  switch (g.nx) {
#define FCM_CASE(v)                                                                                             \
  case v: {                                                                                                     \
    using SH = ColShape<T, v, XPM<T>::value, XTMAX<T>::value>;                                                  \
    dim3 grid((g.nzc + SH::C - 1) / SH::C, g.ny);                                                               \
    if (stress) {                                                                                               \
      hipLaunchKernelGGL((xstress_kernel<T, v>), grid, dim3(SH::THREADS), 0, st, fp, g.ny, g.cpitch, g.nzc, pp, tws.x); \
      break;                                                                                                    \
    }                                                                                                           \
    if constexpr (sizeof(T) == 4 && v >= 64)                                                                    \
      if (pm == 8) {                                                                                            \
        using S8 = ColShape<T, v, 8, XTMAX<T>::value>;                                                          \
        hipLaunchKernelGGL((xproj_kernel<T, v, 8>), dim3((g.nzc + S8::C - 1) / S8::C, g.ny), dim3(S8::THREADS), \
                           0, st, fp, g.ny, g.cpitch, g.nzc, pp, tws.x);                                        \
        break;                                                                                                  \
      }                                                                                                         \
    hipLaunchKernelGGL((xproj_kernel<T, v>), grid, dim3(SH::THREADS), 0, st, fp, g.ny, g.cpitch, g.nzc, pp, tws.x); \
    break;                                                                                                      \
  }
    FCM_FOR_EACH_N(FCM_CASE)
#undef FCM_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t fft3d_forward(const GridGeom& g, const FieldPtrs<T>& fp, int nf, const TwiddleSet<T>& tws, hipStream_t st) {
  hipError_t e = launch_z<T>(g, fp, nf, true, tws, st, nullptr);
  if (e == hipSuccess) e = launch_cols<T>(g, fp, nf, 1, true, tws, st);
  if (e == hipSuccess) e = launch_cols<T>(g, fp, nf, 0, true, tws, st);
  return e;
}

template <typename T>
hipError_t fft3d_inverse(const GridGeom& g, const FieldPtrs<T>& fp, int nf, const TwiddleSet<T>& tws, hipStream_t st) {
  hipError_t e = launch_cols<T>(g, fp, nf, 0, false, tws, st);
  if (e == hipSuccess) e = launch_cols<T>(g, fp, nf, 1, false, tws, st);
  if (e == hipSuccess) e = launch_z<T>(g, fp, nf, false, tws, st, nullptr);
  return e;
}

template <typename T>
hipError_t stokes_solve(const GridGeom& g, const FieldPtrs<T>& fp, const ProjParams& pp, const TwiddleSet<T>& tws,
                        hipStream_t st, bool images) {
  if (images && !g.channel) return hipErrorInvalidValue;
  hipError_t e = launch_z<T>(g, fp, 3, true, tws, st, nullptr);
  if (e == hipSuccess) e = launch_cols<T>(g, fp, 3, 1, true, tws, st, images);
  if (e == hipSuccess) e = launch_xproj<T>(g, fp, false, pp, tws, st);
  if (e == hipSuccess) e = launch_cols<T>(g, fp, 3, 1, false, tws, st);
  if (e == hipSuccess) e = launch_z<T>(g, fp, 3, false, tws, st, nullptr);
  return e;
}

template <typename T>
hipError_t stokes_pass(const GridGeom& g, const FieldPtrs<T>& fp, int pass, const ProjParams& pp,
                       const TwiddleSet<T>& tws, hipStream_t st, bool images) {
  if (images && (!g.channel || pass != 1)) return hipErrorInvalidValue;
  switch (pass) {
    case 0: return launch_z<T>(g, fp, 3, true, tws, st, nullptr);
    case 1: return launch_cols<T>(g, fp, 3, 1, true, tws, st, images);
    case 2: return launch_xproj<T>(g, fp, false, pp, tws, st);
    case 3: return launch_cols<T>(g, fp, 3, 1, false, tws, st);
    case 4: return launch_z<T>(g, fp, 3, false, tws, st, nullptr);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t random_velocity(const GridGeom& g, const FieldPtrs<T>& stress6, bool generate, unsigned long long seed,
                           const double scale[6], const ProjParams& pp, const TwiddleSet<T>& tws, hipStream_t st) {
  StressGen sg{};
  sg.seed = seed;
  for (int i = 0; i < 6; ++i) sg.scale[i] = scale[i];
  sg.comp0 = 0;
  sg.ny = g.ny;
  sg.channel = g.channel;
  hipError_t e = launch_z<T>(g, stress6, 6, true, tws, st, generate ? &sg : nullptr);
  if (e == hipSuccess) e = launch_cols<T>(g, stress6, 6, 1, true, tws, st);
  if (e == hipSuccess) e = launch_xproj<T>(g, stress6, true, pp, tws, st);
  if (e == hipSuccess) e = launch_cols<T>(g, stress6, 3, 1, false, tws, st);
  if (e == hipSuccess) e = launch_z<T>(g, stress6, 3, false, tws, st, nullptr);
  return e;
}

template hipError_t fft3d_forward<double>(const GridGeom&, const FieldPtrs<double>&, int, const TwiddleSet<double>&, hipStream_t);
template hipError_t fft3d_forward<float>(const GridGeom&, const FieldPtrs<float>&, int, const TwiddleSet<float>&, hipStream_t);
template hipError_t fft3d_inverse<double>(const GridGeom&, const FieldPtrs<double>&, int, const TwiddleSet<double>&, hipStream_t);
template hipError_t fft3d_inverse<float>(const GridGeom&, const FieldPtrs<float>&, int, const TwiddleSet<float>&, hipStream_t);
template hipError_t stokes_solve<double>(const GridGeom&, const FieldPtrs<double>&, const ProjParams&, const TwiddleSet<double>&, hipStream_t, bool);
template hipError_t stokes_solve<float>(const GridGeom&, const FieldPtrs<float>&, const ProjParams&, const TwiddleSet<float>&, hipStream_t, bool);
template hipError_t stokes_pass<double>(const GridGeom&, const FieldPtrs<double>&, int, const ProjParams&, const TwiddleSet<double>&, hipStream_t, bool);
template hipError_t stokes_pass<float>(const GridGeom&, const FieldPtrs<float>&, int, const ProjParams&, const TwiddleSet<float>&, hipStream_t, bool);
template hipError_t random_velocity<double>(const GridGeom&, const FieldPtrs<double>&, bool, unsigned long long, const double*, const ProjParams&, const TwiddleSet<double>&, hipStream_t);
template hipError_t random_velocity<float>(const GridGeom&, const FieldPtrs<float>&, bool, unsigned long long, const double*, const ProjParams&, const TwiddleSet<float>&, hipStream_t);

}  // namespace fcm
