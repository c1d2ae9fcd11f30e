// fcm_common.hpp -- shared device helpers for the FCM mobility kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fcm {

// ---------------------------------------------------------------------------
// complex helpers (double2 / float2)
// ---------------------------------------------------------------------------
template <typename T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <typename T> using c2 = typename cplx_t<T>::type;

template <typename T> __host__ __device__ __forceinline__ c2<T> mk(T x, T y) { c2<T> r; r.x = x; r.y = y; return r; }
template <typename V> __host__ __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __host__ __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __host__ __device__ __forceinline__ V cmul(V a, V b) {
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// a * conj(b)
template <typename V> __host__ __device__ __forceinline__ V cmulc(V a, V b) {
  V r; r.x = a.x * b.x + a.y * b.y; r.y = a.y * b.x - a.x * b.y; return r;
}
template <typename V> __host__ __device__ __forceinline__ V cconj(V a) { V r; r.x = a.x; r.y = -a.y; return r; }
template <typename V, typename S> __host__ __device__ __forceinline__ V cscale(V a, S s) { V r; r.x = a.x * s; r.y = a.y * s; return r; }
// multiply by s*i, s = +1 or -1
template <int S, typename V> __host__ __device__ __forceinline__ V cmul_si(V a) {
  V r;
  if (S > 0) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}

// Real-to-complex untangle of Z, the M-point complex FFT of 2M reals packed as
// (even, odd) pairs, for the bin pair (k, M - k), 0 <= k <= M/2:
//   E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / (2i)  (Z indices mod M),
//   X[k] = E + W^k O,  X[M-k] = conj(E - W^k O),  W = exp(-i pi / M),
// the second from E[M-k] = conj E[k], O[M-k] = conj O[k] and W^(M-k) = -conj W^k.
template <typename T>
__device__ __forceinline__ void r2c_untangle_pair(c2<T> zk, c2<T> zmk, c2<T> wk, c2<T>& xk, c2<T>& xmk) {
  const c2<T> zm = cconj(zmk);
  const c2<T> e = cscale(cadd(zk, zm), T(0.5));
  const c2<T> d = csub(zk, zm);
  const c2<T> o = mk<T>(T(0.5) * d.y, T(-0.5) * d.x);  // d / (2i)
  const c2<T> t = cmul(wk, o);
  xk = cadd(e, t);
  xmk = cconj(csub(e, t));
}

// ---------------------------------------------------------------------------
// FFT plans: a length N is split into stages of radix 16, 8, 4, 2, 3 or 5.
// Each thread of a column holds P points; every radix divides P, so each
// thread owns P/R butterflies per stage (Stockham auto-sort formulation).
// ---------------------------------------------------------------------------
struct FftPlan {
  int N = 0, P = 0, T = 0, ns = 0;
  int r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool ok = false;
};

constexpr bool only_235(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

constexpr bool primes_divide(int N, int P) {
  for (int p : {2, 3, 5})
    if (N % p == 0 && P % p != 0) return false;
  return true;
}

// PM caps the points held per thread (register budget of the caller)
constexpr FftPlan make_plan(int N, int PM = 16) {
  FftPlan pl;
  pl.N = N;
  if (N < 2 || !only_235(N)) return pl;
  int P = 0;
  for (int d = PM; d >= 2; --d)
    if (N % d == 0 && primes_divide(N, d)) { P = d; break; }
  if (P == 0) return pl;
  pl.P = P;
  pl.T = N / P;
  int rem = N;
  const int cand[6] = {16, 8, 4, 2, 3, 5};
  while (rem > 1) {
    int pick = 0;
    for (int c : cand)
      if (P % c == 0 && rem % c == 0) { pick = c; break; }
    if (pick == 0 || pl.ns >= 8) return pl;
    pl.r[pl.ns++] = pick;
    rem /= pick;
  }
  pl.ok = true;
  return pl;
}

// inverse plans use the radices in reverse order, so that the register layout
// at the end of a forward transform is the layout an inverse starts from
constexpr FftPlan reverse_plan(FftPlan p) {
  FftPlan q = p;
  for (int i = 0; i < p.ns; ++i) q.r[i] = p.r[p.ns - 1 - i];
  return q;
}

constexpr int plan_ns_before(const FftPlan& p, int s) {
  int v = 1;
  for (int i = 0; i < s; ++i) v *= p.r[i];
  return v;
}

// supported transform lengths (x, y and half-z)
#define FCM_FOR_EACH_N(X) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

inline bool fft_len_supported(int n) {
  switch (n) {
#define FCM_CASE(v) case v:
    FCM_FOR_EACH_N(FCM_CASE)
#undef FCM_CASE
    return true;
    default:
      return false;
  }
}

// ---------------------------------------------------------------------------
// small-radix DFTs, natural-order output; S = -1 forward, +1 inverse
// ---------------------------------------------------------------------------
template <typename T, int R, int S> struct Dft;

template <typename T, int S> struct Dft<T, 2, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    const c2<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <typename T, int S> struct Dft<T, 3, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    const T h = T(0.8660254037844386467637231707529362);
    const c2<T> t1 = cadd(a[1], a[2]);
    const c2<T> t2 = csub(a[1], a[2]);
    const c2<T> m = csub(a[0], cscale(t1, T(0.5)));
    const c2<T> n = cscale(cmul_si<S>(t2), h);
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, n);
    a[2] = csub(m, n);
  }
};

template <typename T, int S> struct Dft<T, 4, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    const c2<T> t0 = cadd(a[0], a[2]);
    const c2<T> t1 = csub(a[0], a[2]);
    const c2<T> t2 = cadd(a[1], a[3]);
    const c2<T> t3 = cmul_si<S>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <typename T, int S> struct Dft<T, 5, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    const T c1 = T(0.3090169943749474241022934171828191);
    const T c2v = T(-0.8090169943749474241022934171828191);
    const T s1 = T(0.9510565162951535721164393333793821);
    const T s2 = T(0.5877852522924731291687059546390728);
    const c2<T> t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
    const c2<T> t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
    const c2<T> a0 = a[0];
    const c2<T> m1 = cadd(a0, cadd(cscale(t1, c1), cscale(t2, c2v)));
    const c2<T> m2 = cadd(a0, cadd(cscale(t1, c2v), cscale(t2, c1)));
    const c2<T> n1 = cmul_si<S>(cadd(cscale(t3, s1), cscale(t4, s2)));
    const c2<T> n2 = cmul_si<S>(csub(cscale(t3, s2), cscale(t4, s1)));
    a[0] = cadd(a0, cadd(t1, t2));
    a[1] = cadd(m1, n1);
    a[4] = csub(m1, n1);
    a[2] = cadd(m2, n2);
    a[3] = csub(m2, n2);
  }
};

// radix 8 = two radix-4 halves (even/odd samples) joined with W8^k
template <typename T, int S> struct Dft<T, 8, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    c2<T> e[4] = {a[0], a[2], a[4], a[6]};
    c2<T> o[4] = {a[1], a[3], a[5], a[7]};
    Dft<T, 4, S>::run(e);
    Dft<T, 4, S>::run(o);
    const T h = T(0.7071067811865475244008443621048490);
    // W8^1 = h(1 + S i), W8^2 = S i, W8^3 = h(-1 + S i)
    const c2<T> o1 = cscale(cadd(o[1], cmul_si<S>(o[1])), h);
    const c2<T> o2 = cmul_si<S>(o[2]);
    const c2<T> o3 = cscale(csub(cmul_si<S>(o[3]), o[3]), h);
    a[0] = cadd(e[0], o[0]); a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);   a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);   a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);   a[7] = csub(e[3], o3);
  }
};

// radix 16 = two radix-8 halves joined with W16^k
template <typename T, int S> struct Dft<T, 16, S> {
  static __device__ __forceinline__ void run(c2<T>* a) {
    c2<T> e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) { e[i] = a[2 * i]; o[i] = a[2 * i + 1]; }
    Dft<T, 8, S>::run(e);
    Dft<T, 8, S>::run(o);
    // W16^k = cos(2 pi k/16) + S i sin(2 pi k/16)
    const T cs[8] = {T(1.0), T(0.9238795325112867561281831893967883), T(0.7071067811865475244008443621048490),
                     T(0.3826834323650897717284599840303989), T(0.0), T(-0.3826834323650897717284599840303989),
                     T(-0.7071067811865475244008443621048490), T(-0.9238795325112867561281831893967883)};
    const T sn[8] = {T(0.0), T(0.3826834323650897717284599840303989), T(0.7071067811865475244008443621048490),
                     T(0.9238795325112867561281831893967883), T(1.0), T(0.9238795325112867561281831893967883),
                     T(0.7071067811865475244008443621048490), T(0.3826834323650897717284599840303989)};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      // W16^0 = 1 and W16^4 = S i take no multiplications
      const c2<T> w = mk<T>(cs[k], T(S) * sn[k]);
      const c2<T> t = (k == 0) ? o[0] : (k == 4) ? cmul_si<S>(o[4]) : cmul(o[k], w);
      a[k] = cadd(e[k], t);
      a[k + 8] = csub(e[k], t);
    }
  }
};

// ---------------------------------------------------------------------------
// Stockham column FFT.  `v` holds this thread's P points in the layout of the
// first radix r0: v[i*r0 + r] = x[(tcol + i*T) + r*N/r0].  On return v holds
// the outputs in the layout of the last radix.  Between stages the data goes
// through `lds` (element e of this column at lds[e*ls + col]).  Every thread
// of the block must call this (it contains barriers).
// tw[m] = exp(-2 pi i m / N), m < N.
// ---------------------------------------------------------------------------
template <typename T, int N, int S, bool REV, int STG, int PM>
struct FftStage {
  static constexpr FftPlan PL = REV ? reverse_plan(make_plan(N, PM)) : make_plan(N, PM);
  static constexpr int P = PL.P;
  static constexpr int TT = PL.T;
  static constexpr int R = PL.r[STG];
  static constexpr int NS = plan_ns_before(PL, STG);
  static constexpr bool LAST = (STG == PL.ns - 1);

  static __device__ __forceinline__ void run(c2<T>* v, c2<T>* lds, int ls, int col, int tcol,
                                             const c2<T>* __restrict__ tw) {
#pragma unroll
    for (int i = 0; i < P / R; ++i) {
      const int j = tcol + i * TT;
      c2<T>* a = v + i * R;
      if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int m = r * k * (N / (NS * R));
          c2<T> w = tw[m];
          if (S > 0) w.y = -w.y;
          a[r] = cmul(a[r], w);
        }
      }
      Dft<T, R, S>::run(a);
    }
    if constexpr (!LAST) {
      constexpr int R2 = PL.r[STG + 1];
#pragma unroll
      for (int i = 0; i < P / R; ++i) {
        const int j = tcol + i * TT;
        const int k = j % NS;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[(base + r * NS) * ls + col] = v[i * R + r];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < P / R2; ++i) {
        const int j = tcol + i * TT;
#pragma unroll
        for (int r = 0; r < R2; ++r) v[i * R2 + r] = lds[(j + r * (N / R2)) * ls + col];
      }
      __syncthreads();
      FftStage<T, N, S, REV, STG + 1, PM>::run(v, lds, ls, col, tcol, tw);
    }
  }
};

template <typename T, int N, int S, bool REV, int PM = 16>
__device__ __forceinline__ void fft_column(c2<T>* v, c2<T>* lds, int ls, int col, int tcol,
                                           const c2<T>* __restrict__ tw) {
  FftStage<T, N, S, REV, 0, PM>::run(v, lds, ls, col, tcol, tw);
}

// The same transform for NC columns per thread (column c of this thread at
// lds column col + c * cs), with one set of barriers for all of them and each
// twiddle factor loaded once for the NC columns.  Element e of a column lies at
// lds[e ls + (e >> SSH) SB + column]: the skew (SB > 0) spreads the exchanges
// of a wavefront over more LDS banks for some shapes.
template <int SSH, int SB>
__device__ __forceinline__ int fft_slot(int e, int ls, int col) {
  return e * ls + (SB != 0 ? (e >> SSH) * SB : 0) + col;
}

template <typename T, int N, int S, bool REV, int STG, int PM, int NC, int SSH = 0, int SB = 0>
struct FftStageB {
  static constexpr FftPlan PL = REV ? reverse_plan(make_plan(N, PM)) : make_plan(N, PM);
  static constexpr int P = PL.P;
  static constexpr int TT = PL.T;
  static constexpr int R = PL.r[STG];
  static constexpr int NS = plan_ns_before(PL, STG);
  static constexpr bool LAST = (STG == PL.ns - 1);

  static __device__ __forceinline__ void run(c2<T> (&v)[NC][P], c2<T>* lds, int ls, int col, int cs, int tcol,
                                             const c2<T>* __restrict__ tw) {
#pragma unroll
    for (int i = 0; i < P / R; ++i) {
      const int j = tcol + i * TT;
      if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const int m = r * k * (N / (NS * R));
          c2<T> w = tw[m];
          if (S > 0) w.y = -w.y;
#pragma unroll
          for (int c = 0; c < NC; ++c) v[c][i * R + r] = cmul(v[c][i * R + r], w);
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) Dft<T, R, S>::run(v[c] + i * R);
    }
    if constexpr (!LAST) {
      constexpr int R2 = PL.r[STG + 1];
#pragma unroll
      for (int i = 0; i < P / R; ++i) {
        const int j = tcol + i * TT;
        const int k = j % NS;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int c = 0; c < NC; ++c) lds[fft_slot<SSH, SB>(base + r * NS, ls, col + c * cs)] = v[c][i * R + r];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < P / R2; ++i) {
        const int j = tcol + i * TT;
#pragma unroll
        for (int r = 0; r < R2; ++r)
#pragma unroll
          for (int c = 0; c < NC; ++c) v[c][i * R2 + r] = lds[fft_slot<SSH, SB>(j + r * (N / R2), ls, col + c * cs)];
      }
      __syncthreads();
      FftStageB<T, N, S, REV, STG + 1, PM, NC, SSH, SB>::run(v, lds, ls, col, cs, tcol, tw);
    }
  }
};

template <typename T, int N, int S, bool REV, int PM, int NC, int SSH = 0, int SB = 0>
__device__ __forceinline__ void fft_columns(c2<T> (&v)[NC][REV ? reverse_plan(make_plan(N, PM)).P : make_plan(N, PM).P],
                                            c2<T>* lds, int ls, int col, int cs, int tcol,
                                            const c2<T>* __restrict__ tw) {
  FftStageB<T, N, S, REV, 0, PM, NC, SSH, SB>::run(v, lds, ls, col, cs, tcol, tw);
}

// element index held in register slot (i, r) for the first radix of a plan
template <int N, bool REV, int PM = 16>
__device__ __forceinline__ int fft_in_index(int tcol, int slot) {
  constexpr FftPlan PL = REV ? reverse_plan(make_plan(N, PM)) : make_plan(N, PM);
  constexpr int R0 = PL.r[0];
  const int i = slot / R0, r = slot % R0;
  return tcol + i * PL.T + r * (N / R0);
}

// element index held in register slot for the last radix of a plan
template <int N, bool REV, int PM = 16>
__device__ __forceinline__ int fft_out_index(int tcol, int slot) {
  constexpr FftPlan PL = REV ? reverse_plan(make_plan(N, PM)) : make_plan(N, PM);
  constexpr int RL = PL.r[PL.ns - 1];
  const int i = slot / RL, r = slot % RL;
  return tcol + i * PL.T + r * (N / RL);
}

// ---------------------------------------------------------------------------
// misc
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ int cyc(int d, int n, int omin) {
  // the unique value congruent to d mod n in [omin, omin + n)
  int t = (d - omin) % n;
  if (t < 0) t += n;
  return t + omin;
}

// Wavefront (64-lane) sum with DPP moves: quad butterflies, row shifts by 4
// and 8, then the row broadcasts 15 and 31 leave the total in lane 63, which
// is read back into every lane.  The addition order is fixed, so the result
// is bitwise repeatable.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xf, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROWMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_f32<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_f32<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f32<0x118, 0xf>(v);  // row_shr:8
  v += dpp_f32<0x142, 0xa>(v);  // row_bcast:15
  v += dpp_f32<0x143, 0xc>(v);  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0xb1, 0xf>(v);
  v += dpp_f64<0x4e, 0xf>(v);
  v += dpp_f64<0x114, 0xf>(v);
  v += dpp_f64<0x118, 0xf>(v);
  v += dpp_f64<0x142, 0xa>(v);
  v += dpp_f64<0x143, 0xc>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Partner exchange of the reduce-scatter below: lane l receives lane l ^ X
// within its row of 16.  X = 1, 2: quad permutations; X = 4, 8: a row shift
// left (lane l reads l + X) into the banks of 4 lanes whose bit X is clear,
// then a row shift right (l - X) into the others.
template <int X>
__device__ __forceinline__ int dpp_xor_row(int v) {
  if constexpr (X == 1) return __builtin_amdgcn_update_dpp(v, v, 0xb1, 0xf, 0xf, false);
  else if constexpr (X == 2) return __builtin_amdgcn_update_dpp(v, v, 0x4e, 0xf, 0xf, false);
  else if constexpr (X == 4) {
    const int r = __builtin_amdgcn_update_dpp(v, v, 0x104, 0xf, 0x5, false);
    return __builtin_amdgcn_update_dpp(r, v, 0x114, 0xf, 0xa, false);
  } else {
    static_assert(X == 8, "partner distance");
    const int r = __builtin_amdgcn_update_dpp(v, v, 0x108, 0xf, 0x3, false);
    return __builtin_amdgcn_update_dpp(r, v, 0x118, 0xf, 0xc, false);
  }
}
template <int X>
__device__ __forceinline__ double dpp_xor_row(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_xor_row<X>((int)(b & 0xffffffffLL));
  const int hi = dpp_xor_row<X>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// one halving step: the lane keeps the half of v selected by its bit X and
// adds the partner's copy of that half
template <int X, int H>
__device__ __forceinline__ void rs_step(const double* v, double* out, bool upper) {
#pragma unroll
  for (int i = 0; i < H; ++i) {
    const double mine = upper ? v[H + i] : v[i];
    const double send = upper ? v[i] : v[H + i];
    out[i] = mine + dpp_xor_row<X>(send);
  }
}

// Reduce-scatter of 16 values over each row of 16 lanes: returns the sum over
// the row of value rs16_index(lane).  Four halving steps (15 exchanges) where
// 16 wavefront sums would take 96; the addition order is fixed.
__device__ __forceinline__ int rs16_index(int lane) {
  return ((lane & 1) << 3) | ((lane & 2) << 1) | ((lane & 4) >> 1) | ((lane & 8) >> 3);
}
__device__ __forceinline__ double row_reduce_scatter16(const double (&v)[16], int lane) {
  double s8[8], s4[4], s2[2], s1[1];
  rs_step<1, 8>(v, s8, (lane & 1) != 0);
  rs_step<2, 4>(s8, s4, (lane & 2) != 0);
  rs_step<4, 2>(s4, s2, (lane & 4) != 0);
  rs_step<8, 1>(s2, s1, (lane & 8) != 0);
  return s1[0];
}

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float warp_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace fcm
