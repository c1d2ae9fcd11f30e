/*
 * fcm_oracle.c -- CPU oracle for the FCM mobility solve (TEST INFRASTRUCTURE).
 *
 * A plain FP64 C restatement of the reference's hot path.  It exists to check
 * the HIP product; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Each function cites the reference lines it
 * follows.  Parallelism (OpenMP) mirrors the reference's MPI x-slab split and
 * never changes a result: every sum is taken in the reference's order or in
 * a fixed-chunk order that does not depend on the thread count.
 *
 * Third-party arithmetic the reference delegates (absent from /root/reference
 * and from this machine, versions unpinned in the reference):
 *   - FFTW 2.x rfftwnd_mpi (2.1.5 is the last 2.x): restated below as an exact
 *     mixed-radix DFT with FFTW's halfcomplex C2R semantics (the imaginary
 *     parts of the kz=0 and kz=nz/2 planes are ignored on the inverse).
 *   - Armadillo inv_sympd / inv: restated as a Cholesky inverse and a
 *     symmetric (Jacobi) pseudo-inverse.
 *   - SPRNG: replaced, as the task statement requires, by a seeded
 *     counter-based SplitMix64 stream; the Gaussian transform is the
 *     Marsaglia polar method of parallel_random_numbers.cpp:8-19.
 */
#include "fcm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846264338327950288

/* ========================================================================= */
/* Random streams                                                            */
/* ========================================================================= */

static const uint64_t GOLDEN = 0x9E3779B97F4A7C15ULL;

static uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Sequential SplitMix64 (Steele, Lea & Flood 2014): state += gamma; mix. */
uint64_t or_splitmix64_next(uint64_t* state) {
  *state += GOLDEN;
  return sm_mix(*state);
}

static double u53(uint64_t u) { return (double)(u >> 11) * (1.0 / 9007199254740992.0); }

/* One N(0,1) draw for stress component `comp` at unpadded node index `node`.
 * Stream: SplitMix64 seeded with key = mix(seed + gamma), read at random
 * access position p = (comp*nnodes + node)*64 + t, t = 0,1,2,... (two
 * uniforms per polar attempt, at most 32 attempts; 0 if all fail).
 * Gaussian: Marsaglia polar method, first variate only, exactly as
 * parallel_random_numbers.cpp:8-19 (w == 0 is rejected as well, so the log
 * is finite). */
double or_stress_normal(uint64_t seed, int comp, uint64_t node, uint64_t nnodes) {
  const uint64_t key = sm_mix(seed + GOLDEN);
  const uint64_t base = ((uint64_t)comp * nnodes + node) * 64ULL;
  for (int t = 0; t < 32; ++t) {
    const double x1 = 2.0 * u53(sm_mix(key + GOLDEN * (base + 2 * (uint64_t)t + 1))) - 1.0;
    const double x2 = 2.0 * u53(sm_mix(key + GOLDEN * (base + 2 * (uint64_t)t + 2))) - 1.0;
    const double w = x1 * x1 + x2 * x2;
    if (w < 1.0 && w > 0.0) {
      return x1 * sqrt((-2.0 * log(w)) / w);
    }
  }
  return 0.0;
}

/* ========================================================================= */
/* 1-D transforms                                                            */
/* ========================================================================= */

/* Brute-force DFT: out[k] = sum_n in[n] exp(sign*2*pi*i*n*k/N).  O(N^2).
 * Used only to pin or_fft1d. */
void or_dft1d(int n, const double* in, double* out, int sign) {
  for (int k = 0; k < n; ++k) {
    long double sr = 0.0L, si = 0.0L;
    for (int j = 0; j < n; ++j) {
      const long m = ((long)j * k) % n;
      const long double ang = (long double)sign * 2.0L * (long double)OR_PI * (long double)m / (long double)n;
      const long double c = cosl(ang), s = sinl(ang);
      sr += in[2 * j] * c - in[2 * j + 1] * s;
      si += in[2 * j] * s + in[2 * j + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
}

typedef struct {
  int n;
  int sign;
  double* w; /* w[j] = exp(sign*2*pi*i*j/n), j < n */
} fft_plan;

static fft_plan make_plan(int n, int sign) {
  fft_plan p;
  p.n = n;
  p.sign = sign;
  p.w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    const long double ang = (long double)sign * 2.0L * (long double)OR_PI * (long double)j / (long double)n;
    p.w[2 * j] = (double)cosl(ang);
    p.w[2 * j + 1] = (double)sinl(ang);
  }
  return p;
}

static void free_plan(fft_plan* p) { free(p->w); p->w = NULL; }

static int small_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

/* Recursive decimation in time.  Input length n with stride `stride` (in
 * complex elements), output contiguous.  Root length N with table w. */
static void fft_rec(const fft_plan* P, int n, const double* in, int stride, double* out) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  const int p = small_factor(n);
  const int m = n / p;
  for (int r = 0; r < p; ++r) fft_rec(P, m, in + 2 * (size_t)r * stride, stride * p, out + 2 * (size_t)r * m);
  const int N = P->n;
  const int step = N / n; /* W_n^e = w[e*step] */
  double yr[64], yi[64];
  double* tr = (p > 64) ? (double*)malloc(sizeof(double) * 2 * (size_t)p) : NULL;
  double* ar = tr ? tr : yr;
  double* ai = tr ? tr + p : yi;
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      const double xr = out[2 * ((size_t)r * m + k)], xi = out[2 * ((size_t)r * m + k) + 1];
      const long e = ((long)r * k) % n;
      const double cr = P->w[2 * e * step], ci = P->w[2 * e * step + 1];
      ar[r] = xr * cr - xi * ci;
      ai[r] = xr * ci + xi * cr;
    }
    for (int t = 0; t < p; ++t) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; ++r) {
        const long e = ((long)r * t) % p; /* W_p^{rt} = W_n^{rt*m} */
        const double cr = P->w[2 * e * m * step], ci = P->w[2 * e * m * step + 1];
        sr += ar[r] * cr - ai[r] * ci;
        si += ar[r] * ci + ai[r] * cr;
      }
      out[2 * ((size_t)t * m + k)] = sr;
      out[2 * ((size_t)t * m + k) + 1] = si;
    }
  }
  free(tr);
}

/* Exact unnormalised DFT, out[k] = sum_n in[n] exp(sign*2*pi*i*n*k/N). */
void or_fft1d(int n, const double* in, double* out, int sign) {
  fft_plan P = make_plan(n, sign);
  fft_rec(&P, n, in, 1, out);
  free_plan(&P);
}

/* ========================================================================= */
/* 3-D transforms with the reference's layouts                               */
/* ========================================================================= */

static int pad_of(const or_cfg* c) { return 2 * (c->nz / 2 + 1); }

/* In-place unnormalised R2C, forward sign exp(-i k.x), output in transposed
 * order [ny][nx][nz/2+1] (fft_solver.cpp:21,32-36: rfftw3d_mpi_create_plan
 * FFTW_REAL_TO_COMPLEX + rfftwnd_mpi(..., FFTW_TRANSPOSED_ORDER)). */
void or_fft3d_r2c(const or_cfg* c, double* field) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzc = nz / 2 + 1, pad = pad_of(c);
  double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)nx * ny * nzc);
  fft_plan Pz = make_plan(nz, -1), Py = make_plan(ny, -1), Px = make_plan(nx, -1);
  const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
#pragma omp parallel
  {
    double* a = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
#pragma omp for schedule(static)
    for (int ij = 0; ij < nx * ny; ++ij) {
      const double* row = field + (size_t)ij * pad;
      for (int k = 0; k < nz; ++k) { a[2 * k] = row[k]; a[2 * k + 1] = 0.0; }
      fft_rec(&Pz, nz, a, 1, b);
      memcpy(tmp + 2 * (size_t)ij * nzc, b, sizeof(double) * 2 * (size_t)nzc);
    }
#pragma omp for schedule(static)
    for (int ik = 0; ik < nx * nzc; ++ik) {
      const int i = ik / nzc, kz = ik % nzc;
      for (int j = 0; j < ny; ++j) {
        const size_t s = 2 * (((size_t)i * ny + j) * nzc + kz);
        a[2 * j] = tmp[s]; a[2 * j + 1] = tmp[s + 1];
      }
      fft_rec(&Py, ny, a, 1, b);
      for (int j = 0; j < ny; ++j) {
        const size_t s = 2 * (((size_t)i * ny + j) * nzc + kz);
        tmp[s] = b[2 * j]; tmp[s + 1] = b[2 * j + 1];
      }
    }
#pragma omp for schedule(static)
    for (int jk = 0; jk < ny * nzc; ++jk) {
      const int j = jk / nzc, kz = jk % nzc;
      for (int i = 0; i < nx; ++i) {
        const size_t s = 2 * (((size_t)i * ny + j) * nzc + kz);
        a[2 * i] = tmp[s]; a[2 * i + 1] = tmp[s + 1];
      }
      fft_rec(&Px, nx, a, 1, b);
      for (int i = 0; i < nx; ++i) {
        double* d = field + ((size_t)j * nx + i) * pad + 2 * kz;
        d[0] = b[2 * i]; d[1] = b[2 * i + 1];
      }
    }
    free(a);
    free(b);
  }
  free_plan(&Pz); free_plan(&Py); free_plan(&Px);
  free(tmp);
}

/* In-place unnormalised C2R from transposed order back to [nx][ny][pad]
 * (fft_solver.cpp:22,38-42).  The complex inverse transforms along x and y
 * come first; the last dimension is FFTW's halfcomplex->real transform, which
 * stores no imaginary part for kz = 0 and kz = nz/2, so those are ignored. */
void or_fft3d_c2r(const or_cfg* c, double* field) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, nzc = nz / 2 + 1, pad = pad_of(c);
  double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)nx * ny * nzc); /* [j][i][kz] */
  fft_plan Pz = make_plan(nz, +1), Py = make_plan(ny, +1), Px = make_plan(nx, +1);
  const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
#pragma omp parallel
  {
    double* a = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
    double* b = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
#pragma omp for schedule(static)
    for (int jk = 0; jk < ny * nzc; ++jk) {
      const int j = jk / nzc, kz = jk % nzc;
      for (int i = 0; i < nx; ++i) {
        const double* d = field + ((size_t)j * nx + i) * pad + 2 * kz;
        a[2 * i] = d[0]; a[2 * i + 1] = d[1];
      }
      fft_rec(&Px, nx, a, 1, b);
      for (int i = 0; i < nx; ++i) {
        const size_t s = 2 * (((size_t)j * nx + i) * nzc + kz);
        tmp[s] = b[2 * i]; tmp[s + 1] = b[2 * i + 1];
      }
    }
#pragma omp for schedule(static)
    for (int ik = 0; ik < nx * nzc; ++ik) {
      const int i = ik / nzc, kz = ik % nzc;
      for (int j = 0; j < ny; ++j) {
        const size_t s = 2 * (((size_t)j * nx + i) * nzc + kz);
        a[2 * j] = tmp[s]; a[2 * j + 1] = tmp[s + 1];
      }
      fft_rec(&Py, ny, a, 1, b);
      for (int j = 0; j < ny; ++j) {
        const size_t s = 2 * (((size_t)j * nx + i) * nzc + kz);
        tmp[s] = b[2 * j]; tmp[s + 1] = b[2 * j + 1];
      }
    }
#pragma omp for schedule(static)
    for (int ij = 0; ij < nx * ny; ++ij) {
      const int i = ij / ny, j = ij % ny;
      const double* src = tmp + 2 * (((size_t)j * nx + i) * nzc);
      for (int k = 0; k < nzc; ++k) { a[2 * k] = src[2 * k]; a[2 * k + 1] = src[2 * k + 1]; }
      a[1] = 0.0;                                   /* no Im at kz = 0    */
      if (nz % 2 == 0) a[2 * (nz / 2) + 1] = 0.0;   /* no Im at kz = nz/2 */
      for (int k = nzc; k < nz; ++k) { a[2 * k] = a[2 * (nz - k)]; a[2 * k + 1] = -a[2 * (nz - k) + 1]; }
      fft_rec(&Pz, nz, a, 1, b);
      double* row = field + (size_t)ij * pad;
      for (int k = 0; k < nz; ++k) row[k] = b[2 * k];
      for (int k = nz; k < pad; ++k) row[k] = 0.0;
    }
    free(a);
    free(b);
  }
  free_plan(&Pz); free_plan(&Py); free_plan(&Px);
  free(tmp);
}

/* ========================================================================= */
/* Fluid solver object                                                       */
/* ========================================================================= */

struct or_fcm {
  or_cfg c;
  int pad;
  size_t total;
  double* fld[18];
  double *qx, *qxsq, *qy, *qysq, *qpad, *qpadsq;
  int nb, max_blobs;
  double *gaussx, *gaussy, *gaussz;
  int *indx, *indy, *indz;
};

or_fcm* or_fcm_create(const or_cfg* c, int max_blobs) {
  or_fcm* f = (or_fcm*)calloc(1, sizeof(or_fcm));
  f->c = *c;
  f->pad = pad_of(c);
  f->total = (size_t)c->nx * c->ny * f->pad;
  for (int i = 0; i < 18; ++i) f->fld[i] = (double*)calloc(f->total, sizeof(double));
  f->max_blobs = max_blobs > 0 ? max_blobs : 1;
  const size_t gsz = (size_t)f->max_blobs * c->ng;
  f->gaussx = (double*)calloc(gsz, sizeof(double));
  f->gaussy = (double*)calloc(gsz, sizeof(double));
  f->gaussz = (double*)calloc(gsz, sizeof(double));
  f->indx = (int*)calloc(gsz, sizeof(int));
  f->indy = (int*)calloc(gsz, sizeof(int));
  f->indz = (int*)calloc(gsz, sizeof(int));

  /* Wavenumber tables, fcm_fluid_solver.cpp:71-127.  Nyquist entries are
   * positive (i <= n/2), qpad pairs Re/Im slots with the same kz. */
  const double LX = c->dx * c->nx, LY = c->dx * c->ny, LZ = c->dx * c->nz;
  f->qx = (double*)malloc(sizeof(double) * c->nx);
  f->qxsq = (double*)malloc(sizeof(double) * c->nx);
  f->qy = (double*)malloc(sizeof(double) * c->ny);
  f->qysq = (double*)malloc(sizeof(double) * c->ny);
  f->qpad = (double*)malloc(sizeof(double) * f->pad);
  f->qpadsq = (double*)malloc(sizeof(double) * f->pad);
  double ff = 2.0 * OR_PI / LX;
  for (int i = 0; i < c->nx; ++i) {
    f->qx[i] = (i <= c->nx / 2) ? i * ff : (i - c->nx) * ff;
    f->qxsq[i] = f->qx[i] * f->qx[i];
  }
  ff = 2.0 * OR_PI / LY;
  for (int i = 0; i < c->ny; ++i) {
    f->qy[i] = (i <= c->ny / 2) ? i * ff : (i - c->ny) * ff;
    f->qysq[i] = f->qy[i] * f->qy[i];
  }
  ff = 2.0 * OR_PI / LZ;
  for (int i = 0; i < f->pad; i += 2) {
    f->qpad[i] = 0.5 * i * ff;
    f->qpad[i + 1] = f->qpad[i];
    f->qpadsq[i] = f->qpad[i] * f->qpad[i];
    f->qpadsq[i + 1] = f->qpadsq[i];
  }
  return f;
}

void or_fcm_destroy(or_fcm* f) {
  if (!f) return;
  for (int i = 0; i < 18; ++i) free(f->fld[i]);
  free(f->qx); free(f->qxsq); free(f->qy); free(f->qysq); free(f->qpad); free(f->qpadsq);
  free(f->gaussx); free(f->gaussy); free(f->gaussz);
  free(f->indx); free(f->indy); free(f->indz);
  free(f);
}

int or_fcm_pad(const or_fcm* f) { return f->pad; }
size_t or_fcm_total(const or_fcm* f) { return f->total; }
double* or_fcm_field(or_fcm* f, int which) { return (which >= 0 && which < 18) ? f->fld[which] : NULL; }

void or_fcm_tables(or_fcm* f, double* qx, double* qy, double* qpad) {
  memcpy(qx, f->qx, sizeof(double) * f->c.nx);
  memcpy(qy, f->qy, sizeof(double) * f->c.ny);
  memcpy(qpad, f->qpad, sizeof(double) * f->pad);
}

#define UX f->fld[0]
#define UY f->fld[1]
#define UZ f->fld[2]
#define FX f->fld[3]
#define FY f->fld[4]
#define FZ f->fld[5]
#define SXX f->fld[6]
#define SYY f->fld[7]
#define SZZ f->fld[8]
#define SXY f->fld[9]
#define SXZ f->fld[10]
#define SYZ f->fld[11]
#define UXR f->fld[12]
#define UYR f->fld[13]
#define UZR f->fld[14]
#define UXS f->fld[15]
#define UYS f->fld[16]
#define UZS f->fld[17]

static int wrap_index(int g, int n) { return g - n * (int)floor((double)g / (double)n); }

/* fcm_fluid_solver.cpp:131-165.  Y is nb x 3 row-major (positions already
 * wrapped into [0,L)). */
void or_fcm_set_up_gaussians(or_fcm* f, const double* Y, int nb) {
  const or_cfg* c = &f->c;
  const int ng = c->ng, ngdh = ng / 2;
  const double anorm = 1.0 / (sqrt(2.0) * c->a);
  const double anorm2 = 2.0 * c->a * c->a / OR_PI;
  if (nb > f->max_blobs) {
    const size_t gsz = (size_t)nb * ng;
    f->max_blobs = nb;
    f->gaussx = (double*)realloc(f->gaussx, gsz * sizeof(double));
    f->gaussy = (double*)realloc(f->gaussy, gsz * sizeof(double));
    f->gaussz = (double*)realloc(f->gaussz, gsz * sizeof(double));
    f->indx = (int*)realloc(f->indx, gsz * sizeof(int));
    f->indy = (int*)realloc(f->indy, gsz * sizeof(int));
    f->indz = (int*)realloc(f->indz, gsz * sizeof(int));
  }
  f->nb = nb;
#pragma omp parallel for schedule(static)
  for (int np = 0; np < nb; ++np) {
    const int xc = (int)(Y[3 * np] / c->dx);
    const int yc = (int)(Y[3 * np + 1] / c->dx);
    const int zc = (int)(Y[3 * np + 2] / c->dx);
    for (int i = 0; i < ng; ++i) {
      const size_t o = (size_t)np * ng + i;
      const int xg = xc - ngdh + (i + 1);
      f->indx[o] = wrap_index(xg, c->nx);
      double xx = xg * c->dx - Y[3 * np];
      f->gaussx[o] = anorm * exp(-xx * xx / anorm2);
      const int yg = yc - ngdh + (i + 1);
      f->indy[o] = wrap_index(yg, c->ny);
      xx = yg * c->dx - Y[3 * np + 1];
      f->gaussy[o] = anorm * exp(-xx * xx / anorm2);
      const int zg = zc - ngdh + (i + 1);
      f->indz[o] = wrap_index(zg, c->nz);
      xx = zg * c->dx - Y[3 * np + 2];
      f->gaussz[o] = anorm * exp(-xx * xx / anorm2);
    }
  }
}

void or_fcm_get_gaussians(const or_fcm* f, double* gx, double* gy, double* gz, int* ix, int* iy, int* iz) {
  const size_t n = (size_t)f->nb * f->c.ng;
  memcpy(gx, f->gaussx, n * sizeof(double));
  memcpy(gy, f->gaussy, n * sizeof(double));
  memcpy(gz, f->gaussz, n * sizeof(double));
  memcpy(ix, f->indx, n * sizeof(int));
  memcpy(iy, f->indy, n * sizeof(int));
  memcpy(iz, f->indz, n * sizeof(int));
}

/* fcm_fluid_solver.cpp:552-600 (fully periodic branch).  Each thread owns an
 * x-slab and walks every blob in order, exactly as one MPI rank does with its
 * slab filter at :570, so every node's sum is taken in blob order. */
/* Slip-channel image system on the force fields (fcm_fluid_solver.cpp:602-630):
 * fy vanishes on the wall rows j = 0 and j = ny/2, and every row j > ny/2 is
 * replaced by the mirror image of row ny - j, fx and fz even, fy odd.  Whatever
 * spreading deposited in those rows (blobs whose stencil wraps through y = 0 or
 * reaches past ny/2) is discarded. */
static void channel_image(or_fcm* f) {
  const or_cfg* c = &f->c;
  const int ny = c->ny, half = ny / 2, pad = f->pad;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c->nx; ++i) {
    const size_t x_index = (size_t)i * ny * pad;
    for (int k = 0; k < pad; ++k) {
      FY[x_index + k] = 0.0;
      FY[x_index + (size_t)half * pad + k] = 0.0;
      for (int j = half + 1; j < ny; ++j) {
        const size_t index = x_index + (size_t)j * pad + k;
        const size_t image = x_index + (size_t)(ny - j) * pad + k;
        FX[index] = FX[image];
        FY[index] = -FY[image];
        FZ[index] = FZ[image];
      }
    }
  }
}

void or_fcm_spread(or_fcm* f, const double* F) {
  const or_cfg* c = &f->c;
  const int ng = c->ng, ny = c->ny, pad = f->pad, nb = f->nb;
  memset(FX, 0, f->total * sizeof(double));
  memset(FY, 0, f->total * sizeof(double));
  memset(FZ, 0, f->total * sizeof(double));
#pragma omp parallel
  {
    int nth = 1, tid = 0;
#ifdef _OPENMP
    nth = omp_get_num_threads();
    tid = omp_get_thread_num();
#endif
    const int x0 = (int)((long)c->nx * tid / nth), x1 = (int)((long)c->nx * (tid + 1) / nth);
    for (int np = 0; np < nb; ++np) {
      for (int i = 0; i < ng; ++i) {
        const int ii = f->indx[(size_t)np * ng + i];
        if (ii < x0 || ii >= x1) continue;
        const size_t x_index = (size_t)ii * ny * pad;
        for (int j = 0; j < ng; ++j) {
          const size_t y_index = (size_t)f->indy[(size_t)np * ng + j] * pad;
          for (int k = 0; k < ng; ++k) {
            const size_t index = x_index + y_index + f->indz[(size_t)np * ng + k];
            const double temp = f->gaussx[(size_t)np * ng + i] * f->gaussy[(size_t)np * ng + j] * f->gaussz[(size_t)np * ng + k];
            const int id = 3 * np;
            FX[index] += F[id] * temp;
            FY[index] += F[id + 1] * temp;
            FZ[index] += F[id + 2] * temp;
          }
        }
      }
    }
  }
  if (c->slip_channel) channel_image(f);
}

/* Fourier-space Stokes projection over every stored real slot, Nyquist modes
 * included (fcm_fluid_solver.cpp:653-684). */
static void project(or_fcm* f, const double* f1a, const double* f2a, const double* f3a,
                    double* u1, double* u2, double* u3) {
  const or_cfg* c = &f->c;
  const int nx = c->nx, ny = c->ny, pad = f->pad;
  const double totalpts = (double)nx * ny * c->nz;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; ++j) {
    const double q2 = f->qy[j];
    const size_t y_index = (size_t)j * nx * pad;
    for (int i = 0; i < nx; ++i) {
      const double q1 = f->qx[i];
      const double qq = f->qysq[j] + f->qxsq[i] + 1e-16;
      const size_t x_index = (size_t)i * pad;
      for (int k = 0; k < pad; ++k) {
        const size_t index = x_index + y_index + k;
        const double q3 = f->qpad[k];
        const double fa = f1a[index], fb = f2a[index], fc = f3a[index];
        const double norm = 1.0 / (qq + f->qpadsq[k]);
        const double kdotf = (q1 * fa + q2 * fb + q3 * fc) * norm;
        u1[index] = norm * (fa - q1 * (kdotf)) / (c->mu * totalpts);
        u2[index] = norm * (fb - q2 * (kdotf)) / (c->mu * totalpts);
        u3[index] = norm * (fc - q3 * (kdotf)) / (c->mu * totalpts);
      }
    }
  }
}

/* fcm_fluid_solver.cpp:636-701 */
void or_fcm_stokes(or_fcm* f) {
  or_fft3d_r2c(&f->c, FX);
  or_fft3d_r2c(&f->c, FY);
  or_fft3d_r2c(&f->c, FZ);
  FX[0] = FX[1] = FY[0] = FY[1] = FZ[0] = FZ[1] = 0.0;
  project(f, FX, FY, FZ, UX, UY, UZ);
  UX[0] = UX[1] = UY[0] = UY[1] = UZ[0] = UZ[1] = 0.0;
  or_fft3d_c2r(&f->c, UX);
  or_fft3d_c2r(&f->c, UY);
  or_fft3d_c2r(&f->c, UZ);
}

/* fcm_fluid_solver.cpp:363-424 (one process: the all-reduce is the identity). */
static void interp_from(or_fcm* f, const double* u1, const double* u2, const double* u3, double* V) {
  const or_cfg* c = &f->c;
  const int ng = c->ng, ny = c->ny, pad = f->pad, nb = f->nb;
  const double norm = c->dx * c->dx * c->dx;
#pragma omp parallel for schedule(static)
  for (int np = 0; np < nb; ++np) {
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int i = 0; i < ng; ++i) {
      const size_t x_index = (size_t)f->indx[(size_t)np * ng + i] * ny * pad;
      for (int j = 0; j < ng; ++j) {
        const int jj = f->indy[(size_t)np * ng + j];
        if (c->slip_channel && jj >= ny / 2) continue;  /* truncated Gaussians, :387-413 */
        const size_t y_index = (size_t)jj * pad;
        for (int k = 0; k < ng; ++k) {
          const size_t index = x_index + y_index + f->indz[(size_t)np * ng + k];
          const double temp = f->gaussx[(size_t)np * ng + i] * f->gaussy[(size_t)np * ng + j] * f->gaussz[(size_t)np * ng + k] * norm;
          v0 += u1[index] * temp;
          v1 += u2[index] * temp;
          v2 += u3[index] * temp;
        }
      }
    }
    V[3 * np] = v0;
    V[3 * np + 1] = v1;
    V[3 * np + 2] = v2;
  }
}

void or_fcm_interp(or_fcm* f, double* V) { interp_from(f, UX, UY, UZ, V); }
void or_fcm_random_interp(or_fcm* f, double* V) { interp_from(f, UXR, UYR, UZR, V); }
/* calculate_shear_blob_velocities, fcm_fluid_solver.cpp:488-550 */
void or_fcm_shear_interp(or_fcm* f, double* V) { interp_from(f, UXS, UYS, UZS, V); }

void or_fcm_mobility(or_fcm* f, const double* F, double* V) {
  or_fcm_spread(f, F);
  or_fcm_stokes(f);
  or_fcm_interp(f, V);
}

/* Fill the six stress fields from the seeded stream with the variances of
 * fcm_fluid_solver.cpp:169-170,221-226: var1 = sqrt(2 kT mu/(dx^3 dt)),
 * var2 = sqrt(2) var1; order sxx syy szz (var2), sxy sxz syz (var1).
 * Component c of the stream maps to fields in that order.  Padding slots are
 * set to zero (the R2C transform never reads them). */
void or_fcm_fill_stress(or_fcm* f, uint64_t seed) {
  const or_cfg* c = &f->c;
  const double var1 = sqrt(2.0 * c->kT * c->mu / (c->dx * c->dx * c->dx * c->dt));
  const double var2 = sqrt(2.0) * var1;
  double* dst[6] = {SXX, SYY, SZZ, SXY, SXZ, SYZ};
  const double scale[6] = {var2, var2, var2, var1, var1, var1};
  const uint64_t nnodes = (uint64_t)c->nx * c->ny * c->nz;
  const int ny = c->ny, half = ny / 2;
  for (int s = 0; s < 6; ++s) {
    double* d = dst[s];
    const int odd = (s == 3 || s == 5);  /* sxy, syz change sign under y -> -y */
#pragma omp parallel for schedule(static)
    for (int ij = 0; ij < c->nx * ny; ++ij) {
      /* slip channel (fcm_fluid_solver.cpp:186-217): the wall rows j = 0 and
       * j = ny/2 carry sqrt(2) times the variance-scaled draw and no sxy, syz;
       * a row j > ny/2 is the image of row ny - j, sxy and syz negated */
      const int j = ij % ny;
      int src = ij;
      double sc = scale[s];
      if (c->slip_channel) {
        if (j == 0 || j == half) {
          sc = odd ? 0.0 : sqrt(2.0) * scale[s];
        } else if (j > half) {
          src = ij - j + (ny - j);
          if (odd) sc = -sc;
        }
      }
      for (int k = 0; k < f->pad; ++k) {
        d[(size_t)ij * f->pad + k] =
            (k < c->nz && sc != 0.0) ? sc * or_stress_normal(seed, s, (uint64_t)src * c->nz + k, nnodes) : 0.0;
      }
    }
  }
}

/* fcm_fluid_solver.cpp:236-359 (fully periodic branch), from the six stress
 * fields currently stored. */
void or_fcm_random_field(or_fcm* f) {
  const or_cfg* c = &f->c;
  const int nx = c->nx, ny = c->ny, pad = f->pad;
  or_fft3d_r2c(c, SXX);
  or_fft3d_r2c(c, SXY);
  or_fft3d_r2c(c, SXZ);
  or_fft3d_r2c(c, SYY);
  or_fft3d_r2c(c, SZZ);
  or_fft3d_r2c(c, SYZ);
  /* f = i q.S with q1,q2,q3 zeroed on the Nyquist indices (:243-292). */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; ++j) {
    double q2 = f->qy[j];
    if (j == ny / 2) q2 = 0.0;
    const size_t y_index = (size_t)j * nx * pad;
    for (int i = 0; i < nx; ++i) {
      double q1 = f->qx[i];
      if (i == nx / 2) q1 = 0.0;
      const size_t x_index = (size_t)i * pad;
      for (int k = 0; k < pad; k += 2) {
        const size_t index = x_index + y_index + k, ip = index + 1;
        double q3 = f->qpad[k];
        if (k == pad - 2) q3 = 0.0;
        FX[index] = -q1 * SXX[ip] - q2 * SXY[ip] - q3 * SXZ[ip];
        FY[index] = -q1 * SXY[ip] - q2 * SYY[ip] - q3 * SYZ[ip];
        FZ[index] = -q1 * SXZ[ip] - q2 * SYZ[ip] - q3 * SZZ[ip];
        FX[ip] = q1 * SXX[index] + q2 * SXY[index] + q3 * SXZ[index];
        FY[ip] = q1 * SXY[index] + q2 * SYY[index] + q3 * SYZ[index];
        FZ[ip] = q1 * SXZ[index] + q2 * SYZ[index] + q3 * SZZ[index];
      }
    }
  }
  FX[0] = FX[1] = FY[0] = FY[1] = FZ[0] = FZ[1] = 0.0;
  project(f, FX, FY, FZ, UX, UY, UZ);
  UX[0] = UX[1] = UY[0] = UY[1] = UZ[0] = UZ[1] = 0.0;
  or_fft3d_c2r(c, UX);
  or_fft3d_c2r(c, UY);
  or_fft3d_c2r(c, UZ);
  memcpy(UXR, UX, f->total * sizeof(double));
  memcpy(UYR, UY, f->total * sizeof(double));
  memcpy(UZR, UZ, f->total * sizeof(double));
}

/* produce_shear_velocity_field, fcm_fluid_solver.cpp:703-840: the force
 * fx = scale kT sin(2 pi kappa j / ny) / a on every real node (padding zero),
 * fy = fz = 0, the channel images, then the Stokes solve of
 * produce_deterministic_velocity_field; the result is kept in ux..uz_shear
 * (the force and velocity fields are left as the solve leaves them). */
void or_fcm_shear_field(or_fcm* f, double scale, double kappa) {
  const or_cfg* c = &f->c;
  const int ny = c->ny, pad = f->pad;
  memset(FX, 0, f->total * sizeof(double));
  memset(FY, 0, f->total * sizeof(double));
  memset(FZ, 0, f->total * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < c->nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const double v = scale * c->kT * sin(kappa * 2.0 * OR_PI * (double)j / (double)ny) / c->a;
      double* row = FX + ((size_t)i * ny + j) * pad;
      for (int k = 0; k < c->nz; ++k) row[k] = v;
    }
  if (c->slip_channel) channel_image(f);
  or_fcm_stokes(f);
  memcpy(UXS, UX, f->total * sizeof(double));
  memcpy(UYS, UY, f->total * sizeof(double));
  memcpy(UZS, UZ, f->total * sizeof(double));
}

/* main.cpp:371-433: the x velocity averaged over x and z on every y row, of
 * the deterministic field currently in ux plus the shear field */
void or_fcm_x_velocity_profile(or_fcm* f, double* prof) {
  const or_cfg* c = &f->c;
  const int ny = c->ny, pad = f->pad;
  for (int j = 0; j < ny; ++j) {
    double u = 0.0;
    for (int i = 0; i < c->nx; ++i) {
      const size_t row = ((size_t)i * ny + j) * pad;
      for (int k = 0; k < c->nz; ++k) u += UX[row + k] + UXS[row + k];
    }
    prof[j] = u / (double)(c->nx * c->nz);
  }
}

/* ========================================================================= */
/* Rigid bodies                                                              */
/* ========================================================================= */

/* quaternion.cpp:61-103; R row-major, R[3*r+c]. */
void or_quat_matrix(const double q[4], double R[9]) {
  const double s = q[0], v0 = q[1], v1 = q[2], v2 = q[3];
  double M[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double t = 2.0 * v0 * v0; M[1][1] -= t; M[2][2] -= t;
  t = 2.0 * v1 * v1; M[0][0] -= t; M[2][2] -= t;
  t = 2.0 * v2 * v2; M[0][0] -= t; M[1][1] -= t;
  t = 2.0 * v0 * v1; M[1][0] += t; M[0][1] += t;
  t = 2.0 * v0 * v2; M[2][0] += t; M[0][2] += t;
  t = 2.0 * v1 * v2; M[1][2] += t; M[2][1] += t;
  t = 2.0 * s * v2; M[1][0] += t; M[0][1] -= t;
  t = 2.0 * s * v1; M[2][0] -= t; M[0][2] += t;
  t = 2.0 * s * v0; M[2][1] += t; M[1][2] -= t;
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc) R[3 * r + cc] = M[r][cc];
}

static void matvec3(const double R[9], const double* v, double* o) {
  o[0] = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  o[1] = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  o[2] = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
}

static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

/* collection_of_rigid_bodies.cpp:155-173 */
void or_fill_blob_pos(const or_cfg* c, int NB, int nbpb, const double* blob_ref,
                      const double* x, const double* q, double* Y) {
  const double L[3] = {c->dx * c->nx, c->dx * c->ny, c->dx * c->nz};
  for (int n = 0; n < NB; ++n) {
    double R[9];
    or_quat_matrix(q + 4 * n, R);
    for (int m = 0; m < nbpb; ++m) {
      double r[3];
      matvec3(R, blob_ref + 3 * m, r);
      for (int d = 0; d < 3; ++d) {
        const double p = x[3 * n + d] + r[d];
        Y[3 * ((size_t)n * nbpb + m) + d] = p - L[d] * floor(p / L[d]);
      }
    }
  }
}

/* collection_of_rigid_bodies.cpp:707-731 */
void or_mul_K(int NB, int nbpb, const double* blob_ref, const double* q, const double* in6, double* out3) {
  for (int i = 0; i < NB; ++i) {
    double R[9];
    or_quat_matrix(q + 4 * i, R);
    for (int j = 0; j < nbpb; ++j) {
      double r[3], w[3];
      matvec3(R, blob_ref + 3 * j, r);
      cross3(in6 + 6 * i + 3, r, w);
      const size_t jp = 3 * ((size_t)nbpb * i + j);
      for (int d = 0; d < 3; ++d) out3[jp + d] = in6[6 * i + d] + w[d];
    }
  }
}

/* collection_of_rigid_bodies.cpp:733-757 */
void or_mul_KT(int NB, int nbpb, const double* blob_ref, const double* q, const double* in3, double* out6) {
  for (int i = 0; i < NB; ++i) {
    double R[9], acc[6] = {0, 0, 0, 0, 0, 0};
    or_quat_matrix(q + 4 * i, R);
    for (int j = 0; j < nbpb; ++j) {
      double r[3], w[3];
      matvec3(R, blob_ref + 3 * j, r);
      const double* l = in3 + 3 * ((size_t)nbpb * i + j);
      cross3(r, l, w);
      for (int d = 0; d < 3; ++d) { acc[d] += l[d]; acc[3 + d] += w[d]; }
    }
    for (int d = 0; d < 6; ++d) out6[6 * i + d] = acc[d];
  }
}

/* ========================================================================= */
/* Small dense algebra (restating Armadillo inv_sympd / inv)                 */
/* ========================================================================= */

/* Inverse of an SPD matrix by Cholesky (lower triangle read), result exactly
 * symmetric.  Returns 0 on success, 1 if not positive definite. */
/* ------------------------------------------------------------------------ */
/* time-step pieces                                                          */
/* ------------------------------------------------------------------------ */
void or_barrier_forces(const or_cfg* c, int Nb, int nbpb, const double* Y, double cutoff, double* F) {
  const double L[3] = {c->dx * c->nx, c->dx * c->ny, c->dx * c->nz};
  const double U0 = 20.0 * c->kT, b = 0.5 * c->a, two_a = 2.0 * c->a;  /* linked_list.cpp:367-368 */
  const double cut2 = cutoff * cutoff;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nb; ++i) {
    double f[3] = {0.0, 0.0, 0.0};
    const int bi = i / nbpb;
    for (int j = 0; j < Nb; ++j) {
      if (j / nbpb == bi) continue;  /* blobs of one rigid body do not repel (:215) */
      double d[3];
      for (int k = 0; k < 3; ++k) {
        const double dd = Y[3 * i + k] - Y[3 * j + k];
        d[k] = dd - L[k] * trunc(2.0 * dd / L[k]);
      }
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(r2 < cut2) || r2 == 0.0) continue;
      const double r = sqrt(r2);
      const double mag = r < two_a ? U0 / b : U0 * exp((two_a - r) / b) / b;  /* :377 */
      for (int k = 0; k < 3; ++k) f[k] += mag * (d[k] / r);
    }
    if (c->slip_channel) {
      /* linear wall springs of range 2.2a (collection_of_rigid_bodies.cpp:640-662) */
      const double Lw = 2.2 * c->a, Kw = (40.0 * c->kT / c->a) / Lw, y = Y[3 * i + 1];
      if (y <= Lw) f[1] += Kw * (Lw - y);
      else if ((0.5 * L[1] - y <= Lw) && (y <= 0.5 * L[1])) f[1] -= Kw * (Lw - 0.5 * L[1] + y);
    }
    for (int k = 0; k < 3; ++k) F[3 * i + k] = f[k];
  }
}

/* linked_list::initialise / populate / apply_barrier_forces
 * (linked_list.cpp:23-70, 80-102, 192-281): an ncell^3 cell grid, cell of a
 * blob floor(ncell (Y/L)) per axis (:74-78), head/list chains filled in blob
 * order, then for every cell each of its blobs against the later blobs of
 * its own chain (raw separation, :205-227) and against every blob of the 13
 * half-shell cells (nearest image, :231-275), at any distance; each pair's
 * force is added to the first blob and subtracted from the second, in the
 * reference's loop order.  Wall springs as or_barrier_forces. */
static int ll_cell_id(int i, int j, int k, int n) {
  const int im = (int)fmod(i + n, n), jm = (int)fmod(j + n, n), km = (int)fmod(k + n, n);
  return im + jm * n + km * n * n;  /* get_cell_id, :14-22 */
}

void or_barrier_forces_cells(const or_cfg* c, int Nb, int nbpb, const double* Y, int ncell, double* F) {
  const double L[3] = {c->dx * c->nx, c->dx * c->ny, c->dx * c->nz};
  const double U0 = 20.0 * c->kT, b = 0.5 * c->a, two_a = 2.0 * c->a;
  const int n = ncell, ncells = n * n * n;
  int* head = (int*)malloc(sizeof(int) * ncells);
  int* list = (int*)malloc(sizeof(int) * (Nb > 0 ? Nb : 1));
  int* map = (int*)malloc(sizeof(int) * 13 * ncells);
  /* the half-shell map, :44-68 */
  static const int off[13][3] = {{1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {-1, 1, 0}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
                                 {-1, 1, -1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1}, {0, 0, 1}};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        const int me = ll_cell_id(i, j, k, n);
        for (int m = 0; m < 13; ++m) map[13 * me + m] = ll_cell_id(i + off[m][0], j + off[m][1], k + off[m][2], n);
      }
  for (int i = 0; i < ncells; ++i) head[i] = -1;
  for (int i = 0; i < Nb; ++i) {
    const int cx = (int)floor(n * (Y[3 * i] / L[0]));
    const int cy = (int)floor(n * (Y[3 * i + 1] / L[1]));
    const int cz = (int)floor(n * (Y[3 * i + 2] / L[2]));
    const int cell = cx + cy * n + cz * n * n;  /* find_containing_cell, :74-78 */
    list[i] = head[cell];
    head[cell] = i;
  }
  for (int i = 0; i < 3 * Nb; ++i) F[i] = 0.0;
  for (int cell = 0; cell < ncells; ++cell) {
    for (int b1 = head[cell]; b1 != -1; b1 = list[b1]) {
      const int body1 = b1 / nbpb;
      for (int b2 = list[b1]; b2 != -1; b2 = list[b2]) {
        if (b2 / nbpb == body1) continue;
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = Y[3 * b1 + k] - Y[3 * b2 + k];
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (r == 0.0) continue;
        const double mag = r < two_a ? U0 / b : U0 * exp((two_a - r) / b) / b;
        for (int k = 0; k < 3; ++k) {
          F[3 * b1 + k] += mag * (d[k] / r);
          F[3 * b2 + k] -= mag * (d[k] / r);
        }
      }
      for (int m = 0; m < 13; ++m)
        for (int b2 = head[map[13 * cell + m]]; b2 != -1; b2 = list[b2]) {
          if (b2 / nbpb == body1) continue;
          double d[3];
          for (int k = 0; k < 3; ++k) {
            const double dd = Y[3 * b1 + k] - Y[3 * b2 + k];
            d[k] = dd - L[k] * trunc(2.0 * dd / L[k]);
          }
          const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          if (r == 0.0) continue;
          const double mag = r < two_a ? U0 / b : U0 * exp((two_a - r) / b) / b;
          for (int k = 0; k < 3; ++k) {
            F[3 * b1 + k] += mag * (d[k] / r);
            F[3 * b2 + k] -= mag * (d[k] / r);
          }
        }
    }
  }
  if (c->slip_channel) {
    const double Lw = 2.2 * c->a, Kw = (40.0 * c->kT / c->a) / Lw;
    for (int i = 0; i < Nb; ++i) {
      const double y = Y[3 * i + 1];
      if (y <= Lw) F[3 * i + 1] += Kw * (Lw - y);
      else if ((0.5 * L[1] - y <= Lw) && (y <= 0.5 * L[1])) F[3 * i + 1] -= Kw * (Lw - 0.5 * L[1] + y);
    }
  }
  free(head);
  free(list);
  free(map);
}

void or_lie_exp(const double u[3], double q[4]) {
  const double theta = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  if (theta < 1e-10) {
    q[0] = 1.0; q[1] = q[2] = q[3] = 0.0;
    return;
  }
  const double cs = cos(0.5 * theta), sn = sin(0.5 * theta);
  q[0] = cs;
  for (int k = 0; k < 3; ++k) q[1 + k] = sn * u[k] / theta;
}

void or_quat_mul(const double p[4], const double q[4], double out[4]) {
  const double s = p[0] * q[0] - (p[1] * q[1] + p[2] * q[2] + p[3] * q[3]);
  double v[3], cr[3];
  cross3(p + 1, q + 1, cr);
  for (int k = 0; k < 3; ++k) v[k] = p[0] * q[1 + k] + q[0] * p[1 + k] + cr[k];
  out[0] = s;
  for (int k = 0; k < 3; ++k) out[1 + k] = v[k];
}

void or_end_of_step_update(const or_cfg* c, int NB, const double* V6, double s, double* x, double* q) {
  const double L[3] = {c->dx * c->nx, c->dx * c->ny, c->dx * c->nz};
  for (int n = 0; n < NB; ++n) {
    double w[3], e[4], qn[4];
    for (int k = 0; k < 3; ++k) {
      double v = x[3 * n + k] + s * V6[6 * n + k];
      x[3 * n + k] = v - L[k] * floor(v / L[k]);
      w[k] = s * V6[6 * n + 3 + k];
    }
    or_lie_exp(w, e);
    or_quat_mul(e, q + 4 * n, qn);
    for (int k = 0; k < 4; ++k) q[4 * n + k] = qn[k];
  }
}

int or_inv_sympd(int n, const double* A, double* Ainv) {
  double* L = (double*)calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(s > 0.0)) { free(L); return 1; }
    const double d = sqrt(s);
    L[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = t / d;
    }
  }
  /* Solve L L^T X = I column by column. */
  double* col = (double*)malloc(sizeof(double) * n);
  for (int cidx = 0; cidx < n; ++cidx) {
    for (int i = 0; i < n; ++i) {
      double t = (i == cidx) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) t -= L[(size_t)i * n + k] * col[k];
      col[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = col[i];
      for (int k = i + 1; k < n; ++k) t -= L[(size_t)k * n + i] * col[k];
      col[i] = t / L[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) Ainv[(size_t)i * n + cidx] = col[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = 0.5 * (Ainv[(size_t)i * n + j] + Ainv[(size_t)j * n + i]);
      Ainv[(size_t)i * n + j] = Ainv[(size_t)j * n + i] = s;
    }
  free(col);
  free(L);
  return 0;
}

/* Moore-Penrose pseudo-inverse of a symmetric matrix by cyclic Jacobi
 * eigen-decomposition; eigenvalues with |lambda| <= rtol*max|lambda| are
 * treated as zero.  Equals the inverse for a well-conditioned matrix.
 * Returns the number of eigenvalues dropped. */
int or_pinv_sym(int n, const double* A, double* Ainv, double rtol) {
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* v = (double*)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n * n; ++i) a[i] = A[i];
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) a[(size_t)i * n + j] = a[(size_t)j * n + i] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int i = 0; i < n; ++i) {
      dia += a[(size_t)i * n + i] * a[(size_t)i * n + i];
      for (int j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    }
    if (off <= 1e-32 * dia || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = cs * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = cs * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = cs * vkp - sn * vkq;
          v[(size_t)k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < n; ++i) lmax = fmax(lmax, fabs(a[(size_t)i * n + i]));
  int dropped = 0;
  for (int i = 0; i < n * n; ++i) Ainv[i] = 0.0;
  for (int k = 0; k < n; ++k) {
    const double lam = a[(size_t)k * n + k];
    if (fabs(lam) <= rtol * lmax) { ++dropped; continue; }
    const double il = 1.0 / lam;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) Ainv[(size_t)i * n + j] += v[(size_t)i * n + k] * il * v[(size_t)j * n + k];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = 0.5 * (Ainv[(size_t)i * n + j] + Ainv[(size_t)j * n + i]);
      Ainv[(size_t)i * n + j] = Ainv[(size_t)j * n + i] = s;
    }
  free(a);
  free(v);
  return dropped;
}

/* Dense LU with partial pivoting; returns 1 if singular. */
int or_solve_dense(int n, const double* A, const double* b, double* x) {
  double* M = (double*)malloc(sizeof(double) * (size_t)n * n);
  memcpy(M, A, sizeof(double) * (size_t)n * n);
  memcpy(x, b, sizeof(double) * n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(M[(size_t)i * n + k]) > fabs(M[(size_t)p * n + k])) p = i;
    if (M[(size_t)p * n + k] == 0.0) { free(M); return 1; }
    if (p != k) {
      for (int j = 0; j < n; ++j) { const double t = M[(size_t)k * n + j]; M[(size_t)k * n + j] = M[(size_t)p * n + j]; M[(size_t)p * n + j] = t; }
      const double t = x[k]; x[k] = x[p]; x[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      const double l = M[(size_t)i * n + k] / M[(size_t)k * n + k];
      for (int j = k; j < n; ++j) M[(size_t)i * n + j] -= l * M[(size_t)k * n + j];
      x[i] -= l * x[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = x[i];
    for (int j = i + 1; j < n; ++j) t -= M[(size_t)i * n + j] * x[j];
    x[i] = t / M[(size_t)i * n + i];
  }
  free(M);
  return 0;
}

/* ========================================================================= */
/* Reference matrices, fcm_fluid_solver.cpp:842-881                          */
/* ========================================================================= */

/* rcross(x) v = v x x, rcross.cpp:5-9; K(3m..,0..2)=I, K(3m..,3..5)=rcross(r_m). */
static void build_Kref(int nbpb, const double* r, double* K /* 3nbpb x 6 */) {
  for (int m = 0; m < nbpb; ++m) {
    const double* x = r + 3 * m;
    double* K0 = K + (size_t)(3 * m) * 6;
    double* K1 = K0 + 6;
    double* K2 = K1 + 6;
    K0[0] = 1; K0[1] = 0; K0[2] = 0; K0[3] = 0;     K0[4] = x[2];  K0[5] = -x[1];
    K1[0] = 0; K1[1] = 1; K1[2] = 0; K1[3] = -x[2]; K1[4] = 0;     K1[5] = x[0];
    K2[0] = 0; K2[1] = 0; K2[2] = 1; K2[3] = x[1];  K2[4] = -x[0]; K2[5] = 0;
  }
}

/* Single body at the box centre with identity orientation
 * (collection_of_rigid_bodies.cpp:141-149, periodic branch); 3*nbpb unit-force
 * mobility solves give Msingle (:853-864); Minv = inv_sympd(Msingle) (:866);
 * N = inv(K^T Minv K) (:877), taken as a pseudo-inverse so that collinear
 * rods (axial spin in ker K) are handled.  Returns 0, or 1 if Msingle is not
 * SPD. */
int or_ref_matrices(or_fcm* f, int nbpb, const double* blob_ref, double* Msingle, double* Minv, double* N) {
  const or_cfg* c = &f->c;
  const int n3 = 3 * nbpb;
  /* collection_of_rigid_bodies.cpp:139-147: box centre, or y = LY/4 (mid-channel) */
  const double x[3] = {0.5 * c->dx * c->nx, (c->slip_channel ? 0.25 : 0.5) * c->dx * c->ny, 0.5 * c->dx * c->nz};
  const double q[4] = {1.0, 0.0, 0.0, 0.0};
  double* Y = (double*)malloc(sizeof(double) * n3);
  or_fill_blob_pos(c, 1, nbpb, blob_ref, x, q, Y);
  or_fcm_set_up_gaussians(f, Y, nbpb);
  double* e = (double*)calloc(n3, sizeof(double));
  double* V = (double*)malloc(sizeof(double) * n3);
  for (int n = 0; n < n3; ++n) {
    if (n > 0) e[n - 1] = 0.0;
    e[n] = 1.0;
    or_fcm_mobility(f, e, V);
    for (int i = 0; i < n3; ++i) Msingle[(size_t)i * n3 + n] = V[i];
  }
  int status = or_inv_sympd(n3, Msingle, Minv);
  double* K = (double*)malloc(sizeof(double) * n3 * 6);
  build_Kref(nbpb, blob_ref, K);
  double* MK = (double*)malloc(sizeof(double) * n3 * 6);
  double KtMK[36];
  for (int i = 0; i < n3; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < n3; ++k) s += Minv[(size_t)i * n3 + k] * K[(size_t)k * 6 + j];
      MK[(size_t)i * 6 + j] = s;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < n3; ++k) s += K[(size_t)k * 6 + i] * MK[(size_t)k * 6 + j];
      KtMK[i * 6 + j] = s;
    }
  or_pinv_sym(6, KtMK, N, 1e-9);
  free(Y); free(e); free(V); free(K); free(MK);
  return status;
}

/* ========================================================================= */
/* GMRES saddle-point solve, gmres_solver.cpp                                */
/* ========================================================================= */

/* Fixed-chunk dot product: the result does not depend on the thread count. */
static double dotp(size_t n, const double* a, const double* b) {
  enum { NCH = 64 };
  double part[NCH];
#pragma omp parallel for schedule(static)
  for (int ch = 0; ch < NCH; ++ch) {
    const size_t lo = n * ch / NCH, hi = n * (ch + 1) / NCH;
    double s = 0.0;
    for (size_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[ch] = s;
  }
  double s = 0.0;
  for (int ch = 0; ch < NCH; ++ch) s += part[ch];
  return s;
}

/* gmres_solver.cpp:72-83 */
void or_saddle_matvec(or_fcm* f, const or_bodies* b, const double* in, double* out) {
  const int nb = b->NB * b->nbpb;
  const size_t n3 = 3 * (size_t)nb;
  or_fcm_mobility(f, in, out);
  double* kv = (double*)malloc(sizeof(double) * n3);
  or_mul_K(b->NB, b->nbpb, b->blob_ref, b->q, in + n3, kv);
  for (size_t i = 0; i < n3; ++i) out[i] -= kv[i];
  or_mul_KT(b->NB, b->nbpb, b->blob_ref, b->q, in, out + n3);
  for (size_t i = 0; i < 6 * (size_t)b->NB; ++i) out[n3 + i] = -out[n3 + i];
  free(kv);
}

/* gmres_solver.cpp:85-146, rotating the reference matrices per body exactly
 * as the reference does. */
void or_apply_Pinv(const or_bodies* b, const double* in, double* out) {
  const int nbpb = b->nbpb, n3 = 3 * nbpb, NB = b->NB;
  const size_t nlam = 3 * (size_t)NB * nbpb;
#pragma omp parallel
  {
    double* Minv = (double*)malloc(sizeof(double) * (size_t)n3 * n3);
    double* K = (double*)malloc(sizeof(double) * (size_t)n3 * 6);
    double* t = (double*)malloc(sizeof(double) * n3);
    double* t2 = (double*)malloc(sizeof(double) * n3);
#pragma omp for schedule(static)
    for (int n = 0; n < NB; ++n) {
      double Q[9];
      or_quat_matrix(b->q + 4 * n, Q);
      /* K with rotated positions (:97-115) */
      double* rr = (double*)malloc(sizeof(double) * n3);
      for (int m = 0; m < nbpb; ++m) matvec3(Q, b->blob_ref + 3 * m, rr + 3 * m);
      build_Kref(nbpb, rr, K);
      free(rr);
      /* Minv blocks (:119-129) */
      memcpy(Minv, b->Minv, sizeof(double) * (size_t)n3 * n3);
      for (int m = 0; m < nbpb; ++m)
        for (int l = m; l < nbpb; ++l) {
          double B[9], T[9];
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) B[3 * r + cc] = Minv[(size_t)(3 * m + r) * n3 + 3 * l + cc];
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k) s += Q[3 * r + k] * B[3 * k + cc];
              T[3 * r + cc] = s;
            }
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k) s += T[3 * r + k] * Q[3 * cc + k];
              Minv[(size_t)(3 * m + r) * n3 + 3 * l + cc] = s;
            }
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
              Minv[(size_t)(3 * l + r) * n3 + 3 * m + cc] = Minv[(size_t)(3 * m + cc) * n3 + 3 * l + r];
        }
      /* N blocks (:131-134) */
      double Nn[36];
      memcpy(Nn, b->N, sizeof(Nn));
      for (int bi = 0; bi < 2; ++bi)
        for (int bj = 0; bj < 2; ++bj) {
          double B[9], T[9];
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) B[3 * r + cc] = Nn[(3 * bi + r) * 6 + 3 * bj + cc];
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k) s += Q[3 * r + k] * B[3 * k + cc];
              T[3 * r + cc] = s;
            }
          for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc) {
              double s = 0.0;
              for (int k = 0; k < 3; ++k) s += T[3 * r + k] * Q[3 * cc + k];
              Nn[(3 * bi + r) * 6 + 3 * bj + cc] = s;
            }
        }
      /* Schur solve (:137-140) */
      const double* inl = in + (size_t)n3 * n;
      const double* inv = in + nlam + 6 * (size_t)n;
      for (int i = 0; i < n3; ++i) {
        double s = 0.0;
        for (int k = 0; k < n3; ++k) s += Minv[(size_t)i * n3 + k] * inl[k];
        t[i] = s;
      }
      double sv[6], ov[6];
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < n3; ++k) s += K[(size_t)k * 6 + j] * t[k];
        sv[j] = -inv[j] - s;
      }
      for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += Nn[i * 6 + k] * sv[k];
        ov[i] = s;
      }
      for (int i = 0; i < n3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += K[(size_t)i * 6 + k] * ov[k];
        t2[i] = s + inl[i];
      }
      double* outl = out + (size_t)n3 * n;
      for (int i = 0; i < n3; ++i) {
        double s = 0.0;
        for (int k = 0; k < n3; ++k) s += Minv[(size_t)i * n3 + k] * t2[k];
        outl[i] = s;
      }
      for (int j = 0; j < 6; ++j) out[nlam + 6 * (size_t)n + j] = ov[j];
    }
    free(Minv); free(K); free(t); free(t2);
  }
}

/* gmres_solver.cpp:198-210: modified Gram-Schmidt of the new vector qn against
 * the iter basis vectors in Q (n each), h_i = <qn, q_i> then qn -= h_i q_i in
 * turn, hcol[iter] = ||qn|| and qn normalised when that is > 0. */
static double mgs_orth(size_t n, const double* Q, double* qn, int iter, double* hcol) {
  for (int i = 0; i < iter; ++i) {
    const double h = dotp(n, qn, Q + n * i);
    hcol[i] = h;
    const double* qi = Q + n * i;
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)n; ++r) qn[r] -= h * qi[r];
  }
  const double q_norm = sqrt(dotp(n, qn, qn));
  hcol[iter] = q_norm;
  if (q_norm > 0.0) {
#pragma omp parallel for schedule(static)
    for (long r = 0; r < (long)n; ++r) qn[r] /= q_norm;
  }
  return q_norm;
}

double or_gmres_orth_seconds(size_t n, int iter) {
  double* Q = (double*)malloc(sizeof(double) * n * (size_t)iter);
  double* qn = (double*)malloc(sizeof(double) * n);
  double* hcol = (double*)malloc(sizeof(double) * (iter + 1));
#pragma omp parallel for schedule(static)
  for (long r = 0; r < (long)n; ++r) {
    for (int i = 0; i < iter; ++i) Q[(size_t)i * n + r] = sin(0.37 * (double)r + 1.3 * i);
    qn[r] = cos(0.11 * (double)r);
  }
  const double t0 = omp_get_wtime();
  mgs_orth(n, Q, qn, iter, hcol);
  const double t1 = omp_get_wtime();
  free(Q); free(qn); free(hcol);
  return t1 - t0;
}

/* gmres_solver.cpp:148-306: unrestarted, right-preconditioned GMRES with
 * modified Gram-Schmidt, x0 = 0.  Returns 0 converged, 1 iteration cap hit,
 * 2 zero right-hand side (solution 0), 3 exact breakdown (solution exact).
 * resid_hist (max_iter entries, may be NULL) receives |beta_j|/||RHS||. */
int or_gmres(or_fcm* f, const or_bodies* b, const double* rhs, double* soln,
             double tol, int max_iter_in, int* iters, double* resid_hist) {
  const size_t nlam = 3 * (size_t)b->NB * b->nbpb;
  const size_t n = nlam + 6 * (size_t)b->NB;
  const int max_iter = (int)((size_t)max_iter_in < n ? (size_t)max_iter_in : n);
  const double norm_of_RHS = sqrt(dotp(n, rhs, rhs));
  if (iters) *iters = 0;
  if (norm_of_RHS == 0.0) {
    memset(soln, 0, sizeof(double) * n);
    return 2;
  }
  double* Q = (double*)calloc(n * (size_t)(max_iter + 1), sizeof(double));
  double* H = (double*)calloc((size_t)(max_iter + 1) * (max_iter + 1), sizeof(double));
  double* beta = (double*)calloc(max_iter + 1, sizeof(double));
  double* CS = (double*)calloc(max_iter + 1, sizeof(double));
  double* SN = (double*)calloc(max_iter + 1, sizeof(double));
  double* z = (double*)malloc(sizeof(double) * n);
#define HH(i, j) H[(size_t)(i) * (max_iter + 1) + (j)]
  for (size_t i = 0; i < n; ++i) Q[i] = rhs[i] / norm_of_RHS;
  beta[0] = norm_of_RHS;
  int status = 1, breakdown = 0;
  for (int iter = 1; iter <= max_iter; ++iter) {
    double* qn = Q + n * iter;
    or_apply_Pinv(b, Q + n * (iter - 1), z);
    or_saddle_matvec(f, b, z, qn);
    double* hcol = (double*)malloc(sizeof(double) * (iter + 1));
    const double q_norm = mgs_orth(n, Q, qn, iter, hcol);
    for (int i = 0; i <= iter; ++i) HH(i, iter - 1) = hcol[i];
    free(hcol);
    if (!(q_norm > 0.0)) breakdown = 1;
    for (int i = 1; i < iter; ++i) {
      const double temp = CS[i - 1] * HH(i - 1, iter - 1) + SN[i - 1] * HH(i, iter - 1);
      HH(i, iter - 1) = CS[i - 1] * HH(i, iter - 1) - SN[i - 1] * HH(i - 1, iter - 1);
      HH(i - 1, iter - 1) = temp;
    }
    if (HH(iter, iter - 1) == 0.0) {
      CS[iter - 1] = 1.0;
      SN[iter - 1] = 0.0;
    } else if (fabs(HH(iter, iter - 1)) > fabs(HH(iter - 1, iter - 1))) {
      const double temp = HH(iter - 1, iter - 1) / HH(iter, iter - 1);
      SN[iter - 1] = (2.0 * (double)(HH(iter, iter - 1) > 0) - 1.0) / sqrt(1.0 + temp * temp);
      CS[iter - 1] = temp * SN[iter - 1];
    } else {
      const double temp = HH(iter, iter - 1) / HH(iter - 1, iter - 1);
      CS[iter - 1] = (2.0 * (double)(HH(iter - 1, iter - 1) > 0) - 1.0) / sqrt(1.0 + temp * temp);
      SN[iter - 1] = temp * CS[iter - 1];
    }
    HH(iter - 1, iter - 1) = CS[iter - 1] * HH(iter - 1, iter - 1) + SN[iter - 1] * HH(iter, iter - 1);
    HH(iter, iter - 1) = 0.0;
    beta[iter] = -SN[iter - 1] * beta[iter - 1];
    beta[iter - 1] = CS[iter - 1] * beta[iter - 1];
    const double relative_error = fabs(beta[iter]) / norm_of_RHS;
    if (resid_hist) resid_hist[iter - 1] = relative_error;
    if (relative_error < tol || iter == max_iter || breakdown) {
      double* y = (double*)calloc(iter, sizeof(double));
      y[iter - 1] = beta[iter - 1] / HH(iter - 1, iter - 1);
      for (int i = iter - 2; i >= 0; --i) {
        y[i] = beta[i];
        for (int j = iter - 1; j > i; --j) y[i] -= HH(i, j) * y[j];
        y[i] /= HH(i, i);
      }
      double* s = (double*)calloc(n, sizeof(double));
      for (int i = 0; i < iter; ++i) {
        const double* qi = Q + n * i;
        const double yi = y[i];
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)n; ++r) s[r] += yi * qi[r];
      }
      or_apply_Pinv(b, s, soln);
      free(s);
      free(y);
      if (iters) *iters = iter;
      status = breakdown ? 3 : (relative_error < tol ? 0 : 1);
      break;
    }
  }
#undef HH
  free(Q); free(H); free(beta); free(CS); free(SN); free(z);
  return status;
}
