// fcm_dense.hpp -- small dense FP64 algebra on the host for the single-body
// reference matrices (replaces the reference's Armadillo inv_sympd / inv,
// fcm_fluid_solver.cpp:866,877).
#pragma once

#include <cmath>
#include <vector>

namespace fcm {
namespace dense {

// A^{-1} for symmetric positive definite A (n x n, row-major) via Cholesky;
// the result is exactly symmetric.  Returns 0, or 1 if A is not SPD.
inline int inv_spd(int n, const double* A, double* Ai) {
  std::vector<double> L((size_t)n * n, 0.0), col(n);
  for (int j = 0; j < n; ++j) {
    double s = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(s > 0.0)) return 1;
    const double d = std::sqrt(s);
    L[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = t / d;
    }
  }
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {
      double t = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) t -= L[(size_t)i * n + k] * col[k];
      col[i] = t / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = col[i];
      for (int k = i + 1; k < n; ++k) t -= L[(size_t)k * n + i] * col[k];
      col[i] = t / L[(size_t)i * n + i];
    }
    for (int i = 0; i < n; ++i) Ai[(size_t)i * n + c] = col[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = 0.5 * (Ai[(size_t)i * n + j] + Ai[(size_t)j * n + i]);
      Ai[(size_t)i * n + j] = Ai[(size_t)j * n + i] = s;
    }
  return 0;
}

// Moore-Penrose pseudo-inverse of a symmetric matrix by cyclic Jacobi
// rotations; eigenvalues below rtol * max |eigenvalue| count as zero.
// Returns the number of eigenvalues dropped.
inline int pinv_sym(int n, const double* A, double* Ai, double rtol) {
  std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = 0.5 * (A[(size_t)i * n + j] + A[(size_t)j * n + i]);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int i = 0; i < n; ++i) {
      dia += a[(size_t)i * n + i] * a[(size_t)i * n + i];
      for (int j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    }
    if (off == 0.0 || off <= 1e-32 * dia) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double x = a[(size_t)k * n + p], y = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = cs * x - sn * y;
          a[(size_t)k * n + q] = sn * x + cs * y;
        }
        for (int k = 0; k < n; ++k) {
          const double x = a[(size_t)p * n + k], y = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = cs * x - sn * y;
          a[(size_t)q * n + k] = sn * x + cs * y;
        }
        for (int k = 0; k < n; ++k) {
          const double x = v[(size_t)k * n + p], y = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = cs * x - sn * y;
          v[(size_t)k * n + q] = sn * x + cs * y;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < n; ++i) lmax = std::fmax(lmax, std::fabs(a[(size_t)i * n + i]));
  for (int i = 0; i < n * n; ++i) Ai[i] = 0.0;
  int dropped = 0;
  for (int k = 0; k < n; ++k) {
    const double lam = a[(size_t)k * n + k];
    if (std::fabs(lam) <= rtol * lmax) { ++dropped; continue; }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) Ai[(size_t)i * n + j] += v[(size_t)i * n + k] * v[(size_t)j * n + k] / lam;
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = 0.5 * (Ai[(size_t)i * n + j] + Ai[(size_t)j * n + i]);
      Ai[(size_t)i * n + j] = Ai[(size_t)j * n + i] = s;
    }
  return dropped;
}

// N = pinv(K^T Minv K), K(3m..,0..2) = I, K(3m..,3..5) = rcross(r_m) with
// rcross(x) v = v x x (rcross.cpp:5-9; fcm_fluid_solver.cpp:868-877)
inline int ref_N(int nbpb, const double* ref, const double* Minv, double* N) {
  const int n3 = 3 * nbpb;
  std::vector<double> K((size_t)n3 * 6, 0.0), MK((size_t)n3 * 6, 0.0);
  for (int m = 0; m < nbpb; ++m) {
    const double* x = ref + 3 * m;
    double* K0 = &K[(size_t)(3 * m) * 6];
    double* K1 = K0 + 6;
    double* K2 = K1 + 6;
    K0[0] = 1; K0[4] = x[2];  K0[5] = -x[1];
    K1[1] = 1; K1[3] = -x[2]; K1[5] = x[0];
    K2[2] = 1; K2[3] = x[1];  K2[4] = -x[0];
  }
  for (int i = 0; i < n3; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < n3; ++k) s += Minv[(size_t)i * n3 + k] * K[(size_t)k * 6 + j];
      MK[(size_t)i * 6 + j] = s;
    }
  double KtMK[36];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < n3; ++k) s += K[(size_t)k * 6 + i] * MK[(size_t)k * 6 + j];
      KtMK[i * 6 + j] = s;
    }
  return pinv_sym(6, KtMK, N, 1e-9);
}

}  // namespace dense
}  // namespace fcm
