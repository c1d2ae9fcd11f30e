// fcm_synth.cpp -- seeded synthetic rigid-rod configurations (host only).
//
// A rod of nbpb spheres of radius a has body-frame offsets
// r_m = (2a (m - (nbpb-1)/2), 0, 0).  Centres are uniform in the box; a trial
// placement is rejected when a blob lies closer than 2a (minimum image) to a
// blob of an already placed rod -- the overlap criterion of the reference's
// cell list (linked_list.cpp:104-190).  After max_attempts failures the last
// trial is kept and counted as overlapping.  Orientations are normalised
// 4-Gaussian quaternions (quaternion.cpp:10-17); forces and torques are
// N(0,1) with the torque component along the rod axis removed (the axial
// spin of a collinear rod is in the kernel of K).
//
// Random numbers: one sequential SplitMix64 stream (Steele, Lea & Flood,
// 2014) seeded with `seed`; uniforms (u >> 11) 2^-53; Gaussians by the
// Marsaglia polar method, first variate (parallel_random_numbers.cpp:8-19).
#include "../../include/fcm_hip.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Stream {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    for (;;) {
      const double x1 = 2.0 * uniform() - 1.0, x2 = 2.0 * uniform() - 1.0;
      const double w = x1 * x1 + x2 * x2;
      if (w < 1.0 && w > 0.0) return x1 * std::sqrt(-2.0 * std::log(w) / w);
    }
  }
};

void quat_matrix(const double* q, double* R) {
  const double s = q[0], v0 = q[1], v1 = q[2], v2 = q[3];
  R[0] = 1 - 2 * (v1 * v1 + v2 * v2); R[1] = 2 * (v0 * v1 - s * v2);     R[2] = 2 * (v0 * v2 + s * v1);
  R[3] = 2 * (v0 * v1 + s * v2);     R[4] = 1 - 2 * (v0 * v0 + v2 * v2); R[5] = 2 * (v1 * v2 - s * v0);
  R[6] = 2 * (v0 * v2 - s * v1);     R[7] = 2 * (v1 * v2 + s * v0);     R[8] = 1 - 2 * (v0 * v0 + v1 * v1);
}

struct CellList {
  int nc[3];
  double cs[3], L[3];
  std::vector<std::vector<int>> cells;
  std::vector<double> pos;  // 3 per stored blob
  CellList(const double* Lb, double cut) {
    for (int d = 0; d < 3; ++d) {
      L[d] = Lb[d];
      nc[d] = std::max(1, (int)std::floor(L[d] / cut));
      cs[d] = L[d] / nc[d];
    }
    cells.resize((size_t)nc[0] * nc[1] * nc[2]);
  }
  int cell_of(const double* p, int* c) const {
    for (int d = 0; d < 3; ++d) {
      c[d] = (int)std::floor(p[d] / cs[d]);
      if (c[d] >= nc[d]) c[d] = nc[d] - 1;
      if (c[d] < 0) c[d] = 0;
    }
    return (c[0] * nc[1] + c[1]) * nc[2] + c[2];
  }
  bool near(const double* p, double cut2) const {
    int c[3];
    cell_of(p, c);
    const int rx = nc[0] >= 3 ? 1 : nc[0] - 1, ry = nc[1] >= 3 ? 1 : nc[1] - 1, rz = nc[2] >= 3 ? 1 : nc[2] - 1;
    for (int i = -rx; i <= rx; ++i)
      for (int j = -ry; j <= ry; ++j)
        for (int k = -rz; k <= rz; ++k) {
          const int a = (c[0] + i + nc[0]) % nc[0], b = (c[1] + j + nc[1]) % nc[1], e = (c[2] + k + nc[2]) % nc[2];
          for (int id : cells[((size_t)a * nc[1] + b) * nc[2] + e]) {
            double r2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              double dd = p[d] - pos[3 * (size_t)id + d];
              dd -= L[d] * std::round(dd / L[d]);
              r2 += dd * dd;
            }
            if (r2 < cut2) return true;
          }
        }
    return false;
  }
  void add(const double* p) {
    int c[3];
    const int cell = cell_of(p, c);
    const int id = (int)(pos.size() / 3);
    pos.insert(pos.end(), p, p + 3);
    cells[cell].push_back(id);
  }
};

}  // namespace

extern "C" int fcm_synth_rods(uint64_t seed, int NB, int nbpb, double a, const double L[3], int max_attempts,
                              double* blob_ref, double* x, double* q, double* F6, int* n_overlapping) {
  if (NB < 1 || nbpb < 1 || !(a > 0) || !L || !blob_ref || !x || !q) return FCM_ERR_ARG;
  for (int m = 0; m < nbpb; ++m) {
    blob_ref[3 * m] = 2.0 * a * (m - 0.5 * (nbpb - 1));
    blob_ref[3 * m + 1] = 0.0;
    blob_ref[3 * m + 2] = 0.0;
  }
  Stream rng{seed};
  CellList cl(L, 2.0 * a);
  const double cut2 = 4.0 * a * a;
  int overl = 0;
  std::vector<double> trial(3 * (size_t)nbpb);
  for (int n = 0; n < NB; ++n) {
    bool ok = false;
    double xc[3], qc[4];
    for (int att = 0; att < std::max(1, max_attempts) && !ok; ++att) {
      for (int d = 0; d < 3; ++d) xc[d] = rng.uniform() * L[d];
      double nq = 0.0;
      for (int i = 0; i < 4; ++i) { qc[i] = rng.normal(); nq += qc[i] * qc[i]; }
      nq = std::sqrt(nq);
      for (int i = 0; i < 4; ++i) qc[i] /= nq;
      double R[9];
      quat_matrix(qc, R);
      ok = true;
      for (int m = 0; m < nbpb; ++m) {
        const double* r = blob_ref + 3 * m;
        for (int d = 0; d < 3; ++d) {
          const double p = xc[d] + R[3 * d] * r[0] + R[3 * d + 1] * r[1] + R[3 * d + 2] * r[2];
          trial[3 * m + d] = p - L[d] * std::floor(p / L[d]);
        }
        if (ok && cl.near(&trial[3 * m], cut2)) ok = false;
      }
    }
    if (!ok) ++overl;
    for (int m = 0; m < nbpb; ++m) cl.add(&trial[3 * m]);
    for (int d = 0; d < 3; ++d) x[3 * n + d] = xc[d];
    for (int i = 0; i < 4; ++i) q[4 * n + i] = qc[i];
  }
  if (F6) {
    for (int n = 0; n < NB; ++n) {
      double* f = F6 + 6 * (size_t)n;
      for (int i = 0; i < 6; ++i) f[i] = rng.normal();
      double R[9];
      quat_matrix(q + 4 * n, R);
      const double e[3] = {R[0], R[3], R[6]};  // body x axis in the world frame
      const double tdot = f[3] * e[0] + f[4] * e[1] + f[5] * e[2];
      for (int d = 0; d < 3; ++d) f[3 + d] -= tdot * e[d];
    }
  }
  if (n_overlapping) *n_overlapping = overl;
  return FCM_OK;
}
