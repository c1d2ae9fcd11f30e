// fcm_noalloc.hip -- the barrier forces of ll_barrier_forces for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The cell list (each blob's cell, per-cell counts, starts and cursors, the
// blobs sorted by cell) lives in one static device array instead of buffers
// the library allocates, so the entry point calls no hipMalloc, hipFree or
// allocating library routine.  Every GPU the library runs on holds its own
// copy of the array until the process ends; its address on a device is taken
// once, with that device current, by hipGetSymbolAddress, and only that
// address is handed to kernels.  The kernels are the ones ll_barrier_forces
// runs, on the same cells, so the forces are bitwise the same.  Every part of
// the array a call reads is written earlier in the same call by a kernel (the
// counts are cleared by a fill kernel, not by a memset), so nothing carries
// over from the last call.  Calls that overlap on one device share the array
// and race.
#include "fcm_handle.hpp"

#include <mutex>

namespace {

constexpr size_t SCRATCH_INTS = (size_t)FCM_NOALLOC_SCRATCH_BYTES / sizeof(int);
static __device__ __align__(256) int scratch[SCRATCH_INTS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[MAX_DEVICES] = {};

// the array's address on the current device, resolved on the first call there
int scratch_on_current_device(int** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fcm_host::fail(FCM_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fcm_host::fail(FCM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_ptr[dev];
  return FCM_OK;
}

}  // namespace

extern "C" {

int ll_barrier_forces_noalloc(const double* Y, int nb, int nbpb, const double L[3], double a, double kT,
                              double cutoff, int cells_per_dim, int slip_channel, double* F, void* stream) {
  using fcm_host::fail;
  if (!Y || !F || !L || nb < 0 || nbpb < 1) return fail(FCM_ERR_ARG, "bad arguments");
  if (!(L[0] > 0) || !(L[1] > 0) || !(L[2] > 0) || !(a > 0) || !(cutoff > 0))
    return fail(FCM_ERR_ARG, "box, radius and cutoff must be positive");
  if (cells_per_dim != 0 && (cells_per_dim < 3 || cells_per_dim > 1024))
    return fail(FCM_ERR_ARG, "cells_per_dim must be 0 or in 3..1024");
  if (slip_channel != 0 && slip_channel != 1) return fail(FCM_ERR_ARG, "slip_channel must be 0 or 1");
  const StericCells cg = steric_cells(L, cutoff, cells_per_dim);
  const size_t ncells = (size_t)cg.nc[0] * cg.nc[1] * cg.nc[2];
  // cell_of[nb], sorted[nb], count[ncells], start[ncells + 1], cursor[ncells]
  const size_t need = 2 * (size_t)nb + 3 * ncells + 1;
  if (need > SCRATCH_INTS)
    return fail(FCM_ERR_CAPACITY, "the cell list needs " + std::to_string(need * sizeof(int)) +
                                      " bytes; the static scratch holds " +
                                      std::to_string((size_t)FCM_NOALLOC_SCRATCH_BYTES));
  if (nb == 0) return FCM_OK;
  int* ptr = nullptr;
  if (int rc = scratch_on_current_device(&ptr)) return rc;
  StericWork w;
  w.cell_of = ptr;
  w.sorted = ptr + nb;
  w.count = ptr + 2 * (size_t)nb;
  w.start = w.count + ncells;
  w.cursor = w.start + ncells + 1;
  w.cap_cells = (int)ncells;
  HIPCHK(barrier_forces(Y, nb, nbpb, cg, cutoff, cells_per_dim > 0, kT, a, slip_channel != 0, w, F,
                        static_cast<hipStream_t>(stream)));
  return FCM_OK;
}

}  // extern "C"
