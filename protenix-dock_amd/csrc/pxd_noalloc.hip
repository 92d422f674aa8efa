// pxd_noalloc.hip — pxd_lbfgs_batch_noalloc: the batched L-BFGS entry point
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).  Its work-queue heads and L-BFGS histories
// live in a file-scope device array, `scratch`, of a fixed size; each device
// the library runs on holds its own copy until the process ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "pxd_internal.hpp"
#include "pxd_kernels.hpp"
#include "pxdock.h"

using namespace pxd;

extern "C" pxd_status pxd_set_error_(pxd_status st, const char* msg);

namespace {

constexpr size_t kScratchBytes = size_t(PXD_NOALLOC_SCRATCH_MB) << 20;
constexpr size_t kHeadBytes = 256;  // queue heads, one int per size class
constexpr int kMaxDevices = 64;

}  // namespace

// The scratch array.  Its name is a host-side stand-in in host code: the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call made
// with that device current.
pxd_status scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return pxd_set_error_(PXD_ERR_RANGE, "lbfgs_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return pxd_set_error_(PXD_ERR_HIP, e != hipSuccess ? hipGetErrorString(e)
                                                         : "lbfgs_noalloc: no scratch address");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return PXD_OK;
}

}  // namespace

// Declared in pxdock.h.
extern "C" pxd_status pxd_lbfgs_batch_noalloc(const pxd_receptor* r, const pxd_grid* g,
                                              const pxd_ligand_batch* b,
                                              const pxd_lbfgs_params* prm, int32_t n_poses,
                                              const int32_t* pose_lig,
                                              const int32_t* pose_atom_off, float* xyz,
                                              float* tors, double* energy, int32_t* nevals,
                                              int32_t* converged, void* stream) {
  LbfgsArgs a;
  int n_slots = 0;
  pxd_status st = lbfgs_prepare(r, g, b, prm, n_poses, pose_lig, pose_atom_off, xyz, tors, energy,
                                nevals, converged, &a, &n_slots);
  if (st != PXD_OK || n_poses == 0) return st;
  // as many work slots as the tuning asks for and the scratch holds
  const size_t per_slot = lbfgs_slot_floats(prm->history) * sizeof(float);
  const size_t fit = (kScratchBytes - kHeadBytes) / per_slot;
  if (fit < 1) return pxd_set_error_(PXD_ERR_RANGE, "lbfgs_noalloc: one history slot exceeds the scratch");
  n_slots = (int)std::min<size_t>((size_t)n_slots, fit);
  void* base = nullptr;
  st = scratch_address(&base);
  if (st != PXD_OK) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  a.counter = static_cast<int*>(base);
  a.hist = reinterpret_cast<float*>(static_cast<unsigned char*>(base) + kHeadBytes);
  // queue heads start at zero: a fill kernel (no memset node, no allocation)
  launch_fill_int(a.counter, kLbfgsClasses, 0, s);
  launch_lbfgs(r->dev, b->dev, grid_device(g), a, n_slots, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return pxd_set_error_(PXD_ERR_HIP, hipGetErrorString(e));
  return PXD_OK;
}
