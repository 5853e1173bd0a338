// The allocation-free coarse solve: msfem_coarse_solve_device_noalloc, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  It runs the same
// kernels as msfem_coarse_solve_device (the multigrid-preconditioned CG, or
// Jacobi-PCG for sizes that do not coarsen) with their scratch in a fixed
// device array instead of context-owned memory.
//
// `scratch` is a file-scope device array of kNoAllocScratchBytes (16 MiB); each
// GPU the library runs on holds its own copy until the process ends.  Its
// device address is looked up once per device with hipGetSymbolAddress (from
// msfem_ctx_create, with that device current) and only that address is passed
// to kernels; the host-side name is never used as a device pointer.  Every call
// on one GPU shares that GPU's array: calls that overlap on one GPU (on
// different streams) race.  The kernels write every scratch value before they
// read it, so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "msfem.h"
#include "msfem_kernels.hpp"

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

} // namespace

namespace msfem {

// Looks up (once) this device's copy of `scratch`; the device must be current.
hipError_t noalloc_register_device(int device)
{
  if (device < 0 || device >= kMaxDevices)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[device])
    return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess)
    return e;
  g_ptr[device] = (double *)p;
  return hipSuccess;
}

double *noalloc_scratch(int device)
{
  if (device < 0 || device >= kMaxDevices)
    return nullptr;
  std::lock_guard<std::mutex> lock(g_mu);
  return g_ptr[device];
}

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

} // namespace msfem
