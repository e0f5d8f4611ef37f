// ugk_noalloc.hip -- the fixed device workspace of the solve entry point that
// allocates nothing (ufem_sys_solve_noalloc): one file-scope array per GPU,
// holding the Krylov vectors, the reduction partials and the solver state of
// a Jacobi PCG solve of up to UGK_NOALLOC_MAX_N unknowns.
//
// Host code never uses the array's name as a device address: the address is
// taken once per device with hipGetSymbolAddress (that device current) and
// handed out from here.  Every call on one GPU shares that GPU's array and
// finds the last call's contents in it; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "ugk.h"

#define UGK_NOALLOC_DOUBLES (6L * UGK_NOALLOC_MAX_N + UGK_PART_DOUBLES + UGK_NOALLOC_STATE_DOUBLES)

static __device__ __align__(256) double scratch[UGK_NOALLOC_DOUBLES];

namespace {
const int kMaxDevices = 64;
std::mutex g_lock;
void *g_ptr[kMaxDevices];
}  // namespace

int ugk_hip_fail(const char *what, hipError_t e);   // ugk_kernels.hip

extern "C" int ugk_noalloc_workspace(void **ptr, size_t *bytes)
{
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return ugk_hip_fail("hipGetDevice", e);
  if (dev < 0 || dev >= kMaxDevices) return ugk_hip_fail("no-alloc workspace: device index", hipErrorInvalidDevice);
  std::lock_guard<std::mutex> hold(g_lock);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return ugk_hip_fail("hipGetSymbolAddress(no-alloc workspace)", e);
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  *bytes = sizeof(double) * (size_t)UGK_NOALLOC_DOUBLES;
  return 0;
}
