// The fixed device workspace of the allocation-free per-individual fitness
// entry point (gpuga_fitness_compute_plain_noalloc).
//
// The workspace is a file-scope device array, so it exists once per GPU the
// library runs on, for the life of the process, and no call ever allocates
// or frees device memory for it.  Host code reaches it only through the
// device address hipGetSymbolAddress returns (the array's name is not a
// device address in host code); the address is looked up once per device,
// with that device current, and kept.
//
// Every call on one GPU shares that GPU's array: two calls that overlap on
// one GPU (two host threads, or two streams) race on it.  It keeps the last
// call's contents; the kernels that use it write every element they read
// before reading it, so nothing needs clearing between calls.
#include <mutex>
#include <vector>

#include "common.h"
#include "noalloc.h"

namespace gpuga {

namespace {

static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

std::mutex g_mu;
std::vector<void*> g_ptr; // device address of scratch, per device index

} // namespace

void* static_workspace()
{
  int dev = 0;
  GPUGA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_mu);
  if ((size_t)dev >= g_ptr.size()) g_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    GPUGA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail("the fixed device workspace has no device address");
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

} // namespace gpuga
