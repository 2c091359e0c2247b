// Fixed device scratch of the entry points that allocate nothing (rbc_solver_solve_noalloc).
//
// The work arrays of such a call live in one file-scope device array per GPU, held by the
// library until the process ends.  Host code never uses the array's name as an address:
// the array's device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and kept; every kernel argument and copy uses that address.  Calls
// on one GPU share the array and find the previous call's contents in it; calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "noalloc.hpp"

namespace rbc {
namespace noalloc {

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_mutex;
std::vector<double*> g_ptr;  // per device: the array's device address (nullptr: not looked up)
}  // namespace

double* scratch_ptr() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_mutex);
  if ((int)g_ptr.size() <= dev) g_ptr.resize((size_t)dev + 1, nullptr);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[dev] = static_cast<double*>(p);
  }
  return g_ptr[dev];
}

}  // namespace noalloc
}  // namespace rbc
