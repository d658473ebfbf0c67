// The static device workspace of the no-allocation entry points, and the per-device lookup of
// its address. In host code the array's name is only a stand-in for the symbol: every kernel
// argument, copy or fill uses the address returned by noalloc_workspace(), never the name.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.h"

namespace ExaDG
{
namespace
{
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex    g_mutex;
double *      g_address[kMaxDevices] = {};
}  // namespace

double * noalloc_workspace()
{
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if(dev < 0 || dev >= kMaxDevices)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "device index beyond the no-allocation workspace table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if(g_address[dev] == nullptr)
  {
    void * ptr = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if(ptr == nullptr)
      EXADG_THROW(EXADG_ERR_HIP, "no device address for the no-allocation workspace");
    g_address[dev] = static_cast<double *>(ptr);
  }
  return g_address[dev];
}
}  // namespace ExaDG
