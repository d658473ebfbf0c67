// Streaming vector kernels (HBM-bound): diagonal inversion and precision conversion.
#include <hip/hip_runtime.h>

#include "common.h"
#include "laplace_operator.h"

namespace ExaDG
{
namespace
{
template<typename Number>
__global__ void __launch_bounds__(256) invert_diagonal_kernel(Number * d, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * blockDim.x;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
  {
    Number const v = d[i];
    d[i]           = (fabs(v) > Number(1e-10)) ? Number(1) / v : Number(1);
  }
}

inline int grid_for(int64_t n)
{
  int64_t const b = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 8192);
}
}  // namespace

namespace Poisson
{
template<typename Number>
void invert_diagonal_device(Number * d, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((invert_diagonal_kernel<Number>), dim3(grid_for(n)), dim3(256), 0, s, d, n);
  HIP_CHECK(hipGetLastError());
}

template void invert_diagonal_device<double>(double *, int64_t, hipStream_t);
template void invert_diagonal_device<float>(float *, int64_t, hipStream_t);
}  // namespace Poisson

}  // namespace ExaDG
