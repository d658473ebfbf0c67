// Block-Jacobi inversion and application (see block_kernels.h).
#include <hip/hip_runtime.h>

#include <atomic>

#include "block_kernels.h"
#include "line_sweeps.h"

namespace ExaDG
{
namespace
{
// One workgroup per cell: the m x m block in LDS, Gauss-Jordan elimination in place (the blocks
// are SPD, so the pivots stay positive and no pivoting is needed). Step k: save column k, scale
// row k by 1 / pivot (its diagonal becoming 1 / pivot), then eliminate column k from every other
// row (its column-k entry becoming -a_ik / pivot).
template<typename Number>
__global__ void __launch_bounds__(256) block_invert_kernel(Number * blocks, int m, int64_t n_cells)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      A    = reinterpret_cast<Number *>(smem_raw);
  Number *      col  = A + (size_t)m * m;
  int64_t const cell = blockIdx.x;
  if(cell >= n_cells)
    return;
  int const mm  = m * m;
  Number *  src = blocks + cell * mm;
  for(int e = threadIdx.x; e < mm; e += blockDim.x)
    A[e] = src[e];
  __syncthreads();
  for(int k = 0; k < m; ++k)
  {
    for(int i = threadIdx.x; i < m; i += blockDim.x)
      col[i] = A[i * m + k];
    __syncthreads();
    Number const inv_piv = Number(1) / col[k];
    for(int j = threadIdx.x; j < m; j += blockDim.x)
      A[k * m + j] = (j == k ? Number(1) : A[k * m + j]) * inv_piv;
    __syncthreads();
    for(int e = threadIdx.x; e < mm; e += blockDim.x)
    {
      int const i = e / m, j = e - i * m;
      if(i == k)
        continue;
      A[e] = (j == k ? Number(0) : A[e]) - col[i] * A[k * m + j];
    }
    __syncthreads();
  }
  for(int e = threadIdx.x; e < mm; e += blockDim.x)
    src[e] = A[e];
}

// Several cells per workgroup, one lane per row: lane e of a cell computes
// z_e = sum_j B_je r_j (B symmetric; column j of the block is contiguous, so the lanes of a cell
// read consecutive addresses), r staged in LDS.
template<typename Number>
__global__ void __launch_bounds__(256) block_apply_kernel(BlockApplyArgs<Number> a, int lanes, int cpb)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      r      = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / lanes;
  int const     e      = threadIdx.x - slot * lanes;
  int64_t const cell   = (int64_t)blockIdx.x * cpb + slot;
  int const     m      = a.m;
  bool const    active = slot < cpb && cell < a.n_cells && e < m;
  Number *      rc     = r + (slot < cpb ? slot : 0) * lanes;
  if(active)
  {
    int64_t const g = cell * m + e;
    rc[e]           = a.t ? a.b[g] - a.t[g] : a.b[g];
  }
  __syncthreads();
  if(!active)
    return;
  Number const * B = a.blocks + cell * (int64_t)m * m;
  Number         z = 0;
  for(int j = 0; j < m; ++j)
    z += B[(int64_t)j * m + e] * rc[j];
  int64_t const g = cell * m + e;
  if(a.variant == 0)
    a.dst[g] = a.f2 * z;
  else
  {
    Number v = (Number(1) + a.f1) * a.x[g] + a.f2 * z;
    if(a.variant == 2)
      v -= a.f1 * a.x_old[g];
    a.dst[g] = v;
  }
}
// Fast-diagonalization inverse of the affine cell blocks: one group of n^2 lanes per cell,
// six 1D sweeps in LDS (V^T along x, y, z, the division by the eigenvalue sums, V along z, y, x).
template<typename Number, int N>
__global__ void __launch_bounds__(256) fdm_apply_kernel(BlockApplyArgs<Number> a, int cpb)
{
  using namespace sweeps;
  constexpr int N2 = N * N, N3 = N2 * N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      sm     = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / N2;
  int const     lt     = threadIdx.x - slot * N2;
  int64_t const cell   = (int64_t)blockIdx.x * cpb + slot;
  bool const    active = slot < cpb && cell < a.n_cells;
  Number *      U      = sm + (active ? slot : 0) * 2 * N3;
  Number *      T      = U + N3;
  int           cfg[3] = {0, 0, 0};
  if(active)
  {
#pragma unroll
    for(int d = 0; d < 3; ++d)
      cfg[d] = a.fdm_cfg[cell * 3 + d];
    for(int e = lt; e < N3; e += N2)
    {
      int64_t const g = cell * N3 + e;
      U[e]            = a.t ? a.b[g] - a.t[g] : a.b[g];
    }
  }
  __syncthreads();
  if(active)
    sweep<Number, N, true, false>(a.fdm_vec + cfg[0] * N2, U, T, 0, lt);
  __syncthreads();
  if(active)
    sweep<Number, N, true, false>(a.fdm_vec + cfg[1] * N2, T, U, 1, lt);
  __syncthreads();
  if(active)
  {
    // z-line of this lane: transform, divide by the eigenvalue sums, transform back
    int const      base = line_base<N>(2, lt);
    Number const * vz   = a.fdm_vec + cfg[2] * N2;
    Number const   lx   = a.fdm_val[cfg[0] * N + lt % N];
    Number const   ly   = a.fdm_val[cfg[1] * N + lt / N];
    Number         x[N], y[N];
#pragma unroll
    for(int i = 0; i < N; ++i)
      x[i] = U[base + i * N2];
    mat_line<Number, N, true>(vz, x, y);
#pragma unroll
    for(int c = 0; c < N; ++c)
      y[c] /= lx + ly + a.fdm_val[cfg[2] * N + c];
    mat_line<Number, N, false>(vz, y, x);
#pragma unroll
    for(int i = 0; i < N; ++i)
      U[base + i * N2] = x[i];
  }
  __syncthreads();
  if(active)
    sweep<Number, N, false, false>(a.fdm_vec + cfg[1] * N2, U, T, 1, lt);
  __syncthreads();
  if(active)
    sweep<Number, N, false, false>(a.fdm_vec + cfg[0] * N2, T, U, 0, lt);
  __syncthreads();
  if(active)
    for(int e = lt; e < N3; e += N2)
    {
      int64_t const g = cell * N3 + e;
      Number const  z = U[e];
      if(a.variant == 0)
        a.dst[g] = a.f2 * z;
      else
      {
        Number v = (Number(1) + a.f1) * a.x[g] + a.f2 * z;
        if(a.variant == 2)
          v -= a.f1 * a.x_old[g];
        a.dst[g] = v;
      }
    }
}

template<typename Number, int N>
hipError_t launch_fdm_n(BlockApplyArgs<Number> const & a, hipStream_t s)
{
  int const     cpb = 256 / (N * N);
  size_t const  lds = sizeof(Number) * cpb * 2 * N * N * N;
  int64_t const nb  = (a.n_cells + cpb - 1) / cpb;
  hipLaunchKernelGGL((fdm_apply_kernel<Number, N>), dim3((unsigned)nb), dim3(256), lds, s, a, cpb);
  return hipGetLastError();
}
}  // namespace

template<typename Number>
hipError_t launch_block_invert(Number * blocks, int m, int64_t n_cells, hipStream_t s)
{
  if(m < 1 || m > kMaxDenseBlock)
    return hipErrorInvalidValue;
  size_t const lds = sizeof(Number) * ((size_t)m * m + m);
  // set once; concurrent first calls may both set it, which is harmless
  static std::atomic<bool> attr{false};
  if(lds > 65536 && !attr.load(std::memory_order_acquire))
  {
    hipError_t e = hipFuncSetAttribute((void const *)block_invert_kernel<Number>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(sizeof(Number) * ((size_t)kMaxDenseBlock * kMaxDenseBlock +
                                                               kMaxDenseBlock)));
    if(e != hipSuccess)
      return e;
    attr.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((block_invert_kernel<Number>), dim3((unsigned)n_cells), dim3(256), lds, s,
                     blocks, m, n_cells);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_block_apply(BlockApplyArgs<Number> const & a, hipStream_t s)
{
  if(a.fdm_vec)
    switch(a.n1d)
    {
      case 2: return launch_fdm_n<Number, 2>(a, s);
      case 3: return launch_fdm_n<Number, 3>(a, s);
      case 4: return launch_fdm_n<Number, 4>(a, s);
      case 5: return launch_fdm_n<Number, 5>(a, s);
      case 6: return launch_fdm_n<Number, 6>(a, s);
      case 7: return launch_fdm_n<Number, 7>(a, s);
      case 8: return launch_fdm_n<Number, 8>(a, s);
      case 9: return launch_fdm_n<Number, 9>(a, s);
      default: return hipErrorInvalidValue;
    }
  if(a.m < 1 || a.m > 256)
    return hipErrorInvalidValue;
  int const     lanes = (a.m + 63) / 64 * 64;
  int const     cpb   = 256 / lanes;
  int64_t const nb    = (a.n_cells + cpb - 1) / cpb;
  size_t const  lds   = sizeof(Number) * cpb * lanes;
  hipLaunchKernelGGL((block_apply_kernel<Number>), dim3((unsigned)nb), dim3(256), lds, s, a, lanes,
                     cpb);
  return hipGetLastError();
}

template hipError_t launch_block_invert<double>(double *, int, int64_t, hipStream_t);
template hipError_t launch_block_invert<float>(float *, int, int64_t, hipStream_t);
template hipError_t launch_block_apply<double>(BlockApplyArgs<double> const &, hipStream_t);
template hipError_t launch_block_apply<float>(BlockApplyArgs<float> const &, hipStream_t);

}  // namespace ExaDG
