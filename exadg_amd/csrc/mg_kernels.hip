// Multigrid transfer and coarse-level kernels (see mg_kernels.h).
#include <hip/hip_runtime.h>

#include "../../include/exadg_amd.h"
#include "mg_kernels.h"

namespace ExaDG
{
namespace
{
// ---------------------------------------------------------------- p-transfer
//
// One cell per group of nf^2 lanes, three sum-factorized 1D interpolation sweeps through LDS.
// PROLONG: input nc^3 -> output nf^3 with M = P; else input nf^3 -> output nc^3 with M = P^T.
template<typename Number, bool PROLONG>
__global__ void __launch_bounds__(256)
  p_transfer_kernel(Number * dst, Number const * src, Number const * P, int nf, int nc,
                    int64_t n_cells, int cpb)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      sm     = reinterpret_cast<Number *>(smem_raw);
  int const     nf3    = nf * nf * nf;
  int const     tpc    = nf * nf;
  int const     slot   = threadIdx.x / tpc;
  int const     lt     = threadIdx.x - slot * tpc;
  int64_t const cell   = (int64_t)blockIdx.x * cpb + slot;
  bool const    active = slot < cpb && cell < n_cells;
  Number *      b0     = sm + (active ? slot : 0) * 2 * nf3;
  Number *      b1     = b0 + nf3;
  int const     nin = PROLONG ? nc : nf, nout = PROLONG ? nf : nc;
  int const     nin3 = nin * nin * nin, nout3 = nout * nout * nout;
  auto          M = [&](int a, int i) { return PROLONG ? P[a * nc + i] : P[i * nc + a]; };

  if(active)
    for(int e = lt; e < nin3; e += tpc)
      b0[e] = src[cell * nin3 + e];
  __syncthreads();
  if(active)
    for(int task = lt; task < nin * nin; task += tpc)
    {
      int const j = task % nin, k = task / nin;
      for(int a = 0; a < nout; ++a)
      {
        Number s = 0;
        for(int i = 0; i < nin; ++i)
          s += M(a, i) * b0[i + nin * (j + nin * k)];
        b1[a + nout * (j + nin * k)] = s;
      }
    }
  __syncthreads();
  if(active)
    for(int task = lt; task < nout * nin; task += tpc)
    {
      int const a = task % nout, k = task / nout;
      for(int b = 0; b < nout; ++b)
      {
        Number s = 0;
        for(int j = 0; j < nin; ++j)
          s += M(b, j) * b1[a + nout * (j + nin * k)];
        b0[a + nout * (b + nout * k)] = s;
      }
    }
  __syncthreads();
  if(active)
    for(int task = lt; task < nout * nout; task += tpc)
    {
      int const a = task % nout, b = task / nout;
      for(int c = 0; c < nout; ++c)
      {
        Number s = 0;
        for(int k = 0; k < nin; ++k)
          s += M(c, k) * b0[a + nout * (b + nout * k)];
        dst[cell * nout3 + a + nout * (b + nout * c)] += s;
      }
    }
}

// Q1 operator as an assembled 27-point stencil: W[j * n + node] couples node with its
// neighbour at offset (j % 3 - 1, (j / 3) % 3 - 1, j / 9 - 1) (periodic wrap). Weights to
// constrained or missing neighbours are zero and those reads are clamped in bounds; a
// constrained node's row is the identity. One thread per node, W read plane by plane
// (consecutive nodes, consecutive addresses).
template<typename Number>
__device__ __forceinline__ Number q1_stencil_row(Number const * __restrict__ W,
                                                 Number const * __restrict__ src,
                                                 Q1Topology const & q, int64_t node, int const * nd,
                                                 int64_t n)
{
  int idx[3][3];
#pragma unroll
  for(int d = 0; d < 3; ++d)
#pragma unroll
    for(int o = 0; o < 3; ++o)
    {
      int c = nd[d] + o - 1;
      if(q.periodic[d])
        c = c < 0 ? c + q.nn[d] : (c >= q.nn[d] ? c - q.nn[d] : c);
      else
        c = c < 0 ? 0 : (c >= q.nn[d] ? q.nn[d] - 1 : c);
      idx[d][o] = c;
    }
  Number acc = 0;
#pragma unroll
  for(int j = 0; j < 27; ++j)
  {
    int64_t const nb =
      idx[0][j % 3] + (int64_t)q.nn[0] * (idx[1][(j / 3) % 3] + (int64_t)q.nn[1] * idx[2][j / 9]);
    acc += W[j * n + node] * src[nb];
  }
  return acc;
}

template<typename Number>
__global__ void __launch_bounds__(256)
  q1_stencil_kernel(Number * __restrict__ dst, Number const * __restrict__ src,
                    Number const * __restrict__ W, Q1Topology q, int64_t n)
{
  int64_t const node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(node >= n)
    return;
  int nd[3];
  nd[0]           = (int)(node % q.nn[0]);
  int64_t const r = node / q.nn[0];
  nd[1]           = (int)(r % q.nn[1]);
  nd[2]           = (int)(r / q.nn[1]);
  dst[node]       = q1_stencil_row(W, src, q, node, nd, n);
}

// one Chebyshev step on the Q1 level with the operator application fused in:
//   x_new = (1 + f1) x - f1 x_old + f2 D^{-1} (b - A x)     (variant 2)
//   x_new = (1 + f1) x + f2 D^{-1} (b - A x)                (variant 1)
template<typename Number>
__global__ void __launch_bounds__(256)
  q1_chebyshev_kernel(Number * __restrict__ x_new, Number const * __restrict__ x,
                      Number const * __restrict__ x_old, Number const * __restrict__ b,
                      Number const * __restrict__ dinv, Number const * __restrict__ W,
                      Q1Topology q, Number f1, Number f2, int variant, int64_t n)
{
  int64_t const node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(node >= n)
    return;
  int nd[3];
  nd[0]           = (int)(node % q.nn[0]);
  int64_t const r = node / q.nn[0];
  nd[1]           = (int)(r % q.nn[1]);
  nd[2]           = (int)(r / q.nn[1]);
  Number const t  = q1_stencil_row(W, x, q, node, nd, n);
  Number       v  = (Number(1) + f1) * x[node] + f2 * dinv[node] * (b[node] - t);
  if(variant == 2)
    v -= f1 * x_old[node];
  x_new[node] = v;
}

inline int blocks_for(int64_t n) { return (int)((n + 255) / 256); }
}  // namespace

template<typename Number>
hipError_t launch_p_prolongate_add(Number * dst_f, Number const * src_c, Number const * P, int nf,
                                   int nc, int64_t n_cells, hipStream_t s)
{
  int const    cpb = 256 / (nf * nf);
  size_t const lds = sizeof(Number) * cpb * 2 * nf * nf * nf;
  hipLaunchKernelGGL((p_transfer_kernel<Number, true>), dim3((n_cells + cpb - 1) / cpb), dim3(256),
                     lds, s, dst_f, src_c, P, nf, nc, n_cells, cpb);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_p_restrict_add(Number * dst_c, Number const * src_f, Number const * P, int nf,
                                 int nc, int64_t n_cells, hipStream_t s)
{
  int const    cpb = 256 / (nf * nf);
  size_t const lds = sizeof(Number) * cpb * 2 * nf * nf * nf;
  hipLaunchKernelGGL((p_transfer_kernel<Number, false>), dim3((n_cells + cpb - 1) / cpb), dim3(256),
                     lds, s, dst_c, src_f, P, nf, nc, n_cells, cpb);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_q1_stencil(Number * dst, Number const * src, Number const * W,
                             Q1Topology const & q, hipStream_t s)
{
  int64_t const n = (int64_t)q.nn[0] * q.nn[1] * q.nn[2];
  hipLaunchKernelGGL((q1_stencil_kernel<Number>), dim3(blocks_for(n)), dim3(256), 0, s, dst, src, W,
                     q, n);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_q1_chebyshev(Number * x_new, Number const * x, Number const * x_old,
                               Number const * b, Number const * dinv, Number const * W,
                               Q1Topology const & q, Number f1, Number f2, bool has_old,
                               hipStream_t s)
{
  int64_t const n = (int64_t)q.nn[0] * q.nn[1] * q.nn[2];
  hipLaunchKernelGGL((q1_chebyshev_kernel<Number>), dim3(blocks_for(n)), dim3(256), 0, s, x_new, x,
                     x_old, b, dinv, W, q, f1, f2, has_old ? 2 : 1, n);
  return hipGetLastError();
}

#define EXADG_INST(T)                                                                              \
  template hipError_t launch_p_prolongate_add<T>(T *, T const *, T const *, int, int, int64_t,     \
                                                 hipStream_t);                                     \
  template hipError_t launch_p_restrict_add<T>(T *, T const *, T const *, int, int, int64_t,       \
                                               hipStream_t);                                       \
  template hipError_t launch_q1_stencil<T>(T *, T const *, T const *, Q1Topology const &,          \
                                           hipStream_t);                                           \
  template hipError_t launch_q1_chebyshev<T>(T *, T const *, T const *, T const *, T const *,      \
                                             T const *, Q1Topology const &, T, T, bool, hipStream_t);
EXADG_INST(double)
EXADG_INST(float)
#undef EXADG_INST

}  // namespace ExaDG
