// Device helpers for sum-factorized kernels that hold an n^3 tensor (x fastest) in LDS and
// process it as n^2 lines along one direction, one line per lane.
#pragma once
#include <hip/hip_runtime.h>

namespace ExaDG
{
namespace sweeps
{
// first index of line `line` along direction d (the two other coordinates, lower fastest)
template<int N>
__device__ __forceinline__ int line_base(int d, int line)
{
  int const e0 = line % N, e1 = line / N;
  return d == 0 ? N * (e0 + N * e1) : (d == 1 ? e0 + N * N * e1 : e0 + N * e1);
}

template<int N>
__device__ __forceinline__ constexpr int line_stride(int d)
{
  return d == 0 ? 1 : (d == 1 ? N : N * N);
}

// y[a] = sum_i M(a, i) x[i] with M(a, i) = M[a N + i], or M[i N + a] when T
template<typename Number, int N, bool T>
__device__ __forceinline__ void mat_line(Number const * __restrict__ M, Number const (&x)[N],
                                         Number (&y)[N])
{
#pragma unroll
  for(int a = 0; a < N; ++a)
  {
    Number s = 0;
#pragma unroll
    for(int i = 0; i < N; ++i)
      s += (T ? M[i * N + a] : M[a * N + i]) * x[i];
    y[a] = s;
  }
}

// out (this lane's line along d) = M applied to in; ACC adds to out. in and out may be the same
// array: the lane reads its whole line before writing it.
template<typename Number, int N, bool T, bool ACC>
__device__ __forceinline__ void sweep(Number const * __restrict__ M, Number const * in, Number * out,
                                      int d, int line)
{
  int const base = line_base<N>(d, line), st = line_stride<N>(d);
  Number    x[N], y[N];
#pragma unroll
  for(int i = 0; i < N; ++i)
    x[i] = in[base + i * st];
  mat_line<Number, N, T>(M, x, y);
#pragma unroll
  for(int a = 0; a < N; ++a)
    out[base + a * st] = ACC ? out[base + a * st] + y[a] : y[a];
}
}  // namespace sweeps
}  // namespace ExaDG
