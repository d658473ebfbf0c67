// Relative L2 error of a DG solution against the manufactured sine solution, the quantity
// the reference's sine application prints (calculate_error, error_calculation.cpp:38-103,
// QGauss(k+3)): one workgroup per cell interpolates the nodal values to the ne^3 points by
// sum factorization through LDS and accumulates the two weighted sums.
#include <hip/hip_runtime.h>

#include "error_kernels.h"

namespace ExaDG
{
namespace
{
constexpr int kThreads = 128;

__global__ void __launch_bounds__(kThreads)
  error_sine_kernel(double * partial, double const * u, double omega, int n, int ne,
                    double const * Se, double const * w, BoxTopology topo, double lx, double ly,
                    double lz, double hx, double hy, double hz, double const * geo)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *      a   = reinterpret_cast<double *>(smem_raw);  // ne^3
  double *      b   = a + ne * ne * ne;                      // ne^3
  __shared__ double red[2][kThreads];
  int64_t const c   = blockIdx.x;
  int const     n3 = n * n * n, ne3 = ne * ne * ne;
  // nodal values into a (n^3 entries)
  for(int e = threadIdx.x; e < n3; e += kThreads)
    a[e] = u[c * n3 + e];
  __syncthreads();
  // x sweep: b(qx, j, k) = sum_i Se[qx][i] a(i, j, k)   (ne x n x n)
  for(int e = threadIdx.x; e < ne * n * n; e += kThreads)
  {
    int const qx = e % ne, r = e / ne;
    double    s = 0;
    for(int i = 0; i < n; ++i)
      s += Se[qx * n + i] * a[i + n * r];
    b[e] = s;
  }
  __syncthreads();
  // y sweep: a(qx, qy, k) = sum_j Se[qy][j] b(qx, j, k)   (ne x ne x n)
  for(int e = threadIdx.x; e < ne * ne * n; e += kThreads)
  {
    int const qx = e % ne, qy = (e / ne) % ne, k = e / (ne * ne);
    double    s = 0;
    for(int j = 0; j < n; ++j)
      s += Se[qy * n + j] * b[qx + ne * (j + n * k)];
    a[e] = s;
  }
  __syncthreads();
  // z sweep and the error at every point
  int cc[3];
  cell_coords(topo, c, cc);
  double e2 = 0, x2 = 0;
  for(int e = threadIdx.x; e < ne3; e += kThreads)
  {
    int const qx = e % ne, qy = (e / ne) % ne, qz = e / (ne * ne);
    double    uq = 0;
    for(int k = 0; k < n; ++k)
      uq += Se[qz * n + k] * a[qx + ne * (qy + ne * k)];
    double x, y, z, jxw;
    if(geo)
    {
      double const * g = geo + c * 4 * ne3;
      x                = g[e];
      y                = g[ne3 + e];
      z                = g[2 * ne3 + e];
      jxw              = g[3 * ne3 + e];
    }
    else
    {
      x   = lx + hx * (cc[0] + w[ne + qx]);
      y   = ly + hy * (cc[1] + w[ne + qy]);
      z   = lz + hz * (cc[2] + w[ne + qz]);
      jxw = hx * hy * hz * w[qx] * w[qy] * w[qz];
    }
    double const ex = sin(omega * x) * sin(omega * y) * sin(omega * z);
    double const d  = uq - ex;
    e2 += jxw * d * d;
    x2 += jxw * ex * ex;
  }
  red[0][threadIdx.x] = e2;
  red[1][threadIdx.x] = x2;
  __syncthreads();
  for(int s = kThreads / 2; s > 0; s >>= 1)
  {
    if((int)threadIdx.x < s)
    {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if(threadIdx.x < 2)
    partial[2 * c + threadIdx.x] = red[threadIdx.x][0];
}

// block b sums rows [b*rows, (b+1)*rows) of the n x 2 array into scratch[2b..2b+1]
__global__ void __launch_bounds__(256)
  sum_pairs_kernel(double const * p, int64_t n, int64_t rows, double * out)
{
  __shared__ double red[2][256];
  int64_t const r0 = blockIdx.x * rows, r1 = min(n, r0 + rows);
  double        s0 = 0, s1 = 0;
  for(int64_t r = r0 + threadIdx.x; r < r1; r += 256)
  {
    s0 += p[2 * r];
    s1 += p[2 * r + 1];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  __syncthreads();
  for(int s = 128; s > 0; s >>= 1)
  {
    if((int)threadIdx.x < s)
    {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if(threadIdx.x < 2)
    out[2 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}
}  // namespace

hipError_t launch_error_sine(double * partial, double const * u, double omega, int n, int ne,
                             double const * Se, double const * w, BoxTopology const & topo,
                             double const * lower, double const * h, double const * geo,
                             int64_t n_cells, hipStream_t s)
{
  if(n < 2 || n > 9 || ne < n || ne > 12 || n_cells < 1)
    return hipErrorInvalidValue;
  size_t const lds = sizeof(double) * 2 * ne * ne * ne;
  hipLaunchKernelGGL(error_sine_kernel, dim3(n_cells), dim3(kThreads), lds, s, partial, u, omega, n,
                     ne, Se, w, topo, lower[0], lower[1], lower[2], h[0], h[1], h[2], geo);
  return hipGetLastError();
}

hipError_t launch_sum_pairs(double const * partial, int64_t n_rows, double * scratch, double * out,
                            hipStream_t s)
{
  // first stage: 1024 blocks (scratch holds 2 x 1024), second stage: one block
  int64_t const rows = (n_rows + 1023) / 1024;
  int const     g    = (int)((n_rows + rows - 1) / rows);
  hipLaunchKernelGGL(sum_pairs_kernel, dim3(g), dim3(256), 0, s, partial, n_rows, rows, scratch);
  hipLaunchKernelGGL(sum_pairs_kernel, dim3(1), dim3(256), 0, s, scratch, (int64_t)g, (int64_t)g,
                     out);
  return hipGetLastError();
}

}  // namespace ExaDG
