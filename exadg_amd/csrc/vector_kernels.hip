// Streaming vector kernels (HBM-bound): reductions with wave-level shuffles, CG and
// Chebyshev updates, diagonal inversion, precision conversion.
#include <hip/hip_runtime.h>

#include "common.h"
#include "laplace_operator.h"
#include "vector_ops.h"

namespace ExaDG
{
namespace
{
constexpr int kBlock = 256;

inline int grid_for(int64_t n, int cap = 8192)
{
  int64_t const b = (n + kBlock - 1) / kBlock;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), cap);
}

// Wave64 sum of one double per lane by DPP row permutations: two quad swaps, the half-row
// mirror and the row mirror leave every lane of a row of 16 holding that row's sum (each step
// adds a lane's value to its partner's, so both partners hold the same bits); the four row
// sums are then read from lanes 0, 16, 32, 48 and added in that order. Deterministic.
template<int Ctrl>
__device__ __forceinline__ double dpp_move(double v)
{
  int const lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), Ctrl, 0xf, 0xf, true);
  int const hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), Ctrl, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
  int const lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int const hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v)
{
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  v += dpp_move<0x140>(v);  // row_mirror
  return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
}

// block-wide sum of one double per thread (DPP within each wave, then LDS across the waves);
// callers separate two uses by a barrier
__device__ __forceinline__ double block_sum(double v)
{
  __shared__ double wsum[kBlock / 64];
  v              = wave_sum(v);
  int const wave = threadIdx.x >> 6;
  if((threadIdx.x & 63) == 0)
    wsum[wave] = v;
  __syncthreads();
  double r = 0;
  if(threadIdx.x == 0)
    for(int w = 0; w < kBlock / 64; ++w)
      r += wsum[w];
  return r;  // valid on thread 0
}

// One-launch deterministic reduction: every block writes its partial sum(s), and the block
// that finishes last (ticket from a device-scope counter) adds the partials in block order,
// writes the result slot(s) and resets the counter to zero for the next reduction.
// K = 0: out0 = sum a*b; K = 1: out0 = sum a*b and out1 = sum a*a. b may be narrower than a
// (the FP32 result of a mixed-precision preconditioner): its entries widen exactly to FP64.
template<typename Number, int K, typename NumberB = Number>
__global__ void __launch_bounds__(kBlock)
  reduce_fused(Number const * a, NumberB const * b, int64_t n, double * partials,
               unsigned int * counter, double * out0, double * out1)
{
  double        s0 = 0, s1 = 0;
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    double const x = (double)a[i];
    double const y = (double)b[i];
    s0 += x * y;
    if constexpr(K == 1)
      s1 += x * x;
  }
  __shared__ bool last;
  double const    r0 = block_sum(s0);
  __syncthreads();
  double r1 = 0;
  if constexpr(K == 1)
  {
    r1 = block_sum(s1);
    __syncthreads();
  }
  if(threadIdx.x == 0)
  {
    partials[blockIdx.x] = r0;
    if constexpr(K == 1)
      partials[Reducer::kMaxPartials + blockIdx.x] = r1;
    __threadfence();
    last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if(!last)
    return;
  __threadfence();
  double volatile const * pv = partials;
  double                  t0 = 0, t1 = 0;
  for(int i = threadIdx.x; i < (int)gridDim.x; i += kBlock)
  {
    t0 += pv[i];
    if constexpr(K == 1)
      t1 += pv[Reducer::kMaxPartials + i];
  }
  t0 = block_sum(t0);
  __syncthreads();
  if constexpr(K == 1)
    t1 = block_sum(t1);
  if(threadIdx.x == 0)
  {
    *out0 = t0;
    if constexpr(K == 1)
      *out1 = t1;
    *counter = 0u;
  }
}

__global__ void __launch_bounds__(kBlock)
  reduce_stage2(double const * partials, int n_partials, double * out)
{
  double s = 0;
  for(int i = threadIdx.x; i < n_partials; i += kBlock)
    s += partials[i];
  double const r = block_sum(s);
  if(threadIdx.x == 0)
    *out = r;
}

template<typename Number>
__global__ void __launch_bounds__(kBlock) set_zero_kernel(Number * x, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    x[i] = Number(0);
}

// x[i] = uniform[-1,1] from the SplitMix64 mixer applied to (seed, i): a counter-based
// generator, so any device, grid or host loop produces the same vector
template<typename Number>
__global__ void __launch_bounds__(kBlock) random_uniform_kernel(Number * x, int64_t n, uint64_t seed)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z          = z ^ (z >> 31);
    x[i]       = (Number)(2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0);
  }
}

template<typename To, typename From>
__global__ void __launch_bounds__(kBlock) convert_kernel(To * dst, From const * src, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    dst[i] = (To)src[i];
}

template<typename Number>
__global__ void __launch_bounds__(kBlock)
  axpby_kernel(Number * y, Number a, Number const * x, Number b, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    y[i] = a * x[i] + b * y[i];
}

template<typename Number>
__global__ void __launch_bounds__(kBlock)
  cg_update_kernel(Number * x, Number * r, Number const * p, Number const * v, double const * num,
                   double const * den, int64_t n)
{
  Number const  alpha  = (Number)(*num / *den);
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    x[i] += alpha * p[i];
    r[i] -= alpha * v[i];
  }
}

// the CG update that also writes the new residual rounded to the preconditioner's precision
template<typename Number, typename Narrow>
__global__ void __launch_bounds__(kBlock)
  cg_update_narrow_kernel(Number * x, Number * r, Narrow * r_narrow, Number const * p,
                          Number const * v, double const * num, double const * den, int64_t n)
{
  Number const  alpha  = (Number)(*num / *den);
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    x[i] += alpha * p[i];
    Number const ri = r[i] - alpha * v[i];
    r[i]            = ri;
    r_narrow[i]     = (Narrow)ri;
  }
}

template<typename Number, typename NumberZ = Number>
__global__ void __launch_bounds__(kBlock)
  cg_direction_kernel(Number * p, NumberZ const * z, double const * num, double const * den,
                      int64_t n)
{
  Number const  beta   = (Number)(*num / *den);
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    p[i] = (Number)z[i] + beta * p[i];
}

template<typename Number>
__global__ void __launch_bounds__(kBlock)
  scale_kernel(Number * z, Number const * d, Number const * r, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    z[i] = d[i] * r[i];
}

template<typename Number>
__global__ void __launch_bounds__(kBlock)
  chebyshev_kernel(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                   Number const * t, Number const * dinv, Number f1, Number f2, int variant,
                   int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    if(variant == 0)
      x_new[i] = f2 * dinv[i] * b[i];
    else if(variant == 1)
      x_new[i] = (Number(1) + f1) * x[i] + f2 * dinv[i] * (b[i] - t[i]);
    else
      x_new[i] = (Number(1) + f1) * x[i] - f1 * x_old[i] + f2 * dinv[i] * (b[i] - t[i]);
  }
}

template<typename Number>
__global__ void __launch_bounds__(kBlock) shift_kernel(Number * x, double const * sum, int64_t n)
{
  Number const  mean   = (Number)(*sum / (double)n);
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    x[i] -= mean;
}

template<typename Number>
__global__ void __launch_bounds__(kBlock) sum_stage1(Number const * a, int64_t n, double * partials)
{
  double        s      = 0;
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    s += (double)a[i];
  double const r = block_sum(s);
  if(threadIdx.x == 0)
    partials[blockIdx.x] = r;
}

template<typename Number>
__global__ void __launch_bounds__(kBlock) invert_diagonal_kernel(Number * d, int64_t n)
{
  int64_t const stride = (int64_t)gridDim.x * kBlock;
  for(int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
  {
    Number const v = d[i];
    d[i]           = (fabs(v) > Number(1e-10)) ? Number(1) / v : Number(1);
  }
}
}  // namespace

Reducer::Reducer()
{
  own_.resize(kDeviceDoubles);
  partials_ = own_.data();
  result_   = partials_ + 2 * kMaxPartials;
  counter_  = reinterpret_cast<unsigned int *>(result_ + kSlots);
  HIP_CHECK(hipMemset(counter_, 0, sizeof(unsigned int)));
  HIP_CHECK(hipHostMalloc((void **)&host_, kSlots * sizeof(double), hipHostMallocDefault));
  pinned_ = true;
}

Reducer::Reducer(double * device_memory)
{
  if(device_memory == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null reduction workspace");
  partials_ = device_memory;
  result_   = partials_ + 2 * kMaxPartials;
  counter_  = reinterpret_cast<unsigned int *>(result_ + kSlots);
  host_     = host_local_;
}

Reducer::~Reducer()
{
  if(pinned_ && host_)
    (void)hipHostFree(host_);
}

void Reducer::clear_counter(hipStream_t s)
{
  hipLaunchKernelGGL((set_zero_kernel<double>), dim3(1), dim3(kBlock), 0, s,
                     reinterpret_cast<double *>(counter_), int64_t(1));
  HIP_CHECK(hipGetLastError());
}

double const * Reducer::fetch(hipStream_t s, int n_slots)
{
  HIP_CHECK(hipMemcpyAsync(host_, result_, n_slots * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return host_;
}

namespace vec
{
template<typename Number>
void set_zero(Number * x, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((set_zero_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, x, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void random_uniform(Number * x, int64_t n, uint64_t seed, hipStream_t s)
{
  hipLaunchKernelGGL((random_uniform_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, x, n,
                     seed);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void copy(Number * dst, Number const * src, int64_t n, hipStream_t s)
{
  HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(Number), hipMemcpyDeviceToDevice, s));
}

template<typename To, typename From>
void convert(To * dst, From const * src, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((convert_kernel<To, From>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void dot(Reducer & red, int slot, Number const * a, Number const * b, int64_t n, hipStream_t s)
{
  int const g = grid_for(n, Reducer::kMaxPartials);
  hipLaunchKernelGGL((reduce_fused<Number, 0>), dim3(g), dim3(kBlock), 0, s, a, b, n,
                     red.partials(), red.counter(), red.result() + slot, nullptr);
  HIP_CHECK(hipGetLastError());
}

template<typename Number, typename NumberB>
void norm2_and_dot(Reducer & red, int slot_aa, int slot_ab, Number const * a, NumberB const * b,
                   int64_t n, hipStream_t s)
{
  int const g = grid_for(n, Reducer::kMaxPartials);
  hipLaunchKernelGGL((reduce_fused<Number, 1, NumberB>), dim3(g), dim3(kBlock), 0, s, a, b, n,
                     red.partials(), red.counter(), red.result() + slot_ab, red.result() + slot_aa);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void axpby(Number * y, Number a, Number const * x, Number b, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((axpby_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, y, a, x, b, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void cg_update(Number * x, Number * r, Number const * p, Number const * v, double const * num,
               double const * den, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((cg_update_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, x, r, p, v,
                     num, den, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number, typename NumberZ>
void cg_direction(Number * p, NumberZ const * z, double const * num, double const * den, int64_t n,
                  hipStream_t s)
{
  hipLaunchKernelGGL((cg_direction_kernel<Number, NumberZ>), dim3(grid_for(n)), dim3(kBlock), 0, s,
                     p, z, num, den, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number, typename Narrow>
void cg_update_narrow(Number * x, Number * r, Narrow * r_narrow, Number const * p, Number const * v,
                      double const * num, double const * den, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((cg_update_narrow_kernel<Number, Narrow>), dim3(grid_for(n)), dim3(kBlock), 0,
                     s, x, r, r_narrow, p, v, num, den, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void scale(Number * z, Number const * d, Number const * r, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((scale_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, z, d, r, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void chebyshev_update(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * t, Number const * dinv, Number f1, Number f2, bool first,
                      bool has_old, int64_t n, hipStream_t s)
{
  int const variant = first ? 0 : (has_old ? 2 : 1);
  hipLaunchKernelGGL((chebyshev_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, x_new, x,
                     x_old, b, t, dinv, f1, f2, variant, n);
  HIP_CHECK(hipGetLastError());
}

template<typename Number>
void subtract_mean(Reducer & red, Number * x, int64_t n, hipStream_t s)
{
  int const g = grid_for(n, Reducer::kMaxPartials);
  hipLaunchKernelGGL((sum_stage1<Number>), dim3(g), dim3(kBlock), 0, s, x, n, red.partials());
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(kBlock), 0, s, red.partials(), g,
                     red.result() + 7);
  hipLaunchKernelGGL((shift_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, x,
                     red.result() + 7, n);
  HIP_CHECK(hipGetLastError());
}

#define EXADG_INST(T)                                                                            \
  template void set_zero<T>(T *, int64_t, hipStream_t);                                          \
  template void random_uniform<T>(T *, int64_t, uint64_t, hipStream_t);                          \
  template void copy<T>(T *, T const *, int64_t, hipStream_t);                                   \
  template void dot<T>(Reducer &, int, T const *, T const *, int64_t, hipStream_t);              \
  template void norm2_and_dot<T>(Reducer &, int, int, T const *, T const *, int64_t, hipStream_t); \
  template void axpby<T>(T *, T, T const *, T, int64_t, hipStream_t);                            \
  template void cg_update<T>(T *, T *, T const *, T const *, double const *, double const *,     \
                             int64_t, hipStream_t);                                              \
  template void cg_direction<T>(T *, T const *, double const *, double const *, int64_t,         \
                                hipStream_t);                                                    \
  template void scale<T>(T *, T const *, T const *, int64_t, hipStream_t);                       \
  template void chebyshev_update<T>(T *, T const *, T const *, T const *, T const *, T const *, T, \
                                    T, bool, bool, int64_t, hipStream_t);                        \
  template void subtract_mean<T>(Reducer &, T *, int64_t, hipStream_t);
EXADG_INST(double)
EXADG_INST(float)
#undef EXADG_INST
template void convert<float, double>(float *, double const *, int64_t, hipStream_t);
template void convert<double, float>(double *, float const *, int64_t, hipStream_t);
template void norm2_and_dot<double, float>(Reducer &, int, int, double const *, float const *,
                                           int64_t, hipStream_t);
template void cg_direction<double, float>(double *, float const *, double const *, double const *,
                                          int64_t, hipStream_t);
template void cg_update_narrow<double, float>(double *, double *, float *, double const *,
                                              double const *, double const *, double const *,
                                              int64_t, hipStream_t);
}  // namespace vec

namespace Poisson
{
template<typename Number>
void invert_diagonal_device(Number * d, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL((invert_diagonal_kernel<Number>), dim3(grid_for(n)), dim3(kBlock), 0, s, d, n);
  HIP_CHECK(hipGetLastError());
}

template void invert_diagonal_device<double>(double *, int64_t, hipStream_t);
template void invert_diagonal_device<float>(float *, int64_t, hipStream_t);
}  // namespace Poisson

}  // namespace ExaDG
