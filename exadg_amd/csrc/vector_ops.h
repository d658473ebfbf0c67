// Device vector and the streaming (HBM-bound) vector kernels of the solver layer:
// reductions, AXPY-type updates of CG and Chebyshev, precision conversion.
// Counterpart of dealii::LinearAlgebra::distributed::Vector operations used by SolverCG
// and PreconditionChebyshev (iterative_solvers_dealii_wrapper.h:136-221,
// chebyshev_smoother.h:37-227); one GPU, so no ghost exchange.
#pragma once
#include <cstdint>

#include "common.h"

namespace ExaDG
{
template<typename Number>
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n) { reinit(n); }
  void reinit(int64_t n)
  {
    if((int64_t)buf_.size() != n)
      buf_.resize((size_t)n);
  }
  int64_t        size() const { return (int64_t)buf_.size(); }
  Number *       data() { return buf_.data(); }
  Number const * data() const { return buf_.data(); }

private:
  DeviceBuffer<Number> buf_;
};

// Reduction workspace: partial sums per block, the final scalars and the ticket counter of the
// one-launch reduction (device), plus a host mirror for the values the host needs
// (convergence checks). Reductions on one Reducer must be stream-ordered (they share the
// partials and the counter).
class Reducer
{
public:
  static constexpr int kMaxPartials = 1024;
  static constexpr int kSlots       = 8;
  // device doubles a Reducer needs: 2 kMaxPartials partials, kSlots results, one word for
  // the counter
  static constexpr int kDeviceDoubles = 2 * kMaxPartials + kSlots + 1;
  // owning: allocates the device buffers and a pinned host mirror
  Reducer();
  // non-owning view over kDeviceDoubles doubles of caller device memory (256-byte aligned);
  // allocates nothing. The counter word must be zero before the first reduction
  // (clear_counter). fetch() then copies through pageable memory held by the object.
  explicit Reducer(double * device_memory);
  ~Reducer();
  Reducer(Reducer const &) = delete;
  Reducer & operator=(Reducer const &) = delete;
  double * partials() { return partials_; }
  double * result() { return result_; }  // device, kSlots doubles
  unsigned int * counter() { return counter_; }  // device, zero between reductions
  // zero the counter word with a fill kernel (graph-capturable; no memset node)
  void clear_counter(hipStream_t s);
  // copy the result slots to the host and wait
  double const * fetch(hipStream_t s, int n_slots = kSlots);

private:
  DeviceBuffer<double> own_;
  double *             partials_ = nullptr;
  double *             result_   = nullptr;
  unsigned int *       counter_  = nullptr;
  double *             host_     = nullptr;  // pinned (owning mode) or host_local_
  bool                 pinned_   = false;
  double               host_local_[kSlots] = {};
};

namespace vec
{
template<typename Number>
void set_zero(Number * x, int64_t n, hipStream_t s);
// x[i] = uniform[-1,1] by SplitMix64 on (seed, i), reproducible on any grid
template<typename Number>
void random_uniform(Number * x, int64_t n, uint64_t seed, hipStream_t s);
template<typename Number>
void copy(Number * dst, Number const * src, int64_t n, hipStream_t s);
template<typename To, typename From>
void convert(To * dst, From const * src, int64_t n, hipStream_t s);
// result[slot] = sum a_i b_i (accumulated in FP64, deterministic two-stage reduction)
template<typename Number>
void dot(Reducer & red, int slot, Number const * a, Number const * b, int64_t n, hipStream_t s);
// result[slot_aa] = sum a_i^2 and result[slot_ab] = sum a_i b_i in one pass (b may be FP32
// with a FP64: its entries widen exactly)
template<typename Number, typename NumberB>
void norm2_and_dot(Reducer & red, int slot_aa, int slot_ab, Number const * a, NumberB const * b,
                   int64_t n, hipStream_t s);
// y = a x + b y
template<typename Number>
void axpby(Number * y, Number a, Number const * x, Number b, int64_t n, hipStream_t s);
// x += alpha p; r -= alpha v; alpha read from device scalar (ratio num/den of two slots)
template<typename Number>
void cg_update(Number * x, Number * r, Number const * p, Number const * v, double const * num,
               double const * den, int64_t n, hipStream_t s);
// p = z + (num/den) p (z may be FP32 with p FP64)
template<typename Number, typename NumberZ>
void cg_direction(Number * p, NumberZ const * z, double const * num, double const * den, int64_t n,
                  hipStream_t s);
// cg_update that also writes r_narrow = r rounded to Narrow (the input of a FP32 preconditioner)
template<typename Number, typename Narrow>
void cg_update_narrow(Number * x, Number * r, Narrow * r_narrow, Number const * p, Number const * v,
                      double const * num, double const * den, int64_t n, hipStream_t s);
// z = d .* r
template<typename Number>
void scale(Number * z, Number const * d, Number const * r, int64_t n, hipStream_t s);
// Chebyshev step (deal.II PreconditionChebyshev vector_updates):
//   iteration 0: x = f2 * D^{-1} b  (x_old untouched)
//   else       : x_new = (1+f1) x - f1 x_old + f2 D^{-1} (b - t); x_old <- x_new (caller swaps)
template<typename Number>
void chebyshev_update(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * t, Number const * dinv, Number f1, Number f2, bool first,
                      bool has_old, int64_t n, hipStream_t s);
// x -= mean(x) (singular operators, coarse_grid_solvers.h:161-162); uses red slot 7
template<typename Number>
void subtract_mean(Reducer & red, Number * x, int64_t n, hipStream_t s);
}  // namespace vec

}  // namespace ExaDG
