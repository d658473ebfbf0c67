// Launch interface of the SIPG Laplace kernels (host-callable, no HIP types beyond streams).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
enum class ApplyMode : int
{
  vmult     = 0,  // dst = A src
  vmult_add = 1,  // dst += A src
  diagonal  = 2,  // dst[cell, i] = (A_KK)_{ii}: unit-vector probing of the cell-local block
  // Chebyshev-Jacobi step fused into the epilogue (src = x, dst = x_new):
  // x_new = (1 + f1) x - f1 x_old (if has_old) + f2 dinv (b - A x); x_old may alias dst
  chebyshev = 3
};

// Table layout (Number): general kernel: [S n*n | Dq n*n | d0 n | d1 n];
// Cartesian kernel: per direction d = 0,1,2 a block [h_d M n*n | K/h_d n*n | d0/h_d n | d1/h_d n].
template<typename Number>
struct SipgArgs
{
  Number const * src;
  Number *       dst;
  Number const * tables;
  Number const * cell_metric;  // general kernel only
  Number const * face_data;    // general kernel only
  Number const * tau_face;
  BoxTopology    topo;
  int64_t        n_cells;
  int            probe;      // diagonal mode: local index of the unit vector
  int            tile_rows;  // traversal: y-rows per tile (one tile per XCD)
  // chebyshev mode only
  Number const * x_old   = nullptr;
  Number const * rhs     = nullptr;
  Number const * dinv    = nullptr;
  Number         f1      = 0;
  Number         f2      = 0;
  int            has_old = 0;
};

// degree 1..8; mode as above. Returns hipError of the launch.
template<typename Number>
hipError_t launch_sipg_general(int degree, ApplyMode mode, SipgArgs<Number> const & args,
                               hipStream_t stream);
template<typename Number>
hipError_t launch_sipg_cartesian(int degree, ApplyMode mode, SipgArgs<Number> const & args,
                                 hipStream_t stream);

}  // namespace ExaDG
