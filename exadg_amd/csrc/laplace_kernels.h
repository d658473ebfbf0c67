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

// Table layout (Number, SipgArgs::tab): general kernel: [S n*n | Dq n*n | d0 n | d1 n];
// Cartesian kernel: [M n*n | K n*n | d0 n | d1 n] on the unit interval plus h per direction.
template<typename Number>
struct SipgArgs
{
  // 1D tables [A n*n | B n*n | c n | e n]. The Cartesian kernel reads them from this by-value
  // copy, i.e. as uniform scalar kernel-argument loads (unit-interval mass M, stiffness K,
  // end derivatives d0, d1, scaled per direction by h or 1/h inside the kernel); the general
  // kernel (S, Dq, d0, d1) indexes them at run time and reads the device copy `tables`.
  static constexpr int kMaxTable = 2 * 9 * 9 + 2 * 9;
  Number         tab[kMaxTable];
  Number         h[6];  // Cartesian kernel: h_x, h_y, h_z, 1/h_x, 1/h_y, 1/h_z
  Number const * tables;  // general kernel: the same tables in device memory
  Number const * src;
  Number *       dst;
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
