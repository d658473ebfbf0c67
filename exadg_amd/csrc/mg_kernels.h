// Multigrid kernels: polynomial (p) transfer between DG levels on the same mesh and the
// continuous Q1 operator (the c-transfer and the other continuous levels are in cg_kernels.h).
//
// Reference: MultigridTransfer (solvers_and_preconditioners/multigrid/transfer.cpp:28-96) ->
// MGTwoLevelTransfer::reinit_polynomial_transfer; the continuous coarse level of the hybrid
// multigrid (multigrid_preconditioner_base.cpp:191-226) with Dirichlet constraints
// (multigrid/constraints.h:109-194) acting as identity rows (operator_base.cpp:299-309).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
// Q1 node numbering of the structured box: (n_d + 1) nodes per non-periodic direction,
// n_d nodes per periodic direction (node n_d is node 0).
struct Q1Topology
{
  int nc[3];
  int nn[3];        // nodes per direction
  int periodic[3];
  int dirichlet[6]; // side constrained (Dirichlet, non-periodic)
};

// dst_f += (P x P x P) src_c per cell, P = nf x nc interpolation matrix (row-major)
template<typename Number>
hipError_t launch_p_prolongate_add(Number * dst_f, Number const * src_c, Number const * P, int nf,
                                   int nc, int64_t n_cells, hipStream_t s);
// dst_c += (P^T x P^T x P^T) src_f per cell
template<typename Number>
hipError_t launch_p_restrict_add(Number * dst_c, Number const * src_f, Number const * P, int nf,
                                 int nc, int64_t n_cells, hipStream_t s);
// Q1 operator as an assembled 27-point stencil (W: 27 planes of n_nodes weights, plane j for
// the neighbour offset (j%3-1, (j/3)%3-1, j/9-1), periodic wrap; constrained rows identity)
template<typename Number>
hipError_t launch_q1_stencil(Number * dst, Number const * src, Number const * W,
                             Q1Topology const & q, hipStream_t s);
// one Chebyshev-Jacobi step with the stencil application fused:
// x_new = (1+f1) x - f1 x_old (if has_old) + f2 dinv (b - A x)
template<typename Number>
hipError_t launch_q1_chebyshev(Number * x_new, Number const * x, Number const * x_old,
                               Number const * b, Number const * dinv, Number const * W,
                               Q1Topology const & q, Number f1, Number f2, bool has_old,
                               hipStream_t s);

}  // namespace ExaDG
