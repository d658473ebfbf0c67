// Right-hand side assembly on the device (FP64): the volume term int f v and the boundary terms
// the reference moves to the right-hand side, from point data at the cell and boundary-face
// Gauss points. RHSOperator::evaluate_add (operators/rhs_operator.cpp:52-95) and
// OperatorBase::rhs_add via boundary_face_loop_inhom_operator (operator_base.cpp:507-554,
// 1436-1491) with the inhomogeneous mirror values of weak_boundary_conditions.h:39-47,142-160.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
// Per-cell and per-boundary-face point records, SoA per cell / face (FP64):
//   cell: kRhsCellRec x n^3: x, y, z, JxW
//   boundary face: kRhsFaceRec x n^2: x, y, z, the outward unit normal (3), JxW of the face,
//   a = J^{-1} n (3, so that dv/dn = sum_j a_j dv/dxi_j), tau_F
constexpr int kRhsCellRec = 4;
constexpr int kRhsFaceRec = 11;

struct RhsArgs
{
  double *       b;
  double const * cell_geo;
  double const * face_geo;
  double const * f;         // values at the cell points (NULL: zero); unused for the sine family
  double const * boundary;  // g / h at the boundary points (NULL: zero); unused for the sine family
  double const * tables;    // S (n x n) | Sd (n x n) | d0 (n) | d1 (n)
  BoxTopology    topo;
  int64_t        side_base[7];  // first boundary face of each side (side-major), total last
  int            n;
  int            sine;   // 1: f = 3 omega^2 u, g = u, h = grad u . n for u = prod sin(omega x_d)
  double         omega;
  int64_t        n_cells;
};

hipError_t launch_rhs_assemble(RhsArgs const & a, hipStream_t s);

}  // namespace ExaDG
