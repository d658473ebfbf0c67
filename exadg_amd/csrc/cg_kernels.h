// Kernels of the continuous (FE_Q) and h-coarsened multigrid levels: the continuous Q_p Laplace
// (cell integrals plus assembly through the node map), the DG_p <-> Q_p c-transfer at any
// degree, the separable h- and p-transfers between continuous levels, and the DG h-transfer
// between a mesh and its 2x-coarsened parent.
//
// Reference: MultigridPreconditionerBase::initialize_levels / initialize_transfer_operators
// (solvers_and_preconditioners/multigrid/multigrid_preconditioner_base.cpp:97-323),
// MultigridTransfer (multigrid/transfer.cpp:28-96) over deal.II's MGTwoLevelTransfer, the
// continuous operator of OperatorBase with Dirichlet constraints as identity rows
// (operators/operator_base.cpp:299-309, multigrid/constraints.h:109-194).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
// Continuous FE_Q(p) numbering of the structured box (GL support points, x fastest):
// nn[d] = nc[d] p + 1 nodes in a non-periodic direction and nc[d] p in a periodic one (node
// nc[d] p is node 0). Nodes on Dirichlet sides are constrained: they read as zero and their
// rows act as the identity.
struct CGTopology
{
  int nc[3];
  int p;
  int nn[3];
  int periodic[3];
  int dirichlet[6];
};

CGTopology make_cg_topology(BoxTopology const & t, int p);

__host__ __device__ inline bool cg_constrained(CGTopology const & q, int const * g)
{
  for(int d = 0; d < 3; ++d)
    if(!q.periodic[d] &&
       ((g[d] == 0 && q.dirichlet[2 * d]) || (g[d] == q.nn[d] - 1 && q.dirichlet[2 * d + 1])))
      return true;
  return false;
}

__host__ __device__ inline int64_t cg_node_id(CGTopology const & q, int const * g)
{
  return g[0] + (int64_t)q.nn[0] * (g[1] + (int64_t)q.nn[1] * g[2]);
}

__host__ __device__ inline void cg_node_coords(CGTopology const & q, int64_t node, int * g)
{
  g[0]            = (int)(node % q.nn[0]);
  int64_t const r = node / q.nn[0];
  g[1]            = (int)(r % q.nn[1]);
  g[2]            = (int)(r / q.nn[1]);
}

// one direction: the cells sharing node index g and its local index in each (1 or 2 pairs)
__host__ __device__ inline int cg_sharing_1d(CGTopology const & q, int d, int g, int * cell,
                                             int * local)
{
  int const c = g / q.p, a = g - c * q.p;
  if(a != 0)
  {
    cell[0]  = c;
    local[0] = a;
    return 1;
  }
  int cnt = 0;
  if(c < q.nc[d])  // the cell starting at this node (absent at the non-periodic upper end)
  {
    cell[cnt]  = c;
    local[cnt] = 0;
    ++cnt;
  }
  if(c >= 1 || q.periodic[d])  // the cell ending at this node
  {
    cell[cnt]  = c >= 1 ? c - 1 : q.nc[d] - 1;
    local[cnt] = q.p;
    ++cnt;
  }
  return cnt;
}

// 1D sparse operator in a global numbering: row r holds `width` (index, weight) pairs; unused
// slots carry index 0 and weight 0
template<typename Number>
struct LineOperator
{
  int            rows  = 0;
  int            width = 0;
  int const *    idx   = nullptr;
  Number const * w     = nullptr;
};

// dst_dg (n^3 per cell, lexicographic) += continuous node values (constrained nodes read 0)
template<typename Number>
hipError_t launch_c_prolongate_add(Number * dst_dg, Number const * src_cg, BoxTopology const & t,
                                   CGTopology const & q, hipStream_t s);
// continuous node += sum over the cells sharing it of their DG values (constrained nodes skip)
template<typename Number>
hipError_t launch_c_restrict_add(Number * dst_cg, Number const * src_dg, BoxTopology const & t,
                                 CGTopology const & q, hipStream_t s);
// out = L applied along direction d of the lexicographic array `in` (dims_in, x fastest); the
// output has dims_in with dims_in[d] replaced by L.rows. With mask_in, inputs that are
// constrained nodes of q_in read as zero; with mask_out, outputs that are constrained nodes of
// q_out are left untouched. accumulate: out += instead of out =.
template<typename Number>
hipError_t launch_line_apply(Number * out, Number const * in, int const * dims_in, int d,
                             LineOperator<Number> const & L, CGTopology const * q_in,
                             CGTopology const * q_out, bool accumulate, hipStream_t s);
// DG h-transfer: every fine cell of tf += (Ph[bz] x Ph[by] x Ph[bx]) parent values, with b the
// cell's position in its parent and Ph[b] (n x n, row-major: fine node, coarse node) the child
// interpolation matrices
template<typename Number>
hipError_t launch_h_prolongate_dg_add(Number * dst_f, Number const * src_c, Number const * Ph, int n,
                                      BoxTopology const & tf, BoxTopology const & tc, hipStream_t s);
// every coarse cell of tc += sum over its 8 children of the transposed interpolation
template<typename Number>
hipError_t launch_h_restrict_dg_add(Number * dst_c, Number const * src_f, Number const * Ph, int n,
                                    BoxTopology const & tf, BoxTopology const & tc, hipStream_t s);

// Continuous Q_p cell integrals: the n^3 local results of every cell (the cell-local action
// of grad u . grad v, or the cell-local diagonal) into `out` (DG layout)
template<typename Number>
struct CGCellArgs
{
  CGTopology     q;
  BoxTopology    t;
  // n = p+1 tables, row-major: S (Gauss point x node), Dq (collocation derivative on the Gauss
  // points), w (Gauss weights), then for the diagonal S.S, SD.SD and S.SD elementwise with SD
  // the nodal basis derivative at the Gauss points
  Number const * tables = nullptr;
  // curved cells: per cell 6 x n^3 values of G = JxW J^{-1} J^{-T} (00,11,22,01,02,12); null
  // on affine meshes, where G = diag(hscale) times the weight product
  Number const * metric = nullptr;
  Number         hscale[3];
  Number const * src = nullptr;
  Number *       out = nullptr;
  int64_t        n_cells;
};
template<typename Number>
hipError_t launch_cg_cell(CGCellArgs<Number> const & a, bool diagonal, hipStream_t s);

// Assembly of the cell results per continuous node (sum over the cells sharing it), with the
// constrained rows as the identity, and an epilogue:
//   vmult    : dst = A x
//   residual : dst = b - A x
//   chebyshev: dst = (1+f1) x - f1 x_old (has_old) + f2 dinv (b - A x)
//   diagonal : dst = diag(A) (1 on constrained rows)
enum class CGAssembleMode : int
{
  vmult     = 0,
  residual  = 1,
  chebyshev = 2,
  diagonal  = 3
};
template<typename Number>
struct CGAssembleArgs
{
  CGTopology     q;
  BoxTopology    t;
  Number const * cell_values = nullptr;
  Number *       dst         = nullptr;
  Number const * x           = nullptr;
  Number const * b           = nullptr;
  Number const * x_old       = nullptr;
  Number const * dinv        = nullptr;
  Number         f1 = 0, f2 = 0;
  int            has_old = 0;
};
template<typename Number>
hipError_t launch_cg_assemble(CGAssembleArgs<Number> const & a, CGAssembleMode mode, hipStream_t s);

}  // namespace ExaDG
