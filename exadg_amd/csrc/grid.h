// Structured hexahedral box with DeformedCubeManifold + MappingQ geometry, and the
// per-cell / per-face metric data the SIPG kernels stream from HBM.
//
// Reference: grid/deformed_cube_manifold.h:47-157 (push_forward, vertex placement),
// grid/grid_utilities.h:52-58 (MappingQ on Gauss-Lobatto support points),
// grid/periodic_box.h:35-87, operators/interior_penalty_parameter.h:43-128.
#pragma once
#include <cstdint>
#include <vector>

#include "common.h"
#include "rhs_kernels.h"
#include "shape_info.h"
#include "topology.h"

namespace ExaDG
{
class StructuredGrid
{
public:
  StructuredGrid(exadg_mesh_desc const & mesh, exadg_bc_desc const & bc);

  exadg_mesh_desc const & mesh() const { return mesh_; }
  exadg_bc_desc const &   bc() const { return bc_; }
  // the same box with every cells-per-direction count halved `times` times (h-multigrid
  // levels); same bounds, deformation, mapping and boundary types
  StructuredGrid coarsened(int times) const;
  BoxTopology const &     topology() const { return topo_; }
  int64_t                 n_cells() const { return n_cells_; }
  bool                    affine() const { return affine_; }
  double                  h(int d) const { return h_[d]; }

  void push_forward(double const * chart, double * x) const;

  // Position x[3][np] and Jacobian J[9][np] (J[i*3+j] = dx_i/dxi_j) of cell c at the tensor
  // points (px x py x pz, x fastest) of the reference cell.
  void evaluate(int64_t c, std::vector<double> const & px, std::vector<double> const & py,
                std::vector<double> const & pz, double * x, double * J) const;

private:
  exadg_mesh_desc     mesh_;
  exadg_bc_desc       bc_;
  BoxTopology         topo_;
  int64_t             n_cells_;
  bool                affine_;
  double              h_[3];
  std::vector<double> map_nodes_;  // GL support points of MappingQ(mapping_degree)
};

// Metric data at the quadrature points of one FE degree on a grid, FP64 on the host.
struct GeometryData
{
  int n = 0;
  // per cell, 6 components x n^3 points: symmetric G = JxW J^{-1} J^{-T}
  // (00,11,22,01,02,12) -- curved meshes only
  std::vector<double> cell_metric;
  // per face, 7 components x n^2 points: a_minus = J_minus^{-1} n, a_plus = J_plus^{-1} n,
  // JxW, with n the unit normal pointing in +xi_d -- curved meshes only
  std::vector<double> face_data;
  std::vector<double> tau_cell;  // surface/volume per cell
  std::vector<double> tau_face;  // IP (k+1)^2 max(tau-,tau+), boundary IP (k+1)^2 tau
};

GeometryData compute_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape,
                              double ip_factor, bool store_metric);

// Point data of the weak divergence (right-hand side of the pressure step) on any mesh, FP64:
// per cell 12 x n^3 values: the position x (3) and C = JxW J^{-1} (9, C[a][k] = JxW dxi_a/dx_k)
// at the Gauss points; per face 6 x n^2 values: x (3) and JxW n (3), n the unit normal in +xi_d
// direction, at the face Gauss points (t1 fastest), faces numbered as in BoxTopology.
struct DivergenceGeometry
{
  std::vector<double> cell;
  std::vector<double> face;
};
DivergenceGeometry compute_divergence_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape);

// Point data of the right-hand side assembly (rhs_kernels.h records) on any mesh, FP64: per
// cell x and JxW at the Gauss points; per boundary face of the non-periodic sides x, the
// outward unit normal, JxW, J^{-1} n and tau_F at the face Gauss points. Boundary faces are
// numbered side-major (0:x- 1:x+ 2:y- 3:y+ 4:z- 5:z+), on a side lexicographic in the two
// tangential cell coordinates (lower direction fastest); side_base[s] is a side's first face.
struct RhsGeometry
{
  std::vector<double> cell, face;
  int64_t             side_base[7] = {0, 0, 0, 0, 0, 0, 0};
};
RhsGeometry compute_rhs_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape,
                                 std::vector<double> const & tau_face);

}  // namespace ExaDG
