// Weak DG divergence of a discrete velocity on the pressure space of a Laplace operator: the
// counterpart of IncNS::DivergenceOperator<dim, Number> (incompressible_navier_stokes/
// spatial_discretization/operators/divergence_operator.h:96-200) with the weak formulation,
// integration by parts, central flux and boundary data, FP64. The velocity lives in
// FE_DGQ(degree_u)^3 on the Laplace operator's grid: per cell three component blocks of
// (degree_u+1)^3 nodal values on Gauss-Lobatto points; QGauss(degree_u+1) throughout.
#pragma once
#include <memory>

#include "common.h"
#include "divergence_kernels.h"
#include "grid.h"
#include "shape_info.h"

namespace ExaDG
{
class DivergenceOperator
{
public:
  DivergenceOperator(std::shared_ptr<StructuredGrid const> grid, int degree_p, int degree_u);

  int     degree_u() const { return degree_u_; }
  int     degree_p() const { return degree_p_; }
  int64_t n_velocity() const { return grid_->n_cells() * 3 * nu3_; }
  int64_t n_pressure() const { return grid_->n_cells() * np3_; }

  // DivergenceOperator::apply (homogeneous operator): dst (pressure) = B src (velocity)
  void apply(double * dst, double const * src, hipStream_t s) const;
  // nodal interpolation of the Fourier field (exadg_fourier_modes rows) at time t
  void interpolate_fourier(double const * modes_host, int n_modes, double t, double * u,
                           hipStream_t s) const;
  // positions of the velocity nodes: 3 x (n_cells (degree_u+1)^3), component-major
  void velocity_points(double * x, hipStream_t s) const;

private:
  void ensure_nodes() const;  // positions of the velocity nodes, computed on first use

  std::shared_ptr<StructuredGrid const> grid_;
  int                                   degree_p_, degree_u_, nu3_, np3_;
  DeviceBuffer<double>                  tables_, cell_geo_, face_geo_;
  mutable DeviceBuffer<double>          nodes_, modes_;
};
}  // namespace ExaDG
