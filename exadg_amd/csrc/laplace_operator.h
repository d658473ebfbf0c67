// Host-side SIPG Laplace operator: the MI355X counterpart of ExaDG's
// Poisson::LaplaceOperator<dim=3, Number, 1> on top of OperatorBase
// (poisson/spatial_discretization/laplace_operator.h:227-351, operators/operator_base.h:105-756).
// Vectors are device pointers of length n(); everything is stream ordered.
#pragma once
#include <memory>

#include "common.h"
#include "grid.h"
#include "error_kernels.h"
#include "guess_kernels.h"
#include "block_kernels.h"
#include "laplace_kernels.h"
#include "shape_info.h"
#include "vector_ops.h"

namespace ExaDG
{
namespace Poisson
{
// LaplaceOperatorData / LaplaceKernelData (laplace_operator.h:38-45,212-225)
struct LaplaceOperatorData
{
  int    degree    = 1;
  double IP_factor = 1.0;
};

class LaplaceOperatorBase
{
public:
  virtual ~LaplaceOperatorBase() = default;
  virtual int     precision() const                                         = 0;
  virtual int     degree() const                                            = 0;
  virtual int64_t m() const                                                 = 0;
  virtual int64_t n_cells() const                                           = 0;
  virtual bool    is_affine() const                                         = 0;
  virtual void    vmult(void * dst, void const * src, hipStream_t s) const     = 0;
  virtual void    vmult_add(void * dst, void const * src, hipStream_t s) const = 0;
  virtual void    calculate_diagonal(void * diag, hipStream_t s) const         = 0;
  virtual void    calculate_inverse_diagonal(void * diag, hipStream_t s) const = 0;
  virtual void    rhs_sine(double omega, double * b, hipStream_t s) const      = 0;
  // right-hand side from point values (f at the cell Gauss points, g / h at the boundary-face
  // Gauss points; NULL = zero) and the points themselves (RhsGeometry numbering)
  virtual void rhs(double const * f, double const * boundary, double * b, hipStream_t s) const = 0;
  virtual void rhs_n_points(int64_t & n_cell_points, int64_t & n_boundary_points) const     = 0;
  virtual void rhs_points(double * x_cell, double * x_boundary, hipStream_t s) const        = 0;
  // weak DG divergence (central flux, no-slip walls) of a Fourier velocity field; mean
  // removed when the operator is singular
  virtual void rhs_divergence_fourier(double const * modes_host, int n_modes, double t, double * b,
                                      hipStream_t s) const = 0;
  virtual bool is_singular() const                                             = 0;
  // relative L2 error of u (FP64, device) against prod_d sin(omega x_d), QGauss(k+3)
  // (calculate_error, error_calculation.cpp:38-103); FP64 operators only
  virtual double error_sine(double omega, double const * u, hipStream_t s) const = 0;
  virtual std::shared_ptr<StructuredGrid const> grid_ptr() const               = 0;
  virtual double                                ip_factor() const              = 0;
};

template<typename Number>
class LaplaceOperator;

// The inverses of the cells' diagonal blocks A_KK (block Jacobi), in one of two exact forms:
// on affine meshes by fast diagonalization (the block is the Kronecker sum
// L_x (x) M (x) M + M (x) L_y (x) M + M (x) M (x) L_z of 1D operators that carry the cell's own face
// terms; any degree), else as dense n^3 x n^3 inverses (degree <= 4).
template<typename Number>
class CellBlockInverse
{
public:
  explicit CellBlockInverse(LaplaceOperator<Number> const & op, hipStream_t s);
  // dst = f2 B (b - t) (variant 0; t may be null) or a Chebyshev update (a.variant 1, 2);
  // the block fields of `a` are filled here
  void apply(BlockApplyArgs<Number> a, hipStream_t s) const;
  bool fast_diagonalization() const { return fdm_; }

private:
  bool                 fdm_     = false;
  int                  m_       = 0, n1d_ = 0;
  int64_t              n_cells_ = 0;
  DeviceBuffer<Number> blocks_, vec_, val_;
  DeviceBuffer<int>    cfg_;
};

template<typename Number>
class LaplaceOperator : public LaplaceOperatorBase
{
public:
  LaplaceOperator(std::shared_ptr<StructuredGrid const> grid, LaplaceOperatorData const & data);

  int     precision() const override { return sizeof(Number) == 8 ? EXADG_FP64 : EXADG_FP32; }
  int     degree() const override { return data_.degree; }
  int64_t m() const override { return grid_->n_cells() * n3_; }
  int64_t n() const { return m(); }
  int64_t n_cells() const override { return grid_->n_cells(); }
  bool    is_affine() const override { return grid_->affine(); }

  // OperatorBase::vmult / vmult_add / apply (operator_base.cpp:156-178, 264-310)
  void vmult(Number * dst, Number const * src, hipStream_t s) const;
  void vmult_add(Number * dst, Number const * src, hipStream_t s) const;
  // OperatorBase::calculate_diagonal / calculate_inverse_diagonal (operator_base.cpp:249-262,
  // 608-646): exact diagonal by unit-vector probing of the cell-local block
  void calculate_diagonal(Number * diag, hipStream_t s) const;
  void calculate_inverse_diagonal(Number * diag, hipStream_t s) const;
  // OperatorBase::add_block_diagonal_matrices through the cell-based loop
  // (cell_based_loop_block_diagonal, operator_base.cpp:649-680): the n^3 x n^3 diagonal block
  // A_KK of every cell (cell integral plus its faces' self-coupling), column j of cell c at
  // blocks[(c n^3 + j) n^3 ...]; one unit-vector probe per local index
  void calculate_block_diagonal(Number * blocks, hipStream_t s) const;
  // the inverses of those blocks (exact; blocks of at most 125 x 125, i.e. degree <= 4)
  void calculate_inverse_block_diagonal(Number * blocks, hipStream_t s) const;
  // the 1D eigen-decompositions of the affine cell blocks for CellBlockInverse: per distinct
  // configuration of a cell's two faces along a direction, V (n x n, row = node) and the
  // eigenvalues l of L v = l h M v with V^T h M V = I; per cell three configuration indices
  void fast_diagonalization_tables(std::vector<double> & vec, std::vector<double> & val,
                                   std::vector<int> & cfg) const;
  // x_new = (1 + f1) x - f1 x_old (has_old) + f2 dinv (b - A x), the operator application and
  // the Chebyshev-Jacobi update in one kernel (PreconditionChebyshev's inner step,
  // chebyshev_smoother.h:37-227); x_new may alias x_old, not x
  void chebyshev_step(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * dinv, Number f1, Number f2, bool has_old,
                      hipStream_t s) const;
  // dst = b - A src in one kernel (the residual in the vmult epilogue); dst may alias b
  void residual(Number * dst, Number const * src, Number const * b, hipStream_t s) const;
  // coarse += (R x R x R)(b - A src) per cell in one kernel, R = P^T with P the (k+1) x nc
  // prolongation matrix (device, row-major) to a DG level of degree nc-1 on the same cells;
  // returns true (the fused form exists for both kernels)
  bool residual_restrict(Number * coarse, Number const * src, Number const * b, Number const * P,
                         int nc, hipStream_t s) const;

  void vmult(void * dst, void const * src, hipStream_t s) const override
  {
    vmult(static_cast<Number *>(dst), static_cast<Number const *>(src), s);
  }
  void vmult_add(void * dst, void const * src, hipStream_t s) const override
  {
    vmult_add(static_cast<Number *>(dst), static_cast<Number const *>(src), s);
  }
  void calculate_diagonal(void * d, hipStream_t s) const override
  {
    calculate_diagonal(static_cast<Number *>(d), s);
  }
  void calculate_inverse_diagonal(void * d, hipStream_t s) const override
  {
    calculate_inverse_diagonal(static_cast<Number *>(d), s);
  }
  // Poisson::Operator::rhs for the manufactured sine solution (FP64 output), on the device
  void rhs_sine(double omega, double * b, hipStream_t s) const override;
  void rhs(double const * f, double const * boundary, double * b, hipStream_t s) const override;
  void rhs_n_points(int64_t & n_cell_points, int64_t & n_boundary_points) const override;
  void rhs_points(double * x_cell, double * x_boundary, hipStream_t s) const override;
  void rhs_divergence_fourier(double const * modes_host, int n_modes, double t, double * b,
                              hipStream_t s) const override;
  bool   is_singular() const override;
  double error_sine(double omega, double const * u, hipStream_t s) const override;

  StructuredGrid const & grid() const { return *grid_; }
  std::shared_ptr<StructuredGrid const> grid_ptr() const override { return grid_; }
  double                                ip_factor() const override { return data_.IP_factor; }
  ShapeInfo1D const &    shape() const { return shape_; }
  GeometryData const &   geometry_host() const { return geo_; }

private:
  void apply(ApplyMode mode, Number * dst, Number const * src, int probe, hipStream_t s) const;
  SipgArgs<Number> make_args(Number * dst, Number const * src, int probe) const;
  void             launch(ApplyMode mode, SipgArgs<Number> const & a, hipStream_t s) const;
  // point data of the right-hand side, computed on first use
  void rhs_launch(double const * f, double const * boundary, int sine, double omega, double * b,
                  hipStream_t s) const;
  void ensure_rhs_geometry() const;

  std::shared_ptr<StructuredGrid const> grid_;
  LaplaceOperatorData                   data_;
  ShapeInfo1D                           shape_;
  GeometryData                          geo_;  // host copy (tau, and metric for curved)
  int                                   n3_;
  std::vector<Number>                   tab_;  // 1D tables (kernel-argument copy)
  Number                                h_[6] = {1, 1, 1, 1, 1, 1};
  DeviceBuffer<Number>                  tables_, cell_metric_, face_data_, tau_face_;
  DeviceBuffer<double>                  diag1d_;              // affine: M_ii, K_ii, d0, d1 (FP64)
  DeviceBuffer<double>                  diag_tab_;            // curved: S^2, D^2, S D, l'(0), l'(1)
  DeviceBuffer<double>                  div_tables_;          // S, Sd, weights|points (FP64)
  DeviceBuffer<double>                  rhs_tables_;          // S, Sd, d0, d1 (FP64)
  mutable DeviceBuffer<double>          rhs_cell_geo_, rhs_face_geo_;
  int64_t                               rhs_side_base_[7] = {0, 0, 0, 0, 0, 0, 0};
  mutable DeviceBuffer<double>          div_cell_geo_, div_face_geo_;  // curved meshes
  // error norm: Se (ne x n) | weights | points of QGauss(k+3); curved meshes: x and JxW per point
  DeviceBuffer<double>                  err_tables_;
  mutable DeviceBuffer<double>          err_geo_, err_partial_, err_scratch_;
  mutable DeviceBuffer<double>          modes_;
  mutable std::unique_ptr<Reducer>      reducer_;
};

std::unique_ptr<LaplaceOperatorBase> create_laplace_operator(exadg_mesh_desc const & mesh,
                                                             exadg_bc_desc const & bc, int degree,
                                                             int precision, double ip_factor);


// dst[i] = 1/d[i] with |d| <= 1e-10 -> 1 (utilities/invert_diagonal.h:35-46)
template<typename Number>
void invert_diagonal_device(Number * d, int64_t n, hipStream_t s);

}  // namespace Poisson
}  // namespace ExaDG
