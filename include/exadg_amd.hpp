/*
 * exadg_amd.hpp -- header-only C++ binding of the C ABI in exadg_amd.h, the layer an ExaDG
 * build links against. It restores the reference's C++ conventions on top of the status codes:
 *
 *  - errors become exceptions (AssertThrow(..., ExcMessage) in the reference,
 *    solvers/iterative_solvers_dealii_wrapper.h:210-211 for the NaN check); exadg_amd::Error
 *    carries the status code;
 *  - handles are owned by RAII objects (the reference holds operators and preconditioners by
 *    shared_ptr, poisson/spatial_discretization/operator.h:233-236);
 *  - the method names of OperatorBase (vmult, vmult_add, calculate_inverse_diagonal, rhs, m/n;
 *    operators/operator_base.h:175-315), PreconditionerBase (vmult, update;
 *    preconditioners/preconditioner_base.h:33-67) and Krylov::SolverBase::solve returning the
 *    iteration count (iterative_solvers_dealii_wrapper.h:44-95).
 *
 * Vectors stay raw device pointers of the operator's precision; the caller owns them.
 */
#ifndef EXADG_AMD_HPP
#define EXADG_AMD_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "exadg_amd.h"

namespace exadg_amd
{
class Error : public std::runtime_error
{
public:
  Error(int code, std::string const & msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

private:
  int code_;
};

inline void check(int status)
{
  if(status != EXADG_OK)
    throw Error(status, std::string("exadg status ") + std::to_string(status) + ": " +
                          exadg_last_error());
}

// Poisson::LaplaceOperator<3, Number, 1> (laplace_operator.h:227-351) on the device
class LaplaceOperator
{
public:
  LaplaceOperator(exadg_mesh_desc const & mesh, exadg_bc_desc const & bc, int degree,
                  int precision = EXADG_FP64, double ip_factor = 1.0)
  {
    check(exadg_laplace_create(&mesh, &bc, degree, precision, ip_factor, &h_));
    int64_t n = 0;
    check(exadg_laplace_n_dofs(h_, &n));
    n_ = n;
    check(exadg_laplace_info(h_, &n_cells_, &degree_, &precision_, &affine_));
  }
  ~LaplaceOperator() { exadg_laplace_destroy(h_); }
  LaplaceOperator(LaplaceOperator const &) = delete;
  LaplaceOperator & operator=(LaplaceOperator const &) = delete;

  int64_t m() const { return n_; }
  int64_t n() const { return n_; }
  int64_t n_cells() const { return n_cells_; }
  int     degree() const { return degree_; }
  int     precision() const { return precision_; }
  bool    is_affine() const { return affine_ != 0; }

  void vmult(void * dst, void const * src, void * stream = nullptr) const
  {
    check(exadg_laplace_vmult(h_, dst, src, stream));
  }
  void vmult_add(void * dst, void const * src, void * stream = nullptr) const
  {
    check(exadg_laplace_vmult_add(h_, dst, src, stream));
  }
  void calculate_inverse_diagonal(void * inv_diag, void * stream = nullptr) const
  {
    check(exadg_laplace_calculate_inverse_diagonal(h_, inv_diag, stream));
  }
  // Poisson::Operator::rhs for the manufactured solution prod_d sin(omega x_d)
  void rhs_sine(double omega, double * b, void * stream = nullptr) const
  {
    check(exadg_laplace_rhs_sine(h_, omega, b, stream));
  }
  double error_sine(double omega, double const * u, void * stream = nullptr) const
  {
    double e = 0;
    check(exadg_laplace_error_sine(h_, omega, u, &e, stream));
    return e;
  }
  exadg_laplace * handle() const { return h_; }

private:
  exadg_laplace * h_ = nullptr;
  int64_t         n_ = 0, n_cells_ = 0;
  int             degree_ = 0, precision_ = 0, affine_ = 0;
};

// The factory product of Operator::setup_preconditioner_and_solver (operator.cpp:306-398)
class Preconditioner
{
public:
  // kind: EXADG_PRECOND_*; mg is read for EXADG_PRECOND_MULTIGRID. The operator must outlive
  // the preconditioner (point Jacobi refers to it).
  Preconditioner(LaplaceOperator const & op, int kind, exadg_mg_desc const * mg = nullptr,
                 void * stream = nullptr)
  {
    check(exadg_preconditioner_create(op.handle(), kind, mg, &h_, stream));
  }
  ~Preconditioner() { exadg_preconditioner_destroy(h_); }
  Preconditioner(Preconditioner const &) = delete;
  Preconditioner & operator=(Preconditioner const &) = delete;

  void vmult(double * dst, double const * src, void * stream = nullptr) const
  {
    check(exadg_preconditioner_vmult(h_, dst, src, stream));
  }
  void update(void * stream = nullptr) { check(exadg_preconditioner_update(h_, stream)); }
  exadg_preconditioner * handle() const { return h_; }

private:
  exadg_preconditioner * h_ = nullptr;
};

// MultigridData defaults of the reference for degree k (Bisect p-sequence, FP32 V-cycle,
// Chebyshev(5, 20, 20), coarse CG 1e-3)
inline exadg_mg_desc default_multigrid_data(int degree)
{
  exadg_mg_desc d{};
  check(exadg_mg_desc_default(degree, &d));
  return d;
}

// Krylov::KrylovSolver<..., SolverCG>::solve: returns the iteration count and throws on
// NaN/Inf, on the iteration limit and on a failed final true-residual check, like the
// reference's AssertThrow after SolverControl (iterative_solvers_dealii_wrapper.h:200-221)
inline unsigned int solve(LaplaceOperator const & op, Preconditioner const * prec, double const * b,
                          double * x, exadg_solver_data const & data,
                          exadg_solve_info * info_out = nullptr, void * stream = nullptr)
{
  exadg_solve_info info{};
  int const        rc = exadg_cg_solve(op.handle(), prec ? prec->handle() : nullptr, b, x, &data,
                                       &info, stream);
  if(info_out)
    *info_out = info;
  check(rc);
  return (unsigned int)info.iterations;
}

}  // namespace exadg_amd

#endif /* EXADG_AMD_HPP */
