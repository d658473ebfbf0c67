// Outer FP64 preconditioned CG on the device: Krylov::KrylovSolver::solve with deal.II's
// SolverCG and ReductionControl (solvers/iterative_solvers_dealii_wrapper.h:44-246,
// solver_data.h:91-146). Scalars (alpha, beta) stay on the device; the host reads one
// residual norm per iteration for the stopping test and the NaN check (:210-211).
#pragma once
#include "laplace_operator.h"
#include "multigrid.h"
#include "vector_ops.h"

namespace ExaDG
{
namespace Krylov
{
struct SolverData
{
  int    max_iter = 10000;
  double abs_tol  = 1e-20;
  double rel_tol  = 1e-10;
  // 0: ||r|| <= max(abs, rel ||r0||) (ReductionControl); 1: relative to ||b||
  int rel_mode = 0;
};

struct SolveInfo
{
  int    iterations = 0;
  double r0 = 0, r_final = 0, b_norm = 0;
  bool   converged = false, not_finite = false;
  // final FP64 true residual ||b - A x|| and whether it is within 10x the stopping tolerance
  // (the recursively updated residual drifts from it when the preconditioner is inexact)
  double r_true = 0, tolerance = 0;
  bool   true_residual_ok = true;
  // performance metrics (iterative_solvers_dealii_wrapper.h:63-78)
  double rho = 0, n10 = 0;
};

// the CG loop on caller-provided work vectors r, z, p, v (length A.m(); z unused without a
// preconditioner) and reduction workspace; allocates nothing itself
SolveInfo cg_solve_in(Poisson::LaplaceOperator<double> const & A, PreconditionerBase * P,
                      double const * b, double * x, SolverData const & data, double * r, double * z,
                      double * p, double * v, Reducer & red, hipStream_t s);

class SolverCG
{
public:
  SolveInfo solve(Poisson::LaplaceOperator<double> const & A, PreconditionerBase * P,
                  double const * b, double * x, SolverData const & data, hipStream_t s);

private:
  PcgWorkspace<double> ws_;
  Reducer              red_;
};
}  // namespace Krylov
}  // namespace ExaDG
