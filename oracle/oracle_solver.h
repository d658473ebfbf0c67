// CPU oracle solvers (TEST INFRASTRUCTURE ONLY): PCG, Jacobi, Chebyshev, p-multigrid.
//
// Reference files followed:
//   CG                : include/exadg/solvers_and_preconditioners/solvers/
//                       iterative_solvers_dealii_wrapper.h:136-221 (deal.II SolverCG, ReductionControl)
//   Jacobi            : .../preconditioners/jacobi_preconditioner.h:33-82, utilities/invert_diagonal.h:35-46
#pragma once
#include <functional>
#include <vector>

#include "oracle_core.h"

namespace oracle
{
struct SolveInfo
{
  int    iterations = 0;
  double r0         = 0;
  double r_final    = 0;
  bool   converged  = false;
};

using Apply = std::function<void(double *, double const *)>;

// deal.II SolverCG: r = b - Ax, z = P r, p = z; x += alpha p; r -= alpha Ap; stop on ||r||
SolveInfo pcg(int64_t n, Apply const & A, Apply const * P, double const * b, double * x,
              double rel_tol, double abs_tol, int max_iter);

// invert_diagonal: |d| <= 1e-10 -> 1
void invert_diagonal(std::vector<double> & d);

struct JacobiCG
{
  LaplaceOperator const & op;
  std::vector<double>     inv_diag;
  bool                    use_jacobi;
  JacobiCG(LaplaceOperator const & op, bool use_jacobi);
  SolveInfo solve(double const * b, double * x, double rel_tol, double abs_tol, int max_iter);
};

}  // namespace oracle
