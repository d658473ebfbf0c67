// CPU oracle multigrid (TEST INFRASTRUCTURE ONLY): hybrid p-multigrid DG_k -> ... -> DG_1 ->
// continuous Q1 with Chebyshev-Jacobi smoothing, FP64.
//
// Reference files followed:
//   V-cycle            : solvers_and_preconditioners/multigrid/multigrid_algorithm.h:87-108,172-240
//   levels             : multigrid/multigrid_preconditioner_base.cpp:97-323
//   transfers          : multigrid/transfer.cpp:28-96 (deal.II MGTwoLevelTransfer)
//   smoother           : multigrid/smoothers/chebyshev_smoother.h:37-227 (deal.II
//                        PreconditionChebyshev: eigenvalue estimate by CG/Lanczos, recurrence)
//   coarse solvers     : multigrid/coarse_grid_solvers.h:61-321,
//                        utilities/compute_eigenvalues.h:38-88
//   continuous Q1 level: multigrid/constraints.h:109-194, operators/operator_base.cpp:299-309
#pragma once
#include <memory>
#include <vector>

#include "oracle_core.h"
#include "oracle_solver.h"

namespace oracle
{
// continuous Q1 Laplace on the structured box, Dirichlet nodes as identity rows
struct Q1Operator
{
  Mesh                mesh;
  int                 nn[3];
  int64_t             n_nodes;
  std::vector<double> K;          // per cell 8x8 element stiffness (QGauss(2))
  std::vector<char>   constrained;
  explicit Q1Operator(MeshDesc const & d);
  int64_t node_of(int const * cc, int v) const;  // node of vertex v (bits x,y,z) of cell cc
  void    vmult(double * dst, double const * src) const;
  void    diagonal(double * d) const;
};

struct Chebyshev
{
  Apply               A;
  std::vector<double> inv_diag;
  int                 degree = 5;
  double              theta = 0, delta = 0, lambda_max = 0, lambda_min = 0;
  int                 cg_iterations = 0;
  // smoother setup (deal.II PreconditionChebyshev::estimate_eigenvalues)
  void setup_smoother(Apply const & A, std::vector<double> inv_diag, std::vector<char> const & mask,
                      int degree, double range, int eig_its);
  // MGCoarseChebyshev::update
  void setup_coarse(Apply const & A, std::vector<double> inv_diag, double rel_tol, bool singular);
  void vmult(std::vector<double> & x, std::vector<double> const & b) const;
  void step(std::vector<double> & x, std::vector<double> const & b) const;
};

struct MGDesc
{
  std::vector<int> degrees;       // fine to coarse
  bool             coarse_q1 = true;
  int              smoother_degree = 5;
  double           smoothing_range = 20;
  int              eig_its = 20;
  int              coarse_solver = 0;  // 0 CG+Jacobi, 1 Chebyshev
  double           coarse_rel_tol = 1e-3;
  int              coarse_max_iter = 10000;
  bool             singular = false;
};

struct Multigrid
{
  struct Level
  {
    int                                     degree = 1;
    bool                                    dg     = true;
    int64_t                                 n      = 0;
    std::unique_ptr<LaplaceOperator>        dg_op;
    std::unique_ptr<Q1Operator>             q1_op;
    Apply                                   A;
    std::vector<double>                     inv_diag;
    Chebyshev                               smoother;
    std::vector<double>                     P;  // nf x nc to the next coarser DG level
    mutable std::vector<double>             defect, sol, t;
  };
  MeshDesc           mesh;
  MGDesc             desc;
  std::vector<Level> levels;  // 0 = coarsest
  Chebyshev          coarse_cheb;
  mutable int        last_coarse_its = 0;

  Multigrid(MeshDesc const & mesh, MGDesc const & desc, double ip_factor);
  void vmult(double * dst, double const * src) const;
  void prolongate_add(int l, double * fine, double const * coarse) const;
  void restrict_add(int l, double * coarse, double const * fine) const;

private:
  void v_cycle(int l) const;
};

// PCG recording alpha/beta (SolverCG with the eigenvalue slot)
int pcg_record(int64_t n, Apply const & A, Apply const & P, double const * b, double * x,
               double tol_abs, int max_iter, bool singular, std::vector<double> * alphas,
               std::vector<double> * betas);
void tridiagonal_extreme_eigenvalues(std::vector<double> const & alpha,
                                     std::vector<double> const & beta, double & lmin, double & lmax);

}  // namespace oracle
