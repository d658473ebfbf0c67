// CPU oracle multigrid (TEST INFRASTRUCTURE ONLY): hybrid multigrid over h-levels (coarser
// meshes), p-levels (lower degrees) and c-levels (continuous FE_Q) with Chebyshev-Jacobi
// smoothing, FP64.
//
// Reference files followed:
//   V-cycle            : solvers_and_preconditioners/multigrid/multigrid_algorithm.h:87-108,172-240
//   levels             : multigrid/multigrid_preconditioner_base.cpp:97-323
//   transfers          : multigrid/transfer.cpp:28-96 (deal.II MGTwoLevelTransfer)
//   smoother           : multigrid/smoothers/chebyshev_smoother.h:37-227 (deal.II
//                        PreconditionChebyshev: eigenvalue estimate by CG/Lanczos, recurrence)
//   coarse solvers     : multigrid/coarse_grid_solvers.h:61-321,
//                        utilities/compute_eigenvalues.h:38-88
//   continuous levels : multigrid/constraints.h:109-194, operators/operator_base.cpp:299-309
//   level list (types) : multigrid/multigrid_preconditioner_base.cpp:97-323,
//                        multigrid/multigrid_parameters.h:39-65
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

// Continuous FE_Q(p) node numbering of a structured box (GL support points): (n_d p + 1) nodes
// per non-periodic direction, n_d p per periodic one; Dirichlet sides constrained.
struct CGSpace
{
  MeshDesc          d;
  int               p = 1;
  int               nn[3];
  int64_t           n_nodes = 0;
  std::vector<char> constrained;
  CGSpace() = default;
  CGSpace(MeshDesc const & d, int p);
  // global node of local node (i, j, l) (lexicographic index `local`) of cell c
  int64_t node(int64_t c, int local) const;
  // number of cells sharing node g
  int valence(int64_t g) const;
};

// continuous Q_p Laplace on the structured box: the cell integrals of the SIPG operator (its
// face terms vanish on continuous functions), assembled over cells through the node map;
// Dirichlet nodes read as zero and act as identity rows
struct CGOperator
{
  CGSpace                          space;
  std::unique_ptr<LaplaceOperator> cells;  // same mesh and degree; only cell_loop is used
  CGOperator(MeshDesc const & d, int p, double ip_factor);
  int64_t n() const { return space.n_nodes; }
  void    vmult(double * dst, double const * src) const;
  void    diagonal(double * d) const;
};

// BlockJacobiPreconditioner (block_jacobi_preconditioner.h), matrix-based: the cells' diagonal
// blocks A_KK factorized by Cholesky once, applied by forward and back substitution
struct BlockJacobi
{
  int                 m       = 0;
  int64_t             n_cells = 0;
  std::vector<double> L;  // per cell the lower Cholesky factor, row-major m x m
  explicit BlockJacobi(LaplaceOperator const & op);
  void vmult(double * dst, double const * src) const;
};

struct Chebyshev
{
  Apply               A;
  std::vector<double> inv_diag;
  // inner preconditioner other than point Jacobi (block Jacobi); empty: inv_diag
  Apply               P;
  int                 degree = 5;
  double              theta = 0, delta = 0, lambda_max = 0, lambda_min = 0;
  int                 cg_iterations = 0;
  // smoother setup (deal.II PreconditionChebyshev::estimate_eigenvalues)
  void setup_smoother(Apply const & A, std::vector<double> inv_diag, std::vector<char> const & mask,
                      int degree, double range, int eig_its, Apply const & P = Apply());
  // MGCoarseChebyshev::update
  void setup_coarse(Apply const & A, std::vector<double> inv_diag, double rel_tol, bool singular);
  void vmult(std::vector<double> & x, std::vector<double> const & b) const;
  void step(std::vector<double> & x, std::vector<double> const & b) const;
};

// MultigridType (multigrid_parameters.h:39-57), numbered as there; 0 = the legacy choice
// p then c (pcMG) when coarse_q1, else pMG
enum MGType
{
  mg_legacy = 0, mg_hMG, mg_chMG, mg_hcMG, mg_cMG, mg_pMG, mg_cpMG, mg_pcMG, mg_hpMG, mg_chpMG,
  mg_hcpMG, mg_hpcMG, mg_phMG, mg_cphMG, mg_pchMG, mg_phcMG
};

struct LevelInfo
{
  int  h      = 0;  // mesh level: 0 coarsest
  int  degree = 1;
  bool dg     = true;
};

// MultigridPreconditionerBase::initialize_levels: levels coarse to fine from the type, the
// p-sequence (degrees fine to coarse, degrees[0] = operator degree) and the number of meshes
std::vector<LevelInfo> mg_level_list(int type, std::vector<int> const & degrees, int n_h_levels,
                                     bool coarse_q1);

struct MGDesc
{
  std::vector<int> degrees;       // fine to coarse
  bool             coarse_q1 = true;
  int              type = mg_legacy;
  int              n_h_levels = 1;  // meshes of the h-levels (finest included)
  int              smoother_degree = 5;
  double           smoothing_range = 20;
  int              eig_its = 20;
  int              coarse_solver = 0;  // 0 CG+Jacobi, 1 Chebyshev
  double           coarse_rel_tol = 1e-3;
  int              coarse_max_iter = 10000;
  bool             singular = false;
  int              smoother_preconditioner = 0;  // 0 point Jacobi, 1 block Jacobi, 2 additive
                                                 // Schwarz (= block Jacobi on DG levels)
};

struct Multigrid
{
  struct Level
  {
    int                                     h      = 0;
    int                                     degree = 1;
    bool                                    dg     = true;
    int64_t                                 n      = 0;
    MeshDesc                                mesh;
    CGSpace                                 space;  // continuous levels
    std::unique_ptr<LaplaceOperator>        dg_op;
    std::unique_ptr<Q1Operator>             q1_op;
    std::unique_ptr<CGOperator>             cg_op;
    std::unique_ptr<BlockJacobi>            block;
    Apply                                   A;
    std::vector<double>                     inv_diag;
    Chebyshev                               smoother;
    // transfer to the next coarser level: 'p' (degree), 'h' (mesh) or 'c' (DG -> continuous)
    char                                    transfer = 0;
    // 1D interpolation matrices (n_fine x n_coarse, row-major): P for 'p'; Ph[b] for the
    // child with bit b of 'h' (fine coordinate (b + xi)/2 in the parent)
    std::vector<double>                     P, Ph[2];
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
