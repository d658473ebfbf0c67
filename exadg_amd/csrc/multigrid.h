// Hybrid multigrid preconditioner over h-levels (2x-coarsened meshes), p-levels (lower
// degrees) and c-levels (continuous FE_Q), in any of the reference's MultigridType orders
// (default: DG_k -> ... -> DG_1 -> continuous Q1), with Chebyshev-Jacobi smoothing, run in FP32
// or FP64 inside an FP64 outer Krylov solver.
//
// Reference interfaces: MultigridPreconditionerBase::vmult/update
// (solvers_and_preconditioners/multigrid/multigrid_preconditioner_base.cpp:55-95,673-683),
// MultigridAlgorithm::vmult/v_cycle (multigrid/multigrid_algorithm.h:87-108,172-240),
// ChebyshevSmoother (multigrid/smoothers/chebyshev_smoother.h:37-227) over deal.II's
// PreconditionChebyshev, MGCoarseKrylov / MGCoarseChebyshev (multigrid/coarse_grid_solvers.h),
// MultigridData (multigrid/multigrid_parameters.h:39-313), PreconditionerBase
// (preconditioners/preconditioner_base.h:33-67).
#pragma once
#include <memory>
#include <vector>

#include "block_kernels.h"
#include "cg_kernels.h"
#include "common.h"
#include "grid.h"
#include "laplace_operator.h"
#include "mg_kernels.h"
#include "vector_ops.h"

namespace ExaDG
{
// PreconditionerBase<double>: the FP64 "apply_inverse" seen by the outer solver
class PreconditionerBase
{
public:
  virtual ~PreconditionerBase() = default;
  virtual void vmult(double * dst, double const * src, hipStream_t s) = 0;
  virtual void update(hipStream_t s)                                  = 0;
  // dst = S src with S the finest-level smoother (zero initial guess); false if none
  virtual bool smoother_vmult(double * /*dst*/, double const * /*src*/, hipStream_t /*s*/)
  {
    return false;
  }
  // true if vmult may allocate device memory or HIP objects (graph capture on first use)
  virtual bool vmult_allocates() const { return false; }
  // For an outer FP64 Krylov solver around a FP32 preconditioner: narrow_input() is the FP32
  // vector that vmult_narrow reads, and vmult_narrow applies the preconditioner to that vector
  // as it stands and returns its FP32 result (valid until the next call). vmult(dst, src) equals
  // rounding src into narrow_input(), vmult_narrow and widening the result into dst. nullptr
  // when the preconditioner has no FP32 form.
  virtual float * narrow_input() { return nullptr; }
  virtual float const * vmult_narrow(hipStream_t /*s*/) { return nullptr; }
};

// BlockJacobiPreconditioner (block_jacobi_preconditioner.h) on the FP64 operator: the exact
// inverses of the cells' diagonal blocks (matrix-based variant, operator_base.cpp:789-815); on DG
// spaces it is also the AdditiveSchwarzPreconditioner (operator_base.cpp:2002-2009, 2099-2103)
class BlockJacobiPreconditioner : public PreconditionerBase
{
public:
  explicit BlockJacobiPreconditioner(Poisson::LaplaceOperator<double> const & op);
  void vmult(double * dst, double const * src, hipStream_t s) override;
  void update(hipStream_t s) override;

private:
  Poisson::LaplaceOperator<double> const &                   op_;
  std::unique_ptr<Poisson::CellBlockInverse<double>> inv_;
};

// point Jacobi (jacobi_preconditioner.h:33-82)
class JacobiPreconditioner : public PreconditionerBase
{
public:
  explicit JacobiPreconditioner(Poisson::LaplaceOperator<double> const & op);
  void vmult(double * dst, double const * src, hipStream_t s) override;
  void update(hipStream_t s) override;

private:
  Poisson::LaplaceOperator<double> const & op_;
  DeviceVector<double>                     inv_diag_;
};

enum class CoarseSolver : int
{
  cg        = 0,  // MGCoarseKrylov: CG + point Jacobi to a relative tolerance
  chebyshev = 1   // MGCoarseChebyshev: Chebyshev-Jacobi with degree from the tolerance
};

// MultigridType (multigrid_parameters.h:39-57), numbered as there; legacy = pcMG when
// coarse_q1 is set, else pMG
enum class MultigridType : int
{
  legacy = 0, hMG, chMG, hcMG, cMG, pMG, cpMG, pcMG, hpMG, chpMG, hcpMG, hpcMG, phMG, cphMG, pchMG,
  phcMG
};

// one multigrid level: mesh (h = 0 coarsest), degree, discontinuous or continuous
struct MGLevelInfo
{
  int  h      = 0;
  int  degree = 1;
  bool dg     = true;
};

// the level list of MultigridPreconditionerBase::initialize_levels
// (multigrid_preconditioner_base.cpp:97-323), coarse to fine, from the type, the p-sequence
// (degrees fine to coarse) and the number of meshes of the h-levels
std::vector<MGLevelInfo> multigrid_levels(MultigridType type, std::vector<int> const & degrees,
                                          int n_h_levels, bool coarse_q1);

// PreconditionerSmoother (multigrid_parameters.h:161-167): the inner preconditioner of the
// Chebyshev smoother
enum class SmootherPreconditioner : int
{
  point_jacobi     = 0,
  block_jacobi     = 1,  // DG levels only, as in the reference
  additive_schwarz = 2   // on DG levels the block-Jacobi inverse (operator_base.cpp:2002-2009)
};

// point Jacobi (inv_diag) or block Jacobi (the inverse cell blocks) as a level
// preconditioner z = P r
template<typename Number>
struct JacobiData
{
  Number const *                             inv_diag = nullptr;
  Poisson::CellBlockInverse<Number> const * block    = nullptr;
  void apply(Number * z, Number const * r, int64_t n, hipStream_t s) const;
};

struct MultigridData
{
  std::vector<int> degrees = {3, 1};     // p-sequence, fine to coarse (degrees[0] = k)
  bool             coarse_q1 = true;     // legacy type: append a continuous Q1 level
  MultigridType    type      = MultigridType::legacy;
  // meshes of the h-levels, the finest included (types with h-coarsening); 0: halve while
  // every cells-per-direction count is even and the result keeps at least 2
  int              n_h_levels = 0;
  int              smoother_degree = 5;  // SmootherData::iterations
  SmootherPreconditioner smoother_preconditioner = SmootherPreconditioner::point_jacobi;
  double           smoothing_range = 20; // SmootherData::smoothing_range
  int              eig_cg_iterations = 20;
  CoarseSolver     coarse_solver   = CoarseSolver::cg;
  double           coarse_rel_tol  = 1e-3;
  int              coarse_max_iter = 10000;
  bool             operator_is_singular = false;
  // replay the V-cycle as a HIP graph (only when it has no host synchronization, i.e. with
  // the Chebyshev coarse solver); one graph per (dst, src) pair, rebuilt by update()
  bool capture_graph = false;
};

// work vectors of a PCG on one level (reused across solves)
template<typename Number>
struct PcgWorkspace
{
  DeviceVector<Number> r, z, p, v;
  void reinit(int64_t n)
  {
    r.reinit(n);
    z.reinit(n);
    p.reinit(n);
    v.reinit(n);
  }
};

template<typename Number>
class LevelOperator
{
public:
  virtual ~LevelOperator() = default;
  virtual int64_t n() const                                                 = 0;
  virtual void    vmult(Number * dst, Number const * src, hipStream_t s) const = 0;
  virtual void    inverse_diagonal(Number * d, hipStream_t s) const           = 0;
  virtual bool    continuous() const { return false; }
  // node numbering of a continuous level (null for DG levels)
  virtual CGTopology const * cg_topology() const { return nullptr; }
  // the inverse cell blocks of a DG level (block Jacobi)
  virtual std::unique_ptr<Poisson::CellBlockInverse<Number>> block_inverse(hipStream_t /*s*/) const
  {
    EXADG_THROW(EXADG_ERR_ARGUMENT, "Block Jacobi only implemented for DG!");
  }
  // dst = b - A src; operators with a fused epilogue override this (the default applies A and
  // then updates, so it needs dst != b)
  virtual void residual(Number * dst, Number const * src, Number const * b, hipStream_t s) const;
  // coarse += R (b - A src) with the DG p-restriction R = P^T per cell (P: (k+1) x nc) in one
  // pass; false if this operator has no fused form
  virtual bool residual_restrict(Number * /*coarse*/, Number const * /*src*/, Number const * /*b*/,
                                 Number const * /*P*/, int /*nc*/, hipStream_t /*s*/) const
  {
    return false;
  }
  // mask of constrained entries (1 constrained, 0 free) on the host; empty if none
  virtual std::vector<char> constrained_mask() const { return {}; }
  // x_new = (1+f1) x - f1 x_old (if has_old) + f2 dinv (b - A x) in one pass; false if this
  // operator has no fused form (the caller then applies A and updates separately)
  virtual bool chebyshev_step(Number * /*x_new*/, Number const * /*x*/, Number const * /*x_old*/,
                              Number const * /*b*/, Number const * /*dinv*/, Number /*f1*/,
                              Number /*f2*/, bool /*has_old*/, hipStream_t /*s*/) const
  {
    return false;
  }
};

// continuous Q1 operator with strong Dirichlet constraints on the structured box
template<typename Number>
class Q1Operator : public LevelOperator<Number>
{
public:
  explicit Q1Operator(StructuredGrid const & grid);
  int64_t n() const override { return n_nodes_; }
  void    vmult(Number * dst, Number const * src, hipStream_t s) const override;
  void    inverse_diagonal(Number * d, hipStream_t s) const override;
  bool    continuous() const override { return true; }
  std::vector<char>  constrained_mask() const override;
  bool chebyshev_step(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * dinv, Number f1, Number f2, bool has_old,
                      hipStream_t s) const override;
  Q1Topology const & q1_topology() const { return q_; }
  CGTopology const * cg_topology() const override { return &cg_; }
  std::vector<double> const & element_matrices() const { return mats_host_; }

private:
  BoxTopology          topo_;
  Q1Topology           q_;
  CGTopology           cg_;
  int64_t              n_nodes_;
  std::vector<double>  mats_host_;  // 8x8 element matrices (one shared one if affine)
  DeviceBuffer<Number> stencil_;    // assembled 27-point stencil, 27 planes of n_nodes
  DeviceBuffer<Number> diag_;       // sum over cells of K_vv (1 on constrained rows)
};

// continuous Q_p Laplace (p >= 2) on the structured box: the cell integrals of the SIPG form
// (its face terms vanish on continuous functions), summed over the cells sharing a node;
// Dirichlet nodes read as zero and act as identity rows (operator_base.cpp:299-309)
template<typename Number>
class CGOperator : public LevelOperator<Number>
{
public:
  CGOperator(StructuredGrid const & grid, int degree);
  int64_t n() const override { return n_nodes_; }
  void    vmult(Number * dst, Number const * src, hipStream_t s) const override;
  void    inverse_diagonal(Number * d, hipStream_t s) const override;
  void    residual(Number * dst, Number const * src, Number const * b, hipStream_t s) const override;
  bool    continuous() const override { return true; }
  CGTopology const * cg_topology() const override { return &q_; }
  bool chebyshev_step(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                      Number const * dinv, Number f1, Number f2, bool has_old,
                      hipStream_t s) const override;

private:
  CGCellArgs<Number> cell_args(Number const * src, bool diagonal) const;
  void               assemble(CGAssembleArgs<Number> a, CGAssembleMode mode, hipStream_t s) const;

  BoxTopology                  topo_;
  CGTopology                   q_;
  int64_t                      n_nodes_ = 0, n_cells_ = 0;
  int                          n3_      = 0;
  Number                       hscale_[3];
  DeviceBuffer<Number>         tables_, metric_;
  mutable DeviceVector<Number> cell_values_;
};

// Chebyshev polynomial smoother with point-Jacobi inner preconditioner (deal.II
// PreconditionChebyshev semantics: vmult assumes a zero initial guess, step does not)
template<typename Number>
class ChebyshevSmoother
{
public:
  void setup(LevelOperator<Number> const & op, int degree, double smoothing_range,
             int eig_cg_iterations, hipStream_t s,
             SmootherPreconditioner prec = SmootherPreconditioner::point_jacobi);
  // coarse-grid variant: interval [lambda_min, 1.1 lambda_max] from a CG estimate to 1e-5
  void setup_coarse(LevelOperator<Number> const & op, double rel_tol, bool singular, hipStream_t s);
  // x: current solution buffer, x_alt: second buffer; both may be swapped on return
  void vmult(Number *& x, Number *& x_alt, Number const * b, hipStream_t s) const;
  void step(Number *& x, Number *& x_alt, Number const * b, hipStream_t s) const;

  double lambda_max_estimate = 0, lambda_min_estimate = 0, theta = 0, delta = 0;
  int    degree = 0, cg_iterations = 0;

private:
  // x_new = (1+f1) x - f1 x_old (has_old) + f2 D^{-1} (b - A x), fused where the operator can
  void apply_step(Number * x_new, Number const * x, Number const * x_old, Number const * b,
                  Number f1, Number f2, bool has_old, hipStream_t s) const;

  JacobiData<Number> jacobi_data() const;

  LevelOperator<Number> const *                      op_ = nullptr;
  DeviceVector<Number>                               inv_diag_;
  std::unique_ptr<Poisson::CellBlockInverse<Number>> block_;  // block Jacobi
  mutable DeviceVector<Number>                       t_;
};

template<typename Number>
class MultigridPreconditioner : public PreconditionerBase
{
public:
  MultigridPreconditioner(std::shared_ptr<StructuredGrid const> grid, MultigridData const & data,
                          double ip_factor, hipStream_t s);
  ~MultigridPreconditioner() override;
  MultigridPreconditioner(MultigridPreconditioner const &) = delete;
  MultigridPreconditioner & operator=(MultigridPreconditioner const &) = delete;
  // true if vmult replays captured graphs
  bool uses_graph() const
  {
    return data_.capture_graph && data_.coarse_solver == CoarseSolver::chebyshev;
  }
  void vmult(double * dst, double const * src, hipStream_t s) override;
  void update(hipStream_t s) override;
  bool smoother_vmult(double * dst, double const * src, hipStream_t s) override;
  bool vmult_allocates() const override { return uses_graph(); }
  float *       narrow_input() override;
  float const * vmult_narrow(hipStream_t s) override;

  int n_levels() const { return (int)levels_.size(); }
  MGLevelInfo level_info(int l) const { return {levels_[l].h, levels_[l].degree, levels_[l].dg}; }
  // level l: 0 = coarsest
  LevelOperator<Number> const & level_operator(int l) const { return *levels_[l].op; }
  ChebyshevSmoother<Number> const & smoother(int l) const { return levels_[l].smoother; }
  ChebyshevSmoother<Number> const & coarse_chebyshev() const { return coarse_cheb_; }
  int                              last_coarse_iterations() const { return last_coarse_its_; }
  // components for tests: transfer between level l-1 and l
  void prolongate_and_add(int l, Number * dst_fine, Number const * src_coarse, hipStream_t s) const;
  void restrict_and_add(int l, Number * dst_coarse, Number const * src_fine, hipStream_t s) const;

private:
  struct Level
  {
    std::unique_ptr<LevelOperator<Number>> op;
    int                                    h      = 0;
    int                                    degree = 1;
    bool                                   dg     = true;
    std::shared_ptr<StructuredGrid const>  grid;
    DeviceVector<Number>                   defect, sol_a, sol_b, t;
    Number *                               sol = nullptr, *sol_alt = nullptr;
    ChebyshevSmoother<Number>              smoother;
    // transfer to the next coarser level: 'p', 'h' or 'c'
    char                                   transfer = 0;
    // DG: p-transfer matrix (n_fine x n_coarse) or the two child matrices of the h-transfer
    DeviceBuffer<Number>                   P;
    // continuous p/h-transfer: per direction a prolongation and a restriction line operator
    // (indices and weights), and two work arrays for the separable passes
    DeviceBuffer<int>                      line_idx;
    DeviceBuffer<Number>                   line_w;
    LineOperator<Number>                   line_prolong[3], line_restrict[3];
    DeviceVector<Number>                   work_a, work_b;
  };
  void v_cycle(int l, hipStream_t s);
  void coarse_solve(hipStream_t s);
  // one V-cycle from src (NULL: the finest defect as it stands) into dst (NULL: none); returns
  // the finest level's solution vector of this cycle
  Number const * run(double * dst, double const * src, hipStream_t s);
  void           vmult_eager(double * dst, double const * src, hipStream_t s);
  void           clear_graphs();

  struct Graph
  {
    double *        dst;
    double const *  src;
    hipGraphExec_t  exec;
    Number const *  out;  // the finest solution vector the captured cycle ends in
  };
  std::vector<Graph> graphs_;
  hipStream_t        capture_stream_ = nullptr;
  int                eager_calls_    = 0;

  std::shared_ptr<StructuredGrid const> grid_;
  MultigridData                         data_;
  std::vector<Level>                    levels_;
  ChebyshevSmoother<Number>             coarse_cheb_;
  DeviceVector<Number>                  coarse_inv_diag_, cr_;
  PcgWorkspace<Number>                  coarse_ws_;
  Reducer                               red_;
  int                                   last_coarse_its_ = 0;
};

std::unique_ptr<PreconditionerBase> create_multigrid(std::shared_ptr<StructuredGrid const> grid,
                                                     MultigridData const & data, double ip_factor,
                                                     int precision, hipStream_t s);

// Lanczos tridiagonal of CG (diag 1/alpha_k + beta_{k-1}/alpha_{k-1}, off sqrt(beta_k)/alpha_k):
// smallest and largest eigenvalue (SolverCG's eigenvalue slot)
void cg_tridiagonal_extreme_eigenvalues(std::vector<double> const & alpha,
                                        std::vector<double> const & beta, double & lmin,
                                        double & lmax);

}  // namespace ExaDG
