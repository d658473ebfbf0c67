/*
 * exadg_amd.h -- C ABI of the MI355X-native SIPG Laplace hot path (ExaDG drop-in).
 *
 * Every entry point returns an int status (EXADG_OK = 0) and never throws. Vectors are
 * DEVICE pointers (HBM) of length n_dofs in the precision the handle was created with,
 * unless stated otherwise. Calls are ordered on the HIP stream passed as `stream`
 * (a hipStream_t; NULL = the null stream). Handles are not thread-safe: one handle must
 * not be used by two host threads at once (the reference operators mutate `mutable`
 * state in const methods, operator_base.h:473, laplace_operator.h:206).
 *
 * DoF layout: cells lexicographic over the structured box (x fastest), then the
 * (k+1)^3 nodal values of FE_DGQ(k) on Gauss-Lobatto points, x fastest.
 */
#ifndef EXADG_AMD_H
#define EXADG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
enum exadg_status
{
  EXADG_OK               = 0,
  EXADG_ERR_ARGUMENT     = 1, /* bad argument (sizes, aliasing, unsupported degree, ...)   */
  EXADG_ERR_HIP          = 2, /* HIP runtime error                                          */
  EXADG_ERR_NOT_FINITE   = 3, /* NaN/Inf detected (AssertThrow(std::isfinite) in the ref.) */
  EXADG_ERR_NO_CONVERGE  = 4, /* iteration limit reached (SolverControl::NoConvergence)     */
  EXADG_ERR_NOT_READY    = 5, /* object needs update() before use                           */
  EXADG_ERR_INTERNAL     = 6,
  EXADG_ERR_TRUE_RESIDUAL = 7 /* solver converged on its recursive residual, but the final FP64
                                 true residual ||b - A x|| exceeds 10x the tolerance          */
};

enum exadg_precision
{
  EXADG_FP64 = 0,
  EXADG_FP32 = 1
};

enum exadg_boundary_type
{
  EXADG_BC_DIRICHLET = 0,
  EXADG_BC_NEUMANN   = 1
};

/* Message of the last error on the calling thread (never NULL). */
const char * exadg_last_error(void);

/* ---------------------------------------------------------------- mesh / operator */

/* Structured hexahedral box, the grids the reference builds with
 * GridGenerator::subdivided_hyper_cube (sine/application.h:206) and create_periodic_box
 * (grid/periodic_box.h:35-87), optionally deformed by DeformedCubeManifold
 * (grid/deformed_cube_manifold.h:47-157) and described by MappingQ(mapping_degree)
 * (grid/grid_utilities.h:52-58). deformation == 0 gives an affine Cartesian mesh. */
typedef struct exadg_mesh_desc
{
  int32_t n_cells[3];
  double  lower[3];
  double  upper[3];
  double  deformation;
  int32_t frequency;
  int32_t mapping_degree;
  int32_t periodic[3];
} exadg_mesh_desc;

/* Boundary descriptor (poisson/user_interface/boundary_descriptor.h:66-80): the type of
 * each box side 0:x-, 1:x+, 2:y-, 3:y+, 4:z-, 5:z+ (ignored for periodic directions). */
typedef struct exadg_bc_desc
{
  int32_t type[6];
} exadg_bc_desc;

typedef struct exadg_laplace exadg_laplace;

/* Replaces LaplaceOperator::initialize (laplace_operator.h:250-254) together with the
 * MatrixFree / mapping setup of Poisson::Operator::setup (operator.cpp:261-284) and the
 * penalty computation IP::calculate_penalty_parameter (interior_penalty_parameter.h:43-99).
 * degree 1..8; precision EXADG_FP64 or EXADG_FP32; ip_factor = LaplaceKernelData::IP_factor. */
int exadg_laplace_create(const exadg_mesh_desc * mesh, const exadg_bc_desc * bc, int degree,
                         int precision, double ip_factor, exadg_laplace ** out);
int exadg_laplace_destroy(exadg_laplace * op);

/* OperatorBase::m() / n() (operator_base.h:187-191). */
int exadg_laplace_n_dofs(const exadg_laplace * op, int64_t * n_dofs);
/* Number of cells, degree, precision, and whether the affine (Cartesian) kernel is used. */
int exadg_laplace_info(const exadg_laplace * op, int64_t * n_cells, int * degree, int * precision,
                       int * is_affine);

/* OperatorBase::vmult(dst, src) (operator_base.h:175, operator_base.cpp:156-168 ->
 * apply :264-310): dst = A src for the homogeneous SIPG operator. dst is overwritten;
 * dst == src is rejected with EXADG_ERR_ARGUMENT. */
int exadg_laplace_vmult(exadg_laplace * op, void * dst, const void * src, void * stream);
/* OperatorBase::vmult_add (operator_base.h:178): dst += A src. */
int exadg_laplace_vmult_add(exadg_laplace * op, void * dst, const void * src, void * stream);

/* OperatorBase::calculate_inverse_diagonal (operator_base.cpp:249-262): diag(A) inverted
 * with invert_diagonal's guard |d| <= 1e-10 -> 1 (invert_diagonal.h:35-46). */
int exadg_laplace_calculate_inverse_diagonal(exadg_laplace * op, void * inv_diag, void * stream);
/* diag(A) without inversion (OperatorBase::calculate_diagonal, operator_base.cpp:608-616). */
int exadg_laplace_calculate_diagonal(exadg_laplace * op, void * diag, void * stream);

/* Poisson::Operator::rhs (operator.cpp:414-423) for the manufactured solution
 * u = prod_d sin(omega x_d): f = 3 omega^2 u, Dirichlet g = u, Neumann h = grad(u).n.
 * b is always FP64. (sine/application.h:32-90 with omega = 3 pi.) Assembled on the device;
 * the first call computes the mesh's right-hand-side point data. */
int exadg_laplace_rhs_sine(exadg_laplace * op, double omega, double * b, void * stream);

/* Right-hand side for general data. The reference evaluates dealii::Function objects at the
 * quadrature points (FunctionEvaluator, functions_and_boundary_conditions/evaluate_functions.h:
 * 151-198); here the caller evaluates its functions at the points this library names and passes
 * the values.
 *
 * Points: the (k+1)^3 Gauss points of every cell (cells lexicographic, points x fastest), and
 * the (k+1)^2 Gauss points of every boundary face of the non-periodic sides: sides in the order
 * 0:x- 1:x+ 2:y- 3:y+ 4:z- 5:z+, on a side the faces lexicographic in the two tangential cell
 * coordinates (lower direction fastest), on a face the points with the lower tangential
 * direction fastest. */
int exadg_laplace_rhs_n_points(const exadg_laplace * op, int64_t * n_cell_points,
                               int64_t * n_boundary_points);
/* Physical coordinates of those points: FP64 device arrays, component-major (all x, then all y,
 * then all z), 3 * n_cell_points and 3 * n_boundary_points doubles; either may be NULL. */
int exadg_laplace_rhs_points(exadg_laplace * op, double * x_cell, double * x_boundary, void * stream);
/* OperatorBase::rhs (operator_base.cpp:507-554): RHSOperator::evaluate_add (operators/
 * rhs_operator.cpp:52-95) plus the inhomogeneous boundary terms moved to the right-hand side
 * (boundary_face_loop_inhom_operator, operator_base.cpp:1436-1491):
 *   b = int f v + sum_{Dirichlet faces} int (-g dv/dn + 2 tau g v) + sum_{Neumann faces} int h v
 * with f at the cell points and, at the boundary points, g on Dirichlet sides and h = grad(u).n
 * on Neumann sides (FP64 device arrays; NULL = zero). b (FP64) is overwritten. */
int exadg_laplace_rhs(exadg_laplace * op, const double * f, const double * boundary, double * b,
                      void * stream);


/* ---------------------------------------------------------------- preconditioners / solver */

enum exadg_preconditioner_kind
{
  EXADG_PRECOND_NONE             = 0,
  EXADG_PRECOND_POINT_JACOBI     = 1,
  EXADG_PRECOND_MULTIGRID        = 2,
  EXADG_PRECOND_BLOCK_JACOBI     = 3, /* exact inverse cell blocks (curved: degree <= 4)    */
  EXADG_PRECOND_ADDITIVE_SCHWARZ = 4  /* on the DG space: the block-Jacobi inverse          */
};

/* PreconditionerSmoother (multigrid_parameters.h:161-167): the inner preconditioner of the
 * Chebyshev smoother. BLOCK_JACOBI and ADDITIVE_SCHWARZ need DG levels, of degree <= 4 on
 * curved meshes (the reference refuses block Jacobi on continuous levels; additive Schwarz on
 * continuous levels is not implemented here). */
enum exadg_smoother_preconditioner
{
  EXADG_SMOOTHER_POINT_JACOBI     = 0,
  EXADG_SMOOTHER_BLOCK_JACOBI     = 1,
  EXADG_SMOOTHER_ADDITIVE_SCHWARZ = 2
};

enum exadg_coarse_solver
{
  EXADG_COARSE_CG        = 0, /* MGCoarseKrylov: CG + point Jacobi to coarse_rel_tol       */
  EXADG_COARSE_CHEBYSHEV = 1  /* MGCoarseChebyshev: degree chosen for coarse_rel_tol        */
};

/* MultigridType (multigrid_parameters.h:39-57), numbered as there. EXADG_MG_LEGACY selects
 * pcMG when coarse_q1 = 1 and pMG otherwise. */
enum exadg_mg_type
{
  EXADG_MG_LEGACY = 0,
  EXADG_MG_HMG    = 1,
  EXADG_MG_CHMG   = 2,
  EXADG_MG_HCMG   = 3,
  EXADG_MG_CMG    = 4,
  EXADG_MG_PMG    = 5,
  EXADG_MG_CPMG   = 6,
  EXADG_MG_PCMG   = 7,
  EXADG_MG_HPMG   = 8,
  EXADG_MG_CHPMG  = 9,
  EXADG_MG_HCPMG  = 10,
  EXADG_MG_HPCMG  = 11,
  EXADG_MG_PHMG   = 12,
  EXADG_MG_CPHMG  = 13,
  EXADG_MG_PCHMG  = 14,
  EXADG_MG_PHCMG  = 15
};

/* MultigridData (solvers_and_preconditioners/multigrid/multigrid_parameters.h:39-313) of the
 * hybrid multigrid. dg_degrees is the p-sequence, fine to coarse (dg_degrees[0] = operator
 * degree, strictly decreasing); mg_type orders the h-, p- and c-levels as
 * MultigridPreconditionerBase::initialize_levels does (multigrid_preconditioner_base.cpp:97-323):
 * c-levels are continuous FE_Q of the same degree, h-levels the box with every cell count
 * halved (n_h_levels meshes including the finest; 0 = halve while every count is even and keeps
 * at least 2 cells). With EXADG_MG_LEGACY the last degree must be 1 when coarse_q1 = 1.
 * precision is the V-cycle's floating-point type (MultigridNumber). */
typedef struct exadg_mg_desc
{
  int32_t n_dg_levels;
  int32_t dg_degrees[16];
  int32_t coarse_q1;
  int32_t precision;
  int32_t smoother_degree;   /* Chebyshev degree, SmootherData::iterations (5)        */
  double  smoothing_range;   /* SmootherData::smoothing_range (20)                    */
  int32_t eig_cg_iterations; /* SmootherData::iterations_eigenvalue_estimation (20)   */
  int32_t coarse_solver;     /* exadg_coarse_solver                                   */
  double  coarse_rel_tol;    /* CoarseGridData solver_data.rel_tol (1e-3)             */
  int32_t coarse_max_iter;
  int32_t operator_is_singular;
  int32_t capture_graph;     /* replay the V-cycle as a HIP graph when it has no host
                                synchronization (Chebyshev coarse solver); default 0: on
                                MI355X the replay measured no faster than eager launches */
  int32_t mg_type;           /* exadg_mg_type (default EXADG_MG_LEGACY)                */
  int32_t n_h_levels;        /* meshes of the h-levels, 0 = automatic (default)        */
  int32_t smoother_preconditioner; /* exadg_smoother_preconditioner (default point Jacobi) */
} exadg_mg_desc;

/* Defaults of the reference for degree k: PSequenceType::Bisect (p -> max(p/2,1),
 * multigrid_preconditioner_base.cpp:202-216), c-transfer to continuous Q1 after p-coarsening,
 * FP32 V-cycle (multigrid_preconditioner_base.h:60), Chebyshev(5, 20, 20), coarse CG 1e-3. */
int exadg_mg_desc_default(int degree, exadg_mg_desc * out);

typedef struct exadg_preconditioner exadg_preconditioner;

/* The preconditioner factory of Poisson::Operator::setup_preconditioner_and_solver
 * (operator.cpp:306-398) for None / PointJacobi / BlockJacobi / AdditiveSchwarz / Multigrid;
 * `mg` is read for MULTIGRID only. BlockJacobi (block_jacobi_preconditioner.h) applies the exact
 * inverses of the cells' diagonal blocks A_KK: on affine meshes by fast diagonalization of
 * their Kronecker form (any degree), on curved meshes as dense inverses computed through the
 * cell-based loop at setup (matrix-based variant, operator_base.cpp:789-815; degree <= 4).
 * The operator must be FP64 (the outer solver's precision). Setup includes the diagonals,
 * the Chebyshev eigenvalue estimates and the coarse solver (PreconditionerBase::update).
 * Lifetime: a POINT_JACOBI, BLOCK_JACOBI or ADDITIVE_SCHWARZ preconditioner refers to `op` (as
 * the reference's JacobiPreconditioner and BlockJacobiPreconditioner hold a reference to their
 * operator, jacobi_preconditioner.h:33-82), so `op` must outlive it; a MULTIGRID one shares the
 * operator's grid and owns its levels. */
int exadg_preconditioner_create(exadg_laplace * op, int kind, const exadg_mg_desc * mg,
                                exadg_preconditioner ** out, void * stream);
/* PreconditionerBase::vmult (preconditioner_base.h:33-67): dst = P^{-1} src, FP64 vectors. */
int exadg_preconditioner_vmult(exadg_preconditioner * p, double * dst, const double * src,
                               void * stream);
/* x <- x + S (b - A x), S the finest-level Chebyshev-Jacobi smoother of a MULTIGRID
 * preconditioner applied with a zero initial guess (PreconditionChebyshev::vmult on the
 * residual, in the V-cycle's precision; residual and update in FP64). No counterpart in the
 * reference: it polishes an initial guess (the learned one) before the CG solve. */
int exadg_preconditioner_smooth(exadg_laplace * op, exadg_preconditioner * p, const double * b,
                                double * x, void * stream);
/* The multigrid level list, coarse (0) to fine: mesh index h (0 = coarsest mesh), degree,
 * discontinuous (1) or continuous (0), and size; arrays of at least max_levels entries (any
 * may be NULL); *n_levels receives the count. */
int exadg_preconditioner_mg_levels(const exadg_preconditioner * p, int max_levels, int * n_levels,
                                   int * h, int * degree, int * dg, int64_t * n);
/* PreconditionerBase::update: recompute diagonals / eigenvalue estimates. */
int exadg_preconditioner_update(exadg_preconditioner * p, void * stream);
int exadg_preconditioner_destroy(exadg_preconditioner * p);
/* Multigrid diagnostics: number of levels and, per level (0 = coarsest), the CG estimate of
 * lambda_max(D^{-1}A) used by the Chebyshev smoother (coarse level: the coarse solver's). */
int exadg_preconditioner_mg_info(const exadg_preconditioner * p, int * n_levels,
                                 double * lambda_max, int * chebyshev_degree,
                                 int * last_coarse_iterations);

/* SolverData (solvers_and_preconditioners/solvers/solver_data.h:91-146) */
typedef struct exadg_solver_data
{
  int32_t max_iter;
  double  abs_tol;
  double  rel_tol;
  int32_t rel_mode; /* 0: ||r|| <= max(abs, rel ||r0||) (ReductionControl); 1: rel ||b|| */
} exadg_solver_data;

typedef struct exadg_solve_info
{
  int32_t iterations;
  int32_t converged;
  int32_t not_finite;
  double  r0;
  double  r_final;
  double  b_norm;
  double  rho; /* average convergence rate (iterative_solvers_dealii_wrapper.h:63-78) */
  double  n10; /* iterations per 10 orders of magnitude */
  double  r_true;    /* final true residual ||b - A x||, FP64 (one residual kernel + norm)  */
  double  tolerance; /* the stopping tolerance max(abs_tol, rel_tol * ||r0|| or ||b||)      */
} exadg_solve_info;

/* Krylov::KrylovSolver::solve with CG (iterative_solvers_dealii_wrapper.h:136-221) on the FP64
 * operator, preconditioned by `prec` (NULL = none; it must have been created for `op` or for an
 * operator of the same size, else EXADG_ERR_ARGUMENT); x holds the initial guess on entry.
 * Returns EXADG_ERR_NOT_FINITE on NaN/Inf (the NaN check at :210-211), EXADG_ERR_NO_CONVERGE
 * when max_iter is hit, and EXADG_ERR_TRUE_RESIDUAL when the final FP64 true residual
 * ||b - A x|| exceeds 10x the tolerance although the recursive residual met it (info is filled
 * in every case; converged = 0 in all three). */
int exadg_cg_solve(exadg_laplace * op, exadg_preconditioner * prec, const double * b, double * x,
                   const exadg_solver_data * data, exadg_solve_info * info, void * stream);

/* Largest operator (DoF) that exadg_cg_solve_noalloc accepts. */
#define EXADG_NOALLOC_MAX_DOFS (INT64_C(1) << 22)

/* No counterpart in the reference: exadg_cg_solve for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc). Same arguments, results and
 * status codes as exadg_cg_solve, and the same kernels in the same order (so the same x), but
 * the CG work vectors and the reduction buffers live in a static device array that each GPU
 * holds from library load until the process ends (4 * EXADG_NOALLOC_MAX_DOFS + 2057 doubles,
 * about 128 MiB); the call allocates nothing. Rejected with EXADG_ERR_ARGUMENT: an operator
 * with more than EXADG_NOALLOC_MAX_DOFS DoF, and a multigrid preconditioner created with
 * capture_graph (its first replay captures a graph). The preconditioner and operator must have
 * been created (and updated) beforehand. Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (two streams, two threads) race and must be serialized by the
 * caller. */
int exadg_cg_solve_noalloc(exadg_laplace * op, exadg_preconditioner * prec, const double * b,
                           double * x, const exadg_solver_data * data, exadg_solve_info * info,
                           void * stream);

/* Relative L2 error of u (FP64, device) against u* = prod_d sin(omega x_d) with QGauss(k+3)
 * on the mapped cells: the number the reference's sine application prints
 * (ErrorCalculator::calculate_error, postprocessor/error_calculation.cpp:38-103). FP64
 * operators; the first call on a curved mesh computes its point data. */
int exadg_laplace_error_sine(exadg_laplace * op, double omega, const double * u, double * rel_error,
                             void * stream);

/* ---------------------------------------------------------------- pressure sequences */

/* Seeded solenoidal Fourier velocity field u(x,t) = sum_m A_m d_m cos(2 pi n_m.x + phi_m + omega_m t)
 * over the integer shell k_min <= |n_m| <= k_max (one of each pair +-n): d_m = (n_m/|n_m|) x xi_m
 * with xi_m uniform on the unit sphere, A_m = |n_m|^{-5/3}, phi_m uniform in [0, 2 pi),
 * omega_m = |n_m|. modes (host): 10 doubles per mode, n(3) d(3) A phi omega 0; with
 * modes == NULL only *n_modes (the shell size) is returned. */
int exadg_fourier_modes(uint64_t seed, int k_min, int k_max, int max_modes, double * modes,
                        int * n_modes);

/* Right-hand side of the pressure step: the weak DG divergence with central flux and
 * no-slip walls (divergence_operator.h:63-68, divergence_operator.cpp:168-298) of the Fourier
 * field at time t, mean removed when the operator is singular (no Dirichlet side). Affine and
 * curved meshes (the first call on a curved mesh computes its point data); b is FP64
 * (device). */
int exadg_laplace_rhs_divergence_fourier(exadg_laplace * op, const double * modes, int n_modes,
                                         double t, double * b, void * stream);

/* v[i] = uniform[-1,1] from SplitMix64 on (seed, i) (device, FP64 or FP32 by `precision`): the
 * source vector of the operator-throughput measurement. The reference driver applies the
 * operator to src = 1 (Driver::apply_operator, poisson/driver.cpp:190-199); a seeded
 * counter-based vector gives the same input on the host and the device at any size. */
int exadg_vector_random(void * v, int64_t n, int precision, uint64_t seed, void * stream);

/* v -= mean(v) (FP64, device): the pressure level of a singular problem. */
int exadg_vector_subtract_mean(double * v, int64_t n, void * stream);

/* ---------------------------------------------------------------- pressure step (INS dual splitting) */

/* The weak DG divergence of a discrete velocity on the pressure space of `op`:
 * IncNS::DivergenceOperator (incompressible_navier_stokes/spatial_discretization/operators/
 * divergence_operator.{h,cpp}, :168-298) with the weak formulation, the central flux
 * (divergence_operator.h:63-68) and boundary data. The velocity lives in FE_DGQ(degree_u)^3 on
 * the same mesh: per cell three component blocks (x, y, z) of (degree_u+1)^3 nodal values on
 * Gauss-Lobatto points; QGauss(degree_u+1). Boundaries follow the pressure boundary types of
 * `op`: a pressure Neumann side is a wall at rest (velocity Dirichlet, u+ = -u-, no flux), a
 * pressure Dirichlet side an outflow (velocity Neumann, u+ = u-). FP64. The operator shares the
 * mesh of `op` but not its lifetime. */
typedef struct exadg_divergence exadg_divergence;
int exadg_divergence_create(const exadg_laplace * op, int degree_u, exadg_divergence ** out);
int exadg_divergence_destroy(exadg_divergence * div);
int exadg_divergence_n_dofs(const exadg_divergence * div, int64_t * n_velocity,
                            int64_t * n_pressure);
/* DivergenceOperator::apply (homogeneous operator): dst (pressure) = B src (velocity). */
int exadg_divergence_apply(exadg_divergence * div, double * dst, const double * src, void * stream);
/* Positions of the velocity nodes, component-major: 3 * n_cells * (degree_u+1)^3 doubles. */
int exadg_divergence_velocity_points(exadg_divergence * div, double * x, void * stream);
/* Nodal interpolation of the Fourier field of exadg_fourier_modes (|n_d| <= 8) at time t into
 * velocity DoFs u (a synthetic intermediate velocity for pressure sequences). */
int exadg_divergence_interpolate_fourier(exadg_divergence * div, const double * modes, int n_modes,
                                         double t, double * u, void * stream);
/* TimeIntBDFDualSplitting::pressure_step (time_int_bdf_dual_splitting.cpp:478-560) for walls at
 * rest and no body force: rhs = scale * B velocity (scale = -gamma0 / dt, rhs_pressure :536-545),
 * mean removed when the pressure problem is singular; p_np = sum_i beta_i history[i]
 * (ExtrapolationConstants of `order` 1..4, newest first) or, with order 0, p_np as passed
 * (pressure_last_iter); then the CG solve with `prec` as exadg_cg_solve, and the mean of p_np
 * removed again for a singular problem (adjust_pressure_level_if_undefined). rhs (n_dofs of op)
 * returns the right-hand side. Same status codes as exadg_cg_solve. */
int exadg_pressure_step(exadg_laplace * op, exadg_preconditioner * prec, exadg_divergence * div,
                        const double * velocity, double scale, int order,
                        const double * const * history, double * p_np, double * rhs,
                        const exadg_solver_data * data, exadg_solve_info * info, void * stream);

/* Initial guess by extrapolation, x0 = sum_i beta_i p^{n-i} with the constants of
 * ExtrapolationConstants::set_constant_time_step (extrapolation_constants.cpp:55-93), as used for
 * the pressure at time_int_bdf_dual_splitting.cpp:489-501. history: host array of `order`
 * device pointers, newest first. */
int exadg_extrapolate(int order, const double * const * history, double * x0, int64_t n,
                      void * stream);

/* Learned initial guess (no counterpart in the reference; it replaces the extrapolated guess
 * above, which remains its fallback). A per-cell MLP, FP32, of `width` neurons in two hidden
 * ReLU layers. Input on the cell's (k+1)^3 DoFs: the backward differences of orders
 * 0..history-1 of p^{n-1}, p^{n-2}, ... and b^n, each divided by its scale. Output: the
 * correction of the order-`history` extrapolation divided by s_t. */
typedef struct exadg_guess exadg_guess;
int exadg_guess_create(int degree, int history, int width, uint64_t seed, exadg_guess ** out);
int exadg_guess_destroy(exadg_guess * g);
int exadg_guess_n_params(const exadg_guess * g, int * n_params, int * n_features, int * n_outputs);
/* copy the flat FP32 parameters [W1 | b1 | W2 | b2 | W3 | b3] (W row-major, out x in) to / from
 * n_params floats of device memory; stream-ordered, returns after the copy */
int exadg_guess_get_params(const exadg_guess * g, float * dst, void * stream);
int exadg_guess_set_params(exadg_guess * g, const float * src, void * stream);
/* scales [s_0 .. s_{history-1}, s_b, s_t] (n = history + 2): the RMS of the j-th backward
 * difference of the history, of the rhs and of the extrapolation error; train() sets them */
int exadg_guess_set_scales(exadg_guess * g, const double * scales, int n);
int exadg_guess_get_scales(const exadg_guess * g, double * scales, int n);
/* mean squared error over B samples of features X (B x n_features) and targets T (B x n_outputs),
 * device FP32; gradient (n_params floats) copied to grad if not NULL. */
int exadg_guess_loss_and_grad(exadg_guess * g, const float * X, const float * T, int B, float * grad,
                              double * loss, void * stream);
/* Training losses of the learned guess, each normalized per step by the same quantity of the
 * extrapolation error (1 = as good as the extrapolation): L2 = sum of squared DoF errors
 * (per-cell minibatches), ENERGY = e^T A e and RESIDUAL = |A e|^2 (whole field per step, A the
 * FP64 operator of the sequence). */
enum exadg_guess_loss
{
  EXADG_GUESS_LOSS_L2       = 0,
  EXADG_GUESS_LOSS_ENERGY   = 1,
  EXADG_GUESS_LOSS_RESIDUAL = 2
};

/* Adam training on a stored FP64 sequence P[t][n_dofs] (solutions), Bv[t][n_dofs] (right-hand
 * sides), device, step-major, over steps [t_begin + history, t_end). op: the FP64 operator
 * (NULL allowed for L2). The first call on a network sets the scales from the same steps.
 * loss = mean normalized loss over the last epoch. */
int exadg_guess_train(exadg_guess * g, exadg_laplace * op, int loss_kind, const double * P,
                      const double * Bv, int64_t n_dofs, int t_begin, int t_end, int epochs,
                      int batch_cells, double lr, uint64_t seed, double * loss, void * stream);
int exadg_guess_evaluate(exadg_guess * g, exadg_laplace * op, int loss_kind, const double * P,
                         const double * Bv, int64_t n_dofs, int t_begin, int t_end, int batch_cells,
                         double * loss, void * stream);
/* x0 = extrapolation + network correction; *used_fallback = 1 if a non-finite network output
 * forced the plain extrapolation. history: host array of `history` device pointers (FP64). */
int exadg_guess_predict(exadg_guess * g, const double * const * history, const double * rhs,
                        double * x0, int64_t n_cells, int * used_fallback, void * stream);

/* ---------------------------------------------------------------- snapshots (restart files) */

/* What a snapshot holds besides its vectors: the parameters the discretization was written with
 * (DeserializationParameters, time_integration/restart.h:240-300: degree, mapping degree, mesh),
 * the time and step of the time integrator, and caller data (e.g. the learned guess's scales). */
typedef struct exadg_snapshot_info
{
  int32_t         version;   /* set by the writer */
  int32_t         precision; /* EXADG_FP64 / EXADG_FP32 of the vectors */
  int32_t         n_vectors;
  int64_t         n;         /* entries per vector */
  int32_t         degree;    /* 0 if the vectors belong to no discretization */
  exadg_mesh_desc mesh;
  double          time;
  int64_t         step;
  double          user[16];
} exadg_snapshot_info;

/* Write info->n_vectors vectors of info->n entries (device pointers if on_device, else host
 * pointers) to `path`. As the reference's restart files (restart.h:48-83), older snapshots are
 * kept by renaming path -> path.old.1 -> ... up to n_snapshots_keep files in all. The file is
 * written under a temporary name first and carries a checksum. Synchronizes the stream. */
int exadg_snapshot_write(const char * path, const exadg_snapshot_info * info,
                         const void * const * vectors, int on_device, int n_snapshots_keep,
                         void * stream);
/* Read a snapshot's header (read_deserialization_parameters, restart.h:270-300). */
int exadg_snapshot_read_info(const char * path, exadg_snapshot_info * info);
/* Read a snapshot into n_vectors vectors of n entries of `precision` (device pointers if
 * on_device). EXADG_ERR_ARGUMENT if the snapshot holds other counts, sizes or precision, or fails
 * its checksum; the vectors are written only after the whole file is verified. The header is
 * returned in info_out (may be NULL). Synchronizes the stream. */
int exadg_snapshot_read(const char * path, exadg_snapshot_info * info_out, void * const * vectors,
                        int n_vectors, int64_t n, int precision, int on_device, void * stream);

#ifdef __cplusplus
}
#endif

#endif /* EXADG_AMD_H */
