// Block-Jacobi kernels: the exact inverse of every cell's diagonal block A_KK of the SIPG
// operator (cell integral plus the self-coupling of the cell's faces, neighbours zero), stored as
// one dense n^3 x n^3 matrix per cell, and its application, optionally with the Chebyshev update
// of the smoother fused into the epilogue.
//
// Reference: BlockJacobiPreconditioner (solvers_and_preconditioners/preconditioners/
// block_jacobi_preconditioner.h), OperatorBase::add_block_diagonal_matrices /
// apply_inverse_block_diagonal_matrix_based / calculate_lu_factorization_block_jacobi
// (operators/operator_base.cpp:649-729, 789-815), AdditiveSchwarzPreconditioner on DG spaces
// (additive_schwarz_preconditioner.h; operator_base.cpp:2002-2009, 2099-2103: the block-Jacobi
// inverse), ChebyshevSmoother with PreconditionerSmoother::BlockJacobi (chebyshev_smoother.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ExaDG
{
// largest block (n^3) the dense path holds: degree 4 (125 x 125 doubles in LDS while inverting)
constexpr int kMaxDenseBlock = 125;

// In place: every m x m block of `blocks` (n_cells of them, symmetric positive definite) is
// replaced by its inverse (Gauss-Jordan without pivoting in LDS). m <= kMaxDenseBlock.
template<typename Number>
hipError_t launch_block_invert(Number * blocks, int m, int64_t n_cells, hipStream_t s);

// dst = B r per cell, where r = b - t (t may be NULL: r = b) and B the inverse cell block;
// variant 0: dst = f2 B r; 1: dst = (1+f1) x + f2 B r; 2: dst = (1+f1) x - f1 x_old + f2 B r.
// B is either dense (blocks: m x m per cell, symmetric) or, on affine cells, given by fast
// diagonalization: B = (V_z x V_y x V_x) diag(1 / (l_x + l_y + l_z)) (V_z x V_y x V_x)^T with
// the 1D generalized eigenvectors V (n x n, row = node) and eigenvalues l of the cell's three
// 1D configurations (fdm_cfg: 3 indices per cell into fdm_vec / fdm_val).
template<typename Number>
struct BlockApplyArgs
{
  Number const * blocks  = nullptr;
  Number const * fdm_vec = nullptr;
  Number const * fdm_val = nullptr;
  int const *    fdm_cfg = nullptr;
  int            n1d     = 0;
  Number const * b      = nullptr;
  Number const * t      = nullptr;
  Number const * x      = nullptr;
  Number const * x_old  = nullptr;
  Number *       dst    = nullptr;
  Number         f1 = 0, f2 = 1;
  int            variant = 0;
  int            m       = 0;
  int64_t        n_cells = 0;
};
template<typename Number>
hipError_t launch_block_apply(BlockApplyArgs<Number> const & a, hipStream_t s);

}  // namespace ExaDG
