// Launch interface of the SIPG Laplace kernels (host-callable, no HIP types beyond streams).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "topology.h"

namespace ExaDG
{
enum class ApplyMode : int
{
  vmult     = 0,  // dst = A src
  vmult_add = 1,  // dst += A src
  diagonal  = 2,  // dst[cell, i] = (A_KK)_{ii}: unit-vector probing of the cell-local block
  // Chebyshev-Jacobi step fused into the epilogue (src = x, dst = x_new):
  // x_new = (1 + f1) x - f1 x_old (if has_old) + f2 dinv (b - A x); x_old may alias dst
  chebyshev = 3,
  // residual in the epilogue: dst = rhs - A src (dst may alias rhs, not src)
  residual = 4,
  // residual and p-restriction in the epilogue: per cell
  // dst_coarse += (R x R x R)(rhs - A src), R = P^T with P the n x nc prolongation matrix;
  // dst is the coarse DG vector of the same cells (nc^3 values per cell)
  residual_restrict = 5
};

// Centrosymmetric n x n matrices (A_ij = A_{n-1-i,n-1-j}, such as the 1D mass and stiffness
// on symmetric nodes) are stored for n >= 5 in even-odd form, which needs ~n^2/2 products per
// application instead of n^2: with m = n/2 and m_e = m + n%2,
//   Ae (m x m_e): (A_ij + A_{i,n-1-j})/2 for j < m, A_im for j = m (odd n),
//   Ao (m x m):   (A_ij - A_{i,n-1-j})/2,
//   Am (m_e, odd n only): the middle row A_mj, j <= m.
// The Cartesian kernel packs for n >= 5; the general kernel for n >= 6 (at n = 5 its saving
// is 4% of the instructions and the FP32 instantiation measured slower with it).
constexpr int  eo_size(int n, bool packed)
{
  return packed ? (n / 2) * (n / 2 + n % 2) + (n / 2) * (n / 2) + (n % 2 ? n / 2 + 1 : 0) : n * n;
}
constexpr bool cs_packed(int n) { return n >= 5; }
constexpr int  cs_size(int n) { return eo_size(n, cs_packed(n)); }
constexpr bool gen_packed(int n) { return n >= 6; }
constexpr int  gen_size(int n) { return eo_size(n, gen_packed(n)); }
// Offsets in the Cartesian kernel's table: [M | K | d0 n | d1 n], M and K of cs_size(n) each
template<int N>
struct CartTable
{
  static constexpr int M = 0, K = cs_size(N), D0 = 2 * cs_size(N), D1 = D0 + N;
};

// Table layout (Number, SipgArgs::tab): general kernel: [S n*n | Dq n*n | d0 n | d1 n], and in
// device memory (`tables`) followed by [l(0) n | l(1) n | l'(0) n | l'(1) n] of the Lagrange
// polynomials on the Gauss points and by the four sweep matrices S, Dq, S^T, Dq^T in even-odd
// form (GenTable). S is centrosymmetric and Dq skew-centrosymmetric (A_ij = -A_{n-1-i,n-1-j})
// on the symmetric Gauss points; the skew form stores Ae and Ao as above and, for odd n, the
// middle row's coefficients of the odd part, A_mj for j < m, padded to m_e entries.
// Cartesian kernel: CartTable, on the unit interval, plus h per direction.
template<int N>
struct GenTable
{
  static constexpr int S = 0, DQ = N * N, D0 = 2 * N * N, D1 = D0 + N, L0 = D1 + N, L1 = L0 + N,
                       G0 = L1 + N, G1 = G0 + N, SE = G1 + N, DE = SE + gen_size(N),
                       STE = DE + gen_size(N), DTE = STE + gen_size(N), SIZE = DTE + gen_size(N);
};
template<typename Number>
struct SipgArgs
{
  // 1D tables. The Cartesian kernel reads them from this by-value copy, i.e. as uniform scalar
  // kernel-argument loads (unit-interval mass M, stiffness K, end derivatives d0, d1, scaled
  // per direction inside the kernel); the general kernel (S, Dq, d0, d1) indexes them at run
  // time and reads the device copy `tables`.
  static constexpr int kMaxTable = 2 * 9 * 9 + 2 * 9;
  Number         tab[kMaxTable];
  Number         h[6];  // Cartesian kernel: h_x, h_y, h_z, 1/h_x, 1/h_y, 1/h_z
  Number const * tables;  // general kernel: the same tables in device memory
  Number const * src;
  Number *       dst;
  Number const * cell_metric;  // general kernel only
  Number const * face_data;    // general kernel only
  Number const * tau_face;
  BoxTopology    topo;
  int64_t        n_cells;
  int            probe;      // diagonal mode: local index of the unit vector
  // diagonal mode: 0 stores (A_KK)_{probe,probe} at dst[cell n^3 + probe]; 1 stores the whole
  // response A_KK e_probe as column `probe` of the cell's n^3 x n^3 block,
  // dst[(cell n^3 + probe) n^3 + e] (the block is symmetric, so this is also row `probe`)
  int            block = 0;
  int            tile_rows;  // traversal: y-rows per tile (one tile per XCD)
  // chebyshev and residual modes
  Number const * x_old   = nullptr;
  Number const * rhs     = nullptr;
  Number const * dinv    = nullptr;
  Number         f1      = 0;
  Number         f2      = 0;
  int            has_old = 0;
  // residual_restrict mode: P (n x nc, row-major) and nc
  Number const * restrict_P  = nullptr;
  int            restrict_nc = 0;
};

// degree 1..8; mode as above. Returns hipError of the launch.
template<typename Number>
hipError_t launch_sipg_general(int degree, ApplyMode mode, SipgArgs<Number> const & args,
                               hipStream_t stream);
// Diagonal of the affine SIPG operator in closed form, one thread per DoF: the cell block is
// L_x (x) h_y M (x) h_z M + h_x M (x) L_y (x) h_z M + h_x M (x) h_y M (x) L_z with the 1D
// cell-local line operators L_d (the cell's own face terms, neighbours zero, or the cell's own
// values across a periodic face of a one-cell direction), so its diagonal is a sum of three
// products of 1D diagonals. diag1d: [M_ii | K_ii | d0 | d1] of the unit interval (4n doubles).
template<typename Number>
hipError_t launch_cartesian_diagonal(Number * diag, double const * diag1d, Number const * tau_face,
                                     BoxTopology const & topo, double const (&h)[3], int n,
                                     int64_t n_cells, hipStream_t stream);

// Diagonal of the curved SIPG operator, sum-factorized, one group of n^2 lanes per cell:
// the cell term sum_q G_ab(q) d_a phi_i d_b phi_i as six three-sweep contractions of the metric
// with products of the 1D tables, plus the own-side face terms of the face-layer DoFs,
// c w (tau phi_i^2 - phi_i a.grad phi_i) with c = 1 (interior, neighbour zero), 2 (Dirichlet
// mirror), 0 (Neumann). tab1d: [SS | DD | SD] (n x n each, [q][i]: S^2, D^2, S D with S the
// Gauss interpolation and D = Dq S) then [l'(0) of node 0, l'(1) of node n-1]. Meshes with a
// one-cell periodic direction are not handled (the cell is its own neighbour).
template<typename Number>
hipError_t launch_general_diagonal(Number * diag, double const * tab1d, SipgArgs<Number> const & a,
                                   int degree, hipStream_t stream);

template<typename Number>
hipError_t launch_sipg_cartesian(int degree, ApplyMode mode, SipgArgs<Number> const & args,
                                 hipStream_t stream);

}  // namespace ExaDG
