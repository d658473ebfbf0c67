// CPU oracle for the SIPG Laplace hot path of ExaDG. TEST INFRASTRUCTURE ONLY.
//
// This is a plain FP64 C++/OpenMP restatement of the reference algorithm, used as the
// parity checker for the HIP product and as the timed CPU baseline. Only tests/, the
// smoke check and the CPU-baseline leg of bench.py may load it; the product never does.
//
// Reference files followed (paths relative to the ExaDG source tree):
//   operator / kernels : include/exadg/poisson/spatial_discretization/laplace_operator.{h,cpp}
//   weak BCs           : include/exadg/poisson/spatial_discretization/weak_boundary_conditions.h
//   penalty            : include/exadg/operators/interior_penalty_parameter.h:43-128
//   loops              : include/exadg/operators/operator_base.cpp:264-310, 1349-1491
//   RHS                : include/exadg/operators/rhs_operator.cpp:52-95
//   geometry           : include/exadg/grid/deformed_cube_manifold.h:47-157,
//                        include/exadg/grid/grid_utilities.h:52-58 (MappingQ)
//   error              : include/exadg/postprocessor/error_calculation.cpp:38-103
//   FE / quadrature    : include/exadg/operators/finite_element.h:56-61, quadrature.h:43-46
#pragma once
#include <cstdint>
#include <vector>

namespace oracle
{
// 1D data of FE_DGQ(k) on Gauss-Lobatto points of [0,1] with QGauss(k+1).
struct Tables1D
{
  int                 n = 0;       // k+1
  std::vector<double> gl;          // n GL support points
  std::vector<double> gp, gw;      // n Gauss points / weights on [0,1]
  std::vector<double> S;           // S[q*n+i]  = phi_i(gp_q)
  std::vector<double> Sd;          // Sd[q*n+i] = phi_i'(gp_q)
  std::vector<double> v0, v1;      // phi_i(0), phi_i(1)
  std::vector<double> d0, d1;      // phi_i'(0), phi_i'(1)
};

void gauss_legendre(int n, std::vector<double> & x, std::vector<double> & w);  // on [0,1]
void gauss_lobatto(int n, std::vector<double> & x);                            // on [0,1]
double lagrange(std::vector<double> const & nodes, int i, double x);
double lagrange_deriv(std::vector<double> const & nodes, int i, double x);
Tables1D make_tables(int degree);

// Structured box [lower,upper] subdivided into n_cells per direction, optionally deformed
// by the DeformedCubeManifold map and described by a MappingQ(mapping_degree) on GL support
// points. Boundary sides: 0:x-,1:x+,2:y-,3:y+,4:z-,5:z+; type 0 Dirichlet, 1 Neumann.
struct MeshDesc
{
  int    n_cells[3];
  double lower[3];
  double upper[3];
  double deformation;
  int    frequency;
  int    mapping_degree;
  int    periodic[3];
  int    boundary_type[6];
};

struct Mesh
{
  MeshDesc d;
  int64_t  n_cells;
  bool     affine;  // no deformation: J = diag(h) exactly
  double   h[3];

  explicit Mesh(MeshDesc const & desc);
  int64_t cell_index(int cx, int cy, int cz) const
  {
    return cx + (int64_t)d.n_cells[0] * (cy + (int64_t)d.n_cells[1] * cz);
  }
  void cell_coords(int64_t c, int & cx, int & cy, int & cz) const;
  // neighbour across face (dir, side); -1 on a non-periodic boundary
  int64_t neighbor(int64_t c, int dir, int side) const;
  // push_forward of DeformedCubeManifold (deformed_cube_manifold.h:47-59)
  void push_forward(double const * p, double * x) const;
  // physical position and Jacobian J[i*3+j] = dx_i/dxi_j at reference point xi of cell c
  void map_point(int64_t c, double const * xi, double * x, double * J) const;

  std::vector<double> map_gl;     // GL support points of the mapping on [0,1]
};

void invert3(double const * J, double * Jinv, double & det);

// Tensor evaluation: out[a + ma*(b + mb*c)] = sum A[a,i] B[b,j] C[c,l] in[i + n*(j + n*l)]
void tensor_apply(int ma, int mb, int mc, int n, double const * A, double const * B,
                  double const * C, double const * in, double * out);
// transpose: out[i,j,l] += sum_abc A[a,i] B[b,j] C[c,l] in[a,b,c]
void tensor_apply_T(int ma, int mb, int mc, int n, double const * A, double const * B,
                    double const * C, double const * in, double * out);

struct LaplaceOperator
{
  Mesh     mesh;
  int      degree, n, n3;
  double   ip_factor;
  Tables1D t;

  // cell geometry at the n^3 Gauss points: Jinv (9) and JxW (1) per point
  std::vector<double> cell_jinv, cell_jxw;
  std::vector<double> tau_cell;  // surface/volume per cell (interior_penalty_parameter.h:43-99)

  LaplaceOperator(MeshDesc const & desc, int degree, double ip_factor);
  int64_t n_dofs() const { return mesh.n_cells * n3; }

  double penalty_factor() const { return ip_factor * (degree + 1.0) * (degree + 1.0); }

  // geometry on face (dir, side) of cell c at the n^2 face Gauss points (a fastest):
  // Jinv (9), JxW (surface element * weights), outward unit normal (3), position (3)
  void face_geometry(int64_t c, int dir, int side, double * jinv, double * jxw, double * normal,
                     double * xq) const;
  void compute_face_geometry(int64_t c, int dir, int side, double * jinv, double * jxw,
                             double * normal, double * xq) const;
  std::vector<double> face_cache;  // curved meshes only: per cell face, n^2 x (Jinv,JxW,n,x)

  // dst = A src (homogeneous operator); the loop of operator_base.cpp:264-310
  void vmult(double * dst, double const * src) const;
  void cell_loop(double * dst, double const * src) const;
  void face_loop(double * dst, double const * src) const;
  void boundary_face_loop(double * dst, double const * src) const;

  // evaluate value and reference gradient of cell c on face (dir, side) at face Gauss points
  void face_evaluate(double const * u, int dir, int side, double * val, double * gref) const;
  void face_integrate(int dir, int side, double const * val, double const * gref,
                      double * out) const;

  // b for u* = prod sin(omega x_d): f = 3 omega^2 u*, Dirichlet g = u*, Neumann h = grad u*.n
  void rhs_sine(double omega, double * b) const;
  // weak DG divergence of u = sum_m A_m d_m cos(2 pi n_m.x + phi_m + omega_m t) (modes: 10
  // doubles each, n d A phi omega pad) with central flux on interior/periodic faces and zero flux
  // on walls (no slip): b_i = -(grad phi_i, u)_K + sum_F ({u}.n_K, phi_i)_F; mean removed when
  // no side is Dirichlet (divergence_operator.cpp:168-298, rhs of the pressure Poisson step)
  void rhs_divergence_fourier(double const * modes, int n_modes, double t, double * b) const;
  // relative L2 error vs prod sin(omega x_d) with QGauss(k+3)
  double error_sine(double omega, double const * u) const;
  // exact diagonal by unit-vector probing of the local cell/face matrices
  void diagonal(double * d) const;
  // the n^3 x n^3 diagonal block A_KK of every cell (cell term plus the self-coupling of its
  // faces, neighbours zero), column j of cell c at blocks[(c n^3 + j) n^3 ...]
  void block_diagonal(double * blocks) const;
  // both by one probing loop (either pointer may be null)
  void probe_blocks(double * dg, double * blocks) const;
};

// b = weak DG divergence of the discrete velocity u (FE_DGQ(degree_u)^3: per cell three
// component blocks of (degree_u+1)^3 GL nodal values), tested with the pressure space of op,
// QGauss(degree_u+1), central flux; boundary: pressure Neumann sides are walls at rest (no
// flux), pressure Dirichlet sides carry u-.n (oracle_divergence.cpp)
void divergence_apply(LaplaceOperator const & op, int degree_u, double const * u, double * b);

// SplitMix64 counter-based uniform[-1,1] of (seed, index)
double splitmix_uniform(uint64_t seed, uint64_t index);

}  // namespace oracle
