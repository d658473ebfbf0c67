// 1D shape data of FE_DGQ(k) on Gauss-Lobatto points with QGauss(k+1) -- the tables the
// reference obtains from deal.II's ShapeInfo for FE_DGQ (operators/finite_element.h:56-61)
// and QGauss (operators/quadrature.h:43-46). Host-side, FP64.
#pragma once
#include <vector>

namespace ExaDG
{
struct ShapeInfo1D
{
  int                 n = 0;  // k+1 nodes = k+1 Gauss points
  std::vector<double> nodes;  // Gauss-Lobatto support points on [0,1]
  std::vector<double> qp, qw; // Gauss points and weights on [0,1]
  // row-major n x n tables (row = quadrature point or output index)
  std::vector<double> S;      // S[q*n+i] = phi_i(qp_q): nodal -> Gauss values
  std::vector<double> Dq;     // Dq[q*n+p] = l_p'(qp_q): collocation derivative on Gauss points
  std::vector<double> mass;   // M[i*n+j] = int phi_i phi_j (exact with QGauss(k+1))
  std::vector<double> stiff;  // K[i*n+j] = int phi_i' phi_j'
  std::vector<double> d0, d1; // phi_i'(0), phi_i'(1)
};

void gauss_points(int n, std::vector<double> & x, std::vector<double> & w);
void gauss_lobatto_points(int n, std::vector<double> & x);
// barycentric Lagrange evaluation of all n basis functions (and derivatives) at x
void lagrange_basis(std::vector<double> const & nodes, double x, double * val, double * der);
ShapeInfo1D make_shape_info(int degree);

}  // namespace ExaDG
