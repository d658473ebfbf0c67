// Right-hand side of the Poisson problem for the manufactured solution u = prod sin(omega x_d):
// RHSOperator::evaluate_add (operators/rhs_operator.cpp:52-95) for f = 3 omega^2 u, plus the
// inhomogeneous boundary face integrals moved to the right-hand side with factor -1
// (OperatorBase::rhs_add, operator_base.cpp:507-554, boundary_face_loop_inhom_operator
// :1436-1491, mirror values of weak_boundary_conditions.h:39-47,142-160):
//   Dirichlet g: b += int (-g dv/dn + 2 tau g v),  Neumann h: b += int h v.
// Host-side FP64, once per solve.
#include <cmath>

#include "laplace_operator.h"

namespace ExaDG
{
namespace Poisson
{
namespace
{
// out[i + n (j + n k)] += sum_abc A[a][i] B[b][j] C[c][k] in[a + ma (b + mb c)]
void integrate3(int ma, int mb, int mc, int n, double const * A, double const * B, double const * C,
                double const * in, double * out)
{
  std::vector<double> w2((size_t)ma * mb * n), w1((size_t)ma * n * n);
  for(int k = 0; k < n; ++k)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int c = 0; c < mc; ++c)
          s += C[c * n + k] * in[a + ma * (b + mb * c)];
        w2[a + ma * (b + mb * k)] = s;
      }
  for(int k = 0; k < n; ++k)
    for(int j = 0; j < n; ++j)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int b = 0; b < mb; ++b)
          s += B[b * n + j] * w2[a + ma * (b + mb * k)];
        w1[a + ma * (j + n * k)] = s;
      }
  for(int k = 0; k < n; ++k)
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < n; ++i)
      {
        double s = 0;
        for(int a = 0; a < ma; ++a)
          s += A[a * n + i] * w1[a + ma * (j + n * k)];
        out[i + n * (j + n * k)] += s;
      }
}

void inv3(double const * J, double * Ji, double & det)
{
  det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
        J[2] * (J[3] * J[7] - J[4] * J[6]);
  double const r = 1.0 / det;
  Ji[0]          = (J[4] * J[8] - J[5] * J[7]) * r;
  Ji[1]          = (J[2] * J[7] - J[1] * J[8]) * r;
  Ji[2]          = (J[1] * J[5] - J[2] * J[4]) * r;
  Ji[3]          = (J[5] * J[6] - J[3] * J[8]) * r;
  Ji[4]          = (J[0] * J[8] - J[2] * J[6]) * r;
  Ji[5]          = (J[2] * J[3] - J[0] * J[5]) * r;
  Ji[6]          = (J[3] * J[7] - J[4] * J[6]) * r;
  Ji[7]          = (J[1] * J[6] - J[0] * J[7]) * r;
  Ji[8]          = (J[0] * J[4] - J[1] * J[3]) * r;
}
}  // namespace

void assemble_rhs_sine(StructuredGrid const & grid, ShapeInfo1D const & shape,
                       GeometryData const & geo, double omega, std::vector<double> & b)
{
  int const           n = shape.n, n2 = n * n, n3 = n * n * n;
  int64_t const       nc = grid.n_cells();
  BoxTopology const & topo = grid.topology();
  b.assign((size_t)nc * n3, 0.0);
  // derivative of the nodal basis at the Gauss points, Sd = Dq S
  std::vector<double> Sd(n2, 0.0);
  for(int q = 0; q < n; ++q)
    for(int i = 0; i < n; ++i)
      for(int p = 0; p < n; ++p)
        Sd[q * n + i] += shape.Dq[q * n + p] * shape.S[p * n + i];
  std::vector<double> v0(n, 0.0), v1(n, 0.0);
  v0[0]     = 1.0;
  v1[n - 1] = 1.0;
  std::vector<double> const one0 = {0.0}, one1 = {1.0};

  parallel_range(nc, [&](int64_t cb, int64_t ce) {
    std::vector<double> x(3 * n3), J(9 * n3), fq(n3), xf(3 * n2), Jf(9 * n2), rv(n2), rg(3 * n2);
    for(int64_t c = cb; c < ce; ++c)
    {
      double * out = &b[(size_t)c * n3];
      grid.evaluate(c, shape.qp, shape.qp, shape.qp, x.data(), J.data());
      for(int qz = 0; qz < n; ++qz)
        for(int qy = 0; qy < n; ++qy)
          for(int qx = 0; qx < n; ++qx)
          {
            int const q = qx + n * (qy + n * qz);
            double    Jq[9], Ji[9], det;
            for(int k = 0; k < 9; ++k)
              Jq[k] = J[k * n3 + q];
            inv3(Jq, Ji, det);
            double const u = std::sin(omega * x[q]) * std::sin(omega * x[n3 + q]) *
                             std::sin(omega * x[2 * n3 + q]);
            fq[q] = 3.0 * omega * omega * u * det * shape.qw[qx] * shape.qw[qy] * shape.qw[qz];
          }
      integrate3(n, n, n, n, shape.S.data(), shape.S.data(), shape.S.data(), fq.data(), out);

      int cc[3];
      cell_coords(topo, c, cc);
      for(int d = 0; d < 3; ++d)
        for(int s = 0; s < 2; ++s)
        {
          if(neighbor_id(topo, cc, d, s) >= 0)
            continue;
          int const                   type = topo.bc[2 * d + s];
          std::vector<double> const * p[3] = {&shape.qp, &shape.qp, &shape.qp};
          p[d]                             = s ? &one1 : &one0;
          grid.evaluate(c, *p[0], *p[1], *p[2], xf.data(), Jf.data());
          double const tau = geo.tau_face[face_id(topo, cc, d, s)];
          for(int q = 0; q < n2; ++q)
          {
            double Jq[9], Ji[9], det;
            for(int k = 0; k < 9; ++k)
              Jq[k] = Jf[k * n2 + q];
            inv3(Jq, Ji, det);
            double nv[3] = {Ji[d * 3 + 0], Ji[d * 3 + 1], Ji[d * 3 + 2]};
            double const nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            double const sg  = s ? 1.0 : -1.0;
            for(int e = 0; e < 3; ++e)
              nv[e] *= sg / nrm;
            double const w   = det * nrm * shape.qw[q % n] * shape.qw[q / n];
            double const px = xf[q], py = xf[n2 + q], pz = xf[2 * n2 + q];
            double       aK[3];
            for(int j = 0; j < 3; ++j)
              aK[j] = Ji[j * 3 + 0] * nv[0] + Ji[j * 3 + 1] * nv[1] + Ji[j * 3 + 2] * nv[2];
            double const s0 = std::sin(omega * px), s1 = std::sin(omega * py), s2 = std::sin(omega * pz);
            if(type == EXADG_BC_DIRICHLET)
            {
              double const g = s0 * s1 * s2;
              rv[q]          = 2.0 * tau * g * w;
              for(int j = 0; j < 3; ++j)
                rg[j * n2 + q] = -g * w * aK[j];
            }
            else
            {
              double const gr[3] = {omega * std::cos(omega * px) * s1 * s2,
                                    omega * s0 * std::cos(omega * py) * s2,
                                    omega * s0 * s1 * std::cos(omega * pz)};
              rv[q] = (gr[0] * nv[0] + gr[1] * nv[1] + gr[2] * nv[2]) * w;
              for(int j = 0; j < 3; ++j)
                rg[j * n2 + q] = 0.0;
            }
          }
          // value test: along d the end value row, tangentially S; gradient test component j:
          // along d the end derivative row (j == d) or value row, tangentially Sd for j
          int m[3] = {n, n, n};
          m[d]     = 1;
          double const * vrow = s ? v1.data() : v0.data();
          double const * drow = s ? shape.d1.data() : shape.d0.data();
          double const * ops[3];
          for(int k = 0; k < 3; ++k)
            ops[k] = (k == d) ? vrow : shape.S.data();
          integrate3(m[0], m[1], m[2], n, ops[0], ops[1], ops[2], rv.data(), out);
          for(int j = 0; j < 3; ++j)
          {
            for(int k = 0; k < 3; ++k)
              ops[k] = (k == d) ? (j == d ? drow : vrow) : (j == k ? Sd.data() : shape.S.data());
            integrate3(m[0], m[1], m[2], n, ops[0], ops[1], ops[2], rg.data() + j * n2, out);
          }
        }
    }
  });
}

}  // namespace Poisson
}  // namespace ExaDG
