#include "shape_info.h"

#include <cmath>
#include <stdexcept>

namespace ExaDG
{
namespace
{
// Legendre P_m(t) for m = N and N-1 together with P_N'(t)
void legendre_pair(int N, double t, double & pN, double & pNm1, double & dpN)
{
  double a = 1.0, b = t;
  if(N == 0)
  {
    pN = 1.0, pNm1 = 0.0, dpN = 0.0;
    return;
  }
  for(int m = 2; m <= N; ++m)
  {
    double const c = ((2.0 * m - 1.0) * t * b - (m - 1.0) * a) / m;
    a              = b;
    b              = c;
  }
  pN   = b;
  pNm1 = a;
  dpN  = N * (pNm1 - t * pN) / (1.0 - t * t);
}
}  // namespace

void gauss_points(int n, std::vector<double> & x, std::vector<double> & w)
{
  x.resize(n);
  w.resize(n);
  for(int i = 0; i < n; ++i)
  {
    // Newton from the Chebyshev-like initial guess, on [-1,1], descending
    double t = std::cos(M_PI * (4.0 * i + 3.0) / (4.0 * n + 2.0));
    double pN, pNm1, dpN;
    for(int it = 0; it < 100; ++it)
    {
      legendre_pair(n, t, pN, pNm1, dpN);
      double const step = pN / dpN;
      t -= step;
      if(std::fabs(step) <= 1e-16 * std::fabs(t) + 1e-300)
        break;
    }
    legendre_pair(n, t, pN, pNm1, dpN);
    x[n - 1 - i] = 0.5 + 0.5 * t;
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dpN * dpN);  // 2/((1-t^2)P'^2) halved for [0,1]
  }
}

void gauss_lobatto_points(int n, std::vector<double> & x)
{
  x.resize(n);
  if(n == 1)
  {
    x[0] = 0.5;
    return;
  }
  int const N = n - 1;
  x[0]        = 0.0;
  x[N]        = 1.0;
  for(int i = 1; i < N; ++i)
  {
    // interior points: roots of P_N'(t); Newton on P_N' with P_N'' from the Legendre ODE
    double t = -std::cos(M_PI * i / N);
    for(int it = 0; it < 200; ++it)
    {
      double pN, pNm1, dpN;
      legendre_pair(N, t, pN, pNm1, dpN);
      double const d2p  = (2.0 * t * dpN - N * (N + 1.0) * pN) / (1.0 - t * t);
      double const step = dpN / d2p;
      t -= step;
      if(std::fabs(step) <= 1e-16)
        break;
    }
    x[i] = 0.5 + 0.5 * t;
  }
}

void lagrange_basis(std::vector<double> const & nodes, double x, double * val, double * der)
{
  int const n = (int)nodes.size();
  for(int j = 0; j < n; ++j)
  {
    // product form for the value, logarithmic derivative summed term by term
    double v = 1.0, d = 0.0;
    for(int k = 0; k < n; ++k)
    {
      if(k == j)
        continue;
      double const f = 1.0 / (nodes[j] - nodes[k]);
      d              = d * (x - nodes[k]) * f + v * f;
      v *= (x - nodes[k]) * f;
    }
    val[j] = v;
    if(der)
      der[j] = d;
  }
}

ShapeInfo1D make_shape_info(int degree)
{
  if(degree < 1 || degree > 15)
    throw std::invalid_argument("degree out of range");
  ShapeInfo1D s;
  int const   n = degree + 1;
  s.n           = n;
  gauss_lobatto_points(n, s.nodes);
  gauss_points(n, s.qp, s.qw);
  s.S.assign(n * n, 0.0);
  s.Dq.assign(n * n, 0.0);
  std::vector<double> Sd(n * n);
  for(int q = 0; q < n; ++q)
  {
    lagrange_basis(s.nodes, s.qp[q], &s.S[q * n], &Sd[q * n]);
    std::vector<double> v(n);
    lagrange_basis(s.qp, s.qp[q], v.data(), &s.Dq[q * n]);
  }
  s.mass.assign(n * n, 0.0);
  s.stiff.assign(n * n, 0.0);
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      for(int q = 0; q < n; ++q)
      {
        s.mass[i * n + j] += s.qw[q] * s.S[q * n + i] * s.S[q * n + j];
        s.stiff[i * n + j] += s.qw[q] * Sd[q * n + i] * Sd[q * n + j];
      }
  s.d0.resize(n);
  s.d1.resize(n);
  std::vector<double> v(n);
  lagrange_basis(s.nodes, 0.0, v.data(), s.d0.data());
  lagrange_basis(s.nodes, 1.0, v.data(), s.d1.data());
  return s;
}

}  // namespace ExaDG
