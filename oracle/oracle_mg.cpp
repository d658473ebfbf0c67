// CPU oracle multigrid (TEST INFRASTRUCTURE ONLY). See oracle_mg.h for the reference files.
#include "oracle_mg.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>

namespace oracle
{
// ---------------------------------------------------------------- eigenvalues

void tridiagonal_extreme_eigenvalues(std::vector<double> const & alpha,
                                     std::vector<double> const & beta, double & lmin, double & lmax)
{
  // Lanczos matrix of CG, eigenvalues by the implicit QL algorithm with Wilkinson shifts
  int const n = (int)alpha.size();
  if(n == 0)
  {
    lmin = lmax = 1.0;
    return;
  }
  std::vector<double> d(n), e(n, 0.0);
  for(int k = 0; k < n; ++k)
  {
    d[k] = 1.0 / alpha[k] + (k > 0 ? beta[k - 1] / alpha[k - 1] : 0.0);
    if(k + 1 < n)
      e[k] = std::sqrt(beta[k]) / alpha[k];
  }
  for(int l = 0; l < n; ++l)
  {
    int iter = 0, m;
    do
    {
      for(m = l; m < n - 1; ++m)
      {
        double const dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if(std::fabs(e[m]) <= 1e-15 * dd)
          break;
      }
      if(m != l)
      {
        if(++iter > 200)
          throw std::runtime_error("QL did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g        = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int    i;
        for(i = m - 1; i >= l; --i)
        {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if(r == 0.0)
          {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g        = c * r - b;
        }
        if(r == 0.0 && i >= l)
          continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while(m != l);
  }
  lmin = *std::min_element(d.begin(), d.end());
  lmax = *std::max_element(d.begin(), d.end());
}

static double dotv(int64_t n, double const * a, double const * b)
{
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for(int64_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

static void subtract_mean(std::vector<double> & v)
{
  double s = 0;
  for(double x : v)
    s += x;
  s /= (double)v.size();
  for(double & x : v)
    x -= s;
}

int pcg_record(int64_t n, Apply const & A, Apply const & P, double const * b, double * x,
               double tol_abs, int max_iter, bool singular, std::vector<double> * alphas,
               std::vector<double> * betas)
{
  // zero initial guess
  std::vector<double> r(b, b + n), z(n), p(n), v(n);
  std::fill(x, x + n, 0.0);
  if(singular)
    subtract_mean(r);
  P(z.data(), r.data());
  double rho = dotv(n, r.data(), z.data());
  if(std::sqrt(dotv(n, r.data(), r.data())) <= tol_abs)
    return 0;
  p      = z;
  int it = 0;
  for(it = 1; it <= max_iter; ++it)
  {
    A(v.data(), p.data());
    double const alpha = rho / dotv(n, p.data(), v.data());
    for(int64_t i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * v[i];
    }
    if(singular)
      subtract_mean(r);
    P(z.data(), r.data());
    double const rho_new = dotv(n, r.data(), z.data());
    double const rn      = std::sqrt(dotv(n, r.data(), r.data()));
    double const beta    = rho_new / rho;
    if(alphas)
      alphas->push_back(alpha);
    if(betas)
      betas->push_back(beta);
    if(rn <= tol_abs)
      break;
    rho = rho_new;
    for(int64_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  return std::min(it, max_iter);
}

// ---------------------------------------------------------------- Q1 operator

Q1Operator::Q1Operator(MeshDesc const & d) : mesh(d)
{
  for(int k = 0; k < 3; ++k)
    nn[k] = d.n_cells[k] + (d.periodic[k] ? 0 : 1);
  n_nodes = (int64_t)nn[0] * nn[1] * nn[2];
  constrained.assign(n_nodes, 0);
  for(int64_t i = 0; i < n_nodes; ++i)
  {
    int const c[3] = {(int)(i % nn[0]), (int)((i / nn[0]) % nn[1]), (int)(i / ((int64_t)nn[0] * nn[1]))};
    for(int k = 0; k < 3; ++k)
      if(!d.periodic[k] && ((c[k] == 0 && d.boundary_type[2 * k] == 0) ||
                            (c[k] == d.n_cells[k] && d.boundary_type[2 * k + 1] == 0)))
        constrained[i] = 1;
  }
  std::vector<double> gp, gw;
  gauss_legendre(2, gp, gw);
  K.assign((size_t)mesh.n_cells * 64, 0.0);
#pragma omp parallel for schedule(static)
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    double * Kc = &K[(size_t)c * 64];
    for(int q = 0; q < 8; ++q)
    {
      double const xi[3] = {gp[q & 1], gp[(q >> 1) & 1], gp[(q >> 2) & 1]};
      double const w     = gw[q & 1] * gw[(q >> 1) & 1] * gw[(q >> 2) & 1];
      double       x[3], J[9], Ji[9], det;
      mesh.map_point(c, xi, x, J);
      invert3(J, Ji, det);
      double g[8][3];
      for(int a = 0; a < 8; ++a)
      {
        double gr[3];
        for(int e = 0; e < 3; ++e)
        {
          double v = 1;
          for(int k = 0; k < 3; ++k)
          {
            int const bit = (a >> k) & 1;
            v *= (k == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[k] : 1.0 - xi[k]);
          }
          gr[e] = v;
        }
        for(int i = 0; i < 3; ++i)
          g[a][i] = Ji[0 * 3 + i] * gr[0] + Ji[1 * 3 + i] * gr[1] + Ji[2 * 3 + i] * gr[2];
      }
      for(int a = 0; a < 8; ++a)
        for(int b = 0; b < 8; ++b)
          Kc[a * 8 + b] += det * w * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
    }
  }
}

int64_t Q1Operator::node_of(int const * cc, int v) const
{
  int nd[3];
  for(int k = 0; k < 3; ++k)
  {
    int c = cc[k] + ((v >> k) & 1);
    if(mesh.d.periodic[k] && c == mesh.d.n_cells[k])
      c = 0;
    nd[k] = c;
  }
  return nd[0] + (int64_t)nn[0] * (nd[1] + (int64_t)nn[1] * nd[2]);
}

void Q1Operator::vmult(double * dst, double const * src) const
{
  // cell-wise assembly of the action (serial scatter), constrained rows identity
  std::fill(dst, dst + n_nodes, 0.0);
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    int cc[3];
    mesh.cell_coords(c, cc[0], cc[1], cc[2]);
    int64_t nodes[8];
    double  u[8];
    for(int v = 0; v < 8; ++v)
    {
      nodes[v] = node_of(cc, v);
      u[v]     = constrained[nodes[v]] ? 0.0 : src[nodes[v]];
    }
    double const * Kc = &K[(size_t)c * 64];
    for(int a = 0; a < 8; ++a)
    {
      double s = 0;
      for(int b = 0; b < 8; ++b)
        s += Kc[a * 8 + b] * u[b];
      dst[nodes[a]] += s;
    }
  }
  for(int64_t i = 0; i < n_nodes; ++i)
    if(constrained[i])
      dst[i] = src[i];
}

void Q1Operator::diagonal(double * dg) const
{
  std::fill(dg, dg + n_nodes, 0.0);
  for(int64_t c = 0; c < mesh.n_cells; ++c)
  {
    int cc[3];
    mesh.cell_coords(c, cc[0], cc[1], cc[2]);
    for(int a = 0; a < 8; ++a)
      dg[node_of(cc, a)] += K[(size_t)c * 64 + a * 9];
  }
  for(int64_t i = 0; i < n_nodes; ++i)
    if(constrained[i])
      dg[i] = 1.0;
}

// ---------------------------------------------------------------- Chebyshev

static Apply jacobi(std::vector<double> const & inv_diag)
{
  return [&inv_diag](double * dst, double const * src) {
    int64_t const n = (int64_t)inv_diag.size();
    for(int64_t i = 0; i < n; ++i)
      dst[i] = inv_diag[i] * src[i];
  };
}

void Chebyshev::setup_smoother(Apply const & A_, std::vector<double> dinv,
                               std::vector<char> const & mask, int degree_, double range,
                               int eig_its)
{
  A        = A_;
  inv_diag = std::move(dinv);
  degree   = degree_;
  int64_t const n = (int64_t)inv_diag.size();
  // deal.II initial vector: (i mod 11) minus mean, constrained entries zero
  std::vector<double> b(n), x(n);
  for(int64_t i = 0; i < n; ++i)
    b[i] = (double)(i % 11);
  subtract_mean(b);
  for(int64_t i = 0; i < n; ++i)
    if(!mask.empty() && mask[i])
      b[i] = 0;
  std::vector<double> al, be;
  double const        nb = std::sqrt(dotv(n, b.data(), b.data()));
  cg_iterations = pcg_record(n, A, jacobi(inv_diag), b.data(), x.data(), 1e-2 * nb, eig_its, false,
                             &al, &be);
  tridiagonal_extreme_eigenvalues(al, be, lambda_min, lambda_max);
  double const max_ev = 1.2 * lambda_max;
  double const lo     = max_ev / range;
  theta               = 0.5 * (max_ev + lo);
  delta               = 0.5 * (max_ev - lo);
}

void Chebyshev::setup_coarse(Apply const & A_, std::vector<double> dinv, double rel_tol,
                             bool singular)
{
  A                  = A_;
  inv_diag           = std::move(dinv);
  int64_t const n    = (int64_t)inv_diag.size();
  std::vector<double> b(n), x(n);
  srand(1);
  for(int64_t i = 0; i < n; ++i)
    b[i] = (double)rand() / RAND_MAX;
  if(singular)
    subtract_mean(b);
  std::vector<double> al, be;
  double const        nb = std::sqrt(dotv(n, b.data(), b.data()));
  cg_iterations = pcg_record(n, A, jacobi(inv_diag), b.data(), x.data(), 1e-5 * nb, 10000, singular,
                             &al, &be);
  tridiagonal_extreme_eigenvalues(al, be, lambda_min, lambda_max);
  double const factor = 1.1;
  double const max_ev = factor * lambda_max;
  double const range  = lambda_max / lambda_min * factor;
  double const sigma  = (1. - std::sqrt(1. / range)) / (1. + std::sqrt(1. / range));
  degree = std::max(1, (int)(std::log(1. / rel_tol + std::sqrt(1. / rel_tol / rel_tol - 1.)) /
                             std::log(1. / sigma)));
  double const lo = max_ev / range;
  theta           = 0.5 * (max_ev + lo);
  delta           = 0.5 * (max_ev - lo);
}

void Chebyshev::vmult(std::vector<double> & x, std::vector<double> const & b) const
{
  // deal.II PreconditionChebyshev::vmult: zero initial guess
  int64_t const       n = (int64_t)b.size();
  std::vector<double> xo(n, 0.0), t(n);
  for(int64_t i = 0; i < n; ++i)
    x[i] = inv_diag[i] * b[i] / theta;
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int it = 1; it < degree; ++it)
  {
    A(t.data(), x.data());
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    for(int64_t i = 0; i < n; ++i)
    {
      double const xn = (1. + f1) * x[i] - (it > 1 ? f1 * xo[i] : 0.0) + f2 * inv_diag[i] * (b[i] - t[i]);
      xo[i]           = x[i];
      x[i]            = xn;
    }
  }
}

void Chebyshev::step(std::vector<double> & x, std::vector<double> const & b) const
{
  int64_t const       n = (int64_t)b.size();
  std::vector<double> xo(n), t(n);
  A(t.data(), x.data());
  for(int64_t i = 0; i < n; ++i)
  {
    xo[i] = x[i];
    x[i]  = x[i] + inv_diag[i] * (b[i] - t[i]) / theta;
  }
  if(degree < 2 || std::fabs(delta) < 1e-40)
    return;
  double rhok = delta / theta, sigma = theta / delta;
  for(int it = 1; it < degree; ++it)
  {
    A(t.data(), x.data());
    double const rhokp = 1. / (2. * sigma - rhok);
    double const f1 = rhokp * rhok, f2 = 2. * rhokp / delta;
    rhok            = rhokp;
    for(int64_t i = 0; i < n; ++i)
    {
      double const xn = (1. + f1) * x[i] - f1 * xo[i] + f2 * inv_diag[i] * (b[i] - t[i]);
      xo[i]           = x[i];
      x[i]            = xn;
    }
  }
}

// ---------------------------------------------------------------- multigrid

Multigrid::Multigrid(MeshDesc const & m, MGDesc const & d, double ip) : mesh(m), desc(d)
{
  int const nl = (int)d.degrees.size() + (d.coarse_q1 ? 1 : 0);
  levels.resize(nl);
  int l = 0;
  if(d.coarse_q1)
  {
    levels[0].q1_op  = std::make_unique<Q1Operator>(m);
    levels[0].dg     = false;
    levels[0].degree = 1;
    levels[0].n      = levels[0].q1_op->n_nodes;
    Q1Operator * q   = levels[0].q1_op.get();
    levels[0].A      = [q](double * dst, double const * src) { q->vmult(dst, src); };
    levels[0].inv_diag.resize(levels[0].n);
    q->diagonal(levels[0].inv_diag.data());
    invert_diagonal(levels[0].inv_diag);
    ++l;
  }
  for(int i = (int)d.degrees.size() - 1; i >= 0; --i, ++l)
  {
    Level & L = levels[l];
    L.degree  = d.degrees[i];
    L.dg_op   = std::make_unique<LaplaceOperator>(m, L.degree, ip);
    L.n       = L.dg_op->n_dofs();
    LaplaceOperator * op = L.dg_op.get();
    L.A                  = [op](double * dst, double const * src) { op->vmult(dst, src); };
    L.inv_diag.resize(L.n);
    op->diagonal(L.inv_diag.data());
    invert_diagonal(L.inv_diag);
  }
  for(int k = 0; k < nl; ++k)
  {
    Level & L = levels[k];
    L.defect.assign(L.n, 0.0);
    L.sol.assign(L.n, 0.0);
    L.t.assign(L.n, 0.0);
    if(k > 0)
    {
      std::vector<char> mask;
      L.smoother.setup_smoother(L.A, L.inv_diag, mask, d.smoother_degree, d.smoothing_range,
                                d.eig_its);
      if(levels[k - 1].dg)
      {
        int const           nf = L.degree + 1, nc = levels[k - 1].degree + 1;
        std::vector<double> xf, xc;
        gauss_lobatto(nf, xf);
        gauss_lobatto(nc, xc);
        L.P.resize(nf * nc);
        for(int a = 0; a < nf; ++a)
          for(int i = 0; i < nc; ++i)
            L.P[a * nc + i] = lagrange(xc, i, xf[a]);
      }
    }
  }
  if(d.coarse_solver == 1)
    coarse_cheb.setup_coarse(levels[0].A, levels[0].inv_diag, d.coarse_rel_tol, d.singular);
}

void Multigrid::prolongate_add(int l, double * fine, double const * coarse) const
{
  Level const & F = levels[l];
  Level const & C = levels[l - 1];
  if(C.dg)
  {
    int const nf = F.degree + 1, nc = C.degree + 1;
    int const nf3 = nf * nf * nf, nc3 = nc * nc * nc;
#pragma omp parallel for schedule(static)
    for(int64_t c = 0; c < mesh.n_cells[0] * (int64_t)mesh.n_cells[1] * mesh.n_cells[2]; ++c)
    {
      std::vector<double> out(nf3);
      tensor_apply(nf, nf, nf, nc, F.P.data(), F.P.data(), F.P.data(), coarse + c * nc3, out.data());
      for(int i = 0; i < nf3; ++i)
        fine[c * nf3 + i] += out[i];
    }
  }
  else
  {
    Q1Operator const & q  = *C.q1_op;
    int64_t const      nc = q.mesh.n_cells;
    for(int64_t c = 0; c < nc; ++c)
    {
      int cc[3];
      q.mesh.cell_coords(c, cc[0], cc[1], cc[2]);
      for(int v = 0; v < 8; ++v)
      {
        int64_t const node = q.node_of(cc, v);
        if(!q.constrained[node])
          fine[c * 8 + v] += coarse[node];
      }
    }
  }
}

void Multigrid::restrict_add(int l, double * coarse, double const * fine) const
{
  Level const & F = levels[l];
  Level const & C = levels[l - 1];
  if(C.dg)
  {
    int const nf = F.degree + 1, nc = C.degree + 1;
    int const nf3 = nf * nf * nf, nc3 = nc * nc * nc;
#pragma omp parallel for schedule(static)
    for(int64_t c = 0; c < mesh.n_cells[0] * (int64_t)mesh.n_cells[1] * mesh.n_cells[2]; ++c)
      tensor_apply_T(nf, nf, nf, nc, F.P.data(), F.P.data(), F.P.data(), fine + c * nf3,
                     coarse + c * nc3);
  }
  else
  {
    Q1Operator const & q  = *C.q1_op;
    int64_t const      nc = q.mesh.n_cells;
    for(int64_t c = 0; c < nc; ++c)
    {
      int cc[3];
      q.mesh.cell_coords(c, cc[0], cc[1], cc[2]);
      for(int v = 0; v < 8; ++v)
      {
        int64_t const node = q.node_of(cc, v);
        if(!q.constrained[node])
          coarse[node] += fine[c * 8 + v];
      }
    }
  }
}

void Multigrid::v_cycle(int l) const
{
  Level const & L = levels[l];
  if(l == 0)
  {
    if(desc.coarse_solver == 1)
    {
      coarse_cheb.vmult(L.sol, L.defect);
      last_coarse_its = coarse_cheb.degree;
      return;
    }
    std::vector<double> r = L.defect;
    if(desc.singular)
      subtract_mean(r);
    double const r0  = std::sqrt(dotv(L.n, r.data(), r.data()));
    double const tol = std::max(1e-12, desc.coarse_rel_tol * r0);
    last_coarse_its  = pcg_record(L.n, L.A, jacobi(L.inv_diag), r.data(), L.sol.data(), tol,
                                  desc.coarse_max_iter, desc.singular, nullptr, nullptr);
    return;
  }
  L.smoother.vmult(L.sol, L.defect);
  L.A(L.t.data(), L.sol.data());
  for(int64_t i = 0; i < L.n; ++i)
    L.t[i] = L.defect[i] - L.t[i];
  restrict_add(l, levels[l - 1].defect.data(), L.t.data());
  v_cycle(l - 1);
  prolongate_add(l, L.sol.data(), levels[l - 1].sol.data());
  L.smoother.step(L.sol, L.defect);
}

void Multigrid::vmult(double * dst, double const * src) const
{
  int const top = (int)levels.size() - 1;
  for(int l = 0; l < top; ++l)
    std::fill(levels[l].defect.begin(), levels[l].defect.end(), 0.0);
  std::copy(src, src + levels[top].n, levels[top].defect.begin());
  v_cycle(top);
  std::copy(levels[top].sol.begin(), levels[top].sol.end(), dst);
}

}  // namespace oracle
