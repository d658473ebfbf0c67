// CPU oracle (TEST INFRASTRUCTURE ONLY): SIPG Laplace operator, geometry, RHS, error.
// See oracle_core.h for the reference files this restates.
#include "oracle_core.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace oracle
{
// ---------------------------------------------------------------- 1D tables

static void legendre(int n, double t, double & p, double & dp)
{
  // P_n(t) and P_n'(t) by the three-term recurrence
  double p0 = 1.0, p1 = t;
  if(n == 0)
  {
    p  = 1.0;
    dp = 0.0;
    return;
  }
  for(int m = 1; m < n; ++m)
  {
    double const p2 = ((2 * m + 1) * t * p1 - m * p0) / (m + 1);
    p0              = p1;
    p1              = p2;
  }
  p  = p1;
  dp = n * (t * p1 - p0) / (t * t - 1.0);
}

void gauss_legendre(int n, std::vector<double> & x, std::vector<double> & w)
{
  // QGauss(n) on [0,1] (operators/quadrature.h:43-46): roots of P_n by Newton iteration
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for(int i = 0; i < n; ++i)
  {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for(int it = 0; it < 100; ++it)
    {
      double p, dp;
      legendre(n, t, p, dp);
      double const dt = p / dp;
      t -= dt;
      if(std::fabs(dt) < 1e-17)
        break;
    }
    double p, dp;
    legendre(n, t, p, dp);
    // ascending order on [0,1]
    x[n - 1 - i] = 0.5 * (t + 1.0);
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dp * dp);
  }
}

void gauss_lobatto(int n, std::vector<double> & x)
{
  // support points of FE_DGQ(k) (finite_element.h:56-61): 0, 1 and the roots of P'_{n-1}
  x.assign(n, 0.0);
  int const N = n - 1;
  if(N == 0)
  {
    x[0] = 0.5;
    return;
  }
  for(int i = 0; i < n; ++i)
  {
    double t = std::cos(M_PI * i / N);
    for(int it = 0; it < 200; ++it)
    {
      // P_N and P_{N-1}; Newton update on (1-t^2) P_N'(t) written via the recurrence
      double p0 = 1.0, p1 = t;
      for(int m = 2; m <= N; ++m)
      {
        double const p2 = ((2 * m - 1) * t * p1 - (m - 1) * p0) / m;
        p0              = p1;
        p1              = p2;
      }
      double const pN = (N == 1) ? t : p1, pNm1 = (N == 1) ? 1.0 : p0;
      double const dt = (t * pN - pNm1) / (n * pN);
      t -= dt;
      if(std::fabs(dt) < 1e-17)
        break;
    }
    x[N - i] = 0.5 * (t + 1.0);
  }
  x[0] = 0.0;
  x[N] = 1.0;
}

double lagrange(std::vector<double> const & nodes, int i, double x)
{
  double r = 1.0;
  for(size_t j = 0; j < nodes.size(); ++j)
    if((int)j != i)
      r *= (x - nodes[j]) / (nodes[i] - nodes[j]);
  return r;
}

double lagrange_deriv(std::vector<double> const & nodes, int i, double x)
{
  double s = 0.0;
  for(size_t m = 0; m < nodes.size(); ++m)
  {
    if((int)m == i)
      continue;
    double r = 1.0 / (nodes[i] - nodes[m]);
    for(size_t j = 0; j < nodes.size(); ++j)
      if((int)j != i && j != m)
        r *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    s += r;
  }
  return s;
}

Tables1D make_tables(int degree)
{
  Tables1D t;
  int const n = degree + 1;
  t.n         = n;
  gauss_lobatto(n, t.gl);
  gauss_legendre(n, t.gp, t.gw);
  t.S.resize(n * n);
  t.Sd.resize(n * n);
  for(int q = 0; q < n; ++q)
    for(int i = 0; i < n; ++i)
    {
      t.S[q * n + i]  = lagrange(t.gl, i, t.gp[q]);
      t.Sd[q * n + i] = lagrange_deriv(t.gl, i, t.gp[q]);
    }
  t.v0.resize(n);
  t.v1.resize(n);
  t.d0.resize(n);
  t.d1.resize(n);
  for(int i = 0; i < n; ++i)
  {
    t.v0[i] = (i == 0) ? 1.0 : 0.0;  // nodal basis on GL points: exact endpoint values
    t.v1[i] = (i == n - 1) ? 1.0 : 0.0;
    t.d0[i] = lagrange_deriv(t.gl, i, 0.0);
    t.d1[i] = lagrange_deriv(t.gl, i, 1.0);
  }
  return t;
}

// ---------------------------------------------------------------- mesh / geometry

Mesh::Mesh(MeshDesc const & desc) : d(desc)
{
  n_cells = (int64_t)d.n_cells[0] * d.n_cells[1] * d.n_cells[2];
  affine  = (d.deformation == 0.0);
  for(int k = 0; k < 3; ++k)
    h[k] = (d.upper[k] - d.lower[k]) / d.n_cells[k];
  if(!affine)
  {
    if(d.mapping_degree < 1)
      throw std::runtime_error("mapping_degree must be >= 1");
    gauss_lobatto(d.mapping_degree + 1, map_gl);
  }
}

void Mesh::cell_coords(int64_t c, int & cx, int & cy, int & cz) const
{
  cx = (int)(c % d.n_cells[0]);
  int64_t r = c / d.n_cells[0];
  cy        = (int)(r % d.n_cells[1]);
  cz        = (int)(r / d.n_cells[1]);
}

int64_t Mesh::neighbor(int64_t c, int dir, int side) const
{
  int cc[3];
  cell_coords(c, cc[0], cc[1], cc[2]);
  cc[dir] += side ? 1 : -1;
  if(cc[dir] < 0 || cc[dir] >= d.n_cells[dir])
  {
    if(!d.periodic[dir])
      return -1;
    cc[dir] = (cc[dir] + d.n_cells[dir]) % d.n_cells[dir];
  }
  return cell_index(cc[0], cc[1], cc[2]);
}

void Mesh::push_forward(double const * p, double * x) const
{
  // DeformedCubeManifold::push_forward: x_d = p_d + delta * prod_e sin(f pi (p_e-left)/(right-left))
  double s = d.deformation;
  for(int e = 0; e < 3; ++e)
    s *= std::sin(d.frequency * M_PI * (p[e] - d.lower[e]) / (d.upper[e] - d.lower[e]));
  for(int e = 0; e < 3; ++e)
    x[e] = p[e] + s;
}

void Mesh::map_point(int64_t c, double const * xi, double * x, double * J) const
{
  int cc[3];
  cell_coords(c, cc[0], cc[1], cc[2]);
  if(affine)
  {
    for(int i = 0; i < 3; ++i)
    {
      x[i] = d.lower[i] + h[i] * (cc[i] + xi[i]);
      for(int j = 0; j < 3; ++j)
        J[i * 3 + j] = (i == j) ? h[i] : 0.0;
    }
    return;
  }
  // MappingQ(m-1) with GL support points placed by the chart manifold. All refined vertices
  // and higher-order support points of a ChartManifold cell are push_forward of the chart-space
  // (undeformed) points, so X_abc = push_forward(cartesian GL point).
  int const            m = (int)map_gl.size();
  std::vector<double>  psi[3], dpsi[3];
  for(int k = 0; k < 3; ++k)
  {
    psi[k].resize(m);
    dpsi[k].resize(m);
    for(int a = 0; a < m; ++a)
    {
      psi[k][a]  = lagrange(map_gl, a, xi[k]);
      dpsi[k][a] = lagrange_deriv(map_gl, a, xi[k]);
    }
  }
  for(int i = 0; i < 3; ++i)
  {
    x[i] = 0.0;
    for(int j = 0; j < 3; ++j)
      J[i * 3 + j] = 0.0;
  }
  for(int cz = 0; cz < m; ++cz)
    for(int cy = 0; cy < m; ++cy)
      for(int cx = 0; cx < m; ++cx)
      {
        double p[3] = {d.lower[0] + h[0] * (cc[0] + map_gl[cx]),
                       d.lower[1] + h[1] * (cc[1] + map_gl[cy]),
                       d.lower[2] + h[2] * (cc[2] + map_gl[cz])};
        double X[3];
        push_forward(p, X);
        double const w  = psi[0][cx] * psi[1][cy] * psi[2][cz];
        double const g0 = dpsi[0][cx] * psi[1][cy] * psi[2][cz];
        double const g1 = psi[0][cx] * dpsi[1][cy] * psi[2][cz];
        double const g2 = psi[0][cx] * psi[1][cy] * dpsi[2][cz];
        for(int i = 0; i < 3; ++i)
        {
          x[i] += w * X[i];
          J[i * 3 + 0] += g0 * X[i];
          J[i * 3 + 1] += g1 * X[i];
          J[i * 3 + 2] += g2 * X[i];
        }
      }
}

void invert3(double const * J, double * Ji, double & det)
{
  det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
        J[2] * (J[3] * J[7] - J[4] * J[6]);
  double const id = 1.0 / det;
  Ji[0]           = (J[4] * J[8] - J[5] * J[7]) * id;
  Ji[1]           = (J[2] * J[7] - J[1] * J[8]) * id;
  Ji[2]           = (J[1] * J[5] - J[2] * J[4]) * id;
  Ji[3]           = (J[5] * J[6] - J[3] * J[8]) * id;
  Ji[4]           = (J[0] * J[8] - J[2] * J[6]) * id;
  Ji[5]           = (J[2] * J[3] - J[0] * J[5]) * id;
  Ji[6]           = (J[3] * J[7] - J[4] * J[6]) * id;
  Ji[7]           = (J[1] * J[6] - J[0] * J[7]) * id;
  Ji[8]           = (J[0] * J[4] - J[1] * J[3]) * id;
}

// ---------------------------------------------------------------- tensor kernels

static constexpr int kMaxTensor = 16 * 16 * 16;

void tensor_apply(int ma, int mb, int mc, int n, double const * A, double const * B,
                  double const * C, double const * in, double * out)
{
  // three 1D contractions (sum factorization), x then y then z
  double t1[kMaxTensor], t2[kMaxTensor];
  if(ma * n * n > kMaxTensor || ma * mb * n > kMaxTensor)
    throw std::runtime_error("tensor too large");
  for(int l = 0; l < n; ++l)
    for(int j = 0; j < n; ++j)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int i = 0; i < n; ++i)
          s += A[a * n + i] * in[i + n * (j + n * l)];
        t1[a + ma * (j + n * l)] = s;
      }
  for(int l = 0; l < n; ++l)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int j = 0; j < n; ++j)
          s += B[b * n + j] * t1[a + ma * (j + n * l)];
        t2[a + ma * (b + mb * l)] = s;
      }
  for(int c = 0; c < mc; ++c)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int l = 0; l < n; ++l)
          s += C[c * n + l] * t2[a + ma * (b + mb * l)];
        out[a + ma * (b + mb * c)] = s;
      }
}

void tensor_apply_T(int ma, int mb, int mc, int n, double const * A, double const * B,
                    double const * C, double const * in, double * out)
{
  double t2[kMaxTensor], t1[kMaxTensor];
  if(ma * n * n > kMaxTensor || ma * mb * n > kMaxTensor)
    throw std::runtime_error("tensor too large");
  for(int l = 0; l < n; ++l)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int c = 0; c < mc; ++c)
          s += C[c * n + l] * in[a + ma * (b + mb * c)];
        t2[a + ma * (b + mb * l)] = s;
      }
  for(int l = 0; l < n; ++l)
    for(int j = 0; j < n; ++j)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int b = 0; b < mb; ++b)
          s += B[b * n + j] * t2[a + ma * (b + mb * l)];
        t1[a + ma * (j + n * l)] = s;
      }
  for(int l = 0; l < n; ++l)
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < n; ++i)
      {
        double s = 0;
        for(int a = 0; a < ma; ++a)
          s += A[a * n + i] * t1[a + ma * (j + n * l)];
        out[i + n * (j + n * l)] += s;
      }
}

// ---------------------------------------------------------------- operator

LaplaceOperator::LaplaceOperator(MeshDesc const & desc, int degree_, double ip)
  : mesh(desc), degree(degree_), n(degree_ + 1), n3((degree_ + 1) * (degree_ + 1) * (degree_ + 1)),
    ip_factor(ip), t(make_tables(degree_))
{
  int64_t const nc = mesh.n_cells;
  cell_jinv.resize((size_t)nc * n3 * 9);
  cell_jxw.resize((size_t)nc * n3);
  tau_cell.resize(nc);

  if(!mesh.affine)
  {
    // curved cells: evaluate the face mappings once (n^2 points x 16 values per cell face)
    int const n2 = n * n;
    face_cache.resize((size_t)nc * 6 * n2 * 16);
#pragma omp parallel for schedule(dynamic, 16)
    for(int64_t c = 0; c < nc; ++c)
    {
      std::vector<double> fj(9 * n2), fw(n2), fn(3 * n2), fx(3 * n2);
      for(int f = 0; f < 6; ++f)
      {
        compute_face_geometry(c, f / 2, f % 2, fj.data(), fw.data(), fn.data(), fx.data());
        double * p = &face_cache[((size_t)c * 6 + f) * n2 * 16];
        for(int q = 0; q < n2; ++q)
        {
          for(int e = 0; e < 9; ++e)
            p[16 * q + e] = fj[9 * q + e];
          p[16 * q + 9] = fw[q];
          for(int e = 0; e < 3; ++e)
          {
            p[16 * q + 10 + e] = fn[3 * q + e];
            p[16 * q + 13 + e] = fx[3 * q + e];
          }
        }
      }
    }
  }

#pragma omp parallel for schedule(dynamic, 16)
  for(int64_t c = 0; c < nc; ++c)
  {
    double vol = 0;
    for(int qz = 0; qz < n; ++qz)
      for(int qy = 0; qy < n; ++qy)
        for(int qx = 0; qx < n; ++qx)
        {
          int const    q     = qx + n * (qy + n * qz);
          double const xi[3] = {t.gp[qx], t.gp[qy], t.gp[qz]};
          double       x[3], J[9], Ji[9], det;
          mesh.map_point(c, xi, x, J);
          invert3(J, Ji, det);
          for(int e = 0; e < 9; ++e)
            cell_jinv[((size_t)c * n3 + q) * 9 + e] = Ji[e];
          double const jxw              = det * t.gw[qx] * t.gw[qy] * t.gw[qz];
          cell_jxw[(size_t)c * n3 + q] = jxw;
          vol += jxw;
        }
    // penalty parameter (interior_penalty_parameter.h:43-99): QGauss(k+1) on cell and faces,
    // factor 1 on non-periodic boundary faces and 1/2 otherwise
    double surf = 0;
    std::vector<double> fj(n * n * 9), fw(n * n), fn(n * n * 3), fx(n * n * 3);
    for(int dir = 0; dir < 3; ++dir)
      for(int side = 0; side < 2; ++side)
      {
        face_geometry(c, dir, side, fj.data(), fw.data(), fn.data(), fx.data());
        double const factor = (mesh.neighbor(c, dir, side) < 0) ? 1.0 : 0.5;
        for(int q = 0; q < n * n; ++q)
          surf += fw[q] * factor;
      }
    tau_cell[c] = surf / vol;
  }
}

void LaplaceOperator::face_geometry(int64_t c, int dir, int side, double * jinv, double * jxw,
                                    double * normal, double * xq) const
{
  if(face_cache.empty())
  {
    compute_face_geometry(c, dir, side, jinv, jxw, normal, xq);
    return;
  }
  int const      n2 = n * n;
  double const * p  = &face_cache[((size_t)c * 6 + 2 * dir + side) * n2 * 16];
  for(int q = 0; q < n2; ++q)
  {
    for(int e = 0; e < 9; ++e)
      jinv[9 * q + e] = p[16 * q + e];
    jxw[q] = p[16 * q + 9];
    for(int e = 0; e < 3; ++e)
    {
      normal[3 * q + e] = p[16 * q + 10 + e];
      xq[3 * q + e]     = p[16 * q + 13 + e];
    }
  }
}

void LaplaceOperator::compute_face_geometry(int64_t c, int dir, int side, double * jinv,
                                            double * jxw, double * normal, double * xq) const
{
  int const t1 = (dir == 0) ? 1 : 0;
  int const t2 = (dir == 2) ? 1 : 2;
  for(int b = 0; b < n; ++b)
    for(int a = 0; a < n; ++a)
    {
      int const q = a + n * b;
      double    xi[3];
      xi[dir] = side ? 1.0 : 0.0;
      xi[t1]  = t.gp[a];
      xi[t2]  = t.gp[b];
      double J[9], Ji[9], det;
      mesh.map_point(c, xi, xq + 3 * q, J);
      invert3(J, Ji, det);
      // physical normal ~ J^{-T} nhat, nhat = +-e_dir; surface element det*|J^{-T} nhat|
      double nv[3];
      double const sgn = side ? 1.0 : -1.0;
      for(int i = 0; i < 3; ++i)
        nv[i] = sgn * Ji[dir * 3 + i];  // (J^{-T})_{i,dir} = Jinv_{dir,i}
      double const nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      for(int i = 0; i < 3; ++i)
        normal[3 * q + i] = nv[i] / nrm;
      jxw[q] = det * nrm * t.gw[a] * t.gw[b];
      for(int e = 0; e < 9; ++e)
        jinv[9 * q + e] = Ji[e];
    }
}

// 1D operator rows for a face evaluation: along dir a 1xn row, along tangentials nxn
void LaplaceOperator::face_evaluate(double const * u, int dir, int side, double * val,
                                    double * gref) const
{
  double const *vrow = side ? t.v1.data() : t.v0.data(), *drow = side ? t.d1.data() : t.d0.data();
  int           m[3] = {n, n, n};
  m[dir]             = 1;
  double const *ops_v[3], *ops_g[3][3];
  for(int k = 0; k < 3; ++k)
    ops_v[k] = (k == dir) ? vrow : t.S.data();
  for(int j = 0; j < 3; ++j)
    for(int k = 0; k < 3; ++k)
    {
      if(k == dir)
        ops_g[j][k] = (j == dir) ? drow : vrow;
      else
        ops_g[j][k] = (j == k) ? t.Sd.data() : t.S.data();
    }
  tensor_apply(m[0], m[1], m[2], n, ops_v[0], ops_v[1], ops_v[2], u, val);
  for(int j = 0; j < 3; ++j)
    tensor_apply(m[0], m[1], m[2], n, ops_g[j][0], ops_g[j][1], ops_g[j][2], u, gref + j * n * n);
}

void LaplaceOperator::face_integrate(int dir, int side, double const * val, double const * gref,
                                     double * out) const
{
  double const *vrow = side ? t.v1.data() : t.v0.data(), *drow = side ? t.d1.data() : t.d0.data();
  int           m[3] = {n, n, n};
  m[dir]             = 1;
  double const *ops_v[3], *ops_g[3][3];
  for(int k = 0; k < 3; ++k)
    ops_v[k] = (k == dir) ? vrow : t.S.data();
  for(int j = 0; j < 3; ++j)
    for(int k = 0; k < 3; ++k)
    {
      if(k == dir)
        ops_g[j][k] = (j == dir) ? drow : vrow;
      else
        ops_g[j][k] = (j == k) ? t.Sd.data() : t.S.data();
    }
  tensor_apply_T(m[0], m[1], m[2], n, ops_v[0], ops_v[1], ops_v[2], val, out);
  for(int j = 0; j < 3; ++j)
    tensor_apply_T(m[0], m[1], m[2], n, ops_g[j][0], ops_g[j][1], ops_g[j][2], gref + j * n * n,
                   out);
}

void LaplaceOperator::cell_loop(double * dst, double const * src) const
{
  // OperatorBase::cell_loop + LaplaceOperator::do_cell_integral: submit_gradient(get_gradient)
  int64_t const nc = mesh.n_cells;
#pragma omp parallel
  {
    std::vector<double> g(3 * n3), out(n3);
    double const *      ops[3][3];
    for(int j = 0; j < 3; ++j)
      for(int k = 0; k < 3; ++k)
        ops[j][k] = (j == k) ? t.Sd.data() : t.S.data();
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      double const * u = src + c * n3;
      for(int j = 0; j < 3; ++j)
        tensor_apply(n, n, n, n, ops[j][0], ops[j][1], ops[j][2], u, g.data() + j * n3);
      for(int q = 0; q < n3; ++q)
      {
        double const * Ji  = &cell_jinv[((size_t)c * n3 + q) * 9];
        double const   jxw = cell_jxw[(size_t)c * n3 + q];
        double         gp[3], f[3];
        for(int i = 0; i < 3; ++i)  // physical gradient J^{-T} grad_ref
          gp[i] = Ji[0 * 3 + i] * g[q] + Ji[1 * 3 + i] * g[n3 + q] + Ji[2 * 3 + i] * g[2 * n3 + q];
        for(int j = 0; j < 3; ++j)  // back to reference: J^{-1} (JxW grad)
          f[j] = jxw * (Ji[j * 3 + 0] * gp[0] + Ji[j * 3 + 1] * gp[1] + Ji[j * 3 + 2] * gp[2]);
        for(int j = 0; j < 3; ++j)
          g[j * n3 + q] = f[j];
      }
      std::fill(out.begin(), out.end(), 0.0);
      for(int j = 0; j < 3; ++j)
        tensor_apply_T(n, n, n, n, ops[j][0], ops[j][1], ops[j][2], g.data() + j * n3, out.data());
      for(int i = 0; i < n3; ++i)
        dst[c * n3 + i] += out[i];
    }
  }
}

void LaplaceOperator::face_loop(double * dst, double const * src) const
{
  // OperatorBase::face_loop + do_face_integral (laplace_operator.cpp:139-163), one visit per
  // interior (or periodic) face; faces ordered along pencils so that pencils run in parallel
  int const n2 = n * n;
  double const tau_factor = penalty_factor();
  for(int dir = 0; dir < 3; ++dir)
  {
    int const t1 = (dir == 0) ? 1 : 0, t2 = (dir == 2) ? 1 : 2;
    int const np = mesh.d.n_cells[t1] * mesh.d.n_cells[t2];
#pragma omp parallel
    {
      std::vector<double> vm(n2), vp(n2), gm(3 * n2), gp(3 * n2);
      std::vector<double> jm(9 * n2), jp(9 * n2), wm(n2), wp(n2), nm(3 * n2), np_(3 * n2),
        xm(3 * n2), xp(3 * n2);
      std::vector<double> rvm(n2), rgm(3 * n2), rvp(n2), rgp(3 * n2), om(n * n2), op(n * n2);
#pragma omp for schedule(static)
      for(int pencil = 0; pencil < np; ++pencil)
      {
        int cc[3];
        cc[t1] = pencil % mesh.d.n_cells[t1];
        cc[t2] = pencil / mesh.d.n_cells[t1];
        for(int s = 0; s < mesh.d.n_cells[dir]; ++s)
        {
          cc[dir]          = s;
          int64_t const cm = mesh.cell_index(cc[0], cc[1], cc[2]);
          int64_t const cp = mesh.neighbor(cm, dir, 1);
          if(cp < 0)
            continue;
          face_evaluate(src + cm * n3, dir, 1, vm.data(), gm.data());
          face_evaluate(src + cp * n3, dir, 0, vp.data(), gp.data());
          face_geometry(cm, dir, 1, jm.data(), wm.data(), nm.data(), xm.data());
          face_geometry(cp, dir, 0, jp.data(), wp.data(), np_.data(), xp.data());
          double const tau = std::max(tau_cell[cm], tau_cell[cp]) * tau_factor;
          for(int q = 0; q < n2; ++q)
          {
            double const * Jm = &jm[9 * q];
            double const * Jp = &jp[9 * q];
            double const * nv = &nm[3 * q];
            // J^{-1} n for both sides (normal derivative = (J^{-1} n) . grad_ref)
            double am[3], ap[3];
            for(int j = 0; j < 3; ++j)
            {
              am[j] = Jm[j * 3 + 0] * nv[0] + Jm[j * 3 + 1] * nv[1] + Jm[j * 3 + 2] * nv[2];
              ap[j] = Jp[j * 3 + 0] * nv[0] + Jp[j * 3 + 1] * nv[1] + Jp[j * 3 + 2] * nv[2];
            }
            double const um = vm[q], up = vp[q];
            double const dnm =
              am[0] * gm[q] + am[1] * gm[n2 + q] + am[2] * gm[2 * n2 + q];
            double const dnp =
              ap[0] * gp[q] + ap[1] * gp[n2 + q] + ap[2] * gp[2 * n2 + q];
            double const gflux = -0.5 * (um - up);
            double const vflux = 0.5 * (dnm + dnp) - tau * (um - up);
            double const jxw   = wm[q];
            rvm[q]             = -vflux * jxw;
            rvp[q]             = vflux * jxw;
            for(int j = 0; j < 3; ++j)
            {
              rgm[j * n2 + q] = gflux * jxw * am[j];
              rgp[j * n2 + q] = gflux * jxw * ap[j];
            }
          }
          std::fill(om.begin(), om.end(), 0.0);
          std::fill(op.begin(), op.end(), 0.0);
          face_integrate(dir, 1, rvm.data(), rgm.data(), om.data());
          face_integrate(dir, 0, rvp.data(), rgp.data(), op.data());
          for(int i = 0; i < n3; ++i)
          {
            dst[cm * n3 + i] += om[i];
            dst[cp * n3 + i] += op[i];
          }
        }
      }
    }
  }
}

void LaplaceOperator::boundary_face_loop(double * dst, double const * src) const
{
  // boundary_face_loop_hom_operator + do_boundary_integral with the mirror principle of
  // weak_boundary_conditions.h:39-47,142-160 (homogeneous operator)
  int const     n2 = n * n;
  int64_t const nc = mesh.n_cells;
  double const  tau_factor = penalty_factor();
#pragma omp parallel
  {
    std::vector<double> vm(n2), gm(3 * n2), jm(9 * n2), wm(n2), nm(3 * n2), xm(3 * n2), rv(n2),
      rg(3 * n2), om(n3);
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      for(int dir = 0; dir < 3; ++dir)
        for(int side = 0; side < 2; ++side)
        {
          if(mesh.neighbor(c, dir, side) >= 0)
            continue;
          int const  type = mesh.d.boundary_type[2 * dir + side];
          face_evaluate(src + c * n3, dir, side, vm.data(), gm.data());
          face_geometry(c, dir, side, jm.data(), wm.data(), nm.data(), xm.data());
          double const tau = tau_cell[c] * tau_factor;
          for(int q = 0; q < n2; ++q)
          {
            double const * Jm = &jm[9 * q];
            double const * nv = &nm[3 * q];
            double         am[3];
            for(int j = 0; j < 3; ++j)
              am[j] = Jm[j * 3 + 0] * nv[0] + Jm[j * 3 + 1] * nv[1] + Jm[j * 3 + 2] * nv[2];
            double const um  = vm[q];
            double const dnm = am[0] * gm[q] + am[1] * gm[n2 + q] + am[2] * gm[2 * n2 + q];
            double       up, dnp;
            if(type == 0)
            {  // Dirichlet: u+ = -u-, grad(u+).n = grad(u-).n
              up  = -um;
              dnp = dnm;
            }
            else
            {  // Neumann: u+ = u-, grad(u+).n = -grad(u-).n
              up  = um;
              dnp = -dnm;
            }
            double const gflux = -0.5 * (um - up);
            double const vflux = 0.5 * (dnm + dnp) - tau * (um - up);
            rv[q]              = -vflux * wm[q];
            for(int j = 0; j < 3; ++j)
              rg[j * n2 + q] = gflux * wm[q] * am[j];
          }
          std::fill(om.begin(), om.end(), 0.0);
          face_integrate(dir, side, rv.data(), rg.data(), om.data());
          for(int i = 0; i < n3; ++i)
            dst[c * n3 + i] += om[i];
        }
    }
  }
}

void LaplaceOperator::vmult(double * dst, double const * src) const
{
  int64_t const N = n_dofs();
#pragma omp parallel for schedule(static)
  for(int64_t i = 0; i < N; ++i)
    dst[i] = 0.0;
  cell_loop(dst, src);
  face_loop(dst, src);
  boundary_face_loop(dst, src);
}

static inline double sine_u(double omega, double const * x)
{
  return std::sin(omega * x[0]) * std::sin(omega * x[1]) * std::sin(omega * x[2]);
}

void LaplaceOperator::rhs_sine(double omega, double * b) const
{
  // Operator::rhs (operator.cpp:414-423): RHSOperator (f) plus the inhomogeneous boundary
  // operator moved to the right side with factor -1 (operator_base.cpp:507-554,1436-1491)
  int64_t const nc = mesh.n_cells;
  int const     n2 = n * n;
  double const  tau_factor = penalty_factor();
#pragma omp parallel
  {
    std::vector<double> fq(n3), out(n3), vm(n2), jm(9 * n2), wm(n2), nm(3 * n2), xm(3 * n2),
      rv(n2), rg(3 * n2);
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      for(int qz = 0; qz < n; ++qz)
        for(int qy = 0; qy < n; ++qy)
          for(int qx = 0; qx < n; ++qx)
          {
            int const    q     = qx + n * (qy + n * qz);
            double const xi[3] = {t.gp[qx], t.gp[qy], t.gp[qz]};
            double       x[3], J[9];
            mesh.map_point(c, xi, x, J);
            fq[q] = 3.0 * omega * omega * sine_u(omega, x) * cell_jxw[(size_t)c * n3 + q];
          }
      std::fill(out.begin(), out.end(), 0.0);
      tensor_apply_T(n, n, n, n, t.S.data(), t.S.data(), t.S.data(), fq.data(), out.data());
      for(int dir = 0; dir < 3; ++dir)
        for(int side = 0; side < 2; ++side)
        {
          if(mesh.neighbor(c, dir, side) >= 0)
            continue;
          int const type = mesh.d.boundary_type[2 * dir + side];
          face_geometry(c, dir, side, jm.data(), wm.data(), nm.data(), xm.data());
          double const tau = tau_cell[c] * tau_factor;
          for(int q = 0; q < n2; ++q)
          {
            double const * Jm = &jm[9 * q];
            double const * nv = &nm[3 * q];
            double         am[3];
            for(int j = 0; j < 3; ++j)
              am[j] = Jm[j * 3 + 0] * nv[0] + Jm[j * 3 + 1] * nv[1] + Jm[j * 3 + 2] * nv[2];
            double const * x = &xm[3 * q];
            double         gflux, vflux;
            if(type == 0)
            {  // inhomogeneous Dirichlet: u- = 0, u+ = 2g, normal gradients zero
              double const g = sine_u(omega, x);
              gflux          = -0.5 * (0.0 - 2.0 * g);
              vflux          = 0.0 - tau * (0.0 - 2.0 * g);
            }
            else
            {  // inhomogeneous Neumann: u-=u+=0, grad(u+).n = 2h
              double const s0 = std::sin(omega * x[0]), s1 = std::sin(omega * x[1]),
                           s2 = std::sin(omega * x[2]);
              double const gr[3] = {omega * std::cos(omega * x[0]) * s1 * s2,
                                    omega * s0 * std::cos(omega * x[1]) * s2,
                                    omega * s0 * s1 * std::cos(omega * x[2])};
              double const h     = gr[0] * nv[0] + gr[1] * nv[1] + gr[2] * nv[2];
              gflux              = 0.0;
              vflux              = 0.5 * (0.0 + 2.0 * h);
            }
            // integral of the inhomogeneous operator, then multiplied by -1
            rv[q] = vflux * wm[q];
            for(int j = 0; j < 3; ++j)
              rg[j * n2 + q] = -gflux * wm[q] * am[j];
          }
          // value part: operator submits -vflux; moved to rhs => +vflux. gradient part:
          // operator submits +gflux; moved to rhs => -gflux
          face_integrate(dir, side, rv.data(), rg.data(), out.data());
        }
      for(int i = 0; i < n3; ++i)
        b[c * n3 + i] = out[i];
    }
  }
}

namespace
{
void fourier_velocity(double const * modes, int n_modes, double t, double const * x, double * u)
{
  u[0] = u[1] = u[2] = 0.0;
  for(int m = 0; m < n_modes; ++m)
  {
    double const * q  = modes + 10 * m;
    double const   ph = 2.0 * M_PI * (q[0] * x[0] + q[1] * x[1] + q[2] * x[2]) + q[7] + q[8] * t;
    double const   a  = q[6] * std::cos(ph);
    for(int d = 0; d < 3; ++d)
      u[d] += a * q[3 + d];
  }
}
}  // namespace

void LaplaceOperator::rhs_divergence_fourier(double const * modes, int n_modes, double time,
                                             double * b) const
{
  int64_t const nc = mesh.n_cells;
  int const     n2 = n * n;
  // cell term -(grad phi_i, u): reference-gradient data J^{-1} u JxW, tested with Sd
#pragma omp parallel
  {
    std::vector<double> g(3 * n3), out(n3);
    double const *      ops[3][3];
    for(int j = 0; j < 3; ++j)
      for(int k = 0; k < 3; ++k)
        ops[j][k] = (j == k) ? t.Sd.data() : t.S.data();
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      for(int qz = 0; qz < n; ++qz)
        for(int qy = 0; qy < n; ++qy)
          for(int qx = 0; qx < n; ++qx)
          {
            int const    q     = qx + n * (qy + n * qz);
            double const xi[3] = {t.gp[qx], t.gp[qy], t.gp[qz]};
            double       x[3], J[9], u[3];
            mesh.map_point(c, xi, x, J);
            fourier_velocity(modes, n_modes, time, x, u);
            double const * Ji  = &cell_jinv[((size_t)c * n3 + q) * 9];
            double const   jxw = cell_jxw[(size_t)c * n3 + q];
            for(int j = 0; j < 3; ++j)
              g[j * n3 + q] = -jxw * (Ji[j * 3 + 0] * u[0] + Ji[j * 3 + 1] * u[1] + Ji[j * 3 + 2] * u[2]);
          }
      std::fill(out.begin(), out.end(), 0.0);
      for(int j = 0; j < 3; ++j)
        tensor_apply_T(n, n, n, n, ops[j][0], ops[j][1], ops[j][2], g.data() + j * n3, out.data());
      for(int i = 0; i < n3; ++i)
        b[c * n3 + i] = out[i];
    }
  }
  // face term, one visit per interior/periodic face (upper side of cell c): the flux u.n at
  // the face points goes to c with +, to the neighbour with - (its outward normal is -n).
  // Boundary faces are walls: {u} = 0, no contribution.
  std::vector<double> jm(9 * n2), wm(n2), nm(3 * n2), xm(3 * n2), fv(n2), fneg(n2),
    zero(3 * n2, 0.0), out(n3);
  for(int64_t c = 0; c < nc; ++c)
    for(int dir = 0; dir < 3; ++dir)
    {
      int64_t const nb = mesh.neighbor(c, dir, 1);
      if(nb < 0)
        continue;
      face_geometry(c, dir, 1, jm.data(), wm.data(), nm.data(), xm.data());
      for(int q = 0; q < n2; ++q)
      {
        double u[3];
        fourier_velocity(modes, n_modes, time, &xm[3 * q], u);
        fv[q]   = (u[0] * nm[3 * q] + u[1] * nm[3 * q + 1] + u[2] * nm[3 * q + 2]) * wm[q];
        fneg[q] = -fv[q];
      }
      std::fill(out.begin(), out.end(), 0.0);
      face_integrate(dir, 1, fv.data(), zero.data(), out.data());
      for(int i = 0; i < n3; ++i)
        b[c * n3 + i] += out[i];
      std::fill(out.begin(), out.end(), 0.0);
      face_integrate(dir, 0, fneg.data(), zero.data(), out.data());
      for(int i = 0; i < n3; ++i)
        b[nb * n3 + i] += out[i];
    }
  bool singular = true;
  for(int d = 0; d < 3; ++d)
    if(!mesh.d.periodic[d] && (mesh.d.boundary_type[2 * d] == 0 || mesh.d.boundary_type[2 * d + 1] == 0))
      singular = false;
  if(singular)
  {
    int64_t const nd   = n_dofs();
    double        mean = 0;
    for(int64_t i = 0; i < nd; ++i)
      mean += b[i];
    mean /= (double)nd;
    for(int64_t i = 0; i < nd; ++i)
      b[i] -= mean;
  }
}

double LaplaceOperator::error_sine(double omega, double const * u) const
{
  // calculate_error with QGauss(k+3) (error_calculation.cpp:38-103), relative L2 norm
  int const           ne = degree + 3;
  std::vector<double> ep, ew;
  gauss_legendre(ne, ep, ew);
  std::vector<double> Se(ne * n);
  for(int q = 0; q < ne; ++q)
    for(int i = 0; i < n; ++i)
      Se[q * n + i] = lagrange(t.gl, i, ep[q]);
  int64_t const nc = mesh.n_cells;
  double        err2 = 0, nrm2 = 0;
#pragma omp parallel reduction(+ : err2, nrm2)
  {
    std::vector<double> uq(ne * ne * ne);
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      tensor_apply(ne, ne, ne, n, Se.data(), Se.data(), Se.data(), u + c * n3, uq.data());
      for(int qz = 0; qz < ne; ++qz)
        for(int qy = 0; qy < ne; ++qy)
          for(int qx = 0; qx < ne; ++qx)
          {
            double const xi[3] = {ep[qx], ep[qy], ep[qz]};
            double       x[3], J[9], Ji[9], det;
            mesh.map_point(c, xi, x, J);
            invert3(J, Ji, det);
            double const w  = det * ew[qx] * ew[qy] * ew[qz];
            double const ex = sine_u(omega, x);
            double const e  = uq[qx + ne * (qy + ne * qz)] - ex;
            err2 += w * e * e;
            nrm2 += w * ex * ex;
          }
    }
  }
  return std::sqrt(err2) / std::sqrt(nrm2);
}

void LaplaceOperator::diagonal(double * dg) const { probe_blocks(dg, nullptr); }

void LaplaceOperator::block_diagonal(double * blocks) const { probe_blocks(nullptr, blocks); }

void LaplaceOperator::probe_blocks(double * dg, double * blocks) const
{
  // diag(A) or the whole diagonal block A_KK: cell term plus the self-coupling of every face of
  // the cell (neighbour values zero), probed with unit vectors (MatrixFreeTools::compute_diagonal
  // semantics; OperatorBase::add_block_diagonal_matrices, operator_base.cpp:649-680)
  int64_t const nc         = mesh.n_cells;
  int const     n2         = n * n;
  double const  tau_factor = penalty_factor();
#pragma omp parallel
  {
    std::vector<double> e(n3), g(3 * n3), out(n3), vm(n2), gm(3 * n2), jm(9 * n2), wm(n2),
      nm(3 * n2), xm(3 * n2), rv(n2), rg(3 * n2);
    double const *ops[3][3];
    for(int j = 0; j < 3; ++j)
      for(int k = 0; k < 3; ++k)
        ops[j][k] = (j == k) ? t.Sd.data() : t.S.data();
#pragma omp for schedule(dynamic, 4)
    for(int64_t c = 0; c < nc; ++c)
    {
      // face geometry of all six faces once
      std::vector<double> fj(6 * 9 * n2), fw(6 * n2), fn(6 * 3 * n2), fx(6 * 3 * n2);
      double              ftau[6];
      int                 ftype[6];
      bool                fself[6];  // periodic direction with one cell: the neighbour is c
      for(int f = 0; f < 6; ++f)
      {
        int const dir = f / 2, side = f % 2;
        face_geometry(c, dir, side, &fj[f * 9 * n2], &fw[f * n2], &fn[f * 3 * n2], &fx[f * 3 * n2]);
        int64_t const nb = mesh.neighbor(c, dir, side);
        ftype[f]         = nb >= 0 ? 2 : mesh.d.boundary_type[f];
        fself[f]         = (nb == c);
        ftau[f]          = (nb >= 0 ? std::max(tau_cell[c], tau_cell[nb]) : tau_cell[c]) * tau_factor;
      }
      for(int i = 0; i < n3; ++i)
      {
        std::fill(e.begin(), e.end(), 0.0);
        e[i] = 1.0;
        std::fill(out.begin(), out.end(), 0.0);
        for(int j = 0; j < 3; ++j)
          tensor_apply(n, n, n, n, ops[j][0], ops[j][1], ops[j][2], e.data(), g.data() + j * n3);
        for(int q = 0; q < n3; ++q)
        {
          double const * Ji  = &cell_jinv[((size_t)c * n3 + q) * 9];
          double const   jxw = cell_jxw[(size_t)c * n3 + q];
          double         gp[3], f[3];
          for(int a = 0; a < 3; ++a)
            gp[a] = Ji[0 * 3 + a] * g[q] + Ji[1 * 3 + a] * g[n3 + q] + Ji[2 * 3 + a] * g[2 * n3 + q];
          for(int j = 0; j < 3; ++j)
            f[j] = jxw * (Ji[j * 3 + 0] * gp[0] + Ji[j * 3 + 1] * gp[1] + Ji[j * 3 + 2] * gp[2]);
          for(int j = 0; j < 3; ++j)
            g[j * n3 + q] = f[j];
        }
        for(int j = 0; j < 3; ++j)
          tensor_apply_T(n, n, n, n, ops[j][0], ops[j][1], ops[j][2], g.data() + j * n3, out.data());
        for(int f = 0; f < 6; ++f)
        {
          int const dir = f / 2, side = f % 2;
          if(ftype[f] == 1)
            continue;  // homogeneous Neumann contributes nothing
          face_evaluate(e.data(), dir, side, vm.data(), gm.data());
          std::vector<double> vs(n2, 0.0), gs(3 * n2, 0.0);
          if(fself[f])  // the same cell seen across the periodic face (other side)
            face_evaluate(e.data(), dir, 1 - side, vs.data(), gs.data());
          for(int q = 0; q < n2; ++q)
          {
            double const * Jm = &fj[f * 9 * n2 + 9 * q];
            double const * nv = &fn[f * 3 * n2 + 3 * q];
            double         am[3];
            for(int j = 0; j < 3; ++j)
              am[j] = Jm[j * 3 + 0] * nv[0] + Jm[j * 3 + 1] * nv[1] + Jm[j * 3 + 2] * nv[2];
            double const um  = vm[q];
            double const dnm = am[0] * gm[q] + am[1] * gm[n2 + q] + am[2] * gm[2 * n2 + q];
            double       up = 0, dnp = 0;  // interior: neighbour contributes zero
            if(fself[f])
            {
              int const      fo = 2 * dir + (1 - side);
              double const * Jo = &fj[fo * 9 * n2 + 9 * q];
              double         ao[3];
              for(int j = 0; j < 3; ++j)
                ao[j] = Jo[j * 3 + 0] * nv[0] + Jo[j * 3 + 1] * nv[1] + Jo[j * 3 + 2] * nv[2];
              up  = vs[q];
              dnp = ao[0] * gs[q] + ao[1] * gs[n2 + q] + ao[2] * gs[2 * n2 + q];
            }
            if(ftype[f] == 0)
            {
              up  = -um;
              dnp = dnm;
            }
            double const gflux = -0.5 * (um - up);
            double const vflux = 0.5 * (dnm + dnp) - ftau[f] * (um - up);
            double const w     = fw[f * n2 + q];
            rv[q]              = -vflux * w;
            for(int j = 0; j < 3; ++j)
              rg[j * n2 + q] = gflux * w * am[j];
          }
          face_integrate(dir, side, rv.data(), rg.data(), out.data());
        }
        if(dg)
          dg[c * n3 + i] = out[i];
        if(blocks)  // column i of the cell's block
          std::copy(out.begin(), out.end(), blocks + ((size_t)c * n3 + i) * n3);
      }
    }
  }
}

double splitmix_uniform(uint64_t seed, uint64_t index)
{
  uint64_t z = seed + (index + 1) * 0x9E3779B97F4A7C15ull;
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z          = z ^ (z >> 31);
  return 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

}  // namespace oracle
