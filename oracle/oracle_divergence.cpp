// Weak DG divergence of a discrete velocity (the right-hand side of the pressure step of the
// dual-splitting scheme). TEST INFRASTRUCTURE ONLY (the parity checker of the HIP operator).
//
// Follows incompressible_navier_stokes/spatial_discretization/operators/divergence_operator.cpp
// :168-298 (weak form, do_cell_integral_weak / do_face_integral / do_boundary_integral) with the
// central flux of divergence_operator.h:63-68 and the exterior velocity of
// weak_boundary_conditions.h:84-142 (homogeneous operator: Dirichlet u+ = -u-, Neumann u+ = u-).
// Face-centric: every interior or periodic face is visited once and adds to both cells.
#include <cmath>
#include <vector>

#include "oracle_core.h"

namespace oracle
{
void divergence_apply(LaplaceOperator const & op, int degree_u, double const * u, double * b)
{
  Mesh const &  mesh = op.mesh;
  int const     np = op.n, nu = degree_u + 1, np3 = np * np * np, nu3 = nu * nu * nu;
  int64_t const nc = mesh.n_cells;
  Tables1D const tu = make_tables(degree_u);  // Gauss points of the velocity quadrature
  // pressure basis (GL nodes of degree k_p) at the velocity Gauss points, and at 0 and 1
  std::vector<double> Sp(nu * np), Sdp(nu * np), p0(np), p1(np);
  for(int q = 0; q < nu; ++q)
    for(int i = 0; i < np; ++i)
    {
      Sp[q * np + i]  = lagrange(op.t.gl, i, tu.gp[q]);
      Sdp[q * np + i] = lagrange_deriv(op.t.gl, i, tu.gp[q]);
    }
  for(int i = 0; i < np; ++i)
  {
    p0[i] = lagrange(op.t.gl, i, 0.0);
    p1[i] = lagrange(op.t.gl, i, 1.0);
  }
  for(int64_t i = 0; i < nc * np3; ++i)
    b[i] = 0.0;

  // cell term: -(u, grad q) = -sum_j (J^{-1} u)_j JxW d q / d xi_j
#pragma omp parallel
  {
    std::vector<double> uq(3 * nu3), g(3 * nu3);
#pragma omp for schedule(static)
    for(int64_t c = 0; c < nc; ++c)
    {
      for(int k = 0; k < 3; ++k)
        tensor_apply(nu, nu, nu, nu, tu.S.data(), tu.S.data(), tu.S.data(),
                     u + (c * 3 + k) * nu3, uq.data() + k * nu3);
      for(int qz = 0; qz < nu; ++qz)
        for(int qy = 0; qy < nu; ++qy)
          for(int qx = 0; qx < nu; ++qx)
          {
            int const    q     = qx + nu * (qy + nu * qz);
            double const xi[3] = {tu.gp[qx], tu.gp[qy], tu.gp[qz]};
            double       x[3], J[9], Ji[9], det;
            mesh.map_point(c, xi, x, J);
            invert3(J, Ji, det);
            double const w = det * tu.gw[qx] * tu.gw[qy] * tu.gw[qz];
            for(int j = 0; j < 3; ++j)
              g[j * nu3 + q] = -w * (Ji[j * 3 + 0] * uq[q] + Ji[j * 3 + 1] * uq[nu3 + q] +
                                     Ji[j * 3 + 2] * uq[2 * nu3 + q]);
          }
      for(int j = 0; j < 3; ++j)
        tensor_apply_T(nu, nu, nu, np, j == 0 ? Sdp.data() : Sp.data(), j == 1 ? Sdp.data() : Sp.data(),
                       j == 2 ? Sdp.data() : Sp.data(), g.data() + j * nu3, b + c * np3);
    }
  }

  // face terms at QGauss(k_u+1) face points; value of the velocity on a face from the cell's
  // nodal values (along dir the basis at 0 or 1, tangentially S), pressure tests likewise
  std::vector<double> u0(nu), u1(nu);
  for(int i = 0; i < nu; ++i)
  {
    u0[i] = lagrange(tu.gl, i, 0.0);
    u1[i] = lagrange(tu.gl, i, 1.0);
  }
  int const nf2 = nu * nu;
  auto face_values = [&](int64_t c, int dir, int side, double * val /* 3 x nu^2 */) {
    int m[3] = {nu, nu, nu};
    m[dir]   = 1;
    double const * ops[3];
    for(int k = 0; k < 3; ++k)
      ops[k] = (k == dir) ? (side ? u1.data() : u0.data()) : tu.S.data();
    for(int k = 0; k < 3; ++k)
      tensor_apply(m[0], m[1], m[2], nu, ops[0], ops[1], ops[2], u + (c * 3 + k) * nu3, val + k * nf2);
  };
  auto face_test = [&](int64_t c, int dir, int side, double const * f /* nu^2 */) {
    int m[3] = {nu, nu, nu};
    m[dir]   = 1;
    double const * ops[3];
    for(int k = 0; k < 3; ++k)
      ops[k] = (k == dir) ? (side ? p1.data() : p0.data()) : Sp.data();
    tensor_apply_T(m[0], m[1], m[2], np, ops[0], ops[1], ops[2], f, b + c * np3);
  };
  // geometry of face (dir, side) of cell c at the velocity face points: JxW and outward normal
  auto face_geo = [&](int64_t c, int dir, int side, double * jxw, double * nrm_out) {
    int const t1 = dir == 0 ? 1 : 0, t2 = dir == 2 ? 1 : 2;
    for(int qb = 0; qb < nu; ++qb)
      for(int qa = 0; qa < nu; ++qa)
      {
        int const q = qa + nu * qb;
        double    xi[3];
        xi[dir] = side ? 1.0 : 0.0;
        xi[t1]  = tu.gp[qa];
        xi[t2]  = tu.gp[qb];
        double x[3], J[9], Ji[9], det;
        mesh.map_point(c, xi, x, J);
        invert3(J, Ji, det);
        double       nv[3];
        double const sgn = side ? 1.0 : -1.0;
        for(int i = 0; i < 3; ++i)
          nv[i] = sgn * Ji[dir * 3 + i];
        double const nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        for(int i = 0; i < 3; ++i)
          nrm_out[3 * q + i] = nv[i] / nn;
        jxw[q] = det * nn * tu.gw[qa] * tu.gw[qb];
      }
  };
  std::vector<double> vm(3 * nf2), vp(3 * nf2), jxw(nf2), nrm(3 * nf2), f(nf2), fneg(nf2);
  for(int64_t c = 0; c < nc; ++c)
    for(int dir = 0; dir < 3; ++dir)
    {
      // interior / periodic face above the cell, once: central flux, + to c, - to the neighbour
      int64_t const nb = mesh.neighbor(c, dir, 1);
      if(nb >= 0)
      {
        face_values(c, dir, 1, vm.data());
        face_values(nb, dir, 0, vp.data());
        face_geo(c, dir, 1, jxw.data(), nrm.data());
        for(int q = 0; q < nf2; ++q)
        {
          double s = 0;
          for(int k = 0; k < 3; ++k)
            s += 0.5 * (vm[k * nf2 + q] + vp[k * nf2 + q]) * nrm[3 * q + k];
          f[q]    = s * jxw[q];
          fneg[q] = -f[q];
        }
        face_test(c, dir, 1, f.data());
        face_test(nb, dir, 0, fneg.data());
      }
      // boundary faces of c
      for(int side = 0; side < 2; ++side)
      {
        if(mesh.neighbor(c, dir, side) >= 0)
          continue;
        if(mesh.d.boundary_type[2 * dir + side] != 0)
          continue;  // pressure Neumann: velocity Dirichlet at rest, u+ = -u-, no flux
        // pressure Dirichlet: velocity Neumann, u+ = u-, flux u-.n
        face_values(c, dir, side, vm.data());
        face_geo(c, dir, side, jxw.data(), nrm.data());
        for(int q = 0; q < nf2; ++q)
        {
          double s = 0;
          for(int k = 0; k < 3; ++k)
            s += vm[k * nf2 + q] * nrm[3 * q + k];
          f[q] = s * jxw[q];
        }
        face_test(c, dir, side, f.data());
      }
    }
}
}  // namespace oracle
