#include "grid.h"

#include <cmath>

namespace ExaDG
{
namespace
{
// out[a + ma (b + mb c)] = sum_{ijk} A[a][i] B[b][j] C[c][k] in[i + m (j + m k)]
void contract3(int ma, int mb, int mc, int m, double const * A, double const * B, double const * C,
               double const * in, double * out, std::vector<double> & w1, std::vector<double> & w2)
{
  if(mc == 1 || mb == 1)
  {
    // a face normal to z or y: contract that single point first (m^3), then the plane (m^2 per
    // output line) instead of carrying all m planes through the other contractions
    bool const   zf = mc == 1;
    int const    mo = zf ? mb : mc;           // the other tangential output count
    double const * R  = zf ? C : B;            // the single row of the face direction
    double const * T  = zf ? B : C;            // the tangential matrix of the other direction
    w1.resize((size_t)m * m);
    for(int o = 0; o < m; ++o)
      for(int i = 0; i < m; ++i)
      {
        double s = 0;
        for(int l = 0; l < m; ++l)
          s += R[l] * (zf ? in[i + m * (o + m * l)] : in[i + m * (l + m * o)]);
        w1[i + m * o] = s;
      }
    w2.resize((size_t)ma * m);
    for(int o = 0; o < m; ++o)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int i = 0; i < m; ++i)
          s += A[a * m + i] * w1[i + m * o];
        w2[a + ma * o] = s;
      }
    for(int b = 0; b < mo; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int o = 0; o < m; ++o)
          s += T[b * m + o] * w2[a + ma * o];
        out[a + ma * b] = s;  // (a, b) with the face direction's single index 0
      }
    return;
  }
  w1.resize((size_t)ma * m * m);
  w2.resize((size_t)ma * mb * m);
  for(int k = 0; k < m; ++k)
    for(int j = 0; j < m; ++j)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int i = 0; i < m; ++i)
          s += A[a * m + i] * in[i + m * (j + m * k)];
        w1[a + ma * (j + m * k)] = s;
      }
  for(int k = 0; k < m; ++k)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int j = 0; j < m; ++j)
          s += B[b * m + j] * w1[a + ma * (j + m * k)];
        w2[a + ma * (b + mb * k)] = s;
      }
  for(int c = 0; c < mc; ++c)
    for(int b = 0; b < mb; ++b)
      for(int a = 0; a < ma; ++a)
      {
        double s = 0;
        for(int k = 0; k < m; ++k)
          s += C[c * m + k] * w2[a + ma * (b + mb * k)];
        out[a + ma * (b + mb * c)] = s;
      }
}

double inverse3(double const * J, double * Ji)
{
  double const c00 = J[4] * J[8] - J[5] * J[7];
  double const c01 = J[5] * J[6] - J[3] * J[8];
  double const c02 = J[3] * J[7] - J[4] * J[6];
  double const det = J[0] * c00 + J[1] * c01 + J[2] * c02;
  double const r   = 1.0 / det;
  Ji[0]            = c00 * r;
  Ji[3]            = c01 * r;
  Ji[6]            = c02 * r;
  Ji[1]            = (J[2] * J[7] - J[1] * J[8]) * r;
  Ji[4]            = (J[0] * J[8] - J[2] * J[6]) * r;
  Ji[7]            = (J[1] * J[6] - J[0] * J[7]) * r;
  Ji[2]            = (J[1] * J[5] - J[2] * J[4]) * r;
  Ji[5]            = (J[2] * J[3] - J[0] * J[5]) * r;
  Ji[8]            = (J[0] * J[4] - J[1] * J[3]) * r;
  return det;
}
}  // namespace

StructuredGrid::StructuredGrid(exadg_mesh_desc const & mesh, exadg_bc_desc const & bc)
  : mesh_(mesh), bc_(bc)
{
  for(int d = 0; d < 3; ++d)
  {
    if(mesh.n_cells[d] < 1)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "n_cells must be >= 1 in every direction");
    if(!(mesh.upper[d] > mesh.lower[d]))
      EXADG_THROW(EXADG_ERR_ARGUMENT, "upper must exceed lower in every direction");
    topo_.nc[d]       = mesh.n_cells[d];
    topo_.periodic[d] = mesh.periodic[d] ? 1 : 0;
    topo_.layers[d]   = mesh.n_cells[d] + (topo_.periodic[d] ? 0 : 1);
    h_[d]             = (mesh.upper[d] - mesh.lower[d]) / mesh.n_cells[d];
  }
  for(int s = 0; s < 6; ++s)
  {
    if(bc.type[s] != EXADG_BC_DIRICHLET && bc.type[s] != EXADG_BC_NEUMANN)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "boundary type must be Dirichlet or Neumann");
    topo_.bc[s] = bc.type[s];
  }
  int64_t base = 0;
  for(int d = 0; d < 3; ++d)
  {
    topo_.face_base[d] = base;
    base += (int64_t)topo_.layers[d] * topo_.nc[tangent1(d)] * topo_.nc[tangent2(d)];
  }
  topo_.n_faces = base;
  n_cells_      = (int64_t)mesh.n_cells[0] * mesh.n_cells[1] * mesh.n_cells[2];
  affine_       = mesh.deformation == 0.0;
  if(!affine_)
  {
    if(mesh.mapping_degree < 1 || mesh.mapping_degree > 15)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "mapping_degree must be in [1,15] for deformed meshes");
    gauss_lobatto_points(mesh.mapping_degree + 1, map_nodes_);
  }
}

StructuredGrid StructuredGrid::coarsened(int times) const
{
  exadg_mesh_desc m = mesh_;
  for(int d = 0; d < 3; ++d)
  {
    if(times < 0 || m.n_cells[d] % (1 << times) != 0)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "h-levels need cells-per-direction counts divisible by 2 "
                                      "for every coarsening");
    m.n_cells[d] >>= times;
  }
  return StructuredGrid(m, bc_);
}

void StructuredGrid::push_forward(double const * p, double * x) const
{
  // DeformedCubeManifold::push_forward (deformed_cube_manifold.h:47-59)
  double s = mesh_.deformation;
  for(int e = 0; e < 3; ++e)
    s *= std::sin(mesh_.frequency * M_PI * (p[e] - mesh_.lower[e]) / (mesh_.upper[e] - mesh_.lower[e]));
  for(int e = 0; e < 3; ++e)
    x[e] = p[e] + s;
}

void StructuredGrid::evaluate(int64_t c, std::vector<double> const & px,
                              std::vector<double> const & py, std::vector<double> const & pz,
                              double * x, double * J) const
{
  int cc[3];
  cell_coords(topo_, c, cc);
  int const                         mx = (int)px.size(), my = (int)py.size(), mz = (int)pz.size();
  size_t const                      np = (size_t)mx * my * mz;
  std::vector<double> const * const pts[3] = {&px, &py, &pz};
  if(affine_)
  {
    for(int c2 = 0; c2 < mz; ++c2)
      for(int b = 0; b < my; ++b)
        for(int a = 0; a < mx; ++a)
        {
          size_t const q     = a + (size_t)mx * (b + (size_t)my * c2);
          int const    id[3] = {a, b, c2};
          for(int i = 0; i < 3; ++i)
          {
            x[i * np + q] = mesh_.lower[i] + h_[i] * (cc[i] + (*pts[i])[id[i]]);
            for(int j = 0; j < 3; ++j)
              J[(i * 3 + j) * np + q] = (i == j) ? h_[i] : 0.0;
          }
        }
    return;
  }
  // MappingQ: support points are push_forward of the chart-space GL points of the cell
  // (vertices and all higher-order support points of a ChartManifold cell). The deformation
  // factor of push_forward is a product of one sine per direction, so the m^3 points need 3m
  // sines; the factors are multiplied in push_forward's order, so the points are bit-identical
  // to calling it per point.
  int const                  m = (int)map_nodes_.size();
  thread_local std::vector<double> X;
  X.resize(3 * (size_t)m * m * m);
  double pc[3][16], sn[3][16];
  for(int e = 0; e < 3; ++e)
    for(int i = 0; i < m; ++i)
    {
      pc[e][i] = mesh_.lower[e] + h_[e] * (cc[e] + map_nodes_[i]);
      sn[e][i] = std::sin(mesh_.frequency * M_PI * (pc[e][i] - mesh_.lower[e]) /
                          (mesh_.upper[e] - mesh_.lower[e]));
    }
  for(int k = 0; k < m; ++k)
    for(int j = 0; j < m; ++j)
      for(int i = 0; i < m; ++i)
      {
        double       f = mesh_.deformation;
        f *= sn[0][i];
        f *= sn[1][j];
        f *= sn[2][k];
        size_t const s = i + (size_t)m * (j + (size_t)m * k);
        X[s]                         = pc[0][i] + f;
        X[(size_t)m * m * m + s]     = pc[1][j] + f;
        X[2 * (size_t)m * m * m + s] = pc[2][k] + f;
      }
  thread_local std::vector<double> V[3], D[3];
  for(int d = 0; d < 3; ++d)
  {
    int const mp = (int)pts[d]->size();
    V[d].resize((size_t)mp * m);
    D[d].resize((size_t)mp * m);
    for(int a = 0; a < mp; ++a)
      lagrange_basis(map_nodes_, (*pts[d])[a], &V[d][(size_t)a * m], &D[d][(size_t)a * m]);
  }
  thread_local std::vector<double> w1, w2;
  for(int e = 0; e < 3; ++e)
  {
    double const * Xe = &X[e * (size_t)m * m * m];
    contract3(mx, my, mz, m, V[0].data(), V[1].data(), V[2].data(), Xe, x + e * np, w1, w2);
    contract3(mx, my, mz, m, D[0].data(), V[1].data(), V[2].data(), Xe, J + (e * 3 + 0) * np, w1, w2);
    contract3(mx, my, mz, m, V[0].data(), D[1].data(), V[2].data(), Xe, J + (e * 3 + 1) * np, w1, w2);
    contract3(mx, my, mz, m, V[0].data(), V[1].data(), D[2].data(), Xe, J + (e * 3 + 2) * np, w1, w2);
  }
}

GeometryData compute_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape,
                              double ip_factor, bool store_metric)
{
  GeometryData        g;
  int const           n  = shape.n;
  int const           n2 = n * n, n3 = n * n * n;
  int64_t const       nc = grid.n_cells();
  BoxTopology const & topo = grid.topology();
  g.n                      = n;
  g.tau_cell.resize(nc);
  if(store_metric)
    g.cell_metric.resize((size_t)nc * 6 * n3);

  std::vector<double> const one0 = {0.0}, one1 = {1.0};

  // cells: metric G = JxW J^{-1} J^{-T} and the penalty tau_K = sum_f c_f |f| / |K|
  // (interior_penalty_parameter.h:43-99 with QGauss(k+1) on cells and faces)
  parallel_range(nc, [&](int64_t b, int64_t e) {
    std::vector<double> x(3 * n3), J(9 * n3), xf(3 * n2), Jf(9 * n2);
    for(int64_t c = b; c < e; ++c)
    {
      grid.evaluate(c, shape.qp, shape.qp, shape.qp, x.data(), J.data());
      double vol = 0;
      for(int qz = 0; qz < n; ++qz)
        for(int qy = 0; qy < n; ++qy)
          for(int qx = 0; qx < n; ++qx)
          {
            int const q = qx + n * (qy + n * qz);
            double    Jq[9], Ji[9];
            for(int k = 0; k < 9; ++k)
              Jq[k] = J[k * n3 + q];
            double const det = inverse3(Jq, Ji);
            double const jxw = det * shape.qw[qx] * shape.qw[qy] * shape.qw[qz];
            vol += jxw;
            if(store_metric)
            {
              double * G = &g.cell_metric[(size_t)c * 6 * n3];
              auto     sym = [&](int a, int bb) {
                return jxw * (Ji[a * 3 + 0] * Ji[bb * 3 + 0] + Ji[a * 3 + 1] * Ji[bb * 3 + 1] +
                              Ji[a * 3 + 2] * Ji[bb * 3 + 2]);
              };
              G[0 * n3 + q] = sym(0, 0);
              G[1 * n3 + q] = sym(1, 1);
              G[2 * n3 + q] = sym(2, 2);
              G[3 * n3 + q] = sym(0, 1);
              G[4 * n3 + q] = sym(0, 2);
              G[5 * n3 + q] = sym(1, 2);
            }
          }
      int cc[3];
      cell_coords(topo, c, cc);
      double surf = 0;
      for(int d = 0; d < 3; ++d)
        for(int s = 0; s < 2; ++s)
        {
          std::vector<double> const * p[3] = {&shape.qp, &shape.qp, &shape.qp};
          p[d]                             = s ? &one1 : &one0;
          grid.evaluate(c, *p[0], *p[1], *p[2], xf.data(), Jf.data());
          double const factor = neighbor_id(topo, cc, d, s) < 0 ? 1.0 : 0.5;
          int const    t1 = tangent1(d), t2 = tangent2(d);
          (void)t2;
          for(int q = 0; q < n2; ++q)
          {
            double Jq[9], Ji[9];
            for(int k = 0; k < 9; ++k)
              Jq[k] = Jf[k * n2 + q];
            double const det = inverse3(Jq, Ji);
            double const nrm =
              std::sqrt(Ji[d * 3 + 0] * Ji[d * 3 + 0] + Ji[d * 3 + 1] * Ji[d * 3 + 1] + Ji[d * 3 + 2] * Ji[d * 3 + 2]);
            int const qa = q % n, qb = q / n;
            surf += factor * det * nrm * shape.qw[qa] * shape.qw[qb];
          }
          (void)t1;
        }
      g.tau_cell[c] = surf / vol;
    }
  });

  // faces: tau_F and (curved) the normal-derivative vectors of both sides + JxW
  double const pf = ip_factor * n * n;  // IP::get_penalty_factor: IP (k+1)^2
  g.tau_face.resize(topo.n_faces);
  if(store_metric)
    g.face_data.resize((size_t)topo.n_faces * 7 * n2);
  for(int d = 0; d < 3; ++d)
  {
    int const     t1 = tangent1(d), t2 = tangent2(d);
    int64_t const nf = (int64_t)topo.layers[d] * topo.nc[t1] * topo.nc[t2];
    parallel_range(nf, [&](int64_t b, int64_t e) {
      std::vector<double> xm(3 * n2), Jm(9 * n2), xp(3 * n2), Jp(9 * n2);
      for(int64_t fi = b; fi < e; ++fi)
      {
        int const     fl = (int)(fi % topo.layers[d]);
        int64_t const r  = fi / topo.layers[d];
        int           cc[3];
        cc[t1]              = (int)(r % topo.nc[t1]);
        cc[t2]              = (int)(r / topo.nc[t1]);
        int64_t const f     = topo.face_base[d] + fi;
        bool const    has_m = topo.periodic[d] || fl > 0;
        bool const    has_p = fl < topo.nc[d];
        int           cm[3] = {cc[0], cc[1], cc[2]}, cp[3] = {cc[0], cc[1], cc[2]};
        cm[d]               = (fl == 0) ? topo.nc[d] - 1 : fl - 1;
        cp[d]               = fl;
        int64_t const im = has_m ? cell_id(topo, cm) : -1;
        int64_t const ip = has_p ? cell_id(topo, cp) : -1;
        if(has_m && has_p)
          g.tau_face[f] = pf * std::max(g.tau_cell[im], g.tau_cell[ip]);
        else
          g.tau_face[f] = pf * g.tau_cell[has_m ? im : ip];
        if(!store_metric)
          continue;
        std::vector<double> const * p[3] = {&shape.qp, &shape.qp, &shape.qp};
        if(has_m)
        {
          p[d] = &one1;
          grid.evaluate(im, *p[0], *p[1], *p[2], xm.data(), Jm.data());
        }
        if(has_p)
        {
          p[d] = &one0;
          grid.evaluate(ip, *p[0], *p[1], *p[2], xp.data(), Jp.data());
        }
        double * out = &g.face_data[(size_t)f * 7 * n2];
        for(int q = 0; q < n2; ++q)
        {
          double Jq[9], Jim[9], Jip[9], detm = 0, detp = 0;
          if(has_m)
          {
            for(int k = 0; k < 9; ++k)
              Jq[k] = Jm[k * n2 + q];
            detm = inverse3(Jq, Jim);
          }
          if(has_p)
          {
            for(int k = 0; k < 9; ++k)
              Jq[k] = Jp[k * n2 + q];
            detp = inverse3(Jq, Jip);
          }
          double const * Jr  = has_m ? Jim : Jip;
          double const   det = has_m ? detm : detp;
          double         nv[3] = {Jr[d * 3 + 0], Jr[d * 3 + 1], Jr[d * 3 + 2]};
          double const   nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          for(int i = 0; i < 3; ++i)
            nv[i] /= nrm;
          int const qa = q % n, qb = q / n;
          for(int j = 0; j < 3; ++j)
          {
            out[j * n2 + q] =
              has_m ? Jim[j * 3 + 0] * nv[0] + Jim[j * 3 + 1] * nv[1] + Jim[j * 3 + 2] * nv[2] : 0.0;
            out[(3 + j) * n2 + q] =
              has_p ? Jip[j * 3 + 0] * nv[0] + Jip[j * 3 + 1] * nv[1] + Jip[j * 3 + 2] * nv[2] : 0.0;
          }
          out[6 * n2 + q] = det * nrm * shape.qw[qa] * shape.qw[qb];
        }
      }
    });
  }
  return g;
}

DivergenceGeometry compute_divergence_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape)
{
  DivergenceGeometry  g;
  int const           n = shape.n, n2 = n * n, n3 = n2 * n;
  int64_t const       nc = grid.n_cells();
  BoxTopology const & topo = grid.topology();
  g.cell.resize((size_t)nc * 12 * n3);
  g.face.resize((size_t)topo.n_faces * 6 * n2);
  parallel_range(nc, [&](int64_t b, int64_t e) {
    std::vector<double> x(3 * n3), J(9 * n3);
    for(int64_t c = b; c < e; ++c)
    {
      grid.evaluate(c, shape.qp, shape.qp, shape.qp, x.data(), J.data());
      double * out = &g.cell[(size_t)c * 12 * n3];
      for(int qz = 0; qz < n; ++qz)
        for(int qy = 0; qy < n; ++qy)
          for(int qx = 0; qx < n; ++qx)
          {
            int const q = qx + n * (qy + n * qz);
            double    Jq[9], Ji[9];
            for(int k = 0; k < 9; ++k)
              Jq[k] = J[k * n3 + q];
            double const jxw = inverse3(Jq, Ji) * shape.qw[qx] * shape.qw[qy] * shape.qw[qz];
            for(int d = 0; d < 3; ++d)
              out[d * n3 + q] = x[d * n3 + q];
            for(int k = 0; k < 9; ++k)
              out[(3 + k) * n3 + q] = jxw * Ji[k];
          }
    }
  });
  std::vector<double> const one0 = {0.0}, one1 = {1.0};
  for(int d = 0; d < 3; ++d)
  {
    int const     t1 = tangent1(d), t2 = tangent2(d);
    int64_t const nf = (int64_t)topo.layers[d] * topo.nc[t1] * topo.nc[t2];
    parallel_range(nf, [&](int64_t b, int64_t e) {
      std::vector<double> xf(3 * n2), Jf(9 * n2);
      for(int64_t fi = b; fi < e; ++fi)
      {
        int const     fl = (int)(fi % topo.layers[d]);
        int64_t const r  = fi / topo.layers[d];
        int           cc[3];
        cc[t1]           = (int)(r % topo.nc[t1]);
        cc[t2]           = (int)(r / topo.nc[t1]);
        bool const has_m = topo.periodic[d] || fl > 0;
        // evaluate from the cell below the face (xi_d = 1), or above it on the lower boundary
        cc[d] = has_m ? ((fl == 0) ? topo.nc[d] - 1 : fl - 1) : fl;
        std::vector<double> const * p[3] = {&shape.qp, &shape.qp, &shape.qp};
        p[d]                             = has_m ? &one1 : &one0;
        grid.evaluate(cell_id(topo, cc), *p[0], *p[1], *p[2], xf.data(), Jf.data());
        double * out = &g.face[(size_t)(topo.face_base[d] + fi) * 6 * n2];
        for(int q = 0; q < n2; ++q)
        {
          double Jq[9], Ji[9];
          for(int k = 0; k < 9; ++k)
            Jq[k] = Jf[k * n2 + q];
          double const det = inverse3(Jq, Ji);
          double       nv[3] = {Ji[d * 3 + 0], Ji[d * 3 + 1], Ji[d * 3 + 2]};
          double const nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          int const    qa = q % n, qb = q / n;
          double const jxw = det * nrm * shape.qw[qa] * shape.qw[qb];
          for(int k = 0; k < 3; ++k)
          {
            out[k * n2 + q]       = xf[k * n2 + q];
            out[(3 + k) * n2 + q] = jxw * nv[k] / nrm;
          }
        }
      }
    });
  }
  return g;
}

}  // namespace ExaDG

namespace ExaDG
{
RhsGeometry compute_rhs_geometry(StructuredGrid const & grid, ShapeInfo1D const & shape,
                                 std::vector<double> const & tau_face)
{
  RhsGeometry         g;
  int const           n = shape.n, n2 = n * n, n3 = n2 * n;
  int64_t const       nc = grid.n_cells();
  BoxTopology const & topo = grid.topology();
  g.cell.resize((size_t)nc * kRhsCellRec * n3);
  parallel_range(nc, [&](int64_t b, int64_t e) {
    std::vector<double> x(3 * n3), J(9 * n3);
    for(int64_t c = b; c < e; ++c)
    {
      grid.evaluate(c, shape.qp, shape.qp, shape.qp, x.data(), J.data());
      double * out = &g.cell[(size_t)c * kRhsCellRec * n3];
      for(int q = 0; q < n3; ++q)
      {
        double Jq[9], Ji[9];
        for(int k = 0; k < 9; ++k)
          Jq[k] = J[k * n3 + q];
        double const det = inverse3(Jq, Ji);
        for(int d = 0; d < 3; ++d)
          out[d * n3 + q] = x[d * n3 + q];
        out[3 * n3 + q] = std::abs(det) * shape.qw[q % n] * shape.qw[(q / n) % n] * shape.qw[q / n2];
      }
    }
  });
  // boundary faces, side-major; on a side lexicographic in (t1, t2)
  g.side_base[0] = 0;
  for(int side = 0; side < 6; ++side)
  {
    int const d = side / 2;
    g.side_base[side + 1] =
      g.side_base[side] + (topo.periodic[d] ? 0 : (int64_t)topo.nc[tangent1(d)] * topo.nc[tangent2(d)]);
  }
  g.face.resize((size_t)g.side_base[6] * kRhsFaceRec * n2);
  std::vector<double> const one0 = {0.0}, one1 = {1.0};
  for(int side = 0; side < 6; ++side)
  {
    int const     d = side / 2, s = side % 2, t1 = tangent1(d), t2 = tangent2(d);
    int64_t const nf = g.side_base[side + 1] - g.side_base[side];
    parallel_range(nf, [&](int64_t b, int64_t e) {
      std::vector<double> xf(3 * n2), Jf(9 * n2);
      for(int64_t fi = b; fi < e; ++fi)
      {
        int cc[3];
        cc[t1] = (int)(fi % topo.nc[t1]);
        cc[t2] = (int)(fi / topo.nc[t1]);
        cc[d]  = s ? topo.nc[d] - 1 : 0;
        std::vector<double> const * p[3] = {&shape.qp, &shape.qp, &shape.qp};
        p[d]                             = s ? &one1 : &one0;
        grid.evaluate(cell_id(topo, cc), *p[0], *p[1], *p[2], xf.data(), Jf.data());
        double const tau = tau_face[face_id(topo, cc, d, s)];
        double *     out = &g.face[(size_t)(g.side_base[side] + fi) * kRhsFaceRec * n2];
        for(int q = 0; q < n2; ++q)
        {
          double Jq[9], Ji[9];
          for(int k = 0; k < 9; ++k)
            Jq[k] = Jf[k * n2 + q];
          double const det = inverse3(Jq, Ji);
          // outward unit normal: the gradient of xi_d, oriented away from the cell
          double       nv[3] = {Ji[d * 3 + 0], Ji[d * 3 + 1], Ji[d * 3 + 2]};
          double const nrm = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          for(int k = 0; k < 3; ++k)
            nv[k] *= (s ? 1.0 : -1.0) / nrm;
          for(int k = 0; k < 3; ++k)
          {
            out[k * n2 + q]       = xf[k * n2 + q];
            out[(3 + k) * n2 + q] = nv[k];
            out[(7 + k) * n2 + q] = Ji[k * 3 + 0] * nv[0] + Ji[k * 3 + 1] * nv[1] + Ji[k * 3 + 2] * nv[2];
          }
          out[6 * n2 + q]  = std::abs(det) * nrm * shape.qw[q % n] * shape.qw[q / n];
          out[10 * n2 + q] = tau;
        }
      }
    });
  }
  return g;
}
}  // namespace ExaDG
