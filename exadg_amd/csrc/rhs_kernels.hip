// Right-hand side assembly (K9): one workgroup per cell. The volume data f JxW at the Gauss
// points goes back to the nodal basis by three transposed S sweeps in LDS; each boundary face
// of the cell adds its value test (the face's end node layer) and its gradient test (the end
// derivative row along the face normal direction, or the tangential derivative S_d), from the
// per-point records. Every b entry is written once.
#include <hip/hip_runtime.h>

#include "../../include/exadg_amd.h"
#include "rhs_kernels.h"

namespace ExaDG
{
namespace
{
constexpr int kThreads = 128;

__global__ void __launch_bounds__(kThreads) rhs_assemble_kernel(RhsArgs a)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int const      n = a.n, n2 = n * n, n3 = n2 * n;
  double *       t   = reinterpret_cast<double *>(smem_raw);  // n^3 quadrature data
  double *       tmp = t + n3;                                 // n^3
  double *       o   = tmp + n3;                               // n^3 result
  double *       rv  = o + n3;                                 // n^2 face value test
  double *       rg  = rv + n2;                                // 3 n^2 face gradient test
  double const * S   = a.tables;
  double const * Sd  = S + n2;
  double const * d0  = Sd + n2;
  double const * d1  = d0 + n;
  int64_t const  c   = blockIdx.x;
  double const   om  = a.omega;

  double const * cg = a.cell_geo + c * kRhsCellRec * n3;
  for(int q = threadIdx.x; q < n3; q += kThreads)
  {
    double f = 0;
    if(a.sine)
      f = 3.0 * om * om * sin(om * cg[q]) * sin(om * cg[n3 + q]) * sin(om * cg[2 * n3 + q]);
    else if(a.f)
      f = a.f[c * n3 + q];
    t[q] = f * cg[3 * n3 + q];
  }
  __syncthreads();
  // o = (S^T x S^T x S^T) t: z, then y, then x
  for(int e = threadIdx.x; e < n3; e += kThreads)
  {
    int const i = e % n, j = (e / n) % n, k = e / n2;
    double    s = 0;
    for(int q = 0; q < n; ++q)
      s += S[q * n + k] * t[i + n * (j + n * q)];
    tmp[e] = s;
  }
  __syncthreads();
  for(int e = threadIdx.x; e < n3; e += kThreads)
  {
    int const i = e % n, j = (e / n) % n, k = e / n2;
    double    s = 0;
    for(int q = 0; q < n; ++q)
      s += S[q * n + j] * tmp[i + n * (q + n * k)];
    t[e] = s;
  }
  __syncthreads();
  for(int e = threadIdx.x; e < n3; e += kThreads)
  {
    int const i = e % n, j = (e / n) % n, k = e / n2;
    double    s = 0;
    for(int q = 0; q < n; ++q)
      s += S[q * n + i] * t[q + n * (j + n * k)];
    o[e] = s;
  }
  __syncthreads();

  int cc[3];
  cell_coords(a.topo, c, cc);
  for(int d = 0; d < 3; ++d)
    for(int sd = 0; sd < 2; ++sd)
    {
      if(neighbor_id(a.topo, cc, d, sd) >= 0)
        continue;  // interior or periodic face: no boundary data
      int const      t1 = tangent1(d), t2 = tangent2(d);
      int64_t const  bf = a.side_base[2 * d + sd] + cc[t1] + (int64_t)a.topo.nc[t1] * cc[t2];
      double const * fg = a.face_geo + bf * kRhsFaceRec * n2;
      bool const     dirichlet = a.topo.bc[2 * d + sd] == EXADG_BC_DIRICHLET;
      for(int q = threadIdx.x; q < n2; q += kThreads)
      {
        double const x = fg[q], y = fg[n2 + q], z = fg[2 * n2 + q];
        double const w = fg[6 * n2 + q], tau = fg[10 * n2 + q];
        double       v = 0;
        if(a.sine)
        {
          double const s0 = sin(om * x), s1 = sin(om * y), s2 = sin(om * z);
          if(dirichlet)
            v = s0 * s1 * s2;
          else
            v = om * (cos(om * x) * s1 * s2 * fg[3 * n2 + q] + s0 * cos(om * y) * s2 * fg[4 * n2 + q] +
                      s0 * s1 * cos(om * z) * fg[5 * n2 + q]);
        }
        else if(a.boundary)
          v = a.boundary[bf * n2 + q];
        if(dirichlet)
        {  // int (-g dv/dn + 2 tau g v)
          rv[q] = 2.0 * tau * v * w;
          for(int j = 0; j < 3; ++j)
            rg[j * n2 + q] = -v * w * fg[(7 + j) * n2 + q];
        }
        else
        {  // int h v
          rv[q] = v * w;
          for(int j = 0; j < 3; ++j)
            rg[j * n2 + q] = 0.0;
        }
      }
      __syncthreads();
      double const * drow = sd ? d1 : d0;
      int const      iend = sd ? n - 1 : 0;
      for(int e = threadIdx.x; e < n3; e += kThreads)
      {
        int const ii[3] = {e % n, (e / n) % n, e / n2};
        int const id = ii[d], i1 = ii[t1], i2 = ii[t2];
        double    acc = 0;
        // value test: the end node layer along d, S tangentially
        if(id == iend)
          for(int q2 = 0; q2 < n; ++q2)
          {
            double s = 0;
            for(int q1 = 0; q1 < n; ++q1)
              s += S[q1 * n + i1] * rv[q1 + n * q2];
            acc += S[q2 * n + i2] * s;
          }
        // gradient test component j
        if(dirichlet)
          for(int j = 0; j < 3; ++j)
          {
            double const cd = (j == d) ? drow[id] : (id == iend ? 1.0 : 0.0);
            if(cd == 0.0)
              continue;
            double const * T1 = (j == t1) ? Sd : S;
            double const * T2 = (j == t2) ? Sd : S;
            double         sj = 0;
            for(int q2 = 0; q2 < n; ++q2)
            {
              double s = 0;
              for(int q1 = 0; q1 < n; ++q1)
                s += T1[q1 * n + i1] * rg[j * n2 + q1 + n * q2];
              sj += T2[q2 * n + i2] * s;
            }
            acc += cd * sj;
          }
        o[e] += acc;
      }
      __syncthreads();
    }
  for(int e = threadIdx.x; e < n3; e += kThreads)
    a.b[c * n3 + e] = o[e];
}
}  // namespace

hipError_t launch_rhs_assemble(RhsArgs const & a, hipStream_t s)
{
  if(a.n < 2 || a.n > 9 || a.n_cells < 1 || a.b == nullptr || a.cell_geo == nullptr ||
     (a.side_base[6] > 0 && a.face_geo == nullptr))
    return hipErrorInvalidValue;
  int const    n   = a.n;
  size_t const lds = sizeof(double) * (3 * n * n * n + 4 * n * n);  // <= 23.6 KB at n = 9
  hipLaunchKernelGGL(rhs_assemble_kernel, dim3((unsigned)a.n_cells), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace ExaDG
