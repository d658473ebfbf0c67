// Weak DG divergence of a discrete velocity, one workgroup per cell (divergence_operator.cpp
// :168-298, weak form). The velocity goes to the Gauss points of QGauss(k_u+1) by three S
// sweeps per component in LDS, the reference-gradient data -C u (C = JxW J^{-1}) is tested with
// the pressure basis at those points, and each face adds the central flux of the own end node
// layer and the neighbour's opposite layer, interpolated to the face points. Element-centric:
// every face is evaluated from both cells and every pressure entry is written once.
#include <hip/hip_runtime.h>

#include "../../include/exadg_amd.h"
#include "divergence_kernels.h"

namespace ExaDG
{
namespace
{
constexpr int kThreads = 256;

// out[i + m0 (j + m1 k)] for in[a + n0 (b + n1 c)] with one matrix along one direction:
// along direction `dir` apply M (rows: output index, m_out x m_in, M[o * m_in + a]), the
// other two directions copied
__device__ void sweep(double const * M, int m_out, int m_in, int dir, int const (&ext_in)[3],
                      double const * in, double * out, bool transpose)
{
  int ext_out[3] = {ext_in[0], ext_in[1], ext_in[2]};
  ext_out[dir]   = m_out;
  int const total = ext_out[0] * ext_out[1] * ext_out[2];
  for(int e = threadIdx.x; e < total; e += kThreads)
  {
    int idx[3] = {e % ext_out[0], (e / ext_out[0]) % ext_out[1], e / (ext_out[0] * ext_out[1])};
    int const o = idx[dir];
    double    s = 0;
    for(int a = 0; a < m_in; ++a)
    {
      idx[dir]       = a;
      double const m = transpose ? M[a * m_out + o] : M[o * m_in + a];
      s += m * in[idx[0] + ext_in[0] * (idx[1] + ext_in[1] * idx[2])];
    }
    out[e] = s;
  }
}

__global__ void __launch_bounds__(kThreads) divergence_apply_kernel(DivergenceArgs a)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int const      nu = a.n_u, np = a.n_p, nu2 = nu * nu, nu3 = nu2 * nu, np3 = np * np * np;
  double *       U   = reinterpret_cast<double *>(smem_raw);  // 3 nu^3 velocity, then Gauss data
  double *       T0  = U + 3 * nu3;                            // nu^3
  double *       T1  = T0 + nu3;                               // nu^3
  double *       O   = T1 + nu3;                               // np^3 result
  double *       Fm  = O + np3;                                // 3 nu^2 own face values
  double *       Fp  = Fm + 3 * nu2;                           // 3 nu^2 neighbour face values
  double *       Ft  = Fp + 3 * nu2;                           // nu^2 face flux
  double *       Fs  = Ft + nu2;                               // 6 nu^2 face scratch
  double const * Su  = a.tables;
  double const * Sp  = Su + nu2;
  double const * Sdp = Sp + nu * np;
  int64_t const  c   = blockIdx.x;
  int            cc[3];
  cell_coords(a.topo, c, cc);

  for(int e = threadIdx.x; e < 3 * nu3; e += kThreads)
    U[e] = a.src[c * 3 * nu3 + e];
  __syncthreads();
  int const full[3] = {nu, nu, nu};
  // velocity at the Gauss points, component by component (in place through T0, T1)
  for(int k = 0; k < 3; ++k)
  {
    sweep(Su, nu, nu, 0, full, U + k * nu3, T0, false);
    __syncthreads();
    sweep(Su, nu, nu, 1, full, T0, T1, false);
    __syncthreads();
    sweep(Su, nu, nu, 2, full, T1, T0, false);
    __syncthreads();
    for(int e = threadIdx.x; e < nu3; e += kThreads)
      U[k * nu3 + e] = T0[e];
    __syncthreads();
  }
  // F_j = -sum_k C[j][k] u_k (reference-gradient data), overwriting U
  double const * cg = a.cell_geo + c * 12 * nu3;
  for(int q = threadIdx.x; q < nu3; q += kThreads)
  {
    double const u0 = U[q], u1 = U[nu3 + q], u2 = U[2 * nu3 + q];
    double       f[3];
    for(int j = 0; j < 3; ++j)
      f[j] = -(cg[(3 + 3 * j) * nu3 + q] * u0 + cg[(4 + 3 * j) * nu3 + q] * u1 +
               cg[(5 + 3 * j) * nu3 + q] * u2);
    U[q] = f[0], U[nu3 + q] = f[1], U[2 * nu3 + q] = f[2];
  }
  for(int e = threadIdx.x; e < np3; e += kThreads)
    O[e] = 0.0;
  __syncthreads();
  // O += sum_j (D_j)^T F_j: along direction j the derivative table, otherwise values
  for(int j = 0; j < 3; ++j)
  {
    int       ext[3] = {nu, nu, nu};
    double *  in     = U + j * nu3;
    double *  bufs[2] = {T0, T1};
    double *  cur    = in;
    for(int d = 0; d < 3; ++d)
    {
      double * out = (d == 2) ? T0 : bufs[(d + 1) % 2];
      if(d == 2 && cur == T0)
        out = T1;
      sweep(d == j ? Sdp : Sp, np, nu, d, ext, cur, out, true);
      ext[d] = np;
      __syncthreads();
      cur = out;
    }
    for(int e = threadIdx.x; e < np3; e += kThreads)
      O[e] += cur[e];
    __syncthreads();
  }

  // faces
  for(int d = 0; d < 3; ++d)
    for(int sd = 0; sd < 2; ++sd)
    {
      int64_t const nb = neighbor_id(a.topo, cc, d, sd);
      bool const    interior = nb >= 0;
      if(!interior && a.topo.bc[2 * d + sd] != EXADG_BC_DIRICHLET)
        continue;  // wall at rest: u+ = -u-, the central flux vanishes
      int const t1 = tangent1(d), t2 = tangent2(d);
      // end node layers: own at sd, neighbour at 1 - sd; indices (t1 fastest) -> Fm / Fp
      for(int e = threadIdx.x; e < 3 * nu2; e += kThreads)
      {
        int const k = e / nu2, r = e % nu2, i1 = r % nu, i2 = r / nu;
        int       idx[3];
        idx[t1]   = i1;
        idx[t2]   = i2;
        idx[d]    = sd ? nu - 1 : 0;
        int const own = idx[0] + nu * (idx[1] + nu * idx[2]);
        Fm[e]     = a.src[c * 3 * nu3 + k * nu3 + own];
        if(interior)
        {
          idx[d]      = sd ? 0 : nu - 1;
          int const o2 = idx[0] + nu * (idx[1] + nu * idx[2]);
          Fp[e]        = a.src[nb * 3 * nu3 + k * nu3 + o2];
        }
      }
      __syncthreads();
      // tangential interpolation of the 3 (or 6) layers: along t1, then t2
      int const nl = interior ? 6 : 3;
      for(int e = threadIdx.x; e < nl * nu2; e += kThreads)
      {
        int const  l = e / nu2, r = e % nu2, q1 = r % nu, i2 = r / nu;
        double *   L = (l < 3 ? Fm + l * nu2 : Fp + (l - 3) * nu2);
        double     s = 0;
        for(int i1 = 0; i1 < nu; ++i1)
          s += Su[q1 * nu + i1] * L[i1 + nu * i2];
        Fs[e] = s;
      }
      __syncthreads();
      for(int e = threadIdx.x; e < nl * nu2; e += kThreads)
      {
        int const l = e / nu2, r = e % nu2, q1 = r % nu, q2 = r / nu;
        double    s = 0;
        for(int i2 = 0; i2 < nu; ++i2)
          s += Su[q2 * nu + i2] * Fs[l * nu2 + q1 + nu * i2];
        (l < 3 ? Fm + l * nu2 : Fp + (l - 3) * nu2)[q1 + nu * q2] = s;
      }
      __syncthreads();
      double const * fg = a.face_geo + face_id(a.topo, cc, d, sd) * 6 * nu2;
      double const   sg = sd ? 1.0 : -1.0;  // own outward normal = sg n
      for(int q = threadIdx.x; q < nu2; q += kThreads)
      {
        double f = 0;
        for(int k = 0; k < 3; ++k)
        {
          double const v = interior ? 0.5 * (Fm[k * nu2 + q] + Fp[k * nu2 + q]) : Fm[k * nu2 + q];
          f += v * fg[(3 + k) * nu2 + q];
        }
        Ft[q] = sg * f;
      }
      __syncthreads();
      // pressure test: the end node layer along d, Sp tangentially
      for(int e = threadIdx.x; e < np * np; e += kThreads)
      {
        int const i1 = e % np, i2 = e / np;
        double    s  = 0;
        for(int q2 = 0; q2 < nu; ++q2)
        {
          double t = 0;
          for(int q1 = 0; q1 < nu; ++q1)
            t += Sp[q1 * np + i1] * Ft[q1 + nu * q2];
          s += Sp[q2 * np + i2] * t;
        }
        int idx[3];
        idx[d]  = sd ? np - 1 : 0;
        idx[t1] = i1;
        idx[t2] = i2;
        O[idx[0] + np * (idx[1] + np * idx[2])] += s;
      }
      __syncthreads();
    }
  for(int e = threadIdx.x; e < np3; e += kThreads)
    a.dst[c * np3 + e] = O[e];
}

// one thread per node; per-thread tables e^{2 pi i m x_d}, m = 0..K, in LDS
constexpr int kInterpThreads = 128;

__global__ void __launch_bounds__(kInterpThreads)
  fourier_interpolate_kernel(double * u, double const * __restrict__ x, double const * __restrict__ md,
                             int n_modes, int K, int n3, int64_t n_points)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int const     nk = K + 1;
  double *      E  = reinterpret_cast<double *>(smem_raw) + threadIdx.x * 6 * nk;
  int64_t const p  = (int64_t)blockIdx.x * kInterpThreads + threadIdx.x;
  if(p >= n_points)
    return;
  for(int d = 0; d < 3; ++d)
  {
    double sn, cs;
    sincos(2.0 * M_PI * x[d * n_points + p], &sn, &cs);
    double * e = E + d * 2 * nk;
    e[0]       = 1.0;
    e[1]       = 0.0;
    for(int m = 1; m <= K; ++m)
    {
      e[2 * m]     = e[2 * m - 2] * cs - e[2 * m - 1] * sn;
      e[2 * m + 1] = e[2 * m - 2] * sn + e[2 * m - 1] * cs;
    }
  }
  double v[3] = {0, 0, 0};
  for(int m = 0; m < n_modes; ++m)
  {
    double const * q = md + 8 * m;
    int const      a0 = (int)q[0], a1 = (int)q[1], a2 = (int)q[2];
    double const * ex = E + 2 * (a0 < 0 ? -a0 : a0);
    double const * ey = E + 2 * nk + 2 * (a1 < 0 ? -a1 : a1);
    double const * ez = E + 4 * nk + 2 * (a2 < 0 ? -a2 : a2);
    double const   xr = ex[0], xi = a0 < 0 ? -ex[1] : ex[1];
    double const   yr = ey[0], yi = a1 < 0 ? -ey[1] : ey[1];
    double const   zr0 = ez[0], zi0 = a2 < 0 ? -ez[1] : ez[1];
    double const   ar = xr * yr - xi * yi, ai = xr * yi + xi * yr;
    double const   zr = ar * zr0 - ai * zi0, zi = ar * zi0 + ai * zr0;
    double const   w  = q[6] * zr - q[7] * zi;
    v[0] += w * q[3];
    v[1] += w * q[4];
    v[2] += w * q[5];
  }
  int64_t const c = p / n3, i = p % n3;
  for(int k = 0; k < 3; ++k)
    u[(c * 3 + k) * n3 + i] = v[k];
}
}  // namespace

hipError_t launch_divergence_apply(DivergenceArgs const & a, hipStream_t s)
{
  int const nu = a.n_u, np = a.n_p;
  if(nu < 2 || nu > 9 || np < 2 || np > 9 || a.n_cells < 1 || !a.dst || !a.src || !a.cell_geo ||
     !a.face_geo || !a.tables)
    return hipErrorInvalidValue;
  size_t const lds = sizeof(double) * (5 * nu * nu * nu + np * np * np + 13 * nu * nu);
  if(lds > 65536)
    return hipErrorInvalidValue;  // at most 5 * 729 + 729 + 1053 doubles = 43.6 KB for n <= 9
  hipLaunchKernelGGL(divergence_apply_kernel, dim3((unsigned)a.n_cells), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_fourier_interpolate(double * u, double const * x_nodes, double const * mode_data,
                                      int n_modes, int K, int n3, int64_t n_cells, hipStream_t s)
{
  if(!u || !x_nodes || !mode_data || n_modes < 1 || K < 0 || K > 8 || n3 < 1 || n_cells < 1)
    return hipErrorInvalidValue;
  int64_t const n_points = n_cells * n3;
  size_t const  lds      = sizeof(double) * kInterpThreads * 6 * (K + 1);  // <= 55.3 KB
  int64_t const blocks   = (n_points + kInterpThreads - 1) / kInterpThreads;
  hipLaunchKernelGGL(fourier_interpolate_kernel, dim3((unsigned)blocks), dim3(kInterpThreads), lds, s,
                     u, x_nodes, mode_data, n_modes, K, n3, n_points);
  return hipGetLastError();
}
}  // namespace ExaDG
