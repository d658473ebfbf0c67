// SIPG Laplace operator kernels for CDNA4 (gfx950).
//
// Both kernels are element-centric: one group of (k+1)^2 lanes owns one cell, computes the
// cell integral and all six face integrals seen from that cell (reading the neighbour's
// lines from global memory / L2), and writes each dst entry exactly once -- no atomics.
// A workgroup packs floor(256/(k+1)^2) cells; each cell gets its own LDS slice.
//
// Reference semantics: LaplaceOperator::do_cell_integral / do_face_integral /
// do_boundary_integral (poisson/spatial_discretization/laplace_operator.cpp:129-265),
// LaplaceKernel::calculate_{gradient,value}_flux (laplace_operator.h:180-197) and the
// homogeneous mirror BCs of weak_boundary_conditions.h:39-47,142-160.
#include <hip/hip_runtime.h>

#include "../../include/exadg_amd.h"
#include "laplace_kernels.h"

namespace ExaDG
{
namespace
{
constexpr int kInterior = 2;  // face type besides EXADG_BC_DIRICHLET / EXADG_BC_NEUMANN

constexpr int cells_per_block(int n) { return (256 / (n * n)) > 0 ? 256 / (n * n) : 1; }

// ---------------------------------------------------------------- line helpers

// out[i] = sum_j M[i*N+j] in[j]
template<int N, typename Number>
__device__ __forceinline__ void mat_apply(Number const * __restrict__ M, Number const (&in)[N],
                                          Number (&out)[N])
{
#pragma unroll
  for(int i = 0; i < N; ++i)
  {
    Number s = 0;
#pragma unroll
    for(int j = 0; j < N; ++j)
      s += M[i * N + j] * in[j];
    out[i] = s;
  }
}

// out[j] = sum_i M[i*N+j] in[i]
template<int N, typename Number>
__device__ __forceinline__ void mat_apply_T(Number const * __restrict__ M, Number const (&in)[N],
                                            Number (&out)[N])
{
#pragma unroll
  for(int j = 0; j < N; ++j)
  {
    Number s = 0;
#pragma unroll
    for(int i = 0; i < N; ++i)
      s += M[i * N + j] * in[i];
    out[j] = s;
  }
}

template<int N, typename Number>
__device__ __forceinline__ void load_line(Number const * p, int stride, Number (&v)[N])
{
#pragma unroll
  for(int i = 0; i < N; ++i)
    v[i] = p[i * stride];
}

template<int N, typename Number>
__device__ __forceinline__ void store_line(Number * p, int stride, Number const (&v)[N])
{
#pragma unroll
  for(int i = 0; i < N; ++i)
    p[i * stride] = v[i];
}

template<int N, typename Number>
__device__ __forceinline__ Number dot_line(Number const * __restrict__ a, Number const (&v)[N])
{
  Number s = 0;
#pragma unroll
  for(int i = 0; i < N; ++i)
    s += a[i] * v[i];
  return s;
}

// Base offset and stride of the line along direction d through lane (p,q) of an N^3 block;
// (p,q) index the two other directions in increasing order.
template<int N>
__device__ __forceinline__ void line_geometry(int d, int p, int q, int & base, int & stride)
{
  if(d == 0)
  {
    base   = N * (p + N * q);
    stride = 1;
  }
  else if(d == 1)
  {
    base   = p + N * N * q;
    stride = N;
  }
  else
  {
    base   = p + N * q;
    stride = N * N;
  }
}

// the cell's unit vector in diagonal (probing) mode, else the src line
template<int N, typename Number, ApplyMode mode>
__device__ __forceinline__ void source_line(Number const * __restrict__ src, int64_t cell, int base,
                                            int stride, int probe, Number (&v)[N])
{
  constexpr int N3 = N * N * N;
  if constexpr(mode == ApplyMode::diagonal)
  {
#pragma unroll
    for(int i = 0; i < N; ++i)
      v[i] = (base + i * stride == probe) ? Number(1) : Number(0);
  }
  else
    load_line<N>(src + cell * N3 + base, stride, v);
}

// ================================================================ general (curved) kernel
//
// Cell term by collocation at the Gauss points: u_q = (S x S x S) u, reference gradient
// by the Gauss collocation derivative Dq, symmetric metric G = JxW J^{-1} J^{-T} per point,
// then the transposed sweeps. Face terms per direction: traces and normal derivatives at
// the face nodes of both sides, (S x S) to the face Gauss points, flux with the per-point
// normal-derivative vectors a = J^{-1} n of both sides, transposed back and added to the
// own lines.

template<int N>
constexpr int general_work_size()
{
  return (3 * N * N * N > 16 * N * N) ? 3 * N * N * N : 16 * N * N;
}

template<int N, typename Number, ApplyMode mode>
__device__ __forceinline__ void general_face_direction(int d, SipgArgs<Number> const & a,
                                                       int64_t cell, int const * cc, bool active,
                                                       int lt, Number * A, Number * W)
{
  constexpr int  N2 = N * N, N3 = N * N * N;
  Number const * S  = a.tables;
  Number const * Dq = a.tables + N2;
  Number const * d0 = a.tables + 2 * N2;
  Number const * d1 = d0 + N;
  Number *       E  = W;           // E[(s*4+c)*N2]: c = u-, du-/dxi_d, u+, du+/dxi_d
  Number *       R  = W + 8 * N2;  // R[(s*4+c)*N2]: c = value test, t1, t2, d test coefficients
  int const      t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
  int const      p = lt % N, q = lt / N;

  int64_t nb[2];
  int     type[2];
#pragma unroll
  for(int s = 0; s < 2; ++s)
  {
    nb[s]   = neighbor_id(a.topo, cc, d, s);
    type[s] = nb[s] >= 0 ? kInterior : a.topo.bc[2 * d + s];
  }

  // F1: traces at the face nodes
  if(active)
  {
    int base, stride;
    line_geometry<N>(d, p, q, base, stride);
    Number u[N];
    source_line<N, Number, mode>(a.src, cell, base, stride, a.probe, u);
    E[0 * N2 + lt] = u[0];
    E[1 * N2 + lt] = dot_line<N>(d0, u);
    E[4 * N2 + lt] = u[N - 1];
    E[5 * N2 + lt] = dot_line<N>(d1, u);
    if constexpr(mode != ApplyMode::diagonal)
    {
      if(type[0] == kInterior)
      {
        Number w[N];
        load_line<N>(a.src + nb[0] * N3 + base, stride, w);
        E[2 * N2 + lt] = w[N - 1];
        E[3 * N2 + lt] = dot_line<N>(d1, w);
      }
      if(type[1] == kInterior)
      {
        Number w[N];
        load_line<N>(a.src + nb[1] * N3 + base, stride, w);
        E[6 * N2 + lt] = w[0];
        E[7 * N2 + lt] = dot_line<N>(d0, w);
      }
    }
    else
    {
      // probing the cell-local block: neighbours contribute zero, except a cell that is its
      // own periodic neighbour (one cell along d), whose other end then acts as neighbour
      bool const sl = nb[0] == cell, sr = nb[1] == cell;
      E[2 * N2 + lt] = sl ? u[N - 1] : Number(0);
      E[3 * N2 + lt] = sl ? dot_line<N>(d1, u) : Number(0);
      E[6 * N2 + lt] = sr ? u[0] : Number(0);
      E[7 * N2 + lt] = sr ? dot_line<N>(d0, u) : Number(0);
    }
  }
  __syncthreads();
  // F2/F3: interpolate the 8 face arrays to the face Gauss points (S along t1, then t2)
  if(active)
    for(int task = lt; task < 8 * N; task += N2)
    {
      Number * arr = E + (task / N) * N2;
      int const r  = task % N;
      Number    v[N], o[N];
      load_line<N>(arr + N * r, 1, v);
      mat_apply<N>(S, v, o);
      store_line<N>(arr + N * r, 1, o);
    }
  __syncthreads();
  if(active)
    for(int task = lt; task < 8 * N; task += N2)
    {
      Number * arr = E + (task / N) * N2;
      int const r  = task % N;
      Number    v[N], o[N];
      load_line<N>(arr + r, N, v);
      mat_apply<N>(S, v, o);
      store_line<N>(arr + r, N, o);
    }
  __syncthreads();
  // F4: fluxes at the 2 x N^2 face quadrature points
  if(active)
    for(int task = lt; task < 2 * N2; task += N2)
    {
      int const s = task / N2, qq = task % N2, q1 = qq % N, q2 = qq / N;
      Number *  Rs = R + s * 4 * N2;
      if(type[s] == EXADG_BC_NEUMANN)
      {  // homogeneous Neumann: u+ = u-, du+/dn = -du-/dn -> no contribution
        Rs[qq] = Rs[N2 + qq] = Rs[2 * N2 + qq] = Rs[3 * N2 + qq] = Number(0);
        continue;
      }
      int64_t const  f  = face_id(a.topo, cc, d, s);
      Number const * fd = a.face_data + f * 7 * N2;
      Number         aK[3], aN[3];
#pragma unroll
      for(int j = 0; j < 3; ++j)
      {
        Number const am = fd[j * N2 + qq], ap = fd[(3 + j) * N2 + qq];
        aK[j]           = s ? am : -ap;
        aN[j]           = s ? ap : -am;
      }
      Number const   w  = fd[6 * N2 + qq];
      Number const * Es = E + s * 4 * N2;
      Number         gt1 = 0, gt2 = 0;
#pragma unroll
      for(int m = 0; m < N; ++m)
      {
        gt1 += Dq[q1 * N + m] * Es[m + N * q2];
        gt2 += Dq[q2 * N + m] * Es[q1 + N * m];
      }
      Number const u      = Es[qq];
      Number const dudn_K = aK[d] * Es[N2 + qq] + aK[t1] * gt1 + aK[t2] * gt2;
      Number       uN, dudn_N;
      if(type[s] == kInterior)
      {
        Number g1 = 0, g2 = 0;
#pragma unroll
        for(int m = 0; m < N; ++m)
        {
          g1 += Dq[q1 * N + m] * Es[2 * N2 + m + N * q2];
          g2 += Dq[q2 * N + m] * Es[2 * N2 + q1 + N * m];
        }
        uN     = Es[2 * N2 + qq];
        dudn_N = aN[d] * Es[3 * N2 + qq] + aN[t1] * g1 + aN[t2] * g2;
      }
      else
      {  // homogeneous Dirichlet: u+ = -u-, du+/dn = du-/dn
        uN     = -u;
        dudn_N = dudn_K;
      }
      Number const tau   = a.tau_face[f];
      Number const jump  = u - uN;
      Number const sigma = Number(-0.5) * jump * w;
      Number const mu    = -(Number(0.5) * (dudn_K + dudn_N) - tau * jump) * w;
      Rs[qq]             = mu;
      Rs[N2 + qq]        = sigma * aK[t1];
      Rs[2 * N2 + qq]    = sigma * aK[t2];
      Rs[3 * N2 + qq]    = sigma * aK[d];
    }
  __syncthreads();
  // F5: tangential gradient test functions folded into the value test (Dq^T)
  if(active)
    for(int task = lt; task < 2 * N2; task += N2)
    {
      int const s = task / N2, qq = task % N2, q1 = qq % N, q2 = qq / N;
      Number *  Rs = R + s * 4 * N2;
      Number    acc = Rs[qq];
#pragma unroll
      for(int m = 0; m < N; ++m)
        acc += Dq[m * N + q1] * Rs[N2 + m + N * q2] + Dq[m * N + q2] * Rs[2 * N2 + q1 + N * m];
      Rs[qq] = acc;
    }
  __syncthreads();
  // F6/F7: (S x S)^T back to the face nodes for the value and normal-derivative tests
  if(active)
    for(int task = lt; task < 4 * N; task += N2)
    {
      int const which = task / N, r = task % N;
      Number *  arr   = R + ((which >> 1) * 4 + ((which & 1) ? 3 : 0)) * N2;
      Number    v[N], o[N];
      load_line<N>(arr + N * r, 1, v);
      mat_apply_T<N>(S, v, o);
      store_line<N>(arr + N * r, 1, o);
    }
  __syncthreads();
  if(active)
    for(int task = lt; task < 4 * N; task += N2)
    {
      int const which = task / N, r = task % N;
      Number *  arr   = R + ((which >> 1) * 4 + ((which & 1) ? 3 : 0)) * N2;
      Number    v[N], o[N];
      load_line<N>(arr + r, N, v);
      mat_apply_T<N>(S, v, o);
      store_line<N>(arr + r, N, o);
    }
  __syncthreads();
  // F8: add to the own d-lines
  if(active)
  {
    int base, stride;
    line_geometry<N>(d, p, q, base, stride);
    Number v[N];
    load_line<N>(A + base, stride, v);
#pragma unroll
    for(int s = 0; s < 2; ++s)
    {
      if(type[s] == EXADG_BC_NEUMANN)
        continue;
      Number const * Rs = R + s * 4 * N2;
      Number const   rv = Rs[lt], rd = Rs[3 * N2 + lt];
      Number const * dd = s ? d1 : d0;
#pragma unroll
      for(int i = 0; i < N; ++i)
        v[i] += dd[i] * rd;
      v[s ? N - 1 : 0] += rv;
    }
    store_line<N>(A + base, stride, v);
  }
  __syncthreads();
}

template<int N, typename Number, ApplyMode mode>
__global__ void __launch_bounds__(256) sipg_general_kernel(SipgArgs<Number> a)
{
  constexpr int  N2 = N * N, N3 = N * N * N, WS = general_work_size<N>();
  constexpr int  CPB = cells_per_block(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *       smem = reinterpret_cast<Number *>(smem_raw);
  int const      slot = threadIdx.x / N2;
  int const      lt   = threadIdx.x - slot * N2;
  int64_t const  cell = (int64_t)blockIdx.x * CPB + slot;
  bool const     active = slot < CPB && cell < a.n_cells;
  Number *       A      = smem + (active ? slot : 0) * (N3 + WS);
  Number *       W      = A + N3;
  Number *       Gx = W, *Gy = W + N3, *Gz = W + 2 * N3;
  Number const * S  = a.tables;
  Number const * Dq = a.tables + N2;
  int const      p = lt % N, q = lt / N;
  int            cc[3] = {0, 0, 0};
  if(active)
    cell_coords(a.topo, cell, cc);

  // load the cell values
  if(active)
  {
    if constexpr(mode == ApplyMode::diagonal)
      for(int e = lt; e < N3; e += N2)
        A[e] = (e == a.probe) ? Number(1) : Number(0);
    else
      for(int e = lt; e < N3; e += N2)
        A[e] = a.src[cell * N3 + e];
  }
  __syncthreads();
  // nodal -> Gauss values: x, y sweeps in place
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + N * (p + N * q), 1, v);
    mat_apply<N>(S, v, o);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + p + N2 * q, N, v);
    mat_apply<N>(S, v, o);
    store_line<N>(A + p + N2 * q, N, o);
  }
  __syncthreads();
  // z sweep, then the z derivative at the Gauss points
  if(active)
  {
    Number v[N], o[N], g[N];
    load_line<N>(A + p + N * q, N2, v);
    mat_apply<N>(S, v, o);
    store_line<N>(A + p + N * q, N2, o);
    mat_apply<N>(Dq, o, g);
    store_line<N>(Gz + p + N * q, N2, g);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], g[N];
    load_line<N>(A + N * (p + N * q), 1, v);
    mat_apply<N>(Dq, v, g);
    store_line<N>(Gx + N * (p + N * q), 1, g);
    load_line<N>(A + p + N2 * q, N, v);
    mat_apply<N>(Dq, v, g);
    store_line<N>(Gy + p + N2 * q, N, g);
  }
  __syncthreads();
  // metric: (fx,fy,fz) = G (gx,gy,gz), G symmetric per quadrature point
  if(active)
  {
    Number const * G = a.cell_metric + cell * 6 * N3;
#pragma unroll
    for(int m = 0; m < N; ++m)
    {
      int const    e  = lt + m * N2;
      Number const gx = Gx[e], gy = Gy[e], gz = Gz[e];
      Number const g00 = G[e], g11 = G[N3 + e], g22 = G[2 * N3 + e];
      Number const g01 = G[3 * N3 + e], g02 = G[4 * N3 + e], g12 = G[5 * N3 + e];
      Gx[e]            = g00 * gx + g01 * gy + g02 * gz;
      Gy[e]            = g01 * gx + g11 * gy + g12 * gz;
      Gz[e]            = g02 * gx + g12 * gy + g22 * gz;
    }
  }
  __syncthreads();
  // transposed derivative sweeps summed into A, then S^T back to nodal values
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(Gx + N * (p + N * q), 1, v);
    mat_apply_T<N>(Dq, v, o);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N], w[N];
    load_line<N>(Gy + p + N2 * q, N, v);
    mat_apply_T<N>(Dq, v, o);
    load_line<N>(A + p + N2 * q, N, w);
#pragma unroll
    for(int i = 0; i < N; ++i)
      w[i] += o[i];
    store_line<N>(A + p + N2 * q, N, w);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N], w[N];
    load_line<N>(Gz + p + N * q, N2, v);
    mat_apply_T<N>(Dq, v, o);
    load_line<N>(A + p + N * q, N2, w);
#pragma unroll
    for(int i = 0; i < N; ++i)
      w[i] += o[i];
    mat_apply_T<N>(S, w, o);
    store_line<N>(A + p + N * q, N2, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + p + N2 * q, N, v);
    mat_apply_T<N>(S, v, o);
    store_line<N>(A + p + N2 * q, N, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + N * (p + N * q), 1, v);
    mat_apply_T<N>(S, v, o);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();

  for(int d = 0; d < 3; ++d)
    general_face_direction<N, Number, mode>(d, a, cell, cc, active, lt, A, W);

  if(active)
  {
    if constexpr(mode == ApplyMode::diagonal)
    {
      if(lt == 0)
        a.dst[cell * N3 + a.probe] = A[a.probe];
    }
    else if constexpr(mode == ApplyMode::vmult_add)
      for(int e = lt; e < N3; e += N2)
        a.dst[cell * N3 + e] += A[e];
    else
      for(int e = lt; e < N3; e += N2)
        a.dst[cell * N3 + e] = A[e];
  }
}

// ================================================================ Cartesian kernel
//
// On an affine axis-parallel mesh the SIPG operator is exactly a sum of Kronecker products:
//   A = L_x (x) M_y (x) M_z + M_x (x) L_y (x) M_z + M_x (x) M_y (x) L_z
// with M_d = h_d * M (1D mass, exact with QGauss(k+1)) and L_d the 1D DG Laplacian along d:
// K/h_d plus the consistency/symmetry/penalty face terms of both ends, which couple a line
// to the end value and end derivative of the neighbour's line. Evaluated as
//   out = M_z (M_y a_x + M_x a_y) + M_y M_x a_z,  a_d = L_d(u, neighbours).

template<int N, typename Number, ApplyMode mode>
__device__ __forceinline__ void line_laplacian(Number const * __restrict__ tab, Number const (&u)[N],
                                               Number (&o)[N], int const (&type)[2],
                                               Number const (&tau)[2], Number const * wl,
                                               Number const * wr, int stride, bool self_nb)
{
  Number const * K  = tab + N * N;
  Number const * d0 = tab + 2 * N * N;
  Number const * d1 = d0 + N;
  mat_apply<N>(K, u, o);
  if(type[0] != EXADG_BC_NEUMANN)
  {  // left end, outward normal -e_d
    bool const   dir  = type[0] == EXADG_BC_DIRICHLET;
    Number const f    = dir ? Number(1) : Number(0.5);
    Number const pen  = dir ? Number(2) * tau[0] : tau[0];
    Number const du0  = dot_line<N>(d0, u);
    Number const u0   = u[0];
#pragma unroll
    for(int i = 0; i < N; ++i)
      o[i] += f * d0[i] * u0;
    o[0] += f * du0 + pen * u0;
    if(type[0] == kInterior && (mode != ApplyMode::diagonal || self_nb))
      {
        Number w[N];
        if constexpr(mode == ApplyMode::diagonal)
        {  // one cell along a periodic direction: the cell is its own neighbour
#pragma unroll
          for(int i = 0; i < N; ++i)
            w[i] = u[i];
        }
        else
          load_line<N>(wl, stride, w);
        Number const wn = w[N - 1], dw = dot_line<N>(d1, w);
#pragma unroll
        for(int i = 0; i < N; ++i)
          o[i] -= Number(0.5) * d0[i] * wn;
        o[0] += Number(0.5) * dw - tau[0] * wn;
      }
  }
  if(type[1] != EXADG_BC_NEUMANN)
  {  // right end, outward normal +e_d
    bool const   dir = type[1] == EXADG_BC_DIRICHLET;
    Number const f   = dir ? Number(1) : Number(0.5);
    Number const pen = dir ? Number(2) * tau[1] : tau[1];
    Number const du1 = dot_line<N>(d1, u);
    Number const un  = u[N - 1];
#pragma unroll
    for(int i = 0; i < N; ++i)
      o[i] -= f * d1[i] * un;
    o[N - 1] += pen * un - f * du1;
    if(type[1] == kInterior && (mode != ApplyMode::diagonal || self_nb))
      {
        Number w[N];
        if constexpr(mode == ApplyMode::diagonal)
        {
#pragma unroll
          for(int i = 0; i < N; ++i)
            w[i] = u[i];
        }
        else
          load_line<N>(wr, stride, w);
        Number const w0 = w[0], dw = dot_line<N>(d0, w);
#pragma unroll
        for(int i = 0; i < N; ++i)
          o[i] += Number(0.5) * d1[i] * w0;
        o[N - 1] -= Number(0.5) * dw + tau[1] * w0;
      }
  }
}

template<int N, typename Number, ApplyMode mode>
__global__ void __launch_bounds__(256) sipg_cartesian_kernel(SipgArgs<Number> a)
{
  constexpr int N2 = N * N, N3 = N * N * N, TB = 2 * N * N + 2 * N;
  constexpr int CPB = cells_per_block(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      smem   = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / N2;
  int const     lt     = threadIdx.x - slot * N2;
  int64_t const cell   = (int64_t)blockIdx.x * CPB + slot;
  bool const    active = slot < CPB && cell < a.n_cells;
  Number *      U      = smem + (active ? slot : 0) * 3 * N3;
  Number *      Abuf   = U + N3;
  Number *      Bbuf   = Abuf + N3;
  int const     p = lt % N, q = lt / N;

  int     cc[3] = {0, 0, 0};
  int64_t nb[3][2];
  int     type[3][2];
  Number  tau[3][2];
  if(active)
  {
    cell_coords(a.topo, cell, cc);
#pragma unroll
    for(int d = 0; d < 3; ++d)
#pragma unroll
      for(int s = 0; s < 2; ++s)
      {
        nb[d][s]   = neighbor_id(a.topo, cc, d, s);
        type[d][s] = nb[d][s] >= 0 ? kInterior : a.topo.bc[2 * d + s];
        tau[d][s]  = a.tau_face[face_id(a.topo, cc, d, s)];
      }
    if constexpr(mode == ApplyMode::diagonal)
      for(int e = lt; e < N3; e += N2)
        U[e] = (e == a.probe) ? Number(1) : Number(0);
    else
      for(int e = lt; e < N3; e += N2)
        U[e] = a.src[cell * N3 + e];
  }
  __syncthreads();
  Number const * tx = a.tables;
  Number const * ty = a.tables + TB;
  Number const * tz = a.tables + 2 * TB;

  // X: a_x = L_x u  -> A
  if(active)
  {
    int const base = N * (p + N * q);
    Number    u[N], o[N];
    load_line<N>(U + base, 1, u);
    int const    ty_[2] = {type[0][0], type[0][1]};
    Number const ta_[2] = {tau[0][0], tau[0][1]};
    line_laplacian<N, Number, mode>(tx, u, o, ty_, ta_, a.src + (nb[0][0] >= 0 ? nb[0][0] : 0) * N3 + base,
                                    a.src + (nb[0][1] >= 0 ? nb[0][1] : 0) * N3 + base, 1,
                                    nb[0][0] == cell);
    store_line<N>(Abuf + base, 1, o);
  }
  __syncthreads();
  // Y: A <- M_y a_x, B <- a_y = L_y u
  if(active)
  {
    int const base = p + N2 * q;
    Number    u[N], o[N], v[N], w[N];
    load_line<N>(U + base, N, u);
    int const    ty_[2] = {type[1][0], type[1][1]};
    Number const ta_[2] = {tau[1][0], tau[1][1]};
    line_laplacian<N, Number, mode>(ty, u, o, ty_, ta_, a.src + (nb[1][0] >= 0 ? nb[1][0] : 0) * N3 + base,
                                    a.src + (nb[1][1] >= 0 ? nb[1][1] : 0) * N3 + base, N,
                                    nb[1][0] == cell);
    store_line<N>(Bbuf + base, N, o);
    load_line<N>(Abuf + base, N, v);
    mat_apply<N>(ty, v, w);
    store_line<N>(Abuf + base, N, w);
  }
  __syncthreads();
  // X2: A <- A + M_x a_y
  if(active)
  {
    int const base = N * (p + N * q);
    Number    v[N], w[N], r[N];
    load_line<N>(Bbuf + base, 1, v);
    mat_apply<N>(tx, v, w);
    load_line<N>(Abuf + base, 1, r);
#pragma unroll
    for(int i = 0; i < N; ++i)
      r[i] += w[i];
    store_line<N>(Abuf + base, 1, r);
  }
  __syncthreads();
  // Z: out_partial = M_z A (registers), B <- a_z = L_z u
  Number outp[N];
  if(active)
  {
    int const base = p + N * q;
    Number    u[N], o[N], v[N];
    load_line<N>(U + base, N2, u);
    int const    ty_[2] = {type[2][0], type[2][1]};
    Number const ta_[2] = {tau[2][0], tau[2][1]};
    line_laplacian<N, Number, mode>(tz, u, o, ty_, ta_, a.src + (nb[2][0] >= 0 ? nb[2][0] : 0) * N3 + base,
                                    a.src + (nb[2][1] >= 0 ? nb[2][1] : 0) * N3 + base, N2,
                                    nb[2][0] == cell);
    load_line<N>(Abuf + base, N2, v);
    mat_apply<N>(tz, v, outp);
    store_line<N>(Bbuf + base, N2, o);
  }
  __syncthreads();
  // X3: B <- M_x B
  if(active)
  {
    int const base = N * (p + N * q);
    Number    v[N], w[N];
    load_line<N>(Bbuf + base, 1, v);
    mat_apply<N>(tx, v, w);
    store_line<N>(Bbuf + base, 1, w);
  }
  __syncthreads();
  // Y3: B <- M_y B
  if(active)
  {
    int const base = p + N2 * q;
    Number    v[N], w[N];
    load_line<N>(Bbuf + base, N, v);
    mat_apply<N>(ty, v, w);
    store_line<N>(Bbuf + base, N, w);
  }
  __syncthreads();
  // out = out_partial + B along the z-lines
  if(active)
  {
    int const base = p + N * q;
    Number    v[N];
    load_line<N>(Bbuf + base, N2, v);
    if constexpr(mode == ApplyMode::diagonal)
    {
#pragma unroll
      for(int k = 0; k < N; ++k)
        if(base + k * N2 == a.probe)
          a.dst[cell * N3 + a.probe] = outp[k] + v[k];
    }
    else
    {
      Number * out = a.dst + cell * N3 + base;
#pragma unroll
      for(int k = 0; k < N; ++k)
      {
        if constexpr(mode == ApplyMode::vmult_add)
          out[k * N2] += outp[k] + v[k];
        else
          out[k * N2] = outp[k] + v[k];
      }
    }
  }
}

template<int N, typename Number, ApplyMode mode>
hipError_t launch_general_n(SipgArgs<Number> const & a, hipStream_t stream)
{
  constexpr int CPB    = cells_per_block(N);
  int const     blocks = (int)((a.n_cells + CPB - 1) / CPB);
  size_t const  lds    = sizeof(Number) * CPB * (N * N * N + general_work_size<N>());
  static bool   attr   = false;
  if(!attr && lds > 65536)
  {
    hipError_t e = hipFuncSetAttribute((void const *)sipg_general_kernel<N, Number, mode>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if(e != hipSuccess)
      return e;
    attr = true;
  }
  hipLaunchKernelGGL((sipg_general_kernel<N, Number, mode>), dim3(blocks), dim3(256), lds, stream, a);
  return hipGetLastError();
}

template<int N, typename Number, ApplyMode mode>
hipError_t launch_cartesian_n(SipgArgs<Number> const & a, hipStream_t stream)
{
  constexpr int CPB    = cells_per_block(N);
  int const     blocks = (int)((a.n_cells + CPB - 1) / CPB);
  size_t const  lds    = sizeof(Number) * CPB * 3 * N * N * N;
  hipLaunchKernelGGL((sipg_cartesian_kernel<N, Number, mode>), dim3(blocks), dim3(256), lds, stream,
                     a);
  return hipGetLastError();
}

template<typename Number, ApplyMode mode>
hipError_t dispatch_general(int degree, SipgArgs<Number> const & a, hipStream_t s)
{
  switch(degree)
  {
    case 1: return launch_general_n<2, Number, mode>(a, s);
    case 2: return launch_general_n<3, Number, mode>(a, s);
    case 3: return launch_general_n<4, Number, mode>(a, s);
    case 4: return launch_general_n<5, Number, mode>(a, s);
    case 5: return launch_general_n<6, Number, mode>(a, s);
    case 6: return launch_general_n<7, Number, mode>(a, s);
    case 7: return launch_general_n<8, Number, mode>(a, s);
    case 8: return launch_general_n<9, Number, mode>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template<typename Number, ApplyMode mode>
hipError_t dispatch_cartesian(int degree, SipgArgs<Number> const & a, hipStream_t s)
{
  switch(degree)
  {
    case 1: return launch_cartesian_n<2, Number, mode>(a, s);
    case 2: return launch_cartesian_n<3, Number, mode>(a, s);
    case 3: return launch_cartesian_n<4, Number, mode>(a, s);
    case 4: return launch_cartesian_n<5, Number, mode>(a, s);
    case 5: return launch_cartesian_n<6, Number, mode>(a, s);
    case 6: return launch_cartesian_n<7, Number, mode>(a, s);
    case 7: return launch_cartesian_n<8, Number, mode>(a, s);
    case 8: return launch_cartesian_n<9, Number, mode>(a, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

template<typename Number>
hipError_t launch_sipg_general(int degree, ApplyMode mode, SipgArgs<Number> const & a,
                               hipStream_t s)
{
  switch(mode)
  {
    case ApplyMode::vmult: return dispatch_general<Number, ApplyMode::vmult>(degree, a, s);
    case ApplyMode::vmult_add: return dispatch_general<Number, ApplyMode::vmult_add>(degree, a, s);
    case ApplyMode::diagonal: return dispatch_general<Number, ApplyMode::diagonal>(degree, a, s);
  }
  return hipErrorInvalidValue;
}

template<typename Number>
hipError_t launch_sipg_cartesian(int degree, ApplyMode mode, SipgArgs<Number> const & a,
                                 hipStream_t s)
{
  switch(mode)
  {
    case ApplyMode::vmult: return dispatch_cartesian<Number, ApplyMode::vmult>(degree, a, s);
    case ApplyMode::vmult_add: return dispatch_cartesian<Number, ApplyMode::vmult_add>(degree, a, s);
    case ApplyMode::diagonal: return dispatch_cartesian<Number, ApplyMode::diagonal>(degree, a, s);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_sipg_general<double>(int, ApplyMode, SipgArgs<double> const &, hipStream_t);
template hipError_t launch_sipg_general<float>(int, ApplyMode, SipgArgs<float> const &, hipStream_t);
template hipError_t launch_sipg_cartesian<double>(int, ApplyMode, SipgArgs<double> const &, hipStream_t);
template hipError_t launch_sipg_cartesian<float>(int, ApplyMode, SipgArgs<float> const &, hipStream_t);

}  // namespace ExaDG
