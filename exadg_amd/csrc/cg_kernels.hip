// Continuous-level and h-transfer kernels of the hybrid multigrid (see cg_kernels.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/exadg_amd.h"
#include "cg_kernels.h"
#include "line_sweeps.h"

namespace ExaDG
{
CGTopology make_cg_topology(BoxTopology const & t, int p)
{
  CGTopology q;
  q.p = p;
  for(int d = 0; d < 3; ++d)
  {
    q.nc[d]       = t.nc[d];
    q.periodic[d] = t.periodic[d];
    q.nn[d]       = t.nc[d] * p + (t.periodic[d] ? 0 : 1);
    for(int s = 0; s < 2; ++s)
      q.dirichlet[2 * d + s] = (!t.periodic[d] && t.bc[2 * d + s] == EXADG_BC_DIRICHLET) ? 1 : 0;
  }
  return q;
}

namespace
{
inline int blocks_for(int64_t n) { return (int)((n + 255) / 256); }

using sweeps::line_base;
using sweeps::line_stride;
using sweeps::mat_line;
using sweeps::sweep;

// ---------------------------------------------------------------- c-transfer (any degree)

template<typename Number>
__global__ void __launch_bounds__(256)
  c_prolongate_kernel(Number * dst_dg, Number const * src_cg, BoxTopology t, CGTopology q, int64_t n)
{
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n)
    return;
  int const     n1 = q.p + 1, n3 = n1 * n1 * n1;
  int64_t const cell = i / n3;
  int const     loc  = (int)(i - cell * n3);
  int const     a[3] = {loc % n1, (loc / n1) % n1, loc / (n1 * n1)};
  int           cc[3], g[3];
  cell_coords(t, cell, cc);
#pragma unroll
  for(int d = 0; d < 3; ++d)
  {
    g[d] = cc[d] * q.p + a[d];
    if(q.periodic[d] && g[d] == q.nn[d])
      g[d] = 0;
  }
  if(!cg_constrained(q, g))
    dst_dg[i] += src_cg[cg_node_id(q, g)];
}

template<typename Number>
__global__ void __launch_bounds__(256)
  c_restrict_kernel(Number * dst_cg, Number const * src_dg, BoxTopology t, CGTopology q, int64_t n)
{
  int64_t const node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(node >= n)
    return;
  int g[3];
  cg_node_coords(q, node, g);
  if(cg_constrained(q, g))
    return;
  int const n1 = q.p + 1, n3 = n1 * n1 * n1;
  int       c0[2], l0[2], c1[2], l1[2], c2[2], l2[2];
  int const k0 = cg_sharing_1d(q, 0, g[0], c0, l0);
  int const k1 = cg_sharing_1d(q, 1, g[1], c1, l1);
  int const k2 = cg_sharing_1d(q, 2, g[2], c2, l2);
  Number    acc = 0;
  for(int i2 = 0; i2 < k2; ++i2)
    for(int i1 = 0; i1 < k1; ++i1)
      for(int i0 = 0; i0 < k0; ++i0)
      {
        int const cc[3] = {c0[i0], c1[i1], c2[i2]};
        acc += src_dg[cell_id(t, cc) * n3 + l0[i0] + n1 * (l1[i1] + n1 * l2[i2])];
      }
  dst_cg[node] += acc;
}

// ---------------------------------------------------------------- separable line transfer

template<typename Number>
__global__ void __launch_bounds__(256)
  line_apply_kernel(Number * out, Number const * __restrict__ in, int di0, int di1, int d, int rows,
                    int width, int const * __restrict__ idx, Number const * __restrict__ w,
                    CGTopology q_in, int mask_in, CGTopology q_out, int mask_out, int accumulate,
                    int do0, int do1, int64_t n_out)
{
  int64_t const i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n_out)
    return;
  int o[3];
  o[0]            = (int)(i % do0);
  int64_t const r = i / do0;
  o[1]            = (int)(r % do1);
  o[2]            = (int)(r / do1);
  if(mask_out && cg_constrained(q_out, o))
    return;
  int const row = o[d];
  Number    acc = 0;
  for(int j = 0; j < width; ++j)
  {
    int s[3] = {o[0], o[1], o[2]};
    s[d]     = idx[row * width + j];
    if(mask_in && cg_constrained(q_in, s))
      continue;
    acc += w[row * width + j] * in[s[0] + (int64_t)di0 * (s[1] + (int64_t)di1 * s[2])];
  }
  out[i] = accumulate ? out[i] + acc : acc;
}

// ---------------------------------------------------------------- DG h-transfer

template<typename Number, int N>
__global__ void __launch_bounds__(256)
  h_prolongate_dg_kernel(Number * dst, Number const * __restrict__ src, Number const * __restrict__ Ph,
                         BoxTopology tf, BoxTopology tc, int64_t n_cells, int cpb)
{
  constexpr int N2 = N * N, N3 = N2 * N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      sm     = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / N2;
  int const     lt     = threadIdx.x - slot * N2;
  int64_t const cell   = (int64_t)blockIdx.x * cpb + slot;
  bool const    active = slot < cpb && cell < n_cells;
  Number *      b0     = sm + (active ? slot : 0) * 2 * N3;
  Number *      b1     = b0 + N3;
  int           cc[3] = {0, 0, 0}, pc[3] = {0, 0, 0};
  int64_t       parent = 0;
  if(active)
  {
    cell_coords(tf, cell, cc);
    for(int d = 0; d < 3; ++d)
      pc[d] = cc[d] >> 1;
    parent = cell_id(tc, pc);
    for(int e = lt; e < N3; e += N2)
      b0[e] = src[parent * N3 + e];
  }
  __syncthreads();
  if(active)
    sweep<Number, N, false, false>(Ph + (cc[0] & 1) * N2, b0, b1, 0, lt);
  __syncthreads();
  if(active)
    sweep<Number, N, false, false>(Ph + (cc[1] & 1) * N2, b1, b0, 1, lt);
  __syncthreads();
  if(active)
  {
    int const base = line_base<N>(2, lt);
    Number    x[N], y[N];
#pragma unroll
    for(int i = 0; i < N; ++i)
      x[i] = b0[base + i * N2];
    mat_line<Number, N, false>(Ph + (cc[2] & 1) * N2, x, y);
#pragma unroll
    for(int a = 0; a < N; ++a)
      dst[cell * N3 + base + a * N2] += y[a];
  }
}

template<typename Number, int N>
__global__ void __launch_bounds__(256)
  h_restrict_dg_kernel(Number * dst, Number const * __restrict__ src, Number const * __restrict__ Ph,
                       BoxTopology tf, BoxTopology tc, int64_t n_cells, int cpb)
{
  constexpr int N2 = N * N, N3 = N2 * N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      sm     = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / N2;
  int const     lt     = threadIdx.x - slot * N2;
  int64_t const cell   = (int64_t)blockIdx.x * cpb + slot;  // coarse cell
  bool const    active = slot < cpb && cell < n_cells;
  Number *      b0     = sm + (active ? slot : 0) * 2 * N3;
  Number *      b1     = b0 + N3;
  int           pc[3]  = {0, 0, 0};
  if(active)
    cell_coords(tc, cell, pc);
  Number acc[N];
#pragma unroll
  for(int c = 0; c < N; ++c)
    acc[c] = 0;
  for(int child = 0; child < 8; ++child)
  {
    int const bits[3] = {child & 1, (child >> 1) & 1, (child >> 2) & 1};
    if(active)
    {
      int const     cc[3] = {2 * pc[0] + bits[0], 2 * pc[1] + bits[1], 2 * pc[2] + bits[2]};
      int64_t const fine  = cell_id(tf, cc);
      for(int e = lt; e < N3; e += N2)
        b0[e] = src[fine * N3 + e];
    }
    __syncthreads();
    if(active)
      sweep<Number, N, true, false>(Ph + bits[0] * N2, b0, b1, 0, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, true, false>(Ph + bits[1] * N2, b1, b0, 1, lt);
    __syncthreads();
    if(active)
    {
      int const base = line_base<N>(2, lt);
      Number    x[N], y[N];
#pragma unroll
      for(int i = 0; i < N; ++i)
        x[i] = b0[base + i * N2];
      mat_line<Number, N, true>(Ph + bits[2] * N2, x, y);
#pragma unroll
      for(int a = 0; a < N; ++a)
        acc[a] += y[a];
    }
    __syncthreads();  // b0 is reloaded by the next child
  }
  if(active)
  {
    int const base = line_base<N>(2, lt);
#pragma unroll
    for(int a = 0; a < N; ++a)
      dst[cell * N3 + base + a * N2] += acc[a];
  }
}

// ---------------------------------------------------------------- continuous Q_p cells

template<typename Number, int N, bool DIAG>
__global__ void __launch_bounds__(256) cg_cell_kernel(CGCellArgs<Number> a, int cpb)
{
  constexpr int N2 = N * N, N3 = N2 * N;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *       sm     = reinterpret_cast<Number *>(smem_raw);
  int const      slot   = threadIdx.x / N2;
  int const      lt     = threadIdx.x - slot * N2;
  int64_t const  cell   = (int64_t)blockIdx.x * cpb + slot;
  bool const     active = slot < cpb && cell < a.n_cells;
  Number *       U      = sm + (active ? slot : 0) * 4 * N3;
  Number *       G0     = U + N3;
  Number *       G1     = G0 + N3;
  Number *       G2     = G1 + N3;
  Number const * S      = a.tables;
  Number const * Dq     = S + N2;
  Number const * w      = Dq + N2;
  Number const * SS     = w + N;
  Number const * DD     = SS + N2;
  Number const * SDp    = DD + N2;
  Number const * metric = a.metric ? a.metric + (active ? cell : 0) * 6 * N3 : nullptr;
  int            cc[3]  = {0, 0, 0};
  if(active)
    cell_coords(a.t, cell, cc);

  if constexpr(!DIAG)
  {
    if(active)
      for(int e = lt; e < N3; e += N2)
      {
        int const l[3] = {e % N, (e / N) % N, e / N2};
        int       g[3];
#pragma unroll
        for(int d = 0; d < 3; ++d)
        {
          g[d] = cc[d] * a.q.p + l[d];
          if(a.q.periodic[d] && g[d] == a.q.nn[d])
            g[d] = 0;
        }
        U[e] = cg_constrained(a.q, g) ? Number(0) : a.src[cg_node_id(a.q, g)];
      }
    __syncthreads();
    // values at the Gauss points
    if(active)
      sweep<Number, N, false, false>(S, U, G0, 0, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, false, false>(S, G0, G1, 1, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, false, false>(S, G1, U, 2, lt);
    __syncthreads();
    // reference gradient by collocation derivatives
    if(active)
    {
      sweep<Number, N, false, false>(Dq, U, G0, 0, lt);
      sweep<Number, N, false, false>(Dq, U, G1, 1, lt);
      sweep<Number, N, false, false>(Dq, U, G2, 2, lt);
    }
    __syncthreads();
    // flux G grad u at every point
    if(active)
      for(int e = lt; e < N3; e += N2)
      {
        Number const g0 = G0[e], g1 = G1[e], g2 = G2[e];
        if(metric)
        {
          Number const m00 = metric[e], m11 = metric[N3 + e], m22 = metric[2 * N3 + e];
          Number const m01 = metric[3 * N3 + e], m02 = metric[4 * N3 + e], m12 = metric[5 * N3 + e];
          G0[e] = m00 * g0 + m01 * g1 + m02 * g2;
          G1[e] = m01 * g0 + m11 * g1 + m12 * g2;
          G2[e] = m02 * g0 + m12 * g1 + m22 * g2;
        }
        else
        {
          Number const we = w[e % N] * w[(e / N) % N] * w[e / N2];
          G0[e]           = a.hscale[0] * we * g0;
          G1[e]           = a.hscale[1] * we * g1;
          G2[e]           = a.hscale[2] * we * g2;
        }
      }
    __syncthreads();
    // divergence with the test gradients: U = Dq^T (x) G0 + Dq^T (y) G1 + Dq^T (z) G2
    if(active)
      sweep<Number, N, true, false>(Dq, G0, U, 0, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, true, true>(Dq, G1, U, 1, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, true, true>(Dq, G2, U, 2, lt);
    __syncthreads();
    // back to the nodes
    if(active)
      sweep<Number, N, true, false>(S, U, G0, 0, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, true, false>(S, G0, G1, 1, lt);
    __syncthreads();
    if(active)
      sweep<Number, N, true, false>(S, G1, U, 2, lt);
    __syncthreads();
  }
  else
  {
    // diag_i = sum_q sum_(c,d) G_cd(q) dphi_i/dxi_c (q) dphi_i/dxi_d (q): per metric component a
    // tensor contraction with the elementwise products of the 1D tables
    if(active)
      for(int e = lt; e < N3; e += N2)
        U[e] = 0;
    int const n_comp = metric ? 6 : 3;
    for(int comp = 0; comp < n_comp; ++comp)
    {
      Number const * A0 = comp == 0 ? DD : ((comp == 3 || comp == 4) ? SDp : SS);
      Number const * A1 = comp == 1 ? DD : ((comp == 3 || comp == 5) ? SDp : SS);
      Number const * A2 = comp == 2 ? DD : ((comp == 4 || comp == 5) ? SDp : SS);
      Number const   f  = comp >= 3 ? Number(2) : Number(1);
      if(active)
        for(int e = lt; e < N3; e += N2)
          G0[e] = metric ? f * metric[comp * N3 + e]
                         : a.hscale[comp] * w[e % N] * w[(e / N) % N] * w[e / N2];
      __syncthreads();
      if(active)
        sweep<Number, N, true, false>(A0, G0, G1, 0, lt);
      __syncthreads();
      if(active)
        sweep<Number, N, true, false>(A1, G1, G2, 1, lt);
      __syncthreads();
      if(active)
        sweep<Number, N, true, true>(A2, G2, U, 2, lt);
      __syncthreads();
    }
  }
  if(active)
    for(int e = lt; e < N3; e += N2)
      a.out[cell * N3 + e] = U[e];
}

template<typename Number>
__global__ void __launch_bounds__(256)
  cg_assemble_kernel(CGAssembleArgs<Number> a, int mode, int64_t n)
{
  int64_t const node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(node >= n)
    return;
  int g[3];
  cg_node_coords(a.q, node, g);
  bool const con = cg_constrained(a.q, g);
  Number     val = 0;
  if(!con)
  {
    int const n1 = a.q.p + 1, n3 = n1 * n1 * n1;
    int       c0[2], l0[2], c1[2], l1[2], c2[2], l2[2];
    int const k0 = cg_sharing_1d(a.q, 0, g[0], c0, l0);
    int const k1 = cg_sharing_1d(a.q, 1, g[1], c1, l1);
    int const k2 = cg_sharing_1d(a.q, 2, g[2], c2, l2);
    for(int i2 = 0; i2 < k2; ++i2)
      for(int i1 = 0; i1 < k1; ++i1)
        for(int i0 = 0; i0 < k0; ++i0)
        {
          int const cc[3] = {c0[i0], c1[i1], c2[i2]};
          val += a.cell_values[cell_id(a.t, cc) * n3 + l0[i0] + n1 * (l1[i1] + n1 * l2[i2])];
        }
  }
  if(mode == (int)CGAssembleMode::diagonal)
  {
    a.dst[node] = con ? Number(1) : val;
    return;
  }
  Number const x  = a.x[node];
  Number const Ax = con ? x : val;
  if(mode == (int)CGAssembleMode::vmult)
    a.dst[node] = Ax;
  else if(mode == (int)CGAssembleMode::residual)
    a.dst[node] = a.b[node] - Ax;
  else
  {
    Number v = (Number(1) + a.f1) * x + a.f2 * a.dinv[node] * (a.b[node] - Ax);
    if(a.has_old)
      v -= a.f1 * a.x_old[node];
    a.dst[node] = v;
  }
}

template<typename Number, int N>
hipError_t launch_cg_cell_n(CGCellArgs<Number> const & a, bool diagonal, hipStream_t s)
{
  // cells per block: 256 lanes, at most 64 KiB of LDS (four n^3 arrays per cell)
  int const    cpb  = std::min<int>(256 / (N * N), (int)(65536 / (sizeof(Number) * 4 * N * N * N)));
  size_t const lds  = sizeof(Number) * cpb * 4 * N * N * N;
  int64_t const nb  = (a.n_cells + cpb - 1) / cpb;
  if(diagonal)
    hipLaunchKernelGGL((cg_cell_kernel<Number, N, true>), dim3(nb), dim3(256), lds, s, a, cpb);
  else
    hipLaunchKernelGGL((cg_cell_kernel<Number, N, false>), dim3(nb), dim3(256), lds, s, a, cpb);
  return hipGetLastError();
}

template<typename Number, int N>
hipError_t launch_h_dg_n(Number * dst, Number const * src, Number const * Ph, BoxTopology const & tf,
                         BoxTopology const & tc, bool prolong, hipStream_t s)
{
  int const     cpb = 256 / (N * N);
  size_t const  lds = sizeof(Number) * cpb * 2 * N * N * N;
  BoxTopology const & t  = prolong ? tf : tc;
  int64_t const nc  = (int64_t)t.nc[0] * t.nc[1] * t.nc[2];
  int64_t const nb  = (nc + cpb - 1) / cpb;
  if(prolong)
    hipLaunchKernelGGL((h_prolongate_dg_kernel<Number, N>), dim3(nb), dim3(256), lds, s, dst, src, Ph,
                       tf, tc, nc, cpb);
  else
    hipLaunchKernelGGL((h_restrict_dg_kernel<Number, N>), dim3(nb), dim3(256), lds, s, dst, src, Ph,
                       tf, tc, nc, cpb);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_h_dg(Number * dst, Number const * src, Number const * Ph, int n,
                       BoxTopology const & tf, BoxTopology const & tc, bool prolong, hipStream_t s)
{
  switch(n)
  {
    case 2: return launch_h_dg_n<Number, 2>(dst, src, Ph, tf, tc, prolong, s);
    case 3: return launch_h_dg_n<Number, 3>(dst, src, Ph, tf, tc, prolong, s);
    case 4: return launch_h_dg_n<Number, 4>(dst, src, Ph, tf, tc, prolong, s);
    case 5: return launch_h_dg_n<Number, 5>(dst, src, Ph, tf, tc, prolong, s);
    case 6: return launch_h_dg_n<Number, 6>(dst, src, Ph, tf, tc, prolong, s);
    case 7: return launch_h_dg_n<Number, 7>(dst, src, Ph, tf, tc, prolong, s);
    case 8: return launch_h_dg_n<Number, 8>(dst, src, Ph, tf, tc, prolong, s);
    case 9: return launch_h_dg_n<Number, 9>(dst, src, Ph, tf, tc, prolong, s);
    default: return hipErrorInvalidValue;
  }
}
}  // namespace

template<typename Number>
hipError_t launch_c_prolongate_add(Number * dst_dg, Number const * src_cg, BoxTopology const & t,
                                   CGTopology const & q, hipStream_t s)
{
  int const     n1 = q.p + 1;
  int64_t const n  = (int64_t)t.nc[0] * t.nc[1] * t.nc[2] * n1 * n1 * n1;
  hipLaunchKernelGGL((c_prolongate_kernel<Number>), dim3(blocks_for(n)), dim3(256), 0, s, dst_dg,
                     src_cg, t, q, n);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_c_restrict_add(Number * dst_cg, Number const * src_dg, BoxTopology const & t,
                                 CGTopology const & q, hipStream_t s)
{
  int64_t const n = (int64_t)q.nn[0] * q.nn[1] * q.nn[2];
  hipLaunchKernelGGL((c_restrict_kernel<Number>), dim3(blocks_for(n)), dim3(256), 0, s, dst_cg,
                     src_dg, t, q, n);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_line_apply(Number * out, Number const * in, int const * dims_in, int d,
                             LineOperator<Number> const & L, CGTopology const * q_in,
                             CGTopology const * q_out, bool accumulate, hipStream_t s)
{
  int dout[3] = {dims_in[0], dims_in[1], dims_in[2]};
  dout[d]     = L.rows;
  int64_t const n_out = (int64_t)dout[0] * dout[1] * dout[2];
  CGTopology const none{};
  hipLaunchKernelGGL((line_apply_kernel<Number>), dim3(blocks_for(n_out)), dim3(256), 0, s, out, in,
                     dims_in[0], dims_in[1], d, L.rows, L.width, L.idx, L.w, q_in ? *q_in : none,
                     q_in ? 1 : 0, q_out ? *q_out : none, q_out ? 1 : 0, accumulate ? 1 : 0, dout[0],
                     dout[1], n_out);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_h_prolongate_dg_add(Number * dst_f, Number const * src_c, Number const * Ph, int n,
                                      BoxTopology const & tf, BoxTopology const & tc, hipStream_t s)
{
  return launch_h_dg<Number>(dst_f, src_c, Ph, n, tf, tc, true, s);
}

template<typename Number>
hipError_t launch_h_restrict_dg_add(Number * dst_c, Number const * src_f, Number const * Ph, int n,
                                    BoxTopology const & tf, BoxTopology const & tc, hipStream_t s)
{
  return launch_h_dg<Number>(dst_c, src_f, Ph, n, tf, tc, false, s);
}

template<typename Number>
hipError_t launch_cg_cell(CGCellArgs<Number> const & a, bool diagonal, hipStream_t s)
{
  switch(a.q.p + 1)
  {
    case 2: return launch_cg_cell_n<Number, 2>(a, diagonal, s);
    case 3: return launch_cg_cell_n<Number, 3>(a, diagonal, s);
    case 4: return launch_cg_cell_n<Number, 4>(a, diagonal, s);
    case 5: return launch_cg_cell_n<Number, 5>(a, diagonal, s);
    case 6: return launch_cg_cell_n<Number, 6>(a, diagonal, s);
    case 7: return launch_cg_cell_n<Number, 7>(a, diagonal, s);
    case 8: return launch_cg_cell_n<Number, 8>(a, diagonal, s);
    case 9: return launch_cg_cell_n<Number, 9>(a, diagonal, s);
    default: return hipErrorInvalidValue;
  }
}

template<typename Number>
hipError_t launch_cg_assemble(CGAssembleArgs<Number> const & a, CGAssembleMode mode, hipStream_t s)
{
  int64_t const n = (int64_t)a.q.nn[0] * a.q.nn[1] * a.q.nn[2];
  hipLaunchKernelGGL((cg_assemble_kernel<Number>), dim3(blocks_for(n)), dim3(256), 0, s, a,
                     (int)mode, n);
  return hipGetLastError();
}

#define EXADG_INST(T)                                                                              \
  template hipError_t launch_c_prolongate_add<T>(T *, T const *, BoxTopology const &,              \
                                                 CGTopology const &, hipStream_t);                 \
  template hipError_t launch_c_restrict_add<T>(T *, T const *, BoxTopology const &,                \
                                               CGTopology const &, hipStream_t);                   \
  template hipError_t launch_line_apply<T>(T *, T const *, int const *, int,                       \
                                           LineOperator<T> const &, CGTopology const *,            \
                                           CGTopology const *, bool, hipStream_t);                 \
  template hipError_t launch_h_prolongate_dg_add<T>(T *, T const *, T const *, int,                \
                                                    BoxTopology const &, BoxTopology const &,      \
                                                    hipStream_t);                                  \
  template hipError_t launch_h_restrict_dg_add<T>(T *, T const *, T const *, int,                  \
                                                  BoxTopology const &, BoxTopology const &,        \
                                                  hipStream_t);                                    \
  template hipError_t launch_cg_cell<T>(CGCellArgs<T> const &, bool, hipStream_t);                 \
  template hipError_t launch_cg_assemble<T>(CGAssembleArgs<T> const &, CGAssembleMode, hipStream_t);
EXADG_INST(double)
EXADG_INST(float)
#undef EXADG_INST

}  // namespace ExaDG
