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

#include <atomic>

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

// out = A in for a centrosymmetric (Skew = false) or skew-centrosymmetric (Skew = true) A,
// stored in even-odd form when Packed (see cs_packed / gen_packed), else as the plain row-major
// matrix. Centro: out_i = E_i + O_i, out_{n-1-i} = E_i - O_i, middle output (odd n) from the
// even part; skew: out_{n-1-i} = O_i - E_i and the middle output from the odd part.
template<int N, bool Packed, bool Skew, typename Number>
__device__ __forceinline__ void eo_apply(Number const * __restrict__ T, Number const (&in)[N],
                                         Number (&out)[N])
{
  if constexpr(!Packed)
    mat_apply<N>(T, in, out);
  else
  {
    constexpr int  m = N / 2, me = m + N % 2;
    Number const * Ae = T;
    Number const * Ao = T + m * me;
    Number const * Am = Ao + m * m;
    Number         e[me], o[m];
#pragma unroll
    for(int j = 0; j < m; ++j)
    {
      e[j] = in[j] + in[N - 1 - j];
      o[j] = in[j] - in[N - 1 - j];
    }
    if constexpr(N % 2 == 1)
      e[m] = in[m];
#pragma unroll
    for(int i = 0; i < m; ++i)
    {
      Number E = 0, O = 0;
#pragma unroll
      for(int j = 0; j < me; ++j)
        E += Ae[i * me + j] * e[j];
#pragma unroll
      for(int j = 0; j < m; ++j)
        O += Ao[i * m + j] * o[j];
      out[i]         = E + O;
      out[N - 1 - i] = Skew ? O - E : E - O;
    }
    if constexpr(N % 2 == 1)
    {
      Number s = 0;
      if constexpr(Skew)
      {
#pragma unroll
        for(int j = 0; j < m; ++j)
          s += Am[j] * o[j];
      }
      else
      {
#pragma unroll
        for(int j = 0; j < me; ++j)
          s += Am[j] * e[j];
      }
      out[m] = s;
    }
  }
}

// the Cartesian kernel's centrosymmetric 1D products (CartTable)
template<int N, typename Number>
__device__ __forceinline__ void cs_apply(Number const * __restrict__ T, Number const (&in)[N],
                                         Number (&out)[N])
{
  eo_apply<N, cs_packed(N), false>(T, in, out);
}

// the general kernel's sweeps (GenTable): S and S^T centrosymmetric, Dq and Dq^T skew
template<int N, bool Skew, typename Number>
__device__ __forceinline__ void gen_apply(Number const * __restrict__ T, Number const (&in)[N],
                                          Number (&out)[N])
{
  eo_apply<N, gen_packed(N), Skew>(T, in, out);
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

// A store of an output that no kernel of the same apply reads back: marked non-temporal, so
// that it does not displace from the caches the neighbour cells that other workgroups are
// still reading. Used for the Cartesian vmult (FP64 k = 4: 7% faster on 64^3 cells, 3.5% on
// 32^3; FP32 unchanged). The Chebyshev and residual outputs are read by the next kernel of a
// V-cycle and stay cached (non-temporal there: V-cycle 0.8% slower), and the general kernel
// keeps plain stores (0-3% slower with these).
template<typename Number>
__device__ __forceinline__ void store_stream(Number * p, Number v)
{
  __builtin_nontemporal_store(v, p);
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
// then the transposed sweeps. Face terms between the two: the own traces from the Gauss
// values, the neighbour traces interpolated from the face nodes, flux with the per-point
// normal-derivative vectors a = J^{-1} n of both sides, and the face tests added to the
// Gauss-level residual before the transposed interpolation back to the nodes.

// FP64 at n = 5, 6: one face direction per round, so that the smaller workspace admits a fourth
// workgroup per CU
template<int N, typename Number>
constexpr bool general_single_rounds()
{
  return sizeof(Number) == 8 && (N == 5 || N == 6);
}

// The z derivative of the Gauss values taken after the face phase, so that it is not held
// across the face rounds: it removes the spills of FP32 n = 6 and FP64 n = 6 and lowers the
// register counts (FP64 n = 9 173 -> 132 VGPRs, three waves per SIMD instead of two), 4% for
// the V-cycles of the curved k = 5 hierarchy and 6% for the FP64 k = 5 vmult. FP64 at n = 8
// stays at three waves per SIMD either way and only pays the extra LDS reads (7%), except the
// Chebyshev step, which it brings from two waves to three (k = 7 V-cycle 8% faster).
template<int N, typename Number, ApplyMode mode>
constexpr bool general_late_gz()
{
  return !(sizeof(Number) == 8 && N == 8 && mode != ApplyMode::chebyshev);
}

template<int N, typename Number>
constexpr int general_work_size()
{
  // cell phase: the x and y gradients (the z one stays in registers); face phases: 8 trace
  // arrays per direction, two directions per round (x and y, then z) or one
  constexpr int faces = (general_single_rounds<N, Number>() ? 8 : 16) * N * N;
  return (2 * N * N * N > faces) ? 2 * N * N * N : faces;
}

// Face terms from the cell's values at the Gauss points. The own traces at the
// face Gauss points are one dot product per end of the lane's d-line of U (the Gauss values, in
// A at this point): u = sum_q l_q(0) U_q, du/dxi_d = sum_q l_q'(0) U_q with l_q the Lagrange
// polynomials on the Gauss points; only the neighbour traces (at the face nodes) need the 2D
// interpolation (S x S). The flux: per face point the traces and normal derivatives of both
// sides (a = J^{-1} n from the per-face data), laplace_operator.h:180-197 with the mirror
// values of weak_boundary_conditions.h on boundary faces; the tangential-gradient tests are
// folded into the value test by Dq^T along the face lines.
// The results end in four scalars per direction for the lane's d-line at Gauss indices (p, q):
// fr[d] = (value test, normal-derivative test) at the lower end, then at the upper end; the
// transposed derivative sweeps add l(0) fr[0] + l'(0) fr[1] + l(1) fr[2] + l'(1) fr[3] to that
// line of the Gauss-level residual. Compared with face rounds after the cell phase on nodal
// traces, this drops the 2D interpolation of the own-side traces and all transposed face
// interpolations; the face arrays of direction j live at W + 8 j N^2.
template<int N, typename Number, ApplyMode mode, int ND>
__device__ __forceinline__ void gauss_faces(int const (&dirs)[ND], SipgArgs<Number> const & a,
                                            int64_t cell, int const * cc, bool active, int lt,
                                            Number const * U, Number * W,
                                            Number const (&nbw)[3][2], Number const (&nbdw)[3][2],
                                            Number (&fr)[3][4])
{
  constexpr int  N2 = N * N;
  using T             = GenTable<N>;
  Number const * Dq = a.tables + T::DQ;
  Number const * d0 = a.tables + T::D0;
  Number const * d1 = a.tables + T::D1;
  Number const * SE = a.tables + T::SE;
  Number const * L0 = a.tables + T::L0;
  Number const * L1 = a.tables + T::L1;
  Number const * G0 = a.tables + T::G0;
  Number const * G1 = a.tables + T::G1;
  int const      p = lt % N, q = lt / N;
  int            type[ND][2];
#pragma unroll
  for(int j = 0; j < ND; ++j)
#pragma unroll
    for(int s = 0; s < 2; ++s)
    {
      int64_t const nb = neighbor_id(a.topo, cc, dirs[j], s);
      type[j][s]       = nb >= 0 ? kInterior : a.topo.bc[2 * dirs[j] + s];
    }
  // F1: own traces at the face Gauss points, neighbour traces at the face nodes
  if(active)
#pragma unroll
    for(int j = 0; j < ND; ++j)
    {
      int const d = dirs[j];
      Number *  E = W + 8 * j * N2;
      int       base, stride;
      line_geometry<N>(d, p, q, base, stride);
      Number v[N];
      load_line<N>(U + base, stride, v);
      E[0 * N2 + lt] = dot_line<N>(L0, v);
      E[1 * N2 + lt] = dot_line<N>(G0, v);
      E[4 * N2 + lt] = dot_line<N>(L1, v);
      E[5 * N2 + lt] = dot_line<N>(G1, v);
      Number nl = 0, dnl = 0, nr = 0, dnr = 0;
      if constexpr(mode != ApplyMode::diagonal)
      {
        if(type[j][0] == kInterior)
          nl = nbw[d][0], dnl = nbdw[d][0];
        if(type[j][1] == kInterior)
          nr = nbw[d][1], dnr = nbdw[d][1];
      }
      else
      {
        // probing the cell-local block: neighbours contribute zero, except a cell that is its
        // own periodic neighbour (one cell along d), whose other end then acts as neighbour
        Number u[N];
        source_line<N, Number, mode>(a.src, cell, base, stride, a.probe, u);
        if(neighbor_id(a.topo, cc, d, 0) == cell)
          nl = u[N - 1], dnl = dot_line<N>(d1, u);
        if(neighbor_id(a.topo, cc, d, 1) == cell)
          nr = u[0], dnr = dot_line<N>(d0, u);
      }
      E[2 * N2 + lt] = nl;
      E[3 * N2 + lt] = dnl;
      E[6 * N2 + lt] = nr;
      E[7 * N2 + lt] = dnr;
    }
  __syncthreads();
  // F2/F3: the neighbour arrays (slots 2, 3 of each side) to the face Gauss points
  if(active)
    for(int task = lt; task < 4 * ND * N; task += N2)
    {
      int const which = task / N, r = task % N;
      Number *  arr   = W + (8 * (which >> 2) + 4 * ((which >> 1) & 1) + 2 + (which & 1)) * N2;
      Number    v[N], o[N];
      load_line<N>(arr + N * r, 1, v);
      gen_apply<N, false>(SE, v, o);
      store_line<N>(arr + N * r, 1, o);
    }
  __syncthreads();
  if(active)
    for(int task = lt; task < 4 * ND * N; task += N2)
    {
      int const which = task / N, r = task % N;
      Number *  arr   = W + (8 * (which >> 2) + 4 * ((which >> 1) & 1) + 2 + (which & 1)) * N2;
      Number    v[N], o[N];
      load_line<N>(arr + r, N, v);
      gen_apply<N, false>(SE, v, o);
      store_line<N>(arr + r, N, o);
    }
  __syncthreads();
  // F4: fluxes at the face quadrature points (side s, point lt), held in registers until
  // every lane has read E
  Number res[ND][2][4];
  if(active)
#pragma unroll
    for(int j = 0; j < ND; ++j)
#pragma unroll
      for(int s = 0; s < 2; ++s)
      {
        int const d = dirs[j];
        int const t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        int const qq = lt, q1 = qq % N, q2 = qq / N;
        if(type[j][s] == EXADG_BC_NEUMANN)
        {  // homogeneous Neumann: u+ = u-, du+/dn = -du-/dn -> no contribution
          res[j][s][0] = res[j][s][1] = res[j][s][2] = res[j][s][3] = Number(0);
          continue;
        }
        int64_t const  f  = face_id(a.topo, cc, d, s);
        Number const * fd = a.face_data + f * 7 * N2;
        Number         aK[3], aN[3];
#pragma unroll
        for(int c = 0; c < 3; ++c)
        {
          Number const am = fd[c * N2 + qq], ap = fd[(3 + c) * N2 + qq];
          aK[c]           = s ? am : -ap;
          aN[c]           = s ? ap : -am;
        }
        Number const   w   = fd[6 * N2 + qq];
        Number const * Es  = W + 8 * j * N2 + s * 4 * N2;
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
        if(type[j][s] == kInterior)
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
        res[j][s][0]       = -(Number(0.5) * (dudn_K + dudn_N) - tau * jump) * w;
        res[j][s][1]       = sigma * aK[t1];
        res[j][s][2]       = sigma * aK[t2];
        res[j][s][3]       = sigma * aK[d];
      }
  __syncthreads();
  if(active)
#pragma unroll
    for(int j = 0; j < ND; ++j)
#pragma unroll
      for(int s = 0; s < 2; ++s)
#pragma unroll
        for(int c = 0; c < 4; ++c)
          W[(8 * j + 4 * s + c) * N2 + lt] = res[j][s][c];
  __syncthreads();
  // F5: tangential gradient tests folded into the value test (Dq^T); this lane's results
  if(active)
#pragma unroll
    for(int j = 0; j < ND; ++j)
#pragma unroll
      for(int s = 0; s < 2; ++s)
      {
        int const      qq = lt, q1 = qq % N, q2 = qq / N;
        Number const * R  = W + (8 * j + 4 * s) * N2;
        Number         acc = R[qq];
#pragma unroll
        for(int m = 0; m < N; ++m)
          acc += Dq[m * N + q1] * R[N2 + m + N * q2] + Dq[m * N + q2] * R[2 * N2 + q1 + N * m];
        fr[dirs[j]][2 * s]     = acc;
        fr[dirs[j]][2 * s + 1] = R[3 * N2 + qq];
      }
  __syncthreads();  // W is reused next
}

// Minimum waves per SIMD asked of the compiler for the general kernel, matched to what the
// LDS allows: FP64 at n = 3, 5, 6 four (n = 5, 6 with one face direction per round, whose
// smaller workspace admits a fourth workgroup per CU: n = 5 8% faster, n = 6 4%); FP32 at n = 3
// eight (<= 64 VGPRs, 19 KB of LDS per workgroup), at n = 4..6 seven (<= 72 VGPRs), at n = 7 six
template<int N, typename Number>
constexpr int general_min_waves()
{
  if constexpr(sizeof(Number) == 8)
    return (N == 5 || N == 6) ? 4 : N == 3 ? 4 : 1;
  else
    return N == 3 ? 8 : (N >= 4 && N <= 6) ? 7 : N == 7 ? 6 : 1;
}

template<int N, typename Number, ApplyMode mode>
__global__ void __launch_bounds__(256)
  __attribute__((amdgpu_waves_per_eu(general_min_waves<N, Number>()))) sipg_general_kernel(SipgArgs<Number> a)
{
  constexpr int  N2 = N * N, N3 = N * N * N, WS = general_work_size<N, Number>();
  constexpr int  CPB = cells_per_block(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *       smem = reinterpret_cast<Number *>(smem_raw);
  int const      slot = threadIdx.x / N2;
  int const      lt   = threadIdx.x - slot * N2;
  int const      pos0 = (int)xcd_logical_block(blockIdx.x, gridDim.x) * CPB;
  bool const     active = slot < CPB && pos0 + slot < (int)a.n_cells;
  Number *       A      = smem + (active ? slot : 0) * (N3 + WS);
  Number *       W      = A + N3;
  Number *       Gx = W, *Gy = W + N3;
  Number const * SE  = a.tables + GenTable<N>::SE;
  Number const * DE  = a.tables + GenTable<N>::DE;
  Number const * STE = a.tables + GenTable<N>::STE;
  Number const * DTE = a.tables + GenTable<N>::DTE;
  int const      p = lt % N, q = lt / N;
  int            cc[3] = {0, 0, 0};
  if(active)
    traversal_coords(a.topo, a.tile_rows, pos0, slot, cc);
  int64_t const cell = active ? cell_id(a.topo, cc) : 0;

  // neighbour traces of this lane's line in every direction: value and derivative at the
  // shared face, read up front so that the loads overlap the cell phases
  Number nbw[3][2] = {}, nbdw[3][2] = {};
  auto   load_traces = [&](int d) {
    if constexpr(mode != ApplyMode::diagonal)
      if(active)
      {
        Number const * d0 = a.tables + GenTable<N>::D0;
        Number const * d1 = a.tables + GenTable<N>::D1;
        int            base, stride;
        line_geometry<N>(d, p, q, base, stride);
#pragma unroll
        for(int sd = 0; sd < 2; ++sd)
        {
          int64_t const nbc = neighbor_id(a.topo, cc, d, sd);
          if(nbc >= 0)
          {
            Number w[N];
            load_line<N>(a.src + nbc * N3 + base, stride, w);
            nbw[d][sd]  = sd ? w[0] : w[N - 1];
            nbdw[d][sd] = dot_line<N>(sd ? d0 : d1, w);
          }
        }
      }
  };
  load_traces(0);
  load_traces(1);
  load_traces(2);

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
    gen_apply<N, false>(SE, v, o);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + p + N2 * q, N, v);
    gen_apply<N, false>(SE, v, o);
    store_line<N>(A + p + N2 * q, N, o);
  }
  __syncthreads();
  // z sweep, then the z derivative at the Gauss points: this lane's z-line (points lt + m N2),
  // kept in registers -- the same lane applies the metric at those points below. With
  // late_gz the derivative is taken after the face phase from the z-line in A instead, so that
  // it is not live across the face rounds (the same values)
  constexpr bool late_gz = general_late_gz<N, Number, mode>();
  Number         gz_line[N];
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + p + N * q, N2, v);
    gen_apply<N, false>(SE, v, o);
    store_line<N>(A + p + N * q, N2, o);
    if constexpr(!late_gz)
      gen_apply<N, true>(DE, o, gz_line);
  }
  __syncthreads();
  // the face terms from the Gauss values in A: four result scalars per direction per lane, added
  // by the transposed derivative sweeps below
  Number fr[3][4] = {};
  {
    if constexpr(general_single_rounds<N, Number>())
    {
      int const x[1] = {0}, y[1] = {1}, z[1] = {2};
      gauss_faces<N, Number, mode, 1>(x, a, cell, cc, active, lt, A, W, nbw, nbdw, fr);
      gauss_faces<N, Number, mode, 1>(y, a, cell, cc, active, lt, A, W, nbw, nbdw, fr);
      gauss_faces<N, Number, mode, 1>(z, a, cell, cc, active, lt, A, W, nbw, nbdw, fr);
    }
    else
    {
      int const xy[2] = {0, 1}, z[1] = {2};
      gauss_faces<N, Number, mode, 2>(xy, a, cell, cc, active, lt, A, W, nbw, nbdw, fr);
      gauss_faces<N, Number, mode, 1>(z, a, cell, cc, active, lt, A, W, nbw, nbdw, fr);
    }
  }
  Number const * L0 = a.tables + GenTable<N>::L0;
  Number const * L1 = a.tables + GenTable<N>::L1;
  Number const * G0 = a.tables + GenTable<N>::G0;
  Number const * G1 = a.tables + GenTable<N>::G1;
  // line o of the Gauss-level residual += the face tests of the lane's d-line
  auto add_face_tests = [&](Number(&o)[N], Number const(&f)[4]) {
#pragma unroll
    for(int i = 0; i < N; ++i)
      o[i] += L0[i] * f[0] + G0[i] * f[1] + L1[i] * f[2] + G1[i] * f[3];
  };
  if(active)
  {
    Number v[N], g[N];
    load_line<N>(A + N * (p + N * q), 1, v);
    gen_apply<N, true>(DE, v, g);
    store_line<N>(Gx + N * (p + N * q), 1, g);
    load_line<N>(A + p + N2 * q, N, v);
    gen_apply<N, true>(DE, v, g);
    store_line<N>(Gy + p + N2 * q, N, g);
    if constexpr(late_gz)
    {
      load_line<N>(A + p + N * q, N2, v);
      gen_apply<N, true>(DE, v, gz_line);
    }
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
      Number const gx = Gx[e], gy = Gy[e], gz = gz_line[m];
      Number const g00 = G[e], g11 = G[N3 + e], g22 = G[2 * N3 + e];
      Number const g01 = G[3 * N3 + e], g02 = G[4 * N3 + e], g12 = G[5 * N3 + e];
      Gx[e]            = g00 * gx + g01 * gy + g02 * gz;
      Gy[e]            = g01 * gx + g11 * gy + g12 * gz;
      gz_line[m]       = g02 * gx + g12 * gy + g22 * gz;
    }
  }
  __syncthreads();
  // transposed derivative sweeps summed into A, then S^T back to nodal values
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(Gx + N * (p + N * q), 1, v);
    gen_apply<N, true>(DTE, v, o);
    add_face_tests(o, fr[0]);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N], w[N];
    load_line<N>(Gy + p + N2 * q, N, v);
    gen_apply<N, true>(DTE, v, o);
    add_face_tests(o, fr[1]);
    load_line<N>(A + p + N2 * q, N, w);
#pragma unroll
    for(int i = 0; i < N; ++i)
      w[i] += o[i];
    store_line<N>(A + p + N2 * q, N, w);
  }
  __syncthreads();
  if(active)
  {
    Number o[N], w[N];
    gen_apply<N, true>(DTE, gz_line, o);
    add_face_tests(o, fr[2]);
    load_line<N>(A + p + N * q, N2, w);
#pragma unroll
    for(int i = 0; i < N; ++i)
      w[i] += o[i];
    gen_apply<N, false>(STE, w, o);
    store_line<N>(A + p + N * q, N2, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + p + N2 * q, N, v);
    gen_apply<N, false>(STE, v, o);
    store_line<N>(A + p + N2 * q, N, o);
  }
  __syncthreads();
  if(active)
  {
    Number v[N], o[N];
    load_line<N>(A + N * (p + N * q), 1, v);
    gen_apply<N, false>(STE, v, o);
    store_line<N>(A + N * (p + N * q), 1, o);
  }
  __syncthreads();

  if constexpr(mode == ApplyMode::residual_restrict)
  {
    // r = rhs - A u in A, then the three transposed interpolation sweeps x, y, z of the
    // restriction through the workspace; the coarse cell's values are added to dst
    int const      nc = a.restrict_nc;
    Number const * P  = a.restrict_P;  // P[i nc + c]: fine node i, coarse node c
    Number *       W1 = W;             // nc x N x N
    Number *       W2 = W + N3;        // nc x nc x N
    if(active)
      for(int e = lt; e < N3; e += N2)
        A[e] = a.rhs[cell * N3 + e] - A[e];
    __syncthreads();
    if(active)
    {
      int const j = lt % N, k = lt / N;  // x-line (j, k)
      for(int c = 0; c < nc; ++c)
      {
        Number s = 0;
#pragma unroll
        for(int i = 0; i < N; ++i)
          s += P[i * nc + c] * A[i + N * (j + N * k)];
        W1[c + nc * (j + N * k)] = s;
      }
    }
    __syncthreads();
    if(active && lt < nc * N)
    {
      int const c0 = lt % nc, k = lt / nc;  // y-line (c0, k)
      for(int c1 = 0; c1 < nc; ++c1)
      {
        Number s = 0;
#pragma unroll
        for(int j = 0; j < N; ++j)
          s += P[j * nc + c1] * W1[c0 + nc * (j + N * k)];
        W2[c0 + nc * (c1 + nc * k)] = s;
      }
    }
    __syncthreads();
    if(active && lt < nc * nc)
    {
      int const     c0 = lt % nc, c1 = lt / nc;  // z-line (c0, c1)
      int64_t const nc3 = (int64_t)nc * nc * nc;
      for(int c2 = 0; c2 < nc; ++c2)
      {
        Number s = 0;
#pragma unroll
        for(int k = 0; k < N; ++k)
          s += P[k * nc + c2] * W2[c0 + nc * (c1 + nc * k)];
        a.dst[cell * nc3 + c0 + nc * (c1 + nc * c2)] += s;
      }
    }
    return;
  }
  if(active)
  {
    if constexpr(mode == ApplyMode::diagonal)
    {
      if(a.block)
        for(int e = lt; e < N3; e += N2)
          a.dst[(cell * N3 + a.probe) * N3 + e] = A[e];
      else if(lt == 0)
        a.dst[cell * N3 + a.probe] = A[a.probe];
    }
    else if constexpr(mode == ApplyMode::vmult_add)
      for(int e = lt; e < N3; e += N2)
        a.dst[cell * N3 + e] += A[e];
    else if constexpr(mode == ApplyMode::residual)
      for(int e = lt; e < N3; e += N2)
        a.dst[cell * N3 + e] = a.rhs[cell * N3 + e] - A[e];
    else if constexpr(mode == ApplyMode::chebyshev)
      for(int e = lt; e < N3; e += N2)
      {
        int64_t const g = cell * N3 + e;
        Number        v = (Number(1) + a.f1) * a.src[g] + a.f2 * a.dinv[g] * (a.rhs[g] - A[e]);
        if(a.has_old)
          v -= a.f1 * a.x_old[g];
        a.dst[g] = v;
      }
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
// K/h_d plus, at each line end, the flux terms of laplace_operator.h:180-197 written with
// the exterior trace (value w and d/dx_d derivative dw at the shared face):
//   right end: o_i -= 1/2 d1_i [u], o_{n-1} += -1/2 (du + dw) + tau [u],  [u] = u_{n-1} - w
//   left end : o_i += 1/2 d0_i [u], o_0     +=  1/2 (du + dw) + tau [u],  [u] = u_0 - w
// For boundary faces (w, dw) come from the mirror principle (weak_boundary_conditions.h):
// Dirichlet w = -u, dw = du; Neumann w = u, dw = -du -- one branch-free code path.
// Evaluated as out = M_z (M_y a_x + M_x a_y) + M_y M_x a_z with a_d = L_d(u, traces) in four
// LDS phases (five workgroup barriers): the three a_d from x-, y- and z-lines; M_x a_y, M_x a_z
// (x-lines) and M_y a_x (y-lines); M_y M_x a_z (y-lines) and M_z (...) (z-lines); the sum along
// the z-lines. The neighbour lines are read once, at kernel start, and reduced to their traces.

// LDS strides (x, y, z, cell) per n. n = 4 and 7 are padded to spread the 8-byte line
// accesses of the three sweep orientations over the LDS banks; at n = 6 and 8 padding costs a
// workgroup per CU in FP64 (measured 6-10% slower), so those are dense.
template<int N>
struct LdsLayout
{
  static constexpr int PX = N;
  static constexpr int PY = (N == 4) ? 19 : (N == 7) ? 52 : N * N;
  static constexpr int PC = (N == 4) ? 80 : (N == 7) ? 369 : N * N * N;
};

// a = c L_d u along one line: tab holds the unit-interval K, d0, d1; hinv = c/h_d and tau the
// penalties times c, c being the scale of the Kronecker term. The neighbour derivative traces
// dwnb are unit-interval derivatives as well (same h_d on a uniform box).
template<int N, typename Number>
__device__ __forceinline__ void line_laplacian(Number const * __restrict__ tab, Number hinv,
                                               Number const (&u)[N], Number (&o)[N],
                                               int const (&type)[2], Number const (&tau)[2],
                                               Number const (&wnb)[2], Number const (&dwnb)[2])
{
  Number const * K  = tab + CartTable<N>::K;
  Number const * d0 = tab + CartTable<N>::D0;
  Number const * d1 = tab + CartTable<N>::D1;
  cs_apply<N>(K, u, o);
  Number const du0 = dot_line<N>(d0, u), du1 = dot_line<N>(d1, u);
  // exterior traces: neighbour values or mirror values
  Number const sl  = type[0] == EXADG_BC_DIRICHLET ? Number(-1) : Number(1);
  Number const sr  = type[1] == EXADG_BC_DIRICHLET ? Number(-1) : Number(1);
  Number const wl  = type[0] == kInterior ? wnb[0] : sl * u[0];
  Number const dwl = type[0] == kInterior ? dwnb[0] : -sl * du0;
  Number const wr  = type[1] == kInterior ? wnb[1] : sr * u[N - 1];
  Number const dwr = type[1] == kInterior ? dwnb[1] : -sr * du1;
  Number const jl = Number(0.5) * (u[0] - wl), jr = Number(0.5) * (u[N - 1] - wr);
#pragma unroll
  for(int i = 0; i < N; ++i)
    o[i] = hinv * (o[i] + d0[i] * jl - d1[i] * jr);
  o[0] += hinv * Number(0.5) * (du0 + dwl) + Number(2) * tau[0] * jl;
  o[N - 1] += Number(2) * tau[1] * jr - hinv * Number(0.5) * (du1 + dwr);
}

template<int N, typename Number, ApplyMode mode>
__global__ void __launch_bounds__(256) sipg_cartesian_kernel(SipgArgs<Number> a)
{
  constexpr int N2 = N * N, N3 = N * N * N;
  constexpr int CPB = cells_per_block(N);
  using L           = LdsLayout<N>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Number *      smem   = reinterpret_cast<Number *>(smem_raw);
  int const     slot   = threadIdx.x / N2;
  int const     lt     = threadIdx.x - slot * N2;
  int const     pos0   = (int)xcd_logical_block(blockIdx.x, gridDim.x) * CPB;
  bool const    active = slot < CPB && pos0 + slot < (int)a.n_cells;
  // three LDS buffers per cell: U (the cell's values), A and B; before the X phase A and B
  // stage the two y-neighbour cells
  Number *      U      = smem + (active ? slot : 0) * 3 * L::PC;
  Number *      Abuf   = U + L::PC;
  Number *      Bbuf   = Abuf + L::PC;
  int const     p = lt % N, q = lt / N;
  // LDS line bases of this lane for the x, y and z sweeps
  int const     bx = L::PX * p + L::PY * q, by = p + L::PY * q, bz = p + L::PX * q;

  int    type[3][2];
  Number tau[3][2], wnb[3][2], dwnb[3][2];
  Number own[N];  // this lane's z-line of the cell (lt + m N2), used by the Z phase
  int    cell       = 0;
  bool   in_block_l = false, in_block_r = false;
  if(active)
  {
    int cc[3];
    traversal_coords(a.topo, a.tile_rows, pos0, slot, cc);
    int const nx = a.topo.nc[0], ny = a.topo.nc[1];
    int const stride_c[3] = {1, nx, nx * ny};
    cell                  = cc[0] + nx * (cc[1] + ny * cc[2]);
    int nb[3][2];
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
      int const n_d = a.topo.nc[d], per = a.topo.periodic[d];
      int const t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      int const fbase =
        (int)a.topo.face_base[d] + a.topo.layers[d] * (cc[t1] + a.topo.nc[t1] * cc[t2]);
      nb[d][0] = cc[d] > 0 ? cell - stride_c[d] : (per ? cell + (n_d - 1) * stride_c[d] : -1);
      nb[d][1] = cc[d] < n_d - 1 ? cell + stride_c[d] : (per ? cell - (n_d - 1) * stride_c[d] : -1);
      int const fr = (per && cc[d] + 1 == n_d) ? 0 : cc[d] + 1;
      tau[d][0]    = a.tau_face[fbase + cc[d]];
      tau[d][1]    = a.tau_face[fbase + fr];
#pragma unroll
      for(int s = 0; s < 2; ++s)
        type[d][s] = nb[d][s] >= 0 ? kInterior : a.topo.bc[2 * d + s];
    }
    if constexpr(mode == ApplyMode::diagonal)
    {
#pragma unroll
      for(int m = 0; m < N; ++m)
      {
        own[m]            = (lt + m * N2 == a.probe) ? Number(1) : Number(0);
        U[bz + L::PY * m] = own[m];
      }
      // neighbours contribute zero, except a cell that is its own periodic neighbour
#pragma unroll
      for(int d = 0; d < 3; ++d)
      {
        Number const * tb = a.tab;
        int const      gbase = d == 0 ? N * (p + N * q) : d == 1 ? p + N2 * q : p + N * q;
        int const      gstr  = d == 0 ? 1 : d == 1 ? N : N2;
        Number         w[N];
#pragma unroll
        for(int i = 0; i < N; ++i)
          w[i] = (gbase + i * gstr == a.probe) ? Number(1) : Number(0);
        bool const self = nb[d][0] == cell;
        wnb[d][0]       = self ? w[N - 1] : Number(0);
        dwnb[d][0]      = self ? dot_line<N>(tb + CartTable<N>::D1, w) : Number(0);
        wnb[d][1]       = self ? w[0] : Number(0);
        dwnb[d][1]      = self ? dot_line<N>(tb + CartTable<N>::D0, w) : Number(0);
      }
    }
    else
    {
      // All global loads up front. The own cell and the z- and y-neighbour cells are read in
      // the same pattern (lane lt reads DoFs lt + m N^2: consecutive lanes, consecutive
      // addresses); a lane then holds z-lines, so the z traces are lane-local and the
      // y-neighbour cells go through LDS (A, B) to be read as y-lines. Only x-neighbours at
      // the workgroup's ends are read as strided x-lines; the others come from U.
      Number const * own_p = a.src + (size_t)cell * N3 + lt;
      in_block_l = slot > 0 && cc[0] > 0;
      in_block_r = slot + 1 < CPB && cc[0] + 1 < a.topo.nc[0] && pos0 + slot + 1 < (int)a.n_cells;
      Number zl[N], zr[N], yl[N], yr[N], xl[N], xr[N];
      Number const * pz0 = a.src + (size_t)(nb[2][0] >= 0 ? nb[2][0] : cell) * N3 + lt;
      Number const * pz1 = a.src + (size_t)(nb[2][1] >= 0 ? nb[2][1] : cell) * N3 + lt;
      Number const * py0 = a.src + (size_t)(nb[1][0] >= 0 ? nb[1][0] : cell) * N3 + lt;
      Number const * py1 = a.src + (size_t)(nb[1][1] >= 0 ? nb[1][1] : cell) * N3 + lt;
#pragma unroll
      for(int m = 0; m < N; ++m)
      {
        own[m] = own_p[m * N2];
        zl[m]  = pz0[m * N2];
        zr[m]  = pz1[m * N2];
        yl[m]  = py0[m * N2];
        yr[m]  = py1[m * N2];
      }
      int const xbase = N * (p + N * q);  // this lane's x-line (y = p, z = q)
      {
        Number const * px0 = a.src + (size_t)(nb[0][0] >= 0 ? nb[0][0] : cell) * N3 + xbase;
        Number const * px1 = a.src + (size_t)(nb[0][1] >= 0 ? nb[0][1] : cell) * N3 + xbase;
#pragma unroll
        for(int i = 0; i < N; ++i)
        {
          xl[i] = in_block_l ? Number(0) : px0[i];
          xr[i] = in_block_r ? Number(0) : px1[i];
        }
      }
#pragma unroll
      for(int m = 0; m < N; ++m)
      {
        U[bz + L::PY * m]    = own[m];
        Abuf[bz + L::PY * m] = yl[m];
        Bbuf[bz + L::PY * m] = yr[m];
      }
      // lane-local traces (x ones are replaced from LDS for in-block neighbours)
      Number const * tzb = a.tab;
      Number const * txb = a.tab;
      wnb[2][0]          = zl[N - 1];
      dwnb[2][0]         = dot_line<N>(tzb + CartTable<N>::D1, zl);
      wnb[2][1]          = zr[0];
      dwnb[2][1]         = dot_line<N>(tzb + CartTable<N>::D0, zr);
      wnb[0][0]          = xl[N - 1];
      dwnb[0][0]         = dot_line<N>(txb + CartTable<N>::D1, xl);
      wnb[0][1]          = xr[0];
      dwnb[0][1]         = dot_line<N>(txb + CartTable<N>::D0, xr);
    }
  }
  __syncthreads();
  Number const * tx = a.tab;  // the same unit-interval tables in every direction
  Number const * ty = a.tab;
  Number const * tz = a.tab;
  Number const   hx = a.h[0], hy = a.h[1], hz = a.h[2];
  Number const   hxi = a.h[3], hyi = a.h[4], hzi = a.h[5];

  // Scale factors of the three Kronecker terms: with M_d = h_d M (unit-interval mass M),
  //   out = M_z M_y (c_x L_x u) + M_z M_x (c_y L_y u) + M_y M_x (c_z L_z u),
  // c_x = h_y h_z, c_y = h_x h_z, c_z = h_x h_y, so that every mass sweep is a plain M.
  Number const cx = hy * hz, cy = hx * hz, cz = hx * hy;
  // P1: the remaining neighbour traces, then the three line Laplacians into registers
  Number rx[N], ry[N], rz[N];
  if(active)
  {
    if constexpr(mode != ApplyMode::diagonal)
    {
      // y traces from the staged neighbour cells, read as this lane's y-line (x = p, z = q)
      Number w[N];
      load_line<N>(Abuf + by, L::PX, w);
      wnb[1][0]  = w[N - 1];
      dwnb[1][0] = dot_line<N>(ty + CartTable<N>::D1, w);
      load_line<N>(Bbuf + by, L::PX, w);
      wnb[1][1]  = w[0];
      dwnb[1][1] = dot_line<N>(ty + CartTable<N>::D0, w);
      if(in_block_l)
      {
        load_line<N>(U - 3 * L::PC + bx, 1, w);
        wnb[0][0]  = w[N - 1];
        dwnb[0][0] = dot_line<N>(tx + CartTable<N>::D1, w);
      }
      if(in_block_r)
      {
        load_line<N>(U + 3 * L::PC + bx, 1, w);
        wnb[0][1]  = w[0];
        dwnb[0][1] = dot_line<N>(tx + CartTable<N>::D0, w);
      }
    }
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
      Number const c = d == 0 ? cx : d == 1 ? cy : cz;
      tau[d][0] *= c;
      tau[d][1] *= c;
    }
    Number u[N];
    load_line<N>(U + bx, 1, u);
    line_laplacian<N>(tx, cx * hxi, u, rx, type[0], tau[0], wnb[0], dwnb[0]);
    load_line<N>(U + by, L::PX, u);
    line_laplacian<N>(ty, cy * hyi, u, ry, type[1], tau[1], wnb[1], dwnb[1]);
    line_laplacian<N>(tz, cz * hzi, own, rz, type[2], tau[2], wnb[2], dwnb[2]);
  }
  __syncthreads();  // U, A and B are read across lanes up to here
  if(active)
  {
    store_line<N>(Abuf + bx, 1, rx);
    store_line<N>(Bbuf + by, L::PX, ry);
    store_line<N>(U + bz, L::PY, rz);
  }
  __syncthreads();
  // P2: B <- M_x B and U <- M_x U along x-lines, A <- M_y A along y-lines (each line in place:
  // in one phase every line of a buffer belongs to one lane)
  if(active)
  {
    Number v[N], w[N];
    load_line<N>(Bbuf + bx, 1, v);
    cs_apply<N>(tx + CartTable<N>::M, v, w);
    store_line<N>(Bbuf + bx, 1, w);
    load_line<N>(U + bx, 1, v);
    cs_apply<N>(tx + CartTable<N>::M, v, w);
    store_line<N>(U + bx, 1, w);
    load_line<N>(Abuf + by, L::PX, v);
    cs_apply<N>(ty + CartTable<N>::M, v, w);
    store_line<N>(Abuf + by, L::PX, w);
  }
  __syncthreads();
  // P3: U <- M_y U along y-lines (the finished z-term); outp = M_z (A + B) along z-lines
  Number outp[N];
  if(active)
  {
    Number v[N], w[N];
    load_line<N>(U + by, L::PX, v);
    cs_apply<N>(ty + CartTable<N>::M, v, w);
    store_line<N>(U + by, L::PX, w);
    load_line<N>(Abuf + bz, L::PY, v);
    load_line<N>(Bbuf + bz, L::PY, w);
#pragma unroll
    for(int i = 0; i < N; ++i)
      v[i] += w[i];
    cs_apply<N>(tz + CartTable<N>::M, v, outp);
  }
  __syncthreads();
  if constexpr(mode == ApplyMode::residual_restrict)
  {
    // r = rhs - A u along the lane's z-line, restricted along z in registers, then along x and
    // y through A and B (free after the barrier above); the coarse cell is added to dst
    int const      nc = a.restrict_nc;
    Number const * P  = a.restrict_P;  // P[i nc + c]: fine node i, coarse node c
    if(active)
    {
      Number v[N], r[N];
      load_line<N>(U + bz, L::PY, v);
      size_t const g0 = (size_t)cell * N3 + lt;
#pragma unroll
      for(int k = 0; k < N; ++k)
        r[k] = a.rhs[g0 + k * N2] - (outp[k] + v[k]);
#pragma unroll
      for(int c = 0; c < N; ++c)
        if(c < nc)
        {
          Number s = 0;
#pragma unroll
          for(int k = 0; k < N; ++k)
            s += P[k * nc + c] * r[k];
          Abuf[p + N * q + N2 * c] = s;  // (x, y, c_z)
        }
    }
    __syncthreads();
    if(active && lt < N * nc)
    {
      int const y = lt % N, cz = lt / N;
      for(int cx = 0; cx < nc; ++cx)
      {
        Number s = 0;
#pragma unroll
        for(int x = 0; x < N; ++x)
          s += P[x * nc + cx] * Abuf[x + N * y + N2 * cz];
        Bbuf[cx + nc * (y + N * cz)] = s;
      }
    }
    __syncthreads();
    if(active && lt < nc * nc)
    {
      int const     cx = lt % nc, cz = lt / nc;
      int64_t const nc3 = (int64_t)nc * nc * nc;
      for(int cy = 0; cy < nc; ++cy)
      {
        Number s = 0;
#pragma unroll
        for(int y = 0; y < N; ++y)
          s += P[y * nc + cy] * Bbuf[cx + nc * (y + N * cz)];
        a.dst[(int64_t)cell * nc3 + cx + nc * (cy + nc * cz)] += s;
      }
    }
    return;
  }
  // out = outp + the z-term held in U, along the z-lines
  if constexpr(mode == ApplyMode::diagonal)
  {
    if(active)
    {
      Number v[N];
      load_line<N>(U + bz, L::PY, v);
#pragma unroll
      for(int k = 0; k < N; ++k)
        if(a.block)
          a.dst[((size_t)cell * N3 + a.probe) * N3 + lt + k * N2] = outp[k] + v[k];
        else if(lt + k * N2 == a.probe)
          a.dst[(size_t)cell * N3 + a.probe] = outp[k] + v[k];
    }
  }
  else if(active)
  {
    // direct per-lane stores: lane (p, q) owns the z-line of DoFs lt + k N2 (a workgroup-wide
    // LDS-staged store of the block's contiguous cells measured 10% slower)
    Number v[N];
    load_line<N>(U + bz, L::PY, v);
    size_t const g0 = (size_t)cell * N3 + lt;
#pragma unroll
    for(int k = 0; k < N; ++k)
    {
      Number const ax = outp[k] + v[k];
      size_t const g  = g0 + k * N2;
      if constexpr(mode == ApplyMode::vmult)
        store_stream(&a.dst[g], ax);
      else if constexpr(mode == ApplyMode::vmult_add)
        a.dst[g] += ax;
      else if constexpr(mode == ApplyMode::residual)
        a.dst[g] = a.rhs[g] - ax;
      else
      {
        Number r = (Number(1) + a.f1) * own[k] + a.f2 * a.dinv[g] * (a.rhs[g] - ax);
        if(a.has_old)
          r -= a.f1 * a.x_old[g];
        a.dst[g] = r;
      }
    }
  }
}

template<int N, typename Number, ApplyMode mode>
hipError_t launch_general_n(SipgArgs<Number> const & a, hipStream_t stream)
{
  constexpr int CPB    = cells_per_block(N);
  int const     blocks = (int)((a.n_cells + CPB - 1) / CPB);
  size_t const  lds    = sizeof(Number) * CPB * (N * N * N + general_work_size<N, Number>());
  // set once per instantiation; concurrent first calls may both set it, which is harmless
  static std::atomic<bool> attr{false};
  if(lds > 65536 && !attr.load(std::memory_order_acquire))
  {
    hipError_t e = hipFuncSetAttribute((void const *)sipg_general_kernel<N, Number, mode>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if(e != hipSuccess)
      return e;
    attr.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((sipg_general_kernel<N, Number, mode>), dim3(blocks), dim3(256), lds, stream, a);
  return hipGetLastError();
}

template<int N, typename Number, ApplyMode mode>
hipError_t launch_cartesian_n(SipgArgs<Number> const & a, hipStream_t stream)
{
  constexpr int CPB    = cells_per_block(N);
  int const     blocks = (int)((a.n_cells + CPB - 1) / CPB);
  size_t const  lds    = sizeof(Number) * CPB * 3 * LdsLayout<N>::PC;
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
// diagonal entry i of the 1D cell-local line operator along a direction of width h: the line
// Laplacian of the Cartesian kernel applied to the unit vector e_i, with the exterior trace of
// weak_boundary_conditions.h on boundary faces (type_l/type_r EXADG_BC_*), zero across interior
// faces, or the cell's own opposite end across the periodic faces of a one-cell direction (self)
__device__ inline double line_diagonal(double const * t, int n, int i, double h, int type_l,
                                       int type_r, bool self, double tau_l, double tau_r)
{
  double const * M = t;
  double const * K = t + n;
  double const * d0 = t + 2 * n;
  double const * d1 = t + 3 * n;
  (void)M;
  double const u0 = i == 0 ? 1.0 : 0.0, un = i == n - 1 ? 1.0 : 0.0;
  double const du0 = d0[i], du1 = d1[i];
  double       wl = 0, dwl = 0, wr = 0, dwr = 0;
  if(type_l == EXADG_BC_DIRICHLET)
    wl = -u0, dwl = du0;
  else if(type_l == EXADG_BC_NEUMANN)
    wl = u0, dwl = -du0;
  else if(self)
    wl = un, dwl = du1;
  if(type_r == EXADG_BC_DIRICHLET)
    wr = -un, dwr = du1;
  else if(type_r == EXADG_BC_NEUMANN)
    wr = un, dwr = -du1;
  else if(self)
    wr = u0, dwr = du0;
  double const jl = 0.5 * (u0 - wl), jr = 0.5 * (un - wr);
  double       o  = (K[i] + d0[i] * jl - d1[i] * jr) / h;
  if(i == 0)
    o += 0.5 * (du0 + dwl) / h + 2.0 * tau_l * jl;
  if(i == n - 1)
    o += 2.0 * tau_r * jr - 0.5 * (du1 + dwr) / h;
  return o;
}

template<typename Number>
__global__ void __launch_bounds__(256)
  cartesian_diagonal_kernel(Number * diag, double const * t, Number const * tau_face,
                            BoxTopology topo, double h0, double h1, double h2, int n,
                            int64_t n_dofs)
{
  int64_t const g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(g >= n_dofs)
    return;
  int const     n3   = n * n * n;
  int64_t const cell = g / n3;
  int const     e    = (int)(g - cell * n3);
  int const     id[3] = {e % n, (e / n) % n, e / (n * n)};
  int           cc[3];
  cell_coords(topo, cell, cc);
  double const h[3] = {h0, h1, h2};
  double       L[3], m[3];
#pragma unroll
  for(int d = 0; d < 3; ++d)
  {
    int const  n_d = topo.nc[d], per = topo.periodic[d];
    int const  t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    bool const has_l = cc[d] > 0 || per, has_r = cc[d] < n_d - 1 || per;
    bool const self  = per && n_d == 1;
    int64_t const fbase = topo.face_base[d] + (int64_t)topo.layers[d] * (cc[t1] + (int64_t)topo.nc[t1] * cc[t2]);
    int const     fr    = (per && cc[d] + 1 == n_d) ? 0 : cc[d] + 1;
    int const     tl = has_l ? -1 : topo.bc[2 * d], tr = has_r ? -1 : topo.bc[2 * d + 1];
    L[d] = line_diagonal(t, n, id[d], h[d], tl, tr, self, (double)tau_face[fbase + cc[d]],
                         (double)tau_face[fbase + fr]);
    m[d] = h[d] * t[id[d]];
  }
  diag[g] = (Number)(L[0] * m[1] * m[2] + m[0] * L[1] * m[2] + m[0] * m[1] * L[2]);
}
template<int N, typename Number>
__global__ void __launch_bounds__(256)
  general_diagonal_kernel(Number * diag, double const * __restrict__ t1d, SipgArgs<Number> a)
{
  constexpr int  N2 = N * N, N3 = N * N * N;
  constexpr int  CPB = cells_per_block(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *       smem   = reinterpret_cast<double *>(smem_raw);
  int const      slot   = threadIdx.x / N2;
  int const      lt     = threadIdx.x - slot * N2;
  int const      pos0   = (int)xcd_logical_block(blockIdx.x, gridDim.x) * CPB;
  bool const     active = slot < CPB && pos0 + slot < (int)a.n_cells;
  double *       W1     = smem + (active ? slot : 0) * 2 * N3;
  double *       W2     = W1 + N3;
  double const * SS = t1d, *DD = t1d + N2, *SD = t1d + 2 * N2;
  int const      p = lt % N, q = lt / N;
  int            cc[3] = {0, 0, 0};
  if(active)
    traversal_coords(a.topo, a.tile_rows, pos0, slot, cc);
  int64_t const cell = active ? cell_id(a.topo, cc) : 0;
  double        acc[N];
#pragma unroll
  for(int m = 0; m < N; ++m)
    acc[m] = 0;
  // cell term: component c of G (g00, g11, g22, g01, g02, g12) against tables per direction
  for(int term = 0; term < 6; ++term)
  {
    int const      c  = term;
    double const * T0 = term == 0 ? DD : (term == 3 || term == 4) ? SD : SS;
    double const * T1 = term == 1 ? DD : (term == 3 || term == 5) ? SD : SS;
    double const * T2 = term == 2 ? DD : (term == 4 || term == 5) ? SD : SS;
    double const   f  = term < 3 ? 1.0 : 2.0;
    if(active)
    {
      Number const * G = a.cell_metric + cell * 6 * N3 + (int64_t)c * N3 + N * (p + N * q);
      double         g[N];
#pragma unroll
      for(int q0 = 0; q0 < N; ++q0)
        g[q0] = (double)G[q0];
#pragma unroll
      for(int i0 = 0; i0 < N; ++i0)
      {
        double s = 0;
#pragma unroll
        for(int q0 = 0; q0 < N; ++q0)
          s += T0[q0 * N + i0] * g[q0];
        W1[i0 + N * (p + N * q)] = s;  // (i0, q1 = p, q2 = q)
      }
    }
    __syncthreads();
    if(active)
    {
      double v[N];
#pragma unroll
      for(int q1 = 0; q1 < N; ++q1)
        v[q1] = W1[p + N * (q1 + N * q)];  // (i0 = p, q1, q2 = q)
#pragma unroll
      for(int i1 = 0; i1 < N; ++i1)
      {
        double s = 0;
#pragma unroll
        for(int q1 = 0; q1 < N; ++q1)
          s += T1[q1 * N + i1] * v[q1];
        W2[p + N * (i1 + N * q)] = s;
      }
    }
    __syncthreads();
    if(active)
    {
      double v[N];
#pragma unroll
      for(int q2 = 0; q2 < N; ++q2)
        v[q2] = W2[p + N * (q + N * q2)];  // (i0 = p, i1 = q, q2)
#pragma unroll
      for(int i2 = 0; i2 < N; ++i2)
      {
        double s = 0;
#pragma unroll
        for(int q2 = 0; q2 < N; ++q2)
          s += T2[q2 * N + i2] * v[q2];
        acc[i2] += f * s;
      }
    }
  }
  // face terms of the face-layer DoFs
  if(active)
  {
    double const lp[2] = {t1d[3 * N2], t1d[3 * N2 + 1]};  // l'(0) of node 0, l'(1) of node N-1
#pragma unroll
    for(int d = 0; d < 3; ++d)
#pragma unroll
      for(int s = 0; s < 2; ++s)
      {
        int const end = s ? N - 1 : 0;
        // this lane's DoFs on the face: (i_t1, i_t2) and the slot m of its z-line
        int it1, it2, mfix = -1;
        if(d == 0)
        {
          if(p != end)
            continue;
          it1 = q;  // i1, and i2 runs over m
          it2 = -1;
        }
        else if(d == 1)
        {
          if(q != end)
            continue;
          it1 = p;  // i0, and i2 runs over m
          it2 = -1;
        }
        else
        {
          it1  = p;
          it2  = q;
          mfix = end;
        }
        int64_t const nb = neighbor_id(a.topo, cc, d, s);
        int const     type = nb >= 0 ? kInterior : a.topo.bc[2 * d + s];
        double const  cf   = type == kInterior ? 1.0 : type == EXADG_BC_DIRICHLET ? 2.0 : 0.0;
        if(cf == 0.0)
          continue;
        int const      t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        int64_t const  fid = face_id(a.topo, cc, d, s);
        Number const * fd  = a.face_data + fid * 7 * N2;
        double const   tau = (double)a.tau_face[fid];
        // out(i_t1, i_t2) = sum_{q1,q2} CA SS SS + CB SD SS + CC SS SD
        double out[N];
#pragma unroll
        for(int m = 0; m < N; ++m)
          out[m] = 0;
        for(int q2 = 0; q2 < N; ++q2)
        {
          double rA = 0, rB = 0, rC = 0;  // sum over q1 with the it1 tables
#pragma unroll
          for(int q1 = 0; q1 < N; ++q1)
          {
            int const    qq = q1 + N * q2;
            double const am_d = (double)fd[d * N2 + qq], ap_d = (double)fd[(3 + d) * N2 + qq];
            double const am_1 = (double)fd[t1 * N2 + qq], ap_1 = (double)fd[(3 + t1) * N2 + qq];
            double const am_2 = (double)fd[t2 * N2 + qq], ap_2 = (double)fd[(3 + t2) * N2 + qq];
            double const ad = s ? am_d : -ap_d, a1 = s ? am_1 : -ap_1, a2 = s ? am_2 : -ap_2;
            double const w  = (double)fd[6 * N2 + qq];
            double const CA = cf * w * (tau - ad * lp[s]);
            double const CB = -cf * w * a1;
            double const CC = -cf * w * a2;
            rA += CA * SS[q1 * N + it1];
            rB += CB * SD[q1 * N + it1];
            rC += CC * SS[q1 * N + it1];
          }
          if(mfix >= 0)
            out[mfix] += (rA + rB) * SS[q2 * N + it2] + rC * SD[q2 * N + it2];
          else
#pragma unroll
            for(int m = 0; m < N; ++m)
              out[m] += (rA + rB) * SS[q2 * N + m] + rC * SD[q2 * N + m];
        }
#pragma unroll
        for(int m = 0; m < N; ++m)
          acc[m] += out[m];
      }
#pragma unroll
    for(int m = 0; m < N; ++m)
      diag[cell * N3 + p + N * q + N2 * m] = (Number)acc[m];
  }
}

template<int N, typename Number>
hipError_t launch_general_diagonal_n(Number * diag, double const * t1d, SipgArgs<Number> const & a,
                                     hipStream_t stream)
{
  constexpr int CPB    = cells_per_block(N);
  int const     blocks = (int)((a.n_cells + CPB - 1) / CPB);
  size_t const  lds    = sizeof(double) * CPB * 2 * N * N * N;
  static std::atomic<bool> attr{false};
  if(lds > 65536 && !attr.load(std::memory_order_acquire))
  {
    hipError_t e = hipFuncSetAttribute((void const *)general_diagonal_kernel<N, Number>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if(e != hipSuccess)
      return e;
    attr.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((general_diagonal_kernel<N, Number>), dim3(blocks), dim3(256), lds, stream,
                     diag, t1d, a);
  return hipGetLastError();
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
    case ApplyMode::chebyshev: return dispatch_general<Number, ApplyMode::chebyshev>(degree, a, s);
    case ApplyMode::residual: return dispatch_general<Number, ApplyMode::residual>(degree, a, s);
    case ApplyMode::residual_restrict:
      if(a.restrict_P == nullptr || a.restrict_nc < 1 || a.restrict_nc > degree + 1)
        return hipErrorInvalidValue;
      return dispatch_general<Number, ApplyMode::residual_restrict>(degree, a, s);
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
    case ApplyMode::chebyshev:
      return dispatch_cartesian<Number, ApplyMode::chebyshev>(degree, a, s);
    case ApplyMode::residual: return dispatch_cartesian<Number, ApplyMode::residual>(degree, a, s);
    case ApplyMode::residual_restrict:
      if(a.restrict_P == nullptr || a.restrict_nc < 1 || a.restrict_nc > degree + 1)
        return hipErrorInvalidValue;
      return dispatch_cartesian<Number, ApplyMode::residual_restrict>(degree, a, s);
  }
  return hipErrorInvalidValue;
}

template<typename Number>
hipError_t launch_cartesian_diagonal(Number * diag, double const * diag1d, Number const * tau_face,
                                     BoxTopology const & topo, double const (&h)[3], int n,
                                     int64_t n_cells, hipStream_t stream)
{
  int64_t const n_dofs = n_cells * n * n * n;
  if(n_dofs == 0)
    return hipSuccess;
  int64_t const blocks = (n_dofs + 255) / 256;
  hipLaunchKernelGGL((cartesian_diagonal_kernel<Number>), dim3((unsigned)blocks), dim3(256), 0,
                     stream, diag, diag1d, tau_face, topo, h[0], h[1], h[2], n, n_dofs);
  return hipGetLastError();
}

template<typename Number>
hipError_t launch_general_diagonal(Number * diag, double const * tab1d, SipgArgs<Number> const & a,
                                   int degree, hipStream_t stream)
{
  switch(degree)
  {
    case 1: return launch_general_diagonal_n<2, Number>(diag, tab1d, a, stream);
    case 2: return launch_general_diagonal_n<3, Number>(diag, tab1d, a, stream);
    case 3: return launch_general_diagonal_n<4, Number>(diag, tab1d, a, stream);
    case 4: return launch_general_diagonal_n<5, Number>(diag, tab1d, a, stream);
    case 5: return launch_general_diagonal_n<6, Number>(diag, tab1d, a, stream);
    case 6: return launch_general_diagonal_n<7, Number>(diag, tab1d, a, stream);
    case 7: return launch_general_diagonal_n<8, Number>(diag, tab1d, a, stream);
    case 8: return launch_general_diagonal_n<9, Number>(diag, tab1d, a, stream);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_general_diagonal<double>(double *, double const *,
                                                    SipgArgs<double> const &, int, hipStream_t);
template hipError_t launch_general_diagonal<float>(float *, double const *,
                                                   SipgArgs<float> const &, int, hipStream_t);
template hipError_t launch_cartesian_diagonal<double>(double *, double const *, double const *,
                                                      BoxTopology const &, double const (&)[3], int,
                                                      int64_t, hipStream_t);
template hipError_t launch_cartesian_diagonal<float>(float *, double const *, float const *,
                                                     BoxTopology const &, double const (&)[3], int,
                                                     int64_t, hipStream_t);
template hipError_t launch_sipg_general<double>(int, ApplyMode, SipgArgs<double> const &, hipStream_t);
template hipError_t launch_sipg_general<float>(int, ApplyMode, SipgArgs<float> const &, hipStream_t);
template hipError_t launch_sipg_cartesian<double>(int, ApplyMode, SipgArgs<double> const &, hipStream_t);
template hipError_t launch_sipg_cartesian<float>(int, ApplyMode, SipgArgs<float> const &, hipStream_t);

}  // namespace ExaDG
