// Kernels of the learned initial guess (per-cell MLP, FP32): a tiled GEMM on the matrix cores
// with strided operands, feature gathering from a solution history, ReLU and MSE backward,
// bias gradients, Adam; plus the FP64 extrapolation of the initial guess
// (time_integration/extrapolation_constants.cpp:55-93) and the right-hand side of the
// pressure sequence: the weak DG divergence with central flux (divergence_operator.h:63-68,
// divergence_operator.cpp:190-215) of an analytic Fourier velocity field.
#include <hip/hip_runtime.h>

#include <mutex>

#include "guess_kernels.h"

namespace ExaDG
{
namespace
{
constexpr int TM = 64, TN = 64, TK = 16;

typedef float f32x4 __attribute__((ext_vector_type(4)));

// C(i,j) = alpha * sum_k A(i,k) B(k,j) (+ C) (+ bias_j), optional ReLU, on the matrix cores:
// each of the 4 waves computes a 32 x 32 quadrant of the 64 x 64 tile as 2 x 2 products
// v_mfma_f32_16x16x4f32 (FP32 operands, FP32 accumulation). Lane l supplies A(l%16, l/16) and
// B(l/16, l%16) of a 16 x 4 / 4 x 16 step and holds D(4 (l/16) + r, l%16), r = 0..3.
// A(i,k) = A[i*a_i + k*a_k], B(k,j) = B[k*b_k + j*b_j], C row-major with leading dim ldc.
// With gridDim.z > 1 (split K), block z covers k in [z*kc, min(K, (z+1)*kc)) and writes its
// raw partial product to C + z * M * ldc; splitk_reduce_kernel applies alpha/C/bias/ReLU.
__global__ void __launch_bounds__(256)
  sgemm_kernel(int M, int N, int K, float const * __restrict__ A, int64_t a_i, int64_t a_k,
               float const * __restrict__ B, int64_t b_k, int64_t b_j, float * __restrict__ C,
               int64_t ldc, float alpha, int accumulate, float const * __restrict__ bias, int relu,
               int kc)
{
  __shared__ float As[TK][TM + 1];
  __shared__ float Bs[TK][TN + 1];
  int const tid = threadIdx.x;
  int const i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  int const kb = blockIdx.z * kc, ke = min(K, kb + kc);
  bool const split = gridDim.z > 1;
  if(split)
  {
    C += (int64_t)blockIdx.z * M * ldc;
    alpha = 1.f, accumulate = 0, bias = nullptr, relu = 0;
  }
  int const lane = tid & 63, wave = tid >> 6;
  int const wr = 32 * (wave >> 1), wc = 32 * (wave & 1);  // this wave's quadrant
  int const li = lane & 15, lk = lane >> 4;
  f32x4     acc[2][2];
#pragma unroll
  for(int rb = 0; rb < 2; ++rb)
#pragma unroll
    for(int cb = 0; cb < 2; ++cb)
      acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
  for(int k0 = kb; k0 < ke; k0 += TK)
  {
    // 64x16 tiles of A and 16x64 of B: 4 elements per thread each
    for(int e = tid; e < TM * TK; e += 256)
    {
      int const r = e / TK, c = e % TK;
      int const i = i0 + r, k = k0 + c;
      As[c][r]    = (i < M && k < ke) ? A[i * a_i + k * a_k] : 0.f;
    }
    for(int e = tid; e < TK * TN; e += 256)
    {
      int const r = e / TN, c = e % TN;
      int const k = k0 + r, j = j0 + c;
      Bs[r][c]    = (k < ke && j < N) ? B[k * b_k + j * b_j] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for(int kk = 0; kk < TK; kk += 4)
    {
      float a[2], b[2];
#pragma unroll
      for(int rb = 0; rb < 2; ++rb)
        a[rb] = As[kk + lk][wr + 16 * rb + li];
#pragma unroll
      for(int cb = 0; cb < 2; ++cb)
        b[cb] = Bs[kk + lk][wc + 16 * cb + li];
#pragma unroll
      for(int rb = 0; rb < 2; ++rb)
#pragma unroll
        for(int cb = 0; cb < 2; ++cb)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for(int rb = 0; rb < 2; ++rb)
#pragma unroll
    for(int cb = 0; cb < 2; ++cb)
#pragma unroll
      for(int r = 0; r < 4; ++r)
      {
        int const i = i0 + wr + 16 * rb + 4 * lk + r;
        int const j = j0 + wc + 16 * cb + li;
        if(i >= M || j >= N)
          continue;
        float v = alpha * acc[rb][cb][r];
        if(accumulate)
          v += C[i * ldc + j];
        if(bias)
          v += bias[j];
        if(relu)
          v = v > 0.f ? v : 0.f;
        C[i * ldc + j] = v;
      }
}

__global__ void __launch_bounds__(256)
  splitk_reduce_kernel(int M, int N, int splits, float const * __restrict__ part, float * C,
                       int64_t ldc, float alpha, int accumulate, float const * __restrict__ bias,
                       int relu)
{
  int64_t const total = (int64_t)M * N;
  for(int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
      e += (int64_t)gridDim.x * blockDim.x)
  {
    int const i = (int)(e / N), j = (int)(e % N);
    float     v = 0.f;
    for(int z = 0; z < splits; ++z)
      v += part[((int64_t)z * M + i) * N + j];
    v *= alpha;
    if(accumulate)
      v += C[i * ldc + j];
    if(bias)
      v += bias[j];
    if(relu)
      v = v > 0.f ? v : 0.f;
    C[i * ldc + j] = v;
  }
}

// features of cells [cell0, cell0+B): backward differences of orders 0..H-1 of the history
// (h_1, h_1 - h_2, h_1 - 2 h_2 + h_3, ...), each scaled, then the scaled rhs; targets
// (optional) = (x_t - extrapolation(h)) * inv_t. The differences are an invertible linear map
// of the last H solutions that removes their near-collinearity.
template<typename In>
__global__ void __launch_bounds__(256)
  features_kernel(float * __restrict__ X, float * __restrict__ T, History<In> hist, In const * rhs,
                  In const * target, int history, int n_cell_dofs, int64_t cell0, int B,
                  FeatureScales sc, double beta0, double beta1, double beta2, double beta3)
{
  int64_t const total = (int64_t)B * n_cell_dofs;
  int const     F     = (history + 1) * n_cell_dofs;
  for(int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
      e += (int64_t)gridDim.x * blockDim.x)
  {
    int64_t const b = e / n_cell_dofs;
    int const     j = (int)(e - b * n_cell_dofs);
    int64_t const g = (cell0 + b) * n_cell_dofs + j;
    double        h[4] = {0, 0, 0, 0};
    for(int k = 0; k < history; ++k)
      h[k] = (double)hist.p[k][g];
    double const d[4] = {h[0], h[0] - h[1], h[0] - 2.0 * h[1] + h[2],
                         h[0] - 3.0 * h[1] + 3.0 * h[2] - h[3]};
    for(int k = 0; k < history; ++k)
      X[b * F + k * n_cell_dofs + j] = (float)(d[k] * (double)sc.inv_d[k]);
    X[b * F + history * n_cell_dofs + j] = (float)((double)rhs[g] * (double)sc.inv_b);
    if(T)
    {
      double const ex = beta0 * h[0] + beta1 * h[1] + beta2 * h[2] + beta3 * h[3];
      T[b * n_cell_dofs + j] = (float)(((double)target[g] - ex) * (double)sc.inv_t);
    }
  }
}

// per-block sums of squares of the feature groups (H differences, rhs) and of the targets
__global__ void __launch_bounds__(256)
  feature_moments_kernel(float const * X, float const * T, int history, int n_cell_dofs, int B,
                         double * partial)
{
  __shared__ double red[6][256];
  double            acc[6] = {0, 0, 0, 0, 0, 0};
  int const         F      = (history + 1) * n_cell_dofs;
  int64_t const     total  = (int64_t)B * F;
  for(int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
      e += (int64_t)gridDim.x * blockDim.x)
  {
    int const    grp = (int)((e % F) / n_cell_dofs);
    double const v   = X[e];
    for(int k = 0; k <= 4; ++k)
      acc[k] += (grp == k) ? v * v : 0.0;
  }
  int64_t const nt = (int64_t)B * n_cell_dofs;
  for(int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nt;
      e += (int64_t)gridDim.x * blockDim.x)
    acc[5] += (double)T[e] * T[e];
  for(int k = 0; k < 6; ++k)
    red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if((int)threadIdx.x < w)
      for(int k = 0; k < 6; ++k)
        red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x < 6)
    partial[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// acc[k] += sum_b partial[b*6 + k]; group index of the rhs is `history`, mapped to slot 4
__global__ void __launch_bounds__(64)
  moments_reduce_kernel(double const * partial, int n_blocks, int history, double * acc)
{
  int const k = threadIdx.x;
  if(k >= 6)
    return;
  double s = 0;
  for(int b = 0; b < n_blocks; ++b)
    s += partial[b * 6 + k];
  // slots 0..H-1 differences, slot H rhs (stored at 4), slot 5 target
  int const dst = (k < history) ? k : (k == history ? 4 : (k == 5 ? 5 : -1));
  if(dst >= 0)
    acc[dst] += s;
}

__global__ void __launch_bounds__(256)
  mse_backward_kernel(float const * Y, float const * T, float * dY, int64_t n, float scale,
                      double * partial)
{
  __shared__ double red[256];
  double s = 0;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
  {
    float const d = Y[i] - T[i];
    dY[i]         = 2.f * d * scale;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0)
    partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256)
  sum_partials_kernel(double const * p, int n, double * out, int accumulate)
{
  __shared__ double red[256];
  double s = 0;
  for(int i = threadIdx.x; i < n; i += 256)
    s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0)
    *out = accumulate ? *out + red[0] : red[0];
}

__global__ void __launch_bounds__(256) relu_backward_kernel(float * dH, float const * H, int64_t n)
{
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
    if(H[i] <= 0.f)
      dH[i] = 0.f;
}

// db_j = sum_i dY(i,j): one block per column
__global__ void __launch_bounds__(256) column_sum_kernel(float const * dY, int M, int N, float * db)
{
  __shared__ float red[256];
  int const        j = blockIdx.x;
  float            s = 0;
  for(int i = threadIdx.x; i < M; i += 256)
    s += dY[(int64_t)i * N + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for(int w = 128; w > 0; w >>= 1)
  {
    if((int)threadIdx.x < w)
      red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if(threadIdx.x == 0)
    db[j] = red[0];
}

__global__ void __launch_bounds__(256)
  adam_kernel(float * p, float const * g, float * m, float * v, int64_t n, float lr, float b1,
              float b2, float eps, float c1, float c2)
{
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
  {
    float const gi = g[i];
    float const mi = b1 * m[i] + (1.f - b1) * gi;
    float const vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i]           = mi;
    v[i]           = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// x0 = sum_h beta_h hist_h (+ st * Y if Y != null) (- target if != null); counts non-finite
// entries of Y
__global__ void __launch_bounds__(256)
  combine_kernel(double * x0, History<double> hist, int history, double beta0, double beta1,
                 double beta2, double beta3, float const * Y, double st, double const * target,
                 int64_t n, unsigned int * n_bad)
{
  double const beta[4] = {beta0, beta1, beta2, beta3};
  unsigned int bad     = 0;
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
  {
    double v = 0;
    for(int h = 0; h < history; ++h)
      v += beta[h] * hist.p[h][i];
    if(Y)
    {
      float const y = Y[i];
      if(isfinite(y))
        v += st * (double)y;
      else
        ++bad;
    }
    if(target)
      v -= target[i];
    x0[i] = v;
  }
  if(n_bad && bad)
    atomicAdd(n_bad, bad);
}

// ---------------------------------------------------------------- divergence RHS

// one workgroup per cell: weak DG divergence, central flux, walls u+ = -u- (no slip):
//   b_i = -int_K u . grad(phi_i) + sum_{non-wall faces} int_F u.n_K phi_i
// The field is evaluated separably: per cell and direction d a table
// E_d[m][pos] = exp(i 2 pi m x_d(pos)) for m in [-K, K] at the n Gauss points and the two
// face positions, so u(x) = Re sum_modes d c E_x[n_x] E_y[n_y] E_z[n_z] costs two complex
// products per mode instead of a cosine. Mode data: n_x, n_y, n_z, d[3], c_re, c_im.
__global__ void __launch_bounds__(128)
  divergence_rhs_kernel(double * b, double const * __restrict__ md, int n_modes, int K,
                        BoxTopology topo, double lx, double ly, double lz, double hx, double hy,
                        double hz, int n, double const * S, double const * Sd, double const * w)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int const     n3 = n * n * n, n2 = n * n, np = n + 2, nm = 2 * K + 1;
  double *      f   = reinterpret_cast<double *>(smem_raw);  // 3 * n^3
  double *      o   = f + 3 * n3;                             // n^3 output
  double *      tmp = o + n3;                                 // n^3 scratch
  double *      E   = tmp + n3;                               // 3 * nm * np complex
  int64_t const c   = blockIdx.x;
  int           cc[3];
  cell_coords(topo, c, cc);
  double const h[3]  = {hx, hy, hz};
  double const lo[3] = {lx, ly, lz};
  double const vol   = hx * hy * hz;

  for(int e = threadIdx.x; e < 3 * nm * np; e += blockDim.x)
  {
    int const    d = e / (nm * np), r = e % (nm * np);
    int const    m = r / np - K, pos = r % np;
    double const xi = pos < n ? w[n + pos] : (double)(pos - n);
    double const x  = lo[d] + h[d] * (cc[d] + xi);
    double       sn, cs;
    sincos(2.0 * M_PI * (double)m * x, &sn, &cs);
    E[2 * e]     = cs;
    E[2 * e + 1] = sn;
  }
  __syncthreads();

  // u at the point with table positions (p0, p1, p2)
  auto velocity = [&](int p0, int p1, int p2, double * u) {
    u[0] = u[1] = u[2] = 0;
    for(int m = 0; m < n_modes; ++m)
    {
      double const * q  = md + 8 * m;
      double const * ex = E + 2 * ((0 * nm + (int)q[0] + K) * np + p0);
      double const * ey = E + 2 * ((1 * nm + (int)q[1] + K) * np + p1);
      double const * ez = E + 2 * ((2 * nm + (int)q[2] + K) * np + p2);
      double const   ar = ex[0] * ey[0] - ex[1] * ey[1];
      double const   ai = ex[0] * ey[1] + ex[1] * ey[0];
      double const   zr = ar * ez[0] - ai * ez[1];
      double const   zi = ar * ez[1] + ai * ez[0];
      double const   v  = q[6] * zr - q[7] * zi;
      u[0] += v * q[3];
      u[1] += v * q[4];
      u[2] += v * q[5];
    }
  };

  // volume: f_d(q) = w_q |J| u_d(x_q) / h_d  (reference-gradient test weights)
  for(int q = threadIdx.x; q < n3; q += blockDim.x)
  {
    int const qa = q % n, qb = (q / n) % n, qc = q / n2;
    double    u[3];
    velocity(qa, qb, qc, u);
    double const wq = w[qa] * w[qb] * w[qc] * vol;
    for(int d = 0; d < 3; ++d)
      f[d * n3 + q] = wq * u[d] / h[d];
  }
  for(int e = threadIdx.x; e < n3; e += blockDim.x)
    o[e] = 0;
  __syncthreads();
  // o_i = -sum_d (grad_d test)^T f_d, three transposed sweeps per component
  for(int d = 0; d < 3; ++d)
  {
    double const * Az = d == 2 ? Sd : S;
    double const * Ay = d == 1 ? Sd : S;
    double const * Ax = d == 0 ? Sd : S;
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const a = e % n, bb = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int cq = 0; cq < n; ++cq)
        sum += Az[cq * n + k] * f[d * n3 + a + n * (bb + n * cq)];
      tmp[e] = sum;
    }
    __syncthreads();
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const a = e % n, j = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int bq = 0; bq < n; ++bq)
        sum += Ay[bq * n + j] * tmp[a + n * (bq + n * k)];
      f[d * n3 + e] = sum;  // this component's quadrature data is no longer needed
    }
    __syncthreads();
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const i = e % n, j = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int aq = 0; aq < n; ++aq)
        sum += Ax[aq * n + i] * f[d * n3 + aq + n * (j + n * k)];
      o[e] -= sum;
    }
    __syncthreads();
  }
  // faces: interior / periodic faces carry {u}.n = u.n (continuous field); walls carry 0
  for(int d = 0; d < 3; ++d)
    for(int s = 0; s < 2; ++s)
    {
      if(neighbor_id(topo, cc, d, s) < 0)
        continue;
      int const    t1 = tangent1(d), t2 = tangent2(d);
      double const area = h[t1] * h[t2];
      double const sg   = s ? 1.0 : -1.0;
      for(int q = threadIdx.x; q < n2; q += blockDim.x)
      {
        int const qa = q % n, qb = q / n;
        int       pos[3];
        pos[d]  = n + s;
        pos[t1] = qa;
        pos[t2] = qb;
        double u[3];
        velocity(pos[0], pos[1], pos[2], u);
        tmp[q] = sg * u[d] * w[qa] * w[qb] * area;
      }
      __syncthreads();
      // face layer dofs: o(i,j at layer) += sum_{qa,qb} S(qa,i) S(qb,j) tmp(qa,qb)
      for(int e = threadIdx.x; e < n2; e += blockDim.x)
      {
        int const i = e % n, j = e / n;
        double    sum = 0;
        for(int qb = 0; qb < n; ++qb)
          for(int qa = 0; qa < n; ++qa)
            sum += S[qa * n + i] * S[qb * n + j] * tmp[qa + n * qb];
        int idx[3];
        idx[d]  = s ? n - 1 : 0;
        idx[t1] = i;
        idx[t2] = j;
        o[idx[0] + n * (idx[1] + n * idx[2])] += sum;
      }
      __syncthreads();
    }
  for(int e = threadIdx.x; e < n3; e += blockDim.x)
    b[c * n3 + e] = o[e];
}

// Curved meshes: the same weak divergence from precomputed point data (position and
// C = JxW J^{-1} per Gauss point, position and JxW n per face point). The field is evaluated
// at arbitrary points: per thread a table of exp(i 2 pi m x_d), m = 0..K, in LDS (negative m
// by conjugation), two complex products per mode.
__global__ void __launch_bounds__(128)
  divergence_rhs_curved_kernel(double * b, double const * __restrict__ md, int n_modes, int K,
                               BoxTopology topo, int n, double const * S, double const * Sd,
                               double const * w, double const * __restrict__ cgeo,
                               double const * __restrict__ fgeo)
{
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int const     n3 = n * n * n, n2 = n * n, nk = K + 1;
  double *      f   = reinterpret_cast<double *>(smem_raw);  // 3 * n^3
  double *      o   = f + 3 * n3;
  double *      tmp = o + n3;
  double *      E   = tmp + n3 + threadIdx.x * 6 * nk;      // this thread's table
  int64_t const c   = blockIdx.x;
  int           cc[3];
  cell_coords(topo, c, cc);

  auto velocity = [&](double const * x, double * u) {
#pragma unroll
    for(int d = 0; d < 3; ++d)
    {
      double sn, cs;
      sincos(2.0 * M_PI * x[d], &sn, &cs);
      double * e = E + d * 2 * nk;
      e[0]       = 1.0;
      e[1]       = 0.0;
      for(int m = 1; m <= K; ++m)
      {
        e[2 * m]     = e[2 * m - 2] * cs - e[2 * m - 1] * sn;
        e[2 * m + 1] = e[2 * m - 2] * sn + e[2 * m - 1] * cs;
      }
    }
    u[0] = u[1] = u[2] = 0;
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
      double const   v  = q[6] * zr - q[7] * zi;
      u[0] += v * q[3];
      u[1] += v * q[4];
      u[2] += v * q[5];
    }
  };

  // volume: f_a(q) = sum_k C[a][k] u_k(x_q), then o = -sum_a (grad_a test)^T f_a
  double const * cg = cgeo + c * 12 * n3;
  for(int q = threadIdx.x; q < n3; q += blockDim.x)
  {
    double const x[3] = {cg[q], cg[n3 + q], cg[2 * n3 + q]};
    double       u[3];
    velocity(x, u);
    for(int a = 0; a < 3; ++a)
      f[a * n3 + q] = cg[(3 + 3 * a) * n3 + q] * u[0] + cg[(4 + 3 * a) * n3 + q] * u[1] +
                      cg[(5 + 3 * a) * n3 + q] * u[2];
  }
  for(int e = threadIdx.x; e < n3; e += blockDim.x)
    o[e] = 0;
  __syncthreads();
  for(int d = 0; d < 3; ++d)
  {
    double const * Az = d == 2 ? Sd : S;
    double const * Ay = d == 1 ? Sd : S;
    double const * Ax = d == 0 ? Sd : S;
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const a = e % n, bb = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int cq = 0; cq < n; ++cq)
        sum += Az[cq * n + k] * f[d * n3 + a + n * (bb + n * cq)];
      tmp[e] = sum;
    }
    __syncthreads();
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const a = e % n, j = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int bq = 0; bq < n; ++bq)
        sum += Ay[bq * n + j] * tmp[a + n * (bq + n * k)];
      f[d * n3 + e] = sum;
    }
    __syncthreads();
    for(int e = threadIdx.x; e < n3; e += blockDim.x)
    {
      int const i = e % n, j = (e / n) % n, k = e / n2;
      double    sum = 0;
      for(int aq = 0; aq < n; ++aq)
        sum += Ax[aq * n + i] * f[d * n3 + aq + n * (j + n * k)];
      o[e] -= sum;
    }
    __syncthreads();
  }
  // faces: outward flux u.n_K with n_K = +n on the face above the cell (s = 1), -n below
  for(int d = 0; d < 3; ++d)
    for(int s = 0; s < 2; ++s)
    {
      if(neighbor_id(topo, cc, d, s) < 0)
        continue;
      double const * fg = fgeo + face_id(topo, cc, d, s) * 6 * n2;
      double const   sg = s ? 1.0 : -1.0;
      int const      t1 = tangent1(d), t2 = tangent2(d);
      for(int q = threadIdx.x; q < n2; q += blockDim.x)
      {
        double const x[3] = {fg[q], fg[n2 + q], fg[2 * n2 + q]};
        double       u[3];
        velocity(x, u);
        tmp[q] = sg * (u[0] * fg[3 * n2 + q] + u[1] * fg[4 * n2 + q] + u[2] * fg[5 * n2 + q]);
      }
      __syncthreads();
      for(int e = threadIdx.x; e < n2; e += blockDim.x)
      {
        int const i = e % n, j = e / n;
        double    sum = 0;
        for(int qb = 0; qb < n; ++qb)
          for(int qa = 0; qa < n; ++qa)
            sum += S[qa * n + i] * S[qb * n + j] * tmp[qa + n * qb];
        int idx[3];
        idx[d]  = s ? n - 1 : 0;
        idx[t1] = i;
        idx[t2] = j;
        o[idx[0] + n * (idx[1] + n * idx[2])] += sum;
      }
      __syncthreads();
    }
  for(int e = threadIdx.x; e < n3; e += blockDim.x)
    b[c * n3 + e] = o[e];
}

__global__ void __launch_bounds__(256)
  scale_to_float_kernel(float * dst, double const * src, double coef, int64_t n)
{
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
      i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)(coef * src[i]);
}

__global__ void fill_u32_kernel(unsigned int * p, unsigned int v, int n)
{
  int const i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n)
    p[i] = v;
}

inline int grid_for(int64_t n)
{
  int64_t const b = (n + 255) / 256;
  return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 4096);
}
}  // namespace

hipError_t launch_sgemm(int M, int N, int K, float const * A, int64_t a_i, int64_t a_k,
                        float const * B, int64_t b_k, int64_t b_j, float * C, int64_t ldc,
                        float alpha, bool accumulate, float const * bias, bool relu,
                        float * workspace, int64_t workspace_floats, hipStream_t s)
{
  int const tiles = ((N + TN - 1) / TN) * ((M + TM - 1) / TM);
  // split K when the output tiles alone cannot fill the chip (weight gradients: K = batch)
  int splits = 1;
  if(workspace != nullptr && tiles < 512 && K >= 8 * TK)
  {
    splits = std::min((512 + tiles - 1) / tiles, K / (4 * TK));
    splits = (int)std::min<int64_t>(splits, workspace_floats / ((int64_t)M * N));
    splits = std::max(splits, 1);
  }
  int const kc = splits > 1 ? ((K + splits - 1) / splits + TK - 1) / TK * TK : K;
  splits       = (K + kc - 1) / kc;
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, splits);
  if(splits == 1)
  {
    hipLaunchKernelGGL(sgemm_kernel, grid, dim3(256), 0, s, M, N, K, A, a_i, a_k, B, b_k, b_j, C, ldc,
                       alpha, accumulate ? 1 : 0, bias, relu ? 1 : 0, kc);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(sgemm_kernel, grid, dim3(256), 0, s, M, N, K, A, a_i, a_k, B, b_k, b_j,
                     workspace, (int64_t)N, 1.f, 0, nullptr, 0, kc);
  int64_t const total = (int64_t)M * N;
  int const     g     = (int)std::min<int64_t>((total + 255) / 256, 2048);
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(g), dim3(256), 0, s, M, N, splits, workspace, C, ldc,
                     alpha, accumulate ? 1 : 0, bias, relu ? 1 : 0);
  return hipGetLastError();
}

template<typename In>
hipError_t launch_features(float * X, float * T, History<In> const & hist, In const * rhs,
                           In const * target, int history, int n_cell_dofs, int64_t cell0, int B,
                           FeatureScales const & sc, double const * beta, hipStream_t s)
{
  if(history < 1 || history > 4)
    return hipErrorInvalidValue;
  int64_t const total = (int64_t)B * n_cell_dofs;
  hipLaunchKernelGGL((features_kernel<In>), dim3(grid_for(total)), dim3(256), 0, s, X, T, hist,
                     rhs, target, history, n_cell_dofs, cell0, B, sc, beta[0], beta[1], beta[2],
                     beta[3]);
  return hipGetLastError();
}
template hipError_t launch_features<float>(float *, float *, History<float> const &, float const *,
                                           float const *, int, int, int64_t, int,
                                           FeatureScales const &, double const *, hipStream_t);
template hipError_t launch_features<double>(float *, float *, History<double> const &,
                                            double const *, double const *, int, int, int64_t, int,
                                            FeatureScales const &, double const *, hipStream_t);

hipError_t launch_feature_moments(float const * X, float const * T, int history, int n_cell_dofs,
                                  int B, double * partial, double * acc, hipStream_t s)
{
  if(history < 1 || history > 4)
    return hipErrorInvalidValue;
  int const g = std::min(grid_for((int64_t)B * (history + 1) * n_cell_dofs), 1024);
  hipLaunchKernelGGL(feature_moments_kernel, dim3(g), dim3(256), 0, s, X, T, history, n_cell_dofs,
                     B, partial);
  hipLaunchKernelGGL(moments_reduce_kernel, dim3(1), dim3(64), 0, s, partial, g, history, acc);
  return hipGetLastError();
}

hipError_t launch_mse_backward(float const * Y, float const * T, float * dY, int64_t n, float scale,
                               double * partial, double * loss_sum, bool accumulate, hipStream_t s)
{
  int const g = std::min(grid_for(n), 1024);
  hipLaunchKernelGGL(mse_backward_kernel, dim3(g), dim3(256), 0, s, Y, T, dY, n, scale, partial);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partial, g, loss_sum,
                     accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_fill_u32(unsigned int * p, unsigned int v, int n, hipStream_t s)
{
  hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_relu_backward(float * dH, float const * H, int64_t n, hipStream_t s)
{
  hipLaunchKernelGGL(relu_backward_kernel, dim3(grid_for(n)), dim3(256), 0, s, dH, H, n);
  return hipGetLastError();
}

hipError_t launch_column_sum(float const * dY, int M, int N, float * db, hipStream_t s)
{
  hipLaunchKernelGGL(column_sum_kernel, dim3(N), dim3(256), 0, s, dY, M, N, db);
  return hipGetLastError();
}

hipError_t launch_adam(float * p, float const * g, float * m, float * v, int64_t n, float lr,
                       float b1, float b2, float eps, int step, hipStream_t s)
{
  float const c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

hipError_t launch_combine(double * x0, History<double> const & hist, int history,
                          double const * beta, float const * Y, double st, double const * target,
                          int64_t n, unsigned int * n_bad, hipStream_t s)
{
  hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, s, x0, hist, history,
                     beta[0], beta[1], beta[2], beta[3], Y, st, target, n, n_bad);
  return hipGetLastError();
}

hipError_t launch_scale_to_float(float * dst, double const * src, double coef, int64_t n,
                                 hipStream_t s)
{
  hipLaunchKernelGGL(scale_to_float_kernel, dim3(grid_for(n)), dim3(256), 0, s, dst, src, coef, n);
  return hipGetLastError();
}

hipError_t launch_divergence_rhs(double * b, double const * mode_data, int n_modes, int K,
                                 BoxTopology const & topo, double const * lower, double const * h,
                                 int n, double const * S, double const * Sd, double const * w,
                                 int64_t n_cells, hipStream_t s)
{
  if(n < 1 || n > 9 || K < 0 || K > 16 || n_cells < 1)
    return hipErrorInvalidValue;
  size_t const lds = sizeof(double) * (5 * n * n * n + 6 * (2 * K + 1) * (n + 2));
  hipLaunchKernelGGL(divergence_rhs_kernel, dim3(n_cells), dim3(128), lds, s, b, mode_data, n_modes,
                     K, topo, lower[0], lower[1], lower[2], h[0], h[1], h[2], n, S, Sd, w);
  return hipGetLastError();
}

hipError_t launch_divergence_rhs_curved(double * b, double const * mode_data, int n_modes, int K,
                                        BoxTopology const & topo, int n, double const * S,
                                        double const * Sd, double const * w,
                                        double const * cell_geo, double const * face_geo,
                                        int64_t n_cells, hipStream_t s)
{
  if(n < 1 || n > 9 || K < 0 || K > 16 || n_cells < 1)
    return hipErrorInvalidValue;
  size_t const lds = sizeof(double) * (5 * n * n * n + 128 * 6 * (K + 1));
  // raised under a lock to the largest size asked for so far, so it never decreases
  static std::mutex attr_lock;
  static size_t     attr = 0;
  if(lds > 65536)
  {
    std::lock_guard<std::mutex> guard(attr_lock);
    if(lds > attr)
    {
      hipError_t e = hipFuncSetAttribute((void const *)divergence_rhs_curved_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if(e != hipSuccess)
        return e;
      attr = lds;
    }
  }
  hipLaunchKernelGGL(divergence_rhs_curved_kernel, dim3(n_cells), dim3(128), lds, s, b, mode_data,
                     n_modes, K, topo, n, S, Sd, w, cell_geo, face_geo);
  return hipGetLastError();
}

}  // namespace ExaDG
