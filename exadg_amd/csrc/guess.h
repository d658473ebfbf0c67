// Learned initial guess for a sequence of pressure-Poisson solves (no reference counterpart;
// the reference's guess is the extrapolation p0 = sum_i beta_i p^{n-i} of
// time_int_bdf_dual_splitting.cpp:489-501 with extrapolation_constants.cpp:55-93, which is
// both the fallback and the baseline here).
//
// A per-cell MLP (FP32) maps the last H solutions and the rhs on the cell's DoFs to a correction
// of the order-H extrapolation:
//   p0 = sum_h beta_h p^{n-h} + s_t * (W3 relu(W2 relu(W1 x + b1) + b2) + b3),
//   x  = [D^0 p/s_0, ..., D^{H-1} p/s_{H-1}, b^n/s_b],
// with D^j p the j-th backward difference of (p^{n-1}, p^{n-2}, ...) and every scale the RMS of
// its quantity over the training set (s_t: RMS of the extrapolation error, so the targets have
// unit RMS and a trained loss below 1 is a gain over the extrapolation).
// Trained on the GPU (tiled SGEMM forward/backward, Adam) on solutions the GPU solver
// produced; a non-finite network output falls back to the extrapolation.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "common.h"
#include "guess_kernels.h"
#include "vector_ops.h"

namespace ExaDG
{
// beta of ExtrapolationConstants::set_constant_time_step for orders 1..4
void extrapolation_beta(int order, double * beta);

namespace Poisson
{
class LaplaceOperatorBase;
}

// training losses; each normalized per time step by the same quantity of the extrapolation
// error e_ex = p_ex - p, so that 1 means "as good as the extrapolation"
enum class GuessLoss
{
  l2       = 0,  // sum over DoFs of e^2 (per-cell minibatches)
  energy   = 1,  // e^T A e (whole field per step)
  residual = 2   // |A e|^2 (whole field per step)
};

class GuessNet
{
public:
  GuessNet(int degree, int history, int width, uint64_t seed);

  int     n_params() const { return n_params_; }
  int     n_cell_dofs() const { return n3_; }
  int     n_features() const { return F_; }
  int     width() const { return width_; }
  int     history() const { return history_; }
  // [s_0 .. s_{H-1}, s_b, s_t]
  std::vector<double> scales() const;
  void                set_scales(std::vector<double> const & sc);
  float *             params() { return params_.data(); }
  float const *       params() const { return params_.data(); }
  float *             grads() { return grads_.data(); }

  // loss = mean (Y - T)^2 over B x n3 on given features (device FP32, row-major); fills grads
  double loss_and_grad(float const * X, float const * T, int B, hipStream_t s);
  // Y = MLP(X) for B samples into an internal buffer; returns its device pointer
  float const * forward(float const * X, int B, hipStream_t s);

  // Adam on the stored FP64 sequence P[t][N] (solutions), Bv[t][N] (right-hand sides),
  // step-major, t in [t_begin + H, t_end). l2: shuffled (t, cell-block) minibatches of
  // batch_cells cells; energy / residual: one whole step per Adam step (op: the FP64 operator).
  // The first call (no Adam step taken yet) fits the scales on the same steps. Returns the
  // mean normalized loss of the last epoch.
  double train(Poisson::LaplaceOperatorBase const * op, GuessLoss loss, double const * P,
               double const * Bv, int64_t n_dofs, int t_begin, int t_end, int epochs,
               int batch_cells, float lr, uint64_t seed, hipStream_t s);
  // mean normalized loss over steps [t_begin + H, t_end) without updating
  double evaluate(Poisson::LaplaceOperatorBase const * op, GuessLoss loss, double const * P,
                  double const * Bv, int64_t n_dofs, int t_begin, int t_end, int batch_cells,
                  hipStream_t s);

  // p0 for all cells from the history (newest first, FP64) and rhs; returns the number of
  // non-finite network outputs (> 0 means the extrapolation fallback was used)
  int predict(double * x0, History<double> const & hist, double const * rhs, int64_t n_cells,
              int batch_cells, hipStream_t s);

private:
  void          ensure_workspace(int B);
  FeatureScales feature_scales() const;
  void          backward(float const * X, int B, hipStream_t s);
  void          fit_scales(double const * P, double const * Bv, int64_t n_dofs, int t_begin,
                           int t_end, int batch_cells, hipStream_t s);
  History<double> history_at(double const * P, int64_t n_dofs, int t) const;
  // whole-field loss of step t (unnormalized); with grad, also dY = d loss / d Y * coef
  double field_loss(Poisson::LaplaceOperatorBase const & op, GuessLoss loss, double const * P,
                    double const * Bv, int64_t n_dofs, int t, bool use_net, float grad_coef,
                    hipStream_t s);
  float * W1() { return params_.data(); }
  float * b1() { return W1() + (size_t)width_ * F_; }
  float * W2() { return b1() + width_; }
  float * b2() { return W2() + (size_t)width_ * width_; }
  float * W3() { return b2() + width_; }
  float * b3() { return W3() + (size_t)n3_ * width_; }
  float * gW(float * p) { return grads_.data() + (p - params_.data()); }

  int                        n3_, history_, width_, F_, n_params_;
  double                     beta_[4];
  double                     sd_[4] = {1, 1, 1, 1}, sb_ = 1, st_ = 1;
  int                        ws_B_      = 0;
  int                        adam_step_ = 0;
  DeviceBuffer<float>        params_, grads_, m_, v_;
  DeviceBuffer<float>        X_, T_, H1_, H2_, Y_, dY_, dH1_, dH2_, split_;
  DeviceBuffer<double>       partial_, loss_, moments_;
  DeviceBuffer<double>       e_, r_, q_;
  DeviceBuffer<unsigned int> bad_;
  std::unique_ptr<Reducer>   red_;
};

}  // namespace ExaDG
