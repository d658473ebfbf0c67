#include "guess.h"

#include "laplace_operator.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>

namespace ExaDG
{
void extrapolation_beta(int order, double * beta)
{
  // ExtrapolationConstants::set_constant_time_step (extrapolation_constants.cpp:62-83)
  for(int i = 0; i < 4; ++i)
    beta[i] = 0.0;
  switch(order)
  {
    case 1: beta[0] = 1.0; break;
    case 2: beta[0] = 2.0, beta[1] = -1.0; break;
    case 3: beta[0] = 3.0, beta[1] = -3.0, beta[2] = 1.0; break;
    case 4: beta[0] = 4.0, beta[1] = -6.0, beta[2] = 4.0, beta[3] = -1.0; break;
    default: EXADG_THROW(EXADG_ERR_ARGUMENT, "extrapolation order must be 1..4");
  }
}

namespace
{
uint64_t splitmix(uint64_t & s)
{
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t & s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void check_launch(hipError_t e)
{
  if(e != hipSuccess)
    EXADG_THROW(EXADG_ERR_HIP, hipGetErrorString(e));
}
}  // namespace

GuessNet::GuessNet(int degree, int history, int width, uint64_t seed)
  : history_(history), width_(width)
{
  if(degree < 1 || degree > 8)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "degree must be in [1,8]");
  if(history < 1 || history > 4)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "history must be in [1,4]");
  if(width < 1 || width > 4096)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "width out of range");
  n3_       = (degree + 1) * (degree + 1) * (degree + 1);
  F_        = (history + 1) * n3_;
  n_params_ = width * F_ + width + width * width + width + n3_ * width + n3_;
  extrapolation_beta(history, beta_);
  // uniform(-sqrt(6/fan_in), +) for the hidden layers; the output layer starts small so that
  // the initial guess is close to the extrapolation
  std::vector<float> p((size_t)n_params_, 0.f);
  uint64_t           s = seed;
  auto fill = [&](size_t off, size_t count, double bound) {
    for(size_t i = 0; i < count; ++i)
      p[off + i] = (float)((2.0 * uniform01(s) - 1.0) * bound);
  };
  size_t off = 0;
  fill(off, (size_t)width * F_, std::sqrt(6.0 / F_));
  off += (size_t)width * F_ + width;
  fill(off, (size_t)width * width, std::sqrt(6.0 / width));
  off += (size_t)width * width + width;
  fill(off, (size_t)n3_ * width, 0.1 * std::sqrt(6.0 / width));
  params_.upload(p);
  std::vector<float> z((size_t)n_params_, 0.f);
  grads_.upload(z);
  m_.upload(z);
  v_.upload(z);
  partial_.resize(1024 * 6);
  // split-K partials of the weight gradients: up to 64 slices of the largest weight matrix
  split_.resize((size_t)64 * std::max<int64_t>((int64_t)width * F_, (int64_t)n3_ * width));
  loss_.resize(1);
  bad_.resize(1);
}

void GuessNet::ensure_workspace(int B)
{
  if(B <= ws_B_)
    return;
  ws_B_ = B;
  X_.resize((size_t)B * F_);
  T_.resize((size_t)B * n3_);
  Y_.resize((size_t)B * n3_);
  dY_.resize((size_t)B * n3_);
  H1_.resize((size_t)B * width_);
  H2_.resize((size_t)B * width_);
  dH1_.resize((size_t)B * width_);
  dH2_.resize((size_t)B * width_);
}

float const * GuessNet::forward(float const * X, int B, hipStream_t s)
{
  ensure_workspace(B);
  int const w = width_;
  check_launch(launch_sgemm(B, w, F_, X, F_, 1, W1(), 1, F_, H1_.data(), w, 1.f, false, b1(), true, nullptr, 0, s));
  check_launch(launch_sgemm(B, w, w, H1_.data(), w, 1, W2(), 1, w, H2_.data(), w, 1.f, false, b2(), true, nullptr, 0, s));
  check_launch(launch_sgemm(B, n3_, w, H2_.data(), w, 1, W3(), 1, w, Y_.data(), n3_, 1.f, false, b3(), false, nullptr, 0, s));
  return Y_.data();
}

void GuessNet::backward(float const * X, int B, hipStream_t s)
{
  int const w = width_;
  // output layer
  check_launch(launch_sgemm(n3_, w, B, dY_.data(), 1, n3_, H2_.data(), w, 1, gW(W3()), w, 1.f, false, nullptr, false, split_.data(),
                            (int64_t)split_.size(), s));
  check_launch(launch_column_sum(dY_.data(), B, n3_, gW(b3()), s));
  check_launch(launch_sgemm(B, w, n3_, dY_.data(), n3_, 1, W3(), w, 1, dH2_.data(), w, 1.f, false, nullptr, false, split_.data(),
                            (int64_t)split_.size(), s));
  check_launch(launch_relu_backward(dH2_.data(), H2_.data(), (int64_t)B * w, s));
  // hidden layer 2
  check_launch(launch_sgemm(w, w, B, dH2_.data(), 1, w, H1_.data(), w, 1, gW(W2()), w, 1.f, false, nullptr, false, split_.data(),
                            (int64_t)split_.size(), s));
  check_launch(launch_column_sum(dH2_.data(), B, w, gW(b2()), s));
  check_launch(launch_sgemm(B, w, w, dH2_.data(), w, 1, W2(), w, 1, dH1_.data(), w, 1.f, false, nullptr, false, split_.data(),
                            (int64_t)split_.size(), s));
  check_launch(launch_relu_backward(dH1_.data(), H1_.data(), (int64_t)B * w, s));
  // hidden layer 1
  check_launch(launch_sgemm(w, F_, B, dH1_.data(), 1, w, X, F_, 1, gW(W1()), F_, 1.f, false, nullptr, false, split_.data(),
                            (int64_t)split_.size(), s));
  check_launch(launch_column_sum(dH1_.data(), B, w, gW(b1()), s));
}

double GuessNet::loss_and_grad(float const * X, float const * T, int B, hipStream_t s)
{
  forward(X, B, s);
  int64_t const n = (int64_t)B * n3_;
  check_launch(launch_mse_backward(Y_.data(), T, dY_.data(), n, 1.f / (float)n, partial_.data(),
                                   loss_.data(), false, s));
  backward(X, B, s);
  double h = 0;
  HIP_CHECK(hipMemcpyAsync(&h, loss_.data(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return h / (double)n;
}

std::vector<double> GuessNet::scales() const
{
  std::vector<double> v(sd_, sd_ + history_);
  v.push_back(sb_);
  v.push_back(st_);
  return v;
}

void GuessNet::set_scales(std::vector<double> const & sc)
{
  if((int)sc.size() != history_ + 2)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "expected history + 2 scales");
  for(double v : sc)
    if(!(v > 0.0) || !std::isfinite(v))
      EXADG_THROW(EXADG_ERR_ARGUMENT, "scales must be positive and finite");
  for(int k = 0; k < history_; ++k)
    sd_[k] = sc[k];
  sb_ = sc[history_];
  st_ = sc[history_ + 1];
}

FeatureScales GuessNet::feature_scales() const
{
  FeatureScales f;
  for(int k = 0; k < history_; ++k)
    f.inv_d[k] = (float)(1.0 / sd_[k]);
  f.inv_b = (float)(1.0 / sb_);
  f.inv_t = (float)(1.0 / st_);
  return f;
}

namespace
{
struct Sample
{
  int     t;
  int64_t cell0;
  int     B;
};

std::vector<Sample> make_samples(int t0, int t1, int64_t n_cells, int batch)
{
  std::vector<Sample> v;
  for(int t = t0; t < t1; ++t)
    for(int64_t c = 0; c < n_cells; c += batch)
      v.push_back({t, c, (int)std::min<int64_t>(batch, n_cells - c)});
  return v;
}

void check_range(int history, int n3, int64_t n_dofs, int t_begin, int t_end)
{
  if(t_begin < 0 || t_end - t_begin <= history)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "step range shorter than the history");
  if(n_dofs <= 0 || n_dofs % n3 != 0)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "n_dofs is not a positive multiple of the cell size");
}
}  // namespace

History<double> GuessNet::history_at(double const * P, int64_t n_dofs, int t) const
{
  History<double> h;
  for(int k = 0; k < history_; ++k)
    h.p[k] = P + (size_t)(t - 1 - k) * n_dofs;
  return h;
}

void GuessNet::fit_scales(double const * P, double const * Bv, int64_t n_dofs, int t_begin,
                          int t_end, int batch_cells, hipStream_t s)
{
  int64_t const n_cells = n_dofs / n3_;
  batch_cells           = (int)std::max<int64_t>(1, std::min<int64_t>(batch_cells, n_cells));
  ensure_workspace(batch_cells);
  moments_.resize(6);
  std::vector<double> zero(6, 0.0);
  moments_.upload(zero);
  FeatureScales const unit;
  for(Sample const & sm : make_samples(t_begin + history_, t_end, n_cells, batch_cells))
  {
    check_launch(launch_features<double>(X_.data(), T_.data(), history_at(P, n_dofs, sm.t),
                                         Bv + (size_t)sm.t * n_dofs, P + (size_t)sm.t * n_dofs,
                                         history_, n3_, sm.cell0, sm.B, unit, beta_, s));
    check_launch(launch_feature_moments(X_.data(), T_.data(), history_, n3_, sm.B, partial_.data(),
                                        moments_.data(), s));
  }
  double mom[6];
  HIP_CHECK(hipMemcpyAsync(mom, moments_.data(), sizeof(mom), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  double const cnt = (double)n_dofs * (t_end - t_begin - history_);
  auto rms = [cnt](double v) { return (v > 0 && std::isfinite(v)) ? std::sqrt(v / cnt) : 1.0; };
  for(int k = 0; k < history_; ++k)
    sd_[k] = rms(mom[k]);
  sb_ = rms(mom[4]);
  st_ = rms(mom[5]);
}

double GuessNet::field_loss(Poisson::LaplaceOperatorBase const & op, GuessLoss loss,
                            double const * P, double const * Bv, int64_t n_dofs, int t,
                            bool use_net, float grad_coef, hipStream_t s)
{
  int64_t const n_cells = n_dofs / n3_;
  if(e_.size() != (size_t)n_dofs)
  {
    e_.resize(n_dofs);
    r_.resize(n_dofs);
    if(loss == GuessLoss::residual)
      q_.resize(n_dofs);
  }
  if(loss == GuessLoss::residual && q_.size() != (size_t)n_dofs)
    q_.resize(n_dofs);
  if(!red_)
    red_ = std::make_unique<Reducer>();
  History<double> const h = history_at(P, n_dofs, t);
  float const *         Y = nullptr;
  if(use_net)
  {
    ensure_workspace((int)n_cells);
    check_launch(launch_features<double>(X_.data(), nullptr, h, Bv + (size_t)t * n_dofs, nullptr,
                                         history_, n3_, 0, (int)n_cells, feature_scales(), beta_, s));
    Y = forward(X_.data(), (int)n_cells, s);
  }
  // e = guess - p^t
  check_launch(launch_combine(e_.data(), h, history_, beta_, Y, st_, P + (size_t)t * n_dofs, n_dofs,
                              nullptr, s));
  op.vmult(r_.data(), e_.data(), s);  // r = A e
  double const * g = r_.data();
  if(loss == GuessLoss::energy)
    vec::dot<double>(*red_, 0, e_.data(), r_.data(), n_dofs, s);
  else
  {
    vec::dot<double>(*red_, 0, r_.data(), r_.data(), n_dofs, s);
    op.vmult(q_.data(), r_.data(), s);  // A^T A e (A symmetric)
    g = q_.data();
  }
  if(grad_coef != 0.f)  // d loss / d Y = 2 st (A e or A A e), times coef
    check_launch(launch_scale_to_float(dY_.data(), g, 2.0 * st_ * grad_coef, n_dofs, s));
  return red_->fetch(s, 1)[0];
}

double GuessNet::train(Poisson::LaplaceOperatorBase const * op, GuessLoss loss, double const * P,
                       double const * Bv, int64_t n_dofs, int t_begin, int t_end, int epochs,
                       int batch_cells, float lr, uint64_t seed, hipStream_t s)
{
  check_range(history_, n3_, n_dofs, t_begin, t_end);
  if(loss != GuessLoss::l2)
  {
    if(op == nullptr || op->precision() != EXADG_FP64 || op->m() != n_dofs)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "the operator losses need the FP64 operator of the sequence");
    if(op->n_cells() > (int64_t)std::numeric_limits<int>::max())
      EXADG_THROW(EXADG_ERR_ARGUMENT, "too many cells for one batch");
  }
  int64_t const n_cells = n_dofs / n3_;
  batch_cells           = (int)std::max<int64_t>(1, std::min<int64_t>(batch_cells, n_cells));
  if(adam_step_ == 0)  // the first training call fixes the normalization
    fit_scales(P, Bv, n_dofs, t_begin, t_end, batch_cells, s);
  std::mt19937_64 rng(seed);
  double          last = 0;
  if(loss == GuessLoss::l2)
  {
    ensure_workspace(batch_cells);
    FeatureScales const sc      = feature_scales();
    std::vector<Sample> samples = make_samples(t_begin + history_, t_end, n_cells, batch_cells);
    for(int ep = 0; ep < epochs; ++ep)
    {
      std::shuffle(samples.begin(), samples.end(), rng);
      // the epoch's loss sum accumulates on the device: one host synchronization per epoch
      int64_t cnt   = 0;
      bool    first = true;
      for(Sample const & sm : samples)
      {
        check_launch(launch_features<double>(X_.data(), T_.data(), history_at(P, n_dofs, sm.t),
                                             Bv + (size_t)sm.t * n_dofs, P + (size_t)sm.t * n_dofs,
                                             history_, n3_, sm.cell0, sm.B, sc, beta_, s));
        forward(X_.data(), sm.B, s);
        int64_t const n = (int64_t)sm.B * n3_;
        check_launch(launch_mse_backward(Y_.data(), T_.data(), dY_.data(), n, 1.f / (float)n,
                                         partial_.data(), loss_.data(), !first, s));
        first = false;
        backward(X_.data(), sm.B, s);
        check_launch(launch_adam(params_.data(), grads_.data(), m_.data(), v_.data(), n_params_, lr,
                                 0.9f, 0.999f, 1e-8f, ++adam_step_, s));
        cnt += n;
      }
      double l = 0;
      HIP_CHECK(hipMemcpyAsync(&l, loss_.data(), sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      last = l / (double)std::max<int64_t>(cnt, 1);
      if(!std::isfinite(last))
        EXADG_THROW(EXADG_ERR_NOT_FINITE, "non-finite training loss");
    }
    return last;
  }
  // operator losses: normalization of each step by the extrapolation's loss
  std::vector<int> steps;
  for(int t = t_begin + history_; t < t_end; ++t)
    steps.push_back(t);
  std::vector<double> norm(steps.size());
  for(size_t i = 0; i < steps.size(); ++i)
  {
    norm[i] = field_loss(*op, loss, P, Bv, n_dofs, steps[i], false, 0.f, s);
    if(!(norm[i] > 0.0))
      norm[i] = 1.0;
  }
  std::vector<size_t> order(steps.size());
  std::iota(order.begin(), order.end(), 0);
  for(int ep = 0; ep < epochs; ++ep)
  {
    std::shuffle(order.begin(), order.end(), rng);
    double sum = 0;
    for(size_t i : order)
    {
      double const l =
        field_loss(*op, loss, P, Bv, n_dofs, steps[i], true, (float)(1.0 / norm[i]), s);
      sum += l / norm[i];
      backward(X_.data(), (int)n_cells, s);
      check_launch(launch_adam(params_.data(), grads_.data(), m_.data(), v_.data(), n_params_, lr,
                               0.9f, 0.999f, 1e-8f, ++adam_step_, s));
    }
    last = sum / (double)steps.size();
    if(!std::isfinite(last))
      EXADG_THROW(EXADG_ERR_NOT_FINITE, "non-finite training loss");
  }
  HIP_CHECK(hipStreamSynchronize(s));
  return last;
}

double GuessNet::evaluate(Poisson::LaplaceOperatorBase const * op, GuessLoss loss, double const * P,
                          double const * Bv, int64_t n_dofs, int t_begin, int t_end, int batch_cells,
                          hipStream_t s)
{
  check_range(history_, n3_, n_dofs, t_begin, t_end);
  int64_t const n_cells = n_dofs / n3_;
  if(loss != GuessLoss::l2)
  {
    if(op == nullptr || op->precision() != EXADG_FP64 || op->m() != n_dofs)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "the operator losses need the FP64 operator of the sequence");
    double sum = 0;
    int    cnt = 0;
    for(int t = t_begin + history_; t < t_end; ++t, ++cnt)
    {
      double nrm = field_loss(*op, loss, P, Bv, n_dofs, t, false, 0.f, s);
      double l   = field_loss(*op, loss, P, Bv, n_dofs, t, true, 0.f, s);
      sum += l / (nrm > 0.0 ? nrm : 1.0);
    }
    return sum / std::max(cnt, 1);
  }
  batch_cells = (int)std::max<int64_t>(1, std::min<int64_t>(batch_cells, n_cells));
  ensure_workspace(batch_cells);
  FeatureScales const sc    = feature_scales();
  int64_t             cnt   = 0;
  bool                first = true;
  for(Sample const & sm : make_samples(t_begin + history_, t_end, n_cells, batch_cells))
  {
    check_launch(launch_features<double>(X_.data(), T_.data(), history_at(P, n_dofs, sm.t),
                                         Bv + (size_t)sm.t * n_dofs, P + (size_t)sm.t * n_dofs,
                                         history_, n3_, sm.cell0, sm.B, sc, beta_, s));
    forward(X_.data(), sm.B, s);
    int64_t const n = (int64_t)sm.B * n3_;
    check_launch(launch_mse_backward(Y_.data(), T_.data(), dY_.data(), n, 1.f, partial_.data(),
                                     loss_.data(), !first, s));
    first = false;
    cnt += n;
  }
  double l = 0;
  HIP_CHECK(hipMemcpyAsync(&l, loss_.data(), sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return l / (double)std::max<int64_t>(cnt, 1);
}

int GuessNet::predict(double * x0, History<double> const & hist, double const * rhs, int64_t n_cells,
                      int batch_cells, hipStream_t s)
{
  if(x0 == nullptr || rhs == nullptr)
    EXADG_THROW(EXADG_ERR_ARGUMENT, "null vector");
  for(int k = 0; k < history_; ++k)
    if(hist.p[k] == nullptr)
      EXADG_THROW(EXADG_ERR_ARGUMENT, "null history vector");
  batch_cells = (int)std::max<int64_t>(1, std::min<int64_t>(batch_cells, n_cells));
  ensure_workspace(batch_cells);
  FeatureScales const sc = feature_scales();
  check_launch(launch_fill_u32(bad_.data(), 0u, 1, s));
  for(int64_t c0 = 0; c0 < n_cells; c0 += batch_cells)
  {
    int const B = (int)std::min<int64_t>(batch_cells, n_cells - c0);
    check_launch(launch_features<double>(X_.data(), nullptr, hist, rhs, nullptr, history_, n3_, c0, B,
                                         sc, beta_, s));
    float const *   Y = forward(X_.data(), B, s);
    History<double> ho;
    for(int k = 0; k < history_; ++k)
      ho.p[k] = hist.p[k] + c0 * n3_;
    check_launch(launch_combine(x0 + c0 * n3_, ho, history_, beta_, Y, st_, nullptr,
                                (int64_t)B * n3_, bad_.data(), s));
  }
  unsigned int bad = 0;
  HIP_CHECK(hipMemcpyAsync(&bad, bad_.data(), sizeof(unsigned int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if(bad > 0)  // fallback: the plain extrapolation
    check_launch(launch_combine(x0, hist, history_, beta_, nullptr, 0.0, nullptr, n_cells * n3_,
                                nullptr, s));
  return (int)bad;
}

}  // namespace ExaDG
