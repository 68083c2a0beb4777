// An ensemble of maps of uneven nside on one GPU: one batched engine per
// nside group (operators shared through the per-device cache, so a group's
// level at nside s reuses the operators of every other group's level at s),
// the forward (and the VJP) of all groups composed into ONE HIP graph whose
// groups are independent branches, launched once per evaluation.
#pragma once

#include <memory>
#include <vector>

#include "engine.hpp"

namespace stl {

struct EnsembleGroup {
  int nside = 128, J = 6, batch = 1;
};

class Ensemble {
 public:
  Ensemble(const std::vector<EnsembleGroup>& groups, bool fp64, int device);
  ~Ensemble();
  Ensemble(const Ensemble&) = delete;
  Ensemble& operator=(const Ensemble&) = delete;

  size_t groups() const { return engines_.size(); }
  Engine& engine(size_t g) { return *engines_[g]; }
  // Statistics of every group, concatenated [group][batch][ncoef].
  size_t values() const;
  // x_dev[g]: float64 [B_g, 12 nside_g^2] on the device.
  void forward(const double* const* x_dev, bool moments, cplx* values, double* variance);
  // cot [group][batch][ncoef]; grad_dev[g] as x_dev.
  void vjp(const cplx* cot, double* const* grad_dev);
  double setup_seconds() const { return setup_s_; }

 private:
  struct Cached {
    std::vector<const void*> ptrs;
    bool flag;
    hipGraphExec_t exec;
  };
  template <typename Capture>
  hipGraphExec_t get(std::vector<Cached>& cache, const std::vector<const void*>& ptrs, bool flag, Capture&& capture);

  std::vector<std::unique_ptr<Engine>> engines_;
  hipStream_t s_ = nullptr;
  std::vector<Cached> fwd_, vjp_;
  double setup_s_ = 0.0;
};

}  // namespace stl
