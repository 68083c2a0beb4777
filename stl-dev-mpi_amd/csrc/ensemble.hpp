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
  // batch_nside: as EngineOptions::batch_nside, for every group.
  Ensemble(const std::vector<EnsembleGroup>& groups, bool fp64, int device, int batch_nside = 16);
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
  // One executable graph per group, keyed by the groups' device pointers.
  struct Cached {
    std::vector<const void*> ptrs;
    bool flag;
    std::vector<hipGraphExec_t> execs;
  };
  template <typename Capture>
  const std::vector<hipGraphExec_t>& get(std::vector<Cached>& cache, const std::vector<const void*>& ptrs,
                                         bool flag, Capture&& capture);
  // Each group's graph on its own stream, then all streams joined.
  void launch(const std::vector<hipGraphExec_t>& execs);

  std::vector<std::unique_ptr<Engine>> engines_;
  std::vector<hipStream_t> streams_;
  std::vector<size_t> order_;  // launch order: most pixels first
  std::vector<Cached> fwd_, vjp_;
  double setup_s_ = 0.0;
};

}  // namespace stl
