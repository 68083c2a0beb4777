#include "ensemble.hpp"

#include <algorithm>
#include <chrono>

#include "common.hpp"

namespace stl {

namespace {
constexpr size_t kMaxGraphs = 4;
}

Ensemble::Ensemble(const std::vector<EnsembleGroup>& groups, bool fp64, int device) {
  auto t0 = std::chrono::steady_clock::now();
  STL_CHECK(!groups.empty(), "an ensemble needs at least one group");
  STL_HIP(hipSetDevice(device));
  // Largest nside first: its engine builds the operators the others share.
  engines_.resize(groups.size());
  std::vector<size_t> by_size(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) by_size[g] = g;
  std::stable_sort(by_size.begin(), by_size.end(),
                   [&](size_t a, size_t b) { return groups[a].nside > groups[b].nside; });
  for (size_t g : by_size) {
    EngineOptions o;
    o.nside = groups[g].nside;
    o.J = groups[g].J;
    o.batch = groups[g].batch;
    o.fp64 = fp64;
    o.device = device;
    o.timing = false;
    engines_[g] = Engine::create(o);
  }
  STL_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Ensemble::~Ensemble() {
  for (auto* c : {&fwd_, &vjp_})
    for (auto& e : *c) (void)hipGraphExecDestroy(e.exec);
  if (s_) (void)hipStreamDestroy(s_);
}

size_t Ensemble::values() const {
  size_t n = 0;
  for (auto& e : engines_) n += size_t(e->options().batch) * e->plan().size();
  return n;
}

template <typename Capture>
hipGraphExec_t Ensemble::get(std::vector<Cached>& cache, const std::vector<const void*>& ptrs, bool flag,
                             Capture&& capture) {
  for (auto& c : cache)
    if (c.ptrs == ptrs && c.flag == flag) return c.exec;
  hipGraph_t graph = nullptr;
  STL_HIP(hipGraphCreate(&graph, 0));
  try {
    for (size_t g = 0; g < engines_.size(); ++g) {
      hipGraph_t child = capture(g);
      hipGraphNode_t node;
      hipError_t err = hipGraphAddChildGraphNode(&node, graph, nullptr, 0, child);
      (void)hipGraphDestroy(child);
      STL_HIP(err);
    }
  } catch (...) {
    (void)hipGraphDestroy(graph);
    throw;
  }
  hipGraphExec_t exec = nullptr;
  hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  STL_HIP(err);
  if (cache.size() == kMaxGraphs) {
    (void)hipGraphExecDestroy(cache.front().exec);
    cache.erase(cache.begin());
  }
  cache.push_back(Cached{ptrs, flag, exec});
  return exec;
}

void Ensemble::forward(const double* const* x_dev, bool moments, cplx* values, double* variance) {
  std::vector<const void*> ptrs(x_dev, x_dev + engines_.size());
  hipGraphExec_t exec =
      get(fwd_, ptrs, moments, [&](size_t g) { return engines_[g]->capture_forward(x_dev[g], moments); });
  STL_HIP(hipGraphLaunch(exec, s_));
  STL_HIP(hipStreamSynchronize(s_));
  size_t at = 0;
  for (auto& e : engines_) {
    e->finish_forward(moments, values + at, variance ? variance + at : nullptr);
    at += size_t(e->options().batch) * e->plan().size();
  }
}

void Ensemble::vjp(const cplx* cot, double* const* grad_dev) {
  std::vector<const void*> ptrs(grad_dev, grad_dev + engines_.size());
  size_t at = 0;
  for (auto& e : engines_) {
    e->stage_cotangent(cot + at);
    at += size_t(e->options().batch) * e->plan().size();
  }
  hipGraphExec_t exec = get(vjp_, ptrs, false, [&](size_t g) { return engines_[g]->capture_vjp(grad_dev[g]); });
  STL_HIP(hipGraphLaunch(exec, s_));
  STL_HIP(hipStreamSynchronize(s_));
}

}  // namespace stl
