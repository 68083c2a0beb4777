#include "ensemble.hpp"

#include <algorithm>
#include <chrono>

#include "common.hpp"

namespace stl {

namespace {
constexpr size_t kMaxGraphs = 4;
}

Ensemble::Ensemble(const std::vector<EnsembleGroup>& groups, bool fp64, int device, int batch_nside) {
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
    o.batch_nside = batch_nside;
    engines_[g] = Engine::create(o);
  }
  // Default-priority streams, one per group: the groups' graphs run
  // concurrently (their small levels overlap the others' large kernels),
  // launched largest group first.
  streams_.assign(groups.size(), nullptr);
  order_.resize(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) order_[g] = g;
  auto work = [&](size_t g) { return double(groups[g].batch) * double(groups[g].nside) * groups[g].nside; };
  std::stable_sort(order_.begin(), order_.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
  for (auto& st : streams_) STL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

Ensemble::~Ensemble() {
  for (auto& st : streams_)
    if (st) (void)hipStreamSynchronize(st);
  for (auto* c : {&fwd_, &vjp_})
    for (auto& e : *c)
      for (auto x : e.execs) (void)hipGraphExecDestroy(x);
  for (auto& st : streams_)
    if (st) (void)hipStreamDestroy(st);
}

size_t Ensemble::values() const {
  size_t n = 0;
  for (auto& e : engines_) n += size_t(e->options().batch) * e->plan().size();
  return n;
}

template <typename Capture>
const std::vector<hipGraphExec_t>& Ensemble::get(std::vector<Cached>& cache, const std::vector<const void*>& ptrs,
                                                 bool flag, Capture&& capture) {
  for (auto& c : cache)
    if (c.ptrs == ptrs && c.flag == flag) return c.execs;
  std::vector<hipGraphExec_t> execs;
  try {
    for (size_t g = 0; g < engines_.size(); ++g) {
      hipGraph_t graph = capture(g);
      hipGraphExec_t exec = nullptr;
      hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      STL_HIP(err);
      execs.push_back(exec);
    }
  } catch (...) {
    for (auto x : execs) (void)hipGraphExecDestroy(x);
    throw;
  }
  if (cache.size() == kMaxGraphs) {
    for (auto x : cache.front().execs) (void)hipGraphExecDestroy(x);
    cache.erase(cache.begin());
  }
  cache.push_back(Cached{ptrs, flag, std::move(execs)});
  return cache.back().execs;
}

void Ensemble::launch(const std::vector<hipGraphExec_t>& execs) {
  hipError_t first = hipSuccess;
  for (size_t i = 0; i < execs.size() && first == hipSuccess; ++i) {
    const size_t g = order_[i];
    first = hipGraphLaunch(execs[g], streams_[g]);
  }
  // Join every stream, also after a failed launch, before reporting.
  for (auto& st : streams_) {
    hipError_t err = hipStreamSynchronize(st);
    if (first == hipSuccess) first = err;
  }
  STL_HIP(first);
}

void Ensemble::forward(const double* const* x_dev, bool moments, cplx* values, double* variance) {
  std::vector<const void*> ptrs(x_dev, x_dev + engines_.size());
  launch(get(fwd_, ptrs, moments, [&](size_t g) { return engines_[g]->capture_forward(x_dev[g], moments); }));
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
  launch(get(vjp_, ptrs, false, [&](size_t g) { return engines_[g]->capture_vjp(grad_dev[g]); }));
}

}  // namespace stl
