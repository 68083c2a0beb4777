// Host orchestration of the level sweeps on one GPU (the reference's
// LocalEngine::forward / reverse / map_cotangent, crates/stl-core/src/engine.rs
// :500-779, with the whole sky resident instead of blocks and halos).
//
// Forward, per level n ascending:
//   maps[n]  = D maps[n-1]                          (dec, rows B)
//   z_n      = W maps[n], |z_n| -> modulus row n     (conv, rows B)
//   modulus rows d < n = D modulus[n-1]              (dec, rows n B, inner kL)
//   branch   = W modulus (all d <= n)                (conv, rows (n+1) B, inner kL)
//   Gram sums of V = [branch(d)[l1][l3], z[l3]] and the S1/S2, mean/var sums.
// Reverse, per level n descending (the cotangents of the Gram products are
// the Hermitian matrices H = C + C^H of the coefficient cotangents / N):
//   gV = H V -> branch cotangent, z part           (gv)
//   chain[n] = W^T gbranch + D^T chain[n+1]          (conv adjoint, dec adjoint)
//   gz = z part + S1/S2 terms + |.|' of chain row n  (gz)
//   mg[n] = W^T gz + D^T mg[n+1]                      (conv adjoint, dec adjoint)
// which is the reference's reverse with its Horner chain of D^T over levels.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <unordered_map>

#include "geometry.hpp"
#include "sweeps.hpp"

namespace stl {

std::atomic<size_t>& device_bytes_live() {
  static std::atomic<size_t> v{0};
  return v;
}
std::atomic<size_t>& device_bytes_peak() {
  static std::atomic<size_t> v{0};
  return v;
}

std::vector<Coefficient> build_plan(int J, int L) {
  std::vector<Coefficient> plan;
  plan.push_back(Coefficient{F_MEAN, 0, 0, 0});
  plan.push_back(Coefficient{F_VAR, 0, 0, 0});
  for (int j3 = 0; j3 < J; ++j3) {
    int dg = j3;
    for (int fam : {F_S1, F_S2})
      for (int l = 0; l < L; ++l) {
        Coefficient c{fam, dg, 0, 0};
        c.j1 = j3;
        c.l1 = l;
        plan.push_back(c);
      }
    for (int j2 = 0; j2 <= j3; ++j2) {
      for (int l2 = 0; l2 < L; ++l2)
        for (int l3 = 0; l3 < L; ++l3) {
          Coefficient c{F_S3, dg, 0, 0};
          c.j1 = j2; c.j3 = j3; c.l1 = l2; c.l3 = l3;
          plan.push_back(c);
        }
      for (int j1 = 0; j1 <= j2; ++j1)
        for (int l1 = 0; l1 < L; ++l1)
          for (int l2 = 0; l2 < L; ++l2)
            for (int l3 = 0; l3 < L; ++l3) {
              Coefficient c{F_S4, dg, 0, 0};
              c.j1 = j1; c.j2 = j2; c.j3 = j3; c.l1 = l1; c.l2 = l2; c.l3 = l3;
              plan.push_back(c);
            }
    }
  }
  std::stable_sort(plan.begin(), plan.end(),
                   [](const Coefficient& a, const Coefficient& b) { return a.native_dg < b.native_dg; });
  return plan;
}

std::vector<Coefficient> build_cross_plan(int J, int L) {
  std::vector<Coefficient> plan;
  for (int j3 = 0; j3 < J; ++j3) {
    const int dg = j3;
    for (int fam : {F_S1, F_S2})
      for (int l = 0; l < L; ++l) {
        Coefficient c{fam, dg, 0, 1};
        c.j1 = j3;
        c.l1 = l;
        plan.push_back(c);
      }
    for (int ca = 0; ca < 2; ++ca) {
      const int cb = 1 - ca;
      for (int j2 = 0; j2 <= j3; ++j2) {
        for (int l2 = 0; l2 < L; ++l2)
          for (int l3 = 0; l3 < L; ++l3) {
            Coefficient c{F_S3, dg, ca, cb};
            c.j1 = j2; c.j3 = j3; c.l1 = l2; c.l3 = l3;
            plan.push_back(c);
          }
        for (int j1 = 0; j1 <= j2; ++j1)
          for (int l1 = 0; l1 < L; ++l1)
            for (int l2 = 0; l2 < L; ++l2)
              for (int l3 = 0; l3 < L; ++l3) {
                Coefficient c{F_S4, dg, ca, cb};
                c.j1 = j1; c.j2 = j2; c.j3 = j3; c.l1 = l1; c.l2 = l2; c.l3 = l3;
                plan.push_back(c);
              }
      }
    }
  }
  std::stable_sort(plan.begin(), plan.end(),
                   [](const Coefficient& a, const Coefficient& b) { return a.native_dg < b.native_dg; });
  return plan;
}

namespace {

inline int gram_k(int n) { return (n + 1) * kL + 1; }
inline int gram_e(int n) { int K = gram_k(n); return K * (K + 1) / 2; }
inline int tri_index(int K, int i, int j) { return i * K - i * (i - 1) / 2 + (j - i); }

// Where a coefficient's sum lives.
struct SumSlot {
  enum Kind { MEANVAR, ZSTAT, GRAM, ZCROSS, GRAMX } kind;
  int n = 0;      // level
  int l = 0;      // gauge (S1/S2: l1, Gram: l3)
  int idx = 0;    // meanvar / zstat column, or Gram entry
  bool conj = false;
  int gi = -1, gj = -1;  // Gram operand indices (reverse)
};

// A coefficient of the cross plan: S1 from the S1 cross sums; S2, S3 and S4
// entries (i, j) of the cross Gram block G = V_a V_b^H (channel a of the
// coefficient first: for c1 = 1 the conjugate of entry (j, i)).
SumSlot cross_slot_of(const Coefficient& c) {
  SumSlot s;
  s.n = c.native_dg;
  const int K = gram_k(s.n);
  if (c.family == F_S1) {
    s.kind = SumSlot::ZCROSS;
    s.l = c.l1;
    return s;
  }
  s.kind = SumSlot::GRAMX;
  int i, j;
  if (c.family == F_S2) {
    s.l = c.l1;
    i = j = K - 1;
  } else {
    s.l = c.l3;
    i = c.j1 * kL + c.l1;
    j = c.family == F_S3 ? K - 1 : c.j2 * kL + c.l2;
  }
  if (c.c1 == 0) { s.gi = i; s.gj = j; s.conj = false; }
  else { s.gi = j; s.gj = i; s.conj = true; }
  s.idx = s.gi * K + s.gj;
  return s;
}

SumSlot slot_of(const Coefficient& c) {
  SumSlot s;
  s.n = c.native_dg;
  switch (c.family) {
    case F_MEAN: s.kind = SumSlot::MEANVAR; s.idx = 0; break;
    case F_VAR: s.kind = SumSlot::MEANVAR; s.idx = 1; break;
    case F_S1: s.kind = SumSlot::ZSTAT; s.l = c.l1; s.idx = 0; break;
    case F_S2: s.kind = SumSlot::ZSTAT; s.l = c.l1; s.idx = 1; break;
    case F_S3:
    case F_S4: {
      s.kind = SumSlot::GRAM;
      s.l = c.l3;
      int K = gram_k(s.n);
      int i = c.j1 * kL + c.l1;
      int j = c.family == F_S3 ? K - 1 : c.j2 * kL + c.l2;
      s.gi = i;
      s.gj = j;
      if (i <= j) { s.idx = tri_index(K, i, j); s.conj = false; }
      else { s.idx = tri_index(K, j, i); s.conj = true; }
      break;
    }
  }
  return s;
}

// Convolution variants chosen by timing: (device, fp64, order, kernel, rows).
using TuneKey = std::tuple<int, bool, int, int, int>;
std::map<TuneKey, int>& tune_cache() {
  static std::map<TuneKey, int> c;
  return c;
}
std::mutex& tune_mutex() {
  static std::mutex m;
  return m;
}

// Timer name of a kernel family at a level.
std::string tag(const char* family, int level) { return std::string(family) + "@L" + std::to_string(level); }

// HIP-event timers by kernel family.
class Profiler {
 public:
  void enable(bool on, hipStream_t s) { on_ = on; s_ = s; }
  bool on() const { return on_; }
  int start() {
    if (!on_) return -1;
    hipEvent_t a = get();
    STL_HIP(hipEventRecord(a, s_));
    pending_.push_back({"", a, nullptr, 0.0});
    return int(pending_.size()) - 1;
  }
  void stop(int id, const std::string& name, double bytes) {
    if (id < 0) return;
    hipEvent_t b = get();
    STL_HIP(hipEventRecord(b, s_));
    pending_[id].name = name;
    pending_[id].b = b;
    pending_[id].bytes = bytes;
  }
  // After a stream synchronisation.
  void collect() {
    for (auto& p : pending_) {
      float ms = 0.f;
      STL_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      auto& t = acc_[p.name];
      t.name = p.name;
      t.ms += ms;
      t.launches += 1;
      t.bytes += p.bytes;
      pool_.push_back(p.a);
      pool_.push_back(p.b);
    }
    pending_.clear();
  }
  std::vector<KernelTimer> timers() const {
    std::vector<KernelTimer> v;
    for (auto& kv : acc_) v.push_back(kv.second);
    return v;
  }
  void reset() { acc_.clear(); }
  ~Profiler() {
    for (auto e : pool_) (void)hipEventDestroy(e);
    for (auto& p : pending_) {
      (void)hipEventDestroy(p.a);
      if (p.b) (void)hipEventDestroy(p.b);
    }
  }

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
    double bytes;
  };
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    STL_HIP(hipEventCreate(&e));
    return e;
  }
  bool on_ = false;
  hipStream_t s_ = nullptr;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, KernelTimer> acc_;
};

template <typename T>
class EngineT final : public Engine {
  using C2 = cplx_t<T>;

  // One batched launch of the coarse levels: independent tasks on the device.
  struct Batch {
    DevArray<LevelTask<T>> tasks;
    int ntasks = 0;
    int64_t items = 0;
    size_t lds = 0;
    double bytes = 0.0;  // algorithmic bytes of its tasks
    std::string name;
  };

  // The level fields (forward state and reverse scratch), shared by the
  // engines made by share(); holder: the engine whose forward they hold.
  struct FieldSet {
    int rows = 0;  // field rows the buffers are sized for
    uint64_t holder = 0;
    // Recompute store: z and br hold one level at a time (sized for level 0),
    // and every level's reverse uses gbr, gzp and acc[n & 1].
    bool recompute = false;
    size_t bytes = 0;  // device bytes of the fields and reverse buffers
    std::vector<DevArray<T>> maps, M;
    std::vector<DevArray<C2>> z, br;
    // Reverse buffers: gbr / gzp / acc[0] hold level 0, gbr1 / gzp1 / acc[1]
    // level 1, gbr2 / gzp2 and acc[2 + (n & 1)] the levels n >= 2, so that
    // level 0's and level 1's gV and branch adjoints can run beside the
    // coarser levels' reverse.
    DevArray<C2> gbr, gzp, gz, gbr1, gzp1, gbr2, gzp2;
    DevArray<T> acc[4], mg[2];
    // Reverse buffers of every batched level (from level qc on), so that one
    // launch can hold a step of all of them: gV, z part, chain, gz and map
    // cotangent.
    int qc = 0;
    std::vector<DevArray<C2>> qgbr, qgzp, qgz;
    std::vector<DevArray<T>> qchain, qmg;
  };

 public:
  // fields: the field set of the engine this one shares (nullptr: its own).
  explicit EngineT(const EngineOptions& opt, std::shared_ptr<FieldSet> fields = nullptr)
      : opt_(opt), f_(std::move(fields)) {
    static std::atomic<uint64_t> next_id{1};
    id_ = next_id.fetch_add(1);
    auto t0 = std::chrono::steady_clock::now();
    STL_CHECK(opt.nside >= 1 && (opt.nside & (opt.nside - 1)) == 0, "nside must be a power of two");
    STL_CHECK(opt.nside <= kMaxNside, "nside must be at most 8192 (32-bit pixel ids)");
    int order0 = 0;
    while ((1 << order0) < opt.nside) ++order0;
    STL_CHECK(opt.J >= 1 && opt.J <= order0 + 1, "invalid number of scales");
    STL_CHECK(opt.J <= 12, "at most 12 scales");
    STL_CHECK(opt.batch >= 1, "batch must be positive");
    STL_HIP(hipSetDevice(opt.device));
    STL_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    // Side streams of the split sweeps, default priority as the engine stream.
    STL_HIP(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking));
    STL_HIP(hipStreamCreateWithFlags(&s3_, hipStreamNonBlocking));
    for (auto& ev : ev_) STL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    prof_.enable(opt.timing, s_);
    cross_ = opt.cross;
    plan_ = cross_ ? build_cross_plan(opt.J, kL) : build_plan(opt.J, kL);
    if (!opt.families.empty()) {
      // A plan of some families only (the reference's build_plan families,
      // crates/stl-core/src/plan.rs:80-100), in the full plan's order.
      std::vector<Coefficient> kept;
      for (const auto& c : plan_)
        if (std::find(opt.families.begin(), opt.families.end(), c.family) != opt.families.end()) kept.push_back(c);
      STL_CHECK(!kept.empty(), "no coefficient of the requested families in the plan");
      plan_.swap(kept);
    }
    levels_ = opt.J;
    for (auto& c : plan_) slots_.push_back(cross_ ? cross_slot_of(c) : slot_of(c));
    Bs_ = opt.batch;
    B_ = cross_ ? 2 * opt.batch : opt.batch;
    for (int n = 0; n < levels_; ++n) {
      int order = order0 - n;
      N_.push_back(12 * (int64_t(1) << (2 * order)));
      conv_.push_back(shared_conv<T>(opt.device, order, s_));
      if (n + 1 < levels_) dec_.push_back(shared_dec<T>(opt.device, order, s_));
    }
    geometry_bytes_ = device_bytes_live();
    // Batched from the first level (n >= 2) whose pixels over all the field
    // rows number at most 12 batch_nside^2: below that a level's kernels are
    // launch-bound, above it the tuned per-level kernels do better.
    qc_ = levels_;
    if (opt.batch_nside > 0)
      for (int n = 2; n < levels_; ++n)
        if (int64_t(B_) * N_[n] <= 12 * int64_t(opt.batch_nside) * opt.batch_nside) {
          qc_ = n;
          break;
        }
    // One workgroup per item: the dispatcher balances the items over the CUs,
    // also beside other streams' kernels (a fixed stride over a persistent grid
    // waits for its last workgroup to find a free CU).
    q_workgroups_ = std::numeric_limits<int>::max();
    allocate();
    tune();
    // The timing launches of tune() wrote the (possibly shared) fields.
    f_->holder = id_;
    STL_HIP(hipStreamSynchronize(s_));
    setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  ~EngineT() override {
    for (auto* cache : {&fwd_graphs_, &vjp_graphs_})
      for (auto& e : *cache)
        for (int i = 0; i < e.count; ++i) (void)hipGraphExecDestroy(e.exec[i]);
    for (auto& ev : ev_)
      if (ev) (void)hipEventDestroy(ev);
    if (s3_) (void)hipStreamDestroy(s3_);
    if (s2_) (void)hipStreamDestroy(s2_);
    if (s_) (void)hipStreamDestroy(s_);
  }

  std::unique_ptr<Engine> share(int batch, bool cross, const std::vector<int>& families) override {
    EngineOptions o = opt_;
    o.batch = batch;
    o.cross = cross;
    o.families = families;
    o.conv_variant = opt_.conv_variant;
    return std::unique_ptr<Engine>(new EngineT<T>(o, f_));
  }

  // The reverse and the state copy need this engine's forward in the fields.
  void check_holder() const {
    STL_CHECK(f_->holder == id_, "gradient without the forward sweep of this evaluator (shared fields)");
  }

  const EngineOptions& options() const override { return opt_; }
  const std::vector<Coefficient>& plan() const override { return plan_; }
  int levels() const override { return levels_; }
  int masks() const override { return nmask_; }
  std::vector<int64_t> mask_counts() const override { return mcount_; }

  void set_masks(const double* w, int k) override {
    STL_CHECK(k >= 0 && k <= 64, "at most 64 masks");
    // Membership bits of every pixel at every level: level n + 1 holds the
    // parents of level n (a parent is in a mask when one of its 4 NESTED
    // children is).
    std::vector<std::vector<uint64_t>> bits(levels_);
    std::vector<int64_t> counts(size_t(k) * levels_, 0);
    if (k > 0) {
      STL_CHECK(w != nullptr, "null mask weights");
      const int64_t N0 = N_[0];
      bits[0].assign(size_t(N0), 0);
      for (int q = 0; q < k; ++q)
        for (int64_t p = 0; p < N0; ++p) {
          const double v = w[size_t(q) * N0 + p];
          STL_CHECK(std::isfinite(v), "mask weights must be finite");
          if (v != 0.0) bits[0][p] |= uint64_t(1) << q;
        }
      for (int n = 1; n < levels_; ++n) {
        bits[n].assign(size_t(N_[n]), 0);
        const auto& f = bits[n - 1];
        for (int64_t p = 0; p < N_[n]; ++p) bits[n][p] = f[4 * p] | f[4 * p + 1] | f[4 * p + 2] | f[4 * p + 3];
      }
      for (int n = 0; n < levels_; ++n)
        for (uint64_t b : bits[n])
          for (int q = 0; q < k; ++q) counts[size_t(q) * levels_ + n] += int64_t((b >> q) & 1);
      for (int64_t c : counts) STL_CHECK(c > 0, "empty mask");
    }
    // Classes: the distinct membership patterns of each level, numbered in
    // order of first appearance.
    std::vector<std::vector<uint16_t>> cls(levels_);
    std::vector<std::vector<uint64_t>> pat(levels_);
    if (k > 0)
      for (int n = 0; n < levels_; ++n) {
        std::unordered_map<uint64_t, int> id;
        cls[n].resize(bits[n].size());
        for (size_t p = 0; p < bits[n].size(); ++p) {
          auto it = id.find(bits[n][p]);
          if (it == id.end()) {
            STL_CHECK(pat[n].size() < 65535, "too many distinct mask overlaps");
            it = id.emplace(bits[n][p], int(pat[n].size())).first;
            pat[n].push_back(bits[n][p]);
          }
          cls[n][p] = uint16_t(it->second);
        }
      }
    drop_graphs();
    nmask_ = k;
    Km_ = std::max(k, 1);
    mcount_ = counts;
    mcls_.clear();
    mpat_.clear();
    mC_.assign(levels_, 1);
    mcls_.resize(levels_);
    mpat_.resize(levels_);
    for (int n = 0; n < levels_ && k > 0; ++n) {
      upload(mcls_[n], cls[n]);
      upload(mpat_[n], pat[n]);
      mC_[n] = int(pat[n].size());
    }
    alloc_sums();
  }
  hipStream_t stream() const override { return s_; }
  size_t device_bytes() const override { return device_bytes_live(); }
  int field_store() const override { return f_->recompute ? EngineOptions::kRecompute : EngineOptions::kKeep; }
  size_t field_bytes() const override { return f_->bytes; }
  double setup_seconds() const override { return setup_s_; }
  std::vector<KernelTimer> timers() const override { return prof_.timers(); }
  std::vector<std::array<int, 4>> conv_variants() const override { return conv_var_; }
  std::vector<TuneRecord> tuning() const override { return tune_ms_; }
  void reset_timers() override { prof_.reset(); }

  void forward_device(const double* x_dev, bool moments, cplx* values, double* variance) override {
    f_->holder = id_;
    run_steps(fwd_graphs_, x_dev, moments, forward_steps(x_dev, moments));
    finish_forward(moments, values, variance);
  }

  hipGraph_t capture_forward(const double* x_dev, bool moments) override {
    STL_CHECK(!prof_.on(), "graph capture with per-kernel timing");
    f_->holder = id_;
    return capture([&] { enqueue_forward(x_dev, moments); });
  }

  void stage_cotangent(const cplx* cot) override {
    std::memcpy(hcot_.data(), cot, nvals() * sizeof(double2));
  }

  hipGraph_t capture_vjp(double* grad_dev) override {
    STL_CHECK(!prof_.on(), "graph capture with per-kernel timing");
    check_holder();
    return capture([&] { sweep_reverse(grad_dev); });
  }

  void forward_levels_device(const double* x_dev, cplx* values) override {
    f_->holder = id_;
    size_t at = 0;
    for (int n = 0; n < levels_; ++n) {
      int64_t len = int64_t(B_) * N_[n];
      cast_in(x_dev + at, f_->maps[n].data(), len);
      at += size_t(len);
    }
    sweep_forward(false);
    STL_HIP(hipStreamSynchronize(s_));
    prof_.collect();
    finish_forward(false, values, nullptr);
  }

  void vjp_device(const cplx* cot, double* grad_dev) override {
    check_holder();
    stage_cotangent(cot);
    run_steps(vjp_graphs_, grad_dev, false, reverse_steps(grad_dev));
  }

  // The reverse sweep in phases: the prelude (cotangents onto the sums'
  // operands), each level's gV and branch adjoint (rev_branch), each level's
  // chain, z cotangent and z adjoint (rev_chain), then the map gradient.
  // Level 0's rev_branch touches only level-0 buffers, so it may run beside
  // levels J - 1 .. 1; rev_chain(0) needs both.
  void sweep_reverse(double* grad_dev) {
    rev_prelude();
    rev_levels(0, grad_dev);
  }

  // The reverse of levels J - 1 .. down_to: the batched levels' steps (a
  // launch each for all of them), then the others level by level.
  // grad_dev: level 0's rev_chain writes the map gradient (down_to 0; the
  // batched levels start at 2).
  void rev_levels(int down_to, double* grad_dev = nullptr) {
    int n = levels_ - 1;
    if (batch_on() && qc_ >= down_to) {
      run_batches(qrev_);
      n = qc_ - 1;
    }
    for (; n >= down_to; --n) {
      rev_branch(n);
      rev_chain(n, n == 0 ? grad_dev : nullptr);
    }
  }

  void rev_prelude() {
    const size_t nv = nvals();
    STL_HIP(hipMemcpyAsync(dcot_.data(), hcot_.data(), nv * sizeof(double2), hipMemcpyHostToDevice, s_));
    launch_scatter_cot(nrt_, Bs_, Km_, levels_, int(plan_.size()), drt_.data(), drk_.data(), dnd_.data(),
                       dcot_.data(), ddst_.data(), s_);
    if (nmask_ > 0) {
      // Per-mask operands -> per-class ones (the sum over the masks holding the class).
      for (int n = 0; n < levels_; ++n) {
        const int K = gram_k(n), C = mC_[n];
        const uint64_t* pat = mpat_[n].data();
        launch_mask_combine(Bs_, Km_, kL, C, int64_t(K) * K * 2, pat,
                            reinterpret_cast<const double*>(dH_[n].data()), reinterpret_cast<double*>(Hc_[n].data()),
                            s_);
        launch_mask_combine(Bs_, Km_, kL, C, cross_ ? 2 : 1, pat, ds1_[n].data(), s1c_[n].data(), s_);
        if (!cross_) launch_mask_combine(Bs_, Km_, kL, C, 1, pat, ds2_[n].data(), s2c_[n].data(), s_);
      }
      if (!cross_) {
        launch_mask_combine(Bs_, Km_, 1, mC_[0], 1, mpat_[0].data(), dgm_.data(), gmc_.data(), s_);
        launch_mask_combine(Bs_, Km_, 1, mC_[0], 1, mpat_[0].data(), dgv_.data(), gvc_.data(), s_);
      }
    }
  }

  // Reverse buffers of level n (levels 0 and 1 apart from the others).
  // Level-n buffers.  Keep store: z and branches of every level; the reverse
  // buffers of levels 0 and 1 apart from the others.  Recompute store: one
  // level's z, branches and reverse buffers at a time.
  C2* z_of(int n) { return f_->z[f_->recompute ? 0 : n].data(); }
  C2* br_of(int n) { return f_->br[f_->recompute ? 0 : n].data(); }
  // The batched levels (n >= qc_) have reverse buffers of their own.
  bool batched(int n) const { return n >= qc_ && batch_on(); }
  C2* gbr_of(int n) {
    if (batched(n)) return f_->qgbr[n].data();
    return n == 0 || f_->recompute ? f_->gbr.data() : n == 1 ? f_->gbr1.data() : f_->gbr2.data();
  }
  C2* gzp_of(int n) {
    if (batched(n)) return f_->qgzp[n].data();
    return n == 0 || f_->recompute ? f_->gzp.data() : n == 1 ? f_->gzp1.data() : f_->gzp2.data();
  }
  T* chain_of(int n) {
    if (batched(n)) return f_->qchain[n].data();
    return f_->acc[f_->recompute ? (n & 1) : n < 2 ? n : 2 + (n & 1)].data();
  }
  C2* gz_of(int n) { return batched(n) ? f_->qgz[n].data() : f_->gz.data(); }
  T* mg_of(int n) { return batched(n) ? f_->qmg[n].data() : f_->mg[n % 2].data(); }

  // The batched launches serve the keep store's one-channel, unmasked sweeps.
  bool batch_on() const { return qc_ < levels_ && !cross_ && nmask_ == 0 && !f_->recompute; }

  // Recompute store: level n's z (with its modulus row, the same bits as the
  // forward's) and branches again, from the kept maps and moduli.
  void recompute_level(int n) {
    const int64_t N = N_[n];
    int id = prof_.start();
    launch_conv_fwd<T>(conv_[n]->store.view(), 1, f_->maps[n].data(), B_, z_of(n),
                       f_->M[n].data() + size_t(n) * B_ * N * kL, s_, conv_var_[n][0]);
    prof_.stop(id, tag("recompute_z", n), conv_bytes(n, 1, B_, false) + double(B_) * N * kL * sizeof(T));
    id = prof_.start();
    launch_conv_fwd<T>(conv_[n]->store.view(), kL, f_->M[n].data(), (n + 1) * B_, br_of(n), nullptr, s_,
                       conv_var_[n][1]);
    prof_.stop(id, tag("recompute_branch", n), conv_bytes(n, kL, (n + 1) * B_, false));
  }

  void rev_branch(int n) {
    const int64_t N = N_[n];
    const MaskView mv = mask_view(n);
    if (f_->recompute) recompute_level(n);
    const double2* H = nmask_ > 0 ? Hc_[n].data() : dH_[n].data();
    int id = prof_.start();
    if (cross_) launch_gv_x<T>(n, N, Bs_, br_of(n), z_of(n), H, gbr_of(n), gzp_of(n), mv, s_);
    else launch_gv<T>(n, N, B_, br_of(n), z_of(n), H, gbr_of(n), gzp_of(n), mv, s_);
    // V (branches (n + 1) L x L and z L complex values per pixel) read, gV written.
    prof_.stop(id, tag("gv", n), double(B_) * N * ((n + 1) * kL * kL + kL) * 2 * sizeof(T) * 2);
    conv_adj(n, kL, gbr_of(n), (n + 1) * B_, chain_of(n));
  }

  // Level n's chain (decimated coarser chain added), gz, z adjoint and map
  // cotangent; at level 0 (grad_dev given) also the map gradient.  With a
  // tiled level-0 decimation, level 0's chain (one row per map, read only by
  // gz) and its map cotangent stay in registers: the decimation adjoints
  // compute gz and the gradient in the same pass.
  void rev_chain(int n, double* grad_dev = nullptr) {
    const int64_t N = N_[n];
    const MaskView mv = mask_view(n);
    T* chain = chain_of(n);
    const double* s1 = nmask_ > 0 ? s1c_[n].data() : ds1_[n].data();
    const double* s2 = nmask_ > 0 ? s2c_[n].data() : ds2_[n].data();
    const bool fused = n == 0 && n + 1 < levels_ && !cross_ && dec_[0]->store.view().ttiled;
    if (fused) {
      int id = prof_.start();
      launch_dec_adj_gz<T>(dec_[0]->store.view(), chain_of(1), chain, B_, z_of(0), gzp_of(0), s1, s2, gz_of(0), mv,
                           s_);
      // The decimation's entries, the coarse chain and the branch part in;
      // z, its gV part in and gz out (complex, L per pixel).
      prof_.stop(id, tag("dec_adj_gz", 0),
                 dec_[0]->tslots * (sizeof(T) + 4) + double(B_) * (N_[1] + N) * kL * sizeof(T) +
                     double(B_) * kL * N * (6 * sizeof(T)));
    } else {
      if (n + 1 < levels_) dec_adj(n, kL, chain_of(n + 1), chain, (n + 1) * B_, chain);
      int id = prof_.start();
      if (cross_)
        launch_gz_x<T>(n, N, Bs_, z_of(n), gzp_of(n), chain, reinterpret_cast<const double2*>(s1), gz_of(n),
                       mv, s_);
      else
        launch_gz<T>(n, N, B_, z_of(n), gzp_of(n), chain, s1, s2, gz_of(n), mv, s_);
      // z, its gV part and gz (complex) and the chain row (real), L each per pixel.
      prof_.stop(id, tag("gz", n), double(B_) * kL * N * (7 * sizeof(T)));
    }

    T* mg = mg_of(n);
    conv_adj(n, 1, gz_of(n), B_, mg);
    if (n == 0 && grad_dev != nullptr) {
      if (fused) {
        int id = prof_.start();
        launch_dec_adj_grad<T>(dec_[0]->store.view(), mg_of(1), mg, B_, f_->maps[0].data(),
                               nmask_ > 0 ? gmc_.data() : dgm_.data(), nmask_ > 0 ? gvc_.data() : dgv_.data(),
                               grad_dev, mask_view(0), s_);
        // The decimation's entries, the coarse and fine cotangents and the map in, the FP64 gradient out.
        prof_.stop(id, tag("dec_adj_grad", 0),
                   dec_[0]->tslots * (sizeof(T) + 4) + double(B_) * (N_[1] + 2 * N) * sizeof(T) +
                       double(B_) * N * sizeof(double));
        return;
      }
      if (n + 1 < levels_) dec_adj(n, 1, mg_of(n + 1), mg, B_, mg);
      rev_finish(grad_dev);
      return;
    }
    if (n + 1 < levels_) dec_adj(n, 1, mg_of(n + 1), mg, B_, mg);
  }

  // The map cotangent out in FP64 (with the mean / variance terms, in one pass).
  void rev_finish(double* grad_dev) {
    if (!cross_)
      launch_meanvar_grad<T>(N_[0], B_, f_->maps[0].data(), nmask_ > 0 ? gmc_.data() : dgm_.data(),
                             nmask_ > 0 ? gvc_.data() : dgv_.data(), f_->mg[0].data(), grad_dev, mask_view(0), s_);
    else
      cast_out(f_->mg[0].data(), grad_dev, int64_t(B_) * N_[0]);
  }

  // Staging of apply_operator: the operator's input and output at the
  // engine's precision, each 256-byte aligned.
  static size_t aligned(size_t b) { return (b + 255) / 256 * 256; }
  std::array<size_t, 2> operator_staging(int kind, int n, int rows) const {
    STL_CHECK(n >= 0 && n < levels_, "level out of range");
    STL_CHECK(rows >= 1, "rows must be positive");
    STL_CHECK(kind >= 0 && kind <= 3, "unknown operator kind");
    const size_t N = size_t(N_[n]);
    if (kind == 0) return {size_t(rows) * N * sizeof(T), size_t(rows) * kL * N * sizeof(C2)};
    if (kind == 1) return {size_t(rows) * kL * N * sizeof(C2), size_t(rows) * N * sizeof(T)};
    STL_CHECK(n + 1 < levels_, "no coarser level");
    const size_t Nc = size_t(N_[n + 1]);
    return {size_t(rows) * (kind == 2 ? N : Nc) * sizeof(T), size_t(rows) * (kind == 2 ? Nc : N) * sizeof(T)};
  }

  size_t operator_workspace_bytes(int kind, int n, int rows) const override {
    const auto st = operator_staging(kind, n, rows);
    return aligned(st[0]) + aligned(st[1]);
  }

  void apply_operator(int kind, int n, int rows, const double* in, double* out) override {
    DevArray<unsigned char> ws(operator_workspace_bytes(kind, n, rows));
    apply_operator_in(kind, n, rows, in, out, ws.data(), ws.size());
  }

  void apply_operator_in(int kind, int n, int rows, const double* in, double* out, void* workspace,
                         size_t workspace_bytes) override {
    const auto st = operator_staging(kind, n, rows);
    STL_CHECK(workspace != nullptr && reinterpret_cast<uintptr_t>(workspace) % 256 == 0,
              "operator workspace: null or not 256-byte aligned");
    STL_CHECK(aligned(st[0]) + aligned(st[1]) <= workspace_bytes, "operator workspace too small");
    unsigned char* a = static_cast<unsigned char*>(workspace);
    unsigned char* b = a + aligned(st[0]);
    const int64_t N = N_[n];
    if (kind == 0) {
      T* x = reinterpret_cast<T*>(a);
      C2* y = reinterpret_cast<C2*>(b);
      cast_in(in, x, int64_t(rows) * N);
      launch_conv_fwd<T>(conv_[n]->store.view(), 1, x, rows, y, nullptr, s_);
      cast_out(reinterpret_cast<const T*>(y), out, int64_t(rows) * kL * N * 2);
    } else if (kind == 1) {
      C2* gy = reinterpret_cast<C2*>(a);
      T* gx = reinterpret_cast<T*>(b);
      cast_in(in, reinterpret_cast<T*>(gy), int64_t(rows) * kL * N * 2);
      launch_conv_adj<T>(conv_[n]->store.view(), 1, gy, rows, gx, s_);
      cast_out(gx, out, int64_t(rows) * N);
    } else {
      T* x = reinterpret_cast<T*>(a);
      T* y = reinterpret_cast<T*>(b);
      cast_in(in, x, int64_t(st[0] / sizeof(T)));
      if (kind == 2) launch_dec_fwd<T>(dec_[n]->store.view(), 1, x, rows, y, s_);
      else launch_dec_adj<T>(dec_[n]->store.view(), 1, x, nullptr, rows, y, s_);
      cast_out(y, out, int64_t(st[1] / sizeof(T)));
    }
    STL_HIP(hipStreamSynchronize(s_));
  }

  std::vector<char> save_state() const override {
    check_holder();
    std::vector<char> out;
    auto put = [&](const void* d, size_t bytes) {
      size_t at = out.size();
      out.resize(at + bytes);
      STL_HIP(hipMemcpy(out.data() + at, d, bytes, hipMemcpyDeviceToHost));
    };
    // The recompute store's state is the maps and moduli (z and the branches
    // are recomputed by the reverse).
    for (int n = 0; n < levels_; ++n) {
      put(f_->maps[n].data(), f_->maps[n].bytes());
      put(f_->M[n].data(), f_->M[n].bytes());
      if (!f_->recompute) {
        put(f_->z[n].data(), f_->z[n].bytes());
        put(f_->br[n].data(), f_->br[n].bytes());
      }
    }
    return out;
  }

  void load_state(const std::vector<char>& s) override {
    f_->holder = id_;
    size_t at = 0;
    auto get = [&](void* d, size_t bytes) {
      STL_CHECK(at + bytes <= s.size(), "checkpoint size mismatch");
      STL_HIP(hipMemcpy(d, s.data() + at, bytes, hipMemcpyHostToDevice));
      at += bytes;
    };
    for (int n = 0; n < levels_; ++n) {
      get(f_->maps[n].data(), f_->maps[n].bytes());
      get(f_->M[n].data(), f_->M[n].bytes());
      if (!f_->recompute) {
        get(f_->z[n].data(), f_->z[n].bytes());
        get(f_->br[n].data(), f_->br[n].bytes());
      }
    }
    STL_CHECK(at == s.size(), "checkpoint size mismatch");
  }

 private:
  void allocate() {
    mC_.assign(levels_, 1);
    mcls_.resize(levels_);
    mpat_.resize(levels_);
    if (f_) {
      STL_CHECK(f_->rows >= B_ && int(f_->maps.size()) == levels_,
                "share: more maps than the shared fields hold (or another pyramid)");
      if (f_->recompute || int(f_->qmg.size()) != levels_) qc_ = levels_;
      else qc_ = f_->qc;
      alloc_sums();
      return;
    }
    f_ = std::make_shared<FieldSet>();
    f_->rows = B_;
    f_->recompute = choose_recompute();
    const size_t live0 = device_bytes_live();
    f_->maps.resize(levels_);
    f_->z.resize(levels_);
    f_->M.resize(levels_);
    f_->br.resize(levels_);
    for (int n = 0; n < levels_; ++n) {
      int64_t N = N_[n];
      f_->maps[n].resize(size_t(B_) * N);
      f_->M[n].resize(size_t(n + 1) * B_ * N * kL);
      if (n == 0 || !f_->recompute) {
        f_->z[n].resize(size_t(B_) * kL * N);
        f_->br[n].resize(size_t(n + 1) * B_ * kL * N * kL);
      }
    }
    // Level 0 has the largest branch, z and modulus fields: (n + 1) / 4^n <= 1.
    f_->gbr.resize(f_->br[0].size());
    f_->gzp.resize(f_->z[0].size());
    f_->gz.resize(f_->z[0].size());
    f_->acc[0].resize(f_->M[0].size());
    f_->mg[0].resize(f_->maps[0].size());
    f_->mg[1].resize(f_->maps[0].size());
    if (f_->recompute) {
      if (levels_ > 1) f_->acc[1].resize(f_->M[0].size());
      f_->bytes = device_bytes_live() - live0;
      alloc_sums();
      return;
    }
    if (levels_ > 1) {
      f_->gbr1.resize(f_->br[1].size());
      f_->gzp1.resize(f_->z[1].size());
      f_->acc[1].resize(f_->M[1].size());
    }
    if (levels_ > 2) {
      size_t g2 = 0;
      for (int n = 2; n < levels_; ++n) g2 = std::max(g2, f_->br[n].size());
      f_->gbr2.resize(g2);
      f_->gzp2.resize(f_->z[2].size());
      f_->acc[2].resize(f_->M[2].size());
      f_->acc[3].resize(f_->M[2].size());
    }
    f_->qc = qc_;
    f_->qgbr.resize(levels_);
    f_->qgzp.resize(levels_);
    f_->qgz.resize(levels_);
    f_->qchain.resize(levels_);
    f_->qmg.resize(levels_);
    for (int n = qc_; n < levels_; ++n) {
      f_->qgbr[n].resize(f_->br[n].size());
      f_->qgzp[n].resize(f_->z[n].size());
      f_->qgz[n].resize(f_->z[n].size());
      f_->qchain[n].resize(f_->M[n].size());
      f_->qmg[n].resize(f_->maps[n].size());
    }
    f_->bytes = device_bytes_live() - live0;
    alloc_sums();
  }

  // Bytes of the keep store's fields and reverse buffers (what allocate() makes).
  size_t keep_field_bytes() const {
    auto len = [&](int n) { return n < levels_ ? size_t(N_[n]) : size_t(0); };
    size_t e = 0, c = 0;  // real and complex elements
    for (int n = 0; n < levels_; ++n) {
      e += size_t(B_) * len(n) * (1 + (n + 1) * kL);
      c += size_t(B_) * len(n) * kL * (1 + (n + 1) * kL);
    }
    size_t br2 = 0;
    for (int n = 2; n < levels_; ++n) br2 = std::max(br2, size_t(n + 1) * B_ * kL * len(n) * kL);
    c += size_t(B_) * kL * kL * len(0) + 2 * size_t(B_) * kL * len(0) + 2 * size_t(B_) * kL * kL * len(1) +
         size_t(B_) * kL * len(1) + br2 + size_t(B_) * kL * len(2);
    e += size_t(B_) * kL * (len(0) + 2 * len(1) + 6 * len(2)) + 2 * size_t(B_) * len(0);
    return e * sizeof(T) + c * sizeof(C2);
  }

  // The field store: forced, or (auto) keep when the kept fields take at most
  // half of the free device memory, as the reference's FieldStore::Auto
  // (crates/stl-gpu/src/engine.rs:574-578).
  bool choose_recompute() const {
    if (opt_.fields == EngineOptions::kKeep) return false;
    if (opt_.fields == EngineOptions::kRecompute) return true;
    size_t free_b = 0, total_b = 0;
    STL_HIP(hipMemGetInfo(&free_b, &total_b));
    return keep_field_bytes() > free_b / 2;
  }

  // Sums, reverse operands and statistics staging: [B][Km] rows of each.
  void alloc_sums() {
    gpart_.resize(levels_);
    gpart2_.resize(levels_);
    gsum_.resize(levels_);
    gsum2_.resize(levels_);
    zpart_.resize(levels_);
    zsum_.resize(levels_);
    dH_.resize(levels_);
    ds1_.resize(levels_);
    ds2_.resize(levels_);
    Hc_.resize(levels_);
    s1c_.resize(levels_);
    s2c_.resize(levels_);
    nblk_gram_.clear();
    nblk_z_.clear();
    const size_t rows = size_t(Bs_) * Km_;
    for (int n = 0; n < levels_; ++n) {
      int64_t N = N_[n];
      int K = gram_k(n), E = cross_ ? K * K : gram_e(n);
      int64_t tiles = (N + gram_tile(n) - 1) / gram_tile(n);
      int nb = cross_ ? int(std::min<int64_t>(tiles, std::max<int64_t>(1, 2048 / (int64_t(rows) * kL))))
                      : gram_blocks<T>(n, N, int64_t(rows));
      nblk_gram_.push_back(nb);
      gpart_[n].resize(size_t(nb) * rows * kL * E * 2);
      gpart2_[n].resize(size_t(nb) * rows * kL * E);
      gsum_[n].resize(rows * kL * E * 2);
      gsum2_[n].resize(rows * kL * E);
      int nz = int(std::min<int64_t>((N + 1023) / 1024, 1024));  // ~4 pixels per thread
      nblk_z_.push_back(nz);
      zpart_[n].resize(size_t(std::max(nz, nb)) * rows * kL * 4);
      zsum_[n].resize(rows * kL * 4);
      dH_[n].resize(rows * kL * K * K);
      ds1_[n].resize(rows * kL * (cross_ ? 2 : 1));  // cross: complex S1 cotangents
      ds2_[n].resize(rows * kL);
      const size_t cls = nmask_ > 0 ? size_t(mC_[n]) : 0;
      Hc_[n].resize(size_t(Bs_) * kL * cls * K * K);
      s1c_[n].resize(size_t(Bs_) * kL * cls * (cross_ ? 2 : 1));
      s2c_[n].resize(size_t(Bs_) * kL * cls);
    }
    nblk_mv_ = int(std::min<int64_t>((N_[0] + 1023) / 1024, 1024));
    mvpart_.resize(size_t(nblk_mv_) * rows * 4);
    mvsum_.resize(rows * 4);
    dgm_.resize(rows);
    dgv_.resize(rows);
    gmc_.resize(nmask_ > 0 ? size_t(Bs_) * mC_[0] : 0);
    gvc_.resize(nmask_ > 0 ? size_t(Bs_) * mC_[0] : 0);
    std::vector<double> nd(size_t(Km_) * levels_);
    for (int k = 0; k < Km_; ++k)
      for (int n = 0; n < levels_; ++n)
        nd[size_t(k) * levels_ + n] = nmask_ > 0 ? double(mcount_[size_t(k) * levels_ + n]) : double(N_[n]);
    upload(dnd_, nd);
    build_tables();
    build_batches();
  }

  MaskView mask_view(int n) const {
    MaskView v;
    if (nmask_ > 0) {
      v.cls = mcls_[n].data();
      v.pat = mpat_[n].data();
      v.K = Km_;
      v.C = mC_[n];
    }
    return v;
  }

  size_t nvals() const { return size_t(Bs_) * Km_ * plan_.size(); }

  void drop_graphs() {
    for (auto* cache : {&fwd_graphs_, &vjp_graphs_}) {
      for (auto& e : *cache)
        for (int i = 0; i < e.count; ++i) (void)hipGraphExecDestroy(e.exec[i]);
      cache->clear();
    }
  }

  void cast_in(const double* x, T* out, int64_t n) {
    if (std::is_same<T, double>::value)
      STL_HIP(hipMemcpyAsync(out, x, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s_));
    else
      launch_cast<float, double>(n, x, reinterpret_cast<float*>(out), s_);
  }

  void cast_out(const T* g, double* out, int64_t n) {
    if (std::is_same<T, double>::value)
      STL_HIP(hipMemcpyAsync(out, g, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice, s_));
    else
      launch_cast<double, float>(n, reinterpret_cast<const float*>(g), out, s_);
  }

  void load_maps(const double* x_dev) { cast_in(x_dev, f_->maps[0].data(), int64_t(B_) * N_[0]); }

  void dec_maps(int n) {
    int id = prof_.start();
    launch_dec_fwd<T>(dec_[n - 1]->store.view(), 1, f_->maps[n - 1].data(), B_, f_->maps[n].data(), s_);
    prof_.stop(id, tag("dec_fwd", n), dec_bytes(n - 1, 1, B_));
  }

  double conv_bytes(int n, int NI, int rows, bool adj) const {
    double slots = adj ? conv_[n]->tslots : conv_[n]->slots;
    double coef = slots * (2 * kL * sizeof(T) + 4);
    double in = adj ? double(rows) * kL * N_[n] * NI * 2 * sizeof(T) : double(rows) * N_[n] * NI * sizeof(T);
    double out = adj ? double(rows) * N_[n] * NI * sizeof(T) : double(rows) * kL * N_[n] * NI * 2 * sizeof(T);
    return coef + in + out;
  }

  double dec_bytes(int n, int NI, int rows, bool adj = false) const {
    double slots = adj ? dec_[n]->tslots : dec_[n]->slots;
    // Forward: the fine rows in, the coarse rows out; adjoint: the coarse
    // cotangent and the fine term it is added to in, the fine rows out.
    const double elems = adj ? double(N_[n + 1]) + 2.0 * double(N_[n]) : double(N_[n]) + double(N_[n + 1]);
    return slots * (sizeof(T) + 4) + double(rows) * elems * NI * sizeof(T);
  }

  void conv_adj(int n, int NI, const C2* gy, int rows, T* gx) {
    int id = prof_.start();
    launch_conv_adj<T>(conv_[n]->store.view(), NI, gy, rows, gx, s_, conv_var_[n][NI == 1 ? 2 : 3]);
    prof_.stop(id, tag(NI == 1 ? "conv_adj_z" : "conv_adj_branch", n), conv_bytes(n, NI, rows, true));
  }

  void dec_adj(int n, int NI, const T* gy, const T* add, int rows, T* gx) {
    int id = prof_.start();
    launch_dec_adj<T>(dec_[n]->store.view(), NI, gy, add, rows, gx, s_);
    prof_.stop(id, tag("dec_adj", n), dec_bytes(n, NI, rows, true));
  }

  // The forward sweep in phases: per level, the z convolution with its
  // modulus and the coarser moduli (fwd_z), then the branch convolution and
  // the sums (fwd_sums); then the statistics.  Levels >= 1 need only level
  // 0's fwd_z, so level 0's fwd_sums may run beside them.
  void sweep_forward(bool moments) {
    fwd_levels(0, moments);
    fwd_finish(moments);
  }

  // The forward of levels from .. J - 1 (maps decimated): level by level, and
  // the batched levels' steps a launch each for all of them.
  void fwd_levels(int from, bool moments) {
    for (int n = from; n < levels_; ++n) {
      if (batch_on() && n == qc_) {
        run_batches(qfwd_[moments ? 1 : 0]);
        return;
      }
      fwd_z(n);
      fwd_sums(n, moments);
    }
  }

  void fwd_z(int n) {
    const int64_t N = N_[n];
    int id = prof_.start();
    launch_conv_fwd<T>(conv_[n]->store.view(), 1, f_->maps[n].data(), B_, z_of(n),
                       f_->M[n].data() + size_t(n) * B_ * N * kL, s_, conv_var_[n][0]);
    prof_.stop(id, tag("conv_fwd_z", n), conv_bytes(n, 1, B_, false) + double(B_) * N * kL * sizeof(T));
    if (n > 0) {
      id = prof_.start();
      launch_dec_fwd<T>(dec_[n - 1]->store.view(), kL, f_->M[n - 1].data(), n * B_, f_->M[n].data(), s_);
      prof_.stop(id, tag("dec_fwd", n), dec_bytes(n - 1, kL, n * B_));
    }
  }

  void fwd_sums(int n, bool moments) {
    const int64_t N = N_[n];
    int id = prof_.start();
    launch_conv_fwd<T>(conv_[n]->store.view(), kL, f_->M[n].data(), (n + 1) * B_, br_of(n), nullptr, s_,
                       conv_var_[n][1]);
    prof_.stop(id, tag("conv_fwd_branch", n), conv_bytes(n, kL, (n + 1) * B_, false));
    id = prof_.start();
    int nb = nblk_gram_[n];
    const MaskView mv = mask_view(n);
    const int64_t rows = int64_t(Bs_) * Km_;
    const int K = gram_k(n);
    int64_t m = rows * kL * (cross_ ? K * K : gram_e(n));
    if (cross_)
      launch_gram_x<T>(n, N, Bs_, br_of(n), z_of(n), moments, nb, gpart_[n].data(), gpart2_[n].data(),
                       mv, s_);
    else
      launch_gram<T>(n, N, B_, br_of(n), z_of(n), moments, nb, gpart_[n].data(), gpart2_[n].data(),
                     zpart_[n].data(), mv, s_);
    launch_reduce_parts(gpart_[n].data(), nb, m * 2, gsum_[n].data(), s_);
    if (moments) launch_reduce_parts(gpart2_[n].data(), nb, m, gsum2_[n].data(), s_);
    prof_.stop(id, tag("gram", n), double(rows) * kL * N * (gram_k(n)) * 2 * sizeof(T));
    id = prof_.start();
    // The S1/S2 sums came with the Gram pass (one partial per Gram block);
    // the cross S1 sums take their own pass.
    if (cross_) launch_zcross<T>(N, Bs_, z_of(n), nblk_z_[n], zpart_[n].data(), mv, s_);
    launch_reduce_parts(zpart_[n].data(), cross_ ? nblk_z_[n] : nb, rows * kL * 4, zsum_[n].data(), s_);
    if (n == 0 && !cross_) {
      launch_meanvar<T>(N, B_, f_->maps[0].data(), nblk_mv_, mvpart_.data(), mv, s_);
      launch_reduce_parts(mvpart_.data(), nblk_mv_, rows * 4, mvsum_.data(), s_);
    }
    // The S1/S2 sums come with the Gram pass: this is the reduction of their
    // partials, the cross S1 pass over z and level 0's mean / variance pass.
    double zbytes = double(cross_ ? nblk_z_[n] : nb) * rows * kL * 4 * 8.0 + double(rows) * kL * 4 * 8.0;
    if (cross_) zbytes += double(B_) * kL * N * 2 * sizeof(T);
    if (n == 0 && !cross_) zbytes += double(B_) * N * sizeof(T) + double(nblk_mv_ + 1) * rows * 4 * 8.0;
    prof_.stop(id, tag("zstats", n), zbytes);
  }

  void fwd_finish(bool moments) {
    const size_t nv = nvals();
    launch_assemble(int(plan_.size()), Bs_, Km_, levels_, dstat_.data(), dsrc_.data(), dnd_.data(), moments,
                    dvalues_.data(), dvar_.data(), s_);
    STL_HIP(hipMemcpyAsync(hvalues_.data(), dvalues_.data(), nv * sizeof(double2), hipMemcpyDeviceToHost, s_));
    if (moments)
      STL_HIP(hipMemcpyAsync(hvar_.data(), dvar_.data(), nv * sizeof(double), hipMemcpyDeviceToHost, s_));
  }

  void enqueue_forward(const double* x_dev, bool moments) {
    load_maps(x_dev);
    for (int n = 1; n < levels_; ++n) dec_maps(n);
    sweep_forward(moments);
  }

  // After the forward's stream work has completed: statistics out of the pinned staging.
  void finish_forward(bool moments, cplx* values, double* variance) override {
    const size_t nv = nvals();
    std::memcpy(values, hvalues_.data(), nv * sizeof(double2));
    if (moments && variance) std::memcpy(variance, hvar_.data(), nv * sizeof(double));
  }

  // Executable graphs of a sweep, captured at the first call with these
  // arguments (kernel arguments are fixed in a graph, so they are keyed by
  // the device pointer and the flag) and replayed afterwards.
  struct GraphEntry {
    const void* ptr;
    bool flag;
    std::array<hipGraphExec_t, 8> exec{};  // one graph, or the steps of a split sweep
    int count = 0;
  };
  static constexpr size_t kMaxGraphs = 8;
  static constexpr int kMaxSteps = 8;

  template <typename Body>
  hipGraph_t capture(Body&& body) {
    hipGraph_t graph = nullptr;
    STL_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      body();
    } catch (...) {
      if (hipStreamEndCapture(s_, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
      throw;
    }
    STL_HIP(hipStreamEndCapture(s_, &graph));
    return graph;
  }

  template <typename Body>
  hipGraphExec_t instantiate(Body&& body) {
    hipGraph_t graph = capture(body);
    hipGraphExec_t exec = nullptr;
    hipError_t err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    STL_HIP(err);
    return exec;
  }

  GraphEntry* cached(std::vector<GraphEntry>& cache, const void* ptr, bool flag) {
    for (auto& e : cache)
      if (e.ptr == ptr && e.flag == flag) return &e;
    return nullptr;
  }

  GraphEntry* remember(std::vector<GraphEntry>& cache, const GraphEntry& entry) {
    if (cache.size() == kMaxGraphs) {
      for (int i = 0; i < cache.front().count; ++i) (void)hipGraphExecDestroy(cache.front().exec[i]);
      cache.erase(cache.begin());
    }
    cache.push_back(entry);
    return &cache.back();
  }

  // A sweep as a list of steps, each a graph of its own (captured at the
  // first call with these arguments): step i runs on stream `on` (0: the
  // engine stream, 1 and 2: side streams) after the earlier steps of that
  // stream and the steps listed in `after`.  With timers the steps run in
  // list order on the engine stream, so the list is a valid sequential order.
  struct Step {
    std::function<void()> body;
    int on;
    std::vector<int> after;
  };

  void run_steps(std::vector<GraphEntry>& cache, const void* ptr, bool flag, const std::vector<Step>& steps) {
    STL_CHECK(!steps.empty() && steps.size() <= size_t(kMaxSteps), "sweep: bad step list");
    if (prof_.on()) {
      for (const auto& st : steps) st.body();
      STL_HIP(hipStreamSynchronize(s_));
      prof_.collect();
        return;
    }
    GraphEntry* e = cached(cache, ptr, flag);
    if (!e) {
      GraphEntry n{ptr, flag};
      try {
        for (const auto& st : steps) n.exec[n.count++] = instantiate(st.body);
      } catch (...) {
        for (int i = 0; i < n.count; ++i) (void)hipGraphExecDestroy(n.exec[i]);
        throw;
      }
      e = remember(cache, n);
    }
    const hipStream_t streams[3] = {s_, s2_, s3_};
    hipError_t err = hipSuccess;
    for (size_t i = 0; i < steps.size() && err == hipSuccess; ++i) {
      const hipStream_t st = streams[steps[i].on];
      for (int d : steps[i].after)
        if (err == hipSuccess) err = hipStreamWaitEvent(st, ev_[d], 0);
      if (err == hipSuccess) err = hipGraphLaunch(e->exec[i], st);
      if (err == hipSuccess) err = hipEventRecord(ev_[i], st);
    }
    // Every stream drained before any error is reported.
    const hipError_t e3 = hipStreamSynchronize(s3_), e2 = hipStreamSynchronize(s2_), e1 = hipStreamSynchronize(s_);
    STL_HIP(err);
    STL_HIP(e3);
    STL_HIP(e2);
    STL_HIP(e1);
  }

  // The forward's steps.  After level n's z convolution (and the moduli it
  // completes), levels > n need nothing more from level n, and level n's
  // branch convolution and sums need nothing from them: levels 0 and 1's
  // sums run on side streams beside the coarser levels.
  std::vector<Step> forward_steps(const double* x_dev, bool moments) {
    // The maps' decimations are not needed by level 0: they follow the fork.
    auto first = [this, x_dev] {
      load_maps(x_dev);
      fwd_z(0);
    };
    auto dec_all = [this] {
      for (int n = 1; n < levels_; ++n) dec_maps(n);
    };
    auto coarser = [this, moments](int from) { return [this, moments, from] { fwd_levels(from, moments); }; };
    auto sums = [this, moments](int n) { return [this, moments, n] { fwd_sums(n, moments); }; };
    auto finish = [this, moments] { fwd_finish(moments); };
    // The recompute store holds one level's z and branches: the levels in turn.
    if (f_->recompute)
      return {{[this, x_dev, moments] { enqueue_forward(x_dev, moments); }, 0, {}}};
    if (levels_ == 1) return {{[first, sums, finish] { first(); sums(0)(); finish(); }, 0, {}}};
    if (levels_ == 2)
      return {{first, 0, {}},
              {sums(0), 1, {0}},
              {[dec_all, coarser] {
                 dec_all();
                 coarser(1)();
               },
               0, {}},
              {finish, 0, {1}}};
    return {{first, 0, {}},
            {sums(0), 1, {0}},
            {[this, dec_all] {
               dec_all();
               fwd_z(1);
             },
             0, {}},
            {sums(1), 2, {2}},
            {coarser(2), 0, {}},
            {finish, 0, {1, 3}}};
  }

  // The reverse's steps.  Level n's gV and branch adjoint read only level-n
  // fields and the prelude's operands, and write level-n buffers: levels 0
  // and 1's run on side streams beside the coarser levels; their chains
  // (which need the coarser chain) follow on the engine stream.
  std::vector<Step> reverse_steps(double* grad_dev) {
    auto coarser = [this](int down_to) { return [this, down_to] { rev_levels(down_to); }; };
    auto last = [this, grad_dev] { rev_chain(0, grad_dev); };
    auto prelude = [this] { rev_prelude(); };
    if (levels_ == 1 || f_->recompute) return {{[this, grad_dev] { sweep_reverse(grad_dev); }, 0, {}}};
    if (levels_ == 2)
      return {{prelude, 0, {}}, {[this] { rev_branch(0); }, 1, {0}}, {coarser(1), 0, {}}, {last, 0, {1}}};
    return {{prelude, 0, {}},        {[this] { rev_branch(0); }, 1, {0}}, {[this] { rev_branch(1); }, 2, {0}},
            {coarser(2), 0, {}},     {[this] { rev_chain(1); }, 0, {2}},  {last, 0, {1}}};
  }

  // Convolution kernel variant per level and kernel (forward z, forward
  // branch, adjoint z, adjoint branch): forced, or the fastest of the
  // candidates timed on this level's operator and buffers (a candidate must
  // beat the default by 3%), remembered for the process.
  void tune() {
    conv_var_.assign(levels_, {kConvTiled, kConvTiled, kConvTiled, kConvTiled});
    for (int n = 0; n < levels_; ++n) {
      ConvOp<T> op = conv_[n]->store.view();
      for (int kind = 0; kind < 4; ++kind) {
        const bool adj = kind >= 2;
        const int NI = kind % 2 == 0 ? 1 : kL, rows = kind % 2 == 0 ? B_ : (n + 1) * B_;
        std::vector<int> cands{kConvTiled};
        if (op.tiled && (!adj || conv_adj_has_variant<T>(NI, kConvFaces))) cands.push_back(kConvFaces);
        if (op.tiled && op.order >= 4) cands.push_back(kConvTiled8);
        if (op.tiled && op.order >= 5) cands.push_back(kConvTiled32);
        if (op.tiled && op.order >= 5 &&
            (adj ? conv_adj_has_variant<T>(NI, kConvFaces32) : conv_fwd_has_variant<T>(NI, kConvFaces32)))
          cands.push_back(kConvFaces32);
        if (op.order <= 6) cands.push_back(kConvSplit);
        if (opt_.conv_variant >= 0) {
          bool ok = std::find(cands.begin(), cands.end(), opt_.conv_variant) != cands.end();
          conv_var_[n][kind] = ok ? opt_.conv_variant : kConvTiled;
          continue;
        }
        const TuneKey key{opt_.device, sizeof(T) == 8, op.order, kind, rows};
        {
          std::lock_guard<std::mutex> lock(tune_mutex());
          auto it = tune_cache().find(key);
          if (it != tune_cache().end()) {
            conv_var_[n][kind] = it->second;
            continue;
          }
        }
        auto timed = [&](int v) {
          if (!adj) {
            T* x = NI == 1 ? f_->maps[n].data() : f_->M[n].data();
            C2* y = NI == 1 ? z_of(n) : br_of(n);
            T* mod = NI == 1 ? f_->M[n].data() + size_t(n) * B_ * N_[n] * kL : nullptr;
            return time_launch([&] { launch_conv_fwd<T>(op, NI, x, rows, y, mod, s_, v); });
          }
          const C2* gy = NI == 1 ? f_->gz.data() : f_->gbr.data();
          T* gx = NI == 1 ? f_->mg[0].data() : f_->acc[0].data();
          return time_launch([&] { launch_conv_adj<T>(op, NI, gy, rows, gx, s_, v); });
        };
        // Two interleaved rounds, each candidate's best: no bias from the
        // order of the candidates or from the first touches of the buffers.
        std::vector<double> ms(cands.size(), 1e30);
        for (int round = 0; round < 2; ++round)
          for (size_t c = 0; c < cands.size(); ++c) ms[c] = std::min(ms[c], timed(cands[c]));
        double best = ms[0];
        int pick = kConvTiled;
        for (size_t c = 1; c < cands.size(); ++c)
          if (ms[c] < 0.97 * ms[0] && ms[c] < best) {
            best = ms[c];
            pick = cands[c];
          }
        conv_var_[n][kind] = pick;
        for (size_t c = 0; c < cands.size(); ++c) tune_ms_.push_back(TuneRecord{n, kind, cands[c], ms[c]});
        std::lock_guard<std::mutex> lock(tune_mutex());
        tune_cache()[key] = pick;
      }
    }
  }

  // Mean duration of 2 launches after one warm-up (HIP events on the engine stream).
  template <typename F>
  double time_launch(F&& launch) {
    hipEvent_t a, b;
    STL_HIP(hipEventCreate(&a));
    STL_HIP(hipEventCreate(&b));
    launch();
    STL_HIP(hipEventRecord(a, s_));
    for (int i = 0; i < 2; ++i) launch();
    STL_HIP(hipEventRecord(b, s_));
    STL_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    STL_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms / 2.0;
  }

  // The batched launches of a sweep, in order, on the engine stream.
  void run_batches(const std::vector<Batch>& batches) {
    for (const Batch& b : batches) {
      int id = prof_.start();
      launch_batch<T>(b.tasks.data(), b.ntasks, b.items, b.lds, q_workgroups_, s_);
      prof_.stop(id, b.name, b.bytes);
    }
  }

 public:
  int batch_level() const override { return batch_on() ? qc_ : levels_; }

 private:
  // The batched launches of the coarse levels (qc_ .. J - 1): every step of
  // the sweep that the levels can take at once is one launch holding that
  // step for all of them; the decimation chains, whose levels depend on each
  // other, take one launch per level.  Forward (without and with moments):
  // z convolutions | moduli decimation, level by level | branch
  // convolutions | Gram sums | their reductions.  Reverse: gV | branch
  // adjoints | their mirror corrections | chain decimation, level by level
  // down | gz | z adjoints | their mirror corrections | map-cotangent
  // decimation, level by level down.  Pointers are this engine's buffers,
  // fixed once allocated.
  void build_batches() {
    if (!batch_on()) return;
    auto base = [&](int op, int n, int ny, int64_t nx) {
      LevelTask<T> t;
      t.op = op;
      t.n = n;
      t.N = N_[n];
      t.B = B_;
      t.ny = ny;
      t.nx = nx;
      return t;
    };
    auto blocks = [](int64_t n, int per) { return (n + per - 1) / per; };
    const int64_t rows = int64_t(Bs_) * Km_;
    struct Step {
      std::vector<LevelTask<T>> tasks;
      double bytes = 0.0;
      std::string name;
    };
    auto finish = [&](std::vector<Batch>& out, std::vector<Step>& steps) {
      out.clear();
      for (auto& st : steps) {
        if (st.tasks.empty()) continue;
        out.emplace_back();
        Batch& b = out.back();
        b.items = batch_layout<T>(st.tasks.data(), int(st.tasks.size()));
        b.lds = batch_lds_bytes<T>(st.tasks.data(), int(st.tasks.size()));
        b.ntasks = int(st.tasks.size());
        b.bytes = st.bytes;
        b.name = st.name;
        upload(b.tasks, st.tasks);
      }
    };
    const std::string at = "@L" + std::to_string(qc_);
    for (int mom = 0; mom < 2; ++mom) {
      std::vector<Step> steps;
      Step zs{{}, 0.0, "batch_conv_fwd_z" + at}, bs{{}, 0.0, "batch_conv_fwd_branch" + at},
          gs{{}, 0.0, "batch_gram" + at}, rs{{}, 0.0, "batch_reduce" + at};
      std::vector<Step> decs;
      for (int n = qc_; n < levels_; ++n) {
        const int64_t N = N_[n];
        const ConvOp<T> conv = conv_[n]->store.view();
        LevelTask<T> z = base(kQConvFwd1, n, B_, blocks(N * 8, 256));
        z.conv = conv;
        z.in0 = f_->maps[n].data();
        z.out0 = f_->z[n].data();
        z.out1 = f_->M[n].data() + size_t(n) * B_ * N * kL;
        zs.tasks.push_back(z);
        zs.bytes += conv_bytes(n, 1, B_, false) + double(B_) * N * kL * sizeof(T);
        const DecOp<T> dec = dec_[n - 1]->store.view();
        LevelTask<T> d = base(kQDecFwd4, n, n * B_, blocks(dec.Rc * 4, 256));
        d.dec = dec;
        d.in0 = f_->M[n - 1].data();
        d.out0 = f_->M[n].data();
        decs.push_back(Step{{d}, dec_bytes(n - 1, kL, n * B_), tag("batch_dec_fwd", n)});
        LevelTask<T> b = base(kQConvFwd4, n, (n + 1) * B_, blocks(N * 8, 256));
        b.conv = conv;
        b.in0 = f_->M[n].data();
        b.out0 = f_->br[n].data();
        bs.tasks.push_back(b);
        bs.bytes += conv_bytes(n, kL, (n + 1) * B_, false);
        const int K = gram_k(n), E = gram_e(n), nb = nblk_gram_[n];
        LevelTask<T> g = base(kQGram, n, int(rows) * kL, nb);
        g.in0 = f_->br[n].data();
        g.in1 = f_->z[n].data();
        g.out0 = gpart_[n].data();
        g.out1 = gpart2_[n].data();
        g.out2 = zpart_[n].data();
        g.tpx = gram_tile(n);
        g.bs = K <= 13 ? 2 : 4;
        g.moments = mom;
        gs.tasks.push_back(g);
        gs.bytes += double(rows) * kL * N * K * 2 * sizeof(T);
        auto reduce = [&](const double* part, int64_t m, double* out) {
          LevelTask<T> r = base(kQReduce, n, 1, blocks(m, 64));
          r.in0 = part;
          r.nblk = nb;
          r.m = m;
          r.out0 = out;
          rs.tasks.push_back(r);
          rs.bytes += (double(nb) + 1.0) * m * 8.0;
        };
        reduce(gpart_[n].data(), rows * kL * E * 2, gsum_[n].data());
        if (mom) reduce(gpart2_[n].data(), rows * kL * E, gsum2_[n].data());
        reduce(zpart_[n].data(), rows * kL * 4, zsum_[n].data());
      }
      steps.push_back(std::move(zs));
      for (auto& d : decs) steps.push_back(std::move(d));
      steps.push_back(std::move(bs));
      steps.push_back(std::move(gs));
      steps.push_back(std::move(rs));
      finish(qfwd_[mom], steps);
    }
    std::vector<Step> steps;
    Step vs{{}, 0.0, "batch_gv" + at}, as{{}, 0.0, "batch_conv_adj_branch" + at}, fs{{}, 0.0, "batch_fix" + at};
    Step zs{{}, 0.0, "batch_gz" + at}, za{{}, 0.0, "batch_conv_adj_z" + at}, zf{{}, 0.0, "batch_fix_z" + at};
    std::vector<Step> chain, mgs;
    for (int n = levels_ - 1; n >= qc_; --n) {
      const int64_t N = N_[n];
      const ConvOp<T> conv = conv_[n]->store.view();
      LevelTask<T> v = base(kQGv, n, B_ * kL, blocks(N, 32));
      v.in0 = f_->br[n].data();
      v.in1 = f_->z[n].data();
      v.in2 = dH_[n].data();
      v.out0 = f_->qgbr[n].data();
      v.out1 = f_->qgzp[n].data();
      vs.tasks.push_back(v);
      vs.bytes += double(B_) * N * ((n + 1) * kL * kL + kL) * 2 * sizeof(T) * 2;
      LevelTask<T> a = base(kQConvAdj4, n, (n + 1) * B_, blocks(N * 8, 256));
      a.conv = conv;
      a.in0 = f_->qgbr[n].data();
      a.out0 = f_->qchain[n].data();
      as.tasks.push_back(a);
      as.bytes += conv_bytes(n, kL, (n + 1) * B_, true);
      if (conv.nfix > 0) {
        LevelTask<T> f = base(kQFix4, n, (n + 1) * B_, blocks(int64_t(conv.nfix) * kL, 256));
        f.conv = conv;
        f.in0 = f_->qgbr[n].data();
        f.out0 = f_->qchain[n].data();
        fs.tasks.push_back(f);
      }
      if (n + 1 < levels_) {
        const DecOp<T> dec = dec_[n]->store.view();
        LevelTask<T> c = base(kQDecAdj4, n, (n + 1) * B_, blocks(dec.Rf * 4, 256));
        c.dec = dec;
        c.in0 = f_->qchain[n + 1].data();
        c.in1 = f_->qchain[n].data();
        c.out0 = f_->qchain[n].data();
        chain.push_back(Step{{c}, dec_bytes(n, kL, (n + 1) * B_, true), tag("batch_dec_adj", n)});
      }
      LevelTask<T> z = base(kQGz, n, B_, blocks(N, 256));
      z.in0 = f_->z[n].data();
      z.in1 = f_->qgzp[n].data();
      z.in2 = f_->qchain[n].data();
      z.s1 = ds1_[n].data();
      z.s2 = ds2_[n].data();
      z.out0 = f_->qgz[n].data();
      zs.tasks.push_back(z);
      zs.bytes += double(B_) * kL * N * (7 * sizeof(T));
      LevelTask<T> za1 = base(kQConvAdj1, n, B_, blocks(N * 8, 256));
      za1.conv = conv;
      za1.in0 = f_->qgz[n].data();
      za1.out0 = f_->qmg[n].data();
      za.tasks.push_back(za1);
      za.bytes += conv_bytes(n, 1, B_, true);
      if (conv.nfix > 0) {
        LevelTask<T> f = base(kQFix1, n, B_, blocks(conv.nfix, 256));
        f.conv = conv;
        f.in0 = f_->qgz[n].data();
        f.out0 = f_->qmg[n].data();
        zf.tasks.push_back(f);
      }
      if (n + 1 < levels_) {
        const DecOp<T> dec = dec_[n]->store.view();
        LevelTask<T> m = base(kQDecAdj1, n, B_, blocks(dec.Rf * 4, 256));
        m.dec = dec;
        m.in0 = f_->qmg[n + 1].data();
        m.in1 = f_->qmg[n].data();
        m.out0 = f_->qmg[n].data();
        mgs.push_back(Step{{m}, dec_bytes(n, 1, B_, true), tag("batch_dec_adj_mg", n)});
      }
    }
    steps.push_back(std::move(vs));
    steps.push_back(std::move(as));
    steps.push_back(std::move(fs));
    for (auto& c : chain) steps.push_back(std::move(c));
    steps.push_back(std::move(zs));
    steps.push_back(std::move(za));
    steps.push_back(std::move(zf));
    for (auto& m : mgs) steps.push_back(std::move(m));
    finish(qrev_, steps);
  }

  // Device tables of the statistics assembly and of the cotangent scatter.
  void build_tables() {
    const int np = int(plan_.size());
    std::vector<const double*> src{mvsum_.data()};
    std::vector<double*> dst{dgm_.data(), dgv_.data()};
    for (int n = 0; n < levels_; ++n) {
      src.push_back(gsum_[n].data());
      src.push_back(gsum2_[n].data());
      src.push_back(zsum_[n].data());
      dst.push_back(reinterpret_cast<double*>(dH_[n].data()));
      dst.push_back(ds1_[n].data());
      dst.push_back(ds2_[n].data());
    }
    std::vector<StatEntry> st(np);
    std::vector<RevTarget> targets;
    std::vector<int> hbase(levels_), s1base(levels_), s2base(levels_);
    targets.push_back(RevTarget{0, 0, 0, 0, 0, 1});
    targets.push_back(RevTarget{1, 0, 0, 0, 0, 1});
    for (int n = 0; n < levels_; ++n) {
      const int K = gram_k(n);
      hbase[n] = int(targets.size());
      for (int e = 0; e < kL * K * K; ++e) targets.push_back(RevTarget{2 + 3 * n, 1, 0, 0, 2 * e, 2 * kL * K * K});
      s1base[n] = int(targets.size());
      for (int l = 0; l < kL; ++l)
        targets.push_back(cross_ ? RevTarget{3 + 3 * n, 1, 0, 0, 2 * l, 2 * kL} : RevTarget{3 + 3 * n, 0, 0, 0, l, kL});
      s2base[n] = int(targets.size());
      for (int l = 0; l < kL; ++l) targets.push_back(RevTarget{4 + 3 * n, 0, 0, 0, l, kL});
    }
    std::vector<std::vector<RevTerm>> lists(targets.size());
    for (int i = 0; i < np; ++i) {
      const SumSlot& sl = slots_[i];
      const int n = sl.n;
      StatEntry& e = st[i];
      e = StatEntry{};
      switch (sl.kind) {
        case SumSlot::MEANVAR: {
          e.arr = 0; e.re = sl.idx; e.im = -1; e.bs = 4;
          e.sq_arr = 0; e.sq = 2 + sl.idx; e.sbs = 4; e.level = 0;
          lists[sl.idx].push_back(RevTerm{i, 0, 0, 0, sl.idx == 0 ? 1.0 : 2.0});
          break;
        }
        case SumSlot::ZSTAT: {
          e.arr = 3 + 3 * n; e.re = sl.l * 4 + sl.idx; e.im = -1; e.bs = kL * 4;
          e.sq_arr = 3 + 3 * n; e.sq = sl.l * 4 + 2 + sl.idx; e.sbs = kL * 4; e.level = n;
          lists[(sl.idx == 0 ? s1base[n] : s2base[n]) + sl.l].push_back(
              RevTerm{i, 0, n, 0, sl.idx == 0 ? 1.0 : 2.0});
          break;
        }
        case SumSlot::ZCROSS: {
          e.arr = 3 + 3 * n; e.re = sl.l * 4; e.im = sl.l * 4 + 1; e.bs = kL * 4;
          e.sq_arr = 3 + 3 * n; e.sq = sl.l * 4 + 2; e.sbs = kL * 4; e.level = n;
          lists[s1base[n] + sl.l].push_back(RevTerm{i, 0, n, 0, 1.0});
          break;
        }
        case SumSlot::GRAMX: {
          const int K = gram_k(n), E = K * K;
          const int64_t el = int64_t(sl.l) * E + sl.idx;
          e.arr = 1 + 3 * n; e.re = 2 * el; e.im = 2 * el + 1; e.bs = int64_t(kL) * E * 2; e.conj = sl.conj;
          e.sq_arr = 2 + 3 * n; e.sq = el; e.sbs = int64_t(kL) * E; e.level = n;
          lists[hbase[n] + (sl.l * K + sl.gi) * K + sl.gj].push_back(RevTerm{i, sl.conj ? 1 : 0, n, 0, 1.0});
          break;
        }
        case SumSlot::GRAM: {
          const int K = gram_k(n), E = gram_e(n);
          const int64_t el = int64_t(sl.l) * E + sl.idx;
          e.arr = 1 + 3 * n; e.re = 2 * el; e.im = 2 * el + 1; e.bs = int64_t(kL) * E * 2; e.conj = sl.conj;
          e.sq_arr = 2 + 3 * n; e.sq = el; e.sbs = int64_t(kL) * E; e.level = n;
          lists[hbase[n] + (sl.l * K + sl.gi) * K + sl.gj].push_back(RevTerm{i, 0, n, 0, 1.0});
          lists[hbase[n] + (sl.l * K + sl.gj) * K + sl.gi].push_back(RevTerm{i, 1, n, 0, 1.0});
          break;
        }
      }
    }
    std::vector<RevTerm> terms;
    for (size_t t = 0; t < targets.size(); ++t) {
      targets[t].first = int32_t(terms.size());
      targets[t].count = int32_t(lists[t].size());
      terms.insert(terms.end(), lists[t].begin(), lists[t].end());
    }
    nrt_ = int(targets.size());
    upload(dstat_, st);
    upload(dsrc_, src);
    upload(drt_, targets);
    upload(drk_, terms);
    upload(ddst_, dst);
    const size_t nv = nvals();
    dvalues_.resize(nv);
    dvar_.resize(nv);
    dcot_.resize(nv);
    hvalues_.resize(nv);
    hvar_.resize(nv);
    hcot_.resize(nv);
  }

  template <typename U>
  void upload(DevArray<U>& d, const std::vector<U>& h) {
    d.resize(std::max<size_t>(h.size(), 1));
    if (!h.empty()) STL_HIP(hipMemcpy(d.data(), h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
  }

  int qc_ = 0;  // first batched level (levels_: none)
  int q_workgroups_ = 1;
  std::vector<Batch> qfwd_[2], qrev_;  // forward without / with moments, reverse

  EngineOptions opt_;
  hipStream_t s_ = nullptr;
  hipStream_t s2_ = nullptr, s3_ = nullptr;  // the split sweeps' side streams
  std::array<hipEvent_t, kMaxSteps> ev_{};     // one per step of a split sweep
  std::vector<std::array<int, 4>> conv_var_;
  std::vector<TuneRecord> tune_ms_;
  std::vector<GraphEntry> fwd_graphs_, vjp_graphs_;
  Profiler prof_;
  std::vector<Coefficient> plan_;
  std::vector<SumSlot> slots_;
  // B_: field rows (maps, or 2 per pair with cross); Bs_: statistics rows (maps / pairs).
  int levels_ = 0, B_ = 1, Bs_ = 1;
  bool cross_ = false;
  std::vector<int64_t> N_;
  std::vector<std::shared_ptr<const SharedConv<T>>> conv_;
  std::vector<std::shared_ptr<const SharedDec<T>>> dec_;
  size_t geometry_bytes_ = 0;
  double setup_s_ = 0.0;
  std::shared_ptr<FieldSet> f_;
  uint64_t id_ = 0;
  // reductions
  std::vector<int> nblk_gram_, nblk_z_;
  int nblk_mv_ = 1;
  std::vector<DevArray<double>> gpart_, gpart2_, gsum_, gsum2_, zpart_, zsum_;
  DevArray<double> mvpart_, mvsum_;
  // reverse
  std::vector<DevArray<double2>> dH_;
  std::vector<DevArray<double>> ds1_, ds2_;
  DevArray<double> dgm_, dgv_;
  // masks: count, rows per map (max(masks, 1)), per-level pixel classes and
  // their membership patterns, pixel counts [Km][levels] (host and device),
  // and the per-class reverse operands
  int nmask_ = 0, Km_ = 1;
  std::vector<int> mC_;
  std::vector<DevArray<uint16_t>> mcls_;
  std::vector<DevArray<uint64_t>> mpat_;
  std::vector<int64_t> mcount_;
  DevArray<double> dnd_;
  std::vector<DevArray<double2>> Hc_;
  std::vector<DevArray<double>> s1c_, s2c_;
  DevArray<double> gmc_, gvc_;
  // device statistics assembly and cotangent scatter
  int nrt_ = 0;
  DevArray<StatEntry> dstat_;
  DevArray<const double*> dsrc_;
  DevArray<RevTarget> drt_;
  DevArray<RevTerm> drk_;
  DevArray<double*> ddst_;
  DevArray<double2> dvalues_, dcot_;
  DevArray<double> dvar_;
  PinnedArray<double2> hvalues_, hcot_;
  PinnedArray<double> hvar_;
};

}  // namespace

std::unique_ptr<Engine> Engine::create(const EngineOptions& opt) {
  if (opt.fp64) return std::unique_ptr<Engine>(new EngineT<double>(opt));
  return std::unique_ptr<Engine>(new EngineT<float>(opt));
}

}  // namespace stl
