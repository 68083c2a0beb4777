// The ScatCov engine on one GPU: the whole sky, every level resident in HBM.
// Mirrors the reference's Evaluator (crates/stl-core/src/engine.rs:254-271):
// forward(x, moments) -> (values, variances), gradient(cotangent) -> map
// cotangent, for the complete one-channel plan (fewer = false, dg(j) = j,
// gauge cosmo, kernel 5, L = 4).
#pragma once

#include <array>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

namespace stl {

typedef std::complex<double> cplx;

enum Family { F_MEAN = 0, F_VAR = 1, F_S1 = 2, F_S2 = 3, F_S3 = 4, F_S4 = 5 };

// One coefficient of the plan (crates/stl-core/src/plan.rs:345-366); -1 = None.
struct Coefficient {
  int family, native_dg, c1, c2;
  int j1 = -1, j2 = -1, j3 = -1, l1 = -1, l2 = -1, l3 = -1;
};

// The complete plan of one channel, sorted by native level (plan.rs:378-456).
std::vector<Coefficient> build_plan(int J, int L);
// The cross terms of two channels (c1 != c2 of the two-channel plan, the
// reference's plan_of(cross = true), crates/stl-bridge/src/lib.rs:238-245):
// per scale S1 and S2 of (0, 1), then S3 and S4 of (0, 1) and of (1, 0).
std::vector<Coefficient> build_cross_plan(int J, int L);

struct EngineOptions {
  int nside = 128;
  int J = 7;
  int batch = 1;
  bool cross = false;    // map pairs [batch][2][12 nside^2] and the cross plan
  std::vector<int> families;  // plan of these families only (empty: all)
  bool fp64 = true;      // false: FP32 bulk fields and operators, FP64 sums
  int device = 0;
  bool timing = false;   // per-kernel HIP-event timers
  // Convolution kernel variant of the LDS-tiled levels: -1 chooses per level
  // and kernel by timing at creation (cached per device, precision, level and
  // shape); 0, 1, 2 force one where the kernel has it (see sweeps.hpp).
  int conv_variant = -1;
  // What the forward keeps for the reverse (the reference's FieldStore,
  // crates/stl-gpu/src/engine.rs:172-186): keep stores every level's z and
  // branches; recompute stores the maps and moduli only and recomputes each
  // level's z and branches in the reverse (two more convolutions per level,
  // one level's z, branch and cotangent buffers instead of all of them);
  // auto keeps when the kept fields take at most half of the free device
  // memory after the operators are built.
  enum Fields { kKeep = 0, kRecompute = 1, kAuto = 2 };
  int fields = kAuto;
  // Levels (from level 2 on) whose pixels over all the maps of the batch
  // number at most 12 batch_nside^2 run batched: each step of a sweep is one
  // launch for all of them (the decimation chains one per level) instead of a
  // launch per kernel and level; 0 switches the batching off.  Used with the
  // keep store, one-channel plans and no masks.
  int batch_nside = 16;
};

// One timed candidate of the convolution-variant choice: level, kernel
// (0 forward z, 1 forward branch, 2 adjoint z, 3 adjoint branch), variant, ms.
struct TuneRecord {
  int level, kernel, variant;
  double ms;
};

// Per-kernel-family timing (HIP events on the engine stream).
struct KernelTimer {
  std::string name;
  double ms = 0.0;
  long launches = 0;
  double bytes = 0.0;  // algorithmic bytes of the timed launches
};

class Engine {
 public:
  virtual ~Engine() = default;
  static std::unique_ptr<Engine> create(const EngineOptions& opt);

  // Another engine on this one's level fields and operators (the reference's
  // GpuEvaluator::share, crates/stl-gpu/src/engine.rs:1909-1916): its own
  // batch (at most this engine's field rows), cross flag and families, no
  // device memory beyond its sums.  Engines sharing fields take turns: a
  // VJP (or state copy) needs the last forward in the fields to be its own.
  // One caller thread for all of them.
  virtual std::unique_ptr<Engine> share(int batch, bool cross, const std::vector<int>& families) = 0;
  virtual const EngineOptions& options() const = 0;
  virtual const std::vector<Coefficient>& plan() const = 0;
  virtual int levels() const = 0;
  // Averages over k masks from the next forward on (the reference's
  // set_masks, crates/stl-core/src/engine.rs:392-404 with Masks::from_weights,
  // mask.rs:43-70): w [k][12 nside^2] host indicators (nonzero: in), level
  // n + 1 holding the parents of level n; k = 0: the whole sky.  Values,
  // variances and cotangents become [B][k][ncoef].  At most 64 masks.
  virtual void set_masks(const double* w, int k) = 0;
  // Masks in use (0: the whole sky) and their pixel counts [k][levels].
  virtual int masks() const = 0;
  virtual std::vector<int64_t> mask_counts() const = 0;
  // Forward on maps [B][12 nside^2] ([B][2][12 nside^2] with cross) in
  // device memory (float64); values and variances ([B][K][ncoef],
  // K = max(masks, 1)) to host memory.
  virtual void forward_device(const double* x_dev, bool moments, cplx* values, double* variance) = 0;
  // Forward on maps given at every level (concatenated, device memory), no decimation of level 0.
  virtual void forward_levels_device(const double* x_dev, cplx* values) = 0;
  // Map cotangent [B][12 nside^2] (device, float64) of Re sum(conj(cot) values) after forward.
  virtual void vjp_device(const cplx* cot, double* grad_dev) = 0;
  // The stream work of forward_device / vjp_device as an uninstantiated HIP
  // graph (the caller owns it), for composition with other engines' graphs:
  // run it, then finish_forward copies the statistics out; stage_cotangent
  // must precede every run of a captured VJP.  Not with per-kernel timing.
  virtual hipGraph_t capture_forward(const double* x_dev, bool moments) = 0;
  virtual void finish_forward(bool moments, cplx* values, double* variance) = 0;
  virtual void stage_cotangent(const cplx* cot) = 0;
  virtual hipGraph_t capture_vjp(double* grad_dev) = 0;
  virtual hipStream_t stream() const = 0;
  virtual size_t device_bytes() const = 0;
  // The field store in use (EngineOptions::kKeep or kRecompute) and the bytes of the level
  // fields and reverse buffers.
  virtual int field_store() const = 0;
  virtual size_t field_bytes() const = 0;
  // First level of the batched launches (levels() when there are none).
  virtual int batch_level() const = 0;
  virtual double setup_seconds() const = 0;
  virtual std::vector<KernelTimer> timers() const = 0;
  // Chosen convolution variant per level: [level][forward z, forward branch,
  // adjoint z, adjoint branch].
  virtual std::vector<std::array<int, 4>> conv_variants() const = 0;
  // Candidate timings of the variant choice made by this engine (empty when
  // forced or taken from an earlier engine of the process).
  virtual std::vector<TuneRecord> tuning() const = 0;
  virtual void reset_timers() = 0;
  // One level's operator applied to float64 device arrays (kind 0 conv forward
  // [rows][N] -> [rows][4][N] complex, 1 its adjoint, 2 decimation to level + 1,
  // 3 its adjoint); for operator-level tests.
  virtual void apply_operator(int kind, int level, int rows, const double* in_dev, double* out_dev) = 0;
  // Device bytes of staging apply_operator needs for (kind, level, rows), and the same
  // operator application with that staging carved from a caller's device workspace
  // (256-byte aligned, at least that many bytes): no allocation.
  virtual size_t operator_workspace_bytes(int kind, int level, int rows) const = 0;
  virtual void apply_operator_in(int kind, int level, int rows, const double* in_dev, double* out_dev,
                                 void* workspace, size_t workspace_bytes) = 0;
  // Host copy of the forward state (maps, z, moduli, branches) and its restore.
  virtual std::vector<char> save_state() const = 0;
  virtual void load_state(const std::vector<char>& s) = 0;
};

}  // namespace stl
