// C ABI (include/stl_bridge.h): the reference's stlp_* surface
// (crates/stl-bridge/src/lib.rs:257-741) over the MI355X engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/stl_bridge.h"
#include "device_buffer.hpp"
#include "engine.hpp"
#include "ensemble.hpp"
#include "lbfgs.hpp"
#include "mapio.hpp"
#include "vectors.hpp"

namespace {

// Version of this C ABI and its JSON (stlp_info).
const char* kVersion = "mi355x-2026-12-r1";

thread_local std::string g_error;
std::atomic<uint64_t> g_next_id{1};

template <typename R, typename F>
R guard(R fail, F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_error = e.what();
  } catch (...) {
    g_error = "native failure";
  }
  return fail;
}

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

// Flat JSON objects: string, number, bool and null values.
struct Json {
  std::map<std::string, std::string> kv;  // raw scalar text (strings unquoted)
  std::map<std::string, bool> is_string;

  static Json parse(const char* text) {
    if (!text) fail("null JSON input");
    Json j;
    const char* p = text;
    auto ws = [&] { while (*p && std::isspace(static_cast<unsigned char>(*p))) ++p; };
    auto str = [&]() -> std::string {
      if (*p != '"') fail("valid JSON: expected a string");
      ++p;
      std::string s;
      while (*p && *p != '"') {
        if (*p == '\\' && p[1]) { s += p[1]; p += 2; }
        else s += *p++;
      }
      if (*p != '"') fail("valid JSON: unterminated string");
      ++p;
      return s;
    };
    ws();
    if (*p != '{') fail("valid JSON: expected an object");
    ++p;
    ws();
    if (*p == '}') return j;
    while (true) {
      ws();
      std::string key = str();
      ws();
      if (*p != ':') fail("valid JSON: expected ':'");
      ++p;
      ws();
      if (*p == '"') {
        j.kv[key] = str();
        j.is_string[key] = true;
      } else {
        const char* s = p;
        while (*p && *p != ',' && *p != '}' && !std::isspace(static_cast<unsigned char>(*p))) ++p;
        j.kv[key] = std::string(s, p);
        j.is_string[key] = false;
      }
      ws();
      if (*p == ',') { ++p; continue; }
      if (*p == '}') break;
      fail("valid JSON: expected ',' or '}'");
    }
    return j;
  }

  bool has(const std::string& k) const { return kv.count(k) && kv.at(k) != "null"; }
  long integer(const std::string& k, long def) const {
    if (!has(k)) return def;
    char* end = nullptr;
    long v = std::strtol(kv.at(k).c_str(), &end, 10);
    if (*end) fail(k + " must be an integer");
    return v;
  }
  double real(const std::string& k, double def) const {
    if (!has(k)) return def;
    char* end = nullptr;
    double v = std::strtod(kv.at(k).c_str(), &end);
    if (*end) fail(k + " must be a number");
    return v;
  }
  bool boolean(const std::string& k, bool def) const {
    if (!has(k)) return def;
    if (kv.at(k) == "true") return true;
    if (kv.at(k) == "false") return false;
    fail(k + " must be a boolean");
  }
  std::string string(const std::string& k, const std::string& def) const {
    if (!has(k)) return def;
    if (!is_string.at(k)) fail(k + " must be a string");
    return kv.at(k);
  }
};

struct Identity {
  uint64_t owner, generation, member, objective;
  bool operator==(const Identity& o) const {
    return owner == o.owner && generation == o.generation && member == o.member && objective == o.objective;
  }
};

const char* family_name(int f) {
  static const char* names[] = {"mean", "var", "S1", "S2", "S3", "S4"};
  return names[f];
}

std::string json_opt(int v) { return v < 0 ? "null" : std::to_string(v); }

}  // namespace

// Static device scratch of the allocation-free entry point
// (stlx_apply_operator_noalloc): the staging of one operator application, at
// most kScratchBytes.  Each device holds its own copy from the library's load
// until the process ends; host code uses only the per-device address from
// hipGetSymbolAddress (scratch_address), never the symbol itself.  Calls that
// overlap on one device share it and race.
namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;
}
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// Address of `scratch` on a device, looked up once per device with that device
// current (at handle creation, the library's set-up of a device).
void* scratch_address(int device) {
  static std::mutex lock;
  static std::map<int, void*> addresses;
  std::lock_guard<std::mutex> guard_lock(lock);
  auto it = addresses.find(device);
  if (it != addresses.end()) return it->second;
  int current = 0;
  STL_HIP(hipGetDevice(&current));
  STL_HIP(hipSetDevice(device));
  void* ptr = nullptr;
  const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  STL_HIP(hipSetDevice(current));
  if (err != hipSuccess || ptr == nullptr)
    throw std::runtime_error(std::string("scratch address: ") + hipGetErrorString(err));
  addresses.emplace(device, ptr);
  return ptr;
}
}  // namespace

struct StlHandle {
  std::unique_ptr<stl::Engine> engine;
  uint64_t id = 0;
  int nside = 0, batch = 1, J = 0, channels = 1;
  size_t npix = 0;
  std::optional<Identity> current;
  stl::DevArray<double> xbuf, gbuf;
  std::string precision;
  size_t ncoef() const { return engine->plan().size(); }
  // Averages per map: one per mask, 1 without masks.
  size_t nmask_rows() const { return size_t(std::max(engine->masks(), 1)); }
  // Values / cotangents: [batch][K][ncoef].
  size_t nvals() const { return size_t(batch) * nmask_rows() * ncoef(); }
  // Maps [batch][channels][12 nside^2] (channels 2: the pairs of the cross plan).
  size_t input_len() const { return size_t(batch) * channels * npix; }
  Identity identity(uint64_t g, uint64_t m, uint64_t o) const { return Identity{id, g, m, o}; }
};

struct StlSnapshot {
  Identity identity;
  std::vector<char> state;
};

struct StlEnsemble {
  std::unique_ptr<stl::Ensemble> ensemble;
  bool forwarded = false;
};

// "families": comma-separated family names (mean, var, S1 .. S4); absent: all.
static std::vector<int> parse_families(const Json& v) {
  std::vector<int> out;
  if (!v.has("families")) return out;
  std::string text = v.string("families", "");
  size_t at = 0;
  while (at <= text.size()) {
    size_t comma = text.find(',', at);
    std::string name = text.substr(at, comma == std::string::npos ? std::string::npos : comma - at);
    name.erase(0, name.find_first_not_of(' '));
    name.erase(name.find_last_not_of(' ') + 1);
    int f = -1;
    for (int k = 0; k < 6; ++k)
      if (name == family_name(k)) f = k;
    if (f < 0) fail("unknown family \"" + name + "\"");
    out.push_back(f);
    if (comma == std::string::npos) break;
    at = comma + 1;
  }
  return out;
}

extern "C" {

const char* stlp_last_error(void) { return g_error.c_str(); }

void stlp_string_free(char* s) { std::free(s); }

StlHandle* stlp_new(const char* config) {
  return guard<StlHandle*>(nullptr, [&]() -> StlHandle* {
    Json v = Json::parse(config);
    long nside = v.integer("nside", 128);
    if (nside <= 0 || (nside & (nside - 1))) fail("nside must be a power of two");
    if (nside > stl::kMaxNside) fail("nside must be at most 8192 (32-bit pixel ids)");
    long order = 0;
    while ((1L << order) < nside) ++order;
    long j = v.integer("j", 7);
    if (j <= 0 || j > order + 1) fail("invalid number of scales");
    long batch = v.integer("batch", 1);
    if (batch <= 0) fail("batch must be positive");
    if (v.integer("kernel", 5) != 5) fail("kernel: only 5 is implemented");
    if (v.integer("orientations", 4) != 4) fail("orientations: only 4 is implemented");
    if (v.integer("block_pixels", 4096) <= 0) fail("block_pixels must be positive");
    std::string backend = v.string("backend", "gpu");
    if (backend == "cpu") fail("backend cpu is not provided by this library (GPU engine only)");
    if (backend != "gpu") fail("backend must be cpu or gpu, not \"" + backend + "\"");
    if (v.integer("spin", 0) != 0) fail("spin 2 is not implemented");
    std::string precision = v.string("precision", "fp64");
    if (precision != "fp64" && precision != "mixed") fail("precision must be fp64 or mixed");
    stl::EngineOptions opt;
    opt.nside = int(nside);
    opt.J = int(j);
    opt.batch = int(batch);
    opt.cross = v.boolean("cross", false);
    opt.families = parse_families(v);
    opt.fp64 = precision == "fp64";
    opt.device = int(v.integer("device", 0));
    opt.timing = v.boolean("timing", false);
    // The reference's field store key (lib.rs:277, :340-341): auto | keep | recompute.
    const std::string fields = v.string("fields", "auto");
    if (fields == "keep") opt.fields = stl::EngineOptions::kKeep;
    else if (fields == "recompute") opt.fields = stl::EngineOptions::kRecompute;
    else if (fields == "auto") opt.fields = stl::EngineOptions::kAuto;
    else fail("fields: auto, keep or recompute");
    opt.conv_variant = int(v.integer("conv_variant", -1));
    opt.batch_nside = int(v.integer("batch_nside", 16));
    if (opt.batch_nside < 0) fail("batch_nside must be nonnegative");
    if (opt.conv_variant < -1 || opt.conv_variant > 5) fail("conv_variant must be -1 or 0 to 5");
    auto h = std::make_unique<StlHandle>();
    h->engine = stl::Engine::create(opt);
    (void)scratch_address(opt.device);
    h->id = g_next_id.fetch_add(1);
    h->nside = int(nside);
    h->batch = int(batch);
    h->channels = opt.cross ? 2 : 1;
    h->J = int(j);
    h->npix = size_t(12) * nside * nside;
    h->precision = precision;
    return h.release();
  });
}

void stlp_free(StlHandle* h) { delete h; }

StlHandle* stlx_share(StlHandle* h, const char* config) {
  return guard<StlHandle*>(nullptr, [&]() -> StlHandle* {
    if (!h) fail("null handle");
    Json v = Json::parse(config ? config : "{}");
    const long batch = v.integer("batch", h->batch);
    if (batch <= 0) fail("batch must be positive");
    const bool cross = v.boolean("cross", h->channels == 2);
    auto s = std::make_unique<StlHandle>();
    s->engine = h->engine->share(int(batch), cross, parse_families(v));
    s->id = g_next_id.fetch_add(1);
    s->nside = h->nside;
    s->batch = int(batch);
    s->channels = cross ? 2 : 1;
    s->J = h->J;
    s->npix = h->npix;
    s->precision = h->precision;
    return s.release();
  });
}

char* stlp_info(StlHandle* h) {
  return guard<char*>(nullptr, [&]() -> char* {
    if (!h) fail("null handle");
    std::ostringstream o;
    o.precision(17);
    o << "{\"version\": \"" << kVersion << "\", \"nside\": " << h->nside << ", \"j\": " << h->J
      << ", \"batch\": " << h->batch << ", \"channels\": " << h->channels
      << ", \"cross\": " << (h->channels == 2 ? "true" : "false") << ", \"fewer\": false, \"backend\": \"gpu\""
      << ", \"precision\": \"" << h->precision << "\", \"dtype\": \""
      << (h->precision == "fp64" ? "float64/complex128" : "float32 fields and operators, float64 sums") << "\""
      << ", \"masks\": " << h->engine->masks() << ", \"spin\": 0, \"conv\": \"folded gather, 4-fold face symmetry\", \"fields\": \""
      << (h->engine->field_store() == stl::EngineOptions::kRecompute ? "recompute" : "keep") << "\""
      << ", \"batch_level\": " << h->engine->batch_level();
    o << ", \"device_memory\": {\"total_bytes\": " << h->engine->device_bytes()
      << ", \"peak_bytes\": " << stl::device_bytes_peak() << ", \"fields_bytes\": " << h->engine->field_bytes()
      << "}";
    o << ", \"setup_seconds\": " << h->engine->setup_seconds();
    o << ", \"conv_variants\": [";
    {
      const auto cv = h->engine->conv_variants();
      for (size_t n = 0; n < cv.size(); ++n)
        o << (n ? ", " : "") << "[" << cv[n][0] << ", " << cv[n][1] << ", " << cv[n][2] << ", " << cv[n][3] << "]";
    }
    o << "], \"conv_tuning\": [";
    {
      const auto tr = h->engine->tuning();
      for (size_t i = 0; i < tr.size(); ++i)
        o << (i ? ", " : "") << "[" << tr[i].level << ", " << tr[i].kernel << ", " << tr[i].variant << ", "
          << tr[i].ms << "]";
    }
    o << "]";
    o << ", \"timers\": {";
    bool first = true;
    for (auto& t : h->engine->timers()) {
      o << (first ? "" : ", ") << "\"" << t.name << "\": {\"ms\": " << t.ms << ", \"launches\": " << t.launches
        << ", \"bytes\": " << t.bytes << "}";
      first = false;
    }
    o << "}, \"plan\": [";
    const auto& plan = h->engine->plan();
    for (size_t i = 0; i < plan.size(); ++i) {
      const auto& c = plan[i];
      o << (i ? ", " : "") << "{\"index\": " << i << ", \"family\": \"" << family_name(c.family)
        << "\", \"native_dg\": " << c.native_dg << ", \"c1\": " << c.c1 << ", \"c2\": " << c.c2
        << ", \"j1\": " << json_opt(c.j1) << ", \"j2\": " << json_opt(c.j2) << ", \"j3\": " << json_opt(c.j3)
        << ", \"l1\": " << json_opt(c.l1) << ", \"l2\": " << json_opt(c.l2) << ", \"l3\": " << json_opt(c.l3)
        << "}";
    }
    o << "]}";
    std::string s = o.str();
    char* out = static_cast<char*>(std::malloc(s.size() + 1));
    std::memcpy(out, s.c_str(), s.size() + 1);
    return out;
  });
}

static void write_values(const std::vector<stl::cplx>& v, double* out) {
  for (size_t i = 0; i < v.size(); ++i) {
    out[2 * i] = v[i].real();
    out[2 * i + 1] = v[i].imag();
  }
}

static int forward_impl(StlHandle* h, const double* x_dev, size_t n, uint64_t g, uint64_t m, uint64_t o,
                        bool moments, double* out, double* var) {
  if (!h) fail("null handle");
  if (n != h->input_len()) fail("input shape mismatch");
  if (!out) fail("null output");
  if (moments && !var) fail("null variance output");
  h->current.reset();
  std::vector<stl::cplx> values(h->nvals());
  std::vector<double> variance(moments ? values.size() : 0);
  h->engine->forward_device(x_dev, moments, values.data(), moments ? variance.data() : nullptr);
  write_values(values, out);
  if (moments) std::memcpy(var, variance.data(), variance.size() * sizeof(double));
  h->current = h->identity(g, m, o);
  return 0;
}

int stlp_forward(StlHandle* h, const double* x, size_t n, uint64_t g, uint64_t m, uint64_t o, bool moments,
                 double* out, double* var) {
  return guard(-1, [&] {
    if (!h) fail("null handle");
    if (n != h->input_len()) fail("input shape mismatch");
    if (!x) fail("null input");
    h->xbuf.resize(n);
    STL_HIP(hipMemcpyAsync(h->xbuf.data(), x, n * sizeof(double), hipMemcpyHostToDevice, h->engine->stream()));
    return forward_impl(h, h->xbuf.data(), n, g, m, o, moments, out, var);
  });
}

int stlp_forward_device(StlHandle* h, uint64_t x_dev, size_t n, uint64_t g, uint64_t m, uint64_t o, bool moments,
                        double* out, double* var) {
  return guard(-1, [&] {
    if (!x_dev) fail("null device input");
    return forward_impl(h, reinterpret_cast<const double*>(x_dev), n, g, m, o, moments, out, var);
  });
}

static std::vector<stl::cplx> read_cot(StlHandle* h, const double* cot, size_t n, uint64_t g, uint64_t m,
                                       uint64_t o) {
  if (!h) fail("null handle");
  if (!h->current || !(*h->current == h->identity(g, m, o)))
    fail("stale, wrong-member or wrong-objective forward state");
  if (n != h->nvals()) fail("cotangent shape mismatch");
  if (!cot) fail("null cotangent");
  std::vector<stl::cplx> c(n);
  for (size_t i = 0; i < n; ++i) c[i] = stl::cplx(cot[2 * i], cot[2 * i + 1]);
  return c;
}

int stlp_vjp(StlHandle* h, const double* cot, size_t n, uint64_t g, uint64_t m, uint64_t o, double* out) {
  return guard(-1, [&] {
    auto c = read_cot(h, cot, n, g, m, o);
    if (!out) fail("null output");
    h->gbuf.resize(h->input_len());
    h->engine->vjp_device(c.data(), h->gbuf.data());
    STL_HIP(hipMemcpy(out, h->gbuf.data(), h->gbuf.bytes(), hipMemcpyDeviceToHost));
    return 0;
  });
}

int stlp_vjp_device(StlHandle* h, const double* cot, size_t n, uint64_t g, uint64_t m, uint64_t o,
                    uint64_t out_dev) {
  return guard(-1, [&] {
    auto c = read_cot(h, cot, n, g, m, o);
    if (!out_dev) fail("null device output");
    h->engine->vjp_device(c.data(), reinterpret_cast<double*>(out_dev));
    return 0;
  });
}

int stlp_native_pixels(StlHandle* h, size_t* out) {
  return guard(-1, [&] {
    if (!h || !out) fail("null argument");
    // [K][ncoef]: the pixels of mask k at the coefficient's level (lib.rs:456-472).
    const auto& plan = h->engine->plan();
    const int K = h->engine->masks(), levels = h->engine->levels();
    const auto counts = h->engine->mask_counts();
    for (int k = 0; k < std::max(K, 1); ++k)
      for (size_t i = 0; i < plan.size(); ++i) {
        const int d = plan[i].native_dg;
        out[size_t(k) * plan.size() + i] =
            K > 0 ? size_t(counts[size_t(k) * levels + d])
                  : size_t(12) * (size_t(h->nside) >> d) * (size_t(h->nside) >> d);
      }
    return 0;
  });
}

int stlp_forward_levels(StlHandle* h, const double* x, size_t n, double* out) {
  return guard(-1, [&] {
    if (!h || !x || !out) fail("null argument");
    size_t total = 0;
    for (int d = 0; d < h->engine->levels(); ++d)
      total += size_t(h->batch) * h->channels * 12 * (size_t(h->nside) >> d) * (size_t(h->nside) >> d);
    if (n != total) fail("input shape mismatch");
    h->current.reset();
    stl::DevArray<double> buf(n);
    STL_HIP(hipMemcpy(buf.data(), x, n * sizeof(double), hipMemcpyHostToDevice));
    std::vector<stl::cplx> values(h->nvals());
    h->engine->forward_levels_device(buf.data(), values.data());
    write_values(values, out);
    return 0;
  });
}

StlSnapshot* stlp_checkpoint(StlHandle* h) {
  return guard<StlSnapshot*>(nullptr, [&]() -> StlSnapshot* {
    if (!h) fail("null handle");
    if (!h->current) fail("checkpoint requires a forward state");
    auto s = std::make_unique<StlSnapshot>();
    s->identity = *h->current;
    s->state = h->engine->save_state();
    return s.release();
  });
}

void stlp_checkpoint_free(StlSnapshot* s) { delete s; }

size_t stlp_checkpoint_bytes(const StlSnapshot* s) { return s ? s->state.size() : 0; }

int stlp_restore(StlHandle* h, const StlSnapshot* s, uint64_t g, uint64_t m, uint64_t o) {
  return guard(-1, [&] {
    if (!h || !s) fail("null argument");
    if (!(s->identity == h->identity(g, m, o)))
      fail("stale, wrong-member, wrong-objective or foreign checkpoint");
    h->engine->load_state(s->state);
    h->current = s->identity;
    return 0;
  });
}

int stlp_set_masks(StlHandle* h, const double* weights, size_t k) {
  return guard(-1, [&] {
    if (!h) fail("null handle");
    h->current.reset();
    if (k > 0 && !weights) fail("null mask weights");
    if (k > 64) fail("at most 64 masks");
    h->engine->set_masks(weights, int(k));
    return 0;
  });
}

int stlp_spin_fields(StlHandle*, const double*, size_t, size_t, double*, double*) {
  g_error = "not implemented: spin 2";
  return -1;
}

static stl::LbfgsOptions lbfgs_options(const Json& v) {
  stl::LbfgsOptions opt;
  opt.lr = v.real("lr", 1.0);
  opt.max_iter = int(v.integer("max_iter", 20));
  opt.max_eval = v.has("max_eval") ? int(v.integer("max_eval", 0)) : -1;
  opt.tolerance_grad = v.real("tolerance_grad", 1e-12);
  opt.tolerance_change = v.real("tolerance_change", 1e-15);
  opt.history_size = int(v.integer("history_size", 10));
  if (!(opt.history_size > 0 && opt.lr > 0.0 && std::isfinite(opt.lr) && opt.max_iter >= 0))
    fail("invalid L-BFGS options");
  return opt;
}

int64_t stlp_lbfgs(double* x, size_t n, const char* options, stlp_callback callback, void* context) {
  return guard<int64_t>(-1, [&]() -> int64_t {
    Json v = Json::parse(options);
    stl::LbfgsOptions opt = lbfgs_options(v);
    if (!x || !callback) fail("null argument");
    stl::HostBackend be;
    std::vector<double> xv(x, x + n);
    auto f = [&](const std::vector<double>& p, std::vector<double>& g) {
      g.assign(n, 0.0);
      double loss = callback(p.data(), n, g.data(), context);
      bool ok = std::isfinite(loss);
      for (size_t i = 0; ok && i < n; ++i) ok = std::isfinite(g[i]);
      if (!ok) fail("non-finite objective or gradient (or Python callback exception)");
      return loss;
    };
    stl::Lbfgs<stl::HostBackend> lb(be, opt);
    size_t evals = lb.minimize(xv, f);
    std::copy(xv.begin(), xv.end(), x);
    return int64_t(evals);
  });
}

int64_t stlx_synthesize(StlHandle* h, uint64_t x_dev, size_t n, const double* target, const double* weights,
                        const char* options, double* loss_history, size_t cap) {
  return guard<int64_t>(-1, [&]() -> int64_t {
    if (!h || !x_dev || !target) fail("null argument");
    if (n != h->input_len()) fail("input shape mismatch");
    Json v = Json::parse(options);
    stl::LbfgsOptions opt = lbfgs_options(v);
    const bool ensemble = v.boolean("ensemble", false);
    // "history": "evaluations" (the loss of every evaluation) or "iterations"
    // (the start's loss, then the loss at each iteration's accepted step).
    const std::string record = v.string("history", "evaluations");
    if (record != "evaluations" && record != "iterations") fail("history: evaluations or iterations");
    const bool per_iteration = record == "iterations";
    if (loss_history && per_iteration)
      for (size_t i = 0; i < cap; ++i) loss_history[i] = std::nan("");
    h->current.reset();
    const size_t nc = h->nmask_rows() * h->ncoef(), nv = size_t(h->batch) * nc;
    std::vector<stl::cplx> values(nv), cot(nv);
    size_t evals = 0;
    stl::VectorOps ops(int64_t(n), h->engine->stream());
    {
      stl::VectorOps::Vec x = ops.make();
      STL_HIP(hipMemcpyAsync(x.p, reinterpret_cast<const double*>(x_dev), n * sizeof(double),
                             hipMemcpyDeviceToDevice, h->engine->stream()));
      // The engine replays graphs keyed by device pointer: evaluate at the
      // handle's fixed staging buffers, not at the pooled vectors.
      hipStream_t st = h->engine->stream();
      h->xbuf.resize(n);
      h->gbuf.resize(n);
      auto f = [&](const stl::VectorOps::Vec& p, stl::VectorOps::Vec& g) {
        STL_HIP(hipMemcpyAsync(h->xbuf.data(), p.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        h->engine->forward_device(h->xbuf.data(), false, values.data(), nullptr);
        double loss = 0.0;
        if (ensemble) {
          // Loss on the mean over the members; each member gets 2 w r / B
          // (synthesis.rs:376-470 without the scatter term).
          const double b = double(h->batch);
          for (size_t i = 0; i < nc; ++i) {
            stl::cplx mean = 0.0;
            for (int m = 0; m < h->batch; ++m) mean += values[size_t(m) * nc + i];
            const double w = weights ? weights[i] : 1.0;
            const stl::cplx r = mean / b - stl::cplx(target[2 * i], target[2 * i + 1]);
            loss += w * std::norm(r);
            for (int m = 0; m < h->batch; ++m) cot[size_t(m) * nc + i] = (2.0 * w / b) * r;
          }
        } else {
          for (size_t k = 0; k < nv; ++k) {
            double w = weights ? weights[k % nc] : 1.0;
            stl::cplx r = values[k] - stl::cplx(target[2 * k], target[2 * k + 1]);
            double a = std::abs(r);
            loss += w * (a * a);
            cot[k] = (2.0 * w) * r;
          }
        }
        if (!std::isfinite(loss)) fail("non-finite loss");
        h->engine->vjp_device(cot.data(), h->gbuf.data());
        STL_HIP(hipMemcpyAsync(g.p, h->gbuf.data(), n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (loss_history && !per_iteration && evals < cap) loss_history[evals] = loss;
        ++evals;
        return loss;
      };
      stl::Lbfgs<stl::VectorOps> lb(ops, opt);
      if (loss_history && per_iteration)
        lb.on_iteration = [&](int it, double loss) {
          if (size_t(it) < cap) loss_history[it] = loss;
        };
      lb.minimize(x, f);
      STL_HIP(hipMemcpyAsync(reinterpret_cast<double*>(x_dev), x.p, n * sizeof(double), hipMemcpyDeviceToDevice,
                             h->engine->stream()));
      STL_HIP(hipStreamSynchronize(h->engine->stream()));
    }
    return int64_t(evals);
  });
}

StlEnsemble* stlx_ensemble_new(const int32_t* nside, const int32_t* j, const int32_t* batch, size_t groups,
                               const char* options) {
  return guard<StlEnsemble*>(nullptr, [&]() -> StlEnsemble* {
    if (!nside || !j || !batch || groups == 0) fail("an ensemble needs at least one group");
    Json v = Json::parse(options ? options : "{}");
    std::string precision = v.string("precision", "mixed");
    if (precision != "fp64" && precision != "mixed") fail("precision must be fp64 or mixed");
    int64_t batch_nside = v.integer("batch_nside", 16);
    if (batch_nside < 0 || batch_nside > (int64_t(1) << 20)) fail("batch_nside must be nonnegative");
    std::vector<stl::EnsembleGroup> gs(groups);
    for (size_t g = 0; g < groups; ++g) {
      if (nside[g] < 1 || (nside[g] & (nside[g] - 1))) fail("nside must be a power of two");
      if (nside[g] > stl::kMaxNside) fail("nside must be at most 8192 (32-bit pixel ids)");
      int order = 0;
      while ((1 << order) < nside[g]) ++order;
      if (j[g] < 1 || j[g] > order + 1) fail("j must satisfy 1 <= j <= log2(nside) + 1");
      if (batch[g] < 1) fail("batch must be positive");
      gs[g] = stl::EnsembleGroup{nside[g], j[g], batch[g]};
    }
    auto e = std::make_unique<StlEnsemble>();
    e->ensemble = std::make_unique<stl::Ensemble>(gs, precision == "fp64", int(v.integer("device", 0)),
                                                  int(batch_nside));
    return e.release();
  });
}

void stlx_ensemble_free(StlEnsemble* e) { delete e; }

int stlx_ensemble_ncoef(StlEnsemble* e, size_t* out) {
  return guard<int>(-1, [&] {
    if (!e || !out) fail("null argument");
    for (size_t g = 0; g < e->ensemble->groups(); ++g) out[g] = e->ensemble->engine(g).plan().size();
    return 0;
  });
}

int stlx_ensemble_forward_device(StlEnsemble* e, const uint64_t* x_dev, size_t groups, bool moments, double* out,
                                 double* var) {
  return guard<int>(-1, [&] {
    if (!e || !x_dev || !out || (moments && !var)) fail("null argument");
    if (groups != e->ensemble->groups()) fail("one device map array per group expected");
    std::vector<const double*> x(groups);
    for (size_t g = 0; g < groups; ++g) {
      if (!x_dev[g]) fail("null device map");
      x[g] = reinterpret_cast<const double*>(x_dev[g]);
    }
    e->forwarded = false;
    e->ensemble->forward(x.data(), moments, reinterpret_cast<stl::cplx*>(out), moments ? var : nullptr);
    e->forwarded = true;
    return 0;
  });
}

int stlx_ensemble_vjp_device(StlEnsemble* e, const double* cot, size_t n, const uint64_t* grad_dev, size_t groups) {
  return guard<int>(-1, [&] {
    if (!e || !cot || !grad_dev) fail("null argument");
    if (!e->forwarded) fail("no forward state: call stlx_ensemble_forward_device first");
    if (n != e->ensemble->values()) fail("cotangent shape mismatch");
    if (groups != e->ensemble->groups()) fail("one device gradient array per group expected");
    std::vector<double*> g(groups);
    for (size_t k = 0; k < groups; ++k) {
      if (!grad_dev[k]) fail("null device gradient");
      g[k] = reinterpret_cast<double*>(grad_dev[k]);
    }
    e->ensemble->vjp(reinterpret_cast<const stl::cplx*>(cot), g.data());
    return 0;
  });
}

int stlx_apply_operator_noalloc(StlHandle* h, int kind, int level, int rows, uint64_t in_dev, uint64_t out_dev) {
  return guard(-1, [&] {
    if (!h || !in_dev || !out_dev) fail("null argument");
    const size_t need = h->engine->operator_workspace_bytes(kind, level, rows);
    if (need > kScratchBytes)
      fail("stlx_apply_operator_noalloc: the operator's staging needs " + std::to_string(need) +
           " bytes, more than the " + std::to_string(kScratchBytes) + "-byte static scratch");
    void* ws = scratch_address(h->engine->options().device);
    h->engine->apply_operator_in(kind, level, rows, reinterpret_cast<const double*>(in_dev),
                                 reinterpret_cast<double*>(out_dev), ws, kScratchBytes);
    return 0;
  });
}

int stlx_apply_operator(StlHandle* h, int kind, int level, int rows, uint64_t in_dev, uint64_t out_dev) {
  return guard(-1, [&] {
    if (!h || !in_dev || !out_dev) fail("null argument");
    h->engine->apply_operator(kind, level, rows, reinterpret_cast<const double*>(in_dev),
                              reinterpret_cast<double*>(out_dev));
    return 0;
  });
}

// ---------------------------------------------------------------------------
// Map files and initial maps (crates/stl-mpi/src/io.rs, crates/stl-cli/src/lib.rs:523-540).
// ---------------------------------------------------------------------------

int stlx_map_len(const char* path, size_t* out) {
  return guard(-1, [&] {
    if (!path || !out) fail("null argument");
    *out = stl::io::map_len(path);
    return 0;
  });
}

int stlx_read_map(const char* path, double* out, size_t n) {
  return guard(-1, [&] {
    if (!path || (!out && n)) fail("null argument");
    std::vector<double> v = stl::io::read_map(path);
    if (v.size() != n)
      fail(std::string(path) + ": " + std::to_string(v.size()) + " values, buffer of " + std::to_string(n));
    std::memcpy(out, v.data(), n * sizeof(double));
    return 0;
  });
}

int stlx_read_map_range(const char* path, size_t start, size_t end, double* out) {
  return guard(-1, [&] {
    if (!path || (!out && end > start)) fail("null argument");
    std::vector<double> v = stl::io::read_map_range(path, start, end);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return 0;
  });
}

int stlx_read_npy(const char* path, double* out, size_t n, size_t* shape, size_t* ndim) {
  return guard(-1, [&] {
    if (!path || !ndim) fail("null argument");
    std::vector<size_t> dims;
    std::vector<double> v = stl::io::read_npy(path, &dims);
    if (!out) {
      // Size query: shape (when room) and rank only.
      if (shape)
        for (size_t i = 0; i < dims.size() && i < *ndim; ++i) shape[i] = dims[i];
      *ndim = dims.size();
      return 0;
    }
    if (v.size() != n)
      fail(std::string(path) + ": " + std::to_string(v.size()) + " values, buffer of " + std::to_string(n));
    std::memcpy(out, v.data(), n * sizeof(double));
    if (shape)
      for (size_t i = 0; i < dims.size() && i < *ndim; ++i) shape[i] = dims[i];
    *ndim = dims.size();
    return 0;
  });
}

int stlx_write_npy(const char* path, const double* values, const size_t* shape, size_t ndim) {
  return guard(-1, [&] {
    if (!path || (!shape && ndim)) fail("null argument");
    std::vector<size_t> dims(shape, shape + ndim);
    size_t n = 1;
    for (size_t d : dims) n *= d;
    if (!values && n) fail("null argument");
    stl::io::write_npy(path, values, dims);
    return 0;
  });
}

int stlx_write_map(const char* path, const double* values, size_t n, const char* column, bool dbl) {
  return guard(-1, [&] {
    if (!path || (!values && n)) fail("null argument");
    std::string p(path);
    auto dot = p.rfind('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    if (ext == "fits" || ext == "fit") stl::io::write_healpix_fits(p, values, n, column ? column : "VALUE", dbl);
    else stl::io::write_npy(p, values, {n});
    return 0;
  });
}

int stlx_sliced_create(const char* path, size_t n, const char* column, bool dbl) {
  return guard(-1, [&] {
    if (!path) fail("null argument");
    stl::io::sliced_create(path, n, column ? column : "VALUE", dbl);
    return 0;
  });
}

int stlx_sliced_write(const char* path, bool dbl, size_t start, const double* values, size_t n) {
  return guard(-1, [&] {
    if (!path || (!values && n)) fail("null argument");
    stl::io::sliced_write(path, dbl, start, values, n);
    return 0;
  });
}

int stlx_pixel_normal(uint64_t seed, int64_t first, size_t n, double* out) {
  return guard(-1, [&] {
    if (!out && n) fail("null argument");
    for (size_t i = 0; i < n; ++i) out[i] = stl::io::pixel_normal(seed, first + int64_t(i));
    return 0;
  });
}

int stlx_stream_copy(uint64_t src_dev, uint64_t dst_dev, size_t bytes, int reps, double* ms) {
  return guard(-1, [&] {
    if (!src_dev || !dst_dev || !ms) fail("null argument");
    hipStream_t s = nullptr;
    STL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    try {
      *ms = stl::stream_copy_ms(reinterpret_cast<const void*>(src_dev), reinterpret_cast<void*>(dst_dev), bytes,
                                reps, s);
    } catch (...) {
      (void)hipStreamDestroy(s);
      throw;
    }
    STL_HIP(hipStreamDestroy(s));
    return 0;
  });
}

// ---------------------------------------------------------------------------
// The reference's other handles: present, not implemented.
// ---------------------------------------------------------------------------

#define STL_NI_PTR(name, ...)               \
  void* name(__VA_ARGS__) {                 \
    g_error = "not implemented: " #name;    \
    return nullptr;                         \
  }
#define STL_NI_INT(name, ...)               \
  int name(__VA_ARGS__) {                   \
    g_error = "not implemented: " #name;    \
    return -1;                              \
  }

STL_NI_PTR(stlp_patch_new, const char*)
void stlp_patch_free(void*) {}
char* stlp_patch_info(void*) {
  g_error = "not implemented: stlp_patch_info";
  return nullptr;
}
STL_NI_INT(stlp_pixel_vectors, size_t, const int64_t*, size_t, double*)
STL_NI_INT(stlp_patch_select, void*, const double*, size_t, double, double, size_t*)
STL_NI_INT(stlp_patch_ids, void*, size_t, int64_t*)
STL_NI_INT(stlp_patch_core_pixels, void*, size_t*)
STL_NI_INT(stlp_patch_forward, void*, const double*, size_t, uint64_t, bool, double*, double*)
STL_NI_INT(stlp_patch_vjp, void*, const double*, size_t, uint64_t, double*)
STL_NI_INT(stlp_patch_forward_device, void*, uint64_t, size_t, uint64_t, bool, double*, double*)
STL_NI_INT(stlp_patch_vjp_device, void*, const double*, size_t, uint64_t, uint64_t)
STL_NI_PTR(stlp_spec_new, const char*)
void stlp_spec_free(void*) {}
char* stlp_spec_info(void*) {
  g_error = "not implemented: stlp_spec_info";
  return nullptr;
}
STL_NI_INT(stlp_spec_set_masks, void*, const double*, size_t)
STL_NI_INT(stlp_spec_forward, void*, const double*, size_t, uint64_t, uint64_t, uint64_t, double*)
STL_NI_INT(stlp_spec_vjp, void*, const double*, size_t, uint64_t, uint64_t, uint64_t, double*)
STL_NI_PTR(stlp_unet_new, const char*)
void stlp_unet_free(void*) {}
char* stlp_unet_info(void*) {
  g_error = "not implemented: stlp_unet_info";
  return nullptr;
}
STL_NI_INT(stlp_unet_get_params, void*, float*)
STL_NI_INT(stlp_unet_set_params, void*, const float*, size_t)
STL_NI_INT(stlp_unet_get_grads, void*, float*)
STL_NI_INT(stlp_unet_zero_grad, void*)
STL_NI_INT(stlp_unet_adam, void*, double)
STL_NI_INT(stlp_unet_set_cond, void*, const float*, size_t)
STL_NI_INT(stlp_unet_apply, void*, const double*, size_t, double*, double*, double*)
STL_NI_INT(stlp_train_set_masks, void*, const double*, size_t)
STL_NI_INT(stlp_train_set_common, void*, const double*)
STL_NI_INT(stlp_train_set_field, void*, uint32_t, const double*, const double*, const double*, const double*,
           const double*, const double*, size_t, double)
STL_NI_INT(stlp_train_bias, void*, uint32_t, const uint64_t*, const double*, size_t)
STL_NI_INT(stlp_train_loss, void*, uint32_t, bool, double*)
STL_NI_INT(stlp_train_step, void*, uint32_t, double, double*)
STL_NI_INT(stlp_train_live, void*, uint32_t, const uint64_t*, const double*, size_t, const uint64_t*, const double*,
           size_t, bool, double, double*)
char* stlp_train_lbfgs(void*, const uint32_t*, size_t, const char*) {
  g_error = "not implemented: stlp_train_lbfgs";
  return nullptr;
}
STL_NI_INT(stlp_train_set_joint, void*, const double*, const double*, uint32_t, uint32_t, const char*, double)
STL_NI_INT(stlp_train_front, void*, const double*, size_t, bool, double*)

#undef STL_NI_PTR
#undef STL_NI_INT

}  // extern "C"
