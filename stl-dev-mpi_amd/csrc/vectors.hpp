// Device FP64 vectors of one length for the L-BFGS iteration (a back end of
// lbfgs.hpp), with a pool of buffers so that the iteration allocates nothing
// after its first steps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "device_buffer.hpp"

namespace stl {

class VectorOps {
 public:
  // Elements per reduction block: 1536 blocks for the 12.6 M doubles of an
  // nside-1024 map, enough to fill the GPU (bandwidth-bound reductions).
  static constexpr int64_t kChunk = int64_t(1) << 13;

  // A pooled device vector of n doubles; returns its buffer to the pool.
  class Vec {
   public:
    Vec() = default;
    Vec(VectorOps* o, double* p) : ops(o), p(p) {}
    Vec(const Vec&) = delete;
    Vec& operator=(const Vec&) = delete;
    Vec(Vec&& v) noexcept : ops(v.ops), p(v.p) { v.p = nullptr; }
    Vec& operator=(Vec&& v) noexcept {
      if (this != &v) {
        if (p) ops->release(p);
        ops = v.ops;
        p = v.p;
        v.p = nullptr;
      }
      return *this;
    }
    ~Vec() {
      if (p) ops->release(p);
    }
    VectorOps* ops = nullptr;
    double* p = nullptr;
  };

  VectorOps(int64_t n, hipStream_t s);
  ~VectorOps();
  VectorOps(const VectorOps&) = delete;
  VectorOps& operator=(const VectorOps&) = delete;

  int64_t size() const { return n_; }
  Vec make();
  Vec copy(const Vec& a);
  Vec neg(const Vec& a);
  Vec sub(const Vec& a, const Vec& b);
  Vec scale(const Vec& a, double s);
  Vec fma_new(const Vec& d, double t, const Vec& x);
  void fma_into(Vec& out, const Vec& d, double t, const Vec& x);
  void axpy(Vec& p, double a, const Vec& u);
  double dot(const Vec& a, const Vec& b);
  // p += a u, then dot(w, p) of the updated p: the values of axpy and dot, in one pass.
  double axpy_dot(Vec& p, double a, const Vec& u, const Vec& w);
  double abs_sum(const Vec& a);
  double abs_max(const Vec& a);
  double scaled_abs_max(const Vec& d, double t);

 private:
  double* acquire();
  void release(double* p);
  double reduce(int kind, const double* a, const double* b, double t);
  double gather(int kind);

  int64_t n_ = 0, nchunks_ = 0;
  hipStream_t s_ = nullptr;
  std::vector<double*> free_;
  DevArray<double> partial_;
  double* host_ = nullptr;  // pinned: the per-block partials come back by a device copy
};

// Mean ms of one copy of `bytes` by a hand-written stream kernel (the
// attainable HBM rate, against which the engine's kernels are read).
double stream_copy_ms(const void* src, void* dst, size_t bytes, int reps, hipStream_t s);

}  // namespace stl
