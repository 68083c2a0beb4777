// FP64 vector operations of the device L-BFGS (element-wise kernels with
// fused multiply-adds like torch's add_(u, alpha); reductions by fixed chunks
// of 2^16 elements, one workgroup per chunk with a fixed tree, chunk results
// combined on the host in chunk order: deterministic for a given length).
#include "vectors.hpp"

#include <algorithm>
#include <vector>

namespace stl {

namespace {

constexpr int kThreads = 256;

inline unsigned blocks(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

__global__ void k_neg(int64_t n, const double* __restrict__ a, double* __restrict__ o) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) o[i] = -a[i];
}

__global__ void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) o[i] = a[i] - b[i];
}

__global__ void k_scale(int64_t n, const double* __restrict__ a, double s, double* __restrict__ o) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) o[i] = a[i] * s;
}

// o = d * t + x (fused)
__global__ void k_fma(int64_t n, const double* __restrict__ d, double t, const double* x, double* o) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) o[i] = fma(d[i], t, x[i]);
}

// kind 0: sum a*b, 1: sum |a|, 2: max |a|, 3: max |a * t|
__global__ __launch_bounds__(kThreads) void k_chunk_reduce(int64_t n, int kind, const double* __restrict__ a,
                                                           const double* __restrict__ b, double t,
                                                           double* __restrict__ out) {
  __shared__ double sh[kThreads];
  int64_t c0 = int64_t(blockIdx.x) * VectorOps::kChunk;
  int64_t c1 = min(n, c0 + VectorOps::kChunk);
  double acc = 0.0;
  for (int64_t i = c0 + threadIdx.x; i < c1; i += kThreads) {
    double v;
    switch (kind) {
      case 0: v = a[i] * b[i]; acc += v; break;
      case 1: acc += fabs(a[i]); break;
      case 2: acc = fmax(acc, fabs(a[i])); break;
      default: acc = fmax(acc, fabs(a[i] * t)); break;
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double x = sh[threadIdx.x], y = sh[threadIdx.x + w];
      sh[threadIdx.x] = kind <= 1 ? x + y : fmax(x, y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

}  // namespace

VectorOps::VectorOps(int64_t n, hipStream_t s) : n_(n), s_(s) {
  nchunks_ = (n + kChunk - 1) / kChunk;
  partial_.resize(size_t(nchunks_));
  host_.resize(size_t(nchunks_));
}

VectorOps::~VectorOps() {
  for (double* p : free_) (void)hipFree(p);
}

double* VectorOps::acquire() {
  if (!free_.empty()) {
    double* p = free_.back();
    free_.pop_back();
    return p;
  }
  double* p = nullptr;
  STL_HIP(hipMalloc(reinterpret_cast<void**>(&p), size_t(n_) * sizeof(double)));
  return p;
}

void VectorOps::release(double* p) {
  if (p) free_.push_back(p);
}

VectorOps::Vec VectorOps::make() { return Vec(this, acquire()); }

VectorOps::Vec VectorOps::copy(const Vec& a) {
  Vec o = make();
  STL_HIP(hipMemcpyAsync(o.p, a.p, size_t(n_) * sizeof(double), hipMemcpyDeviceToDevice, s_));
  return o;
}

VectorOps::Vec VectorOps::neg(const Vec& a) {
  Vec o = make();
  k_neg<<<blocks(n_), kThreads, 0, s_>>>(n_, a.p, o.p);
  STL_HIP(hipGetLastError());
  return o;
}

VectorOps::Vec VectorOps::sub(const Vec& a, const Vec& b) {
  Vec o = make();
  k_sub<<<blocks(n_), kThreads, 0, s_>>>(n_, a.p, b.p, o.p);
  STL_HIP(hipGetLastError());
  return o;
}

VectorOps::Vec VectorOps::scale(const Vec& a, double s) {
  Vec o = make();
  k_scale<<<blocks(n_), kThreads, 0, s_>>>(n_, a.p, s, o.p);
  STL_HIP(hipGetLastError());
  return o;
}

VectorOps::Vec VectorOps::fma_new(const Vec& d, double t, const Vec& x) {
  Vec o = make();
  fma_into(o, d, t, x);
  return o;
}

void VectorOps::fma_into(Vec& out, const Vec& d, double t, const Vec& x) {
  k_fma<<<blocks(n_), kThreads, 0, s_>>>(n_, d.p, t, x.p, out.p);
  STL_HIP(hipGetLastError());
}

void VectorOps::axpy(Vec& p, double a, const Vec& u) {
  k_fma<<<blocks(n_), kThreads, 0, s_>>>(n_, u.p, a, p.p, p.p);
  STL_HIP(hipGetLastError());
}

double VectorOps::reduce(int kind, const double* a, const double* b, double t) {
  k_chunk_reduce<<<unsigned(nchunks_), kThreads, 0, s_>>>(n_, kind, a, b, t, partial_.data());
  STL_HIP(hipGetLastError());
  STL_HIP(hipMemcpyAsync(host_.data(), partial_.data(), size_t(nchunks_) * sizeof(double), hipMemcpyDeviceToHost,
                         s_));
  STL_HIP(hipStreamSynchronize(s_));
  double r = 0.0;
  for (double v : host_) r = kind <= 1 ? r + v : std::max(r, v);
  return r;
}

double VectorOps::dot(const Vec& a, const Vec& b) { return reduce(0, a.p, b.p, 0.0); }
double VectorOps::abs_sum(const Vec& a) { return reduce(1, a.p, nullptr, 0.0); }
double VectorOps::abs_max(const Vec& a) { return reduce(2, a.p, nullptr, 0.0); }
double VectorOps::scaled_abs_max(const Vec& d, double t) { return reduce(3, d.p, nullptr, t); }

}  // namespace stl
