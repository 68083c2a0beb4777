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

// Stream copy of n 16-byte words: each workgroup moves 4 x 256 consecutive
// words (16 KB), every lane 4 words 4 KB apart, loads before stores.
__global__ __launch_bounds__(kThreads) void k_stream_copy(int64_t n, const float4* __restrict__ src,
                                                          float4* __restrict__ dst) {
  const int64_t base = int64_t(blockIdx.x) * 4 * kThreads + threadIdx.x;
  float4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + int64_t(k) * kThreads;
    if (i < n) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + int64_t(k) * kThreads;
    if (i < n) dst[i] = v[k];
  }
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

// p = u * a + p (as k_fma), then the per-chunk sums of w * p over the
// updated p, with k_chunk_reduce's chunks, strides and tree (kind 0): the
// same values as the two kernels, in one pass.
__global__ __launch_bounds__(kThreads) void k_axpy_dot(int64_t n, double a, const double* __restrict__ u,
                                                       const double* __restrict__ w, double* __restrict__ p,
                                                       double* __restrict__ out) {
  __shared__ double sh[kThreads];
  int64_t c0 = int64_t(blockIdx.x) * VectorOps::kChunk;
  int64_t c1 = min(n, c0 + VectorOps::kChunk);
  double acc = 0.0;
  for (int64_t i = c0 + threadIdx.x; i < c1; i += kThreads) {
    const double pn = fma(u[i], a, p[i]);
    p[i] = pn;
    double v = w[i] * pn;
    acc += v;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int k = kThreads / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

}  // namespace

VectorOps::VectorOps(int64_t n, hipStream_t s) : n_(n), s_(s) {
  nchunks_ = (n + kChunk - 1) / kChunk;
  partial_.resize(size_t(nchunks_));
  STL_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_), size_t(std::max<int64_t>(nchunks_, 1)) * sizeof(double),
                        hipHostMallocDefault));
}

VectorOps::~VectorOps() {
  for (double* p : free_) (void)hipFree(p);
  if (host_) (void)hipHostFree(host_);
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
  return gather(kind);
}

// The chunk partials on the host, combined in chunk order (sum for kinds 0
// and 1, max otherwise).
double VectorOps::gather(int kind) {
  STL_HIP(hipMemcpyAsync(host_, partial_.data(), size_t(nchunks_) * sizeof(double), hipMemcpyDeviceToHost, s_));
  STL_HIP(hipStreamSynchronize(s_));
  double r = 0.0;
  for (int64_t c = 0; c < nchunks_; ++c) r = kind <= 1 ? r + host_[c] : std::max(r, host_[c]);
  return r;
}

double VectorOps::dot(const Vec& a, const Vec& b) { return reduce(0, a.p, b.p, 0.0); }

double VectorOps::axpy_dot(Vec& p, double a, const Vec& u, const Vec& w) {
  k_axpy_dot<<<unsigned(nchunks_), kThreads, 0, s_>>>(n_, a, u.p, w.p, p.p, partial_.data());
  STL_HIP(hipGetLastError());
  return gather(0);
}
double VectorOps::abs_sum(const Vec& a) { return reduce(1, a.p, nullptr, 0.0); }
double VectorOps::abs_max(const Vec& a) { return reduce(2, a.p, nullptr, 0.0); }
double VectorOps::scaled_abs_max(const Vec& d, double t) { return reduce(3, d.p, nullptr, t); }

// Mean milliseconds of one stream copy of `bytes` (a multiple of 16, both
// addresses 16-byte aligned) over `reps` copies after one warm-up, HIP events
// on the stream the copies run on.
double stream_copy_ms(const void* src, void* dst, size_t bytes, int reps, hipStream_t s) {
  STL_CHECK(bytes % 16 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0 &&
                reinterpret_cast<uintptr_t>(dst) % 16 == 0 && reps >= 1,
            "stream copy: 16-byte multiples and alignment, reps >= 1");
  const int64_t n = int64_t(bytes / 16);
  const unsigned grid = unsigned((n + 4 * kThreads - 1) / (4 * kThreads));
  auto copy = [&] {
    k_stream_copy<<<grid, kThreads, 0, s>>>(n, static_cast<const float4*>(src), static_cast<float4*>(dst));
    STL_HIP(hipGetLastError());
  };
  hipEvent_t a, b;
  STL_HIP(hipEventCreate(&a));
  STL_HIP(hipEventCreate(&b));
  copy();
  STL_HIP(hipEventRecord(a, s));
  for (int r = 0; r < reps; ++r) copy();
  STL_HIP(hipEventRecord(b, s));
  STL_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  STL_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return double(ms) / reps;
}

}  // namespace stl
