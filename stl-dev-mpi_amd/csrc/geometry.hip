// One-off device build of the pyramid operators (FP64 geometry), built for the
// rep faces 0, 4, 8; the convolution is then kept for the faces 0 and 4 (the
// south faces are mirrors of the north ones) plus a correction list of the
// mirrored columns.  Compiled with -ffp-contract=off so that products and
// sums round like the reference's scalar code.
//
// Convolution (reference conv.rs:73-135 + dia.rs:213-238): per rep row and
// gauge, the kL x 25 stencil taps are interpolated on the ring scheme
// (4 cells each, renormalised over the cells found -- all of them on the full
// sky); the kernel is then folded into one coefficient per distinct cell read
// with a nonzero weight, per gauge and (re, im), summed over taps then
// bilinear neighbours like DiaGeometry::from_conv.  Slots are sorted by cell id.
//
// Decimation (decimation.rs:450-561): per rep coarse row, the candidates are
// the children of the 2-ring (5 x 5 extended-frame window) of the coarse cell,
// or every fine pixel below coarse nside 4; Gaussian-cone weights within
// radius = 3 fine pixels, sigma = radius / 2, haversine distance, l1 rows,
// zeros dropped, columns ascending.
//
// Transposes: every forward entry (row p, cell c) of a rep row is moved by the
// one face shift that brings c into a rep face; the entries of a rep cell are
// sorted by their window offset (then row id) for the convolution, by row id
// for the decimation, which fixes the summation order of the adjoints.
#include "geometry.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "healpix_device.hpp"

namespace stl {

namespace {

constexpr int kTaps = 25;  // kernel 5 x 5
constexpr int kEntries = kL * kTaps * 4;
constexpr int kDecCand = 256;

__global__ void k_gridvec(int ksz, int64_t nside, double* out) {
  int p = threadIdx.x;
  if (p >= ksz * ksz) return;
  // stencil.rs:33-47
  int64_t half = ksz / 2;
  double ns = double(nside);
  double alpha_pix = sqrt(4.0 * hp::kPi / (12.0 * ns * ns));
  int64_t x = (p % ksz) - half, y = (p / ksz) - half;
  double dtheta = sqrt(double(x * x + y * y)) * alpha_pix;
  double dphi = atan2(double(y), double(x));
  out[p * 3 + 0] = sin(dtheta) * cos(dphi);
  out[p * 3 + 1] = sin(dtheta) * sin(dphi);
  out[p * 3 + 2] = cos(dtheta);
}

// Thread per (rep row of the chunk, gauge): 25 taps x 4 (cell, weight).
__global__ void k_taps(int order, int64_t row0, int64_t nrows, const double* __restrict__ gridvec,
                       int32_t* __restrict__ tcell, double* __restrict__ tw) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nrows * kL) return;
  int64_t i = t / kL;
  int g = int(t % kL);
  int64_t pix = rep_pixel(row0 + i, order);
  double theta, phi;
  hp::pix2ang_nest(pix, order, theta, phi);
  // centres rounded through degrees (conv.rs:89-92)
  theta = hp::deg2rad(90.0 - hp::rad2deg(hp::kPi / 2.0 - theta));
  phi = hp::deg2rad(hp::rad2deg(phi));
  double r[9];
  hp::rotation_cosmo(theta, phi, g, kL, r);
  for (int p = 0; p < kTaps; ++p) {
    double v0 = gridvec[p * 3], v1 = gridvec[p * 3 + 1], v2 = gridvec[p * 3 + 2];
    double d[3] = {r[0] * v0 + r[1] * v1 + r[2] * v2, r[3] * v0 + r[4] * v1 + r[5] * v2,
                   r[6] * v0 + r[7] * v1 + r[8] * v2};
    int64_t pc[4];
    double w[4];
    hp::interp_weights_nested(d, order, pc, w);
    // Binding on the full sky: every cell is found; weights renormalised by their sum.
    double colsum = w[0] + w[1] + w[2] + w[3];
    if (colsum == 0.0) {
      w[0] = 1.0;
      colsum = w[0] + w[1] + w[2] + w[3];
    }
    double den = colsum > 1e-12 ? colsum : 1e-12;
    size_t base = ((size_t(i) * kL + g) * kTaps + p) * 4;
    for (int n = 0; n < 4; ++n) {
      tcell[base + n] = int32_t(pc[n]);
      tw[base + n] = w[n] / den;
    }
  }
}

// One wavefront per rep row: distinct cells with a nonzero weight, sorted, and
// their folded coefficients.  WRITE = false only counts the slots.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_fold(int order, int64_t row0, int64_t nrows, int64_t R,
                                              const int32_t* __restrict__ tcell, const double* __restrict__ tw,
                                              const double* __restrict__ ker, uint8_t* __restrict__ cnt,
                                              int32_t* __restrict__ cell, int16_t* __restrict__ off,
                                              double* __restrict__ coef64, float* __restrict__ coef32,
                                              int* __restrict__ err, int* __restrict__ untiled) {
  __shared__ int32_t ec[4][kEntries];
  __shared__ double ew[4][kEntries];
  __shared__ int32_t slots[4][kMaxSlots];
  __shared__ int32_t sorted[4][kMaxSlots];
  __shared__ int nsl[4];
  int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  int64_t i = int64_t(blockIdx.x) * 4 + wv;
  bool active = i < nrows;
  if (lane == 0) nsl[wv] = 0;
  for (int e = lane; e < kEntries; e += 64) {
    ec[wv][e] = active ? tcell[size_t(i) * kEntries + e] : -1;
    ew[wv][e] = active ? tw[size_t(i) * kEntries + e] : 0.0;
  }
  __syncthreads();
  for (int e = lane; e < kEntries; e += 64) {
    if (ew[wv][e] == 0.0) continue;
    int32_t c = ec[wv][e];
    bool first = true;
    for (int e2 = 0; e2 < e; ++e2)
      if (ec[wv][e2] == c && ew[wv][e2] != 0.0) { first = false; break; }
    if (first) {
      int pos = atomicAdd(&nsl[wv], 1);
      if (pos < kMaxSlots) slots[wv][pos] = c;
    }
  }
  __syncthreads();
  int ns = nsl[wv];
  if (ns > kMaxSlots) {
    if (lane == 0) atomicOr(err, 1);
    ns = 0;
  }
  if (lane < ns) {
    int32_t c = slots[wv][lane];
    int rank = 0;
    for (int j = 0; j < ns; ++j) rank += slots[wv][j] < c ? 1 : 0;
    sorted[wv][rank] = c;
  }
  __syncthreads();
  if (!active) return;
  int64_t r = row0 + i;
  if (lane == 0) cnt[r] = uint8_t(ns);
  if (!WRITE || lane >= ns) return;
  int32_t c = sorted[wv][lane];
  double acc[kL][2];
  for (int g = 0; g < kL; ++g) {
    acc[g][0] = 0.0;
    acc[g][1] = 0.0;
    for (int p = 0; p < kTaps; ++p)
      for (int n = 0; n < 4; ++n) {
        int e = (g * kTaps + p) * 4 + n;
        double w = ew[wv][e];
        if (ec[wv][e] == c && w != 0.0) {
          acc[g][0] += ker[p] * w;
          acc[g][1] += ker[kTaps + p] * w;
        }
      }
  }
  size_t at = size_t(lane) * R + r;
  cell[at] = c;
  int dx = 0, dy = 0;
  bool ok = hp::rel_offset(rep_pixel(r, order), c, order, dx, dy);
  if (!ok || abs(dx) > kWin || abs(dy) > kWin) {
    atomicOr(untiled, 1);
    dx = dy = 0;
  }
  off[at] = pack_off(dx, dy);
  for (int g = 0; g < kL; ++g)
    for (int o = 0; o < 2; ++o) {
      if (coef64) coef64[(at * kL + g) * 2 + o] = acc[g][o];
      if (coef32) coef32[(at * kL + g) * 2 + o] = float(acc[g][o]);
    }
}

// One wavefront per rep coarse row of the decimation from order_f.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_dec_rows(int order_f, int64_t Rc, double radius, double sigma,
                                                  uint8_t* __restrict__ cnt, int32_t* __restrict__ col,
                                                  double* __restrict__ v64, float* __restrict__ v32, int S,
                                                  int* __restrict__ err) {
  __shared__ int32_t cand[4][kDecCand];
  __shared__ int32_t srt[4][kDecCand];
  __shared__ double w[4][kDecCand];
  __shared__ int ncand[4];
  __shared__ double scale[4];
  int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  int64_t r = int64_t(blockIdx.x) * 4 + wv;
  bool active = r < Rc;
  int order_c = order_f - 1;
  int64_t nside_c = int64_t(1) << order_c;
  int64_t c = rep_pixel(active ? r : 0, order_c);
  if (lane == 0) ncand[wv] = 0;
  __syncthreads();
  if (active) {
    if (nside_c >= 4) {
      if (lane < 25) {
        int64_t x, y;
        int f;
        hp::nest2xyf(c, order_c, x, y, f);
        int64_t v = hp::ext_pix(f, x + (lane % 5) - 2, y + (lane / 5) - 2, order_c);
        if (v >= 0) {
          int pos = atomicAdd(&ncand[wv], 4);
          for (int q = 0; q < 4; ++q) cand[wv][pos + q] = int32_t(4 * v + q);
        }
      }
    } else {
      int64_t nf = 12 * (int64_t(1) << (2 * order_f));
      for (int64_t f = lane; f < nf; f += 64) cand[wv][f] = int32_t(f);
      if (lane == 0) ncand[wv] = int(nf);
    }
  }
  __syncthreads();
  int n = active ? ncand[wv] : 0;
  for (int j = lane; j < n; j += 64) {
    int32_t v = cand[wv][j];
    int rank = 0;
    for (int q = 0; q < n; ++q) rank += cand[wv][q] < v ? 1 : 0;
    srt[wv][rank] = v;
  }
  __syncthreads();
  double lon_c = 0.0, lat_c = 0.0;
  if (active) hp::centre_lonlat(c, order_c, lon_c, lat_c);
  for (int j = lane; j < n; j += 64) {
    double lon_f, lat_f;
    hp::centre_lonlat(srt[wv][j], order_f, lon_f, lat_f);
    double gamma = hp::haversine(lat_c, lon_c, lat_f, lon_f);
    double t = gamma / sigma;
    w[wv][j] = gamma > radius ? 0.0 : exp(-0.5 * (t * t));
  }
  __syncthreads();
  if (lane == 0 && active) {
    double row_sum = 0.0;
    for (int j = 0; j < n; ++j) row_sum += w[wv][j];
    if (row_sum <= 0.0 && n > 0) {
      double u = 1.0 / double(n);
      for (int j = 0; j < n; ++j) w[wv][j] = u;
      row_sum = 0.0;
      for (int j = 0; j < n; ++j) row_sum += w[wv][j];
    }
    scale[wv] = row_sum > 1e-300 ? row_sum : 1e-300;
  }
  __syncthreads();
  if (!active) return;
  int nnz = 0;
  for (int j = 0; j < n; ++j) nnz += w[wv][j] != 0.0 ? 1 : 0;
  if (lane == 0) {
    if (nnz > S || nnz > 255) atomicOr(err, 2);
    cnt[r] = uint8_t(nnz > 255 ? 255 : nnz);
  }
  if (!WRITE || nnz > S) return;
  for (int j = lane; j < n; j += 64) {
    if (w[wv][j] == 0.0) continue;
    int pos = 0;
    for (int q = 0; q < j; ++q) pos += w[wv][q] != 0.0 ? 1 : 0;
    size_t at = size_t(pos) * Rc + r;
    col[at] = srt[wv][j];
    double v = w[wv][j] / scale[wv];
    if (v64) v64[at] = v;
    if (v32) v32[at] = float(v);
  }
}

// Offset of every forward decimation entry from its coarse row's first child
// (fine extended frame), and of every transposed entry's coarse row from the
// fine cell's parent (coarse extended frame); out-of-bound entries clear the
// tiled flag of their form.
__global__ void k_dec_off(int64_t Rc, int S, int order_f, const uint8_t* __restrict__ cnt,
                          const int32_t* __restrict__ col, int16_t* __restrict__ off, int* __restrict__ bad) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(S) * Rc) return;
  int s = int(t / Rc);
  int64_t r = t % Rc;
  if (s >= cnt[r]) {
    off[t] = 0;
    return;
  }
  int dx = 0, dy = 0;
  bool ok = hp::rel_offset(4 * rep_pixel(r, order_f - 1), col[t], order_f, dx, dy);
  if (!ok || dx < kDecLo || dx > kDecHi || dy < kDecLo || dy > kDecHi) {
    atomicOr(bad, 1);
    dx = dy = 0;
  }
  off[t] = pack_off(dx, dy);
}

__global__ void k_dec_toff(int64_t Rf, int Tn, int order_f, const uint8_t* __restrict__ tcnt,
                           const int32_t* __restrict__ trow, int16_t* __restrict__ toff, int* __restrict__ bad) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(Tn) * Rf) return;
  int s = int(t / Rf);
  int64_t q = t % Rf;
  if (s >= tcnt[q]) {
    toff[t] = 0;
    return;
  }
  int dx = 0, dy = 0;
  bool ok = hp::rel_offset(rep_pixel(q, order_f) >> 2, trow[t], order_f - 1, dx, dy);
  if (!ok || dx < -kDecT || dx > kDecT || dy < -kDecT || dy > kDecT) {
    atomicOr(bad, 2);
    dx = dy = 0;
  }
  toff[t] = pack_off(dx, dy);
}

__global__ void k_max_u8(int64_t n, const uint8_t* __restrict__ a, int* __restrict__ out) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int m = 0;
  for (int64_t i = t; i < n; i += int64_t(gridDim.x) * blockDim.x) m = max(m, int(a[i]));
  atomicMax(out, m);
}

__global__ void k_max_i32(int64_t n, const int* __restrict__ a, int* __restrict__ out) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int m = 0;
  for (int64_t i = t; i < n; i += int64_t(gridDim.x) * blockDim.x) m = max(m, a[i]);
  atomicMax(out, m);
}

// Target rep cell of forward entry (s, r): the face shift k0 bringing its cell
// into a rep face.
__device__ inline void entry_target(int32_t c, int order_col, int& k0, int64_t& qr) {
  int fc = int(int64_t(c) >> (2 * order_col));
  k0 = (4 - (fc & 3)) & 3;
  qr = rep_index(rot_pix(c, k0, order_col), order_col);
}

__global__ void k_tcount(int64_t R, int S, const uint8_t* __restrict__ cnt, const int32_t* __restrict__ cell,
                         int order_col, int* __restrict__ tc) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(S) * R) return;
  int s = int(t / R);
  int64_t r = t % R;
  if (s >= cnt[r]) return;
  int k0;
  int64_t qr;
  entry_target(cell[t], order_col, k0, qr);
  atomicAdd(&tc[qr], 1);
}

// C = kL * 2 with the gauge permutation (convolution) or 1 (decimation).
template <typename T, int C>
__global__ void k_tfill(int64_t R, int S, const uint8_t* __restrict__ cnt, const int32_t* __restrict__ cell,
                        const T* __restrict__ coef, int order_row, int order_col, int64_t Rq,
                        int* __restrict__ fill, int32_t* __restrict__ trow, T* __restrict__ tcoef) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(S) * R) return;
  int s = int(t / R);
  int64_t r = t % R;
  if (s >= cnt[r]) return;
  int k0;
  int64_t qr;
  entry_target(cell[t], order_col, k0, qr);
  int slot = atomicAdd(&fill[qr], 1);
  size_t at = size_t(slot) * Rq + qr;
  trow[at] = int32_t(rot_pix(rep_pixel(r, order_row), k0, order_row));
  if (C == 1) {
    tcoef[at] = coef[t];
  } else {
    for (int g = 0; g < kL; ++g) {
      int gs = gauge_src(k0, g);
      T re = coef[(size_t(t) * kL + gs) * 2];
      T im = coef[(size_t(t) * kL + gs) * 2 + 1];
      tcoef[(at * kL + g) * 2] = re;
      tcoef[(at * kL + g) * 2 + 1] = gauge_conj(k0, g) ? -im : im;
    }
  }
}

// Convolution transpose: row ids, their offsets in the cell's frame, and the
// coefficients in gauge planes [kL][Tn][Rq][2].
template <typename T>
__global__ void k_conv_tfill(int64_t R, int S, const uint8_t* __restrict__ cnt, const int32_t* __restrict__ cell,
                             const T* __restrict__ coef, int order, int Tn, int* __restrict__ fill,
                             int32_t* __restrict__ trow, int16_t* __restrict__ toff, T* __restrict__ tcoef,
                             int* __restrict__ untiled) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(S) * R) return;
  int s = int(t / R);
  int64_t r = t % R;
  if (s >= cnt[r]) return;
  int k0;
  int64_t qr;
  entry_target(cell[t], order, k0, qr);
  int slot = atomicAdd(&fill[qr], 1);
  size_t at = size_t(slot) * R + qr;
  int64_t p0 = rot_pix(rep_pixel(r, order), k0, order);
  int64_t q = rot_pix(cell[t], k0, order);
  trow[at] = int32_t(p0);
  int dx = 0, dy = 0;
  bool ok = hp::rel_offset(q, p0, order, dx, dy);
  if (!ok || abs(dx) > kWin || abs(dy) > kWin) {
    atomicOr(untiled, 1);
    dx = dy = 0;
  }
  toff[at] = pack_off(dx, dy);
  for (int g = 0; g < kL; ++g) {
    int gs = gauge_src(k0, g);
    T re = coef[(size_t(t) * kL + gs) * 2];
    T im = coef[(size_t(t) * kL + gs) * 2 + 1];
    size_t o = ((size_t(g) * Tn + slot) * R + qr) * 2;
    tcoef[o] = re;
    tcoef[o + 1] = gauge_conj(k0, g) ? -im : im;
  }
}

// Sort key of a transposed entry: its window offset, dy then dx (the lanes of
// a tile then read neighbouring LDS words at each slot), then the row id
// (unique; it orders the entries of untiled levels, whose offsets may repeat).
__device__ inline int64_t tkey(int16_t off, int32_t row) {
  return (int64_t((off_dy(off) + 64) * 128 + off_dx(off) + 64) << 32) | uint32_t(row);
}

template <typename T>
__global__ void k_conv_tsort(int64_t R, int Tn, const int* __restrict__ tc, uint8_t* __restrict__ tcnt,
                             int32_t* __restrict__ trow, int16_t* __restrict__ toff, T* __restrict__ tcoef) {
  int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (q >= R) return;
  int n = tc[q];
  tcnt[q] = uint8_t(n);
  auto cidx = [&](int g, int slot, int o) { return ((size_t(g) * Tn + slot) * R + q) * 2 + o; };
  for (int a = 1; a < n; ++a) {
    int32_t key = trow[size_t(a) * R + q];
    int16_t ko = toff[size_t(a) * R + q];
    const int64_t k64 = tkey(ko, key);
    T v[2 * kL];
    for (int g = 0; g < kL; ++g)
      for (int o = 0; o < 2; ++o) v[g * 2 + o] = tcoef[cidx(g, a, o)];
    int b = a - 1;
    while (b >= 0 && tkey(toff[size_t(b) * R + q], trow[size_t(b) * R + q]) > k64) {
      trow[size_t(b + 1) * R + q] = trow[size_t(b) * R + q];
      toff[size_t(b + 1) * R + q] = toff[size_t(b) * R + q];
      for (int g = 0; g < kL; ++g)
        for (int o = 0; o < 2; ++o) tcoef[cidx(g, b + 1, o)] = tcoef[cidx(g, b, o)];
      --b;
    }
    trow[size_t(b + 1) * R + q] = key;
    toff[size_t(b + 1) * R + q] = ko;
    for (int g = 0; g < kL; ++g)
      for (int o = 0; o < 2; ++o) tcoef[cidx(g, b + 1, o)] = v[g * 2 + o];
  }
}

template <typename T, int C>
__global__ void k_tsort(int64_t Rq, const int* __restrict__ tc, uint8_t* __restrict__ tcnt,
                        int32_t* __restrict__ trow, T* __restrict__ tcoef) {
  int64_t q = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (q >= Rq) return;
  int n = tc[q];
  tcnt[q] = uint8_t(n);
  for (int a = 1; a < n; ++a) {
    int32_t key = trow[size_t(a) * Rq + q];
    T v[C];
    for (int k = 0; k < C; ++k) v[k] = tcoef[(size_t(a) * Rq + q) * C + k];
    int b = a - 1;
    while (b >= 0 && trow[size_t(b) * Rq + q] > key) {
      trow[size_t(b + 1) * Rq + q] = trow[size_t(b) * Rq + q];
      for (int k = 0; k < C; ++k) tcoef[(size_t(b + 1) * Rq + q) * C + k] = tcoef[(size_t(b) * Rq + q) * C + k];
      --b;
    }
    trow[size_t(b + 1) * Rq + q] = key;
    for (int k = 0; k < C; ++k) tcoef[(size_t(b + 1) * Rq + q) * C + k] = v[k];
  }
}

inline unsigned blocks_for(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

int device_max_u8(const uint8_t* a, int64_t n, hipStream_t s) {
  DevArray<int> out(1);
  out.zero(s);
  k_max_u8<<<std::min<int64_t>(1024, blocks_for(n, 256)), 256, 0, s>>>(n, a, out.data());
  int h = 0;
  STL_HIP(hipMemcpyAsync(&h, out.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  STL_HIP(hipStreamSynchronize(s));
  return h;
}

int device_max_i32(const int* a, int64_t n, hipStream_t s) {
  DevArray<int> out(1);
  out.zero(s);
  k_max_i32<<<std::min<int64_t>(1024, blocks_for(n, 256)), 256, 0, s>>>(n, a, out.data());
  int h = 0;
  STL_HIP(hipMemcpyAsync(&h, out.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  STL_HIP(hipStreamSynchronize(s));
  return h;
}

void check_err(DevArray<int>& err, hipStream_t s, const char* what) {
  int h = 0;
  STL_HIP(hipMemcpyAsync(&h, err.data(), sizeof(int), hipMemcpyDeviceToHost, s));
  STL_HIP(hipStreamSynchronize(s));
  STL_CHECK(h == 0, std::string(what) + ": slot capacity exceeded");
}

// Transpose of the forward ELL (cnt, cell, coef) into the store's (tcnt, trow, tcoef).
template <typename T, int C>
void build_transpose(int64_t R, int S, const DevArray<uint8_t>& cnt, const DevArray<int32_t>& cell,
                     const DevArray<T>& coef, int order_row, int order_col, DevArray<uint8_t>& tcnt,
                     DevArray<int32_t>& trow, DevArray<T>& tcoef, int& Tn, hipStream_t s) {
  int64_t Rq = 3 * (int64_t(1) << (2 * order_col));
  DevArray<int> tc(Rq);
  tc.zero(s);
  int64_t total = int64_t(S) * R;
  k_tcount<<<blocks_for(total, 256), 256, 0, s>>>(R, S, cnt.data(), cell.data(), order_col, tc.data());
  STL_HIP(hipGetLastError());
  Tn = device_max_i32(tc.data(), Rq, s);
  STL_CHECK(Tn <= 255, "transposed operator: too many rows per cell");
  trow.resize(size_t(Tn) * Rq);
  tcoef.resize(size_t(Tn) * Rq * C);
  trow.zero(s);
  tcoef.zero(s);
  tcnt.resize(Rq);
  DevArray<int> fill(Rq);
  fill.zero(s);
  k_tfill<T, C><<<blocks_for(total, 256), 256, 0, s>>>(R, S, cnt.data(), cell.data(), coef.data(), order_row,
                                                       order_col, Rq, fill.data(), trow.data(), tcoef.data());
  STL_HIP(hipGetLastError());
  k_tsort<T, C><<<blocks_for(Rq, 128), 128, 0, s>>>(Rq, tc.data(), tcnt.data(), trow.data(), tcoef.data());
  STL_HIP(hipGetLastError());
  STL_HIP(hipStreamSynchronize(s));
}

template <typename T>
void build_conv_transpose(int order, int64_t R, int S, const DevArray<uint8_t>& cnt, const DevArray<int32_t>& cell,
                          const DevArray<T>& coef, ConvStore<T>& st, DevArray<int>& untiled, hipStream_t s) {
  DevArray<int> tc(R);
  tc.zero(s);
  int64_t total = int64_t(S) * R;
  k_tcount<<<blocks_for(total, 256), 256, 0, s>>>(R, S, cnt.data(), cell.data(), order, tc.data());
  STL_HIP(hipGetLastError());
  int Tn = device_max_i32(tc.data(), R, s);
  STL_CHECK(Tn <= 255, "transposed operator: too many rows per cell");
  st.Tn = Tn;
  st.trow.resize(size_t(Tn) * R);
  st.toff.resize(size_t(Tn) * R);
  st.tcoef.resize(size_t(kL) * Tn * R * 2);
  st.trow.zero(s);
  st.toff.zero(s);
  st.tcoef.zero(s);
  st.tcnt.resize(R);
  DevArray<int> fill(R);
  fill.zero(s);
  k_conv_tfill<T><<<blocks_for(total, 256), 256, 0, s>>>(R, S, cnt.data(), cell.data(), coef.data(), order, Tn,
                                                         fill.data(), st.trow.data(), st.toff.data(), st.tcoef.data(),
                                                         untiled.data());
  STL_HIP(hipGetLastError());
  k_conv_tsort<T><<<blocks_for(R, 128), 128, 0, s>>>(R, Tn, tc.data(), st.tcnt.data(), st.trow.data(),
                                                      st.toff.data(), st.tcoef.data());
  STL_HIP(hipGetLastError());
  STL_HIP(hipStreamSynchronize(s));
}

// Rows of the equator ring of face 4 (local (i, nside - 1 - i), rep index
// npf + local): slot counts, cells and coefficients, [i][S] and [i][S][8].
template <typename T>
__global__ void k_gather_ring(int order, int64_t R, int S, const uint8_t* __restrict__ cnt,
                              const int32_t* __restrict__ cell, const T* __restrict__ coef, uint8_t* __restrict__ ocnt,
                              int32_t* __restrict__ ocell, T* __restrict__ ocoef) {
  const int64_t nside = int64_t(1) << order;
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= nside * S) return;
  int64_t i = t / S;
  int s = int(t % S);
  int64_t r = nside * nside + hp::spread_bits(i) + (hp::spread_bits(nside - 1 - i) << 1);
  if (s == 0) ocnt[i] = cnt[r];
  size_t at = size_t(s) * R + r;
  ocell[t] = cell[at];
  for (int k = 0; k < 2 * kL; ++k) ocoef[size_t(t) * 2 * kL + k] = coef[at * 2 * kL + k];
}

// First R2 rows of every plane of a [planes][R3][width] array.
template <typename E>
void keep_rows(DevArray<E>& a, size_t planes, int64_t R3, int64_t R2, size_t width, hipStream_t s) {
  DevArray<E> b(planes * size_t(R2) * width);
  for (size_t p = 0; p < planes; ++p)
    STL_HIP(hipMemcpyAsync(b.data() + p * size_t(R2) * width, a.data() + p * size_t(R3) * width,
                           size_t(R2) * width * sizeof(E), hipMemcpyDeviceToDevice, s));
  STL_HIP(hipStreamSynchronize(s));
  a = std::move(b);
}

// Correction list of the mirrored columns (ConvOp::fix_*) from the equator-ring
// rows of the faces 4-7: for each of their entries (row p, cell c) with c in a
// south face, + A[p, c]; with c in a north face, - conj(A[p, c]) at M(c) (what
// the mirrored column of M(c) holds for p, its own mirror).  Entries of one
// (cell, row) are summed; exact zeros are dropped.
template <typename T>
void build_mirror_fix(int order, int64_t R3, ConvStore<T>& st, hipStream_t s) {
  const int64_t nside = int64_t(1) << order, npf = nside * nside;
  const int S = st.S;
  DevArray<uint8_t> dcnt(nside);
  DevArray<int32_t> dcell(size_t(nside) * S);
  DevArray<T> dcoef(size_t(nside) * S * 2 * kL);
  k_gather_ring<T><<<blocks_for(nside * S, 256), 256, 0, s>>>(order, R3, S, st.cnt.data(), st.cell.data(),
                                                              st.coef.data(), dcnt.data(), dcell.data(), dcoef.data());
  STL_HIP(hipGetLastError());
  std::vector<uint8_t> hcnt(nside);
  std::vector<int32_t> hcell(dcell.size());
  std::vector<T> hcoef(dcoef.size());
  STL_HIP(hipMemcpyAsync(hcnt.data(), dcnt.data(), dcnt.bytes(), hipMemcpyDeviceToHost, s));
  STL_HIP(hipMemcpyAsync(hcell.data(), dcell.data(), dcell.bytes(), hipMemcpyDeviceToHost, s));
  STL_HIP(hipMemcpyAsync(hcoef.data(), dcoef.data(), dcoef.bytes(), hipMemcpyDeviceToHost, s));
  STL_HIP(hipStreamSynchronize(s));
  std::map<std::pair<int32_t, int32_t>, std::array<double, 2 * kL>> acc;  // (cell, row) -> delta
  for (int k = 0; k < 4; ++k)
    for (int64_t i = 0; i < nside; ++i) {
      int64_t loc = 0;
      for (int b = 0; b < order; ++b) loc |= (((i >> b) & 1) << (2 * b)) | ((((nside - 1 - i) >> b) & 1) << (2 * b + 1));
      const int32_t row = int32_t((4 + k) * npf + loc);
      for (int sl = 0; sl < hcnt[i]; ++sl) {
        const size_t t = size_t(i) * S + sl;
        const int64_t c = rot_pix(hcell[t], k, order);
        const int fc = int(c / npf);
        if (fc >= 4 && fc < 8) continue;
        const bool south = fc >= 8;
        auto& d = acc[{int32_t(south ? c : mirror_pix(c, order)), row}];
        for (int g = 0; g < kL; ++g) {
          const int gs = gauge_src(k, g);
          double re = double(hcoef[t * 2 * kL + 2 * gs]);
          double im = double(hcoef[t * 2 * kL + 2 * gs + 1]);
          if (gauge_conj(k, g)) im = -im;
          // South: + A.  North: - conj(A).
          d[2 * g] += south ? re : -re;
          d[2 * g + 1] += im;
        }
      }
    }
  std::vector<int32_t> fcell, fptr, frow;
  std::vector<T> fcoef;
  for (const auto& kv : acc) {
    bool zero = true;
    for (double v : kv.second) zero = zero && v == 0.0;
    if (zero) continue;
    if (fcell.empty() || fcell.back() != kv.first.first) {
      fcell.push_back(kv.first.first);
      fptr.push_back(int32_t(frow.size()));
    }
    frow.push_back(kv.first.second);
    for (double v : kv.second) fcoef.push_back(T(v));
  }
  fptr.push_back(int32_t(frow.size()));
  st.nfix = int(fcell.size());
  if (st.nfix == 0) return;
  st.fix_cell.resize(fcell.size());
  st.fix_ptr.resize(fptr.size());
  st.fix_row.resize(frow.size());
  st.fix_coef.resize(fcoef.size());
  STL_HIP(hipMemcpyAsync(st.fix_cell.data(), fcell.data(), st.fix_cell.bytes(), hipMemcpyHostToDevice, s));
  STL_HIP(hipMemcpyAsync(st.fix_ptr.data(), fptr.data(), st.fix_ptr.bytes(), hipMemcpyHostToDevice, s));
  STL_HIP(hipMemcpyAsync(st.fix_row.data(), frow.data(), st.fix_row.bytes(), hipMemcpyHostToDevice, s));
  STL_HIP(hipMemcpyAsync(st.fix_coef.data(), fcoef.data(), st.fix_coef.bytes(), hipMemcpyHostToDevice, s));
  STL_HIP(hipStreamSynchronize(s));
}

double host_sum_u8(const DevArray<uint8_t>& cnt) {
  std::vector<uint8_t> h(cnt.size());
  STL_HIP(hipMemcpy(h.data(), cnt.data(), cnt.bytes(), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (auto v : h) s += v;
  return s;
}

// Union slots of a tiled level's stored rows, per block of 64 consecutive rep
// rows (an aligned 8 x 8 NESTED block: the rows of one wave in the tiled
// kernels, also in the mirrored tiles' lane order).  Slot s of every row of
// block w is made to hold the block's s-th window offset: the offsets of the
// block's rows in (dy, dx) order, each as many times as one row has it, so
// that a wave's lanes read LDS words at one uniform offset from their own
// positions, with the offset and the slot count scalar loads.  A row without
// an entry at a slot's offset gets a zero coefficient there (and id 0).  The
// entries of one offset keep their order.  Pass 1 (uids == nullptr) writes
// the block's slot count bcnt[w]; pass 2 fills the new arrays of Su slots.
// kTrans: the transposed layout (coefficients in gauge planes [kL][Sn][R][2],
// entries sorted by offset); else the forward one ([Sn][R][kL][2]).
constexpr int kWinSide = 2 * kWin + 1, kWinPos = kWinSide * kWinSide;

__device__ inline int win_pos(int16_t o) { return (off_dy(o) + kWin) * kWinSide + off_dx(o) + kWin; }

template <typename T, bool kTrans>
__host__ __device__ inline size_t coef_index(int g, int s, int64_t q, int Sn, int64_t R) {
  return kTrans ? ((size_t(g) * Sn + s) * R + q) * 2 : ((size_t(s) * R + q) * kL + g) * 2;
}

template <typename T, bool kTrans>
__global__ __launch_bounds__(64) void k_union_slots(int64_t R, int Sn, const uint8_t* __restrict__ cnt,
                                                    const int32_t* __restrict__ ids,
                                                    const int16_t* __restrict__ off, const T* __restrict__ coef,
                                                    int Su, int32_t* __restrict__ bcnt, int32_t* __restrict__ boff,
                                                    uint8_t* __restrict__ ucnt, int32_t* __restrict__ uids,
                                                    int16_t* __restrict__ uoff, T* __restrict__ ucoef) {
  __shared__ int mult[kWinPos], first[kWinPos + 1];
  __shared__ int16_t uo[256];
  const int64_t w = blockIdx.x;
  const int64_t q = w * 64 + threadIdx.x;
  for (int i = threadIdx.x; i < kWinPos; i += 64) mult[i] = 0;
  __syncthreads();
  const int n = cnt[q];
  // Occurrence k of entry s: the row's earlier entries at the same offset.
  auto occurrence = [&](int s, int pos) {
    int k = 0;
    for (int t = 0; t < s; ++t) k += win_pos(off[size_t(t) * R + q]) == pos;
    return k;
  };
  for (int s = 0; s < n; ++s) {
    const int pos = win_pos(off[size_t(s) * R + q]);
    atomicMax(&mult[pos], occurrence(s, pos) + 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int u = 0;
    for (int i = 0; i < kWinPos; ++i) {
      first[i] = u;
      u += mult[i];
    }
    first[kWinPos] = u;
    bcnt[w] = u;
  }
  __syncthreads();
  if (uids == nullptr) return;
  const int u = first[kWinPos];
  // Offsets of the block's slots; zero slots past its count.
  for (int s = threadIdx.x; s < Su; s += 64) {
    int16_t o = 0;
    if (s < u) {
      int i = 0;
      while (first[i + 1] <= s) ++i;
      o = pack_off(i % kWinSide - kWin, i / kWinSide - kWin);
    }
    uo[s] = o;
    boff[w * Su + s] = o;
  }
  __syncthreads();
  for (int s = 0; s < Su; ++s) {
    const size_t at = size_t(s) * R + q;
    uids[at] = 0;
    uoff[at] = uo[s];
    for (int g = 0; g < kL; ++g) {
      const size_t c = coef_index<T, kTrans>(g, s, q, Su, R);
      ucoef[c] = T(0);
      ucoef[c + 1] = T(0);
    }
  }
  for (int s = 0; s < n; ++s) {
    const size_t at = size_t(s) * R + q;
    const int pos = win_pos(off[at]);
    const int slot = first[pos] + occurrence(s, pos);
    uids[size_t(slot) * R + q] = ids[at];
    for (int g = 0; g < kL; ++g) {
      const size_t c = coef_index<T, kTrans>(g, s, q, Sn, R), d = coef_index<T, kTrans>(g, slot, q, Su, R);
      ucoef[d] = coef[c];
      ucoef[d + 1] = coef[c + 1];
    }
  }
  ucnt[q] = uint8_t(u);
}

__global__ void k_lin_off32(int64_t n, const int32_t* __restrict__ off, int pitch, int32_t* __restrict__ lin) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) lin[i] = off_dy(int16_t(off[i])) * pitch + off_dx(int16_t(off[i]));
}

// Re-slots (cnt, ids, off, coef) of Sn slots over R rows into union slots
// (capacity a multiple of 4: the tiled loops run whole chunks of 4 slots);
// returns the new capacity.  bcnt: the blocks' counts; lin[i]: the blocks'
// offsets in linear form for window pitches pitch[i].
template <typename T, bool kTrans>
int union_slots(int64_t R, int Sn, DevArray<uint8_t>& cnt, DevArray<int32_t>& ids, DevArray<int16_t>& off,
                DevArray<T>& coef, DevArray<int32_t>& bcnt, const int (&pitch)[3], DevArray<int32_t>* const (&lin)[3],
                hipStream_t s) {
  STL_CHECK(R % 64 == 0, "union slots: blocks of 64 rows");
  const int64_t nb = R / 64;
  bcnt.resize(size_t(nb));
  k_union_slots<T, kTrans><<<unsigned(nb), 64, 0, s>>>(R, Sn, cnt.data(), ids.data(), off.data(), coef.data(), 0,
                                                       bcnt.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  STL_HIP(hipGetLastError());
  const int umax = device_max_i32(bcnt.data(), nb, s);
  STL_CHECK(umax >= Sn && umax <= 255, "union slots: bad slot count");
  const int Su = (umax + 3) / 4 * 4;
  DevArray<uint8_t> ucnt{static_cast<size_t>(R)};
  DevArray<int32_t> uids(size_t(Su) * R);
  DevArray<int16_t> uoff(size_t(Su) * R);
  DevArray<T> ucoef(size_t(kL) * Su * R * 2);
  DevArray<int32_t> boff(size_t(nb) * Su);
  k_union_slots<T, kTrans><<<unsigned(nb), 64, 0, s>>>(R, Sn, cnt.data(), ids.data(), off.data(), coef.data(), Su,
                                                       bcnt.data(), boff.data(), ucnt.data(), uids.data(),
                                                       uoff.data(), ucoef.data());
  STL_HIP(hipGetLastError());
  for (int i = 0; i < 3; ++i) {
    lin[i]->resize(boff.size());
    k_lin_off32<<<blocks_for(int64_t(boff.size()), 256), 256, 0, s>>>(int64_t(boff.size()), boff.data(), pitch[i],
                                                                     lin[i]->data());
    STL_HIP(hipGetLastError());
  }
  STL_HIP(hipStreamSynchronize(s));
  cnt = std::move(ucnt);
  ids = std::move(uids);
  off = std::move(uoff);
  coef = std::move(ucoef);
  return Su;
}

// Rows and columns of the faces 0 and 4 only (the faces 8-11 are mirrors),
// and the union slots of the tiled adjoints.
template <typename T>
void mirror_store(int order, ConvStore<T>& st, hipStream_t s) {
  const int64_t npf = int64_t(1) << (2 * order), R3 = 3 * npf, R2 = 2 * npf;
  build_mirror_fix<T>(order, R3, st, s);
  keep_rows(st.cnt, 1, R3, R2, 1, s);
  keep_rows(st.cell, size_t(st.S), R3, R2, 1, s);
  keep_rows(st.off, size_t(st.S), R3, R2, 1, s);
  keep_rows(st.coef, size_t(st.S), R3, R2, 2 * kL, s);
  keep_rows(st.tcnt, 1, R3, R2, 1, s);
  keep_rows(st.trow, size_t(st.Tn), R3, R2, 1, s);
  keep_rows(st.toff, size_t(st.Tn), R3, R2, 1, s);
  keep_rows(st.tcoef, size_t(kL) * st.Tn, R3, R2, 2, s);
  st.slots = host_sum_u8(st.cnt);
  st.tslots = host_sum_u8(st.tcnt);
  if (st.tiled) {
    // Forward windows: pitch SW = TS + 2 kWin; the adjoints': lds_pitch(TS)
    // of their complex elements.
    st.S = union_slots<T, false>(R2, st.S, st.cnt, st.cell, st.off, st.coef, st.fbcnt,
                                 {32 + 2 * kWin, 16 + 2 * kWin, 8 + 2 * kWin}, {&st.fbo32, &st.fbo16, &st.fbo8}, s);
    st.Tn = union_slots<T, true>(R2, st.Tn, st.tcnt, st.trow, st.toff, st.tcoef, st.tbcnt,
                                 {lds_pitch(32, 2 * sizeof(T)), lds_pitch(16, 2 * sizeof(T)), lds_pitch(8, 2 * sizeof(T))},
                                 {&st.tbo32, &st.tbo16, &st.tbo8}, s);
  }
}

}  // namespace

std::vector<double> wavelet_kernel_f32(int ksz) {
  // wavelet.rs:121-145 then the f32 round trip of Kernel::shared.
  const double pi = 3.14159265358979323846;
  const double sigma = 1.0;
  int P = ksz * ksz;
  std::vector<double> re(P), im(P);
  for (int i = 0; i < ksz; ++i)
    for (int j = 0; j < ksz; ++j) {
      double x = double(j) - (double(ksz) - 1.0) / 2.0;
      double y = double(i) - (double(ksz) - 1.0) / 2.0;
      double mother = std::exp(-(x * x + y * y) / (2.0 * sigma * sigma));
      double angle = 0.5 * pi * x;
      re[i * ksz + j] = std::cos(angle) * mother;
      im[i * ksz + j] = std::sin(angle) * mother;
    }
  double sr = 0.0, si = 0.0;
  for (int p = 0; p < P; ++p) { sr += re[p]; si += im[p]; }
  double mr = sr / double(P), mi = si / double(P);
  for (int p = 0; p < P; ++p) { re[p] -= mr; im[p] -= mi; }
  double l1 = 0.0;
  for (int p = 0; p < P; ++p) l1 += std::hypot(re[p], im[p]);
  std::vector<double> out(2 * P);
  for (int p = 0; p < P; ++p) {
    out[p] = double(float(re[p] / l1));
    out[P + p] = double(float(im[p] / l1));
  }
  return out;
}

void build_conv(int order, const std::vector<double>& ker, ConvStore<double>* c64, ConvStore<float>* c32,
                hipStream_t s) {
  STL_CHECK(int(ker.size()) == 2 * kTaps, "kernel 5 only");
  int64_t nside = int64_t(1) << order;
  int64_t npf = nside * nside;
  int64_t R = 3 * npf;
  DevArray<double> gridvec(kTaps * 3), dker(2 * kTaps);
  k_gridvec<<<1, 64, 0, s>>>(5, nside, gridvec.data());
  STL_HIP(hipMemcpyAsync(dker.data(), ker.data(), ker.size() * sizeof(double), hipMemcpyHostToDevice, s));
  int64_t chunk = std::min<int64_t>(R, int64_t(1) << 18);
  DevArray<int32_t> tcell(size_t(chunk) * kEntries);
  DevArray<double> tw(size_t(chunk) * kEntries);
  DevArray<uint8_t> cnt(R);
  DevArray<int> err(1), untiled(1);
  err.zero(s);
  untiled.zero(s);
  // Pass 1: slot counts.
  for (int64_t r0 = 0; r0 < R; r0 += chunk) {
    int64_t nr = std::min(chunk, R - r0);
    k_taps<<<blocks_for(nr * kL, 128), 128, 0, s>>>(order, r0, nr, gridvec.data(), tcell.data(), tw.data());
    k_fold<false><<<blocks_for(nr, 4), 256, 0, s>>>(order, r0, nr, R, tcell.data(), tw.data(), dker.data(),
                                                    cnt.data(), nullptr, nullptr, nullptr, nullptr, err.data(),
                                                    untiled.data());
    STL_HIP(hipGetLastError());
  }
  check_err(err, s, "convolution geometry");
  int S = device_max_u8(cnt.data(), R, s);
  DevArray<int32_t> cell(size_t(S) * R);
  DevArray<int16_t> off(size_t(S) * R);
  DevArray<double> coef64;
  DevArray<float> coef32;
  cell.zero(s);
  off.zero(s);
  if (c64) { coef64.resize(size_t(S) * R * 2 * kL); coef64.zero(s); }
  if (c32) { coef32.resize(size_t(S) * R * 2 * kL); coef32.zero(s); }
  // Pass 2: cells and coefficients.
  for (int64_t r0 = 0; r0 < R; r0 += chunk) {
    int64_t nr = std::min(chunk, R - r0);
    k_taps<<<blocks_for(nr * kL, 128), 128, 0, s>>>(order, r0, nr, gridvec.data(), tcell.data(), tw.data());
    k_fold<true><<<blocks_for(nr, 4), 256, 0, s>>>(order, r0, nr, R, tcell.data(), tw.data(), dker.data(),
                                                   cnt.data(), cell.data(), off.data(),
                                                   c64 ? coef64.data() : nullptr, c32 ? coef32.data() : nullptr,
                                                   err.data(), untiled.data());
    STL_HIP(hipGetLastError());
  }
  check_err(err, s, "convolution geometry");
  tcell.release();
  tw.release();
  if (c64) {
    c64->order = order;
    c64->S = S;
    build_conv_transpose<double>(order, R, S, cnt, cell, coef64, *c64, untiled, s);
  }
  if (c32) {
    c32->order = order;
    c32->S = S;
    build_conv_transpose<float>(order, R, S, cnt, cell, coef32, *c32, untiled, s);
  }
  int untiled_h = 0;
  STL_HIP(hipMemcpy(&untiled_h, untiled.data(), sizeof(int), hipMemcpyDeviceToHost));
  bool tiled = order >= 3 && untiled_h == 0;
  if (c64) c64->tiled = tiled;
  if (c32) c32->tiled = tiled;
  // Forward arrays: shared ids, owned per store.
  if (c64 && c32) {
    c32->cnt.resize(R);
    c32->cell.resize(cell.size());
    c32->off.resize(off.size());
    STL_HIP(hipMemcpyAsync(c32->cnt.data(), cnt.data(), cnt.bytes(), hipMemcpyDeviceToDevice, s));
    STL_HIP(hipMemcpyAsync(c32->cell.data(), cell.data(), cell.bytes(), hipMemcpyDeviceToDevice, s));
    STL_HIP(hipMemcpyAsync(c32->off.data(), off.data(), off.bytes(), hipMemcpyDeviceToDevice, s));
    c32->coef = std::move(coef32);
    c64->cnt = std::move(cnt);
    c64->cell = std::move(cell);
    c64->off = std::move(off);
    c64->coef = std::move(coef64);
  } else if (c64) {
    c64->cnt = std::move(cnt);
    c64->cell = std::move(cell);
    c64->off = std::move(off);
    c64->coef = std::move(coef64);
  } else if (c32) {
    c32->cnt = std::move(cnt);
    c32->cell = std::move(cell);
    c32->off = std::move(off);
    c32->coef = std::move(coef32);
  }
  STL_HIP(hipStreamSynchronize(s));
  if (c64) mirror_store(order, *c64, s);
  if (c32) mirror_store(order, *c32, s);
}

void build_dec(int order_f, DecStore<double>* d64, DecStore<float>* d32, hipStream_t s) {
  STL_CHECK(order_f >= 1, "decimation needs nside >= 2");
  const double pi = 3.14159265358979323846;
  int64_t nside_f = int64_t(1) << order_f;
  int64_t Rc = 3 * (int64_t(1) << (2 * (order_f - 1)));
  // decimation.rs:478-484: radius of 3 fine pixels (CPython degrees/radians), sigma = radius / 2.
  double pix_deg = std::sqrt(4.0 * pi / double(12 * nside_f * nside_f)) * (180.0 / pi);
  double radius = (3.0 * pix_deg) * (pi / 180.0);
  double sigma = radius / 2.0;
  DevArray<uint8_t> cnt(Rc);
  DevArray<int> err(1);
  err.zero(s);
  k_dec_rows<false><<<blocks_for(Rc, 4), 256, 0, s>>>(order_f, Rc, radius, sigma, cnt.data(), nullptr, nullptr,
                                                      nullptr, 255, err.data());
  STL_HIP(hipGetLastError());
  check_err(err, s, "decimation");
  int S = device_max_u8(cnt.data(), Rc, s);
  DevArray<int32_t> col(size_t(S) * Rc);
  DevArray<double> v64;
  DevArray<float> v32;
  col.zero(s);
  if (d64) { v64.resize(size_t(S) * Rc); v64.zero(s); }
  if (d32) { v32.resize(size_t(S) * Rc); v32.zero(s); }
  k_dec_rows<true><<<blocks_for(Rc, 4), 256, 0, s>>>(order_f, Rc, radius, sigma, cnt.data(), col.data(),
                                                     d64 ? v64.data() : nullptr, d32 ? v32.data() : nullptr, S,
                                                     err.data());
  STL_HIP(hipGetLastError());
  check_err(err, s, "decimation");
  if (d64) {
    d64->order_f = order_f;
    d64->S = S;
    build_transpose<double, 1>(Rc, S, cnt, col, v64, order_f - 1, order_f, d64->tcnt, d64->trow, d64->tval,
                               d64->Tn, s);
  }
  if (d32) {
    d32->order_f = order_f;
    d32->S = S;
    build_transpose<float, 1>(Rc, S, cnt, col, v32, order_f - 1, order_f, d32->tcnt, d32->trow, d32->tval, d32->Tn,
                              s);
  }
  // Offsets for the LDS-tiled kernels (coarse nside >= 8 forward, fine nside
  // >= 16 adjoint), shared by both precisions.
  DevArray<int16_t> off, toff;
  int tiled_bad = 3;
  const int Tn = d64 ? d64->Tn : d32->Tn;
  const DevArray<uint8_t>& tcnt = d64 ? d64->tcnt : d32->tcnt;
  const DevArray<int32_t>& trow = d64 ? d64->trow : d32->trow;
  if (order_f - 1 >= 3) {
    DevArray<int> bad(1);
    bad.zero(s);
    off.resize(size_t(S) * Rc);
    k_dec_off<<<blocks_for(int64_t(S) * Rc, 256), 256, 0, s>>>(Rc, S, order_f, cnt.data(), col.data(), off.data(),
                                                               bad.data());
    STL_HIP(hipGetLastError());
    int64_t Rf = 3 * (int64_t(1) << (2 * order_f));
    toff.resize(size_t(Tn) * Rf);
    k_dec_toff<<<blocks_for(int64_t(Tn) * Rf, 256), 256, 0, s>>>(Rf, Tn, order_f, tcnt.data(), trow.data(),
                                                                 toff.data(), bad.data());
    STL_HIP(hipGetLastError());
    STL_HIP(hipMemcpyAsync(&tiled_bad, bad.data(), sizeof(int), hipMemcpyDeviceToHost, s));
    STL_HIP(hipStreamSynchronize(s));
  }
  for (auto* d : {d64 ? &d64->tiled : nullptr, d32 ? &d32->tiled : nullptr})
    if (d) *d = !(tiled_bad & 1);
  for (auto* d : {d64 ? &d64->ttiled : nullptr, d32 ? &d32->ttiled : nullptr})
    if (d) *d = !(tiled_bad & 2) && order_f >= 4;
  if (d64 && d32) {
    d32->off.resize(off.size());
    d32->toff.resize(toff.size());
    if (off.size()) STL_HIP(hipMemcpyAsync(d32->off.data(), off.data(), off.bytes(), hipMemcpyDeviceToDevice, s));
    if (toff.size()) STL_HIP(hipMemcpyAsync(d32->toff.data(), toff.data(), toff.bytes(), hipMemcpyDeviceToDevice, s));
    d64->off = std::move(off);
    d64->toff = std::move(toff);
  } else if (d64) {
    d64->off = std::move(off);
    d64->toff = std::move(toff);
  } else {
    d32->off = std::move(off);
    d32->toff = std::move(toff);
  }
  if (d64 && d32) {
    d32->cnt.resize(Rc);
    d32->col.resize(col.size());
    STL_HIP(hipMemcpyAsync(d32->cnt.data(), cnt.data(), cnt.bytes(), hipMemcpyDeviceToDevice, s));
    STL_HIP(hipMemcpyAsync(d32->col.data(), col.data(), col.bytes(), hipMemcpyDeviceToDevice, s));
    d32->val = std::move(v32);
    d64->cnt = std::move(cnt);
    d64->col = std::move(col);
    d64->val = std::move(v64);
  } else if (d64) {
    d64->cnt = std::move(cnt);
    d64->col = std::move(col);
    d64->val = std::move(v64);
  } else if (d32) {
    d32->cnt = std::move(cnt);
    d32->col = std::move(col);
    d32->val = std::move(v32);
  }
  STL_HIP(hipStreamSynchronize(s));
}

namespace {

double sum_u8(const DevArray<uint8_t>& cnt) {
  std::vector<uint8_t> h(cnt.size());
  STL_HIP(hipMemcpy(h.data(), cnt.data(), cnt.bytes(), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (auto v : h) s += v;
  return s;
}

std::mutex& cache_mutex() {
  static std::mutex m;
  return m;
}

// (device, fp64, order) -> operators still held by some engine.
template <typename Op>
std::map<std::tuple<int, bool, int>, std::weak_ptr<const Op>>& cache() {
  static std::map<std::tuple<int, bool, int>, std::weak_ptr<const Op>> c;
  return c;
}

}  // namespace

template <typename T>
std::shared_ptr<const SharedConv<T>> shared_conv(int device, int order, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(cache_mutex());
  auto key = std::make_tuple(device, std::is_same<T, double>::value, order);
  auto& c = cache<SharedConv<T>>();
  if (auto hit = c[key].lock()) return hit;
  auto op = std::make_shared<SharedConv<T>>();
  static const std::vector<double> ker = wavelet_kernel_f32(5);
  if constexpr (std::is_same<T, double>::value) build_conv(order, ker, &op->store, nullptr, stream);
  else build_conv(order, ker, nullptr, &op->store, stream);
  op->slots = op->store.slots;
  op->tslots = op->store.tslots;
  c[key] = op;
  return op;
}

template <typename T>
std::shared_ptr<const SharedDec<T>> shared_dec(int device, int order_f, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(cache_mutex());
  auto key = std::make_tuple(device, std::is_same<T, double>::value, order_f);
  auto& c = cache<SharedDec<T>>();
  if (auto hit = c[key].lock()) return hit;
  auto op = std::make_shared<SharedDec<T>>();
  if constexpr (std::is_same<T, double>::value) build_dec(order_f, &op->store, nullptr, stream);
  else build_dec(order_f, nullptr, &op->store, stream);
  op->slots = sum_u8(op->store.cnt);
  op->tslots = sum_u8(op->store.tcnt);
  c[key] = op;
  return op;
}

template std::shared_ptr<const SharedConv<double>> shared_conv<double>(int, int, hipStream_t);
template std::shared_ptr<const SharedConv<float>> shared_conv<float>(int, int, hipStream_t);
template std::shared_ptr<const SharedDec<double>> shared_dec<double>(int, int, hipStream_t);
template std::shared_ptr<const SharedDec<float>> shared_dec<float>(int, int, hipStream_t);

}  // namespace stl
