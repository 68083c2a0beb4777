// Forward-sweep and VJP kernels of the ScatCov pyramid.  Every kernel is a
// gather (no floating-point atomics), so results are deterministic; sums over
// the sky are FP64 with a fixed order (per-thread, then a fixed LDS tree or
// group order, then blocks in order).
//
// Threads of the convolution / decimation kernels are (rep row, face shift k):
// the 4 lanes of a rep row share its coefficients (one fetch per wavefront),
// each applies them to its own face through the face-rotation symmetry
// (gauge g reads the rep coefficient of gauge g ^ 2(k & 1), conjugated per
// gauge_conj).  Convolution rows come in 3 npf "virtual" rows: the stored
// rows of the faces 0 and 4, then the mirrors of face 0's rows (conjugated
// coefficients, cells and outputs at their mirror pixels).
#include "sweeps.hpp"

#include <algorithm>

namespace stl {

namespace {

inline unsigned nblocks(int64_t n, int t) { return unsigned((n + t - 1) / t); }

// Gram entries per thread when there are more entries than threads.

template <typename T>
struct Load;

template <>
struct Load<double> {
  __device__ static inline void coef8(const double* p, double c[8]) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double2 v = q[i];
      c[2 * i] = v.x;
      c[2 * i + 1] = v.y;
    }
  }
  template <int NI>
  __device__ static inline void vec(const double* p, double v[NI]) {
    if constexpr (NI == 4) {
      const double2* q = reinterpret_cast<const double2*>(p);
      double2 a = q[0], b = q[1];
      v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else {
#pragma unroll
      for (int j = 0; j < NI; ++j) v[j] = p[j];
    }
  }
};

template <>
struct Load<float> {
  __device__ static inline void coef8(const float* p, float c[8]) {
    const float4* q = reinterpret_cast<const float4*>(p);
    float4 a = q[0], b = q[1];
    c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w;
    c[4] = b.x; c[5] = b.y; c[6] = b.z; c[7] = b.w;
  }
  template <int NI>
  __device__ static inline void vec(const float* p, float v[NI]) {
    if constexpr (NI == 4) {
      float4 a = *reinterpret_cast<const float4*>(p);
      v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
#pragma unroll
      for (int j = 0; j < NI; ++j) v[j] = p[j];
    }
  }
};

template <typename T>
__device__ inline cplx_t<T> make_c(T re, T im) {
  cplx_t<T> v;
  v.x = re;
  v.y = im;
  return v;
}

// Coefficients of face shift k from the rep coefficients c[g][o].
template <typename T>
__device__ inline void permute(int k, const T c[8], T a[8]) {
  bool swap = k & 1;
#pragma unroll
  for (int g = 0; g < kL; ++g) {
    int gs = swap ? (g ^ 2) : g;
    T re = swap ? c[(g ^ 2) * 2] : c[g * 2];
    T im = swap ? c[(g ^ 2) * 2 + 1] : c[g * 2 + 1];
    (void)gs;
    a[g * 2] = re;
    a[g * 2 + 1] = gauge_conj(k, g) ? -im : im;
  }
}

// y[row][g][p][j] = sum_s A[p, cell_s][g] x[row][cell_s][j]
template <typename T, int NI>
__global__ __launch_bounds__(256) void k_conv_fwd(ConvOp<T> op, const T* __restrict__ x,
                                                  cplx_t<T>* __restrict__ y, T* __restrict__ mod) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= 12 * op.npf) return;
  const bool m = (t >> 2) >= op.R;
  int64_t r = m ? (t >> 2) - op.R : t >> 2;
  int k = int(t & 3);
  int row = blockIdx.y;
  int64_t N = 12 * op.npf;
  const T* xr = x + size_t(row) * N * NI;
  T acc[NI][kL][2];
#pragma unroll
  for (int j = 0; j < NI; ++j)
#pragma unroll
    for (int g = 0; g < kL; ++g) acc[j][g][0] = acc[j][g][1] = T(0);
  int cnt = op.cnt[r];
  for (int s = 0; s < cnt; ++s) {
    size_t at = size_t(s) * op.R + r;
    int64_t ck = rot_pix(op.cell[at], k, op.order);
    if (m) ck = mirror_pix(ck, op.order);
    T c[8], a[8];
    Load<T>::coef8(op.coef + at * 8, c);
    permute<T>(k, c, a);
    T xv[NI];
    Load<T>::template vec<NI>(xr + size_t(ck) * NI, xv);
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int g = 0; g < kL; ++g) {
        acc[j][g][0] += a[g * 2] * xv[j];
        acc[j][g][1] += a[g * 2 + 1] * xv[j];
      }
  }
  int64_t p = rot_pix(rep_pixel(r, op.order), k, op.order);
  if (m) p = mirror_pix(p, op.order);
#pragma unroll
  for (int g = 0; g < kL; ++g)
#pragma unroll
    for (int j = 0; j < NI; ++j)
      y[((size_t(row) * kL + g) * N + p) * NI + j] = make_c<T>(acc[j][g][0], m ? -acc[j][g][1] : acc[j][g][1]);
  if (mod != nullptr && NI == 1) {
#pragma unroll
    for (int g = 0; g < kL; ++g) mod[(size_t(row) * N + p) * kL + g] = hypot(acc[0][g][0], acc[0][g][1]);
  }
}

// gx[row][q][j] = sum_rows p reading q sum_g Re(conj-free pairing) A[p, q][g] . gy[row][g][p][j]
template <typename T, int NI>
__global__ __launch_bounds__(256) void k_conv_adj(ConvOp<T> op, const cplx_t<T>* __restrict__ gy,
                                                  T* __restrict__ gx) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= 12 * op.npf) return;
  const bool m = (t >> 2) >= op.R;
  int64_t q = m ? (t >> 2) - op.R : t >> 2;
  int k = int(t & 3);
  int row = blockIdx.y;
  int64_t N = 12 * op.npf;
  const cplx_t<T>* gr = gy + size_t(row) * kL * N * NI;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
  int cnt = op.tcnt[q];
  for (int s = 0; s < cnt; ++s) {
    size_t at = size_t(s) * op.R + q;
    int64_t pk = rot_pix(op.trow[at], k, op.order);
    if (m) pk = mirror_pix(pk, op.order);
    T c[8], a[8];
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      cplx_t<T> v = reinterpret_cast<const cplx_t<T>*>(op.tcoef)[(size_t(g) * op.Tn + s) * op.R + q];
      c[2 * g] = v.x;
      c[2 * g + 1] = m ? -v.y : v.y;
    }
    permute<T>(k, c, a);
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      const cplx_t<T>* src = gr + (size_t(g) * N + pk) * NI;
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        cplx_t<T> v = src[j];
        acc[j] = fma(a[g * 2 + 1], v.y, fma(a[g * 2], v.x, acc[j]));
      }
    }
  }
  int64_t qk = rot_pix(rep_pixel(q, op.order), k, op.order);
  if (m) qk = mirror_pix(qk, op.order);
#pragma unroll
  for (int j = 0; j < NI; ++j) gx[(size_t(row) * N + qk) * NI + j] = acc[j];
}

// The same gathers with the slots of one output spread over LP consecutive
// lanes, for levels too small to fill the GPU: each thread's chain of
// dependent loads is LP times shorter.  The LP partial sums are added by a
// fixed butterfly (deterministic).
// Bodies of the 256-thread kernels take their block (bx, row) as arguments, so
// that the work-queue kernel of the coarse levels can run them as items.
template <typename T, int NI, int LP>
__device__ __forceinline__ void conv_fwd_split_body(const ConvOp<T>& op, const T* __restrict__ x,
                                                    cplx_t<T>* __restrict__ y, T* __restrict__ mod, int64_t bx,
                                                    int row) {
  const int64_t o = (bx * 256 + threadIdx.x) / LP;
  const int lane = int(threadIdx.x % LP);
  const bool active = o < 12 * op.npf;
  const bool mir = active && (o >> 2) >= op.R;
  const int64_t r = active ? (mir ? (o >> 2) - op.R : o >> 2) : 0;
  const int k = int(o & 3);
  const int64_t N = 12 * op.npf;
  const T* xr = x + size_t(row) * N * NI;
  T acc[NI][kL][2];
#pragma unroll
  for (int j = 0; j < NI; ++j)
#pragma unroll
    for (int g = 0; g < kL; ++g) acc[j][g][0] = acc[j][g][1] = T(0);
  const int cnt = active ? op.cnt[r] : 0;
  for (int s = lane; s < cnt; s += LP) {
    size_t at = size_t(s) * op.R + r;
    int64_t ck = rot_pix(op.cell[at], k, op.order);
    if (mir) ck = mirror_pix(ck, op.order);
    T c[8], a[8];
    Load<T>::coef8(op.coef + at * 8, c);
    permute<T>(k, c, a);
    T xv[NI];
    Load<T>::template vec<NI>(xr + size_t(ck) * NI, xv);
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int g = 0; g < kL; ++g) {
        acc[j][g][0] += a[g * 2] * xv[j];
        acc[j][g][1] += a[g * 2 + 1] * xv[j];
      }
  }
#pragma unroll
  for (int m = LP / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int g = 0; g < kL; ++g) {
        acc[j][g][0] += __shfl_xor(acc[j][g][0], m);
        acc[j][g][1] += __shfl_xor(acc[j][g][1], m);
      }
  if (!active || lane != 0) return;
  int64_t p = rot_pix(rep_pixel(r, op.order), k, op.order);
  if (mir) p = mirror_pix(p, op.order);
#pragma unroll
  for (int g = 0; g < kL; ++g)
#pragma unroll
    for (int j = 0; j < NI; ++j)
      y[((size_t(row) * kL + g) * N + p) * NI + j] = make_c<T>(acc[j][g][0], mir ? -acc[j][g][1] : acc[j][g][1]);
  if (mod != nullptr && NI == 1) {
#pragma unroll
    for (int g = 0; g < kL; ++g) mod[(size_t(row) * N + p) * kL + g] = hypot(acc[0][g][0], acc[0][g][1]);
  }
}

template <typename T, int NI, int LP>
__global__ __launch_bounds__(256) void k_conv_fwd_split(ConvOp<T> op, const T* __restrict__ x,
                                                        cplx_t<T>* __restrict__ y, T* __restrict__ mod) {
  conv_fwd_split_body<T, NI, LP>(op, x, y, mod, blockIdx.x, blockIdx.y);
}

template <typename T, int NI, int LP>
__device__ __forceinline__ void conv_adj_split_body(const ConvOp<T>& op, const cplx_t<T>* __restrict__ gy,
                                                    T* __restrict__ gx, int64_t bx, int row) {
  const int64_t o = (bx * 256 + threadIdx.x) / LP;
  const int lane = int(threadIdx.x % LP);
  const bool active = o < 12 * op.npf;
  const bool mir = active && (o >> 2) >= op.R;
  const int64_t q = active ? (mir ? (o >> 2) - op.R : o >> 2) : 0;
  const int k = int(o & 3);
  const int64_t N = 12 * op.npf;
  const cplx_t<T>* gr = gy + size_t(row) * kL * N * NI;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
  const int cnt = active ? op.tcnt[q] : 0;
  for (int s = lane; s < cnt; s += LP) {
    size_t at = size_t(s) * op.R + q;
    int64_t pk = rot_pix(op.trow[at], k, op.order);
    if (mir) pk = mirror_pix(pk, op.order);
    T c[8], a[8];
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      cplx_t<T> v = reinterpret_cast<const cplx_t<T>*>(op.tcoef)[(size_t(g) * op.Tn + s) * op.R + q];
      c[2 * g] = v.x;
      c[2 * g + 1] = mir ? -v.y : v.y;
    }
    permute<T>(k, c, a);
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      const cplx_t<T>* src = gr + (size_t(g) * N + pk) * NI;
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        cplx_t<T> v = src[j];
        acc[j] = fma(a[g * 2 + 1], v.y, fma(a[g * 2], v.x, acc[j]));
      }
    }
  }
#pragma unroll
  for (int m = LP / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += __shfl_xor(acc[j], m);
  if (!active || lane != 0) return;
  int64_t qk = rot_pix(rep_pixel(q, op.order), k, op.order);
  if (mir) qk = mirror_pix(qk, op.order);
#pragma unroll
  for (int j = 0; j < NI; ++j) gx[(size_t(row) * N + qk) * NI + j] = acc[j];
}

template <typename T, int NI, int LP>
__global__ __launch_bounds__(256) void k_conv_adj_split(ConvOp<T> op, const cplx_t<T>* __restrict__ gy,
                                                        T* __restrict__ gx) {
  conv_adj_split_body<T, NI, LP>(op, gy, gx, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// LDS-tiled convolution (levels with nside >= 8).  A workgroup is one TS x TS
// tile of one face (an aligned range of TS^2 NESTED ids of a stored face, its
// shift k and mirror flag); the inputs of the tile plus a halo of kWin are
// staged in LDS from the face's extended frame, and every slot reads LDS at
// its stored (dx, dy) offset.  A mirrored tile stages, at every position, the
// mirror of the pixel the north tile stages there, so the same offsets read
// the mirrored cells.
// ---------------------------------------------------------------------------

// The even bits of v packed (local NESTED index -> x; of v >> 1 -> y), and the
// inverse spread, on 32-bit words: local indices of nside <= 8192 faces (the
// engine's limit) fit.
__device__ inline uint32_t compress_bits(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0F0F0F0Fu;
  v = (v | (v >> 4)) & 0x00FF00FFu;
  return (v | (v >> 8)) & 0x0000FFFFu;
}

__device__ inline uint32_t spread_bits(uint32_t v) {
  v &= 0x0000FFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  return (v | (v << 1)) & 0x55555555u;
}

// HEALPix C++ neighbour faces / swaps by direction 4 + dx + 3 dy (dia.rs:27-41).
static __constant__ const int8_t kNbFaceT[9][12] = {
    {8, 9, 10, 11, -1, -1, -1, -1, 10, 11, 8, 9},
    {5, 6, 7, 4, 8, 9, 10, 11, 9, 10, 11, 8},
    {-1, -1, -1, -1, 5, 6, 7, 4, -1, -1, -1, -1},
    {4, 5, 6, 7, 11, 8, 9, 10, 11, 8, 9, 10},
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11},
    {1, 2, 3, 0, 0, 1, 2, 3, 5, 6, 7, 4},
    {-1, -1, -1, -1, 7, 4, 5, 6, -1, -1, -1, -1},
    {3, 0, 1, 2, 3, 0, 1, 2, 4, 5, 6, 7},
    {2, 3, 0, 1, -1, -1, -1, -1, 0, 1, 2, 3},
};
static __constant__ const uint8_t kNbSwapT[9][3] = {{0, 0, 3}, {0, 0, 6}, {0, 0, 0}, {0, 0, 5}, {0, 0, 0},
                                                     {5, 0, 0}, {0, 0, 0}, {6, 0, 0}, {3, 0, 0}};

// Pixel at extended coordinates of a face (-nside <= ex, ey < 2 nside), -1 in the
// quadrant missing at a vertex of three faces.
__device__ inline int64_t ext_pixel(int face, int ex, int ey, int order) {
  int nside = 1 << order;
  int x = ex, y = ey, nb = 4;
  if (x < 0) { x += nside; nb -= 1; }
  else if (x >= nside) { x -= nside; nb += 1; }
  if (y < 0) { y += nside; nb -= 3; }
  else if (y >= nside) { y -= nside; nb += 3; }
  int f = kNbFaceT[nb][face];
  if (f < 0) return -1;
  uint8_t bits = kNbSwapT[nb][face >> 2];
  if (bits & 1) x = nside - x - 1;
  if (bits & 2) y = nside - y - 1;
  if (bits & 4) { int t = x; x = y; y = t; }
  return (int64_t(f) << (2 * order)) + int64_t(spread_bits(uint32_t(x)) | (spread_bits(uint32_t(y)) << 1));
}

struct TileInfo {
  int face, fr, k, x0, y0;  // face = 4 fr + k: the north / equatorial face staged
  bool m;                   // mirrored (outputs in face 8 + k)
  int64_t base;             // local NESTED index of the first pixel of the tile
};

// Workgroup -> (rep tile, face shift k, row).  The 4 * rows workgroups of one
// rep tile read the same coefficients; workgroups are dispatched round-robin
// over the 8 XCDs (each with its own L2), so they are numbered to land on the
// same XCD: in each group of 8 * ways consecutive workgroups, workgroup b runs
// tile group * 8 + b % 8 and way (b / 8) % ways.
constexpr int kXcds = 8;

// Tile unit u of 3 ntile -> (stored face fr, tile t, mirror m): units
// 2 ntile .. 3 ntile are the equatorial tiles; below, a north tile and its
// mirror (the same coefficients) are kXcds units apart in groups of 2 kXcds,
// so that their workgroups run back to back on one XCD.
__device__ inline void unit_tile(int64_t u, int64_t ntile, int& fr, int64_t& t, bool& m) {
  if (u >= 2 * ntile) {
    fr = 1;
    t = u - 2 * ntile;
    m = false;
    return;
  }
  fr = 0;
  if (ntile % kXcds == 0) {
    t = (u / (2 * kXcds)) * kXcds + u % kXcds;
    m = ((u / kXcds) & 1) != 0;
  } else {
    t = u >> 1;
    m = (u & 1) != 0;
  }
}

// Pixel staged at extended coordinates (ex, ey) of a tile: face's pixel, or
// its mirror.
__device__ inline int64_t staged_pixel(int face, bool m, int ex, int ey, int order) {
  int64_t p = ext_pixel(face, ex, ey, order);
  return (m && p >= 0) ? mirror_pix(p, order) : p;
}

// staged_pixel of a position inside the face (0 <= ex, ey < nside): the
// NESTED index without the neighbour-face cases, on 32-bit words.
__device__ inline int64_t inner_pixel(int face, bool m, int ex, int ey, int order) {
  const uint32_t loc = spread_bits(uint32_t(ex)) | (spread_bits(uint32_t(ey)) << 1);
  if (!m) return (int64_t(face) << (2 * order)) + loc;
  const uint32_t v = loc ^ ((1u << (2 * order)) - 1u);
  return (int64_t(mirror_face(face)) << (2 * order)) + (((v & 0x55555555u) << 1) | ((v >> 1) & 0x55555555u));
}

// Whether a tile's window (kWin around the TS x TS tile at (x0, y0)) lies
// inside its face: then every staged pixel is an inner_pixel.
template <int TS>
__device__ inline bool window_inside(int x0, int y0, int order) {
  const int nside = 1 << order;
  return x0 >= kWin && y0 >= kWin && x0 + TS + kWin <= nside && y0 + TS + kWin <= nside;
}

// Pixel staged at a window position: the inner form when the whole window is
// inside the face (a workgroup-uniform choice).
__device__ inline int64_t window_pixel(bool inside, int face, bool m, int ex, int ey, int order) {
  return inside ? inner_pixel(face, m, ex, ey, order) : staged_pixel(face, m, ex, ey, order);
}

// Staging position pi of a tile's window (SW x SW) -> its LDS index.  A
// mirrored tile walks the window's columns (iy fastest): they are the rows of
// the mirrored face, so that consecutive lanes read consecutive pixels along x
// there, like the north tiles.
template <int SW>
__device__ inline int stage_index(bool m, int pi) {
  return m ? (pi % SW) * SW + pi / SW : pi;
}

// Lane li of a tile writes output pixel tile_out + li.  In a mirrored tile it
// is the mirror of rep-local pixel base + rep_lane(li) (the mirror of base + l
// is the mirrored tile's base plus the mirror of l within the tile), so the
// lane reads that pixel's row and LDS position, and the writes stay
// consecutive.
template <int TS>
__device__ inline int rep_lane(bool m, int li) {
  constexpr int kTileOrder = TS == 32 ? 5 : (TS == 16 ? 4 : 3);
  static_assert((1 << kTileOrder) == TS, "tile side");
  return m ? int(mirror_loc(li, kTileOrder)) : li;
}

template <int TS>
__device__ inline int64_t tile_out(int face, bool m, int64_t base, int order) {
  const int64_t npf = int64_t(1) << (2 * order);
  return m ? int64_t(face + 8) * npf + (mirror_loc(base, order) & ~int64_t(TS * TS - 1))
           : int64_t(face) * npf + base;
}

// MIRROR: tile units of the convolution (unit_tile); otherwise unit u is tile
// u % ntile of the stored face u / ntile (decimation: faces 0, 4, 8).
template <int TS, bool MIRROR>
__device__ inline TileInfo tile_info(int order, int rows, int& row) {
  TileInfo t;
  const int64_t npf = int64_t(1) << (2 * order);
  const int64_t ntile = npf / (TS * TS);
  const int64_t tiles = 3 * ntile;
  const int ways = 4 * rows;
  const int64_t full = (tiles / kXcds) * kXcds * ways;
  const int64_t b = blockIdx.x;
  int64_t tr;
  int j;
  if (b < full) {
    int64_t rem = b % (int64_t(kXcds) * ways);
    tr = (b / (int64_t(kXcds) * ways)) * kXcds + rem % kXcds;
    j = int(rem / kXcds);
  } else {
    int64_t rem = b - full;
    tr = (tiles / kXcds) * kXcds + rem / ways;
    j = int(rem % ways);
  }
  t.k = j & 3;
  row = j >> 2;
  int64_t tile;
  if (MIRROR) {
    unit_tile(tr, ntile, t.fr, tile, t.m);
  } else {
    t.fr = int(tr / ntile);
    tile = tr % ntile;
    t.m = false;
  }
  t.base = tile * (TS * TS);
  t.face = 4 * t.fr + t.k;
  t.x0 = int(compress_bits(uint32_t(t.base)));
  t.y0 = int(compress_bits(uint32_t(t.base >> 1)));
  return t;
}

// Union slots of the wave's block of 64 rep rows or cells (tile rows fr,
// tile base, lane's row or cell rl): the block's window offsets in the linear
// form of the TS tiles' windows, the same in every lane, read with scalar
// loads, and the block's slot count.
struct BlockSlots {
  const int32_t* off;
  int cnt;
};
__device__ inline int64_t wave_block(int64_t npf, int fr, int64_t base, int rl) {
  return (int64_t(fr) * npf + base) / 64 + (__builtin_amdgcn_readfirstlane(rl) >> 6);
}
// Transposed slots (the adjoints).
template <int TS, typename T>
__device__ inline BlockSlots block_slots(const ConvOp<T>& op, int fr, int64_t base, int rl) {
  const int64_t blk = wave_block(op.npf, fr, base, rl);
  const int32_t* lin = TS == 32 ? op.tbo32 : (TS == 16 ? op.tbo16 : op.tbo8);
  return BlockSlots{lin + blk * op.Tn, op.tbcnt[blk]};
}
// Forward slots.
template <int TS, typename T>
__device__ inline BlockSlots fwd_block_slots(const ConvOp<T>& op, int fr, int64_t base, int rl) {
  const int64_t blk = wave_block(op.npf, fr, base, rl);
  const int32_t* lin = TS == 32 ? op.fbo32 : (TS == 16 ? op.fbo16 : op.fbo8);
  return BlockSlots{lin + blk * op.S, op.fbcnt[blk]};
}

// A 32 x 32 tile is a 16-wave workgroup: two of them per CU need 8 waves per
// SIMD, i.e. at most 64 VGPRs; at 7 waves only one fits.
template <typename T, int TS>
constexpr int tile32_waves() { return TS == 32 && sizeof(T) == 4 ? 8 : 1; }

// Stage COUNT window positions (position -> pixel by `pix`, -1 outside the
// sphere's extension, and its LDS index) into LDS planes of PLANE elements,
// NI planes per position.  The thread's pixel ids come first (outside a
// face's interior they read the neighbour tables, whose waits would otherwise
// also wait for data loads in flight), then every data load is issued
// (missing pixels read pixel 0 of the row and store 0) before the first LDS
// write, so their latencies overlap.
template <typename T, int NI, int COUNT, int PLANE, int NT, typename Pix>
__device__ __forceinline__ void stage_window(const T* __restrict__ xr, T* xs, Pix pix) {
  constexpr int NIT = (COUNT + NT - 1) / NT;
  int64_t px[NIT];
  int at[NIT];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const int idx = int(threadIdx.x) + it * NT;
    const bool in = idx < COUNT;
    int lp = 0;
    const int64_t p = pix(in ? idx : 0, lp);
    px[it] = in ? p : -1;
    at[it] = in ? lp : -1;
  }
  T v[NIT][NI];
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    const size_t src = size_t(px[it] >= 0 ? px[it] : 0) * NI;
#pragma unroll
    for (int j = 0; j < NI; ++j) v[it][j] = xr[src + j];
  }
#pragma unroll
  for (int it = 0; it < NIT; ++it)
    if (at[it] >= 0)
#pragma unroll
      for (int j = 0; j < NI; ++j) xs[j * PLANE + at[it]] = px[it] >= 0 ? v[it][j] : T(0);
}

template <typename T, int NI, int TS>
__global__ __launch_bounds__(TS * TS) __attribute__((amdgpu_waves_per_eu(tile32_waves<T, TS>())))
void k_conv_fwd_tiled(ConvOp<T> op, const T* __restrict__ x, int rows,
                                                           cplx_t<T>* __restrict__ y, T* __restrict__ mod) {
  constexpr int SW = TS + 2 * kWin, SA = SW * SW;
  // Planar [j][pixel]: consecutive pixels in consecutive banks.
  __shared__ T xs[NI * SA];
  int row;
  const TileInfo ti = tile_info<TS, true>(op.order, rows, row);
  const bool inside = window_inside<TS>(ti.x0, ti.y0, op.order);
  const int64_t N = 12 * op.npf;
  const T* xr = x + size_t(row) * N * NI;
  for (int i = threadIdx.x; i < SA; i += TS * TS) {
    const int idx = stage_index<SW>(ti.m, i);
    int64_t p = window_pixel(inside, ti.face, ti.m, ti.x0 - kWin + idx % SW, ti.y0 - kWin + idx / SW, op.order);
#pragma unroll
    for (int j = 0; j < NI; ++j) xs[j * SA + idx] = p >= 0 ? xr[size_t(p) * NI + j] : T(0);
  }
  __syncthreads();
  const int li = threadIdx.x, rl = rep_lane<TS>(ti.m, li);
  const int pos = (int(compress_bits(uint32_t(rl >> 1))) + kWin) * SW + int(compress_bits(uint32_t(rl))) + kWin;
  const int64_t r = int64_t(ti.fr) * op.npf + ti.base + rl;
  // FP32: rep-frame sums, permuted to the face's gauges at the end (the same
  // products and additions); FP64 permutes each slot's coefficients (faster
  // there: more loads in flight).
  constexpr bool kRepSums = sizeof(T) == 4;
  T acc[NI][kL][2];
#pragma unroll
  for (int j = 0; j < NI; ++j)
#pragma unroll
    for (int g = 0; g < kL; ++g) acc[j][g][0] = acc[j][g][1] = T(0);
  const BlockSlots bs = fwd_block_slots<TS>(op, ti.fr, ti.base, rl);
#pragma unroll 4
  for (int s = 0; s < bs.cnt; ++s) {
    size_t at = size_t(s) * op.R + r;
    int src = pos + bs.off[s];
    T c[8], a[8];
    Load<T>::coef8(op.coef + at * 8, c);
    if constexpr (kRepSums) {
#pragma unroll
      for (int i = 0; i < 8; ++i) a[i] = c[i];
    } else {
      permute<T>(ti.k, c, a);
    }
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      T xv = xs[j * SA + src];
#pragma unroll
      for (int g = 0; g < kL; ++g) {
        acc[j][g][0] += a[g * 2] * xv;
        acc[j][g][1] += a[g * 2 + 1] * xv;
      }
    }
  }
  const int64_t p = tile_out<TS>(ti.face, ti.m, ti.base, op.order) + li;
#pragma unroll
  for (int j = 0; j < NI; ++j) {
    T sum[8], out[8];
#pragma unroll
    for (int g = 0; g < kL; ++g) { sum[g * 2] = acc[j][g][0]; sum[g * 2 + 1] = ti.m ? -acc[j][g][1] : acc[j][g][1]; }
    if constexpr (kRepSums) {
      permute<T>(ti.k, sum, out);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) out[i] = sum[i];
    }
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      y[((size_t(row) * kL + g) * N + p) * NI + j] = make_c<T>(out[g * 2], out[g * 2 + 1]);
      if (mod != nullptr && NI == 1) mod[(size_t(row) * N + p) * kL + g] = hypot(out[g * 2], out[g * 2 + 1]);
    }
  }
}

// Workgroup -> (rep tile, face group of FW faces, row) for the multi-face
// kernels: the 4 / FW * rows workgroups of one rep tile are numbered to land
// on one XCD as in tile_info.
template <int TS, int FW>
__device__ inline void tile_group(int order, int rows, int64_t& tr, int& fg, int& row) {
  // tr: tile unit (unit_tile).
  const int64_t npf = int64_t(1) << (2 * order);
  const int64_t tiles = 3 * (npf / (TS * TS));
  const int ways = (4 / FW) * rows;
  const int64_t full = (tiles / kXcds) * kXcds * ways;
  const int64_t b = blockIdx.x;
  int j;
  if (b < full) {
    int64_t rem = b % (int64_t(kXcds) * ways);
    tr = (b / (int64_t(kXcds) * ways)) * kXcds + rem % kXcds;
    j = int(rem / kXcds);
  } else {
    int64_t rem = b - full;
    tr = (tiles / kXcds) * kXcds + rem / ways;
    j = int(rem % ways);
  }
  fg = j % (4 / FW);
  row = j / (4 / FW);
}

// Forward convolution of FW rotated faces of one rep tile per workgroup:
// every slot's coefficients are loaded once and applied to the FW faces'
// staged inputs (faces k = fg * FW + f).
template <typename T, int NI, int TS, int FW, int kSlotUnroll = TS == 32 ? 4 : 2>
__global__ __launch_bounds__(TS * TS) void k_conv_fwd_faces(ConvOp<T> op, const T* __restrict__ x, int rows,
                                                           cplx_t<T>* __restrict__ y, T* __restrict__ mod) {
  constexpr int SW = TS + 2 * kWin, SA = SW * SW;
  // Planar [face][j][pixel].
  __shared__ T xs[FW * NI * SA];
  int64_t tr;
  int fg, row;
  tile_group<TS, FW>(op.order, rows, tr, fg, row);
  const int64_t ntile = op.npf / (TS * TS);
  int fr;
  int64_t tile;
  bool m;
  unit_tile(tr, ntile, fr, tile, m);
  const int64_t base = tile * (TS * TS);
  const int x0 = int(compress_bits(uint32_t(base))), y0 = int(compress_bits(uint32_t(base >> 1)));
  const bool inside = window_inside<TS>(x0, y0, op.order);
  const int64_t N = 12 * op.npf;
  const T* xr = x + size_t(row) * N * NI;
  for (int idx = threadIdx.x; idx < FW * SA; idx += TS * TS) {
    const int f = idx / SA, pi = stage_index<SW>(m, idx % SA);
    const int face = 4 * fr + fg * FW + f;
    int64_t p = window_pixel(inside, face, m, x0 - kWin + pi % SW, y0 - kWin + pi / SW, op.order);
#pragma unroll
    for (int j = 0; j < NI; ++j) xs[(f * NI + j) * SA + pi] = p >= 0 ? xr[size_t(p) * NI + j] : T(0);
  }
  __syncthreads();
  const int li = threadIdx.x, rl = rep_lane<TS>(m, li);
  const int pos = (int(compress_bits(uint32_t(rl >> 1))) + kWin) * SW + int(compress_bits(uint32_t(rl))) + kWin;
  const int64_t r = int64_t(fr) * op.npf + base + rl;
  T acc[FW][NI][kL][2];
#pragma unroll
  for (int f = 0; f < FW; ++f)
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
      for (int g = 0; g < kL; ++g) acc[f][j][g][0] = acc[f][j][g][1] = T(0);
  const BlockSlots bs = fwd_block_slots<TS>(op, fr, base, rl);
  // The mirror's conjugation, on the coefficients: 4 products per slot for
  // all the faces (at the outputs it would hold more registers across the loop).
  const T sg = m ? T(-1) : T(1);
#pragma unroll kSlotUnroll
  for (int s = 0; s < bs.cnt; ++s) {
    size_t at = size_t(s) * op.R + r;
    int src = pos + bs.off[s];
    T c[8];
    Load<T>::coef8(op.coef + at * 8, c);
#pragma unroll
    for (int g = 0; g < kL; ++g) c[2 * g + 1] *= sg;
#pragma unroll
    for (int f = 0; f < FW; ++f) {
      T a[8];
      // FW = 4: one face group, the shift is the (unrolled) face index.
      permute<T>(FW == 4 ? f : fg * FW + f, c, a);
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        T xv = xs[(f * NI + j) * SA + src];
#pragma unroll
        for (int g = 0; g < kL; ++g) {
          acc[f][j][g][0] += a[g * 2] * xv;
          acc[f][j][g][1] += a[g * 2 + 1] * xv;
        }
      }
    }
  }
#pragma unroll
  for (int f = 0; f < FW; ++f) {
    const int64_t p = tile_out<TS>(4 * fr + fg * FW + f, m, base, op.order) + li;
#pragma unroll
    for (int g = 0; g < kL; ++g)
#pragma unroll
      for (int j = 0; j < NI; ++j)
        y[((size_t(row) * kL + g) * N + p) * NI + j] = make_c<T>(acc[f][j][g][0], acc[f][j][g][1]);
    if (mod != nullptr && NI == 1) {
#pragma unroll
      for (int g = 0; g < kL; ++g) mod[(size_t(row) * N + p) * kL + g] = hypot(acc[f][0][g][0], acc[f][0][g][1]);
    }
  }
}

// Element at byte offset lane_bytes of a [planes][R] array's plane: a uniform
// plane base plus a loop-invariant 32-bit lane offset.
template <typename E>
__device__ inline E plane_load(const E* plane, uint32_t lane_bytes) {
  return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(plane) + lane_bytes);
}

// The NP planes (SA elements apart) of an adjoint's LDS window at element
// offset src: planes 0 .. NP/2 - 1 from one base and the rest from a second,
// NP/2 planes up, when the whole stack is beyond the 64 KB reach of an LDS
// read's immediate offset (one address add per base per slot instead of one
// per far plane).  The second base is made opaque so that the compiler does
// not fold it back into the first.
template <int NP, int SA, typename E>
struct WindowPlanes {
  static constexpr bool kSplit = size_t(NP) * SA * sizeof(E) > 65536;
  int hi_pos;
  __device__ explicit WindowPlanes(int pos) : hi_pos(pos + (kSplit ? NP / 2 * SA : 0)) {
    // Not volatile: a statement with side effects would count as a memory
    // write and turn the scalar loads of the slot offsets into vector loads.
    if constexpr (kSplit) asm("" : "+v"(hi_pos));
  }
  template <typename F>
  __device__ void read(const E* w, int pos, int off, F&& out) const {
    const E* lo = w + pos + off;
    if constexpr (kSplit) {
      const E* hi = w + hi_pos + off;
#pragma unroll
      for (int p = 0; p < NP; ++p) out(p, p < NP / 2 ? lo[p * SA] : hi[(p - NP / 2) * SA]);
    } else {
#pragma unroll
      for (int p = 0; p < NP; ++p) out(p, lo[p * SA]);
    }
  }
};

// The slot loop of the tiled adjoints, software-pipelined in chunks of U
// slots: the next chunk's NC coefficients per slot (load(s0, u, k) for slot
// s0 + u, vector loads) are issued before the current chunk's products run (body(offset,
// coefficients)), so a wave keeps a chunk of loads in flight behind its
// arithmetic.  The window offsets (scalar loads) are read at the start of
// their own chunk: a scalar load in flight would make every LDS wait of the
// chunk a full wait.  The union slot capacity cap is a multiple of 4 with
// zero coefficients past a block's count, so whole chunks run.  The
// prefetches are unconditional (slots clamped to the capacity, the last
// one unused) and each sits in the same block as the products that use it:
// one issued on some paths only would make the compiler wait for every load
// before the current chunk's products, and one followed by a branch would be
// moved past it, next to its use.
template <int U, int NC, typename C, typename Load, typename Body>
__device__ __forceinline__ void pipelined_slots(const int32_t* __restrict__ off, int cnt, int cap, Load load,
                                                Body body) {
  const int nch = (cnt + U - 1) / U;
  C a0[U][NC], a1[U][NC];
  auto fetch = [&](C (&a)[U][NC], int s0) {
    // A chunk lies wholly below the capacity or wholly past the count.
    const int s0c = min(s0, cap - U);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < NC; ++k) a[u][k] = load(s0c, u, k);
  };
  auto run = [&](const C (&a)[U][NC], int s0) {
    int32_t o[U];
#pragma unroll
    for (int u = 0; u < U; ++u) o[u] = off[s0 + u];
#pragma unroll
    for (int u = 0; u < U; ++u) body(o[u], a[u]);
  };
  if (nch == 0) return;
  fetch(a0, 0);
  int c = 0;
  for (; c + 1 < nch; c += 2) {
    fetch(a1, (c + 1) * U);
    run(a0, c * U);
    fetch(a0, (c + 2) * U);
    run(a1, (c + 1) * U);
  }
  if (c < nch) run(a0, c * U);
}

// Window positions staged per batch by the adjoints (every load of a batch
// is in flight before its LDS writes): values of at most 48 VGPRs with a
// budget (512 / waves per SIMD) of 128 or more, budget / 8 below.
template <typename E, int PER, int VALS, int BUDGET>
constexpr int stage_batch() {
  constexpr int regs = VALS * int(sizeof(E)) / 4;
  constexpr int cap = BUDGET >= 128 ? 48 : BUDGET / 8;
  constexpr int ch = regs >= cap ? 1 : cap / regs;
  return ch < PER ? ch : PER;
}

// Staging order of an adjoint's window (SW x SW positions per face): blocks
// of 8 x 4 positions, x fastest, one block per 32 consecutive staging slots,
// so that 32 lanes write 4 rows of 8 LDS words (conflict-free at the pitches
// of lds_pitch) and read two aligned 4 x 4 NESTED blocks of the face, in a
// mirrored tile as well.  Slots past the window's edge stage nothing.
template <int SW>
struct StageGrid {
  static constexpr int kBx = (SW + 7) / 8, kBy = (SW + 3) / 4;
  static constexpr int kSlots = kBx * kBy * 32;  // per face
  // Window position (wx, wy) of slot k of a face; false past the edge.
  __device__ static bool at(int k, int& wx, int& wy) {
    const int b = k >> 5, t = k & 31;
    wx = (b % kBx) * 8 + (t & 7);
    wy = (b / kBx) * 4 + (t >> 3);
    return wx < SW && wy < SW;
  }
};

// GP gauges of the cotangent staged per pass.
// RW rows (input maps) per workgroup share each slot's offset and
// coefficient loads (rows a multiple of RW).
template <int RW>
constexpr int adj_rows_waves() { return RW == 2 ? 4 : 1; }
template <typename T, int NI, int TS, int GP, int RW = 1>
__global__ __launch_bounds__(TS * TS) __attribute__((amdgpu_waves_per_eu(
    RW == 1 ? tile32_waves<T, TS>() : adj_rows_waves<RW>())))
void k_conv_adj_tiled(ConvOp<T> op, const cplx_t<T>* __restrict__ gy, int rows,
                                                           T* __restrict__ gx) {
  // Window of SW x SW positions, planes of SW rows of pitch PW.
  constexpr int SW = TS + 2 * kWin, PW = lds_pitch(TS, sizeof(cplx_t<T>)), SA = SW * PW;
  // Planar [r][gg][j][pixel]: consecutive pixels in consecutive banks.
  __shared__ cplx_t<T> gs[RW * GP * NI * SA];
  int rowg;
  const TileInfo ti = tile_info<TS, true>(op.order, rows / RW, rowg);
  const bool inside = window_inside<TS>(ti.x0, ti.y0, op.order);
  const int row = rowg * RW;
  const int64_t N = 12 * op.npf;
  const cplx_t<T>* gr = gy + size_t(row) * kL * N * NI;
  const int li = threadIdx.x, rl = rep_lane<TS>(ti.m, li);
  const int pos = (int(compress_bits(uint32_t(rl >> 1))) + kWin) * PW + int(compress_bits(uint32_t(rl))) + kWin;
  const int64_t q = int64_t(ti.fr) * op.npf + ti.base + rl;
  const BlockSlots bs = block_slots<TS>(op, ti.fr, ti.base, rl);
  const WindowPlanes<RW * GP * NI, SA, cplx_t<T>> wp(pos);
  const cplx_t<T>* tc = reinterpret_cast<const cplx_t<T>*>(op.tcoef);
  const uint32_t qoc = uint32_t(q) * uint32_t(sizeof(cplx_t<T>));
  // Pixels (or -1; -2 past the window) of the window positions this thread
  // stages, the same in every pass; their LDS indexes are recomputed (cheaper
  // than the registers).
  using Grid = StageGrid<SW>;
  constexpr int kPer = (Grid::kSlots + TS * TS - 1) / (TS * TS);
  int32_t spix[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = threadIdx.x + i * TS * TS;
    int wx, wy;
    const bool in = idx < Grid::kSlots && Grid::at(idx, wx, wy);
    spix[i] = in ? int32_t(window_pixel(inside, ti.face, ti.m, ti.x0 - kWin + wx, ti.y0 - kWin + wy, op.order)) : -2;
  }
  T acc[RW][NI];
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[r][j] = T(0);
  for (int pass = 0; pass < kL / GP; ++pass) {
    // A face of odd shift k reads gauge g from coefficient plane g ^ 2: it
    // walks its gauges in that order, so that the 4 faces of a rep tile
    // (consecutive workgroups on one XCD) read the same planes in the same
    // pass and share them in L2.
    const int g0 = GP == kL ? 0 : ((pass * GP) ^ (2 * (ti.k & 1)));
    __syncthreads();
    // In batches of CH positions: every load of a batch is issued (missing
    // pixels read pixel 0 and store 0) before its LDS writes.
    constexpr int kWaves = RW == 1 ? tile32_waves<T, TS>() : adj_rows_waves<RW>();
    constexpr int CH = stage_batch<cplx_t<T>, kPer, GP * RW * NI, 512 / kWaves>();
#pragma unroll
    for (int i0 = 0; i0 < kPer; i0 += CH) {
      cplx_t<T> sv[CH][GP][RW][NI];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        if (i0 + c >= kPer) break;
        const int64_t p = spix[i0 + c] >= 0 ? spix[i0 + c] : 0;
#pragma unroll
        for (int gg = 0; gg < GP; ++gg)
#pragma unroll
          for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int j = 0; j < NI; ++j) sv[c][gg][r][j] = gr[((size_t(r) * kL + g0 + gg) * N + p) * NI + j];
      }
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const int i = i0 + c;
        if (i >= kPer) break;
        if (spix[i] == -2) continue;
        const bool ok = spix[i] >= 0;
        int wx, wy;
        Grid::at(threadIdx.x + i * TS * TS, wx, wy);
        const int pi = wy * PW + wx;
#pragma unroll
        for (int gg = 0; gg < GP; ++gg) {
          // The face shift's (and the mirror's) conjugation of a gauge is
          // applied to the staged cotangent: a.y (-v.y) instead of (-a.y) v.y,
          // the same products.
          const bool cj = gauge_conj(ti.k, g0 + gg) != ti.m;
#pragma unroll
          for (int r = 0; r < RW; ++r)
#pragma unroll
            for (int j = 0; j < NI; ++j) {
              cplx_t<T> v = ok ? sv[c][gg][r][j] : make_c<T>(T(0), T(0));
              if (cj) v.y = -v.y;
              gs[((r * GP + gg) * NI + j) * SA + pi] = v;
            }
        }
      }
    }
    __syncthreads();
    // Chunks of 2 slots when every slot loads 4 gauges' coefficients.
    pipelined_slots<GP >= 4 ? 2 : 4, GP, cplx_t<T>>(
        bs.off, bs.cnt, op.Tn,
        [&](int s0, int u, int gg) {
          return plane_load(tc + (size_t(gauge_src(ti.k, g0 + gg)) * op.Tn + size_t(s0)) * op.R + size_t(u) * op.R,
                            qoc);
        },
        [&](int32_t o, const cplx_t<T>(&a)[GP]) {
          // Planes [r][gg][j].
          cplx_t<T> v[RW * GP * NI];
          wp.read(gs, pos, o, [&](int p, cplx_t<T> x) { v[p] = x; });
#pragma unroll
          for (int gg = 0; gg < GP; ++gg)
#pragma unroll
            for (int r = 0; r < RW; ++r)
#pragma unroll
              for (int j = 0; j < NI; ++j) {
                const cplx_t<T> x = v[(r * GP + gg) * NI + j];
                acc[r][j] = fma(a[gg].y, x.y, fma(a[gg].x, x.x, acc[r][j]));
              }
        });
  }
  const int64_t p = tile_out<TS>(ti.face, ti.m, ti.base, op.order) + li;
#pragma unroll
  for (int r = 0; r < RW; ++r)
#pragma unroll
    for (int j = 0; j < NI; ++j) gx[(size_t(row + r) * N + p) * NI + j] = acc[r][j];
}

// Adjoint convolution of FW rotated faces of one rep tile per workgroup.  In
// pass P every face reads coefficient planes P * GP .. P * GP + GP - 1 (face k
// takes gauge plane ^ 2 (k & 1) from them, conjugated by its shift), so each
// plane entry is loaded once for the FW faces.
// Slot loop unrolled 4 times on 32 x 32 tiles (A/B on MI355X: deeper
// unrolling, or pinning the 4-face z adjoint to 64 VGPRs for two workgroups
// per CU, was slower), twice on smaller tiles.
// Waves per SIMD that the LDS window of an adjoint leaves room for (160 KB
// per CU): the register budget the scheduler may use.
template <typename T, int NI, int TS, int FW, int GP>
constexpr int adj_lds_waves() {
  constexpr int SW = TS + 2 * kWin;
  constexpr int bytes = FW * GP * NI * SW * lds_pitch(TS, sizeof(cplx_t<T>)) * int(sizeof(cplx_t<T>));
  constexpr int wgs = (160 * 1024) / bytes;
  constexpr int waves = wgs * TS * TS / 256;
  return waves < 1 ? 1 : (waves > 8 ? 8 : waves);
}
template <typename T, int NI, int TS, int FW, int GP, int kSlotUnroll = TS == 32 ? 4 : 2, int PF = 0>
__global__ __launch_bounds__(TS * TS) __attribute__((amdgpu_waves_per_eu(adj_lds_waves<T, NI, TS, FW, GP>(),
                                                                         adj_lds_waves<T, NI, TS, FW, GP>())))
void k_conv_adj_faces(ConvOp<T> op, const cplx_t<T>* __restrict__ gy,
                                                           int rows, T* __restrict__ gx) {
  // Window of SW x SW positions, planes of SW rows of pitch PW.
  constexpr int SW = TS + 2 * kWin, PW = lds_pitch(TS, sizeof(cplx_t<T>)), SA = SW * PW;
  // Planar [face][gg][j][pixel].
  __shared__ cplx_t<T> gs[FW * GP * NI * SA];
  int64_t tr;
  int fg, row;
  tile_group<TS, FW>(op.order, rows, tr, fg, row);
  const int64_t ntile = op.npf / (TS * TS);
  int fr;
  int64_t tile;
  bool m;
  unit_tile(tr, ntile, fr, tile, m);
  const int64_t base = tile * (TS * TS);
  const int x0 = int(compress_bits(uint32_t(base))), y0 = int(compress_bits(uint32_t(base >> 1)));
  const bool inside = window_inside<TS>(x0, y0, op.order);
  const int64_t N = 12 * op.npf;
  const cplx_t<T>* gr = gy + size_t(row) * kL * N * NI;
  const int li = threadIdx.x, rl = rep_lane<TS>(m, li);
  const int pos = (int(compress_bits(uint32_t(rl >> 1))) + kWin) * PW + int(compress_bits(uint32_t(rl))) + kWin;
  const int64_t q = int64_t(fr) * op.npf + base + rl;
  const BlockSlots bs = block_slots<TS>(op, fr, base, rl);
  const WindowPlanes<FW * GP * NI, SA, cplx_t<T>> wp(pos);
  const cplx_t<T>* tc = reinterpret_cast<const cplx_t<T>*>(op.tcoef);
  const uint32_t qoc = uint32_t(q) * uint32_t(sizeof(cplx_t<T>));
  // Pixels (or -1; -2 past the window) of the window positions (face f,
  // position) this thread stages, the same in every pass.
  using Grid = StageGrid<SW>;
  constexpr int kPer = (FW * Grid::kSlots + TS * TS - 1) / (TS * TS);
  int32_t spix[kPer];
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int idx = threadIdx.x + i * TS * TS;
    const int f = idx / Grid::kSlots;
    int wx, wy;
    const bool in = idx < FW * Grid::kSlots && Grid::at(idx % Grid::kSlots, wx, wy);
    spix[i] = in ? int32_t(window_pixel(inside, 4 * fr + fg * FW + f, m, x0 - kWin + wx, y0 - kWin + wy, op.order))
                 : -2;
  }
  T acc[FW][NI];
#pragma unroll
  for (int f = 0; f < FW; ++f)
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[f][j] = T(0);
  // Cotangent values of window position i (missing pixels read pixel 0) in
  // pass `pass`, and their LDS writes (0 for missing pixels).
  auto stage_load = [&](int i, int pass, cplx_t<T> (&sv)[GP][NI]) {
    const int64_t p = spix[i] >= 0 ? spix[i] : 0;
    const int k = fg * FW + (threadIdx.x + i * TS * TS) / Grid::kSlots;
#pragma unroll
    for (int gg = 0; gg < GP; ++gg) {
      const int g = (pass * GP + gg) ^ (2 * (k & 1));
#pragma unroll
      for (int j = 0; j < NI; ++j) sv[gg][j] = gr[(size_t(g) * N + p) * NI + j];
    }
  };
  auto stage_write = [&](int i, int pass, const cplx_t<T> (&sv)[GP][NI]) {
    const bool ok = spix[i] >= 0;
    const int idx = threadIdx.x + i * TS * TS;
    const int f = idx / Grid::kSlots;
    int wx, wy;
    Grid::at(idx % Grid::kSlots, wx, wy);
    const int pi = wy * PW + wx;
    const int k = fg * FW + f;
#pragma unroll
    for (int gg = 0; gg < GP; ++gg) {
      const int g = (pass * GP + gg) ^ (2 * (k & 1));
      const bool cj = gauge_conj(k, g) != m;
#pragma unroll
      for (int j = 0; j < NI; ++j) {
        cplx_t<T> v = ok ? sv[gg][j] : make_c<T>(T(0), T(0));
        if (cj) v.y = -v.y;
        gs[((f * GP + gg) * NI + j) * SA + pi] = v;
      }
    }
  };
  // The first PF positions of the next pass are loaded into registers before
  // the current pass's slot loop, so that their loads travel while it runs;
  // the other positions are staged between the barriers.
  constexpr int kPF = PF < kPer ? PF : kPer;
  cplx_t<T> pv[kPF > 0 ? kPF : 1][GP][NI];
#pragma unroll
  for (int i = 0; i < kPF; ++i)
    if (spix[i] != -2) stage_load(i, 0, pv[i]);
  for (int pass = 0; pass < kL / GP; ++pass) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPF; ++i)
      if (spix[i] != -2) stage_write(i, pass, pv[i]);
    // In batches of CH positions: every load of a batch is issued before its
    // LDS writes.
    constexpr int CH = stage_batch<cplx_t<T>, kPer, GP * NI, 512 / adj_lds_waves<T, NI, TS, FW, GP>()>();
#pragma unroll
    for (int i0 = kPF; i0 < kPer; i0 += CH) {
      cplx_t<T> sv[CH][GP][NI];
#pragma unroll
      for (int c = 0; c < CH; ++c)
        if (i0 + c < kPer) stage_load(i0 + c, pass, sv[c]);
#pragma unroll
      for (int c = 0; c < CH; ++c)
        if (i0 + c < kPer && spix[i0 + c] != -2) stage_write(i0 + c, pass, sv[c]);
    }
    __syncthreads();
    if (pass + 1 < kL / GP) {
#pragma unroll
      for (int i = 0; i < kPF; ++i)
        if (spix[i] != -2) stage_load(i, pass + 1, pv[i]);
    }
    // Chunks of kSlotUnroll slots; 2 where the window has fewer than 8 planes
    // (the 4-face z adjoint: two workgroups per CU leave 64 VGPRs).
    pipelined_slots<(FW * GP * NI >= 8 ? kSlotUnroll : 2), GP, cplx_t<T>>(
        bs.off, bs.cnt, op.Tn,
        [&](int s0, int u, int gg) {
          return plane_load(tc + (size_t(pass * GP + gg) * op.Tn + size_t(s0)) * op.R + size_t(u) * op.R, qoc);
        },
        [&](int32_t o, const cplx_t<T>(&a)[GP]) {
          // Planes [f][gg][j].
          cplx_t<T> v[FW * GP * NI];
          wp.read(gs, pos, o, [&](int p, cplx_t<T> x) { v[p] = x; });
#pragma unroll
          for (int gg = 0; gg < GP; ++gg)
#pragma unroll
            for (int f = 0; f < FW; ++f)
#pragma unroll
              for (int j = 0; j < NI; ++j) {
                const cplx_t<T> x = v[(f * GP + gg) * NI + j];
                acc[f][j] = fma(a[gg].y, x.y, fma(a[gg].x, x.x, acc[f][j]));
              }
        });
  }
#pragma unroll
  for (int f = 0; f < FW; ++f) {
    const int64_t p = tile_out<TS>(4 * fr + fg * FW + f, m, base, op.order) + li;
#pragma unroll
    for (int j = 0; j < NI; ++j) gx[(size_t(row) * N + p) * NI + j] = acc[f][j];
  }
}

// Correction of the mirrored columns, after the gather: thread per (south cell
// of the list, j), gx += sum over its entries of Re(conj(delta_g) gy[g][row]).
template <typename T, int NI>
__device__ __forceinline__ void conv_adj_fix_body(const ConvOp<T>& op, const cplx_t<T>* __restrict__ gy,
                                                  T* __restrict__ gx, int64_t bx, int row) {
  const int64_t t = bx * 256 + threadIdx.x;
  if (t >= int64_t(op.nfix) * NI) return;
  const int i = int(t / NI), j = int(t % NI);
  const int64_t N = 12 * op.npf;
  const cplx_t<T>* gr = gy + size_t(row) * kL * N * NI;
  const cplx_t<T>* d = reinterpret_cast<const cplx_t<T>*>(op.fix_coef);
  T acc = T(0);
  for (int e = op.fix_ptr[i]; e < op.fix_ptr[i + 1]; ++e) {
    const int64_t p = op.fix_row[e];
#pragma unroll
    for (int g = 0; g < kL; ++g) {
      const cplx_t<T> a = d[size_t(e) * kL + g];
      const cplx_t<T> v = gr[(size_t(g) * N + p) * NI + j];
      acc = fma(a.y, v.y, fma(a.x, v.x, acc));
    }
  }
  gx[(size_t(row) * N + op.fix_cell[i]) * NI + j] += acc;
}

template <typename T, int NI>
__global__ __launch_bounds__(256) void k_conv_adj_fix(ConvOp<T> op, const cplx_t<T>* __restrict__ gy,
                                                      T* __restrict__ gx) {
  conv_adj_fix_body<T, NI>(op, gy, gx, blockIdx.x, blockIdx.y);
}

// Entries of a decimation row loaded before its window is staged (their
// latency then overlaps the staging's).  Slot s reads stored slot
// min(s, capacity - 1); past the row's count that is padding (offset 0,
// inside the window) or the row's last entry, taken with a zero weight.
constexpr int kDecPre = 8;

template <typename T>
__device__ inline void dec_prefetch(const int16_t* __restrict__ off, const T* __restrict__ val, int64_t R, int cap,
                                    int64_t r, int cnt, int16_t (&eo)[kDecPre], T (&ev)[kDecPre]) {
#pragma unroll
  for (int s = 0; s < kDecPre; ++s) {
    const size_t at = size_t(min(s, cap - 1)) * R + r;
    eo[s] = off[at];
    ev[s] = s < cnt ? val[at] : T(0);
  }
}

// Window position of staging index idx in 4 x 4 pixel blocks (SW / 4 blocks
// per side): 16 consecutive indices are one block, whose pixels are 16
// consecutive NESTED ids (the window's corner is a multiple of 4, and the
// neighbour-face maps send aligned blocks to aligned blocks), so a wave's
// loads are 4 contiguous runs.
template <int SW>
__device__ inline void block_pos(int idx, int& wx, int& wy) {
  static_assert(SW % 4 == 0, "window side");
  const int blk = idx >> 4, w = idx & 15;
  wx = 4 * (blk % (SW / 4)) + int(compress_bits(uint32_t(w)));
  wy = 4 * (blk / (SW / 4)) + int(compress_bits(uint32_t(w >> 1)));
}

__device__ inline bool span_inside(int x0, int y0, int w, int order) {
  const int nside = 1 << order;
  return x0 >= 0 && y0 >= 0 && x0 + w <= nside && y0 + w <= nside;
}

// LDS-tiled decimation: a workgroup is a TC x TC tile of coarse pixels of one
// face and one row; the fine window of their candidates, fine offsets
// kDecLo..kDecHi around the tile's children ((2 TC + 8)^2 pixels), is staged
// in LDS (planar [j][pixel]) and every entry reads it at its stored offset.
template <typename T, int NI, int TC>
__global__ __launch_bounds__(TC * TC) void k_dec_fwd_tiled(DecOp<T> op, const T* __restrict__ x, int rows,
                                                           T* __restrict__ y) {
  constexpr int SW = 2 * TC + (kDecHi - kDecLo + 1) - 2, SA = SW * SW;
  static_assert(SW == 2 * TC + 8, "fine window");
  __shared__ T xs[SA * NI];
  int row;
  const TileInfo ti = tile_info<TC, false>(op.order_c, rows, row);
  const int64_t Nf = 12 * op.npf_f, Nc = 12 * op.npf_c;
  const T* xr = x + size_t(row) * Nf * NI;
  const int fx0 = 2 * ti.x0 + kDecLo, fy0 = 2 * ti.y0 + kDecLo;
  const int li = threadIdx.x;
  const int64_t r = int64_t(ti.fr) * op.npf_c + ti.base + li;
  const int cnt = op.cnt[r];
  int16_t eo[kDecPre];
  T ev[kDecPre];
  dec_prefetch<T>(op.off, op.val, op.Rc, op.S, r, cnt, eo, ev);
  const bool inside = span_inside(fx0, fy0, SW, op.order_f);
  stage_window<T, NI, SA, SA, TC * TC>(xr, xs, [&](int idx, int& lp) {
    int wx, wy;
    block_pos<SW>(idx, wx, wy);
    lp = wy * SW + wx;
    return inside ? inner_pixel(ti.face, false, fx0 + wx, fy0 + wy, op.order_f)
                  : ext_pixel(ti.face, fx0 + wx, fy0 + wy, op.order_f);
  });
  __syncthreads();
  const int lx = int(compress_bits(uint32_t(li))), ly = int(compress_bits(uint32_t(li >> 1)));
  const int pos = (2 * ly - kDecLo) * SW + 2 * lx - kDecLo;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
#pragma unroll
  for (int s = 0; s < kDecPre; ++s) {
    const int src = pos + off_dy(eo[s]) * SW + off_dx(eo[s]);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += ev[s] * xs[j * SA + src];
  }
  for (int s = kDecPre; s < cnt; ++s) {
    size_t at = size_t(s) * op.Rc + r;
    int16_t o = op.off[at];
    const T v = op.val[at];
    const int src = pos + off_dy(o) * SW + off_dx(o);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += v * xs[j * SA + src];
  }
  const int64_t p = int64_t(ti.face) * op.npf_c + ti.base + li;
#pragma unroll
  for (int j = 0; j < NI; ++j) y[(size_t(row) * Nc + p) * NI + j] = acc[j];
}

// Its transpose: a TF x TF tile of fine cells gathers from the coarse window
// of their parents +- kDecT ((TF / 2 + 2 kDecT)^2 pixels) staged in LDS.  The
// cell's first entries and the fine term it adds to are loaded before the
// staging.  h = the cell's gather (+ add), rounded to T; p its fine pixel.
template <typename T, int NI, int TF>
__device__ __forceinline__ void dec_adj_tile(const DecOp<T>& op, const T* __restrict__ gy, const T* add, int rows,
                                             T* gs, T (&h)[NI], int64_t& p, int& row) {
  constexpr int SW = TF / 2 + 2 * kDecT, SA = SW * SW;
  const TileInfo ti = tile_info<TF, false>(op.order_f, rows, row);
  const int64_t Nf = 12 * op.npf_f, Nc = 12 * op.npf_c;
  const T* gr = gy + size_t(row) * Nc * NI;
  const int cx0 = ti.x0 / 2 - kDecT, cy0 = ti.y0 / 2 - kDecT;
  const int li = threadIdx.x;
  const int64_t q = int64_t(ti.fr) * op.npf_f + ti.base + li;
  p = int64_t(ti.face) * op.npf_f + ti.base + li;
  const size_t o = (size_t(row) * Nf + p) * NI;
  const int cnt = op.tcnt[q];
  int16_t eo[kDecPre];
  T ev[kDecPre];
  dec_prefetch<T>(op.toff, op.tval, op.Rf, op.Tn, q, cnt, eo, ev);
  T base[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) base[j] = add ? add[o + j] : T(0);
  const bool inside = span_inside(cx0, cy0, SW, op.order_c);
  stage_window<T, NI, SA, SA, TF * TF>(gr, gs, [&](int idx, int& lp) {
    const int wx = idx % SW, wy = idx / SW;
    lp = idx;
    return inside ? inner_pixel(ti.face, false, cx0 + wx, cy0 + wy, op.order_c)
                  : ext_pixel(ti.face, cx0 + wx, cy0 + wy, op.order_c);
  });
  __syncthreads();
  const int lx = int(compress_bits(uint32_t(li))), ly = int(compress_bits(uint32_t(li >> 1)));
  const int pos = ((ly >> 1) + kDecT) * SW + (lx >> 1) + kDecT;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
#pragma unroll
  for (int s = 0; s < kDecPre; ++s) {
    const int src = pos + off_dy(eo[s]) * SW + off_dx(eo[s]);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += ev[s] * gs[j * SA + src];
  }
  for (int s = kDecPre; s < cnt; ++s) {
    size_t at = size_t(s) * op.Rf + q;
    int16_t e = op.toff[at];
    const T v = op.tval[at];
    const int src = pos + off_dy(e) * SW + off_dx(e);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += v * gs[j * SA + src];
  }
#pragma unroll
  for (int j = 0; j < NI; ++j) h[j] = add ? acc[j] + base[j] : acc[j];
}

template <typename T, int NI, int TF>
__global__ __launch_bounds__(TF * TF) void k_dec_adj_tiled(DecOp<T> op, const T* __restrict__ gy, const T* add,
                                                           int rows, T* gx) {
  constexpr int SW = TF / 2 + 2 * kDecT;
  __shared__ T gs[SW * SW * NI];
  T h[NI];
  int64_t p;
  int row;
  dec_adj_tile<T, NI, TF>(op, gy, add, rows, gs, h, p, row);
  const size_t o = (size_t(row) * 12 * op.npf_f + p) * NI;
#pragma unroll
  for (int j = 0; j < NI; ++j) gx[o + j] = h[j];
}

template <typename T, int NI>
__device__ __forceinline__ void dec_fwd_body(const DecOp<T>& op, const T* __restrict__ x, T* __restrict__ y,
                                             int64_t bx, int row) {
  // Wave w of the workgroup = face shift w over 64 consecutive rep rows: a
  // wave's gathers stay in one face, and the 4 waves share the rows' entries.
  const int64_t r = bx * 64 + (threadIdx.x & 63);
  if (r >= op.Rc) return;
  const int k = int(threadIdx.x >> 6);
  int64_t Nf = 12 * op.npf_f, Nc = 12 * op.npf_c;
  const T* xr = x + size_t(row) * Nf * NI;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
  int cnt = op.cnt[r];
  for (int s = 0; s < cnt; ++s) {
    size_t at = size_t(s) * op.Rc + r;
    int64_t fk = rot_pix(op.col[at], k, op.order_f);
    T v = op.val[at];
    T xv[NI];
    Load<T>::template vec<NI>(xr + size_t(fk) * NI, xv);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += v * xv[j];
  }
  int64_t p = rot_pix(rep_pixel(r, op.order_c), k, op.order_c);
#pragma unroll
  for (int j = 0; j < NI; ++j) y[(size_t(row) * Nc + p) * NI + j] = acc[j];
}

template <typename T, int NI>
__global__ __launch_bounds__(256) void k_dec_fwd(DecOp<T> op, const T* __restrict__ x, T* __restrict__ y) {
  dec_fwd_body<T, NI>(op, x, y, blockIdx.x, blockIdx.y);
}

template <typename T, int NI>
__device__ __forceinline__ void dec_adj_body(const DecOp<T>& op, const T* __restrict__ gy, const T* add, T* gx,
                                             int64_t bx, int row) {
  const int64_t q = bx * 64 + (threadIdx.x & 63);
  if (q >= op.Rf) return;
  const int k = int(threadIdx.x >> 6);
  int64_t Nf = 12 * op.npf_f, Nc = 12 * op.npf_c;
  const T* gr = gy + size_t(row) * Nc * NI;
  T acc[NI];
#pragma unroll
  for (int j = 0; j < NI; ++j) acc[j] = T(0);
  int cnt = op.tcnt[q];
  for (int s = 0; s < cnt; ++s) {
    size_t at = size_t(s) * op.Rf + q;
    int64_t ck = rot_pix(op.trow[at], k, op.order_c);
    T v = op.tval[at];
    T gv[NI];
    Load<T>::template vec<NI>(gr + size_t(ck) * NI, gv);
#pragma unroll
    for (int j = 0; j < NI; ++j) acc[j] += v * gv[j];
  }
  int64_t qk = rot_pix(rep_pixel(q, op.order_f), k, op.order_f);
  size_t o = (size_t(row) * Nf + qk) * NI;
#pragma unroll
  for (int j = 0; j < NI; ++j) gx[o + j] = add ? acc[j] + add[o + j] : acc[j];
}

template <typename T, int NI>
__global__ __launch_bounds__(256) void k_dec_adj(DecOp<T> op, const T* __restrict__ gy, const T* add,
                                                 T* gx) {
  dec_adj_body<T, NI>(op, gy, add, gx, blockIdx.x, blockIdx.y);
}

template <typename T>
__device__ inline double2 to_d2(cplx_t<T> v) {
  return make_double2(double(v.x), double(v.y));
}

// Fixed-order block tree of 256 values (4 per thread).
__device__ inline void block_sum4(double v[4], double* sh) {
  int tid = threadIdx.x;
  for (int c = 0; c < 4; ++c) sh[c * 256 + tid] = v[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w)
      for (int c = 0; c < 4; ++c) sh[c * 256 + tid] += sh[c * 256 + tid + w];
    __syncthreads();
  }
  for (int c = 0; c < 4; ++c) v[c] = sh[c * 256];
}

__device__ inline bool in_mask(const MaskView& m, int64_t p, int k) {
  return m.cls == nullptr || ((m.pat[m.cls[p]] >> k) & 1ull) != 0;
}
__device__ inline int mask_class(const MaskView& m, int64_t p) { return m.cls == nullptr ? 0 : int(m.cls[p]); }

// One DPP step of a 64-bit value: each half moves by the same lane pattern
// (CTRL); rows left out of ROW_MASK read 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ inline double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, int(uint32_t(u)), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(uint32_t(u >> 32)), CTRL, ROW_MASK, 0xF, false);
  return __builtin_bit_cast(double, (uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// Sum over the 64 lanes of a wave in a fixed order, all lanes active: quad
// swaps (lane ^ 1, lane ^ 2), the half-row and row mirrors (every lane of a
// row then holds the row's sum), row 0 into row 1 and row 2 into row 3 (row
// broadcast of lane 15), rows 0-1 into row 3 (broadcast of lane 31).  The
// total is in lane 63.
__device__ inline double wave_sum_lane63(double v) {
  v += dpp_f64<0xB1>(v);        // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E>(v);        // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141>(v);       // row_half_mirror
  v += dpp_f64<0x140>(v);       // row_mirror
  v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Gram partial sums: block (tile stride, b * kL + l3).
// Gram products at the levels of small K (all E = K (K + 1) / 2 entries of a
// pixel fit in registers): each thread reads its pixels' V straight from HBM
// (40 bytes per pixel at K = 5, coalesced) and accumulates every entry in
// FP64 (two chained FMAs per real and imaginary part; the diagonal entries'
// imaginary parts are exactly zero and not formed); the workgroup then sums
// each wave's lanes in a fixed DPP order and the 4 wave sums in order
// (deterministic).
template <typename T, int KT, bool MOM, bool MASKED>
__global__ __launch_bounds__(256) void k_gram_reg(int n, int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                                  const cplx_t<T>* __restrict__ z, double* __restrict__ part,
                                                  double* __restrict__ part2, double* __restrict__ zpart,
                                                  MaskView mask) {
  constexpr int E = KT * (KT + 1) / 2;
  constexpr int ND = (KT - 1) / kL;  // branch rows d = 0..n
  __shared__ double red[4][MOM ? 3 : 2][E];
  __shared__ double zred[4][4];
  const int bl = blockIdx.y;  // (b K + k) kL + l3
  const int l3 = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const int tid = threadIdx.x;
  double are[E], aim[E], asq[MOM ? E : 1];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    are[e] = aim[e] = 0.0;
    if (MOM) asq[e] = 0.0;
  }
  // The S1/S2 sums of z (sum |z|, |z|^2, re^2 + im^2, |z|^4) ride along.
  double zs[4] = {0.0, 0.0, 0.0, 0.0};
  const cplx_t<T>* zr = z + (size_t(b) * kL + l3) * N;
  for (int64_t p = blockIdx.x * int64_t(256) + tid; p < N; p += int64_t(gridDim.x) * 256) {
    if (MASKED && !in_mask(mask, p, mk)) continue;
    double2 v[KT];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const cplx_t<T>* src = br + ((size_t(d) * B + b) * kL + l3) * N * kL + size_t(p) * kL;
#pragma unroll
      for (int l1 = 0; l1 < kL; ++l1) v[d * kL + l1] = to_d2<T>(src[l1]);
    }
    v[KT - 1] = to_d2<T>(zr[p]);
    {
      const double h = hypot(v[KT - 1].x, v[KT - 1].y);
      zs[0] += h;
      zs[1] += h * h;
      zs[2] += v[KT - 1].x * v[KT - 1].x + v[KT - 1].y * v[KT - 1].y;
      zs[3] += (h * h) * (h * h);
    }
    int e = 0;
#pragma unroll
    for (int i = 0; i < KT; ++i)
#pragma unroll
      for (int j = i; j < KT; ++j, ++e) {
        if constexpr (MOM) {
          const double re = v[i].x * v[j].x + v[i].y * v[j].y;
          const double im = i == j ? 0.0 : v[i].y * v[j].x - v[i].x * v[j].y;
          are[e] += re;
          aim[e] += im;
          asq[e] += re * re + im * im;
        } else {
          are[e] = fma(v[i].x, v[j].x, are[e]);
          are[e] = fma(v[i].y, v[j].y, are[e]);
          if (i != j) {
            aim[e] = fma(v[i].y, v[j].x, aim[e]);
            aim[e] = fma(-v[i].x, v[j].y, aim[e]);
          }
        }
      }
  }
  (void)n;
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const double a = wave_sum_lane63(are[e]);
    const double c = wave_sum_lane63(aim[e]);
    const double q = MOM ? wave_sum_lane63(asq[e]) : 0.0;
    if (lane == 63) {
      red[w][0][e] = a;
      red[w][1][e] = c;
      if (MOM) red[w][MOM ? 2 : 1][e] = q;
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double a = wave_sum_lane63(zs[c]);
    if (lane == 63) zred[w][c] = a;
  }
  __syncthreads();
  const size_t base = (size_t(blockIdx.x) * gridDim.y + bl) * E;
  for (int e = tid; e < E; e += 256) {
    double a = 0.0, c = 0.0, q = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a += red[k][0][e];
      c += red[k][1][e];
      if (MOM) q += red[k][MOM ? 2 : 1][e];
    }
    part[(base + e) * 2] = a;
    part[(base + e) * 2 + 1] = c;
    if (MOM) part2[base + e] = q;
  }
  if (tid < 4)
    zpart[(size_t(blockIdx.x) * gridDim.y + bl) * 4 + tid] = zred[0][tid] + zred[1][tid] + zred[2][tid] + zred[3][tid];
}

// Gram products at the middle levels, K = 4 R + 1 (R = n + 1 = 3..6 branch
// rows d).  The upper triangle of the K x K Gram is R (R + 1) / 2 blocks of
// 4 x 4 entries between rows a <= b, plus the z column.  Lane class q of a
// wave owns block (a_q, b_q): its S = 64 / P lanes q S .. q S + S - 1 take S
// consecutive pixels, load rows a_q, b_q and z there and accumulate the 16
// block products, the 4 products of row a_q with z (kept from the diagonal
// blocks), z z and the S1/S2 sums of z (kept from class 0), in FP64
// registers: every lane runs the same code on its own block, and the pixel
// loop touches no LDS.  At the end the lanes of each class are summed in
// wave and lane order through LDS (deterministic).
template <int R>
struct GramClasses {
  static constexpr int P = R * (R + 1) / 2;  // classes
  static constexpr int S = 64 / P;           // lanes (pixels) per class and wave
  static constexpr int K = 4 * R + 1, E = K * (K + 1) / 2;
  // per-lane values: block re [16], block im [16], z column re [4], im [4],
  // z z, the 4 z sums; second moments: block [16], z column [4], z z
  static constexpr int GR = 0, GI = 16, ZR = 32, ZI = 36, ZZ = 40, ZS = 41, SQ = 45, ZSQ = 61, ZZSQ = 65;
  static constexpr int nv(bool mom) { return mom ? 66 : 45; }
  __host__ __device__ static constexpr int cls(int a, int b) { return a * R - a * (a - 1) / 2 + (b - a); }
};
constexpr int kGramChunk = 8;  // values per LDS round of the class sums

template <typename T, int R, bool MOM, bool MASKED>
__global__ __launch_bounds__(256) void k_gram_cls(int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                                  const cplx_t<T>* __restrict__ z, double* __restrict__ part,
                                                  double* __restrict__ part2, double* __restrict__ zpart,
                                                  MaskView mask) {
  using G = GramClasses<R>;
  constexpr int P = G::P, S = G::S, K = G::K, E = G::E, NV = G::nv(MOM);
  __shared__ double buf[kGramChunk][256];
  __shared__ double red[P][NV];
  const int bl = blockIdx.y;  // (b K + k) kL + l3
  const int l3 = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int q = lane / S, kk = lane % S;
  const bool live = q < P;
  int qa = 0, rem = live ? q : 0;
  while (rem >= R - qa) {
    rem -= R - qa;
    ++qa;
  }
  const int qb = qa + rem;
  const bool diag = qa == qb;
  const cplx_t<T>* ra = br + ((size_t(qa) * B + b) * kL + l3) * size_t(N) * kL;
  const cplx_t<T>* rb = br + ((size_t(qb) * B + b) * kL + l3) * size_t(N) * kL;
  const cplx_t<T>* zr = z + (size_t(b) * kL + l3) * size_t(N);
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  const int64_t step = int64_t(gridDim.x) * 4 * S;
  struct Pix {
    cplx_t<T> a[4], b[4], z;
  };
  auto fetch = [&](int64_t p, Pix& v) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      v.a[c] = ra[p * kL + c];
      v.b[c] = rb[p * kL + c];
    }
    v.z = zr[p];
  };
  auto body = [&](const Pix& pv, int64_t p) {
    if (MASKED && !in_mask(mask, p, mk)) return;
    double2 va[4], vb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      va[c] = to_d2<T>(pv.a[c]);
      vb[c] = to_d2<T>(pv.b[c]);
    }
    const double2 vz = to_d2<T>(pv.z);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double& gre = acc[G::GR + r * 4 + c];
        double& gim = acc[G::GI + r * 4 + c];
        if constexpr (MOM) {
          const double re = va[r].x * vb[c].x + va[r].y * vb[c].y;
          double im = va[r].y * vb[c].x - va[r].x * vb[c].y;
          if (r == c) im = diag ? 0.0 : im;
          gre += re;
          gim += im;
          acc[G::SQ + r * 4 + c] += re * re + im * im;
        } else {
          gre = fma(va[r].x, vb[c].x, gre);
          gre = fma(va[r].y, vb[c].y, gre);
          gim = fma(va[r].y, vb[c].x, gim);
          gim = fma(-va[r].x, vb[c].y, gim);
        }
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if constexpr (MOM) {
        const double re = va[r].x * vz.x + va[r].y * vz.y;
        const double im = va[r].y * vz.x - va[r].x * vz.y;
        acc[G::ZR + r] += re;
        acc[G::ZI + r] += im;
        acc[G::ZSQ + r] += re * re + im * im;
      } else {
        acc[G::ZR + r] = fma(va[r].x, vz.x, acc[G::ZR + r]);
        acc[G::ZR + r] = fma(va[r].y, vz.y, acc[G::ZR + r]);
        acc[G::ZI + r] = fma(va[r].y, vz.x, acc[G::ZI + r]);
        acc[G::ZI + r] = fma(-va[r].x, vz.y, acc[G::ZI + r]);
      }
    }
    const double zz = vz.x * vz.x + vz.y * vz.y;
    acc[G::ZZ] += zz;
    if constexpr (MOM) acc[G::ZZSQ] += zz * zz;
    const double h = hypot(vz.x, vz.y);
    acc[G::ZS] += h;
    acc[G::ZS + 1] += h * h;
    acc[G::ZS + 2] += zz;
    acc[G::ZS + 3] += (h * h) * (h * h);
  };
  // Two pixel buffers in turn: a pass's values load while the previous
  // pass's products run (addresses clamped, so every load is unconditional).
  const int64_t p0 = live ? (int64_t(blockIdx.x) * 4 + w) * S + kk : N;
  const int64_t np = p0 < N ? (N - 1 - p0) / step + 1 : 0;
  Pix v0, v1;
  if (np > 0) fetch(p0, v0);
  int64_t t = 0;
  for (; t + 1 < np; t += 2) {
    fetch(p0 + (t + 1) * step, v1);
    body(v0, p0 + t * step);
    fetch(std::min(p0 + (t + 2) * step, N - 1), v0);
    body(v1, p0 + (t + 1) * step);
  }
  if (t < np) body(v0, p0 + t * step);
  // Class sums: kGramChunk values of every lane per LDS round; the value of
  // class q is summed over waves 0..3, then lanes q S .. q S + S - 1.
#pragma unroll
  for (int c0 = 0; c0 < NV; c0 += kGramChunk) {
#pragma unroll
    for (int v = 0; v < kGramChunk; ++v)
      if (c0 + v < NV) buf[v][tid] = acc[c0 + v];
    __syncthreads();
    if (tid < P * kGramChunk) {
      const int cq = tid / kGramChunk, v = tid % kGramChunk;
      if (c0 + v < NV) {
        double s = 0.0;
        for (int ww = 0; ww < 4; ++ww)
#pragma unroll
          for (int k = 0; k < S; ++k) s += buf[v][ww * 64 + cq * S + k];
        red[cq][c0 + v] = s;
      }
    }
    __syncthreads();
  }
  const size_t base = (size_t(blockIdx.x) * gridDim.y + bl) * E;
  for (int e = tid; e < E; e += 256) {
    int i = 0, re = e;
    while (re >= K - i) {
      re -= K - i;
      ++i;
    }
    const int j = i + re;
    double vr, vi, vq = 0.0;
    if (j == K - 1) {
      if (i == K - 1) {
        vr = red[0][G::ZZ];
        vi = 0.0;
        if (MOM) vq = red[0][G::ZZSQ];
      } else {
        const int cq = G::cls(i / 4, i / 4), r = i % 4;
        vr = red[cq][G::ZR + r];
        vi = red[cq][G::ZI + r];
        if (MOM) vq = red[cq][G::ZSQ + r];
      }
    } else {
      const int cq = G::cls(i / 4, j / 4), rc = (i % 4) * 4 + j % 4;
      vr = red[cq][G::GR + rc];
      vi = i == j ? 0.0 : red[cq][G::GI + rc];
      if (MOM) vq = red[cq][G::SQ + rc];
    }
    part[(base + e) * 2] = vr;
    part[(base + e) * 2 + 1] = vi;
    if (MOM) part2[base + e] = vq;
  }
  if (tid < 4) zpart[(size_t(blockIdx.x) * gridDim.y + bl) * 4 + tid] = red[0][G::ZS + tid];
}

// Gram products of one (b, l3) per workgroup row.  A tile of V [TPX][KP]
// (columns K..KP-1 zero) is staged in LDS; each thread owns a 4 x 4 block
// (bi <= bj) of entries V_i conj(V_j) and a group of the tile's pixels, so
// every LDS read feeds 4 products.  The values widen exactly to FP64; the
// products and their sums are FP64.  The per-group sums are added in group
// order: deterministic.
// gx_dim / gy_dim: the grid of the launch (tiles are strided by gx_dim, the
// partials indexed by it); smem: the dynamic LDS.
template <typename T, bool MOM, int BS, bool MASKED>
__device__ __forceinline__ void gram_body(int n, int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                          const cplx_t<T>* __restrict__ z, int TPX, double* __restrict__ part,
                                          double* __restrict__ part2, double* __restrict__ zpart,
                                          const MaskView& mask, int bx, int gx_dim, int bl, int gy_dim,
                                          double2* smem) {
  const int K = (n + 1) * kL + 1;
  const int KP = (K + BS - 1) / BS * BS, NR = KP / BS, NBLK = NR * (NR + 1) / 2;
  // The S1/S2 sums of z, accumulated by the threads that stage z.
  double zs[4] = {0.0, 0.0, 0.0, 0.0};
  const int E = K * (K + 1) / 2;
  double2* Vs = smem;
  // bl = (b K + k) kL + l3
  const int l3 = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const int tid = threadIdx.x;
  const int G = 256 / NBLK;
  const int blk = tid % NBLK, grp = tid / NBLK;
  const bool act = grp < G;
  int bi = 0, rem = blk;
  while (rem >= NR - bi) { rem -= NR - bi; ++bi; }
  const int bj = bi + rem;
  double are[BS][BS], aim[BS][BS], asq[BS][BS];
#pragma unroll
  for (int r = 0; r < BS; ++r)
#pragma unroll
    for (int c = 0; c < BS; ++c) are[r][c] = aim[r][c] = asq[r][c] = 0.0;
  const int64_t ntiles = (N + TPX - 1) / TPX;
  for (int64_t tile = bx; tile < ntiles; tile += gx_dim) {
    const int64_t p0 = tile * TPX;
    const int tn = int(min<int64_t>(TPX, N - p0));
    for (int d = 0; d <= n; ++d) {
      const cplx_t<T>* src = br + ((size_t(d) * B + b) * kL + l3) * N * kL + size_t(p0) * kL;
      for (int idx = tid; idx < tn * kL; idx += 256) {
        int p = idx / kL, l1 = idx % kL;
        // Pixels outside the mask stage zeros: their products add nothing.
        Vs[p * KP + d * kL + l1] =
            !MASKED || in_mask(mask, p0 + p, mk) ? to_d2<T>(src[idx]) : make_double2(0.0, 0.0);
      }
    }
    const cplx_t<T>* zt = z + (size_t(b) * kL + l3) * N + p0;
    for (int p = tid; p < tn; p += 256) {
      const double2 v = !MASKED || in_mask(mask, p0 + p, mk) ? to_d2<T>(zt[p]) : make_double2(0.0, 0.0);
      Vs[p * KP + K - 1] = v;
      const double h = hypot(v.x, v.y);
      zs[0] += h;
      zs[1] += h * h;
      zs[2] += v.x * v.x + v.y * v.y;
      zs[3] += (h * h) * (h * h);
    }
    const int pad = KP - K;
    for (int idx = tid; idx < tn * pad; idx += 256) Vs[(idx / pad) * KP + K + idx % pad] = make_double2(0.0, 0.0);
    __syncthreads();
    if (act) {
      for (int p = grp; p < tn; p += G) {
        double2 a[BS], c[BS];
#pragma unroll
        for (int r = 0; r < BS; ++r) {
          a[r] = Vs[p * KP + BS * bi + r];
          c[r] = Vs[p * KP + BS * bj + r];
        }
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
          for (int q = 0; q < BS; ++q) {
            double re = a[r].x * c[q].x + a[r].y * c[q].y;
            double im = a[r].y * c[q].x - a[r].x * c[q].y;
            are[r][q] += re;
            aim[r][q] += im;
            if (MOM) asq[r][q] += re * re + im * im;
          }
      }
    }
    __syncthreads();
  }
  // Sum the pixel groups: one (r, q) entry of every block per round.
  double2* red = smem;                                          // [256]
  double* red2 = reinterpret_cast<double*>(smem + 256);         // [256]
  constexpr int BB = BS * BS;
  double2* fin = smem + 256 + 128;                              // [NBLK][BB]
  double* fin2 = reinterpret_cast<double*>(fin + NBLK * BB);    // [NBLK][BB]
#pragma unroll
  for (int r = 0; r < BS; ++r)
#pragma unroll
    for (int q = 0; q < BS; ++q) {
      red[tid] = make_double2(are[r][q], aim[r][q]);
      if (MOM) red2[tid] = asq[r][q];
      __syncthreads();
      if (tid < NBLK) {
        double2 s = make_double2(0.0, 0.0);
        double s2 = 0.0;
        for (int g = 0; g < G; ++g) {
          s.x += red[g * NBLK + tid].x;
          s.y += red[g * NBLK + tid].y;
          if (MOM) s2 += red2[g * NBLK + tid];
        }
        fin[tid * BB + r * BS + q] = s;
        if (MOM) fin2[tid * BB + r * BS + q] = s2;
      }
      __syncthreads();
    }
  const size_t base = (size_t(bx) * gy_dim + bl) * E;
  for (int e = tid; e < E; e += 256) {
    int i = 0, re = e;
    while (re >= K - i) { re -= K - i; ++i; }
    const int j = i + re;
    const int ib = i / BS, jb = j / BS;
    const int k = (ib * NR - ib * (ib - 1) / 2 + (jb - ib)) * BB + (i % BS) * BS + (j % BS);
    part[(base + e) * 2] = fin[k].x;
    part[(base + e) * 2 + 1] = fin[k].y;
    if (MOM) part2[base + e] = fin2[k];
  }
  __syncthreads();
  block_sum4(zs, reinterpret_cast<double*>(smem));
  if (tid == 0)
    for (int c = 0; c < 4; ++c) zpart[(size_t(bx) * gy_dim + bl) * 4 + c] = zs[c];
}

template <typename T, bool MOM, int BS, bool MASKED>
__global__ __launch_bounds__(256) void k_gram(int n, int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                              const cplx_t<T>* __restrict__ z, int TPX, double* __restrict__ part,
                                              double* __restrict__ part2, double* __restrict__ zpart,
                                              MaskView mask) {
  extern __shared__ double2 smem[];
  gram_body<T, MOM, BS, MASKED>(n, N, B, br, z, TPX, part, part2, zpart, mask, blockIdx.x, gridDim.x, blockIdx.y,
                                gridDim.y, smem);
}


// Cross Gram products of a map pair: G[i][j] = sum_p V_a,i conj(V_b,j), the
// full K x K block between the V of field row 2b (channel a) and of row
// 2b + 1 (channel b), per (b, mask k, l3).  Tiles of both V staged in LDS;
// each thread owns a BS x BS block (bi, bj) of the NR x NR grid and a group
// of the tile's pixels; the group sums are added in group order per entry
// and written straight out: part [nblk][B K kL][K K][2] (+ [K K] |.|^2).
template <typename T, bool MOM, int BS, bool MASKED>
__global__ __launch_bounds__(256) void k_gram_x(int n, int64_t N, int Bf, const cplx_t<T>* __restrict__ br,
                                                const cplx_t<T>* __restrict__ z, int TPX, double* __restrict__ part,
                                                double* __restrict__ part2, MaskView mask) {
  extern __shared__ double2 smem[];
  const int K = (n + 1) * kL + 1;
  const int KP = (K + BS - 1) / BS * BS, NR = KP / BS, NBLK = NR * NR;
  const int E = K * K;
  double2* Va = smem;
  double2* Vb = smem + size_t(TPX) * KP;
  const int bl = blockIdx.y;  // (b K + k) kL + l3
  const int l3 = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const int tid = threadIdx.x;
  const int G = 256 / NBLK;
  const int blk = tid % NBLK, grp = tid / NBLK;
  const bool act = grp < G;
  const int bi = blk / NR, bj = blk % NR;
  double are[BS][BS], aim[BS][BS], asq[BS][BS];
#pragma unroll
  for (int r = 0; r < BS; ++r)
#pragma unroll
    for (int c = 0; c < BS; ++c) are[r][c] = aim[r][c] = asq[r][c] = 0.0;
  const int64_t ntiles = (N + TPX - 1) / TPX;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t p0 = tile * TPX;
    const int tn = int(min<int64_t>(TPX, N - p0));
    for (int h = 0; h < 2; ++h) {
      double2* V = h == 0 ? Va : Vb;
      const int row = 2 * b + h;
      for (int d = 0; d <= n; ++d) {
        const cplx_t<T>* src = br + ((size_t(d) * Bf + row) * kL + l3) * N * kL + size_t(p0) * kL;
        for (int idx = tid; idx < tn * kL; idx += blockDim.x) {
          int p = idx / kL, l1 = idx % kL;
          // Channel a outside the mask stages zeros: its products add nothing.
          V[p * KP + d * kL + l1] = h == 1 || !MASKED || in_mask(mask, p0 + p, mk) ? to_d2<T>(src[idx])
                                                                                     : make_double2(0.0, 0.0);
        }
      }
      const cplx_t<T>* zs = z + (size_t(row) * kL + l3) * N + p0;
      for (int p = tid; p < tn; p += blockDim.x)
        V[p * KP + K - 1] = h == 1 || !MASKED || in_mask(mask, p0 + p, mk) ? to_d2<T>(zs[p]) : make_double2(0.0, 0.0);
      const int pad = KP - K;
      for (int idx = tid; idx < tn * pad; idx += blockDim.x) V[(idx / pad) * KP + K + idx % pad] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (act) {
      for (int p = grp; p < tn; p += G) {
        double2 a[BS], c[BS];
#pragma unroll
        for (int r = 0; r < BS; ++r) {
          a[r] = Va[p * KP + BS * bi + r];
          c[r] = Vb[p * KP + BS * bj + r];
        }
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
          for (int q = 0; q < BS; ++q) {
            double re = a[r].x * c[q].x + a[r].y * c[q].y;
            double im = a[r].y * c[q].x - a[r].x * c[q].y;
            are[r][q] += re;
            aim[r][q] += im;
            if (MOM) asq[r][q] += re * re + im * im;
          }
      }
    }
    __syncthreads();
  }
  double2* red = smem;                                   // [256]
  double* red2 = reinterpret_cast<double*>(smem + 256);  // [256]
  const size_t base = (size_t(blockIdx.x) * gridDim.y + bl) * E;
#pragma unroll
  for (int r = 0; r < BS; ++r)
#pragma unroll
    for (int q = 0; q < BS; ++q) {
      red[tid] = make_double2(are[r][q], aim[r][q]);
      if (MOM) red2[tid] = asq[r][q];
      __syncthreads();
      if (tid < NBLK) {
        const int i = BS * (tid / NR) + r, j = BS * (tid % NR) + q;
        double2 sum = make_double2(0.0, 0.0);
        double s2 = 0.0;
        for (int g = 0; g < G; ++g) {
          sum.x += red[g * NBLK + tid].x;
          sum.y += red[g * NBLK + tid].y;
          if (MOM) s2 += red2[g * NBLK + tid];
        }
        if (i < K && j < K) {
          part[(base + size_t(i) * K + j) * 2] = sum.x;
          part[(base + size_t(i) * K + j) * 2 + 1] = sum.y;
          if (MOM) part2[base + size_t(i) * K + j] = s2;
        }
      }
      __syncthreads();
    }
}

__device__ inline double2 root_normalise(double2 v) {
  const double s = sqrt(hypot(v.x, v.y) + 1e-8);
  return make_double2(v.x / s, v.y / s);
}

// Cotangent of v for the root-normalised operand v / sqrt(|v| + eps) given
// the cotangent g of that operand (the reference's root_modulus_vjp,
// crates/stl-core/src/engine.rs:46-56).
__device__ inline double2 root_modulus_vjp(double2 v, double2 g) {
  const double r = hypot(v.x, v.y);
  const double radial = g.x * v.x + g.y * v.y;
  const double mag = r == 0.0 ? 1.0 : r;
  const double q = sqrt(r + 1e-8);
  const double s = 0.5 * radial / ((r + 1e-8) * q);
  return make_double2(g.x / q - (s * v.x) / mag, g.y / q - (s * v.y) / mag);
}

// S1 cross sums of a map pair per (b, mask k, l): t = rn(z_a) conj(rn(z_b)),
// rn(v) = v / sqrt(|v| + 1e-8): part [nblk][B K kL][4] = (sum re t, sum im t,
// sum |t|^2, 0).
template <typename T>
__global__ __launch_bounds__(256) void k_zcross(int64_t N, const cplx_t<T>* __restrict__ z, double* __restrict__ part,
                                                MaskView mask) {
  __shared__ double sh[4 * 256];
  const int bl = blockIdx.y;
  const int l = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const cplx_t<T>* za = z + (size_t(2 * b) * kL + l) * N;
  const cplx_t<T>* zb = z + (size_t(2 * b + 1) * kL + l) * N;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < N; p += int64_t(gridDim.x) * blockDim.x) {
    if (!in_mask(mask, p, mk)) continue;
    const double2 a = root_normalise(to_d2<T>(za[p])), c = root_normalise(to_d2<T>(zb[p]));
    const double re = a.x * c.x + a.y * c.y, im = a.y * c.x - a.x * c.y;
    v[0] += re;
    v[1] += im;
    v[2] += re * re + im * im;
  }
  block_sum4(v, sh);
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[(size_t(blockIdx.x) * gridDim.y + bl) * 4 + c] = v[c];
}

template <typename T>
__global__ __launch_bounds__(256) void k_zstats(int64_t N, const cplx_t<T>* __restrict__ z, double* __restrict__ part,
                                                MaskView mask) {
  __shared__ double sh[4 * 256];
  const int bl = blockIdx.y;  // (b K + k) kL + l
  const int l = bl % kL, mk = (bl / kL) % mask.K, b = bl / (kL * mask.K);
  const cplx_t<T>* zs = z + (size_t(b) * kL + l) * N;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < N; p += int64_t(gridDim.x) * blockDim.x) {
    if (!in_mask(mask, p, mk)) continue;
    double re = double(zs[p].x), im = double(zs[p].y);
    double h = hypot(re, im);
    v[0] += h;
    v[1] += h * h;
    v[2] += re * re + im * im;
    v[3] += (h * h) * (h * h);
  }
  block_sum4(v, sh);
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[(size_t(blockIdx.x) * gridDim.y + bl) * 4 + c] = v[c];
}

template <typename T>
__global__ __launch_bounds__(256) void k_meanvar(int64_t N, const T* __restrict__ x, double* __restrict__ part,
                                                 MaskView mask) {
  __shared__ double sh[4 * 256];
  const int bl = blockIdx.y;  // b K + k
  const int mk = bl % mask.K, b = bl / mask.K;
  const T* xs = x + size_t(b) * N;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; p < N; p += int64_t(gridDim.x) * blockDim.x) {
    if (!in_mask(mask, p, mk)) continue;
    double a = double(xs[p]);
    double sq = fabs(a) * fabs(a);
    v[0] += a;
    v[1] += sq;
    v[2] += a * a;
    v[3] += sq * sq;
  }
  block_sum4(v, sh);
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[(size_t(blockIdx.x) * gridDim.y + bl) * 4 + c] = v[c];
}

// part [nblk][m] -> out [m]: 64 outputs per workgroup, wave w adding the
// blocks b = w, w + 16, ... in order, then the 16 wave sums in order.
constexpr int kRedWaves = 16;
// The same sums by a 256-thread workgroup (a work-queue item): each wave
// takes the block sequences of 4 of the 16 "waves" above in turn, so every
// output adds the same terms in the same order (the same bits).  lds: 16 x 64 doubles.
__device__ __forceinline__ void reduce_parts_body(const double* __restrict__ part, int nblk, int64_t m,
                                                  double* __restrict__ out, int64_t bx, double* lds) {
  const int lane = threadIdx.x & 63, w4 = threadIdx.x >> 6;
  const int64_t i = bx * 64 + lane;
  for (int v = w4; v < kRedWaves; v += 4) {
    double s = 0.0;
    if (i < m) {
#pragma unroll 8
      for (int b = v; b < nblk; b += kRedWaves) s += part[size_t(b) * m + i];
    }
    lds[v * 64 + lane] = s;
  }
  __syncthreads();
  if (w4 == 0 && i < m) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kRedWaves; ++k) t += lds[k * 64 + lane];
    out[i] = t;
  }
}

__global__ __launch_bounds__(64 * kRedWaves) void k_reduce_parts(const double* __restrict__ part, int nblk,
                                                                 int64_t m, double* __restrict__ out) {
  __shared__ double sh[kRedWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i = blockIdx.x * int64_t(64) + lane;
  double s = 0.0;
  if (i < m) {
#pragma unroll 8
    for (int b = w; b < nblk; b += kRedWaves) s += part[size_t(b) * m + i];
  }
  sh[w][lane] = s;
  __syncthreads();
  if (w == 0 && i < m) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kRedWaves; ++k) t += sh[k][lane];
    out[i] = t;
  }
}

__global__ void k_assemble(int np, int rows, int K, int levels, const StatEntry* __restrict__ entries,
                           const double* const* src, const double* __restrict__ nd, bool moments,
                           double2* __restrict__ values, double* __restrict__ variance) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(np) * rows) return;
  const int r = int(t / np), i = int(t % np);  // r = b K + k
  const StatEntry e = entries[i];
  const double Nd = nd[(r % K) * levels + e.level];
  const double* a = src[e.arr];
  double re = a[e.re + r * e.bs] / Nd;
  double im = e.im < 0 ? 0.0 : a[e.im + r * e.bs] / Nd;
  if (e.conj) im = -im;
  values[t] = make_double2(re, im);
  if (moments && variance != nullptr) {
    double second = e.sq_arr < 0 ? 0.0 : src[e.sq_arr][e.sq + r * e.sbs] / Nd;
    variance[t] = fmax(second - (re * re + im * im), 0.0);
  }
}

__global__ void k_scatter_cot(int ntargets, int rows, int K, int levels, int np, const RevTarget* __restrict__ targets,
                              const RevTerm* __restrict__ terms, const double* __restrict__ nd,
                              const double2* __restrict__ cot, double* const* dst) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(ntargets) * rows) return;
  const int row = int(t / ntargets);  // b K + k
  const RevTarget r = targets[t % ntargets];
  const double* ndk = nd + (row % K) * levels;
  double re = 0.0, im = 0.0;
  for (int q = r.first; q < r.first + r.count; ++q) {
    const RevTerm m = terms[q];
    const double Nd = ndk[m.level];
    double2 g = cot[size_t(row) * np + m.coef];
    re += (m.mult * g.x) / Nd;
    im += (m.mult * (m.conj ? -g.y : g.y)) / Nd;
  }
  double* d = dst[r.arr] + r.off + row * r.bs;
  d[0] = re;
  if (r.cplx) d[1] = im;
}

__global__ void k_mask_combine(int B, int K, int R, int C, int64_t E, const uint64_t* __restrict__ pat,
                               const double* __restrict__ in, double* __restrict__ out) {
  int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (t >= int64_t(B) * R * C * E) return;
  const int64_t e = t % E;
  const int c = int((t / E) % C), r = int((t / (E * C)) % R), b = int(t / (E * C * R));
  const uint64_t bits = pat[c];
  double acc = 0.0;
  for (int k = 0; k < K; ++k)
    if ((bits >> k) & 1ull) acc += in[((size_t(b) * K + k) * R + r) * E + e];
  out[t] = acc;
}

// gV = H V of one pixel: writes the branch cotangents and the z part.
template <typename T, int K>
__device__ __forceinline__ void gv_apply(const double2* Hs, const double2 (&V)[K], int64_t N, int B, int b, int l3,
                                         int64_t p, cplx_t<T>* __restrict__ gbr, cplx_t<T>* __restrict__ gzp) {
#pragma unroll 1
  for (int i = 0; i < K; ++i) {
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      double2 h = Hs[i * K + j];
      re += h.x * V[j].x - h.y * V[j].y;
      im += h.x * V[j].y + h.y * V[j].x;
    }
    if (i < K - 1) {
      int d = i / kL, l1 = i % kL;
      gbr[(((size_t(d) * B + b) * kL + l3) * N + p) * kL + l1] = make_c<T>(T(re), T(im));
    } else {
      gzp[(size_t(b) * kL + l3) * N + p] = make_c<T>(T(re), T(im));
    }
  }
}

// H of the bl row's C mask classes staged in LDS when LDS (C K^2 double2 of
// dynamic shared memory), else read through the caches; the pixel uses the
// matrix of its class.
template <typename T, int NL, bool LDS>
__global__ __launch_bounds__(256) void k_gv(int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                            const cplx_t<T>* __restrict__ z, const double2* __restrict__ H,
                                            cplx_t<T>* __restrict__ gbr, cplx_t<T>* __restrict__ gzp, MaskView mask) {
  constexpr int K = NL * kL + 1;
  extern __shared__ double2 Hsm[];
  int bl = blockIdx.y;
  int b = bl / kL, l3 = bl % kL;
  const double2* Hrow = H + size_t(bl) * mask.C * K * K;
  if (LDS) {
    for (int e = threadIdx.x; e < mask.C * K * K; e += blockDim.x) Hsm[e] = Hrow[e];
    __syncthreads();
  }
  int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p >= N) return;
  const size_t hc = size_t(mask_class(mask, p)) * K * K;
  double2 V[K];
#pragma unroll
  for (int d = 0; d < NL; ++d) {
    const cplx_t<T>* src = br + (((size_t(d) * B + b) * kL + l3) * N + p) * kL;
#pragma unroll
    for (int l1 = 0; l1 < kL; ++l1) V[d * kL + l1] = to_d2<T>(src[l1]);
  }
  V[K - 1] = to_d2<T>(z[(size_t(b) * kL + l3) * N + p]);
  if (LDS) gv_apply<T, K>(Hsm + hc, V, N, B, b, l3, p, gbr, gzp);
  else gv_apply<T, K>(Hrow + hc, V, N, B, b, l3, p, gbr, gzp);
}

// Reverse of the cross Gram products of a map pair: with H [B kL][C][K][K]
// the cotangent of G = V_a V_b^H (per class), gV_a = H V_b and
// gV_b = H^H V_a; writes the branch cotangents and z parts of field rows 2b
// and 2b + 1.  H staged in LDS when LDS.
template <typename T, int NL, bool LDS>
__global__ __launch_bounds__(256) void k_gv_x(int64_t N, int Bf, const cplx_t<T>* __restrict__ br,
                                              const cplx_t<T>* __restrict__ z, const double2* __restrict__ H,
                                              cplx_t<T>* __restrict__ gbr, cplx_t<T>* __restrict__ gzp, MaskView mask) {
  constexpr int K = NL * kL + 1;
  extern __shared__ double2 Hsm[];
  const int bl = blockIdx.y;  // b kL + l3
  const int b = bl / kL, l3 = bl % kL;
  const double2* Hrow = H + size_t(bl) * mask.C * K * K;
  if (LDS) {
    for (int e = threadIdx.x; e < mask.C * K * K; e += blockDim.x) Hsm[e] = Hrow[e];
    __syncthreads();
  }
  const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p >= N) return;
  const double2* Hs = (LDS ? static_cast<const double2*>(Hsm) : Hrow) + size_t(mask_class(mask, p)) * K * K;
  double2 V[K];
  auto load = [&](int row) {
#pragma unroll
    for (int d = 0; d < NL; ++d) {
      const cplx_t<T>* src = br + (((size_t(d) * Bf + row) * kL + l3) * N + p) * kL;
#pragma unroll
      for (int l1 = 0; l1 < kL; ++l1) V[d * kL + l1] = to_d2<T>(src[l1]);
    }
    V[K - 1] = to_d2<T>(z[(size_t(row) * kL + l3) * N + p]);
  };
  auto store = [&](int row, int i, double re, double im) {
    if (i < K - 1) {
      const int d = i / kL, l1 = i % kL;
      gbr[(((size_t(d) * Bf + row) * kL + l3) * N + p) * kL + l1] = make_c<T>(T(re), T(im));
    } else {
      gzp[(size_t(row) * kL + l3) * N + p] = make_c<T>(T(re), T(im));
    }
  };
  // Channel a: gV_a[i] = sum_j H[i][j] V_b[j].
  load(2 * b + 1);
#pragma unroll 1
  for (int i = 0; i < K; ++i) {
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double2 h = Hs[i * K + j];
      re += h.x * V[j].x - h.y * V[j].y;
      im += h.x * V[j].y + h.y * V[j].x;
    }
    store(2 * b, i, re, im);
  }
  // Channel b: gV_b[j] = sum_i conj(H[i][j]) V_a[i].
  load(2 * b);
#pragma unroll 1
  for (int j = 0; j < K; ++j) {
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const double2 h = Hs[i * K + j];
      re += h.x * V[i].x + h.y * V[i].y;
      im += h.x * V[i].y - h.y * V[i].x;
    }
    store(2 * b + 1, j, re, im);
  }
}

// gz of a pair's field row (2b + c, l): gz_part + the modulus cotangent h
// z/|z| of the chain + the S1 cross term root_modulus_vjp(z_c, x) with
// x = g rn(z_other) for channel a and conj(g) rn(z_other) for channel b,
// g = s1x[(b kL + l) C + class] (complex).
template <typename T>
__global__ __launch_bounds__(256) void k_gz_x(int n, int64_t N, int Bf, const cplx_t<T>* __restrict__ z,
                                              const cplx_t<T>* __restrict__ gzp, const T* __restrict__ chain,
                                              const double2* __restrict__ s1x, cplx_t<T>* __restrict__ gz,
                                              MaskView mask) {
  const int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p >= N) return;
  const int rl = blockIdx.y;  // row kL + l
  const int row = rl / kL, l = rl % kL, b = row / 2, c = row % 2;
  const size_t at = size_t(rl) * N + p;
  const double2 zc = to_d2<T>(z[at]);
  const double2 zo = to_d2<T>(z[(size_t(row ^ 1) * kL + l) * N + p]);
  const double mag = hypot(zc.x, zc.y);
  const double den = mag == 0.0 ? 1.0 : mag;
  const double h = double(chain[((size_t(n) * Bf + row) * N + p) * kL + l]);
  double2 g = s1x[(size_t(b) * kL + l) * mask.C + mask_class(mask, p)];
  if (c == 1) g.y = -g.y;
  const double2 o = root_normalise(zo);
  const double2 x = make_double2(g.x * o.x - g.y * o.y, g.x * o.y + g.y * o.x);
  const double2 r = root_modulus_vjp(zc, x);
  const double2 gp = to_d2<T>(gzp[at]);
  gz[at] = make_c<T>(T(gp.x + (h * zc.x) / den + r.x), T(gp.y + (h * zc.y) / den + r.y));
}

// gV = H V at the small levels (few pixels, large K): a workgroup takes 32
// pixels, stages their V and the row's H in LDS, and its 8 groups of 32
// lanes split the K outputs (output i in group i mod 8), so each lane forms
// K / 8 outputs instead of K.  One mask class (unmasked).
constexpr int kGvSmallPix = 32;
// LDS: (K^2 + kGvSmallPix K) double2.
template <typename T, int NL>
__device__ __forceinline__ void gv_small_body(int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                              const cplx_t<T>* __restrict__ z, const double2* __restrict__ H,
                                              cplx_t<T>* __restrict__ gbr, cplx_t<T>* __restrict__ gzp, int64_t bx,
                                              int bl, double2* lds) {
  constexpr int K = NL * kL + 1;
  double2* Hs = lds;
  double2* Vs = lds + K * K;
  const int b = bl / kL, l3 = bl % kL;
  const int64_t p0 = bx * kGvSmallPix;
  for (int e = threadIdx.x; e < K * K; e += blockDim.x) Hs[e] = H[size_t(bl) * K * K + e];
  for (int e = threadIdx.x; e < kGvSmallPix * K; e += blockDim.x) {
    const int px = e / K, j = e % K;
    const int64_t p = p0 + px;
    double2 v = make_double2(0.0, 0.0);
    if (p < N) {
      if (j < K - 1) {
        const int d = j / kL, l1 = j % kL;
        v = to_d2<T>(br[(((size_t(d) * B + b) * kL + l3) * N + p) * kL + l1]);
      } else {
        v = to_d2<T>(z[(size_t(b) * kL + l3) * N + p]);
      }
    }
    Vs[e] = v;
  }
  __syncthreads();
  const int px = threadIdx.x % kGvSmallPix, grp = threadIdx.x / kGvSmallPix;
  const int64_t p = p0 + px;
  if (p >= N) return;
  const double2* V = Vs + px * K;
  for (int i = grp; i < K; i += 256 / kGvSmallPix) {
    double re = 0.0, im = 0.0;
#pragma unroll 8
    for (int j = 0; j < K; ++j) {
      const double2 h = Hs[i * K + j], v = V[j];
      re += h.x * v.x - h.y * v.y;
      im += h.x * v.y + h.y * v.x;
    }
    if (i < K - 1) {
      const int d = i / kL, l1 = i % kL;
      gbr[(((size_t(d) * B + b) * kL + l3) * N + p) * kL + l1] = make_c<T>(T(re), T(im));
    } else {
      gzp[(size_t(b) * kL + l3) * N + p] = make_c<T>(T(re), T(im));
    }
  }
}

template <typename T, int NL>
__global__ __launch_bounds__(256) void k_gv_small(int64_t N, int B, const cplx_t<T>* __restrict__ br,
                                                  const cplx_t<T>* __restrict__ z, const double2* __restrict__ H,
                                                  cplx_t<T>* __restrict__ gbr, cplx_t<T>* __restrict__ gzp) {
  constexpr int K = NL * kL + 1;
  __shared__ double2 lds[K * K + kGvSmallPix * K];
  gv_small_body<T, NL>(N, B, br, z, H, gbr, gzp, blockIdx.x, blockIdx.y, lds);
}

// gz of one orientation from z, its gV part, the modulus cotangent h and the
// S1/S2 cotangents a, c of its row and mask class.
template <typename T>
__device__ inline cplx_t<T> gz_value(cplx_t<T> zv, cplx_t<T> gp, double h, double a, double c) {
  const double zr = double(zv.x), zi = double(zv.y);
  const double mag = hypot(zr, zi);
  const double den = mag == 0.0 ? 1.0 : mag;
  const double re = double(gp.x) + (a * zr) / den + c * zr + (h * zr) / den;
  const double im = double(gp.y) + (a * zi) / den + c * zi + (h * zi) / den;
  return make_c<T>(T(re), T(im));
}

// gz of every orientation of one pixel (thread per (b, p)): z, its gV part
// and gz are [b][l][N] planes; the pixel's chain row ([n][b][N][kL]) is one
// vector load; the S1/S2 cotangents are per (b, l, mask class).
template <typename T>
__device__ __forceinline__ void gz_body(int n, int64_t N, int B, const cplx_t<T>* __restrict__ z,
                                        const cplx_t<T>* __restrict__ gzp, const T* __restrict__ chain,
                                        const double* __restrict__ s1, const double* __restrict__ s2,
                                        cplx_t<T>* __restrict__ gz, const MaskView& mask, int64_t bx, int b) {
  int64_t p = bx * 256 + threadIdx.x;
  if (p >= N) return;
  T hv[kL];
  Load<T>::template vec<kL>(chain + ((size_t(n) * B + b) * N + p) * kL, hv);
  const int cls = mask_class(mask, p);
#pragma unroll
  for (int l = 0; l < kL; ++l) {
    const int bl = b * kL + l;
    const size_t at = size_t(bl) * N + p;
    const size_t sc = size_t(bl) * mask.C + cls;
    gz[at] = gz_value<T>(z[at], gzp[at], double(hv[l]), s1[sc], s2[sc]);
  }
}

// Level 0's chain decimation and gz in one pass (tiled decimation
// operators): a cell's chain row (its branch adjoint part add plus the
// decimated coarser chain gy, one row per map) is read only by gz, so it
// stays in registers.
template <typename T, int TF>
__global__ __launch_bounds__(TF * TF) void k_dec_adj_gz_tiled(DecOp<T> op, const T* __restrict__ gy,
                                                              const T* __restrict__ add, int B,
                                                              const cplx_t<T>* __restrict__ z,
                                                              const cplx_t<T>* __restrict__ gzp,
                                                              const double* __restrict__ s1,
                                                              const double* __restrict__ s2,
                                                              cplx_t<T>* __restrict__ gz, MaskView mask) {
  constexpr int SW = TF / 2 + 2 * kDecT;
  __shared__ T gs[SW * SW * kL];
  T h[kL];
  int64_t p;
  int b;
  dec_adj_tile<T, kL, TF>(op, gy, add, B, gs, h, p, b);
  const int64_t N = 12 * op.npf_f;
  const int cls = mask_class(mask, p);
#pragma unroll
  for (int l = 0; l < kL; ++l) {
    const int bl = b * kL + l;
    const size_t at = size_t(bl) * N + p;
    const size_t sc = size_t(bl) * mask.C + cls;
    gz[at] = gz_value<T>(z[at], gzp[at], double(h[l]), s1[sc], s2[sc]);
  }
}

// Level 0's map-cotangent decimation with the mean / variance terms and the
// FP64 output (the map cotangent completed as by k_meanvar_grad).
template <typename T, int TF>
__global__ __launch_bounds__(TF * TF) void k_dec_adj_grad_tiled(DecOp<T> op, const T* __restrict__ gy,
                                                                const T* __restrict__ add, int B,
                                                                const T* __restrict__ x,
                                                                const double* __restrict__ gmean,
                                                                const double* __restrict__ gvar,
                                                                double* __restrict__ out, MaskView mask) {
  constexpr int SW = TF / 2 + 2 * kDecT;
  __shared__ T gs[SW * SW];
  T h[1];
  int64_t p;
  int b;
  dec_adj_tile<T, 1, TF>(op, gy, add, B, gs, h, p, b);
  const size_t at = size_t(b) * 12 * op.npf_f + p;
  const size_t sc = size_t(b) * mask.C + mask_class(mask, p);
  out[at] = double(T(double(h[0]) + gmean[sc] + gvar[sc] * double(x[at])));
}

template <typename T>
__global__ __launch_bounds__(256) void k_gz(int n, int64_t N, int B, const cplx_t<T>* __restrict__ z,
                                            const cplx_t<T>* __restrict__ gzp, const T* __restrict__ chain,
                                            const double* __restrict__ s1, const double* __restrict__ s2,
                                            cplx_t<T>* __restrict__ gz, MaskView mask) {
  gz_body<T>(n, N, B, z, gzp, chain, s1, s2, gz, mask, blockIdx.x, blockIdx.y);
}

// The map cotangent out (FP64): g0 plus the mean / variance terms, rounded
// to the engine's precision first (as the map cotangent it completes).
template <typename T>
__global__ void k_meanvar_grad(int64_t N, const T* __restrict__ x, const double* __restrict__ gmean,
                               const double* __restrict__ gvar, const T* __restrict__ g0, double* __restrict__ out,
                               MaskView mask) {
  int64_t p = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (p >= N) return;
  int b = blockIdx.y;
  size_t at = size_t(b) * N + p;
  const size_t sc = size_t(b) * mask.C + mask_class(mask, p);
  out[at] = double(T(double(g0[at]) + gmean[sc] + gvar[sc] * double(x[at])));
}

template <typename TO, typename TI>
__global__ void k_cast(int64_t n, const TI* __restrict__ in, TO* __restrict__ out) {
  int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < n) out[i] = TO(in[i]);
}

// ---------------------------------------------------------------------------
// Batched launches of the coarse levels (LevelTask, sweeps.hpp).
// ---------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void level_item(const LevelTask<T>& t, int64_t bx, int by, double2* lds) {
  using C2 = cplx_t<T>;
  const MaskView none;
  switch (t.op) {
    case kQDecFwd1:
      dec_fwd_body<T, 1>(t.dec, static_cast<const T*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQDecFwd4:
      dec_fwd_body<T, kL>(t.dec, static_cast<const T*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQConvFwd1:
      conv_fwd_split_body<T, 1, 8>(t.conv, static_cast<const T*>(t.in0), static_cast<C2*>(t.out0),
                                   static_cast<T*>(t.out1), bx, by);
      break;
    case kQConvFwd4:
      conv_fwd_split_body<T, kL, 8>(t.conv, static_cast<const T*>(t.in0), static_cast<C2*>(t.out0), nullptr, bx,
                                    by);
      break;
    case kQGram: {
      const C2* br = static_cast<const C2*>(t.in0);
      const C2* z = static_cast<const C2*>(t.in1);
      double* part = static_cast<double*>(t.out0);
      double* part2 = static_cast<double*>(t.out1);
      double* zpart = static_cast<double*>(t.out2);
      const int bxi = int(bx);
      if (t.bs == 2) {
        if (t.moments) gram_body<T, true, 2, false>(t.n, t.N, t.B, br, z, t.tpx, part, part2, zpart, none, bxi, int(t.nx), by, t.ny, lds);
        else gram_body<T, false, 2, false>(t.n, t.N, t.B, br, z, t.tpx, part, part2, zpart, none, bxi, int(t.nx), by, t.ny, lds);
      } else {
        if (t.moments) gram_body<T, true, 4, false>(t.n, t.N, t.B, br, z, t.tpx, part, part2, zpart, none, bxi, int(t.nx), by, t.ny, lds);
        else gram_body<T, false, 4, false>(t.n, t.N, t.B, br, z, t.tpx, part, part2, zpart, none, bxi, int(t.nx), by, t.ny, lds);
      }
      break;
    }
    case kQReduce:
      reduce_parts_body(static_cast<const double*>(t.in0), t.nblk, t.m, static_cast<double*>(t.out0), bx,
                        reinterpret_cast<double*>(lds));
      break;
    case kQGv: {
      const C2* br = static_cast<const C2*>(t.in0);
      const C2* z = static_cast<const C2*>(t.in1);
      const double2* H = static_cast<const double2*>(t.in2);
      C2* gbr = static_cast<C2*>(t.out0);
      C2* gzp = static_cast<C2*>(t.out1);
      switch (t.n + 1) {
#define STL_QGV(NL) \
  case NL: gv_small_body<T, NL>(t.N, t.B, br, z, H, gbr, gzp, bx, by, lds); break;
        STL_QGV(1) STL_QGV(2) STL_QGV(3) STL_QGV(4) STL_QGV(5) STL_QGV(6) STL_QGV(7) STL_QGV(8) STL_QGV(9)
        STL_QGV(10) STL_QGV(11) STL_QGV(12)
#undef STL_QGV
        default: break;
      }
      break;
    }
    case kQConvAdj1:
      conv_adj_split_body<T, 1, 8>(t.conv, static_cast<const C2*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQConvAdj4:
      conv_adj_split_body<T, kL, 8>(t.conv, static_cast<const C2*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQFix1:
      conv_adj_fix_body<T, 1>(t.conv, static_cast<const C2*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQFix4:
      conv_adj_fix_body<T, kL>(t.conv, static_cast<const C2*>(t.in0), static_cast<T*>(t.out0), bx, by);
      break;
    case kQDecAdj1:
      dec_adj_body<T, 1>(t.dec, static_cast<const T*>(t.in0), static_cast<const T*>(t.in1), static_cast<T*>(t.out0),
                         bx, by);
      break;
    case kQDecAdj4:
      dec_adj_body<T, kL>(t.dec, static_cast<const T*>(t.in0), static_cast<const T*>(t.in1),
                          static_cast<T*>(t.out0), bx, by);
      break;
    case kQGz:
      gz_body<T>(t.n, t.N, t.B, static_cast<const C2*>(t.in0), static_cast<const C2*>(t.in1),
                 static_cast<const T*>(t.in2), t.s1, t.s2, static_cast<C2*>(t.out0), none, bx, by);
      break;
    default:
      break;
  }
}

// Persistent workgroups take items in queue order; an item runs once the tasks
// its task depends on have completed.  Every workgroup leaves once the items
// are exhausted (the item counter only grows), so the grid always drains.
// Workgroup w runs items w, w + grid, ...: the items of independent tasks,
// each found by binary search over the tasks' first items.
template <typename T>
__global__ __launch_bounds__(256) void k_batch(const LevelTask<T>* __restrict__ tasks, int ntasks, int64_t items) {
  extern __shared__ double2 blds[];
  for (int64_t item = blockIdx.x; item < items; item += gridDim.x) {
    int lo = 0, hi = ntasks - 1;  // the last task whose first item <= item
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (tasks[mid].first <= item) lo = mid;
      else hi = mid - 1;
    }
    const LevelTask<T>& t = tasks[lo];
    const int64_t local = item - t.first;
    level_item<T>(t, local % t.nx, int(local / t.nx), blds);
    __syncthreads();  // the next item reuses the LDS
  }
}

}  // namespace

template <typename T>
int64_t batch_layout(LevelTask<T>* tasks, int ntasks) {
  int64_t first = 0;
  for (int i = 0; i < ntasks; ++i) {
    STL_CHECK(tasks[i].nx >= 1 && tasks[i].ny >= 1, "batch: empty task");
    tasks[i].first = first;
    first += tasks[i].nx * tasks[i].ny;
  }
  STL_CHECK(first < (int64_t(1) << 31), "batch: too many items");
  return first;
}

template <typename T>
size_t batch_lds_bytes(const LevelTask<T>* tasks, int ntasks) {
  size_t lds = 0;
  for (int i = 0; i < ntasks; ++i) {
    const LevelTask<T>& t = tasks[i];
    size_t need = 0;
    if (t.op == kQGram) {
      const int K = (t.n + 1) * kL + 1, BS = t.bs;
      const int KP = (K + BS - 1) / BS * BS, NR = KP / BS, NBLK = NR * (NR + 1) / 2;
      need = std::max(size_t(t.tpx) * KP * sizeof(double2),
                      (256 + 128) * sizeof(double2) + size_t(NBLK) * BS * BS * (sizeof(double2) + sizeof(double)));
      need = std::max(need, size_t(4 * 256) * sizeof(double));
    } else if (t.op == kQReduce) {
      need = size_t(kRedWaves) * 64 * sizeof(double);
    } else if (t.op == kQGv) {
      const int K = (t.n + 1) * kL + 1;
      need = size_t(K * K + kGvSmallPix * K) * sizeof(double2);
    }
    lds = std::max(lds, need);
  }
  return lds;
}

template <typename T>
void launch_batch(const LevelTask<T>* tasks, int ntasks, int64_t items, size_t lds, int workgroups, hipStream_t s) {
  STL_CHECK(ntasks >= 1 && items >= 1 && workgroups >= 1, "batch: nothing to run");
  STL_CHECK(lds <= 64 * 1024, "batch: the tasks need more than 64 KB of LDS");
  const int grid = int(std::min<int64_t>(items, workgroups));
  k_batch<T><<<grid, 256, lds, s>>>(tasks, ntasks, items);
  STL_HIP(hipGetLastError());
}

template <typename T, int TS>
void conv_fwd_tiled(const ConvOp<T>& op, int NI, const T* x, int rows, cplx_t<T>* y, T* mod, hipStream_t s) {
  dim3 grid(unsigned(3 * op.npf / (TS * TS) * 4 * rows));
  if (NI == 1) k_conv_fwd_tiled<T, 1, TS><<<grid, TS * TS, 0, s>>>(op, x, rows, y, mod);
  else k_conv_fwd_tiled<T, kL, TS><<<grid, TS * TS, 0, s>>>(op, x, rows, y, nullptr);
}

template <typename T, int TS>
void conv_adj_tiled(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s) {
  dim3 grid(unsigned(3 * op.npf / (TS * TS) * 4 * rows));
  if constexpr (TS == 32) {
    // 42 rows of pitch 56 (FP32) or 44 (FP64): one gauge per pass for NI = kL
    // (74 KB FP32, 116 KB FP64), 4 (FP32) or 2 (FP64) gauges for NI = 1.
    constexpr int GP1 = sizeof(T) == 8 ? 2 : 4;
    if (NI == 1) {
      k_conv_adj_tiled<T, 1, TS, GP1><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
      return;
    }
    if constexpr (sizeof(T) == 4) {
      // Two rows per workgroup (147 KB of LDS, one workgroup per CU). An odd
      // last row goes to the one-row kernel where the faces hold at least
      // 256 x 256 pixels; below that the second launch costs more than the
      // shared loads save.
      if (rows % 2 == 0 || (rows > 2 && op.npf >= 65536)) {
        const int even = rows & ~1;
        const unsigned per_row = grid.x / unsigned(rows);
        k_conv_adj_tiled<T, kL, TS, 1, 2><<<dim3(per_row * even / 2), TS * TS, 0, s>>>(op, gy, even, gx);
        if (even < rows) {
          const size_t N = size_t(12) * op.npf;
          k_conv_adj_tiled<T, kL, TS, 1><<<dim3(per_row), TS * TS, 0, s>>>(
              op, gy + size_t(even) * kL * N * kL, 1, gx + size_t(even) * N * kL);
        }
        return;
      }
    }
    k_conv_adj_tiled<T, kL, TS, 1><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
  } else {
    constexpr int GP4 = sizeof(T) == 8 ? 1 : 2;  // NI = kL: 47 KB of LDS per pass
    if (NI == 1) k_conv_adj_tiled<T, 1, TS, 4><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
    else k_conv_adj_tiled<T, kL, TS, GP4><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
  }
}

template <typename T, int TS>
void conv_fwd_faces(const ConvOp<T>& op, int NI, const T* x, int rows, cplx_t<T>* y, T* mod, hipStream_t s) {
  // FP64 with NI = kL: 2 faces per workgroup (registers); otherwise 4.
  constexpr int FW4 = sizeof(T) == 8 ? 2 : 4;
  if (NI == 1) {
    dim3 grid(unsigned(3 * op.npf / (TS * TS) * rows));
    k_conv_fwd_faces<T, 1, TS, 4><<<grid, TS * TS, 0, s>>>(op, x, rows, y, mod);
  } else {
    dim3 grid(unsigned(3 * op.npf / (TS * TS) * (4 / FW4) * rows));
    k_conv_fwd_faces<T, kL, TS, FW4><<<grid, TS * TS, 0, s>>>(op, x, rows, y, nullptr);
  }
}

// Multi-face forward on 32 x 32 tiles: NI = 1: 4 faces (28 KB of LDS FP32,
// 56 KB FP64); NI = kL: 2 faces (56 KB, FP32 only: FP64 would need twice the
// accumulator registers).
template <typename T>
void conv_fwd_faces32(const ConvOp<T>& op, int NI, const T* x, int rows, cplx_t<T>* y, T* mod, hipStream_t s) {
  if (NI == 1) {
    dim3 grid(unsigned(3 * op.npf / (32 * 32) * rows));
    k_conv_fwd_faces<T, 1, 32, 4><<<grid, 32 * 32, 0, s>>>(op, x, rows, y, mod);
    return;
  }
  if constexpr (sizeof(T) == 4) {
    dim3 grid(unsigned(3 * op.npf / (32 * 32) * 2 * rows));
    k_conv_fwd_faces<T, kL, 32, 2><<<grid, 32 * 32, 0, s>>>(op, x, rows, y, nullptr);
  }
}

template <typename T>
bool conv_fwd_has_variant(int NI, int variant) {
  if (variant == kConvFaces32) return NI == 1 || sizeof(T) == 4;
  return variant >= kConvTiled && variant < kConvFaces32;
}

template <typename T>
void launch_conv_fwd(const ConvOp<T>& op, int NI, const T* x, int rows, cplx_t<T>* y, T* mod, hipStream_t s,
                     int variant) {
  STL_CHECK(NI == 1 || NI == kL, "conv: unsupported inner width");
  STL_CHECK(conv_fwd_has_variant<T>(NI, variant), "conv: no such variant");
  if (variant == kConvSplit) {
    constexpr int LP = 8;
    dim3 grid(nblocks(12 * op.npf * LP, 256), rows);
    if (NI == 1) k_conv_fwd_split<T, 1, LP><<<grid, 256, 0, s>>>(op, x, y, mod);
    else k_conv_fwd_split<T, kL, LP><<<grid, 256, 0, s>>>(op, x, y, nullptr);
    STL_HIP(hipGetLastError());
    return;
  }
  if (op.tiled) {
    const int TS = op.order >= 4 ? 16 : 8;
    if (variant == kConvTiled32 && op.order >= 5) conv_fwd_tiled<T, 32>(op, NI, x, rows, y, mod, s);
    else if (variant == kConvFaces32 && op.order >= 5) conv_fwd_faces32<T>(op, NI, x, rows, y, mod, s);
    else if (variant == kConvFaces && TS == 16) conv_fwd_faces<T, 16>(op, NI, x, rows, y, mod, s);
    else if (variant == kConvFaces) conv_fwd_faces<T, 8>(op, NI, x, rows, y, mod, s);
    else if (TS == 16 && variant != kConvTiled8) conv_fwd_tiled<T, 16>(op, NI, x, rows, y, mod, s);
    else conv_fwd_tiled<T, 8>(op, NI, x, rows, y, mod, s);
    STL_HIP(hipGetLastError());
    return;
  }
  dim3 grid(nblocks(12 * op.npf, 256), rows);
  if (NI == 1) k_conv_fwd<T, 1><<<grid, 256, 0, s>>>(op, x, y, mod);
  else if (NI == kL) k_conv_fwd<T, kL><<<grid, 256, 0, s>>>(op, x, y, nullptr);
  else throw std::runtime_error("conv: unsupported inner width");
  STL_HIP(hipGetLastError());
}

// Multi-face adjoint within 47 KB of LDS: FP32 NI = kL: 2 faces, 1 gauge per
// pass; NI = 1: 4 faces, 2 (FP32) or 1 (FP64) gauges per pass.  FP64 with
// NI = kL has no such variant (it would need 93 KB).
template <typename T, int TS>
bool conv_adj_faces(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s) {
  if (NI == 1) {
    constexpr int GP1 = sizeof(T) == 8 ? 1 : 2;
    dim3 grid(unsigned(3 * op.npf / (TS * TS) * rows));
    k_conv_adj_faces<T, 1, TS, 4, GP1><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
    return true;
  }
  if constexpr (sizeof(T) == 4) {
    dim3 grid(unsigned(3 * op.npf / (TS * TS) * 2 * rows));
    k_conv_adj_faces<T, kL, TS, 2, 1><<<grid, TS * TS, 0, s>>>(op, gy, rows, gx);
    return true;
  }
  return false;
}

// Multi-face adjoint on 32 x 32 tiles (147 KB of LDS at most, one
// workgroup per CU): NI = kL: 2 faces, 1 gauge per pass (FP32 only); NI = 1:
// 4 faces, 1 gauge per pass (74 KB in FP32, two workgroups per CU).
// Window positions of the next pass prefetched into registers by the 32 x 32
// multi-face adjoints: the FP32 branch adjoint (one workgroup per CU) takes 4
// of its 5, 32 VGPRs.  The FP32 z adjoint runs two workgroups per CU within 64
// VGPRs (a prefetch spills there); the FP64 z adjoint measured no gain.
template <typename T, int NI>
constexpr int adj_prefetch() { return NI == 1 ? 0 : 4; }
template <typename T>
void conv_adj_faces32(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s) {
  if (NI == 1) {
    dim3 grid(unsigned(3 * op.npf / (32 * 32) * rows));
    k_conv_adj_faces<T, 1, 32, 4, 1, 4, adj_prefetch<T, 1>()><<<grid, 32 * 32, 0, s>>>(op, gy, rows, gx);
    return;
  }
  if constexpr (sizeof(T) == 4) {
    dim3 grid(unsigned(3 * op.npf / (32 * 32) * 2 * rows));
    k_conv_adj_faces<T, kL, 32, 2, 1, 4, adj_prefetch<T, kL>()><<<grid, 32 * 32, 0, s>>>(op, gy, rows, gx);
  }
}

template <typename T>
bool conv_adj_has_variant(int NI, int variant) {
  if (variant == kConvFaces || variant == kConvFaces32) return NI == 1 || sizeof(T) == 4;
  return variant >= kConvTiled && variant <= kConvFaces32;
}

template <typename T>
void conv_adj_gather(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s, int variant) {
  if (variant == kConvSplit) {
    constexpr int LP = 8;
    dim3 grid(nblocks(12 * op.npf * LP, 256), rows);
    if (NI == 1) k_conv_adj_split<T, 1, LP><<<grid, 256, 0, s>>>(op, gy, gx);
    else k_conv_adj_split<T, kL, LP><<<grid, 256, 0, s>>>(op, gy, gx);
    STL_HIP(hipGetLastError());
    return;
  }
  if (op.tiled) {
    const int TS = op.order >= 4 ? 16 : 8;
    if (variant == kConvTiled32 && op.order >= 5) conv_adj_tiled<T, 32>(op, NI, gy, rows, gx, s);
    else if (variant == kConvFaces32 && op.order >= 5) conv_adj_faces32<T>(op, NI, gy, rows, gx, s);
    else if (variant == kConvFaces && TS == 16) conv_adj_faces<T, 16>(op, NI, gy, rows, gx, s);
    else if (variant == kConvFaces) conv_adj_faces<T, 8>(op, NI, gy, rows, gx, s);
    else if (TS == 16 && variant != kConvTiled8) conv_adj_tiled<T, 16>(op, NI, gy, rows, gx, s);
    else conv_adj_tiled<T, 8>(op, NI, gy, rows, gx, s);
    STL_HIP(hipGetLastError());
    return;
  }
  dim3 grid(nblocks(12 * op.npf, 256), rows);
  if (NI == 1) k_conv_adj<T, 1><<<grid, 256, 0, s>>>(op, gy, gx);
  else if (NI == kL) k_conv_adj<T, kL><<<grid, 256, 0, s>>>(op, gy, gx);
  else throw std::runtime_error("conv adjoint: unsupported inner width");
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_conv_adj(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s,
                     int variant) {
  STL_CHECK(NI == 1 || NI == kL, "conv adjoint: unsupported inner width");
  STL_CHECK(conv_adj_has_variant<T>(NI, variant), "conv adjoint: no such variant");
  conv_adj_gather(op, NI, gy, rows, gx, s, variant);
  if (op.nfix > 0) {
    dim3 grid(nblocks(int64_t(op.nfix) * NI, 256), rows);
    if (NI == 1) k_conv_adj_fix<T, 1><<<grid, 256, 0, s>>>(op, gy, gx);
    else k_conv_adj_fix<T, kL><<<grid, 256, 0, s>>>(op, gy, gx);
    STL_HIP(hipGetLastError());
  }
}

template <typename T>
void launch_dec_fwd(const DecOp<T>& op, int NI, const T* x, int rows, T* y, hipStream_t s) {
  if (op.tiled) {
    STL_CHECK(NI == 1 || NI == kL, "decimation: unsupported inner width");
    STL_CHECK(op.S >= 1, "decimation: empty operator store");
    if (op.order_c >= 4) {
      dim3 grid(unsigned(op.Rc / 256 * 4 * rows));
      if (NI == 1) k_dec_fwd_tiled<T, 1, 16><<<grid, 256, 0, s>>>(op, x, rows, y);
      else k_dec_fwd_tiled<T, kL, 16><<<grid, 256, 0, s>>>(op, x, rows, y);
    } else {
      dim3 grid(unsigned(op.Rc / 64 * 4 * rows));
      if (NI == 1) k_dec_fwd_tiled<T, 1, 8><<<grid, 64, 0, s>>>(op, x, rows, y);
      else k_dec_fwd_tiled<T, kL, 8><<<grid, 64, 0, s>>>(op, x, rows, y);
    }
    STL_HIP(hipGetLastError());
    return;
  }
  dim3 grid(nblocks(op.Rc * 4, 256), rows);
  if (NI == 1) k_dec_fwd<T, 1><<<grid, 256, 0, s>>>(op, x, y);
  else if (NI == kL) k_dec_fwd<T, kL><<<grid, 256, 0, s>>>(op, x, y);
  else throw std::runtime_error("decimation: unsupported inner width");
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_dec_adj(const DecOp<T>& op, int NI, const T* gy, const T* add, int rows, T* gx, hipStream_t s) {
  if (op.ttiled) {
    STL_CHECK(NI == 1 || NI == kL, "decimation adjoint: unsupported inner width");
    STL_CHECK(op.Tn >= 1, "decimation adjoint: empty operator store");
    dim3 grid(unsigned(op.Rf / 256 * 4 * rows));
    if (NI == 1) k_dec_adj_tiled<T, 1, 16><<<grid, 256, 0, s>>>(op, gy, add, rows, gx);
    else k_dec_adj_tiled<T, kL, 16><<<grid, 256, 0, s>>>(op, gy, add, rows, gx);
    STL_HIP(hipGetLastError());
    return;
  }
  dim3 grid(nblocks(op.Rf * 4, 256), rows);
  if (NI == 1) k_dec_adj<T, 1><<<grid, 256, 0, s>>>(op, gy, add, gx);
  else if (NI == kL) k_dec_adj<T, kL><<<grid, 256, 0, s>>>(op, gy, add, gx);
  else throw std::runtime_error("decimation adjoint: unsupported inner width");
  STL_HIP(hipGetLastError());
}

int gram_tile(int n) {
  int K = (n + 1) * kL + 1;
  int tpx = 24576 / (16 * K);
  return std::max(16, std::min(256, tpx));
}

template <typename T>
int gram_blocks(int n, int64_t N, int64_t rows) {
  const void* fn = nullptr;
  int64_t step = 0;  // pixels per workgroup pass
  switch (n + 1) {
    case 1: fn = reinterpret_cast<const void*>(&k_gram_reg<T, 5, false, false>); step = 256; break;
    case 2: fn = reinterpret_cast<const void*>(&k_gram_reg<T, 9, false, false>); step = 256; break;
#define STL_GB(RR) \
  case RR: fn = reinterpret_cast<const void*>(&k_gram_cls<T, RR, false, false>); step = 4 * GramClasses<RR>::S; break;
    STL_GB(3) STL_GB(4) STL_GB(5) STL_GB(6)
#undef STL_GB
    default: break;
  }
  if (fn == nullptr) {  // the LDS-tiled kernel: tiles of gram_tile(n) pixels
    const int64_t tiles = (N + gram_tile(n) - 1) / gram_tile(n);
    return int(std::min<int64_t>(tiles, std::max<int64_t>(1, 2048 / (rows * kL))));
  }
  int dev = 0, cus = 0, per_cu = 0;
  STL_HIP(hipGetDevice(&dev));
  STL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  STL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
  const int64_t target = int64_t(std::max(per_cu, 1)) * std::max(cus, 1);
  const int64_t nb = std::min((target + rows * kL - 1) / (rows * kL), (N + step - 1) / step);
  return int(std::max<int64_t>(nb, 1));
}


template <typename T>
void launch_gram(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, bool moments, int nblk,
                 double* part, double* part2, double* zpart, const MaskView& mask, hipStream_t s) {
  const int K = (n + 1) * kL + 1;
  dim3 grid(nblk, B * mask.K * kL);
  const bool masked = mask.cls != nullptr;
#define STL_GRAM_REG(KT, MOM)                                                                      \
  {                                                                                                \
    if (masked) k_gram_reg<T, KT, MOM, true><<<grid, 256, 0, s>>>(n, N, B, branch, z, part, part2, zpart, mask);  \
    else k_gram_reg<T, KT, MOM, false><<<grid, 256, 0, s>>>(n, N, B, branch, z, part, part2, zpart, mask);        \
    STL_HIP(hipGetLastError());                                                                    \
    return;                                                                                        \
  }
  if (K == 5 && moments) STL_GRAM_REG(5, true)
  if (K == 5) STL_GRAM_REG(5, false)
  if (K == 9 && !moments) STL_GRAM_REG(9, false)
#undef STL_GRAM_REG
#define STL_GRAM_CLS(RR)                                                                                   \
  case RR:                                                                                                 \
    if (moments) {                                                                                         \
      if (masked) k_gram_cls<T, RR, true, true><<<grid, 256, 0, s>>>(N, B, branch, z, part, part2, zpart, mask);   \
      else k_gram_cls<T, RR, true, false><<<grid, 256, 0, s>>>(N, B, branch, z, part, part2, zpart, mask);         \
    } else {                                                                                               \
      if (masked) k_gram_cls<T, RR, false, true><<<grid, 256, 0, s>>>(N, B, branch, z, part, part2, zpart, mask);  \
      else k_gram_cls<T, RR, false, false><<<grid, 256, 0, s>>>(N, B, branch, z, part, part2, zpart, mask);        \
    }                                                                                                      \
    STL_HIP(hipGetLastError());                                                                            \
    return;
  switch (n + 1) {
    STL_GRAM_CLS(3) STL_GRAM_CLS(4) STL_GRAM_CLS(5) STL_GRAM_CLS(6)
    default: break;
  }
#undef STL_GRAM_CLS
  // Small K: 2 x 2 blocks waste fewer padded products; larger K: 4 x 4 blocks
  // halve the LDS reads per product.
  const int BS = K <= 13 ? 2 : 4;
  const int KP = (K + BS - 1) / BS * BS, NR = KP / BS, NBLK = NR * (NR + 1) / 2;
  STL_CHECK(NBLK <= 256, "gram: too many levels");
  const int tpx = gram_tile(n);
  size_t lds = std::max(size_t(tpx) * KP * sizeof(double2),
                        (256 + 128) * sizeof(double2) + size_t(NBLK) * BS * BS * (sizeof(double2) + sizeof(double)));
  lds = std::max(lds, size_t(4 * 256) * sizeof(double));  // the S1/S2 block sums
#define STL_GRAM(MOM, BSZ)                                                                          \
  {                                                                                                 \
    if (masked) k_gram<T, MOM, BSZ, true><<<grid, 256, lds, s>>>(n, N, B, branch, z, tpx, part, part2, zpart, mask);  \
    else k_gram<T, MOM, BSZ, false><<<grid, 256, lds, s>>>(n, N, B, branch, z, tpx, part, part2, zpart, mask);       \
  }
  if (BS == 2) {
    if (moments) STL_GRAM(true, 2) else STL_GRAM(false, 2)
  } else {
    if (moments) STL_GRAM(true, 4) else STL_GRAM(false, 4)
  }
#undef STL_GRAM
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_gram_x(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, bool moments, int nblk,
                   double* part, double* part2, const MaskView& mask, hipStream_t s) {
  const int K = (n + 1) * kL + 1;
  const int BS = K <= 13 ? 2 : 4;
  const int KP = (K + BS - 1) / BS * BS, NR = KP / BS;
  STL_CHECK(NR * NR <= 256, "cross gram: too many levels");
  const int tpx = gram_tile(n);
  const size_t lds = std::max(size_t(2) * tpx * KP * sizeof(double2), size_t(256) * (sizeof(double2) + sizeof(double)));
  STL_CHECK(lds <= 64 * 1024, "cross gram: LDS tile too large");
  dim3 grid(nblk, B * mask.K * kL);
  const bool masked = mask.cls != nullptr;
  const int Bf = 2 * B;
#define STL_GRAMX(MOM, BSZ)                                                                                         \
  {                                                                                                                 \
    if (masked) k_gram_x<T, MOM, BSZ, true><<<grid, 256, lds, s>>>(n, N, Bf, branch, z, tpx, part, part2, mask);   \
    else k_gram_x<T, MOM, BSZ, false><<<grid, 256, lds, s>>>(n, N, Bf, branch, z, tpx, part, part2, mask);         \
  }
  if (BS == 2) {
    if (moments) STL_GRAMX(true, 2) else STL_GRAMX(false, 2)
  } else {
    if (moments) STL_GRAMX(true, 4) else STL_GRAMX(false, 4)
  }
#undef STL_GRAMX
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_zcross(int64_t N, int B, const cplx_t<T>* z, int nblk, double* part, const MaskView& mask, hipStream_t s) {
  k_zcross<T><<<dim3(nblk, B * mask.K * kL), 256, 0, s>>>(N, z, part, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_gv_x(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, const double2* H,
                 cplx_t<T>* gbranch, cplx_t<T>* gz_part, const MaskView& mask, hipStream_t s) {
  dim3 grid(nblocks(N, 256), B * kL);
  const int K = (n + 1) * kL + 1;
  const size_t hbytes = size_t(mask.C) * K * K * sizeof(double2);
  const bool lds = hbytes <= 64 * 1024;
  const int Bf = 2 * B;
  switch (n + 1) {
#define STL_GVX(NL)                                                                                     \
  case NL:                                                                                              \
    if (lds) k_gv_x<T, NL, true><<<grid, 256, hbytes, s>>>(N, Bf, branch, z, H, gbranch, gz_part, mask); \
    else k_gv_x<T, NL, false><<<grid, 256, 0, s>>>(N, Bf, branch, z, H, gbranch, gz_part, mask);         \
    break;
    STL_GVX(1) STL_GVX(2) STL_GVX(3) STL_GVX(4) STL_GVX(5) STL_GVX(6) STL_GVX(7) STL_GVX(8) STL_GVX(9)
    STL_GVX(10) STL_GVX(11) STL_GVX(12)
#undef STL_GVX
    default: throw std::runtime_error("gv cross: too many levels");
  }
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_gz_x(int n, int64_t N, int B, const cplx_t<T>* z, const cplx_t<T>* gz_part, const T* chain,
                 const double2* s1x, cplx_t<T>* gz, const MaskView& mask, hipStream_t s) {
  k_gz_x<T><<<dim3(nblocks(N, 256), 2 * B * kL), 256, 0, s>>>(n, N, 2 * B, z, gz_part, chain, s1x, gz, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_zstats(int64_t N, int B, const cplx_t<T>* z, int nblk, double* part, const MaskView& mask,
                   hipStream_t s) {
  k_zstats<T><<<dim3(nblk, B * mask.K * kL), 256, 0, s>>>(N, z, part, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_meanvar(int64_t N, int B, const T* x, int nblk, double* part, const MaskView& mask, hipStream_t s) {
  k_meanvar<T><<<dim3(nblk, B * mask.K), 256, 0, s>>>(N, x, part, mask);
  STL_HIP(hipGetLastError());
}

void launch_assemble(int np, int B, int K, int levels, const StatEntry* entries, const double* const* src,
                     const double* nd, bool moments, double2* values, double* variance, hipStream_t s) {
  k_assemble<<<nblocks(int64_t(np) * B * K, 256), 256, 0, s>>>(np, B * K, K, levels, entries, src, nd, moments,
                                                                values, variance);
  STL_HIP(hipGetLastError());
}

void launch_scatter_cot(int ntargets, int B, int K, int levels, int np, const RevTarget* targets,
                        const RevTerm* terms, const double* nd, const double2* cot, double* const* dst,
                        hipStream_t s) {
  k_scatter_cot<<<nblocks(int64_t(ntargets) * B * K, 256), 256, 0, s>>>(ntargets, B * K, K, levels, np, targets,
                                                                         terms, nd, cot, dst);
  STL_HIP(hipGetLastError());
}

void launch_mask_combine(int B, int K, int R, int C, int64_t E, const uint64_t* pat, const double* in, double* out,
                         hipStream_t s) {
  k_mask_combine<<<nblocks(int64_t(B) * R * C * E, 256), 256, 0, s>>>(B, K, R, C, E, pat, in, out);
  STL_HIP(hipGetLastError());
}

void launch_reduce_parts(const double* part, int nblk, int64_t m, double* out, hipStream_t s) {
  k_reduce_parts<<<nblocks(m, 64), 64 * kRedWaves, 0, s>>>(part, nblk, m, out);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_gv(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, const double2* H,
               cplx_t<T>* gbranch, cplx_t<T>* gz_part, const MaskView& mask, hipStream_t s) {
  dim3 grid(nblocks(N, 256), B * kL);
  const int K = (n + 1) * kL + 1;
  if (mask.C == 1 && n >= 3 && N <= 49152) {
    // Small levels: split the K outputs over lanes (k_gv_small).
    dim3 sgrid(unsigned((N + kGvSmallPix - 1) / kGvSmallPix), B * kL);
    switch (n + 1) {
#define STL_GVS(NL) \
  case NL: k_gv_small<T, NL><<<sgrid, 256, 0, s>>>(N, B, branch, z, H, gbranch, gz_part); break;
      STL_GVS(4) STL_GVS(5) STL_GVS(6) STL_GVS(7) STL_GVS(8) STL_GVS(9) STL_GVS(10) STL_GVS(11) STL_GVS(12)
#undef STL_GVS
      default: throw std::runtime_error("gv: too many levels");
    }
    STL_HIP(hipGetLastError());
    return;
  }
  const size_t hbytes = size_t(mask.C) * K * K * sizeof(double2);
  const bool lds = hbytes <= 64 * 1024;
  switch (n + 1) {
#define STL_GV(NL)                                                                                  \
  case NL:                                                                                          \
    if (lds) k_gv<T, NL, true><<<grid, 256, hbytes, s>>>(N, B, branch, z, H, gbranch, gz_part, mask); \
    else k_gv<T, NL, false><<<grid, 256, 0, s>>>(N, B, branch, z, H, gbranch, gz_part, mask);         \
    break;
    STL_GV(1) STL_GV(2) STL_GV(3) STL_GV(4) STL_GV(5) STL_GV(6) STL_GV(7) STL_GV(8) STL_GV(9) STL_GV(10)
    STL_GV(11) STL_GV(12)
#undef STL_GV
    default: throw std::runtime_error("gv: too many levels");
  }
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_gz(int n, int64_t N, int B, const cplx_t<T>* z, const cplx_t<T>* gz_part, const T* chain,
               const double* s1, const double* s2, cplx_t<T>* gz, const MaskView& mask, hipStream_t s) {
  k_gz<T><<<dim3(nblocks(N, 256), B), 256, 0, s>>>(n, N, B, z, gz_part, chain, s1, s2, gz, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_meanvar_grad(int64_t N, int B, const T* x, const double* gmean, const double* gvar, const T* g0,
                         double* out, const MaskView& mask, hipStream_t s) {
  k_meanvar_grad<T><<<dim3(nblocks(N, 256), B), 256, 0, s>>>(N, x, gmean, gvar, g0, out, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_dec_adj_gz(const DecOp<T>& op, const T* gy, const T* add, int B, const cplx_t<T>* z,
                       const cplx_t<T>* gz_part, const double* s1, const double* s2, cplx_t<T>* gz,
                       const MaskView& mask, hipStream_t s) {
  STL_CHECK(op.ttiled && op.Tn >= 1, "decimation adjoint with gz: needs the tiled operator");
  dim3 grid(unsigned(op.Rf / 256 * 4 * B));
  k_dec_adj_gz_tiled<T, 16><<<grid, 256, 0, s>>>(op, gy, add, B, z, gz_part, s1, s2, gz, mask);
  STL_HIP(hipGetLastError());
}

template <typename T>
void launch_dec_adj_grad(const DecOp<T>& op, const T* gy, const T* add, int B, const T* x, const double* gmean,
                         const double* gvar, double* out, const MaskView& mask, hipStream_t s) {
  STL_CHECK(op.ttiled && op.Tn >= 1, "decimation adjoint with the map cotangent: needs the tiled operator");
  dim3 grid(unsigned(op.Rf / 256 * 4 * B));
  k_dec_adj_grad_tiled<T, 16><<<grid, 256, 0, s>>>(op, gy, add, B, x, gmean, gvar, out, mask);
  STL_HIP(hipGetLastError());
}

template <typename TO, typename TI>
void launch_cast(int64_t n, const TI* in, TO* out, hipStream_t s) {
  k_cast<TO, TI><<<nblocks(n, 256), 256, 0, s>>>(n, in, out);
  STL_HIP(hipGetLastError());
}

#define STL_INSTANTIATE(T)                                                                                     \
  template void launch_conv_fwd<T>(const ConvOp<T>&, int, const T*, int, cplx_t<T>*, T*, hipStream_t, int);        \
  template void launch_conv_adj<T>(const ConvOp<T>&, int, const cplx_t<T>*, int, T*, hipStream_t, int);     \
  template bool conv_adj_has_variant<T>(int, int);            \
  template bool conv_fwd_has_variant<T>(int, int);            \
  template void launch_dec_fwd<T>(const DecOp<T>&, int, const T*, int, T*, hipStream_t);                      \
  template void launch_dec_adj<T>(const DecOp<T>&, int, const T*, const T*, int, T*, hipStream_t);            \
  template void launch_gram<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, bool, int, double*,     \
                               double*, double*, const MaskView&, hipStream_t);                               \
  template int gram_blocks<T>(int, int64_t, int64_t);                                                          \
  template void launch_dec_adj_gz<T>(const DecOp<T>&, const T*, const T*, int, const cplx_t<T>*,               \
                                     const cplx_t<T>*, const double*, const double*, cplx_t<T>*,              \
                                     const MaskView&, hipStream_t);                                           \
  template void launch_dec_adj_grad<T>(const DecOp<T>&, const T*, const T*, int, const T*, const double*,      \
                                       const double*, double*, const MaskView&, hipStream_t);                 \
  template void launch_zstats<T>(int64_t, int, const cplx_t<T>*, int, double*, const MaskView&, hipStream_t); \
  template void launch_meanvar<T>(int64_t, int, const T*, int, double*, const MaskView&, hipStream_t);        \
  template void launch_gv<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, const double2*,           \
                             cplx_t<T>*, cplx_t<T>*, const MaskView&, hipStream_t);                           \
  template void launch_gz<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, const T*, const double*,  \
                             const double*, cplx_t<T>*, const MaskView&, hipStream_t);                        \
  template void launch_meanvar_grad<T>(int64_t, int, const T*, const double*, const double*, const T*,        \
                                       double*, const MaskView&, hipStream_t);                                \
  template void launch_gram_x<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, bool, int, double*,   \
                                 double*, const MaskView&, hipStream_t);                                      \
  template void launch_zcross<T>(int64_t, int, const cplx_t<T>*, int, double*, const MaskView&, hipStream_t); \
  template void launch_gv_x<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, const double2*,         \
                               cplx_t<T>*, cplx_t<T>*, const MaskView&, hipStream_t);                         \
  template void launch_gz_x<T>(int, int64_t, int, const cplx_t<T>*, const cplx_t<T>*, const T*,               \
                               const double2*, cplx_t<T>*, const MaskView&, hipStream_t);

STL_INSTANTIATE(double)
STL_INSTANTIATE(float)
#undef STL_INSTANTIATE

#define STL_BATCH_INSTANTIATE(T)                                                                   \
  template int64_t batch_layout<T>(LevelTask<T>*, int);                                           \
  template size_t batch_lds_bytes<T>(const LevelTask<T>*, int);                                   \
  template void launch_batch<T>(const LevelTask<T>*, int, int64_t, size_t, int, hipStream_t);
STL_BATCH_INSTANTIATE(double)
STL_BATCH_INSTANTIATE(float)
#undef STL_BATCH_INSTANTIATE

template void launch_cast<double, float>(int64_t, const float*, double*, hipStream_t);
template void launch_cast<float, double>(int64_t, const double*, float*, hipStream_t);

}  // namespace stl
