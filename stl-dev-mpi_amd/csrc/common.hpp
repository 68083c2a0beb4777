// Shared definitions of the MI355X ScatCov library: error handling, the
// face-rotation symmetry of the pyramid operators and the operator layouts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace stl {

#define STL_HIP(call)                                                                                   \
  do {                                                                                                  \
    hipError_t err_ = (call);                                                                           \
    if (err_ != hipSuccess)                                                                             \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(err_) + " at " + __FILE__ + \
                               ":" + std::to_string(__LINE__) + " (" #call ")");                        \
  } while (0)

#define STL_CHECK(cond, msg)                                    \
  do {                                                          \
    if (!(cond)) throw std::runtime_error(std::string(msg));    \
  } while (0)

// Number of orientations / gauges handled by the kernels (L = 4).
constexpr int kL = 4;
// Complex outputs per conv row: kL gauges x (re, im).
constexpr int kCoefPerSlot = 2 * kL;
// Slot capacity of a conv row (distinct cells read with a nonzero weight:
// at most 49 for kernel 5, L = 4, gauge cosmo, at every nside).
constexpr int kMaxSlots = 64;

// ---------------------------------------------------------------------------
// Face-rotation symmetry.  HEALPix is invariant under the rotation by pi/2
// about the polar axis, which maps face f to (f & ~3) | ((f + 1) & 3) with the
// same in-face (x, y).  The stencil of gauge g at the rotated pixel is that of
// gauge (g + 2) mod 4 at the original one, point-reflected when g + 2 wraps;
// the Gabor kernel satisfies k(-x, -y) = conj(k(x, y)), so a point reflection
// conjugates the folded coefficient.  For a shift of k faces:
//   A[rot^k(p), rot^k(c)][g] = A[p, c][g ^ (2 (k & 1))], conjugated when
//   k = 1 and g < 2, k = 2, or k = 3 and g >= 2.
// Operators are therefore built for the faces 0, 4 and 8 ("rep" rows); the
// decimation is stored that way.
//
// North-south mirror.  z -> -z maps face f < 4 to f + 8 and the in-face
// (x, y) to (nside - 1 - y, nside - 1 - x); it keeps the longitude, so every
// gauge keeps its index, and it reverses orientations, so the folded
// coefficient is conjugated:
//   A[M(p), M(c)][g] = conj(A[p, c][g])   for every row p of the faces 0-3.
// The convolution is therefore stored for the faces 0 and 4 only; rows of the
// faces 8-11 are the mirrors of those of the faces 0-3.  The rows on the
// equator ring (faces 4-7, x + y = nside - 1, z = 0) are their own mirrors but
// not mirror-symmetric, so the columns of south cells they read are not the
// mirrors of the north columns: those entries carry a correction list.
// ---------------------------------------------------------------------------

__host__ __device__ inline int rot_face(int f, int k) { return (f & ~3) | ((f + k) & 3); }

__host__ __device__ inline int64_t rot_pix(int64_t pix, int k, int order) {
  int64_t npf = int64_t(1) << (2 * order);
  int f = int(pix >> (2 * order));
  return int64_t(rot_face(f, k)) * npf + (pix & (npf - 1));
}

__host__ __device__ inline int gauge_src(int k, int g) { return g ^ (2 * (k & 1)); }

__host__ __device__ inline bool gauge_conj(int k, int g) {
  return (k == 1 && g < 2) || k == 2 || (k == 3 && g >= 2);
}

// Rep index of a pixel lying in a rep face (0, 4 or 8): (face / 4) * npf + local.
__host__ __device__ inline int64_t rep_index(int64_t pix, int order) {
  int64_t npf = int64_t(1) << (2 * order);
  return (pix >> (2 * order)) / 4 * npf + (pix & (npf - 1));
}

// Pixel of a rep index: face 4 * (r / npf), local r % npf.
__host__ __device__ inline int64_t rep_pixel(int64_t r, int order) {
  int64_t npf = int64_t(1) << (2 * order);
  return (r / npf) * 4 * npf + (r % npf);
}

// Mirror of a local NESTED index: (x, y) -> (nside - 1 - y, nside - 1 - x) is
// the complement of every bit followed by the exchange of the x and y bits.
__host__ __device__ inline int64_t mirror_loc(int64_t loc, int order) {
  int64_t v = loc ^ ((int64_t(1) << (2 * order)) - 1);
  return ((v & 0x5555555555555555LL) << 1) | ((v >> 1) & 0x5555555555555555LL);
}

__host__ __device__ inline int mirror_face(int f) { return f < 4 ? f + 8 : (f >= 8 ? f - 8 : f); }

__host__ __device__ inline int64_t mirror_pix(int64_t pix, int order) {
  int64_t npf = int64_t(1) << (2 * order);
  return int64_t(mirror_face(int(pix >> (2 * order)))) * npf + mirror_loc(pix & (npf - 1), order);
}

// ---------------------------------------------------------------------------
// Operator layouts (device memory, ELL by slot: entry (s, r) at s * R + r so
// that consecutive rep rows read consecutive memory for one slot).
// ---------------------------------------------------------------------------

// Half-width of the stencil window in the extended face frame: every cell read
// by a row, and every row reading a cell, lies within max-norm kWin of it at
// nside >= 8 (kernel 5, L = 4).
constexpr int kWin = 5;

// Largest nside: pixel ids are stored as int32 (12 nside^2 < 2^31) and the
// kernels' bit interleaving works on 32-bit local indices.
constexpr int kMaxNside = 8192;

// Row pitch (in elements) of the LDS windows of the tiled adjoints, per tile
// side and element size.  gfx950's LDS serves 64-bit reads 32 lanes at a
// time over 64 four-byte banks: the lanes of a wave hold an 8 x 8 block of
// cells, so lanes 0-31 read 4 window rows of 8 elements, which fill the 64
// banks once the pitch is 8 or 24 mod 32 (measured: 187 B/clk per CU at
// pitch 56 against 118 at pitch 44, 8-byte elements at uniform offsets).
// 16-byte elements (FP64 complex) keep pitches 4 or 12 mod 16.
constexpr int lds_pitch(int ts, int elem_bytes = 8) {
  return elem_bytes == 8 ? (ts == 32 ? 56 : (ts == 16 ? 40 : 24)) : (ts == 32 ? 44 : (ts == 16 ? 28 : 20));
}

// Offset (dx, dy) packed as two int8 in an int16.
__host__ __device__ inline int16_t pack_off(int dx, int dy) {
  return int16_t((dx & 0xFF) | ((dy & 0xFF) << 8));
}
__host__ __device__ inline int off_dx(int16_t o) { return int(int8_t(o & 0xFF)); }
__host__ __device__ inline int off_dy(int16_t o) { return int(int8_t((o >> 8) & 0xFF)); }

// Wavelet convolution of one level, complex outputs of kL gauges.  Rows and
// columns are stored for the faces 0 and 4; the "virtual" rows 2 npf .. 3 npf
// of the kernels are the mirrors of face 0's rows (the stored row r - 2 npf,
// written to the face-8 pixel mirror_loc(r - 2 npf)).
template <typename T>
struct ConvOp {
  int order = 0;          // log2 nside of the level
  int64_t npf = 0;        // pixels per face
  int64_t R = 0;          // stored rows = 2 npf
  int S = 0;              // slot capacity of the forward rows
  int Tn = 0;             // slot capacity of the transposed rows
  bool tiled = false;     // every offset within kWin and nside >= 8: LDS-tiled kernels
  const uint8_t* cnt = nullptr;   // [R] slots per rep row
  const int32_t* cell = nullptr;  // [S][R] NESTED cell read (rep row frame)
  const int16_t* off = nullptr;   // [S][R] offset of that cell in the row's extended face frame
  const T* coef = nullptr;        // [S][R][kL][2]
  const uint8_t* tcnt = nullptr;  // [R] rows reading each rep cell
  const int32_t* trow = nullptr;  // [Tn][R] NESTED row id (any face), by window offset (dy, dx)
  const int16_t* toff = nullptr;  // [Tn][R] offset of that row in the cell's extended face frame
  const T* tcoef = nullptr;       // [kL][Tn][R][2] that row's coefficient (its own gauges), gauge planes
  // Tiled levels: the transposed slots are union slots of blocks of 64 cells
  // (an aligned 8 x 8 NESTED block, one wave of the tiled adjoints): slot s
  // of every cell of block w reads the same window offset (dy, dx), the cells
  // without a row there with a zero coefficient; tcnt holds the block's count
  // tbcnt[w].  The offsets in the linear form dy * lds_pitch(TS) + dx of the
  // 32 x 32, 16 x 16 and 8 x 8 tiles' windows.  Null on untiled levels.
  const int32_t* tbcnt = nullptr;  // [R / 64]
  // The same for the forward rows (fbcnt; cnt holds the block's count), in
  // the linear form dy * SW + dx of the windows (SW = TS + 2 kWin).
  const int32_t* fbcnt = nullptr;  // [R / 64]
  const int32_t* fbo32 = nullptr;  // [R / 64][S]
  const int32_t* fbo16 = nullptr;  // [R / 64][S]
  const int32_t* fbo8 = nullptr;   // [R / 64][S]
  const int32_t* tbo32 = nullptr;  // [R / 64][Tn]
  const int32_t* tbo16 = nullptr;  // [R / 64][Tn]
  const int32_t* tbo8 = nullptr;   // [R / 64][Tn]
  // Adjoint correction of the mirrored columns: for each south cell read by an
  // equator-ring row, the true coefficients minus the mirrored ones, summed
  // after the gather (CSR by cell, rows ascending).
  int nfix = 0;
  const int32_t* fix_cell = nullptr;  // [nfix] NESTED cell (faces 8-11)
  const int32_t* fix_ptr = nullptr;   // [nfix + 1]
  const int32_t* fix_row = nullptr;   // [entries] NESTED row (faces 4-7)
  const T* fix_coef = nullptr;        // [entries][kL][2] row's gauges
};

// Decimation from level n (fine) to level n + 1 (coarse), real weights.
template <typename T>
struct DecOp {
  int order_f = 0, order_c = 0;
  int64_t npf_f = 0, npf_c = 0;
  int64_t Rc = 0, Rf = 0;          // rep coarse rows, rep fine cells
  int S = 0, Tn = 0;
  const uint8_t* cnt = nullptr;    // [Rc]
  const int32_t* col = nullptr;    // [S][Rc] fine cell (rep frame), ascending
  const T* val = nullptr;          // [S][Rc]
  const uint8_t* tcnt = nullptr;   // [Rf]
  const int32_t* trow = nullptr;   // [Tn][Rf] coarse row id (any face), ascending
  const T* tval = nullptr;         // [Tn][Rf]
  // LDS-tiled forms: the fine cell's offset from the coarse row's first child
  // (fine extended frame, within [kDecLo, kDecHi]) and the coarse row's offset
  // from the fine cell's parent (coarse extended frame, within +-kDecT).
  bool tiled = false, ttiled = false;
  const int16_t* off = nullptr;    // [S][Rc]
  const int16_t* toff = nullptr;   // [Tn][Rf]
};

// Offset bounds of the decimation entries used by the tiled kernels: the fine
// candidates of a coarse row are the children of its +-2 coarse neighbours.
constexpr int kDecLo = -4, kDecHi = 5, kDecT = 2;

}  // namespace stl
