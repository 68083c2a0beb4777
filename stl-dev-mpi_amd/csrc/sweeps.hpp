// Kernels of the forward sweep and its VJP on the pyramid (templated on the
// bulk precision T = double or float; sky-wide sums always in FP64).
//
// Field layouts (row-major, N pixels of the level, NESTED order):
//   map      [B][N]                    real
//   z        [B][kL][N]                complex (z of gauge l at pixel p)
//   modulus  [(d * B + b)][N][kL]      real, d = 0..n (row d = n holds |z_n|)
//   branch   [(d * B + b)][kL(l3)][N][kL(l1)]  complex: second wavelet of gauge
//            l3 of the modulus of orientation l1
// A convolution maps real rows x [R][N][NI] to complex y [R][kL][N][NI]; a
// decimation maps [R][N_f][NI] to [R][N_c][NI].
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace stl {

template <typename T>
struct Complex2;
template <>
struct Complex2<double> {
  using type = double2;
};
template <>
struct Complex2<float> {
  using type = float2;
};
template <typename T>
using cplx_t = typename Complex2<T>::type;

// Kernel variants of the convolutions: 16 x 16 LDS tiles with one face per
// workgroup (the global gather at levels too small to tile), several rotated
// faces per workgroup (one coefficient load for all of them), 8 x 8 tiles,
// 32 x 32 tiles (1024 threads, smaller halo share; nside >= 32), the
// global gather with each output's slots spread over 8 lanes (small levels),
// or several rotated faces per workgroup on 32 x 32 tiles.
constexpr int kConvTiled = 0, kConvFaces = 1, kConvTiled8 = 2, kConvTiled32 = 3, kConvSplit = 4,
              kConvFaces32 = 5;
template <typename T>
void launch_conv_fwd(const ConvOp<T>& op, int NI, const T* x, int rows, cplx_t<T>* y, T* modulus_out,
                     hipStream_t s, int variant = kConvTiled);
template <typename T>
void launch_conv_adj(const ConvOp<T>& op, int NI, const cplx_t<T>* gy, int rows, T* gx, hipStream_t s,
                     int variant = kConvTiled);
// Whether the adjoint has this variant for the precision and inner width.
template <typename T>
bool conv_adj_has_variant(int NI, int variant);
template <typename T>
bool conv_fwd_has_variant(int NI, int variant);
template <typename T>
void launch_dec_fwd(const DecOp<T>& op, int NI, const T* x, int rows, T* y, hipStream_t s);
template <typename T>
void launch_dec_adj(const DecOp<T>& op, int NI, const T* gy, const T* add, int rows, T* gx, hipStream_t s);

// Masks of one level (the reference's Masks, crates/stl-core/src/mask.rs):
// pixel p belongs to mask k when bit k of pat[cls[p]] is set, cls holding
// the class (distinct membership pattern) of every pixel.  cls == nullptr:
// no masks, every pixel in the one average (K = C = 1).  The reductions
// below run over (b, k) rows; the reverse reads per-class operands.
struct MaskView {
  const uint16_t* cls = nullptr;
  const uint64_t* pat = nullptr;
  int K = 1;  // masks
  int C = 1;  // classes
};

// Pixels per LDS tile of the Gram kernel at level n.
int gram_tile(int n);
// Workgroups per (row, orientation) of launch_gram at level n for `rows`
// (map, mask) rows: as many as the device's CUs hold at once over all rows,
// at most one per pixel pass.  Any count >= 1 is valid for every Gram kernel.
template <typename T>
int gram_blocks(int n, int64_t N, int64_t rows);

// Per-block partial sums of the Gram products V_i conj(V_j), i <= j, of
// V = [branch(d)[l1][l3] for d <= n, l1 < kL; z[l3]] per (b, mask k, l3)
// over the pixels of the mask: part [nblk][B * K * kL][E][2] (+ [E] second
// moments |.|^2 when moments), and in the same pass the S1/S2 sums of z
// (as launch_zstats) to zpart [nblk][B * K * kL][4].
template <typename T>
void launch_gram(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, bool moments, int nblk,
                 double* part, double* part2, double* zpart, const MaskView& mask, hipStream_t s);
// S1/S2 sums of z: part [nblk][B * K * kL][4] = (sum |z|, sum |z|^2, sum re^2+im^2, sum |z|^4).
template <typename T>
void launch_zstats(int64_t N, int B, const cplx_t<T>* z, int nblk, double* part, const MaskView& mask,
                   hipStream_t s);
// mean/var sums of the level-0 map: part [nblk][B * K][4] = (sum x, sum x^2, sum x^2, sum x^4).
template <typename T>
void launch_meanvar(int64_t N, int B, const T* x, int nblk, double* part, const MaskView& mask, hipStream_t s);
// Sums the partials over blocks in block order: part [nblk][m] -> out [m].
void launch_reduce_parts(const double* part, int nblk, int64_t m, double* out, hipStream_t s);

// Statistics from the sky-wide sums, on the device: value i of row r =
// b K + k (map b, mask k) is (src[arr][re + r bs] + i src[arr][im + r bs]) / Nd
// (imaginary part 0 when im < 0, negated when conj), Nd = nd[k levels + level]
// the pixels averaged, and with moments the variance of its terms
// max(src[sq_arr][sq + r sbs] / Nd - |value|^2, 0).
struct StatEntry {
  int32_t arr, sq_arr, conj, level;
  int64_t re, im, bs, sq, sbs;
};
void launch_assemble(int np, int B, int K, int levels, const StatEntry* entries, const double* const* src,
                     const double* nd, bool moments, double2* values, double* variance, hipStream_t s);

// Coefficient cotangents -> the per-level reverse operands (H, s1, s2 and the
// mean / var factors), on the device, in a fixed order: target t of row
// r = b K + k is the sum over its terms q of (mult_q (conj_q ? conj(g) : g)) /
// Nd_q with g = cot[r np + coef_q], Nd_q = nd[k levels + level_q], in the
// terms' order; written to dst[arr][off + r bs] (and off + 1, the imaginary
// part, when cplx).
struct RevTarget {
  int32_t arr, cplx, first, count;
  int64_t off, bs;
};
struct RevTerm {
  int32_t coef, conj, level, pad;
  double mult;
};
void launch_scatter_cot(int ntargets, int B, int K, int levels, int np, const RevTarget* targets,
                        const RevTerm* terms, const double* nd, const double2* cot, double* const* dst,
                        hipStream_t s);
// Per-mask reverse operands -> per-class ones: out [B][R][C][E] = sum over the
// masks k of class c (bit k of pat[c], k ascending) of in [B][K][R][E].
void launch_mask_combine(int B, int K, int R, int C, int64_t E, const uint64_t* pat, const double* in, double* out,
                         hipStream_t s);

// Reverse of the Gram products: gV = H V per pixel and (b, l3), H [B * kL][C][K][K]
// complex (double2), the matrix of the pixel's mask class; writes the branch
// cotangents [(d B + b)][l3][N][l1] and the z part [B][kL][N].
template <typename T>
void launch_gv(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, const double2* H,
               cplx_t<T>* gbranch, cplx_t<T>* gz_part, const MaskView& mask, hipStream_t s);
// gz = gz_part + s1[b,l,c] z/|z| + s2[b,l,c] z + (h z/|z|) with c the pixel's
// mask class and h = modulus cotangent row d = n of the chain stack [(d B + b)][N][kL].
template <typename T>
void launch_gz(int n, int64_t N, int B, const cplx_t<T>* z, const cplx_t<T>* gz_part, const T* chain,
               const double* s1, const double* s2, cplx_t<T>* gz, const MaskView& mask, hipStream_t s);
// Level 0's chain decimation adjoint (tiled operator, one chain row per map:
// chain = add + decimation transpose of gy) followed by launch_gz's gz, the
// chain kept in registers (not written).
template <typename T>
void launch_dec_adj_gz(const DecOp<T>& op, const T* gy, const T* add, int B, const cplx_t<T>* z,
                       const cplx_t<T>* gz_part, const double* s1, const double* s2, cplx_t<T>* gz,
                       const MaskView& mask, hipStream_t s);
// Level 0's map-cotangent decimation adjoint (g0 = add + transpose of gy)
// followed by launch_meanvar_grad into out, g0 not written.
template <typename T>
void launch_dec_adj_grad(const DecOp<T>& op, const T* gy, const T* add, int B, const T* x, const double* gmean,
                         const double* gvar, double* out, const MaskView& mask, hipStream_t s);
// out [B][N] (FP64) = T(g0 + gmean[b,c] + gvar[b,c] x), c the pixel's mask class.
template <typename T>
void launch_meanvar_grad(int64_t N, int B, const T* x, const double* gmean, const double* gvar, const T* g0,
                         double* out, const MaskView& mask, hipStream_t s);
// Cross statistics of map pairs (the reference's cross plan, two channels,
// c1 != c2): B pairs, field rows 2b (channel a) and 2b + 1 (channel b) of
// the field layouts above.
// Full K x K Gram block G[i][j] = sum_p V_a,i conj(V_b,j) per (b, mask k, l3):
// part [nblk][B * K * kL][K * K][2] (+ [K * K] second moments).
template <typename T>
void launch_gram_x(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, bool moments, int nblk,
                   double* part, double* part2, const MaskView& mask, hipStream_t s);
// S1 cross sums (re, im, |.|^2, 0) of rn(z_a) conj(rn(z_b)), rn(v) = v / sqrt(|v| + 1e-8):
// part [nblk][B * K * kL][4].
template <typename T>
void launch_zcross(int64_t N, int B, const cplx_t<T>* z, int nblk, double* part, const MaskView& mask,
                   hipStream_t s);
// gV_a = H V_b, gV_b = H^H V_a with H [B * kL][C][K][K] the cotangent of G (per class).
template <typename T>
void launch_gv_x(int n, int64_t N, int B, const cplx_t<T>* branch, const cplx_t<T>* z, const double2* H,
                 cplx_t<T>* gbranch, cplx_t<T>* gz_part, const MaskView& mask, hipStream_t s);
// gz of both rows of every pair: gz_part + chain modulus term + S1 cross term with the
// complex cotangents s1x [B][kL][C].
template <typename T>
void launch_gz_x(int n, int64_t N, int B, const cplx_t<T>* z, const cplx_t<T>* gz_part, const T* chain,
                 const double2* s1x, cplx_t<T>* gz, const MaskView& mask, hipStream_t s);

// ---------------------------------------------------------------------------
// Batched launches of the coarse levels.  One persistent kernel launch runs a
// list of independent tasks (the same operation at several levels, e.g. the
// branch convolutions of levels 5..8), each split into the work items its own
// kernel would have had as workgroups: a 256-pixel block of one row, a Gram
// block, 64 reduction outputs.  Workgroups walk the items of the whole list in
// a fixed stride (no waits, no atomics); launches follow each other in stream
// order, which orders the dependent steps.  Every item computes its outputs
// completely, in the order of the equivalent kernel, so results do not depend
// on which workgroup runs it.
enum LevelOp : int32_t {
  kQDecFwd1 = 0, kQDecFwd4, kQConvFwd1, kQConvFwd4, kQGram, kQReduce, kQGv, kQConvAdj1, kQConvAdj4, kQFix1,
  kQFix4, kQDecAdj1, kQDecAdj4, kQGz
};
template <typename T>
struct LevelTask {
  int32_t op = 0;
  int32_t ny = 1;           // rows (or Gram row blocks) of the task
  int64_t nx = 0;           // items per row
  int64_t first = 0;        // the task's first item in its list
  ConvOp<T> conv;           // convolution of the level (conv ops)
  DecOp<T> dec;             // decimation from the level (dec ops)
  const void* in0 = nullptr;
  const void* in1 = nullptr;
  const void* in2 = nullptr;
  void* out0 = nullptr;
  void* out1 = nullptr;
  void* out2 = nullptr;
  const double* s1 = nullptr;
  const double* s2 = nullptr;
  int32_t n = 0, B = 0, tpx = 0, bs = 4, moments = 0, nblk = 1;
  int64_t N = 0, m = 0;
};
// Items of a task list (sets every task's `first`), and the dynamic LDS the
// launch needs.
template <typename T>
int64_t batch_layout(LevelTask<T>* tasks, int ntasks);
template <typename T>
size_t batch_lds_bytes(const LevelTask<T>* tasks, int ntasks);
// One launch over the items of independent tasks (on the device), at most
// `workgroups` persistent 256-thread workgroups.
template <typename T>
void launch_batch(const LevelTask<T>* tasks, int ntasks, int64_t items, size_t lds, int workgroups, hipStream_t s);

// Element-wise precision casts (maps in, gradients out).
template <typename TO, typename TI>
void launch_cast(int64_t n, const TI* in, TO* out, hipStream_t s);

}  // namespace stl
