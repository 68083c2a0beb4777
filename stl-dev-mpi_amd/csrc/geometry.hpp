// Device-built pyramid operators: the wavelet convolution of every level
// (folded "kernel x bilinear weight" coefficients per distinct cell) and the
// Gaussian-cone decimation between levels, each with its transpose for the
// gather-form adjoint.  Built once per device and order, stored for the rep
// faces only.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "device_buffer.hpp"

namespace stl {

template <typename T>
struct ConvStore {
  int order = 0;
  int S = 0, Tn = 0;
  bool tiled = false;
  DevArray<uint8_t> cnt, tcnt;
  DevArray<int32_t> cell, trow;
  DevArray<int16_t> off, toff;
  DevArray<T> coef, tcoef;
  DevArray<int32_t> fbcnt, fbo32, fbo16, fbo8;  // ConvOp::fbcnt / fbo* (tiled levels)
  DevArray<int32_t> tbcnt, tbo32, tbo16, tbo8;  // ConvOp::tbcnt / tbo* (tiled levels)
  // Entries stored (forward and transposed), without the union slots' zeros.
  double slots = 0.0, tslots = 0.0;
  // Correction of the mirrored columns (ConvOp::fix_*).
  int nfix = 0;
  DevArray<int32_t> fix_cell, fix_ptr, fix_row;
  DevArray<T> fix_coef;
  ConvOp<T> view() const {
    ConvOp<T> v;
    v.order = order;
    v.npf = int64_t(1) << (2 * order);
    v.R = 2 * v.npf;
    v.S = S;
    v.Tn = Tn;
    v.tiled = tiled;
    v.cnt = cnt.data(); v.cell = cell.data(); v.off = off.data(); v.coef = coef.data();
    v.tcnt = tcnt.data(); v.trow = trow.data(); v.toff = toff.data(); v.tcoef = tcoef.data();
    v.fbcnt = fbcnt.data(); v.fbo32 = fbo32.data(); v.fbo16 = fbo16.data(); v.fbo8 = fbo8.data();
    v.tbcnt = tbcnt.data(); v.tbo32 = tbo32.data(); v.tbo16 = tbo16.data(); v.tbo8 = tbo8.data();
    v.nfix = nfix;
    v.fix_cell = fix_cell.data(); v.fix_ptr = fix_ptr.data(); v.fix_row = fix_row.data();
    v.fix_coef = fix_coef.data();
    return v;
  }
  size_t bytes() const {
    return cnt.bytes() + tcnt.bytes() + cell.bytes() + trow.bytes() + off.bytes() + toff.bytes() + coef.bytes() +
           tcoef.bytes() + fbcnt.bytes() + fbo32.bytes() + fbo16.bytes() + fbo8.bytes() + tbcnt.bytes() + tbo32.bytes() +
           tbo16.bytes() + tbo8.bytes() + fix_cell.bytes() + fix_ptr.bytes() + fix_row.bytes() +
           fix_coef.bytes();
  }
};

template <typename T>
struct DecStore {
  int order_f = 0;
  int S = 0, Tn = 0;
  bool tiled = false, ttiled = false;
  DevArray<uint8_t> cnt, tcnt;
  DevArray<int32_t> col, trow;
  DevArray<int16_t> off, toff;
  DevArray<T> val, tval;
  DecOp<T> view() const {
    DecOp<T> v;
    v.order_f = order_f;
    v.order_c = order_f - 1;
    v.npf_f = int64_t(1) << (2 * order_f);
    v.npf_c = int64_t(1) << (2 * (order_f - 1));
    v.Rf = 3 * v.npf_f;
    v.Rc = 3 * v.npf_c;
    v.S = S;
    v.Tn = Tn;
    v.cnt = cnt.data(); v.col = col.data(); v.val = val.data();
    v.tcnt = tcnt.data(); v.trow = trow.data(); v.tval = tval.data();
    v.tiled = tiled; v.ttiled = ttiled;
    v.off = off.data(); v.toff = toff.data();
    return v;
  }
  size_t bytes() const {
    return cnt.bytes() + tcnt.bytes() + col.bytes() + trow.bytes() + val.bytes() + tval.bytes() + off.bytes() +
           toff.bytes();
  }
};

// Zero-mean l1-normalised Gabor kernel of the wavelet (5 x 5, sigma 1),
// rounded through float like the reference (wavelet.rs:11-35,
// conv.rs:166-172): [re taps..., im taps...], 2 * ksz^2 values.
std::vector<double> wavelet_kernel_f32(int ksz);

// Builds the convolution of the level of `order` (kernel 5, kL gauges,
// gauge cosmo) into the stores whose pointer is non-null.
void build_conv(int order, const std::vector<double>& ker, ConvStore<double>* c64, ConvStore<float>* c32,
                hipStream_t stream);

// Builds the decimation from the level of `order_f` to order_f - 1.
void build_dec(int order_f, DecStore<double>* d64, DecStore<float>* d32, hipStream_t stream);

}  // namespace stl

namespace stl {

// Operators shared by every engine of the process on one device: the
// operators of a level depend only on its order (and the precision), so
// engines of different nside reuse each other's levels.  Built on the
// caller's stream at the first request (and completed before returning);
// freed when the last engine holding them goes.
template <typename T>
struct SharedConv {
  ConvStore<T> store;
  double slots = 0.0, tslots = 0.0;  // stored (transposed) slots summed over rep rows
};
template <typename T>
struct SharedDec {
  DecStore<T> store;
  double slots = 0.0, tslots = 0.0;
};

template <typename T>
std::shared_ptr<const SharedConv<T>> shared_conv(int device, int order, hipStream_t stream);
template <typename T>
std::shared_ptr<const SharedDec<T>> shared_dec(int device, int order_f, hipStream_t stream);

}  // namespace stl
