// CPU FP64 oracle of the ScatCov forward sweep and its VJP on the HEALPix
// NESTED pyramid.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and
// the CPU-baseline leg of the benchmark load this library as the checker;
// the product (libstl_bridge.so) never links, loads or calls it.
//
// It restates, in plain C++ with IEEE double arithmetic (no FMA contraction,
// glibc libm), the reference crate stl-core:
//   healpix.rs    pixel <-> (x, y, face), centres, RING -> NESTED, ring bilinear
//                 interpolation                                   (healpix.rs:53-347)
//   stencil.rs    tangent stencil and per-pixel gauge rotations   (stencil.rs:33-110)
//   wavelet.rs    zero-mean l1-normalised Gabor kernel            (wavelet.rs:11-35)
//   conv.rs       bound stencils, f32-cast kernel, forward/adjoint (conv.rs:53-294)
//   dia.rs        extended face frame (ext_pix, rel_offset)       (dia.rs:27-101)
//   decimation.rs Gaussian-cone decimation rows, CSR, transpose   (decimation.rs:62-219)
//   plan.rs       coefficient enumeration                         (plan.rs:80-158)
//   engine.rs     forward / reverse sweeps, term sums, cotangents (engine.rs:485-806, 1191-1330)
//   mask.rs       masked averages: the mask pyramid of parents and the
//                 masked branches of forward / reverse            (mask.rs:14-70, engine.rs:500-724)
// The sweeps are evaluated on the whole sky, level by level (the reference's
// blocks and halos only bound its memory: its block rows are exact full-sky
// rows).  Partial sums follow the reference order: pixels of a block of
// `block_pixels` in id order, blocks in id order.
//
// Parity status of each part is pinned by the reference fixtures (see the
// tests under tests/test_oracle_*.py); the neighbourhood order of cdshealpix
// is not reproduced (only its sets are used by the reference, sorted).  The
// masks, which no reference fixture covers, are pinned by identities: the
// whole-sky mask, partitions of the sky by base faces, closed forms of mean
// and var over a mask, and central differences (tests/test_oracle_masks.py).

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cplx;

namespace {

const double PI = 3.14159265358979323846;
// torch.rad2deg / deg2rad constants (healpix.rs:12-14).
const double RAD2DEG = 57.295779513082320876798154814105170332405472466564;
const double DEG2RAD = 0.017453292519943295769236907684886127134428718885417;

const int64_t JRLL[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4};
const int64_t JPLL[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7};

inline double rad2deg(double x) { return x * RAD2DEG; }
inline double deg2rad(double x) { return x * DEG2RAD; }

// Python / torch remainder (sign of b) (healpix.rs:33-39).
inline double remainder_py(double a, double b) {
  double m = std::fmod(a, b);
  if (m != 0.0 && ((b < 0.0) != (m < 0.0))) m += b;
  return m;
}

inline int64_t div_euclid(int64_t a, int64_t b) {
  int64_t q = a / b;
  if (a % b < 0) q = b > 0 ? q - 1 : q + 1;
  return q;
}

int order_of(int64_t nside) {
  if (nside <= 0 || (nside & (nside - 1))) throw std::runtime_error("nside must be a power of two");
  int o = 0;
  while ((int64_t(1) << o) < nside) ++o;
  return o;
}

inline int64_t npix(int64_t nside) { return 12 * nside * nside; }

inline int64_t compress_bits(int64_t v) {
  v &= 0x5555555555555555LL;
  v = (v | (v >> 1)) & 0x3333333333333333LL;
  v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0FLL;
  v = (v | (v >> 4)) & 0x00FF00FF00FF00FFLL;
  v = (v | (v >> 8)) & 0x0000FFFF0000FFFFLL;
  return (v | (v >> 16)) & 0x00000000FFFFFFFFLL;
}

inline int64_t spread_bits(int64_t v) {
  v &= 0x00000000FFFFFFFFLL;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFLL;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFLL;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FLL;
  v = (v | (v << 2)) & 0x3333333333333333LL;
  return (v | (v << 1)) & 0x5555555555555555LL;
}

// healpix.rs:75-79
inline void nest2xyf(int64_t pix, int order, int64_t& ix, int64_t& iy, int64_t& face) {
  face = pix >> (2 * order);
  int64_t local = pix & ((int64_t(1) << (2 * order)) - 1);
  ix = compress_bits(local);
  iy = compress_bits(local >> 1);
}

// healpix.rs:83-85
inline int64_t xyf2nest(int64_t ix, int64_t iy, int64_t face, int order) {
  return (face << (2 * order)) + spread_bits(ix) + (spread_bits(iy) << 1);
}

// healpix.rs:88-127 (healpix_base pix2loc)
void xyf2ang(int64_t ix, int64_t iy, int64_t face, int64_t nside, double& theta, double& phi) {
  int64_t jr = JRLL[face] * nside - ix - iy - 1;
  double ns = double(nside);
  double fact2 = 4.0 / (12.0 * ns * ns);
  double fact1 = 2.0 / (3.0 * ns);
  bool north = jr < nside;
  bool south = jr > 3 * nside;
  int64_t nr = north ? jr : (south ? 4 * nside - jr : nside);
  int64_t kshift = (north || south) ? 0 : ((jr - nside) & 1);
  double nrf = double(nr);
  if (north || south) {
    double tmp = nrf * nrf * fact2;
    double z_cap = 1.0 - tmp;
    double sth_cap = std::sqrt(tmp * (2.0 - tmp));
    double theta_cap = std::atan2(sth_cap, z_cap);
    theta = north ? theta_cap : PI - theta_cap;
  } else {
    theta = std::acos((2.0 * ns - double(jr)) * fact1);
  }
  int64_t jp = div_euclid(JPLL[face] * nr + ix - iy + 1 + kshift, 2);
  if (jp > 4 * nr) jp -= 4 * nr;
  if (jp < 1) jp += 4 * nr;
  phi = (double(jp) - (double(kshift) + 1.0) * 0.5) * (0.5 * PI / nrf);
}

void pix2ang_nest(int64_t pix, int64_t nside, double& theta, double& phi) {
  int64_t ix, iy, f;
  nest2xyf(pix, order_of(nside), ix, iy, f);
  xyf2ang(ix, iy, f, nside, theta, phi);
}

// healpix.rs:140-143: centre rounded through degrees.
void centre_lonlat(int64_t pix, int64_t nside, double& lon, double& lat) {
  double theta, phi;
  pix2ang_nest(pix, nside, theta, phi);
  lon = deg2rad(rad2deg(phi));
  lat = deg2rad(90.0 - rad2deg(theta));
}

int64_t isqrt(int64_t v) {
  int64_t r = int64_t(std::sqrt(double(v)));
  if (r * r > v) r -= 1;
  if ((r + 1) * (r + 1) <= v) r += 1;
  return r;
}

// healpix.rs:157-210
int64_t ring2nest(int64_t pix, int order) {
  int64_t nside = int64_t(1) << order;
  int64_t np = npix(nside);
  int64_t ncap = 2 * nside * (nside - 1);
  int64_t nl2 = 2 * nside;
  int64_t iring, iphi, kshift, nr, face;
  if (pix < ncap) {
    int64_t ir = (1 + isqrt(1 + 2 * pix)) >> 1;
    int64_t ip = pix + 1 - 2 * ir * (ir - 1);
    iring = ir; iphi = ip; kshift = 0; nr = ir;
    face = div_euclid(ip - 1, std::max<int64_t>(ir, 1));
  } else if (pix < np - ncap) {
    int64_t ip = pix - ncap;
    int64_t tmp = ip >> (order + 2);
    int64_t ir = tmp + nside;
    int64_t ip2 = ip - tmp * 4 * nside + 1;
    int64_t ire = ir - nside + 1;
    int64_t irm = nl2 + 2 - ire;
    int64_t ifm = (ip2 - div_euclid(ire, 2) + nside - 1) >> order;
    int64_t ifp = (ip2 - div_euclid(irm, 2) + nside - 1) >> order;
    iring = ir; iphi = ip2; kshift = (ir + nside) & 1; nr = nside;
    face = (ifp == ifm) ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8);
  } else {
    int64_t ip = np - pix;
    int64_t ir = (1 + isqrt(2 * ip - 1)) >> 1;
    int64_t ip2 = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
    iring = 2 * nl2 - ir; iphi = ip2; kshift = 0; nr = ir;
    face = 8 + div_euclid(ip2 - 1, std::max<int64_t>(ir, 1));
  }
  int64_t irt = iring - JRLL[face] * nside + 1;
  int64_t ipt = 2 * iphi - JPLL[face] * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  int64_t ix = (ipt - irt) >> 1;
  int64_t iy = (-ipt - irt) >> 1;
  return xyf2nest(ix, iy, face, order);
}

// healpix.rs:216-229
int64_t ring_above(double z, int64_t nside) {
  double az = std::fabs(z);
  double ns = double(nside);
  if (az <= 2.0 / 3.0) return int64_t(std::trunc(ns * (2.0 - 1.5 * z)));
  int64_t iring = int64_t(std::trunc(ns * std::sqrt(3.0 * (1.0 - az))));
  return z > 0.0 ? iring : 4 * nside - iring - 1;
}

// healpix.rs:232-259
void ring_info(int64_t ring, int64_t nside, int64_t& startpix, int64_t& ringpix, double& theta, bool& shifted) {
  int64_t ncap = 2 * nside * (nside - 1);
  int64_t np = npix(nside);
  double ns = double(nside);
  double fact1 = 2.0 / (3.0 * ns);
  double fact2 = 1.0 / (3.0 * ns * ns);
  int64_t northring = ring > 2 * nside ? 4 * nside - ring : ring;
  if (northring < nside) {
    double nrf = double(northring);
    double tmp = nrf * nrf * fact2;
    theta = std::atan2(std::sqrt(tmp * (2.0 - tmp)), 1.0 - tmp);
    ringpix = 4 * northring;
    shifted = true;
    startpix = 2 * northring * (northring - 1);
  } else {
    theta = std::acos(double(2 * nside - northring) * fact1);
    ringpix = 4 * nside;
    shifted = ((northring - nside) & 1) == 0;
    startpix = ncap + (northring - nside) * 4 * nside;
  }
  if (northring != ring) {
    theta = PI - theta;
    startpix = np - startpix - ringpix;
  }
}

// healpix.rs:262-280
void ring_pair(double phi, int64_t ring, int64_t nside, int64_t& p0, int64_t& p1, double& w0, double& w1,
               double& theta_ring) {
  int64_t sp, nr;
  bool shifted;
  ring_info(ring, nside, sp, nr, theta_ring, shifted);
  double dphi = (2.0 * PI) / double(nr);
  double half = shifted ? 0.5 : 0.0;
  double tmp = phi / dphi - half;
  int64_t i1 = int64_t(std::trunc(tmp));
  if (tmp < 0.0) i1 -= 1;
  double ww = (phi - (double(i1) + half) * dphi) / dphi;
  int64_t i2 = i1 + 1;
  if (i1 < 0) i1 += nr;
  if (i2 >= nr) i2 -= nr;
  p0 = sp + i1;
  p1 = sp + i2;
  w0 = 1.0 - ww;
  w1 = ww;
}

// healpix.rs:283-324 (T_Healpix_Base::get_interpol)
void get_interpol_ring(double theta, double phi, int64_t nside, int64_t pix[4], double wgt[4]) {
  int64_t np = npix(nside);
  double z = std::cos(theta);
  int64_t ir1 = ring_above(z, nside);
  int64_t ir2 = ir1 + 1;
  int64_t p[4] = {0, 0, 0, 0};
  double w[4] = {0, 0, 0, 0};
  int64_t a0, a1;
  double b0, b1, theta1, theta2;
  ring_pair(phi, std::max<int64_t>(ir1, 1), nside, a0, a1, b0, b1, theta1);
  if (ir1 > 0) { p[0] = a0; p[1] = a1; w[0] = b0; w[1] = b1; }
  ring_pair(phi, std::min<int64_t>(ir2, 4 * nside - 1), nside, a0, a1, b0, b1, theta2);
  if (ir2 < 4 * nside) { p[2] = a0; p[3] = a1; w[2] = b0; w[3] = b1; }
  bool north = ir1 == 0;
  bool south = ir2 == 4 * nside && !north;
  if (north) {
    double wn = theta / theta2;
    double facn = (1.0 - wn) * 0.25;
    pix[0] = (p[2] + 2) & 3; pix[1] = (p[3] + 2) & 3; pix[2] = p[2]; pix[3] = p[3];
    wgt[0] = facn; wgt[1] = facn; wgt[2] = w[2] * wn + facn; wgt[3] = w[3] * wn + facn;
  } else if (south) {
    double ws = (theta - theta1) / (PI - theta1);
    double facs = ws * 0.25;
    pix[0] = p[0]; pix[1] = p[1];
    pix[2] = ((p[0] + 2) & 3) + np - 4; pix[3] = ((p[1] + 2) & 3) + np - 4;
    wgt[0] = w[0] * (1.0 - ws) + facs; wgt[1] = w[1] * (1.0 - ws) + facs; wgt[2] = facs; wgt[3] = facs;
  } else {
    double wm = (theta - theta1) / (theta2 - theta1);
    for (int i = 0; i < 4; ++i) pix[i] = p[i];
    wgt[0] = w[0] * (1.0 - wm); wgt[1] = w[1] * (1.0 - wm); wgt[2] = w[2] * wm; wgt[3] = w[3] * wm;
  }
}

inline double norm3(const double v[3]) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// healpix.rs:336-347: direction -> (theta, phi) -> degrees -> radians -> ring weights -> NESTED.
void interp_weights_nested(const double v[3], int order, int64_t pix[4], double w[4]) {
  double n = std::max(norm3(v), 1e-12);
  double vn[3] = {v[0] / n, v[1] / n, v[2] / n};
  double theta = std::acos(std::min(std::max(vn[2], -1.0), 1.0));
  double phi = remainder_py(std::atan2(vn[1], vn[0]), 2.0 * PI);
  double lon_deg = rad2deg(phi);
  double lat_deg = 90.0 - rad2deg(theta);
  double theta2 = deg2rad(90.0 - lat_deg);
  double phi2 = remainder_py(deg2rad(lon_deg), 2.0 * PI);
  int64_t rp[4];
  get_interpol_ring(theta2, phi2, int64_t(1) << order, rp, w);
  for (int i = 0; i < 4; ++i) pix[i] = ring2nest(rp[i], order);
}

// ---------------------------------------------------------------------------
// Stencil and rotations (stencil.rs)
// ---------------------------------------------------------------------------

typedef double Mat3[3][3];

// stencil.rs:33-47
std::vector<std::array<double, 3>> local_kernel_grid(int ksz, int64_t nside) {
  int64_t half = ksz / 2;
  double ns = double(nside);
  double alpha_pix = std::sqrt(4.0 * PI / (12.0 * ns * ns));
  std::vector<std::array<double, 3>> out;
  for (int64_t i = 0; i < ksz; ++i)
    for (int64_t j = 0; j < ksz; ++j) {
      int64_t x = j - half, y = i - half;
      double dtheta = std::sqrt(double(x * x + y * y)) * alpha_pix;
      double dphi = std::atan2(double(y), double(x));
      out.push_back({std::sin(dtheta) * std::cos(dphi), std::sin(dtheta) * std::sin(dphi), std::cos(dtheta)});
    }
  return out;
}

inline void matmul3(const Mat3& a, const Mat3& b, Mat3& c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
}

inline void apply3(const Mat3& r, const double v[3], double out[3]) {
  out[0] = r[0][0] * v[0] + r[0][1] * v[1] + r[0][2] * v[2];
  out[1] = r[1][0] * v[0] + r[1][1] * v[1] + r[1][2] * v[2];
  out[2] = r[2][0] * v[0] + r[2][1] * v[1] + r[2][2] * v[2];
}

// stencil.rs:51-88; gauge 0 = cosmo, 1 = phi.
void rotation_matrices(double theta, double phi, int G, int gauge, std::vector<std::array<double, 9>>& out) {
  double ct = std::cos(theta), st = std::sin(theta);
  double cp = std::cos(phi), sp = std::sin(phi);
  Mat3 base = {{cp * ct, -sp, cp * st}, {sp * ct, cp, sp * st}, {-st, 0.0, ct}};
  double n0[3] = {base[0][2], base[1][2], base[2][2]};
  double nn = std::max(norm3(n0), 1e-12);
  double n[3] = {n0[0] / nn, n0[1] / nn, n0[2] / nn};
  double alpha_base, sign_g;
  if (gauge == 0) {
    if (theta > PI / 2.0) { alpha_base = phi; sign_g = -1.0; }
    else { alpha_base = -phi; sign_g = 1.0; }
  } else {
    alpha_base = 0.0; sign_g = 1.0;
  }
  Mat3 skew = {{0.0, -n[2], n[1]}, {n[2], 0.0, -n[0]}, {-n[1], n[0], 0.0}};
  double step = PI / double(G);
  out.resize(G);
  for (int g = 0; g < G; ++g) {
    double alpha = alpha_base + sign_g * (double(g) * step);
    double ca = std::cos(alpha), sa = std::sin(alpha);
    Mat3 rg;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double eye = i == j ? 1.0 : 0.0;
        rg[i][j] = eye * ca + skew[i][j] * sa + n[i] * n[j] * (1.0 - ca);
      }
    Mat3 c;
    matmul3(rg, base, c);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) out[g][i * 3 + j] = c[i][j];
  }
}

// ---------------------------------------------------------------------------
// Wavelet kernel (wavelet.rs:11-35) and its f32 cast (conv.rs:166-172)
// ---------------------------------------------------------------------------

std::vector<double> wavelet_weights(int ksz) {
  const double sigma = 1.0;
  std::vector<cplx> k;
  for (int i = 0; i < ksz; ++i)
    for (int j = 0; j < ksz; ++j) {
      double x = double(j) - (double(ksz) - 1.0) / 2.0;
      double y = double(i) - (double(ksz) - 1.0) / 2.0;
      double mother = std::exp(-(x * x + y * y) / (2.0 * sigma * sigma));
      double angle = 0.5 * PI * x;
      k.emplace_back(std::cos(angle) * mother, std::sin(angle) * mother);
    }
  double n = double(k.size());
  double sr = 0.0, si = 0.0;
  for (auto& v : k) { sr += v.real(); si += v.imag(); }
  double mr = sr / n, mi = si / n;
  for (auto& v : k) v = cplx(v.real() - mr, v.imag() - mi);
  double l1 = 0.0;
  for (auto& v : k) l1 += std::hypot(v.real(), v.imag());
  std::vector<double> out(2 * k.size());
  for (size_t p = 0; p < k.size(); ++p) {
    out[p] = k[p].real() / l1;
    out[k.size() + p] = k[p].imag() / l1;
  }
  return out;
}

// ---------------------------------------------------------------------------
// Extended face frame (dia.rs:27-101)
// ---------------------------------------------------------------------------

const int8_t NB_FACE[9][12] = {
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
const uint8_t NB_SWAP[9][3] = {{0, 0, 3}, {0, 0, 6}, {0, 0, 0}, {0, 0, 5}, {0, 0, 0},
                               {5, 0, 0}, {0, 0, 0}, {6, 0, 0}, {3, 0, 0}};

// Pixel at extended coordinates (ex, ey) of `face`, -1 in the missing quadrant.
int64_t ext_pix(int64_t face, int64_t ex, int64_t ey, int64_t nside, int order) {
  int64_t x = ex, y = ey;
  int nb = 4;
  if (x < 0) { x += nside; nb -= 1; }
  else if (x >= nside) { x -= nside; nb += 1; }
  if (y < 0) { y += nside; nb -= 3; }
  else if (y >= nside) { y -= nside; nb += 3; }
  int f = NB_FACE[nb][face];
  if (f < 0) return -1;
  uint8_t bits = NB_SWAP[nb][face >> 2];
  if (bits & 1) x = nside - x - 1;
  if (bits & 2) y = nside - y - 1;
  if (bits & 4) std::swap(x, y);
  return xyf2nest(x, y, f, order);
}

bool rel_offset(int64_t row, int64_t cell, int64_t nside, int order, int64_t& dx, int64_t& dy) {
  int64_t x, y, f, xq, yq, fq;
  nest2xyf(row, order, x, y, f);
  nest2xyf(cell, order, xq, yq, fq);
  if (fq == f) { dx = xq - x; dy = yq - y; return true; }
  int nb = -1;
  for (int k = 0; k < 9; ++k)
    if (k != 4 && NB_FACE[k][f] == fq) { nb = k; break; }
  if (nb < 0) return false;
  uint8_t bits = NB_SWAP[nb][f >> 2];
  int64_t a = xq, b = yq;
  if (bits & 4) std::swap(a, b);
  if (bits & 2) b = nside - b - 1;
  if (bits & 1) a = nside - a - 1;
  int64_t xdir = nb % 3 - 1, ydir = nb / 3 - 1;
  dx = a + xdir * nside - x;
  dy = b + ydir * nside - y;
  return true;
}

// Cells within k rings (sorted).  The k-ring is the (2k+1)^2 window of the
// extended face frame, the quadrant missing at a vertex of three faces
// excluded; the whole level when k >= nside (neighbours.rs:28-41).
std::vector<int64_t> kth_neighbourhood(int64_t id, int64_t nside, int k) {
  std::vector<int64_t> out;
  if (int64_t(k) >= nside) {
    for (int64_t p = 0; p < npix(nside); ++p) out.push_back(p);
    return out;
  }
  int order = order_of(nside);
  int64_t x, y, f;
  nest2xyf(id, order, x, y, f);
  for (int dy = -k; dy <= k; ++dy)
    for (int dx = -k; dx <= k; ++dx) {
      int64_t c = ext_pix(f, x + dx, y + dy, nside, order);
      if (c >= 0) out.push_back(c);
    }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

// ---------------------------------------------------------------------------
// Bound stencils (conv.rs:73-135) for one output row on a sorted grid; the
// full sky when grid == nullptr.  Output: for (g, p, n) the grid position
// (or NESTED id on the full sky) and renormalised weight; returns complete.
// ---------------------------------------------------------------------------

struct Stencil {
  int ksz, G, gauge, P;
  int64_t nside;
  int order;
  std::vector<std::array<double, 3>> grid_vec;
};

inline int64_t search_sorted(const int64_t* grid, int64_t n, int64_t v) {
  const int64_t* it = std::lower_bound(grid, grid + n, v);
  if (it != grid + n && *it == v) return it - grid;
  return -1;
}

// pos/w laid out [g][p][n] (4 G P entries).
bool bind_row(const Stencil& st, int64_t row_id, const int64_t* grid, int64_t n_grid, int64_t* pos, double* w,
              std::vector<std::array<double, 9>>& rots) {
  int P = st.P, G = st.G;
  int p_ref = (st.ksz / 2) * (st.ksz + 1);
  double theta, phi;
  pix2ang_nest(row_id, st.nside, theta, phi);
  theta = deg2rad(90.0 - rad2deg(PI / 2.0 - theta));
  phi = deg2rad(rad2deg(phi));
  rotation_matrices(theta, phi, G, st.gauge, rots);
  bool complete = true;
  int64_t idx[64][4];
  double wt[64][4];
  for (int g = 0; g < G; ++g) {
    Mat3 r;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) r[i][j] = rots[g][i * 3 + j];
    for (int p = 0; p < P; ++p) {
      double v[3] = {st.grid_vec[p][0], st.grid_vec[p][1], st.grid_vec[p][2]};
      double d[3];
      apply3(r, v, d);
      interp_weights_nested(d, st.order, idx[p], wt[p]);
    }
    for (int p = 0; p < P; ++p) {
      int64_t* ix = idx[p];
      double* wx = wt[p];
      int64_t found[4];
      bool any = false;
      for (int n = 0; n < 4; ++n) {
        found[n] = grid ? search_sorted(grid, n_grid, ix[n]) : ix[n];
        any |= found[n] >= 0;
      }
      if (!any) {  // empty tap: central tap of the same row
        ix = idx[p_ref];
        wx = wt[p_ref];
        for (int n = 0; n < 4; ++n) found[n] = grid ? search_sorted(grid, n_grid, ix[n]) : ix[n];
      }
      double ww[4];
      for (int n = 0; n < 4; ++n) ww[n] = wx[n] * (found[n] >= 0 ? 1.0 : 0.0);
      double colsum = ww[0] + ww[1] + ww[2] + ww[3];
      if (colsum == 0.0) {
        ww[0] = found[0] >= 0 ? 1.0 : 0.0;
        colsum = ww[0] + ww[1] + ww[2] + ww[3];
      }
      for (int n = 0; n < 4; ++n) {
        size_t a = (size_t(g) * P + p) * 4 + n;
        w[a] = ww[n] / std::max(colsum, 1e-12);
        pos[a] = found[n] >= 0 ? found[n] : 0;
        complete &= found[n] >= 0;
      }
    }
  }
  return complete;
}

// ---------------------------------------------------------------------------
// Decimation (decimation.rs)
// ---------------------------------------------------------------------------

inline double py_degrees(double x) { return x * (180.0 / PI); }
inline double py_radians(double x) { return x * (PI / 180.0); }

inline double haversine(double lat1, double lon1, double lat2, double lon2) {
  double dlat = 0.5 * (lat2 - lat1);
  double dlon = 0.5 * (lon2 - lon1);
  double sa = std::sin(dlat), so = std::sin(dlon);
  double a = std::min(std::max(sa * sa + std::cos(lat1) * std::cos(lat2) * (so * so), 0.0), 1.0);
  return 2.0 * std::asin(std::sqrt(a));
}

// Candidate fine pixels of a coarse pixel (decimation.rs:450-461).
std::vector<int64_t> candidates(int64_t c, int64_t nside) {
  int64_t cn = nside / 2;
  std::vector<int64_t> out;
  if (cn >= 4) {
    for (int64_t v : kth_neighbourhood(c, cn, 2))
      for (int i = 0; i < 4; ++i) out.push_back(4 * v + i);
  } else {
    for (int64_t p = 0; p < npix(nside); ++p) out.push_back(p);
  }
  return out;
}

void decim_params(int64_t nside, double& radius, double& sigma) {
  double pix_deg = py_degrees(std::sqrt(4.0 * PI / double(npix(nside))));
  radius = py_radians(3.0 * pix_deg);
  sigma = radius / 2.0;
}

// decimation.rs:516-543 (l1 rows)
std::vector<double> row_weights(int64_t c, const std::vector<int64_t>& fine, int64_t nside, double radius,
                                double sigma) {
  double lon_c, lat_c;
  centre_lonlat(c, nside / 2, lon_c, lat_c);
  std::vector<double> w(fine.size());
  for (size_t i = 0; i < fine.size(); ++i) {
    double lon_f, lat_f;
    centre_lonlat(fine[i], nside, lon_f, lat_f);
    double gamma = haversine(lat_c, lon_c, lat_f, lon_f);
    double t = gamma / sigma;
    w[i] = gamma > radius ? 0.0 : std::exp(-0.5 * (t * t));
  }
  double row_sum = 0.0;
  for (double v : w) row_sum += v;
  if (row_sum <= 0.0 && !w.empty()) {
    double u = 1.0 / double(w.size());
    for (double& v : w) v = u;
    row_sum = 0.0;
    for (double v : w) row_sum += v;
  }
  double scale = std::max(row_sum, 1e-300);
  for (double& v : w) v = v / scale;
  return w;
}

// Full-sky row of coarse cell c (at nside/2) from nside (decimation.rs:549-561).
void decimation_row(int64_t c, int64_t nside, std::vector<int64_t>& cols, std::vector<double>& vals) {
  double radius, sigma;
  decim_params(nside, radius, sigma);
  std::vector<int64_t> fine = candidates(c, nside);
  std::sort(fine.begin(), fine.end());
  fine.erase(std::unique(fine.begin(), fine.end()), fine.end());
  std::vector<double> w = row_weights(c, fine, nside, radius, sigma);
  cols.clear();
  vals.clear();
  for (size_t i = 0; i < fine.size(); ++i)
    if (w[i] != 0.0) { cols.push_back(fine[i]); vals.push_back(w[i]); }
}

struct Csr {
  int64_t n_rows = 0, n_cols = 0;
  std::vector<int64_t> indptr, indices;
  std::vector<double> values;
  // y[r] = sum over the row in column order.
  void apply(const double* x, double* y) const {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_rows; ++r) {
      double s = 0.0;
      for (int64_t q = indptr[r]; q < indptr[r + 1]; ++q) s += values[q] * x[indices[q]];
      y[r] = s;
    }
  }
  Csr transpose() const {
    Csr t;
    t.n_rows = n_cols;
    t.n_cols = n_rows;
    std::vector<int64_t> count(n_cols + 1, 0);
    for (int64_t c : indices) count[c + 1] += 1;
    for (int64_t i = 0; i < n_cols; ++i) count[i + 1] += count[i];
    t.indptr = count;
    std::vector<int64_t> next(count.begin(), count.end());
    t.indices.assign(indices.size(), 0);
    t.values.assign(values.size(), 0.0);
    for (int64_t r = 0; r < n_rows; ++r)
      for (int64_t q = indptr[r]; q < indptr[r + 1]; ++q) {
        int64_t c = indices[q];
        t.indices[next[c]] = r;
        t.values[next[c]] = values[q];
        next[c] += 1;
      }
    return t;
  }
};

// Full-sky decimation from nside to nside/2.
Csr full_decimation(int64_t nside) {
  int64_t nc = npix(nside / 2);
  std::vector<std::vector<int64_t>> cols(nc);
  std::vector<std::vector<double>> vals(nc);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t c = 0; c < nc; ++c) decimation_row(c, nside, cols[c], vals[c]);
  Csr m;
  m.n_rows = nc;
  m.n_cols = npix(nside);
  m.indptr.push_back(0);
  for (int64_t c = 0; c < nc; ++c) {
    m.indices.insert(m.indices.end(), cols[c].begin(), cols[c].end());
    m.values.insert(m.values.end(), vals[c].begin(), vals[c].end());
    m.indptr.push_back(int64_t(m.indices.size()));
  }
  return m;
}

// Partial-grid decimation (decimation.rs:476-512).
void smooth_decimation(const int64_t* ids, int64_t n, int64_t nside, std::vector<int64_t>& coarse, Csr& m) {
  double radius, sigma;
  decim_params(nside, radius, sigma);
  coarse.clear();
  for (int64_t i = 0; i < n; ++i) coarse.push_back(div_euclid(ids[i], 4));
  std::sort(coarse.begin(), coarse.end());
  coarse.erase(std::unique(coarse.begin(), coarse.end()), coarse.end());
  m = Csr();
  m.n_rows = int64_t(coarse.size());
  m.n_cols = n;
  m.indptr.push_back(0);
  for (int64_t c : coarse) {
    std::vector<int64_t> cols;
    for (int64_t v : candidates(c, nside)) {
      if (v < 0) continue;
      int64_t q = search_sorted(ids, n, v);
      if (q >= 0) cols.push_back(q);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    std::vector<int64_t> fine;
    for (int64_t q : cols) fine.push_back(ids[q]);
    std::vector<double> w = row_weights(c, fine, nside, radius, sigma);
    for (size_t i = 0; i < cols.size(); ++i)
      if (w[i] != 0.0) { m.indices.push_back(cols[i]); m.values.push_back(w[i]); }
    m.indptr.push_back(int64_t(m.indices.size()));
  }
}

// ---------------------------------------------------------------------------
// Plan (plan.rs:80-158)
// ---------------------------------------------------------------------------

enum Family { F_MEAN = 0, F_VAR = 1, F_S1 = 2, F_S2 = 3, F_S3 = 4, F_S4 = 5 };

struct Coef {
  int family, native_dg, c1, c2;
  int j1 = -1, j2 = -1, j3 = -1, l1 = -1, l2 = -1, l3 = -1;
};

std::vector<Coef> build_plan(int J, int L, const std::vector<int>& j_to_dg, int C) {
  std::vector<Coef> plan;
  for (int fam : {F_MEAN, F_VAR})
    for (int c = 0; c < C; ++c) plan.push_back(Coef{fam, 0, c, c});
  for (int j3 = 0; j3 < J; ++j3) {
    int dg = j_to_dg[j3];
    for (int c1 = 0; c1 < C; ++c1)
      for (int c2 = c1; c2 < C; ++c2) {
        for (int fam : {F_S1, F_S2})
          for (int l = 0; l < L; ++l) {
            Coef k{fam, dg, c1, c2};
            k.j1 = j3;
            k.l1 = l;
            plan.push_back(k);
          }
        std::vector<std::pair<int, int>> pairs;
        pairs.push_back({c1, c2});
        if (c1 != c2) pairs.push_back({c2, c1});
        for (auto [ca, cb] : pairs)
          for (int j2 = 0; j2 <= j3; ++j2) {
            for (int l2 = 0; l2 < L; ++l2)
              for (int l3 = 0; l3 < L; ++l3) {
                Coef k{F_S3, dg, ca, cb};
                k.j1 = j2; k.j3 = j3; k.l1 = l2; k.l3 = l3;
                plan.push_back(k);
              }
            for (int j1 = 0; j1 <= j2; ++j1)
              for (int l1 = 0; l1 < L; ++l1)
                for (int l2 = 0; l2 < L; ++l2)
                  for (int l3 = 0; l3 < L; ++l3) {
                    Coef k{F_S4, dg, ca, cb};
                    k.j1 = j1; k.j2 = j2; k.j3 = j3; k.l1 = l1; k.l2 = l2; k.l3 = l3;
                    plan.push_back(k);
                  }
          }
      }
  }
  std::stable_sort(plan.begin(), plan.end(), [](const Coef& a, const Coef& b) { return a.native_dg < b.native_dg; });
  return plan;
}

// ---------------------------------------------------------------------------
// Full-sky convolution of one level: geometry per row (rebuilt, or cached),
// forward in the reference order (conv.rs:253-294), adjoint as a serial
// scatter in row order.
// ---------------------------------------------------------------------------

struct LevelConv {
  Stencil st;
  int64_t n = 0;
  bool cache = false;
  std::vector<int32_t> cpos;  // [row][g][p][n]
  std::vector<double> cw;
  size_t per_row() const { return size_t(st.G) * st.P * 4; }

  void init(int64_t nside, int ksz, int G, int gauge, bool cache_geometry) {
    st.ksz = ksz; st.G = G; st.gauge = gauge; st.P = ksz * ksz; st.nside = nside; st.order = order_of(nside);
    st.grid_vec = local_kernel_grid(ksz, nside);
    n = npix(nside);
    cache = cache_geometry;
    if (cache) {
      cpos.assign(size_t(n) * per_row(), 0);
      cw.assign(size_t(n) * per_row(), 0.0);
#pragma omp parallel
      {
        std::vector<int64_t> pos(per_row());
        std::vector<std::array<double, 9>> rots;
#pragma omp for schedule(dynamic, 64)
        for (int64_t k = 0; k < n; ++k) {
          bind_row(st, k, nullptr, 0, pos.data(), cw.data() + size_t(k) * per_row(), rots);
          for (size_t a = 0; a < per_row(); ++a) cpos[size_t(k) * per_row() + a] = int32_t(pos[a]);
        }
      }
    }
  }

  // Geometry of row k into (pos, w).
  void row(int64_t k, int64_t* pos, double* w, std::vector<std::array<double, 9>>& rots) const {
    if (cache) {
      for (size_t a = 0; a < per_row(); ++a) {
        pos[a] = cpos[size_t(k) * per_row() + a];
        w[a] = cw[size_t(k) * per_row() + a];
      }
    } else {
      bind_row(st, k, nullptr, 0, pos, w, rots);
    }
  }

  // x [R][n] real -> y [R][G][n] complex.
  void forward(const double* x, int R, const std::vector<double>& ker, cplx* y) const {
    int G = st.G, P = st.P;
#pragma omp parallel
    {
      std::vector<int64_t> pos(per_row());
      std::vector<double> w(per_row());
      std::vector<std::array<double, 9>> rots;
      std::vector<double> gathered(P);
#pragma omp for schedule(dynamic, 64)
      for (int64_t k = 0; k < n; ++k) {
        row(k, pos.data(), w.data(), rots);
        for (int r = 0; r < R; ++r) {
          const double* xs = x + size_t(r) * n;
          for (int g = 0; g < G; ++g) {
            for (int p = 0; p < P; ++p) {
              double acc = 0.0;
              for (int nn = 0; nn < 4; ++nn) {
                size_t a = (size_t(g) * P + p) * 4 + nn;
                acc += xs[pos[a]] * w[a];
              }
              gathered[p] = acc;
            }
            double out[2];
            for (int o = 0; o < 2; ++o) {
              double s = 0.0;
              for (int p = 0; p < P; ++p) s += gathered[p] * ker[(size_t(g) * 2 + o) * P + p];
              out[o] = 0.0 + s;
            }
            y[(size_t(r) * G + g) * n + k] = cplx(out[0], out[1]);
          }
        }
      }
    }
  }

  // gy [R][G][n] complex -> gx [R][n] real (added).
  void adjoint_add(const cplx* gy, int R, const std::vector<double>& ker, double* gx) const {
    int G = st.G, P = st.P;
    const int64_t chunk = 8192;
    std::vector<int64_t> pos(size_t(chunk) * per_row());
    std::vector<double> w(size_t(chunk) * per_row());
    for (int64_t k0 = 0; k0 < n; k0 += chunk) {
      int64_t k1 = std::min(n, k0 + chunk);
#pragma omp parallel
      {
        std::vector<std::array<double, 9>> rots;
#pragma omp for schedule(dynamic, 64)
        for (int64_t k = k0; k < k1; ++k)
          row(k, pos.data() + size_t(k - k0) * per_row(), w.data() + size_t(k - k0) * per_row(), rots);
      }
      for (int r = 0; r < R; ++r) {
        double* dst = gx + size_t(r) * n;
        for (int64_t k = k0; k < k1; ++k) {
          const int64_t* pk = pos.data() + size_t(k - k0) * per_row();
          const double* wk = w.data() + size_t(k - k0) * per_row();
          for (int g = 0; g < G; ++g) {
            cplx gyv = gy[(size_t(r) * G + g) * n + k];
            for (int p = 0; p < P; ++p) {
              double tap = gyv.real() * ker[(size_t(g) * 2 + 0) * P + p] + gyv.imag() * ker[(size_t(g) * 2 + 1) * P + p];
              for (int nn = 0; nn < 4; ++nn) {
                size_t a = (size_t(g) * P + p) * 4 + nn;
                dst[pk[a]] += tap * wk[a];
              }
            }
          }
        }
      }
    }
  }
};

inline double hypot_norm(cplx z) { return std::hypot(z.real(), z.imag()); }
inline double norm_sqr_exact(cplx z) { double r = hypot_norm(z); return r * r; }
inline double norm_sqr(cplx z) { return z.real() * z.real() + z.imag() * z.imag(); }
inline cplx cdiv_real(cplx z, double d) { return cplx(z.real() / d, z.imag() / d); }
inline cplx cmul(cplx a, cplx b) {  // num-complex product
  return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}
inline cplx cmulr(double s, cplx z) { return cplx(s * z.real(), s * z.imag()); }
inline cplx root_normalise(cplx z) { return cdiv_real(z, std::sqrt(hypot_norm(z) + 1e-8)); }

// AdjointOps.modulus_vjp (local.rs:289-298)
inline cplx modulus_vjp1(cplx z, double g) {
  double mag = hypot_norm(z);
  return cdiv_real(cmulr(g, z), mag == 0.0 ? 1.0 : mag);
}

// root_modulus_vjp (engine.rs:46-56)
inline cplx root_modulus_vjp1(cplx v, cplx gi) {
  double r = hypot_norm(v);
  double radial = cmul(std::conj(gi), v).real();
  double mag = r == 0.0 ? 1.0 : r;
  cplx a = cdiv_real(gi, std::sqrt(r + 1e-8));
  double s = 0.5 * radial / std::pow(r + 1e-8, 1.5);
  cplx b = cdiv_real(cmulr(s, v), mag);
  return a - b;
}

// ---------------------------------------------------------------------------
// Engine (engine.rs): full-sky sweeps.
// ---------------------------------------------------------------------------

struct Engine {
  int64_t nside0;
  int J, L, ksz, gauge, C, B, levels;
  bool fewer;
  std::vector<int> j_to_dg;
  int64_t block_pixels;
  std::vector<Coef> plan;
  std::vector<double> ker;  // [G][2][P], f32-rounded
  std::vector<LevelConv> conv;
  std::vector<Csr> down, up;  // D_d: level d -> d+1, and its transpose
  std::vector<std::vector<double>> maps;           // [B*C][N_d]
  std::map<std::pair<int, int>, std::vector<double>> m;  // (d, n): [B*C*L][N_n]
  std::vector<std::vector<cplx>> zc;                 // cached z_n [B*C][L][N_n]
  // Masks (mask.rs:14-70): mw[k][n][p] = 1 on the pixels of mask k at level
  // n (level n + 1: the parents of level n), counts mcnt[k][n]; none: the
  // whole sky.
  std::vector<std::vector<std::vector<uint8_t>>> mw;
  std::vector<std::vector<int64_t>> mcnt;

  int nmask() const { return mw.empty() ? 1 : int(mw.size()); }

  // Masks::from_weights (mask.rs:43-70): nonzero level-0 weights are in.
  void set_masks(const double* w, int k) {
    mw.clear();
    mcnt.clear();
    for (int q = 0; q < k; ++q) {
      std::vector<uint8_t> lev(size_t(pixels(0)));
      for (int64_t p = 0; p < pixels(0); ++p) {
        double v = w[size_t(q) * pixels(0) + p];
        if (!std::isfinite(v)) throw std::runtime_error("mask weights must be finite");
        lev[p] = v != 0.0;
      }
      std::vector<std::vector<uint8_t>> per;
      std::vector<int64_t> cnt;
      for (int n = 0; n < levels; ++n) {
        if (n > 0) {
          std::vector<uint8_t> up(lev.size() / 4);
          for (size_t p = 0; p < up.size(); ++p) up[p] = lev[4 * p] | lev[4 * p + 1] | lev[4 * p + 2] | lev[4 * p + 3];
          lev.swap(up);
        }
        int64_t c = 0;
        for (uint8_t v : lev) c += v;
        if (c == 0) throw std::runtime_error("empty mask");
        per.push_back(lev);
        cnt.push_back(c);
      }
      mw.push_back(per);
      mcnt.push_back(cnt);
    }
  }

  int64_t nside(int d) const { return nside0 >> d; }
  int64_t pixels(int d) const { return npix(nside(d)); }
  int rows() const { return B * C; }

  void init() {
    levels = 1;
    for (int d : j_to_dg) levels = std::max(levels, d + 1);
    plan = build_plan(J, L, j_to_dg, C);
    std::vector<double> w = wavelet_weights(ksz);
    int P = ksz * ksz;
    ker.assign(size_t(L) * 2 * P, 0.0);
    for (int g = 0; g < L; ++g)
      for (int q = 0; q < 2 * P; ++q) ker[size_t(g) * 2 * P + q] = double(float(w[q]));
  }

  void build_geometry(bool cache_limit_rows) {
    conv.resize(levels);
    for (int d = 0; d < levels; ++d) conv[d].init(nside(d), ksz, L, gauge, pixels(d) <= cache_limit_rows);
    down.clear();
    up.clear();
    for (int d = 0; d + 1 < levels; ++d) {
      down.push_back(full_decimation(nside(d)));
      up.push_back(down.back().transpose());
    }
  }

  void load(const double* x) {
    maps.assign(levels, {});
    maps[0].assign(x, x + size_t(rows()) * pixels(0));
    for (int d = 0; d + 1 < levels; ++d) {
      maps[d + 1].assign(size_t(rows()) * pixels(d + 1), 0.0);
      for (int r = 0; r < rows(); ++r)
        down[d].apply(maps[d].data() + size_t(r) * pixels(d), maps[d + 1].data() + size_t(r) * pixels(d + 1));
    }
    m.clear();
    zc.assign(levels, {});
  }

  const std::vector<cplx>& z_of(int n) {
    if (zc[n].empty()) {
      zc[n].assign(size_t(rows()) * L * pixels(n), cplx());
      conv[n].forward(maps[n].data(), rows(), ker, zc[n].data());
    }
    return zc[n];
  }

  // Modulus of level d <= n on level n: [B*C*L][N_n].
  std::vector<double> modulus(int d, int n) {
    if (d == n) {
      const auto& z = z_of(n);
      std::vector<double> a(z.size());
      for (size_t i = 0; i < z.size(); ++i) a[i] = hypot_norm(z[i]);
      return a;
    }
    return m.at({d, n});
  }

  // Second wavelet of modulus(d, n): [B][C][L(l1)][L(l3)][N_n].
  std::vector<cplx> branch(int d, int n) {
    std::vector<double> md = modulus(d, n);
    std::vector<cplx> br(size_t(rows()) * L * L * pixels(n));
    conv[n].forward(md.data(), rows() * L, ker, br.data());
    return br;
  }

  // Operands (engine.rs:956-984) as full-sky [B][N_n] complex vectors.
  void operands(const Coef& c, int n, std::map<int, std::vector<cplx>>& branches, std::vector<cplx>& a,
                std::vector<cplx>& b) {
    int64_t N = pixels(n);
    auto z_core = [&](int ch, int l, std::vector<cplx>& out) {
      const auto& z = z_of(n);
      out.resize(size_t(B) * N);
      for (int bi = 0; bi < B; ++bi)
        std::memcpy(&out[size_t(bi) * N], &z[((size_t(bi) * C + ch) * L + l) * N], N * sizeof(cplx));
    };
    auto map_core = [&](int level, int ch, std::vector<cplx>& out) {
      out.resize(size_t(B) * pixels(level));
      for (int bi = 0; bi < B; ++bi)
        for (int64_t p = 0; p < pixels(level); ++p)
          out[size_t(bi) * pixels(level) + p] = cplx(maps[level][(size_t(bi) * C + ch) * pixels(level) + p], 0.0);
    };
    auto br_of = [&](int d) -> const std::vector<cplx>& {
      if (!branches.count(d)) branches[d] = branch(d, n);
      return branches[d];
    };
    auto branch_core = [&](int d, int ch, int l1, int l3, std::vector<cplx>& out) {
      const auto& br = br_of(d);
      out.resize(size_t(B) * N);
      for (int bi = 0; bi < B; ++bi)
        std::memcpy(&out[size_t(bi) * N], &br[((((size_t(bi) * C + ch) * L + l1) * L + l3) * N)], N * sizeof(cplx));
    };
    auto modulus_core = [&](int d, int ch, int l1, std::vector<cplx>& out) {
      std::vector<double> md = modulus(d, n);
      out.resize(size_t(B) * N);
      for (int bi = 0; bi < B; ++bi)
        for (int64_t p = 0; p < N; ++p) out[size_t(bi) * N + p] = cplx(md[((size_t(bi) * C + ch) * L + l1) * N + p], 0.0);
    };
    a.clear();
    b.clear();
    switch (c.family) {
      case F_MEAN:
      case F_VAR: map_core(0, c.c1, a); break;
      case F_S1:
      case F_S2: z_core(c.c1, c.l1, a); z_core(c.c2, c.l1, b); break;
      case F_S3: {
        int d1 = j_to_dg[c.j1];
        if (!fewer) { branch_core(d1, c.c1, c.l1, c.l3, a); z_core(c.c2, c.l3, b); }
        else { map_core(n, c.c1, a); branch_core(d1, c.c2, c.l1, c.l3, b); }
        break;
      }
      case F_S4: {
        int d1 = j_to_dg[c.j1], d2 = j_to_dg[c.j2];
        if (fewer) modulus_core(d1, c.c1, c.l1, a);
        else branch_core(d1, c.c1, c.l1, c.l3, a);
        branch_core(d2, c.c2, c.l2, c.l3, b);
        break;
      }
    }
  }

  // term_sums / term_second_sums (engine.rs:1194-1240), blocks of block_pixels;
  // with a mask (sel not null) the pixels of the mask only, over n_total
  // = its pixel count (engine.rs:521-540).
  void partial(const Coef& c, const std::vector<cplx>& a, const std::vector<cplx>& b, int64_t N, cplx* values,
               double* second, const uint8_t* sel = nullptr, double n_total = 0.0) {
    if (!sel) n_total = double(N);
    for (int bi = 0; bi < B; ++bi) {
      const cplx* pa = &a[size_t(bi) * N];
      const cplx* pb = b.empty() ? nullptr : &b[size_t(bi) * N];
      for (int64_t s0 = 0; s0 < N; s0 += block_pixels) {
        int64_t s1 = std::min(N, s0 + block_pixels);
        cplx sum(0.0, 0.0);
        double sum2 = 0.0;
        for (int64_t p = s0; p < s1; ++p) {
          if (sel && !sel[p]) continue;
          cplx t;
          double t2;
          switch (c.family) {
            case F_MEAN: t = pa[p]; t2 = norm_sqr(pa[p]); break;
            case F_VAR: { double v = norm_sqr_exact(pa[p]); t = cplx(v, 0.0); t2 = v * v; break; }
            default:
              if (c.family == F_S2 && c.c1 == c.c2) { double v = norm_sqr_exact(pa[p]); t = cplx(v, 0.0); t2 = v * v; }
              else if (c.family == F_S1 && c.c1 == c.c2) { t = cplx(hypot_norm(pa[p]), 0.0); t2 = norm_sqr(pa[p]); }
              else if (c.family == F_S1) { t = cmul(root_normalise(pa[p]), std::conj(root_normalise(pb[p]))); t2 = norm_sqr(t); }
              else { t = cmul(pa[p], std::conj(pb[p])); t2 = norm_sqr(t); }
          }
          sum = cplx(sum.real() + t.real(), sum.imag() + t.imag());
          sum2 += t2;
        }
        cplx v = cdiv_real(sum, n_total);
        values[bi] = cplx(values[bi].real() + v.real(), values[bi].imag() + v.imag());
        if (second) second[bi] += sum2 / n_total;
      }
    }
  }

  // Forward (engine.rs:500-602); values [B][K][np], second [B][K][np] when
  // moments (K masks, 1 without).
  void forward(cplx* values, double* second) {
    size_t np = plan.size();
    const int K = nmask();
    for (size_t i = 0; i < size_t(B) * K * np; ++i) values[i] = cplx(0.0, 0.0);
    if (second) std::fill(second, second + size_t(B) * K * np, 0.0);
    m.clear();
    for (int d = 0; d < levels; ++d)
      for (int n = d + 1; n < levels; ++n) m[{d, n}].assign(size_t(rows()) * L * pixels(n), 0.0);
    std::vector<cplx> vb(B);
    std::vector<double> sb(B);
    for (int n = 0; n < levels; ++n) {
      std::map<int, std::vector<cplx>> branches;
      for (size_t i = 0; i < np; ++i) {
        if (plan[i].native_dg != n) continue;
        std::vector<cplx> a, b;
        operands(plan[i], n, branches, a, b);
        const int lev = plan[i].family <= F_VAR ? 0 : n;
        for (int k = 0; k < K; ++k) {
          std::fill(vb.begin(), vb.end(), cplx());
          std::fill(sb.begin(), sb.end(), 0.0);
          const uint8_t* sel = mw.empty() ? nullptr : mw[k][lev].data();
          double total = mw.empty() ? 0.0 : double(mcnt[k][lev]);
          partial(plan[i], a, b, pixels(lev), vb.data(), second ? sb.data() : nullptr, sel, total);
          for (int bi = 0; bi < B; ++bi) {
            size_t o = (size_t(bi) * K + k) * np + i;
            values[o] = cplx(values[o].real() + vb[bi].real(), values[o].imag() + vb[bi].imag());
            if (second) second[o] += sb[bi];
          }
        }
      }
      // m[n][n2] = D^(n2-n) |z_n|
      std::vector<double> a = modulus(n, n);
      for (int n2 = n + 1; n2 < levels; ++n2) {
        std::vector<double> next(size_t(rows()) * L * pixels(n2));
        for (int r = 0; r < rows() * L; ++r)
          down[n2 - 1].apply(a.data() + size_t(r) * pixels(n2 - 1), next.data() + size_t(r) * pixels(n2));
        a.swap(next);
        m[{n, n2}] = a;
      }
    }
  }

  // operand_cotangents (engine.rs:1247-1277)
  void operand_cotangents(const Coef& c, const std::vector<cplx>& a, const std::vector<cplx>& b, const cplx* g,
                          int64_t nf, double n_total, std::vector<cplx>& ga, std::vector<cplx>& gb) {
    size_t tot = size_t(B) * nf;
    ga.assign(tot, cplx());
    gb.clear();
    switch (c.family) {
      case F_MEAN:
        for (size_t k = 0; k < tot; ++k) ga[k] = cplx(g[k / nf].real() / n_total, 0.0);
        return;
      case F_VAR:
        for (size_t k = 0; k < tot; ++k) ga[k] = cplx(2.0 * g[k / nf].real() * a[k].real() / n_total, 0.0);
        return;
      default: break;
    }
    if ((c.family == F_S1 || c.family == F_S2) && c.c1 == c.c2) {
      if (c.family == F_S1) {
        for (size_t k = 0; k < tot; ++k) ga[k] = modulus_vjp1(a[k], g[k / nf].real() / n_total);
      } else {
        for (size_t k = 0; k < tot; ++k) ga[k] = cdiv_real(cmulr(2.0 * g[k / nf].real(), a[k]), n_total);
      }
      return;
    }
    gb.assign(tot, cplx());
    if (c.family == F_S1) {
      for (size_t k = 0; k < tot; ++k) {
        cplx gi = g[k / nf];
        cplx x = cdiv_real(cmul(gi, root_normalise(b[k])), n_total);
        cplx y = cdiv_real(cmul(std::conj(gi), root_normalise(a[k])), n_total);
        ga[k] = root_modulus_vjp1(a[k], x);
        gb[k] = root_modulus_vjp1(b[k], y);
      }
      return;
    }
    for (size_t k = 0; k < tot; ++k) {
      cplx gi = g[k / nf];
      ga[k] = cdiv_real(cmul(gi, b[k]), n_total);
      gb[k] = cdiv_real(cmul(std::conj(gi), a[k]), n_total);
    }
  }

  // Reverse (engine.rs:610-724, 1062-1174): map cotangents of every level
  // for the cotangents gradient [B][K][np].
  std::vector<std::vector<double>> reverse(const cplx* gradient) {
    size_t np = plan.size();
    const int K = nmask();
    std::vector<bool> active(np, false);
    for (size_t i = 0; i < np; ++i)
      for (int r = 0; r < B * K; ++r) active[i] = active[i] || gradient[r * np + i] != cplx(0.0, 0.0);
    std::map<std::pair<int, int>, std::vector<double>> gm;
    for (auto& kv : m) gm[kv.first].assign(kv.second.size(), 0.0);
    std::vector<std::vector<double>> mg(levels);
    for (int d = 0; d < levels; ++d) mg[d].assign(size_t(rows()) * pixels(d), 0.0);
    for (int n = levels - 1; n >= 0; --n) {
      int64_t N = pixels(n);
      const auto& z = z_of(n);
      std::vector<cplx> gz(z.size(), cplx());
      std::map<int, std::vector<cplx>> branches;
      std::vector<std::pair<int, std::vector<cplx>>> gbranch;
      std::vector<std::pair<int, std::vector<double>>> gmod;
      auto gbranch_slot = [&](int d) -> std::vector<cplx>& {
        for (auto& e : gbranch)
          if (e.first == d) return e.second;
        gbranch.push_back({d, std::vector<cplx>(size_t(rows()) * L * L * N, cplx())});
        return gbranch.back().second;
      };
      auto gmod_slot = [&](int d) -> std::vector<double>& {
        for (auto& e : gmod)
          if (e.first == d) return e.second;
        gmod.push_back({d, std::vector<double>(size_t(rows()) * L * N, 0.0)});
        return gmod.back().second;
      };
      auto add_z = [&](int ch, int l, const std::vector<cplx>& g) {
        for (int bi = 0; bi < B; ++bi)
          for (int64_t p = 0; p < N; ++p) {
            cplx& t = gz[((size_t(bi) * C + ch) * L + l) * N + p];
            t = cplx(t.real() + g[size_t(bi) * N + p].real(), t.imag() + g[size_t(bi) * N + p].imag());
          }
      };
      auto add_branch = [&](int d, int ch, int l1, int l3, const std::vector<cplx>& g) {
        auto& gb = gbranch_slot(d);
        for (int bi = 0; bi < B; ++bi)
          for (int64_t p = 0; p < N; ++p) {
            cplx& t = gb[((((size_t(bi) * C + ch) * L + l1) * L + l3) * N) + p];
            t = cplx(t.real() + g[size_t(bi) * N + p].real(), t.imag() + g[size_t(bi) * N + p].imag());
          }
      };
      auto add_map = [&](int level, int ch, const std::vector<cplx>& g) {
        int64_t Nl = pixels(level);
        for (int bi = 0; bi < B; ++bi)
          for (int64_t p = 0; p < Nl; ++p) mg[level][(size_t(bi) * C + ch) * Nl + p] += g[size_t(bi) * Nl + p].real();
      };
      auto add_modulus_core = [&](int d, int ch, int l1, const std::vector<cplx>& g) {
        auto& gmv = gmod_slot(d);
        for (int bi = 0; bi < B; ++bi)
          for (int64_t p = 0; p < N; ++p) gmv[((size_t(bi) * C + ch) * L + l1) * N + p] += g[size_t(bi) * N + p].real();
      };
      std::vector<cplx> gcoef(B);
      for (size_t i = 0; i < np; ++i) {
        const Coef& c = plan[i];
        if (c.native_dg != n || !active[i]) continue;
        std::vector<cplx> a, b, ga, gb;
        operands(c, n, branches, a, b);
        int64_t nf = (c.family <= F_VAR) ? pixels(0) : N;
        if (mw.empty()) {
          for (int bi = 0; bi < B; ++bi) gcoef[bi] = gradient[bi * np + i];
          operand_cotangents(c, a, b, gcoef.data(), nf, double(nf), ga, gb);
        } else {
          // Operand cotangents of every mask on its pixels, added
          // (engine.rs:640-662).
          const int lev = c.family <= F_VAR ? 0 : n;
          std::vector<cplx> gak, gbk;
          for (int k = 0; k < K; ++k) {
            for (int bi = 0; bi < B; ++bi) gcoef[bi] = gradient[(size_t(bi) * K + k) * np + i];
            operand_cotangents(c, a, b, gcoef.data(), nf, double(mcnt[k][lev]), gak, gbk);
            const uint8_t* sel = mw[k][lev].data();
            if (ga.empty()) ga.assign(gak.size(), cplx());
            if (!gbk.empty() && gb.empty()) gb.assign(gbk.size(), cplx());
            for (size_t q = 0; q < gak.size(); ++q) {
              if (!sel[q % size_t(nf)]) continue;
              ga[q] = cplx(ga[q].real() + gak[q].real(), ga[q].imag() + gak[q].imag());
              if (!gbk.empty()) gb[q] = cplx(gb[q].real() + gbk[q].real(), gb[q].imag() + gbk[q].imag());
            }
          }
        }
        switch (c.family) {
          case F_MEAN:
          case F_VAR: add_map(0, c.c1, ga); break;
          case F_S1:
          case F_S2:
            add_z(c.c1, c.l1, ga);
            if (c.c1 != c.c2) add_z(c.c2, c.l1, gb);
            break;
          case F_S3: {
            int d1 = j_to_dg[c.j1];
            if (fewer) { add_map(n, c.c1, ga); add_branch(d1, c.c2, c.l1, c.l3, gb); }
            else { add_branch(d1, c.c1, c.l1, c.l3, ga); add_z(c.c2, c.l3, gb); }
            break;
          }
          case F_S4: {
            int d1 = j_to_dg[c.j1];
            if (fewer) add_modulus_core(d1, c.c1, c.l1, ga);
            else add_branch(d1, c.c1, c.l1, c.l3, ga);
            add_branch(j_to_dg[c.j2], c.c2, c.l2, c.l3, gb);
            break;
          }
        }
      }
      // finish_reverse (engine.rs:1117-1174)
      for (auto& e : gbranch) {
        auto& slot = gmod_slot(e.first);
        conv[n].adjoint_add(e.second.data(), rows() * L, ker, slot.data());
      }
      for (auto& e : gmod) {
        if (e.first == n) {
          for (size_t k = 0; k < gz.size(); ++k) {
            cplx v = modulus_vjp1(z[k], e.second[k]);
            gz[k] = cplx(gz[k].real() + v.real(), gz[k].imag() + v.imag());
          }
        } else {
          auto& t = gm[{e.first, n}];
          for (size_t k = 0; k < t.size(); ++k) t[k] += e.second[k];
        }
      }
      // the block's own chain to coarser levels: sum_n2 (D^T)^(n2-n) gm[n][n2]
      if (n + 1 < levels) {
        std::vector<double> gf(size_t(rows()) * L * N, 0.0);
        for (int n2 = n + 1; n2 < levels; ++n2) {
          std::vector<double> g = gm[{n, n2}];
          for (int lev = n2 - 1; lev >= n; --lev) {
            std::vector<double> nx(size_t(rows()) * L * pixels(lev));
            for (int r = 0; r < rows() * L; ++r)
              up[lev].apply(g.data() + size_t(r) * pixels(lev + 1), nx.data() + size_t(r) * pixels(lev));
            g.swap(nx);
          }
          for (size_t k = 0; k < gf.size(); ++k) gf[k] += g[k];
        }
        for (size_t k = 0; k < gz.size(); ++k) {
          cplx v = modulus_vjp1(z[k], gf[k]);
          gz[k] = cplx(gz[k].real() + v.real(), gz[k].imag() + v.imag());
        }
      }
      conv[n].adjoint_add(gz.data(), rows(), ker, mg[n].data());
    }
    return mg;
  }

  // map_cotangent + band cotangents (engine.rs:730-779)
  void analysis_vjp(std::vector<std::vector<double>> mg, std::vector<std::vector<double>>& bands) {
    for (int d = levels - 1; d >= 1; --d) {
      std::vector<double> t(size_t(rows()) * pixels(d - 1));
      for (int r = 0; r < rows(); ++r)
        up[d - 1].apply(mg[d].data() + size_t(r) * pixels(d), t.data() + size_t(r) * pixels(d - 1));
      for (size_t k = 0; k < t.size(); ++k) mg[d - 1][k] += t[k];
    }
    bands.assign(1, mg[0]);
    for (int d = 0; d + 1 < levels; ++d) {
      const auto& fine = bands.back();
      std::vector<double> coarse(size_t(rows()) * pixels(d + 1), 0.0);
      for (int r = 0; r < rows(); ++r)
        for (int64_t p = 0; p < pixels(d); ++p)
          coarse[size_t(r) * pixels(d + 1) + p / 4] += fine[size_t(r) * pixels(d) + p];
      bands.push_back(coarse);
    }
  }
};

thread_local std::string g_error;

template <class F>
int guarded(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface for the Python checker (ctypes).
// ---------------------------------------------------------------------------

extern "C" {

const char* or_last_error() { return g_error.c_str(); }

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#endif
}

int or_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_nest2xyf(int64_t pix, int order, int64_t* out3) { nest2xyf(pix, order, out3[0], out3[1], out3[2]); }
int64_t or_xyf2nest(int64_t ix, int64_t iy, int64_t f, int order) { return xyf2nest(ix, iy, f, order); }
void or_xyf2ang(int64_t ix, int64_t iy, int64_t f, int64_t nside, double* out2) { xyf2ang(ix, iy, f, nside, out2[0], out2[1]); }
void or_centre_lonlat(int64_t pix, int64_t nside, double* out2) { centre_lonlat(pix, nside, out2[0], out2[1]); }
int64_t or_ring2nest(int64_t pix, int order) { return ring2nest(pix, order); }
void or_get_interpol_ring(double theta, double phi, int64_t nside, int64_t* pix, double* w) {
  get_interpol_ring(theta, phi, nside, pix, w);
}
void or_interp_weights_nested(const double* v, int order, int64_t* pix, double* w) {
  interp_weights_nested(v, order, pix, w);
}
void or_rotation_matrices(double theta, double phi, int G, int gauge, double* out) {
  std::vector<std::array<double, 9>> r;
  rotation_matrices(theta, phi, G, gauge, r);
  for (int g = 0; g < G; ++g)
    for (int k = 0; k < 9; ++k) out[g * 9 + k] = r[g][k];
}
void or_local_kernel_grid(int ksz, int64_t nside, double* out) {
  auto v = local_kernel_grid(ksz, nside);
  for (size_t p = 0; p < v.size(); ++p)
    for (int k = 0; k < 3; ++k) out[p * 3 + k] = v[p][k];
}
void or_wavelet_weights(int ksz, double* out) {
  auto w = wavelet_weights(ksz);
  std::copy(w.begin(), w.end(), out);
}
int64_t or_ext_pix(int64_t face, int64_t ex, int64_t ey, int64_t nside) { return ext_pix(face, ex, ey, nside, order_of(nside)); }
int or_rel_offset(int64_t row, int64_t cell, int64_t nside, int64_t* out2) {
  return rel_offset(row, cell, nside, order_of(nside), out2[0], out2[1]) ? 1 : 0;
}
int or_kth_neighbourhood(int64_t id, int64_t nside, int k, int64_t* out, int cap) {
  auto v = kth_neighbourhood(id, nside, k);
  int n = int(std::min<size_t>(v.size(), size_t(cap)));
  std::copy(v.begin(), v.begin() + n, out);
  return int(v.size());
}

// Bound stencils of the grid rows on the sorted grid (conv.rs ConvGeometry::new),
// reference layout ((n*G+g)*K+k)*P+p.
int or_conv_geometry(const int64_t* grid, int64_t K, int64_t nside, int ksz, int G, int gauge, int64_t* pos, double* w,
                     uint8_t* complete) {
  return guarded([&] {
    Stencil st;
    st.ksz = ksz; st.G = G; st.gauge = gauge; st.P = ksz * ksz; st.nside = nside; st.order = order_of(nside);
    st.grid_vec = local_kernel_grid(ksz, nside);
    int P = st.P;
#pragma omp parallel
    {
      std::vector<int64_t> pr(size_t(4) * G * P);
      std::vector<double> wr(size_t(4) * G * P);
      std::vector<std::array<double, 9>> rots;
#pragma omp for schedule(dynamic, 16)
      for (int64_t k = 0; k < K; ++k) {
        complete[k] = bind_row(st, grid[k], grid, K, pr.data(), wr.data(), rots) ? 1 : 0;
        for (int g = 0; g < G; ++g)
          for (int p = 0; p < P; ++p)
            for (int n = 0; n < 4; ++n) {
              size_t a = ((size_t(n) * G + g) * K + k) * P + p;
              pos[a] = pr[(size_t(g) * P + p) * 4 + n];
              w[a] = wr[(size_t(g) * P + p) * 4 + n];
            }
      }
    }
  });
}

// Partial-grid decimation; returns the number of entries (or -1), fills up to cap.
int64_t or_smooth_decimation(const int64_t* ids, int64_t n, int64_t nside, int64_t* coarse, int64_t* n_coarse,
                             int64_t* rows, int64_t* cols, double* vals, int64_t cap) {
  int64_t out = -1;
  int rc = guarded([&] {
    std::vector<int64_t> co;
    Csr m;
    smooth_decimation(ids, n, nside, co, m);
    *n_coarse = int64_t(co.size());
    for (size_t i = 0; i < co.size() && int64_t(i) < cap; ++i) coarse[i] = co[i];
    int64_t e = 0;
    for (int64_t r = 0; r < m.n_rows; ++r)
      for (int64_t q = m.indptr[r]; q < m.indptr[r + 1]; ++q, ++e)
        if (e < cap) { rows[e] = r; cols[e] = m.indices[q]; vals[e] = m.values[q]; }
    out = e;
  });
  return rc == 0 ? out : -1;
}

// Full-sky decimation rows of the coarse cells [c0, c1) from nside: CSR over those rows.
int64_t or_decimation_rows(int64_t nside, int64_t c0, int64_t c1, int64_t* indptr, int64_t* cols, double* vals, int64_t cap) {
  int64_t e = 0;
  indptr[0] = 0;
  std::vector<int64_t> cc;
  std::vector<double> vv;
  for (int64_t c = c0; c < c1; ++c) {
    decimation_row(c, nside, cc, vv);
    for (size_t i = 0; i < cc.size(); ++i, ++e)
      if (e < cap) { cols[e] = cc[i]; vals[e] = vv[i]; }
    indptr[c - c0 + 1] = e;
  }
  return e;
}

// Plan rows [family, native_dg, c1, c2, j1, j2, j3, l1, l2, l3] (-1 for None).
int or_plan(int J, int L, const int* j_to_dg, int C, int* out, int cap) {
  std::vector<int> dg(j_to_dg, j_to_dg + J);
  auto plan = build_plan(J, L, dg, C);
  for (size_t i = 0; i < plan.size() && int(i) < cap; ++i) {
    const Coef& c = plan[i];
    int row[10] = {c.family, c.native_dg, c.c1, c.c2, c.j1, c.j2, c.j3, c.l1, c.l2, c.l3};
    std::copy(row, row + 10, out + 10 * i);
  }
  return int(plan.size());
}

void* or_engine_new(int64_t nside, int J, int L, int ksz, int gauge, int C, int B, int fewer, const int* j_to_dg,
                    int64_t block_pixels, int64_t cache_rows) {
  Engine* e = nullptr;
  int rc = guarded([&] {
    order_of(nside);
    e = new Engine();
    e->nside0 = nside; e->J = J; e->L = L; e->ksz = ksz; e->gauge = gauge; e->C = C; e->B = B;
    e->fewer = fewer != 0;
    e->j_to_dg.assign(j_to_dg, j_to_dg + J);
    e->block_pixels = block_pixels > 0 ? block_pixels : 4096;
    e->init();
    e->build_geometry(cache_rows);
  });
  if (rc != 0) { delete e; return nullptr; }
  return e;
}

void or_engine_free(void* h) { delete static_cast<Engine*>(h); }
int or_engine_levels(void* h) { return static_cast<Engine*>(h)->levels; }
int or_engine_ncoef(void* h) { return int(static_cast<Engine*>(h)->plan.size()); }

int or_engine_load(void* h, const double* x) {
  return guarded([&] { static_cast<Engine*>(h)->load(x); });
}

// k masks [k][12 nside^2] (nonzero: in), or k = 0: the whole sky.
int or_engine_set_masks(void* h, const double* w, int k) {
  return guarded([&] { static_cast<Engine*>(h)->set_masks(w, k); });
}

// Pixels of mask k at level n.
int64_t or_engine_mask_count(void* h, int k, int n) {
  Engine* e = static_cast<Engine*>(h);
  if (e->mw.empty()) return e->pixels(n);
  return e->mcnt[k][n];
}

// values: [B][K][np] interleaved (re, im); second: [B][K][np] or null.
int or_engine_forward(void* h, double* values, double* second) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    std::vector<cplx> v(size_t(e->B) * e->nmask() * e->plan.size());
    e->forward(v.data(), second);
    for (size_t i = 0; i < v.size(); ++i) { values[2 * i] = v[i].real(); values[2 * i + 1] = v[i].imag(); }
  });
}

// cot: [B][K][np] interleaved; maps_grad: concatenation over levels of [B*C][N_d].
int or_engine_reverse(void* h, const double* cot, double* maps_grad) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    std::vector<cplx> g(size_t(e->B) * e->nmask() * e->plan.size());
    for (size_t i = 0; i < g.size(); ++i) g[i] = cplx(cot[2 * i], cot[2 * i + 1]);
    auto mg = e->reverse(g.data());
    size_t at = 0;
    for (auto& v : mg) { std::copy(v.begin(), v.end(), maps_grad + at); at += v.size(); }
  });
}

// From the concatenated per-level map cotangents: level-0 gradient and bands (both concatenated).
int or_engine_analysis_vjp(void* h, const double* maps_grad, double* bands) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    std::vector<std::vector<double>> mg(e->levels), out;
    size_t at = 0;
    for (int d = 0; d < e->levels; ++d) {
      size_t len = size_t(e->rows()) * e->pixels(d);
      mg[d].assign(maps_grad + at, maps_grad + at + len);
      at += len;
    }
    e->analysis_vjp(mg, out);
    at = 0;
    for (auto& v : out) { std::copy(v.begin(), v.end(), bands + at); at += v.size(); }
  });
}

int or_engine_get_map(void* h, int d, double* out) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    std::copy(e->maps.at(d).begin(), e->maps[d].end(), out);
  });
}

int or_engine_get_m(void* h, int d, int n, double* out) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    const auto& v = e->m.at({d, n});
    std::copy(v.begin(), v.end(), out);
  });
}

// z_n [B*C][L][N_n] interleaved.
int or_engine_get_z(void* h, int n, double* out) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    const auto& z = e->z_of(n);
    for (size_t i = 0; i < z.size(); ++i) { out[2 * i] = z[i].real(); out[2 * i + 1] = z[i].imag(); }
  });
}

// One full-sky level convolution x [R][N] -> y [R][L][N] interleaved.
int or_engine_conv(void* h, int n, const double* x, int R, double* y) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    std::vector<cplx> out(size_t(R) * e->L * e->pixels(n));
    e->conv.at(n).forward(x, R, e->ker, out.data());
    for (size_t i = 0; i < out.size(); ++i) { y[2 * i] = out[i].real(); y[2 * i + 1] = out[i].imag(); }
  });
}

// Its adjoint gy [R][L][N] interleaved -> gx [R][N].
int or_engine_conv_adjoint(void* h, int n, const double* gy, int R, double* gx) {
  return guarded([&] {
    Engine* e = static_cast<Engine*>(h);
    size_t len = size_t(R) * e->L * e->pixels(n);
    std::vector<cplx> g(len);
    for (size_t i = 0; i < len; ++i) g[i] = cplx(gy[2 * i], gy[2 * i + 1]);
    std::fill(gx, gx + size_t(R) * e->pixels(n), 0.0);
    e->conv.at(n).adjoint_add(g.data(), R, e->ker, gx);
  });
}

// Kernel [L][2][P] (f32-rounded values).
void or_engine_kernel(void* h, double* out) {
  Engine* e = static_cast<Engine*>(h);
  std::copy(e->ker.begin(), e->ker.end(), out);
}

// Bound stencil of one full-sky row: pos/w [G][P][4].
int or_full_row(int64_t nside, int ksz, int G, int gauge, int64_t row, int64_t* pos, double* w) {
  return guarded([&] {
    Stencil st;
    st.ksz = ksz; st.G = G; st.gauge = gauge; st.P = ksz * ksz; st.nside = nside; st.order = order_of(nside);
    st.grid_vec = local_kernel_grid(ksz, nside);
    std::vector<std::array<double, 9>> rots;
    bind_row(st, row, nullptr, 0, pos, w, rots);
  });
}

}  // extern "C"
