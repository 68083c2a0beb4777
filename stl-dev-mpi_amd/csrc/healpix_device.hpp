// HEALPix NESTED geometry on the device, FP64, for the one-off build of the
// pyramid operators.  Same formulas and operation order as the reference's
// stl-core healpix.rs / stencil.rs / decimation.rs (cited per function), so
// that the operators match the reference's to rounding; acos/atan2/... are
// HIP's own device libm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace stl {
namespace hp {

constexpr double kPi = 3.14159265358979323846;
// torch.rad2deg / deg2rad constants (healpix.rs:12-14).
constexpr double kRad2Deg = 57.295779513082320876798154814105170332405472466564;
constexpr double kDeg2Rad = 0.017453292519943295769236907684886127134428718885417;

__device__ inline double rad2deg(double x) { return x * kRad2Deg; }
__device__ inline double deg2rad(double x) { return x * kDeg2Rad; }

// healpix.rs:33-39 (torch.remainder)
__device__ inline double remainder_py(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0 && ((b < 0.0) != (m < 0.0))) m += b;
  return m;
}

__host__ __device__ inline int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}

// Base-face row / column offsets (healpix_base jrll, jpll).
__host__ __device__ inline int64_t jrll(int f) { return 2 + f / 4; }
__host__ __device__ inline int64_t jpll(int f) { return 2 * (f & 3) + ((f >> 2) == 1 ? 0 : 1); }

__host__ __device__ inline int64_t compress_bits(int64_t v) {
  v &= 0x5555555555555555LL;
  v = (v | (v >> 1)) & 0x3333333333333333LL;
  v = (v | (v >> 2)) & 0x0F0F0F0F0F0F0F0FLL;
  v = (v | (v >> 4)) & 0x00FF00FF00FF00FFLL;
  v = (v | (v >> 8)) & 0x0000FFFF0000FFFFLL;
  return (v | (v >> 16)) & 0x00000000FFFFFFFFLL;
}

__host__ __device__ inline int64_t spread_bits(int64_t v) {
  v &= 0x00000000FFFFFFFFLL;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFLL;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFLL;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FLL;
  v = (v | (v << 2)) & 0x3333333333333333LL;
  return (v | (v << 1)) & 0x5555555555555555LL;
}

__device__ inline void nest2xyf(int64_t pix, int order, int64_t& ix, int64_t& iy, int& face) {
  face = int(pix >> (2 * order));
  int64_t local = pix & ((int64_t(1) << (2 * order)) - 1);
  ix = compress_bits(local);
  iy = compress_bits(local >> 1);
}

__host__ __device__ inline int64_t xyf2nest(int64_t ix, int64_t iy, int face, int order) {
  return (int64_t(face) << (2 * order)) + spread_bits(ix) + (spread_bits(iy) << 1);
}

// Pixel centre (healpix.rs:88-127).
__device__ inline void xyf2ang(int64_t ix, int64_t iy, int face, int64_t nside, double& theta, double& phi) {
  int64_t jr = jrll(face) * nside - ix - iy - 1;
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
    double sth_cap = sqrt(tmp * (2.0 - tmp));
    double theta_cap = atan2(sth_cap, z_cap);
    theta = north ? theta_cap : kPi - theta_cap;
  } else {
    theta = acos((2.0 * ns - double(jr)) * fact1);
  }
  int64_t jp = floordiv(jpll(face) * nr + ix - iy + 1 + kshift, 2);
  if (jp > 4 * nr) jp -= 4 * nr;
  if (jp < 1) jp += 4 * nr;
  phi = (double(jp) - (double(kshift) + 1.0) * 0.5) * (0.5 * kPi / nrf);
}

__device__ inline void pix2ang_nest(int64_t pix, int order, double& theta, double& phi) {
  int64_t ix, iy;
  int f;
  nest2xyf(pix, order, ix, iy, f);
  xyf2ang(ix, iy, f, int64_t(1) << order, theta, phi);
}

// Centre (lon, lat) rounded through degrees (healpix.rs:140-143).
__device__ inline void centre_lonlat(int64_t pix, int order, double& lon, double& lat) {
  double theta, phi;
  pix2ang_nest(pix, order, theta, phi);
  lon = deg2rad(rad2deg(phi));
  lat = deg2rad(90.0 - rad2deg(theta));
}

__host__ __device__ inline int64_t isqrt(int64_t v) {
  int64_t r = int64_t(sqrt(double(v)));
  if (r * r > v) r -= 1;
  if ((r + 1) * (r + 1) <= v) r += 1;
  return r;
}

// RING -> NESTED (healpix.rs:157-210).
__host__ __device__ inline int64_t ring2nest(int64_t pix, int order) {
  int64_t nside = int64_t(1) << order;
  int64_t np = 12 * nside * nside;
  int64_t ncap = 2 * nside * (nside - 1);
  int64_t nl2 = 2 * nside;
  int64_t iring, iphi, kshift, nr;
  int face;
  if (pix < ncap) {
    int64_t ir = (1 + isqrt(1 + 2 * pix)) >> 1;
    int64_t ip = pix + 1 - 2 * ir * (ir - 1);
    iring = ir; iphi = ip; kshift = 0; nr = ir;
    face = int(floordiv(ip - 1, ir > 1 ? ir : 1));
  } else if (pix < np - ncap) {
    int64_t ip = pix - ncap;
    int64_t tmp = ip >> (order + 2);
    int64_t ir = tmp + nside;
    int64_t ip2 = ip - tmp * 4 * nside + 1;
    int64_t ire = ir - nside + 1;
    int64_t irm = nl2 + 2 - ire;
    int64_t ifm = (ip2 - floordiv(ire, 2) + nside - 1) >> order;
    int64_t ifp = (ip2 - floordiv(irm, 2) + nside - 1) >> order;
    iring = ir; iphi = ip2; kshift = (ir + nside) & 1; nr = nside;
    face = int((ifp == ifm) ? (ifp | 4) : (ifp < ifm ? ifp : ifm + 8));
  } else {
    int64_t ip = np - pix;
    int64_t ir = (1 + isqrt(2 * ip - 1)) >> 1;
    int64_t ip2 = 4 * ir + 1 - (ip - 2 * ir * (ir - 1));
    iring = 2 * nl2 - ir; iphi = ip2; kshift = 0; nr = ir;
    face = 8 + int(floordiv(ip2 - 1, ir > 1 ? ir : 1));
  }
  int64_t irt = iring - jrll(face) * nside + 1;
  int64_t ipt = 2 * iphi - jpll(face) * nr - kshift - 1;
  if (ipt >= nl2) ipt -= 8 * nside;
  int64_t ix = (ipt - irt) >> 1;
  int64_t iy = (-ipt - irt) >> 1;
  return xyf2nest(ix, iy, face, order);
}

// healpix.rs:216-229
__device__ inline int64_t ring_above(double z, int64_t nside) {
  double az = fabs(z);
  double ns = double(nside);
  if (az <= 2.0 / 3.0) return int64_t(trunc(ns * (2.0 - 1.5 * z)));
  int64_t iring = int64_t(trunc(ns * sqrt(3.0 * (1.0 - az))));
  return z > 0.0 ? iring : 4 * nside - iring - 1;
}

// healpix.rs:232-259
__device__ inline void ring_info(int64_t ring, int64_t nside, int64_t& startpix, int64_t& ringpix, double& theta,
                                 bool& shifted) {
  int64_t ncap = 2 * nside * (nside - 1);
  int64_t np = 12 * nside * nside;
  double ns = double(nside);
  double fact1 = 2.0 / (3.0 * ns);
  double fact2 = 1.0 / (3.0 * ns * ns);
  int64_t northring = ring > 2 * nside ? 4 * nside - ring : ring;
  if (northring < nside) {
    double nrf = double(northring);
    double tmp = nrf * nrf * fact2;
    theta = atan2(sqrt(tmp * (2.0 - tmp)), 1.0 - tmp);
    ringpix = 4 * northring;
    shifted = true;
    startpix = 2 * northring * (northring - 1);
  } else {
    theta = acos(double(2 * nside - northring) * fact1);
    ringpix = 4 * nside;
    shifted = ((northring - nside) & 1) == 0;
    startpix = ncap + (northring - nside) * 4 * nside;
  }
  if (northring != ring) {
    theta = kPi - theta;
    startpix = np - startpix - ringpix;
  }
}

// healpix.rs:262-280
__device__ inline void ring_pair(double phi, int64_t ring, int64_t nside, int64_t& p0, int64_t& p1, double& w0,
                                 double& w1, double& theta_ring) {
  int64_t sp, nr;
  bool shifted;
  ring_info(ring, nside, sp, nr, theta_ring, shifted);
  double dphi = (2.0 * kPi) / double(nr);
  double half = shifted ? 0.5 : 0.0;
  double tmp = phi / dphi - half;
  int64_t i1 = int64_t(trunc(tmp));
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

// Ring bilinear interpolation, T_Healpix_Base::get_interpol (healpix.rs:283-324).
__device__ inline void get_interpol_ring(double theta, double phi, int64_t nside, int64_t pix[4], double wgt[4]) {
  int64_t np = 12 * nside * nside;
  double z = cos(theta);
  int64_t ir1 = ring_above(z, nside);
  int64_t ir2 = ir1 + 1;
  int64_t p[4] = {0, 0, 0, 0};
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t a0, a1;
  double b0, b1, theta1, theta2;
  ring_pair(phi, ir1 > 1 ? ir1 : 1, nside, a0, a1, b0, b1, theta1);
  if (ir1 > 0) { p[0] = a0; p[1] = a1; w[0] = b0; w[1] = b1; }
  ring_pair(phi, ir2 < 4 * nside - 1 ? ir2 : 4 * nside - 1, nside, a0, a1, b0, b1, theta2);
  if (ir2 < 4 * nside) { p[2] = a0; p[3] = a1; w[2] = b0; w[3] = b1; }
  bool north = ir1 == 0;
  bool south = ir2 == 4 * nside && !north;
  if (north) {
    double wn = theta / theta2;
    double facn = (1.0 - wn) * 0.25;
    pix[0] = (p[2] + 2) & 3; pix[1] = (p[3] + 2) & 3; pix[2] = p[2]; pix[3] = p[3];
    wgt[0] = facn; wgt[1] = facn; wgt[2] = w[2] * wn + facn; wgt[3] = w[3] * wn + facn;
  } else if (south) {
    double ws = (theta - theta1) / (kPi - theta1);
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

// Four NESTED cells and weights of a direction, through degrees (healpix.rs:336-347).
__device__ inline void interp_weights_nested(const double v[3], int order, int64_t pix[4], double w[4]) {
  double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double n = nrm > 1e-12 ? nrm : 1e-12;
  double vn0 = v[0] / n, vn1 = v[1] / n, vn2 = v[2] / n;
  double c = vn2 < -1.0 ? -1.0 : (vn2 > 1.0 ? 1.0 : vn2);
  double theta = acos(c);
  double phi = remainder_py(atan2(vn1, vn0), 2.0 * kPi);
  double lon_deg = rad2deg(phi);
  double lat_deg = 90.0 - rad2deg(theta);
  double theta2 = deg2rad(90.0 - lat_deg);
  double phi2 = remainder_py(deg2rad(lon_deg), 2.0 * kPi);
  int64_t rp[4];
  get_interpol_ring(theta2, phi2, int64_t(1) << order, rp, w);
  for (int i = 0; i < 4; ++i) pix[i] = ring2nest(rp[i], order);
}

// Rotation of gauge g taking the polar stencil to (theta, phi), gauge cosmo
// (stencil.rs:51-88): rows of the 3x3 matrix into r[9].
__device__ inline void rotation_cosmo(double theta, double phi, int g, int G, double r[9]) {
  double ct = cos(theta), st = sin(theta);
  double cp = cos(phi), sp = sin(phi);
  double base[3][3] = {{cp * ct, -sp, cp * st}, {sp * ct, cp, sp * st}, {-st, 0.0, ct}};
  double n0[3] = {base[0][2], base[1][2], base[2][2]};
  double nn = sqrt(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
  nn = nn > 1e-12 ? nn : 1e-12;
  double n[3] = {n0[0] / nn, n0[1] / nn, n0[2] / nn};
  double alpha_base, sign_g;
  if (theta > kPi / 2.0) { alpha_base = phi; sign_g = -1.0; }
  else { alpha_base = -phi; sign_g = 1.0; }
  double skew[3][3] = {{0.0, -n[2], n[1]}, {n[2], 0.0, -n[0]}, {-n[1], n[0], 0.0}};
  double step = kPi / double(G);
  double alpha = alpha_base + sign_g * (double(g) * step);
  double ca = cos(alpha), sa = sin(alpha);
  double rg[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double eye = i == j ? 1.0 : 0.0;
      rg[i][j] = eye * ca + skew[i][j] * sa + n[i] * n[j] * (1.0 - ca);
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i * 3 + j] = rg[i][0] * base[0][j] + rg[i][1] * base[1][j] + rg[i][2] * base[2][j];
}

// HEALPix C++ nb_facearray / nb_swaparray by direction 4 + dx + 3 dy (dia.rs:27-41).
static __constant__ const int8_t kNbFace[9][12] = {
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
static __constant__ const uint8_t kNbSwap[9][3] = {{0, 0, 3}, {0, 0, 6}, {0, 0, 0}, {0, 0, 5}, {0, 0, 0},
                                             {5, 0, 0}, {0, 0, 0}, {6, 0, 0}, {3, 0, 0}};

// Pixel at extended coordinates of a face, -1 in the quadrant missing at a
// vertex of three faces (dia.rs:46-77).  Needs -nside <= ex, ey < 2 nside.
__device__ inline int64_t ext_pix(int face, int64_t ex, int64_t ey, int order) {
  int64_t nside = int64_t(1) << order;
  int64_t x = ex, y = ey;
  int nb = 4;
  if (x < 0) { x += nside; nb -= 1; }
  else if (x >= nside) { x -= nside; nb += 1; }
  if (y < 0) { y += nside; nb -= 3; }
  else if (y >= nside) { y -= nside; nb += 3; }
  int f = kNbFace[nb][face];
  if (f < 0) return -1;
  uint8_t bits = kNbSwap[nb][face >> 2];
  if (bits & 1) x = nside - x - 1;
  if (bits & 2) y = nside - y - 1;
  if (bits & 4) { int64_t t = x; x = y; y = t; }
  return xyf2nest(x, y, f, order);
}

// Offset (dx, dy) of `cell` from `row` in the extended frame of row's face;
// false when cell lies on neither that face nor one of its neighbours (dia.rs:79-101).
__device__ inline bool rel_offset(int64_t row, int64_t cell, int order, int& dx, int& dy) {
  int64_t nside = int64_t(1) << order;
  int64_t x, y, xq, yq;
  int f, fq;
  nest2xyf(row, order, x, y, f);
  nest2xyf(cell, order, xq, yq, fq);
  if (fq == f) {
    dx = int(xq - x);
    dy = int(yq - y);
    return true;
  }
  int nb = -1;
  for (int k = 0; k < 9; ++k)
    if (k != 4 && kNbFace[k][f] == fq) { nb = k; break; }
  if (nb < 0) return false;
  uint8_t bits = kNbSwap[nb][f >> 2];
  int64_t a = xq, b = yq;
  if (bits & 4) { int64_t t = a; a = b; b = t; }
  if (bits & 2) b = nside - b - 1;
  if (bits & 1) a = nside - a - 1;
  int64_t xdir = nb % 3 - 1, ydir = nb / 3 - 1;
  dx = int(a + xdir * nside - x);
  dy = int(b + ydir * nside - y);
  return true;
}

// Haversine distance (decimation.rs:464-470).
__device__ inline double haversine(double lat1, double lon1, double lat2, double lon2) {
  double dlat = 0.5 * (lat2 - lat1);
  double dlon = 0.5 * (lon2 - lon1);
  double sa = sin(dlat), so = sin(dlon);
  double a = sa * sa + cos(lat1) * cos(lat2) * (so * so);
  a = a < 0.0 ? 0.0 : (a > 1.0 ? 1.0 : a);
  return 2.0 * asin(sqrt(a));
}

}  // namespace hp
}  // namespace stl
