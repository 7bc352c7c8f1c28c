// gxt_device.h -- device functions of the step loop for gfx950.
//
// Physics state is templated on R: double for the FP64 form, float for the
// precision form (table interpolation, rejection sampling, helix in FP32;
// positions, safety, boundary distances and tallies stay FP64).  The
// algorithms restate the reference's leaf functions (file:line citations are
// relative to GeantExascalePilot/source/Geant/) with the fixes of the
// reference's hard-coded values listed beside each function; the CPU oracle
// implements the same contract.
#pragma once

#include <hip/hip_runtime.h>

#include "gxt_types.h"

namespace gxt {

constexpr double kMe = 0.510998910;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kFineStructure = 1.0 / 137.0;
constexpr double kB2C = -0.299792458;  // MeV/c per (tesla mm)
constexpr double kHuge = 1.0e300;
constexpr double kAvogadro = 6.02214179e+23;
constexpr double kClassicRe = 2.8179402894e-12;

// ------------------------------------------------------------------ math
// FP64 log and exp, and what is built on them, computed the same way by the
// CPU oracle so that both give the same bits (samplers, table bins and
// interaction lengths call them every step; a last-ulp difference between
// two libms grows into a different shower in a strong field).
//  - log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh f,
//    f = (m - 1)/(m + 1), by its series to f^21 (|f| <= 0.172, truncation
//    below 1e-18), plus e ln 2 with ln 2 in two parts (the first of 29 bits,
//    so e ln2_hi is exact);
//  - exp: Cody and Waite's reduction x = k ln 2 + r, |r| <= ln2/2, the
//    Taylor series of e^r to r^13, scaled by 2^k (ldexp, exact);
//  - log1p: log(u) - ((u - 1) - x)/u with u = 1 + x (Kahan's correction);
//  - expm1: the Taylor series to x^14 for |x| < ln2/2, else exp(x) - 1;
//  - pow(x, y) = exp(y log x) (x > 0), log10 x = log x / ln 10.
// Every operation is a correctly rounded IEEE one.
constexpr double kLn2Hi = 0x1.62e42ffp-1;
constexpr double kLn2Lo = -0x1.718432a1b0e26p-35;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kInvLn10 = 0x1.bcb7b1526e50ep-2;
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
__device__ __forceinline__ double log_cw(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);
  if (m < kSqrtHalf) {
    m *= 2.0;
    e -= 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double w = f * f;
  const double t =
      w * (2.0 / 3.0 +
           w * (2.0 / 5.0 +
                w * (2.0 / 7.0 +
                     w * (2.0 / 9.0 +
                          w * (2.0 / 11.0 +
                               w * (2.0 / 13.0 + w * (2.0 / 15.0 + w * (2.0 / 17.0 + w * (2.0 / 19.0 + w * (2.0 / 21.0))))))))));
  const double lm = 2.0 * f + f * t;
  const double de = (double)e;
  return de * kLn2Hi + (de * kLn2Lo + lm);
}
__device__ __forceinline__ double exp_cw(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double k = rint(x * kInvLn2);
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  const double p =
      1.0 +
      r * (1.0 +
           r * (1.0 / 2.0 +
                r * (1.0 / 6.0 +
                     r * (1.0 / 24.0 +
                          r * (1.0 / 120.0 +
                               r * (1.0 / 720.0 +
                                    r * (1.0 / 5040.0 +
                                         r * (1.0 / 40320.0 +
                                              r * (1.0 / 362880.0 +
                                                   r * (1.0 / 3628800.0 +
                                                        r * (1.0 / 39916800.0 +
                                                             r * (1.0 / 479001600.0 + r * (1.0 / 6227020800.0)))))))))))));
  return ldexp(p, (int)k);
}
__device__ __forceinline__ double log1p_cw(double x) {
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  if (!(u > 0.0) || u == HUGE_VAL) return log_cw(u);
  return log_cw(u) - ((u - 1.0) - x) / u;
}
__device__ __forceinline__ double expm1_cw(double x) {
  if (!(fabs(x) < 0.34657359027997264)) return exp_cw(x) - 1.0;
  return x * (1.0 +
              x * (1.0 / 2.0 +
                   x * (1.0 / 6.0 +
                        x * (1.0 / 24.0 +
                             x * (1.0 / 120.0 +
                                  x * (1.0 / 720.0 +
                                       x * (1.0 / 5040.0 +
                                            x * (1.0 / 40320.0 +
                                                 x * (1.0 / 362880.0 +
                                                      x * (1.0 / 3628800.0 +
                                                           x * (1.0 / 39916800.0 +
                                                                x * (1.0 / 479001600.0 +
                                                                     x * (1.0 / 6227020800.0 +
                                                                          x * (1.0 / 87178291200.0))))))))))))));
}
__device__ __forceinline__ double pow_cw(double x, double y) { return exp_cw(y * log_cw(x)); }
__device__ __forceinline__ double log10_cw(double x) { return log_cw(x) * kInvLn10; }

__device__ __forceinline__ double m_log(double x) { return log_cw(x); }
__device__ __forceinline__ float m_log(float x) { return logf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp_cw(x); }
__device__ __forceinline__ float m_exp(float x) { return expf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float m_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return pow_cw(x, y); }
__device__ __forceinline__ double m_expm1(double x) { return expm1_cw(x); }
__device__ __forceinline__ float m_expm1(float x) { return expm1f(x); }
__device__ __forceinline__ float m_pow(float x, float y) { return powf(x, y); }
// FP64 sine and cosine, computed the same way by the CPU oracle so that both
// give the same bits (a helix turns by sin/cos every step, and a last-ulp
// difference between two libms grows into a different shower).  Cody and
// Waite's additive reduction x = k pi/2 + r, with k = rint(x 2/pi) and pi/2
// taken off in three parts (the first two hold 33 bits, so k P1 and k P2 are
// exact for |k| < 2^20); then the Taylor series of sin to r^15 and of cos to
// r^16 on |r| <= pi/4 (truncation below 1e-16 relative), and the quadrant
// k mod 4.  Every operation is a correctly rounded IEEE one.
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;  // 2/pi
constexpr double kPio2P1 = 0x1.921fb544p+0;            // pi/2 to 33 bits
constexpr double kPio2P2 = 0x1.0b4611a6p-34;           // the next 33 bits
constexpr double kPio2P3 = 0x1.3198a2e037073p-69;      // the rest, to 53 bits
__device__ __forceinline__ void sincos_cw(double x, double &s, double &c) {
  if (!(fabs(x) <= 1e15)) {
    s = x - x;
    c = x - x;
    return;
  }
  const double k = rint(x * kTwoOverPi);
  const double r = ((x - k * kPio2P1) - k * kPio2P2) - k * kPio2P3;
  const double z = r * r;
  const double ps =
      -1.0 / 6.0 +
      z * (1.0 / 120.0 +
           z * (-1.0 / 5040.0 +
                z * (1.0 / 362880.0 + z * (-1.0 / 39916800.0 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0))))));
  const double pc =
      1.0 / 24.0 +
      z * (-1.0 / 720.0 +
           z * (1.0 / 40320.0 +
                z * (-1.0 / 3628800.0 + z * (1.0 / 479001600.0 + z * (-1.0 / 87178291200.0 + z * (1.0 / 20922789888000.0))))));
  const double sr = r + r * z * ps;
  const double cr = (1.0 - 0.5 * z) + z * z * pc;
  double q = fmod(k, 4.0);
  if (q < 0.0) q += 4.0;
  if (q == 0.0) {
    s = sr;
    c = cr;
  } else if (q == 1.0) {
    s = cr;
    c = -sr;
  } else if (q == 2.0) {
    s = -sr;
    c = -cr;
  } else {
    s = -cr;
    c = sr;
  }
}
__device__ __forceinline__ double m_sin(double x) {
  double s, c;
  sincos_cw(x, s, c);
  return s;
}
__device__ __forceinline__ float m_sin(float x) { return sinf(x); }
__device__ __forceinline__ double m_cos(double x) {
  double s, c;
  sincos_cw(x, s, c);
  return c;
}
__device__ __forceinline__ float m_cos(float x) { return cosf(x); }
__device__ __forceinline__ double m_fabs(double x) { return fabs(x); }
__device__ __forceinline__ float m_fabs(float x) { return fabsf(x); }
__device__ __forceinline__ double m_log10(double x) { return log10_cw(x); }

template <class R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <class R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }

// ------------------------------------------------------------------ MRG32k3a
// L'Ecuyer's recurrence (random/MRG.hpp:35-45, random/src/MRG32k3a.cpp:123-163)
// in FP64 arithmetic; the state is held as 32-bit integers (half the registers
// of doubles; each use converts exactly). a*s < 2^53 is exact; q = floor(p *
// (1/m)) is exact up to one unit, which the two fix-ups absorb, so the residue
// is the exact integer residue and the output (one FP64 multiply by norm) is
// bit-identical to the reference's host branch.
constexpr double kM1 = 4294967087.0, kM2 = 4294944443.0;
constexpr double kInvM1 = 1.0 / 4294967087.0, kInvM2 = 1.0 / 4294944443.0;
constexpr double kA12 = 1403580.0, kA13n = 810728.0, kA21 = 527612.0, kA23n = 1370589.0;
constexpr double kNorm = 1.0 / (4294967087.0 + 1.0);

struct Rng {
  uint32_t s[6];
  __device__ __forceinline__ void load(const uint32_t *plane, int64_t cap, int64_t i) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = plane[k * cap + i];
  }
  __device__ __forceinline__ void store(uint32_t *plane, int64_t cap, int64_t i) const {
#pragma unroll
    for (int k = 0; k < 6; ++k) plane[k * cap + i] = s[k];
  }
  __device__ __forceinline__ double uniform() {
    double p1 = kA12 * (double)s[1] - kA13n * (double)s[0];
    double k1 = floor(p1 * kInvM1);
    p1 -= k1 * kM1;
    if (p1 < 0.0) p1 += kM1;
    if (p1 >= kM1) p1 -= kM1;
    s[0] = s[1];
    s[1] = s[2];
    s[2] = (uint32_t)p1;
    double p2 = kA21 * (double)s[5] - kA23n * (double)s[3];
    double k2 = floor(p2 * kInvM2);
    p2 -= k2 * kM2;
    if (p2 < 0.0) p2 += kM2;
    if (p2 >= kM2) p2 -= kM2;
    s[3] = s[4];
    s[4] = s[5];
    s[5] = (uint32_t)p2;
    return (p1 > p2) ? (p1 - p2) * kNorm : (p1 - p2 + kM1) * kNorm;
  }
  template <class R> __device__ __forceinline__ R u() { return (R)uniform(); }
};

// a * b mod m for a, b < m < 2^32, exactly: the 64-bit product wraps
// correctly, the FP64 quotient estimate (below 2^32) is within one of the
// true quotient, and one correction each way restores the residue (no 64-bit
// division; the operands stay 32-bit, so each conversion is one instruction).
__device__ __forceinline__ uint32_t mulmod32(uint32_t a, uint32_t b, uint32_t m, double invm) {
  const uint64_t p = (uint64_t)a * b;
  const uint32_t q = (uint32_t)floor((double)a * (double)b * invm);
  int64_t r = (int64_t)(p - (uint64_t)q * m);
  if (r < 0) r += (int64_t)m;
  if (r >= (int64_t)m) r -= (int64_t)m;
  return (uint32_t)r;
}

// (M v) mod m for 32-bit residues (M row-major 3x3).
__device__ __forceinline__ void matvec_mod(const uint32_t M[9], uint32_t v[3], uint32_t m, double invm) {
  uint32_t r[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    uint64_t acc = (uint64_t)mulmod32(M[i * 3], v[0], m, invm) + mulmod32(M[i * 3 + 1], v[1], m, invm) +
                   mulmod32(M[i * 3 + 2], v[2], m, invm);
    if (acc >= m) acc -= m;
    if (acc >= m) acc -= m;
    r[i] = (uint32_t)acc;
  }
  v[0] = r[0];
  v[1] = r[1];
  v[2] = r[2];
}

// One jump-table row: three 16-byte loads of a 48-byte aligned row.
__device__ __forceinline__ void jump_row(const uint32_t *T, int c, int k, int d, uint32_t M[9]) {
  const uint4 *row = reinterpret_cast<const uint4 *>(T + ((c * 8 + k) * 256 + d) * kJumpRow);
  const uint4 a = row[0], b = row[1], e = row[2];
  M[0] = a.x;
  M[1] = a.y;
  M[2] = a.z;
  M[3] = a.w;
  M[4] = b.x;
  M[5] = b.y;
  M[6] = b.z;
  M[7] = b.w;
  M[8] = e.x;
}

// Substream idx of the stream seeded with s: (A^(2^76))^idx s, as the
// product over the 8 bytes d_k of idx of A^(d_k 2^(76+8k)) (powers of one
// matrix commute).  T[((c * 8 + k) * 256 + d) * kJumpRow + i], i < 9, holds
// A_c^(d 2^(76+8k)) mod m_c, the identity for d = 0.  The loop is unrolled
// so the rows' loads, which depend only on idx, issue ahead of the products.
__device__ __forceinline__ void jump_substream(const uint32_t *T, const uint32_t s[6], uint64_t idx,
                                               uint32_t v1[3], uint32_t v2[3]) {
  v1[0] = s[0];
  v1[1] = s[1];
  v1[2] = s[2];
  v2[0] = s[3];
  v2[1] = s[4];
  v2[2] = s[5];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int d = (int)((idx >> (8 * k)) & 255u);
    uint32_t M1[9], M2[9];
    jump_row(T, 0, k, d, M1);
    jump_row(T, 1, k, d, M2);
    matvec_mod(M1, v1, 4294967087u, 1.0 / 4294967087.0);
    matvec_mod(M2, v2, 4294944443u, 1.0 / 4294944443.0);
  }
}

__device__ __forceinline__ void substream_state(const DevParams *P, uint64_t idx, uint32_t out[6]) {
  uint32_t v1[3], v2[3];
  jump_substream(P->jump8, P->seed, idx, v1, v2);
  out[0] = v1[0];
  out[1] = v1[1];
  out[2] = v1[2];
  out[3] = v2[0];
  out[4] = v2[1];
  out[5] = v2[2];
}

// Substream index of a secondary: two draws of the creating track, 32 bits each.
__device__ __forceinline__ uint64_t child_substream(Rng &r) {
  const uint64_t hi = (uint64_t)(r.uniform() * 4294967296.0);
  const uint64_t lo = (uint64_t)(r.uniform() * 4294967296.0);
  return (hi << 32) | lo;
}

// ------------------------------------------------------------------ tables
template <class R> struct NodeType;
template <> struct NodeType<double> {
  using T = Node64;
  __device__ static const Node64 *global(const DevParams *P) { return P->nd; }
};
template <> struct NodeType<float> {
  using T = Node32;
  __device__ static const Node32 *global(const DevParams *P) { return P->nf; }
};
// Where a kernel reads physics-vector nodes: the node array in device memory
// (DevTable::off), or with kLds its LDS copy (DevTable::loff).
template <class R, bool kLds = false>
struct NodesT {
  using T = typename NodeType<R>::T;
  const T *v;
  const int *hints;
  __device__ NodesT(const DevParams *P) : v(NodeType<R>::global(P)), hints(P->hints) {}
  __device__ NodesT(const T *lds, const int *h) : v(lds), hints(h) {}
  __device__ __forceinline__ const T &at(const DevTable &t, int i) const { return v[(kLds ? t.loff : t.off) + i]; }
};
template <class R> using Nodes = NodesT<R, false>;

template <class R> __device__ __forceinline__ R tdl(const DevTable &t);
template <> __device__ __forceinline__ double tdl<double>(const DevTable &t) { return t.dl; }
template <> __device__ __forceinline__ float tdl<float>(const DevTable &t) { return t.dlf; }
template <class R> __device__ __forceinline__ R tbase(const DevTable &t);
template <> __device__ __forceinline__ double tbase<double>(const DevTable &t) { return t.base; }
template <> __device__ __forceinline__ float tbase<float>(const DevTable &t) { return t.basef; }
// first and last node of a vector: (x0, y0), (xn, yn)
template <class R> __device__ __forceinline__ void tends(const DevTable &t, R &x0, R &y0, R &xn, R &yn);
template <> __device__ __forceinline__ void tends<double>(const DevTable &t, double &x0, double &y0, double &xn,
                                                          double &yn) {
  x0 = t.x0;
  y0 = t.y0;
  xn = t.xn;
  yn = t.yn;
}
template <> __device__ __forceinline__ void tends<float>(const DevTable &t, float &x0, float &y0, float &xn,
                                                         float &yn) {
  x0 = t.x0f;
  y0 = t.y0f;
  xn = t.xnf;
  yn = t.ynf;
}

// Cubic spline between two adjacent nodes (ProxyPhysicsVector.cu:396-429).
template <class R, class V>
__device__ __forceinline__ R spline(const V &lo, const V &hi, R e) {
  const R x1 = lo.x, x2 = hi.x;
  const R delta = x2 - x1;
  const R a = (x2 - e) / delta;
  const R b = (e - x1) / delta;
  return a * lo.y + b * hi.y + ((a * a * a - a) * lo.d2 + (b * b * b - b) * hi.d2) * delta * delta / R(6.0);
}

// ProxyPhysicsVector::Value on a log-binned vector (ProxyPhysicsVector.cu:388-394,469-483).
template <class R, class NS>
__device__ __forceinline__ R pvec_value(const NS &N, const DevTable &t, R e, R le) {
  const int n = t.n;
  R x0, y0, xn, yn;
  tends<R>(t, x0, y0, xn, yn);
  if (e <= x0) return y0;
  if (e >= xn) return yn;
  int bin = (int)(le / tdl<R>(t) - tbase<R>(t));
  if (bin > n - 2) bin = n - 2;
  if (bin < 0) bin = 0;
  auto lo = N.at(t, bin);
  auto hi = N.at(t, bin + 1);
  if (bin > 0 && e < lo.x) {
    hi = lo;
    lo = N.at(t, bin - 1);
  }
  return spline<R>(lo, hi, e);
}

// The same with log(e) computed here (le: m_log(e), which a step's lookups
// at one energy share).
template <class R, class NS>
__device__ __forceinline__ R pvec_value(const NS &N, const DevTable &t, R e) {
  R x0, y0, xn, yn;
  tends<R>(t, x0, y0, xn, yn);
  if (e <= x0) return y0;
  if (e >= xn) return yn;
  return pvec_value(N, t, e, m_log(e));
}

// InvRangeValue (ProxyPhysicsVector.cu:491-524): binary search (the reference calls the
// log-bin Value() on this non-log vector).
template <class R, class NS>
__device__ __forceinline__ R pvec_inv_value(const NS &N, const DevTable &t, R r) {
  const int n = t.n;
  R x0, y0, xn, yn;
  tends<R>(t, x0, y0, xn, yn);
  if (r <= x0) return y0;
  if (r >= xn) return yn;
  // The binary search ends at the last node below r, j with x[j] < r <= x[j+1].
  // The same node is reached from the hint of the log2 bin one below r's
  // (the FP32 log2 is off by far less than a bin), by stepping up while the
  // next node is below r: about one dependent load instead of log2(n).
  int k = (int)floorf((__log2f((float)r) - t.hlog0) * t.hscale) - 1;
  k = k < 0 ? 0 : (k >= t.nhint ? t.nhint - 1 : k);
  int j = N.hints[t.hint + k];
  while (j > 0 && !(N.at(t, j).x < r)) --j;
  while (j < n - 2 && N.at(t, j + 1).x < r) ++j;
  return spline<R>(N.at(t, j), N.at(t, j + 1), r);
}

// Seltzer-Berger grid: x[32], y[57], v[57][32] at offset sb.
template <class R> __device__ __forceinline__ const R *sb_base(const DevParams *P, int off);
template <> __device__ __forceinline__ const double *sb_base<double>(const DevParams *P, int off) { return P->sbd + off; }
template <> __device__ __forceinline__ const float *sb_base<float>(const DevParams *P, int off) { return P->sbf + off; }

// Bin search of Proxy2DVector.cpp:93-139.
template <class R>
__device__ __forceinline__ int sb_bin(const R *a, int n, R z) {
  if (z < a[1]) return 0;
  if (z >= a[n - 2]) return n - 2;
  int lo = 0, hi = n - 2;
  while (lo <= hi) {
    const int mid = (lo + hi) / 2;
    if (z < a[mid])
      hi = mid - 1;
    else
      lo = mid + 1;
  }
  return hi;
}

template <class R> __device__ __forceinline__ const R *sb_xgrid(const DevParams *P);
template <> __device__ __forceinline__ const double *sb_xgrid<double>(const DevParams *P) { return P->sbx; }
template <> __device__ __forceinline__ const float *sb_xgrid<float>(const DevParams *P) { return P->sbxf; }
template <class R> __device__ __forceinline__ const R *sb_ygrid(const DevParams *P);
template <> __device__ __forceinline__ const double *sb_ygrid<double>(const DevParams *P) { return P->sby; }
template <> __device__ __forceinline__ const float *sb_ygrid<float>(const DevParams *P) { return P->sbyf; }

// Proxy2DVector::Value (Proxy2DVector.cpp:37-67), split in two: the part that
// depends only on y (its bin and the two value rows), which a rejection loop
// over x at fixed electron energy computes once, and the lookup at x.  The
// grid comes from DevParams (an LDS copy in the step kernels); the values
// from the element's table g (x[32], y[57], v[57][32]).
template <class R>
struct SbRow {
  const R *v1, *v2;  // value rows iy and iy + 1
  R y, y1, y2;
};

template <class R>
__device__ __forceinline__ SbRow<R> sb_row(const DevParams *P, const R *g, R y) {
  const R *gy = sb_ygrid<R>(P);
  if (y < gy[0]) y = gy[0];
  else if (y > gy[56]) y = gy[56];
  const int iy = sb_bin(gy, 57, y);
  SbRow<R> r;
  r.v1 = g + 32 + 57 + iy * 32;
  r.v2 = r.v1 + 32;
  r.y = y;
  r.y1 = gy[iy];
  r.y2 = gy[iy + 1];
  return r;
}

template <class R>
__device__ __forceinline__ R sb_at(const DevParams *P, const SbRow<R> &r, R x) {
  const R *gx = sb_xgrid<R>(P);
  if (x < gx[0]) x = gx[0];
  else if (x > gx[31]) x = gx[31];
  const int ix = sb_bin(gx, 32, x);
  const R x1 = gx[ix], x2 = gx[ix + 1], y = r.y, y1 = r.y1, y2 = r.y2;
  const R v11 = r.v1[ix], v12 = r.v1[ix + 1];
  const R v21 = r.v2[ix], v22 = r.v2[ix + 1];
  return ((y2 - y) * (v11 * (x2 - x) + v12 * (x - x1)) + ((y - y1) * (v21 * (x2 - x) + v22 * (x - x1)))) /
         ((x2 - x1) * (y2 - y1));
}

template <class R>
__device__ __forceinline__ R sb_value(const DevParams *P, const R *g, R x, R y) {
  return sb_at(P, sb_row(P, g, y), x);
}

// Merged Sandia interval of material m at ek (keV): j with B[j-1] <= ek <
// B[j] over its breakpoints B (DevMaterial::sx_*).  The search starts from
// the hint of the log2 bin one below ek's (the FP32 log2 is off by far less
// than a bin) and steps to the interval: one or two dependent loads, where a
// walk over each element's rows took up to its number of intervals.
__device__ __forceinline__ int sandia_interval(const DevParams *P, const DevMaterial &m, double ek) {
  const double *B = P->sx + m.sx_b;
  const int nb = m.sx_n;
  int k = (int)floorf((__log2f((float)ek) - m.sx_hlog0) * m.sx_hscale) - 1;
  k = k < 0 ? 0 : (k >= m.sx_nhint ? m.sx_nhint - 1 : k);
  int j = P->hints[m.sx_hint + k];
  while (j > 0 && B[j - 1] > ek) --j;
  while (j < nb && B[j] <= ek) ++j;
  return j;
}

// Sandia photo-absorption per atom (mm2) of an element with coefficients c
// (a1..a4 of its row at ek, in keV).
__device__ __forceinline__ double sandia_atom(const DevElement &el, const double *c, double ek) {
  const double inv = 1.0 / ek;
  const double s = inv * (c[0] + inv * (c[1] + inv * (c[2] + inv * c[3])));
  return (el.Z / el.zoa) / kAvogadro * s * 100.0;
}

// Coefficients of the elements of m at ek: element i's at the result + 4 i.
__device__ __forceinline__ const double *sandia_coefs(const DevParams *P, const DevMaterial &m, double ek) {
  return P->sx + m.sx_c + sandia_interval(P, m, ek) * m.nel * 4;
}

__device__ __forceinline__ double sandia_macro(const DevParams *P, const DevMaterial &m, double ekin) {
  const double ek = ekin * 1.0e3;
  const double *c = sandia_coefs(P, m, ek);
  double s = 0;
  for (int i = 0; i < m.nel; ++i) s += m.el[i].natoms * sandia_atom(m.el[i], c + 4 * i, ek);
  return s;
}

// Heitler two-photon annihilation cross-section per electron (mm2).
template <class R>
__device__ __forceinline__ R heitler_xs(R ekin) {
  const R tau = ekin / R(kMe);
  const R gam = tau + R(1.0);
  const R gamma2 = gam * gam;
  const R bg2 = tau * (tau + R(2.0));
  const R bg = m_sqrt(bg2);
  return R(kPi * kClassicRe * kClassicRe) *
         ((gamma2 + R(4.0) * gam + R(1.0)) * m_log(gam + bg) - (gam + R(3.0)) * bg) / (bg2 * (gam + R(1.0)));
}

// Macroscopic cross-section (1/mm) of discrete process p of a species
// (ProxyEmProcess::GetLambda and the processes' GetLambda) with the track's
// couple, Compton above 1 MeV and photoelectric from LambdaPrim/E.
template <class R, class NS>
__device__ __forceinline__ R macro_xs(const DevParams *P, const NS &N, const DevMaterial &m, int species, int p, R e,
                                      R le) {
  if (species == 0) {
    if (p == 0) {
      if ((double)e < m.compt_split) return pvec_value(N, m.tab[GXT_TAB_LAMBDA_COMPT], e, le);
      return pvec_value(N, m.tab[GXT_TAB_LAMBDAPRIM_COMPT], e, le) / e;
    }
    if (p == 1) return pvec_value(N, m.tab[GXT_TAB_LAMBDA_CONV], e, le);
    if ((double)e >= m.phot_emin) return pvec_value(N, m.tab[GXT_TAB_LAMBDAPRIM_PHOT], e, le) / e;
    return (R)sandia_macro(P, m, (double)e);
  }
  const bool pos = species == 2;
  if (p == 0) return pvec_value(N, m.tab[pos ? GXT_TAB_LAMBDA_BREM_EP : GXT_TAB_LAMBDA_BREM_EM], e, le);
  if (p == 1) return pvec_value(N, m.tab[pos ? GXT_TAB_LAMBDA_IONI_EP : GXT_TAB_LAMBDA_IONI_EM], e, le);
  return R(m.ne) * heitler_xs(e);
}
template <class R, class NS>
__device__ __forceinline__ R macro_xs(const DevParams *P, const NS &N, const DevMaterial &m, int species, int p, R e) {
  return macro_xs(P, N, m, species, p, e, m_log(e));
}

__device__ __forceinline__ int nproc(int species) { return species == 1 ? 2 : 3; }

// ------------------------------------------------------------------ vectors
template <class R>
__device__ __forceinline__ void unit3(R &a, R &b, R &c) {
  const R m2 = a * a + b * b + c * c;
  if (m2 > R(0)) {
    const R inv = R(1) / m_sqrt(m2);
    a *= inv;
    b *= inv;
    c *= inv;
  }
}

// Math::RotateToLabFrame (core/math_wrappers.hpp:194-212).
template <class R>
__device__ __forceinline__ void rotate_to_lab(R &u, R &v, R &w, R u1, R u2, R u3) {
  R up = u1 * u1 + u2 * u2;
  if (up > R(0)) {
    up = m_sqrt(up);
    const R px = u, py = v, pz = w;
    u = (u3 * px * u1 - u2 * py) * (R(1) / up) + u1 * pz;
    v = (u3 * px * u2 + u1 * py) * (R(1) / up) + u2 * pz;
    w = -up * px + u3 * pz;
  } else if (u3 < R(0)) {
    u = -u;
    w = -w;
  }
}

// ------------------------------------------------------------------ samplers
template <class R>
struct Sec {
  int species;
  R e;
  R d[3];
};

template <class R>
struct Interaction {
  R ekin;  // primary after the interaction (0 = killed)
  R d[3];
  R edep;
  int nsec;
  Sec<R> sec[2];
};

template <class R>
__device__ __forceinline__ void ia_init(Interaction<R> &o, R e, const R d[3], R ekin_out) {
  o.nsec = 0;
  o.edep = R(0);
  o.ekin = ekin_out;
  o.d[0] = d[0];
  o.d[1] = d[1];
  o.d[2] = d[2];
  (void)e;
}

// Element choice for compounds (the reference hard-codes Z): kind 0 = n Z(Z+1), kind 1 = partial Sandia.
template <class R>
__device__ __forceinline__ int select_element(const DevParams *P, const DevMaterial &m, int kind, R e, Rng &rng) {
  if (m.nel == 1) return 0;
  double w[kMaxElem], tot = 0;
  const double ek = (double)e * 1.0e3;
  const double *c = kind == 0 ? nullptr : sandia_coefs(P, m, ek);
  for (int i = 0; i < m.nel; ++i) {
    w[i] = kind == 0 ? m.el[i].wbrem : m.el[i].natoms * sandia_atom(m.el[i], c + 4 * i, ek);
    tot += w[i];
  }
  const double r = rng.uniform() * tot;
  double acc = 0;
  for (int i = 0; i < m.nel - 1; ++i) {
    acc += w[i];
    if (r < acc) return i;
  }
  return m.nel - 1;
}

// Klein-Nishina (proxy/ProxyKleinNishina.hpp:96-166).
template <class R>
__device__ __forceinline__ void sample_compton(R e0, const R d0[3], Rng &rng, Interaction<R> &o) {
  const R lowE = R(100.0e-6);
  ia_init(o, e0, d0, e0);
  if (e0 < lowE) return;
  const R E0_m = e0 / R(kMe);
  const R eps0 = R(1) / (R(1) + R(2) * E0_m);
  const R eps0sq = eps0 * eps0;
  // 1 - eps0^2 and 1 - eps without cancellation (eps -> 1 is forward
  // scattering): 1 - eps0 = 2 E0_m eps0; 1 - eps = -expm1(ln eps) on the
  // first branch, (1 - eps^2) / (1 + eps) on the second
  const R omeps0sq = R(2) * E0_m * eps0 * (R(1) + eps0);
  const R alpha1 = -m_log(eps0);
  const R alpha2 = R(0.5) * omeps0sq;
  R eps, omeps, epssq, onecost, sint2, greject;
  do {
    if (alpha1 / (alpha1 + alpha2) > rng.template u<R>()) {
      const R le = -alpha1 * rng.template u<R>();
      eps = m_exp(le);
      omeps = -m_expm1(le);
      epssq = eps * eps;
    } else {
      const R u = rng.template u<R>();
      epssq = eps0sq + omeps0sq * u;
      eps = m_sqrt(epssq);
      omeps = omeps0sq * (R(1) - u) / (R(1) + eps);
    }
    onecost = omeps / (eps * E0_m);
    sint2 = onecost * (R(2) - onecost);
    greject = R(1) - eps * sint2 / (R(1) + epssq);
  } while (greject < rng.template u<R>());
  const R sinT = sint2 < R(0) ? R(0) : m_sqrt(sint2);
  const R phi = R(kTwoPi) * rng.template u<R>();
  const R e1 = eps * e0;
  R g0 = sinT * m_cos(phi), g1 = sinT * m_sin(phi), g2 = R(1) - onecost;
  rotate_to_lab(g0, g1, g2, d0[0], d0[1], d0[2]);
  const R te = e0 - e1;
  if (e1 > lowE) {
    o.ekin = e1;
    o.d[0] = g0;
    o.d[1] = g1;
    o.d[2] = g2;
    Sec<R> &el = o.sec[0];
    o.nsec = 1;
    el.species = 1;
    el.e = te;
    el.d[0] = e0 * d0[0] - e1 * g0;
    el.d[1] = e0 * d0[1] - e1 * g1;
    el.d[2] = e0 * d0[2] - e1 * g2;
    unit3(el.d[0], el.d[1], el.d[2]);
  } else {
    o.ekin = R(0);
    o.edep = e1 + te;
  }
}

template <class R> __device__ __forceinline__ R screen1(R x) {
  return x > R(1) ? R(42.24) - R(8.368) * m_log(x + R(0.952)) : R(42.392) - x * (R(7.796) - R(1.961) * x);
}
template <class R> __device__ __forceinline__ R screen2(R x) {
  return x > R(1) ? R(42.24) - R(8.368) * m_log(x + R(0.952)) : R(41.405) - x * (R(5.828) - R(0.8945) * x);
}
template <class R> __device__ __forceinline__ R coulomb_factor(R Z) {
  const R k1 = R(0.0083), k2 = R(0.20206), k3 = R(0.0020), k4 = R(0.0369);
  const R az1 = R(kFineStructure) * Z;
  const R az2 = az1 * az1;
  const R az4 = az2 * az2;
  return (k1 * az4 + k2 + R(1) / (R(1) + az2)) * az2 - (k3 * az4 + k4) * az4;
}

// Bethe-Heitler (proxy/ProxyBetheHeitler.hpp:64-182): element from the material, Tsai
// angle used as an angle.
template <class R>
__device__ __forceinline__ void sample_pair(const DevParams *P, const DevMaterial &m, R e0, const R d0[3], Rng &rng,
                                   Interaction<R> &o) {
  ia_init(o, e0, d0, e0);
  const R eps0 = R(kMe) / e0;
  if (eps0 > R(0.5)) return;
  const int ie = select_element(P, m, 0, e0, rng);
  const R Z = (R)m.el[ie].Z;
  R eps;
  if (e0 < R(2.0)) {
    eps = eps0 + (R(0.5) - eps0) * rng.template u<R>();
  } else {
    const R logZ3 = m_log(Z) / R(3.0);
    R FZ = R(8.0) * logZ3;
    if (e0 > R(50.0)) FZ += R(8.0) * coulomb_factor(Z);
    const R Z3 = m_pow(Z, R(1.0) / R(3.0));
    const R screenfac = R(136.0) * eps0 / Z3;
    const R screenmax = m_exp((R(42.24) - FZ) / R(8.368)) - R(0.952);
    const R screenmin = rmin(R(4.0) * screenfac, screenmax);
    const R eps1 = R(0.5) - R(0.5) * m_sqrt(R(1) - screenmin / screenmax);
    const R epsmin = rmax(eps0, eps1);
    const R epsrange = R(0.5) - epsmin;
    const R F10 = screen1(screenmin) - FZ;
    const R F20 = screen2(screenmin) - FZ;
    const R NormF1 = rmax(F10 * epsrange * epsrange, R(0));
    const R NormF2 = rmax(R(1.5) * F20, R(0));
    R greject;
    do {
      if (NormF1 / (NormF1 + NormF2) > rng.template u<R>()) {
        eps = R(0.5) - epsrange * m_pow(rng.template u<R>(), R(0.333333));
        const R sv = screenfac / (eps * (R(1) - eps));
        greject = (screen1(sv) - FZ) / F10;
      } else {
        eps = epsmin + epsrange * rng.template u<R>();
        const R sv = screenfac / (eps * (R(1) - eps));
        greject = (screen2(sv) - FZ) / F20;
      }
    } while (greject < rng.template u<R>());
  }
  const R eTot = (rng.template u<R>() > R(0.5)) ? (R(1) - eps) * e0 : eps * e0;
  const R pTot = e0 - eTot;
  const R r1 = rng.template u<R>();
  const R r2 = rng.template u<R>();
  const R aa0 = -m_log(r1 * r2);
  const R u = (R(0.25) > rng.template u<R>()) ? aa0 / R(0.625) : aa0 / R(1.875);
  const R phi = R(kTwoPi) * rng.template u<R>();
  const R cosp = m_cos(phi), sinp = m_sin(phi);
  const R the = u * R(kMe) / eTot, thp = u * R(kMe) / pTot;
  const R ste = m_sin(the), cte = m_cos(the);
  const R stp = m_sin(thp), ctp = m_cos(thp);
  Sec<R> &el = o.sec[0];
  el.species = 1;
  el.e = eTot - R(kMe);
  el.d[0] = ste * cosp;
  el.d[1] = ste * sinp;
  el.d[2] = cte;
  rotate_to_lab(el.d[0], el.d[1], el.d[2], d0[0], d0[1], d0[2]);
  Sec<R> &po = o.sec[1];
  po.species = 2;
  po.e = pTot - R(kMe);
  po.d[0] = -stp * cosp;
  po.d[1] = -stp * sinp;
  po.d[2] = ctp;
  rotate_to_lab(po.d[0], po.d[1], po.d[2], d0[0], d0[1], d0[2]);
  o.nsec = 2;
  o.ekin = R(0);
}

// Sauter-Gavrila (proxy/ProxySauterGavrila.hpp:66-130): element by partial Sandia
// cross-section, first shell bound below E, binding energy deposited.
template <class R>
__device__ __forceinline__ void sample_photo(const DevParams *P, const DevMaterial &m, R e0, const R d0[3], Rng &rng,
                                    Interaction<R> &o) {
  ia_init(o, e0, d0, R(0));
  const int ie = select_element(P, m, 1, e0, rng);
  const DevElement &el = m.el[ie];
  R bind = R(0);
  for (int i = 0; i < el.nshell; ++i) {
    const R b = (R)P->shell_e[el.shell + i];
    if (e0 >= b) {
      bind = b;
      break;
    }
  }
  const R te = e0 - bind;
  const R tau = e0 / R(kMe);
  R cost;
  if (tau > R(50.0)) {
    cost = R(1);
  } else {
    const R gamma = tau + R(1);
    const R beta = m_sqrt(tau * (tau + R(2))) / gamma;
    const R A = (R(1) - beta) / beta;
    const R Ap2 = A + R(2);
    const R B = R(0.5) * beta * gamma * (gamma - R(1)) * (gamma - R(2));
    const R grej = R(2) * (R(1) + A * B) / A;
    R z, g;
    do {
      const R q = rng.template u<R>();
      z = R(2) * A * (R(2) * q + Ap2 * m_sqrt(q)) / (Ap2 * Ap2 - R(4) * q);
      g = (R(2) - z) * (R(1) / (A + z) + B);
    } while (g < rng.template u<R>() * grej);
    cost = R(1) - z;
  }
  const R sint = m_sqrt((R(1) + cost) * (R(1) - cost));
  const R phi = R(kTwoPi) * rng.template u<R>();
  Sec<R> &e = o.sec[0];
  e.species = 1;
  e.e = te;
  e.d[0] = sint * m_cos(phi);
  e.d[1] = sint * m_sin(phi);
  e.d[2] = cost;
  rotate_to_lab(e.d[0], e.d[1], e.d[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  o.edep = bind;
}

// Seltzer-Berger (proxy/ProxySeltzerBerger.hpp:204-315): element from the material, its
// own grid, Migdal factor, gamma cut of the couple, electron kinetic energy.
template <class R>
__device__ __forceinline__ void sample_brems(const DevParams *P, const DevMaterial &m, R t0, const R d0[3], Rng &rng,
                                    Interaction<R> &o) {
  ia_init(o, t0, d0, t0);
  const R cut = rmin((R)m.cut_gamma, t0);
  const R emax = t0;
  if (cut >= emax) return;
  const int ie = select_element(P, m, 0, t0, rng);
  const R *sb = sb_base<R>(P, m.el[ie].sb);
  const R etot = t0 + R(kMe);
  const R dcorr = (R)m.migdal * etot * etot;
  const R xmin = m_log(cut * cut + dcorr);
  const R xmax = m_log(emax * emax + dcorr);
  const R y = m_log(t0);
  const R x0 = cut / t0;
  const SbRow<R> row = sb_row(P, sb, y);
  R vmax = sb_at(P, row, x0) * R(1.02);
  const R epeak = R(300.0), elow = R(10.0e-3);
  if (x0 < R(0.97) && (t0 > epeak || t0 < elow)) {
    const R ylim = rmin(sb_value(P, sb, R(0.97), R(4.0) * m_log(R(10.0))), R(1.1) * sb_at(P, row, R(0.97)));
    if (ylim > vmax) vmax = ylim;
  }
  if (x0 < R(0.05)) vmax *= R(1.2);
  R k, v;
  do {
    const R r = rng.template u<R>();
    R x = m_exp(xmin + r * (xmax - xmin)) - dcorr;
    if (x < R(0)) x = R(0);
    k = m_sqrt(x);
    v = sb_at(P, row, k / t0);
  } while (v < vmax * rng.template u<R>());
  const R u0 = rng.template u<R>();
  const R c = R(4.0) - R(8.0) * u0;
  const R signc = c < R(0) ? R(-1) : R(1);
  const R a = m_fabs(c);
  const R delta = R(0.5) * (m_sqrt(R(4.0) + a * a) + a);
  // cos = cofA - 1/cofA with cofA = -sign(c) delta^(1/3) is -sign(c) x, where
  // x = d - 1/d (d = delta^(1/3)) solves x^3 + 3x = a; so 1 - x =
  // (4 - a) / (4 + x + x^2), and 4 - a = 8 min(u0, 1 - u0): neither 1 - cos
  // nor 1 + cos cancels
  const R d = m_pow(delta, R(1.0) / R(3.0));
  const R x = d - R(1) / d;
  const R omx = R(8) * rmin(u0, R(1) - u0) / (R(4) + x + x * x);
  const R opx = R(1) + x;
  const R cost0 = -signc * x;
  const R omc0 = signc > R(0) ? opx : omx;
  const R opc0 = signc > R(0) ? omx : opx;
  // the boost to the lab, (cos + beta) / (1 + cos beta), in forms without
  // cancellation near the forward direction: 1 - beta =
  // 1 / (gam (gam + sqrt(tau (tau + 2)))) and 1 - cos' = (1 - cos)(1 - beta) /
  // (1 + cos beta); beta itself rounds to 1 in FP32 above about 1 GeV
  const R tau = t0 / R(kMe);
  const R gam = tau + R(1);
  const R bgam = m_sqrt(tau * (tau + R(2)));
  const R omb = R(1) / (gam * (gam + bgam));
  const R omc = omc0 * omb / (opc0 - cost0 * omb);
  const R cost = R(1) - omc;
  const R sint = m_sqrt(rmax(omc * (R(2) - omc), R(0)));
  const R phi = R(kTwoPi) * rng.template u<R>();
  Sec<R> &g = o.sec[0];
  g.species = 0;
  g.e = k;
  g.d[0] = sint * m_cos(phi);
  g.d[1] = sint * m_sin(phi);
  g.d[2] = cost;
  rotate_to_lab(g.d[0], g.d[1], g.d[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  const R ptot = m_sqrt(t0 * (etot + R(kMe)));
  o.ekin = t0 - k;
  o.d[0] = ptot * d0[0] - k * g.d[0];
  o.d[1] = ptot * d0[1] - k * g.d[1];
  o.d[2] = ptot * d0[2] - k * g.d[2];
  unit3(o.d[0], o.d[1], o.d[2]);
}

// Moller / Bhabha (proxy/ProxyMollerScattering.hpp:92-159, electron cut of the couple;
// the Bhabha branch from Geant4's G4MollerBhabhaModel, absent in the reference).
template <class R>
__device__ __forceinline__ void sample_ioni(const DevMaterial &m, bool positron, R t0, const R d0[3], Rng &rng,
                                   Interaction<R> &o) {
  ia_init(o, t0, d0, t0);
  const R tmin = (R)m.cut_e;
  const R tmax = positron ? t0 : R(0.5) * t0;
  if (tmin >= tmax) return;
  const R energy = t0 + R(kMe);
  const R xmin = tmin / t0, xmax = tmax / t0;
  const R gam = energy / R(kMe);
  const R gamma2 = gam * gam;
  R x, z, grej;
  if (!positron) {
    const R gg = (R(2) * gam - R(1)) / gamma2;
    R y = R(1) - xmax;
    grej = R(1) - gg * xmax + xmax * xmax * (R(1) - gg + (R(1) - gg * y) / (y * y));
    do {
      const R r = rng.template u<R>();
      x = xmin * xmax / (xmin * (R(1) - r) + xmax * r);
      y = R(1) - x;
      z = R(1) - gg * x + x * x * (R(1) - gg + (R(1) - gg * y) / (y * y));
    } while (grej * rng.template u<R>() > z);
  } else {
    const R beta2 = R(1) - R(1) / gamma2;
    const R yb = R(1) / (R(1) + gam);
    const R y2 = yb * yb;
    const R y12 = R(1) - R(2) * yb;
    const R b1 = R(2) - y2;
    const R b2 = y12 * (R(3) + y2);
    const R y122 = y12 * y12;
    const R b4 = y122 * y12;
    const R b3 = b4 + y122;
    const R xm2 = xmax * xmax;
    grej = R(1) + (xm2 * xm2 * b4 - xmin * xmin * xmin * b3 + xm2 * b2 - xmin * b1) * beta2;
    do {
      const R r = rng.template u<R>();
      x = xmin * xmax / (xmin * (R(1) - r) + xmax * r);
      const R x2 = x * x;
      z = R(1) + (x2 * x2 * b4 - x * x2 * b3 + x2 * b2 - x * b1) * beta2;
    } while (grej * rng.template u<R>() > z);
  }
  const R dT = x * t0;
  const R ptot = m_sqrt(t0 * (t0 + R(2) * R(kMe)));
  const R pd = m_sqrt(dT * (dT + R(2) * R(kMe)));
  R cost = dT * (energy + R(kMe)) / (pd * ptot);
  if (cost > R(1)) cost = R(1);
  const R sint = m_sqrt((R(1) - cost) * (R(1) + cost));
  const R phi = R(kTwoPi) * rng.template u<R>();
  Sec<R> &d = o.sec[0];
  d.species = 1;
  d.e = dT;
  d.d[0] = sint * m_cos(phi);
  d.d[1] = sint * m_sin(phi);
  d.d[2] = cost;
  rotate_to_lab(d.d[0], d.d[1], d.d[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  o.ekin = t0 - dT;
  o.d[0] = ptot * d0[0] - pd * d.d[0];
  o.d[1] = ptot * d0[1] - pd * d.d[1];
  o.d[2] = ptot * d0[2] - pd * d.d[2];
  unit3(o.d[0], o.d[1], o.d[2]);
}

// e+ annihilation in flight (Heitler; Geant4's G4eeToTwoGammaModel, absent in
// the reference).
template <class R>
__device__ __forceinline__ void sample_annihil(R t0, const R d0[3], Rng &rng, Interaction<R> &o) {
  ia_init(o, t0, d0, R(0));
  const R tau = t0 / R(kMe);
  const R gam = tau + R(1);
  const R tau2 = tau + R(2);
  const R sqgrate = m_sqrt(tau / tau2) * R(0.5);
  const R sqg2m1 = m_sqrt(tau * tau2);
  // 1/2 - sqgrate without cancellation at high energy: (1 - s) / 2 with
  // s = sqrt(tau / tau2) is 1 / (tau2 (1 + s))
  const R epsmin = R(1) / (tau2 * (R(1) + R(2) * sqgrate));
  const R epsmax = R(0.5) + sqgrate;
  const R epsqot = epsmax / epsmin;
  R eps, greject;
  do {
    eps = epsmin * m_exp(m_log(epsqot) * rng.template u<R>());
    greject = R(1) - eps + (R(2) * gam * eps - R(1)) / (eps * tau2 * tau2);
  } while (greject < rng.template u<R>());
  R cost = (eps * tau2 - R(1)) / (eps * sqg2m1);
  if (cost > R(1)) cost = R(1);
  if (cost < R(-1)) cost = R(-1);
  const R sint = m_sqrt((R(1) + cost) * (R(1) - cost));
  const R phi = R(kTwoPi) * rng.template u<R>();
  const R etot = t0 + R(2) * R(kMe);
  Sec<R> &g1 = o.sec[0];
  g1.species = 0;
  g1.e = eps * etot;
  g1.d[0] = sint * m_cos(phi);
  g1.d[1] = sint * m_sin(phi);
  g1.d[2] = cost;
  rotate_to_lab(g1.d[0], g1.d[1], g1.d[2], d0[0], d0[1], d0[2]);
  Sec<R> &g2 = o.sec[1];
  g2.species = 0;
  g2.e = (R(1) - eps) * etot;
  const R pp = m_sqrt(t0 * (t0 + R(2) * R(kMe)));
  g2.d[0] = pp * d0[0] - g1.e * g1.d[0];
  g2.d[1] = pp * d0[1] - g1.e * g1.d[1];
  g2.d[2] = pp * d0[2] - g1.e * g1.d[2];
  unit3(g2.d[0], g2.d[1], g2.d[2]);
  o.nsec = 2;
}

template <class R>
__device__ __forceinline__ void sample_model(const DevParams *P, const DevMaterial &m, int model, R e, const R d[3], Rng &rng,
                                    Interaction<R> &o) {
  switch (model) {
    case GXT_MODEL_KLEIN_NISHINA: sample_compton(e, d, rng, o); break;
    case GXT_MODEL_BETHE_HEITLER: sample_pair(P, m, e, d, rng, o); break;
    case GXT_MODEL_SAUTER_GAVRILA: sample_photo(P, m, e, d, rng, o); break;
    case GXT_MODEL_SELTZER_BERGER_EM:
    case GXT_MODEL_SELTZER_BERGER_EP: sample_brems(P, m, e, d, rng, o); break;
    case GXT_MODEL_MOLLER: sample_ioni(m, false, e, d, rng, o); break;
    case GXT_MODEL_BHABHA: sample_ioni(m, true, e, d, rng, o); break;
    default: sample_annihil(e, d, rng, o); break;
  }
}

// Samplers of the photon queues (0 Compton, 1 conversion, 2 photoelectric).
template <class R>
__device__ __forceinline__ void sample_neutral(const DevParams *P, const DevMaterial &m, int q, R e, const R d[3],
                                               Rng &rng, Interaction<R> &o) {
  if (q == 0)
    sample_compton(e, d, rng, o);
  else if (q == 1)
    sample_pair(P, m, e, d, rng, o);
  else
    sample_photo(P, m, e, d, rng, o);
}

// Samplers of the e-/e+ queues (3, 5 bremsstrahlung, 4 Moller, 6 Bhabha, 7 annihilation).
template <class R>
__device__ __forceinline__ void sample_charged(const DevParams *P, const DevMaterial &m, int q, R e, const R d[3],
                                               Rng &rng, Interaction<R> &o) {
  if (q == 3 || q == 5)
    sample_brems(P, m, e, d, rng, o);
  else if (q == 4)
    sample_ioni(m, false, e, d, rng, o);
  else if (q == 6)
    sample_ioni(m, true, e, d, rng, o);
  else
    sample_annihil(e, d, rng, o);
}

// queue index -> species, process, model
__device__ __forceinline__ int queue_species(int q) { return q < 3 ? 0 : (q < 5 ? 1 : 2); }
__device__ __forceinline__ int queue_proc(int q) { return q < 3 ? q : (q < 5 ? q - 3 : q - 5); }
__device__ __forceinline__ int queue_of(int species, int p) { return species == 0 ? p : (species == 1 ? 3 + p : 5 + p); }
__device__ __forceinline__ int queue_model(int q) {
  // gamma: KN, BH, SG; e-: SB, Moller; e+: SB, Bhabha, annihilation
  return q == 0 ? GXT_MODEL_KLEIN_NISHINA
                : q == 1 ? GXT_MODEL_BETHE_HEITLER
                         : q == 2 ? GXT_MODEL_SAUTER_GAVRILA
                                  : q == 3 ? GXT_MODEL_SELTZER_BERGER_EM
                                           : q == 4 ? GXT_MODEL_MOLLER
                                                    : q == 5 ? GXT_MODEL_SELTZER_BERGER_EP
                                                             : q == 6 ? GXT_MODEL_BHABHA : GXT_MODEL_ANNIHIL;
}

// ------------------------------------------------------------------ continuous loss
// ProxyIonization::AlongStepGPIL (ProxyIonization.hpp:108-128): production-cut
// length for finR and Geant4's (2 - finR/range) form.
template <class R, class NS>
__device__ __forceinline__ R along_limit(const NS &N, const DevMaterial &m, bool pos, R e, R le, R mine, R &range0) {
  R range = pvec_value(N, m.tab[pos ? GXT_TAB_RANGE_EP : GXT_TAB_RANGE_EM], e, le);
  range0 = range;
  if (e < mine) range *= e / mine;
  const R finR = rmin(R(10.0), (R)m.cut_length);
  return range > finR ? range * R(0.2) + finR * R(0.8) * (R(2.0) - finR / range) : range;
}

// ProxyIonization::AlongStepDoIt (ProxyIonization.hpp:130-168): binary-search
// inverse range.
template <class R, class NS>
__device__ __forceinline__ R along_doit(const NS &N, const DevMaterial &m, bool pos, R e, R le, R range, R step,
                                        R mine, R &eloss) {
  if (step >= range || e <= mine) {
    eloss = e;
    return R(0);
  }
  const R dedx = pvec_value(N, m.tab[pos ? GXT_TAB_DEDX_EP : GXT_TAB_DEDX_EM], e, le);
  R loss = step * dedx;
  if (loss > e * R(0.01))
    loss = e - pvec_inv_value(N, m.tab[pos ? GXT_TAB_INVRANGE_EP : GXT_TAB_INVRANGE_EM], range - step);
  R fin = e - loss;
  if (fin < mine) {
    loss += fin;
    fin = R(0);
  }
  eloss = loss < R(0) ? R(0) : loss;
  return fin;
}

// ------------------------------------------------------------------ geometry
// Analytic slab/box navigator (replaces the VecGeom navigation of
// geometry/NavigationInterface.hpp:42-98).  Positions and distances are FP64.
__device__ __forceinline__ int locate(const DevGeometry &g, const double p[3], double d0, double d1, double d2) {
  const double ax = fabs(p[0]), ay = fabs(p[1]);
  if (ax > g.hx || ay > g.hy) return -1;
  if (ax == g.hx && p[0] * d0 > 0) return -1;
  if (ay == g.hy && p[1] * d1 > 0) return -1;
  const int nv = g.nvol;
  const double z = p[2];
  if (z < g.z[0] || z > g.z[nv]) return -1;
  if (z == g.z[0] && d2 < 0) return -1;
  if (z == g.z[nv] && d2 > 0) return -1;
  int lo = 0, hi = nv - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (g.z[mid] <= z)
      lo = mid;
    else
      hi = mid - 1;
  }
  if (z == g.z[lo] && lo > 0 && d2 < 0) --lo;
  return lo;
}

__device__ __forceinline__ double dist_out(const DevGeometry &g, int v, const double p[3], double d0, double d1,
                                           double d2, int &face) {
  double s = kHuge;
  face = -1;
  if (d2 > 0) {
    s = (g.z[v + 1] - p[2]) / d2;
    face = 1;
  } else if (d2 < 0) {
    s = (g.z[v] - p[2]) / d2;
    face = 0;
  }
  if (d0 != 0) {
    const double sx = ((d0 > 0 ? g.hx : -g.hx) - p[0]) / d0;
    if (sx < s) {
      s = sx;
      face = 2;
    }
  }
  if (d1 != 0) {
    const double sy = ((d1 > 0 ? g.hy : -g.hy) - p[1]) / d1;
    if (sy < s) {
      s = sy;
      face = 3;
    }
  }
  return s < 0 ? 0.0 : s;
}

__device__ __forceinline__ double safety(const DevGeometry &g, int v, const double p[3]) {
  double s = fmin(p[2] - g.z[v], g.z[v + 1] - p[2]);
  s = fmin(s, g.hx - fabs(p[0]));
  s = fmin(s, g.hy - fabs(p[1]));
  return s < 0 ? 0.0 : s;
}

// ------------------------------------------------------- volume-tree navigator
// Boxes placed by translations inside each other (VecGeom's UnplacedBox and
// PlaceDaughter as the reference builds its calorimeter,
// proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:225-325), with
// VecGeom's surface tolerance (NavigationInterface.hpp:66).  A point within
// kNavTol of a face belongs to the box only if its direction enters through
// that face.
constexpr double kNavTol = 1e-9;
constexpr int kNavMaxDepth = 32;
constexpr int kNavGridMin = 8;  // daughters from which a volume has a daughter index

__device__ __forceinline__ bool nav_contains(const NavNode &b, const double p[3], const double v[3]) {
  for (int k = 0; k < 3; ++k) {
    const double x = p[k] - b.c[k];
    const double a = fabs(x);
    if (a > b.h[k] + kNavTol) return false;
    if (a >= b.h[k] - kNavTol && x * v[k] >= 0.0) return false;
  }
  return true;
}

// Deepest touchable below t (inclusive) containing the point; t contains it.
// The first daughter (lowest index) holding the point is taken: with a
// daughter index only the candidates of the point's bin are tested, which
// include every daughter that can hold it, in the same order.
__device__ __forceinline__ int nav_descend(const DevNav &g, int t, const double p[3], const double v[3]) {
  for (int depth = 0; depth < kNavMaxDepth; ++depth) {
    const NavNode &m = g.node[t];
    int in = -1;
    if (m.count >= kNavGridMin && g.grid[m.volume].nb > 0) {
      const NavGrid &gr = g.grid[m.volume];
      const int b = nav_bin(p[gr.ax] - m.c[gr.ax], m.h[gr.ax], gr.inv, gr.nb);
      for (int k = g.idx[gr.off + b]; k < g.idx[gr.off + b + 1]; ++k) {
        const int d = m.first + g.idx[k];
        if (nav_contains(g.node[d], p, v)) {
          in = d;
          break;
        }
      }
    } else {
      for (int d = m.first; d < m.first + m.count; ++d) {
        if (nav_contains(g.node[d], p, v)) {
          in = d;
          break;
        }
      }
    }
    if (in < 0) break;
    t = in;
  }
  return t;
}

// GlobalLocator::LocateGlobalPoint: -1 outside the world (touchable 0).
__device__ __forceinline__ int nav_locate(const DevNav &g, const double p[3], const double v[3]) {
  if (!nav_contains(g.node[0], p, v)) return -1;
  return nav_descend(g, 0, p, v);
}

// Starting from touchable t, the touchable holding the point: up through the
// parents until one contains it, then down (the relocation after a boundary
// crossing).
__device__ __forceinline__ int nav_relocate(const DevNav &g, int t, const double p[3], const double v[3]) {
  for (int depth = 0; depth <= kNavMaxDepth && t >= 0; ++depth) {
    if (nav_contains(g.node[t], p, v)) return nav_descend(g, t, p, v);
    t = g.node[t].parent;
  }
  return -1;
}

// Distance along v from p to the box b from outside (the slab method):
// kHuge unless the ray passes more than kNavTol through b.
__device__ __forceinline__ double nav_dist_in(const NavNode &b, const double p[3], const double v[3]) {
  double tmin = -kHuge, tmax = kHuge;
  for (int k = 0; k < 3; ++k) {
    const double x = p[k] - b.c[k];
    if (v[k] == 0.0) {
      if (fabs(x) >= b.h[k] - kNavTol) return kHuge;
      continue;
    }
    const double t1 = (-b.h[k] - x) / v[k], t2 = (b.h[k] - x) / v[k];
    tmin = fmax(tmin, fmin(t1, t2));
    tmax = fmin(tmax, fmax(t1, t2));
  }
  const double t0 = fmax(tmin, 0.0);
  return tmax > t0 + kNavTol ? t0 : kHuge;
}

// Distance along v from p inside the box b to its surface.
__device__ __forceinline__ double nav_dist_out(const NavNode &b, const double p[3], const double v[3]) {
  double s = kHuge;
  for (int k = 0; k < 3; ++k) {
    const double x = p[k] - b.c[k];
    if (v[k] > 0.0)
      s = fmin(s, (b.h[k] - x) / v[k]);
    else if (v[k] < 0.0)
      s = fmin(s, (-b.h[k] - x) / v[k]);
  }
  return s < 0.0 ? 0.0 : s;
}

// Distance along v from p in touchable t to leaving t or entering one of its
// daughters, whichever is first (ties leave t, then take the lowest
// daughter); enter = that daughter, or -1.
__device__ __forceinline__ double nav_boundary(const DevNav &g, int t, const double p[3], const double v[3],
                                               int &enter) {
  const NavNode *node = g.node;
  const NavNode &m = node[t];
  double s = nav_dist_out(m, p, v);
  enter = -1;
  for (int d = m.first; d < m.first + m.count; ++d) {
    const double sd = nav_dist_in(node[d], p, v);
    if (sd < s) {
      s = sd;
      enter = d;
    }
  }
  return s;
}

// Isotropic safety of p in touchable t: the distance to t's faces and to the
// nearest daughter box.
__device__ __forceinline__ double nav_safety(const DevNav &g, int t, const double p[3]) {
  const NavNode *node = g.node;
  const NavNode &m = node[t];
  double sm = kHuge;
  for (int k = 0; k < 3; ++k) sm = fmin(sm, m.h[k] - fabs(p[k] - m.c[k]));
  double dmin = kHuge;
  for (int d = m.first; d < m.first + m.count; ++d) {
    const NavNode &b = node[d];
    double q = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double e = fabs(p[k] - b.c[k]) - b.h[k];
      if (e > 0.0) q += e * e;
    }
    dmin = fmin(dmin, q);
  }
  if (dmin < kHuge) sm = fmin(sm, sqrt(dmin));
  return sm < 0.0 ? 0.0 : sm;
}

// The boundary step from p in touchable t (FindNextBoundaryAndStep,
// NavigationInterface.hpp:64-98): nav_boundary's distance, the touchable
// holding p + s v, and the isotropic safety.
__device__ __forceinline__ int nav_step(const DevNav &g, int t, const double p[3], const double v[3],
                                        double &snext, double &safety) {
  int enter;
  const double s = nav_boundary(g, t, p, v, enter);
  safety = nav_safety(g, t, p);
  snext = s;
  const double q[3] = {p[0] + s * v[0], p[1] + s * v[1], p[2] + s * v[2]};
  return nav_relocate(g, enter >= 0 ? enter : t, q, v);
}

// NavigationInterface::FindNextBoundary (NavigationInterface.hpp:64-98) on
// top of nav_step: the boundary distance limited to the physics step pstep,
// the safety shortcut, the boundary flag and the 2 kNavTol push.
__device__ __forceinline__ int nav_find_next_boundary(const DevNav &g, int t, const double p[3], const double v[3],
                                                      double pstep, double &snext, double &safety, int &boundary) {
  double d;
  int next = nav_step(g, t, p, v, d, safety);
  if (!(d < pstep)) {
    d = pstep;
    next = t;
  }
  if (safety > pstep) {
    snext = pstep;
    boundary = 0;
    return t;
  }
  boundary = d < pstep ? 1 : 0;
  snext = fmax(2.0 * kNavTol, d + 2.0 * kNavTol);
  return next;
}

// ConstFieldHelixStepper::DoStep (magneticfield/ConstFieldHelixStepper.hpp:164-239)
// in mm/MeV/tesla; the direction is renormalised (FieldPropagationHandler.cpp:323-330).
// The arithmetic is in R; the position update is accumulated in FP64.
template <class R>
__device__ __forceinline__ void helix_step(const double B[3], double bmag, R charge, R mom, R step, double pos[3], R dir[3]) {
  const R ub[3] = {(R)(B[0] / bmag), (R)(B[1] / bmag), (R)(B[2] / bmag)};
  const R uvb = dir[0] * ub[0] + dir[1] * ub[1] + dir[2] * ub[2];
  R dt2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] - uvb * uvb;
  if (dt2 < R(0)) dt2 = R(0);
  const R sinvb = m_sqrt(dt2) + R(1.0e-30);
  const R Rad = mom * sinvb / (R(kB2C) * charge * (R)bmag);
  R vx[3] = {dir[0] - uvb * ub[0], dir[1] - uvb * ub[1], dir[2] - uvb * ub[2]};
  unit3(vx[0], vx[1], vx[2]);
  const R cvb[3] = {vx[1] * ub[2] - vx[2] * ub[1], vx[2] * ub[0] - vx[0] * ub[2], vx[0] * ub[1] - vx[1] * ub[0]};
  const R phi = -step * charge * (R)bmag * R(kB2C) / mom;
  const R sphi = m_sin(phi), cphi = m_cos(phi);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    pos[i] = pos[i] + (double)(Rad * (cphi - R(1)) * cvb[i]) - (double)(Rad * sphi * vx[i]) +
             (double)(step * uvb * ub[i]);
    dir[i] = uvb * ub[i] + cphi * sinvb * vx[i] + sphi * sinvb * cvb[i];
  }
  unit3(dir[0], dir[1], dir[2]);
}

enum { kPhysics = 0, kBoundary = 1, kEscaped = 2, kLimit = 3 };

// Propagation inside the current volume to pstep or the boundary
// (FieldPropagationHandler.cpp:87-360; LinearPropagationHandler.cpp:13-63;
// driver loop test_BasicCpuTransport.cpp:319-336).
template <class R>
__device__ __forceinline__ int propagate(const DevParams *P, int species, R e, double pos[3], R dir[3], int &vol, double pstep,
                                double &step) {
  const DevGeometry &g = P->geo;
  if (species == 0 || P->bmag == 0.0) {
    int face;
    const double d0 = dir[0], d1 = dir[1], d2 = dir[2];
    const double sn = dist_out(g, vol, pos, d0, d1, d2, face);
    if (pstep < sn) {
      pos[0] += pstep * d0;
      pos[1] += pstep * d1;
      pos[2] += pstep * d2;
      step = pstep;
      return kPhysics;
    }
    pos[0] += sn * d0;
    pos[1] += sn * d1;
    pos[2] += sn * d2;
    step = sn;
    if (face == 0 || face == 1) {
      const int nv = face == 1 ? vol + 1 : vol - 1;
      pos[2] = g.z[face == 1 ? vol + 1 : vol];
      if (nv < 0 || nv >= g.nvol) return kEscaped;
      vol = nv;
      return kBoundary;
    }
    return kEscaped;
  }
  const R q = species == 1 ? R(-1) : R(1);
  const R mom = m_sqrt(e * (e + R(2) * R(kMe)));
  const R ub[3] = {(R)(P->B[0] / P->bmag), (R)(P->B[1] / P->bmag), (R)(P->B[2] / P->bmag)};
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = safety(g, vol, pos);
    if (remaining <= saf) {
      helix_step<R>(P->B, P->bmag, q, mom, (R)remaining, pos, dir);
      step += remaining;
      return kPhysics;
    }
    int face;
    const double slin = dist_out(g, vol, pos, (double)dir[0], (double)dir[1], (double)dir[2], face);
    const R pl = mom * (dir[0] * ub[0] + dir[1] * ub[1] + dir[2] * ub[2]);
    R pt2 = mom * mom - pl * pl;
    if (pt2 < R(0)) pt2 = R(0);
    const R pt = m_sqrt(pt2);
    const double sl = pt > R(0) ? (double)(R(2) * m_sqrt(R(0.1) / (R(0.299792458) * (R)P->bmag / pt))) : kHuge;
    if (slin <= sl && slin <= remaining) {
      // aimed step: cross the aimed face by construction, on its plane
      const double s = fmax(slin, 1.0e-6);
      helix_step<R>(P->B, P->bmag, q, mom, (R)s, pos, dir);
      step += s;
      if (face == 0 || face == 1) {
        const int nv = face == 1 ? vol + 1 : vol - 1;
        pos[2] = g.z[face == 1 ? vol + 1 : vol];
        if (nv < 0 || nv >= g.nvol || fabs(pos[0]) > g.hx || fabs(pos[1]) > g.hy) return kEscaped;
        vol = nv;
        return kBoundary;
      }
      return kEscaped;
    }
    const double s = fmin(remaining, fmax(sl, saf));
    helix_step<R>(P->B, P->bmag, q, mom, (R)s, pos, dir);
    step += s;
    remaining -= s;
    const int v2 = locate(g, pos, (double)dir[0], (double)dir[1], (double)dir[2]);
    if (v2 != vol) {
      if (v2 < 0) return kEscaped;
      vol = v2;
      return kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

// Volume of a starting point (-1: outside the setup), in either form.
__device__ __forceinline__ int geo_locate(const DevGeometry &g, const double p[3], const double d[3]) {
  if (g.tree) return nav_locate(g.nav, p, d);
  return locate(g, p, d[0], d[1], d[2]);
}

// Tally cell of volume v (DevGeometry): v itself in the slab form, cell[v]
// for a touchable of the tree form.  kGeo 0: slab, 1: tree, 2: read g.tree.
template <int kGeo>
__device__ __forceinline__ int geo_cell(const DevGeometry &g, int v) {
  if constexpr (kGeo == 0)
    return v;
  else if constexpr (kGeo == 1)
    return g.cell[v];
  else
    return g.tree ? g.cell[v] : v;
}

// Propagation in the tree form: along the straight line to pstep or the next
// boundary (the touchable beyond it found by relocation from the daughter
// entered, or from the touchable left); in a uniform field the helix in
// sub-steps chosen as in propagate, from the navigator's safety and
// straight-line distance, a sub-step ending the step when the point it
// reaches lies in another touchable.  A track in vacuum passes pstep = kHuge.
template <class R>
__device__ __forceinline__ int propagate_tree(const DevParams *P, int species, R e, double pos[3], R dir[3], int &vol,
                                              double pstep, double &step) {
  const DevNav &g = P->geo.nav;
  if (species == 0 || P->bmag == 0.0) {
    const double v[3] = {(double)dir[0], (double)dir[1], (double)dir[2]};
    int enter;
    const double sn = nav_boundary(g, vol, pos, v, enter);
    if (pstep < sn) {
      pos[0] += pstep * v[0];
      pos[1] += pstep * v[1];
      pos[2] += pstep * v[2];
      step = pstep;
      return kPhysics;
    }
    pos[0] += sn * v[0];
    pos[1] += sn * v[1];
    pos[2] += sn * v[2];
    step = sn;
    vol = nav_relocate(g, enter >= 0 ? enter : vol, pos, v);
    return vol < 0 ? kEscaped : kBoundary;
  }
  const R q = species == 1 ? R(-1) : R(1);
  const R mom = m_sqrt(e * (e + R(2) * R(kMe)));
  const R ub[3] = {(R)(P->B[0] / P->bmag), (R)(P->B[1] / P->bmag), (R)(P->B[2] / P->bmag)};
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = nav_safety(g, vol, pos);
    if (remaining <= saf) {
      helix_step<R>(P->B, P->bmag, q, mom, (R)remaining, pos, dir);
      step += remaining;
      return kPhysics;
    }
    int enter;
    const double v[3] = {(double)dir[0], (double)dir[1], (double)dir[2]};
    const double slin = nav_boundary(g, vol, pos, v, enter);
    const R pl = mom * (dir[0] * ub[0] + dir[1] * ub[1] + dir[2] * ub[2]);
    R pt2 = mom * mom - pl * pl;
    if (pt2 < R(0)) pt2 = R(0);
    const R pt = m_sqrt(pt2);
    const double sl = pt > R(0) ? (double)(R(2) * m_sqrt(R(0.1) / (R(0.299792458) * (R)P->bmag / pt))) : kHuge;
    // aimed at the boundary when the straight line reaches it within the
    // sagitta-limited length
    const double s = (slin <= sl && slin <= remaining) ? fmax(slin, 1.0e-6) : fmin(remaining, fmax(sl, saf));
    helix_step<R>(P->B, P->bmag, q, mom, (R)s, pos, dir);
    step += s;
    remaining -= s;
    const double w[3] = {(double)dir[0], (double)dir[1], (double)dir[2]};
    const int v2 = nav_relocate(g, vol, pos, w);
    if (v2 != vol) {
      vol = v2;
      return v2 < 0 ? kEscaped : kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

// ------------------------------------------------------------------ multiple scattering
// e-/e+ multiple scattering on the first transport mean free path lambda1,
// tabulated as E^2 / lambda1 (the LambdaMod1.msc tables; the reference
// reserves the process, ProxyProcessIndex.cuh:27, but has no model):
//  - step limit max(0.04 max(range, lambda1), 0.6 safety, 1e-5 mm);
//  - geometric path z = lambda1 (1 - exp(-t/lambda1)) for a true path t
//    (Lewis' mean longitudinal displacement, loss over the step neglected),
//    and its inverse when geometry ends the step;
//  - deflection: x = 1 - cos(theta) from p(x) ~ 1/(x + a)^2 on [0, 2] with a
//    set so that <cos theta> = exp(-t/lambda1) (Lewis' first moment),
//    azimuth uniform; no lateral displacement.
constexpr double kMscFacRange = 0.04, kMscFacSafety = 0.6, kMscTmin = 1.0e-5;
constexpr double kMscNone = 1.0e30;  // lambda1 where a material has no msc table

template <class R, class NS>
__device__ __forceinline__ double msc_lambda1(const NS &N, const DevMaterial &m, bool pos, R e, R le) {
  const DevTable &t = m.tab[pos ? GXT_TAB_MSC_EP : GXT_TAB_MSC_EM];
  if (t.n == 0) return kMscNone;
  const R v = pvec_value(N, t, e, le);
  return v > R(0) ? (double)(e * e / v) : kMscNone;
}

// (expm1 and log1p keep full precision for short steps)
__device__ __forceinline__ double msc_geom(double t, double lam) { return -lam * expm1_cw(-t / lam); }

// true path of a geometric path z, at most tmax
__device__ __forceinline__ double msc_true(double z, double lam, double tmax) {
  const double x = z / lam;
  if (x >= 1.0) return tmax;
  return fmin(-lam * log1p_cw(-x), tmax);
}

// Screening of the deflection x = 1 - cos(theta) ~ 1/(x + a)^2 on [0, 2],
// as u = 2/a, for a mean 1 - cos(theta) of 1 - q, q = exp(-tau):
// 1 - q = mean(a) = a (a + 2) ln(1 + 2/a) / 2 - a, which rises from 0 (a -> 0)
// to 1 (a -> inf).  For q >= 0.05 (a below ~12): twelve Newton steps on ln a
// (d mean / d ln a = a ((a + 1) ln(1 + 2/a) - 2) > 0, steps at most e^2) from
// the small-a form a (ln(2/a) - 1) or the large-a form q = 2/(3a).  For
// q < 0.05 the direct form cancels; q = sum_{j>=1} 2 (-1)^(j+1) u^j /
// ((j + 1)(j + 2)) (from the series of ln(1 + u)) is solved for u < 0.17 by
// twelve Newton steps from u = 3q with 20 terms.
__device__ __attribute__((noinline)) double msc_u(double tau) {
  const double q = exp_cw(-tau);
  if (q >= 0.05) {
    const double mean = -expm1_cw(-tau);
    double a = mean < 0.5 ? mean / fmax(log_cw(2.0 / mean) - 1.0, 0.5) : 2.0 / (3.0 * q);
    for (int k = 0; k < 12; ++k) {
      const double L = log1p_cw(2.0 / a);
      const double f = 0.5 * a * (a + 2.0) * L - a - mean;
      const double df = a * ((a + 1.0) * L - 2.0);
      const double du = fmin(fmax(-f / df, -2.0), 2.0);
      a *= exp_cw(du);
    }
    return 2.0 / a;
  }
  double u = 3.0 * q;
  for (int k = 0; k < 12; ++k) {
    double h = 0.0, dh = 0.0;
    for (int j = 20; j >= 1; --j) {
      const double c = (j % 2 ? 2.0 : -2.0) / ((j + 1.0) * (j + 2.0));
      dh = dh * u + j * c;
      h = h * u + c;
    }
    h *= u;
    u = fmin(fmax(u - (h - q) / dh, 0.0), 0.2);
  }
  return u;
}

// Deflection after a true path of tau transport mean free paths (two draws):
// x = 2 r1 / (1 + u (1 - r1)) inverts the distribution's CDF.
template <class R>
__device__ __forceinline__ void msc_deflect(double tau, R dir[3], Rng &rng) {
  const double r1 = rng.uniform();
  const double r2 = rng.uniform();
  const double u = msc_u(tau);
  const double x = 2.0 * r1 / (1.0 + u * (1.0 - r1));
  const double cost = 1.0 - x;
  const double sint = sqrt(fmax(x * (2.0 - x), 0.0));
  const double phi = kTwoPi * r2;
  R uu = (R)(sint * m_cos(phi)), v = (R)(sint * m_sin(phi)), w = (R)cost;
  rotate_to_lab<R>(uu, v, w, dir[0], dir[1], dir[2]);
  dir[0] = uu;
  dir[1] = v;
  dir[2] = w;
  unit3(dir[0], dir[1], dir[2]);
}

// ------------------------------------------------------------------ general field
// Field of the cylindrical (r, z) map at (x, y, z) of the map frame
// (ScalarCMSmagField::GetFieldValue / GetFieldValueRZ,
// examples/magneticfield/simplifiedCMS/inc/Geant/ScalarCMSmagField.h:247-331):
// r and z clamped to the grid, bilinear weights from the position inside the
// cell, cylindrical components turned with the azimuth.  The cell index is
// clamped to the last cell so the upper grid edge reads its own nodes, and the
// axis (r = 0) takes azimuth 0.  The cell fractions are quotients, so nodes
// and cell-edge midpoints of a grid with representable spacing come out exact.
__device__ __forceinline__ void map_field(const DevField &f, double x, double y, double z, double B[3]) {
  const double r = sqrt(x * x + y * y);
  const double rmax = f.r0 + (double)(f.nr - 1) * f.dr;
  const double zmax = f.z0 + (double)(f.nz - 1) * f.dz;
  const double rc = fmin(fmax(r, f.r0), rmax);
  const double zc = fmin(fmax(z, f.z0), zmax);
  const double fr = (rc - f.r0) / f.dr;
  const double fz = (zc - f.z0) / f.dz;
  const int ir = min((int)fr, f.nr - 2);
  const int iz = min((int)fz, f.nz - 2);
  const double tr = fr - (double)ir, tz = fz - (double)iz;
  const double w00 = (1.0 - tr) * (1.0 - tz), w01 = (1.0 - tr) * tz;
  const double w10 = tr * (1.0 - tz), w11 = tr * tz;
  const double *p = f.v + ((int64_t)ir * f.nz + iz) * 3;
  const double *q = p + (int64_t)f.nz * 3;
  double b[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) b[c] = w00 * p[c] + w01 * p[3 + c] + w10 * q[c] + w11 * q[3 + c];
  double cs = 1.0, sn = 0.0;
  if (r > 0.0) {
    cs = x / r;
    sn = y / r;
  }
  B[0] = b[0] * cs - b[1] * sn;
  B[1] = b[0] * sn + b[1] * cs;
  B[2] = b[2];
}

// Field (tesla) at a point of the setup frame (FieldLookup::GetFieldValue,
// FieldPropagationHandler.hpp:27-39): the map at pos + origin, or uniform.
__device__ __forceinline__ void field_at(const DevField &f, const double pos[3], double B[3]) {
  if (f.v) {
    map_field(f, pos[0] + f.origin[0], pos[1] + f.origin[1], pos[2] + f.origin[2], B);
  } else {
    B[0] = f.b[0];
    B[1] = f.b[1];
    B[2] = f.b[2];
  }
}

// Cash-Karp embedded Runge-Kutta 4(5) tableau (J. R. Cash, A. H. Karp, ACM
// Trans. Math. Softw. 16 (1990) 201-222); the fifth-order solution is
// propagated, fifth minus fourth order is the error estimate.
namespace ck {
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 3.0 / 10.0, a42 = -9.0 / 10.0, a43 = 6.0 / 5.0;
constexpr double a51 = -11.0 / 54.0, a52 = 5.0 / 2.0, a53 = -70.0 / 27.0, a54 = 35.0 / 27.0;
constexpr double a61 = 1631.0 / 55296.0, a62 = 175.0 / 512.0, a63 = 575.0 / 13824.0, a64 = 44275.0 / 110592.0,
                 a65 = 253.0 / 4096.0;
constexpr double b1 = 37.0 / 378.0, b3 = 250.0 / 621.0, b4 = 125.0 / 594.0, b6 = 512.0 / 1771.0;
constexpr double e1 = b1 - 2825.0 / 27648.0, e3 = b3 - 18575.0 / 48384.0, e4 = b4 - 13525.0 / 55296.0,
                 e5 = -277.0 / 14336.0, e6 = b6 - 1.0 / 4.0;
}  // namespace ck

// Stage derivative of the Lorentz equation in arc length: f = kq (u x B(x)),
// kq = -q kB2C / p (1/(tesla mm)).
template <class R>
__device__ __forceinline__ void lorentz(const DevField &F, R kq, const double x[3], const R u[3], R f[3]) {
  double B[3];
  field_at(F, x, B);
  const R b0 = (R)B[0], b1 = (R)B[1], b2 = (R)B[2];
  f[0] = kq * (u[1] * b2 - u[2] * b1);
  f[1] = kq * (u[2] * b0 - u[0] * b2);
  f[2] = kq * (u[0] * b1 - u[1] * b0);
}

// One Cash-Karp step of length h from (pos, dir); f1 is the derivative at the
// start.  Adds |error estimate| of position and direction to ex, eu.
template <class R>
__device__ __forceinline__ void ck45_step(const DevField &F, R kq, double h, const R f1[3], double pos[3], R dir[3],
                                          double &ex, double &eu) {
  const R hr = (R)h;
  double x[3];
  R u2[3], u3[3], u4[3], u5[3], u6[3];
  R f2[3], f3[3], f4[3], f5[3], f6[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c] + h * (ck::a21 * (double)dir[c]);
    u2[c] = dir[c] + hr * (R(ck::a21) * f1[c]);
  }
  lorentz<R>(F, kq, x, u2, f2);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c] + h * (ck::a31 * (double)dir[c] + ck::a32 * (double)u2[c]);
    u3[c] = dir[c] + hr * (R(ck::a31) * f1[c] + R(ck::a32) * f2[c]);
  }
  lorentz<R>(F, kq, x, u3, f3);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c] + h * (ck::a41 * (double)dir[c] + ck::a42 * (double)u2[c] + ck::a43 * (double)u3[c]);
    u4[c] = dir[c] + hr * (R(ck::a41) * f1[c] + R(ck::a42) * f2[c] + R(ck::a43) * f3[c]);
  }
  lorentz<R>(F, kq, x, u4, f4);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c] + h * (ck::a51 * (double)dir[c] + ck::a52 * (double)u2[c] + ck::a53 * (double)u3[c] +
                         ck::a54 * (double)u4[c]);
    u5[c] = dir[c] + hr * (R(ck::a51) * f1[c] + R(ck::a52) * f2[c] + R(ck::a53) * f3[c] + R(ck::a54) * f4[c]);
  }
  lorentz<R>(F, kq, x, u5, f5);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c] + h * (ck::a61 * (double)dir[c] + ck::a62 * (double)u2[c] + ck::a63 * (double)u3[c] +
                         ck::a64 * (double)u4[c] + ck::a65 * (double)u5[c]);
    u6[c] = dir[c] + hr * (R(ck::a61) * f1[c] + R(ck::a62) * f2[c] + R(ck::a63) * f3[c] + R(ck::a64) * f4[c] +
                           R(ck::a65) * f5[c]);
  }
  lorentz<R>(F, kq, x, u6, f6);
  double ex2 = 0.0, eu2 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double dx = h * (ck::e1 * (double)dir[c] + ck::e3 * (double)u3[c] + ck::e4 * (double)u4[c] +
                           ck::e5 * (double)u5[c] + ck::e6 * (double)u6[c]);
    const R du = hr * (R(ck::e1) * f1[c] + R(ck::e3) * f3[c] + R(ck::e4) * f4[c] + R(ck::e5) * f5[c] +
                       R(ck::e6) * f6[c]);
    ex2 += dx * dx;
    eu2 += (double)(du * du);
    pos[c] = pos[c] + h * (ck::b1 * (double)dir[c] + ck::b3 * (double)u3[c] + ck::b4 * (double)u4[c] +
                           ck::b6 * (double)u6[c]);
    dir[c] = dir[c] + hr * (R(ck::b1) * f1[c] + R(ck::b3) * f3[c] + R(ck::b4) * f4[c] + R(ck::b6) * f6[c]);
  }
  unit3(dir[0], dir[1], dir[2]);
  ex += sqrt(ex2);
  eu += sqrt(eu2);
}

// Sub-steps of the RK45 stepper within one propagation sub-step of the
// transport, and within one call of the stepper at most.
constexpr int kRkMaxSub = 16;
constexpr int kRkMaxN = 65537;

// Advances a charged track by s with the stepper (kStep, or F.stepper when
// kStep is -1): a helix in the start-point field B (|B| = bm; straight line
// where it vanishes), or ceil(s k bm / GXT_RK_ANGLE) equal Cash-Karp steps,
// k = 0.299792458 / p.
template <class R, int kStep = -1>
__device__ __forceinline__ void field_advance(const DevField &F, const double B[3], double bm, R q, R mom, double s,
                                              double pos[3], R dir[3], double &ex, double &eu) {
  const bool rk = kStep < 0 ? F.stepper == GXT_STEPPER_RK45 : kStep == GXT_STEPPER_RK45;
  if (rk) {
    // sub-steps: at most kRkMaxN, whatever the inputs (an infinite turn included)
    const double turn = s * (-kB2C) * bm / (double)mom;
    const int n = turn > GXT_RK_ANGLE ? (int)fmin(ceil(turn / GXT_RK_ANGLE), (double)kRkMaxN) : 1;
    const double h = s / (double)n;
    const R kq = -q * R(kB2C) / mom;
    R f1[3];
    const R u0[3] = {dir[0], dir[1], dir[2]};
    {
      const R b0 = (R)B[0], b1 = (R)B[1], b2 = (R)B[2];
      f1[0] = kq * (u0[1] * b2 - u0[2] * b1);
      f1[1] = kq * (u0[2] * b0 - u0[0] * b2);
      f1[2] = kq * (u0[0] * b1 - u0[1] * b0);
    }
    for (int k = 0; k < n; ++k) {
      if (k > 0) lorentz<R>(F, kq, pos, dir, f1);
      ck45_step<R>(F, kq, h, f1, pos, dir, ex, eu);
    }
  } else if (bm > 0.0) {
    helix_step<R>(B, bm, q, mom, (R)s, pos, dir);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[c] += s * (double)dir[c];
  }
}

// Propagation of an e-/e+ in a general field (a map, or the RK45 stepper):
// the same sub-step choice as propagate() with the field looked up at the
// start of every sub-step (FieldPropagationHandler::Propagate,
// FieldPropagationHandler.cpp:93-111: step min(pstep or snext, max(SafeLength
// in the local field, safety))).  RK45 sub-steps are also kept within
// kRkMaxSub Cash-Karp steps.  kStep: the stepper (GXT_STEPPER_*), fixed at
// compile time so each kernel carries one.
template <class R, int kStep>
__device__ __forceinline__ int propagate_general(const DevParams *P, int species, R e, double pos[3], R dir[3],
                                                 int &vol, double pstep, double &step) {
  const DevGeometry &g = P->geo;
  const DevField &F = P->field;
  const R q = species == 1 ? R(-1) : R(1);
  const R mom = m_sqrt(e * (e + R(2) * R(kMe)));
  constexpr bool rk = kStep == GXT_STEPPER_RK45;
  double ex = 0.0, eu = 0.0;
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = safety(g, vol, pos);
    double B[3];
    field_at(F, pos, B);
    const double bm = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    const double smax = rk && bm > 0.0 ? kRkMaxSub * GXT_RK_ANGLE * (double)mom / ((-kB2C) * bm) : kHuge;
    if (remaining <= saf && remaining <= smax) {
      field_advance<R, kStep>(F, B, bm, q, mom, remaining, pos, dir, ex, eu);
      step += remaining;
      return kPhysics;
    }
    int face;
    const double slin = dist_out(g, vol, pos, (double)dir[0], (double)dir[1], (double)dir[2], face);
    double sl = kHuge;
    if (bm > 0.0) {
      const R pl = mom * (dir[0] * (R)(B[0] / bm) + dir[1] * (R)(B[1] / bm) + dir[2] * (R)(B[2] / bm));
      R pt2 = mom * mom - pl * pl;
      if (pt2 < R(0)) pt2 = R(0);
      const R pt = m_sqrt(pt2);
      if (pt > R(0)) sl = (double)(R(2) * m_sqrt(R(0.1) / (R(0.299792458) * (R)bm / pt)));
    }
    if (slin <= sl && slin <= remaining && slin <= smax) {
      // aimed step: cross the aimed face by construction, on its plane
      const double s = fmax(slin, 1.0e-6);
      field_advance<R, kStep>(F, B, bm, q, mom, s, pos, dir, ex, eu);
      step += s;
      if (face == 0 || face == 1) {
        const int nv = face == 1 ? vol + 1 : vol - 1;
        pos[2] = g.z[face == 1 ? vol + 1 : vol];
        if (nv < 0 || nv >= g.nvol || fabs(pos[0]) > g.hx || fabs(pos[1]) > g.hy) return kEscaped;
        vol = nv;
        return kBoundary;
      }
      return kEscaped;
    }
    const double s = fmin(fmin(remaining, fmax(sl, saf)), smax);
    field_advance<R, kStep>(F, B, bm, q, mom, s, pos, dir, ex, eu);
    step += s;
    remaining -= s;
    const int v2 = locate(g, pos, (double)dir[0], (double)dir[1], (double)dir[2]);
    if (v2 != vol) {
      if (v2 < 0) return kEscaped;
      vol = v2;
      return kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

}  // namespace gxt
