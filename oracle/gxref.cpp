// gxref.cpp -- CPU oracle of the EM particle-transport step loop (FP64, scalar).
//
// TEST INFRASTRUCTURE ONLY: the checker for libgxt.so and the reported CPU
// baseline.  It restates, function by function, the reference's algorithms
// (GeantExascalePilot; file:line citations below are relative to
// source/Geant/) under the physics contract shared with the GPU path.  Each
// deviation from the reference is a fix of a hard-coded value or defect of
// the reference, stated in the comment of the function that makes it.
//
// Parity status of each part (how it is pinned) is listed in the repository's
// design notes; the spline second-derivative routine and the samplers' exact
// draw-by-draw outputs have no reference-run golden vectors (the reference
// cannot be built here) and are pinned by closed forms, invariants and
// distribution tests instead.
#include "gxref.h"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace gxref {

// ---------------------------------------------------------------- constants
// CLHEP values as in proxy/ProxyPhysicalConstants.hpp:23-78 (mm, MeV).
constexpr double kMe = 0.510998910;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kClassicRe = 2.8179402894e-12;      // mm
constexpr double kComptonLength = 3.8615926459e-10;  // mm (hbar c / m_e c^2)
constexpr double kAvogadro = 6.02214179e+23;
constexpr double kFineStructure = 1.0 / 137.0;        // as ProxyEmModel.hpp:117
constexpr double kB2C = -0.299792458;                 // MeV/c per (tesla mm)
constexpr double kHuge = 1.0e300;

static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

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
static inline double log_cw(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int e;
  double m = std::frexp(x, &e);
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
static inline double exp_cw(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double k = std::rint(x * kInvLn2);
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
  return std::ldexp(p, (int)k);
}
static inline double log1p_cw(double x) {
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  if (!(u > 0.0) || u == HUGE_VAL) return log_cw(u);
  return log_cw(u) - ((u - 1.0) - x) / u;
}
static inline double expm1_cw(double x) {
  if (!(std::fabs(x) < 0.34657359027997264)) return exp_cw(x) - 1.0;
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
static inline double pow_cw(double x, double y) { return exp_cw(y * log_cw(x)); }
static inline double log10_cw(double x) { return log_cw(x) * kInvLn10; }

// ---------------------------------------------------------------- MRG32k3a
// L'Ecuyer's combined MRG (random/MRG.hpp:35-45, random/src/MRG32k3a.cpp:123-163).
// The reference carries the residues in FP64 and reduces with floor(p/m)
// (host branch :132,146); that quotient is always exact, so the exact 64-bit
// integer residue below yields the same state words and, through one FP64
// multiply by norm, the same output bits.
constexpr int64_t kM1 = 4294967087LL, kM2 = 4294944443LL;
constexpr int64_t kA12 = 1403580, kA13n = 810728, kA21 = 527612, kA23n = 1370589;
constexpr double kNorm = 1.0 / (4294967087.0 + 1.0);

static inline double mrg_uniform(uint32_t s[6]) {
  int64_t p1 = (kA12 * (int64_t)s[1] - kA13n * (int64_t)s[0]) % kM1;
  if (p1 < 0) p1 += kM1;
  s[0] = s[1];
  s[1] = s[2];
  s[2] = (uint32_t)p1;
  int64_t p2 = (kA21 * (int64_t)s[5] - kA23n * (int64_t)s[3]) % kM2;
  if (p2 < 0) p2 += kM2;
  s[3] = s[4];
  s[4] = s[5];
  s[5] = (uint32_t)p2;
  return (p1 > p2) ? (double)(p1 - p2) * kNorm : (double)(p1 - p2 + kM1) * kNorm;
}

// Jump matrices P[c][j] = A_c^(2^(76+j)) mod m_c (the reference's A1p76/A2p76
// at j = 0 and A1p127/A2p127 at j = 51, random/MRG.hpp:50-69), built by
// repeated squaring of the one-step matrices A1p0/A2p0 (MRG.hpp:71-75), as
// MatTwoPowModM does (random/src/MRG32k3a.cpp:171-313).
struct JumpTables {
  uint64_t P[2][64][9];
  JumpTables() {
    const uint64_t m[2] = {(uint64_t)kM1, (uint64_t)kM2};
    uint64_t A[2][9] = {{0, 1, 0, 0, 0, 1, (uint64_t)(kM1 - kA13n), (uint64_t)kA12, 0},
                        {0, 1, 0, 0, 0, 1, (uint64_t)(kM2 - kA23n), 0, (uint64_t)kA21}};
    for (int c = 0; c < 2; ++c) {
      uint64_t M[9];
      std::memcpy(M, A[c], sizeof(M));
      for (int k = 0; k < 76 + 63; ++k) {
        if (k >= 76) std::memcpy(P[c][k - 76], M, sizeof(M));
        uint64_t R[9];
        matmul(M, M, R, m[c]);
        std::memcpy(M, R, sizeof(M));
      }
      std::memcpy(P[c][63], M, sizeof(M));
    }
  }
  static void matmul(const uint64_t *a, const uint64_t *b, uint64_t *r, uint64_t m) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        uint64_t acc = 0;
        for (int k = 0; k < 3; ++k) acc = (acc + (a[i * 3 + k] * b[k * 3 + j]) % m) % m;
        r[i * 3 + j] = acc;
      }
  }
};
static const JumpTables &jumps() {
  static JumpTables t;
  return t;
}

static inline void matvec(const uint64_t *M, uint32_t *v, uint64_t m) {
  uint64_t r[3];
  for (int i = 0; i < 3; ++i) {
    uint64_t acc = 0;
    for (int k = 0; k < 3; ++k) acc += (M[i * 3 + k] * (uint64_t)v[k]) % m;
    r[i] = acc % m;
  }
  v[0] = (uint32_t)r[0];
  v[1] = (uint32_t)r[1];
  v[2] = (uint32_t)r[2];
}

// State of substream `idx` of the stream seeded with seed[6]: (A^(2^76))^idx
// applied to the seed, i.e. idx applications of SetNextSubstream
// (random/src/MRG32k3a.cpp:61-67) done by binary powering.
static void substream_state(const uint32_t seed[6], uint64_t idx, uint32_t out[6]) {
  const JumpTables &J = jumps();
  for (int i = 0; i < 6; ++i) out[i] = seed[i];
  for (int j = 0; j < 64; ++j)
    if ((idx >> j) & 1ULL) {
      matvec(J.P[0][j], out, (uint64_t)kM1);
      matvec(J.P[1][j], out + 3, (uint64_t)kM2);
    }
}

// Substream index of a secondary: two draws from the creating track's stream,
// 32 bits each (the upper half of [0,1) resolution of one MRG32k3a output).
static uint64_t child_substream(uint32_t s[6]) {
  uint64_t hi = (uint64_t)(mrg_uniform(s) * 4294967296.0);
  uint64_t lo = (uint64_t)(mrg_uniform(s) * 4294967296.0);
  return (hi << 32) | lo;
}

// ---------------------------------------------------------------- tables
// One physics vector (proxy/ProxyPhysicsVector.cuh:21-110).
struct PVec {
  int type = 0;
  std::vector<double> x, y, d2;
  double dl = 0, base = 0;
};

// ASCII format of data/table/*.asc (ProxyPhysicsTable.cu:41-88,
// ProxyPhysicsVector.cu:89-128): nvec, then per vector: type, "emin emax n",
// size, size pairs.
static bool read_asc(const std::string &path, std::vector<PVec> &out) {
  std::ifstream f(path);
  if (!f) return false;
  int nvec = 0;
  f >> nvec;
  if (!f || nvec <= 0) return false;
  out.resize(nvec);
  for (int v = 0; v < nvec; ++v) {
    PVec &p = out[v];
    double emin, emax;
    int nn, siz;
    f >> p.type >> emin >> emax >> nn >> siz;
    if (!f || siz < 5) return false;
    p.x.resize(siz);
    p.y.resize(siz);
    for (int i = 0; i < siz; ++i) f >> p.x[i] >> p.y[i];
    if (!f) return false;
  }
  return true;
}

// Second derivatives with the reference's not-a-knot elimination
// (ProxyPhysicsVector.cu:308-386, after Kvasov 2000).  The steps are kept
// as the reference has them, including the use of d[n-3] in the last forward
// row, because the tables are interpolated with these derivatives.
static void fill_d2(PVec &p) {
  const std::vector<double> &x = p.x, &y = p.y;
  const int n = (int)x.size() - 1;
  std::vector<double> d(n + 1, 0.0), u(n, 0.0);
  u[1] = ((y[2] - y[1]) / (x[2] - x[1]) - (y[1] - y[0]) / (x[1] - x[0]));
  u[1] = 6.0 * u[1] * (x[2] - x[1]) / ((x[2] - x[0]) * (x[2] - x[0]));
  d[1] = (2.0 * x[1] - x[0] - x[2]) / (2.0 * x[2] - x[0] - x[1]);
  for (int i = 2; i < n - 1; ++i) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double q = sig * d[i - 1] + 2.0;
    d[i] = (sig - 1.0) / q;
    double t = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * t / (x[i + 1] - x[i - 1])) - sig * u[i - 1] / q;
  }
  double sig = (x[n - 1] - x[n - 2]) / (x[n] - x[n - 2]);
  double q = sig * d[n - 3] + 2.0;
  double t = (y[n] - y[n - 1]) / (x[n] - x[n - 1]) - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  u[n - 1] = 6.0 * sig * t / (x[n] - x[n - 2]) - (2.0 * sig - 1.0) * u[n - 2] / q;
  q = (1.0 + sig) + (2.0 * sig - 1.0) * d[n - 2];
  d[n - 1] = u[n - 1] / q;
  for (int k = n - 2; k > 1; --k) d[k] *= (d[k + 1] - u[k] * (x[k + 1] - x[k - 1]) / (x[k + 1] - x[k]));
  d[n] = (d[n - 1] - (1.0 - sig) * d[n - 2]) / sig;
  sig = 1.0 - ((x[2] - x[1]) / (x[2] - x[0]));
  d[1] *= (d[2] - u[1] / (1.0 - sig));
  d[0] = (d[1] - sig * d[2]) / (1.0 - sig);
  p.d2 = d;
  // log-binning parameters (ProxyPhysicsVector.cu:124-125)
  p.dl = std::log(x[1] / x[0]);
  p.base = std::log(x[0]) / p.dl;
}

static inline double spline(const PVec &p, double e, int bin) {
  // ProxyPhysicsVector.cu:396-429
  const double x1 = p.x[bin], x2 = p.x[bin + 1];
  const double delta = x2 - x1;
  const double a = (x2 - e) / delta;
  const double b = (e - x1) / delta;
  return a * p.y[bin] + b * p.y[bin + 1] +
         ((a * a * a - a) * p.d2[bin] + (b * b * b - b) * p.d2[bin + 1]) * delta * delta / 6.0;
}

// Value() for log-binned vectors (ProxyPhysicsVector.cu:388-394,469-483):
// clamped outside [x0, xN]; bin from log(E)/dl - base, one step back if E
// lies below the node.  The bin is also clamped to [0, N-2] (the reference
// relies on E < xN to stay in range).
static double pvec_value(const PVec &p, double e) {
  const int N = (int)p.x.size();
  if (N == 0) return 0.0;
  if (e <= p.x[0]) return p.y[0];
  if (e >= p.x[N - 1]) return p.y[N - 1];
  int bin = (int)(log_cw(e) / p.dl - p.base);
  if (bin > N - 2) bin = N - 2;
  if (bin < 0) bin = 0;
  if (bin > 0 && e < p.x[bin]) --bin;
  return spline(p, e, bin);
}

// InvRangeValue (ProxyPhysicsVector.cu:491-524): binary search bin location
// for the non-log inverse-range vector (the reference calls the log-bin
// Value() on it, ProxyIonization.hpp:100-104).
static double pvec_inv_value(const PVec &p, double r) {
  const int N = (int)p.x.size();
  if (N == 0) return 0.0;
  if (r <= p.x[0]) return p.y[0];
  if (r >= p.x[N - 1]) return p.y[N - 1];
  int n1 = 0, n2 = N / 2, n3 = N - 1;
  while (n1 != n3 - 1) {
    if (r > p.x[n2])
      n1 = n2;
    else
      n3 = n2;
    n2 = n1 + (n3 - n1 + 1) / 2;
  }
  return spline(p, r, n1);
}

// Seltzer-Berger 2D grid (proxy/Proxy2DVector.hpp:53-55, Proxy2DVector.cpp).
struct SB2D {
  double x[32], y[57], v[57][32];
};

static bool read_sb(const std::string &path, SB2D &t) {
  std::ifstream f(path);
  if (!f) return false;
  int k, nx, ny;
  f >> k >> nx >> ny;
  if (!f || nx != 32 || ny != 57) return false;
  for (int i = 0; i < 32; ++i) f >> t.x[i];
  for (int j = 0; j < 57; ++j) f >> t.y[j];
  for (int j = 0; j < 57; ++j)
    for (int i = 0; i < 32; ++i) f >> t.v[j][i];
  return (bool)f;
}

// Bin search of Proxy2DVector.cpp:93-139 (lower node index, clamped to n-2).
static int sb_bin(const double *a, int n, double z) {
  if (z < a[1]) return 0;
  if (z >= a[n - 2]) return n - 2;
  int lo = 0, hi = n - 2;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (z < a[mid])
      hi = mid - 1;
    else
      lo = mid + 1;
  }
  return hi;
}

// Proxy2DVector::Value (Proxy2DVector.cpp:37-67): clamp, then bilinear.
static double sb_value(const SB2D &t, double x, double y) {
  if (x < t.x[0]) x = t.x[0];
  else if (x > t.x[31]) x = t.x[31];
  if (y < t.y[0]) y = t.y[0];
  else if (y > t.y[56]) y = t.y[56];
  const int ix = sb_bin(t.x, 32, x), iy = sb_bin(t.y, 57, y);
  const double x1 = t.x[ix], x2 = t.x[ix + 1], y1 = t.y[iy], y2 = t.y[iy + 1];
  const double v11 = t.v[iy][ix], v12 = t.v[iy][ix + 1], v21 = t.v[iy + 1][ix], v22 = t.v[iy + 1][ix + 1];
  return ((y2 - y) * (v11 * (x2 - x) + v12 * (x - x1)) + ((y - y1) * (v21 * (x2 - x) + v22 * (x - x1)))) /
         ((x2 - x1) * (y2 - y1));
}

// ---------------------------------------------------------------- materials
struct Element {
  int Z;
  double A;         // g/mole
  double natoms;    // atoms per mm3
  double wmass;     // mass fraction
  double zoa;       // Sandia Z/A ratio
  std::vector<double> sandia;  // rows of 5: edge(keV), a1..a4 (cm2 keV^n / g)
  std::vector<double> shells;  // binding energies (MeV), innermost first
  SB2D sb;
};

struct Material {
  std::string name;
  double density;      // g/cm3
  double ne;           // electrons per mm3
  double migdal;       // dielectric suppression factor k_p^2 = migdal * E_tot^2
  double cut_gamma, cut_e, cut_ep;  // production thresholds (MeV)
  double cut_length;   // mm
  std::vector<Element> el;
  PVec tab[GXT_NTAB];
};

static const char *kTableFile[GXT_NTAB] = {
    "Lambda.compt.gamma.asc", "LambdaPrim.compt.gamma.asc", "Lambda.conv.gamma.asc",
    "LambdaPrim.phot.gamma.asc", "Lambda.eBrem.e-.asc",      "Lambda.eIoni.e-.asc",
    "DEDX.eIoni.e-.asc",         "Range.eIoni.e-.asc",        "InverseRange.eIoni.e-.asc",
    "Lambda.eBrem.e+.asc",       "Lambda.eIoni.e+.asc",       "DEDX.eIoni.e+.asc",
    "Range.eIoni.e+.asc",        "InverseRange.eIoni.e+.asc", "LambdaMod1.msc.e-.asc",
    "LambdaMod1.msc.e+.asc"};

// kind of scaling with density: 0 = proportional (cross-sections, dE/dx,
// inverse transport mean free path), 1 = inverse (range), 2 = inverse on x
// (inverse range)
static const int kTableScale[GXT_NTAB] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 0, 0, 0, 1, 2, 0, 0};

struct ElemSpec {
  int Z;
  double A;
  double count;
};

static bool load_sandia(const std::string &dir, Element &e) {
  std::ifstream f(dir + "/sandia.dat");
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int z, nint;
    double zoa;
    is >> z >> nint >> zoa;
    std::vector<double> rows(5 * nint);
    for (int i = 0; i < nint; ++i) {
      std::getline(f, line);
      std::istringstream r(line);
      for (int k = 0; k < 5; ++k) r >> rows[5 * i + k];
    }
    if (z == e.Z) {
      e.sandia = rows;
      e.zoa = zoa;
      return true;
    }
  }
  return false;
}

static bool load_shells(const std::string &dir, Element &e) {
  std::ifstream f(dir + "/shells.dat");
  if (!f) return false;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int z, n;
    is >> z >> n;
    if (z != e.Z) continue;
    e.shells.resize(n);
    for (int i = 0; i < n; ++i) {
      double b;
      is >> b;
      e.shells[i] = b * 1.0e-6;  // eV -> MeV
    }
    return true;
  }
  return false;
}

static bool read_cuts(const std::string &path, int couple, double &g, double &em, double &ep) {
  // cut.dat is particle-major (gamma, e-, e+, proton) x couples, energies in
  // keV (ProxyDataManager.cu:83-122 / .cuh:53-54).
  std::ifstream f(path);
  if (!f) return false;
  std::string tag;
  int ncouple;
  f >> tag >> ncouple;
  std::vector<double> len(4 * ncouple), en(4 * ncouple);
  for (int i = 0; i < 4 * ncouple; ++i) f >> len[i] >> en[i];
  if (!f) return false;
  g = en[0 * ncouple + couple] * 1e-3;
  em = en[1 * ncouple + couple] * 1e-3;
  ep = en[2 * ncouple + couple] * 1e-3;
  return true;
}

// Build a material: tables from `tabdir` vector `couple`, scaled from the
// tabulated density to `density`.
static int build_material(const std::string &data, int which, double lar_density, Material &m) {
  std::string tabdir;
  int couple = 0;
  double tab_density = 0;
  std::vector<ElemSpec> comp;
  if (which == GXT_MAT_PB) {
    m.name = "G4_Pb";
    tabdir = data + "/g4_pb_lar";
    couple = 1;
    tab_density = m.density = 11.35;
    comp = {{82, 207.217, 1}};
  } else if (which == GXT_MAT_LAR) {
    m.name = "G4_lAr";
    tabdir = data + "/g4_pb_lar";
    couple = 2;
    tab_density = 1.396;
    m.density = lar_density > 0 ? lar_density : 1.396;
    comp = {{18, 39.9477, 1}};
  } else if (which == GXT_MAT_PBWO4) {
    m.name = "PbWO4";
    tabdir = data + "/pbwo4";
    couple = 0;
    tab_density = m.density = 8.28;
    comp = {{8, 15.9994, 4}, {74, 183.842, 1}, {82, 207.217, 1}};
  } else {
    return fail(GXT_EINVAL, "unknown material");
  }
  double amass = 0;
  for (auto &c : comp) amass += c.count * c.A;
  m.ne = 0;
  for (auto &c : comp) {
    Element e;
    e.Z = c.Z;
    e.A = c.A;
    e.wmass = c.count * c.A / amass;
    e.natoms = m.density * kAvogadro * e.wmass / c.A * 1.0e-3;  // per mm3
    m.ne += e.natoms * c.Z;
    if (!load_sandia(data, e)) return fail(GXT_EDATA, "sandia.dat: no Z=" + std::to_string(c.Z));
    if (!load_shells(data, e)) return fail(GXT_EDATA, "shells.dat: no Z=" + std::to_string(c.Z));
    if (!read_sb(data + "/brem_SB/br" + std::to_string(c.Z), e.sb))
      return fail(GXT_EDATA, "brem_SB/br" + std::to_string(c.Z));
    m.el.push_back(e);
  }
  // G4SeltzerBergerModel: densityFactor = n_e * 4 pi r_e lambda_e^2
  m.migdal = m.ne * 4.0 * kPi * kClassicRe * kComptonLength * kComptonLength;
  m.cut_length = 0.7;
  if (!read_cuts(tabdir + "/cut.dat", couple, m.cut_gamma, m.cut_e, m.cut_ep))
    return fail(GXT_EDATA, tabdir + "/cut.dat");
  const double f = m.density / tab_density;
  for (int t = 0; t < GXT_NTAB; ++t) {
    std::vector<PVec> vecs;
    // the multiple-scattering tables are optional (none for PbWO4)
    if ((t == GXT_TAB_MSC_EM || t == GXT_TAB_MSC_EP) && !std::ifstream(tabdir + "/" + kTableFile[t])) {
      m.tab[t] = PVec();
      continue;
    }
    if (!read_asc(tabdir + "/" + kTableFile[t], vecs) || couple >= (int)vecs.size())
      return fail(GXT_EDATA, tabdir + "/" + kTableFile[t]);
    PVec p = vecs[couple];
    if (f != 1.0) {
      for (size_t i = 0; i < p.x.size(); ++i) {
        if (kTableScale[t] == 0) p.y[i] *= f;
        else if (kTableScale[t] == 1) p.y[i] /= f;
        else p.x[i] /= f;
      }
    }
    fill_d2(p);
    m.tab[t] = p;
  }
  return GXT_OK;
}

// Sandia photo-absorption cross-section of one element (per atom, mm2),
// following G4SandiaTable::GetSandiaCofPerAtom's interval choice (last
// interval whose edge is <= E; below the first edge the first interval).
static double sandia_atom(const Element &e, double ekin) {
  const double ek = ekin * 1.0e3;  // keV
  const int nint = (int)e.sandia.size() / 5;
  int row = nint - 1;
  while (row > 0 && ek < e.sandia[5 * row]) --row;
  const double *c = &e.sandia[5 * row];
  const double inv = 1.0 / ek;
  const double s = inv * (c[1] + inv * (c[2] + inv * (c[3] + inv * c[4])));  // cm2/g
  // per atom: (A/N_A) * sigma/rho with A = Z / (Z/A) as in GetSandiaCofPerAtom; cm2 -> mm2
  return (e.Z / e.zoa) / kAvogadro * s * 100.0;
}

static double sandia_macro(const Material &m, double ekin) {
  double s = 0;
  for (const Element &e : m.el) s += e.natoms * sandia_atom(e, ekin);
  return s;
}

// Heitler two-photon annihilation cross-section per electron (mm2)
// (G4eeToTwoGammaModel::ComputeCrossSectionPerElectron; the reference has no
// positron process).
static double heitler_xs(double ekin) {
  const double tau = ekin / kMe;
  const double gam = tau + 1.0;
  const double gamma2 = gam * gam;
  const double bg2 = tau * (tau + 2.0);
  const double bg = std::sqrt(bg2);
  return kPi * kClassicRe * kClassicRe *
         ((gamma2 + 4.0 * gam + 1.0) * log_cw(gam + bg) - (gam + 3.0) * bg) / (bg2 * (gam + 1.0));
}

// ---------------------------------------------------------------- geometry
// Slab form: volumes v between the planes z[v] and z[v + 1].  Tree form
// (GXT_GEOM_TREE): a track's volume is a touchable of the context's
// navigator and cell[touchable] its tally cell (the touchables with a
// material in touchable order, then one vacuum cell, slot -1); slot and
// visible are per cell.
struct Geometry {
  int type;
  double hx, hy;
  std::vector<double> z;      // plane positions, z[0] .. z[nvol]
  std::vector<int> slot;      // material slot per volume (cell)
  std::vector<int> visible;   // volume counts toward visible energy
  bool tree = false;
  std::vector<int> cell;      // tree form: touchable -> cell
  int nvol() const { return (int)slot.size(); }
};

static void build_geometry(const gxt_config &c, Geometry &g) {
  g.type = c.geometry;
  g.hx = g.hy = 0.5 * c.box_xy;
  g.z.clear();
  g.slot.clear();
  g.visible.clear();
  g.tree = false;
  g.cell.clear();
  if (c.geometry == GXT_GEOM_TREE) return;  // cells built with the navigator
  if (c.geometry == GXT_GEOM_BOX) {
    g.z = {0.0, c.box_depth};
    g.slot = {0};
    g.visible = {1};
  } else {
    const double period = c.absorber_mm + c.gap_mm;
    for (int l = 0; l < c.n_layers; ++l) {
      g.z.push_back(l * period);
      g.z.push_back(l * period + c.absorber_mm);
      g.slot.push_back(0);
      g.slot.push_back(1);
      g.visible.push_back(0);
      g.visible.push_back(1);
    }
    g.z.push_back(c.n_layers * period);
  }
}

// Volume containing a point; points on a plane belong to the volume the
// direction points into.  -1 = outside the setup.
static int locate(const Geometry &g, const double p[3], const double d[3]) {
  const double ax = std::fabs(p[0]), ay = std::fabs(p[1]);
  if (ax > g.hx || ay > g.hy) return -1;
  if (ax == g.hx && p[0] * d[0] > 0) return -1;
  if (ay == g.hy && p[1] * d[1] > 0) return -1;
  const int nv = g.nvol();
  const double z = p[2];
  if (z < g.z[0] || z > g.z[nv]) return -1;
  if (z == g.z[0] && d[2] < 0) return -1;
  if (z == g.z[nv] && d[2] > 0) return -1;
  // largest k with z[k] <= z, k in [0, nv-1]
  int lo = 0, hi = nv - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (g.z[mid] <= z)
      lo = mid;
    else
      hi = mid - 1;
  }
  if (z == g.z[lo] && lo > 0 && d[2] < 0) --lo;
  return lo;
}

// Straight-line distance to the boundary of volume v (an analytic
// slab/box navigator in place of the reference's VecGeom navigation).  face: 0 = -z, 1 = +z, 2 = x side, 3 = y side.
static double dist_out(const Geometry &g, int v, const double p[3], const double d[3], int &face) {
  double s = kHuge;
  face = -1;
  if (d[2] > 0) {
    s = (g.z[v + 1] - p[2]) / d[2];
    face = 1;
  } else if (d[2] < 0) {
    s = (g.z[v] - p[2]) / d[2];
    face = 0;
  }
  if (d[0] != 0) {
    const double sx = ((d[0] > 0 ? g.hx : -g.hx) - p[0]) / d[0];
    if (sx < s) {
      s = sx;
      face = 2;
    }
  }
  if (d[1] != 0) {
    const double sy = ((d[1] > 0 ? g.hy : -g.hy) - p[1]) / d[1];
    if (sy < s) {
      s = sy;
      face = 3;
    }
  }
  return s < 0 ? 0.0 : s;
}

static double safety(const Geometry &g, int v, const double p[3]) {
  double s = std::min(p[2] - g.z[v], g.z[v + 1] - p[2]);
  s = std::min(s, g.hx - std::fabs(p[0]));
  s = std::min(s, g.hy - std::fabs(p[1]));
  return s < 0 ? 0.0 : s;
}

// FP64 sine and cosine with the same operations as the device build, so
// that both give the same bits: Cody and Waite's additive reduction
// x = k pi/2 + r (k = rint(x 2/pi); pi/2 in three parts, the first two of 33
// bits so that k P1 and k P2 are exact for |k| < 2^20), the Taylor series of
// sin to r^15 and of cos to r^16 on |r| <= pi/4, and the quadrant k mod 4.
// The split of pi/2 and 2/pi are the correctly rounded values (checked
// against a 60-digit pi in the tests).
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
constexpr double kPio2P1 = 0x1.921fb544p+0;
constexpr double kPio2P2 = 0x1.0b4611a6p-34;
constexpr double kPio2P3 = 0x1.3198a2e037073p-69;
static inline void sincos_cw(double x, double &s, double &c) {
  if (!(std::fabs(x) <= 1e15)) {
    s = x - x;
    c = x - x;
    return;
  }
  const double k = std::rint(x * kTwoOverPi);
  const double r = ((x - k * kPio2P1) - k * kPio2P2) - k * kPio2P3;
  const double z = r * r;
  // 1/n! as exact quotients of exact integers
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
  double q = std::fmod(k, 4.0);
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
static inline double sin_cw(double x) {
  double s, c;
  sincos_cw(x, s, c);
  return s;
}
static inline double cos_cw(double x) {
  double s, c;
  sincos_cw(x, s, c);
  return c;
}

// Math::RotateToLabFrame (core/math_wrappers.hpp:194-212).
static inline void rotate_to_lab(double &u, double &v, double &w, double u1, double u2, double u3) {
  double up = u1 * u1 + u2 * u2;
  if (up > 0.0) {
    up = std::sqrt(up);
    const double px = u, py = v, pz = w;
    u = (u3 * px * u1 - u2 * py) * (1.0 / up) + u1 * pz;
    v = (u3 * px * u2 + u1 * py) * (1.0 / up) + u2 * pz;
    w = -up * px + u3 * pz;
  } else if (u3 < 0.0) {
    u = -u;
    w = -w;
  }
}

static inline void unit3(double *a) {
  const double m2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  if (m2 > 0) {
    const double inv = 1.0 / std::sqrt(m2);
    a[0] *= inv;
    a[1] *= inv;
    a[2] *= inv;
  }
}

// ConstFieldHelixStepper::DoStep (magneticfield/ConstFieldHelixStepper.hpp:164-239)
// in mm/MeV/tesla, followed by the direction normalisation of
// FieldPropagationHandler::PropagateInVolume (FieldPropagationHandler.cpp:323-330).
static void helix_step(const double B[3], double bmag, double charge, double mom, double step,
                       double pos[3], double dir[3]) {
  const double ub[3] = {B[0] / bmag, B[1] / bmag, B[2] / bmag};
  const double uvb = dir[0] * ub[0] + dir[1] * ub[1] + dir[2] * ub[2];
  double dt2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] - uvb * uvb;
  if (dt2 < 0.0) dt2 = 0.0;
  const double sinvb = std::sqrt(dt2) + 1.0e-30;
  const double R = mom * sinvb / (kB2C * charge * bmag);
  double vx[3] = {dir[0] - uvb * ub[0], dir[1] - uvb * ub[1], dir[2] - uvb * ub[2]};
  unit3(vx);
  const double cvb[3] = {vx[1] * ub[2] - vx[2] * ub[1], vx[2] * ub[0] - vx[0] * ub[2],
                         vx[0] * ub[1] - vx[1] * ub[0]};
  const double phi = -step * charge * bmag * kB2C / mom;
  const double sphi = sin_cw(phi), cphi = cos_cw(phi);
  for (int i = 0; i < 3; ++i) {
    pos[i] = pos[i] + R * (cphi - 1.0) * cvb[i] - R * sphi * vx[i] + step * uvb * ub[i];
    dir[i] = uvb * ub[i] + cphi * sinvb * vx[i] + sphi * sinvb * cvb[i];
  }
  unit3(dir);
}

// ---------------------------------------------------------------- field map
// Cylindrical (r, z) field map (ScalarCMSmagField,
// examples/magneticfield/simplifiedCMS/inc/Geant/ScalarCMSmagField.h): nodes
// r-major, (Br, Bphi, Bz) in tesla per node.
struct FieldMap {
  int nr = 0, nz = 0;
  double r0 = 0, z0 = 0, dr = 0, dz = 0, bmax = 0;
  std::vector<double> v;
};

// ScalarCMSmagField::ReadVectorData (ScalarCMSmagField.h:174-206): one node
// per line, "R(mm)= r phi(deg)= phi Z(mm)= z Bz(tesla)= bz Br(tesla)= br
// Bphi(tesla)= bphi".  The grid (181 x 161, 50 mm x 200 mm from (0, -16 m) in
// the reference) is taken from the file: nz = the lines sharing the first r,
// spacings from the first differences; a line off the grid, an unparsable
// line or a count that is not nr * nz is an error.
static int read_field_map(const char *path, FieldMap &m) {
  std::ifstream in(path ? path : "");
  if (!path || !in) return fail(GXT_EDATA, std::string("cannot open field map ") + (path ? path : "(null)"));
  std::vector<double> r, z, b;
  std::string line;
  long lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    char l0[32], l1[32], l2[32], l3[32], l4[32], l5[32], tail[8];
    double rv, pv, zv, bz, br, bp;
    const int k = std::sscanf(line.c_str(), "%31s %lf %31s %lf %31s %lf %31s %lf %31s %lf %31s %lf %7s", l0, &rv, l1,
                              &pv, l2, &zv, l3, &bz, l4, &br, l5, &bp, tail);
    if (k != 12 || std::strcmp(l0, "R(mm)=") || std::strcmp(l1, "phi(deg)=") || std::strcmp(l2, "Z(mm)=") ||
        std::strcmp(l3, "Bz(tesla)=") || std::strcmp(l4, "Br(tesla)=") || std::strcmp(l5, "Bphi(tesla)=") ||
        !std::isfinite(rv) || !std::isfinite(zv) || !std::isfinite(bz) || !std::isfinite(br) || !std::isfinite(bp))
      return fail(GXT_EDATA, "malformed field-map line " + std::to_string(lineno));
    r.push_back(rv);
    z.push_back(zv);
    b.insert(b.end(), {br, bp, bz});
  }
  const size_t n = r.size();
  size_t nz = 1;
  while (nz < n && r[nz] == r[0]) ++nz;
  if (n < 4 || nz < 2 || nz == n || n % nz) return fail(GXT_EDATA, "field map is not a regular (r, z) grid");
  const size_t nr = n / nz;
  const double dr = r[nz] - r[0], dz = z[1] - z[0];
  if (!(dr > 0) || !(dz > 0) || r[0] < 0) return fail(GXT_EDATA, "field-map grid must start at r >= 0 and ascend");
  for (size_t i = 0; i < n; ++i) {
    const double wr = r[0] + (double)(i / nz) * dr, wz = z[0] + (double)(i % nz) * dz;
    if (std::fabs(r[i] - wr) > 1e-9 * std::max(std::fabs(wr), dr) ||
        std::fabs(z[i] - wz) > 1e-9 * std::max(std::fabs(wz), dz))
      return fail(GXT_EDATA, "field-map node " + std::to_string(i + 1) + " is off the regular grid");
  }
  m.nr = (int)nr;
  m.nz = (int)nz;
  m.r0 = r[0];
  m.z0 = z[0];
  m.dr = dr;
  m.dz = dz;
  m.bmax = 0;
  for (size_t i = 0; i < n; ++i)
    m.bmax = std::max(m.bmax, std::sqrt(b[3 * i] * b[3 * i] + b[3 * i + 1] * b[3 * i + 1] + b[3 * i + 2] * b[3 * i + 2]));
  m.v.swap(b);
  return GXT_OK;
}

// ScalarCMSmagField::GetFieldValueRZ / GetFieldValue (ScalarCMSmagField.h:247-331):
// r and z clamped to the grid; bilinear area weights of the cell (here as
// products of the fractional cell position, the same weights divided by the
// cell area); Bx = Br cos - Bphi sin, By = Br sin + Bphi cos with the point's
// azimuth.  Fixed here: the cell index is clamped to the last cell (the
// reference reads past the grid at r = rmax or z = zmax), and the azimuth on
// the axis is 0 (the reference leaves 1/r unset there).
static void map_field(const FieldMap &f, double x, double y, double z, double B[3]) {
  const double r = std::sqrt(x * x + y * y);
  const double rmax = f.r0 + (double)(f.nr - 1) * f.dr;
  const double zmax = f.z0 + (double)(f.nz - 1) * f.dz;
  const double rc = std::fmin(std::fmax(r, f.r0), rmax);
  const double zc = std::fmin(std::fmax(z, f.z0), zmax);
  const double fr = (rc - f.r0) / f.dr;
  const double fz = (zc - f.z0) / f.dz;
  const int ir = std::min((int)fr, f.nr - 2);
  const int iz = std::min((int)fz, f.nz - 2);
  const double tr = fr - (double)ir, tz = fz - (double)iz;
  const double w00 = (1.0 - tr) * (1.0 - tz), w01 = (1.0 - tr) * tz;
  const double w10 = tr * (1.0 - tz), w11 = tr * tz;
  const double *p = f.v.data() + ((size_t)ir * f.nz + iz) * 3;
  const double *q = p + (size_t)f.nz * 3;
  double b[3];
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

// The field a charged track sees: the map at pos + origin, or uniform.
struct Field {
  const FieldMap *map = nullptr;
  double b[3] = {0, 0, 0};
  double origin[3] = {0, 0, 0};
  int stepper = GXT_STEPPER_HELIX;
  void at(const double pos[3], double B[3]) const {
    if (map) {
      map_field(*map, pos[0] + origin[0], pos[1] + origin[1], pos[2] + origin[2], B);
    } else {
      B[0] = b[0];
      B[1] = b[1];
      B[2] = b[2];
    }
  }
};

// Cash-Karp embedded Runge-Kutta 4(5) (J. R. Cash, A. H. Karp, ACM Trans.
// Math. Softw. 16 (1990) 201-222) on dx/ds = u, du/ds = kq (u x B(x)),
// kq = -q kB2C / p: six stages, the fifth-order solution propagated, the
// difference to the fourth-order one as the error estimate.
static const double kCkA[6][5] = {{0, 0, 0, 0, 0},
                                  {1.0 / 5.0, 0, 0, 0, 0},
                                  {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0},
                                  {3.0 / 10.0, -9.0 / 10.0, 6.0 / 5.0, 0, 0},
                                  {-11.0 / 54.0, 5.0 / 2.0, -70.0 / 27.0, 35.0 / 27.0, 0},
                                  {1631.0 / 55296.0, 175.0 / 512.0, 575.0 / 13824.0, 44275.0 / 110592.0, 253.0 / 4096.0}};
static const double kCkB5[6] = {37.0 / 378.0, 0, 250.0 / 621.0, 125.0 / 594.0, 0, 512.0 / 1771.0};
static const double kCkB4[6] = {2825.0 / 27648.0, 0, 18575.0 / 48384.0, 13525.0 / 55296.0, 277.0 / 14336.0, 1.0 / 4.0};

static void lorentz(const Field &F, double kq, const double x[3], const double u[3], double f[3]) {
  double B[3];
  F.at(x, B);
  f[0] = kq * (u[1] * B[2] - u[2] * B[1]);
  f[1] = kq * (u[2] * B[0] - u[0] * B[2]);
  f[2] = kq * (u[0] * B[1] - u[1] * B[0]);
}

static void ck45_step(const Field &F, double kq, double h, const double f1[3], double pos[3], double dir[3], double &ex,
                      double &eu) {
  double u[6][3], f[6][3];
  for (int c = 0; c < 3; ++c) {
    u[0][c] = dir[c];
    f[0][c] = f1[c];
  }
  for (int s = 1; s < 6; ++s) {
    double x[3];
    for (int c = 0; c < 3; ++c) {
      // sums in stage order, as the restatement's arithmetic must follow one fixed order
      double sx = kCkA[s][0] * u[0][c], su = kCkA[s][0] * f[0][c];
      for (int j = 1; j < s; ++j) {
        sx += kCkA[s][j] * u[j][c];
        su += kCkA[s][j] * f[j][c];
      }
      x[c] = pos[c] + h * sx;
      u[s][c] = dir[c] + h * su;
    }
    lorentz(F, kq, x, u[s], f[s]);
  }
  double ex2 = 0, eu2 = 0;
  for (int c = 0; c < 3; ++c) {
    const int st[5] = {0, 2, 3, 4, 5};  // stages with a nonzero b5 - b4 (stage 2's is 0)
    double dx = (kCkB5[0] - kCkB4[0]) * u[0][c], du = (kCkB5[0] - kCkB4[0]) * f[0][c];
    for (int k = 1; k < 5; ++k) {
      dx += (kCkB5[st[k]] - kCkB4[st[k]]) * u[st[k]][c];
      du += (kCkB5[st[k]] - kCkB4[st[k]]) * f[st[k]][c];
    }
    dx *= h;
    du *= h;
    ex2 += dx * dx;
    eu2 += du * du;
    const int sb[4] = {0, 2, 3, 5};  // stages with a nonzero b5
    double sx = kCkB5[0] * u[0][c], su = kCkB5[0] * f[0][c];
    for (int k = 1; k < 4; ++k) {
      sx += kCkB5[sb[k]] * u[sb[k]][c];
      su += kCkB5[sb[k]] * f[sb[k]][c];
    }
    pos[c] = pos[c] + h * sx;
    dir[c] = dir[c] + h * su;
  }
  unit3(dir);
  ex += std::sqrt(ex2);
  eu += std::sqrt(eu2);
}

static const int kRkMaxSub = 16;
static const int kRkMaxN = 65537;

// One propagation sub-step of length s: helix in the start-point field B, or
// RK45 in ceil(s k |B| / GXT_RK_ANGLE) equal Cash-Karp steps (k = 0.299792458 / p).
static void field_advance(const Field &F, const double B[3], double bm, double q, double mom, double s, double pos[3],
                          double dir[3], double &ex, double &eu) {
  if (F.stepper == GXT_STEPPER_RK45) {
    const double turn = s * (-kB2C) * bm / mom;
    const int n = turn > GXT_RK_ANGLE ? (int)std::fmin(std::ceil(turn / GXT_RK_ANGLE), (double)kRkMaxN) : 1;
    const double h = s / (double)n;
    const double kq = -q * kB2C / mom;
    double f1[3] = {kq * (dir[1] * B[2] - dir[2] * B[1]), kq * (dir[2] * B[0] - dir[0] * B[2]),
                    kq * (dir[0] * B[1] - dir[1] * B[0])};
    for (int k = 0; k < n; ++k) {
      if (k > 0) lorentz(F, kq, pos, dir, f1);
      ck45_step(F, kq, h, f1, pos, dir, ex, eu);
    }
  } else if (bm > 0.0) {
    helix_step(B, bm, q, mom, s, pos, dir);
  } else {
    for (int c = 0; c < 3; ++c) pos[c] += s * dir[c];
  }
}

// ---------------------------------------------------------------- physics
struct Secondary {
  int species;
  double ekin;
  double dir[3];
};

struct Interaction {
  double ekin;      // primary after the interaction (0: killed)
  double dir[3];
  double edep;      // local deposit
  int nsec;
  Secondary sec[2];
};

struct Physics {
  std::vector<Material> mat;  // material slots
  double tracking_cut;
};

// Macroscopic cross-section (1/mm) of discrete process p
// (ProxyEmProcess::GetLambda via ProxyCompton.hpp:72-76, ProxyConversion.hpp:73-77,
// ProxyPhotoElectric.hpp:73-77, ProxyBremsstrahlung.hpp:72-76, ProxyIonization.hpp:74-78)
// with the track's couple, Compton above 1 MeV from LambdaPrim/E,
// photoelectric LambdaPrim/E and Sandia below the table.
static double macro_xs(const Material &m, int species, int p, double e) {
  if (species == 0) {
    if (p == 0) {
      const PVec &lo = m.tab[GXT_TAB_LAMBDA_COMPT];
      if (e < lo.x.back()) return pvec_value(lo, e);
      return pvec_value(m.tab[GXT_TAB_LAMBDAPRIM_COMPT], e) / e;
    }
    if (p == 1) return pvec_value(m.tab[GXT_TAB_LAMBDA_CONV], e);
    const PVec &pe = m.tab[GXT_TAB_LAMBDAPRIM_PHOT];
    if (e >= pe.x[0]) return pvec_value(pe, e) / e;
    return sandia_macro(m, e);
  }
  const bool pos = (species == 2);
  if (p == 0) return pvec_value(m.tab[pos ? GXT_TAB_LAMBDA_BREM_EP : GXT_TAB_LAMBDA_BREM_EM], e);
  if (p == 1) return pvec_value(m.tab[pos ? GXT_TAB_LAMBDA_IONI_EP : GXT_TAB_LAMBDA_IONI_EM], e);
  return m.ne * heitler_xs(e);
}

static inline int nproc(int species) { return species == 1 ? 2 : 3; }

// Element choice (the reference hard-codes Z = 10): one draw, only for
// compounds.  weights: 0 = n Z(Z+1) (pair, brems), 1 = partial Sandia (photo).
static int select_element(const Material &m, int kind, double e, uint32_t s[6]) {
  const int ne = (int)m.el.size();
  if (ne == 1) return 0;
  double w[8], tot = 0;
  for (int i = 0; i < ne; ++i) {
    const Element &el = m.el[i];
    w[i] = kind == 0 ? el.natoms * el.Z * (el.Z + 1.0) : el.natoms * sandia_atom(el, e);
    tot += w[i];
  }
  const double r = mrg_uniform(s) * tot;
  double acc = 0;
  for (int i = 0; i < ne - 1; ++i) {
    acc += w[i];
    if (r < acc) return i;
  }
  return ne - 1;
}

// Klein-Nishina (proxy/ProxyKleinNishina.hpp:96-166).
static void sample_compton(const Material &, double e0, const double d0[3], uint32_t s[6], Interaction &o) {
  const double lowE = 100.0e-6;
  o.nsec = 0;
  o.edep = 0;
  o.ekin = e0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  if (e0 < lowE) return;
  const double E0_m = e0 / kMe;
  const double eps0 = 1.0 / (1.0 + 2.0 * E0_m);
  const double eps0sq = eps0 * eps0;
  // 1 - eps0^2 and 1 - eps in forms without cancellation (the reference
  // writes 1 - eps, ProxyKleinNishina.hpp:124): 1 - eps0 = 2 E0_m eps0;
  // 1 - eps = -expm1(ln eps), or (1 - eps^2) / (1 + eps)
  const double omeps0sq = 2.0 * E0_m * eps0 * (1.0 + eps0);
  const double alpha1 = -log_cw(eps0);
  const double alpha2 = 0.5 * omeps0sq;
  double eps, omeps, epssq, onecost, sint2, greject;
  do {
    if (alpha1 / (alpha1 + alpha2) > mrg_uniform(s)) {
      const double le = -alpha1 * mrg_uniform(s);
      eps = exp_cw(le);
      omeps = -expm1_cw(le);
      epssq = eps * eps;
    } else {
      const double u = mrg_uniform(s);
      epssq = eps0sq + omeps0sq * u;
      eps = std::sqrt(epssq);
      omeps = omeps0sq * (1.0 - u) / (1.0 + eps);
    }
    onecost = omeps / (eps * E0_m);
    sint2 = onecost * (2.0 - onecost);
    greject = 1.0 - eps * sint2 / (1.0 + epssq);
  } while (greject < mrg_uniform(s));
  const double sinT = sint2 < 0.0 ? 0.0 : std::sqrt(sint2);
  const double phi = kTwoPi * mrg_uniform(s);
  const double e1 = eps * e0;
  double gd[3] = {sinT * cos_cw(phi), sinT * sin_cw(phi), 1.0 - onecost};
  rotate_to_lab(gd[0], gd[1], gd[2], d0[0], d0[1], d0[2]);
  const double te = e0 - e1;
  if (e1 > lowE) {
    o.ekin = e1;
    o.dir[0] = gd[0];
    o.dir[1] = gd[1];
    o.dir[2] = gd[2];
    Secondary &el = o.sec[o.nsec++];
    el.species = 1;
    el.ekin = te;
    el.dir[0] = e0 * d0[0] - e1 * gd[0];
    el.dir[1] = e0 * d0[1] - e1 * gd[1];
    el.dir[2] = e0 * d0[2] - e1 * gd[2];
    unit3(el.dir);
  } else {
    o.ekin = 0;
    o.edep = e1 + te;
  }
}

// Bethe-Heitler screening functions (ProxyBetheHeitler.hpp:184-210).
static inline double screen1(double x) {
  return x > 1.0 ? 42.24 - 8.368 * log_cw(x + 0.952) : 42.392 - x * (7.796 - 1.961 * x);
}
static inline double screen2(double x) {
  return x > 1.0 ? 42.24 - 8.368 * log_cw(x + 0.952) : 41.405 - x * (5.828 - 0.8945 * x);
}
// ProxyEmModel::ComputeCoulombFactor (proxy/ProxyEmModel.hpp:114-128).
static inline double coulomb_factor(double Z) {
  const double k1 = 0.0083, k2 = 0.20206, k3 = 0.0020, k4 = 0.0369;
  const double az1 = kFineStructure * Z;
  const double az2 = az1 * az1;
  const double az4 = az2 * az2;
  return (k1 * az4 + k2 + 1.0 / (1.0 + az2)) * az2 - (k3 * az4 + k4) * az4;
}

// Bethe-Heitler parameters above 2 MeV (ProxyBetheHeitler.hpp:80-104): the
// nuclear term FZ = 8 ln(Z)/3 (+ 8 f_c(Z) above 50 MeV) and the screening
// variable's factor, delta = screenfac / (eps (1 - eps)).
static inline void bh_setup(int Z, double e0, double &FZ, double &screenfac) {
  const double logZ3 = log_cw(1.0 * Z) / 3.0;
  FZ = 8.0 * logZ3;
  if (e0 > 50.0) FZ += 8.0 * coulomb_factor(Z);
  const double Z3 = pow_cw(1.0 * Z, 1.0 / 3.0);
  screenfac = 136.0 * (kMe / e0) / Z3;
}

// Unnormalised energy-sharing density the pair sampler draws eps from above
// 2 MeV: (eps - 1/2)^2 max(F1, 0) + max(F2, 0) / 2 with F1 = screen1(delta) - FZ,
// F2 = screen2(delta) - FZ (the two rejection branches of ProxyBetheHeitler.hpp:
// 106-126 with their weights NormF1, NormF2).
static double bh_density(int Z, double e0, double eps) {
  double FZ, screenfac;
  bh_setup(Z, e0, FZ, screenfac);
  const double sv = screenfac / (eps * (1.0 - eps));
  const double F1 = std::max(screen1(sv) - FZ, 0.0), F2 = std::max(screen2(sv) - FZ, 0.0);
  return (eps - 0.5) * (eps - 0.5) * F1 + 0.5 * F2;
}

// Bethe-Heitler pair production (proxy/ProxyBetheHeitler.hpp:64-182) with
// the element chosen from the material and the Tsai angle used as an
// angle, cos/sin of u*m/E (the reference uses u*m/E as the cosine).
static void sample_pair(const Material &m, double e0, const double d0[3], uint32_t s[6], Interaction &o) {
  o.nsec = 0;
  o.edep = 0;
  o.ekin = e0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  const double eps0 = kMe / e0;
  if (eps0 > 0.5) return;
  const int ie = select_element(m, 0, e0, s);
  const int Z = m.el[ie].Z;
  double eps;
  if (e0 < 2.0) {
    eps = eps0 + (0.5 - eps0) * mrg_uniform(s);
  } else {
    double FZ, screenfac;
    bh_setup(Z, e0, FZ, screenfac);
    const double screenmax = exp_cw((42.24 - FZ) / 8.368) - 0.952;
    const double screenmin = std::min(4.0 * screenfac, screenmax);
    const double eps1 = 0.5 - 0.5 * std::sqrt(1.0 - screenmin / screenmax);
    const double epsmin = std::max(eps0, eps1);
    const double epsrange = 0.5 - epsmin;
    const double F10 = screen1(screenmin) - FZ;
    const double F20 = screen2(screenmin) - FZ;
    const double NormF1 = std::max(F10 * epsrange * epsrange, 0.0);
    const double NormF2 = std::max(1.5 * F20, 0.0);
    double greject;
    do {
      if (NormF1 / (NormF1 + NormF2) > mrg_uniform(s)) {
        eps = 0.5 - epsrange * pow_cw(mrg_uniform(s), 0.333333);
        const double sv = screenfac / (eps * (1.0 - eps));
        greject = (screen1(sv) - FZ) / F10;
      } else {
        eps = epsmin + epsrange * mrg_uniform(s);
        const double sv = screenfac / (eps * (1.0 - eps));
        greject = (screen2(sv) - FZ) / F20;
      }
    } while (greject < mrg_uniform(s));
  }
  const double eTot = (mrg_uniform(s) > 0.5) ? (1.0 - eps) * e0 : eps * e0;
  const double pTot = e0 - eTot;
  const double r1 = mrg_uniform(s);
  const double r2 = mrg_uniform(s);
  const double aa0 = -log_cw(r1 * r2);
  const double u = (0.25 > mrg_uniform(s)) ? aa0 / 0.625 : aa0 / 1.875;
  const double phi = kTwoPi * mrg_uniform(s);
  const double cosp = cos_cw(phi), sinp = sin_cw(phi);
  const double the = u * kMe / eTot, thp = u * kMe / pTot;
  const double ste = sin_cw(the), cte = cos_cw(the);
  const double stp = sin_cw(thp), ctp = cos_cw(thp);
  Secondary &el = o.sec[0];
  el.species = 1;
  el.ekin = eTot - kMe;
  el.dir[0] = ste * cosp;
  el.dir[1] = ste * sinp;
  el.dir[2] = cte;
  rotate_to_lab(el.dir[0], el.dir[1], el.dir[2], d0[0], d0[1], d0[2]);
  Secondary &po = o.sec[1];
  po.species = 2;
  po.ekin = pTot - kMe;
  po.dir[0] = -stp * cosp;
  po.dir[1] = -stp * sinp;
  po.dir[2] = ctp;
  rotate_to_lab(po.dir[0], po.dir[1], po.dir[2], d0[0], d0[1], d0[2]);
  o.nsec = 2;
  o.ekin = 0;
}

// Sauter-Gavrila photo-electron (proxy/ProxySauterGavrila.hpp:66-130), with the
// element chosen by partial Sandia cross-section, the first shell whose
// binding energy is below E (the reference's shell loop runs past the
// ionisable shells) and the binding energy deposited locally.
static void sample_photo(const Material &m, double e0, const double d0[3], uint32_t s[6], Interaction &o) {
  o.nsec = 0;
  o.edep = 0;
  o.ekin = 0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  const int ie = select_element(m, 1, e0, s);
  const Element &el = m.el[ie];
  double bind = 0.0;
  for (double b : el.shells)
    if (e0 >= b) {
      bind = b;
      break;
    }
  const double te = e0 - bind;
  const double tau = e0 / kMe;
  double cost;
  if (tau > 50.0) {
    cost = 1.0;
  } else {
    const double gamma = tau + 1.0;
    const double beta = std::sqrt(tau * (tau + 2.0)) / gamma;
    const double A = (1.0 - beta) / beta;
    const double Ap2 = A + 2.0;
    const double B = 0.5 * beta * gamma * (gamma - 1.0) * (gamma - 2.0);
    const double grej = 2.0 * (1.0 + A * B) / A;
    double z, g;
    do {
      const double q = mrg_uniform(s);
      z = 2.0 * A * (2.0 * q + Ap2 * std::sqrt(q)) / (Ap2 * Ap2 - 4.0 * q);
      g = (2.0 - z) * (1.0 / (A + z) + B);
    } while (g < mrg_uniform(s) * grej);
    cost = 1.0 - z;
  }
  const double sint = std::sqrt((1.0 + cost) * (1.0 - cost));
  const double phi = kTwoPi * mrg_uniform(s);
  Secondary &e = o.sec[0];
  e.species = 1;
  e.ekin = te;
  e.dir[0] = sint * cos_cw(phi);
  e.dir[1] = sint * sin_cw(phi);
  e.dir[2] = cost;
  rotate_to_lab(e.dir[0], e.dir[1], e.dir[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  o.edep = bind;
}

// Seltzer-Berger bremsstrahlung (proxy/ProxySeltzerBerger.hpp:204-286) with
// the couple's gamma production cut, the element chosen from the
// material, the element's own grid (the reference reads [Z] for data
// stored at [Z-1]), the Migdal dielectric factor of the material (the
// reference leaves densityFactor 0), and the DipBust angle driven by the
// electron's kinetic energy (ProxySeltzerBerger.hpp:288-315).
static void sample_brems(const Material &m, double t0, const double d0[3], uint32_t s[6], Interaction &o) {
  o.nsec = 0;
  o.edep = 0;
  o.ekin = t0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  const double cut = std::min(m.cut_gamma, t0);
  const double emax = t0;
  if (cut >= emax) return;
  const int ie = select_element(m, 0, t0, s);
  const SB2D &sb = m.el[ie].sb;
  const double etot = t0 + kMe;
  const double dcorr = m.migdal * etot * etot;
  const double xmin = log_cw(cut * cut + dcorr);
  const double xmax = log_cw(emax * emax + dcorr);
  const double y = log_cw(t0);
  const double x0 = cut / t0;
  double vmax = sb_value(sb, x0, y) * 1.02;
  const double epeak = 300.0, elow = 10.0e-3;
  if (x0 < 0.97 && (t0 > epeak || t0 < elow)) {
    const double ylim = std::min(sb_value(sb, 0.97, 4.0 * log_cw(10.0)), 1.1 * sb_value(sb, 0.97, y));
    if (ylim > vmax) vmax = ylim;
  }
  if (x0 < 0.05) vmax *= 1.2;
  double k, v;
  do {
    const double r = mrg_uniform(s);
    double x = exp_cw(xmin + r * (xmax - xmin)) - dcorr;
    if (x < 0.0) x = 0.0;
    k = std::sqrt(x);
    v = sb_value(sb, k / t0, y);
  } while (v < vmax * mrg_uniform(s));
  // DipBust direction
  const double u0 = mrg_uniform(s);
  const double c = 4.0 - 8.0 * u0;
  const double signc = c < 0.0 ? -1.0 : 1.0;
  const double a = std::fabs(c);
  const double delta = 0.5 * (std::sqrt(4.0 + a * a) + a);
  // cos = cofA - 1/cofA with cofA = -sign(c) delta^(1/3) (ProxySeltzerBerger.hpp:295-310)
  // is -sign(c) x, where x = d - 1/d (d = delta^(1/3)) solves x^3 + 3x = a:
  // 1 - x = (4 - a) / (4 + x + x^2) and 4 - a = 8 min(u0, 1 - u0)
  const double d = pow_cw(delta, 1.0 / 3.0);
  const double x = d - 1.0 / d;
  const double omx = 8.0 * std::min(u0, 1.0 - u0) / (4.0 + x + x * x);
  const double opx = 1.0 + x;
  const double cost0 = -signc * x;
  const double omc0 = signc > 0.0 ? opx : omx;
  const double opc0 = signc > 0.0 ? omx : opx;
  // the boost to the lab (ProxySeltzerBerger.hpp:295-310 writes (cos + beta) /
  // (1 + cos beta)) in forms without cancellation near the forward
  // direction: 1 - beta = 1 / (gam (gam + sqrt(tau (tau + 2)))) and
  // 1 - cos' = (1 - cos)(1 - beta) / (1 + cos beta)
  const double tau = t0 / kMe;
  const double gam = tau + 1.0;
  const double bgam = std::sqrt(tau * (tau + 2.0));
  const double omb = 1.0 / (gam * (gam + bgam));
  const double omc = omc0 * omb / (opc0 - cost0 * omb);
  const double cost = 1.0 - omc;
  const double sint = std::sqrt(std::max(omc * (2.0 - omc), 0.0));
  const double phi = kTwoPi * mrg_uniform(s);
  Secondary &g = o.sec[0];
  g.species = 0;
  g.ekin = k;
  g.dir[0] = sint * cos_cw(phi);
  g.dir[1] = sint * sin_cw(phi);
  g.dir[2] = cost;
  rotate_to_lab(g.dir[0], g.dir[1], g.dir[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  const double ptot = std::sqrt(t0 * (etot + kMe));
  o.ekin = t0 - k;
  o.dir[0] = ptot * d0[0] - k * g.dir[0];
  o.dir[1] = ptot * d0[1] - k * g.dir[1];
  o.dir[2] = ptot * d0[2] - k * g.dir[2];
  unit3(o.dir);
}

// Moller (e-) / Bhabha (e+) delta rays above the couple's electron
// production cut (proxy/ProxyMollerScattering.hpp:92-159 for Moller with the
// cut from the couple; the Bhabha branch from the published
// G4MollerBhabhaModel cross-section, absent in the reference).
static void sample_ioni(const Material &m, bool positron, double t0, const double d0[3], uint32_t s[6],
                        Interaction &o) {
  o.nsec = 0;
  o.edep = 0;
  o.ekin = t0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  const double tmin = m.cut_e;
  const double tmax = positron ? t0 : 0.5 * t0;
  if (tmin >= tmax) return;
  const double energy = t0 + kMe;
  const double xmin = tmin / t0, xmax = tmax / t0;
  const double gam = energy / kMe;
  const double gamma2 = gam * gam;
  double x, z, grej;
  if (!positron) {
    const double gg = (2.0 * gam - 1.0) / gamma2;
    double y = 1.0 - xmax;
    grej = 1.0 - gg * xmax + xmax * xmax * (1.0 - gg + (1.0 - gg * y) / (y * y));
    do {
      const double r = mrg_uniform(s);
      x = xmin * xmax / (xmin * (1.0 - r) + xmax * r);
      y = 1.0 - x;
      z = 1.0 - gg * x + x * x * (1.0 - gg + (1.0 - gg * y) / (y * y));
    } while (grej * mrg_uniform(s) > z);
  } else {
    const double beta2 = 1.0 - 1.0 / gamma2;
    const double yb = 1.0 / (1.0 + gam);
    const double y2 = yb * yb;
    const double y12 = 1.0 - 2.0 * yb;
    const double b1 = 2.0 - y2;
    const double b2 = y12 * (3.0 + y2);
    const double y122 = y12 * y12;
    const double b4 = y122 * y12;
    const double b3 = b4 + y122;
    const double xm2 = xmax * xmax;
    grej = 1.0 + (xm2 * xm2 * b4 - xmin * xmin * xmin * b3 + xm2 * b2 - xmin * b1) * beta2;
    do {
      const double r = mrg_uniform(s);
      x = xmin * xmax / (xmin * (1.0 - r) + xmax * r);
      const double x2 = x * x;
      z = 1.0 + (x2 * x2 * b4 - x * x2 * b3 + x2 * b2 - x * b1) * beta2;
    } while (grej * mrg_uniform(s) > z);
  }
  const double dT = x * t0;
  const double ptot = std::sqrt(t0 * (t0 + 2.0 * kMe));
  const double pd = std::sqrt(dT * (dT + 2.0 * kMe));
  double cost = dT * (energy + kMe) / (pd * ptot);
  if (cost > 1.0) cost = 1.0;
  const double sint = std::sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = kTwoPi * mrg_uniform(s);
  Secondary &d = o.sec[0];
  d.species = 1;
  d.ekin = dT;
  d.dir[0] = sint * cos_cw(phi);
  d.dir[1] = sint * sin_cw(phi);
  d.dir[2] = cost;
  rotate_to_lab(d.dir[0], d.dir[1], d.dir[2], d0[0], d0[1], d0[2]);
  o.nsec = 1;
  o.ekin = t0 - dT;
  o.dir[0] = ptot * d0[0] - pd * d.dir[0];
  o.dir[1] = ptot * d0[1] - pd * d.dir[1];
  o.dir[2] = ptot * d0[2] - pd * d.dir[2];
  unit3(o.dir);
}

// e+ annihilation in flight into two photons (Heitler; published
// G4eeToTwoGammaModel algorithm, absent in the reference).
static void sample_annihil(const Material &, double t0, const double d0[3], uint32_t s[6], Interaction &o) {
  o.nsec = 0;
  o.edep = 0;
  o.ekin = 0;
  o.dir[0] = d0[0];
  o.dir[1] = d0[1];
  o.dir[2] = d0[2];
  const double tau = t0 / kMe;
  const double gam = tau + 1.0;
  const double tau2 = tau + 2.0;
  const double sqgrate = std::sqrt(tau / tau2) * 0.5;
  const double sqg2m1 = std::sqrt(tau * tau2);
  // 1/2 - sqgrate (G4eeToTwoGammaModel's form) without cancellation at high
  // energy: (1 - s) / 2 with s = sqrt(tau / tau2) is 1 / (tau2 (1 + s))
  const double epsmin = 1.0 / (tau2 * (1.0 + 2.0 * sqgrate));
  const double epsmax = 0.5 + sqgrate;
  const double epsqot = epsmax / epsmin;
  double eps, greject;
  do {
    eps = epsmin * exp_cw(log_cw(epsqot) * mrg_uniform(s));
    greject = 1.0 - eps + (2.0 * gam * eps - 1.0) / (eps * tau2 * tau2);
  } while (greject < mrg_uniform(s));
  double cost = (eps * tau2 - 1.0) / (eps * sqg2m1);
  if (cost > 1.0) cost = 1.0;
  if (cost < -1.0) cost = -1.0;
  const double sint = std::sqrt((1.0 + cost) * (1.0 - cost));
  const double phi = kTwoPi * mrg_uniform(s);
  const double etot = t0 + 2.0 * kMe;
  Secondary &g1 = o.sec[0];
  g1.species = 0;
  g1.ekin = eps * etot;
  g1.dir[0] = sint * cos_cw(phi);
  g1.dir[1] = sint * sin_cw(phi);
  g1.dir[2] = cost;
  rotate_to_lab(g1.dir[0], g1.dir[1], g1.dir[2], d0[0], d0[1], d0[2]);
  Secondary &g2 = o.sec[1];
  g2.species = 0;
  g2.ekin = (1.0 - eps) * etot;
  const double pp = std::sqrt(t0 * (t0 + 2.0 * kMe));
  g2.dir[0] = pp * d0[0] - g1.ekin * g1.dir[0];
  g2.dir[1] = pp * d0[1] - g1.ekin * g1.dir[1];
  g2.dir[2] = pp * d0[2] - g1.ekin * g1.dir[2];
  unit3(g2.dir);
  o.nsec = 2;
}

static void sample_model(const Material &m, int model, double e, const double d[3], uint32_t s[6],
                         Interaction &o) {
  switch (model) {
    case GXT_MODEL_KLEIN_NISHINA: sample_compton(m, e, d, s, o); break;
    case GXT_MODEL_BETHE_HEITLER: sample_pair(m, e, d, s, o); break;
    case GXT_MODEL_SAUTER_GAVRILA: sample_photo(m, e, d, s, o); break;
    case GXT_MODEL_SELTZER_BERGER_EM:
    case GXT_MODEL_SELTZER_BERGER_EP: sample_brems(m, e, d, s, o); break;
    case GXT_MODEL_MOLLER: sample_ioni(m, false, e, d, s, o); break;
    case GXT_MODEL_BHABHA: sample_ioni(m, true, e, d, s, o); break;
    default: sample_annihil(m, e, d, s, o); break;
  }
}

static inline int model_of(int species, int p) {
  static const int tab[3][3] = {{GXT_MODEL_KLEIN_NISHINA, GXT_MODEL_BETHE_HEITLER, GXT_MODEL_SAUTER_GAVRILA},
                                {GXT_MODEL_SELTZER_BERGER_EM, GXT_MODEL_MOLLER, -1},
                                {GXT_MODEL_SELTZER_BERGER_EP, GXT_MODEL_BHABHA, GXT_MODEL_ANNIHIL}};
  return tab[species][p];
}

// Continuous-loss step limit (ProxyIonization.hpp:108-128) with the
// production-cut length for finR and Geant4's (2 - finR/range) form
// (the reference writes 2*finR/range, which exceeds the range near finR).
static double along_limit(const Material &m, bool pos, double e, double mine) {
  double range = pvec_value(m.tab[pos ? GXT_TAB_RANGE_EP : GXT_TAB_RANGE_EM], e);
  if (e < mine) range *= e / mine;
  const double finR = std::min(10.0, m.cut_length);
  return range > finR ? range * 0.2 + finR * 0.8 * (2.0 - finR / range) : range;
}

// ---------------------------------------------------------------- multiple scattering
// e-/e+ multiple scattering on the first transport mean free path lambda1 of
// the reference's LambdaMod1.msc tables (stored as E^2 / lambda1, the table
// convention of Geant4's G4VMscModel).  The reference reserves the process
// (ProxyProcessIndex.cuh:27) but has no model; this one is:
//  - step limit max(0.04 max(range, lambda1), 0.6 safety, 1e-5 mm) (the
//    "UseSafety" form of Geant4's Urban model);
//  - geometric path z = lambda1 (1 - exp(-t/lambda1)) for a true path t (Lewis'
//    mean longitudinal displacement, energy loss over the step neglected), and
//    its inverse when geometry ends the step;
//  - deflection: x = 1 - cos(theta) from p(x) ~ 1/(x + a)^2 on [0, 2] (the
//    screened-Rutherford form) with a set so that <cos theta> = exp(-t/lambda1)
//    exactly (Lewis' first moment), azimuth uniform; no lateral displacement.
constexpr double kMscFacRange = 0.04, kMscFacSafety = 0.6, kMscTmin = 1.0e-5;
constexpr double kMscNone = 1.0e30;  // lambda1 where a material has no msc table

static double msc_lambda1(const Material &m, bool pos, double e) {
  const PVec &p = m.tab[pos ? GXT_TAB_MSC_EP : GXT_TAB_MSC_EM];
  if (p.x.empty()) return kMscNone;
  const double v = pvec_value(p, e);
  return v > 0.0 ? e * e / v : kMscNone;
}

// (expm1 and log1p keep full precision for short steps)
static double msc_geom(double t, double lam) { return -lam * expm1_cw(-t / lam); }

// true path of a geometric path z, at most tmax
static double msc_true(double z, double lam, double tmax) {
  const double x = z / lam;
  if (x >= 1.0) return tmax;
  return std::min(-lam * log1p_cw(-x), tmax);
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
static double msc_u(double tau) {
  const double q = exp_cw(-tau);
  if (q >= 0.05) {
    const double mean = -expm1_cw(-tau);
    double a = mean < 0.5 ? mean / std::max(log_cw(2.0 / mean) - 1.0, 0.5) : 2.0 / (3.0 * q);
    for (int k = 0; k < 12; ++k) {
      const double L = log1p_cw(2.0 / a);
      const double f = 0.5 * a * (a + 2.0) * L - a - mean;
      const double df = a * ((a + 1.0) * L - 2.0);
      const double du = std::min(std::max(-f / df, -2.0), 2.0);
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
    h *= u;  // h = sum c_j u^j, dh = sum j c_j u^(j-1)
    u = std::min(std::max(u - (h - q) / dh, 0.0), 0.2);
  }
  return u;
}

// Deflection after a true path of tau transport mean free paths (two draws):
// x = 2 r1 / (1 + u (1 - r1)) inverts the distribution's CDF.
static void msc_deflect(double tau, double dir[3], uint32_t rng[6]) {
  const double r1 = mrg_uniform(rng);
  const double r2 = mrg_uniform(rng);
  const double u = msc_u(tau);
  const double x = 2.0 * r1 / (1.0 + u * (1.0 - r1));
  const double cost = 1.0 - x;
  const double sint = std::sqrt(std::max(x * (2.0 - x), 0.0));
  const double phi = kTwoPi * r2;
  double uu = sint * cos_cw(phi), v = sint * sin_cw(phi), w = cost;
  rotate_to_lab(uu, v, w, dir[0], dir[1], dir[2]);
  dir[0] = uu;
  dir[1] = v;
  dir[2] = w;
  unit3(dir);
}

// Continuous energy loss (ProxyIonization.hpp:130-168) with the tracking cut
// as minimum energy and the binary-search inverse range.
static double along_doit(const Material &m, bool pos, double e, double step, double mine, double &eloss) {
  const double range = pvec_value(m.tab[pos ? GXT_TAB_RANGE_EP : GXT_TAB_RANGE_EM], e);
  if (step >= range || e <= mine) {
    eloss = e;
    return 0.0;
  }
  double dedx = pvec_value(m.tab[pos ? GXT_TAB_DEDX_EP : GXT_TAB_DEDX_EM], e);
  double loss = step * dedx;
  if (loss > e * 0.01) loss = e - pvec_inv_value(m.tab[pos ? GXT_TAB_INVRANGE_EP : GXT_TAB_INVRANGE_EM], range - step);
  double fin = e - loss;
  if (fin < mine) {
    loss += fin;
    fin = 0.0;
  }
  eloss = loss < 0.0 ? 0.0 : loss;
  return fin;
}

// ---------------------------------------------------------------- transport
struct Track {
  int species, event, vol, nsteps;
  double e;
  double pos[3], dir[3];
  double nil[3];
  uint32_t rng[6];
  uint64_t uid;
};

struct EventTally {
  double edep = 0, visible = 0, leak = 0, clen = 0;
  int64_t nsec[3] = {0, 0, 0};
  int64_t nsteps = 0, npairs = 0, nannihil = 0;
};

struct Worker {
  std::vector<double> ez, er, ev;
  std::vector<int64_t> spectra;
  std::vector<double> trace;
  int64_t steps = 0, killed = 0, maxlive = 0;
};

enum { kPhysics = 0, kBoundary = 1, kEscaped = 2, kLimit = 3 };

}  // namespace gxref

using namespace gxref;

struct gxref_nav {
  struct Touchable {
    double c[3], h[3];
    int parent, volume;
    std::vector<int> daughters;
  };
  std::vector<Touchable> t;
};

namespace gxref {
namespace nav {

constexpr double kTol = 1e-9;

static bool inside(const gxref_nav::Touchable &b, const double p[3], const double v[3]) {
  for (int k = 0; k < 3; ++k) {
    const double x = p[k] - b.c[k];
    if (std::fabs(x) > b.h[k] + kTol) return false;                  // outside
    if (std::fabs(x) >= b.h[k] - kTol && !(x * v[k] < 0.0)) return false;  // on the face, not entering
  }
  return true;
}

static int down(const gxref_nav &g, int cur, const double p[3], const double v[3]) {
  for (int d : g.t[cur].daughters)
    if (inside(g.t[d], p, v)) return down(g, d, p, v);
  return cur;
}

static int relocate(const gxref_nav &g, int cur, const double p[3], const double v[3]) {
  while (cur >= 0 && !inside(g.t[cur], p, v)) cur = g.t[cur].parent;
  return cur < 0 ? -1 : down(g, cur, p, v);
}

// UnplacedBox DistanceToIn: entry into the slab intersection, if the ray
// runs more than the tolerance inside the box.
static double to_in(const gxref_nav::Touchable &b, const double p[3], const double v[3]) {
  double enter = -kHuge, leave = kHuge;
  for (int k = 0; k < 3; ++k) {
    const double x = p[k] - b.c[k];
    if (v[k] == 0.0) {
      if (std::fabs(x) >= b.h[k] - kTol) return kHuge;  // parallel, outside the slab or on its face
      continue;
    }
    const double a = (-b.h[k] - x) / v[k], c = (b.h[k] - x) / v[k];
    enter = std::max(enter, std::min(a, c));
    leave = std::min(leave, std::max(a, c));
  }
  enter = std::max(enter, 0.0);
  return leave > enter + kTol ? enter : kHuge;
}

// UnplacedBox DistanceToOut from inside.
static double to_out(const gxref_nav::Touchable &b, const double p[3], const double v[3]) {
  double s = kHuge;
  for (int k = 0; k < 3; ++k) {
    if (v[k] == 0.0) continue;
    const double x = p[k] - b.c[k];
    s = std::min(s, ((v[k] > 0.0 ? b.h[k] : -b.h[k]) - x) / v[k]);
  }
  return std::max(s, 0.0);
}

static double safety(const gxref_nav &g, int cur, const double p[3]) {
  const gxref_nav::Touchable &m = g.t[cur];
  double s = kHuge;
  for (int k = 0; k < 3; ++k) s = std::min(s, m.h[k] - std::fabs(p[k] - m.c[k]));
  double q2 = kHuge;
  for (int d : m.daughters) {
    double q = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double e = std::fabs(p[k] - g.t[d].c[k]) - g.t[d].h[k];
      if (e > 0.0) q += e * e;
    }
    q2 = std::min(q2, q);
  }
  if (q2 < kHuge) s = std::min(s, std::sqrt(q2));
  return std::max(s, 0.0);
}

}  // namespace nav
}  // namespace gxref

struct gxref_ctx {
  gxt_config cfg;
  std::string data;
  Physics phys;
  Geometry geo;
  double bfield[3], bmag;
  FieldMap fmap;
  Field field;          // map or uniform field seen by e-/e+ on the general path
  bool general = false; // e-/e+ propagate through propagate_general (a map, or RK45)
  bool msc = false;     // e-/e+ multiple scattering
  std::unique_ptr<gxref_nav> nav;  // tree form of the geometry
  uint32_t seed[6];
  std::vector<EventTally> events;
  Worker tot;
};

namespace gxref {

// Tally cell of a track's volume.
static int cell_of(const gxref_ctx &c, int vol) { return c.geo.tree ? c.geo.cell[vol] : vol; }

// Volume of a starting point, -1 outside the setup.
static int geo_locate(const gxref_ctx &c, const double p[3], const double d[3]) {
  if (!c.geo.tree) return locate(c.geo, p, d);
  return nav::inside(c.nav->t[0], p, d) ? nav::down(*c.nav, 0, p, d) : -1;
}

static void deposit(const gxref_ctx &c, Worker &w, EventTally &ev, double e, const double p[3], int vol) {
  if (e <= 0.0) return;
  vol = cell_of(c, vol);
  ev.edep += e;
  if (c.geo.visible[vol]) ev.visible += e;
  w.ev[vol] += e;
  int iz = (int)std::floor(p[2] / c.cfg.zbin_mm);
  iz = std::min(std::max(iz, 0), c.cfg.n_zbins - 1);
  w.ez[iz] += e;
  const double r = std::sqrt(p[0] * p[0] + p[1] * p[1]);
  int ir = (int)(r / c.cfg.rbin_mm);
  ir = std::min(std::max(ir, 0), c.cfg.n_rbins - 1);
  w.er[ir] += e;
}

static void push_secondary(const gxref_ctx &c, Worker &w, EventTally &ev, std::vector<Track> &stack,
                           const Track &parent, uint32_t s[6], int species, double ekin, const double d[3]);

// Positron at rest: two back-to-back 511 keV photons, isotropic (two draws
// from the stream s), each pushed like any secondary.
static void annihilate_at_rest(const gxref_ctx &c, Worker &w, EventTally &ev, std::vector<Track> &stack,
                               const Track &t, uint32_t s[6]) {
  const double cost = 2.0 * mrg_uniform(s) - 1.0;
  const double sint = std::sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = kTwoPi * mrg_uniform(s);
  const double d1[3] = {sint * cos_cw(phi), sint * sin_cw(phi), cost};
  const double d2[3] = {-d1[0], -d1[1], -d1[2]};
  ev.nannihil++;
  push_secondary(c, w, ev, stack, t, s, 0, kMe, d1);
  push_secondary(c, w, ev, stack, t, s, 0, kMe, d2);
}

static void push_secondary(const gxref_ctx &c, Worker &w, EventTally &ev, std::vector<Track> &stack,
                           const Track &parent, uint32_t s[6], int species, double ekin, const double d[3]) {
  if (ekin < c.phys.tracking_cut) {
    deposit(c, w, ev, ekin, parent.pos, parent.vol);
    if (species == 2) annihilate_at_rest(c, w, ev, stack, parent, s);
    return;
  }
  Track t;
  t.species = species;
  t.event = parent.event;
  t.vol = parent.vol;
  t.nsteps = 0;
  t.e = ekin;
  for (int i = 0; i < 3; ++i) {
    t.pos[i] = parent.pos[i];
    t.dir[i] = d[i];
    t.nil[i] = -1.0;
  }
  t.uid = child_substream(s);
  substream_state(c.seed, t.uid, t.rng);
  ev.nsec[species]++;
  double lg = log10_cw(ekin);
  int b = (int)std::floor((lg + 3.0) * 8.0);
  if (b >= 0 && b < 64) w.spectra[species * 64 + b]++;
  stack.push_back(t);
}

static void trace_step(const gxref_ctx &c, Worker &w, const Track &t, int proc, double edep) {
  if (c.cfg.trace_tracks <= 0 || t.event >= c.cfg.trace_tracks || t.nsteps > c.cfg.trace_steps) return;
  double rec[12];
  std::memcpy(&rec[0], &t.uid, 8);
  rec[1] = t.nsteps;
  rec[2] = t.species;
  rec[3] = proc;
  rec[4] = t.e;
  rec[5] = t.pos[0];
  rec[6] = t.pos[1];
  rec[7] = t.pos[2];
  rec[8] = t.dir[0];
  rec[9] = t.dir[1];
  rec[10] = t.dir[2];
  rec[11] = edep;
  w.trace.insert(w.trace.end(), rec, rec + 12);
}

// Straight or helical propagation inside the current volume up to pstep or
// the boundary (FieldPropagationHandler::Propagate/PropagateInVolume,
// FieldPropagationHandler.cpp:87-360; LinearPropagationHandler::Propagate,
// core/src/LinearPropagationHandler.cpp:13-63; driver loop
// proxy/test/test_BasicCpuTransport.cpp:319-336).
// Propagation of an e-/e+ in a general field (a map, or the RK45 stepper):
// the sub-step choice of the uniform case below with the field looked up at
// the start of every sub-step (FieldPropagationHandler::Propagate,
// FieldPropagationHandler.cpp:93-111: step min(pstep or snext, max(SafeLength
// in the local field, safety))); RK45 sub-steps also kept within kRkMaxSub
// Cash-Karp steps.
static int propagate_general(const gxref_ctx &c, Track &t, double pstep, double &step) {
  const Geometry &g = c.geo;
  const Field &F = c.field;
  const double q = t.species == 1 ? -1.0 : 1.0;
  const double mom = std::sqrt(t.e * (t.e + 2.0 * kMe));
  const bool rk = F.stepper == GXT_STEPPER_RK45;
  double ex = 0.0, eu = 0.0;
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = safety(g, t.vol, t.pos);
    double B[3];
    F.at(t.pos, B);
    const double bm = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    const double smax = rk && bm > 0.0 ? kRkMaxSub * GXT_RK_ANGLE * mom / ((-kB2C) * bm) : kHuge;
    if (remaining <= saf && remaining <= smax) {
      field_advance(F, B, bm, q, mom, remaining, t.pos, t.dir, ex, eu);
      step += remaining;
      return kPhysics;
    }
    int face;
    const double slin = dist_out(g, t.vol, t.pos, t.dir, face);
    double sl = kHuge;
    if (bm > 0.0) {
      const double pl = mom * (t.dir[0] * (B[0] / bm) + t.dir[1] * (B[1] / bm) + t.dir[2] * (B[2] / bm));
      double pt2 = mom * mom - pl * pl;
      if (pt2 < 0.0) pt2 = 0.0;
      const double pt = std::sqrt(pt2);
      if (pt > 0.0) sl = 2.0 * std::sqrt(0.1 / (0.299792458 * bm / pt));
    }
    if (slin <= sl && slin <= remaining && slin <= smax) {
      const double s = std::max(slin, 1.0e-6);
      field_advance(F, B, bm, q, mom, s, t.pos, t.dir, ex, eu);
      step += s;
      if (face == 0 || face == 1) {
        const int nv = face == 1 ? t.vol + 1 : t.vol - 1;
        t.pos[2] = g.z[face == 1 ? t.vol + 1 : t.vol];
        if (nv < 0 || nv >= g.nvol() || std::fabs(t.pos[0]) > g.hx || std::fabs(t.pos[1]) > g.hy) return kEscaped;
        t.vol = nv;
        return kBoundary;
      }
      return kEscaped;
    }
    const double s = std::min(std::min(remaining, std::max(sl, saf)), smax);
    field_advance(F, B, bm, q, mom, s, t.pos, t.dir, ex, eu);
    step += s;
    remaining -= s;
    const int v2 = locate(g, t.pos, t.dir);
    if (v2 != t.vol) {
      if (v2 < 0) return kEscaped;
      t.vol = v2;
      return kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

// Tree form: the straight line to pstep or the next boundary (the distance
// to leaving the touchable or entering a daughter, the first daughter on a
// tie; the touchable beyond found by relocation from the daughter entered or
// the touchable left), or in a uniform field the helix in sub-steps chosen as
// in the slab form below from the navigator's safety and straight-line
// distance, a sub-step ending the step when its end point relocates to
// another touchable (FieldPropagationHandler.cpp:93-111 with
// NavigationInterface.hpp:64-98).
static int propagate_tree(const gxref_ctx &c, Track &t, double pstep, double &step) {
  const gxref_nav &g = *c.nav;
  auto boundary = [&](int &target) {
    double s = nav::to_out(g.t[t.vol], t.pos, t.dir);
    target = t.vol;
    for (int d : g.t[t.vol].daughters) {
      const double sd = nav::to_in(g.t[d], t.pos, t.dir);
      if (sd < s) {
        s = sd;
        target = d;
      }
    }
    return s;
  };
  if (t.species == 0 || c.bmag == 0.0) {
    int target;
    const double sn = boundary(target);
    if (pstep < sn) {
      for (int i = 0; i < 3; ++i) t.pos[i] += pstep * t.dir[i];
      step = pstep;
      return kPhysics;
    }
    for (int i = 0; i < 3; ++i) t.pos[i] += sn * t.dir[i];
    step = sn;
    t.vol = nav::relocate(g, target, t.pos, t.dir);
    return t.vol < 0 ? kEscaped : kBoundary;
  }
  const double q = t.species == 1 ? -1.0 : 1.0;
  const double mom = std::sqrt(t.e * (t.e + 2.0 * kMe));
  const double ub[3] = {c.bfield[0] / c.bmag, c.bfield[1] / c.bmag, c.bfield[2] / c.bmag};
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = nav::safety(g, t.vol, t.pos);
    if (remaining <= saf) {
      helix_step(c.bfield, c.bmag, q, mom, remaining, t.pos, t.dir);
      step += remaining;
      return kPhysics;
    }
    int target;
    const double slin = boundary(target);
    const double pl = mom * (t.dir[0] * ub[0] + t.dir[1] * ub[1] + t.dir[2] * ub[2]);
    double pt2 = mom * mom - pl * pl;
    if (pt2 < 0.0) pt2 = 0.0;
    const double pt = std::sqrt(pt2);
    const double sl = pt > 0.0 ? 2.0 * std::sqrt(0.1 / (0.299792458 * c.bmag / pt)) : kHuge;
    const double s = slin <= sl && slin <= remaining ? std::max(slin, 1.0e-6) : std::min(remaining, std::max(sl, saf));
    helix_step(c.bfield, c.bmag, q, mom, s, t.pos, t.dir);
    step += s;
    remaining -= s;
    const int v2 = nav::relocate(g, t.vol, t.pos, t.dir);
    if (v2 != t.vol) {
      t.vol = v2;
      return v2 < 0 ? kEscaped : kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

static int propagate(const gxref_ctx &c, Track &t, double pstep, double &step) {
  const Geometry &g = c.geo;
  const bool charged = t.species != 0;
  if (g.tree) return propagate_tree(c, t, pstep, step);
  if (charged && c.general) return propagate_general(c, t, pstep, step);
  if (!charged || c.bmag == 0.0) {
    int face;
    const double sn = dist_out(g, t.vol, t.pos, t.dir, face);
    if (pstep < sn) {
      for (int i = 0; i < 3; ++i) t.pos[i] += pstep * t.dir[i];
      step = pstep;
      return kPhysics;
    }
    for (int i = 0; i < 3; ++i) t.pos[i] += sn * t.dir[i];
    step = sn;
    if (face == 0 || face == 1) {
      const int nv = face == 1 ? t.vol + 1 : t.vol - 1;
      t.pos[2] = g.z[face == 1 ? t.vol + 1 : t.vol];
      if (nv < 0 || nv >= g.nvol()) return kEscaped;
      t.vol = nv;
      return kBoundary;
    }
    return kEscaped;
  }
  // charged particle in a uniform field
  const double q = t.species == 1 ? -1.0 : 1.0;
  const double mom = std::sqrt(t.e * (t.e + 2.0 * kMe));
  const double ub[3] = {c.bfield[0] / c.bmag, c.bfield[1] / c.bmag, c.bfield[2] / c.bmag};
  double remaining = pstep;
  step = 0.0;
  for (int it = 0; it < 64; ++it) {
    const double saf = safety(g, t.vol, t.pos);
    if (remaining <= saf) {
      helix_step(c.bfield, c.bmag, q, mom, remaining, t.pos, t.dir);
      step += remaining;
      return kPhysics;
    }
    int face;
    const double slin = dist_out(g, t.vol, t.pos, t.dir, face);
    // SafeLength (FieldPropagationHandler.hpp:65-77): 2 sqrt(eps / curvature), eps = 0.1 mm
    const double pl = mom * (t.dir[0] * ub[0] + t.dir[1] * ub[1] + t.dir[2] * ub[2]);
    double pt2 = mom * mom - pl * pl;
    if (pt2 < 0.0) pt2 = 0.0;
    const double pt = std::sqrt(pt2);
    const double sl = pt > 0.0 ? 2.0 * std::sqrt(0.1 / (0.299792458 * c.bmag / pt)) : kHuge;
    if (slin <= sl && slin <= remaining) {
      // The straight line reaches the boundary within the sagitta-limited
      // length: step along the helix by that length and cross the aimed face
      // by construction, the track placed on the plane (the curved path
      // deviates from the chord by less than the 0.1 mm deflection tolerance).
      const double s = std::max(slin, 1.0e-6);
      helix_step(c.bfield, c.bmag, q, mom, s, t.pos, t.dir);
      step += s;
      if (face == 0 || face == 1) {
        const int nv = face == 1 ? t.vol + 1 : t.vol - 1;
        t.pos[2] = g.z[face == 1 ? t.vol + 1 : t.vol];
        if (nv < 0 || nv >= g.nvol() || std::fabs(t.pos[0]) > g.hx || std::fabs(t.pos[1]) > g.hy) return kEscaped;
        t.vol = nv;
        return kBoundary;
      }
      return kEscaped;
    }
    const double s = std::min(remaining, std::max(sl, saf));
    helix_step(c.bfield, c.bmag, q, mom, s, t.pos, t.dir);
    step += s;
    remaining -= s;
    const int v2 = locate(g, t.pos, t.dir);
    if (v2 != t.vol) {
      if (v2 < 0) return kEscaped;
      t.vol = v2;
      return kBoundary;
    }
    if (remaining <= 0.0) return kPhysics;
  }
  return kLimit;
}

// One step of one track (OneStep/InnerStep, proxy/test/test_BasicCpuTransport.cpp:358-497).
static void one_step(gxref_ctx &c, Worker &w, EventTally &ev, std::vector<Track> &stack, Track &t, bool &alive) {
  // tree form: a vacuum volume (slot -1) has no physics; the track flies to
  // the next boundary and keeps its interaction lengths
  const int slot = c.geo.slot[cell_of(c, t.vol)];
  const bool vac = slot < 0;
  const Material &m = c.phys.mat[vac ? 0 : slot];
  const double mine = c.phys.tracking_cut;
  const int np = vac ? 0 : nproc(t.species);
  double xs[3] = {0, 0, 0};
  double best = kHuge;
  int sel = -1;
  // PostStep GPIL (ProxyEmProcess.hpp:122-151)
  for (int p = 0; p < np; ++p) {
    xs[p] = macro_xs(m, t.species, p, t.e);
    if (t.nil[p] <= 0.0) t.nil[p] = -log_cw(mrg_uniform(t.rng));
    const double sp = xs[p] > 0.0 ? t.nil[p] / xs[p] : kHuge;
    if (sp < best) {
      best = sp;
      sel = p;
    }
  }
  const bool charged = t.species != 0;
  if (charged && !vac) {
    const double al = along_limit(m, t.species == 2, t.e, mine);
    if (al < best) {
      best = al;
      sel = -1;
    }
  }
  // multiple scattering: its step limit, then the geometric length of the
  // true path (the step is the true path for interaction lengths and loss)
  double lam1 = kMscNone;
  if (charged && c.msc && !vac) {
    lam1 = msc_lambda1(m, t.species == 2, t.e);
    if (lam1 < kMscNone) {
      const double range = pvec_value(m.tab[t.species == 2 ? GXT_TAB_RANGE_EP : GXT_TAB_RANGE_EM], t.e);
      const double saf = c.geo.tree ? nav::safety(*c.nav, t.vol, t.pos) : safety(c.geo, t.vol, t.pos);
      const double tlim = std::max(std::max(kMscFacRange * std::max(range, lam1), kMscFacSafety * saf), kMscTmin);
      if (tlim < best) {
        best = tlim;
        sel = -1;
      }
    }
  }
  const bool msc_on = lam1 < kMscNone;
  const double pre[3] = {t.pos[0], t.pos[1], t.pos[2]};
  const int prevol = t.vol;
  double gstep = 0.0;
  const int status = propagate(c, t, msc_on ? msc_geom(best, lam1) : best, gstep);
  const double step = msc_on ? (status == kPhysics ? best : msc_true(gstep, lam1, best)) : gstep;
  for (int p = 0; p < np; ++p)
    if (xs[p] > 0.0) t.nil[p] -= step * xs[p];
  t.nsteps++;
  ev.nsteps++;
  w.steps++;
  double edep = 0.0;
  if (charged && !vac) {
    ev.clen += step;
    {
      double loss = 0.0;
      t.e = along_doit(m, t.species == 2, t.e, step, mine, loss);
      const double mid[3] = {0.5 * (pre[0] + t.pos[0]), 0.5 * (pre[1] + t.pos[1]), 0.5 * (pre[2] + t.pos[2])};
      deposit(c, w, ev, loss, mid, prevol);
      edep += loss;
    }
    if (msc_on && status != kEscaped && t.e > 0.0 && step > 0.0) msc_deflect(step / lam1, t.dir, t.rng);
  }
  int tproc = status == kPhysics ? sel : (status == kBoundary ? -2 : (status == kEscaped ? -3 : -1));
  if (status == kPhysics && sel < 0) tproc = -1;
  if (status == kEscaped) {
    ev.leak += t.e;
    trace_step(c, w, t, tproc, edep);
    alive = false;
    return;
  }
  if (t.e <= 0.0) {
    trace_step(c, w, t, tproc, edep);
    if (t.species == 2) annihilate_at_rest(c, w, ev, stack, t, t.rng);
    alive = false;
    return;
  }
  if (status == kPhysics && sel >= 0) {
    t.nil[sel] = -1.0;
    Interaction o;
    sample_model(m, model_of(t.species, sel), t.e, t.dir, t.rng, o);
    if (sel == 1 && t.species == 0 && o.nsec == 2) ev.npairs++;
    if (t.species == 2 && sel == 2) ev.nannihil++;
    deposit(c, w, ev, o.edep, t.pos, t.vol);
    edep += o.edep;
    t.e = o.ekin;
    t.dir[0] = o.dir[0];
    t.dir[1] = o.dir[1];
    t.dir[2] = o.dir[2];
    trace_step(c, w, t, tproc, edep);
    for (int k = 0; k < o.nsec; ++k)
      push_secondary(c, w, ev, stack, t, t.rng, o.sec[k].species, o.sec[k].ekin, o.sec[k].dir);
    if (t.e <= 0.0) {
      alive = false;
      return;
    }
  } else {
    trace_step(c, w, t, tproc, edep);
  }
  if (t.e < mine) {
    deposit(c, w, ev, t.e, t.pos, t.vol);
    if (t.species == 2) annihilate_at_rest(c, w, ev, stack, t, t.rng);
    alive = false;
    return;
  }
  if (t.nsteps >= c.cfg.max_steps) {
    deposit(c, w, ev, t.e, t.pos, t.vol);
    w.killed++;
    alive = false;
  }
}

// Transport every track of an event's stack, and all it creates, to the end
// (depth first).
static void drain(gxref_ctx &c, Worker &w, EventTally &ev, std::vector<Track> &stack) {
  while (!stack.empty()) {
    Track cur = stack.back();
    stack.pop_back();
    bool alive = true;
    while (alive) one_step(c, w, ev, stack, cur, alive);
    if ((int64_t)stack.size() > w.maxlive) w.maxlive = (int64_t)stack.size();
  }
}

// The track a primary starts as: no interaction lengths yet, its substream's
// initial state, located from its position and direction (vol < 0: outside).
static Track primary_track(const gxref_ctx &c, const gxt_primary &p) {
  Track t;
  t.species = p.species;
  t.event = p.event;
  t.nsteps = 0;
  t.e = p.ekin;
  for (int i = 0; i < 3; ++i) {
    t.pos[i] = p.pos[i];
    t.dir[i] = p.dir[i];
    t.nil[i] = -1.0;
  }
  t.uid = p.substream;
  substream_state(c.seed, t.uid, t.rng);
  t.vol = geo_locate(c, t.pos, t.dir);
  return t;
}

static void run_event(gxref_ctx &c, Worker &w, const gxt_primary &p) {
  EventTally &ev = c.events[p.event];
  std::vector<Track> stack;
  const Track t = primary_track(c, p);
  if (t.vol < 0) {
    ev.leak += t.e;
    return;
  }
  stack.push_back(t);
  drain(c, w, ev, stack);
}

// A basket track in flight (gxt_track) as the loop's track: its own
// interaction lengths, RNG state, step count and volume (located when -1).
static Track basket_track(const gxref_ctx &c, const gxt_track &in) {
  Track t;
  t.species = in.species;
  t.event = in.event;
  t.nsteps = in.nsteps;
  t.e = in.ekin;
  for (int i = 0; i < 3; ++i) {
    t.pos[i] = in.pos[i];
    t.dir[i] = in.dir[i];
    t.nil[i] = in.nil[i];
  }
  for (int k = 0; k < 6; ++k) t.rng[k] = in.rng[k];
  t.uid = in.substream;
  t.vol = in.volume >= 0 ? in.volume : geo_locate(c, t.pos, t.dir);
  return t;
}

static gxt_track to_basket(const Track &t) {
  gxt_track o;
  o.species = t.species;
  o.event = t.event;
  o.ekin = t.e;
  for (int i = 0; i < 3; ++i) {
    o.pos[i] = t.pos[i];
    o.dir[i] = t.dir[i];
    o.nil[i] = t.nil[i];
  }
  for (int k = 0; k < 6; ++k) o.rng[k] = t.rng[k];
  o.volume = t.vol;
  o.nsteps = t.nsteps;
  o.substream = t.uid;
  return o;
}

}  // namespace gxref

namespace gxref {

// The tree form (GXT_GEOM_TREE): the navigator, and the tally cells of the
// touchables (with a material: in touchable order, slot 0 absorber, 1 gap,
// visible; vacuum: the last cell, slot -1).  At most 127 cells with a material.
static int build_tree(const gxt_config &cfg, gxref_ctx &c) {
  if (!cfg.tree_volumes || cfg.n_tree_volumes <= 0)
    return fail(GXT_EINVAL, "tree geometry: bad volume or placement arrays");
  for (int i = 0; i < cfg.n_tree_volumes; ++i) {
    const int m = cfg.tree_volumes[i].material;
    if (m != -1 && m != cfg.absorber && m != cfg.gap)
      return fail(GXT_EINVAL, "tree geometry: a volume's material is neither the absorber, the gap nor vacuum");
  }
  if (cfg.field_map || cfg.field_stepper != GXT_STEPPER_HELIX)
    return fail(GXT_EINVAL, "tree geometry: uniform field only (no field map, HELIX stepper)");
  gxref_nav *g = nullptr;
  const int rc = gxref_nav_create(cfg.tree_volumes, cfg.n_tree_volumes, cfg.tree_placements, cfg.n_tree_placements,
                                  cfg.tree_world, &g);
  if (rc) return rc;
  c.nav.reset(g);
  Geometry &geo = c.geo;
  geo.tree = true;
  geo.cell.assign(g->t.size(), -1);
  int k = 0;
  for (size_t t = 0; t < g->t.size(); ++t) {
    const int m = cfg.tree_volumes[g->t[t].volume].material;
    if (m < 0) continue;
    if (k >= 127) return fail(GXT_EINVAL, "tree geometry: more than 127 touchables with a material");
    geo.cell[t] = k++;
    geo.slot.push_back(m == cfg.gap ? 1 : 0);
    geo.visible.push_back(m == cfg.gap ? 1 : 0);
  }
  for (size_t t = 0; t < g->t.size(); ++t)
    if (geo.cell[t] < 0) geo.cell[t] = k;
  geo.slot.push_back(-1);
  geo.visible.push_back(0);
  c.cfg.tree_volumes = nullptr;  // the caller's arrays are not kept
  c.cfg.tree_placements = nullptr;
  return GXT_OK;
}

}  // namespace gxref

// ---------------------------------------------------------------- C API
extern "C" {

const char *gxref_last_error(void) { return g_err.c_str(); }

int gxref_create(const gxt_config *cfg, gxref_ctx **out) {
  if (!cfg || !out) return fail(GXT_EINVAL, "null argument");
  auto *c = new gxref_ctx();
  c->cfg = *cfg;
  c->data = cfg->data_dir ? cfg->data_dir : "";
  for (int i = 0; i < 6; ++i) c->seed[i] = cfg->seed;
  c->phys.tracking_cut = cfg->tracking_cut;
  build_geometry(*cfg, c->geo);
  if (cfg->geometry == GXT_GEOM_TREE) {
    const int rc = build_tree(*cfg, *c);
    if (rc) {
      delete c;
      return rc;
    }
  }
  std::vector<int> mats;
  if (cfg->geometry == GXT_GEOM_BOX)
    mats = {cfg->material};
  else
    mats = {cfg->absorber, cfg->gap};
  for (int mid : mats) {
    Material m;
    const int rc = build_material(c->data, mid, cfg->lar_density, m);
    if (rc) {
      delete c;
      return rc;
    }
    c->phys.mat.push_back(std::move(m));
  }
  c->bfield[0] = cfg->field[0];
  c->bfield[1] = cfg->field[1];
  c->bfield[2] = cfg->field[2];
  c->bmag = std::sqrt(cfg->field[0] * cfg->field[0] + cfg->field[1] * cfg->field[1] + cfg->field[2] * cfg->field[2]);
  if (cfg->field_stepper != GXT_STEPPER_HELIX && cfg->field_stepper != GXT_STEPPER_RK45) {
    delete c;
    return fail(GXT_EINVAL, "bad field_stepper");
  }
  if (cfg->field_map) {
    const int rc = read_field_map(cfg->field_map, c->fmap);
    if (rc) {
      delete c;
      return rc;
    }
    c->field.map = &c->fmap;
  }
  c->cfg.field_map = nullptr;
  for (int k = 0; k < 3; ++k) {
    c->field.b[k] = cfg->field[k];
    c->field.origin[k] = cfg->field_origin[k];
  }
  c->field.stepper = cfg->field_stepper;
  c->general = cfg->field_map != nullptr || cfg->field_stepper == GXT_STEPPER_RK45;
  c->msc = cfg->msc != 0;
  if (c->msc)
    for (const Material &mm : c->phys.mat)
      if (mm.tab[GXT_TAB_MSC_EM].x.empty() || mm.tab[GXT_TAB_MSC_EP].x.empty()) {
        delete c;
        return fail(GXT_EINVAL, "multiple scattering needs the LambdaMod1.msc tables (none for PbWO4)");
      }
  c->events.assign(std::max(cfg->max_events, 1), EventTally());
  c->tot.ez.assign(cfg->n_zbins, 0.0);
  c->tot.er.assign(cfg->n_rbins, 0.0);
  c->tot.ev.assign(c->geo.nvol(), 0.0);
  c->tot.spectra.assign(3 * 64, 0);
  *out = c;
  return GXT_OK;
}

int gxref_destroy(gxref_ctx *c) {
  delete c;
  return GXT_OK;
}

// ProxyHepEvtGenerator::GenerateOneEvent (proxy/src/ProxyHepEvtGenerator.cpp:48-132),
// repeated per event: read NHEP entries, mark the daughters (ISTHEP > 0) of
// the entries with daughters as -1 (:76-90), and turn the marked entries
// without daughters into primaries (:97-120).  Two fixes: stable top-level
// entries (ISTHEP = 1, no daughters, nobody's daughter) are primaries too,
// as in Geant4's G4HEPEvtInterface; the mass column is used as the mass
// (:109 squares it).  Species other than gamma/e-/e+ are counted, not made.
int gxref_read_hepevt(const char *path, int max_events, const double vertex[3], uint64_t first, gxt_primary *out,
                      int capacity, int *n_primaries, int *n_events, int *n_unsupported) {
  if (!path || !n_primaries || !n_events || !n_unsupported) return fail(GXT_EINVAL, "bad arguments");
  std::ifstream f(path);
  if (!f) return fail(GXT_EDATA, "cannot open file");
  struct Hep {
    int ISTHEP, IDHEP, JDAHEP1, JDAHEP2;
    double PHEP1, PHEP2, PHEP3, PHEP4;
  };
  int nprim = 0, nev = 0, nunsup = 0;
  while (max_events < 0 || nev < max_events) {
    int NHEP = 0;
    if (!(f >> NHEP)) break;
    if (NHEP < 0) return fail(GXT_EDATA, "negative entry count");
    std::vector<Hep> hepg(NHEP);
    for (int i = 0; i < NHEP; ++i)
      if (!(f >> hepg[i].ISTHEP >> hepg[i].IDHEP >> hepg[i].JDAHEP1 >> hepg[i].JDAHEP2 >> hepg[i].PHEP1 >>
            hepg[i].PHEP2 >> hepg[i].PHEP3 >> hepg[i].PHEP4))
        return fail(GXT_EDATA, "truncated event");
    std::vector<char> top(NHEP);
    for (int i = 0; i < NHEP; ++i) top[i] = hepg[i].ISTHEP == 1;
    for (int i = 0; i < NHEP; ++i) {
      if (hepg[i].JDAHEP1 > 0) {
        const int jda1 = hepg[i].JDAHEP1 - 1, jda2 = hepg[i].JDAHEP2 - 1;
        if (jda2 < jda1 || jda2 >= NHEP) return fail(GXT_EDATA, "daughter range outside the event");
        for (int j = jda1; j <= jda2; ++j)
          if (hepg[j].ISTHEP > 0) {
            hepg[j].ISTHEP = -1;
            top[j] = 0;
          }
      }
    }
    for (int i = 0; i < NHEP; ++i) {
      if (!((hepg[i].ISTHEP == -1 || top[i]) && hepg[i].JDAHEP1 == 0)) continue;
      const double px = hepg[i].PHEP1, py = hepg[i].PHEP2, pz = hepg[i].PHEP3;
      const double p = std::sqrt(px * px + py * py + pz * pz);
      const int id = hepg[i].IDHEP;
      const int species = id == 22 ? 0 : (id == 11 ? 1 : (id == -11 ? 2 : -1));
      if (species < 0 || !(p > 0.0)) {
        ++nunsup;
        continue;
      }
      if (out) {
        if (nprim >= capacity) {
          *n_primaries = nprim;
          return fail(GXT_ECAPACITY, "more primaries than capacity");
        }
        gxt_primary &q = out[nprim];
        std::memset(&q, 0, sizeof(q));
        const double pm = p * 1000.0, mass = hepg[i].PHEP4 * 1000.0;
        q.species = species;
        q.event = nev;
        q.ekin = pm * pm / (std::sqrt(pm * pm + mass * mass) + mass);
        q.dir[0] = px / p;
        q.dir[1] = py / p;
        q.dir[2] = pz / p;
        for (int k = 0; k < 3; ++k) q.pos[k] = vertex ? vertex[k] : 0.0;
        q.substream = first + (uint64_t)nprim;
      }
      ++nprim;
    }
    ++nev;
  }
  *n_primaries = nprim;
  *n_events = nev;
  *n_unsupported = nunsup;
  return GXT_OK;
}

// ProxyEgammaGun::GenerateOneEvent (proxy/src/ProxyEgammaGun.cpp:43-85): per
// particle a gamma with probability fGammaFraction else an electron, p, eta
// and phi uniform, theta = 2 atan(exp(-eta)), Ekin = p^2 / (sqrt(p^2+m^2)+m);
// the draws come from substream n_total + first + i.
int gxref_egamma_gun(uint32_t seed, int n_events, int per_event, double gamma_fraction, double pmin, double pmax,
                     double eta_min, double eta_max, double phi_min, double phi_max, const double vertex[3],
                     uint64_t first, uint64_t n_total, gxt_primary *out) {
  if (!out || n_events < 0 || per_event <= 0 || !(pmax >= pmin) || !(pmin > 0.0)) return fail(GXT_EINVAL, "bad gun");
  uint32_t s6[6];
  for (int i = 0; i < 6; ++i) s6[i] = seed;
  for (int i = 0; i < n_events * per_event; ++i) {
    uint32_t st[6];
    substream_state(s6, n_total + first + (uint64_t)i, st);
    const int pid = mrg_uniform(st) < gamma_fraction ? 22 : 11;
    const double charge = pid == 22 ? 0.0 : -1.0;
    const double p = pmin + (pmax - pmin) * mrg_uniform(st);
    const double mass = kMe * charge * charge;
    const double eta = eta_min + (eta_max - eta_min) * mrg_uniform(st);
    const double phi = phi_min + (phi_max - phi_min) * mrg_uniform(st);
    const double theta = 2.0 * std::atan(std::exp(-eta));
    gxt_primary &q = out[i];
    std::memset(&q, 0, sizeof(q));
    q.species = pid == 22 ? 0 : 1;
    q.event = i / per_event;
    q.ekin = p * p / (std::sqrt(p * p + mass * mass) + mass);
    q.dir[0] = std::sin(theta) * std::cos(phi);
    q.dir[1] = std::sin(theta) * std::sin(phi);
    q.dir[2] = std::cos(theta);
    for (int k = 0; k < 3; ++k) q.pos[k] = vertex ? vertex[k] : 0.0;
    q.substream = first + (uint64_t)i;
  }
  return GXT_OK;
}

int gxref_make_primaries(uint32_t seed, int species, double emin, double emax, double spot_mm, int n,
                         uint64_t first, uint64_t n_total, gxt_primary *out) {
  if (n < 0 || !out) return fail(GXT_EINVAL, "bad primaries request");
  uint32_t s6[6];
  for (int i = 0; i < 6; ++i) s6[i] = seed;
  for (int i = 0; i < n; ++i) {
    gxt_primary &p = out[i];
    p.species = species;
    p.event = i;
    p.substream = first + (uint64_t)i;
    p.ekin = emin;
    p.pos[0] = p.pos[1] = p.pos[2] = 0.0;
    p.dir[0] = p.dir[1] = 0.0;
    p.dir[2] = 1.0;
    if (emax > emin || spot_mm > 0.0) {
      uint32_t st[6];
      substream_state(s6, n_total + first + (uint64_t)i, st);
      const double u0 = mrg_uniform(st), u1 = mrg_uniform(st), u2 = mrg_uniform(st);
      p.ekin = emin + (emax - emin) * u0;
      p.pos[0] = (u1 - 0.5) * spot_mm;
      p.pos[1] = (u2 - 0.5) * spot_mm;
    }
  }
  return GXT_OK;
}

int gxref_run(gxref_ctx *c, const gxt_primary *prim, int n, int nthreads) {
  if (!c || (n > 0 && !prim)) return fail(GXT_EINVAL, "null argument");
  for (int i = 0; i < n; ++i)
    if (prim[i].event < 0 || prim[i].event >= (int)c->events.size()) return fail(GXT_EINVAL, "event index out of range");
  if (nthreads < 1) nthreads = 1;
  // events are independent; group primaries by event so one worker owns an event
  std::vector<std::vector<int>> byev(c->events.size());
  for (int i = 0; i < n; ++i) byev[prim[i].event].push_back(i);
  std::vector<int> order;
  for (size_t e = 0; e < byev.size(); ++e)
    if (!byev[e].empty()) order.push_back((int)e);
  std::atomic<size_t> next(0);
  std::vector<Worker> ws(nthreads);
  for (auto &w : ws) {
    w.ez.assign(c->cfg.n_zbins, 0.0);
    w.er.assign(c->cfg.n_rbins, 0.0);
    w.ev.assign(c->geo.nvol(), 0.0);
    w.spectra.assign(3 * 64, 0);
  }
  auto work = [&](int id) {
    Worker &w = ws[id];
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= order.size()) break;
      for (int i : byev[order[k]]) run_event(*c, w, prim[i]);
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
    for (auto &t : th) t.join();
  }
  for (auto &w : ws) {
    for (size_t i = 0; i < w.ez.size(); ++i) c->tot.ez[i] += w.ez[i];
    for (size_t i = 0; i < w.er.size(); ++i) c->tot.er[i] += w.er[i];
    for (size_t i = 0; i < w.ev.size(); ++i) c->tot.ev[i] += w.ev[i];
    for (size_t i = 0; i < w.spectra.size(); ++i) c->tot.spectra[i] += w.spectra[i];
    c->tot.trace.insert(c->tot.trace.end(), w.trace.begin(), w.trace.end());
    c->tot.steps += w.steps;
    c->tot.killed += w.killed;
    c->tot.maxlive = std::max(c->tot.maxlive, w.maxlive);
  }
  return GXT_OK;
}

// Continue a basket of tracks in flight (gxt_submit_tracks): the tracks of
// an event go on one stack, which is transported to the end.
int gxref_run_tracks(gxref_ctx *c, const gxt_track *trk, int n, int nthreads) {
  if (!c || (n > 0 && !trk)) return fail(GXT_EINVAL, "null argument");
  for (int i = 0; i < n; ++i)
    if (trk[i].event < 0 || trk[i].event >= (int)c->events.size()) return fail(GXT_EINVAL, "event index out of range");
  if (nthreads < 1) nthreads = 1;
  std::vector<std::vector<int>> byev(c->events.size());
  for (int i = 0; i < n; ++i) byev[trk[i].event].push_back(i);
  std::vector<int> order;
  for (size_t e = 0; e < byev.size(); ++e)
    if (!byev[e].empty()) order.push_back((int)e);
  std::atomic<size_t> next(0);
  std::vector<Worker> ws(nthreads);
  for (auto &w : ws) {
    w.ez.assign(c->cfg.n_zbins, 0.0);
    w.er.assign(c->cfg.n_rbins, 0.0);
    w.ev.assign(c->geo.nvol(), 0.0);
    w.spectra.assign(3 * 64, 0);
  }
  auto work = [&](int id) {
    Worker &w = ws[id];
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= order.size()) break;
      EventTally &ev = c->events[order[k]];
      std::vector<Track> stack;
      for (int i : byev[order[k]]) {
        const Track t = basket_track(*c, trk[i]);
        if (t.vol < 0)
          ev.leak += t.e;
        else
          stack.push_back(t);
      }
      drain(*c, w, ev, stack);
    }
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; ++i) th.emplace_back(work, i);
    for (auto &t : th) t.join();
  }
  for (auto &w : ws) {
    for (size_t i = 0; i < w.ez.size(); ++i) c->tot.ez[i] += w.ez[i];
    for (size_t i = 0; i < w.er.size(); ++i) c->tot.er[i] += w.er[i];
    for (size_t i = 0; i < w.ev.size(); ++i) c->tot.ev[i] += w.ev[i];
    for (size_t i = 0; i < w.spectra.size(); ++i) c->tot.spectra[i] += w.spectra[i];
    c->tot.steps += w.steps;
    c->tot.killed += w.killed;
    c->tot.maxlive = std::max(c->tot.maxlive, w.maxlive);
  }
  return GXT_OK;
}

// Run each event for at most steps_per_event track-steps (one thread, its
// primaries in order, each depth first), then hand what is left over as a
// basket: the track in flight, its stack and the primaries not yet started,
// in that order.  The event tallies hold the steps run here.
int gxref_run_partial(gxref_ctx *c, const gxt_primary *prim, int n, int64_t steps_per_event, gxt_track *out,
                      int capacity, int *n_out) {
  if (!c || (n > 0 && !prim) || !n_out || (capacity > 0 && !out) || steps_per_event < 0)
    return fail(GXT_EINVAL, "bad run_partial arguments");
  for (int i = 0; i < n; ++i)
    if (prim[i].event < 0 || prim[i].event >= (int)c->events.size()) return fail(GXT_EINVAL, "event index out of range");
  std::vector<std::vector<int>> byev(c->events.size());
  for (int i = 0; i < n; ++i) byev[prim[i].event].push_back(i);
  Worker w;
  w.ez.assign(c->cfg.n_zbins, 0.0);
  w.er.assign(c->cfg.n_rbins, 0.0);
  w.ev.assign(c->geo.nvol(), 0.0);
  w.spectra.assign(3 * 64, 0);
  std::vector<gxt_track> basket;
  for (size_t e = 0; e < byev.size(); ++e) {
    EventTally &ev = c->events[e];
    int64_t left = steps_per_event;
    std::vector<Track> stack;
    for (size_t k = 0; k < byev[e].size(); ++k) {
      const Track t0 = primary_track(*c, prim[byev[e][k]]);
      if (left <= 0) {
        // not started: handed over as it starts (interaction lengths -1)
        if (t0.vol < 0)
          ev.leak += t0.e;
        else
          basket.push_back(to_basket(t0));
        continue;
      }
      if (t0.vol < 0) {
        ev.leak += t0.e;
        continue;
      }
      stack.push_back(t0);
      while (!stack.empty() && left > 0) {
        Track cur = stack.back();
        stack.pop_back();
        bool alive = true;
        while (alive && left > 0) {
          one_step(*c, w, ev, stack, cur, alive);
          --left;
        }
        if (alive) basket.push_back(to_basket(cur));
      }
      for (auto it = stack.rbegin(); it != stack.rend(); ++it) basket.push_back(to_basket(*it));
      stack.clear();
    }
  }
  for (size_t i = 0; i < w.ez.size(); ++i) c->tot.ez[i] += w.ez[i];
  for (size_t i = 0; i < w.er.size(); ++i) c->tot.er[i] += w.er[i];
  for (size_t i = 0; i < w.ev.size(); ++i) c->tot.ev[i] += w.ev[i];
  for (size_t i = 0; i < w.spectra.size(); ++i) c->tot.spectra[i] += w.spectra[i];
  c->tot.steps += w.steps;
  c->tot.killed += w.killed;
  *n_out = (int)basket.size();
  if ((int)basket.size() > capacity) return fail(GXT_ECAPACITY, "basket larger than capacity");
  std::copy(basket.begin(), basket.end(), out);
  return GXT_OK;
}

int gxref_volume_bins(gxref_ctx *c, int *n_bins, int32_t *cell) {
  if (!c || !n_bins) return fail(GXT_EINVAL, "bad arguments");
  *n_bins = c->geo.nvol();
  if (cell)
    for (size_t t = 0; t < c->geo.cell.size(); ++t) cell[t] = c->geo.cell[t];
  return GXT_OK;
}

int gxref_get_profiles(gxref_ctx *c, double *ez, double *er, double *ev) {
  if (!c) return fail(GXT_EINVAL, "null ctx");
  if (ez) std::copy(c->tot.ez.begin(), c->tot.ez.end(), ez);
  if (er) std::copy(c->tot.er.begin(), c->tot.er.end(), er);
  if (ev) std::copy(c->tot.ev.begin(), c->tot.ev.end(), ev);
  return GXT_OK;
}

int gxref_get_event_tallies(gxref_ctx *c, gxt_event_tally *out, int n) {
  if (!c || !out) return fail(GXT_EINVAL, "null argument");
  for (int i = 0; i < n && i < (int)c->events.size(); ++i) {
    const EventTally &e = c->events[i];
    gxt_event_tally &o = out[i];
    o.edep = e.edep;
    o.visible = e.visible;
    o.leak = e.leak;
    o.charged_length = e.clen;
    for (int k = 0; k < 3; ++k) o.n_secondaries[k] = e.nsec[k];
    o.n_steps = e.nsteps;
    o.n_pairs = e.npairs;
    o.n_annihil = e.nannihil;
  }
  return GXT_OK;
}

int gxref_get_spectra(gxref_ctx *c, int64_t *counts) {
  if (!c || !counts) return fail(GXT_EINVAL, "null argument");
  std::copy(c->tot.spectra.begin(), c->tot.spectra.end(), counts);
  return GXT_OK;
}

int gxref_get_stats(gxref_ctx *c, gxt_stats *o) {
  if (!c || !o) return fail(GXT_EINVAL, "null argument");
  std::memset(o, 0, sizeof(*o));
  o->track_steps = c->tot.steps;
  o->max_live = c->tot.maxlive;
  o->killed_by_limit = c->tot.killed;
  return GXT_OK;
}

int gxref_get_trace(gxref_ctx *c, double *out, int64_t max_records, int64_t *n_out) {
  if (!c || !n_out) return fail(GXT_EINVAL, "null argument");
  const int64_t n = (int64_t)c->tot.trace.size() / 12;
  *n_out = n;
  if (out) std::copy(c->tot.trace.begin(), c->tot.trace.begin() + 12 * std::min(n, max_records), out);
  return GXT_OK;
}

int gxref_rng_init(const uint32_t seed[6], const uint64_t *idx, int n, uint32_t *states) {
  if (!seed || !idx || !states || n < 0) return fail(GXT_EINVAL, "null argument");
  for (int i = 0; i < n; ++i) substream_state(seed, idx[i], states + 6 * i);
  return GXT_OK;
}

int gxref_rng_draw(uint32_t *states, int n, int k, double *out) {
  if (!states || !out || n < 0 || k < 0) return fail(GXT_EINVAL, "null argument");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) out[(size_t)i * k + j] = mrg_uniform(states + 6 * i);
  return GXT_OK;
}

int gxref_rng_jump_matrices(uint64_t *out) {
  const JumpTables &J = jumps();
  for (int c = 0; c < 2; ++c)
    for (int j = 0; j < 64; ++j)
      for (int k = 0; k < 9; ++k) out[(c * 64 + j) * 9 + k] = J.P[c][j][k];
  return GXT_OK;
}

static int slot_ok(gxref_ctx *c, int slot) { return c && slot >= 0 && slot < (int)c->phys.mat.size(); }

int gxref_table_value(gxref_ctx *c, int slot, int table, const double *x, double *out, int n) {
  if (!slot_ok(c, slot) || table < 0 || table >= GXT_NTAB) return fail(GXT_EINVAL, "bad table request");
  const PVec &p = c->phys.mat[slot].tab[table];
  const bool inv = (table == GXT_TAB_INVRANGE_EM || table == GXT_TAB_INVRANGE_EP);
  for (int i = 0; i < n; ++i) out[i] = inv ? pvec_inv_value(p, x[i]) : pvec_value(p, x[i]);
  return GXT_OK;
}

int gxref_table_nodes(gxref_ctx *c, int slot, int table, double *x, double *y, double *d2, int *n) {
  if (!slot_ok(c, slot) || table < 0 || table >= GXT_NTAB || !n) return fail(GXT_EINVAL, "bad table request");
  const PVec &p = c->phys.mat[slot].tab[table];
  const int cnt = (int)p.x.size();
  if (x && y && d2 && *n >= cnt) {
    std::copy(p.x.begin(), p.x.end(), x);
    std::copy(p.y.begin(), p.y.end(), y);
    std::copy(p.d2.begin(), p.d2.end(), d2);
  }
  *n = cnt;
  return GXT_OK;
}

int gxref_macro_xs(gxref_ctx *c, int slot, int species, int process, const double *e, double *out, int n) {
  if (!slot_ok(c, slot) || species < 0 || species > 2 || process < 0 || process >= nproc(species))
    return fail(GXT_EINVAL, "bad cross-section request");
  for (int i = 0; i < n; ++i) out[i] = macro_xs(c->phys.mat[slot], species, process, e[i]);
  return GXT_OK;
}

int gxref_sandia_xs(gxref_ctx *c, int slot, const double *e, double *out, int n) {
  if (!slot_ok(c, slot)) return fail(GXT_EINVAL, "bad slot");
  for (int i = 0; i < n; ++i) out[i] = sandia_macro(c->phys.mat[slot], e[i]);
  return GXT_OK;
}

int gxref_sample(gxref_ctx *c, int model, int slot, gxt_sample_io *io) {
  if (!slot_ok(c, slot) || !io || model < 0 || model > GXT_MODEL_ANNIHIL) return fail(GXT_EINVAL, "bad sample request");
  const int n = io->n;
  for (int i = 0; i < n; ++i) {
    const double d[3] = {io->dir[i], io->dir[n + i], io->dir[2 * n + i]};
    Interaction o;
    sample_model(c->phys.mat[slot], model, io->ekin[i], d, io->rng + 6 * i, o);
    io->out_ekin[i] = o.ekin;
    io->out_dir[i] = o.dir[0];
    io->out_dir[n + i] = o.dir[1];
    io->out_dir[2 * n + i] = o.dir[2];
    io->out_edep[i] = o.edep;
    io->nsec[i] = o.nsec;
    for (int k = 0; k < 2; ++k) {
      const bool has = k < o.nsec;
      io->sec_species[2 * i + k] = has ? o.sec[k].species : -1;
      io->sec_ekin[2 * i + k] = has ? o.sec[k].ekin : 0.0;
      for (int j = 0; j < 3; ++j) io->sec_dir[6 * i + 3 * k + j] = has ? o.sec[k].dir[j] : 0.0;
    }
  }
  return GXT_OK;
}

int gxref_helix_step(const double field[3], const double *charge, const double *mom, const double *step,
                     double *pos, double *dir, int n) {
  const double bmag = std::sqrt(field[0] * field[0] + field[1] * field[1] + field[2] * field[2]);
  if (bmag <= 0.0) return fail(GXT_EINVAL, "zero field");
  for (int i = 0; i < n; ++i) {
    double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    double d[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    helix_step(field, bmag, charge[i], mom[i], step[i], p, d);
    for (int k = 0; k < 3; ++k) {
      pos[k * n + i] = p[k];
      dir[k * n + i] = d[k];
    }
  }
  return GXT_OK;
}

int gxref_rotate_to_lab(double *u, double *v, double *w, const double *u1, const double *u2, const double *u3,
                        int n) {
  for (int i = 0; i < n; ++i) rotate_to_lab(u[i], v[i], w[i], u1[i], u2[i], u3[i]);
  return GXT_OK;
}

// The loop's straight-line navigation of the analytic geometry
// (NavigationInterface::FindNextBoundary, geometry/NavigationInterface.hpp:64-98,
// for slabs and boxes), same contract as gxt_slab_distance.
int gxref_slab_distance(gxref_ctx *c, const double *pos, const double *dir, const double *pstep, int32_t *vol,
                        double *snext, int32_t *next, double *saf, int32_t *boundary, int n) {
  if (!c || !pos || !dir || !vol || !snext || !next || !saf || n < 0) return fail(GXT_EINVAL, "bad slab request");
  if (c->geo.tree) return fail(GXT_EINVAL, "slab request on a volume-tree geometry");
  const Geometry &g = c->geo;
  for (int i = 0; i < n; ++i)
    if (vol[i] >= g.nvol()) return fail(GXT_EINVAL, "volume out of range");
  for (int i = 0; i < n; ++i) {
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    const double d[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    int v = vol[i] >= 0 ? vol[i] : locate(g, p, d);
    vol[i] = v;
    if (v < 0) {
      snext[i] = 0.0;
      saf[i] = 0.0;
      next[i] = -1;
      if (boundary) boundary[i] = 1;
      continue;
    }
    int face;
    const double s = dist_out(g, v, p, d, face);
    int nv = -1;
    if (face == 0 || face == 1) {
      nv = face == 1 ? v + 1 : v - 1;
      if (nv >= g.nvol()) nv = -1;
    }
    snext[i] = s;
    next[i] = nv;
    saf[i] = safety(g, v, p);
    if (boundary) boundary[i] = pstep ? !(pstep[i] < s) : 1;
  }
  return GXT_OK;
}

// The pair sampler's energy-sharing density above 2 MeV (test hook).
int gxref_bh_density(int Z, double e0, const double *eps, double *out, int n) {
  if (Z < 1 || Z > 100 || !(e0 >= 2.0)) return fail(GXT_EINVAL, "bad Bethe-Heitler density request");
  for (int i = 0; i < n; ++i) out[i] = bh_density(Z, e0, eps[i]);
  return GXT_OK;
}

// Along-step pieces of the oracle's e-/e+ step (test hook): the continuous-loss
// step limit of kinetic energy e[i], and the energy after a step of step[i]
// with its loss (ProxyIonization.hpp:108-168 under the physics contract).
int gxref_along_step(gxref_ctx *c, int slot, int species, const double *e, const double *step, double *limit,
                     double *fin, double *eloss, int n) {
  if (!slot_ok(c, slot) || species < 1 || species > 2) return fail(GXT_EINVAL, "bad along-step request");
  const Material &m = c->phys.mat[slot];
  const double mine = c->phys.tracking_cut;
  for (int i = 0; i < n; ++i) {
    limit[i] = along_limit(m, species == 2, e[i], mine);
    fin[i] = along_doit(m, species == 2, e[i], step[i], mine, eloss[i]);
  }
  return GXT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- field entry points
struct gxref_field_map {
  FieldMap m;
};

int gxref_field_map_load(const char *path, gxref_field_map **out) {
  if (!out) return fail(GXT_EINVAL, "null output");
  auto *f = new gxref_field_map();
  const int rc = read_field_map(path, f->m);
  if (rc) {
    delete f;
    return rc;
  }
  *out = f;
  return GXT_OK;
}

int gxref_field_map_free(gxref_field_map *map) {
  delete map;
  return GXT_OK;
}

int gxref_field_map_info(const gxref_field_map *map, int *nr, int *nz, double *r0, double *z0, double *dr, double *dz,
                         double *bmax) {
  if (!map) return fail(GXT_EINVAL, "null field map");
  const FieldMap &m = map->m;
  *nr = m.nr;
  *nz = m.nz;
  *r0 = m.r0;
  *z0 = m.z0;
  *dr = m.dr;
  *dz = m.dz;
  *bmax = m.bmax;
  return GXT_OK;
}

int gxref_field_map_value(const gxref_field_map *map, const double *pos, double *b, int n) {
  if (!map || !pos || !b || n <= 0) return fail(GXT_EINVAL, "bad field-map value arguments");
  for (int i = 0; i < n; ++i) {
    double B[3];
    map_field(map->m, pos[i], pos[n + i], pos[2 * n + i], B);
    for (int k = 0; k < 3; ++k) b[k * n + i] = B[k];
  }
  return GXT_OK;
}

int gxref_field_step(const gxref_field_map *map, const double field[3], const double origin[3], int stepper,
                     const double *charge, const double *mom, const double *step, double *pos, double *dir,
                     double *err, int n) {
  if ((!map && !field) || !charge || !mom || !step || !pos || !dir || n <= 0)
    return fail(GXT_EINVAL, "bad field-step arguments");
  if (stepper != GXT_STEPPER_HELIX && stepper != GXT_STEPPER_RK45) return fail(GXT_EINVAL, "bad stepper");
  Field F;
  F.map = map ? &map->m : nullptr;
  for (int k = 0; k < 3; ++k) {
    F.b[k] = field ? field[k] : 0.0;
    F.origin[k] = origin ? origin[k] : 0.0;
  }
  F.stepper = stepper;
  for (int i = 0; i < n; ++i) {
    double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    double d[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    double B[3];
    F.at(p, B);
    const double bm = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    double ex = 0.0, eu = 0.0;
    field_advance(F, B, bm, charge[i], mom[i], step[i], p, d, ex, eu);
    for (int k = 0; k < 3; ++k) {
      pos[k * n + i] = p[k];
      dir[k * n + i] = d[k];
    }
    if (err) {
      err[i] = ex;
      err[n + i] = eu;
    }
  }
  return GXT_OK;
}

// ---------------------------------------------------------------- msc test hooks
int gxref_msc_deflect(double tau, uint32_t seed, int n, double *cos_theta, double *phi) {
  if (!(tau > 0.0) || n <= 0 || !cos_theta) return fail(GXT_EINVAL, "bad msc deflect arguments");
  const uint32_t s6[6] = {seed, seed, seed, seed, seed, seed};
  for (int i = 0; i < n; ++i) {
    uint32_t st[6];
    substream_state(s6, (uint64_t)i, st);
    double d[3] = {0.0, 0.0, 1.0};
    msc_deflect(tau, d, st);
    cos_theta[i] = d[2];
    if (phi) phi[i] = std::atan2(d[1], d[0]);
  }
  return GXT_OK;
}

int gxref_msc_lengths(const double *t, const double *lam, int n, double *z, double *t_back) {
  if (!t || !lam || !z || !t_back || n <= 0) return fail(GXT_EINVAL, "bad msc length arguments");
  for (int i = 0; i < n; ++i) {
    z[i] = msc_geom(t[i], lam[i]);
    t_back[i] = msc_true(z[i], lam[i], 1.0e300);
  }
  return GXT_OK;
}

int gxref_msc_screening(const double *tau, int n, double *u) {
  if (!tau || !u || n <= 0) return fail(GXT_EINVAL, "bad msc screening arguments");
  for (int i = 0; i < n; ++i) u[i] = msc_u(tau[i]);
  return GXT_OK;
}

// ------------------------------------------------------- volume-tree navigator
// VecGeom's navigation over placed boxes (the geometry library the reference
// links, an un-vendored submodule at source/VecGeom; its call sites are
// geometry/NavigationInterface.hpp:42-98 and the locate of
// proxy/test/test_BasicCpuTransport.cpp:572), restated for boxes placed by
// translations (proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:
// 225-325 builds the reference calorimeter from exactly these):
//  - locate: descend from the world into the first daughter holding the point;
//  - next boundary: the least of the distance out of the current box and the
//    distances into each daughter (UnplacedBox's slab method), then relocate
//    the boundary point (up through the mothers, then down);
//  - safety: the isotropic distance to the current box's surface or the
//    nearest daughter.
// The surface tolerance is VecGeom's kTolerance, 1e-9 mm
// (NavigationInterface.hpp:66); a point within it of a face belongs to the box
// only if its direction enters through that face.

extern "C" {

int gxref_nav_create(const gxt_box_volume *vol, int nvol, const gxt_placement *pl, int npl, int world,
                     gxref_nav **out) {
  using gxref::fail;
  if (!vol || nvol <= 0 || npl < 0 || (npl > 0 && !pl) || world < 0 || world >= nvol || !out)
    return fail(GXT_EINVAL, "bad navigator arguments");
  for (int i = 0; i < nvol; ++i)
    for (int k = 0; k < 3; ++k)
      if (!(vol[i].half[k] >= 1e-6) || !std::isfinite(vol[i].half[k]) || vol[i].material < -1 ||
          vol[i].material > GXT_MAT_PBWO4)
        return fail(GXT_EINVAL, "bad volume");
  for (int i = 0; i < npl; ++i) {
    const gxt_placement &p = pl[i];
    if (p.mother < 0 || p.mother >= nvol || p.daughter < 0 || p.daughter >= nvol) return fail(GXT_EINVAL, "bad index");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(p.translation[k]) ||
          std::fabs(p.translation[k]) + vol[p.daughter].half[k] > vol[p.mother].half[k] + gxref::nav::kTol)
        return fail(GXT_EINVAL, "daughter outside its mother");
  }
  // overlaps: every pair of daughters of one mother
  for (int i = 0; i < npl; ++i)
    for (int j = i + 1; j < npl; ++j) {
      if (pl[i].mother != pl[j].mother) continue;
      int k = 0;
      while (k < 3 && std::fabs(pl[i].translation[k] - pl[j].translation[k]) <
                          vol[pl[i].daughter].half[k] + vol[pl[j].daughter].half[k] - gxref::nav::kTol)
        ++k;
      if (k == 3) return fail(GXT_EINVAL, "overlapping daughters");
    }
  gxref_nav *g = new gxref_nav();
  std::vector<int> level;
  gxref_nav::Touchable w;
  for (int k = 0; k < 3; ++k) {
    w.c[k] = 0.0;
    w.h[k] = vol[world].half[k];
  }
  w.parent = -1;
  w.volume = world;
  g->t.push_back(w);
  level.push_back(0);
  // breadth-first: touchables of one level in order of their mothers, the
  // daughters of one touchable in placement order
  for (size_t cur = 0; cur < g->t.size(); ++cur) {
    for (int i = 0; i < npl; ++i) {
      if (pl[i].mother != g->t[cur].volume) continue;
      if (level[cur] >= 31 || g->t.size() >= 65535) {
        delete g;
        return fail(GXT_EINVAL, "volume tree too deep or too large");
      }
      gxref_nav::Touchable d;
      for (int k = 0; k < 3; ++k) {
        d.c[k] = g->t[cur].c[k] + pl[i].translation[k];
        d.h[k] = vol[pl[i].daughter].half[k];
      }
      d.parent = (int)cur;
      d.volume = pl[i].daughter;
      g->t[cur].daughters.push_back((int)g->t.size());
      g->t.push_back(d);
      level.push_back(level[cur] + 1);
    }
  }
  *out = g;
  return GXT_OK;
}

int gxref_nav_free(gxref_nav *g) {
  delete g;
  return GXT_OK;
}

int gxref_nav_info(const gxref_nav *g, int *n, int32_t *volume, int32_t *parent, double *centre) {
  if (!g) return gxref::fail(GXT_EINVAL, "null navigator");
  if (n) *n = (int)g->t.size();
  for (size_t i = 0; i < g->t.size(); ++i) {
    if (volume) volume[i] = g->t[i].volume;
    if (parent) parent[i] = g->t[i].parent;
    if (centre)
      for (int k = 0; k < 3; ++k) centre[3 * i + k] = g->t[i].c[k];
  }
  return GXT_OK;
}

int gxref_nav_locate(const gxref_nav *g, const double *pos, const double *dir, int32_t *out, int n) {
  if (!g || !pos || !dir || !out || n <= 0) return gxref::fail(GXT_EINVAL, "bad locate arguments");
  for (int i = 0; i < n; ++i) {
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]}, v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    out[i] = gxref::nav::inside(g->t[0], p, v) ? gxref::nav::down(*g, 0, p, v) : -1;
  }
  return GXT_OK;
}

int gxref_nav_step(const gxref_nav *g, const int32_t *cur, const double *pos, const double *dir, double *snext,
                   int32_t *next, double *safety, int n) {
  if (!g || !cur || !pos || !dir || !snext || !next || !safety || n <= 0)
    return gxref::fail(GXT_EINVAL, "bad step arguments");
  for (int i = 0; i < n; ++i) {
    if (cur[i] < 0 || cur[i] >= (int)g->t.size()) return gxref::fail(GXT_EINVAL, "touchable out of range");
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]}, v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    const gxref_nav::Touchable &m = g->t[cur[i]];
    double s = gxref::nav::to_out(m, p, v);
    int target = cur[i];
    for (int d : m.daughters) {
      const double sd = gxref::nav::to_in(g->t[d], p, v);
      if (sd < s) {
        s = sd;
        target = d;
      }
    }
    const double q[3] = {p[0] + s * v[0], p[1] + s * v[1], p[2] + s * v[2]};
    snext[i] = s;
    next[i] = gxref::nav::relocate(*g, target, q, v);
    safety[i] = gxref::nav::safety(*g, cur[i], p);
  }
  return GXT_OK;
}

}  // extern "C"

// NavigationInterface::FindNextBoundary (geometry/NavigationInterface.hpp:64-98):
// the boundary distance within the physics step, the safety shortcut, the
// boundary flag and the push of 2 x 1e-9 mm past the surface.
extern "C" int gxref_nav_find_next_boundary(const gxref_nav *g, const int32_t *cur, const double *pos,
                                            const double *dir, const double *pstep, double *snext, int32_t *next,
                                            double *safety, int32_t *boundary, int n) {
  if (!pstep || !boundary) return gxref::fail(GXT_EINVAL, "bad find-next-boundary arguments");
  const int rc = gxref_nav_step(g, cur, pos, dir, snext, next, safety, n);
  if (rc) return rc;
  const double tol = gxref::nav::kTol;
  for (int i = 0; i < n; ++i) {
    const double ps = pstep[i];
    double s = snext[i];
    if (!(s < ps)) {  // no boundary within the proposed step
      s = ps;
      next[i] = cur[i];
    }
    if (safety[i] > ps) {
      snext[i] = ps;
      boundary[i] = 0;
      next[i] = cur[i];
      continue;
    }
    boundary[i] = s < ps;
    snext[i] = std::max(2.0 * tol, s + 2.0 * tol);
  }
  return GXT_OK;
}

extern "C" int gxref_sincos(const double *x, double *s, double *c, int n) {
  if (!x || !s || !c || n <= 0) return gxref::fail(GXT_EINVAL, "bad sincos arguments");
  for (int i = 0; i < n; ++i) gxref::sincos_cw(x[i], s[i], c[i]);
  return GXT_OK;
}

// The shared FP64 functions, for their accuracy tests: which = 0 log, 1 exp,
// 2 log1p, 3 expm1, 4 pow(x, y), 5 log10.
extern "C" int gxref_mathfn(int which, const double *x, const double *y, double *out, int n) {
  if (!x || !out || n <= 0 || which < 0 || which > 5 || (which == 4 && !y))
    return gxref::fail(GXT_EINVAL, "bad mathfn arguments");
  for (int i = 0; i < n; ++i) {
    switch (which) {
      case 0: out[i] = gxref::log_cw(x[i]); break;
      case 1: out[i] = gxref::exp_cw(x[i]); break;
      case 2: out[i] = gxref::log1p_cw(x[i]); break;
      case 3: out[i] = gxref::expm1_cw(x[i]); break;
      case 4: out[i] = gxref::pow_cw(x[i], y[i]); break;
      default: out[i] = gxref::log10_cw(x[i]); break;
    }
  }
  return GXT_OK;
}
