// gxt_data.cpp -- host side loading of the physics data into the device layout.
//
// Reads the Geant4-dumped tables (format of ProxyPhysicsTable.cu:41-88 /
// ProxyPhysicsVector.cu:89-128), the production cuts (cut.dat), the
// Seltzer-Berger grids (ProxySeltzerBerger.hpp:487-526 format), the Sandia
// coefficients and the shell binding energies, scales the tables to the
// requested density and computes the spline second derivatives with the
// reference's elimination (ProxyPhysicsVector.cu:308-386).
#include "gxt_data.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

namespace gxt {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kAvogadro = 6.02214179e+23;
constexpr double kClassicRe = 2.8179402894e-12;
constexpr double kComptonLength = 3.8615926459e-10;

const char *kFiles[GXT_NTAB] = {"Lambda.compt.gamma.asc", "LambdaPrim.compt.gamma.asc", "Lambda.conv.gamma.asc",
                                "LambdaPrim.phot.gamma.asc", "Lambda.eBrem.e-.asc", "Lambda.eIoni.e-.asc",
                                "DEDX.eIoni.e-.asc", "Range.eIoni.e-.asc", "InverseRange.eIoni.e-.asc",
                                "Lambda.eBrem.e+.asc", "Lambda.eIoni.e+.asc", "DEDX.eIoni.e+.asc",
                                "Range.eIoni.e+.asc", "InverseRange.eIoni.e+.asc", "LambdaMod1.msc.e-.asc",
                                "LambdaMod1.msc.e+.asc"};
// density scaling: 0 multiply y, 1 divide y, 2 divide x
const int kScale[GXT_NTAB] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 0, 0, 0, 1, 2, 0, 0};
// tables a material may lack (then zero nodes): the multiple-scattering ones
bool optional_table(int t) { return t == GXT_TAB_MSC_EM || t == GXT_TAB_MSC_EP; }

struct Vec {
  std::vector<double> x, y, d2;
};

bool parse_tables(const std::string &path, std::vector<Vec> &out) {
  std::ifstream in(path);
  if (!in) return false;
  int count = 0;
  in >> count;
  if (!in || count <= 0) return false;
  out.assign(count, Vec());
  for (int v = 0; v < count; ++v) {
    int type = 0, nodes = 0, size = 0;
    double lo = 0, hi = 0;
    in >> type >> lo >> hi >> nodes >> size;
    if (!in || size < 5) return false;
    out[v].x.resize(size);
    out[v].y.resize(size);
    for (int i = 0; i < size; ++i) in >> out[v].x[i] >> out[v].y[i];
    if (!in) return false;
  }
  return true;
}

// Second derivatives: the reference's not-a-knot tridiagonal elimination,
// step for step (the interpolated values depend on these exact numbers).
void second_derivatives(Vec &v) {
  const std::vector<double> &x = v.x, &y = v.y;
  const int n = (int)x.size() - 1;
  std::vector<double> d(n + 1, 0.0), u(n, 0.0);
  u[1] = ((y[2] - y[1]) / (x[2] - x[1]) - (y[1] - y[0]) / (x[1] - x[0]));
  u[1] = 6.0 * u[1] * (x[2] - x[1]) / ((x[2] - x[0]) * (x[2] - x[0]));
  d[1] = (2.0 * x[1] - x[0] - x[2]) / (2.0 * x[2] - x[0] - x[1]);
  for (int i = 2; i < n - 1; ++i) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double p = sig * d[i - 1] + 2.0;
    d[i] = (sig - 1.0) / p;
    const double dd = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * dd / (x[i + 1] - x[i - 1])) - sig * u[i - 1] / p;
  }
  double sig = (x[n - 1] - x[n - 2]) / (x[n] - x[n - 2]);
  double p = sig * d[n - 3] + 2.0;
  const double dd = (y[n] - y[n - 1]) / (x[n] - x[n - 1]) - (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
  u[n - 1] = 6.0 * sig * dd / (x[n] - x[n - 2]) - (2.0 * sig - 1.0) * u[n - 2] / p;
  p = (1.0 + sig) + (2.0 * sig - 1.0) * d[n - 2];
  d[n - 1] = u[n - 1] / p;
  for (int k = n - 2; k > 1; --k) d[k] *= (d[k + 1] - u[k] * (x[k + 1] - x[k - 1]) / (x[k + 1] - x[k]));
  d[n] = (d[n - 1] - (1.0 - sig) * d[n - 2]) / sig;
  sig = 1.0 - ((x[2] - x[1]) / (x[2] - x[0]));
  d[1] *= (d[2] - u[1] / (1.0 - sig));
  d[0] = (d[1] - sig * d[2]) / (1.0 - sig);
  v.d2 = d;
}

bool find_sandia(const std::string &dir, int Z, std::vector<double> &rows, double &zoa) {
  std::ifstream in(dir + "/sandia.dat");
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream hs(line);
    int z = 0, nint = 0;
    double r = 0;
    hs >> z >> nint >> r;
    std::vector<double> tmp(5 * nint);
    for (int i = 0; i < nint; ++i) {
      if (!std::getline(in, line)) return false;
      std::istringstream rs(line);
      for (int k = 0; k < 5; ++k) rs >> tmp[5 * i + k];
    }
    if (z == Z) {
      rows = tmp;
      zoa = r;
      return true;
    }
  }
  return false;
}

bool find_shells(const std::string &dir, int Z, std::vector<double> &b) {
  std::ifstream in(dir + "/shells.dat");
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream s(line);
    int z = 0, n = 0;
    s >> z >> n;
    if (z != Z) continue;
    b.resize(n);
    for (int i = 0; i < n; ++i) {
      double ev = 0;
      s >> ev;
      b[i] = ev * 1.0e-6;
    }
    return true;
  }
  return false;
}

bool read_sb(const std::string &path, std::vector<double> &g) {
  std::ifstream in(path);
  if (!in) return false;
  int k = 0, nx = 0, ny = 0;
  in >> k >> nx >> ny;
  if (!in || nx != 32 || ny != 57) return false;
  g.resize(32 + 57 + 57 * 32);
  for (size_t i = 0; i < g.size(); ++i) in >> g[i];
  return (bool)in;
}

bool read_cut_file(const std::string &path, int couple, double &g, double &em) {
  std::ifstream in(path);
  if (!in) return false;
  std::string tag;
  int nc = 0;
  in >> tag >> nc;
  std::vector<double> len(4 * nc), en(4 * nc);
  for (int i = 0; i < 4 * nc; ++i) in >> len[i] >> en[i];
  if (!in || couple >= nc) return false;
  g = en[couple] * 1e-3;
  em = en[nc + couple] * 1e-3;
  return true;
}

struct Comp {
  int Z;
  double A;
  double count;
};

}  // namespace

// Merged Sandia intervals of material m (DevMaterial::sx_*): an element's
// photo-absorption at E uses its row r = the last with edge[r] <= E (r >= 1),
// else row 0 (ProxySandiaData's table, as the reference's GetSandiaCofPerAtom
// reads it).  Between two consecutive edges of the union of all elements'
// edges every element's row is fixed, so one search over the union gives all
// rows, and the per-element arithmetic is unchanged.
static int build_sandia_intervals(HostData &hd, DevMaterial &m, std::string &err) {
  std::vector<double> B;
  for (int i = 0; i < m.nel; ++i) {
    const DevElement &e = m.el[i];
    const double *rows = hd.sandia.data() + e.sandia;
    for (int r = 1; r < e.nint; ++r) {
      if (!(rows[5 * r] > rows[5 * (r - 1)])) {
        err = "sandia.dat: edges of Z=" + std::to_string(e.Z) + " not ascending";
        return 1;
      }
      B.push_back(rows[5 * r]);
    }
  }
  std::sort(B.begin(), B.end());
  B.erase(std::unique(B.begin(), B.end()), B.end());
  const int nb = (int)B.size();
  m.sx_b = (int)hd.sx.size();
  m.sx_n = nb;
  hd.sx.insert(hd.sx.end(), B.begin(), B.end());
  m.sx_c = (int)hd.sx.size();
  for (int j = 0; j <= nb; ++j) {
    for (int i = 0; i < m.nel; ++i) {
      const DevElement &e = m.el[i];
      const double *rows = hd.sandia.data() + e.sandia;
      int r = 0;
      if (j > 0)
        for (int k = 1; k < e.nint; ++k)
          if (rows[5 * k] <= B[j - 1]) r = k;
      for (int k = 1; k <= 4; ++k) hd.sx.push_back(rows[5 * r + k]);
    }
  }
  // hint k: the interval of u_k = 2^(l0 + k / hscale), the lower end of bin k
  const int K = 128;
  const double l0 = nb ? std::log2(B.front()) : 0.0, l1 = nb ? std::log2(B.back()) : 1.0;
  m.sx_hint = (int)hd.hints.size();
  m.sx_nhint = K;
  m.sx_hlog0 = (float)l0;
  m.sx_hscale = (float)(K / std::max(l1 - l0, 1e-3));
  for (int k = 0; k < K; ++k) {
    const double u = std::exp2(l0 + k / (double)m.sx_hscale);
    int j = 0;
    while (j < nb && B[j] <= u) ++j;
    hd.hints.push_back(j);
  }
  return 0;
}

int load_material(const std::string &data, int which, double lar_density, HostData &hd, int slot, std::string &err) {
  std::string dir;
  int couple = 0;
  double tab_rho = 0, rho = 0;
  std::vector<Comp> comp;
  if (which == GXT_MAT_PB) {
    dir = data + "/g4_pb_lar";
    couple = 1;
    tab_rho = rho = 11.35;
    comp = {{82, 207.217, 1}};
  } else if (which == GXT_MAT_LAR) {
    dir = data + "/g4_pb_lar";
    couple = 2;
    tab_rho = 1.396;
    rho = lar_density > 0 ? lar_density : 1.396;
    comp = {{18, 39.9477, 1}};
  } else if (which == GXT_MAT_PBWO4) {
    dir = data + "/pbwo4";
    couple = 0;
    tab_rho = rho = 8.28;
    comp = {{8, 15.9994, 4}, {74, 183.842, 1}, {82, 207.217, 1}};
  } else {
    err = "unknown material id";
    return GXT_EINVAL;
  }
  DevMaterial &m = hd.params.mat[slot];
  std::memset(&m, 0, sizeof(m));
  double amass = 0;
  for (const Comp &c : comp) amass += c.count * c.A;
  m.nel = (int)comp.size();
  m.ne = 0;
  for (int i = 0; i < m.nel; ++i) {
    const Comp &c = comp[i];
    DevElement &e = m.el[i];
    e.Z = c.Z;
    const double w = c.count * c.A / amass;
    e.natoms = rho * kAvogadro * w / c.A * 1.0e-3;
    e.wbrem = e.natoms * c.Z * (c.Z + 1.0);
    m.ne += e.natoms * c.Z;
    std::vector<double> rows, shells, grid;
    if (!find_sandia(data, c.Z, rows, e.zoa)) {
      err = "sandia.dat lacks Z=" + std::to_string(c.Z);
      return GXT_EDATA;
    }
    if (!find_shells(data, c.Z, shells)) {
      err = "shells.dat lacks Z=" + std::to_string(c.Z);
      return GXT_EDATA;
    }
    if (!read_sb(data + "/brem_SB/br" + std::to_string(c.Z), grid)) {
      err = "cannot read brem_SB/br" + std::to_string(c.Z);
      return GXT_EDATA;
    }
    // the kernels keep one copy of the (k / T, ln T) grid, which the
    // Seltzer-Berger files share
    DevParams &P = hd.params;
    if (!hd.sb_grid) {
      for (int k = 0; k < 32; ++k) P.sbx[k] = grid[k], P.sbxf[k] = (float)grid[k];
      for (int k = 0; k < 57; ++k) P.sby[k] = grid[32 + k], P.sbyf[k] = (float)grid[32 + k];
      hd.sb_grid = true;
    } else {
      for (int k = 0; k < 32 + 57; ++k)
        if (grid[k] != (k < 32 ? P.sbx[k] : P.sby[k - 32])) {
          err = "brem_SB/br" + std::to_string(c.Z) + ": its grid differs from the other elements'";
          return GXT_EDATA;
        }
    }
    e.sandia = (int)hd.sandia.size();
    e.nint = (int)rows.size() / 5;
    hd.sandia.insert(hd.sandia.end(), rows.begin(), rows.end());
    e.shell = (int)hd.shells.size();
    e.nshell = (int)shells.size();
    hd.shells.insert(hd.shells.end(), shells.begin(), shells.end());
    e.sb = (int)hd.sb.size();
    hd.sb.insert(hd.sb.end(), grid.begin(), grid.end());
  }
  if (build_sandia_intervals(hd, m, err)) return GXT_EDATA;
  m.migdal = m.ne * 4.0 * kPi * kClassicRe * kComptonLength * kComptonLength;
  m.cut_length = 0.7;
  if (!read_cut_file(dir + "/cut.dat", couple, m.cut_gamma, m.cut_e)) {
    err = "cannot read " + dir + "/cut.dat";
    return GXT_EDATA;
  }
  const double f = rho / tab_rho;
  for (int t = 0; t < GXT_NTAB; ++t) {
    std::vector<Vec> vecs;
    if (optional_table(t) && !std::ifstream(dir + "/" + kFiles[t])) {
      DevTable &dt = m.tab[t];
      std::memset(&dt, 0, sizeof(dt));
      dt.off = (int)hd.x.size();
      dt.hint = -1;
      continue;
    }
    if (!parse_tables(dir + "/" + kFiles[t], vecs) || couple >= (int)vecs.size()) {
      err = "cannot read " + dir + "/" + kFiles[t];
      return GXT_EDATA;
    }
    Vec v = vecs[couple];
    if (f != 1.0)
      for (size_t i = 0; i < v.x.size(); ++i) {
        if (kScale[t] == 0) v.y[i] *= f;
        else if (kScale[t] == 1) v.y[i] /= f;
        else v.x[i] /= f;
      }
    second_derivatives(v);
    DevTable &dt = m.tab[t];
    dt.off = (int)hd.x.size();
    dt.n = (int)v.x.size();
    dt.dl = std::log(v.x[1] / v.x[0]);
    dt.base = std::log(v.x[0]) / dt.dl;
    dt.dlf = (float)dt.dl;
    dt.basef = (float)dt.base;
    dt.x0 = v.x.front();
    dt.y0 = v.y.front();
    dt.xn = v.x.back();
    dt.yn = v.y.back();
    dt.x0f = (float)dt.x0;
    dt.y0f = (float)dt.y0;
    dt.xnf = (float)dt.xn;
    dt.ynf = (float)dt.yn;
    dt.hint = -1;
    dt.nhint = 0;
    dt.hlog0 = 0.0f;
    dt.hscale = 0.0f;
    if (t == GXT_TAB_INVRANGE_EM || t == GXT_TAB_INVRANGE_EP) {
      // hint k: last node below u_k = x[0] 2^(k / hscale); the device starts
      // its search there, one bin below r's, and steps to the last node below r
      const int K = kInvRangeHints;
      const int n = (int)v.x.size();
      const double l0 = std::log2(v.x[0]), l1 = std::log2(v.x[n - 1]);
      dt.hint = (int)hd.hints.size();
      dt.nhint = K;
      dt.hlog0 = (float)l0;
      dt.hscale = (float)(K / (l1 - l0));
      for (int k = 0; k < K; ++k) {
        const double u = std::exp2(l0 + k / (double)dt.hscale);
        int j = 0;
        while (j + 1 < n - 1 && v.x[j + 1] < u) ++j;
        hd.hints.push_back(j);
      }
    }
    hd.x.insert(hd.x.end(), v.x.begin(), v.x.end());
    hd.y.insert(hd.y.end(), v.y.begin(), v.y.end());
    hd.d2.insert(hd.d2.end(), v.d2.begin(), v.d2.end());
  }
  m.compt_split = hd.x[m.tab[GXT_TAB_LAMBDA_COMPT].off + m.tab[GXT_TAB_LAMBDA_COMPT].n - 1];
  m.phot_emin = hd.x[m.tab[GXT_TAB_LAMBDAPRIM_PHOT].off];
  return GXT_OK;
}

void build_geometry(const gxt_config &c, DevGeometry &g) {
  std::memset(&g, 0, sizeof(g));
  if (c.geometry == GXT_GEOM_TREE) return;  // built from the navigator by the engine
  g.hx = g.hy = 0.5 * c.box_xy;
  if (c.geometry == GXT_GEOM_BOX) {
    g.nvol = 1;
    g.z[0] = 0.0;
    g.z[1] = c.box_depth;
    g.slot[0] = 0;
    g.visible[0] = 1;
  } else {
    const double period = c.absorber_mm + c.gap_mm;
    g.nvol = 2 * c.n_layers;
    for (int l = 0; l < c.n_layers; ++l) {
      g.z[2 * l] = l * period;
      g.z[2 * l + 1] = l * period + c.absorber_mm;
      g.slot[2 * l] = 0;
      g.slot[2 * l + 1] = 1;
      g.visible[2 * l] = 0;
      g.visible[2 * l + 1] = 1;
    }
    g.z[2 * c.n_layers] = c.n_layers * period;
  }
}

// A_c^(2^(76+j)) mod m_c, j in [0, 64), by repeated squaring of the
// one-step matrices of the two MRG32k3a components.
void build_jump_matrices(uint32_t out[2][64][9]) {
  const uint64_t m[2] = {4294967087ULL, 4294944443ULL};
  const uint64_t A[2][9] = {{0, 1, 0, 0, 0, 1, 4294967087ULL - 810728ULL, 1403580ULL, 0},
                            {0, 1, 0, 0, 0, 1, 4294944443ULL - 1370589ULL, 0, 527612ULL}};
  for (int c = 0; c < 2; ++c) {
    uint64_t M[9];
    std::memcpy(M, A[c], sizeof(M));
    for (int k = 0; k < 76 + 64; ++k) {
      if (k >= 76)
        for (int i = 0; i < 9; ++i) out[c][k - 76][i] = (uint32_t)M[i];
      uint64_t S[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          uint64_t acc = 0;
          for (int l = 0; l < 3; ++l) acc = (acc + (M[i * 3 + l] * M[l * 3 + j]) % m[c]) % m[c];
          S[i * 3 + j] = acc;
        }
      std::memcpy(M, S, sizeof(M));
    }
  }
}

void HostMrg::init(uint32_t seed, uint64_t idx) {
  static uint32_t J[2][64][9];
  static std::once_flag once;
  std::call_once(once, [] { build_jump_matrices(J); });
  const uint64_t m[2] = {4294967087ULL, 4294944443ULL};
  for (int i = 0; i < 6; ++i) s[i] = seed;
  for (int j = 0; j < 64; ++j)
    if ((idx >> j) & 1ULL)
      for (int c = 0; c < 2; ++c) {
        uint64_t r[3];
        for (int a = 0; a < 3; ++a) {
          uint64_t acc = 0;
          for (int b = 0; b < 3; ++b) acc += ((uint64_t)J[c][j][a * 3 + b] * s[3 * c + b]) % m[c];
          r[a] = acc % m[c];
        }
        for (int a = 0; a < 3; ++a) s[3 * c + a] = r[a];
      }
}

double HostMrg::uniform() {
  int64_t p1 = (1403580LL * (int64_t)s[1] - 810728LL * (int64_t)s[0]) % 4294967087LL;
  if (p1 < 0) p1 += 4294967087LL;
  s[0] = s[1];
  s[1] = s[2];
  s[2] = (uint64_t)p1;
  int64_t p2 = (527612LL * (int64_t)s[5] - 1370589LL * (int64_t)s[3]) % 4294944443LL;
  if (p2 < 0) p2 += 4294944443LL;
  s[3] = s[4];
  s[4] = s[5];
  s[5] = (uint64_t)p2;
  const double norm = 1.0 / (4294967087.0 + 1.0);
  return (p1 > p2) ? (double)(p1 - p2) * norm : (double)(p1 - p2 + 4294967087LL) * norm;
}

void build_jump_tables(const uint32_t jump[2][64][9], std::vector<uint32_t> &out) {
  const uint64_t m[2] = {4294967087ULL, 4294944443ULL};
  out.assign((size_t)2 * 8 * 256 * kJumpRow, 0);
  for (int c = 0; c < 2; ++c)
    for (int k = 0; k < 8; ++k) {
      const uint32_t *J = jump[c][8 * k];
      uint64_t M[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      for (int d = 0; d < 256; ++d) {
        uint32_t *o = &out[((size_t)(c * 8 + k) * 256 + d) * kJumpRow];
        for (int i = 0; i < 9; ++i) o[i] = (uint32_t)M[i];
        uint64_t S[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            uint64_t acc = 0;
            for (int l = 0; l < 3; ++l) acc = (acc + (M[i * 3 + l] * J[l * 3 + j]) % m[c]) % m[c];
            S[i * 3 + j] = acc;
          }
        std::memcpy(M, S, sizeof(M));
      }
    }
}

}  // namespace gxt
