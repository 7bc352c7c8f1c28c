// gxt_field.cpp -- cylindrical (r, z) field map: reader, device copy and the
// field entry points of the C ABI (gxt_field_map_*, gxt_field_step).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gxt.h"
#include "gxt_data.h"
#include "gxt_field.h"
#include "gxt_kernels.h"

namespace gxt {

namespace {

// One node line: "R(mm)= r phi(deg)= phi Z(mm)= z Bz(tesla)= bz Br(tesla)= br
// Bphi(tesla)= bphi" (the column layout ScalarCMSmagField::ReadVectorData
// reads, ScalarCMSmagField.h:185-193).
bool parse_node(const std::string &line, double &r, double &z, double &br, double &bphi, double &bz) {
  static const char *const kLabels[6] = {"R(mm)=", "phi(deg)=", "Z(mm)=", "Bz(tesla)=", "Br(tesla)=", "Bphi(tesla)="};
  std::istringstream in(line);
  double val[6];
  for (int k = 0; k < 6; ++k) {
    std::string lab, num;
    if (!(in >> lab >> num) || lab != kLabels[k]) return false;
    char *end = nullptr;
    val[k] = std::strtod(num.c_str(), &end);
    if (end == num.c_str() || *end != '\0' || !std::isfinite(val[k])) return false;
  }
  std::string extra;
  if (in >> extra) return false;
  r = val[0];
  z = val[2];
  bz = val[3];
  br = val[4];
  bphi = val[5];
  return true;
}

bool on_grid(double x, double x0, double dx, int i) {
  const double want = x0 + (double)i * dx;
  return std::fabs(x - want) <= 1e-9 * std::max(std::fabs(want), std::fabs(dx));
}

}  // namespace

int read_field_map(const char *path, gxt_field_map &m) {
  if (!path) return set_error(GXT_EINVAL, "null field-map path");
  std::ifstream in(path);
  if (!in) return set_error(GXT_EDATA, std::string("cannot open field map ") + path);
  std::vector<double> r, z, b;
  std::string line;
  int64_t lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    double rv, zv, br, bp, bz;
    if (!parse_node(line, rv, zv, br, bp, bz))
      return set_error(GXT_EDATA, "malformed field-map line " + std::to_string(lineno) + " in " + path);
    r.push_back(rv);
    z.push_back(zv);
    b.push_back(br);
    b.push_back(bp);
    b.push_back(bz);
  }
  const int64_t n = (int64_t)r.size();
  int64_t nz = 1;
  while (nz < n && r[nz] == r[0]) ++nz;
  if (n < 4 || nz < 2 || nz == n || n % nz != 0)
    return set_error(GXT_EDATA, std::string("field map is not a regular r-major (r, z) grid: ") + path);
  const int64_t nr = n / nz;
  if (nr > (1 << 20) || nz > (1 << 20)) return set_error(GXT_EDATA, "field map too large");
  const double r0 = r[0], z0 = z[0], dr = r[nz] - r[0], dz = z[1] - z[0];
  if (!(dr > 0) || !(dz > 0) || r0 < 0)
    return set_error(GXT_EDATA, std::string("field-map grid must start at r >= 0 and ascend: ") + path);
  for (int64_t i = 0; i < n; ++i) {
    if (!on_grid(r[i], r0, dr, (int)(i / nz)) || !on_grid(z[i], z0, dz, (int)(i % nz)))
      return set_error(GXT_EDATA, "field-map node " + std::to_string(i + 1) + " is off the regular grid in " + path);
  }
  double bmax = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double *p = &b[3 * i];
    bmax = std::max(bmax, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
  }
  m.nr = (int)nr;
  m.nz = (int)nz;
  m.r0 = r0;
  m.z0 = z0;
  m.dr = dr;
  m.dz = dz;
  m.bmax = bmax;
  m.v.swap(b);
  return GXT_OK;
}

int upload_field_map(gxt_field_map &m) {
  hipError_t e = hipGetDevice(&m.device);
  if (e == hipSuccess) e = hipMalloc(&m.dev, m.v.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(m.dev, m.v.data(), m.v.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (m.dev) (void)hipFree(m.dev);
    m.dev = nullptr;
    return set_error(GXT_EHIP, std::string("field-map upload: ") + hipGetErrorString(e));
  }
  return GXT_OK;
}

void release_field_map(gxt_field_map &m) {
  if (m.dev) (void)hipFree(m.dev);
  m.dev = nullptr;
}

DevField dev_field(const gxt_field_map *m, const double b[3], const double origin[3], int stepper) {
  DevField f;
  std::memset(&f, 0, sizeof(f));
  if (m) {
    f.v = m->dev;
    f.nr = m->nr;
    f.nz = m->nz;
    f.r0 = m->r0;
    f.z0 = m->z0;
    f.dr = m->dr;
    f.dz = m->dz;
  }
  for (int k = 0; k < 3; ++k) {
    f.b[k] = b ? b[k] : 0.0;
    f.origin[k] = origin ? origin[k] : 0.0;
  }
  f.stepper = stepper;
  f.general = (m != nullptr || stepper == GXT_STEPPER_RK45) ? 1 : 0;
  return f;
}

}  // namespace gxt

using namespace gxt;

namespace {

template <class T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(T)); }
};

int hip_fail(hipError_t e, const char *what) { return set_error(GXT_EHIP, std::string(what) + ": " + hipGetErrorString(e)); }

// A map's nodes live on the device that was current when it was loaded.
int check_device(const gxt_field_map *m) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(GXT_EHIP, "hipGetDevice failed");
  if (dev != m->device)
    return set_error(GXT_EINVAL, "field map loaded on device " + std::to_string(m->device) + ", current device is " +
                                     std::to_string(dev));
  return GXT_OK;
}

}  // namespace

extern "C" {

int gxt_field_map_load(const char *path, gxt_field_map **out) {
  if (!path || !out) return set_error(GXT_EINVAL, "bad field-map load arguments");
  gxt_field_map *m = new gxt_field_map();
  int rc = read_field_map(path, *m);
  if (rc == GXT_OK) rc = upload_field_map(*m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return GXT_OK;
}

int gxt_field_map_free(gxt_field_map *map) {
  if (!map) return set_error(GXT_EINVAL, "null field map");
  release_field_map(*map);
  delete map;
  return GXT_OK;
}

int gxt_field_map_info(const gxt_field_map *m, int *nr, int *nz, double *r0, double *z0, double *dr, double *dz,
                       double *bmax) {
  if (!m) return set_error(GXT_EINVAL, "null field map");
  if (nr) *nr = m->nr;
  if (nz) *nz = m->nz;
  if (r0) *r0 = m->r0;
  if (z0) *z0 = m->z0;
  if (dr) *dr = m->dr;
  if (dz) *dz = m->dz;
  if (bmax) *bmax = m->bmax;
  return GXT_OK;
}

int gxt_field_map_value(const gxt_field_map *m, const double *pos, double *b, int n) {
  if (!m || !m->dev || !pos || !b || n <= 0 || n > kMaxSoaPoints)
    return set_error(GXT_EINVAL, "bad field-map value arguments (1 <= n <= 2^29)");
  if (const int rc = check_device(m)) return rc;
  const DevField f = dev_field(m, nullptr, nullptr, GXT_STEPPER_HELIX);
  DevBuf<double> dp, db;
  hipError_t e = dp.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = db.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(dp.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_field_value(f, dp.p, db.p, n, 0);
  if (e == hipSuccess) e = hipMemcpy(b, db.p, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
  return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_field_map_value");
}

int gxt_field_step(const gxt_field_map *m, const double field[3], const double origin[3], int stepper, int precision,
                   const double *charge, const double *mom, const double *step, double *pos, double *dir, double *err,
                   int n) {
  if ((!m && !field) || !charge || !mom || !step || !pos || !dir || n <= 0 || n > kMaxSoaPoints)
    return set_error(GXT_EINVAL, "bad field-step arguments");
  if (m && !m->dev) return set_error(GXT_EINVAL, "field map not on a device");
  if (m)
    if (const int rc = check_device(m)) return rc;
  if (stepper != GXT_STEPPER_HELIX && stepper != GXT_STEPPER_RK45) return set_error(GXT_EINVAL, "bad stepper");
  if (precision != GXT_FP64 && precision != GXT_FP32) return set_error(GXT_EINVAL, "bad precision");
  // the RK45 sub-step count ceil(step 0.3 |B| / (p GXT_RK_ANGLE)) is bounded with the largest |B|
  const double bmax =
      m ? m->bmax : std::sqrt(field[0] * field[0] + field[1] * field[1] + field[2] * field[2]);
  for (int i = 0; i < n; ++i) {
    if (!(mom[i] >= 1e-30) || !std::isfinite(mom[i]) || !(step[i] >= 0) || !std::isfinite(step[i]) ||
        (charge[i] != 1.0 && charge[i] != -1.0))
      return set_error(GXT_EINVAL,
                       "field step " + std::to_string(i) + ": need finite momentum >= 1e-30, step >= 0, charge +-1");
    if (stepper == GXT_STEPPER_RK45 && step[i] * 0.299792458 * bmax / mom[i] > 65536.0 * GXT_RK_ANGLE)
      return set_error(GXT_EINVAL, "field step " + std::to_string(i) + " needs more than 2^16 RK45 sub-steps");
  }
  const DevField f = dev_field(m, field, origin, stepper);
  DevBuf<double> dq, dm, ds, dp, dd, de;
  const size_t n1 = (size_t)n * sizeof(double), n3 = 3 * n1;
  hipError_t e = dq.alloc(n);
  if (e == hipSuccess) e = dm.alloc(n);
  if (e == hipSuccess) e = ds.alloc(n);
  if (e == hipSuccess) e = dp.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dd.alloc(3 * (size_t)n);
  if (e == hipSuccess && err) e = de.alloc(2 * (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(dq.p, charge, n1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dm.p, mom, n1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ds.p, step, n1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dp.p, pos, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd.p, dir, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_field_step(f, precision, dq.p, dm.p, ds.p, dp.p, dd.p, de.p, n, 0);
  if (e == hipSuccess) e = hipMemcpy(pos, dp.p, n3, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(dir, dd.p, n3, hipMemcpyDeviceToHost);
  if (e == hipSuccess && err) e = hipMemcpy(err, de.p, 2 * n1, hipMemcpyDeviceToHost);
  return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_field_step");
}

}  // extern "C"
