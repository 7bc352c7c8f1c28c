// gxt_nav.cpp -- volume tree of placed boxes: validation, flattening into
// touchables (breadth-first, global coordinates), device copy and the
// navigation entry points of the C ABI (gxt_nav_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gxt.h"
#include "gxt_data.h"
#include "gxt_kernels.h"
#include "gxt_types.h"

struct gxt_navigator {
  int device = -1;
  std::vector<gxt::NavNode> node;  // touchables, breadth-first from the world
  std::vector<gxt::NavGrid> grid;  // daughter index per volume
  std::vector<int32_t> idx;
  gxt::NavNode *dev = nullptr;     // one allocation: nodes, grids, index
  gxt::DevNav view() const {
    gxt::DevNav g;
    g.node = dev;
    g.grid = reinterpret_cast<const gxt::NavGrid *>(dev + node.size());
    g.idx = reinterpret_cast<const int32_t *>(g.grid + grid.size());
    g.ntouch = (int)node.size();
    g.pad = 0;
    return g;
  }
  size_t bytes() const {
    return node.size() * sizeof(gxt::NavNode) + grid.size() * sizeof(gxt::NavGrid) + idx.size() * sizeof(int32_t);
  }
  // the allocation: whole 16-byte words, which the LDS form copies
  size_t bytes_padded() const { return (bytes() + 15) & ~(size_t)15; }
  int lds_bytes() const { return bytes_padded() <= (size_t)gxt::kNavLdsMax ? (int)bytes_padded() : 0; }
};

using namespace gxt;

namespace gxt {
DevNav nav_device_view(const gxt_navigator *nav) { return nav->view(); }
const NavNode *nav_host_nodes(const gxt_navigator *nav, int *n) {
  *n = (int)nav->node.size();
  return nav->node.data();
}
}  // namespace gxt

namespace {

constexpr int kMaxTouchables = 65535;
constexpr int kMaxLevel = 31;  // the world is level 0
constexpr double kTol = 1e-9;  // the device navigator's surface tolerance (mm)
constexpr double kMinHalf = 1e-6;

bool finite3(const double *a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

// The daughters of one volume overlap if their boxes intersect by more than
// the tolerance on every axis.  Sweep along the axis where the centres spread
// most (the stacking axis of a layered setup), so that only neighbours along
// it are compared.
int check_overlaps(const gxt_box_volume *vol, const std::vector<const gxt_placement *> &pl, int mother) {
  const size_t n = pl.size();
  if (n < 2) return GXT_OK;
  int ax = 0;
  double best = -1;
  for (int k = 0; k < 3; ++k) {
    double lo = pl[0]->translation[k], hi = lo;
    for (const gxt_placement *q : pl) {
      lo = std::min(lo, q->translation[k]);
      hi = std::max(hi, q->translation[k]);
    }
    if (hi - lo > best) {
      best = hi - lo;
      ax = k;
    }
  }
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  auto lo = [&](size_t i) { return pl[i]->translation[ax] - vol[pl[i]->daughter].half[ax]; };
  auto hi = [&](size_t i) { return pl[i]->translation[ax] + vol[pl[i]->daughter].half[ax]; };
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return lo(a) < lo(b); });
  for (size_t i = 0; i < n; ++i) {
    const gxt_placement *a = pl[order[i]];
    for (size_t j = i + 1; j < n && lo(order[j]) < hi(order[i]) - kTol; ++j) {
      const gxt_placement *b = pl[order[j]];
      bool overlap = true;
      for (int k = 0; k < 3 && overlap; ++k)
        overlap = std::fabs(a->translation[k] - b->translation[k]) <
                  vol[a->daughter].half[k] + vol[b->daughter].half[k] - kTol;
      if (overlap)
        return set_error(GXT_EINVAL, "daughters " + std::to_string(a->daughter) + " and " +
                                         std::to_string(b->daughter) + " of volume " + std::to_string(mother) +
                                         " overlap");
    }
  }
  return GXT_OK;
}

int build_tree(const gxt_box_volume *vol, int nvol, const gxt_placement *pl, int npl, int world,
               std::vector<NavNode> &node, std::vector<std::vector<const gxt_placement *>> &daughters) {
  for (int i = 0; i < nvol; ++i) {
    const gxt_box_volume &v = vol[i];
    if (!finite3(v.half) || !(v.half[0] >= kMinHalf && v.half[1] >= kMinHalf && v.half[2] >= kMinHalf))
      return set_error(GXT_EINVAL, "volume " + std::to_string(i) + ": half-lengths must be finite and >= 1e-6 mm");
    if (v.material < -1 || v.material > GXT_MAT_PBWO4)
      return set_error(GXT_EINVAL, "volume " + std::to_string(i) + ": material must be a GXT_MAT_* or -1");
  }
  daughters.assign(nvol, {});
  for (int i = 0; i < npl; ++i) {
    const gxt_placement &p = pl[i];
    if (p.mother < 0 || p.mother >= nvol || p.daughter < 0 || p.daughter >= nvol || !finite3(p.translation))
      return set_error(GXT_EINVAL, "placement " + std::to_string(i) + ": volume index out of range or translation "
                                                                      "not finite");
    const double *hm = vol[p.mother].half, *hd = vol[p.daughter].half;
    for (int k = 0; k < 3; ++k)
      if (std::fabs(p.translation[k]) + hd[k] > hm[k] + kTol)
        return set_error(GXT_EINVAL, "placement " + std::to_string(i) + ": daughter " + std::to_string(p.daughter) +
                                         " is not inside its mother " + std::to_string(p.mother));
    daughters[p.mother].push_back(&p);
  }
  for (int m = 0; m < nvol; ++m)
    if (const int rc = check_overlaps(vol, daughters[m], m)) return rc;
  node.clear();
  std::vector<int> level;
  NavNode w;
  std::memset(&w, 0, sizeof(w));
  for (int k = 0; k < 3; ++k) w.h[k] = vol[world].half[k];
  w.parent = -1;
  w.volume = world;
  node.push_back(w);
  level.push_back(0);
  for (size_t t = 0; t < node.size(); ++t) {
    const int v = node[t].volume;
    const int first = (int)node.size();
    if (!daughters[v].empty() && level[t] == kMaxLevel)
      return set_error(GXT_EINVAL, "volume tree is deeper than 32 levels or has a placement cycle");
    if (node.size() + daughters[v].size() > (size_t)kMaxTouchables)
      return set_error(GXT_EINVAL, "volume tree has more than 65535 touchables");
    for (const gxt_placement *p : daughters[v]) {
      NavNode d;
      std::memset(&d, 0, sizeof(d));
      for (int k = 0; k < 3; ++k) {
        d.c[k] = node[t].c[k] + p->translation[k];
        d.h[k] = vol[p->daughter].half[k];
      }
      d.parent = (int32_t)t;
      d.volume = p->daughter;
      node.push_back(d);
      level.push_back(level[t] + 1);
    }
    node[t].first = first;
    node[t].count = (int)daughters[v].size();
  }
  return GXT_OK;
}

constexpr int kGridMin = 8;       // as the device's kNavGridMin
constexpr int kGridMaxBins = 4096;

// Axis along which the centres of a volume's daughters spread most (-1: they
// do not spread).
int spread_axis(const std::vector<const gxt_placement *> &pl) {
  int ax = -1;
  double best = 0.0;
  for (int k = 0; k < 3; ++k) {
    double lo = pl[0]->translation[k], hi = lo;
    for (const gxt_placement *q : pl) {
      lo = std::min(lo, q->translation[k]);
      hi = std::max(hi, q->translation[k]);
    }
    if (hi - lo > best) {
      best = hi - lo;
      ax = k;
    }
  }
  return ax;
}

// Daughter index (the device's nav_descend tests only the candidates of a
// point's bin) for each volume with at least kGridMin daughters.  A
// daughter's extent along the axis, widened by a margin far above the
// rounding of the device's coordinate differences, is placed in bins with the
// device's own nav_bin; a volume whose index would hold more than 16 entries
// per daughter keeps the plain scan.
void build_index(const gxt_box_volume *vol, int nvol, const std::vector<std::vector<const gxt_placement *>> &dau,
                 double world_extent, std::vector<NavGrid> &grid, std::vector<int32_t> &idx) {
  grid.assign(nvol, NavGrid{0.0, 0, 0, 0, 0});
  idx.clear();
  const double margin = 1e-6 + 1e-12 * world_extent;
  for (int v = 0; v < nvol; ++v) {
    const auto &pl = dau[v];
    const int n = (int)pl.size();
    if (n < kGridMin) continue;
    const int ax = spread_axis(pl);
    if (ax < 0) continue;
    // bins about as wide as the mean daughter extent along the axis
    const double h = vol[v].half[ax];
    double mean = 0.0;
    for (const gxt_placement *q : pl) mean += 2.0 * vol[q->daughter].half[ax] / n;
    const double want = std::ceil(2.0 * h / mean);
    const int nb = (int)std::max(1.0, std::min(want, (double)std::min(kGridMaxBins, 2 * n)));
    const double inv = (double)nb / (2.0 * h);
    std::vector<int> b0(n), b1(n);
    size_t total = 0;
    for (int j = 0; j < n; ++j) {
      const double t = pl[j]->translation[ax], hd = vol[pl[j]->daughter].half[ax];
      b0[j] = nav_bin(t - hd - margin, h, inv, nb);
      b1[j] = nav_bin(t + hd + margin, h, inv, nb);
      total += (size_t)(b1[j] - b0[j] + 1);
    }
    if (total > 16 * (size_t)n) continue;
    std::vector<std::vector<int32_t>> bins(nb);
    for (int j = 0; j < n; ++j)
      for (int b = b0[j]; b <= b1[j]; ++b) bins[b].push_back(j);
    const size_t base = idx.size();
    idx.resize(base + nb + 1);
    for (int b = 0; b < nb; ++b) {
      idx[base + b] = (int32_t)idx.size();
      idx.insert(idx.end(), bins[b].begin(), bins[b].end());
    }
    idx[base + nb] = (int32_t)idx.size();
    grid[v] = NavGrid{inv, ax, nb, (int32_t)base, 0};
  }
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(T)); }
};

int hip_fail(hipError_t e, const char *what) { return set_error(GXT_EHIP, std::string(what) + ": " + hipGetErrorString(e)); }

int check_device(const gxt_navigator *nav) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return set_error(GXT_EHIP, "hipGetDevice failed");
  if (dev != nav->device)
    return set_error(GXT_EINVAL, "navigator built on device " + std::to_string(nav->device) +
                                     ", current device is " + std::to_string(dev));
  return GXT_OK;
}

// Finite positions and unit directions (to 1e-6) for n points.
int check_points(const double *pos, const double *dir, int n) {
  for (int i = 0; i < n; ++i) {
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    const double v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    if (!finite3(p) || !finite3(v) || std::fabs(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - 1.0) > 1e-6)
      return set_error(GXT_EINVAL, "point " + std::to_string(i) + ": need a finite position and a unit direction");
  }
  return GXT_OK;
}

// Where the n arrays of a query live: 1 when every one is device memory of
// the current device, 0 when none is, -1 when they are mixed.
int residence(const void *const *ptr, int n, int dev) {
  int ondev = 0;
  for (int i = 0; i < n; ++i) {
    hipPointerAttribute_t at;
    std::memset(&at, 0, sizeof(at));
    const hipError_t e = hipPointerGetAttributes(&at, ptr[i]);
    if (e != hipSuccess) {
      (void)hipGetLastError();  // plain host memory: the query's error is not kept
      continue;
    }
    if (at.type == hipMemoryTypeDevice && at.device == dev) ++ondev;
  }
  return ondev == n ? 1 : ondev == 0 ? 0 : -1;
}

}  // namespace

extern "C" {

int gxt_nav_create(const gxt_box_volume *volumes, int n_volumes, const gxt_placement *placements, int n_placements,
                   int world, gxt_navigator **out) {
  if (!volumes || n_volumes <= 0 || n_placements < 0 || (n_placements > 0 && !placements) || world < 0 ||
      world >= n_volumes || !out)
    return set_error(GXT_EINVAL, "bad navigator arguments");
  gxt_navigator *nav = new gxt_navigator();
  std::vector<std::vector<const gxt_placement *>> daughters;
  int rc = build_tree(volumes, n_volumes, placements, n_placements, world, nav->node, daughters);
  if (rc == GXT_OK) {
    const double *wh = volumes[world].half;
    build_index(volumes, n_volumes, daughters, std::max(wh[0], std::max(wh[1], wh[2])), nav->grid, nav->idx);
    // nodes, grids and index in one allocation (the grids start 64 B-aligned after the nodes)
    std::vector<char> blob(nav->bytes_padded(), 0);
    char *q = blob.data();
    std::memcpy(q, nav->node.data(), nav->node.size() * sizeof(NavNode));
    q += nav->node.size() * sizeof(NavNode);
    std::memcpy(q, nav->grid.data(), nav->grid.size() * sizeof(NavGrid));
    q += nav->grid.size() * sizeof(NavGrid);
    if (!nav->idx.empty()) std::memcpy(q, nav->idx.data(), nav->idx.size() * sizeof(int32_t));
    hipError_t e = hipGetDevice(&nav->device);
    if (e == hipSuccess) e = hipMalloc(&nav->dev, blob.size());
    if (e == hipSuccess) e = hipMemcpy(nav->dev, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "navigator upload");
  }
  if (rc) {
    if (nav->dev) (void)hipFree(nav->dev);
    delete nav;
    return rc;
  }
  *out = nav;
  return GXT_OK;
}

int gxt_nav_free(gxt_navigator *nav) {
  if (!nav) return set_error(GXT_EINVAL, "null navigator");
  if (nav->dev) (void)hipFree(nav->dev);
  delete nav;
  return GXT_OK;
}

int gxt_nav_info(const gxt_navigator *nav, int *n_touchables, int32_t *volume, int32_t *parent, double *centre) {
  if (!nav) return set_error(GXT_EINVAL, "null navigator");
  const size_t n = nav->node.size();
  if (n_touchables) *n_touchables = (int)n;
  for (size_t t = 0; t < n; ++t) {
    const NavNode &d = nav->node[t];
    if (volume) volume[t] = d.volume;
    if (parent) parent[t] = d.parent;
    if (centre)
      for (int k = 0; k < 3; ++k) centre[3 * t + k] = d.c[k];
  }
  return GXT_OK;
}

int gxt_nav_locate(const gxt_navigator *nav, const double *pos, const double *dir, int32_t *touchable, int n) {
  if (!nav || !nav->dev || !pos || !dir || !touchable || n <= 0 || n > kMaxSoaPoints)
    return set_error(GXT_EINVAL, "bad locate arguments (1 <= n <= 2^29)");
  if (const int rc = check_device(nav)) return rc;
  const void *arrays[3] = {pos, dir, touchable};
  const int res = residence(arrays, 3, nav->device);
  if (res < 0) return set_error(GXT_EINVAL, "locate arrays must be all host or all device memory");
  if (res == 1) {
    const hipError_t e = launch_nav_locate(nav->view(), nav->lds_bytes(), pos, dir, touchable, n, 0);
    return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_locate");
  }
  if (const int rc = check_points(pos, dir, n)) return rc;
  const size_t n3 = 3 * (size_t)n * sizeof(double);
  DevBuf<double> dp, dd;
  DevBuf<int32_t> dt;
  hipError_t e = dp.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dd.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dt.alloc(n);
  if (e == hipSuccess) e = hipMemcpy(dp.p, pos, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd.p, dir, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_nav_locate(nav->view(), nav->lds_bytes(), dp.p, dd.p, dt.p, n, 0);
  if (e == hipSuccess) e = hipMemcpy(touchable, dt.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost);
  return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_locate");
}

int gxt_nav_step(const gxt_navigator *nav, const int32_t *touchable, const double *pos, const double *dir,
                 double *snext, int32_t *next, double *safety, int n) {
  if (!nav || !nav->dev || !touchable || !pos || !dir || !snext || !next || !safety || n <= 0 || n > kMaxSoaPoints)
    return set_error(GXT_EINVAL, "bad step arguments");
  if (const int rc = check_device(nav)) return rc;
  const int nt = (int)nav->node.size();
  const void *arrays[6] = {touchable, pos, dir, snext, next, safety};
  const int res = residence(arrays, 6, nav->device);
  if (res < 0) return set_error(GXT_EINVAL, "step arrays must be all host or all device memory");
  if (res == 1) {
    const hipError_t e = launch_nav_step(nav->view(), nav->lds_bytes(), touchable, pos, dir, snext, next, safety, n, 0);
    return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_step");
  }
  for (int i = 0; i < n; ++i)
    if (touchable[i] < 0 || touchable[i] >= nt)
      return set_error(GXT_EINVAL, "point " + std::to_string(i) + ": touchable out of range");
  if (const int rc = check_points(pos, dir, n)) return rc;
  const size_t n1 = (size_t)n * sizeof(double), n3 = 3 * n1, ni = (size_t)n * sizeof(int32_t);
  DevBuf<double> dp, dd, ds, dsaf;
  DevBuf<int32_t> dt, dn;
  hipError_t e = dp.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dd.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = ds.alloc(n);
  if (e == hipSuccess) e = dsaf.alloc(n);
  if (e == hipSuccess) e = dt.alloc(n);
  if (e == hipSuccess) e = dn.alloc(n);
  if (e == hipSuccess) e = hipMemcpy(dp.p, pos, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd.p, dir, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt.p, touchable, ni, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_nav_step(nav->view(), nav->lds_bytes(), dt.p, dp.p, dd.p, ds.p, dn.p, dsaf.p, n, 0);
  if (e == hipSuccess) e = hipMemcpy(snext, ds.p, n1, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(next, dn.p, ni, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(safety, dsaf.p, n1, hipMemcpyDeviceToHost);
  return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_step");
}

}  // extern "C"

extern "C" int gxt_nav_find_next_boundary(const gxt_navigator *nav, const int32_t *touchable, const double *pos,
                                          const double *dir, const double *pstep, double *snext, int32_t *next,
                                          double *safety, int32_t *boundary, int n) {
  if (!nav || !nav->dev || !touchable || !pos || !dir || !pstep || !snext || !next || !safety || !boundary || n <= 0 ||
      n > kMaxSoaPoints)
    return set_error(GXT_EINVAL, "bad find-next-boundary arguments");
  if (const int rc = check_device(nav)) return rc;
  const int nt = (int)nav->node.size();
  const void *arrays[8] = {touchable, pos, dir, pstep, snext, next, safety, boundary};
  const int res = residence(arrays, 8, nav->device);
  if (res < 0) return set_error(GXT_EINVAL, "find-next-boundary arrays must be all host or all device memory");
  if (res == 1) {
    const hipError_t e =
        launch_nav_find_next_boundary(nav->view(), nav->lds_bytes(), touchable, pos, dir, pstep, snext, next, safety, boundary, n, 0);
    return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_find_next_boundary");
  }
  for (int i = 0; i < n; ++i) {
    if (touchable[i] < 0 || touchable[i] >= nt)
      return set_error(GXT_EINVAL, "point " + std::to_string(i) + ": touchable out of range");
    if (!(pstep[i] >= 0.0) || !(pstep[i] <= 1e300))
      return set_error(GXT_EINVAL, "point " + std::to_string(i) + ": pstep must be finite and >= 0");
  }
  if (const int rc = check_points(pos, dir, n)) return rc;
  const size_t n1 = (size_t)n * sizeof(double), n3 = 3 * n1, ni = (size_t)n * sizeof(int32_t);
  DevBuf<double> dp, dd, dps, ds, dsaf;
  DevBuf<int32_t> dt, dn, db;
  hipError_t e = dp.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dd.alloc(3 * (size_t)n);
  if (e == hipSuccess) e = dps.alloc(n);
  if (e == hipSuccess) e = ds.alloc(n);
  if (e == hipSuccess) e = dsaf.alloc(n);
  if (e == hipSuccess) e = dt.alloc(n);
  if (e == hipSuccess) e = dn.alloc(n);
  if (e == hipSuccess) e = db.alloc(n);
  if (e == hipSuccess) e = hipMemcpy(dp.p, pos, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dd.p, dir, n3, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dps.p, pstep, n1, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt.p, touchable, ni, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_nav_find_next_boundary(nav->view(), nav->lds_bytes(), dt.p, dp.p, dd.p, dps.p, ds.p, dn.p, dsaf.p, db.p, n, 0);
  if (e == hipSuccess) e = hipMemcpy(snext, ds.p, n1, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(next, dn.p, ni, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(safety, dsaf.p, n1, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(boundary, db.p, ni, hipMemcpyDeviceToHost);
  return e == hipSuccess ? GXT_OK : hip_fail(e, "gxt_nav_find_next_boundary");
}
