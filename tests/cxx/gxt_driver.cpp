// gxt_driver.cpp -- a plain C++ caller of libgxt.so through include/gxt.h,
// the way the reference's C++ host code would bind it: no Python, no torch.
//
//   gxt_driver DATA_DIR block N   10 GeV e- x N into the 30 cm PbWO4 block
//                                 (seed 777); prints one line per event
//   gxt_driver DATA_DIR nav       the reference calorimeter as a volume
//                                 tree; FindNextBoundary for fixed tracks
//   gxt_driver DATA_DIR basket FILE N
//                                 a basket of tracks in flight (raw
//                                 gxt_track records in FILE, N events)
//                                 continued in the PbWO4 block; one line
//                                 per event
//   gxt_driver DATA_DIR slab      gxt_slab_distance of fixed points in the
//                                 PbWO4 block; one line per point
//   gxt_driver DATA_DIR tree N    10 GeV e- x N through the reference
//                                 calorimeter as the transport's volume
//                                 tree; one line per event
//
// Test infrastructure: tests/test_gpu_cxx_driver.py compares its output with
// the Python binding and the oracle.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/gxt.h"

static int fail(const char *what) {
  std::fprintf(stderr, "%s: %s\n", what, gxt_last_error());
  return 1;
}

// The 30 cm PbWO4 block, seed 777, n event slots.
static gxt_ctx *block_context(const char *data_dir, int n) {
  gxt_config cfg;
  if (gxt_default_config(&cfg) != GXT_OK) return nullptr;
  cfg.geometry = GXT_GEOM_BOX;
  cfg.material = GXT_MAT_PBWO4;
  cfg.box_xy = 300.0;
  cfg.box_depth = 300.0;
  cfg.seed = 777;
  cfg.n_zbins = 30;
  cfg.n_rbins = 15;
  cfg.max_events = n;
  cfg.data_dir = data_dir;
  gxt_ctx *ctx = nullptr;
  if (gxt_create(&cfg, &ctx) != GXT_OK) return nullptr;
  return ctx;
}

static void print_events(const std::vector<gxt_event_tally> &t) {
  for (size_t i = 0; i < t.size(); ++i)
    std::printf("event %d %.17g %.17g %.17g %lld %lld %lld %lld\n", (int)i, t[i].edep, t[i].visible, t[i].leak,
                (long long)t[i].n_steps, (long long)t[i].n_secondaries[0], (long long)t[i].n_secondaries[1],
                (long long)t[i].n_secondaries[2]);
}

static int run_block(const char *data_dir, int n) {
  gxt_ctx *ctx = block_context(data_dir, n);
  if (!ctx) return fail("gxt_create");
  std::vector<gxt_primary> prim(n);
  if (gxt_make_primaries(777, 1, 10000.0, 10000.0, 0.0, n, 0, n, prim.data()) != GXT_OK)
    return fail("gxt_make_primaries");
  std::vector<gxt_event_tally> t(n);
  if (gxt_submit(ctx, prim.data(), n, 1) != GXT_OK || gxt_run(ctx, 0) != GXT_OK ||
      gxt_get_event_tallies(ctx, 1, t.data(), n) != GXT_OK) {
    std::fprintf(stderr, "transport: %s\n", gxt_ctx_error(ctx));
    gxt_destroy(ctx);
    return 1;
  }
  print_events(t);
  gxt_destroy(ctx);
  return 0;
}

// A basket handed over by a CPU transport: raw gxt_track records.
static int run_basket(const char *data_dir, const char *path, int n_events) {
  std::FILE *f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    return 1;
  }
  std::vector<gxt_track> trk;
  gxt_track rec;
  while (std::fread(&rec, sizeof rec, 1, f) == 1) trk.push_back(rec);
  std::fclose(f);
  gxt_ctx *ctx = block_context(data_dir, n_events);
  if (!ctx) return fail("gxt_create");
  std::vector<gxt_event_tally> t(n_events);
  if (gxt_submit_tracks(ctx, trk.data(), (int)trk.size(), n_events, 3) != GXT_OK || gxt_run(ctx, 0) != GXT_OK ||
      gxt_get_event_tallies(ctx, 3, t.data(), n_events) != GXT_OK) {
    std::fprintf(stderr, "basket: %s\n", gxt_ctx_error(ctx));
    gxt_destroy(ctx);
    return 1;
  }
  std::printf("tracks %d\n", (int)trk.size());
  print_events(t);
  gxt_destroy(ctx);
  return 0;
}

// gxt_slab_distance in the block: from inside, on the faces, from outside.
static int run_slab(const char *data_dir) {
  gxt_ctx *ctx = block_context(data_dir, 1);
  if (!ctx) return fail("gxt_create");
  const int n = 6;
  const double px[n] = {0.0, 10.0, 150.0, -149.5, 0.0, 200.0};
  const double py[n] = {0.0, -20.0, 0.0, 3.0, 0.0, 0.0};
  const double pz[n] = {0.0, 150.0, 30.0, 299.0, 300.0, 10.0};
  const double dx[n] = {0.0, 0.6, 1.0, -0.6, 0.0, -1.0};
  const double dy[n] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double dz[n] = {1.0, 0.8, 0.0, 0.8, -1.0, 0.0};
  const double pstep[n] = {5.0, 1000.0, 1.0, 0.25, 1e20, 3.0};
  std::vector<double> pos(3 * n), dir(3 * n), snext(n), safety(n);
  std::vector<int32_t> vol(n, -1), next(n), boundary(n);
  for (int i = 0; i < n; ++i) {
    pos[i] = px[i], pos[n + i] = py[i], pos[2 * n + i] = pz[i];
    dir[i] = dx[i], dir[n + i] = dy[i], dir[2 * n + i] = dz[i];
  }
  if (gxt_slab_distance(ctx, pos.data(), dir.data(), pstep, vol.data(), snext.data(), next.data(), safety.data(),
                        boundary.data(), n) != GXT_OK) {
    std::fprintf(stderr, "slab: %s\n", gxt_ctx_error(ctx));
    gxt_destroy(ctx);
    return 1;
  }
  for (int i = 0; i < n; ++i)
    std::printf("point %d %d %.17g %d %.17g %d\n", i, vol[i], snext[i], next[i], safety[i], boundary[i]);
  gxt_destroy(ctx);
  return 0;
}

// ProxyDetectorConstruction::CreateGeometry's calorimeter
// (proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:225-325):
// world (vacuum) > calo > 50 layers along x > 10 mm Pb + 10 mm lAr, 1 m square.
static void reference_calorimeter(std::vector<gxt_box_volume> &vol, std::vector<gxt_placement> &pl) {
  const int nl = 50;
  const double th[2] = {10.0, 10.0}, yz = 1000.0, layer = th[0] + th[1], calo = nl * layer;
  vol.resize(5);
  const double halves[5][3] = {{0.6 * calo, 0.6 * yz, 0.6 * yz},
                               {0.5 * layer, 0.5 * yz, 0.5 * yz},
                               {0.5 * calo, 0.5 * yz, 0.5 * yz},
                               {0.5 * th[0], 0.5 * yz, 0.5 * yz},
                               {0.5 * th[1], 0.5 * yz, 0.5 * yz}};
  const int mats[5] = {-1, -1, -1, GXT_MAT_PB, GXT_MAT_LAR};
  for (int v = 0; v < 5; ++v) {
    std::memcpy(vol[v].half, halves[v], sizeof(halves[v]));
    vol[v].material = mats[v];
    vol[v].pad = 0;
  }
  pl.clear();
  double start = -0.5 * layer;
  for (int i = 0; i < 2; ++i) {
    pl.push_back(gxt_placement{1, 3 + i, {start + 0.5 * th[i], 0.0, 0.0}});
    start += th[i];
  }
  start = -0.5 * calo;
  for (int i = 0; i < nl; ++i) {
    pl.push_back(gxt_placement{2, 1, {start + 0.5 * layer, 0.0, 0.0}});
    start += layer;
  }
  pl.push_back(gxt_placement{0, 2, {0.0, 0.0, 0.0}});
}

static int run_nav() {
  std::vector<gxt_box_volume> vol;
  std::vector<gxt_placement> pl;
  reference_calorimeter(vol, pl);
  gxt_navigator *nav = nullptr;
  if (gxt_nav_create(vol.data(), (int)vol.size(), pl.data(), (int)pl.size(), 0, &nav) != GXT_OK)
    return fail("gxt_nav_create");
  int nt = 0;
  gxt_nav_info(nav, &nt, nullptr, nullptr, nullptr);
  std::printf("touchables %d\n", nt);
  // tracks as [3n] SoA: along +x through the first slabs, oblique, on faces, outside the calorimeter
  const int n = 6;
  const double px[n] = {-495.0, -495.0, -490.0, 0.0, 13.7, 550.0};
  const double py[n] = {0.0, 0.0, 0.0, 499.0, -20.0, 0.0};
  const double pz[n] = {0.0, 0.0, 0.0, 0.0, 7.5, 0.0};
  const double dx[n] = {1.0, 1.0, 1.0, 0.6, -0.48, -1.0};
  const double dy[n] = {0.0, 0.0, 0.0, 0.8, 0.6, 0.0};
  const double dz[n] = {0.0, 0.0, 0.0, 0.0, 0.64, 0.0};
  const double pstep[n] = {1.0, 7.0, 0.0, 1e20, 3.0, 1e20};
  std::vector<double> pos(3 * n), dir(3 * n);
  for (int i = 0; i < n; ++i) {
    pos[i] = px[i], pos[n + i] = py[i], pos[2 * n + i] = pz[i];
    dir[i] = dx[i], dir[n + i] = dy[i], dir[2 * n + i] = dz[i];
  }
  std::vector<int32_t> t(n), next(n), boundary(n);
  std::vector<double> snext(n), safety(n);
  if (gxt_nav_locate(nav, pos.data(), dir.data(), t.data(), n) != GXT_OK) return fail("gxt_nav_locate");
  if (gxt_nav_find_next_boundary(nav, t.data(), pos.data(), dir.data(), pstep, snext.data(), next.data(),
                                 safety.data(), boundary.data(), n) != GXT_OK)
    return fail("gxt_nav_find_next_boundary");
  for (int i = 0; i < n; ++i)
    std::printf("track %d %d %.17g %d %.17g %d\n", i, t[i], snext[i], next[i], safety[i], boundary[i]);
  gxt_nav_free(nav);
  return 0;
}

// Transport through that calorimeter as the context's geometry
// (GXT_GEOM_TREE): n events of one 10 GeV e- each from (0.5, 0.25, 0.75) mm
// along (2, 1, 2)/3, substream = event; one line per event.
static int run_tree(const char *data_dir, int n) {
  std::vector<gxt_box_volume> vol;
  std::vector<gxt_placement> pl;
  reference_calorimeter(vol, pl);
  gxt_config cfg;
  if (gxt_default_config(&cfg) != GXT_OK) return fail("gxt_default_config");
  cfg.geometry = GXT_GEOM_TREE;
  cfg.tree_volumes = vol.data();
  cfg.tree_placements = pl.data();
  cfg.n_tree_volumes = (int)vol.size();
  cfg.n_tree_placements = (int)pl.size();
  cfg.tree_world = 0;
  cfg.absorber = GXT_MAT_PB;
  cfg.gap = GXT_MAT_LAR;
  cfg.max_events = n;
  cfg.data_dir = data_dir;
  gxt_ctx *ctx = nullptr;
  if (gxt_create(&cfg, &ctx) != GXT_OK) return fail("gxt_create");
  int nbins = 0;
  if (gxt_volume_bins(ctx, &nbins, nullptr) != GXT_OK) return fail("gxt_volume_bins");
  std::printf("bins %d\n", nbins);
  std::vector<gxt_primary> prim(n);
  for (int i = 0; i < n; ++i) {
    prim[i].species = 1;
    prim[i].event = i;
    prim[i].ekin = 10000.0;
    prim[i].pos[0] = 0.5, prim[i].pos[1] = 0.25, prim[i].pos[2] = 0.75;
    prim[i].dir[0] = 2.0 / 3.0, prim[i].dir[1] = 1.0 / 3.0, prim[i].dir[2] = 2.0 / 3.0;
    prim[i].substream = (uint64_t)i;
  }
  std::vector<gxt_event_tally> t(n);
  if (gxt_submit(ctx, prim.data(), n, 1) != GXT_OK || gxt_run(ctx, 0) != GXT_OK ||
      gxt_get_event_tallies(ctx, 1, t.data(), n) != GXT_OK) {
    std::fprintf(stderr, "transport: %s\n", gxt_ctx_error(ctx));
    gxt_destroy(ctx);
    return 1;
  }
  print_events(t);
  gxt_destroy(ctx);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s DATA_DIR block N | nav | basket FILE N | slab | tree N\n", argv[0]);
    return 2;
  }
  if (!std::strcmp(argv[2], "block")) return run_block(argv[1], argc > 3 ? std::atoi(argv[3]) : 16);
  if (!std::strcmp(argv[2], "nav")) return run_nav();
  if (!std::strcmp(argv[2], "basket") && argc > 4) return run_basket(argv[1], argv[3], std::atoi(argv[4]));
  if (!std::strcmp(argv[2], "slab")) return run_slab(argv[1]);
  if (!std::strcmp(argv[2], "tree")) return run_tree(argv[1], argc > 3 ? std::atoi(argv[3]) : 8);
  std::fprintf(stderr, "unknown mode %s\n", argv[2]);
  return 2;
}
