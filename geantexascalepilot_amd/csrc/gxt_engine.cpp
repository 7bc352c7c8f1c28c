// gxt_engine.cpp -- host driver of the step loop and the C ABI of libgxt.so.
//
// One gxt_ctx owns one HIP stream, the device copy of the physics data, a
// double-buffered SoA track pool, the interaction queues and the tallies.
// Jobs (groups of primaries) are admitted into the shared pool by estimated
// footprint; the device step loop runs several iterations between host
// synchronisations; a job completes when its live-track counter reaches 0.
// This replaces ProxyDeviceManager (proxy/src/ProxyDeviceManager.cpp:41-276)
// and the host wave loop DoStep<> (proxy/test/test_BasicCpuTransport.cpp:517-550).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gxt.h"
#include "gxt_data.h"
#include "gxt_field.h"
#include "gxt_kernels.h"

using namespace gxt;

namespace {

thread_local std::string g_last_error;

int set_err(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

}  // namespace

// the library's last-error message, set by every failing entry point
int gxt::set_error(int code, const std::string &msg) { return set_err(code, msg); }

namespace {

// Largest table copy the transports stage in LDS (beside 11 KB of wave
// histograms per workgroup).
constexpr int kLdsTablesMax = 40 * 1024;

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return set_err(GXT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
hipError_t dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  return hipMalloc((void **)p, n * sizeof(T));
}

// Device buffer of a component entry point, freed when it goes out of scope
// (also on the early returns of HIP_TRY).  n == 0: no buffer, p null.
template <class T>
struct DevBuf {
  T *p = nullptr;
  hipError_t err = hipSuccess;
  explicit DevBuf(size_t n) {
    if (n) err = hipMalloc((void **)&p, n * sizeof(T));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// ------------------------------------------------ gxt_sample_noalloc staging
// The input and output arrays of up to GXT_SAMPLE_NOALLOC_MAX tracks of
// gxt_sample_noalloc live in one file-scope device array, `scratch` (a copy
// on every GPU the library runs on, held until the process ends: 11.3 MB).
// Its device address is looked up once per device, with that device current,
// when a context is created; host code uses only that address (the array's
// name is a host-side stand-in).  Calls that overlap on one GPU share it and
// race.
constexpr int kNoAllocSampleTracks = GXT_SAMPLE_NOALLOC_MAX;
constexpr int kMaxDevices = 64;
constexpr size_t stage_align(size_t x) { return (x + 255) & ~(size_t)255; }
// Byte offsets of the staged arrays of n tracks (each 256-byte aligned).
struct SampleStaging {
  size_t ek, dir, rng, oek, odir, oed, nsec, ssp, sek, sdir, total;
  constexpr explicit SampleStaging(size_t n)
      : ek(0),
        dir(stage_align(ek + 8 * n)),
        rng(stage_align(dir + 24 * n)),
        oek(stage_align(rng + 24 * n)),
        odir(stage_align(oek + 8 * n)),
        oed(stage_align(odir + 24 * n)),
        nsec(stage_align(oed + 8 * n)),
        ssp(stage_align(nsec + 4 * n)),
        sek(stage_align(ssp + 8 * n)),
        sdir(stage_align(sek + 16 * n)),
        total(stage_align(sdir + 48 * n)) {}
};
constexpr size_t kScratchBytes = SampleStaging(kNoAllocSampleTracks).total;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Record scratch's address on `device` (the current device), once.
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return set_err(GXT_EHIP, "device index beyond the staging table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return GXT_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(GXT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return GXT_OK;
}

struct DeviceData {
  DevParams *P = nullptr;
  Node64 *nd = nullptr;
  Node32 *nf = nullptr;
  double *sbd = nullptr, *sandia = nullptr, *sx = nullptr;
  float *sbf = nullptr;
  uint32_t *jump8 = nullptr;
  int *hints = nullptr;
  void release() {
    (void)hipFree(hints);
    (void)hipFree(P);
    (void)hipFree(jump8);
    (void)hipFree(nd);
    (void)hipFree(nf);
    (void)hipFree(sbd);
    (void)hipFree(sandia);
    (void)hipFree(sx);
    (void)hipFree(sbf);
    *this = DeviceData();
  }
};

// Upload tables, grids and parameters; fills hd.params' device pointers.
int upload(HostData &hd, DeviceData &dd) {
  auto tof = [](const std::vector<double> &v) {
    std::vector<float> f(v.size());
    for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
    return f;
  };
  auto put = [](auto **dst, const auto &v) -> hipError_t {
    hipError_t e = dalloc(dst, v.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
  };
  std::vector<Node64> n64(hd.x.size());
  std::vector<Node32> n32(hd.x.size());
  for (size_t i = 0; i < hd.x.size(); ++i) {
    n64[i] = Node64{hd.x[i], hd.y[i], hd.d2[i], 0.0};
    n32[i] = Node32{(float)hd.x[i], (float)hd.y[i], (float)hd.d2[i], 0.0f};
  }
  HIP_TRY(put(&dd.nd, n64));
  HIP_TRY(put(&dd.nf, n32));
  HIP_TRY(put(&dd.sbd, hd.sb));
  HIP_TRY(put(&dd.sbf, tof(hd.sb)));
  HIP_TRY(put(&dd.sandia, hd.sandia));
  HIP_TRY(put(&dd.sx, hd.sx));
  HIP_TRY(put(&dd.hints, hd.hints));
  DevParams &p = hd.params;
  p.hints = dd.hints;
  uint32_t jump[2][64][9];
  build_jump_matrices(jump);
  std::vector<uint32_t> j8;
  build_jump_tables(jump, j8);
  HIP_TRY(put(&dd.jump8, j8));
  p.nsandia = (int)hd.sandia.size();
  p.nshells = (int)hd.shells.size();
  if (hd.shells.size() > (size_t)kMaxShells) return set_err(GXT_EDATA, "more shell energies than DevParams holds");
  for (size_t i = 0; i < hd.shells.size(); ++i) p.shell_e[i] = hd.shells[i];
  p.jump8 = dd.jump8;
  p.nd = dd.nd;
  p.nf = dd.nf;
  p.sbd = dd.sbd;
  p.sbf = dd.sbf;
  p.sandia = dd.sandia;
  p.sx = dd.sx;
  HIP_TRY(dalloc(&dd.P, 1));
  HIP_TRY(hipMemcpy(dd.P, &p, sizeof(p), hipMemcpyHostToDevice));
  return GXT_OK;
}

template <class R>
struct DevPool {
  Pool<R> p{};
  int alloc(int64_t cap, bool trace) {
    p.cap = cap;
    HIP_TRY(dalloc(&p.x, cap));
    HIP_TRY(dalloc(&p.y, cap));
    HIP_TRY(dalloc(&p.z, cap));
    HIP_TRY(dalloc(&p.dx, cap));
    HIP_TRY(dalloc(&p.dy, cap));
    HIP_TRY(dalloc(&p.dz, cap));
    HIP_TRY(dalloc(&p.e, cap));
    HIP_TRY(dalloc(&p.nil0, cap));
    HIP_TRY(dalloc(&p.nil1, cap));
    HIP_TRY(dalloc(&p.nil2, cap));
    HIP_TRY(dalloc(&p.rng, 6 * cap));
    HIP_TRY(dalloc(&p.event, cap));
    HIP_TRY(dalloc(&p.meta, cap));
    HIP_TRY(dalloc(&p.nsteps, cap));
    HIP_TRY(dalloc(&p.aed, cap));
    HIP_TRY(dalloc(&p.avis, cap));
    HIP_TRY(dalloc(&p.alen, cap));
    if (trace) {
      HIP_TRY(dalloc(&p.uid, cap));
      HIP_TRY(dalloc(&p.tedep, cap));
    }
    return GXT_OK;
  }
  void release() {
    void *ptrs[] = {p.x, p.y, p.z, p.dx, p.dy, p.dz, p.e, p.nil0, p.nil1, p.nil2, p.rng, p.event, p.meta,
                    p.nsteps, p.aed, p.avis, p.alen, p.uid, p.tedep};
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
    p = Pool<R>{};
  }
};

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

struct Job {
  uint64_t id = 0;
  std::vector<gxt_primary> prims;   // sorted by event (a job of primaries)
  gxt_primary *dprims = nullptr;    // device copy of prims (made at submission, freed when the job ends)
  std::vector<gxt_track> tracks;    // sorted by event (a job of tracks in flight)
  gxt_track *dtracks = nullptr;     // device copy of tracks
  std::vector<int> ev_first;        // entries [ev_first[e] .. ev_first[e+1]) belong to event e
  std::vector<double> ev_gev;       // summed primary energy of each event (GeV)
  std::vector<int> ev_max;          // highest live-track count seen for an event in flight
  std::vector<char> ev_state;       // 0 pending, 1 in flight, 2 finished
  int n_events = 0;
  int next_event = 0;               // first event not yet injected
  int n_done = 0;
  int ev_base = -1;                 // first event slot of the job
  bool started = false;
  double submit_ms = 0, admit_ms = 0, done_ms = 0;
  int64_t admit_iter = 0, done_iter = 0;  // step-loop iterations launched at admission / at the finding snapshot
  std::vector<gxt_event_tally> tallies;
  int64_t steps = 0;
};

}  // namespace

struct gxt_ctx {
  gxt_config cfg{};
  std::string data_dir;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  StepStreams ss{};  // main = stream, side stream; fork, join and entry events
  HostData hd;
  DeviceData dd;
  int precision = GXT_FP64;
  int64_t cap = 0;
  bool trace = false;
  DevPool<double> pd[2];
  DevPool<float> pf[2];
  int parity = 0;
  int grid = 1024;
  int32_t *queues = nullptr;
  int32_t *newlist = nullptr;
  Counters *cnt = nullptr;
  Counters *hcnt = nullptr;  // pinned
  DevEvent *ev = nullptr;
  int *ev_live = nullptr;    // live tracks of event slot e at ev_live[e * kLiveStride]
  int *hev_live = nullptr;   // pinned: ev_live[e * kLiveStride] at [e], the host's snapshot
  Counters *dhcnt = nullptr;  // device addresses of hcnt and hev_live
  int *dhev_live = nullptr;
  double *rows = nullptr;
  unsigned long long *spec = nullptr;
  int64_t row_stride = 0;
  double *trace_buf = nullptr;
  int64_t trace_cap = 0;
  hipEvent_t snap = nullptr;      // recorded after a round's counter snapshot
  size_t ev_snap = 0;             // profiling entries recorded before that snapshot
  int64_t snap_iter = 0;          // step-loop iterations launched before that snapshot
  int gen = 0;                    // counter generation of the next iteration's current pool
  int snap_gen = 0;               // the same when the snapshot was queued
  bool side_behind = true;        // main holds work the e-/e+ chain must follow (launch_side_sync)
  // non-uniform field
  gxt_field_map fmap;
  std::string fmap_path;
  // tree form of the geometry: the caller's volumes and placements (copied),
  // the navigator and the touchable -> tally cell map on the device
  std::vector<gxt_box_volume> tree_vol;
  std::vector<gxt_placement> tree_pl;
  gxt_navigator *nav = nullptr;
  int32_t *cell = nullptr;
  int transport_mode = 0;  // e-/e+ transport kernel (StepArgs::transport_mode)
  int inner_steps = 4;     // StepArgs::inner_steps
  // jobs
  std::mutex mu;
  std::deque<std::unique_ptr<Job>> jobs;   // submitted, not finished (FIFO)
  std::vector<std::unique_ptr<Job>> done_unpolled;
  std::map<uint64_t, std::unique_ptr<Job>> finished;
  std::vector<char> ev_used;
  double committed = 0;             // summed footprint of events in flight
  double tracks_per_gev = 300.0;    // estimated peak live tracks per GeV of an event (learned)
  double seen_sum = 0;              // summed peak per GeV of the events past their maximum
  int64_t seen_events = 0;          // events past their maximum seen so far
  std::vector<char> ev_counted;     // event slot already counted in seen_events
  // kernel timing (profile_kernels)
  bool prof = false;
  int prof_level = 1;
  std::vector<std::array<hipEvent_t, kStepEvents>> ev_free, ev_pending;
  gxt_kernel_stats kst{};
  // stats
  int64_t iterations = 0;
  int64_t max_live = 0;
  double wall_ms = 0;
  std::chrono::steady_clock::time_point t0;
  // serving
  std::thread server;
  std::atomic<bool> serving{false};
  std::atomic<bool> stop{false};
  std::atomic<int> serve_rc{GXT_OK};  // first error of the serving thread
  std::mutex err_mu;                  // guards err (the serving thread writes it)

  StepArgs args() {
    StepArgs a;
    a.P = dd.P;
    a.cnt = cnt;
    a.ev = ev;
    a.ev_live = ev_live;
    a.queues = queues;
    a.newlist = newlist;
    a.rows.bins = rows;
    a.rows.spec = spec;
    a.rows.stride = row_stride;
    a.rows.nrows = grid;
    a.trace = trace_buf;
    a.trace_cap = trace_cap;
    a.transport_mode = transport_mode;
    a.inner_steps = inner_steps;
    a.gen = gen;
    // the transports stage their tables in LDS when they take at most
    // kLdsTablesMax bytes
    const int nb = precision == GXT_FP32 ? (int)sizeof(Node32) : (int)sizeof(Node64);
    for (int l = 0; l < 2; ++l) {
      const int bytes = hd.params.lds_nodes[l] * nb;
      // (the e-/e+ tables in LDS do not include the multiple-scattering ones)
      const bool lds_ok = l == 0 || transport_mode == 0 || (transport_mode == 4 && !hd.params.msc);
      a.lds_bytes[l] = (bytes > 0 && bytes <= kLdsTablesMax && lds_ok) ? bytes : 0;
    }
    return a;
  }

  int fail(int code, const std::string &m) {
    {
      std::lock_guard<std::mutex> lk(err_mu);
      err = m;
    }
    return set_err(code, m);
  }

  std::string error() {
    std::lock_guard<std::mutex> lk(err_mu);
    return err;
  }
};

namespace {

// Live tracks of the pool of counter generation g.
int64_t live_tracks(const Counters &h, int g) {
  int64_t n = 0;
  for (int s = 0; s < kStripes; ++s) n += (int64_t)h.ctr(g, kPool0, s) + h.ctr(g, kPool1, s);
  return n;
}

// Estimated peak live tracks of an event, from its primaries' energy.
// Peak live tracks expected of event e: 16 per primary plus the learned
// number per GeV.
double event_peak(const gxt_ctx *c, const Job *j, int e) {
  return 16.0 * (j->ev_first[e + 1] - j->ev_first[e]) + c->tracks_per_gev * j->ev_gev[e];
}

// An event in flight whose live count has fallen below 70 % of the highest
// seen, once that reached 64 tracks, is past its shower maximum.
bool past_peak(int live, int mx) { return mx >= 64 && live < 0.7 * mx; }

// Live tracks event e may still hold at once: past its maximum, a margin
// over its current count; before it, its expected peak (or more, if it
// already exceeds it).
double event_footprint(const gxt_ctx *c, const Job *j, int e, int live) {
  const double peak = event_peak(c, j, e);
  if (past_peak(live, j->ev_max[e])) return std::min(peak, 1.25 * live + 16.0);
  return std::max(peak, 1.25 * live);
}

int alloc_events(gxt_ctx *c, int n) {
  const int m = (int)c->ev_used.size();
  int run = 0;
  for (int i = 0; i < m; ++i) {
    run = c->ev_used[i] ? 0 : run + 1;
    if (run == n) {
      const int b = i - n + 1;
      for (int k = b; k <= i; ++k) c->ev_used[k] = 1;
      return b;
    }
  }
  return -1;
}

// Admit events, oldest job first, while their estimated footprint fits the
// pool budget; inject their primaries from the job's device copy, one kernel
// launch per job.  Events are admitted whole (all primaries of an event
// together) and in order, so a job's admitted primaries are contiguous.
int admit(gxt_ctx *c) {
  const double budget = 0.8 * (double)c->cap;
  struct Range {
    const gxt_primary *prim;  // a job of primaries, or
    const gxt_track *trk;     // of tracks in flight
    int n, event_base;
  };
  std::vector<Range> ranges;
  std::vector<std::pair<int, int>> zero_ranges;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    int active = 0;
    for (auto &jp : c->jobs) {
      Job *j = jp.get();
      if (!j->started) {
        if (active >= c->cfg.max_jobs) break;
        const int eb = alloc_events(c, j->n_events);
        if (eb < 0) break;
        j->ev_base = eb;
        j->started = true;
        j->admit_ms = now_ms(c->t0);
        j->admit_iter = c->iterations;
        zero_ranges.emplace_back(eb, j->n_events);
        std::fill(c->ev_counted.begin() + eb, c->ev_counted.begin() + eb + j->n_events, 0);
      }
      ++active;
      bool full = false;
      const int e0 = j->next_event;
      while (j->next_event < j->n_events) {
        const int e = j->next_event;
        const double fp = event_peak(c, j, e);
        if (c->committed > 0 && c->committed + fp > budget) {
          full = true;
          break;
        }
        j->ev_state[e] = 1;
        c->committed += fp;
        j->next_event++;
      }
      if (j->next_event > e0) {
        const int p0 = j->ev_first[e0], p1 = j->ev_first[j->next_event];
        ranges.push_back(Range{j->dprims ? j->dprims + p0 : nullptr, j->dtracks ? j->dtracks + p0 : nullptr,
                               p1 - p0, j->ev_base});
      }
      if (full) break;
    }
  }
  // everything below is queued on the stream behind the work in flight (the
  // jobs' primaries were copied to the device at submission, earlier on the
  // same stream); the slots of a new job belong to events without tracks, so
  // the look-ahead iterations still running do not touch them
  if (!zero_ranges.empty() || !ranges.empty()) c->side_behind = true;
  for (auto &r : zero_ranges) {
    HIP_TRY(hipMemsetAsync(c->ev + r.first, 0, r.second * sizeof(DevEvent), c->stream));
    HIP_TRY(hipMemsetAsync(c->ev_live + (size_t)r.first * kLiveStride, 0, (size_t)r.second * kLiveStride * sizeof(int),
                           c->stream));
  }
  const StepArgs a = c->args();
  for (const Range &r : ranges) {
    hipError_t e;
    if (c->precision == GXT_FP32) {
      PoolArgs<float> pa{c->pf[c->parity].p, c->pf[1 - c->parity].p};
      e = r.prim ? launch_inject<float>(a, pa, r.prim, r.n, r.event_base, c->stream)
                 : launch_inject<float>(a, pa, r.trk, r.n, r.event_base, c->stream);
    } else {
      PoolArgs<double> pa{c->pd[c->parity].p, c->pd[1 - c->parity].p};
      e = r.prim ? launch_inject<double>(a, pa, r.prim, r.n, r.event_base, c->stream)
                 : launch_inject<double>(a, pa, r.trk, r.n, r.event_base, c->stream);
    }
    if (e != hipSuccess) return c->fail(GXT_EHIP, std::string("inject: ") + hipGetErrorString(e));
  }
  return GXT_OK;
}

int iterate(gxt_ctx *c, int k) {
  StepArgs a = c->args();
  // the e-/e+ chain follows what main holds that it must see (injections,
  // cleared event slots; a snapshot, which must see no iteration in
  // progress); otherwise it follows only the photon chains
  if (c->side_behind) {
    const hipError_t se = launch_side_sync(c->ss);
    if (se != hipSuccess) return c->fail(GXT_EHIP, std::string("step: ") + hipGetErrorString(se));
    c->side_behind = false;
  }
  for (int i = 0; i < k; ++i) {
    a.gen = c->gen;
    hipError_t e;
    hipEvent_t *evs = nullptr;
    if (c->prof) {
      if (c->ev_free.empty()) {
        // level 2 times only the photon interactions and the e-/e+ transport
        // (kernels 1 and 3: events 2, 3, 6, 7); the events only time kernels,
        // so they need no system-scope fence
        std::array<hipEvent_t, kStepEvents> q{};
        for (int k = 0; k < kStepEvents; ++k)
          if (c->prof_level == 1 || k / 2 == 1 || k / 2 == 3)
            HIP_TRY(hipEventCreateWithFlags(&q[k], hipEventDisableSystemFence));
        c->ev_free.push_back(q);
      }
      c->ev_pending.push_back(c->ev_free.back());
      c->ev_free.pop_back();
      evs = c->ev_pending.back().data();
    }
    if (c->precision == GXT_FP32) {
      PoolArgs<float> pa{c->pf[c->parity].p, c->pf[1 - c->parity].p};
      e = launch_step<float>(a, pa, c->grid, c->ss, evs);
    } else {
      PoolArgs<double> pa{c->pd[c->parity].p, c->pd[1 - c->parity].p};
      e = launch_step<double>(a, pa, c->grid, c->ss, evs);
    }
    if (e != hipSuccess) return c->fail(GXT_EHIP, std::string("step: ") + hipGetErrorString(e));
    c->parity = 1 - c->parity;
    c->gen = gen_next(c->gen);
    ++c->iterations;
  }
  return GXT_OK;
}

// Read the snapshot of the counters and per-event live counts (complete:
// the caller waited for its event); finish events and jobs.
int collect(gxt_ctx *c, bool &idle) {
  for (size_t qi = 0; qi < c->ev_snap; ++qi) {
    auto &q = c->ev_pending[qi];
    // kernel j ran between events 2j and 2j + 1
    float t[kStepKernels] = {};
    for (int j = 0; j < kStepKernels; ++j)
      if (q[2 * j] && q[2 * j + 1]) HIP_TRY(hipEventElapsedTime(&t[j], q[2 * j], q[2 * j + 1]));
    c->kst.transport_neutral_ms += t[0];
    c->kst.interact_neutral_ms += t[1];
    c->kst.transport_charged_ms += t[3];
    c->kst.interact_charged_ms += t[4];
    c->kst.derive_ms += t[2] + t[5];
    c->kst.launches++;
    c->ev_free.push_back(q);
  }
  c->ev_pending.erase(c->ev_pending.begin(), c->ev_pending.begin() + c->ev_snap);
  c->ev_snap = 0;
  const Counters &h = *c->hcnt;
  if (h.overflow) return c->fail(GXT_ECAPACITY, "track pool or queue overflow (increase pool_capacity)");
  const int64_t live = live_tracks(h, c->snap_gen);
  c->max_live = std::max<int64_t>(c->max_live, live);
  std::vector<Job *> fin;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    // learn the peak per GeV from the events past their maximum or finished:
    // once 32 have been seen, the estimate is their mean ratio plus 30 %
    // (for peaks that fall between synchronisations).  The budget's 20 %
    // headroom covers the spread of single events; the sum over the events
    // in flight follows the mean
    for (auto &jp : c->jobs) {
      Job *j = jp.get();
      if (!j->started) break;
      for (int e = 0; e < j->next_event; ++e) {
        if (j->ev_state[e] != 1) continue;
        const int lv = c->hev_live[j->ev_base + e];
        j->ev_max[e] = std::max(j->ev_max[e], lv);
        const int mx = j->ev_max[e];
        char &counted = c->ev_counted[j->ev_base + e];
        if (!counted && (lv == 0 || past_peak(lv, mx)) && j->ev_gev[e] > 0) {
          counted = 1;
          c->seen_events++;
          c->seen_sum += std::max(0.0, (mx - 16.0 * (j->ev_first[e + 1] - j->ev_first[e])) / j->ev_gev[e]);
        }
      }
    }
    if (c->seen_events >= 32) c->tracks_per_gev = std::max(1.3 * c->seen_sum / c->seen_events, 1.0);
    c->committed = 0;
    for (auto &jp : c->jobs) {
      Job *j = jp.get();
      if (!j->started) break;
      for (int e = 0; e < j->next_event; ++e) {
        if (j->ev_state[e] != 1) continue;
        const int lv = c->hev_live[j->ev_base + e];
        if (lv == 0) {
          j->ev_state[e] = 2;
          j->n_done++;
        } else {
          c->committed += event_footprint(c, j, e, lv);
        }
      }
      if (j->n_done == j->n_events) fin.push_back(j);
    }
  }
  for (Job *j : fin) {
    std::vector<DevEvent> evs(j->n_events);
    HIP_TRY(hipMemcpy(evs.data(), c->ev + j->ev_base, evs.size() * sizeof(DevEvent), hipMemcpyDeviceToHost));
    j->tallies.resize(j->n_events);
    j->steps = 0;
    for (int i = 0; i < j->n_events; ++i) {
      const DevEvent &d = evs[i];
      gxt_event_tally &t = j->tallies[i];
      t.edep = d.edep;
      t.visible = d.visible;
      t.leak = d.leak;
      t.charged_length = d.clen;
      for (int k = 0; k < 3; ++k) t.n_secondaries[k] = (int64_t)d.nsec[k];
      t.n_steps = (int64_t)d.nsteps;
      t.n_pairs = (int64_t)d.npairs;
      t.n_annihil = (int64_t)d.nannihil;
      j->steps += t.n_steps;
    }
    j->done_ms = now_ms(c->t0);
    j->done_iter = c->snap_iter;
    std::lock_guard<std::mutex> lk(c->mu);
    for (int k = 0; k < j->n_events; ++k) c->ev_used[j->ev_base + k] = 0;
    for (auto it = c->jobs.begin(); it != c->jobs.end(); ++it)
      if (it->get() == j) {
        std::unique_ptr<Job> owned = std::move(*it);
        c->jobs.erase(it);
        // its injections ran before the snapshot that showed it finished
        if (owned->dprims) HIP_TRY(hipFreeAsync(owned->dprims, c->stream));
        if (owned->dtracks) HIP_TRY(hipFreeAsync(owned->dtracks, c->stream));
        owned->dprims = nullptr;
        owned->dtracks = nullptr;
        owned->prims.clear();
        owned->prims.shrink_to_fit();
        owned->tracks.clear();
        owned->tracks.shrink_to_fit();
        c->done_unpolled.emplace_back(new Job(*owned));  // completion order for gxt_poll
        c->finished[owned->id] = std::move(owned);
        break;
      }
  }
  bool any_started = false;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    for (auto &jp : c->jobs)
      if (jp->started) any_started = true;
    idle = c->jobs.empty();
  }
  if (!any_started && live != 0)
    return c->fail(GXT_ESTATE, "tracks alive with no active job (live-count bookkeeping error)");
  return GXT_OK;
}

// Iterations queued behind a round's counter snapshot: the device runs them
// while the host reads the snapshot and plans the next admission.
constexpr int kLookahead = 1;

// One scheduling round of k iterations: admit, run k - kLookahead iterations,
// snapshot the counters, queue the look-ahead iterations, then collect from
// the snapshot.  Admission thus sees counts kLookahead iterations old; each
// event in flight reserves its expected peak, which covers the growth.
int round_trip(gxt_ctx *c, int k, bool &idle) {
  int rc = admit(c);
  if (rc) return rc;
  bool any = false;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    for (auto &jp : c->jobs)
      if (jp->started) any = true;
  }
  if (any) {
    rc = iterate(c, std::max(k - kLookahead, 1));
    if (rc) return rc;
  }
  // one kernel writes the snapshot into pinned memory; the event's
  // system-scope release makes it visible to the host
  HIP_TRY(launch_snapshot(c->cnt, c->gen, c->ev_live, c->cfg.max_events, c->dhcnt, c->dhev_live, c->stream));
  HIP_TRY(hipEventRecord(c->snap, c->stream));
  c->side_behind = true;
  c->ev_snap = c->ev_pending.size();
  c->snap_iter = c->iterations;
  c->snap_gen = c->gen;
  if (any) {
    rc = iterate(c, kLookahead);
    if (rc) return rc;
  }
  HIP_TRY(hipEventSynchronize(c->snap));
  return collect(c, idle);
}

int iters_per_sync(gxt_ctx *c) {
  return live_tracks(*c->hcnt, c->snap_gen) > 100000 ? 8 : 4;
}

void serve_loop(gxt_ctx *c) {
  (void)hipSetDevice(c->device);
  while (!c->stop.load()) {
    bool idle = true;
    const int rc = round_trip(c, iters_per_sync(c), idle);
    if (rc) {
      c->serve_rc = rc;
      break;
    }
    if (idle) std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
}

void destroy_ctx(gxt_ctx *c) {
  if (c->serving.load()) {
    c->stop = true;
    if (c->server.joinable()) c->server.join();
  }
  // look-ahead iterations may still be in flight
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 2; ++i) {
    c->pd[i].release();
    c->pf[i].release();
  }
  for (auto &q : c->ev_free)
    for (auto &x : q)
      if (x) (void)hipEventDestroy(x);
  for (auto &q : c->ev_pending)
    for (auto &x : q)
      if (x) (void)hipEventDestroy(x);
  c->dd.release();
  release_field_map(c->fmap);
  if (c->nav) (void)gxt_nav_free(c->nav);
  (void)hipFree(c->cell);
  if (c->snap) (void)hipEventDestroy(c->snap);
  for (auto &jp : c->jobs) {
    if (jp->dprims) (void)hipFree(jp->dprims);
    if (jp->dtracks) (void)hipFree(jp->dtracks);
  }
  (void)hipFree(c->queues);
  (void)hipFree(c->newlist);
  (void)hipFree(c->cnt);
  (void)hipHostFree(c->hcnt);
  (void)hipFree(c->ev);
  (void)hipFree(c->ev_live);
  (void)hipHostFree(c->hev_live);
  (void)hipFree(c->rows);
  (void)hipFree(c->spec);
  (void)hipFree(c->trace_buf);
  if (c->ss.fork) (void)hipEventDestroy(c->ss.fork);
  if (c->ss.join) (void)hipEventDestroy(c->ss.join);
  if (c->ss.entry) (void)hipEventDestroy(c->ss.entry);
  if (c->ss.side) (void)hipStreamDestroy(c->ss.side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// The tree form of the geometry (GXT_GEOM_TREE): the navigator on the
// context's device, and the tally cells: the touchables with a material in
// touchable order (slot 0 absorber, 1 gap, which is visible), then one
// vacuum cell (slot -1).
int build_tree_geometry(gxt_ctx *c, const gxt_config &cfg, DevGeometry &g) {
  c->tree_vol.assign(cfg.tree_volumes, cfg.tree_volumes + cfg.n_tree_volumes);
  c->tree_pl.assign(cfg.tree_placements, cfg.tree_placements + cfg.n_tree_placements);
  c->cfg.tree_volumes = c->tree_vol.data();
  c->cfg.tree_placements = c->tree_pl.empty() ? nullptr : c->tree_pl.data();
  int rc = gxt_nav_create(c->tree_vol.data(), (int)c->tree_vol.size(), c->cfg.tree_placements,
                          (int)c->tree_pl.size(), cfg.tree_world, &c->nav);
  if (rc) return rc;
  int nt = 0;
  const NavNode *node = nav_host_nodes(c->nav, &nt);
  std::vector<int32_t> cell(nt);
  std::memset(g.z, 0, sizeof(g.z));
  g.hx = g.hy = 0.0;
  int k = 0;
  for (int t = 0; t < nt; ++t) {
    const int m = c->tree_vol[node[t].volume].material;
    if (m < 0) continue;
    if (k >= kMaxVol - 1)
      return set_err(GXT_EINVAL, "tree geometry: more than " + std::to_string(kMaxVol - 1) +
                                     " touchables with a material (one tally cell each)");
    cell[t] = k;
    // a material that is both absorber and gap takes the gap's slot
    g.slot[k] = m == cfg.gap ? 1 : 0;
    g.visible[k] = m == cfg.gap ? 1 : 0;
    ++k;
  }
  for (int t = 0; t < nt; ++t)
    if (c->tree_vol[node[t].volume].material < 0) cell[t] = k;
  g.slot[k] = -1;
  g.visible[k] = 0;
  g.nvol = k + 1;
  g.tree = 1;
  HIP_TRY(dalloc(&c->cell, (size_t)nt));
  HIP_TRY(hipMemcpy(c->cell, cell.data(), (size_t)nt * sizeof(int32_t), hipMemcpyHostToDevice));
  g.cell = c->cell;
  g.nav = nav_device_view(c->nav);
  return GXT_OK;
}

int check_config(const gxt_config *c) {
  if (!c) return set_err(GXT_EINVAL, "null config");
  if (c->geometry != GXT_GEOM_BOX && c->geometry != GXT_GEOM_CALO && c->geometry != GXT_GEOM_TREE)
    return set_err(GXT_EINVAL, "bad geometry");
  if (c->geometry == GXT_GEOM_TREE) {
    if (!c->tree_volumes || c->n_tree_volumes <= 0 || c->n_tree_placements < 0 ||
        (c->n_tree_placements > 0 && !c->tree_placements) || c->tree_world < 0 ||
        c->tree_world >= c->n_tree_volumes)
      return set_err(GXT_EINVAL, "tree geometry: bad volume or placement arrays");
    for (int i = 0; i < c->n_tree_volumes; ++i) {
      const int m = c->tree_volumes[i].material;
      if (m != -1 && m != c->absorber && m != c->gap)
        return set_err(GXT_EINVAL, "tree geometry: a volume's material is neither the absorber, the gap nor vacuum");
    }
    if (c->field_map || c->field_stepper != GXT_STEPPER_HELIX)
      return set_err(GXT_EINVAL, "tree geometry: uniform field only (no field map, HELIX stepper)");
  }
  if (c->geometry == GXT_GEOM_CALO && (c->n_layers <= 0 || 2 * c->n_layers > kMaxVol))
    return set_err(GXT_EINVAL, "n_layers out of range");
  if (c->n_zbins <= 0 || c->n_zbins > kMaxZBins || c->n_rbins <= 0 || c->n_rbins > kMaxRBins)
    return set_err(GXT_EINVAL, "tally bins out of range");
  if (c->zbin_mm <= 0 || c->rbin_mm <= 0) return set_err(GXT_EINVAL, "bad tally bin width");
  if (c->pool_capacity < 1024 || c->pool_capacity > ((int64_t)1 << 27)) return set_err(GXT_EINVAL, "pool_capacity out of range");
  if (c->max_events <= 0 || c->max_jobs <= 0 || c->max_steps <= 0) return set_err(GXT_EINVAL, "bad limits");
  if (c->precision != GXT_FP64 && c->precision != GXT_FP32) return set_err(GXT_EINVAL, "bad precision");
  if (!(c->tracking_cut > 0)) return set_err(GXT_EINVAL, "tracking_cut must be > 0");
  if (!c->data_dir) return set_err(GXT_EINVAL, "data_dir not set");
  if (c->field_stepper != GXT_STEPPER_HELIX && c->field_stepper != GXT_STEPPER_RK45)
    return set_err(GXT_EINVAL, "bad field_stepper");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(c->field[k]) || !std::isfinite(c->field_origin[k])) return set_err(GXT_EINVAL, "bad field");
  return GXT_OK;
}

// Library-wide parameters for the RNG component entry points.
DeviceData g_rng_dd;
std::mutex g_rng_mu;

int rng_params(const DevParams **out) {
  std::lock_guard<std::mutex> lk(g_rng_mu);
  if (!g_rng_dd.P) {
    DevParams p;
    std::memset(&p, 0, sizeof(p));
    uint32_t jump[2][64][9];
    build_jump_matrices(jump);
    std::vector<uint32_t> j8;
    build_jump_tables(jump, j8);
    HIP_TRY(dalloc(&g_rng_dd.jump8, j8.size()));
    HIP_TRY(hipMemcpy(g_rng_dd.jump8, j8.data(), j8.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    p.jump8 = g_rng_dd.jump8;
    HIP_TRY(dalloc(&g_rng_dd.P, 1));
    HIP_TRY(hipMemcpy(g_rng_dd.P, &p, sizeof(p), hipMemcpyHostToDevice));
  }
  *out = g_rng_dd.P;
  return GXT_OK;
}

}  // namespace

namespace {

// Queue job j whose entries (primaries or tracks, src[0..n)) span n_events
// events: sort them by event, index the events, copy the entries to the
// device on the context's stream (ahead of any admission of the job, which
// follows it in stream order; the host copy stays alive until the job ends).
template <class In>
int queue_job(gxt_ctx *c, std::vector<In> &host, In *&dev, std::unique_ptr<Job> j, const In *src, int n,
              int n_events) {
  host.assign(src, src + n);
  j->n_events = n_events;
  std::stable_sort(host.begin(), host.end(), [](const In &a, const In &b) { return a.event < b.event; });
  j->ev_first.assign(j->n_events + 1, 0);
  for (const In &p : host) j->ev_first[p.event + 1]++;
  for (int e = 0; e < j->n_events; ++e) j->ev_first[e + 1] += j->ev_first[e];
  j->ev_gev.assign(j->n_events, 0.0);
  j->ev_max.assign(j->n_events, 0);
  for (const In &p : host) j->ev_gev[p.event] += p.ekin * 1.0e-3;
  j->ev_state.assign(j->n_events, 0);
  j->submit_ms = now_ms(c->t0);
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->finished.count(j->id)) return c->fail(GXT_EINVAL, "job id already used");
  for (const auto &q : c->jobs)
    if (q->id == j->id) return c->fail(GXT_EINVAL, "job id already used");
  const size_t bytes = host.size() * sizeof(In);
  HIP_TRY(hipMallocAsync((void **)&dev, bytes, c->stream));
  const hipError_t ce = hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c->stream);
  if (ce != hipSuccess) {
    (void)hipFreeAsync(dev, c->stream);
    dev = nullptr;
    return set_err(GXT_EHIP, std::string("submit copy: ") + hipGetErrorString(ce));
  }
  c->jobs.push_back(std::move(j));
  return GXT_OK;
}

// A kinematic state a job may carry: finite energy > 0, finite position,
// unit direction.
template <class In>
std::string check_kinematics(const In &p) {
  if (p.species < 0 || p.species > 2) return "bad species";
  if (!(p.ekin > 0) || !std::isfinite(p.ekin)) return "ekin must be finite and > 0";
  if (p.event < 0) return "negative event index";
  bool finite = true;
  for (int k = 0; k < 3; ++k) finite = finite && std::isfinite(p.pos[k]) && std::isfinite(p.dir[k]);
  if (!finite) return "non-finite position or direction";
  const double nn = p.dir[0] * p.dir[0] + p.dir[1] * p.dir[1] + p.dir[2] * p.dir[2];
  if (!(std::fabs(nn - 1.0) <= 2e-6)) return "direction is not a unit vector";
  return "";
}

// ------------------------------------------------------------ samplers
// Inputs of gxt_sample checked before any device work: finite kinetic energy
// > 0, unit direction, a valid MRG32k3a state for every track.
int check_sample(gxt_ctx *c, int model, int slot, const gxt_sample_io *io) {
  if (!c || !io || io->n <= 0 || io->n > kMaxSoaPoints || model < 0 || model > GXT_MODEL_ANNIHIL || slot < 0 ||
      slot >= (c->cfg.geometry == GXT_GEOM_BOX ? 1 : 2))
    return set_err(GXT_EINVAL, "bad sample arguments");
  if (!io->ekin || !io->dir || !io->rng || !io->out_ekin || !io->out_dir || !io->out_edep || !io->nsec ||
      !io->sec_species || !io->sec_ekin || !io->sec_dir)
    return set_err(GXT_EINVAL, "sample: null array");
  const int n = io->n;
  for (int i = 0; i < n; ++i) {
    const double d0 = io->dir[i], d1 = io->dir[n + i], d2 = io->dir[2 * n + i];
    const double nn = d0 * d0 + d1 * d1 + d2 * d2;
    const uint32_t *r = io->rng + 6 * (size_t)i;
    bool ok = io->ekin[i] > 0 && std::isfinite(io->ekin[i]) && std::fabs(nn - 1.0) <= 2e-6;
    ok = ok && !(r[0] == 0 && r[1] == 0 && r[2] == 0) && !(r[3] == 0 && r[4] == 0 && r[5] == 0);
    for (int k = 0; k < 3; ++k) ok = ok && r[k] < 4294967087u && r[3 + k] < 4294944443u;
    if (!ok) return set_err(GXT_EINVAL, "sample: bad energy, direction or RNG state at " + std::to_string(i));
  }
  return GXT_OK;
}

// Copy the inputs to the device arrays of s, sample, copy the outputs back;
// everything on the context's stream.
int run_sample(gxt_ctx *c, int model, int slot, int precision, gxt_sample_io *io, const SampleArgs &s) {
  const size_t n = (size_t)io->n;
  hipStream_t st = c->stream;
  HIP_TRY(hipMemcpyAsync(const_cast<double *>(s.ekin), io->ekin, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(const_cast<double *>(s.dir), io->dir, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s.rng, io->rng, 6 * n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(launch_sample(c->dd.P, precision, slot, model, s, st));
  HIP_TRY(hipMemcpyAsync(io->rng, s.rng, 6 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->out_ekin, s.out_ekin, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->out_dir, s.out_dir, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->out_edep, s.out_edep, n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->nsec, s.nsec, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->sec_species, s.sec_species, 2 * n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->sec_ekin, s.sec_ekin, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(io->sec_dir, s.sec_dir, 6 * n * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return GXT_OK;
}

}  // namespace

extern "C" {

const char *gxt_last_error(void) { return g_last_error.c_str(); }
// A copy per calling thread: the serving thread may replace the context's
// message while the caller reads it.
const char *gxt_ctx_error(gxt_ctx *ctx) {
  static thread_local std::string copy;
  copy = ctx ? ctx->error() : std::string("null context");
  return copy.c_str();
}

int gxt_default_config(gxt_config *c) {
  if (!c) return set_err(GXT_EINVAL, "null config");
  std::memset(c, 0, sizeof(*c));
  c->geometry = GXT_GEOM_BOX;
  c->material = GXT_MAT_PB;
  c->box_xy = 200.0;
  c->box_depth = 200.0;
  c->n_layers = 50;
  c->absorber = GXT_MAT_PB;
  c->gap = GXT_MAT_LAR;
  c->absorber_mm = 2.3;
  c->gap_mm = 5.7;
  c->lar_density = 1.396;
  c->tracking_cut = 1.0e-3;
  c->seed = 12345;
  c->precision = GXT_FP64;
  c->zbin_mm = 10.0;
  c->rbin_mm = 10.0;
  c->n_zbins = 20;
  c->n_rbins = 20;
  c->pool_capacity = 2000000;
  c->max_steps = 100000;
  c->max_events = 1000;
  c->max_jobs = 64;
  return GXT_OK;
}

int gxt_create(const gxt_config *cfg, gxt_ctx **out) {
  if (!out) return set_err(GXT_EINVAL, "null output");
  int rc = check_config(cfg);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(GXT_EHIP, "no HIP device");
  std::unique_ptr<gxt_ctx, void (*)(gxt_ctx *)> c(new gxt_ctx(), destroy_ctx);
  c->cfg = *cfg;
  c->data_dir = cfg->data_dir;
  c->cfg.data_dir = c->data_dir.c_str();
  c->precision = cfg->precision;
  // whole stripes
  c->cap = (cfg->pool_capacity + kStripes - 1) / kStripes * kStripes;
  c->trace = cfg->trace_tracks > 0;
  c->prof = cfg->profile_kernels > 0;
  c->prof_level = cfg->profile_kernels > 1 ? 2 : 1;
  c->t0 = std::chrono::steady_clock::now();
  HIP_TRY(hipGetDevice(&c->device));
  rc = scratch_setup(c->device);
  if (rc) return rc;
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&c->ss.side, hipStreamNonBlocking));
  // the chains' events order kernels of this device only: the end of a
  // kernel already makes its writes visible to the device's later kernels,
  // so they skip the system-scope fence (a cache write-back each)
  const unsigned sync_flags = hipEventDisableTiming | hipEventDisableSystemFence;
  HIP_TRY(hipEventCreateWithFlags(&c->ss.fork, sync_flags));
  HIP_TRY(hipEventCreateWithFlags(&c->ss.join, sync_flags));
  HIP_TRY(hipEventCreateWithFlags(&c->ss.entry, sync_flags));
  HIP_TRY(hipEventCreateWithFlags(&c->snap, hipEventDisableTiming));
  c->ss.main = c->stream;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
  c->grid = std::max(cus, 1) * 4;
  // physics data
  DevParams &P = c->hd.params;
  std::memset(&P, 0, sizeof(P));
  std::vector<int> mats = cfg->geometry == GXT_GEOM_BOX ? std::vector<int>{cfg->material}
                                                        : std::vector<int>{cfg->absorber, cfg->gap};
  for (size_t s = 0; s < mats.size(); ++s) {
    std::string err;
    rc = load_material(c->data_dir, mats[s], cfg->lar_density, c->hd, (int)s, err);
    if (rc) return set_err(rc, err);
  }
  build_geometry(*cfg, P.geo);
  if (cfg->geometry == GXT_GEOM_TREE) {
    rc = build_tree_geometry(c.get(), *cfg, P.geo);
    if (rc) return rc;
  }
  P.B[0] = cfg->field[0];
  P.B[1] = cfg->field[1];
  P.B[2] = cfg->field[2];
  P.bmag = std::sqrt(P.B[0] * P.B[0] + P.B[1] * P.B[1] + P.B[2] * P.B[2]);
  if (cfg->field_map) {
    c->fmap_path = cfg->field_map;
    c->cfg.field_map = c->fmap_path.c_str();
    rc = read_field_map(cfg->field_map, c->fmap);
    if (rc == GXT_OK) rc = upload_field_map(c->fmap);
    if (rc) return rc;
  }
  P.field = dev_field(cfg->field_map ? &c->fmap : nullptr, cfg->field, cfg->field_origin, cfg->field_stepper);
  P.msc = cfg->msc ? 1 : 0;
  if (P.msc)
    for (size_t s = 0; s < mats.size(); ++s)
      if (P.mat[s].tab[GXT_TAB_MSC_EM].n == 0 || P.mat[s].tab[GXT_TAB_MSC_EP].n == 0)
        return set_err(GXT_EINVAL, "multiple scattering needs the LambdaMod1.msc tables (none for PbWO4)");
  c->transport_mode = cfg->geometry == GXT_GEOM_TREE
                           ? 4
                           : (P.field.general ? (P.field.stepper == GXT_STEPPER_RK45 ? 3 : 2) : (P.msc ? 1 : 0));
  // Steps per transport launch for tracks that neither interact nor end (a
  // track's results do not depend on it).  Measured: in the block 2 is
  // fastest (10 GeV e- in PbWO4; 1: -2 %, 3: -1 %, 4: -2 %, 6: -4 %); in the
  // layered calorimeter, where most steps end on a thin layer's face, 4 is
  // 12 % faster than 2.
  c->inner_steps = cfg->geometry == GXT_GEOM_BOX ? 2 : 4;
  P.tracking_cut = cfg->tracking_cut;
  P.zbin = cfg->zbin_mm;
  P.rbin = cfg->rbin_mm;
  P.nz = cfg->n_zbins;
  P.nr = cfg->n_rbins;
  P.max_steps = cfg->max_steps;
  P.trace_events = cfg->trace_tracks > 0 ? cfg->trace_tracks : 0;
  P.trace_steps = cfg->trace_steps;
  for (int i = 0; i < 6; ++i) P.seed[i] = cfg->seed;
  // LDS layouts of the transports' tables: each slot's photon (e-/e+) tables
  // are one contiguous node span (tables are stored in GXT_TAB order)
  {
    const int first[2] = {GXT_TAB_LAMBDA_COMPT, GXT_TAB_LAMBDA_BREM_EM};
    const int last[2] = {GXT_TAB_LAMBDAPRIM_PHOT, GXT_TAB_INVRANGE_EP};
    for (size_t s = 0; s < mats.size(); ++s)
      for (int t = 0; t < GXT_NTAB; ++t) P.mat[s].tab[t].loff = -1;
    for (int l = 0; l < 2; ++l) {
      int base = 0;
      for (size_t s = 0; s < mats.size(); ++s) {
        DevMaterial &m = P.mat[s];
        for (int t = first[l]; t < last[l]; ++t)
          if (m.tab[t + 1].off != m.tab[t].off + m.tab[t].n) return set_err(GXT_EDATA, "tables not contiguous");
        const int src = m.tab[first[l]].off;
        P.lds_src[l][s] = src;
        P.lds_cnt[l][s] = m.tab[last[l]].off + m.tab[last[l]].n - src;
        for (int t = first[l]; t <= last[l]; ++t) m.tab[t].loff = base + (m.tab[t].off - src);
        base += P.lds_cnt[l][s];
      }
      P.lds_nodes[l] = base;
    }
  }
  rc = upload(c->hd, c->dd);
  if (rc) return rc;
  // pools, queues, counters, tallies
  for (int i = 0; i < 2; ++i) {
    rc = c->precision == GXT_FP32 ? c->pf[i].alloc(c->cap, c->trace) : c->pd[i].alloc(c->cap, c->trace);
    if (rc) return rc;
  }
  HIP_TRY(dalloc(&c->queues, (size_t)kNumQueues * c->cap));
  HIP_TRY(dalloc(&c->newlist, (size_t)2 * c->cap));
  HIP_TRY(dalloc(&c->cnt, 1));
  HIP_TRY(hipMemset(c->cnt, 0, sizeof(Counters)));
  HIP_TRY(hipHostMalloc((void **)&c->hcnt, sizeof(Counters), hipHostMallocDefault));
  std::memset(c->hcnt, 0, sizeof(Counters));
  HIP_TRY(dalloc(&c->ev, cfg->max_events));
  HIP_TRY(hipMemset(c->ev, 0, sizeof(DevEvent) * cfg->max_events));
  HIP_TRY(dalloc(&c->ev_live, (size_t)cfg->max_events * kLiveStride));
  HIP_TRY(hipMemset(c->ev_live, 0, sizeof(int) * (size_t)cfg->max_events * kLiveStride));
  HIP_TRY(hipHostMalloc((void **)&c->hev_live, sizeof(int) * cfg->max_events, hipHostMallocDefault));
  std::memset(c->hev_live, 0, sizeof(int) * cfg->max_events);
  HIP_TRY(hipHostGetDevicePointer((void **)&c->dhcnt, c->hcnt, 0));
  HIP_TRY(hipHostGetDevicePointer((void **)&c->dhev_live, c->hev_live, 0));
  c->row_stride = cfg->n_zbins + cfg->n_rbins + P.geo.nvol;
  HIP_TRY(dalloc(&c->rows, (size_t)c->grid * c->row_stride));
  HIP_TRY(hipMemset(c->rows, 0, sizeof(double) * c->grid * c->row_stride));
  HIP_TRY(dalloc(&c->spec, (size_t)c->grid * 3 * kSpecBins));
  HIP_TRY(hipMemset(c->spec, 0, sizeof(unsigned long long) * c->grid * 3 * kSpecBins));
  if (c->trace) {
    c->trace_cap = 4000000;
    HIP_TRY(dalloc(&c->trace_buf, (size_t)12 * c->trace_cap));
  }
  c->ev_used.assign(cfg->max_events, 0);
  c->ev_counted.assign(cfg->max_events, 0);
  HIP_TRY(hipDeviceSynchronize());
  *out = c.release();
  return GXT_OK;
}

int gxt_destroy(gxt_ctx *ctx) {
  if (!ctx) return set_err(GXT_EINVAL, "null context");
  destroy_ctx(ctx);
  return GXT_OK;
}

double gxt_now_ms(gxt_ctx *ctx) { return ctx ? now_ms(ctx->t0) : 0.0; }

int gxt_submit(gxt_ctx *c, const gxt_primary *prim, int n, uint64_t job_id) {
  if (!prim || n <= 0) return set_err(GXT_EINVAL, "bad submit arguments");
  // every primary is checked before anything is queued: a NaN position would
  // be located in volume 0, a non-unit direction would scale every distance,
  // an infinite energy would hold back admission for good
  int maxev = -1;
  for (int i = 0; i < n; ++i) {
    const std::string bad = check_kinematics(prim[i]);
    if (!bad.empty()) return set_err(GXT_EINVAL, bad + " (primary " + std::to_string(i) + ")");
    maxev = std::max(maxev, prim[i].event);
  }
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (maxev >= c->cfg.max_events) return c->fail(GXT_ECAPACITY, "job has more events than max_events");
  auto j = std::make_unique<Job>();
  j->id = job_id;
  Job *jr = j.get();
  return queue_job(c, jr->prims, jr->dprims, std::move(j), prim, n, maxev + 1);
}

int gxt_submit_tracks(gxt_ctx *c, const gxt_track *trk, int n, int n_events, uint64_t job_id) {
  if (!trk || n <= 0 || n_events <= 0) return set_err(GXT_EINVAL, "bad submit_tracks arguments");
  // what a track carries is checked first, then what depends on the context
  // (the step limit, the geometry)
  for (int i = 0; i < n; ++i) {
    const gxt_track &t = trk[i];
    std::string bad = check_kinematics(t);
    if (bad.empty() && t.event >= n_events) bad = "event index >= n_events";
    if (bad.empty() && !(std::isfinite(t.nil[0]) && std::isfinite(t.nil[1]) && std::isfinite(t.nil[2])))
      bad = "non-finite number of interaction lengths";
    if (bad.empty() && t.nsteps < 0) bad = "negative nsteps";
    if (bad.empty()) {
      const bool z1 = t.rng[0] == 0 && t.rng[1] == 0 && t.rng[2] == 0;
      const bool z2 = t.rng[3] == 0 && t.rng[4] == 0 && t.rng[5] == 0;
      bool range = true;
      for (int k = 0; k < 3; ++k) range = range && t.rng[k] < 4294967087u && t.rng[3 + k] < 4294944443u;
      if (z1 || z2 || !range) bad = "invalid MRG32k3a state";
    }
    if (!bad.empty()) return set_err(GXT_EINVAL, bad + " (track " + std::to_string(i) + ")");
  }
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (n_events > c->cfg.max_events) return c->fail(GXT_ECAPACITY, "job has more events than max_events");
  const DevGeometry &g = c->hd.params.geo;
  for (int i = 0; i < n; ++i) {
    const gxt_track &t = trk[i];
    std::string bad;
    if (t.nsteps >= c->cfg.max_steps) bad = "nsteps >= max_steps";
    if (bad.empty() && t.volume != -1 && g.tree) {
      int nt = 0;
      const NavNode *node = nav_host_nodes(c->nav, &nt);
      if (t.volume < 0 || t.volume >= nt) {
        bad = "volume out of range";
      } else {
        // inside the touchable's box, within the navigator's surface tolerance
        for (int k = 0; k < 3; ++k)
          if (!(std::fabs(t.pos[k] - node[t.volume].c[k]) <= node[t.volume].h[k] + 1e-9))
            bad = "position outside its volume";
      }
    } else if (bad.empty() && t.volume != -1) {
      if (t.volume < 0 || t.volume >= g.nvol)
        bad = "volume out of range";
      else if (!(std::fabs(t.pos[0]) <= g.hx && std::fabs(t.pos[1]) <= g.hy && t.pos[2] >= g.z[t.volume] &&
                 t.pos[2] <= g.z[t.volume + 1]))
        bad = "position outside its volume";
    }
    if (!bad.empty()) return c->fail(GXT_EINVAL, bad + " (track " + std::to_string(i) + ")");
  }
  auto j = std::make_unique<Job>();
  j->id = job_id;
  Job *jr = j.get();
  return queue_job(c, jr->tracks, jr->dtracks, std::move(j), trk, n, n_events);
}

int gxt_run(gxt_ctx *c, int64_t max_iterations) {
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (c->serving.load()) return c->fail(GXT_ESTATE, "gxt_run while serving");
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipEventRecord(e0, c->stream));
  const int64_t start = c->iterations;
  int rc = GXT_OK;
  for (;;) {
    bool idle = false;
    rc = round_trip(c, iters_per_sync(c), idle);
    if (rc || idle) break;
    if (max_iterations > 0 && c->iterations - start >= max_iterations) break;
  }
  (void)hipEventRecord(e1, c->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  c->wall_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int gxt_poll(gxt_ctx *c, gxt_job_result *out, int max, int *n_out) {
  if (!c || !n_out || (max > 0 && !out)) return set_err(GXT_EINVAL, "bad poll arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  int n = 0;
  while (n < max && !c->done_unpolled.empty()) {
    std::unique_ptr<Job> j = std::move(c->done_unpolled.front());
    c->done_unpolled.erase(c->done_unpolled.begin());
    gxt_job_result &r = out[n++];
    std::memset(&r, 0, sizeof(r));
    r.job_id = j->id;
    r.n_events = j->n_events;
    for (const gxt_event_tally &t : j->tallies) {
      r.edep += t.edep;
      r.visible += t.visible;
    }
    r.submit_ms = j->submit_ms;
    r.admit_ms = j->admit_ms;
    r.done_ms = j->done_ms;
    r.iterations = j->done_iter - j->admit_iter;
    r.track_steps = j->steps;
  }
  *n_out = n;
  // an error of the serving thread is reported here, to the polling thread
  const int rc = c->serve_rc.load();
  if (rc != GXT_OK) return set_err(rc, c->error());
  return GXT_OK;
}

int gxt_get_event_tallies(gxt_ctx *c, uint64_t job_id, gxt_event_tally *out, int n) {
  if (!c || !out) return set_err(GXT_EINVAL, "bad arguments");
  std::lock_guard<std::mutex> lk(c->mu);
  auto it = c->finished.find(job_id);
  if (it == c->finished.end()) return c->fail(GXT_ENOTFOUND, "job not finished or unknown");
  const Job &j = *it->second;
  for (int i = 0; i < n && i < j.n_events; ++i) out[i] = j.tallies[i];
  return GXT_OK;
}

int gxt_get_profiles(gxt_ctx *c, double *ez, double *er, double *ev) {
  if (!c) return set_err(GXT_EINVAL, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<double> rows((size_t)c->grid * c->row_stride);
  HIP_TRY(hipMemcpy(rows.data(), c->rows, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
  const int nz = c->cfg.n_zbins, nr = c->cfg.n_rbins, nv = c->hd.params.geo.nvol;
  if (ez) std::fill(ez, ez + nz, 0.0);
  if (er) std::fill(er, er + nr, 0.0);
  if (ev) std::fill(ev, ev + nv, 0.0);
  for (int g = 0; g < c->grid; ++g) {
    const double *r = rows.data() + (size_t)g * c->row_stride;
    for (int i = 0; i < nz; ++i)
      if (ez) ez[i] += r[i];
    for (int i = 0; i < nr; ++i)
      if (er) er[i] += r[nz + i];
    for (int i = 0; i < nv; ++i)
      if (ev) ev[i] += r[nz + nr + i];
  }
  return GXT_OK;
}

int gxt_volume_bins(gxt_ctx *c, int *n_bins, int32_t *cell) {
  if (!c || !n_bins) return set_err(GXT_EINVAL, "bad arguments");
  *n_bins = c->hd.params.geo.nvol;
  if (cell && c->cell) {
    int nt = 0;
    (void)nav_host_nodes(c->nav, &nt);
    HIP_TRY(hipMemcpy(cell, c->cell, (size_t)nt * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  return GXT_OK;
}

int gxt_get_spectra(gxt_ctx *c, int64_t *counts) {
  if (!c || !counts) return set_err(GXT_EINVAL, "bad arguments");
  HIP_TRY(hipStreamSynchronize(c->stream));
  std::vector<unsigned long long> s((size_t)c->grid * 3 * kSpecBins);
  HIP_TRY(hipMemcpy(s.data(), c->spec, s.size() * sizeof(s[0]), hipMemcpyDeviceToHost));
  std::fill(counts, counts + 3 * kSpecBins, 0);
  for (int g = 0; g < c->grid; ++g)
    for (int i = 0; i < 3 * kSpecBins; ++i) counts[i] += (int64_t)s[(size_t)g * 3 * kSpecBins + i];
  return GXT_OK;
}

int gxt_get_stats(gxt_ctx *c, gxt_stats *o) {
  if (!c || !o) return set_err(GXT_EINVAL, "bad arguments");
  HIP_TRY(hipStreamSynchronize(c->stream));
  // ~100 KB (three generations of striped counters): on the heap, not the caller's stack
  auto hp = std::make_unique<Counters>();
  Counters &h = *hp;
  HIP_TRY(hipMemcpy(&h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  o->track_steps = (int64_t)h.total(kStatSteps);
  o->iterations = c->iterations;
  o->max_live = c->max_live;
  o->killed_by_limit = (int64_t)h.total(kStatKilled);
  o->wall_ms = c->wall_ms;
  return GXT_OK;
}

int gxt_get_kernel_stats(gxt_ctx *c, gxt_kernel_stats *o) {
  if (!c || !o) return set_err(GXT_EINVAL, "bad arguments");
  HIP_TRY(hipStreamSynchronize(c->stream));
  // ~100 KB (three generations of striped counters): on the heap, not the caller's stack
  auto hp = std::make_unique<Counters>();
  Counters &h = *hp;
  HIP_TRY(hipMemcpy(&h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  *o = c->kst;
  o->track_steps = (int64_t)h.total(kStatSteps);
  o->charged_steps = (int64_t)h.total(kStatStepsCharged);
  o->neutral_steps = (int64_t)(h.total(kStatSteps) - h.total(kStatStepsCharged));
  o->interactions = (int64_t)h.total(kStatInteractions);
  o->created = (int64_t)h.total(kStatCreated);
  o->neutral_loaded = (int64_t)h.total(kStatLoaded0);
  o->charged_loaded = (int64_t)h.total(kStatLoaded1);
  o->neutral_interactions = (int64_t)h.total(kStatInteractions0);
  o->neutral_created = (int64_t)h.total(kStatCreated0);
  // x, y, z (FP64) + dir, E, 3 interaction lengths (R) + 6 RNG words + event, meta, nsteps
  // + 3 FP64 per-track tallies
  const int r = c->precision == GXT_FP32 ? 4 : 8;
  o->record_bytes = 3 * 8 + 7 * r + 6 * 4 + 3 * 4 + 3 * 8;
  return GXT_OK;
}

int gxt_reset_tallies(gxt_ctx *c) {
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (c->serving.load()) return c->fail(GXT_ESTATE, "reset while serving");
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemset(c->rows, 0, sizeof(double) * c->grid * c->row_stride));
  HIP_TRY(hipMemset(c->spec, 0, sizeof(unsigned long long) * c->grid * 3 * kSpecBins));
  // ~100 KB (three generations of striped counters): on the heap, not the caller's stack
  auto hp = std::make_unique<Counters>();
  Counters &h = *hp;
  HIP_TRY(hipMemcpy(&h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  std::memset(h.stat, 0, sizeof(h.stat));
  h.trace_n = 0;
  c->kst = gxt_kernel_stats{};
  HIP_TRY(hipMemcpy(c->cnt, &h, sizeof(h), hipMemcpyHostToDevice));
  c->iterations = 0;
  c->max_live = 0;
  std::lock_guard<std::mutex> lk(c->mu);
  c->finished.clear();
  c->done_unpolled.clear();
  return GXT_OK;
}

int gxt_get_trace(gxt_ctx *c, double *out, int64_t max_records, int64_t *n_out) {
  if (!c || !n_out) return set_err(GXT_EINVAL, "bad arguments");
  HIP_TRY(hipStreamSynchronize(c->stream));
  // ~100 KB (three generations of striped counters): on the heap, not the caller's stack
  auto hp = std::make_unique<Counters>();
  Counters &h = *hp;
  HIP_TRY(hipMemcpy(&h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  const int64_t n = std::min<int64_t>((int64_t)h.trace_n, c->trace_cap);
  *n_out = n;
  if (out && n > 0 && c->trace_buf)
    HIP_TRY(hipMemcpy(out, c->trace_buf, sizeof(double) * 12 * std::min(n, max_records), hipMemcpyDeviceToHost));
  if ((int64_t)h.trace_n > c->trace_cap) return c->fail(GXT_ECAPACITY, "trace buffer overflowed");
  return GXT_OK;
}

int gxt_serve_start(gxt_ctx *c) {
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (c->serving.load()) return c->fail(GXT_ESTATE, "already serving");
  c->stop = false;
  c->serve_rc = GXT_OK;
  c->serving = true;
  c->server = std::thread(serve_loop, c);
  return GXT_OK;
}

int gxt_serve_stop(gxt_ctx *c) {
  if (!c) return set_err(GXT_EINVAL, "null context");
  if (!c->serving.load()) return c->fail(GXT_ESTATE, "not serving");
  c->stop = true;
  if (c->server.joinable()) c->server.join();
  c->serving = false;
  // the last round's look-ahead iterations
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int rc = c->serve_rc.load();
  if (rc != GXT_OK) return set_err(rc, c->error());
  return GXT_OK;
}

int gxt_make_primaries(uint32_t seed, int species, double emin, double emax, double spot_mm, int n, uint64_t first,
                       uint64_t n_total, gxt_primary *out) {
  if (n < 0 || !out || species < 0 || species > 2) return set_err(GXT_EINVAL, "bad primaries request");
  for (int i = 0; i < n; ++i) {
    gxt_primary &p = out[i];
    std::memset(&p, 0, sizeof(p));
    p.species = species;
    p.event = i;
    p.substream = first + (uint64_t)i;
    p.ekin = emin;
    p.dir[2] = 1.0;
    if (emax > emin || spot_mm > 0.0) {
      HostMrg r;
      r.init(seed, n_total + first + (uint64_t)i);
      const double u0 = r.uniform(), u1 = r.uniform(), u2 = r.uniform();
      p.ekin = emin + (emax - emin) * u0;
      p.pos[0] = (u1 - 0.5) * spot_mm;
      p.pos[1] = (u2 - 0.5) * spot_mm;
    }
  }
  return GXT_OK;
}

// ---- component entry points
int gxt_rng_init(const uint32_t seed[6], const uint64_t *idx, int n, uint32_t *states) {
  if (!seed || !idx || !states || n <= 0) return set_err(GXT_EINVAL, "bad rng_init arguments");
  const DevParams *P = nullptr;
  int rc = rng_params(&P);
  if (rc) return rc;
  uint32_t *dseed = nullptr, *dout = nullptr;
  uint64_t *didx = nullptr;
  HIP_TRY(dalloc(&dseed, 6));
  HIP_TRY(dalloc(&didx, n));
  HIP_TRY(dalloc(&dout, 6 * (size_t)n));
  HIP_TRY(hipMemcpy(dseed, seed, 6 * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(didx, idx, n * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIP_TRY(launch_rng_init(P, dseed, didx, n, dout, 0));
  HIP_TRY(hipMemcpy(states, dout, 6 * (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  (void)hipFree(dseed);
  (void)hipFree(didx);
  (void)hipFree(dout);
  return GXT_OK;
}

int gxt_rng_draw(uint32_t *states, int n, int k, double *out) {
  if (!states || !out || n <= 0 || k < 0) return set_err(GXT_EINVAL, "bad rng_draw arguments");
  uint32_t *ds = nullptr;
  double *dout = nullptr;
  HIP_TRY(dalloc(&ds, 6 * (size_t)n));
  HIP_TRY(dalloc(&dout, (size_t)n * k));
  HIP_TRY(hipMemcpy(ds, states, 6 * (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(launch_rng_draw(ds, n, k, dout, 0));
  HIP_TRY(hipMemcpy(out, dout, (size_t)n * k * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(states, ds, 6 * (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  (void)hipFree(ds);
  (void)hipFree(dout);
  return GXT_OK;
}

static int nslots(gxt_ctx *c) { return c->cfg.geometry == GXT_GEOM_BOX ? 1 : 2; }

int gxt_table_value(gxt_ctx *c, int slot, int table, int precision, const double *x, double *out, int n) {
  if (!c || !x || !out || n <= 0 || slot < 0 || slot >= nslots(c) || table < 0 || table >= GXT_NTAB)
    return set_err(GXT_EINVAL, "bad table_value arguments");
  double *dx = nullptr, *dy = nullptr;
  HIP_TRY(dalloc(&dx, n));
  HIP_TRY(dalloc(&dy, n));
  HIP_TRY(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_table_value(c->dd.P, precision, slot, table, dx, dy, n, c->stream));
  HIP_TRY(hipMemcpyAsync(out, dy, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(dx);
  (void)hipFree(dy);
  return GXT_OK;
}

int gxt_macro_xs(gxt_ctx *c, int slot, int species, int process, int precision, const double *e, double *out, int n) {
  if (!c || !e || !out || n <= 0 || slot < 0 || slot >= nslots(c) || species < 0 || species > 2 || process < 0 ||
      process >= (species == 1 ? 2 : 3))
    return set_err(GXT_EINVAL, "bad macro_xs arguments");
  double *de = nullptr, *dy = nullptr;
  HIP_TRY(dalloc(&de, n));
  HIP_TRY(dalloc(&dy, n));
  HIP_TRY(hipMemcpy(de, e, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_macro_xs(c->dd.P, precision, slot, species, process, de, dy, n, c->stream));
  HIP_TRY(hipMemcpyAsync(out, dy, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(de);
  (void)hipFree(dy);
  return GXT_OK;
}

int gxt_sample(gxt_ctx *c, int model, int slot, int precision, gxt_sample_io *io) {
  const int rc = check_sample(c, model, slot, io);
  if (rc) return rc;
  const size_t n = (size_t)io->n;
  DevBuf<double> ek(n), dir(3 * n), oek(n), odir(3 * n), oed(n), sek(2 * n), sdir(6 * n);
  DevBuf<uint32_t> rng(6 * n);
  DevBuf<int32_t> nsec(n), ssp(2 * n);
  for (hipError_t e : {ek.err, dir.err, oek.err, odir.err, oed.err, sek.err, sdir.err, rng.err, nsec.err, ssp.err})
    HIP_TRY(e);
  SampleArgs s{io->n, ek.p, dir.p, rng.p, oek.p, odir.p, oed.p, nsec.p, ssp.p, sek.p, sdir.p};
  return run_sample(c, model, slot, precision, io, s);
}

int gxt_sample_noalloc(gxt_ctx *c, int model, int slot, int precision, gxt_sample_io *io) {
  const int rc = check_sample(c, model, slot, io);
  if (rc) return rc;
  if (io->n > kNoAllocSampleTracks)
    return c->fail(GXT_EINVAL, "gxt_sample_noalloc: n = " + std::to_string(io->n) + " exceeds its staging limit of " +
                                   std::to_string(kNoAllocSampleTracks) + " tracks");
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = static_cast<unsigned char *>(g_scratch_ptr[c->device]);
  }
  if (!base) return c->fail(GXT_ESTATE, "gxt_sample_noalloc: no staging address for this device");
  const SampleStaging L(io->n);
  SampleArgs s{io->n,
               reinterpret_cast<double *>(base + L.ek),
               reinterpret_cast<double *>(base + L.dir),
               reinterpret_cast<uint32_t *>(base + L.rng),
               reinterpret_cast<double *>(base + L.oek),
               reinterpret_cast<double *>(base + L.odir),
               reinterpret_cast<double *>(base + L.oed),
               reinterpret_cast<int32_t *>(base + L.nsec),
               reinterpret_cast<int32_t *>(base + L.ssp),
               reinterpret_cast<double *>(base + L.sek),
               reinterpret_cast<double *>(base + L.sdir)};
  return run_sample(c, model, slot, precision, io, s);
}

int gxt_slab_distance(gxt_ctx *c, const double *pos, const double *dir, const double *pstep, int32_t *vol,
                      double *snext, int32_t *next, double *safety, int32_t *boundary, int n) {
  if (!c || !pos || !dir || !vol || !snext || !next || !safety || n <= 0 || n > kMaxSoaPoints)
    return set_err(GXT_EINVAL, "bad slab_distance arguments (1 <= n <= 2^29)");
  if (c->hd.params.geo.tree)
    return c->fail(GXT_EINVAL, "slab_distance: the context's geometry is a volume tree (use the gxt_nav_* calls)");
  const int nvol = c->hd.params.geo.nvol;
  for (int i = 0; i < n; ++i) {
    const double d[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    const double nn = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (!std::isfinite(pos[i]) || !std::isfinite(pos[n + i]) || !std::isfinite(pos[2 * n + i]) ||
        !(std::fabs(nn - 1.0) <= 2e-6))
      return c->fail(GXT_EINVAL, "slab_distance: non-finite position or non-unit direction at " + std::to_string(i));
    if (vol[i] >= nvol) return c->fail(GXT_EINVAL, "slab_distance: volume out of range at " + std::to_string(i));
    if (pstep && !(pstep[i] >= 0.0 && std::isfinite(pstep[i])))
      return c->fail(GXT_EINVAL, "slab_distance: bad pstep at " + std::to_string(i));
  }
  DevBuf<double> dp(3 * (size_t)n), dd(3 * (size_t)n), ds(n), dsaf(n), dps(pstep ? n : 0);
  DevBuf<int32_t> dv(n), dn(n), db(boundary ? n : 0);
  for (hipError_t e : {dp.err, dd.err, ds.err, dsaf.err, dps.err, dv.err, dn.err, db.err})
    HIP_TRY(e);
  HIP_TRY(hipMemcpyAsync(dp.p, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dd.p, dir, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(dv.p, vol, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  if (pstep) HIP_TRY(hipMemcpyAsync(dps.p, pstep, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_slab_distance(c->dd.P, dp.p, dd.p, pstep ? dps.p : nullptr, dv.p, ds.p, dn.p, dsaf.p,
                               boundary ? db.p : nullptr, n, c->stream));
  HIP_TRY(hipMemcpyAsync(vol, dv.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(snext, ds.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(next, dn.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(safety, dsaf.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (boundary)
    HIP_TRY(hipMemcpyAsync(boundary, db.p, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GXT_OK;
}

int gxt_helix_step(const double field[3], int precision, const double *charge, const double *mom, const double *step,
                   double *pos, double *dir, int n) {
  if (!field || !charge || !mom || !step || !pos || !dir || n <= 0 || n > kMaxSoaPoints)
    return set_err(GXT_EINVAL, "bad helix arguments (1 <= n <= 2^29)");
  const double bm = std::sqrt(field[0] * field[0] + field[1] * field[1] + field[2] * field[2]);
  if (!(bm > 0)) return set_err(GXT_EINVAL, "zero field");
  double *dq = nullptr, *dm = nullptr, *ds = nullptr, *dp = nullptr, *dd = nullptr;
  HIP_TRY(dalloc(&dq, n));
  HIP_TRY(dalloc(&dm, n));
  HIP_TRY(dalloc(&ds, n));
  HIP_TRY(dalloc(&dp, 3 * (size_t)n));
  HIP_TRY(dalloc(&dd, 3 * (size_t)n));
  HIP_TRY(hipMemcpy(dq, charge, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dm, mom, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(ds, step, n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dp, pos, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dd, dir, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(launch_helix(field, precision, dq, dm, ds, dp, dd, n, 0));
  HIP_TRY(hipMemcpy(pos, dp, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dir, dd, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  void *ptrs[] = {dq, dm, ds, dp, dd};
  for (void *p : ptrs) (void)hipFree(p);
  return GXT_OK;
}

}  // extern "C"
