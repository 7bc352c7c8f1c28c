// gxt_kernels.hip -- gfx950 kernels of the EM step loop and of the
// component entry points.
//
// The step loop is a wavefront design.  Each iteration runs
//   k_transport : every live track (neutral region [0, nN), charged region
//                 [cap-nC, cap) of the current pool) does its PostStep GPIL,
//                 propagation, along-step loss; tracks that reach a discrete
//                 interaction are written back in place and their slot is
//                 appended, by wave ballot, to the queue of their
//                 (species, process); the rest go straight to the next pool;
//   k_interact  : the concatenated queues are sampled, so a wave runs one
//                 sampler; survivors and secondaries are appended to the next
//                 pool by wave prefix sums (one atomic per wave and region);
//   k_derive    : the RNG states of the tracks created this iteration.
// The counters rotate over three generations (Counters): the next pool's
// counts are the following iteration's current ones as they stand, and the
// photon transport clears the generation after next.
// The pool, the queues and the new-track list are striped (kStripes stripes
// with their own counters; workgroup b appends to stripe b % kStripes) so no
// counter takes the atomics of the whole grid.  Histogram tallies are
// no-return FP64 atomics into per-workgroup rows; per-event sums are FP64
// atomics on the event's record.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "gxt_device.h"
#include "gxt_kernels.h"

namespace gxt {

constexpr int kBlock = 256;
// Minimum waves per SIMD the step kernels are compiled for (register budget).
#ifndef GXT_STEP_WAVES
#define GXT_STEP_WAVES 2
#endif



// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane() { return (int)__lane_id(); }

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int l = lane();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int t = __shfl_up(v, d, 64);
    if (l >= d) v += t;
  }
  return v;
}

// Reserve n consecutive slots per lane from *counter with one atomic per
// wave; returns the lane's first slot.  Must be reached by all 64 lanes.
__device__ __forceinline__ int wave_reserve(int n, int *counter) {
  const int incl = wave_incl_scan(n);
  const int total = __shfl(incl, 63, 64);
  int base = 0;
  if (total > 0) {
    if (lane() == 63) base = atomicAdd(counter, total);
    base = __shfl(base, 63, 64);
  }
  return base + incl - n;
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *ctr, int pred) {
  const unsigned long long m = __ballot(pred);
  if (m) {
    const int leader = __ffsll((long long)m) - 1;
    if (lane() == leader) atomicAdd(ctr, (unsigned long long)__popcll(m));
  }
}

// Stripe the calling workgroup appends to.
__device__ __forceinline__ int my_stripe() { return (int)(blockIdx.x % kStripes); }

// Item counts of one region over the stripes, as a prefix, with the mapping
// from a global item index to (stripe, offset in the stripe).
struct StripeSpan {
  int pre[kStripes + 1];
  __device__ __forceinline__ int total() const { return pre[kStripes]; }
  // (selects over the unrolled prefix: a lane-dependent index into pre[]
  // would move it out of registers)
  __device__ __forceinline__ void locate(int i, int &s, int &off) const {
    s = 0;
    int base = 0;
#pragma unroll
    for (int k = 1; k < kStripes; ++k) {
      const bool ge = i >= pre[k];
      s = ge ? k : s;
      base = ge ? pre[k] : base;
    }
    off = i - base;
  }
  // pre[] from each lane k < kStripes's count v (lanes >= kStripes pass 0):
  // a wave scan, read back into uniform registers.  The caller loads one
  // count per lane, so the wave has all its counter loads in flight at once
  // (a loop over the stripes' counters compiles to scalar loads that wait
  // for each other in groups of eight).  Must be reached by all 64 lanes.
  __device__ __forceinline__ void from_lanes(int v) {
    static_assert(kStripes <= 64, "one stripe per lane");
    const int incl = wave_incl_scan(v);
    pre[0] = 0;
#pragma unroll
    for (int k = 0; k < kStripes; ++k) pre[k + 1] = __builtin_amdgcn_readlane(incl, k);
  }
};

// Two-phase wave reservation.  issue() sends the wave's one atomic and get()
// uses its result; issuing several before the first get() keeps their
// atomics in flight together.  (vmcnt counts loads, stores and atomics in
// issue order, so a returning atomic issued after stores also waits for
// them: the kernels issue their reservations before their stores.)
struct WaveReserve {
  int n, incl, base;
  __device__ __forceinline__ void issue(int cnt, int *ctr) {
    n = cnt;
    incl = wave_incl_scan(cnt);
    const int total = __shfl(incl, 63, 64);
    base = 0;
    if (total > 0 && lane() == 63) base = atomicAdd(ctr, total);
  }
  __device__ __forceinline__ int get() const { return __shfl(base, 63, 64) + incl - n; }
};

// Reserve an entry of queue q (q < 0: none; q in [Q0, Q0 + NQ)) of the
// workgroup's stripe, one atomic per queue present in the wave, all issued
// before any result is used.  Returns the entry's index in queues[], -1 for
// none, -2 when the stripe of the queue is full.
template <int Q0, int NQ>
__device__ __forceinline__ int64_t queue_reserve(int q, Counters *cnt, int g, int64_t cap, int64_t sc) {
  const unsigned long long lt = (lane() == 0) ? 0ULL : ((~0ULL) >> (64 - lane()));
  const int st = my_stripe();
  unsigned long long m[NQ];
  int base[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    m[k] = __ballot(q == Q0 + k);
    base[k] = 0;
    if (m[k] && lane() == __ffsll((long long)m[k]) - 1) base[k] = atomicAdd(&cnt->ctr(g, kQ0 + Q0 + k, st), (int)__popcll(m[k]));
  }
  int64_t pos = -1;
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    if (m[k]) {
      const int b = __shfl(base[k], __ffsll((long long)m[k]) - 1, 64);
      if (q == Q0 + k) {
        const int64_t idx = (int64_t)b + __popcll(m[k] & lt);
        pos = idx < sc ? (int64_t)(Q0 + k) * cap + (int64_t)st * sc + idx : -2;
      }
    }
  }
  return pos;
}

// ------------------------------------------------------------------ tallies
// Row of the calling workgroup: z bins, r bins, volumes; then its spectrum row.
__device__ __forceinline__ double *tally_row(const TallyRows &r) {
  return r.bins + (int64_t)(blockIdx.x % r.nrows) * r.stride;
}

// Histograms of one wave in LDS.  Deposits of a wave's lanes often share a
// bin (a homogeneous block has one volume; the shower core fills the first
// radial bins), which serialises global atomics in L2; LDS atomics resolve
// those conflicts, and at the end of the kernel the wave adds its non-zero
// bins to the workgroup's row with no-return atomics.  Wave-private: no
// barrier.
struct WaveHist {
  double z[kMaxZBins];
  double r[kMaxRBins];
  double v[kMaxVol];
  unsigned int spec[3 * kSpecBins];
};

__device__ __forceinline__ void wh_clear(WaveHist &h) {
  const int l = lane();
  for (int i = l; i < kMaxZBins; i += 64) h.z[i] = 0.0;
  for (int i = l; i < kMaxRBins; i += 64) h.r[i] = 0.0;
  for (int i = l; i < kMaxVol; i += 64) h.v[i] = 0.0;
  for (int i = l; i < 3 * kSpecBins; i += 64) h.spec[i] = 0u;
}

__device__ __forceinline__ void wh_flush(const StepArgs &a, WaveHist &h) {
  // the wave's LDS atomics are complete before it reads the bins back:
  // s_waitcnt lgkmcnt(0) (vmcnt and expcnt left at their maxima)
  __builtin_amdgcn_s_waitcnt(0xC07F);
  const int l = lane();
  const int nz = a.P->nz, nr = a.P->nr, nv = a.P->geo.nvol;
  double *row = tally_row(a.rows);
  for (int i = l; i < nz; i += 64)
    if (h.z[i] != 0.0) unsafeAtomicAdd(&row[i], h.z[i]);
  for (int i = l; i < nr; i += 64)
    if (h.r[i] != 0.0) unsafeAtomicAdd(&row[nz + i], h.r[i]);
  for (int i = l; i < nv; i += 64)
    if (h.v[i] != 0.0) unsafeAtomicAdd(&row[nz + nr + i], h.v[i]);
  unsigned long long *srow = a.rows.spec + (int64_t)(blockIdx.x % a.rows.nrows) * 3 * kSpecBins;
  for (int i = l; i < 3 * kSpecBins; i += 64)
    if (h.spec[i]) atomicAdd(&srow[i], (unsigned long long)h.spec[i]);
}

__device__ __forceinline__ void spectrum(WaveHist &h, int species, double e) {
  const int b = (int)floor((m_log10(e) + 3.0) * 8.0);
  if (b >= 0 && b < kSpecBins) atomicAdd(&h.spec[species * kSpecBins + b], 1u);
}

// ------------------------------------------------------------------ pool I/O
// The three interaction lengths are named members so that, with the process
// loops unrolled, every access is static and stays in registers.
template <class R>
struct TrackRegs {
  double pos[3];
  R dir[3];
  R e;
  R nil0, nil1, nil2;
  Rng rng;
  int event, vol, species, nsteps;
  double aed, avis, alen;
  uint64_t uid;
  // value selects (a store through a selected member address would force the
  // whole struct out of registers)
  __device__ __forceinline__ void set_nil(int p, R v) {
    nil0 = p == 0 ? v : nil0;
    nil1 = p == 1 ? v : nil1;
    nil2 = p == 2 ? v : nil2;
  }
};

template <class R>
__device__ __forceinline__ void load_track(const Pool<R> &p, int64_t i, TrackRegs<R> &t, bool trace) {
  t.pos[0] = p.x[i];
  t.pos[1] = p.y[i];
  t.pos[2] = p.z[i];
  t.dir[0] = p.dx[i];
  t.dir[1] = p.dy[i];
  t.dir[2] = p.dz[i];
  t.e = p.e[i];
  t.nil0 = p.nil0[i];
  t.nil1 = p.nil1[i];
  t.nil2 = p.nil2[i];
  t.rng.load(p.rng, p.cap, i);
  t.event = p.event[i];
  const int meta = p.meta[i];
  t.vol = meta & 0xffff;
  t.species = meta >> 16;
  t.nsteps = p.nsteps[i];
  t.aed = p.aed[i];
  t.avis = p.avis[i];
  t.alen = p.alen[i];
  t.uid = trace ? p.uid[i] : 0;
}

// kInPlace: the track goes back to the slot it was loaded from, so its event
// is already there; kPhotonTransport additionally leaves out what a photon's
// transport step never changes (direction, energy, track tallies).
template <class R, bool kInPlace = false, bool kPhotonTransport = false>
__device__ __forceinline__ void store_track(const Pool<R> &p, int64_t i, const TrackRegs<R> &t, bool trace) {
  p.x[i] = t.pos[0];
  p.y[i] = t.pos[1];
  p.z[i] = t.pos[2];
  if (!kPhotonTransport) {
    p.dx[i] = t.dir[0];
    p.dy[i] = t.dir[1];
    p.dz[i] = t.dir[2];
    p.e[i] = t.e;
  }
  p.nil0[i] = t.nil0;
  p.nil1[i] = t.nil1;
  p.nil2[i] = t.nil2;
  t.rng.store(p.rng, p.cap, i);
  if (!kInPlace) p.event[i] = t.event;
  p.meta[i] = t.vol | (t.species << 16);
  p.nsteps[i] = t.nsteps;
  if (!kPhotonTransport) {
    p.aed[i] = t.aed;
    p.avis[i] = t.avis;
    p.alen[i] = t.alen;
  }
  if (trace) p.uid[i] = t.uid;
}

// A track created by an interaction or an annihilation at rest, at the
// position, volume and event of its parent.  Its RNG state is substream uid;
// RNG words 1 and 2 hold uid until k_derive replaces them with that state.
template <class R>
__device__ inline void store_new(const DevParams *P, const Pool<R> &p, int64_t i, const TrackRegs<R> &par,
                                 int species, R e, R d0, R d1, R d2, uint64_t uid, bool trace) {
  p.rng[p.cap + i] = (uint32_t)uid;
  p.rng[2 * p.cap + i] = (uint32_t)(uid >> 32);
  p.x[i] = par.pos[0];
  p.y[i] = par.pos[1];
  p.z[i] = par.pos[2];
  p.dx[i] = d0;
  p.dy[i] = d1;
  p.dz[i] = d2;
  p.e[i] = e;
  p.nil0[i] = R(-1);
  p.nil1[i] = R(-1);
  p.nil2[i] = R(-1);
  p.event[i] = par.event;
  p.meta[i] = par.vol | (species << 16);
  p.nsteps[i] = 0;
  p.aed[i] = 0.0;
  p.avis[i] = 0.0;
  p.alen[i] = 0.0;
  if (trace) p.uid[i] = uid;
}

// Slot of item idx of a region in stripe st of a pool (neutral from the
// bottom of the stripe, charged from its top); -1 when it would leave the
// stripe.
__device__ __forceinline__ int64_t out_slot(int region, int st, int idx, int64_t sc) {
  if (idx < 0 || (int64_t)idx >= sc) return -1;
  return region == 0 ? (int64_t)st * sc + idx : (int64_t)(st + 1) * sc - 1 - idx;
}

// Pool slots per stripe.
template <class R>
__device__ __forceinline__ int64_t stripe_cap(const Pool<R> &p) {
  return p.cap / kStripes;
}

// A deposit held back until the kernel's stores: the track's accumulators
// are updated at once, the workgroup's histogram row by flush().  Every add()
// of one Dep is at the same point.
struct Dep {
  double e = 0.0;
  int iz = 0, ir = 0, vol = 0;
  // v: the track's volume (kGeo as geo_cell)
  template <int kGeo = 2, class R>
  __device__ __forceinline__ void add(const DevParams *P, TrackRegs<R> &t, double de, const double p[3], int v) {
    if (de <= 0.0) return;
    v = geo_cell<kGeo>(P->geo, v);
    t.aed += de;
    if (P->geo.visible[v]) t.avis += de;
    if (e == 0.0) {
      int z = (int)floor(p[2] / P->zbin);
      iz = z < 0 ? 0 : (z >= P->nz ? P->nz - 1 : z);
      const double r = sqrt(p[0] * p[0] + p[1] * p[1]);
      int rr = (int)(r / P->rbin);
      ir = rr < 0 ? 0 : (rr >= P->nr ? P->nr - 1 : rr);
      vol = v;
    }
    e += de;
  }
  __device__ __forceinline__ void flush(WaveHist &h) const {
    if (e <= 0.0) return;
    unsafeAtomicAdd(&h.v[vol], e);
    unsafeAtomicAdd(&h.z[iz], e);
    unsafeAtomicAdd(&h.r[ir], e);
  }
};

// A track dies: its tallies go to its event (one set of atomics per track).
template <class R>
__device__ __forceinline__ void retire(const StepArgs &a, const TrackRegs<R> &t) {
  DevEvent &E = a.ev[t.event];
  if (t.aed != 0.0) unsafeAtomicAdd(&E.edep, t.aed);
  if (t.avis != 0.0) unsafeAtomicAdd(&E.visible, t.avis);
  if (t.alen != 0.0) unsafeAtomicAdd(&E.clen, t.alen);
  atomicAdd(&E.nsteps, (unsigned long long)t.nsteps);
}

template <class R>
__device__ __forceinline__ void trace_rec(const StepArgs &a, const TrackRegs<R> &t, int proc, double edep) {
  if (t.event >= a.P->trace_events || t.nsteps > a.P->trace_steps) return;
  const unsigned long long k = atomicAdd(&a.cnt->trace_n, 1ULL);
  if ((int64_t)k >= a.trace_cap) return;
  double *r = a.trace + 12 * k;
  uint64_t u = t.uid;
  double ub;
  __builtin_memcpy(&ub, &u, 8);
  r[0] = ub;
  r[1] = t.nsteps;
  r[2] = t.species;
  r[3] = proc;
  r[4] = (double)t.e;
  r[5] = t.pos[0];
  r[6] = t.pos[1];
  r[7] = t.pos[2];
  r[8] = (double)t.dir[0];
  r[9] = (double)t.dir[1];
  r[10] = (double)t.dir[2];
  r[11] = edep;
}

// Write one new track into the workgroup's stripe of the next pool (region 0
// neutral, 1 charged) and its slot into entry nk of the stripe's new-track
// list.
template <class R>
__device__ __forceinline__ void emit_new(const StepArgs &a, const Pool<R> &nxt, int idx, int &nk,
                                         const TrackRegs<R> &par, int species, R e, R d0, R d1, R d2, uint64_t uid,
                                         WaveHist &wh, bool trace) {
  const int64_t sc = stripe_cap(nxt);
  const int st = my_stripe();
  const int64_t s = out_slot(species == 0 ? 0 : 1, st, idx, sc);
  if (s >= 0 && nk >= 0 && (int64_t)nk < sc) {
    store_new(a.P, nxt, s, par, species, e, d0, d1, d2, uid, trace);
    a.newlist[(int64_t)st * sc + nk] = (int32_t)s;
  } else {
    atomicOr(&a.cnt->overflow, 1);
  }
  ++nk;
  atomicAdd(&a.ev[par.event].nsec[species], 1ULL);
  spectrum(wh, species, (double)e);
}

// Positron at rest: two back-to-back 511 keV photons, isotropic (2 direction
// draws, then 2 substream draws per photon), at neutral slots idx, idx+1.
template <class R>
__device__ __forceinline__ void emit_at_rest(const StepArgs &a, const Pool<R> &nxt, int idx, int &nk,
                                             TrackRegs<R> &t, WaveHist &wh, bool trace) {
  const double cost = 2.0 * t.rng.uniform() - 1.0;
  const double sint = sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = kTwoPi * t.rng.uniform();
  const R d0 = (R)(sint * m_cos(phi)), d1 = (R)(sint * m_sin(phi)), d2 = (R)cost;
  const uint64_t u1 = child_substream(t.rng);
  emit_new(a, nxt, idx, nk, t, 0, (R)kMe, d0, d1, d2, u1, wh, trace);
  const uint64_t u2 = child_substream(t.rng);
  emit_new(a, nxt, idx + 1, nk, t, 0, (R)kMe, -d0, -d1, -d2, u2, wh, trace);
}

// PostStep GPIL of process p (ProxyEmProcess.hpp:122-151): resample the
// number of interaction lengths when exhausted, propose nil / lambda.
template <int p, class R, class NS>
__device__ __forceinline__ void gpil(const DevParams *P, const NS &N, const DevMaterial &m, int species, R e,
                                     R le, R &nil, R &xs, Rng &rng, double &bestd, int &sel) {
  xs = macro_xs(P, N, m, species, p, e, le);
  if (nil <= R(0)) nil = -m_log(rng.template u<R>());
  const double sp = xs > R(0) ? (double)(nil / xs) : kHuge;
  if (sp < bestd) {
    bestd = sp;
    sel = p;
  }
}

// ------------------------------------------------------------------ step loop kernels
// The state a primary starts from: no interaction lengths yet, the
// substream's initial RNG state, located from its position and direction.
template <class R>
__device__ __forceinline__ void inject_state(const DevParams *P, const StepArgs &a, const gxt_primary &p,
                                             TrackRegs<R> &t) {
  t.nil0 = t.nil1 = t.nil2 = R(-1);
  t.nsteps = 0;
  t.vol = geo_locate(P->geo, t.pos, p.dir);
  if (t.vol >= 0) substream_state(P, t.uid, t.rng.s);
}

// A track in flight continues from its own interaction lengths, RNG state,
// volume (located when not given) and step count; the event's step tally
// is credited with -nsteps now, so that the track's full count added when it
// dies leaves only the steps taken here (modulo 2^64 arithmetic).
template <class R>
__device__ __forceinline__ void inject_state(const DevParams *P, const StepArgs &a, const gxt_track &p,
                                             TrackRegs<R> &t) {
  t.nil0 = (R)p.nil[0];
  t.nil1 = (R)p.nil[1];
  t.nil2 = (R)p.nil[2];
  t.nsteps = p.nsteps;
  t.vol = p.volume >= 0 ? p.volume : geo_locate(P->geo, t.pos, p.dir);
  if (t.vol >= 0) {
#pragma unroll
    for (int k = 0; k < 6; ++k) t.rng.s[k] = p.rng[k];
    if (p.nsteps > 0) atomicAdd(&a.ev[t.event].nsteps, (unsigned long long)(-(long long)p.nsteps));
  }
}

// Entries In (gxt_primary or gxt_track) become live tracks of the current
// pool; their event slots are entry.event + event_base.
template <class R, class In>
__global__ __launch_bounds__(kBlock) void k_inject(StepArgs a, PoolArgs<R> pa, const In *prim, int n,
                                                   int event_base) {
  const DevParams *P = a.P;
  const Pool<R> &cur = pa.cur;
  const bool trace = P->trace_events > 0;
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < n;
    TrackRegs<R> t;
    int region = -1;
    if (active) {
      const In &p = prim[i];
      t.species = p.species;
      t.event = p.event + event_base;
      t.e = (R)p.ekin;
      for (int k = 0; k < 3; ++k) {
        t.pos[k] = p.pos[k];
        t.dir[k] = (R)p.dir[k];
      }
      t.aed = t.avis = t.alen = 0.0;
      t.uid = p.substream;
      inject_state(P, a, p, t);
      if (t.vol < 0) {
        unsafeAtomicAdd(&a.ev[t.event].leak, p.ekin);
      } else {
        region = t.species == 0 ? 0 : 1;
        atomicAdd(&a.ev_live[(int64_t)t.event * kLiveStride], 1);
      }
    }
    const int st = my_stripe();
    const int b0 = wave_reserve(region == 0 ? 1 : 0, &a.cnt->ctr(a.gen, kPool0, st));
    const int b1 = wave_reserve(region == 1 ? 1 : 0, &a.cnt->ctr(a.gen, kPool1, st));
    if (region >= 0) {
      const int64_t s = out_slot(region, st, region == 0 ? b0 : b1, stripe_cap(cur));
      if (s >= 0)
        store_track(cur, s, t, trace);
      else
        atomicOr(&a.cnt->overflow, 1);
    }
  }
}

// One kernel per region: kCharged = false steps the photons of [0, nN),
// kCharged = true the electrons and positrons of [cap - nC, cap); the species
// class is a compile-time constant so each variant carries only its own code.
// kMode 0: straight lines and the exact uniform-field helix, no multiple
// scattering; 1: the same propagation with e-/e+ multiple scattering; 2: e-/e+
// in a field map with the local helix, 3: with the RK45 stepper (map or
// uniform field), both through propagate_general and with or without
// multiple scattering; 4: the tree form of the geometry (volume-tree
// navigation, vacuum volumes), straight lines and the uniform-field helix,
// with or without multiple scattering.  Each mode is its own kernel, so the
// simpler ones keep their register budgets.
// The plain e-/e+ transport (no field map, RK45 or multiple scattering) fits
// 3 waves per SIMD with a 28 B/lane spill; that measured 1.6 % faster on the
// 10 GeV e- / PbWO4 batch than 2 waves without spill.  The other modes spill
// heavily at 3 and keep 2.
// With kLds the workgroup first copies its species' physics vectors (one node
// span per material slot, DevParams::lds_*) into dynamic LDS, where every
// lookup of the launch is served.
extern __shared__ uint4 tab_lds[];

// The workgroup's copy of DevParams in LDS, in 8-byte words: every thread
// of the (kBlock-thread) workgroup issues all its loads before its stores, so
// the copy costs one memory latency rather than one per pass of a strided
// loop.  params_load, then params_store, then a barrier (after any other
// staging, so that one barrier serves all).
constexpr int kParamWords = (int)(sizeof(DevParams) / 8);
constexpr int kParamPer = (kParamWords + kBlock - 1) / kBlock;
struct ParamWords {
  unsigned long long v[kParamPer];
};

__device__ __forceinline__ void params_load(const DevParams *P, ParamWords &w) {
  static_assert(sizeof(DevParams) % 8 == 0, "DevParams is copied in 8-byte words");
  const unsigned long long *src = reinterpret_cast<const unsigned long long *>(P);
#pragma unroll
  for (int j = 0; j < kParamPer; ++j) {
    const int k = j * kBlock + (int)threadIdx.x;
    if (k < kParamWords) w.v[j] = src[k];
  }
}

__device__ __forceinline__ const DevParams *params_store(DevParams &sP, const ParamWords &w) {
  unsigned long long *dst = reinterpret_cast<unsigned long long *>(&sP);
#pragma unroll
  for (int j = 0; j < kParamPer; ++j) {
    const int k = j * kBlock + (int)threadIdx.x;
    if (k < kParamWords) dst[k] = w.v[j];
  }
  return &sP;
}

// n 16-byte words from src to LDS dst by the workgroup's threads, four loads
// in flight per thread before their stores.
__device__ __forceinline__ void copy_words16(uint4 *dst, const uint4 *src, int n) {
  for (int k = (int)threadIdx.x; k < n; k += 4 * kBlock) {
    const int k1 = k + kBlock, k2 = k + 2 * kBlock, k3 = k + 3 * kBlock;
    const uint4 v0 = src[k];
    const uint4 v1 = k1 < n ? src[k1] : v0;
    const uint4 v2 = k2 < n ? src[k2] : v0;
    const uint4 v3 = k3 < n ? src[k3] : v0;
    dst[k] = v0;
    if (k1 < n) dst[k1] = v1;
    if (k2 < n) dst[k2] = v2;
    if (k3 < n) dst[k3] = v3;
  }
}

// The transport's species tables into dynamic LDS (no barrier: the caller's
// barrier after params_store covers both).
template <class R, bool kLds>
__device__ __forceinline__ NodesT<R, kLds> stage_tables(const DevParams *P, int layout) {
  using T = typename NodeType<R>::T;
  if constexpr (kLds) {
    constexpr int w = sizeof(T) / sizeof(uint4);
    const uint4 *src = reinterpret_cast<const uint4 *>(NodeType<R>::global(P));
    int dst = 0;
    for (int s = 0; s < kMaxSlots; ++s) {
      const int n16 = P->lds_cnt[layout][s] * w;
      copy_words16(tab_lds + dst, src + (int64_t)P->lds_src[layout][s] * w, n16);
      dst += n16;
    }
    return NodesT<R, true>(reinterpret_cast<const T *>(tab_lds), P->hints);
  } else {
    return NodesT<R, false>(P);
  }
}

// Register budgets (waves per SIMD), as measured on the 10 GeV e- / PbWO4
// batch: the plain FP32 e-/e+ transport at 3 (165 VGPRs); the FP64 one at
// 2 (190 VGPRs), 3 % faster than at 3, where it spilled 14 registers.
template <class R, bool kCharged, int kMode = 0, bool kLds = false>
__global__ __launch_bounds__(kBlock, (kCharged && kMode == 0 && sizeof(R) == 4) ? 3 : GXT_STEP_WAVES) void k_transport(
    StepArgs a, PoolArgs<R> pa) {
  const Pool<R> &cur = pa.cur;
  const Pool<R> &nxt = pa.nxt;
  // counts are clamped to the stripes so a corrupted counter cannot index past them
  const int64_t sc = stripe_cap(cur);
  const int st = my_stripe();
  const int g1 = gen_next(a.gen);
  StripeSpan span;
  bool bad;
  {
    // lane k < kStripes: stripe k's photon and e-/e+ counts
    const int l = lane();
    int64_t r0 = 0, r1 = 0;
    if (l < kStripes) {
      r0 = a.cnt->ctr(a.gen, kPool0, l);
      r1 = a.cnt->ctr(a.gen, kPool1, l);
    }
    const int64_t n0 = min(max(r0, (int64_t)0), sc);
    const int64_t n1 = min(max(r1, (int64_t)0), sc - n0);
    span.from_lanes((int)(kCharged ? n1 : n0));
    bad = __ballot(r0 < 0 || r1 < 0 || r0 + r1 > sc) != 0ULL;
  }
  // injected primaries can fill a stripe from both ends: report, never overlap
  if (bad && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&a.cnt->overflow, 1);
  // after an overflow (set by an earlier iteration, which ended before this
  // one began) the pool is dropped whole: a slot whose store was skipped is
  // never read; the host reports the overflow at its next snapshot
  const int total = a.cnt->overflow ? 0 : span.total();
  // the photon transport clears the generation the next iteration fills (no
  // kernel of this iteration uses it; the next starts after this one); after
  // the counter loads, which stay scalar loads
  if (!kCharged && blockIdx.x == 0)
    for (int i = threadIdx.x; i < kStripes * kNumKinds; i += blockDim.x)
      a.cnt->ctr(gen_next(g1), i % kNumKinds, i / kNumKinds) = 0;
  // a workgroup without tracks (the grid is fixed, the count is on the
  // device) leaves before it stages anything
  if ((int64_t)blockIdx.x * blockDim.x >= total) return;
  __shared__ WaveHist whs[kBlock / 64];
  __shared__ DevParams sP;
  WaveHist &wh = whs[threadIdx.x >> 6];
  wh_clear(wh);
  ParamWords pw;
  params_load(a.P, pw);
  const NodesT<R, kLds> N = stage_tables<R, kLds>(a.P, kCharged ? 1 : 0);
  const DevParams *P = params_store(sP, pw);
  __syncthreads();
  const bool trace = P->trace_events > 0;
  const R mine = (R)P->tracking_cut;
  constexpr int kG = kMode == 4 ? 1 : 0;  // geo_cell form
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < total;
    TrackRegs<R> t;
    int q = -1;
    int slot = 0;
    bool survive = false, annihilate = false, dies = false, killed = false, escaped = false;
    int nsub = 0;  // steps taken by this track in this launch
    double edep = 0.0;
    if (active) {
      int is, off;
      span.locate(i, is, off);
      slot = (int)out_slot(kCharged ? 1 : 0, is, off, sc);
      load_track(cur, slot, t, trace);
      if (!kCharged) t.species = 0;
      // species as a value the compiler can bound: 0 here, or 1 / 2 there
      const int sp = kCharged ? (t.species == 2 ? 2 : 1) : 0;
      const bool charged = kCharged;
      // a track keeps stepping while it neither interacts nor ends, up to
      // a.inner_steps steps: each step is the same as one iteration of the loop
      for (;;) {
      survive = false;
      edep = 0.0;
      Dep dA, dB;  // along-step loss (mid-step, pre-step volume); local deposit at the end point
      // material slot of the track's volume; -1 (tree form only): vacuum, no
      // physics, the track flies to the next boundary
      const int mslot = P->geo.slot[geo_cell<kG>(P->geo, t.vol)];
      const bool vac = kMode == 4 && mslot < 0;
      const DevMaterial &m = P->mat[vac ? 0 : mslot];
      R xs0 = R(0), xs1 = R(0), xs2 = R(0);
      double bestd = kHuge;
      int sel = -1;
      double lam1 = kMscNone;
      // the step's table lookups share log(E) and the e-/e+ range at E
      const R le = m_log(t.e);
      R range = R(0);
      if (!vac) {
      gpil<0>(P, N, m, sp, t.e, le, t.nil0, xs0, t.rng, bestd, sel);
      gpil<1>(P, N, m, sp, t.e, le, t.nil1, xs1, t.rng, bestd, sel);
      if (sp != 1) gpil<2>(P, N, m, sp, t.e, le, t.nil2, xs2, t.rng, bestd, sel);
      if (charged) {
        const double al = (double)along_limit(N, m, sp == 2, t.e, le, mine, range);
        if (al < bestd) {
          bestd = al;
          sel = -1;
        }
      }
      // multiple scattering: its step limit; then the geometric length of the
      // true path is propagated (the step is the true path for interaction
      // lengths and loss)
      if constexpr (kMode != 0) {
        if (charged && P->msc) {
          lam1 = msc_lambda1<R>(N, m, sp == 2, t.e, le);
          if (lam1 < kMscNone) {
            const double saf = kMode == 4 ? nav_safety(P->geo.nav, t.vol, t.pos) : safety(P->geo, t.vol, t.pos);
            const double tlim = fmax(fmax(kMscFacRange * fmax((double)range, lam1), kMscFacSafety * saf), kMscTmin);
            if (tlim < bestd) {
              bestd = tlim;
              sel = -1;
            }
          }
        }
      }
      }
      const double pre[3] = {t.pos[0], t.pos[1], t.pos[2]};
      const int prevol = t.vol;
      double step = 0.0;
      int status;
      if constexpr (kMode != 0) {
        const bool msc_on = lam1 < kMscNone;
        const double request = msc_on ? msc_geom(bestd, lam1) : bestd;
        double gstep = 0.0;
        if constexpr (kMode == 4)
          status = propagate_tree<R>(P, sp, t.e, t.pos, t.dir, t.vol, request, gstep);
        else if constexpr (kMode == 2)
          status = propagate_general<R, GXT_STEPPER_HELIX>(P, sp, t.e, t.pos, t.dir, t.vol, request, gstep);
        else if constexpr (kMode == 3)
          status = propagate_general<R, GXT_STEPPER_RK45>(P, sp, t.e, t.pos, t.dir, t.vol, request, gstep);
        else
          status = propagate<R>(P, sp, t.e, t.pos, t.dir, t.vol, request, gstep);
        step = msc_on ? (status == kPhysics ? bestd : msc_true(gstep, lam1, bestd)) : gstep;
      } else {
        status = propagate<R>(P, sp, t.e, t.pos, t.dir, t.vol, bestd, step);
      }
      if (xs0 > R(0)) t.nil0 -= (R)step * xs0;
      if (xs1 > R(0)) t.nil1 -= (R)step * xs1;
      if (xs2 > R(0)) t.nil2 -= (R)step * xs2;
      t.nsteps++;
      if (charged && !vac) {
        t.alen += step;
        R loss = R(0);
        t.e = along_doit(N, m, t.species == 2, t.e, le, range, (R)step, mine, loss);
        const double mid[3] = {0.5 * (pre[0] + t.pos[0]), 0.5 * (pre[1] + t.pos[1]), 0.5 * (pre[2] + t.pos[2])};
        dA.add<kG>(P, t, (double)loss, mid, prevol);
        edep += (double)loss;
        if constexpr (kMode != 0) {
          if (lam1 < kMscNone && status != kEscaped && t.e > R(0) && step > 0.0)
            msc_deflect<R>(step / lam1, t.dir, t.rng);
        }
      }
      int tproc = status == kPhysics ? sel : (status == kBoundary ? -2 : (status == kEscaped ? -3 : -1));
      if (status == kPhysics && sel < 0) tproc = -1;
      if (status == kEscaped) {
        if (trace) trace_rec(a, t, tproc, edep);
        escaped = true;
        dies = true;
      } else if (t.e <= R(0)) {
        if (trace) trace_rec(a, t, tproc, edep);
        annihilate = t.species == 2;
        dies = !annihilate;
      } else if (status == kPhysics && sel >= 0) {
        t.set_nil(sel, R(-1));
        q = queue_of(t.species, sel);
      } else {
        if (trace) trace_rec(a, t, tproc, edep);
        if (t.e < mine) {
          dB.add<kG>(P, t, (double)t.e, t.pos, t.vol);
          annihilate = t.species == 2;
          dies = !annihilate;
        } else if (t.nsteps >= P->max_steps) {
          dB.add<kG>(P, t, (double)t.e, t.pos, t.vol);
          killed = true;
          dies = true;
        } else {
          survive = true;
        }
      }
      // positron at rest: its two photons are made by the interaction kernel
      if (annihilate) q = kQueueAtRest;
      // the step's deposits go to the wave's LDS histograms now (LDS
      // atomics do not hold up the global reservations below)
      dA.flush(wh);
      dB.flush(wh);
      ++nsub;
      if (!survive || nsub >= a.inner_steps) break;
      }
    }
    // returning atomics first, all in flight together: the survivor's slot in
    // the next pool and the queue entry
    WaveReserve rs;
    rs.issue(survive ? 1 : 0, &a.cnt->ctr(g1, kCharged ? kPool1 : kPool0, st));
    const int64_t qpos = queue_reserve<kCharged ? 3 : 0, kCharged ? kNumQueues - 3 : 3>(q, a.cnt, g1, cur.cap, sc);
    const int bs = rs.get();
    // then stores and no-return atomics
    if (active) {
      if (escaped) unsafeAtomicAdd(&a.ev[t.event].leak, (double)t.e);
      if (killed) atomicAdd(&a.cnt->stat[my_stripe()][kStatKilled], 1ULL);
      if (dies) {
        retire(a, t);
        atomicAdd(&a.ev_live[(int64_t)t.event * kLiveStride], -1);
      }
      if (q >= 0) {
        // interacting (or stopped) track: back in place for the interaction kernel
        store_track<R, true, !kCharged>(cur, slot, t, trace);
        if (trace && q != kQueueAtRest) cur.tedep[slot] = edep;
      }
      if (survive) {
        const int64_t s = out_slot(kCharged ? 1 : 0, st, bs, sc);
        if (s >= 0)
          store_track(nxt, s, t, trace);
        else
          atomicOr(&a.cnt->overflow, 1);
      }
      if (qpos >= 0)
        a.queues[qpos] = slot;
      else if (qpos == -2)
        atomicOr(&a.cnt->overflow, 1);
    }
    {
      // the wave's steps and loaded track records, one atomic each
      const int ns = wave_incl_scan(nsub);
      const int tot = __shfl(ns, 63, 64);
      const unsigned long long nload = (unsigned long long)__popcll(__ballot(active));
      if (tot > 0 && lane() == 63) {
        unsigned long long *sp = a.cnt->stat[st];
        atomicAdd(&sp[kStatSteps], (unsigned long long)tot);
        if (kCharged) atomicAdd(&sp[kStatStepsCharged], (unsigned long long)tot);
        atomicAdd(&sp[kCharged ? kStatLoaded1 : kStatLoaded0], nload);
      }
    }
  }
  wh_flush(a, wh);
}

// Fate of a secondary: pushed (>= tracking cut), or deposited (and, for a
// positron, annihilated at rest).
template <class R>
__device__ __forceinline__ int sec_fate(const DevParams *P, TrackRegs<R> &t, Dep &dp, const Sec<R> &s, R mine,
                                        int &outN, int &outC, int &nannih) {
  if (s.e < mine) {
    dp.add(P, t, (double)s.e, t.pos, t.vol);
    if (s.species == 2) {
      ++nannih;
      outN += 2;
      return 2;
    }
    return 0;
  }
  if (s.species == 0)
    outN++;
  else
    outC++;
  return 1;
}

// Interactions compiled for at least 2 waves per SIMD: FP64 takes 161 VGPRs
// (3 waves fit); bounded to 3 it took 150 and measured 1 % slower beside the
// 2-wave FP64 e-/e+ transport (the chains share the CUs).
template <class R, bool kCharged>
__global__ __launch_bounds__(kBlock, GXT_STEP_WAVES) void k_interact(StepArgs a, PoolArgs<R> pa) {
  const Pool<R> &cur = pa.cur;
  const Pool<R> &nxt = pa.nxt;
  const int64_t sc = stripe_cap(cur);
  const int st = my_stripe();
  // photon queues 0-2 in one variant, e-/e+ queues 3-8 in the other; the
  // variant's (queue, stripe) segments are walked in order through a prefix
  // of their lengths (clamped to the stripe capacity)
  constexpr int kQ0q = kCharged ? 3 : 0;
  constexpr int kNQ = kCharged ? kNumQueues - 3 : 3;
  constexpr int kSeg = kNQ * kStripes;
  __shared__ int seg[kSeg + 1];
  if (threadIdx.x < 64) {
    // wave 0: one counter per lane and 64-counter chunk, all loads in flight
    // at once (a loop of dependent scalar loads took one memory round trip
    // per eight), then a prefix over the lanes, chunk after chunk
    constexpr int kChunks = (kSeg + 63) / 64;
    int v[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int k = c * 64 + (int)threadIdx.x;
      v[c] = k < kSeg ? a.cnt->ctr(gen_next(a.gen), kQ0 + kQ0q + k / kStripes, k % kStripes) : 0;
    }
    int carry = 0;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int k = c * 64 + (int)threadIdx.x;
      const int incl = wave_incl_scan((int)min((int64_t)max(v[c], 0), sc)) + carry;
      if (k < kSeg) seg[k + 1] = incl;
      carry = __shfl(incl, 63, 64);
    }
    if (threadIdx.x == 0) seg[0] = 0;
  }
  __syncthreads();
  const int total = seg[kSeg];
  // a workgroup without items leaves before it stages anything
  if ((int64_t)blockIdx.x * blockDim.x >= total) return;
  __shared__ WaveHist whs[kBlock / 64];
  WaveHist &wh = whs[threadIdx.x >> 6];
  wh_clear(wh);
  __shared__ DevParams sP;
  ParamWords pw;
  params_load(a.P, pw);
  const DevParams *P = params_store(sP, pw);
  __syncthreads();
  const bool trace = P->trace_events > 0;
  const R mine = (R)P->tracking_cut;
  // this variant's new-track list
  a.newlist += kCharged ? cur.cap : 0;
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < total;
    TrackRegs<R> t;
    Interaction<R> o;
    o.nsec = 0;
    int fate0 = 0, fate1 = 0;  // 0 deposited, 1 pushed, 2 annihilated at rest
    bool survive = false, prim_rest = false, killed = false, pair = false;
    int outN = 0, outC = 0, nannih = 0;
    Dep dp;  // every deposit of an interaction is at the track's position
    if (active) {
      // segment of item i: binary search over the prefix (seg[0] = 0 <= i < total)
      int lo = 0, hi = kSeg;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (i >= seg[mid])
          lo = mid;
        else
          hi = mid;
      }
      const int q = kQ0q + lo / kStripes;
      const int qs = lo % kStripes;
      const int slot = a.queues[(int64_t)q * cur.cap + (int64_t)qs * sc + (i - seg[lo])];
      load_track(cur, slot, t, trace);
      if (!kCharged) t.species = 0;
      if (kCharged && q == kQueueAtRest) {
        // positron stopped by the continuous loss or the tracking cut
        ++nannih;
        prim_rest = true;
        outN = 2;
        t.e = R(0);
      } else {
        // an interacting track is in a material (the tree form's vacuum has no physics)
        const DevMaterial &m = P->mat[max(P->geo.slot[geo_cell<2>(P->geo, t.vol)], 0)];
        if (kCharged)
          sample_charged<R>(P, m, q, t.e, t.dir, t.rng, o);
        else
          sample_neutral<R>(P, m, q, t.e, t.dir, t.rng, o);
        pair = q == 1 && o.nsec == 2;
        if (q == 7) ++nannih;
        dp.add(P, t, (double)o.edep, t.pos, t.vol);
        t.e = o.ekin;
        t.dir[0] = o.d[0];
        t.dir[1] = o.d[1];
        t.dir[2] = o.d[2];
        if (trace) trace_rec(a, t, queue_proc(q), cur.tedep[slot] + (double)o.edep);
        if (o.nsec > 0) fate0 = sec_fate(P, t, dp, o.sec[0], mine, outN, outC, nannih);
        if (o.nsec > 1) fate1 = sec_fate(P, t, dp, o.sec[1], mine, outN, outC, nannih);
      }
      if (t.e > R(0)) {
        if (t.e < mine) {
          dp.add(P, t, (double)t.e, t.pos, t.vol);
          if (t.species == 2) {
            ++nannih;
            prim_rest = true;
            outN += 2;
          }
        } else if (t.nsteps >= P->max_steps) {
          dp.add(P, t, (double)t.e, t.pos, t.vol);
          killed = true;
        } else {
          survive = true;
          if (t.species == 0)
            outN++;
          else
            outC++;
        }
      }
    }
    // returning atomics first, all in flight together; then the stores
    const int made = outN + outC - (survive ? 1 : 0);
    WaveReserve rN, rC, rK;
    const int g1 = gen_next(a.gen);
    rN.issue(outN, &a.cnt->ctr(g1, kPool0, st));
    rC.issue(outC, &a.cnt->ctr(g1, kPool1, st));
    rK.issue(made, &a.cnt->ctr(g1, kCharged ? kNew1 : kNew0, st));
    const int bN = rN.get(), bC = rC.get();
    int nk = rK.get();
    if (active) {
      dp.flush(wh);
      if (pair) atomicAdd(&a.ev[t.event].npairs, 1ULL);
      if (nannih) atomicAdd(&a.ev[t.event].nannihil, (unsigned long long)nannih);
      if (killed) atomicAdd(&a.cnt->stat[my_stripe()][kStatKilled], 1ULL);
      if (!survive) retire(a, t);
      if (outN + outC != 1) atomicAdd(&a.ev_live[(int64_t)t.event * kLiveStride], outN + outC - 1);
      // new tracks in the oracle's draw order: secondary 0, secondary 1, then
      // the primary's annihilation at rest; the survivor last
      int kN = bN, kC = bC;
      if (o.nsec > 0) {
        const Sec<R> &s = o.sec[0];
        if (fate0 == 1) {
          const uint64_t u = child_substream(t.rng);
          emit_new(a, nxt, s.species == 0 ? kN++ : kC++, nk, t, s.species, s.e, s.d[0], s.d[1], s.d[2], u,
                   wh, trace);
        } else if (fate0 == 2) {
          emit_at_rest(a, nxt, kN, nk, t, wh, trace);
          kN += 2;
        }
      }
      if (o.nsec > 1) {
        const Sec<R> &s = o.sec[1];
        if (fate1 == 1) {
          const uint64_t u = child_substream(t.rng);
          emit_new(a, nxt, s.species == 0 ? kN++ : kC++, nk, t, s.species, s.e, s.d[0], s.d[1], s.d[2], u,
                   wh, trace);
        } else if (fate1 == 2) {
          emit_at_rest(a, nxt, kN, nk, t, wh, trace);
          kN += 2;
        }
      }
      if (prim_rest) {
        emit_at_rest(a, nxt, kN, nk, t, wh, trace);
        kN += 2;
      }
      if (survive) {
        const int64_t s = t.species == 0 ? out_slot(0, st, kN, sc) : out_slot(1, st, kC, sc);
        if (s >= 0)
          store_track(nxt, s, t, trace);
        else
          atomicOr(&a.cnt->overflow, 1);
      }
    }
    wave_add_u64(&a.cnt->stat[st][kStatInteractions], active);
    if (!kCharged) wave_add_u64(&a.cnt->stat[st][kStatInteractions0], active);
    {
      const int tot = __shfl(rK.incl, 63, 64);
      if (tot > 0 && lane() == 63) {
        atomicAdd(&a.cnt->stat[st][kStatCreated], (unsigned long long)tot);
        if (!kCharged) atomicAdd(&a.cnt->stat[st][kStatCreated0], (unsigned long long)tot);
      }
    }
  }
  wh_flush(a, wh);
}

// RNG state of the tracks created this iteration: substream uid of the seed,
// one MRG32k3a component per lane, two lanes per track (the jump is 8
// table-driven matrix-vector products per component, so it runs here,
// without divergence and at full occupancy, rather than in the interaction
// kernels).
template <class R, int kList>
__global__ __launch_bounds__(kBlock) void k_derive(StepArgs a, PoolArgs<R> pa) {
  const Pool<R> &nxt = pa.nxt;
  const int64_t sc = stripe_cap(nxt);
  a.newlist += kList * nxt.cap;
  StripeSpan span;
  {
    const int l = lane();
    const int64_t r = l < kStripes ? a.cnt->ctr(gen_next(a.gen), kNew0 + kList, l) : 0;
    span.from_lanes((int)min(max(r, (int64_t)0), sc));
  }
  const int n = span.total();
  // two adjacent lanes of a wave per track, one MRG32k3a component each;
  // both read the substream index (RNG words 1 and 2) before either writes
  // its words: the loads complete before the products that use them
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < 2 * n; j += gridDim.x * blockDim.x) {
    const int i = j >> 1, c = j & 1;
    int is, off;
    span.locate(i, is, off);
    const int64_t s = a.newlist[(int64_t)is * sc + off];
    if (s < 0 || s >= nxt.cap) continue;
    const uint64_t uid = (uint64_t)nxt.rng[nxt.cap + s] | ((uint64_t)nxt.rng[2 * nxt.cap + s] << 32);
    const uint32_t m = c ? 4294944443u : 4294967087u;
    const double invm = c ? 1.0 / 4294944443.0 : 1.0 / 4294967087.0;
    uint32_t v[3] = {a.P->seed[3 * c], a.P->seed[3 * c + 1], a.P->seed[3 * c + 2]};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      uint32_t M[9];
      jump_row(a.P->jump8, c, k, (int)((uid >> (8 * k)) & 255u), M);
      matvec_mod(M, v, m, invm);
    }
    nxt.rng[(int64_t)(3 * c) * nxt.cap + s] = v[0];
    nxt.rng[(int64_t)(3 * c + 1) * nxt.cap + s] = v[1];
    nxt.rng[(int64_t)(3 * c + 2) * nxt.cap + s] = v[2];
  }
}

// ------------------------------------------------------------------ component kernels
__global__ void k_rng_init(const DevParams *P, const uint32_t *seed, const uint64_t *idx, int n, uint32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t v1[3], v2[3];
  jump_substream(P->jump8, seed, idx[i], v1, v2);
  out[6 * i + 0] = v1[0];
  out[6 * i + 1] = v1[1];
  out[6 * i + 2] = v1[2];
  out[6 * i + 3] = v2[0];
  out[6 * i + 4] = v2[1];
  out[6 * i + 5] = v2[2];
}

__global__ void k_rng_draw(uint32_t *states, int n, int k, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng r;
  for (int j = 0; j < 6; ++j) r.s[j] = states[6 * i + j];
  for (int j = 0; j < k; ++j) out[(int64_t)i * k + j] = r.uniform();
  for (int j = 0; j < 6; ++j) states[6 * i + j] = r.s[j];
}

// Straight-line navigation of the analytic geometry for n points (the
// locate, dist_out and safety of the step loop).
__global__ void k_slab_distance(const DevParams *P, const double *pos, const double *dir, const double *pstep,
                                int32_t *vol, double *snext, int32_t *next, double *saf, int32_t *boundary, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevGeometry &g = P->geo;
  const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
  const double d0 = dir[i], d1 = dir[n + i], d2 = dir[2 * n + i];
  int v = vol[i];
  if (v < 0) v = locate(g, p, d0, d1, d2);
  vol[i] = v;
  if (v < 0) {
    snext[i] = 0.0;
    saf[i] = 0.0;
    next[i] = -1;
    if (boundary) boundary[i] = 1;
    return;
  }
  int face;
  const double s = dist_out(g, v, p, d0, d1, d2, face);
  int nv = -1;
  if (face == 0 || face == 1) {
    nv = face == 1 ? v + 1 : v - 1;
    if (nv >= g.nvol) nv = -1;
  }
  snext[i] = s;
  next[i] = nv;
  saf[i] = safety(g, v, p);
  if (boundary) boundary[i] = pstep ? !(pstep[i] < s) : 1;
}

template <class R>
__global__ void k_table_value(const DevParams *P, int slot, int table, const double *x, double *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Nodes<R> N(P);
  const DevTable &t = P->mat[slot].tab[table];
  const bool inv = table == GXT_TAB_INVRANGE_EM || table == GXT_TAB_INVRANGE_EP;
  out[i] = (double)(inv ? pvec_inv_value(N, t, (R)x[i]) : pvec_value(N, t, (R)x[i]));
}

template <class R>
__global__ void k_macro_xs(const DevParams *P, int slot, int species, int proc, const double *e, double *out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Nodes<R> N(P);
  out[i] = (double)macro_xs(P, N, P->mat[slot], species, proc, (R)e[i]);
}

template <class R>
__global__ void k_sample(const DevParams *P, int slot, int model, SampleArgs s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = s.n;
  if (i >= n) return;
  Rng r;
  for (int j = 0; j < 6; ++j) r.s[j] = s.rng[6 * i + j];
  const R d[3] = {(R)s.dir[i], (R)s.dir[n + i], (R)s.dir[2 * n + i]};
  Interaction<R> o;
  sample_model<R>(P, P->mat[slot], model, (R)s.ekin[i], d, r, o);
  for (int j = 0; j < 6; ++j) s.rng[6 * i + j] = r.s[j];
  s.out_ekin[i] = (double)o.ekin;
  s.out_dir[i] = (double)o.d[0];
  s.out_dir[n + i] = (double)o.d[1];
  s.out_dir[2 * n + i] = (double)o.d[2];
  s.out_edep[i] = (double)o.edep;
  s.nsec[i] = o.nsec;
  for (int k = 0; k < 2; ++k) {
    const bool has = k < o.nsec;
    s.sec_species[2 * i + k] = has ? o.sec[k].species : -1;
    s.sec_ekin[2 * i + k] = has ? (double)o.sec[k].e : 0.0;
    for (int j = 0; j < 3; ++j) s.sec_dir[6 * i + 3 * k + j] = has ? (double)o.sec[k].d[j] : 0.0;
  }
}

template <class R>
__global__ void k_helix(double b0, double b1, double b2, const double *charge, const double *mom, const double *step,
                        double *pos, double *dir, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double B[3] = {b0, b1, b2};
  const double bmag = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
  double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
  R d[3] = {(R)dir[i], (R)dir[n + i], (R)dir[2 * n + i]};
  helix_step<R>(B, bmag, (R)charge[i], (R)mom[i], (R)step[i], p, d);
  for (int k = 0; k < 3; ++k) {
    pos[k * n + i] = p[k];
    dir[k * n + i] = (double)d[k];
  }
}

__global__ void k_field_value(DevField F, const double *pos, double *b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double B[3];
  map_field(F, pos[i], pos[n + i], pos[2 * n + i], B);
  for (int k = 0; k < 3; ++k) b[k * n + i] = B[k];
}

// Volume-tree queries.  With kLds the workgroup first copies the whole tree
// (nodes, grids and daughter index: lds_bytes, a multiple of 16, read from
// one padded allocation) into LDS, where the node lookups of all its lanes are
// served; the grid then strides over the queries.  Without it the lookups go
// to the tree in device memory.
extern __shared__ uint4 nav_lds[];

template <bool kLds>
__device__ __forceinline__ DevNav nav_view(const DevNav &g, int lds_bytes) {
  if (!kLds) return g;
  const uint4 *src = reinterpret_cast<const uint4 *>(g.node);
  for (int k = threadIdx.x; k < lds_bytes / 16; k += blockDim.x) nav_lds[k] = src[k];
  __syncthreads();
  const char *base = reinterpret_cast<const char *>(g.node);
  const char *lds = reinterpret_cast<const char *>(nav_lds);
  DevNav l = g;
  l.node = reinterpret_cast<const NavNode *>(lds);
  l.grid = reinterpret_cast<const NavGrid *>(lds + (reinterpret_cast<const char *>(g.grid) - base));
  l.idx = reinterpret_cast<const int32_t *>(lds + (reinterpret_cast<const char *>(g.idx) - base));
  return l;
}

template <bool kLds>
__global__ __launch_bounds__(kBlock) void k_nav_locate(DevNav g0, int lds_bytes, const double *pos, const double *dir,
                                                       int32_t *out, int n) {
  const DevNav g = nav_view<kLds>(g0, lds_bytes);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    const double v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    out[i] = nav_locate(g, p, v);
  }
}

// A touchable outside [0, g.ntouch) gives next = -2 and NaN distances.
template <bool kLds>
__global__ __launch_bounds__(kBlock) void k_nav_step(DevNav g0, int lds_bytes, const int32_t *touchable,
                                                     const double *pos, const double *dir, double *snext,
                                                     int32_t *next, double *safety, int n) {
  const DevNav g = nav_view<kLds>(g0, lds_bytes);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int t = touchable[i];
    if (t < 0 || t >= g.ntouch) {
      next[i] = -2;
      snext[i] = __builtin_nan("");
      safety[i] = __builtin_nan("");
      continue;
    }
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    const double v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    double s, saf;
    next[i] = nav_step(g, t, p, v, s, saf);
    snext[i] = s;
    safety[i] = saf;
  }
}

// A touchable outside [0, g.ntouch) or a pstep that is negative or not
// finite gives next = -2 and NaN distances.
template <bool kLds>
__global__ __launch_bounds__(kBlock) void k_nav_find_next_boundary(DevNav g0, int lds_bytes, const int32_t *touchable,
                                                                   const double *pos, const double *dir,
                                                                   const double *pstep, double *snext, int32_t *next,
                                                                   double *safety, int32_t *boundary, int n) {
  const DevNav g = nav_view<kLds>(g0, lds_bytes);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int t = touchable[i];
    const double ps = pstep[i];
    if (t < 0 || t >= g.ntouch || !(ps >= 0.0) || !(ps <= 1e300)) {
      next[i] = -2;
      snext[i] = __builtin_nan("");
      safety[i] = __builtin_nan("");
      boundary[i] = 0;
      continue;
    }
    const double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
    const double v[3] = {dir[i], dir[n + i], dir[2 * n + i]};
    double s, saf;
    int b;
    next[i] = nav_find_next_boundary(g, t, p, v, ps, s, saf, b);
    snext[i] = s;
    safety[i] = saf;
    boundary[i] = b;
  }
}

template <class R>
__global__ void k_field_step(DevField F, const double *charge, const double *mom, const double *step, double *pos,
                             double *dir, double *err, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p[3] = {pos[i], pos[n + i], pos[2 * n + i]};
  R d[3] = {(R)dir[i], (R)dir[n + i], (R)dir[2 * n + i]};
  double B[3];
  field_at(F, p, B);
  const double bm = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  double ex = 0.0, eu = 0.0;
  field_advance<R>(F, B, bm, (R)charge[i], (R)mom[i], step[i], p, d, ex, eu);
  for (int k = 0; k < 3; ++k) {
    pos[k * n + i] = p[k];
    dir[k * n + i] = (double)d[k];
  }
  if (err) {
    err[i] = ex;
    err[n + i] = eu;
  }
}

// ------------------------------------------------------------------ launchers
static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// A step kernel launch; with events (timing only) they take the kernel's own
// start and end times from its dispatch, so no marker packet sits between
// two kernels of a chain.
template <class K, class... Args>
static inline void launch_k(K kernel, int grid, size_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                            Args... args) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), (uint32_t)lds, st, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), lds, st, args...);
}

template <class R>
hipError_t launch_step(const StepArgs &a, const PoolArgs<R> &pa, int grid, const StepStreams &s, hipEvent_t *ev) {
  // The photon chain (transport, interactions, RNG derivation) runs on
  // s.main and the e-/e+ chain on s.side: they read disjoint regions and
  // queues and write the next pool, queues and tallies only through atomics,
  // so they overlap and each fills the other's tail.  Both chains of this
  // iteration follow both of the previous one: each stream waits (queued at
  // the end of the previous launch) for the event the other recorded after
  // its chain.
  hipEvent_t none[2 * kStepKernels] = {};
  hipEvent_t *E = ev ? ev : none;
  // the RNG derivation (80 VGPRs, a chain of dependent modular products per
  // lane) runs twice the others' workgroups per CU: more waves to hide that
  // latency (measured +0.9 % FP64, +1.3 % FP32; half the grid -3.9 %)
  const int dgrid = 2 * grid;
  // the FP64 interactions (168 VGPRs, 3 waves per SIMD) launch the 3
  // workgroups per CU that fit (+0.5 % FP64, measured three times; FP32
  // -0.7 %, so it keeps 4)
  const int igrid = sizeof(R) == 8 ? grid * 3 / 4 : grid;
  const size_t l0 = a.lds_bytes[0], l1 = a.lds_bytes[1];
  if (a.transport_mode == 4 && l0 > 0)
    launch_k(k_transport<R, false, 4, true>, grid, l0, s.main, E[0], E[1], a, pa);
  else if (a.transport_mode == 4)
    launch_k(k_transport<R, false, 4>, grid, 0, s.main, E[0], E[1], a, pa);
  else if (l0 > 0)
    launch_k(k_transport<R, false, 0, true>, grid, l0, s.main, E[0], E[1], a, pa);
  else
    launch_k(k_transport<R, false>, grid, 0, s.main, E[0], E[1], a, pa);
  launch_k(k_interact<R, false>, igrid, 0, s.main, E[2], E[3], a, pa);
  launch_k(k_derive<R, 0>, dgrid, 0, s.main, E[4], E[5], a, pa);
  hipError_t e = hipEventRecord(s.fork, s.main);
  if (e != hipSuccess) return e;
  // the FP64 e-/e+ transports without field sub-steps (plain, with msc) fit
  // 2 workgroups per CU and launch only those: a second round of workgroups
  // waiting for a slot had made the plain one 24 % slower (128.5 against
  // 104 us per launch; +4 % events/s, +4 % with msc).  Where per-track work
  // is irregular (field maps, RK45, the tree in a field) the spare
  // workgroups balance it: 2 per CU measured 2-13 % slower there.
  const int egrid = (sizeof(R) == 8 && a.transport_mode <= 1) ? grid / 2 : grid;
  if (a.transport_mode == 4 && l1 > 0)
    launch_k(k_transport<R, true, 4, true>, egrid, l1, s.side, E[6], E[7], a, pa);
  else if (a.transport_mode == 4)
    launch_k(k_transport<R, true, 4>, egrid, 0, s.side, E[6], E[7], a, pa);
  else if (a.transport_mode == 3)
    launch_k(k_transport<R, true, 3>, egrid, 0, s.side, E[6], E[7], a, pa);
  else if (a.transport_mode == 2)
    launch_k(k_transport<R, true, 2>, egrid, 0, s.side, E[6], E[7], a, pa);
  else if (a.transport_mode == 1)
    launch_k(k_transport<R, true, 1>, egrid, 0, s.side, E[6], E[7], a, pa);
  else if (l1 > 0)
    launch_k(k_transport<R, true, 0, true>, egrid, l1, s.side, E[6], E[7], a, pa);
  else
    launch_k(k_transport<R, true>, egrid, 0, s.side, E[6], E[7], a, pa);
  launch_k(k_interact<R, true>, igrid, 0, s.side, E[8], E[9], a, pa);
  launch_k(k_derive<R, 1>, dgrid, 0, s.side, E[10], E[11], a, pa);
  e = hipEventRecord(s.join, s.side);
  // each stream's wait is queued after its own chain: main's later work
  // (the next photon chain, snapshots, injections, synchronisation) follows
  // the e-/e+ chain, the next e-/e+ chain the photon chain
  if (e == hipSuccess) e = hipStreamWaitEvent(s.main, s.join, 0);
  if (e == hipSuccess) e = hipStreamWaitEvent(s.side, s.fork, 0);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_side_sync(const StepStreams &s) {
  hipError_t e = hipEventRecord(s.entry, s.main);
  if (e == hipSuccess) e = hipStreamWaitEvent(s.side, s.entry, 0);
  return e;
}

template <class R, class In>
hipError_t launch_inject(const StepArgs &a, const PoolArgs<R> &pa, const In *prim, int n, int event_base,
                         hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int g = grid_for(n);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL((k_inject<R, In>), dim3(g), dim3(kBlock), 0, st, a, pa, prim, n, event_base);
  return hipGetLastError();
}

template hipError_t launch_step<double>(const StepArgs &, const PoolArgs<double> &, int, const StepStreams &,
                                        hipEvent_t *);
template hipError_t launch_step<float>(const StepArgs &, const PoolArgs<float> &, int, const StepStreams &,
                                       hipEvent_t *);
template hipError_t launch_inject<double, gxt_primary>(const StepArgs &, const PoolArgs<double> &,
                                                       const gxt_primary *, int, int, hipStream_t);
template hipError_t launch_inject<float, gxt_primary>(const StepArgs &, const PoolArgs<float> &, const gxt_primary *,
                                                      int, int, hipStream_t);
template hipError_t launch_inject<double, gxt_track>(const StepArgs &, const PoolArgs<double> &, const gxt_track *,
                                                     int, int, hipStream_t);
template hipError_t launch_inject<float, gxt_track>(const StepArgs &, const PoolArgs<float> &, const gxt_track *, int,
                                                    int, hipStream_t);

// The host's snapshot of the step loop, written straight into its pinned
// buffers: the per-event live counts packed (ev_live[e * kLiveStride] at
// out_live[e]), the track counts of the pool of generation g and the
// overflow flag (the rest of out_cnt is left as it is).
__global__ void k_snapshot(const Counters *cnt, int g, const int *ev_live, int n, Counters *out_cnt,
                           int *out_live) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) out_live[e] = ev_live[(int64_t)e * kLiveStride];
  if (blockIdx.x == 0 && threadIdx.x < 2 * kStripes) {
    const int k = kPool0 + (threadIdx.x & 1), s = threadIdx.x >> 1;
    out_cnt->ctr(g, k, s) = cnt->ctr(g, k, s);
    if (threadIdx.x == 0) out_cnt->overflow = cnt->overflow;
  }
}

hipError_t launch_snapshot(const Counters *cnt, int g, const int *ev_live, int n, Counters *out_cnt, int *out_live,
                           hipStream_t st) {
  const int grid = n > kBlock ? (n + kBlock - 1) / kBlock : 1;
  hipLaunchKernelGGL(k_snapshot, dim3(grid), dim3(kBlock), 0, st, cnt, g, ev_live, n, out_cnt, out_live);
  return hipGetLastError();
}

hipError_t launch_rng_init(const DevParams *P, const uint32_t *seed, const uint64_t *idx, int n, uint32_t *out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_rng_init, dim3(grid_for(n)), dim3(kBlock), 0, st, P, seed, idx, n, out);
  return hipGetLastError();
}

hipError_t launch_rng_draw(uint32_t *states, int n, int k, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_rng_draw, dim3(grid_for(n)), dim3(kBlock), 0, st, states, n, k, out);
  return hipGetLastError();
}

hipError_t launch_table_value(const DevParams *P, int precision, int slot, int table, const double *x, double *out,
                              int n, hipStream_t st) {
  if (precision == GXT_FP32)
    hipLaunchKernelGGL(k_table_value<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, P, slot, table, x, out, n);
  else
    hipLaunchKernelGGL(k_table_value<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, P, slot, table, x, out, n);
  return hipGetLastError();
}

hipError_t launch_macro_xs(const DevParams *P, int precision, int slot, int species, int proc, const double *e,
                           double *out, int n, hipStream_t st) {
  if (precision == GXT_FP32)
    hipLaunchKernelGGL(k_macro_xs<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, P, slot, species, proc, e, out, n);
  else
    hipLaunchKernelGGL(k_macro_xs<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, P, slot, species, proc, e, out, n);
  return hipGetLastError();
}

hipError_t launch_sample(const DevParams *P, int precision, int slot, int model, const SampleArgs &s, hipStream_t st) {
  if (precision == GXT_FP32)
    hipLaunchKernelGGL(k_sample<float>, dim3(grid_for(s.n)), dim3(kBlock), 0, st, P, slot, model, s);
  else
    hipLaunchKernelGGL(k_sample<double>, dim3(grid_for(s.n)), dim3(kBlock), 0, st, P, slot, model, s);
  return hipGetLastError();
}

hipError_t launch_helix(const double B[3], int precision, const double *charge, const double *mom, const double *step,
                        double *pos, double *dir, int n, hipStream_t st) {
  if (precision == GXT_FP32)
    hipLaunchKernelGGL(k_helix<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, B[0], B[1], B[2], charge, mom, step,
                       pos, dir, n);
  else
    hipLaunchKernelGGL(k_helix<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, B[0], B[1], B[2], charge, mom, step,
                       pos, dir, n);
  return hipGetLastError();
}

hipError_t launch_slab_distance(const DevParams *P, const double *pos, const double *dir, const double *pstep,
                                int32_t *vol, double *snext, int32_t *next, double *safety, int32_t *boundary, int n,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_slab_distance, dim3(grid_for(n)), dim3(kBlock), 0, st, P, pos, dir, pstep, vol, snext, next,
                     safety, boundary, n);
  return hipGetLastError();
}

hipError_t launch_field_value(const DevField &F, const double *pos, double *b, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_field_value, dim3(grid_for(n)), dim3(kBlock), 0, st, F, pos, b, n);
  return hipGetLastError();
}

// LDS form (trees of at most kNavLdsMax bytes): as many workgroups as the
// CUs hold at once, striding over the queries (one copy of the tree each and
// no partial second round: for the step, 6 per CU when resident, 8 per CU
// ran 10 % slower, 12 to 32 per CU tied with 6); otherwise one thread per
// query.
template <class K>
static dim3 nav_grid(K kernel, int n, int lds_bytes) {
  const int g = grid_for(n);
  if (lds_bytes == 0) return dim3(g);
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kBlock, lds_bytes) != hipSuccess || per < 1 ||
      ncu < 1) {
    (void)hipGetLastError();
    per = 4;
    ncu = 256;
  }
  const int r = per * ncu;
  return dim3(g < r ? g : r);
}

hipError_t launch_nav_locate(const DevNav &g, int lds_bytes, const double *pos, const double *dir, int32_t *out,
                             int n, hipStream_t st) {
  if (lds_bytes > 0)
    hipLaunchKernelGGL(k_nav_locate<true>, nav_grid(k_nav_locate<true>, n, lds_bytes), dim3(kBlock), lds_bytes, st, g,
                       lds_bytes, pos, dir, out, n);
  else
    hipLaunchKernelGGL(k_nav_locate<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, g, 0, pos, dir, out, n);
  return hipGetLastError();
}

hipError_t launch_nav_step(const DevNav &g, int lds_bytes, const int32_t *touchable, const double *pos,
                           const double *dir, double *snext, int32_t *next, double *safety, int n, hipStream_t st) {
  if (lds_bytes > 0)
    hipLaunchKernelGGL(k_nav_step<true>, nav_grid(k_nav_step<true>, n, lds_bytes), dim3(kBlock), lds_bytes, st, g,
                       lds_bytes, touchable, pos, dir, snext, next, safety, n);
  else
    hipLaunchKernelGGL(k_nav_step<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, g, 0, touchable, pos, dir, snext, next,
                       safety, n);
  return hipGetLastError();
}

hipError_t launch_nav_find_next_boundary(const DevNav &g, int lds_bytes, const int32_t *touchable, const double *pos,
                                        const double *dir, const double *pstep, double *snext, int32_t *next,
                                        double *safety, int32_t *boundary, int n, hipStream_t st) {
  if (lds_bytes > 0)
    hipLaunchKernelGGL(k_nav_find_next_boundary<true>, nav_grid(k_nav_find_next_boundary<true>, n, lds_bytes),
                       dim3(kBlock), lds_bytes, st, g, lds_bytes, touchable, pos, dir, pstep, snext, next, safety,
                       boundary, n);
  else
    hipLaunchKernelGGL(k_nav_find_next_boundary<false>, dim3(grid_for(n)), dim3(kBlock), 0, st, g, 0, touchable, pos,
                       dir, pstep, snext, next, safety, boundary, n);
  return hipGetLastError();
}

hipError_t launch_field_step(const DevField &F, int precision, const double *charge, const double *mom,
                             const double *step, double *pos, double *dir, double *err, int n, hipStream_t st) {
  if (precision == GXT_FP32)
    hipLaunchKernelGGL(k_field_step<float>, dim3(grid_for(n)), dim3(kBlock), 0, st, F, charge, mom, step, pos, dir,
                       err, n);
  else
    hipLaunchKernelGGL(k_field_step<double>, dim3(grid_for(n)), dim3(kBlock), 0, st, F, charge, mom, step, pos, dir,
                       err, n);
  return hipGetLastError();
}

}  // namespace gxt
