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
//   k_advance   : the next pool's counts become the current ones.
// Tallies go to per-workgroup LDS histograms flushed into per-workgroup rows
// (no global atomics on hot bins); per-event sums use FP64 global atomics.
#include <hip/hip_runtime.h>

#include "gxt_device.h"
#include "gxt_kernels.h"

namespace gxt {

constexpr int kBlock = 256;

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

// Append slot to queue q (q < 0: none) with one atomic per wave and queue.
__device__ __forceinline__ void queue_push(int q, int slot, int *qcnt, int32_t *queues, int64_t cap,
                                           int *overflow) {
  const unsigned long long lt = (lane() == 0) ? 0ULL : ((~0ULL) >> (64 - lane()));
#pragma unroll 1
  for (int k = 0; k < kNumQueues; ++k) {
    const unsigned long long m = __ballot(q == k);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane() == leader) base = atomicAdd(&qcnt[k], (int)__popcll(m));
      base = __shfl(base, leader, 64);
      if (q == k) {
        const int64_t idx = (int64_t)base + __popcll(m & lt);
        if (idx < cap)
          queues[(int64_t)k * cap + idx] = slot;
        else
          atomicOr(overflow, 1);
      }
    }
  }
}

// ------------------------------------------------------------------ tallies
struct LdsTally {
  double z[kMaxZBins];
  double r[kMaxRBins];
  double v[kMaxVol];
  unsigned int spec[3 * kSpecBins];
};

__device__ __forceinline__ void lds_clear(LdsTally &t) {
  for (int i = threadIdx.x; i < kMaxZBins; i += blockDim.x) t.z[i] = 0.0;
  for (int i = threadIdx.x; i < kMaxRBins; i += blockDim.x) t.r[i] = 0.0;
  for (int i = threadIdx.x; i < kMaxVol; i += blockDim.x) t.v[i] = 0.0;
  for (int i = threadIdx.x; i < 3 * kSpecBins; i += blockDim.x) t.spec[i] = 0u;
  __syncthreads();
}

// Add the workgroup's histograms to its own row (no atomics: one owner).
__device__ __forceinline__ void lds_flush(const DevParams *P, LdsTally &t, const TallyRows &rows) {
  __syncthreads();
  const int nz = P->nz, nr = P->nr, nv = P->geo.nvol;
  double *row = rows.bins + (int64_t)blockIdx.x * rows.stride;
  for (int i = threadIdx.x; i < nz; i += blockDim.x)
    if (t.z[i] != 0.0) row[i] += t.z[i];
  for (int i = threadIdx.x; i < nr; i += blockDim.x)
    if (t.r[i] != 0.0) row[nz + i] += t.r[i];
  for (int i = threadIdx.x; i < nv; i += blockDim.x)
    if (t.v[i] != 0.0) row[nz + nr + i] += t.v[i];
  unsigned long long *srow = rows.spec + (int64_t)blockIdx.x * 3 * kSpecBins;
  for (int i = threadIdx.x; i < 3 * kSpecBins; i += blockDim.x)
    if (t.spec[i]) srow[i] += t.spec[i];
}

__device__ __forceinline__ void deposit(const DevParams *P, LdsTally &t, DevEvent *ev, int event, double e,
                                        const double p[3], int vol) {
  if (e <= 0.0) return;
  unsafeAtomicAdd(&ev[event].edep, e);
  if (P->geo.visible[vol]) unsafeAtomicAdd(&ev[event].visible, e);
  unsafeAtomicAdd(&t.v[vol], e);
  int iz = (int)floor(p[2] / P->zbin);
  iz = iz < 0 ? 0 : (iz >= P->nz ? P->nz - 1 : iz);
  unsafeAtomicAdd(&t.z[iz], e);
  const double r = sqrt(p[0] * p[0] + p[1] * p[1]);
  int ir = (int)(r / P->rbin);
  ir = ir < 0 ? 0 : (ir >= P->nr ? P->nr - 1 : ir);
  unsafeAtomicAdd(&t.r[ir], e);
}

__device__ __forceinline__ void spectrum(LdsTally &t, int species, double e) {
  const int b = (int)floor((m_log10(e) + 3.0) * 8.0);
  if (b >= 0 && b < kSpecBins) atomicAdd(&t.spec[species * kSpecBins + b], 1u);
}

// ------------------------------------------------------------------ pool I/O
template <class R>
struct TrackRegs {
  double pos[3];
  R dir[3];
  R e;
  R nil[3];
  Rng rng;
  int event, vol, species, nsteps;
  uint64_t uid;
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
  t.nil[0] = p.nil0[i];
  t.nil[1] = p.nil1[i];
  t.nil[2] = p.nil2[i];
  t.rng.load(p.rng, p.cap, i);
  t.event = p.event[i];
  const int meta = p.meta[i];
  t.vol = meta & 0xffff;
  t.species = meta >> 16;
  t.nsteps = p.nsteps[i];
  t.uid = trace ? p.uid[i] : 0;
}

template <class R>
__device__ __forceinline__ void store_track(const Pool<R> &p, int64_t i, const TrackRegs<R> &t, bool trace) {
  p.x[i] = t.pos[0];
  p.y[i] = t.pos[1];
  p.z[i] = t.pos[2];
  p.dx[i] = t.dir[0];
  p.dy[i] = t.dir[1];
  p.dz[i] = t.dir[2];
  p.e[i] = t.e;
  p.nil0[i] = t.nil[0];
  p.nil1[i] = t.nil[1];
  p.nil2[i] = t.nil[2];
  t.rng.store(p.rng, p.cap, i);
  p.event[i] = t.event;
  p.meta[i] = t.vol | (t.species << 16);
  p.nsteps[i] = t.nsteps;
  if (trace) p.uid[i] = t.uid;
}

// A track created by an interaction or an annihilation at rest.
template <class R>
struct NewTrack {
  int species;
  R e;
  R d[3];
  uint64_t uid;
};

template <class R>
__device__ inline void store_new(const DevParams *P, const Pool<R> &p, int64_t i, const double pos[3], int vol,
                                 int event, const NewTrack<R> &n, bool trace) {
  p.x[i] = pos[0];
  p.y[i] = pos[1];
  p.z[i] = pos[2];
  p.dx[i] = n.d[0];
  p.dy[i] = n.d[1];
  p.dz[i] = n.d[2];
  p.e[i] = n.e;
  p.nil0[i] = R(-1);
  p.nil1[i] = R(-1);
  p.nil2[i] = R(-1);
  double s[6];
  substream_state(P, n.uid, s);
#pragma unroll
  for (int k = 0; k < 6; ++k) p.rng[k * p.cap + i] = (uint32_t)s[k];
  p.event[i] = event;
  p.meta[i] = vol | (n.species << 16);
  p.nsteps[i] = 0;
  if (trace) p.uid[i] = n.uid;
}

// Slot of output k of a lane in the next pool (neutral from the bottom,
// charged from the top); -1 when it would leave the pool.
__device__ __forceinline__ int64_t out_slot(int region, int idx, int64_t cap) {
  if (idx < 0 || (int64_t)idx >= cap) return -1;
  return region == 0 ? (int64_t)idx : cap - 1 - (int64_t)idx;
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

// Positron at rest: two back-to-back 511 keV photons (2 direction draws,
// then 2 substream draws per photon).
template <class R>
__device__ __forceinline__ void at_rest(Rng &rng, NewTrack<R> *out, int &n) {
  const double cost = 2.0 * rng.uniform() - 1.0;
  const double sint = sqrt((1.0 - cost) * (1.0 + cost));
  const double phi = kTwoPi * rng.uniform();
  const R d0 = (R)(sint * cos(phi)), d1 = (R)(sint * sin(phi)), d2 = (R)cost;
  out[n].species = 0;
  out[n].e = (R)kMe;
  out[n].d[0] = d0;
  out[n].d[1] = d1;
  out[n].d[2] = d2;
  out[n].uid = child_substream(rng);
  ++n;
  out[n].species = 0;
  out[n].e = (R)kMe;
  out[n].d[0] = -d0;
  out[n].d[1] = -d1;
  out[n].d[2] = -d2;
  out[n].uid = child_substream(rng);
  ++n;
}

// ------------------------------------------------------------------ step loop kernels
template <class R>
__global__ __launch_bounds__(kBlock) void k_inject(StepArgs a, PoolArgs<R> pa, const gxt_primary *prim, int n) {
  const DevParams *P = a.P;
  const Pool<R> &cur = pa.cur;
  const bool trace = P->trace_events > 0;
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < n;
    TrackRegs<R> t;
    int region = -1;
    if (active) {
      const gxt_primary &p = prim[i];
      t.species = p.species;
      t.event = p.event;
      t.e = (R)p.ekin;
      for (int k = 0; k < 3; ++k) {
        t.pos[k] = p.pos[k];
        t.dir[k] = (R)p.dir[k];
        t.nil[k] = R(-1);
      }
      t.nsteps = 0;
      t.uid = p.substream;
      t.vol = locate(P->geo, t.pos, p.dir[0], p.dir[1], p.dir[2]);
      if (t.vol < 0) {
        unsafeAtomicAdd(&a.ev[t.event].leak, p.ekin);
        atomicAdd(&a.job_live[a.ev[t.event].job], -1);
      } else {
        substream_state(P, t.uid, t.rng.s);
        region = t.species == 0 ? 0 : 1;
      }
    }
    const int b0 = wave_reserve(region == 0 ? 1 : 0, &a.cnt->cur[0]);
    const int b1 = wave_reserve(region == 1 ? 1 : 0, &a.cnt->cur[1]);
    if (region >= 0) {
      const int64_t s = out_slot(region, region == 0 ? b0 : b1, cur.cap);
      if (s >= 0)
        store_track(cur, s, t, trace);
      else
        atomicOr(&a.cnt->overflow, 1);
    }
  }
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_transport(StepArgs a, PoolArgs<R> pa) {
  __shared__ LdsTally lt;
  lds_clear(lt);
  const DevParams *P = a.P;
  const Pool<R> &cur = pa.cur;
  const Pool<R> &nxt = pa.nxt;
  const Nodes<R> N(P);
  const bool trace = P->trace_events > 0;
  // counts are clamped to the pool so a corrupted counter cannot index past it
  const int64_t capc = cur.cap;
  const int nN = (int)min((int64_t)max(a.cnt->cur[0], 0), capc);
  const int nC = (int)min((int64_t)max(a.cnt->cur[1], 0), capc - nN);
  const int total = nN + nC;
  const R mine = (R)P->tracking_cut;
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < total;
    TrackRegs<R> t;
    NewTrack<R> nt[2];
    int nnew = 0;
    int q = -1;
    int slot = 0;
    bool survive = false;
    int live_delta = 0;
    if (active) {
      slot = i < nN ? i : (int)(cur.cap - 1 - (i - nN));
      load_track(cur, slot, t, trace);
      const DevMaterial &m = P->mat[P->geo.slot[t.vol]];
      const int np = nproc(t.species);
      R xs[3] = {R(0), R(0), R(0)};
      double bestd = kHuge;
      int sel = -1;
      // PostStep GPIL (ProxyEmProcess.hpp:122-151)
      for (int p = 0; p < np; ++p) {
        xs[p] = macro_xs(P, N, m, t.species, p, t.e);
        if (t.nil[p] <= R(0)) t.nil[p] = -m_log(t.rng.template u<R>());
        const double sp = xs[p] > R(0) ? (double)(t.nil[p] / xs[p]) : kHuge;
        if (sp < bestd) {
          bestd = sp;
          sel = p;
        }
      }
      const bool charged = t.species != 0;
      if (charged) {
        const double al = (double)along_limit(N, m, t.species == 2, t.e, mine);
        if (al < bestd) {
          bestd = al;
          sel = -1;
        }
      }
      const double pre[3] = {t.pos[0], t.pos[1], t.pos[2]};
      const int prevol = t.vol;
      double step = 0.0;
      const int status = propagate<R>(P, t.species, t.e, t.pos, t.dir, t.vol, bestd, step);
      for (int p = 0; p < np; ++p)
        if (xs[p] > R(0)) t.nil[p] -= (R)step * xs[p];
      t.nsteps++;
      atomicAdd(&a.ev[t.event].nsteps, 1ULL);
      double edep = 0.0;
      if (charged) {
        unsafeAtomicAdd(&a.ev[t.event].clen, step);
        R loss = R(0);
        t.e = along_doit(N, m, t.species == 2, t.e, (R)step, mine, loss);
        const double mid[3] = {0.5 * (pre[0] + t.pos[0]), 0.5 * (pre[1] + t.pos[1]), 0.5 * (pre[2] + t.pos[2])};
        deposit(P, lt, a.ev, t.event, (double)loss, mid, prevol);
        edep += (double)loss;
      }
      int tproc = status == kPhysics ? sel : (status == kBoundary ? -2 : (status == kEscaped ? -3 : -1));
      if (status == kPhysics && sel < 0) tproc = -1;
      if (status == kEscaped) {
        unsafeAtomicAdd(&a.ev[t.event].leak, (double)t.e);
        if (trace) trace_rec(a, t, tproc, edep);
        live_delta = -1;
      } else if (t.e <= R(0)) {
        if (trace) trace_rec(a, t, tproc, edep);
        if (t.species == 2) {
          atomicAdd(&a.ev[t.event].nannihil, 1ULL);
          at_rest(t.rng, nt, nnew);
        }
        live_delta = nnew - 1;
      } else if (status == kPhysics && sel >= 0) {
        t.nil[sel] = R(-1);
        q = queue_of(t.species, sel);
        store_track(cur, slot, t, trace);
        if (trace) cur.tedep[slot] = edep;
      } else {
        if (trace) trace_rec(a, t, tproc, edep);
        if (t.e < mine) {
          deposit(P, lt, a.ev, t.event, (double)t.e, t.pos, t.vol);
          if (t.species == 2) {
            atomicAdd(&a.ev[t.event].nannihil, 1ULL);
            at_rest(t.rng, nt, nnew);
          }
          live_delta = nnew - 1;
        } else if (t.nsteps >= P->max_steps) {
          deposit(P, lt, a.ev, t.event, (double)t.e, t.pos, t.vol);
          atomicAdd(&a.cnt->killed, 1ULL);
          live_delta = -1;
        } else {
          survive = true;
        }
      }
    }
    wave_add_u64(&a.cnt->steps, active);
    // emission: survivors and annihilation photons into the next pool
    const int outN = nnew + ((survive && t.species == 0) ? 1 : 0);
    const int outC = (survive && t.species != 0) ? 1 : 0;
    const int bN = wave_reserve(outN, &a.cnt->nxt[0]);
    const int bC = wave_reserve(outC, &a.cnt->nxt[1]);
    if (survive) {
      const int64_t s = t.species == 0 ? out_slot(0, bN, nxt.cap) : out_slot(1, bC, nxt.cap);
      if (s >= 0)
        store_track(nxt, s, t, trace);
      else
        atomicOr(&a.cnt->overflow, 1);
    }
    for (int k = 0; k < nnew; ++k) {
      const int64_t s = out_slot(0, bN + (survive && t.species == 0 ? 1 : 0) + k, nxt.cap);
      if (s >= 0)
        store_new(P, nxt, s, t.pos, t.vol, t.event, nt[k], trace);
      else
        atomicOr(&a.cnt->overflow, 1);
      atomicAdd(&a.ev[t.event].nsec[0], 1ULL);
      spectrum(lt, 0, (double)nt[k].e);
    }
    queue_push(q, slot, a.cnt->q, a.queues, cur.cap, &a.cnt->overflow);
    if (live_delta != 0) atomicAdd(&a.job_live[a.ev[t.event].job], live_delta);
  }
  lds_flush(P, lt, a.rows);
}

template <class R>
__global__ __launch_bounds__(kBlock) void k_interact(StepArgs a, PoolArgs<R> pa) {
  __shared__ LdsTally lt;
  lds_clear(lt);
  const DevParams *P = a.P;
  const Pool<R> &cur = pa.cur;
  const Pool<R> &nxt = pa.nxt;
  const bool trace = P->trace_events > 0;
  int qc[kNumQueues];
  int total = 0;
#pragma unroll
  for (int k = 0; k < kNumQueues; ++k) {
    qc[k] = (int)min((int64_t)max(a.cnt->q[k], 0), cur.cap);
    total += qc[k];
  }
  const R mine = (R)P->tracking_cut;
  for (int base = blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < total; base += gridDim.x * blockDim.x) {
    const int i = base + lane();
    const bool active = i < total;
    TrackRegs<R> t;
    NewTrack<R> nt[3];
    int nnew = 0;
    bool survive = false;
    int live_delta = 0;
    if (active) {
      int q = 0, j = i;
      while (j >= qc[q]) {
        j -= qc[q];
        ++q;
      }
      const int slot = a.queues[(int64_t)q * cur.cap + j];
      load_track(cur, slot, t, trace);
      const DevMaterial &m = P->mat[P->geo.slot[t.vol]];
      Interaction<R> o;
      sample_model<R>(P, m, queue_model(q), t.e, t.dir, t.rng, o);
      if (q == 1 && o.nsec == 2) atomicAdd(&a.ev[t.event].npairs, 1ULL);
      if (q == 7) atomicAdd(&a.ev[t.event].nannihil, 1ULL);
      deposit(P, lt, a.ev, t.event, (double)o.edep, t.pos, t.vol);
      t.e = o.ekin;
      t.dir[0] = o.d[0];
      t.dir[1] = o.d[1];
      t.dir[2] = o.d[2];
      if (trace) trace_rec(a, t, queue_proc(q), cur.tedep[slot] + (double)o.edep);
      for (int k = 0; k < o.nsec; ++k) {
        const Sec<R> &s = o.sec[k];
        if (s.e < mine) {
          deposit(P, lt, a.ev, t.event, (double)s.e, t.pos, t.vol);
          if (s.species == 2) {
            atomicAdd(&a.ev[t.event].nannihil, 1ULL);
            at_rest(t.rng, nt, nnew);
          }
        } else {
          nt[nnew].species = s.species;
          nt[nnew].e = s.e;
          nt[nnew].d[0] = s.d[0];
          nt[nnew].d[1] = s.d[1];
          nt[nnew].d[2] = s.d[2];
          nt[nnew].uid = child_substream(t.rng);
          ++nnew;
        }
      }
      if (t.e > R(0)) {
        if (t.e < mine) {
          deposit(P, lt, a.ev, t.event, (double)t.e, t.pos, t.vol);
          if (t.species == 2) {
            atomicAdd(&a.ev[t.event].nannihil, 1ULL);
            at_rest(t.rng, nt, nnew);
          }
        } else if (t.nsteps >= P->max_steps) {
          deposit(P, lt, a.ev, t.event, (double)t.e, t.pos, t.vol);
          atomicAdd(&a.cnt->killed, 1ULL);
        } else {
          survive = true;
        }
      }
      live_delta = nnew + (survive ? 1 : 0) - 1;
    }
    wave_add_u64(&a.cnt->interactions, active);
    int outN = 0, outC = 0;
    if (survive) (t.species == 0 ? outN : outC)++;
    for (int k = 0; k < nnew; ++k) (nt[k].species == 0 ? outN : outC)++;
    const int bN = wave_reserve(outN, &a.cnt->nxt[0]);
    const int bC = wave_reserve(outC, &a.cnt->nxt[1]);
    int kN = 0, kC = 0;
    if (survive) {
      const int64_t s = t.species == 0 ? out_slot(0, bN + kN++, nxt.cap) : out_slot(1, bC + kC++, nxt.cap);
      if (s >= 0)
        store_track(nxt, s, t, trace);
      else
        atomicOr(&a.cnt->overflow, 1);
    }
    for (int k = 0; k < nnew; ++k) {
      const int sp = nt[k].species;
      const int64_t s = sp == 0 ? out_slot(0, bN + kN++, nxt.cap) : out_slot(1, bC + kC++, nxt.cap);
      if (s >= 0)
        store_new(P, nxt, s, t.pos, t.vol, t.event, nt[k], trace);
      else
        atomicOr(&a.cnt->overflow, 1);
      atomicAdd(&a.ev[t.event].nsec[sp], 1ULL);
      spectrum(lt, sp, (double)nt[k].e);
    }
    {
      const int made = nnew;
      const int incl = wave_incl_scan(made);
      const int tot = __shfl(incl, 63, 64);
      if (tot > 0 && lane() == 63) atomicAdd(&a.cnt->created, (unsigned long long)tot);
    }
    if (live_delta != 0) atomicAdd(&a.job_live[a.ev[t.event].job], live_delta);
  }
  lds_flush(P, lt, a.rows);
}

__global__ void k_advance(Counters *c, int64_t cap) {
  if (threadIdx.x == 0) {
    const int64_t n0 = c->nxt[0], n1 = c->nxt[1];
    if (c->overflow || n0 < 0 || n1 < 0 || n0 + n1 > cap) {
      // an overflowing pool is dropped whole: the host reports GXT_ECAPACITY
      c->overflow = 1;
      c->cur[0] = 0;
      c->cur[1] = 0;
    } else {
      c->cur[0] = (int32_t)n0;
      c->cur[1] = (int32_t)n1;
    }
    c->nxt[0] = 0;
    c->nxt[1] = 0;
    for (int k = 0; k < kNumQueues; ++k) c->q[k] = 0;
  }
}

// ------------------------------------------------------------------ component kernels
__global__ void k_rng_init(const DevParams *P, const uint32_t *seed, const uint64_t *idx, int n, uint32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t v1[3] = {seed[0], seed[1], seed[2]};
  uint64_t v2[3] = {seed[3], seed[4], seed[5]};
  const uint64_t k = idx[i];
  for (int j = 0; j < 64; ++j)
    if ((k >> j) & 1ULL) {
      matvec_mod(P->jump[0][j], v1, 4294967087ULL);
      matvec_mod(P->jump[1][j], v2, 4294944443ULL);
    }
  out[6 * i + 0] = (uint32_t)v1[0];
  out[6 * i + 1] = (uint32_t)v1[1];
  out[6 * i + 2] = (uint32_t)v1[2];
  out[6 * i + 3] = (uint32_t)v2[0];
  out[6 * i + 4] = (uint32_t)v2[1];
  out[6 * i + 5] = (uint32_t)v2[2];
}

__global__ void k_rng_draw(uint32_t *states, int n, int k, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Rng r;
  for (int j = 0; j < 6; ++j) r.s[j] = (double)states[6 * i + j];
  for (int j = 0; j < k; ++j) out[(int64_t)i * k + j] = r.uniform();
  for (int j = 0; j < 6; ++j) states[6 * i + j] = (uint32_t)r.s[j];
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
  for (int j = 0; j < 6; ++j) r.s[j] = (double)s.rng[6 * i + j];
  const R d[3] = {(R)s.dir[i], (R)s.dir[n + i], (R)s.dir[2 * n + i]};
  Interaction<R> o;
  sample_model<R>(P, P->mat[slot], model, (R)s.ekin[i], d, r, o);
  for (int j = 0; j < 6; ++j) s.rng[6 * i + j] = (uint32_t)r.s[j];
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

// ------------------------------------------------------------------ launchers
static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

template <class R>
hipError_t launch_step(const StepArgs &a, const PoolArgs<R> &pa, int grid, hipStream_t st, hipEvent_t *ev) {
  // ev (optional): 4 events recorded before/after the transport and interaction kernels
  if (ev) (void)hipEventRecord(ev[0], st);
  hipLaunchKernelGGL(k_transport<R>, dim3(grid), dim3(kBlock), 0, st, a, pa);
  if (ev) (void)hipEventRecord(ev[1], st);
  hipLaunchKernelGGL(k_interact<R>, dim3(grid), dim3(kBlock), 0, st, a, pa);
  if (ev) (void)hipEventRecord(ev[2], st);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, a.cnt, pa.cur.cap);
  if (ev) (void)hipEventRecord(ev[3], st);
  return hipGetLastError();
}

template <class R>
hipError_t launch_inject(const StepArgs &a, const PoolArgs<R> &pa, const gxt_primary *prim, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int g = grid_for(n);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_inject<R>, dim3(g), dim3(kBlock), 0, st, a, pa, prim, n);
  return hipGetLastError();
}

template hipError_t launch_step<double>(const StepArgs &, const PoolArgs<double> &, int, hipStream_t, hipEvent_t *);
template hipError_t launch_step<float>(const StepArgs &, const PoolArgs<float> &, int, hipStream_t, hipEvent_t *);
template hipError_t launch_inject<double>(const StepArgs &, const PoolArgs<double> &, const gxt_primary *, int,
                                          hipStream_t);
template hipError_t launch_inject<float>(const StepArgs &, const PoolArgs<float> &, const gxt_primary *, int,
                                         hipStream_t);

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

}  // namespace gxt
