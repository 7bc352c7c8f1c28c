// gxt_kernels.h -- kernel argument blocks and host-side launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "gxt_types.h"

namespace gxt {

// Per-workgroup tally rows: bins[row * stride + (z | nz + r | nz + nr + vol)]
// and spec[row * 3 * kSpecBins + species * kSpecBins + bin].
// Per-workgroup histogram rows: workgroup b adds into row b % nrows
// (z bins, r bins, volumes; spectra: 3 species x kSpecBins).
struct TallyRows {
  double *bins;
  unsigned long long *spec;
  int64_t stride;
  int nrows;
  int pad;
};

struct StepArgs {
  const DevParams *P;
  Counters *cnt;
  DevEvent *ev;
  int *ev_live;      // live tracks of event slot e at ev_live[e * kLiveStride]
  int32_t *queues;  // kNumQueues x cap
  int32_t *newlist; // next-pool slots of the tracks created this iteration: 2 lists of cap
                    // (photon interactions, e-/e+ interactions)
  TallyRows rows;
  double *trace;
  int64_t trace_cap;
  int transport_mode;  // e-/e+ transport: 0 plain, 1 with multiple scattering, 2 field map with the
                       // local helix, 3 RK45 stepper
  int inner_steps;     // steps a track may take in one transport launch while it neither
                       // interacts nor ends (>= 1)
  int lds_bytes[2];    // > 0: the photon (0) and e-/e+ (1) transports stage their tables
                       // (DevParams::lds_*) in that much dynamic LDS; 0: they read them in place
  int gen;             // counter generation of the current pool (Counters)
};

template <class R>
struct PoolArgs {
  Pool<R> cur;
  Pool<R> nxt;
};

struct SampleArgs {
  int n;
  const double *ekin;
  const double *dir;
  uint32_t *rng;
  double *out_ekin;
  double *out_dir;
  double *out_edep;
  int32_t *nsec;
  int32_t *sec_species;
  double *sec_ekin;
  double *sec_dir;
};

// Streams of the step loop: main carries the photon chain, side the e-/e+
// chain.  After its chain of an iteration each stream records its event
// (main: fork, side: join) and waits for the other's, so an iteration's
// chains both follow the previous iteration's, and the two waits run side
// by side.  launch_side_sync makes side follow all work queued on main so
// far (injections, snapshots) through the event entry.
struct StepStreams {
  hipStream_t main, side;
  hipEvent_t fork, join, entry;
};

// One step-loop iteration of kStepKernels kernels: photon transport,
// interactions and RNG derivation of their new tracks (main), then the same
// for e-/e+ (side).  With ev (kStepEvents events, null entries not
// recorded) kernel j takes its start and end times into ev[2j], ev[2j + 1].
constexpr int kStepKernels = 6;
constexpr int kStepEvents = 2 * kStepKernels;
template <class R>
hipError_t launch_step(const StepArgs &a, const PoolArgs<R> &pa, int grid, const StepStreams &s,
                       hipEvent_t *ev = nullptr);
hipError_t launch_side_sync(const StepStreams &s);
// Injects n primaries or tracks in flight (In = gxt_primary, gxt_track) from
// device memory; event slots are prim[i].event + event_base.
template <class R, class In>
hipError_t launch_inject(const StepArgs &a, const PoolArgs<R> &pa, const In *prim, int n, int event_base,
                         hipStream_t st);

// The host's snapshot, into pinned host memory (device addresses out_cnt,
// out_live): out_live[e] = ev_live[e * kLiveStride] for e < n, the pool
// counts of counter generation g and the overflow flag into out_cnt.
hipError_t launch_snapshot(const Counters *cnt, int g, const int *ev_live, int n, Counters *out_cnt, int *out_live,
                           hipStream_t st);
hipError_t launch_rng_init(const DevParams *P, const uint32_t *seed, const uint64_t *idx, int n, uint32_t *out,
                           hipStream_t st);
hipError_t launch_rng_draw(uint32_t *states, int n, int k, double *out, hipStream_t st);
hipError_t launch_table_value(const DevParams *P, int precision, int slot, int table, const double *x, double *out,
                              int n, hipStream_t st);
hipError_t launch_macro_xs(const DevParams *P, int precision, int slot, int species, int proc, const double *e,
                           double *out, int n, hipStream_t st);
hipError_t launch_sample(const DevParams *P, int precision, int slot, int model, const SampleArgs &s, hipStream_t st);
hipError_t launch_helix(const double B[3], int precision, const double *charge, const double *mom, const double *step,
                        double *pos, double *dir, int n, hipStream_t st);
// Straight-line navigation of the analytic geometry (gxt_slab_distance);
// pstep and boundary may be null.
hipError_t launch_slab_distance(const DevParams *P, const double *pos, const double *dir, const double *pstep,
                                int32_t *vol, double *snext, int32_t *next, double *safety, int32_t *boundary, int n,
                                hipStream_t st);
hipError_t launch_field_value(const DevField &F, const double *pos, double *b, int n, hipStream_t st);
hipError_t launch_field_step(const DevField &F, int precision, const double *charge, const double *mom,
                             const double *step, double *pos, double *dir, double *err, int n, hipStream_t st);

// Volume-tree queries over n points (pos, dir: [3n] SoA) in the tree g (the
// step kernels check touchable[i] against g.ntouch).  lds_bytes > 0: the
// tree is one allocation of lds_bytes (a multiple of 16, at most 48 KiB)
// starting at g.node, which each workgroup copies into LDS; 0: it is read in
// place.
hipError_t launch_nav_locate(const DevNav &g, int lds_bytes, const double *pos, const double *dir, int32_t *out,
                             int n, hipStream_t st);
hipError_t launch_nav_step(const DevNav &g, int lds_bytes, const int32_t *touchable, const double *pos,
                           const double *dir, double *snext, int32_t *next, double *safety, int n, hipStream_t st);
hipError_t launch_nav_find_next_boundary(const DevNav &g, int lds_bytes, const int32_t *touchable, const double *pos,
                                        const double *dir, const double *pstep, double *snext, int32_t *next,
                                        double *safety, int32_t *boundary, int n, hipStream_t st);

}  // namespace gxt
