// gxt_types.h -- host/device layout of the physics data, geometry, track
// pool and counters of the MI355X step loop.
#pragma once

#include <stdint.h>

#include "../../include/gxt.h"

namespace gxt {

constexpr int kMaxVol = 128;      // volumes (calorimeter: 2 per layer)
constexpr int kMaxElem = 3;       // elements per material
constexpr int kMaxSlots = 2;      // materials per setup
constexpr int kMaxZBins = 64;
constexpr int kMaxRBins = 64;
constexpr int kNumQueues = 9;     // (species, process) interaction queues + positrons at rest
constexpr int kQueueAtRest = 8;
constexpr int kSpecBins = 64;     // secondary spectrum bins per species
constexpr int kInvRangeHints = 256;  // search-hint bins per inverse-range vector
constexpr int kMaxShells = 128;   // shell binding energies over the elements of the slots
// The pool, the interaction queues and the new-track list are split into
// kStripes stripes of cap / kStripes slots, each with its own counters;
// workgroup b appends to stripe b % kStripes, so each counter takes 1/kStripes
// of the atomics (the workgroups of one stripe share an XCD under the
// round-robin dispatch of workgroups to XCDs).  Measured on the 10 GeV e- /
// PbWO4 batch: 16 stripes 2 % faster than 12 and 3 % faster than 32 (whose
// per-stripe prefixes cost more than the atomics they spread), 5 % faster
// than 8 in FP64 and 10 % in FP32.
constexpr int kStripes = 16;
// Words per row of the substream jump tables: a 3x3 matrix padded to 48 bytes.
constexpr int kJumpRow = 12;

// One node of a physics vector: abscissa (energy or range), value, second
// derivative; a spline lookup reads two adjacent nodes, 64 B (32 B in FP32).
struct alignas(32) Node64 {
  double x, y, d2, pad;
};
struct alignas(16) Node32 {
  float x, y, d2, pad;
};

// One physics vector inside the node array (nodes [off, off+n)).
struct DevTable {
  int off;
  int n;
  int loff;  // first node in the LDS copy of its kernel's tables (DevParams::lds_*), -1: none
  int lpad;
  double dl, base;  // log-binning: bin = log(E)/dl - base
  double x0, xn, y0, yn;  // first and last node (the clamps of a lookup)
  float dlf, basef;
  float x0f, xnf, y0f, ynf;
  // search hints of a non-log vector (inverse range): hints[hint + k] is the
  // last node below x[0] 2^(k / hscale), k < nhint (hint < 0: none)
  int hint, nhint;
  float hlog0, hscale;  // log2 x[0]; hint bins per octave
};

struct DevElement {
  int Z;
  int sb;          // offset of the element's Seltzer-Berger grid (x[32], y[57], v[57][32])
  int sandia;      // offset of its Sandia rows (5 doubles each)
  int nint;        // number of Sandia intervals
  int shell;       // offset of its shell binding energies (MeV)
  int nshell;
  double natoms;   // atoms per mm3
  double zoa;      // Sandia Z/A ratio
  double wbrem;    // n Z (Z+1): element weight for pair and brems
};

struct DevMaterial {
  double ne;          // electrons per mm3
  double migdal;      // k_p^2 = migdal * E_tot^2
  double cut_gamma;   // gamma production threshold (MeV)
  double cut_e;       // electron production threshold (MeV)
  double cut_length;  // production cut (mm)
  double compt_split; // Lambda.compt upper edge: above it LambdaPrim.compt / E
  double phot_emin;   // LambdaPrim.phot lower edge: below it Sandia
  int nel;
  int pad;
  DevElement el[kMaxElem];
  DevTable tab[GXT_NTAB];
  // Sandia photo-absorption by merged interval: the sx_n breakpoints
  // sx[sx_b .. sx_b + sx_n) (keV, ascending) are every element's interval
  // edges but its first; interval j (sx[sx_b + j - 1] <= E < sx[sx_b + j])
  // holds each element's coefficients a1..a4 at sx[sx_c + (j * nel + i) * 4].
  // A search starts at hints[sx_hint + k], the interval of the lower end of
  // log2 bin k (sx_hscale bins per octave from sx_hlog0).
  int sx_b, sx_c, sx_n, sx_hint, sx_nhint;
  float sx_hlog0, sx_hscale;
};

struct NavNode;
struct NavGrid;

// A volume tree on the device.
struct DevNav {
  const NavNode *node;
  const NavGrid *grid;   // per volume
  const int32_t *idx;
  int ntouch;
  int pad;
};

// The transport geometry.  Slab form (tree = 0): nvol volumes between the
// planes z[v] and z[v + 1], |x| <= hx, |y| <= hy; a track's volume is v.
// Tree form (tree = 1): a track's volume is a touchable of nav, and
// cell[touchable] its tally cell; cells [0, nvol - 1) are the touchables
// with a material, in touchable order, and cell nvol - 1 is vacuum (slot -1).
// slot[] and visible[] are per cell (in the slab form cells are volumes).
struct DevGeometry {
  int nvol;
  int tree;
  double hx, hy;
  double z[kMaxVol + 1];
  int slot[kMaxVol];
  int visible[kMaxVol];
  const int32_t *cell;
  DevNav nav;
};

// Magnetic field seen by the general (non-uniform or Runge-Kutta)
// propagation: a cylindrical (r, z) map, or the uniform field b.
struct DevField {
  const double *v;         // (Br, Bphi, Bz) in tesla at (ir * nz + iz) * 3; null: uniform b
  int nr, nz;
  double r0, z0, dr, dz;   // first node and spacing (mm)
  double origin[3];        // setup origin in the map frame (mm)
  double b[3];             // uniform field (tesla) when v is null
  int stepper;             // GXT_STEPPER_*
  int general;             // e-/e+ take the general propagation (a map, or RK45)
};

// One placed box of a volume tree (a touchable), in global coordinates: boxes
// are placed by translations only.  Touchables are numbered breadth-first, so
// the daughters of a touchable are [first, first + count).
struct alignas(64) NavNode {
  double c[3];     // centre (mm)
  double h[3];     // half-lengths (mm)
  int32_t parent;  // -1 for the world
  int32_t first;
  int32_t count;
  int32_t volume;  // index into the caller's volume array
};

// Daughter index of a volume with many daughters: nb bins of its extent
// along axis ax (mother frame, bin = (x + h) inv); the daughters that may hold
// a point of bin b are ordinals idx[idx[off + b] .. idx[off + b + 1]), in
// ascending order.  nb = 0: no index (the daughters are scanned).
struct NavGrid {
  double inv;
  int32_t ax, nb, off, pad;
};

// Bin of mother-frame coordinate x in a daughter index (clamped; NaN -> 0).
// The same function places the daughters' extents when the index is built:
// it is monotone in x, so a point inside a daughter's extent falls in one of
// the daughter's bins.
__host__ __device__ inline int nav_bin(double x, double h, double inv, int nb) {
  const double u = (x + h) * inv;
  if (!(u >= 0.0)) return 0;
  if (u >= (double)nb) return nb - 1;
  return (int)u;
}

// Most points per call of the entry points that take [3n] structure-of-arrays
// (their kernels index them with 32-bit k n + i).
constexpr int kMaxSoaPoints = 1 << 29;

// Largest tree (nodes, grids and index) the navigation kernels copy into LDS.
constexpr int kNavLdsMax = 48 * 1024;

// Everything a kernel needs besides the pool: lives in one device buffer.
struct DevParams {
  DevMaterial mat[kMaxSlots];
  DevGeometry geo;
  double B[3];
  double bmag;
  DevField field;
  double tracking_cut;
  double zbin, rbin;
  int nz, nr;
  int max_steps;
  int trace_events;   // events [0, trace_events) are traced
  int trace_steps;
  int msc;            // e-/e+ multiple scattering
  uint32_t seed[6];
  int nsandia, nshells;     // doubles in the Sandia and shell arrays
  const uint32_t *jump8;    // A_c^(d 2^(76+8k)) mod m_c at ((c * 8 + k) * 256 + d) * kJumpRow
  const Node64 *nd;             // physics-vector nodes, FP64
  const Node32 *nf;             // FP32 mirror
  const double *sbd;            // Seltzer-Berger grids, FP64
  const float *sbf;             // FP32 mirror
  const double *sandia;         // Sandia rows (edge keV, a1..a4)
  const double *sx;             // merged Sandia breakpoints and coefficients (DevMaterial::sx_*)
  const int *hints;             // search hints of the inverse-range vectors
  // the Seltzer-Berger grid every element shares: reduced photon energy k / T
  // (32 nodes) and ln(T / MeV) (57 nodes), FP64 and FP32; the step kernels
  // read them from their LDS copy of DevParams
  double sbx[32], sby[57];
  float sbxf[32], sbyf[57];
  // shell binding energies (MeV) of every element of the slots, innermost
  // first (DevElement::shell, nshell), read where DevParams is
  double shell_e[kMaxShells];
  // LDS copies of the physics vectors a transport kernel reads: layout 0 the
  // photon tables (GXT_TAB_LAMBDA_COMPT .. GXT_TAB_LAMBDAPRIM_PHOT), layout 1
  // the e-/e+ tables (GXT_TAB_LAMBDA_BREM_EM .. GXT_TAB_INVRANGE_EP), one
  // contiguous node span per material slot: nodes [lds_src[l][s], +
  // lds_cnt[l][s]) of the node array, in slot order
  int lds_src[2][kMaxSlots];
  int lds_cnt[2][kMaxSlots];
  int lds_nodes[2];
};

// Structure-of-arrays track pool.  R is the precision of the physics state
// (double, or float in the precision form); positions stay FP64.
template <class R>
struct Pool {
  int64_t cap;
  double *x, *y, *z;
  R *dx, *dy, *dz;
  R *e;
  R *nil0, *nil1, *nil2;
  uint32_t *rng;      // 6 planes of cap words
  int32_t *event;     // event slot
  int32_t *meta;      // volume | species << 16
  int32_t *nsteps;
  double *aed, *avis, *alen;  // per-track tallies (deposit, visible deposit, charged length),
                              // added to the event once, when the track dies
  uint64_t *uid;      // substream index (written only while tracing)
  double *tedep;      // step deposit carried to the interaction kernel (tracing only)
};

// Device counters of the step loop, in three generations.  Generation g
// holds, per stripe, the photon and e-/e+ track counts of a pool (kPool0,
// kPool1) and the queue and new-track-list counts of the iteration that
// fills that pool.  Iteration i reads the pool of generation g = i mod 3,
// fills generation g + 1 (mod 3) and clears generation g + 2, which the
// next iteration fills: no kernel between iterations moves counts from one
// set to another, so each chain of the loop waits only for the other
// chain's last kernel.  The counters of kind k of stripe s in generation g
// are addressed ctr(g, k, s); each stripe's counters sit on cache lines no
// other stripe uses, so the returning atomics of one stripe's workgroups
// never wait behind another stripe's.  Statistics are per stripe as well
// (summed by the host).
enum CounterKind { kPool0 = 0, kPool1 = 1, kQ0 = 2, kNew0 = 2 + kNumQueues, kNew1 = 3 + kNumQueues,
                   kNumKinds = 4 + kNumQueues };
constexpr int kGens = 3;
constexpr int kKindStride = 32;      // ints between the counters of one stripe
constexpr int kStripeStride = 512;  // ints between stripes
__host__ __device__ constexpr int gen_next(int g) { return g == kGens - 1 ? 0 : g + 1; }
enum StatKind { kStatSteps = 0, kStatStepsCharged, kStatKilled, kStatInteractions, kStatCreated, kStatLoaded0,
                kStatLoaded1, kStatInteractions0, kStatCreated0, kNumStats };
struct alignas(256) Counters {
  int32_t c[kGens * kStripes * kStripeStride];
  unsigned long long stat[kStripes][16];  // 128 B per stripe
  int32_t overflow;        // a pool or queue write would have gone out of range
  int32_t pad[31];
  unsigned long long trace_n; // trace records written
  __host__ __device__ int32_t &ctr(int g, int kind, int s) {
    return c[(g * kStripes + s) * kStripeStride + kind * kKindStride];
  }
  __host__ __device__ int32_t ctr(int g, int kind, int s) const {
    return c[(g * kStripes + s) * kStripeStride + kind * kKindStride];
  }
  __host__ __device__ unsigned long long total(int k) const {
    unsigned long long t = 0;
    for (int s = 0; s < kStripes; ++s) t += stat[s][k];
    return t;
  }
};

// Live-track count of event e at ev_live[e * kLiveStride]: one 128-byte line
// per event, so the +/- 1 atomics of one event's tracks never queue behind
// another event's.
constexpr int kLiveStride = 32;

// Per-event tallies on the device (FP64, atomics), one or more whole 128-byte
// lines per event for the same reason.
struct alignas(128) DevEvent {
  double edep, visible, leak, clen;
  unsigned long long nsec[3];
  unsigned long long nsteps, npairs, nannihil;
  int pad0, pad1;
};

}  // namespace gxt
