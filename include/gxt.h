/*
 * gxt.h -- C ABI of the MI355X-native EM particle-transport step loop (libgxt.so).
 *
 * Every entry point returns int: 0 on success, a negative GXT_E* code on
 * failure (the message is available from gxt_last_error / gxt_ctx_error).
 * No exception crosses this boundary.  Host buffers are owned by the caller;
 * all device memory (track pool, RNG states, tables, tallies) is owned by the
 * library.  A gxt_ctx is driven by one host thread and owns one HIP stream;
 * gxt_table_value, gxt_macro_xs and gxt_sample use a context's physics data,
 * gxt_rng_* and gxt_helix_step need no context (current HIP device).
 *
 * Units are CLHEP: mm, MeV, ns; magnetic field in tesla.
 * Species codes: 0 = gamma, 1 = electron, 2 = positron.
 * Entry points that take n points as [3n] structure-of-arrays (x[0..n),
 * y[0..n), z[0..n)) accept 1 <= n <= 2^29 per call.
 */
#ifndef GXT_H
#define GXT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GXT_OK 0
#define GXT_EINVAL (-1)    /* bad argument                              */
#define GXT_EHIP (-2)      /* HIP runtime error                         */
#define GXT_EDATA (-3)     /* physics data file missing or malformed    */
#define GXT_ECAPACITY (-4) /* track pool or job table full              */
#define GXT_ESTATE (-5)    /* call not valid in the current state       */
#define GXT_ENOTFOUND (-6) /* unknown job id                            */

/* ---- configuration ----------------------------------------------------- */
/* A volume of a box geometry (GXT_GEOM_TREE, and the navigator entry points
 * below) and the placement of one volume inside another. */
typedef struct gxt_box_volume {
  double half[3];      /* half-lengths (mm)                         */
  int32_t material;    /* GXT_MAT_* or -1 (vacuum)                  */
  int32_t pad;
} gxt_box_volume;
typedef struct gxt_placement {
  int32_t mother;      /* index into the volume array               */
  int32_t daughter;    /* index into the volume array               */
  double translation[3]; /* daughter centre in the mother frame (mm) */
} gxt_placement;

#define GXT_GEOM_BOX 0  /* homogeneous block: |x|,|y| <= xy/2, 0 <= z <= depth */
#define GXT_GEOM_CALO 1 /* z-stacked layers of absorber + gap                  */
#define GXT_GEOM_TREE 2 /* boxes placed inside each other (gxt_config::tree_*)  */

#define GXT_MAT_PB 0    /* G4_Pb 11.35 g/cm3 (reference tables, couple 1)    */
#define GXT_MAT_LAR 1   /* G4_lAr (reference tables, couple 2, density-scaled) */
#define GXT_MAT_PBWO4 2 /* PbWO4 8.28 g/cm3 (generated tables)                */

#define GXT_FP64 0
#define GXT_FP32 1 /* precision form: FP32 tables/sampling/helix, FP64 position+tallies */

/* Charged-track steppers.  HELIX: a helix in the field at the start of each
 * propagation sub-step (the reference's ConstFieldHelixStepper driven by
 * FieldLookup, FieldPropagationHandler.cpp:101-111,247-253; exact in a
 * uniform field).  RK45: Cash-Karp embedded Runge-Kutta 4(5) on the Lorentz
 * equation, the field evaluated at every stage, sub-steps of at most
 * GXT_RK_ANGLE radians of turning (the integrator the reference reaches
 * through GUFieldPropagator::DoStep under ENABLE_MORE_COMPLEX_FIELD,
 * FieldPropagationHandler.cpp:201-208, and checks the helix against in
 * magneticfield/test/testConstFieldStepper.cpp). */
#define GXT_STEPPER_HELIX 0
#define GXT_STEPPER_RK45 1
#define GXT_RK_ANGLE 0.25

typedef struct gxt_config {
  int32_t geometry;         /* GXT_GEOM_*                                           */
  int32_t material;         /* box material (GXT_MAT_*)                             */
  double box_xy;            /* full transverse size of the block or layers (mm)     */
  double box_depth;         /* box depth along z (mm)                               */
  int32_t n_layers;         /* calorimeter layers                                   */
  int32_t absorber;         /* calorimeter absorber material (GXT_MAT_*)            */
  int32_t gap;              /* calorimeter gap material (GXT_MAT_*)                 */
  int32_t pad0;
  double absorber_mm;       /* absorber thickness per layer (mm)                    */
  double gap_mm;            /* gap thickness per layer (mm)                         */
  double lar_density;       /* g/cm3 used for lAr (tables scaled from 1.396)        */
  double field[3];          /* uniform magnetic field (tesla)                       */
  double tracking_cut;      /* MeV; tracks below it deposit locally                 */
  uint32_t seed;            /* MRG32k3a seed, replicated over the 6 state words     */
  int32_t precision;        /* GXT_FP64 / GXT_FP32                                  */
  double zbin_mm;           /* longitudinal tally bin (mm)                          */
  double rbin_mm;           /* radial tally bin (mm)                                */
  int32_t n_zbins;
  int32_t n_rbins;
  int64_t pool_capacity;    /* live-track pool size (tracks)                        */
  int32_t max_steps;        /* per-track step limit (track killed beyond it)        */
  int32_t max_events;       /* per-event tally slots                                */
  int32_t max_jobs;         /* concurrent job slots                                 */
  int32_t trace_tracks;     /* >0: record the first trace_steps steps of tracks     */
  int32_t trace_steps;
  int32_t profile_kernels;  /* 1: time every step-loop launch with HIP events; 2: only the
                               photon interaction and e-/e+ transport launches (the two
                               largest kernels).  The events take each kernel's start and
                               end from its own dispatch (no packets between kernels). */
  const char *data_dir;     /* directory holding the physics data files             */
  const char *field_map;    /* NULL: the uniform field above; else a cylindrical (r, z)
                               field-map file (see gxt_field_map_load), which replaces it */
  double field_origin[3];   /* position of the setup origin in the map frame (mm)   */
  int32_t field_stepper;    /* GXT_STEPPER_* for e-/e+ (a uniform field with HELIX
                               takes the exact constant-field path)                  */
  int32_t msc;              /* 1: multiple scattering of e-/e+ (see GXT_TAB_MSC_EM);
                               0: none, as in the reference's physics list           */
  /* GXT_GEOM_TREE: the volumes and placements of a box geometry, as
   * gxt_nav_create takes them (the reference's calorimeter is built this way,
   * proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:225-325).  A
   * volume's material is the absorber, the gap or -1 (vacuum: no physics, a
   * track flies to the next boundary; its interaction lengths wait for the
   * next material).  Deposits in gap volumes count as visible.  Per-volume
   * tallies: one bin per touchable with a material, in touchable order (at
   * most 127), then one bin for vacuum.  Uniform field only (no field_map,
   * HELIX stepper); multiple scattering allowed.  Arrays are copied by
   * gxt_create. */
  const gxt_box_volume *tree_volumes;
  const gxt_placement *tree_placements;
  int32_t n_tree_volumes;
  int32_t n_tree_placements;
  int32_t tree_world;       /* index of the world volume                           */
  int32_t pad1;
} gxt_config;

/* One primary particle. */
typedef struct gxt_primary {
  int32_t species;      /* 0 gamma, 1 e-, 2 e+                       */
  int32_t event;        /* event index inside its job                 */
  double ekin;          /* kinetic energy (MeV)                       */
  double pos[3];        /* mm                                         */
  double dir[3];        /* unit vector                                */
  uint64_t substream;   /* MRG32k3a substream index of this primary   */
} gxt_primary;

/* Per-event tallies (FP64). */
typedef struct gxt_event_tally {
  double edep;          /* total energy deposited in the setup (MeV)  */
  double visible;       /* deposited in the gap (calorimeter) or the block (box) */
  double leak;          /* kinetic energy leaving the setup           */
  double charged_length;/* summed e-/e+ step length in material (mm)  */
  int64_t n_secondaries[3]; /* secondaries created per species        */
  int64_t n_steps;      /* track-steps of this event                  */
  int64_t n_pairs;      /* gamma conversions                          */
  int64_t n_annihil;    /* positron annihilations (flight + rest)     */
} gxt_event_tally;

/* Run statistics. */
typedef struct gxt_stats {
  int64_t track_steps;      /* total track-steps                       */
  int64_t iterations;       /* step-loop iterations                    */
  int64_t max_live;         /* peak live tracks                        */
  int64_t killed_by_limit;  /* tracks killed by max_steps              */
  double wall_ms;           /* device time of the last gxt_run          */
} gxt_stats;

/* Kernel timings of the step loop, accumulated while profile_kernels is set. */
typedef struct gxt_kernel_stats {
  double transport_neutral_ms; /* summed durations of the photon transport launches    */
  double transport_charged_ms; /* summed durations of the e-/e+ transport launches     */
  double interact_neutral_ms;  /* summed durations of the photon interaction launches  */
  double interact_charged_ms;  /* summed durations of the e-/e+ interaction launches   */
  double derive_ms;       /* summed durations of the new-track RNG derivation       */
  double advance_ms;      /* 0: the step loop has no counter-advance launch (its
                             counters rotate over generations); kept for layout */
  int64_t launches;       /* step-loop iterations timed (one launch of each kernel) */
  int64_t track_steps;    /* tracks stepped (one per live track per iteration)     */
  int64_t neutral_steps;  /* of which photon steps                                 */
  int64_t charged_steps;  /* of which e-/e+ steps                                  */
  int64_t interactions;   /* tracks sampled by the interaction kernel              */
  int64_t created;        /* tracks created by the interaction kernel              */
  int32_t record_bytes;   /* bytes of one track record in the pool                 */
  int32_t pad;
  int64_t neutral_loaded; /* track records read by the photon transport launches (a
                             track takes up to several steps per launch)          */
  int64_t charged_loaded; /* track records read by the e-/e+ transport launches   */
  int64_t neutral_interactions; /* of interactions, photons sampled               */
  int64_t neutral_created;      /* of created, tracks the photon interactions made */
} gxt_kernel_stats;

/* Completed job, as returned by gxt_poll. */
typedef struct gxt_job_result {
  uint64_t job_id;
  int32_t n_events;
  int32_t pad;
  double edep;            /* summed over the job's events (MeV)      */
  double visible;
  double submit_ms;       /* host clock, ms since context creation  */
  double admit_ms;
  double done_ms;
  int64_t track_steps;
  int64_t iterations;     /* step-loop iterations from the job's admission to the
                             counter snapshot that found its last event finished */
} gxt_job_result;

typedef struct gxt_ctx gxt_ctx;

/* ---- context ----------------------------------------------------------- */
/* Replaces ProxyDeviceManager<T>::ProxyDeviceManager + Initialize
 * (proxy/ProxyDeviceManager.hpp:33-41, proxy/src/ProxyDeviceManager.cpp:41-69):
 * loads the physics tables (ProxyDataManager.cu:27-122), uploads them and
 * allocates the track pool. */
int gxt_create(const gxt_config *cfg, gxt_ctx **out);
int gxt_destroy(gxt_ctx *ctx);
const char *gxt_last_error(void);
const char *gxt_ctx_error(gxt_ctx *ctx);

/* Fill a gxt_config with the defaults (tracking cut 1 keV, 1 cm bins, ...). */
int gxt_default_config(gxt_config *cfg);

/* ---- primary source ---------------------------------------------------- */
/* Pencil beam / front-face spot gun (replaces ProxyEgammaGun::GenerateOneEvent,
 * proxy/src/ProxyEgammaGun.cpp:43-85).  Primary i gets substream first+i;
 * when emax > emin or spot_mm > 0 its energy and (x,y) come from three draws
 * of substream n_total + first + i of the same seed, in that order. */
int gxt_make_primaries(uint32_t seed, int species, double emin, double emax, double spot_mm,
                       int n, uint64_t first, uint64_t n_total, gxt_primary *out);

/* HEPEVT event file -> primaries (replaces ProxyHepEvtGenerator::GenerateOneEvent,
 * proxy/src/ProxyHepEvtGenerator.cpp:48-132, which reads one event per call).
 * Reads up to max_events events (< 0: all).  Each event is NHEP followed by NHEP
 * lines "ISTHEP IDHEP JDAHEP1 JDAHEP2 PX PY PZ MASS" (GeV).  The primaries of
 * an event are its entries without daughters (JDAHEP1 = 0) that are either a
 * daughter of an entry with daughters (the reference's rule) or stable at top
 * level (ISTHEP = 1; the reference drops these, Geant4's G4HEPEvtInterface
 * keeps them).  Ekin = p^2 / (sqrt(p^2 + m^2) + m) with the mass as given (the
 * reference squares it first, ProxyHepEvtGenerator.cpp:109).  Gammas, e- and
 * e+ (PDG 22, 11, -11) become primaries at vertex (mm) with substreams first,
 * first+1, ...; other species and zero momenta are counted in n_unsupported.
 * With out == NULL only the counts are returned.  A truncated event or a
 * daughter range outside its event is GXT_EDATA. */
int gxt_read_hepevt(const char *path, int max_events, const double vertex[3], uint64_t first,
                    gxt_primary *out, int capacity, int *n_primaries, int *n_events,
                    int *n_unsupported);

/* Egamma gun (ProxyEgammaGun::GenerateOneEvent, proxy/src/ProxyEgammaGun.cpp:43-85):
 * n_events x per_event particles, gamma with probability gamma_fraction else
 * e-, momentum uniform in [pmin, pmax] (MeV), pseudorapidity uniform in
 * [eta_min, eta_max], azimuth uniform in [phi_min, phi_max], from vertex (mm).
 * Particle i gets substream first+i; its four draws (type, momentum, eta,
 * phi, in the reference's order) come from substream n_total + first + i. */
int gxt_egamma_gun(uint32_t seed, int n_events, int per_event, double gamma_fraction, double pmin,
                   double pmax, double eta_min, double eta_max, double phi_min, double phi_max,
                   const double vertex[3], uint64_t first, uint64_t n_total, gxt_primary *out);

/* ---- batch transport --------------------------------------------------- */
/* Queue a job of n primaries (replaces ProxyDeviceManager::UploadTrackState,
 * ProxyDeviceManager.cpp:152-204).  The primaries are copied to device memory
 * at once, stream-ordered on the context's stream (prim may be reused after
 * the call); the job is admitted into the shared pool when its estimated
 * footprint fits. */
int gxt_submit(gxt_ctx *ctx, const gxt_primary *prim, int n, uint64_t job_id);
/* One track in flight, in the form the reference's CPU transport would hand a
 * basket over to the device at its "memory transfer to GPU" point of DoStep<>
 * (proxy/test/test_BasicCpuTransport.cpp:527-531): the parts of TrackState
 * (track/TrackState.hpp:177-232) that decide how the track continues, plus
 * its own RNG state (per-track substreams; the reference keeps the RNG in
 * each process object). */
typedef struct gxt_track {
  int32_t species;      /* 0 gamma, 1 e-, 2 e+                                      */
  int32_t event;        /* event index inside its job                               */
  double ekin;          /* kinetic energy (MeV), fPhysicsState.fEkin                */
  double pos[3];        /* mm, fPos                                                 */
  double dir[3];        /* unit vector, fDir                                        */
  double nil[3];        /* interaction lengths left to the next discrete interaction of
                           each process of the species (gamma: Compton, conversion,
                           photoelectric; e-: brems, ionisation; e+: brems, ionisation,
                           annihilation), fPhysicsProcessState[p].fNumOfInteractLengthLeft
                           (TrackState.hpp:126-135); <= 0: sampled afresh at the next step,
                           as after a DoIt (ProxyEmProcess.hpp:122-163)      */
  uint32_t rng[6];      /* the track's MRG32k3a state: words 0-2 < m1, 3-5 < m2,
                           neither triple all zero                                  */
  int32_t volume;       /* volume index of the analytic geometry (fGeometryState), or -1:
                           located from pos and dir like a primary                  */
  int32_t nsteps;       /* steps taken so far (fHistoryState.fNsteps); the step limit
                           max_steps counts them; the event's step tally counts only the
                           steps taken here                                         */
  uint64_t substream;   /* the track's substream index (its identity in step traces) */
} gxt_track;

/* Queue a job of n tracks in flight (a basket of a mid-shower CPU transport)
 * over n_events events: every track and all its descendants are transported
 * to the end; the job's event tallies are those of this continuation (the
 * deposits, steps and secondaries before the hand-off stay with the caller).
 * A track continues exactly as it would have in the loop: with its own nil,
 * RNG state, volume and step count.  Every track is checked (species, event
 * in [0, n_events), finite ekin > 0, finite position, unit direction, finite
 * nil, a valid RNG state, 0 <= nsteps < max_steps, a volume of the geometry
 * that holds the position) before anything is queued.  No reference
 * counterpart beyond the hand-off point cited above. */
int gxt_submit_tracks(gxt_ctx *ctx, const gxt_track *tracks, int n, int n_events, uint64_t job_id);
/* Run the step loop until every submitted job has finished (replaces
 * ProxyDeviceManager::DoStep/DoGPIL, ProxyDeviceManager.cpp:206-276, and the
 * host wave loop DoStep<>, proxy/test/test_BasicCpuTransport.cpp:517-550).
 * max_iterations <= 0 means no limit. */
int gxt_run(gxt_ctx *ctx, int64_t max_iterations);
/* Completed jobs since the last call (up to max). */
int gxt_poll(gxt_ctx *ctx, gxt_job_result *out, int max, int *n_out);
/* Tallies (replaces ProxyDeviceManager::DownloadTrackData,
 * ProxyDeviceManager.cpp:206-259).  Arrays sized n_zbins, n_rbins and the
 * number of volume bins (gxt_volume_bins). */
int gxt_get_profiles(gxt_ctx *ctx, double *edep_z, double *edep_r, double *edep_volume);
/* Volume bins of gxt_get_profiles: 1 for the box, 2*n_layers for the
 * calorimeter, the tree form's cells (touchables with a material, then
 * vacuum).  cell (optional, one entry per touchable): each touchable's bin.
 * No counterpart in the reference, whose scoring is a placeholder
 * (proxy/test/test_BasicCpuTransport.cpp:479-483). */
int gxt_volume_bins(gxt_ctx *ctx, int *n_bins, int32_t *cell);
int gxt_get_event_tallies(gxt_ctx *ctx, uint64_t job_id, gxt_event_tally *out, int n);
/* Energy histogram of created secondaries: 3 species x 64 bins of log10(E/MeV)
 * over [-3, 5). */
int gxt_get_spectra(gxt_ctx *ctx, int64_t *counts);
int gxt_get_stats(gxt_ctx *ctx, gxt_stats *out);
int gxt_get_kernel_stats(gxt_ctx *ctx, gxt_kernel_stats *out);
int gxt_reset_tallies(gxt_ctx *ctx);
/* Step trace of traced tracks: records of 12 doubles
 * {track uid (bits), step, species, process, ekin, x, y, z, dx, dy, dz, edep}.
 * process: -1 transport only, -2 boundary, -3 escaped, otherwise the discrete
 * process index of the species. */
int gxt_get_trace(gxt_ctx *ctx, double *out, int64_t max_records, int64_t *n_out);

/* ---- serving form ------------------------------------------------------ */
/* Start/stop the serving loop: a host thread that keeps the device step loop
 * running, admits queued jobs as pool space frees up, and completes jobs when
 * their last track dies.  gxt_submit stays non-blocking while serving.  An
 * error of the serving loop (e.g. GXT_ECAPACITY) stops it; gxt_poll and
 * gxt_serve_stop then return its code, with the message in gxt_ctx_error and,
 * on the calling thread, gxt_last_error. */
int gxt_serve_start(gxt_ctx *ctx);
int gxt_serve_stop(gxt_ctx *ctx);
/* Host-clock milliseconds since the context was created. */
double gxt_now_ms(gxt_ctx *ctx);

/* ---- component entry points (batched, SoA, host buffers) --------------- */
/* MRG32k3a substream states: out[6*i..6*i+5] = state of substream idx[i] of the
 * stream seeded with seed[0..5] (replaces MRG32k3a::SetNextSubstream /
 * Initialize, random/src/MRG32k3a.cpp:61-67,73-105). */
int gxt_rng_init(const uint32_t seed[6], const uint64_t *substreams, int n, uint32_t *states);
/* k uniforms from each of n states; out[i*k + j]; states advanced in place
 * (replaces MRG32k3a::Kernel, random/src/MRG32k3a.cpp:123-163). */
int gxt_rng_draw(uint32_t *states, int n, int k, double *out);

/* Physics-table lookup on the device (replaces ProxyPhysicsVector::Value and
 * InvRangeValue, proxy/src/ProxyPhysicsVector.cu:469-524, reached through
 * ProxyDataManager::GetTable()->Value(couple, E)).  table: GXT_TAB_*. */
#define GXT_TAB_LAMBDA_COMPT 0
#define GXT_TAB_LAMBDAPRIM_COMPT 1
#define GXT_TAB_LAMBDA_CONV 2
#define GXT_TAB_LAMBDAPRIM_PHOT 3
#define GXT_TAB_LAMBDA_BREM_EM 4
#define GXT_TAB_LAMBDA_IONI_EM 5
#define GXT_TAB_DEDX_EM 6
#define GXT_TAB_RANGE_EM 7
#define GXT_TAB_INVRANGE_EM 8
#define GXT_TAB_LAMBDA_BREM_EP 9
#define GXT_TAB_LAMBDA_IONI_EP 10
#define GXT_TAB_DEDX_EP 11
#define GXT_TAB_RANGE_EP 12
#define GXT_TAB_INVRANGE_EP 13
/* multiple scattering: E^2 / lambda1 (MeV^2/mm), lambda1 the first transport
 * mean free path (LambdaMod1.msc.e-/e+, ProxyPhysicsTableIndex.hpp:39-40);
 * present for Pb and lAr, absent (zero nodes) for PbWO4 */
#define GXT_TAB_MSC_EM 14
#define GXT_TAB_MSC_EP 15
#define GXT_NTAB 16
int gxt_table_value(gxt_ctx *ctx, int material_slot, int table, int precision, const double *x,
                    double *out, int n);
/* Macroscopic cross-section (1/mm) of discrete process p of a species at the
 * given kinetic energies, as the step loop's PostStep GPIL uses it
 * (replaces ProxyEmProcess::GetLambda, proxy/ProxyEmProcess.hpp:89-91). */
int gxt_macro_xs(gxt_ctx *ctx, int material_slot, int species, int process, int precision,
                 const double *ekin, double *out, int n);

/* Final-state samplers (replace ProxyEmModel<T>::SampleSecondaries,
 * proxy/ProxyEmModel.hpp:57-82, for T = KleinNishina, BetheHeitler,
 * SauterGavrila, SeltzerBerger, MollerScattering, and add Bhabha and e+
 * annihilation).  Inputs per track: kinetic energy, direction (3 SoA arrays),
 * RNG state (6 words, advanced in place).  Outputs per track: primary's new
 * kinetic energy and direction (ekin 0 = killed), local deposit, and up to
 * two secondaries (species, ekin, direction); nsec[i] = number produced.
 * model: GXT_MODEL_*. */
#define GXT_MODEL_KLEIN_NISHINA 0
#define GXT_MODEL_BETHE_HEITLER 1
#define GXT_MODEL_SAUTER_GAVRILA 2
#define GXT_MODEL_SELTZER_BERGER_EM 3
#define GXT_MODEL_MOLLER 4
#define GXT_MODEL_SELTZER_BERGER_EP 5
#define GXT_MODEL_BHABHA 6
#define GXT_MODEL_ANNIHIL 7
typedef struct gxt_sample_io {
  int n;
  const double *ekin;        /* [n]      */
  const double *dir;         /* [3n] SoA: dx[n], dy[n], dz[n] */
  uint32_t *rng;             /* [6n] AoS: state of track i at rng[6i] */
  double *out_ekin;          /* [n]      */
  double *out_dir;           /* [3n] SoA */
  double *out_edep;          /* [n]      */
  int32_t *nsec;             /* [n]      */
  int32_t *sec_species;      /* [2n]: secondary j of track i at [2i+j] */
  double *sec_ekin;          /* [2n]     */
  double *sec_dir;           /* [6n]: 3 components of secondary j of track i at [6i+3j] */
} gxt_sample_io;
int gxt_sample(gxt_ctx *ctx, int model, int material_slot, int precision, gxt_sample_io *io);
/* gxt_sample that allocates no device memory: the inputs and outputs are
 * staged through one fixed device array of the library (one per GPU, held
 * until the process ends), for up to GXT_SAMPLE_NOALLOC_MAX tracks per call
 * (a larger n is GXT_EINVAL).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (serving loops and harnesses
 * that forbid hipMalloc).  Every call on a GPU shares that array, so calls
 * that overlap on one GPU race. */
#define GXT_SAMPLE_NOALLOC_MAX 65536
int gxt_sample_noalloc(gxt_ctx *ctx, int model, int material_slot, int precision, gxt_sample_io *io);

/* Constant-field helix step (replaces ConstFieldHelixStepper::DoStep /
 * DoStepArr, magneticfield/ConstFieldHelixStepper.hpp:57-86,164-302, in
 * mm/MeV/tesla instead of cm/GeV/kilogauss).  pos, dir: [3n] SoA, updated in place. */
int gxt_helix_step(const double field[3], int precision, const double *charge,
                   const double *momentum, const double *step, double *pos, double *dir, int n);

/* Straight-line navigation of the transport loop's analytic geometry (the
 * context's block, or its z-stacked layers): NavigationInterface::
 * FindNextBoundary (geometry/NavigationInterface.hpp:64-98) and the volume
 * bookkeeping of IsSameLocation (magneticfield/src/FieldPropagationHandler.cpp:
 * 362-397) for the slab/box navigator the step loop runs.  For n points
 * (pos, dir: [3n] SoA host arrays, dir unit vectors):
 *   vol[i]      in: the volume holding the point (a track's current volume),
 *                   or < 0 to locate it (a point on a plane belongs to the
 *                   volume its direction enters); out: the volume used, -1
 *                   when the point is outside the setup;
 *   snext[i]    distance along dir to the volume's boundary;
 *   next[i]     volume entered there (-1: the track leaves the setup);
 *   safety[i]   distance to the volume's boundary in any direction;
 *   boundary[i] (may be NULL) 1 when a straight step of pstep[i] reaches the
 *               boundary (snext <= pstep; pstep NULL: always 1), the loop's
 *               choice between a physics step and a boundary crossing.
 * Outside the setup: snext = safety = 0, next = -1, boundary = 1.  Volumes are
 * numbered along z from 0 (the calorimeter: absorber 2l, gap 2l+1 of layer
 * l).  GXT_EINVAL for non-finite or non-unit input, vol[i] >= the number of
 * volumes, or a negative or non-finite pstep. */
int gxt_slab_distance(gxt_ctx *ctx, const double *pos, const double *dir, const double *pstep, int32_t *vol,
                      double *snext, int32_t *next, double *safety, int32_t *boundary, int n);

/* ---- non-uniform field ------------------------------------------------- */
/* Cylindrical (r, z) field map on a regular grid, resident in device memory
 * of the current HIP device.  The file has one node per line,
 *   "R(mm)= r phi(deg)= 0 Z(mm)= z Bz(tesla)= bz Br(tesla)= br Bphi(tesla)= bphi",
 * r-major (all z of the first r, then the next r), r and z ascending with
 * constant spacing (the CMS 2015 map: 181 x 161 nodes, r <= 9 m, |z| <= 16 m).
 * Replaces ScalarCMSmagField::ReadVectorData
 * (examples/magneticfield/simplifiedCMS/inc/Geant/ScalarCMSmagField.h:174-206),
 * which assumes the 181 x 161 CMS grid; this reader takes the grid from the
 * file and rejects irregular or truncated files with GXT_EDATA. */
typedef struct gxt_field_map gxt_field_map;
int gxt_field_map_load(const char *path, gxt_field_map **out);
int gxt_field_map_free(gxt_field_map *map);
/* Grid: nr x nz nodes from (r0, z0) with spacing (dr, dz) in mm; bmax: the
 * largest |B| of the nodes (tesla). */
int gxt_field_map_info(const gxt_field_map *map, int *nr, int *nz, double *r0, double *z0, double *dr,
                       double *dz, double *bmax);
/* gxt_field_map_value and gxt_field_step need the map's device to be the
 * current HIP device (GXT_EINVAL otherwise). */
/* Cartesian field (tesla) at n points of the map frame: bilinear in (r, z)
 * inside the grid cell, r and z clamped to the grid, cylindrical components
 * turned with the point's azimuth (on the axis: Bx = Br, By = Bphi).
 * Replaces ScalarCMSmagField::GetFieldValue / GetFieldValueRZ
 * (ScalarCMSmagField.h:208-331).  pos, b: [3n] SoA. */
int gxt_field_map_value(const gxt_field_map *map, const double *pos, double *b, int n);
/* Advances n charged tracks by step[i] (mm) with the stepper the transport
 * uses for one propagation sub-step: HELIX, a helix in the field at the start
 * point; RK45, ceil(step k |B(start)| / GXT_RK_ANGLE) equal Cash-Karp steps
 * (k = 0.299792458 / momentum), direction renormalised after each.  The field
 * is the map at pos + origin (map != NULL) or the uniform field[] (map ==
 * NULL).  pos, dir: [3n] SoA, updated in place.  err (NULL or [2n]): summed
 * Cash-Karp error estimates, position (mm) at [i] and direction at [n+i]
 * (0 for HELIX).  Steps of more than 2^16 RK45 sub-steps are rejected. */
int gxt_field_step(const gxt_field_map *map, const double field[3], const double origin[3], int stepper,
                   int precision, const double *charge, const double *momentum, const double *step,
                   double *pos, double *dir, double *err, int n);

/* ---- volume-tree navigation ---------------------------------------------- */
/* A geometry of box volumes placed inside each other by translations, as the
 * reference builds its calorimeter with VecGeom (UnplacedBox, LogicalVolume,
 * PlaceDaughter; proxy/test/BasicCpuTransport/ProxyDetectorConstruction.cpp:
 * 225-325), navigated on the device.  Replaces, for box geometries, the
 * navigation calls of geometry/NavigationInterface.hpp:42-98 (FindNextBoundary,
 * IsSameLocation) and VecGeom's locate.  A volume's daughters must lie inside
 * it and must not overlap each other (checked when the navigator is built).
 * Touchables (placed instances) are numbered breadth-first from the world (0).
 * The transport loop uses the same navigator for GXT_GEOM_TREE, and the
 * analytic slab navigator for the block and the z-stacked calorimeter. */
/* gxt_box_volume and gxt_placement: see the configuration section. */
typedef struct gxt_navigator gxt_navigator;
/* world: the volume placed at the origin; at most 65535 touchables, 32
 * levels.  GXT_EINVAL for half-lengths below 1e-6 mm, a daughter outside its
 * mother, daughters that overlap by more than 1e-9 mm, a placement cycle or
 * out-of-range indices. */
int gxt_nav_create(const gxt_box_volume *volumes, int n_volumes, const gxt_placement *placements, int n_placements,
                   int world, gxt_navigator **out);
int gxt_nav_free(gxt_navigator *nav);
/* Touchables, and for touchable t: its volume, parent touchable and global
 * centre (3 doubles); any output may be NULL. */
int gxt_nav_info(const gxt_navigator *nav, int *n_touchables, int32_t *volume, int32_t *parent, double *centre);
/* Queries take n points (pos, dir: [3n] SoA, dir unit vectors).  Arrays in
 * host memory are checked (finite, unit, touchables in range: GXT_EINVAL),
 * copied and the call returns with the results.  Arrays that are all device
 * memory of the navigator's device are used in place: the kernel is launched
 * on the null stream and the call returns without waiting; a touchable out of
 * range then gives next = -2 and NaN distances.  Mixed arrays: GXT_EINVAL.
 *
 * Locate: the deepest touchable containing each point (-1 outside the world).
 * A point within 1e-9 mm of a face belongs to the box only if its direction
 * enters through that face. */
int gxt_nav_locate(const gxt_navigator *nav, const double *pos, const double *dir, int32_t *touchable, int n);
/* Step, for points inside touchable[i]: the distance along dir to the next
 * boundary (leaving the touchable or entering a daughter; a daughter the ray
 * passes through by no more than 1e-9 mm is not entered), the touchable
 * holding the boundary point with that direction (-1: outside the world), and
 * the safety (distance to the nearest boundary in any direction). */
int gxt_nav_step(const gxt_navigator *nav, const int32_t *touchable, const double *pos, const double *dir,
                 double *snext, int32_t *next, double *safety, int n);
/* NavigationInterface::FindNextBoundary (geometry/NavigationInterface.hpp:
 * 64-98) for n tracks with proposed physics steps pstep[i] >= 0: with d the
 * boundary distance of gxt_nav_step limited to pstep,
 *   safety > pstep:  snext = pstep, boundary = 0;
 *   otherwise:       boundary = (d < pstep), snext = max(2e-9, d + 2e-9)
 *                    (the push past the surface);
 * next is the touchable beyond the boundary (boundary = 1) or touchable[i];
 * safety is the isotropic safety, >= 0 (the reference leaves the track's
 * previous safety in place in the first case).  Same array rules as
 * gxt_nav_step; a negative or non-finite pstep is GXT_EINVAL for host
 * arrays and gives next = -2, NaN distances for device arrays. */
int gxt_nav_find_next_boundary(const gxt_navigator *nav, const int32_t *touchable, const double *pos,
                               const double *dir, const double *pstep, double *snext, int32_t *next, double *safety,
                               int32_t *boundary, int n);

#ifdef __cplusplus
}
#endif

#endif /* GXT_H */
