// gxref.h -- CPU oracle of the EM particle-transport step loop.
//
// TEST INFRASTRUCTURE ONLY.  This is a scalar FP64 C++ restatement of the
// reference's step loop and leaf algorithms (GeantExascalePilot,
// source/Geant/{random,proxy,magneticfield,core,track}) under the physics
// contract documented in the repository design notes.  Only tests/, the
// smoke check and bench.py's CPU-baseline leg may load it, and only as the
// checker or as the reported CPU baseline; the product library libgxt.so
// never links, loads or calls it.
//
// The C entry points mirror include/gxt.h with a gxref_ prefix so one test
// harness drives both.
#ifndef GXREF_H
#define GXREF_H

#include "../include/gxt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gxref_ctx gxref_ctx;

int gxref_create(const gxt_config *cfg, gxref_ctx **out);
int gxref_destroy(gxref_ctx *ctx);
const char *gxref_last_error(void);

int gxref_make_primaries(uint32_t seed, int species, double emin, double emax, double spot_mm, int n,
                         uint64_t first, uint64_t n_total, gxt_primary *out);
/* Event front end, same contract as gxt_read_hepevt / gxt_egamma_gun. */
int gxref_read_hepevt(const char *path, int max_events, const double vertex[3], uint64_t first, gxt_primary *out,
                      int capacity, int *n_primaries, int *n_events, int *n_unsupported);
int gxref_egamma_gun(uint32_t seed, int n_events, int per_event, double gamma_fraction, double pmin, double pmax,
                     double eta_min, double eta_max, double phi_min, double phi_max, const double vertex[3],
                     uint64_t first, uint64_t n_total, gxt_primary *out);

/* Transport n primaries of one job (events numbered by gxt_primary.event)
 * with nthreads host threads (events handed out dynamically). */
int gxref_run(gxref_ctx *ctx, const gxt_primary *prim, int n, int nthreads);
/* Continue a basket of tracks in flight (gxt_submit_tracks). */
int gxref_run_tracks(gxref_ctx *ctx, const gxt_track *tracks, int n, int nthreads);
/* Run each event for at most steps_per_event track-steps, then return what
 * is left as a basket (test infrastructure for the hand-off identity:
 * partial + continuation = the whole run).  *n_out is the basket size even
 * when it exceeds capacity (GXT_ECAPACITY). */
int gxref_run_partial(gxref_ctx *ctx, const gxt_primary *prim, int n, int64_t steps_per_event, gxt_track *out,
                      int capacity, int *n_out);
int gxref_get_profiles(gxref_ctx *ctx, double *edep_z, double *edep_r, double *edep_volume);
/* Volume bins of gxref_get_profiles (gxt_volume_bins); cell optional, one per touchable. */
int gxref_volume_bins(gxref_ctx *ctx, int *n_bins, int32_t *cell);
int gxref_get_event_tallies(gxref_ctx *ctx, gxt_event_tally *out, int n);
int gxref_get_spectra(gxref_ctx *ctx, int64_t *counts);
int gxref_get_stats(gxref_ctx *ctx, gxt_stats *out);
int gxref_get_trace(gxref_ctx *ctx, double *out, int64_t max_records, int64_t *n_out);

int gxref_rng_init(const uint32_t seed[6], const uint64_t *substreams, int n, uint32_t *states);
int gxref_rng_draw(uint32_t *states, int n, int k, double *out);
/* Jump matrices A^(2^(76+j)) for j in [0,64) of both components, row-major,
 * out[(c*64 + j)*9 + r*3 + col]. */
int gxref_rng_jump_matrices(uint64_t *out);

int gxref_table_value(gxref_ctx *ctx, int material_slot, int table, const double *x, double *out, int n);
/* Second derivatives of a table (for checks against an independent spline). */
int gxref_table_nodes(gxref_ctx *ctx, int material_slot, int table, double *x, double *y, double *d2,
                      int *n);
int gxref_macro_xs(gxref_ctx *ctx, int material_slot, int species, int process, const double *ekin,
                   double *out, int n);
int gxref_sample(gxref_ctx *ctx, int model, int material_slot, gxt_sample_io *io);
int gxref_helix_step(const double field[3], const double *charge, const double *momentum,
                     const double *step, double *pos, double *dir, int n);
/* Multiple-scattering pieces of the oracle's e-/e+ step (test hooks; the
 * library has them only inside its transport kernel):
 * cos(theta) of n deflections of (0,0,1) after tau transport mean free paths,
 * deflection i drawn from substream i of seed (two draws each);
 * z(t) and t(z(t)) for the geometric-path conversion;
 * u = 2/a of the screening a after tau transport mean free paths. */
int gxref_msc_deflect(double tau, uint32_t seed, int n, double *cos_theta, double *phi);
int gxref_msc_lengths(const double *t, const double *lambda1, int n, double *z, double *t_back);
int gxref_msc_screening(const double *tau, int n, double *u);
/* Field map and field stepping, same contract as gxt_field_map_* / gxt_field_step (FP64). */
typedef struct gxref_field_map gxref_field_map;
int gxref_field_map_load(const char *path, gxref_field_map **out);
int gxref_field_map_free(gxref_field_map *map);
int gxref_field_map_info(const gxref_field_map *map, int *nr, int *nz, double *r0, double *z0, double *dr,
                         double *dz, double *bmax);
int gxref_field_map_value(const gxref_field_map *map, const double *pos, double *b, int n);
int gxref_field_step(const gxref_field_map *map, const double field[3], const double origin[3], int stepper,
                     const double *charge, const double *momentum, const double *step, double *pos, double *dir,
                     double *err, int n);
int gxref_rotate_to_lab(double *u, double *v, double *w, const double *u1, const double *u2,
                        const double *u3, int n);
/* Volume-tree navigation, same contract as gxt_nav_* (no device). */
typedef struct gxref_nav gxref_nav;
int gxref_nav_create(const gxt_box_volume *volumes, int n_volumes, const gxt_placement *placements, int n_placements,
                     int world, gxref_nav **out);
int gxref_nav_free(gxref_nav *nav);
int gxref_nav_info(const gxref_nav *nav, int *n_touchables, int32_t *volume, int32_t *parent, double *centre);
int gxref_nav_locate(const gxref_nav *nav, const double *pos, const double *dir, int32_t *touchable, int n);
int gxref_nav_step(const gxref_nav *nav, const int32_t *touchable, const double *pos, const double *dir,
                   double *snext, int32_t *next, double *safety, int n);
int gxref_nav_find_next_boundary(const gxref_nav *nav, const int32_t *touchable, const double *pos, const double *dir,
                                 const double *pstep, double *snext, int32_t *next, double *safety, int32_t *boundary,
                                 int n);
/* The FP64 log, exp, log1p, expm1, pow, log10 (which = 0..5) the oracle and
 * the device build share. */
int gxref_mathfn(int which, const double *x, const double *y, double *out, int n);
/* The FP64 sine and cosine the oracle and the device build share. */
int gxref_sincos(const double *x, double *s, double *c, int n);
/* Sandia photo-absorption macroscopic cross-section (1/mm). */
int gxref_sandia_xs(gxref_ctx *ctx, int material_slot, const double *ekin, double *out, int n);
/* The pair sampler's unnormalised energy-sharing density at eps[i] for a
 * photon of e0 >= 2 MeV on element Z (test hook). */
int gxref_bh_density(int Z, double e0, const double *eps, double *out, int n);
/* Straight-line navigation of the analytic geometry, same contract as gxt_slab_distance. */
int gxref_slab_distance(gxref_ctx *ctx, const double *pos, const double *dir, const double *pstep, int32_t *vol,
                        double *snext, int32_t *next, double *safety, int32_t *boundary, int n);
/* Along-step pieces of the oracle's e-/e+ step (test hook; the library has them
 * only inside its transport kernel): the continuous-loss step limit at e[i],
 * and the kinetic energy after a step of step[i] with its loss. */
int gxref_along_step(gxref_ctx *ctx, int material_slot, int species, const double *e, const double *step,
                     double *limit, double *fin, double *eloss, int n);

#ifdef __cplusplus
}
#endif
#endif
