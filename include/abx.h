/*
 * abx.h - C ABI of the MI355X multigroup transport path.
 *
 * The boundary is the reference's Transporter::transport() plus the mesh /
 * scalar tally readback (include/simulation/transporter.hpp:39-87,
 * include/simulation/tallies.hpp:38-186).  A caller describes the problem once
 * in the flat form below (what parse_input_file() builds in the reference,
 * src/parser.cpp:77-137), then hands banks of particles to abx_transport()
 * and reads tallies back between batches / generations.
 *
 * Conventions
 *  - plain C structs, caller-owned host buffers, no torch types;
 *  - every entry point returns ABX_OK (0) or a negative ABX_E* code; the
 *    reference's fatal_error() sites (src/error.cpp:36-45) become codes and
 *    abx_last_error() gives the message;
 *  - one host thread per abx_problem; calls are synchronous on the problem's
 *    own HIP stream;
 *  - indices, not ids: tokens, fills and universe references are 0-based
 *    indices into the arrays of the description (ids are kept for output).
 */
#ifndef ABX_H
#define ABX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ABX_OK 0
#define ABX_E_INVALID (-1)     /* bad description or argument */
#define ABX_E_HIP (-2)         /* HIP runtime failure */
#define ABX_E_LOST (-3)        /* particle lost (fatal_error in reference) */
#define ABX_E_MAJORANT (-4)    /* Sigma_t exceeded the majorant */
#define ABX_E_CAPACITY (-5)    /* output bank / secondary stack too small */
#define ABX_E_NO_DEVICE (-6)   /* no gfx950 device */
#define ABX_E_BUDGET (-7)      /* serving: request does not fit HBM budget */

/* ---- geometry (include/geometry/) ------------------------------------ */

/* surface types: src/{x,y,z}plane.cpp, src/plane.cpp, src/{x,y,z}cylinder.cpp,
 * src/sphere.cpp, src/cylinder.cpp */
enum abx_surface_type {
  ABX_XPLANE = 0, /* p = {x0}            */
  ABX_YPLANE = 1, /* p = {y0}            */
  ABX_ZPLANE = 2, /* p = {z0}            */
  ABX_PLANE = 3,  /* p = {A, B, C, D}    */
  ABX_XCYL = 4,   /* p = {y0, z0, R}     */
  ABX_YCYL = 5,   /* p = {x0, z0, R}     */
  ABX_ZCYL = 6,   /* p = {x0, y0, R}     */
  ABX_SPHERE = 7, /* p = {x0, y0, z0, R} */
  ABX_CYL = 8     /* p = {x0, y0, z0, u0, v0, w0, R}: Cylinder, src/cylinder.cpp:30-57.
                     The reference's surface is the quadric
                     a x^2 + b y^2 + c z^2 = R^2 with a = 1 - u^2, b = 1 - v^2,
                     c = 1 - w^2 of the normalised axis (u, v, w): a circular
                     cylinder when the axis is a coordinate axis (stored as the
                     x/y/z-cylinder it equals bit for bit), the same quadric
                     evaluated as the reference does for any other axis. */
};

/* boundary types: include/geometry/surfaces/surface.hpp:34 (same order) */
enum abx_boundary { ABX_VACUUM = 0, ABX_REFLECTIVE = 1, ABX_NORMAL = 2 };

typedef struct abx_surface {
  int32_t type;     /* abx_surface_type */
  int32_t boundary; /* abx_boundary */
  uint32_t id;
  int32_t pad_;
  double p[8];
} abx_surface;

/* Cell region tokens in reverse Polish notation (src/cell.cpp:248-294):
 * +-(surface index + 1) for half-spaces, or one of the operators below
 * (include/geometry/cell.hpp:42-48). */
#define ABX_OP_UNIN (2147483647 - 4)
#define ABX_OP_INTR (2147483647 - 3)
#define ABX_OP_COMP (2147483647 - 2)

enum abx_fill { ABX_FILL_MATERIAL = 0, ABX_FILL_UNIVERSE = 1 };

typedef struct abx_cell {
  uint32_t id;
  int32_t fill_type; /* abx_fill */
  int32_t fill;      /* material index or universe index */
  int32_t rpn_offset; /* into abx_problem_desc.rpn */
  int32_t rpn_len;
} abx_cell;

enum abx_universe_type { ABX_CELL_UNIVERSE = 0, ABX_RECT_LATTICE = 1, ABX_HEX_LATTICE = 2 };

/* hexagonal lattice orientation (include/geometry/hex_lattice.hpp:37, same
 * order; the deck's `top: pointy | flat`, pointy by default) */
enum abx_hex_top { ABX_HEX_POINTY = 0, ABX_HEX_FLAT = 1 };

typedef struct abx_universe {
  uint32_t id;
  int32_t type;         /* abx_universe_type */
  /* cell universe (src/cell_universe.cpp) */
  int32_t cells_offset; /* into abx_problem_desc.universe_cells */
  int32_t n_cells;
  /* rectilinear lattice (src/rect_lattice.cpp:33-52); `origin` is the
   * lattice centre as in the deck, elements are x-major, y fastest:
   * index = nz*Nx*Ny + nx*Ny + ny (src/rect_lattice.cpp:300).
   * hexagonal lattice (src/hex_lattice.cpp:30-56): shape = {rings, Nz, -},
   * pitch = {flat-to-flat pitch, axial pitch, -}, origin = centre of the
   * middle tile; elements are the deck's list of rings*(3 rings - 3) + 1
   * tiles per axial layer, layer by layer, each layer in rows of axial r
   * ascending and q ascending within a row, tiles outside the rings
   * skipped (src/hex_lattice.cpp:219-243). */
  int32_t shape[3];
  int32_t elems_offset; /* into abx_problem_desc.lattice_elems; -1 = dummy */
  int32_t outer;        /* outer universe index or -1 */
  int32_t hex_top;      /* abx_hex_top (hexagonal lattices) */
  double pitch[3];
  double origin[3];
} abx_universe;

/* ---- multigroup materials (src/mg_nuclide.cpp:579-922) ----------------- */
/* One nuclide per material at 1 atom/b-cm (src/material.cpp:51-63).  All
 * arrays point at the raw deck values; derived tables are built by the
 * library exactly as MGNuclide's constructor does. */
typedef struct abx_material {
  uint32_t id;
  int32_t n_delayed;          /* delayed groups (0 = none) */
  const double* total;        /* [G] */
  const double* absorption;   /* [G] */
  const double* scatter;      /* [G][G] row = incoming group */
  const double* fission;      /* [G] or NULL */
  const double* nu_prompt;    /* [G] or NULL ("nu" in the deck) */
  const double* nu_delayed;   /* [G] or NULL */
  const double* chi;          /* [G][G] (a single row may be replicated) or NULL */
  const double* delayed_probabilities; /* [n_delayed] */
  const double* delayed_constants;     /* [n_delayed] */
  /* optional piecewise-linear scattering cosine tables per (g_in, g_out)
   * (MGAngleDistribution, include/materials/mg_angle_distribution.hpp);
   * NULL = isotropic everywhere.  angle_offsets has G*G+1 entries into the
   * mu/pdf/cdf arrays. */
  const int32_t* angle_offsets;
  const double* angle_mu;
  const double* angle_pdf;
  const double* angle_cdf;
} abx_material;

/* ---- sources (src/source.cpp:44-90) ------------------------------------ */
enum abx_spatial { ABX_SRC_POINT = 0, ABX_SRC_BOX = 1 };
enum abx_direction { ABX_DIR_ISOTROPIC = 0, ABX_DIR_MONO = 1 };

typedef struct abx_source {
  int32_t spatial;      /* abx_spatial */
  int32_t direction;    /* abx_direction */
  int32_t fissile_only;
  int32_t pad_;
  double low[3];        /* point position, or box low corner */
  double high[3];       /* box high corner */
  double dir[3];        /* mono-directional direction */
  double energy;        /* mono-energetic energy (MeV) */
  double weight;        /* source weight (picks among sources) */
} abx_source;

/* ---- mesh tallies (src/mesh_tally.cpp, src/collision_mesh_tally.cpp,
 *      src/track_length_mesh_tally.cpp) ---------------------------------- */
enum abx_estimator { ABX_EST_COLLISION = 0, ABX_EST_TRACK_LENGTH = 1 };
enum abx_quantity {
  ABX_Q_FLUX = 0,
  ABX_Q_ELASTIC = 1,
  ABX_Q_ABSORPTION = 2,
  ABX_Q_FISSION = 3,
  ABX_Q_TOTAL = 4,
  ABX_Q_REAL_FLUX = 6,
  ABX_Q_IMG_FLUX = 7
};

typedef struct abx_mesh_tally {
  int32_t estimator; /* abx_estimator */
  int32_t quantity;  /* abx_quantity */
  int32_t shape[3];  /* Nx, Ny, Nz */
  int32_t n_ebounds; /* energy bin edges (>= 2) */
  double low[3];
  double high[3];
  const double* ebounds;
} abx_mesh_tally;

/* ---- settings (src/settings.cpp:34-114) -------------------------------- */
enum abx_mode {
  ABX_K_EIGENVALUE = 0,
  ABX_FIXED_SOURCE = 1,
  ABX_MODIFIED_FIXED_SOURCE = 2
};
enum abx_tracking {
  ABX_SURFACE_TRACKING = 0,
  ABX_DELTA_TRACKING = 1,
  ABX_CARTER_TRACKING = 2
};

typedef struct abx_settings {
  int32_t mode;        /* abx_mode */
  int32_t tracking;    /* abx_tracking */
  uint64_t seed;       /* default 19073486328125 */
  uint64_t stride;     /* default 152917 */
  double wgt_cutoff;   /* default 0.25 */
  double wgt_survival; /* default 1.0 */
  double wgt_split;    /* default 2.0 */
  int64_t nparticles;  /* per batch / generation; tallies divide by it */
  int32_t ngroups;
  int32_t precision;   /* abx_precision */
  const double* energy_bounds;   /* [ngroups + 1] */
  const double* sample_xs_ratio; /* [ngroups], carter tracking only */
} abx_settings;

/* plain: FP64 throughout.  split: cross sections, directions and flight
 * distances in FP32; positions, surface tests and tallies FP64. */
enum abx_precision { ABX_FP64 = 0, ABX_SPLIT = 1 };

/* ---- approximate mesh cancellation (src/approximate_mesh_cancelator.cpp)
 * Applied to each generation's banked particles (k-eigenvalue: before
 * weight normalisation, src/power_iterator.cpp:354-356; modified fixed
 * source: before re-banking). */
typedef struct abx_cancel_mesh {
  double low[3];
  double high[3];
  int32_t shape[3];
  int32_t n_ebounds; /* 0 = no energy binning */
  const double* ebounds;
} abx_cancel_mesh;

/* ExactMGCancelator (src/exact_mg_cancelator.cpp) and BasicExactMGCancelator
 * (src/basic_exact_mg_cancelator.cpp): regional cancellation of each
 * k-eigenvalue generation's fission sites (src/power_iterator.cpp:354-356,
 * 750-777) on a mesh of cancellation regions, split by material, with the
 * transition density from each site's parent's previous collision
 * (delta or carter tracking only, as src/cancelator.cpp:43-73 requires).
 * The uniform particles they create draw from the run's global pcg32
 * (settings::rng, seeded with settings.seed when a run starts). */
enum abx_exact_kind { ABX_EXACT_MG = 1, ABX_BASIC_EXACT_MG = 2 };
/* BasicExactMGCancelator::BetaMode (include/simulation/basic_exact_mg_cancelator.hpp:38) */
enum abx_beta_mode {
  ABX_BETA_ZERO = 0,
  ABX_BETA_MINIMUM = 1,
  ABX_BETA_AVERAGE_F = 2,
  ABX_BETA_AVERAGE_G = 3
};

typedef struct abx_exact_cancel {
  int32_t kind;       /* abx_exact_kind */
  int32_t shape[3];   /* regions in x, y, z */
  double low[3];
  double high[3];
  int32_t n_samples;  /* points per region for <f> and <1/f> ("n-samples", 10) */
  int32_t chi_matrix; /* settings::chi_matrix: a fission spectrum given as a matrix */
  int32_t use_virtual_collisions; /* settings::use_virtual_collisions (exact) */
  int32_t beta;       /* abx_beta_mode (basic) */
  int32_t sobol;      /* basic: points from the Sobol sequence (1) or the rng (0) */
  int32_t n_group_bins;             /* exact: energy regions ("group-bins") */
  const int32_t* group_bin_offsets; /* [n_group_bins + 1] into group_bins */
  const int32_t* group_bins;        /* the groups of each energy region */
} abx_exact_cancel;

/* A fission site's parent information for exact cancellation
 * (BankedParticle::parents_previous_* and Esmp_parent,
 * include/simulation/particle.hpp:51-57), kept beside the ordered bank. */
typedef struct abx_parent_info {
  double r1[3];         /* parents_previous_position */
  double u1[3];         /* parents_previous_direction */
  double E1;            /* parents_previous_previous_energy */
  double E3;            /* parents_previous_energy */
  double Esmp;          /* Esmp_parent */
  int32_t prev_virtual; /* parents_previous_was_virtual */
  int32_t pad_;
} abx_parent_info;

typedef struct abx_problem_desc {
  int32_t n_surfaces;
  const abx_surface* surfaces;
  int32_t n_cells;
  const abx_cell* cells;
  int32_t n_rpn;
  const int32_t* rpn;
  int32_t n_universes;
  const abx_universe* universes;
  int32_t n_universe_cells;
  const int32_t* universe_cells;
  int32_t n_lattice_elems;
  const int32_t* lattice_elems; /* universe index or -1 */
  int32_t root_universe;
  int32_t n_materials;
  const abx_material* materials;
  int32_t n_sources;
  const abx_source* sources;
  int32_t n_tallies;
  const abx_mesh_tally* tallies;
  abx_settings settings;
  const abx_cancel_mesh* cancel; /* NULL = no cancellation */
  const abx_exact_cancel* exact_cancel; /* NULL = none (not with `cancel`) */
} abx_problem_desc;

/* ---- particles at the boundary ----------------------------------------- */

/* A history handed to transport(): the reference's Particle right after
 * source sampling or re-banking (include/simulation/particle.hpp:68-242).
 * rng_state is the CURRENT pcg32 state (source sampling has advanced it). */
typedef struct abx_particle {
  double r[3];
  double u[3];
  double E;
  double wgt;
  double wgt2;
  uint64_t history_id;
  uint64_t family_id;
  uint64_t rng_state;
} abx_particle;

/* A fission daughter: the reference's BankedParticle
 * (include/simulation/particle.hpp:38-66) without the exact-cancellation
 * bookkeeping fields, in parent-bank order then push order. */
typedef struct abx_banked {
  double r[3];
  double u[3];
  double E;
  double wgt;
  double wgt2;
  uint64_t parent_history_id;
  uint64_t parent_daughter_id;
  uint64_t family_id;
} abx_banked;

/* Per-call scalar scores, as sums (divide by nparticles for the per-
 * generation values, src/tallies.cpp:169-174) - ThreadLocalScores in
 * include/simulation/transporter.hpp:52-59. */
typedef struct abx_scores {
  double k_col, k_abs, k_trk, k_tot, leakage, mig;
} abx_scores;

typedef struct abx_diag {
  uint64_t histories;          /* source histories completed */
  uint64_t collisions;         /* real collisions */
  uint64_t virtual_collisions; /* delta / carter rejections */
  uint64_t flights;            /* flight events (all trackers) */
  uint64_t crossings;          /* surface crossings + reflections */
  uint64_t tally_updates;      /* mesh-tally atomic updates */
  uint64_t lost;               /* histories lost at birth (killed) */
  uint64_t secondaries;        /* secondaries resurrected */
  int32_t error;               /* first fatal code seen, 0 if none */
  int32_t pad_;
  uint64_t error_history;      /* history id of the first offender */
} abx_diag;

/* ---- entry points -------------------------------------------------------- */
typedef struct abx_problem abx_problem;

/* Which kernel family transport uses. */
enum abx_engine {
  ABX_ENGINE_EVENT = 0,   /* event-based: SoA banks, one kernel per event kind */
  ABX_ENGINE_HISTORY = 1  /* persistent grid, one history per lane at a time,
                             lanes refilled from a queue (the default) */
};

/* Launch shape of the transport kernels; what the autotuner
 * (abeille_amd/autotune.py) searches. */
typedef struct abx_tuning {
  int32_t block_size;     /* threads per block of the persistent transport
                             kernels: 64, 128 or 256 (default 256) */
  int32_t blocks_per_cu;  /* persistent grid size in resident blocks per CU
                             (0 = as many as fit, the default) */
  int32_t stage_tables;   /* 1 = copy the problem tables into LDS when they
                             fit (default), 0 = read them through L2 */
  int32_t tail_threshold; /* event engine: particles below which the rest of
                             their histories run in-lane (0 = automatic) */
  int32_t sort_every;     /* event engine: sort the collision queue by
                             (material, group) before every n-th collision
                             launch (0 = never, the default) */
} abx_tuning;

const char* abx_last_error(void);
const char* abx_version(void);
/* Wait for all work on the current device (hipDeviceSynchronize). */
int abx_synchronize(void);

/* Replaces parse_input_file()'s object graph (src/parser.cpp:77-137) and the
 * DeltaTracker/CarterTracker majorant set-up (src/delta_tracker.cpp:50-70,
 * src/majorant.cpp:133-176): validates, builds derived tables, stages them in
 * HBM.  The description's arrays may be freed after the call. */
int abx_create(const abx_problem_desc* desc, abx_problem** out);
void abx_destroy(abx_problem* p);
int abx_set_engine(abx_problem* p, int engine);
int abx_set_tuning(abx_problem* p, const abx_tuning* t);
int abx_get_tuning(abx_problem* p, abx_tuning* t);
/* Device time of one transport() of n fresh source histories under the
 * current tuning (the autotuner's probe).  The current generation's tallies
 * are cleared afterwards: call it between runs, not inside one. */
int abx_time_transport(abx_problem* p, size_t n, int converged, double* seconds);

/* Simulation::sample_sources (src/simulation.cpp:55-77): n source particles
 * for histories first_history .. first_history+n-1, sampled on the GPU. */
int abx_sample_sources(abx_problem* p, uint64_t first_history, size_t n,
                       abx_particle* out);

/* Transporter::transport (include/simulation/transporter.hpp:44-47,
 * src/delta_tracker.cpp:72-265, src/surface_tracker.cpp:40-221,
 * src/carter_tracker.cpp:92-294).  `bank` is consumed (read only here);
 * fission daughters (k-eigenvalue / modified fixed source) are written in
 * the reference's order to `fission_out` (capacity `cap`, count in *n_out).
 * `kcol` is Tallies::kcol() of the previous generation; `converged` gates
 * the mesh tallies (Tallies::score_collision / score_flight).  Scores add to
 * the problem's current-generation tallies.  n may be 0 (no sites, no
 * scores).  If the sites outnumber `cap` (or fission_out is NULL), *n_out,
 * *scores and *diag still receive the call's values, the call returns
 * ABX_E_CAPACITY, and the ordered sites stay on the device, where
 * abx_fission_fetch hands them over: the transport is not repeated and
 * nothing is scored twice. */
int abx_transport(abx_problem* p, const abx_particle* bank, size_t n,
                  double kcol, int converged, abx_banked* fission_out,
                  size_t cap, size_t* n_out, abx_scores* scores,
                  abx_diag* diag);

/* The ordered fission sites of the last transport (abx_transport, or the
 * last generation of abx_run) into `out` (capacity cap; count in *n_out) -
 * the vector transport() returns in the reference
 * (include/simulation/transporter.hpp:44-47).  For a caller whose
 * fission_out was too small: ABX_E_CAPACITY again if cap still is. */
int abx_fission_fetch(abx_problem* p, abx_banked* out, size_t cap, size_t* n_out);

/* Transport that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * meaning as abx_transport, except that `d_bank` and `d_fission_out` are
 * DEVICE pointers (on the problem's device), the history kernel runs
 * whatever abx_set_engine chose, and the ordered sites go straight to
 * d_fission_out (abx_fission_fetch does not see them).  Every buffer beyond
 * the caller's bank and output lives in one static device array of
 * ABX_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a call
 * takes at most ABX_NOALLOC_MAX_HISTORIES histories (more: ABX_E_INVALID),
 * and a d_fission_out of 4n + 4096 sites never overflows.  All calls on one
 * GPU share that array: calls that overlap on one GPU (from several threads,
 * or several problems) race, and must be serialised by the caller.  Splitting
 * a bank into consecutive slices of histories gives the same fission bank
 * (the slices' outputs concatenated) and the same tallies as one call. */
#define ABX_NOALLOC_MAX_HISTORIES 32768
#define ABX_NOALLOC_SCRATCH_BYTES (64u << 20)
int abx_transport_noalloc(abx_problem* p, const abx_particle* d_bank, size_t n,
                          double kcol, int converged, abx_banked* d_fission_out,
                          size_t cap, size_t* n_out, abx_scores* scores,
                          abx_diag* diag);

/* MeshTally generation array [Ne][Nx][Ny][Nz] (src/mesh_tally.cpp:106). */
int abx_tally_size(abx_problem* p, int tally, size_t* n_bins);
int abx_tally_fetch(abx_problem* p, int tally, double* gen_out);
/* MeshTally::clear_generation / Tallies::clear_generation */
int abx_tally_clear(abx_problem* p);
/* MeshTally::record_generation(multiplier) of every mesh tally
 * (src/mesh_tally.cpp:121-150): the generation arrays join the running mean
 * and variance that abx_tally_stats reads - for a host driver that calls
 * abx_transport itself (abx_run records on its own).  The generation arrays
 * are kept; abx_tally_clear zeroes them (Tallies::clear_generation). */
int abx_tally_record(abx_problem* p, double multiplier);

/* ApproximateMeshCancelator::add_particle for every particle of `bank`, then
 * perform_cancellation (src/approximate_mesh_cancelator.cpp:97-190), on the
 * GPU; `bank` (host memory) is updated in place.  Needs a cancellation mesh
 * in the description. */
int abx_cancel_bank(abx_problem* p, abx_banked* bank, size_t n);

/* The parent record of each fission site of the last abx_transport, in the
 * order of its fission_out (BankedParticle::parents_previous_* and
 * Esmp_parent, filled in Transporter::transport, src/transporter.cpp:440-456).
 * Needs an exact cancelator in the description. */
int abx_parent_fetch(abx_problem* p, abx_parent_info* out, size_t cap, size_t* n);

/* PowerIterator::perform_regional_cancellation (src/power_iterator.cpp:751-777)
 * with the exact or basic-exact cancelator, on the GPU: every site of `bank`
 * (with its parent record) is added, cancellation is performed and the
 * uniform particles from get_new_particles are appended.  out receives the
 * n sites with their new weights and then the uniform particles; *n_out is
 * their total.  Draws from the problem's settings::rng. */
int abx_exact_cancel_bank(abx_problem* p, const abx_banked* bank, const abx_parent_info* parents,
                          size_t n, abx_banked* out, size_t cap, size_t* n_out);

/* settings::rng back to settings.seed (settings::initialize_global_rng,
 * src/settings.cpp:116-118); abx_create and every abx_run do this. */
int abx_rng_reset(abx_problem* p);
/* The problem's settings::rng state (pcg32 on its default stream), read and
 * written, so that a host driver's own pcg32 object can be the one the
 * cancellation draws from (perform_cancellation(settings::rng) and
 * get_new_particles(settings::rng), src/power_iterator.cpp:766-771). */
int abx_rng_get_state(abx_problem* p, uint64_t* state);
int abx_rng_set_state(abx_problem* p, uint64_t state);

/* ---- device-resident drivers (the reference's Simulation loops) ---------- */

/* Result of a whole run: Tallies' running statistics. */
typedef struct abx_run_stats {
  /* Tallies::{kcol,kabs,ktrk,ktot,leakage,mig_area}_{avg,err}
   * (include/simulation/tallies.hpp:103-151): running mean and standard
   * error of the mean over the recorded generations (src/tallies.cpp:268-280) */
  double kcol_avg, kcol_err;
  double kabs_avg, kabs_err;
  double ktrk_avg, ktrk_err;
  double ktot_avg, ktot_err;
  double leak_avg, leak_err;
  double mig_avg, mig_err;
  int32_t active_generations;
  int32_t generations_run;  /* generations this call ran: fewer than asked
                             * after abx_request_stop or the time budget */
  double transport_seconds; /* GPU time of the transport phase */
  double kernel_seconds;    /* GPU time of the dominant transport kernel */
  int64_t kernel_launches;  /* its launches */
  int64_t kernel_items;     /* events it processed */
  int64_t histories;        /* histories transported (incl. re-banked) */
  abx_diag diag;            /* summed over the run */
} abx_run_stats;

/* FixedSource::run (src/fixed_source.cpp:81-173) / PowerIterator::run
 * (src/power_iterator.cpp:305-429) with banks resident in HBM.  kcol_hist
 * (optional, [ngenerations]) receives k_col per generation.  Mesh-tally mean
 * and std-error arrays follow MeshTally::record_generation
 * (src/mesh_tally.cpp:121-150) and write_tally (:191-197). */
int abx_run(abx_problem* p, int ngenerations, int nignored,
            double* kcol_hist, abx_run_stats* stats);
/* Continue the current run for `ngenerations` more batches / generations
 * (the bank, k and statistics carry over; the first call samples the
 * source).  `nignored` counts inactive generations from here on.  Stats
 * cover only these generations' timing and diagnostics. */
int abx_run_continue(abx_problem* p, int ngenerations, int nignored,
                     double* kcol_hist, abx_run_stats* stats);
int abx_tally_stats(abx_problem* p, int tally, double* avg, double* std_err);

/* Stopping a run between generations.  abx_request_stop may be called from
 * any thread: the run in progress (abx_run, abx_run_continue,
 * abx_run_replicas, abx_run_replicas_continue), or else the next one,
 * stops at its next generation boundary, as the reference does on SIGINT /
 * SIGTERM after the generation's k and tallies are recorded
 * (src/main.cpp:47-75, src/power_iterator.cpp:412-417, 695-716; there is no
 * prompt).  abx_set_max_time gives each run a wall-clock budget in seconds
 * (settings::max_time; <= 0: none): the run stops when the time left is
 * under two mean generation times (PowerIterator::out_of_time,
 * src/power_iterator.cpp:719-737).  The graph-replayed drivers keep one
 * generation queued ahead, which then still completes.  A stopped run's
 * statistics, tallies and bank are those of a run of
 * abx_run_stats.generations_run generations, and kcol_hist holds that many
 * values. */
int abx_request_stop(abx_problem* p);
int abx_set_max_time(abx_problem* p, double seconds);

/* PowerIterator::run (src/power_iterator.cpp:305-429) for `replicas`
 * (1..16) independent chains of one k-eigenvalue problem that differ only
 * in their seed (seeds[r], or NULL for settings.seed + r), run together:
 * every generation is one transport launch over all chains' histories and
 * one chain-aware bookkeeping pass, replayed from captured graphs.  With
 * from_bank == 0, chain r runs exactly the histories abx_run runs with seed
 * seeds[r] (sources 0..nparticles-1, its own k, weight normalisation and
 * history ids).  With from_bank != 0, every chain starts from the bank the
 * problem holds after abx_run / abx_run_continue (a converged source): its
 * particles become histories 0..n-1 of each chain's stream, with their
 * weights and the run's last k.  The
 * mesh tallies and the statistics in *stats are those of the chains' mean
 * per generation (abx_tally_stats, abx_gen_values); kcol_hist (optional,
 * [replicas][ngenerations]) receives every chain's k_col.  The reference
 * runs one chain per process; this is the batched form of its k-eigenvalue
 * loop.  Not with carter tracking or cancellation.  The next abx_run starts
 * a new run. */
int abx_run_replicas(abx_problem* p, int replicas, const uint64_t* seeds, int from_bank,
                     int ngenerations, int nignored, double* kcol_hist, abx_run_stats* stats);
/* Continue the last abx_run_replicas for `ngenerations` more generations
 * (chains, banks, k and statistics carry over, as abx_run_continue does for
 * one chain; `nignored` counts inactive generations from here on).  Stats
 * cover these generations' timing and diagnostics; kcol_hist is
 * [replicas][ngenerations]. */
int abx_run_replicas_continue(abx_problem* p, int ngenerations, int nignored,
                              double* kcol_hist, abx_run_stats* stats);

/* The per-generation vectors of Tallies (k_col_vec, k_abs_vec, k_trk_vec,
 * leak_vec, mig_vec; src/tallies.cpp:159-181), every generation since the
 * last abx_run / abx_batch_add, inactive ones included.  `quantity` is an
 * abx_gen_quantity; count in *n (out == NULL asks for the count only),
 * ABX_E_CAPACITY if cap is too small. */
enum abx_gen_quantity {
  ABX_GEN_KCOL = 0,
  ABX_GEN_KABS = 1,
  ABX_GEN_KTRK = 2,
  ABX_GEN_KTOT = 3,
  ABX_GEN_LEAKAGE = 4,
  ABX_GEN_MIG = 5
};
int abx_gen_values(abx_problem* p, int quantity, double* out, size_t cap, size_t* n);

/* Device bank of the last k-eigenvalue generation in the reference's
 * source-file layout [N][9] = x,y,z,ux,uy,uz,E,w,w2
 * (src/simulation.cpp:137-175). */
int abx_bank_size(abx_problem* p, size_t* n);
int abx_bank_fetch(abx_problem* p, double* out9);
/* PowerIterator::load_source_from_file (src/power_iterator.cpp:60-133): a
 * [n][9] source becomes the first generation's bank - history ids 0..n-1,
 * streams seeded from them, directions normalised - and nparticles (the
 * tally normalisation) becomes the rounded total weight.  Statistics
 * restart; the next abx_run_continue transports this bank. */
int abx_bank_load(abx_problem* p, const double* in9, size_t n);

/* ---- per-history traces (parity) ----------------------------------------- */
/* One record per loop iteration of a history (after the flight / collision
 * decision), for bit-for-bit comparison with the CPU restatement. */
typedef struct abx_trace_rec {
  double r[3];
  double u[3];
  double wgt;
  uint64_t history_id;
  int32_t secondary_id;
  int32_t event;     /* abx_event */
  int32_t cell;      /* cell index, -1 lost */
  int32_t material;  /* material index, -1 */
  int32_t group;
  int32_t rng_calls; /* 32-bit draws consumed by the history so far */
} abx_trace_rec;

enum abx_event {
  ABX_EV_VIRTUAL = 0,   /* delta/carter virtual collision */
  ABX_EV_COLLISION = 1, /* real collision (after physics) */
  ABX_EV_CROSS = 2,     /* surface / lattice crossing */
  ABX_EV_REFLECT = 3,   /* reflective boundary */
  ABX_EV_LEAK = 4,      /* vacuum boundary */
  ABX_EV_DEATH = 5,     /* killed at collision (roulette) */
  ABX_EV_RESURRECT = 6  /* secondary popped */
};

/* Transport with tracing on: at most `cap_per_history` records per history
 * in `trace` ([n][cap_per_history]); n_rec[i] gets the count (may exceed cap,
 * extra records are dropped). */
int abx_transport_trace(abx_problem* p, const abx_particle* bank, size_t n,
                        double kcol, int converged, abx_banked* fission_out,
                        size_t cap, size_t* n_out, abx_scores* scores,
                        abx_trace_rec* trace, int cap_per_history,
                        int32_t* n_rec);

/* ---- batches: many problems per launch ----------------------------------- */
/* Fixed-source and modified-fixed-source problems advanced together.  Each
 * step gives every attached problem one transport call - a new batch of
 * sources (FixedSource::run, src/fixed_source.cpp:81-173) or the next
 * generation of its fission chain (ModifiedFixedSource::run,
 * src/modified_fixed_source.cpp:59-143) - and all those histories share one
 * persistent launch per kernel variant.  Per-problem results equal abx_run
 * on the problem alone up to the summation order of tallies and scores. */
typedef struct abx_batch abx_batch;

/* budget_bytes: device bytes the attached problems may hold (0 = no limit). */
int abx_batch_create(size_t budget_bytes, abx_batch** out);
void abx_batch_destroy(abx_batch* b);
/* Device bytes a problem holds while attached (tables, tallies, banks). */
int abx_problem_footprint(abx_problem* p, size_t* bytes);
/* Attach p to run `nbatches` batches of its nparticles (statistics and
 * tallies restart, as in abx_run).  ABX_E_BUDGET if it does not fit. */
int abx_batch_add(abx_batch* b, abx_problem* p, int nbatches);
/* Device bytes abx_batch_add would count against the budget for p (the
 * problem's footprint plus its side-by-side batches), for callers that
 * share one budget between several batches. */
int abx_batch_need(abx_batch* b, abx_problem* p, int nbatches, size_t* bytes);
/* One step.  Problems whose last batch was recorded are detached and listed
 * in `finished` (capacity cap, count in *n_finished); *n_active receives the
 * number still attached. */
int abx_batch_step(abx_batch* b, abx_problem** finished, int cap, int* n_finished,
                   int* n_active);
/* Caps the histories one step runs over all problems (0 = no cap, the
 * default).  A fixed-source batch larger than its share is then split over
 * several steps - its histories, RNG streams and recorded statistics are
 * unchanged - so that short steps let newly attached problems start soon.
 * Every problem still gets at least 8192 histories per step, and a
 * modified-fixed-source generation always runs whole. */
int abx_batch_set_step_histories(abx_batch* b, int64_t histories);  /* before any abx_batch_add */
/* Launch parameters of the batch engine, for the autotuner: histories per
 * chunk of the batch kernel (a block's unit of work; 0 = automatic, about
 * eight chunks per CU; -f = about f chunks per CU, which scales with the
 * step's histories) and the histories that the side-by-side batches of
 * one fixed-source problem may hold together (0 = 2^23; a problem whose
 * batch alone exceeds it runs one batch per step).  The reference has no
 * counterpart: its batches run one after another (FixedSource::run,
 * src/fixed_source.cpp:81-173). */
int abx_batch_set_params(abx_batch* b, int64_t chunk, int64_t multi_hist);  /* before any add */
/* Lanes per block of the batch kernel: 64, 128 or 256 (default), for the
 * autotuner.  No counterpart in the reference. */
int abx_batch_set_block_size(abx_batch* b, int block_size);
/* A step whose launches have the previous step's shape - the same kernel
 * variants, chunk counts per variant and device arrays, as the steps of a
 * fixed-source problem with many batches do - replays the previous step's
 * transport as a captured hipGraph (on = 1, the default) instead of
 * launching it again; the problems, contexts and chunks are uploaded into
 * those arrays before every step.  Results do not depend on it.  No
 * counterpart in the reference. */
int abx_batch_set_graphs(abx_batch* b, int on);
/* Transport graphs captured and replayed by the batch so far. */
int abx_batch_graph_stats(abx_batch* b, int64_t* captures, int64_t* replays);
/* GPU time of the batch's transport launches so far. */
int abx_batch_kernel_seconds(abx_batch* b, double* seconds);
/* Statistics of a problem's batch-engine run (or of its last abx_run). */
int abx_problem_stats(abx_problem* p, abx_run_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* ABX_H */
