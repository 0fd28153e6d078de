/*
 * edm_hip.h — C ABI of the MI355X (gfx950) implementation of the ED-MRDT
 * per-timestep particle path (libedm_hip.so).
 *
 * The reference has no native boundary of its own: its Python engines call
 * numba-compiled kernels on numpy arrays.  These entry points replace that
 * Python -> numba boundary.  Each comment names the reference code the entry
 * point stands in for (paths relative to the reference repository).
 *
 * Conventions
 *   - Plain C types only: host pointers, sizes, POD structs.  A membrane's
 *     state lives in device memory owned by its handle; host arrays are
 *     copied in/out explicitly.
 *   - Every function returns 0 on success or a negative EDM_E* code; the
 *     message of the last error on the calling thread is edm_last_error().
 *   - Particle arrays passed in or out are in particle-index order (the
 *     reference's order).  Inside the handle they are kept in cell order.
 *   - "hook" draws are the reference's own numpy random arrays, passed in
 *     the order and sizes the reference draws them; without hook draws the
 *     library uses its counter-based Philox streams (seed, step, particle).
 */
#ifndef EDM_HIP_H
#define EDM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDM_OK 0
#define EDM_EINVAL -1   /* bad argument or unsupported configuration      */
#define EDM_EHIP -2     /* a HIP runtime call failed                       */
#define EDM_ESTATE -3   /* call out of order (e.g. finish before begin)    */
#define EDM_ECAPACITY -4 /* a device buffer bound was exceeded              */

/* Reaction kinds, in the order DynamicsEngine.step runs them
 * (dynamics.py:2205-2465). */
#define EDM_RX_BIND 1     /* binding + dissociation (dynamics.py:529-654)   */
#define EDM_RX_CAT 2      /* catalysis (dynamics.py:826-1090)               */
#define EDM_RX_RECRUIT 3  /* recruitment (dynamics.py:1230-1458)            */
#define EDM_RX_DETACH 4   /* detachment (dynamics.py:1462-1555)             */
#define EDM_RX_CONVERT 5  /* conversion (dynamics.py:1566-1657)             */
#define EDM_RX_ADAPT 6    /* adaptation field update (dynamics.py:1094-1112)*/

/* One Hill feedback factor read from a coarse density grid
 * (dynamics.py:1955-1971 _build_fb_info). sensor < 0: no factor. */
typedef struct edm_feedback {
    int32_t sensor;      /* species whose density is read                  */
    int32_t grid;        /* index into edm_params.grids                    */
    int32_t inhibition;  /* 1: 1/(1+r^n); 0: basal + (1-basal) r^n/(1+r^n) */
    int32_t pad;
    double K_half, n_hill, basal;
} edm_feedback;

typedef struct edm_grid {
    int32_t nx, ny;
    double sx, sy;
} edm_grid;

typedef struct edm_reaction {
    int32_t kind;          /* EDM_RX_*                                      */
    int32_t a, b, c;       /* BIND: ra, rb, product; CAT: enzyme, substrate,
                              product; RECRUIT/DETACH: species; CONVERT:
                              from, to                                      */
    double contact;        /* BIND/CAT: effective contact radius (um);
                              RECRUIT with use_adapt = 1: recruit_sigma (um),
                              the Gaussian offset around the chosen cell     */
    double prob;           /* per-step probability: P_bind, P_cat, P_on,
                              P_off, P_convert                              */
    double prob2;          /* BIND: P_dissoc                                */
    int32_t max_recruits;  /* RECRUIT cap per step (index order)            */
    int32_t use_adapt;     /* CAT: multiply the adaptation factor; RECRUIT:
                              1 = density-biased (local) placement, the
                              cell drawn from the CDF of max(rho, 0.01) on
                              fb1's grid (dynamics.py:1116-1217, 2355-2370;
                              needs fb1 and no fb2). Hook draws of such a
                              reaction: u_acc[n], u_pos[n,2], u_loc[n],
                              xi_off[n,2]                                   */
    edm_feedback fb1, fb2; /* RECRUIT may use both (dual feedback)          */
} edm_reaction;

typedef struct edm_params {
    int32_t n;              /* particles, alive + dormant                    */
    int32_t n_species;
    int32_t fp64;           /* 1: FP64 state and arithmetic; 0: FP32         */
    int32_t periodic_x, periodic_y;
    int32_t ncx, ncy;       /* cell list (dynamics.py:109-114)               */
    int32_t use_wca, use_harmonic;
    int32_t dfield;         /* 0: circular regions (n_regions may be 0);
                               1: sinusoidal D(x)                            */
    int32_t n_regions;
    int32_t n_grids;
    int32_t n_reactions;
    int32_t pair_cap;       /* reference pair-list cap min(30N, 5e5); 0 = none */
    int32_t exact_force_order; /* 1: with hook draws in FP64, sum pair forces in
                               the reference's pair order (bit-exact runs);
                               0: always the LDS-tiled kernel              */
    double Lx, Ly, dt;
    double cutoff, csx, csy;
    double max_disp;        /* displacement cap (dynamics.py:1826-1864)      */
    double smooth_w;        /* region sigmoid width (um)                     */
    double sin_amp, sin_kx, sin_ky;
    double adapt_tau, adapt_K, adapt_n; /* adaptation (dynamics.py:2068-2096) */
    int32_t adapt_grid, adapt_species;  /* -1 = no adaptation                */
    uint64_t seed;
    const double *sigma, *epsilon, *k_harm; /* [n_species * n_species]      */
    const double *D_base;                   /* [n_species]                  */
    const double *rcx, *rcy, *rr, *rf;      /* [n_regions]                  */
    const edm_grid *grids;                  /* [n_grids]                    */
    const edm_reaction *reactions;          /* [n_reactions], run order     */
} edm_params;

/* Host-side particle state in particle-index order (particles.py:130-139). */
typedef struct edm_host_state {
    double *positions;     /* [n, 2] row-major                               */
    int32_t *species_id;
    int32_t *bond_partner;
    uint8_t *is_alive, *has_dye, *is_bleached;
    int32_t *dye_type_id;
    int8_t *ctmc_state;
} edm_host_state;

/* Photophysics tables (photophysics.py:796-887), FP64, per fluorophore t. */
typedef struct edm_photo_params {
    int32_t n_fluoro;
    int32_t n_fret_maps;
    double T_step[8][9], T_step_f[8][9];  /* [t][3x3] row-major             */
    double mean_ph[8][2], mean_ph_f[8][2], mean_fr[8][2], mean_fr_f[8][2];
    double tau[8], k_bound[8];
    int32_t gs[8], es[8], t1[8], bl[8];
    /* bond-mode FRET maps (photophysics.py:1044-1070), in config order    */
    int32_t fm_donor[8], fm_acceptor[8], fm_intra[8], fm_proximity[8];
    int32_t fm_target[8];
    double fm_kfret[8];     /* (R0/r)^6 / tau_D  for bond mode               */
    double fm_k_hi[8], fm_k_lo[8], fm_radius[8]; /* proximity mode          */
    int32_t fm_skip[8];     /* 1: map skipped by the reference (tau_D <= 0)  */
    int32_t max_photons;    /* photophysics.py:879-887                       */
} edm_photo_params;

/* Virtual microscope constants (microscope.py:73-391, 650-712, 819-838). */
typedef struct edm_image_params {
    int32_t nx, ny, n_bins, n_ch;
    int32_t R;              /* PSF window half-width in pixels              */
    int32_t n_sigma;
    double pixel, offx, offy, bin_w, sigma_irf;
    double qe, dark_lambda, p_ap;
    double sigma_ch[8];
} edm_image_params;

typedef struct edm_membrane edm_membrane;

const char *edm_last_error(void);
int edm_device_count(int *count);

/* ---- membrane lifetime (ParticleState + DynamicsEngine.__init__) ------ */
int edm_create(const edm_params *p, edm_membrane **out);
int edm_destroy(edm_membrane *m);
int edm_upload_state(edm_membrane *m, const edm_host_state *s);
int edm_download_state(edm_membrane *m, edm_host_state *s);
int edm_set_photophysics(edm_membrane *m, const edm_photo_params *pp);
int edm_set_imaging(edm_membrane *m, const edm_image_params *ip);

/* ---- DynamicsEngine.step (dynamics.py:2150-2465) -----------------------
 * Plain runs: edm_run(m, k) advances k steps with Philox draws, no host
 * synchronisation.  Parity runs split a step in two so the host can draw
 * the reference's arrays once the pair count is known:
 *   edm_step_begin   D(x), cell list and pair count on pre-move positions
 *                    (dynamics.py:2159-2172); *n_pairs is the length of the
 *                    reference's (possibly capped) pair list
 *   edm_step_finish  forces, Ermak-McCammon step and every reaction, taking
 *                    `draws` = the concatenation of the step's numpy arrays
 *                    in the reference's draw order (NULL = Philox)          */
/* Binding keeps its accepted candidate edges in a device buffer: with hook
 * draws it holds one edge per listed pair (grown before the step); with
 * Philox draws 4 per particle.  A step that accepts more edges than that is
 * not truncated silently: the next edm_step_begin, edm_counters,
 * edm_synchronize or edm_download_state returns EDM_ECAPACITY.            */
int edm_step_begin(edm_membrane *m, int64_t *n_pairs);
int edm_step_finish(edm_membrane *m, const double *draws, int64_t n_draws);
int edm_run(edm_membrane *m, int32_t n_steps);

/* Event counters: bindings, dissociations, catalytic, recruited, detached,
 * converted (dynamics.py:2068-2071, 2147-2148), cumulative. */
int edm_counters(edm_membrane *m, int64_t out[6]);
/* Cumulative photons emitted (edm_photons batch lengths and
 * edm_photons_image) and detected (edm_photons_image). */
int edm_photon_counters(edm_membrane *m, int64_t out[2]);

/* ---- Tuning (no reference counterpart) ----------------------------------
 * Knobs of the kernels' work decomposition; every value gives bit-identical
 * results, so an autotuner may time them on the live membrane.
 *   EDM_TUNE_TILE_ROWS  cell rows per strip of the tiled force and catalysis
 *                       kernels, 1..256 (default 64)
 *   EDM_TUNE_CAT_SIDE   catalysis walks its pairs from the enzymes (0) or
 *                       the substrates (1); -1 (default): the side with fewer
 *                       alive particles at upload
 *   EDM_TUNE_CAT_WALK   0 (default): catalysis walks strips of cells staged
 *                       in LDS; 1: runs walk the step's list of their
 *                       enzymes or substrates where they can
 *   EDM_TUNE_CAT_LANES  a power of two L in 1..64 (default 16): lanes that
 *                       walk one entity's candidates in the strip walk; in
 *                       the list walk each wavefront takes 64 / L entities
 *                       at a time and spreads their candidates over its
 *                       lanes
 *   EDM_TUNE_FRET_SEARCH  proximity FRET finds targets through a grid of the
 *                       alive targets (0, default) or by scanning every
 *                       particle for every donor (1), as the reference does
 *   EDM_TUNE_GRAPHS     1: edm_run, edm_run_frame and edm_time_steps
 *                       record stretches of up to 256 steps into HIP graphs
 *                       and replay them (Philox draws, no per-step census;
 *                       other steps, and runs shorter than 4 steps, go
 *                       launch by launch); 0 (default): every step launch
 *                       by launch.  On ROCm 7.2 a replayed step takes
 *                       20-35 us longer than the same launches issued one
 *                       by one, for 4k to 1M particles
 *   EDM_TUNE_FORCE_LISTS  dense membranes (>= 2 alive particles per cell), in
 *                       steps that do not need the reference's in-cell order
 *                       (no hook draws, no pair cap): 1 (default) the cell
 *                       sort orders cells by force sub-row and the forces
 *                       read their candidates through the sub-row starts;
 *                       0 the forces stage strips of cells in LDS.  Both
 *                       visit the same candidates in the same order
 *   EDM_TUNE_GRAPH_REPLAYS  (read only) graph replays so far
 *   EDM_TUNE_GRAPH_STATUS   (read only) 0, or why a recording failed (the
 *                       membrane then steps launch by launch): 1000 + the
 *                       HIP error of hipStreamBeginCapture, 2000 + the EDM
 *                       error of a recorded step, 3000 + the HIP error of
 *                       hipStreamEndCapture, 4000 + that of
 *                       hipGraphInstantiate, 5000 an allocation while
 *                       recording                                          */
#define EDM_TUNE_TILE_ROWS 1
#define EDM_TUNE_CAT_SIDE 2
#define EDM_TUNE_CAT_WALK 3
#define EDM_TUNE_CAT_LANES 4
#define EDM_TUNE_FRET_SEARCH 5
#define EDM_TUNE_GRAPHS 6
#define EDM_TUNE_GRAPH_REPLAYS 7
#define EDM_TUNE_GRAPH_STATUS 8
#define EDM_TUNE_FORCE_LISTS 9
int edm_set_tuning(edm_membrane *m, int32_t key, int64_t value);
int edm_get_tuning(edm_membrane *m, int32_t key, int64_t *value);
/* Reference pair list in reference order (CellList.get_neighbor_pairs,
 * dynamics.py:220-278) for the cell list of the last edm_step_begin. */
int edm_pairs(edm_membrane *m, int32_t *pairs, int64_t cap, int64_t *n_pairs);
/* Density field of a sensor on a grid (compute_density_field,
 * dynamics.py:1668-1703) for the current post-step state. */
int edm_density(edm_membrane *m, int32_t species, int32_t grid, double *out);

/* ---- PhotonSimulator.evolve_one_bd_step (photophysics.py:1116-1197) ----
 * Hook run: update FRET rates, then the PSS kernel with the reference's
 * rng_poisson[n], rng_transition[n] and rng_arrival[n_arr]; the photon
 * batch stays on the device, *n_photons is its length.                   */
int edm_photons(edm_membrane *m, const double *u_pois, const double *u_trans,
                const double *u_arr, int64_t n_arr, int64_t *n_photons);
int edm_photon_batch(edm_membrane *m, int32_t *particle_idx, double *arrival_time,
                     int8_t *channel, uint8_t *is_fret, int64_t cap);
int edm_fret_rates(edm_membrane *m, double *k_fret, int32_t *acceptor_idx);

/* ---- VirtualMicroscope.process_photons (microscope.py:1069-1157) -------
 * Hook run, three calls with the reference's draws:
 *   edm_image_detect  QE filter with rng.random(n_photons); stable sort by
 *                     particle index; *n_surviving returned
 *   edm_image_assign  emitter-centric PSF pixel draw with rng.random(n_surv)
 *   edm_image_bin     IRF jitter rng.normal(0, sigma_irf, n_assigned) and
 *                     TCSPC binning into the frame histogram               */
int edm_image_detect(edm_membrane *m, const double *u_qe, int64_t n, int64_t *n_surviving);
int edm_image_assign(edm_membrane *m, const double *u_pix, int64_t n, int64_t *n_assigned);
int edm_image_bin(edm_membrane *m, const double *jitter, int64_t n);

/* Plain run: PSS emission, QE, PSF pixel draw, IRF jitter and TCSPC binning
 * fused per emitter with Philox draws.  *n_emitted / *n_detected may be NULL. */
int edm_photons_image(edm_membrane *m, int64_t *n_emitted, int64_t *n_detected);

/* ---- VirtualMicroscope.finalize_frame (microscope.py:1159-1203) --------
 * Hook run: per-pixel totals for the afterpulse draws and explicit counts
 * (flat [ch, x, y, bin] indices) from the host's numpy draws.
 * Plain run: edm_frame_noise adds Poisson dark counts and binomial
 * afterpulses on the device.                                             */
int edm_frame_pixel_totals(edm_membrane *m, uint32_t *totals);
int edm_frame_add_counts(edm_membrane *m, const int64_t *flat_idx, int64_t n);
int edm_frame_noise(edm_membrane *m, int64_t *n_dark, int64_t *n_afterpulse);
int edm_frame_download(edm_membrane *m, uint32_t *flim, int64_t n_elems);
int edm_frame_reset(edm_membrane *m);
/* Streamed frames (finalize_frame, microscope.py:1159-1203, for every frame
 * of a run, as pipeline.py:645 keeps them): edm_frame_collect ends the
 * current frame and returns a ticket; the steps that follow write into a
 * second stack while a second stream adds the detector noise (the draws
 * edm_frame_noise would make at this step), compacts the ended stack to its
 * nonzero bins and resets it.  edm_frame_fetch waits for that and returns
 * meta = {dark counts, afterpulses, cumulative detected photons at the
 * frame's end}, *total entries and, when the arrays are given (cap entries
 * each), the (flat index into [n_ch, nx, ny, n_bins], count) pairs in index
 * order; a call with the arrays releases the frame.  One frame may be
 * pending; the photon counts need Philox imaging (edm_photons_image). */
int edm_frame_collect(edm_membrane *m, int32_t *ticket);
int edm_frame_fetch(edm_membrane *m, int32_t ticket, int64_t *meta, uint32_t *flat_index,
                    uint32_t *count, int64_t cap, int64_t *total);

/* ---- timing (HIP events on the handle's stream) ------------------------ */
int edm_synchronize(edm_membrane *m);
/* Runs n_steps plain steps (+ fused imaging every step when image != 0,
 * + a census row when recording with EDM_REC_AUTO), returning total
 * elapsed ms and the summed duration of the force/BD kernel launches
 * measured with events around each launch.  When the steps run as replayed
 * graphs (EDM_TUNE_GRAPHS) the launches are not timed one by one and
 * *force_ms is -1. */
int edm_time_steps(edm_membrane *m, int32_t n_steps, int32_t image, double *total_ms,
                   double *force_ms, double *breakdown_ms /* [8] or NULL */);
/* A timed window over whatever is queued on the handle's stream between two
 * calls, with no synchronisation in between: end = 0 records the opening
 * event; end = 1 records the closing one, waits for it and returns the
 * milliseconds between them.  No reference counterpart (the reference times
 * its loop with time.perf_counter, pipeline.py:555). */
int edm_time_mark(edm_membrane *m, int32_t end, double *elapsed_ms);
/* Plain run of a stretch of a frame: n_steps x (dynamics step, then the
 * fused photon + imaging step when photophysics and imaging are set, then
 * a census row when recording with EDM_REC_AUTO), no host synchronisation
 * (the BD step loop of SimulationPipeline.run, pipeline.py:556-564). */
int edm_run_frame(edm_membrane *m, int32_t n_steps);

/* ---- Caller-side recording (SimulationPipeline, pipeline.py) -----------
 * A census is one device reduction over the membrane giving every count
 * the reference's pipeline takes with numpy after a step
 * (pipeline.py:567-642) or at a frame end (_extract_ground_truth,
 * pipeline.py:357-455).  The counted fields are consecutive int64s from
 * n_bond_alive to species_alive[31].                                     */
typedef struct edm_census {
    int64_t step;              /* dynamics steps completed                    */
    int64_t n_bond_alive;      /* alive particles with a bond partner (:570)  */
    int64_t n_bond;            /* particles with a bond partner (:418)        */
    int64_t n_bleached;        /* particles with a bleached dye (:582, :419)  */
    int64_t n_donor0;          /* dye of type 0, alive, not bleached (:594)   */
    int64_t n_free_dye[8];     /* by dye type: not bleached, no bond partner
                                  (:424-433)                                  */
    int64_t species_alive[32]; /* alive particles by species (:622-625, :439) */
    int64_t events[6];         /* cumulative, as edm_counters                 */
    int64_t photons;           /* cumulative photons emitted (PhotonBatch
                                  lengths of edm_photons / edm_photons_image) */
    double e_sum;              /* EDM_REC_FRET_SUM: sum over the n_donor0
                                  donors of k tau0 / (1 + k tau0), k = the
                                  step's FRET rate, tau0 = tau[0] of the
                                  photophysics tables (:600-605)             */
} edm_census;

#define EDM_REC_FRET_SUM 1 /* rows carry e_sum (proximity-mode FRET)        */
#define EDM_REC_AUTO 2     /* edm_run, edm_run_frame and edm_time_steps
                              append a row after every step                 */

/* Recording: a device ring of `capacity` census rows.  edm_record_step
 * appends one row (no host synchronisation; EDM_ECAPACITY when full);
 * edm_record_read copies the rows out (cap >= edm_record_rows) and empties
 * the ring.  Opening again resizes and empties it. */
int edm_record_open(edm_membrane *m, int32_t capacity, int32_t flags);
int edm_record_step(edm_membrane *m);
int edm_record_rows(edm_membrane *m, int32_t *n_rows);
int edm_record_read(edm_membrane *m, edm_census *rows, int32_t cap, int32_t *n_rows);
int edm_record_close(edm_membrane *m);

/* Frame-end ground truth (_extract_ground_truth, pipeline.py:357-455):
 * FRET rates updated (update_fret_rates, as pipeline.py:649 does first),
 * then for every dye carrier that is not bleached and whose dye type is
 * donor[j] of a map j (later maps win): E = k tau_D[j] / (1 + k tau_D[j])
 * and tau_DA = tau_D[j] / (1 + k tau_D[j]) for k > 0, else E = 0 and
 * tau_DA = tau_D[j]; 0 for every other particle.  fret_E / tau_DA are [n]
 * in particle-index order; any output may be NULL.  The caller passes the
 * maps the reference does not skip (donor found, finite tau_D > 0), at
 * most 8. */
int edm_ground_truth(edm_membrane *m, int32_t n_maps, const int32_t *donor,
                     const double *tau_D, double *fret_E, double *tau_DA,
                     edm_census *census);

/* ---- The reference's free kernels on device arrays ----------------------
 * dynamics.py exports its numba kernels (dynamics.py:38-60, __all__) and
 * callers (the reference's own tests among them) run them on numpy arrays
 * in place.  These entry points are the same kernels on DEVICE arrays in
 * the reference's layouts: positions / forces / grad_D / noise / offsets
 * f64[n, 2] row-major, species and bond partners int32[n], is_alive
 * uint8[n] (0/1), pair lists int32[m, 2], density and adaptation fields
 * f64[nx, ny] row-major, uniforms f64.  They keep the reference kernels'
 * sequential results (pair order, index-order caps, sequential FP64 sums)
 * and mutate in place as the reference does.  `stream` is a hipStream_t
 * (NULL: the null stream).  An entry point with a count output waits for
 * its stream and writes the count to host memory; the others only enqueue.
 * A feedback density pointer of NULL drops that Hill factor (the plain
 * form of the kernel). */

/* Device memory for callers of these kernels that have no allocator of
 * their own (no reference counterpart: plumbing).  The copies are
 * synchronous. */
int edm_device_alloc(int64_t bytes, void **out);
int edm_device_free(void *p);
int edm_copy_to_device(void *dst, const void *src, int64_t bytes);
int edm_copy_to_host(void *dst, const void *src, int64_t bytes);

/* CellList.build + get_neighbor_pairs (dynamics.py:120-175, 177-278): the
 * pairs (i < j, d^2 < cutoff^2) in the reference's order (i ascending,
 * neighbour cells (dcx, dcy) lexicographic, each cell's members in
 * descending index), at most max_pairs stored (the reference's cap is
 * min(30 n, 500000)); *n_pairs = stored, *n_found = all (may be NULL). */
int edm_cell_pairs(const double *positions, int32_t n, int32_t nx, int32_t ny, double cell_sx,
                   double cell_sy, double Lx, double Ly, int32_t periodic_x, int32_t periodic_y,
                   double cutoff, int64_t max_pairs, int32_t *pairs, int64_t *n_pairs,
                   int64_t *n_found, void *stream);
/* No counterpart in the reference: edm_cell_pairs without any device
 * allocation, for callers that cannot allocate (code that owns its memory
 * budget, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as edm_cell_pairs; the build's scratch is one static
 * device array per GPU (20 MiB), so n is at most EDM_NOALLOC_MAX_PARTICLES
 * and nx * ny at most EDM_NOALLOC_MAX_CELLS (larger calls return
 * EDM_EINVAL), and calls that overlap on one GPU race. */
#define EDM_NOALLOC_MAX_PARTICLES (1 << 20)
#define EDM_NOALLOC_MAX_CELLS (1 << 20)
int edm_cell_pairs_noalloc(const double *positions, int32_t n, int32_t nx, int32_t ny,
                           double cell_sx, double cell_sy, double Lx, double Ly,
                           int32_t periodic_x, int32_t periodic_y, double cutoff,
                           int64_t max_pairs, int32_t *pairs, int64_t *n_pairs, int64_t *n_found,
                           void *stream);
/* bd_step (dynamics.py:323-406) */
int edm_bd_step(double *positions, const double *D, const double *forces, const double *grad_D,
                double dt, double Lx, double Ly, int32_t periodic_x, int32_t periodic_y, int32_t n,
                const double *noise, const uint8_t *is_alive, double max_displacement,
                void *stream);
/* apply_diffusion_field_regions_smooth (dynamics.py:441-519) */
int edm_diffusion_field_regions(double *D, double *grad_D, const double *D_base,
                                const int32_t *species_id, const double *positions, int32_t n,
                                const double *region_cx, const double *region_cy,
                                const double *region_r, const double *region_Dfactor,
                                int32_t n_regions, double smoothing_width, void *stream);
/* compute_forces_wca (kind 0, param = epsilon; dynamics.py:666-739) and
 * compute_forces_harmonic (kind 1, param = spring constant;
 * dynamics.py:753-816): forces += pair forces, each particle's additions in
 * pair order; matrices [n_species, n_species] row-major */
int edm_pair_forces(int32_t kind, const double *positions, double *forces,
                    const int32_t *species_id, const uint8_t *is_alive, int32_t n,
                    const int32_t *pairs, int64_t n_pairs, const double *sigma_matrix,
                    const double *param_matrix, int32_t n_species, double Lx, double Ly,
                    int32_t periodic_x, int32_t periodic_y, void *stream);
/* apply_binding_reactions (dynamics.py:529-606) */
int edm_binding(const double *positions, int32_t *species_id, int32_t *bond_partner,
                uint8_t *is_alive, int32_t n, int32_t reactant_a, int32_t reactant_b,
                int32_t product_id, double contact_radius_um, double prob_bind,
                const int32_t *pairs, int64_t n_pairs, const double *rng_uniform, double Lx,
                double Ly, int32_t periodic_x, int32_t periodic_y, int64_t *n_bound,
                void *stream);
/* apply_dissociation (dynamics.py:610-654); the product must differ from
 * both reactants */
int edm_dissociation(int32_t *species_id, int32_t *bond_partner, uint8_t *is_alive,
                     double *positions, int32_t n, int32_t product_id, int32_t reactant_a,
                     int32_t reactant_b, double prob_dissoc, const double *rng_uniform,
                     int64_t *n_dissociated, void *stream);
/* apply_catalytic_reaction, _with_feedback, _with_feedback_and_adaptation
 * (dynamics.py:826-1090); the enzyme must differ from the substrate.  With
 * enzyme == product (autocatalysis, E + S -> 2E) a substrate converted at
 * pair p acts as an enzyme for the pairs after p, as the reference's
 * sequential loop over the pair list leaves it.  adaptation_field
 * [nx_adapt, ny_adapt] (NULL: none) is read at the feedback grid's indices,
 * as dynamics.py:1078 reads it */
int edm_catalysis(const double *positions, int32_t *species_id, const uint8_t *is_alive,
                  int32_t n, int32_t enzyme_id, int32_t substrate_id, int32_t product_id,
                  double contact_radius_um, double prob, const int32_t *pairs, int64_t n_pairs,
                  const double *rng_uniform, double Lx, double Ly, int32_t periodic_x,
                  int32_t periodic_y, const double *density_field, int32_t nx_grid,
                  int32_t ny_grid, double cell_sx, double cell_sy, double K_half, double n_hill,
                  int32_t is_inhibition, double basal_fraction, const double *adaptation_field,
                  int32_t nx_adapt, int32_t ny_adapt, double K_adapt, double n_adapt,
                  int64_t *n_converted, void *stream);
/* apply_membrane_recruitment, _with_feedback, _with_dual_feedback
 * (dynamics.py:1230-1458; density_1 / density_2 NULL as the form needs) and
 * apply_membrane_recruitment_local (dynamics.py:1116-1217; cdf != NULL,
 * with rng_location, rng_offsets [n, 2] and density_1 its grid) */
int edm_recruitment(const int32_t *species_id, uint8_t *is_alive, double *positions, int32_t n,
                    int32_t target_species_id, double prob, double Lx, double Ly,
                    const double *rng_accept, const double *rng_positions, int32_t max_recruits,
                    const double *density_1, int32_t nx_1, int32_t ny_1, double sx_1, double sy_1,
                    double K_1, double n_1, int32_t inh_1, double basal_1, const double *density_2,
                    int32_t nx_2, int32_t ny_2, double sx_2, double sy_2, double K_2, double n_2,
                    int32_t inh_2, double basal_2, const double *cdf, const double *rng_location,
                    const double *rng_offsets, double recruit_sigma, int64_t *n_recruited,
                    void *stream);
/* apply_membrane_detachment(_with_feedback) (dynamics.py:1462-1555) */
int edm_detachment(const int32_t *species_id, uint8_t *is_alive, const double *positions,
                   int32_t n, int32_t target_species_id, double prob, const double *rng_uniform,
                   const double *density_field, int32_t nx_grid, int32_t ny_grid, double cell_sx,
                   double cell_sy, double K_half, double n_hill, int32_t is_inhibition,
                   double basal_fraction, int64_t *n_detached, void *stream);
/* apply_species_conversion(_with_feedback) (dynamics.py:1566-1657) */
int edm_conversion(int32_t *species_id, const uint8_t *is_alive, const double *positions,
                   int32_t n, int32_t from_id, int32_t to_id, double prob,
                   const double *rng_uniform, const double *density_field, int32_t nx_grid,
                   int32_t ny_grid, double cell_sx, double cell_sy, double K_half, double n_hill,
                   int32_t is_inhibition, double basal_fraction, int64_t *n_converted,
                   void *stream);
/* compute_density_field (dynamics.py:1668-1703): density_out += 1/area per
 * alive particle of the species, each cell the reference's sequential sum */
int edm_density_field(const double *positions, const int32_t *species_id, const uint8_t *is_alive,
                      int32_t n, int32_t target_species_id, int32_t nx_grid, int32_t ny_grid,
                      double cell_sx, double cell_sy, double *density_out, void *stream);
/* update_adaptation_field (dynamics.py:1094-1112) */
int edm_adaptation_update(double *adaptation_field, const double *density_field, int32_t nx,
                          int32_t ny, double dt, double tau_w, void *stream);

/* ---- Observables on the device ------------------------------------------
 * The statistics the reference computes on the host from copied state.
 *   edm_stats_open       starts (or restarts) the mean-squared-displacement
 *                        accumulators: from here on every step adds each
 *                        particle's minimum-image displacement (the unwrapping
 *                        of validation.py:276-290, B1)
 *   edm_stats_msd        sum over the alive particles of the species in
 *                        species_mask (bit s = species s) of |r(t) - r(0)|^2,
 *                        and their number (MSD = sum / count)
 *   edm_stats_pair_hist  unordered pairs of alive particles of the species in
 *                        species_mask by distance, bins of r_max / n_bins
 *                        (minimum image): the counts behind g(r) of
 *                        spatial_correlation_ground_truth (analysis.py:
 *                        733-790, which counts each pair twice)
 *   edm_stats_flim       for each pixel of the current frame's channel,
 *                        out[pixel][4] = sum I, sum I cos(phi_k), sum I
 *                        sin(phi_k), sum I t_k (ns) over the TCSPC bins k,
 *                        phi_k = 2 pi harmonic (k + 0.5) / n_bins, t_k the bin
 *                        centre: mean_arrival_time (analysis.py:87-127) and
 *                        phasor_transform (analysis.py:497-566) are ratios
 *                        of these */
int edm_stats_open(edm_membrane *m);
int edm_stats_close(edm_membrane *m);
int edm_stats_msd(edm_membrane *m, uint32_t species_mask, double *sum_r2, int64_t *count);
int edm_stats_pair_hist(edm_membrane *m, uint32_t species_mask, int32_t n_bins, double r_max,
                        uint64_t *counts);
int edm_stats_flim(edm_membrane *m, int32_t channel, int32_t harmonic, double *out);

/* ---- Batched ensemble ---------------------------------------------------
 * Many independent membranes of one model advanced by the same launches
 * (the reference runs replicas one after another: one SimulationPipeline
 * per replica, pipeline.py:489-738).  The model is an edm_params (species,
 * potentials, reaction network with feedback rules, D-field type, dt,
 * displacement cap; n, box and cell fields ignored) plus the photophysics
 * tables and imaging constants; each replica brings its own box, particle
 * state, seed, reaction probabilities, recruitment caps, density-grid
 * shapes, circular D(x) regions and image geometry; the adaptation field
 * (model adapt_* fields) is kept per replica on its own grid.  Philox
 * draws only (keys relative to each replica's first particle and admission
 * step), so a replica evolves exactly as it does alone.  Batched
 * models carry at most one proximity FRET map (each replica bins its
 * targets into its own grid).                                            */
typedef struct edm_replica {
    int32_t n;                  /* particles (alive + dormant)              */
    int32_t ncx, ncy;           /* cell grid (CellList, dynamics.py:109-114) */
    int32_t n_grids;            /* must equal the model's                   */
    double Lx, Ly, csx, csy, cutoff;
    double sin_kx, sin_ky;      /* sinusoidal D(x) wave numbers             */
    uint64_t seed;
    edm_grid grids[4];          /* this replica's density grids             */
    double prob[16];            /* per model reaction, in model order       */
    int32_t max_recruits[16];
    int32_t img_nx, img_ny, img_R, n_reactions;
    double img_pixel, img_offx, img_offy;
    /* circular D(x) regions of this replica (apply_diffusion_field_regions_smooth,
     * dynamics.py:441-519), at most 8; the smoothing width is the model's */
    int32_t n_regions, pad_;
    double reg_cx[8], reg_cy[8], reg_r[8], reg_f[8];
    double prob2[16];           /* per model reaction: dissociation probability
                                 * of a BIND reaction (P_off)               */
} edm_replica;

typedef struct edm_ensemble edm_ensemble;

/* capacity: particles in the shared pool; max_replicas: resident replicas */
int edm_ens_create(const edm_params *model, const edm_photo_params *pp,
                   const edm_image_params *ip, int64_t capacity, int32_t max_replicas,
                   edm_ensemble **out);
int edm_ens_destroy(edm_ensemble *e);
/* admit a replica; its state is in particle-index order; *slot receives its id */
int edm_ens_add(edm_ensemble *e, const edm_replica *r, const edm_host_state *s, int32_t *slot);
int edm_ens_remove(edm_ensemble *e, int32_t slot);
/* n_steps steps of every resident replica (+ photons and TCSPC if image) */
int edm_ens_run(edm_ensemble *e, int32_t n_steps, int32_t image);
/* frame end of one replica: dark counts and afterpulses, copy (flim may be
 * NULL) and reset of its TCSPC stack */
int edm_ens_frame(edm_ensemble *e, int32_t slot, uint32_t *flim, int64_t n_elems,
                  int64_t *n_dark, int64_t *n_afterpulse);
/* frame end of several replicas at once: detector noise and reset, no copy;
 * the noise counts accumulate per replica (edm_ens_noise_counters) */
int edm_ens_frame_end(edm_ensemble *e, const int32_t *slots, int32_t n_slots);
int edm_ens_noise_counters(edm_ensemble *e, int32_t slot, int64_t out[2]);
/* Frame end of several replicas with every frame returned (the reference
 * keeps each frame's FLIMFrame, pipeline.py:645): detector noise, then each
 * stack's nonzero bins as (flat index into [n_ch, nx, ny, n_bins], count)
 * pairs in index order, replica after replica in list order, then the
 * reset.  The pairs travel to pinned host memory on a copy stream while the
 * compute stream runs the next steps; edm_ens_frames_fetch waits for that
 * copy and returns counts[n_slots] (pairs per replica), meta[n_slots][3]
 * (cumulative dark counts, afterpulses and detected photons after the frame
 * end) and the pairs (cap >= *total).  A fetch with flat_index and count
 * both NULL only waits and reports the sizes; one with them copies the pairs
 * out and ends the collect.  At most two collects may wait for their
 * fetch. */
int edm_ens_frames_collect(edm_ensemble *e, const int32_t *slots, int32_t n_slots,
                           int32_t *ticket);
int edm_ens_frames_fetch(edm_ensemble *e, int32_t ticket, int64_t *counts, int64_t *meta,
                         uint32_t *flat_index, uint32_t *count, int64_t cap, int64_t *total);
int edm_ens_download(edm_ensemble *e, int32_t slot, edm_host_state *s);
/* the replica's counters: bindings, dissociations, catalytic, recruited,
 * detached, converted, photons emitted, photons detected */
int edm_ens_counters(edm_ensemble *e, int32_t slot, int64_t out[8]);
/* pooled particles, live cells, resident replicas, particle capacity */
int edm_ens_info(edm_ensemble *e, int64_t out[4]);
/* Tuning of the batched step's work decomposition (no reference
 * counterpart); every value gives bit-identical results, so an autotuner
 * may time them on the live ensemble.  Takes effect at the next step.
 *   EDM_ENS_TUNE_STRIP_PARTICLES  own particles a strip workgroup aims at,
 *                                 32..512 (default 192)
 *   EDM_ENS_TUNE_PAIR_BELOW       columns holding fewer particles are
 *                                 walked two at a time, 0..1024 (default
 *                                 128; 0: never)
 *   EDM_ENS_TUNE_CAT_LANES        a power of two L in 1..64 (default 16):
 *                                 lanes per catalysis entity of the strip
 *                                 walk; the list walk takes 64 / L enzymes
 *                                 of a wavefront at a time
 *   EDM_ENS_TUNE_CAT_WALK         0 (default): catalysis walks strips of
 *                                 cells; 1: runs whose enzymes are pass
 *                                 species walk each alive enzyme's 3x3
 *                                 cells from the pass work list            */
#define EDM_ENS_TUNE_STRIP_PARTICLES 1
#define EDM_ENS_TUNE_PAIR_BELOW 2
#define EDM_ENS_TUNE_CAT_LANES 3
#define EDM_ENS_TUNE_CAT_WALK 4
int edm_ens_set_tuning(edm_ensemble *e, int32_t key, int64_t value);
int edm_ens_get_tuning(edm_ensemble *e, int32_t key, int64_t *value);
int edm_ens_synchronize(edm_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif /* EDM_HIP_H */
