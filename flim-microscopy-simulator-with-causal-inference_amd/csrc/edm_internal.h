// Internal declarations of libedm_hip.so: the membrane handle and the
// kernel-parameter blocks shared by the dynamics and photon translation
// units.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "edm_device.h"
#include "edm_hip.h"

#define EDM_FLAG_ALIVE 1u
#define EDM_FLAG_DYE 2u
#define EDM_FLAG_BLEACHED 4u

// Cell rows per strip tile of the tiled dynamics kernels.
static constexpr int kTileRows = 256;  // most cell rows per strip of the tiled kernels
// work lists: kWlPer slots per thread of a 256-thread build block; list 0
// holds the dye carriers, the others are assigned each step to the reaction
// passes' runs and the catalysis walks
static constexpr int kWlPer = 8;
static constexpr int kWlSlots = kWlPer * 256;
static constexpr int kWlMaxLists = 8;
static constexpr int kWlDye = 0;

// Work-list membership: species bit set and every flag of `flags` set.
struct WlSpec {
    uint32_t species;
    uint32_t flags;
};

// One list as its consumers read it: a (blocks, kWlPer) grid of 256-thread
// blocks; block (b, y) takes entries y * 256 .. y * 256 + 255 of region b.
struct WlView {
    const int32_t* list;
    const uint32_t* cnt;
};

// Entries of this block's slice (0: the block has no work; uniform).
__device__ __forceinline__ int wl_slice(const WlView& V) {
    const int c = (int)V.cnt[blockIdx.x] - (int)(blockIdx.y * blockDim.x);
    return c > 0 ? min(c, (int)blockDim.x) : 0;
}

__device__ __forceinline__ int wl_slot(const WlView& V) {
    return V.list[(size_t)blockIdx.x * kWlSlots + blockIdx.y * blockDim.x + threadIdx.x];
}

// The work list of one run of per-particle passes (starting at reaction ri)
// this step: the particles of the species the run can touch.
struct PassPlan {
    int ri;
    int list;  // or -1: the run sweeps every slot
    bool late;
    WlSpec spec;
};

// How one catalysis run (reactions ri .. ri + nf - 1, one launch) walks its
// pairs this step.
struct CatPlan {
    int ri, nf;
    bool by_sub;  // from the substrates' side
    int list;     // work list of its walk entities, or -1: strip kernel
    bool late;    // an earlier reaction of the step may change its entities:
                  // the list is built just before the run
    WlSpec spec;
    int app_list = -1;  // a later run's work list to which this run appends the
                        // substrates it converts into that list's species
};

namespace edm {

// Force strips enumerate candidates by sub-row only in membranes with at
// least two alive particles per cell at upload: below that the per-cell
// sort costs more than the candidates it removes.  A property of the
// membrane (its own state and cell grid), so a replica of an ensemble makes
// the same choice as the same membrane alone.
inline bool force_sub_rows(int64_t n_alive, int64_t n_cells) {
    return n_cells > 0 && n_alive >= 2 * n_cells;
}

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define EDM_HIP(call)                                           \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return ::edm::hip_fail(_e, #call); \
    } while (0)

// Device buffers of one membrane.  Particle arrays are in slot order (cell
// order of the last cell-list build); `id` maps slot -> particle index and
// `slot_of` maps back.  Positions: (x, y) are the current positions between
// steps and the pre-move positions during a step; (xn, yn) receive the
// post-move positions and are swapped in at the end of the step.
struct DevArrays {
    void *x = nullptr, *y = nullptr, *xn = nullptr, *yn = nullptr, *xt = nullptr, *yt = nullptr;
    int8_t *sp = nullptr, *sp_t = nullptr;
    uint8_t *fl = nullptr, *fl_t = nullptr;
    int32_t *id = nullptr, *id_t = nullptr;
    int32_t *bond = nullptr, *bond_t = nullptr;
    int8_t *ctmc = nullptr, *ctmc_t = nullptr;
    int8_t *dye = nullptr, *dye_t = nullptr;
    int32_t* slot_of = nullptr;
};

// Host-side state that a step changes besides device memory: the buffer
// rotation of the particle arrays and the validity of slot_of and the work
// lists.  A recorded stretch of steps starts from one such state and leaves
// another.
struct StepHost {
    DevArrays a;
    bool slot_of_valid, wl_valid;
    uint32_t wl_built;
};

struct StepGraph {
    StepHost h0, h1;
    int n;
    bool image, stats;
    hipGraphExec_t exec;
};

}  // namespace edm

struct edm_membrane {
    edm_params p{};
    std::vector<edm_grid> grids;
    std::vector<edm_reaction> rx;
    std::vector<int64_t> draw_off;  // per reaction: offset of its draws in a hook step
    int64_t draws_noise = 0;        // 2n
    hipStream_t stream = nullptr;
    int device = 0;
    int n = 0, ncell = 0;
    bool fp64 = true;
    size_t rs = 8;
    edm::DevArrays a;

    // cell list
    uint32_t *key = nullptr, *rank = nullptr, *cell_cnt = nullptr, *cell_start = nullptr,
             *perm = nullptr;
    void* cub_tmp = nullptr;
    size_t cub_tmp_bytes = 0;
    // the cell scan's tile states (flag << 32 | sum), its tile ticket and a
    // flag set when a tile gave up waiting; cleared by k_bin every step
    uint64_t* scan_state = nullptr;
    uint32_t* scan_ticket = nullptr;  // [0] ticket, [1] gave-up flag
    int scan_tiles = 0;

    // parameters (Real arrays)
    void *sig = nullptr, *eps = nullptr, *kh = nullptr, *Dbase = nullptr;
    double rint2 = 0.0;  // largest squared interaction range of any species pair
    void *rcx = nullptr, *rcy = nullptr, *rr = nullptr, *rf = nullptr;

    // density counts: counts[count_off[g * ns + s] + cell]; -1 = not kept
    std::vector<int32_t> count_off;
    int32_t* d_count_off = nullptr;
    int32_t* d_grid_dims = nullptr;  // [n_grids][2]
    void* d_grid_s = nullptr;        // Real [n_grids][2]
    int32_t *counts = nullptr, *delta = nullptr;
    int n_count = 0;
    uint32_t sensor_mask_all = 0;  // species kept on any grid

    // reference pair bookkeeping (parity runs and the faithful pair cap)
    double* rcdf = nullptr;  // CDF of the density-biased recruitment (largest grid)
    uint64_t* pair_cnt = nullptr;  // by particle index
    uint64_t* pair_off = nullptr;  // exclusive scan of pair_cnt
    int64_t n_pairs_found = 0, n_pairs = 0;
    bool need_pair_offsets = false;
    bool has_autocat = false;             // a catalysis whose enzyme is its product
    unsigned long long* tconv = nullptr;  // [n] autocatalysis: pair index of conversion
    // parity steps read the pair count back every step; step loops leave it
    // on the device (the kernels then decide the cap themselves)
    bool pair_sync = true;

    // step state
    uint32_t step = 0;
    bool begun = false;
    unsigned long long* d_ctr = nullptr;  // 8 counters

    // hook draws staging
    double* d_draws = nullptr;
    size_t draws_cap = 0;

    // fused per-particle passes: runs of detach / convert / recruit-flag
    std::vector<std::vector<int>> pass_groups;
    std::vector<int64_t> alive_at_upload;  // alive particles per species at upload
    // slot_of (particle index -> slot) is kept by every gather only when bonds
    // can exist (bond partners are looked up through it); otherwise it is
    // rebuilt on demand before a kernel that reads it (dyn_ensure_slot_of)
    bool keep_slot_of = true;
    bool slot_of_valid = true;
    // tuning (every value gives the same results)
    int tile_rows = 64;         // cell rows per strip of the tiled force / catalysis kernels
    int cat_side = -1;          // catalysis walk: -1 the rarer side, 0 enzymes, 1 substrates
    int cat_walk = 0;           // 1: catalysis runs walk work lists where they can; 0: strips
    int cat_lanes = 16;         // lanes per entity of the strip walk (1, 2, 4, .., 64)
    bool sub_rows = false;      // force strips by sub-row (dense membranes; set at upload)
    // force sub-row lists: cells ordered by (sub-row, index) with sub-row
    // starts, forces read from global memory (dense membranes, steps that do
    // not need the reference's in-cell order)
    int fine_force = 1;
    bool fine_active = false;   // this step's sort wrote fstart
    int fine_nsub = 1;
    double fine_inv_h = 0.0;
    uint32_t* fstart = nullptr;  // [ncell * nsub + 1] sub-row starts
    uint32_t* fkey = nullptr;    // [n] sort keys from k_bin (sub-row << 28 | index)
    size_t fstart_cap = 0;
    // Work lists rebuilt after every sort (valid once a step has run since
    // the last upload): kWlMaxLists lists of wl_nblk regions of kWlSlots
    int32_t* wl = nullptr;       // [kWlMaxLists][wl_nblk * kWlSlots]
    uint32_t* wl_cnt = nullptr;  // [kWlMaxLists][wl_nblk]
    int wl_nblk = 0;
    bool wl_valid = false;
    uint32_t wl_built = 0;  // lists built since the last sort (bit per list)
    // the lists of the step's sort, built by the step's force launch (or by
    // dyn_wl_flush before any other reader): n lists, their specs and the
    // species and flag arrays of this step's slot order
    int wl_pend_n = 0;
    int wl_pend_list[kWlMaxLists];
    WlSpec wl_pend_spec[kWlMaxLists];
    const int8_t* wl_pend_sp = nullptr;
    const uint8_t* wl_pend_fl = nullptr;
    std::vector<CatPlan> cat_plan;    // this step's catalysis runs
    std::vector<PassPlan> pass_plan;  // this step's runs of per-particle passes
    std::vector<int> group_start;    // per reaction: group starting here, else -1
    std::vector<int> pass_deferred;  // per reaction: reads a count it writes

    // reaction scratch
    unsigned long long* edge_key = nullptr;
    int32_t *edge_a = nullptr, *edge_b = nullptr;
    unsigned int* n_edges = nullptr;
    size_t edge_cap = 0;
    unsigned long long* best = nullptr;  // per slot
    int32_t *cand_id = nullptr, *cand_slot = nullptr;
    double *cand_x = nullptr, *cand_y = nullptr;
    unsigned int* n_cand = nullptr;
    double* w_adapt = nullptr;

    // photophysics
    bool has_pp = false;
    edm_photo_params pp{};
    double* d_pp = nullptr;  // packed photophysics tables (see edm_photons.hip)
    int32_t* dye_ids = nullptr;  // particle indices carrying a dye (fixed at upload)
    int n_dye = 0;
    double* k_fret = nullptr;    // by position in dye_ids
    int32_t* acc = nullptr;      // by position in dye_ids: acceptor particle index or -1
    // proximity FRET: the alive targets of a map binned on a grid of cells
    // no smaller than the proximity radius (counts, fill cursors, starts,
    // then the binned targets' positions and slots)
    int fret_search = 0;         // 0: target grid, 1: scan every particle
    int32_t *tg_cnt = nullptr, *tg_fill = nullptr, *tg_start = nullptr;
    int tg_cell_cap = 0;
    void *tg_x = nullptr, *tg_y = nullptr;
    int32_t* tg_u = nullptr;
    int tg_cap = 0;
    int32_t* ph_nd = nullptr;    // by particle index
    int32_t* ph_nf = nullptr;
    int64_t* ph_cnt = nullptr;   // by particle index, then scanned in place
    int64_t* ph_off = nullptr;
    int32_t* ph_pid = nullptr;
    double* ph_t = nullptr;
    int8_t* ph_ch = nullptr;
    uint8_t* ph_fret = nullptr;
    int64_t ph_cap = 0, n_ph = 0;

    // imaging
    bool has_ip = false;
    edm_image_params ip{};
    uint32_t* flim = nullptr;
    size_t flim_elems = 0;
    size_t flim_words = 0;  // the counts and one flag byte per pixel row
    // streamed frames (edm_frame_collect / edm_frame_fetch): a second stack
    // the steps write into while the ended one gets its detector noise, is
    // compacted to its nonzero bins and reset on fr_stream; the entries and
    // the dark / afterpulse / detected counts of the one pending frame
    uint32_t* flim_alt = nullptr;
    hipStream_t fr_stream = nullptr;
    hipEvent_t fr_photons = nullptr, fr_done = nullptr;
    // photons beside the reactions: when no reaction after the forces can
    // change a dye carrier (its alive flag, bond or position) and the FRET
    // rates come from bonds only, the fused photon kernel of a step waits
    // only for the step's force kernel (ev_force) on ph_stream, and the main
    // stream waits for it (ev_photons) before anything queued later
    bool ph_overlap_ok = false;
    bool force_mark = false;  // ev_force recorded by this step's forces
    hipStream_t ph_stream = nullptr;
    hipEvent_t ev_force = nullptr, ev_photons = nullptr;
    uint32_t *fr_blk = nullptr, *fr_idx = nullptr, *fr_cnt = nullptr;
    unsigned long long* fr_meta = nullptr;
    int fr_nwave = 0;
    int32_t fr_ticket = -1, fr_next = 0;
    double *pcx = nullptr, *pcy = nullptr;
    // hook-run photon staging
    int64_t* det_flag = nullptr;
    int64_t* det_pos = nullptr;
    uint32_t *sv_key = nullptr, *sv_key2 = nullptr, *sv_val = nullptr, *sv_val2 = nullptr;
    int64_t n_surv = 0, n_asg = 0;
    int32_t *asg_px = nullptr, *asg_py = nullptr;
    int64_t* asg_flag = nullptr;
    int64_t* asg_pos = nullptr;
    double* asg_t = nullptr;
    int8_t* asg_ch = nullptr;
    int32_t *sv_px = nullptr, *sv_py = nullptr;
    uint32_t* pix_tot = nullptr;
    int64_t* d_scal = nullptr;  // device scalars for reductions

    // census: recording ring, block partials of the reduction, one-row
    // result for ground truth, per-particle FRET ground truth
    edm_census* rec = nullptr;
    int rec_cap = 0, rec_n = 0, rec_flags = 0;
    unsigned long long* cen_part = nullptr;
    double* cen_part_e = nullptr;
    edm_census* cen_row = nullptr;
    double *gt_e = nullptr, *gt_tau = nullptr;

    // observables (edm_stats_*): per particle index, the unwrapped
    // displacement since edm_stats_open and the position after the last step
    double* st_acc = nullptr;  // [n][2]
    void* st_prev = nullptr;   // Real [n][2]
    double* st_part = nullptr;  // block partial sums
    unsigned long long* st_hist = nullptr;
    int st_hist_cap = 0;
    double* st_pix = nullptr;  // per-pixel sums of the FLIM reduction
    size_t st_pix_cap = 0;

    // step-loop graphs (edm_run_frame, edm_time_steps): stretches of steps
    // recorded once and replayed.  d_step holds the first step of a replay;
    // while recording, launches take their step as an offset from it
    // (step_ref).  graph_gen counts the changes a recording depends on
    // (allocations, parameters, tuning); recordings of an older generation
    // are dropped.
    uint32_t* d_step = nullptr;
    bool graph_rec = false;
    uint32_t graph_step0 = 0;
    int use_graphs = 0;
    uint64_t graph_gen = 1;
    uint64_t warm_gen[2] = {0, 0};  // generation after a clean eager step (no image, image)
    bool graph_failed = false;      // a recording failed: run eagerly from then on
    uint64_t graph_replays = 0;
    int graph_status = 0;           // 0, or why the last recording failed
    std::vector<edm::StepGraph> graphs;

    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t mark0 = nullptr, mark1 = nullptr;  // edm_time_mark's window
    bool mark_open = false;
    std::vector<hipEvent_t> kev;
    bool time_force = false;
    double force_ms_acc = 0.0;
};

// Step of the launches being issued: while a graph is being recorded, the
// offset from the replay's first step (read from d_step at run time);
// otherwise the step itself.
static inline edm::StepRef step_ref(const edm_membrane* m) {
    return m->graph_rec ? edm::StepRef{m->d_step, m->step - m->graph_step0}
                         : edm::StepRef{nullptr, m->step};
}

// Pair-list offsets of a step in a step loop: for the faithful pair cap,
// and for autocatalysis, whose conversions follow the reference's pair order.
static inline bool loop_pair_offsets(const edm_membrane* m) {
    return m->p.pair_cap > 0 || m->has_autocat;
}

namespace edm {
int dyn_wl_flush(edm_membrane* m);
}

static inline WlView wl_view(edm_membrane* m, int L) {
    if (m->wl_pend_n) edm::dyn_wl_flush(m);
    if (!m->wl_valid || !((m->wl_built >> L) & 1u)) return WlView{nullptr, nullptr};
    return WlView{m->wl + (size_t)L * m->wl_nblk * kWlSlots,
                  m->wl_cnt + (size_t)L * m->wl_nblk};
}

namespace edm {
// dynamics translation unit
int dyn_create(edm_membrane* m);
int dyn_step_begin(edm_membrane* m);
int dyn_step_finish(edm_membrane* m, const double* d_draws_dev, bool hook);
// EDM_ECAPACITY once a binding step overflowed its edge buffer (synchronises)
int dyn_check_overflow(edm_membrane* m);
// EDM_EHIP once a cell scan gave up waiting (synchronises)
int dyn_check_scan(edm_membrane* m);
int dyn_pairs(edm_membrane* m, int32_t* pairs, int64_t cap, int64_t* n_out);
int dyn_density(edm_membrane* m, int32_t species, int32_t grid, double* out);
int dyn_ensure_slot_of(edm_membrane* m);
int64_t hook_draws_per_step(const edm_membrane* m, int64_t n_pairs);
// photon translation unit
int ph_alloc(edm_membrane* m);
int ph_fret(edm_membrane* m);
int ph_hook(edm_membrane* m, const double* u_pois, const double* u_trans, const double* u_arr,
            int64_t n_arr, int64_t* n_photons);
int ph_fused_image(edm_membrane* m, int64_t* n_emitted, int64_t* n_detected);
int img_alloc(edm_membrane* m);
int img_detect(edm_membrane* m, const double* u_qe, int64_t n, int64_t* n_surv);
int img_assign(edm_membrane* m, const double* u_pix, int64_t n, int64_t* n_asg);
int img_bin(edm_membrane* m, const double* jitter, int64_t n);
int frame_totals(edm_membrane* m, uint32_t* totals);
int frame_add(edm_membrane* m, const int64_t* flat, int64_t n);
int frame_noise(edm_membrane* m, int64_t* n_dark, int64_t* n_ap);
int frame_reset(edm_membrane* m);
int stage_draws(edm_membrane* m, const double* host, int64_t n);
// census translation unit
int rec_step(edm_membrane* m);  // one census row into the ring (no sync)
int frame_collect(edm_membrane* m, int32_t* ticket);
int frame_fetch(edm_membrane* m, int32_t ticket, int64_t* meta, uint32_t* idx, uint32_t* cnt,
                int64_t cap, int64_t* total);
void frames_free(edm_membrane* m);
// n steps (dynamics, then the fused photon / imaging pass when image),
// through recorded graphs where the membrane allows it, else launch by
// launch
int steps_run(edm_membrane* m, int n, bool image);
bool graphs_eligible(const edm_membrane* m);
// a change that recorded step graphs depend on: drop them
void graph_touch(edm_membrane* m);
void graphs_free(edm_membrane* m);
void rec_free(edm_membrane* m);
// observables translation unit
int stats_accumulate(edm_membrane* m);  // after a step (no sync)
void stats_free(edm_membrane* m);
}  // namespace edm
