// Internal declarations of the batched ensemble (many independent
// membranes of one model stepped by the same launches).  Not part of the C
// ABI; the entry points are the edm_ens_* functions of edm_hip.h.
//
// Layout.  All replicas share one particle pool and one cell-key space.
// Replica r owns the cell keys [coff_r, coff_r + ncx_r * ncy_r); the per-step
// counting sort by key therefore keeps every replica's particles contiguous
// and cell-ordered, exactly as a single membrane's are.  Particles of a
// removed replica get the sentinel key (after every live cell) and fall off
// the end of the pool at the next sort.  Every per-replica quantity (box,
// cell grid, density grids, reaction probabilities, seed, image geometry,
// TCSPC stack, counters) lives in the replica's EnsSeg entry; every kernel
// that works on strips of cells reads its replica from the strip table, so
// a replica's results do not depend on which other replicas share the
// launch.
#pragma once

#include "edm_internal.h"

static constexpr int kEnsMaxGrids = 4;
static constexpr int kEnsMaxRx = 16;
static constexpr int kEnsMaxSpecies = 32;

// One replica slot, as the kernels read it.
struct EnsSeg {
    int32_t active;  // 0: free slot
    int32_t kbase;   // first particle index of the replica
    int32_t n;       // particles (alive + dormant)
    int32_t coff;    // first cell key
    int32_t ncx, ncy;
    int32_t px, py;
    double Lx, Ly, csx, csy, cutoff, sin_kx, sin_ky;
    uint32_t seed_lo, seed_hi;
    uint32_t step0;  // ensemble step at admission: draws use step - step0
    int32_t n_grids;
    int32_t gdims[2 * kEnsMaxGrids];  // [g][nx, ny]
    float gs_f[2 * kEnsMaxGrids];     // [g][sx, sy]
    double gs_d[2 * kEnsMaxGrids];
    float ginv_f[kEnsMaxGrids];       // 1 / cell area
    double ginv_d[kEnsMaxGrids];
    int32_t goff[kEnsMaxGrids * kEnsMaxSpecies];  // count offset of (grid, species) or -1
    double prob[kEnsMaxRx];
    int32_t maxr[kEnsMaxRx];
    edm_image_params ip;  // the model's imaging constants with this replica's geometry
    uint32_t* flim;       // this replica's TCSPC stack, then one flag byte per pixel
    int64_t flim_elems;   // row (set by every writer of the row), padded to 256 bytes
    int64_t flim_words;   // 32-bit words of the allocation
    double* rcdf;         // CDF of density-biased recruitment on this replica's grid, or nullptr
    // work lists: this replica's region of the pass list [0] and of the dye
    // list [1] starts at entry wl_off (a multiple of the wavefront size, so
    // that every wave of a list launch belongs to one replica) and holds
    // wl_n entries
    int32_t wl_off[2];
    int32_t wl_n[2];
    int32_t sub_rows;  // force candidates by sub-row (dense replicas)
    int32_t n_regions;  // circular D(x) regions of this replica
    void* reg;          // their centres x, y, radii and D factors (Real[4][8]), or nullptr
    double* w_adapt;    // this replica's adaptation field on its adaptation grid, or nullptr
    double prob2[kEnsMaxRx];  // BIND reactions: dissociation probability
    int32_t* slot_of;   // binding models: slot of each particle (index - kbase)
    // proximity FRET: this replica's grid of the map's targets (cells no
    // narrower than the radius plus a rounding margin) and its first cell in
    // the pooled target-cell space
    int32_t tg_ncx, tg_ncy, tg_off;
    double tg_ix, tg_iy;
};

static constexpr int kEnsMaxRegions = 8;

// The wave's replica and its first entry of list L: every wave of a list
// launch lies inside one replica's region.
struct EnsWave {
    int r, j;  // replica; entry index of lane 0 within the replica's region
};

__device__ __forceinline__ EnsWave ens_wave(const int32_t* __restrict__ wrep,
                                            const EnsSeg* __restrict__ segs, int L) {
    const int w = (int)blockIdx.x * (int)(blockDim.x >> 6) +
                  __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int r = wrep[w];
    return EnsWave{r, 64 * w - segs[r].wl_off[L]};
}

// A run of one live replica's slots [beg, end) (the work-list build's unit).
struct EnsChunk {
    int32_t r, beg, end, pad;
};

// Strip tile of one replica.
struct EnsTile {
    int32_t r, cx, cy0, rows, ncol;  // ncol own columns cx .. cx + ncol - 1
};

struct edm_ensemble {
    edm_params p{};  // the model: species, potentials, reactions, D field, dt
    std::vector<edm_grid> grids;
    std::vector<edm_reaction> rx;
    bool fp64 = false;
    size_t rs = 4;
    hipStream_t stream = nullptr;
    int device = 0;
    int64_t cap = 0;  // particle capacity of the pool
    int rmax = 0;

    edm::DevArrays a;  // pool (bond and slot_of unused)
    uint16_t *rep = nullptr, *rep_t = nullptr;
    uint32_t *key = nullptr, *rank = nullptr, *perm = nullptr;
    uint32_t *cell_cnt = nullptr, *cell_start = nullptr;
    int64_t cell_cap = 0;
    void* cub_tmp = nullptr;
    size_t cub_bytes = 0;
    int32_t *counts = nullptr, *delta = nullptr;
    int64_t count_cap = 0;
    int32_t *cand_id = nullptr, *cand_slot = nullptr;
    double *cand_x = nullptr, *cand_y = nullptr;
    unsigned int* n_cand = nullptr;        // [rmax]
    unsigned long long* ctr = nullptr;     // [rmax][8]
    unsigned long long* frame_out = nullptr;  // [2] dark counts, afterpulses of one frame
    unsigned long long* noise_ctr = nullptr;  // [rmax][2] dark counts, afterpulses (frame_end)
    int32_t* d_slot_list = nullptr;
    size_t slot_list_cap = 0;
    void *sig = nullptr, *eps = nullptr, *kh = nullptr, *Dbase = nullptr;
    double rint2 = 0.0;
    double max_disp_model = 0.0;
    edm_photo_params pp{};
    bool has_pp = false;
    double* d_pp = nullptr;
    edm_image_params ip{};  // model imaging constants

    std::vector<EnsSeg> seg;  // host mirror, [rmax]
    EnsSeg* d_seg = nullptr;
    std::vector<EnsTile> tiles;
    int tile_rows = 64;  // most cell rows of any strip (sizes the strips' dynamic LDS)
    int tile_cols = 1;   // most own columns of any strip
    // tuning (every value gives the same results): own particles a strip
    // aims at, columns holding fewer particles than pair_below are taken
    // two at a time (boxes of 4 columns or more; 0: never), lanes per
    // entity of the catalysis walk
    int strip_parts = 192;
    int pair_below = 128;
    int cat_lanes = 16;
    int cat_walk = 0;  // 1: catalysis runs walk their enzymes over the pass work list
    // binding: bonds in the pool (partner particle indices), candidate edges
    // in per-replica regions (entries 2 * first slot .., 2 * n of them),
    // per replica the edge count and the largest overflow seen, the
    // matching's per-slot minimum key
    bool has_bind = false;
    // autocatalysis: pair-list lengths and offsets per replica slot range
    // (the replica's first slot + particle index - kbase), the pairs'
    // conversion times per slot ("never" between passes); its accepted
    // pairs share the binding edge buffers
    bool has_autocat = false;
    // proximity FRET (one map at most): the map, the pooled target cells
    // (counts, fill cursors, starts) and the binned targets
    int prox_map = -1;
    int64_t n_tcells = 0, tcell_cap = 0;
    int32_t *tg_cnt = nullptr, *tg_fill = nullptr, *tg_start = nullptr, *tg_u = nullptr;
    void *tg_x = nullptr, *tg_y = nullptr;
    void* cub_tmp_tg = nullptr;
    size_t cub_bytes_tg = 0;
    uint64_t *pcnt = nullptr, *poff = nullptr;
    unsigned long long* tconv = nullptr;
    void* cub_tmp64 = nullptr;
    size_t cub_bytes64 = 0;
    unsigned long long* ekey = nullptr;
    int32_t *ea = nullptr, *eb = nullptr;
    unsigned int* ne = nullptr;  // [rmax][2]
    unsigned long long* best = nullptr;
    EnsTile* d_tiles = nullptr;
    int64_t tile_cap = 0;
    int64_t n_total = 0;  // particles in the pool (live replicas first after a sort)
    int64_t n_cells = 0;  // live cell keys; key n_cells is the sentinel
    int64_t n_counts = 0;
    uint32_t gstep = 0;
    bool layout_dirty = true;
    int64_t n_retired = 0;  // particles of removed replicas still in the pool
    int64_t next_id = 0;    // particle indices handed out so far
    std::vector<int64_t> fresh_beg;  // per slot: pool position before its first sort, or -1
    std::vector<int> draining;       // removed slots whose particles are still pooled

    // model-level reaction planning (as for one membrane)
    uint32_t sensor_mask_all = 0;
    std::vector<std::pair<int, int>> kept;  // (grid, species) with density counts
    std::vector<std::vector<int>> pass_groups;
    std::vector<int> group_start, pass_deferred;
    std::vector<std::pair<int, int>> cat_runs;  // (first reaction, reactions) per launch
    std::vector<char> cat_list_ok;              // per run: the list walk may take it
    std::vector<int> cat_run_of;                // per reaction: run it starts, else -1
    // per replica slot and run / group: the kernels' parameter blocks with the
    // replica's probabilities, feedback grids and seed (rebuilt by ens_layout)
    void* d_cats = nullptr;
    size_t cats_bytes = 0;
    void* d_pgrp = nullptr;
    size_t pgrp_bytes = 0;
    // work lists rebuilt by every step's gather over the live replicas: the
    // slots of pass-species particles and of dye carriers, in one region per
    // replica padded to whole waves
    uint32_t pass_species = 0;
    int32_t *wl_pass = nullptr, *wl_dye = nullptr;  // [wl_cap]
    int64_t wl_cap = 0;
    uint32_t* wl_rcnt = nullptr;                     // [rmax][2] entries appended this step
    int32_t *wl_wrep_pass = nullptr, *wl_wrep_dye = nullptr;  // replica of each wave
    std::vector<int64_t> seg_npass, seg_ndye;  // per slot: entries it contributes
    std::vector<std::vector<int64_t>> seg_alive;  // per slot: alive per species at admission
    int64_t wl_npass = 0, wl_ndye = 0;         // padded list lengths (launch sizes)
    std::vector<EnsChunk> wl_chunks;           // the list build's blocks
    EnsChunk* d_wl_chunks = nullptr;
    int64_t wl_chunk_cap = 0;

    // Streamed frames (edm_ens_frames_collect / edm_ens_frames_fetch): two
    // buffers used in turn.  A collect compacts the due stacks into the
    // device buffer and a copy stream moves the entries to the pinned host
    // buffer while the compute stream runs on; the fetch waits for that copy.
    struct FrameBuf {
        int32_t ticket = -1;   // the collect it holds, -1: free
        int32_t* d_slots = nullptr;  // the listed replicas, then their first wavefront entries
        uint32_t* d_blk = nullptr;   // per (replica, wavefront of 64 pixel rows): nonzero
                                     // bins, then offsets
        uint32_t *d_idx = nullptr, *d_cnt = nullptr;
        unsigned long long* d_meta = nullptr;  // per slot: dark, afterpulses, detected
        int64_t slot_cap = 0, cap = 0, blk_cap = 0;
        std::vector<int64_t> wbase;  // per listed replica: its first wavefront entry
        uint32_t* h_blk = nullptr;  // pinned
        uint32_t *h_idx = nullptr, *h_cnt = nullptr;
        unsigned long long* h_meta = nullptr;
        std::vector<int32_t> slots;
        hipEvent_t written = nullptr, copied = nullptr;
    } fbuf[2];
    hipStream_t copy_stream = nullptr;
    int32_t next_ticket = 0;
};

namespace edm {
int ens_create(edm_ensemble* e);
int ens_layout(edm_ensemble* e);
int ens_step(edm_ensemble* e, bool image);
int ens_check_edges(edm_ensemble* e);
int ens_photons(edm_ensemble* e);
int ens_frame_noise(edm_ensemble* e, int slot, int64_t* n_dark, int64_t* n_ap);
int ens_frame_reset(edm_ensemble* e, int slot);
int ens_frame_end(edm_ensemble* e, const int32_t* slots, int32_t n);
int ens_frames_collect(edm_ensemble* e, const int32_t* slots, int32_t n, int32_t* ticket);
int ens_frames_fetch(edm_ensemble* e, int32_t ticket, int64_t* counts, int64_t* meta,
                     uint32_t* idx, uint32_t* cnt, int64_t cap, int64_t* total);
void ens_frames_free(edm_ensemble* e);
void plan_passes(const std::vector<edm_reaction>& rx, uint32_t sensor_mask,
                 std::vector<std::vector<int>>& groups, std::vector<int>& group_start,
                 std::vector<int>& deferred, int max_rec);
uint32_t pass_species_of(const std::vector<edm_reaction>& rx,
                         const std::vector<std::vector<int>>& groups);
}  // namespace edm
