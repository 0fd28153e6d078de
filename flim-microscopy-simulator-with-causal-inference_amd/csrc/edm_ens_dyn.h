// Batched ensemble: dynamics kernels and the step (included at the end of
// edm_dynamics.hip, inside namespace edm).  Every kernel runs the single-
// membrane bodies (force_strip, cat_strip, passes_particle,
// recruit_apply_block) on views of one replica built from its EnsSeg entry,
// so a replica computes exactly what it computes alone.

template <typename Real>
__device__ __forceinline__ Cells<Real> ens_cells(const EnsSeg& g, const Real* x, const Real* y,
                                                 const int8_t* sp, const uint8_t* fl,
                                                 const int32_t* id, const uint32_t* start) {
    Cells<Real> C;
    C.x = x;
    C.y = y;
    C.sp = sp;
    C.fl = fl;
    C.id = id;
    C.start = start + g.coff;
    C.ncx = g.ncx;
    C.ncy = g.ncy;
    C.csx = (Real)g.csx;
    C.csy = (Real)g.csy;
    C.Lx = (Real)g.Lx;
    C.Ly = (Real)g.Ly;
    C.cut2 = (Real)g.cutoff * (Real)g.cutoff;
    C.px = g.px != 0;
    C.py = g.py != 0;
    C.pcnt = nullptr;
    C.poff = nullptr;
    C.npairs = nullptr;
    C.cap_active = false;
    C.cap_dev = false;
    C.cap_gate = 0;
    C.kbase = g.kbase;
    // pair offsets (autocatalysis) are kept per slot range of the replica:
    // particle index i at the replica's first slot + (i - kbase)
    C.poff_base = (int)start[g.coff] - g.kbase;
    C.tile_rows = kTileRows;  // unused: the strip table gives each strip's rows
    C.sub_rows = g.sub_rows != 0;
    return C;
}

template <typename Real>
__device__ __forceinline__ Grids<Real> ens_grids(const EnsSeg& g, int32_t* counts, int32_t* delta,
                                                 int ns) {
    Grids<Real> G;
    G.off = g.goff;
    G.dims = g.gdims;
    if constexpr (sizeof(Real) == 8) {
        G.s = (const Real*)g.gs_d;
        G.inv = (const Real*)g.ginv_d;
    } else {
        G.s = (const Real*)g.gs_f;
        G.inv = (const Real*)g.ginv_f;
    }
    G.counts = counts;
    G.delta = delta;
    G.ns = ns;
    return G;
}

// A model feedback rule bound to replica g's grid.
__device__ __forceinline__ FbDev ens_fb(const FbDev& f, const EnsSeg& g, int ns) {
    FbDev d = f;
    if (f.sensor >= 0 && f.grid >= 0) {
        d.cbase = g.goff[f.grid * ns + f.sensor];
        d.gnx = g.gdims[2 * f.grid];
        d.gny = g.gdims[2 * f.grid + 1];
        d.gsx = g.gs_d[2 * f.grid];
        d.gsy = g.gs_d[2 * f.grid + 1];
        d.ginv = g.ginv_d[f.grid];
    }
    return d;
}

__device__ __forceinline__ Rng ens_rng(const EnsSeg& g) { return Rng{g.seed_lo, g.seed_hi}; }

template <typename Real>
__device__ __forceinline__ StripGeo ens_strip(const EnsSeg& g, const EnsTile& T) {
    StripGeo S;
    S.tile = 0;
    S.cx = T.cx;
    S.cy0 = T.cy0;
    S.rows = min(T.rows, g.ncy - T.cy0);
    S.nrow = S.rows + 2;
    S.ncol = T.ncol;
    S.nscol = T.ncol + 2;
    S.ncells = S.nscol * S.nrow;
    return S;
}

// Cell key of every pooled particle (its replica's cell, offset by the
// replica's first key; the sentinel for removed replicas), run-aggregated
// ranks as in k_bin, and the per-step reset of the density counts.
template <typename Real>
__global__ void k_ens_bin(int n, const Real* __restrict__ x, const Real* __restrict__ y,
                          const uint16_t* __restrict__ rep, const EnsSeg* __restrict__ segs,
                          uint32_t sentinel, uint32_t* __restrict__ key,
                          uint32_t* __restrict__ rank, uint32_t* __restrict__ cnt,
                          int32_t* counts, int n_counts, uint32_t* wl_rcnt, int n_rcnt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_counts) counts[s] = 0;
    if (s < n_rcnt) wl_rcnt[s] = 0;  // work lists are rebuilt by the gather
    const bool live = s < n;
    uint32_t c = 0xffffffffu;
    if (live) {
        const EnsSeg& g = segs[rep[s]];
        if (g.active) {
            const int cx = trunc_cell(x[s], (Real)g.csx, g.ncx);
            const int cy = trunc_cell(y[s], (Real)g.csy, g.ncy);
            c = (uint32_t)(g.coff + cx * g.ncy + cy);
        } else {
            c = sentinel;
        }
        key[s] = c;
    }
    const int lane = lane_id();
    const uint32_t prev = __shfl_up(c, 1);
    const bool head = live && (lane == 0 || prev != c);
    const uint64_t heads = __ballot(head);
    const int nlive = __popcll(__ballot(live));
    const uint64_t upto = heads & ((2ull << lane) - 1ull);
    const int rs = upto ? 63 - __clzll(upto) : lane;
    uint32_t base = 0;
    if (head) {
        const uint64_t after = heads & ~((2ull << lane) - 1ull);
        const int nxt = after ? min(__ffsll((unsigned long long)after) - 1, nlive) : nlive;
        base = atomicAdd(&cnt[c], (uint32_t)(nxt - lane));
    }
    base = __shfl(base, rs);
    if (live) rank[s] = base + (uint32_t)(lane - rs);
}

// kBonds (binding models): the bonds move with their particles, and each
// live replica's slot_of records where its particles went.
template <typename Real, bool kBonds>
__global__ void k_ens_gather(int n, const uint32_t* __restrict__ perm, DevArrays a,
                             const uint16_t* __restrict__ rep, uint16_t* __restrict__ rep_t,
                             const EnsSeg* __restrict__ segs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t s = perm[t];
    if constexpr (kBonds) {
        a.bond_t[t] = a.bond[s];
        const EnsSeg& g = segs[rep[s]];
        if (g.active && g.slot_of) g.slot_of[a.id[s] - g.kbase] = t;
    }
    ((Real*)a.xt)[t] = ((const Real*)a.x)[s];
    ((Real*)a.yt)[t] = ((const Real*)a.y)[s];
    a.sp_t[t] = a.sp[s];
    a.fl_t[t] = a.fl[s];
    a.id_t[t] = a.id[s];
    a.ctmc_t[t] = a.ctmc[s];
    a.dye_t[t] = a.dye[s];
    rep_t[t] = rep[s];
}

// Work lists of the live replicas: block x lists the pass-species slots and
// the dye slots of chunk x (one replica's slots) into that replica's regions.
__global__ __launch_bounds__(256) void k_ens_wl_build(const EnsChunk* __restrict__ chunks,
                                                      const EnsSeg* __restrict__ segs,
                                                      const int8_t* __restrict__ sp,
                                                      const uint8_t* __restrict__ fl,
                                                      uint32_t pass_species, int32_t* wl_pass,
                                                      int32_t* wl_dye, uint32_t* rcnt) {
    const EnsChunk C = chunks[blockIdx.x];
    const EnsSeg& g = segs[C.r];
    wl_build_block(
        C.beg, C.end, [&](int t) { return ((pass_species >> sp[t]) & 1u) != 0u; },
        [&](int t) { return (fl[t] & EDM_FLAG_DYE) != 0; },
        WlRun{wl_pass, rcnt + 2 * C.r, (uint32_t)g.wl_off[0], (uint32_t)g.wl_n[0]},
        WlRun{wl_dye, rcnt + 2 * C.r + 1, (uint32_t)g.wl_off[1], (uint32_t)g.wl_n[1]});
}

template <typename Real>
struct EnsPool {
    const Real *x, *y;
    Real *xn, *yn;
    int8_t* sp;
    uint8_t* fl;
    const int32_t* id;
    const uint32_t* start;
    int32_t *counts, *delta;
    int ns, n_grids;
    int32_t* bond;  // binding models: partner particle index per slot, else nullptr
};

template <typename Real>
__device__ __forceinline__ Post<Real> ens_post(const EnsPool<Real>& Q) {
    Post<Real> P;
    P.xn = Q.xn;
    P.yn = Q.yn;
    P.sp = Q.sp;
    P.fl = Q.fl;
    P.bond = nullptr;
    P.slot_of = nullptr;
    return P;
}

template <typename Real, bool kSub>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 7>))) void k_ens_force(const EnsTile* __restrict__ tiles, int ntiles,
                                                   const EnsSeg* __restrict__ segs,
                                                   EnsPool<Real> Q, ForceParams<Real> F,
                                                   uint32_t keep_mask, uint32_t gstep) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    ForceParams<Real> Fr = F;
    Fr.sin_kx = (Real)g.sin_kx;
    Fr.sin_ky = (Real)g.sin_ky;
    Fr.n_regions = g.n_regions;
    if (g.n_regions > 0) {
        const Real* R = (const Real*)g.reg;
        Fr.rcx = R;
        Fr.rcy = R + kEnsMaxRegions;
        Fr.rr = R + 2 * kEnsMaxRegions;
        Fr.rf = R + 3 * kEnsMaxRegions;
    }
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    force_strip<Real, kSub, kMaxOwnCols>(C, Fr, Q.xn, Q.yn, (const double*)nullptr, ens_rng(g),
                                         gstep - g.step0, G, Q.n_grids, keep_mask,
                                         ens_strip<Real>(g, T));
}

// Catalysis for one run of catalytic reactions; the run's parameter block
// for each replica (probabilities, feedback grid, seed) is precomputed.
template <typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 7>))) void k_ens_cat(const EnsTile* __restrict__ tiles, int ntiles,
                                                 const EnsSeg* __restrict__ segs, EnsPool<Real> Q,
                                                 const CatArgs<Real>* __restrict__ cats,
                                                 int n_runs, int ci, unsigned long long* ctr,
                                                 uint32_t gstep, int32_t* fold_counts,
                                                 int32_t* fold_delta, int fold_n) {
    // the previous run's deferred count updates (this run reads no counts)
    if (fold_n > 0) fold_in_launch(fold_counts, fold_delta, fold_n);
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    const CatArgs<Real>& A = cats[(size_t)T.r * n_runs + ci];
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    cat_strip<Real, kMaxOwnCols>(C, ens_post(Q), A, gstep - g.step0, ctr + 8 * T.r, G, Q.n_grids,
              ens_strip<Real>(g, T));
}

// Catalysis run walked from its enzymes over the pass work list (every
// particle of the pass species, alive or not, in slot order; each wave is
// one replica's).  A wave keeps its alive particles of the run's enzyme
// species and walks them 64 >> group_shift at a time with cat_flat_wave,
// as the single-membrane list walk does.
// Every pair is decided by its own uniform and a substrate converts once
// (compare-and-swap of its species byte), so the outcome is the strip
// kernel's.  Used for runs whose enzyme species are all pass species and
// none of them is the run's substrate or product.
template <typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 7>))) void k_ens_cat_list(int n_waves, const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ wrep,
                                                      const uint32_t* __restrict__ rcnt,
                                                      const EnsSeg* __restrict__ segs,
                                                      EnsPool<Real> Q,
                                                      const CatArgs<Real>* __restrict__ cats,
                                                      int n_runs, int ci,
                                                      unsigned long long* ctr, uint32_t gstep,
                                                      int32_t* fold_counts, int32_t* fold_delta,
                                                      int fold_n) {
    // the previous run's deferred count updates (this run reads no counts)
    if (fold_n > 0) fold_in_launch(fold_counts, fold_delta, fold_n);
    if ((int)blockIdx.x * (int)(blockDim.x >> 6) +
            __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) >= n_waves)
        return;  // whole waves
    const EnsWave W = ens_wave(wrep, segs, 0);
    const EnsSeg& g = segs[W.r];
    const CatArgs<Real>& A = cats[(size_t)W.r * n_runs + ci];
    const int lane = lane_id();
    const int j = W.j + lane;
    int t = 0;
    bool ent = false;
    if (j < min(g.wl_n[0], (int)rcnt[2 * W.r])) {
        t = list[g.wl_off[0] + j];
        if (Q.fl[t] & EDM_FLAG_ALIVE) {
            const int s = Q.sp[t];
#pragma unroll
            for (int k = 0; k < kCatFuse; ++k) ent |= k < A.n_enz && s == A.enz[k];
        }
    }
    const uint64_t mask = __ballot(ent);
    if (!mask) return;  // uniform: no enzyme among the wave's entries
    __shared__ CatFlatLds<Real> sL[4];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    const int hits = cat_flat_wave(C, ens_post(Q), A, gstep - g.step0, false, G, Q.n_grids, t,
                                   mask, 64 >> A.group_shift, sL[threadIdx.x >> 6]);
    unsigned long long v = hits;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane == 0 && v) atomicAdd(&ctr[8 * W.r + 2], v);
}

// A run of per-particle passes over the strip's own particles, with the
// run's parameter block of the strip's replica.
template <typename Real, bool kLocal>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 7>))) void k_ens_passes(int n_waves, const int32_t* __restrict__ list,
                                                    const int32_t* __restrict__ wrep,
                                                    const uint32_t* __restrict__ rcnt,
                                                    const EnsSeg* __restrict__ segs,
                                                    EnsPool<Real> Q,
                                                    const PassGroupDev* __restrict__ pg,
                                                    int n_groups, int gi,
                                                    unsigned long long* ctr, int32_t* cid,
                                                    int32_t* cslot, double* cxo, double* cyo,
                                                    unsigned int* nc, uint32_t gstep) {
    if ((int)blockIdx.x * (int)(blockDim.x >> 6) +
            __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) >= n_waves)
        return;  // whole waves
    const EnsWave W = ens_wave(wrep, segs, 0);
    const EnsSeg& g = segs[W.r];
    const PassGroupDev& grp = pg[(size_t)W.r * n_groups + gi];
    const int j = W.j + lane_id();
    int det = 0, conv = 0;
    Cand<Real> cand{false, 0, (Real)0, (Real)0, 0};
    int t = -1;
    if (j < min(g.wl_n[0], (int)rcnt[2 * W.r])) {
        t = list[g.wl_off[0] + j];
        const int s0 = Q.sp[t];
        if ((grp.species >> s0) & 1u) {
            const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
            passes_particle<Real, kLocal>(t, s0, g.n, Q.id, g.kbase, ens_post(Q), grp,
                                          (const double*)nullptr, ens_rng(g), gstep - g.step0,
                                          (Real)g.Lx, (Real)g.Ly, G, Q.n_grids, cand, det,
                                          conv);
        }
    }
    // the wave is one replica's: its candidate region is that replica's own
    // slot range, its counter the replica's
    const int pbeg = (int)Q.start[g.coff];
    append_candidates(cand, t, cid + pbeg, cslot + pbeg, cxo + pbeg, cyo + pbeg, nc + W.r);
    passes_count<Real>(det, conv, ctr + 8 * W.r);
}

// Binding (dynamics.py:529-654) in the batch.  The candidate edges of a
// replica go to its own region of the edge buffers (entries 2 * first
// slot .., at most 2 n; a larger count is recorded for the host to report),
// one workgroup per replica then matches them greedily in key order, and
// every complex may dissociate, its partner found through the replica's
// slot_of -- the single-membrane bodies on the replica's views.
template <typename Real>
__device__ __forceinline__ Post<Real> ens_post_bonds(const EnsPool<Real>& Q, const EnsSeg& g) {
    Post<Real> P = ens_post(Q);
    P.bond = Q.bond;
    P.slot_of = g.slot_of;
    return P;
}

// Own slots of strip T: its columns' rows cy0 .. cy0 + rows - 1, one
// contiguous range per column.
template <typename Real>
__device__ __forceinline__ void ens_own_range(const Cells<Real>& C, const EnsTile& T, int col,
                                              int& beg, int& end) {
    const int rows = min(T.rows, C.ncy - T.cy0);
    const int c0 = (T.cx + col) * C.ncy + T.cy0;
    beg = (int)C.start[c0];
    end = (int)C.start[c0 + rows];
}

template <typename Real>
__global__ __launch_bounds__(256) void k_ens_bind_edges(const EnsTile* __restrict__ tiles,
                                                        int ntiles,
                                                        const EnsSeg* __restrict__ segs,
                                                        EnsPool<Real> Q, int ra, int rb,
                                                        Real contact2, int rxi,
                                                        unsigned long long* ekey, int32_t* ea,
                                                        int32_t* eb, unsigned int* ne,
                                                        uint32_t gstep) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    const Post<Real> P = ens_post_bonds(Q, g);
    const size_t base = 2 * (size_t)C.start[0];
    const Draws dr{nullptr, 0, 0, 0};
    for (int col = 0; col < T.ncol; ++col) {
        int beg, end;
        ens_own_range(C, T, col, beg, end);
        for (int t = beg + (int)threadIdx.x; t < end; t += (int)blockDim.x)
            bind_edges_one(C, P, t, ra, rb, contact2, (Real)g.prob[rxi], dr, ens_rng(g),
                           gstep - g.step0, (uint32_t)rxi, ekey + base, ea + base, eb + base,
                           ne + 2 * T.r, (unsigned int)(2 * g.n));
    }
}

template <typename Real>
__global__ __launch_bounds__(1024) void k_ens_bind_match(const EnsSeg* __restrict__ segs,
                                                         EnsPool<Real> Q, int ra, int rb,
                                                         int prod,
                                                         const unsigned long long* ekey,
                                                         const int32_t* ea, const int32_t* eb,
                                                         unsigned int* ne,
                                                         unsigned long long* best,
                                                         unsigned long long* ctr) {
    const int r = (int)blockIdx.x;
    const EnsSeg& g = segs[r];
    if (!g.active) return;
    const size_t base = 2 * (size_t)Q.start[g.coff];
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    bind_match_block(ens_post_bonds(Q, g), Q.id, ra, rb, prod, ekey + base, ea + base,
                     eb + base, ne + 2 * r, (unsigned int)(2 * g.n), best, ctr + 8 * r, G,
                     Q.n_grids);
}

template <typename Real>
__global__ __launch_bounds__(256) void k_ens_dissoc(const EnsTile* __restrict__ tiles, int ntiles,
                                                    const EnsSeg* __restrict__ segs,
                                                    EnsPool<Real> Q, int prod, int ra, int rb,
                                                    int rxi, unsigned long long* ctr,
                                                    uint32_t gstep) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    const Post<Real> P = ens_post_bonds(Q, g);
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    const Draws dr{nullptr, 0, 0, 0};
    for (int col = 0; col < T.ncol; ++col) {
        int beg, end;
        ens_own_range(C, T, col, beg, end);
        for (int t0 = beg; t0 < end; t0 += (int)blockDim.x) {  // uniform
            const int t = t0 + (int)threadIdx.x;
            const bool hit = t < end && dissoc_one(t, Q.id, P, prod, ra, rb, (Real)g.prob2[rxi],
                                                   dr, ens_rng(g), gstep - g.step0,
                                                   (uint32_t)rxi, g.kbase, G, Q.n_grids);
            wave_count(&ctr[8 * T.r + 1], hit);
        }
    }
}

// Autocatalysis (enzyme == product) in the batch: each replica's
// pair-list offsets over its slot range, its accepted pairs in its own
// region of the edge buffers, one workgroup per replica resolving the
// conversion order -- the single-membrane bodies on the replica's views.
template <typename Real>
__global__ __launch_bounds__(256) void k_ens_pair_count(const EnsTile* __restrict__ tiles,
                                                        int ntiles,
                                                        const EnsSeg* __restrict__ segs,
                                                        EnsPool<Real> Q, uint64_t* pcnt) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    const Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    for (int col = 0; col < T.ncol; ++col) {
        int beg, end;
        ens_own_range(C, T, col, beg, end);
        for (int t = beg + (int)threadIdx.x; t < end; t += (int)blockDim.x)
            pcnt[C.id[t] + C.poff_base] = pair_count_one(C, t);
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void k_ens_autocat_edges(const EnsTile* __restrict__ tiles,
                                                           int ntiles,
                                                           const EnsSeg* __restrict__ segs,
                                                           EnsPool<Real> Q,
                                                           const CatArgs<Real>* __restrict__ cats,
                                                           int n_runs, int ci,
                                                           const uint64_t* poff,
                                                           unsigned long long* ekey, int32_t* ea,
                                                           int32_t* eb, unsigned int* ne,
                                                           uint32_t gstep) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const EnsTile T = tiles[tile];
    const EnsSeg& g = segs[T.r];
    Cells<Real> C = ens_cells<Real>(g, Q.x, Q.y, Q.sp, Q.fl, Q.id, Q.start);
    C.poff = poff;
    const CatArgs<Real>& A = cats[(size_t)T.r * n_runs + ci];
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    const size_t base = 2 * (size_t)C.start[0];
    for (int col = 0; col < T.ncol; ++col) {
        int beg, end;
        ens_own_range(C, T, col, beg, end);
        for (int t = beg + (int)threadIdx.x; t < end; t += (int)blockDim.x)
            autocat_edges_one(C, ens_post(Q), A, gstep - g.step0, G, t, ekey + base, ea + base,
                              eb + base, ne + 2 * T.r, (unsigned int)(2 * g.n));
    }
}

template <typename Real>
__global__ __launch_bounds__(1024) void k_ens_autocat_resolve(const EnsSeg* __restrict__ segs,
                                                              EnsPool<Real> Q,
                                                              const CatArgs<Real>* __restrict__ cats,
                                                              int n_runs, int ci,
                                                              const unsigned long long* ekey,
                                                              const int32_t* ea, const int32_t* eb,
                                                              unsigned int* ne,
                                                              unsigned long long* tconv,
                                                              unsigned long long* ctr) {
    const int r = (int)blockIdx.x;
    const EnsSeg& g = segs[r];
    if (!g.active) return;
    const size_t base = 2 * (size_t)Q.start[g.coff];
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    autocat_resolve_block(ens_post(Q), cats[(size_t)r * n_runs + ci], ekey + base, ea + base,
                          eb + base, ne + 2 * r, (unsigned int)(2 * g.n), tconv, ctr + 8 * r,
                          G, Q.n_grids);
}

// Index-order recruitment cap of one recruitment reaction, one block per
// replica slot.
template <typename Real>
__global__ __launch_bounds__(1024) void k_ens_recruit(const EnsSeg* __restrict__ segs,
                                                      EnsPool<Real> Q, int rxi, int target,
                                                      const int32_t* cid, const int32_t* cslot,
                                                      const double* cxo, const double* cyo,
                                                      unsigned int* nc, unsigned long long* ctr) {
    const int r = blockIdx.x;
    const EnsSeg& g = segs[r];
    if (!g.active) return;
    const int pbeg = (int)Q.start[g.coff];
    const Grids<Real> G = ens_grids<Real>(g, Q.counts, Q.delta, Q.ns);
    recruit_apply_block(ens_post(Q), target, g.maxr[rxi], g.kbase, g.kbase + g.n - 1, cid + pbeg,
                        cslot + pbeg, cxo + pbeg, cyo + pbeg, nc + r, ctr + 8 * r, G,
                        Q.n_grids);
}

// Adaptation field of every replica (update_adaptation_field,
// dynamics.py:1094-1112): block (x, r) updates cells x * 256 .. of replica
// r's field from its sensor counts, as k_adapt does for one membrane.
template <typename Real>
__global__ void k_ens_adapt(const EnsSeg* __restrict__ segs, int32_t* counts, int grid,
                            int species, double rate, int ns) {
    const EnsSeg& g = segs[blockIdx.y];
    if (!g.active || !g.w_adapt) return;
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= g.gdims[2 * grid] * g.gdims[2 * grid + 1]) return;
    const Grids<Real> G = ens_grids<Real>(g, counts, nullptr, ns);
    const double rho = (double)density_of<Real>(G.counts[G.off[grid * G.ns + species] + c],
                                                G.inv[grid]);
    g.w_adapt[c] += rate * (rho - g.w_adapt[c]);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

// A model feedback rule bound to replica g's grid (host side of ens_fb).
static FbDev ens_fb_host(const edm_feedback& f, const EnsSeg& g, int ns) {
    FbDev d = to_fb_model(f);
    if (f.sensor >= 0 && f.grid >= 0) {
        d.cbase = g.goff[f.grid * ns + f.sensor];
        d.gnx = g.gdims[2 * f.grid];
        d.gny = g.gdims[2 * f.grid + 1];
        d.gsx = g.gs_d[2 * f.grid];
        d.gsy = g.gs_d[2 * f.grid + 1];
        d.ginv = g.ginv_d[f.grid];
    }
    return d;
}

// Pass group gi of the model with replica g's probabilities and grids.
// CDFs of a run's density-biased recruitment, one workgroup per replica slot
// (each from its own counts on its own grid, as k_recruit_cdf does for one
// membrane); slots without one (free, or no such pass) do nothing.
template <typename Real>
__global__ __launch_bounds__(1024) void k_ens_recruit_cdf(const PassGroupDev* __restrict__ pg,
                                                          int n_groups, int gi,
                                                          const int32_t* __restrict__ counts) {
    const PassGroupDev& grp = pg[(size_t)blockIdx.x * n_groups + gi];
    if (grp.n <= 0) return;  // uniform: a free slot's table is empty
    const PassDev& d = grp.p[grp.n - 1];
    if (!d.cdf) return;
    recruit_cdf_block<Real>(counts + d.f1.cbase, d.f1.gnx * d.f1.gny, (Real)d.f1.ginv,
                            d.cdf);
}

static PassGroupDev ens_group_host(const edm_ensemble* e, int gi, const EnsSeg& g) {
    const std::vector<int>& ix = e->pass_groups[gi];
    const int ns = e->p.n_species;
    PassGroupDev grp{};
    grp.n = (int)ix.size();
    for (int k = 0; k < grp.n; ++k) {
        const edm_reaction& q = e->rx[ix[k]];
        PassDev& d = grp.p[k];
        d.kind = q.kind;
        d.a = q.a;
        d.b = q.b;
        d.rxi = ix[k];
        d.prob = g.prob[ix[k]];
        d.f1 = ens_fb_host(q.fb1, g, ns);
        d.f2 = ens_fb_host(q.fb2, g, ns);
        d.draw_off = 0;
        d.deferred = e->pass_deferred[ix[k]];
        uint32_t reach = 1u << q.a;
        for (int j = k - 1; j >= 0; --j) {
            const edm_reaction& qj = e->rx[ix[j]];
            if (qj.kind == EDM_RX_CONVERT && ((reach >> qj.b) & 1u)) reach |= 1u << qj.a;
        }
        d.reach = reach;
        grp.species |= reach;
        if (q.kind == EDM_RX_RECRUIT && q.use_adapt && g.rcdf) {
            d.cdf = g.rcdf;
            d.sigma = q.contact;
        }
    }
    return grp;
}

// Catalysis run (start, nf) of the model with replica g's probabilities,
// feedback grid and seed (the step is set by the kernel).
template <typename Real>
static CatArgs<Real> ens_cat_host(const edm_ensemble* e, int start, int nf, const EnsSeg& g) {
    const edm_reaction& r = e->rx[start];
    CatArgs<Real> A{};
    A.n_enz = nf;
    for (int j = 0; j < nf; ++j) {
        const edm_reaction& rj = e->rx[start + j];
        A.enz[j] = rj.a;
        A.contact2[j] = (Real)rj.contact * (Real)rj.contact;
        A.prob[j] = (Real)g.prob[start + j];
        A.rxi[j] = (uint32_t)(start + j);
        A.doff[j] = 0;
    }
    A.sub = r.b;
    A.prod = r.c;
    A.fb = ens_fb_host(r.fb1, g, e->p.n_species);
    A.w = nullptr;
    A.wnx = A.wny = 1;
    if (r.use_adapt && e->p.adapt_grid >= 0 && g.w_adapt) {
        A.w = g.w_adapt;
        A.wnx = g.gdims[2 * e->p.adapt_grid];
        A.wny = g.gdims[2 * e->p.adapt_grid + 1];
    }
    A.Kw = (Real)e->p.adapt_K;
    A.nw = (Real)e->p.adapt_n;
    A.dr = Draws{nullptr, 0, 0, 0};
    A.rng = Rng{g.seed_lo, g.seed_hi};
    A.step = 0;
    A.deferred = r.fb1.sensor >= 0 && (r.fb1.sensor == r.b || r.fb1.sensor == r.c);
    A.by_substrate = cat_by_substrate(e->seg_alive[&g - e->seg.data()], A.enz, nf, r.b);
    A.group_shift = __builtin_ctz((unsigned)e->cat_lanes);
    return A;
}

template <typename Real>
static int ens_tables_t(edm_ensemble* e) {
    const size_t nr = e->cat_runs.size(), ng = e->pass_groups.size();
    std::vector<CatArgs<Real>> cats(std::max<size_t>(1, (size_t)e->rmax * nr));
    std::vector<PassGroupDev> pg(std::max<size_t>(1, (size_t)e->rmax * ng));
    for (int r = 0; r < e->rmax; ++r) {
        const EnsSeg& g = e->seg[r];
        if (!g.active) continue;
        for (size_t c = 0; c < nr; ++c)
            cats[(size_t)r * nr + c] =
                ens_cat_host<Real>(e, e->cat_runs[c].first, e->cat_runs[c].second, g);
        for (size_t k = 0; k < ng; ++k) pg[(size_t)r * ng + k] = ens_group_host(e, (int)k, g);
    }
    const size_t cb = cats.size() * sizeof(CatArgs<Real>), pb = pg.size() * sizeof(PassGroupDev);
    if (cb > e->cats_bytes) {
        hipFree(e->d_cats);
        EDM_HIP(hipMalloc(&e->d_cats, cb));
        e->cats_bytes = cb;
    }
    if (pb > e->pgrp_bytes) {
        hipFree(e->d_pgrp);
        EDM_HIP(hipMalloc(&e->d_pgrp, pb));
        e->pgrp_bytes = pb;
    }
    EDM_HIP(hipMemcpyAsync(e->d_cats, cats.data(), cb, hipMemcpyHostToDevice, e->stream));
    EDM_HIP(hipMemcpyAsync(e->d_pgrp, pg.data(), pb, hipMemcpyHostToDevice, e->stream));
    EDM_HIP(hipStreamSynchronize(e->stream));
    return 0;
}

template <typename Real>
static int ens_create_t(edm_ensemble* e) {
    const edm_params& p = e->p;
    const int ns = p.n_species;
    const int64_t cap = e->cap;
    const size_t rb = sizeof(Real);
    void** rbufs[6] = {&e->a.x, &e->a.y, &e->a.xn, &e->a.yn, &e->a.xt, &e->a.yt};
    for (auto b : rbufs) EDM_HIP(hipMalloc(b, (size_t)cap * rb + 16));
    if (dalloc(&e->a.sp, cap + 4) || dalloc(&e->a.sp_t, cap + 4) || dalloc(&e->a.fl, cap + 4) ||
        dalloc(&e->a.fl_t, cap + 4) || dalloc(&e->a.id, cap) || dalloc(&e->a.id_t, cap) ||
        dalloc(&e->a.ctmc, cap + 4) || dalloc(&e->a.ctmc_t, cap + 4) ||
        dalloc(&e->a.dye, cap + 4) || dalloc(&e->a.dye_t, cap + 4) || dalloc(&e->rep, cap) ||
        dalloc(&e->rep_t, cap) || dalloc(&e->key, cap) || dalloc(&e->rank, cap) ||
        dalloc(&e->perm, cap) || dalloc(&e->cand_id, cap) || dalloc(&e->cand_slot, cap) ||
        dalloc(&e->cand_x, cap) || dalloc(&e->cand_y, cap) || dalloc(&e->n_cand, e->rmax) ||
        dalloc(&e->ctr, 8 * (size_t)e->rmax) || dalloc(&e->frame_out, 4) ||
        dalloc(&e->noise_ctr, 2 * (size_t)e->rmax) ||
        dalloc(&e->d_seg, e->rmax))
        return EDM_EHIP;
    for (const auto& r : e->rx) {
        e->has_bind = e->has_bind || r.kind == EDM_RX_BIND;
        e->has_autocat = e->has_autocat || (r.kind == EDM_RX_CAT && r.a == r.c);
    }
    if (e->has_bind || e->has_autocat) {
        if (dalloc(&e->ekey, 2 * cap) || dalloc(&e->ea, 2 * cap) || dalloc(&e->eb, 2 * cap) ||
            dalloc(&e->ne, 2 * (size_t)e->rmax))
            return EDM_EHIP;
        EDM_HIP(hipMemset(e->ne, 0, 2 * (size_t)e->rmax * sizeof(unsigned int)));
    }
    if (e->has_bind) {
        if (dalloc(&e->a.bond, cap) || dalloc(&e->a.bond_t, cap) || dalloc(&e->best, cap))
            return EDM_EHIP;
        EDM_HIP(hipMemset(e->a.bond, 0xff, (size_t)cap * sizeof(int32_t)));
    }
    if (e->has_autocat) {
        if (dalloc(&e->pcnt, cap) || dalloc(&e->poff, cap) || dalloc(&e->tconv, cap))
            return EDM_EHIP;
        EDM_HIP(hipMemset(e->tconv, 0xff, (size_t)cap * sizeof(unsigned long long)));
        EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, e->cub_bytes64, e->pcnt, e->poff,
                                                 (int)cap));
        EDM_HIP(hipMalloc(&e->cub_tmp64, std::max<size_t>(e->cub_bytes64, 16)));
    }
    // each replica's region of a list is padded to whole waves
    e->wl_cap = cap + 64 * (int64_t)e->rmax;
    if (dalloc(&e->wl_pass, e->wl_cap) || dalloc(&e->wl_dye, e->wl_cap) ||
        dalloc(&e->wl_wrep_pass, e->wl_cap / 64) || dalloc(&e->wl_wrep_dye, e->wl_cap / 64) ||
        dalloc(&e->wl_rcnt, 2 * (size_t)e->rmax))
        return EDM_EHIP;
    EDM_HIP(hipMemset(e->wl_rcnt, 0, 2 * (size_t)e->rmax * sizeof(uint32_t)));
    EDM_HIP(hipMemset(e->n_cand, 0, (size_t)e->rmax * sizeof(unsigned int)));
    EDM_HIP(hipMemset(e->ctr, 0, (size_t)e->rmax * 8 * sizeof(unsigned long long)));
    EDM_HIP(hipMemset(e->noise_ctr, 0, (size_t)e->rmax * 2 * sizeof(unsigned long long)));
    int rc = 0;
    rc |= upload_real<Real>(&e->sig, p.sigma, (size_t)ns * ns);
    rc |= upload_real<Real>(&e->eps, p.epsilon, (size_t)ns * ns);
    rc |= upload_real<Real>(&e->kh, p.k_harm, (size_t)ns * ns);
    rc |= upload_real<Real>(&e->Dbase, p.D_base, ns);
    if (rc) return EDM_EHIP;
    double r2 = 0.0;
    for (int k = 0; k < ns * ns; ++k) {
        const double sg = p.sigma[k];
        if (!(sg > 0.0)) continue;
        if (p.use_wca && p.epsilon[k] > 0.0) r2 = std::max(r2, sg * sg * 1.2599210498948732);
        if (p.use_harmonic && p.k_harm[k] > 0.0) r2 = std::max(r2, sg * sg);
    }
    e->rint2 = r2 * (1.0 + 1e-6);
    // kept (grid, sensor) pairs, in grid-major order as for one membrane
    const int ng = (int)e->grids.size();
    std::vector<int> keep((size_t)std::max(ng, 1) * ns, 0);
    for (auto& r : e->rx) {
        if (r.fb1.sensor >= 0) keep[(size_t)r.fb1.grid * ns + r.fb1.sensor] = 1;
        if (r.fb2.sensor >= 0) keep[(size_t)r.fb2.grid * ns + r.fb2.sensor] = 1;
    }
    if (p.adapt_grid >= 0) keep[(size_t)p.adapt_grid * ns + p.adapt_species] = 1;
    e->kept.clear();
    for (int g = 0; g < ng; ++g)
        for (int s = 0; s < ns; ++s)
            if (keep[(size_t)g * ns + s]) {
                e->kept.emplace_back(g, s);
                e->sensor_mask_all |= 1u << s;
            }
    plan_passes(e->rx, e->sensor_mask_all, e->pass_groups, e->group_start, e->pass_deferred, 1);
    e->pass_species = pass_species_of(e->rx, e->pass_groups);
    e->seg_npass.assign(e->rmax, 0);
    e->seg_ndye.assign(e->rmax, 0);
    e->seg_alive.assign(e->rmax, {});
    // catalysis runs: back-to-back reactions with one substrate and product and
    // no feedback share a launch (as for one membrane)
    e->cat_runs.clear();
    e->cat_list_ok.clear();
    e->cat_run_of.assign(e->rx.size(), -1);
    for (size_t ri = 0; ri < e->rx.size();) {
        const edm_reaction& r = e->rx[ri];
        if (r.kind != EDM_RX_CAT) {
            ++ri;
            continue;
        }
        // autocatalysis depends on the order of the reactions' pairs: never fused
        const bool plain = r.fb1.sensor < 0 && !r.use_adapt && r.a != r.c;
        int nf = 1;
        while (plain && nf < kCatFuse && ri + nf < e->rx.size()) {
            const edm_reaction& r2 = e->rx[ri + nf];
            if (r2.kind != EDM_RX_CAT || r2.b != r.b || r2.c != r.c || r2.fb1.sensor >= 0 ||
                r2.use_adapt || r2.a == r2.c)
                break;
            ++nf;
        }
        // the list walk needs every enzyme in the pass work list and no
        // enzyme that the run can turn into another species
        bool listable = true;
        for (int j = 0; j < nf; ++j) {
            const int a = e->rx[ri + j].a;
            listable = listable && ((e->pass_species >> a) & 1u) && a != r.b && a != r.c;
        }
        e->cat_run_of[ri] = (int)e->cat_runs.size();
        e->cat_runs.emplace_back((int)ri, nf);
        e->cat_list_ok.push_back(listable ? 1 : 0);
        ri += nf;
    }
    e->seg.assign(e->rmax, EnsSeg{});
    e->fresh_beg.assign(e->rmax, -1);
    EDM_HIP(hipMemset(e->d_seg, 0, (size_t)e->rmax * sizeof(EnsSeg)));
    return 0;
}

int ens_create(edm_ensemble* e) { return e->fp64 ? ens_create_t<double>(e) : ens_create_t<float>(e); }

// Offsets of the live replicas in the cell-key space and the count pool,
// the strip table, and buffer growth.  Runs between steps only.
int ens_layout(edm_ensemble* e) {
    const int ns = e->p.n_species;
    // strip height: about strip_parts (192 by default) own particles per
    // 256-thread workgroup at the live replicas' mean density, with the
    // staged 3 x (rows + 2) cells well inside the LDS capacity (1024
    // particles); results do not depend on it
    e->tile_rows = 1;
    e->tile_cols = 1;
    int64_t coff = 0, cnt = 0, tcoff = 0;
    e->tiles.clear();
    std::vector<int32_t> wrep[2];
    for (int r = 0; r < e->rmax; ++r) {
        EnsSeg& g = e->seg[r];
        if (!g.active) continue;
        const int64_t nl[2] = {e->seg_npass[r], e->seg_ndye[r]};
        for (int L = 0; L < 2; ++L) {
            g.wl_off[L] = (int32_t)(64 * wrep[L].size());
            g.wl_n[L] = (int32_t)nl[L];
            wrep[L].insert(wrep[L].end(), (size_t)((nl[L] + 63) / 64), (int32_t)r);
        }
        g.coff = (int32_t)coff;
        coff += (int64_t)g.ncx * g.ncy;
        if (e->prox_map >= 0) {
            g.tg_off = (int32_t)tcoff;
            tcoff += (int64_t)g.tg_ncx * g.tg_ncy;
        }
        for (int k = 0; k < kEnsMaxGrids * kEnsMaxSpecies; ++k) g.goff[k] = -1;
        for (auto& kp : e->kept) {
            g.goff[kp.first * ns + kp.second] = (int32_t)cnt;
            cnt += (int64_t)g.gdims[2 * kp.first] * g.gdims[2 * kp.first + 1];
        }
        // strips of about strip_parts own particles: a column that holds
        // fewer than pair_below is taken together with its neighbour (boxes
        // of 4 columns or more), a long one is cut into equal runs of rows;
        // at most ~700 particles (1024 with room for fluctuations) staged
        // per strip
        {
            const double dens = std::max((double)g.n / ((double)g.ncx * g.ncy), 1e-3);
            const double col_parts = dens * g.ncy;
            const int ncol = (col_parts < (double)e->pair_below && g.ncx >= 4) ? 2 : 1;
            int ty = std::max(1, (int)std::lround(col_parts * ncol / (double)e->strip_parts));
            ty = std::max(ty, (int)std::ceil((ncol + 2) * (g.ncy + 2) * dens / 700.0));
            ty = std::max(ty, (g.ncy + kTileRows - 1) / kTileRows);
            ty = std::min(ty, g.ncy);
            const int rows = (g.ncy + ty - 1) / ty;
            e->tile_rows = std::max(e->tile_rows, rows);
            e->tile_cols = std::max(e->tile_cols, ncol);
            for (int cx = 0; cx < g.ncx; cx += ncol)
                for (int cy0 = 0; cy0 < g.ncy; cy0 += rows)
                    e->tiles.push_back(EnsTile{r, cx, cy0, rows, std::min(ncol, g.ncx - cx)});
        }
    }
    if (coff >= (int64_t)INT32_MAX - 2 || cnt >= (int64_t)INT32_MAX) {
        set_error("edm_ensemble: too many cells for 32-bit keys");
        return EDM_ECAPACITY;
    }
    e->n_cells = coff;
    if (e->prox_map >= 0) {
        if (tcoff >= (int64_t)INT32_MAX - 2) {
            set_error("edm_ensemble: too many proximity-FRET target cells");
            return EDM_ECAPACITY;
        }
        e->n_tcells = tcoff;
        if (tcoff + 1 > e->tcell_cap) {
            hipFree(e->tg_cnt);
            hipFree(e->tg_fill);
            hipFree(e->tg_start);
            hipFree(e->cub_tmp_tg);
            e->tg_cnt = e->tg_fill = e->tg_start = nullptr;
            e->cub_tmp_tg = nullptr;
            e->tcell_cap = std::max<int64_t>(tcoff + 1, e->tcell_cap * 3 / 2);
            if (dalloc(&e->tg_cnt, e->tcell_cap) || dalloc(&e->tg_fill, e->tcell_cap) ||
                dalloc(&e->tg_start, e->tcell_cap))
                return EDM_EHIP;
            e->cub_bytes_tg = 0;
            EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, e->cub_bytes_tg, e->tg_cnt,
                                                     e->tg_start, (int)e->tcell_cap));
            EDM_HIP(hipMalloc(&e->cub_tmp_tg, std::max<size_t>(e->cub_bytes_tg, 16)));
        }
    }
    e->n_counts = cnt;
    if (coff + 2 > e->cell_cap) {
        hipFree(e->cell_cnt);
        hipFree(e->cell_start);
        hipFree(e->cub_tmp);
        e->cell_cap = std::max<int64_t>(coff + 2, e->cell_cap * 3 / 2);
        if (dalloc(&e->cell_cnt, e->cell_cap) || dalloc(&e->cell_start, e->cell_cap))
            return EDM_EHIP;
        EDM_HIP(hipMemset(e->cell_cnt, 0, (size_t)e->cell_cap * sizeof(uint32_t)));
        size_t b = 0;
        EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, e->cell_cnt, e->cell_start,
                                                 (int)e->cell_cap, e->stream));
        e->cub_bytes = b + 256;
        EDM_HIP(hipMalloc(&e->cub_tmp, e->cub_bytes));
    }
    if (cnt > e->count_cap) {
        hipFree(e->counts);
        hipFree(e->delta);
        e->count_cap = std::max<int64_t>(cnt, e->count_cap * 3 / 2);
        if (dalloc(&e->counts, e->count_cap) || dalloc(&e->delta, e->count_cap)) return EDM_EHIP;
        EDM_HIP(hipMemset(e->delta, 0, (size_t)e->count_cap * sizeof(int32_t)));
        EDM_HIP(hipMemset(e->counts, 0, (size_t)e->count_cap * sizeof(int32_t)));
    }
    if ((int64_t)e->tiles.size() > e->tile_cap) {
        hipFree(e->d_tiles);
        e->tile_cap = std::max<int64_t>((int64_t)e->tiles.size(), e->tile_cap * 3 / 2);
        if (dalloc(&e->d_tiles, e->tile_cap)) return EDM_EHIP;
    }
    if (!e->tiles.empty())
        EDM_HIP(hipMemcpyAsync(e->d_tiles, e->tiles.data(), e->tiles.size() * sizeof(EnsTile),
                               hipMemcpyHostToDevice, e->stream));
    // chunks of the live replicas' slot ranges for the list build: after the
    // sort the live replicas occupy consecutive slot ranges in replica order,
    // whose lengths (alive and dormant particles) do not change
    e->wl_chunks.clear();
    for (int64_t r = 0, beg = 0; r < e->rmax; ++r) {
        if (!e->seg[r].active) continue;
        const int64_t end = beg + e->seg[r].n;
        if (e->seg_npass[r] + e->seg_ndye[r] > 0)
            for (int64_t b = beg; b < end; b += kWlSlots)
                e->wl_chunks.push_back(EnsChunk{(int32_t)r, (int32_t)b,
                                                (int32_t)std::min<int64_t>(end, b + kWlSlots), 0});
        beg = end;
    }
    if ((int64_t)e->wl_chunks.size() > e->wl_chunk_cap) {
        hipFree(e->d_wl_chunks);
        e->wl_chunk_cap = std::max<int64_t>((int64_t)e->wl_chunks.size(), e->wl_chunk_cap * 3 / 2);
        if (dalloc(&e->d_wl_chunks, e->wl_chunk_cap)) return EDM_EHIP;
    }
    if (!e->wl_chunks.empty())
        EDM_HIP(hipMemcpyAsync(e->d_wl_chunks, e->wl_chunks.data(),
                               e->wl_chunks.size() * sizeof(EnsChunk), hipMemcpyHostToDevice,
                               e->stream));
    e->wl_npass = 64 * (int64_t)wrep[0].size();
    e->wl_ndye = 64 * (int64_t)wrep[1].size();
    if (e->wl_npass > e->wl_cap || e->wl_ndye > e->wl_cap) {
        set_error("edm_ensemble: work lists exceed their capacity");
        return EDM_ECAPACITY;
    }
    int32_t* wdev[2] = {e->wl_wrep_pass, e->wl_wrep_dye};
    for (int L = 0; L < 2; ++L)
        if (!wrep[L].empty())
            EDM_HIP(hipMemcpyAsync(wdev[L], wrep[L].data(), wrep[L].size() * sizeof(int32_t),
                                   hipMemcpyHostToDevice, e->stream));
    EDM_HIP(hipMemcpyAsync(e->d_seg, e->seg.data(), (size_t)e->rmax * sizeof(EnsSeg),
                           hipMemcpyHostToDevice, e->stream));
    EDM_HIP(hipStreamSynchronize(e->stream));
    if (int rc = e->fp64 ? ens_tables_t<double>(e) : ens_tables_t<float>(e)) return rc;
    e->layout_dirty = false;
    return 0;
}

template <typename Real>
static int ens_step_t(edm_ensemble* e) {
    if (e->layout_dirty)
        if (int rc = ens_layout(e)) return rc;
    hipStream_t st = e->stream;
    const edm_params& p = e->p;
    const int n = (int)e->n_total;
    const int ntiles = (int)e->tiles.size();
    const uint32_t sentinel = (uint32_t)e->n_cells;
    if (n > 0) {
        const int64_t ncnt = e->n_cells + 2;  // live keys, sentinel, scan end
        hipLaunchKernelGGL(k_ens_bin<Real>,
                           dim3(nblk(std::max<int64_t>({n, e->n_counts, 2 * (int64_t)e->rmax}))),
                           dim3(kBlock), 0, st, n, (const Real*)e->a.x, (const Real*)e->a.y,
                           (const uint16_t*)e->rep, (const EnsSeg*)e->d_seg, sentinel, e->key,
                           e->rank, e->cell_cnt, e->counts, (int)e->n_counts, e->wl_rcnt,
                           2 * e->rmax);
        size_t tb = e->cub_bytes;
        EDM_HIP(hipcub::DeviceScan::ExclusiveSum(e->cub_tmp, tb, e->cell_cnt, e->cell_start,
                                                 (int)ncnt, st));
        hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(kBlock), 0, st, n, e->key, e->rank,
                           e->cell_start, e->perm);
        // live cells only: the sentinel bucket holds every particle of the
        // removed replicas, in no particular order, and is just cleared
        if (e->n_cells > 0)
            hipLaunchKernelGGL(k_sort_cells, dim3(nblk(e->n_cells)), dim3(kBlock), 0, st,
                               (int)e->n_cells, e->cell_start, e->perm, e->cell_cnt,
                               (const int32_t*)e->a.id, 1);
        else
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, 1,
                               (unsigned int*)(e->cell_cnt + e->n_cells));
        hipLaunchKernelGGL((e->has_bind ? k_ens_gather<Real, true> : k_ens_gather<Real, false>),
                           dim3(nblk(n)), dim3(kBlock), 0, st, n, (const uint32_t*)e->perm, e->a,
                           (const uint16_t*)e->rep, e->rep_t, (const EnsSeg*)e->d_seg);
        // the gather wrote the new order into the _t arrays (swapped below)
        if (!e->wl_chunks.empty())
            hipLaunchKernelGGL(k_ens_wl_build, dim3((unsigned)e->wl_chunks.size()), dim3(256), 0,
                               st, (const EnsChunk*)e->d_wl_chunks, (const EnsSeg*)e->d_seg,
                               (const int8_t*)e->a.sp_t, (const uint8_t*)e->a.fl_t,
                               e->pass_species, e->wl_pass, e->wl_dye, e->wl_rcnt);
        std::swap(e->a.x, e->a.xt);
        std::swap(e->a.y, e->a.yt);
        std::swap(e->a.sp, e->a.sp_t);
        std::swap(e->a.fl, e->a.fl_t);
        std::swap(e->a.id, e->a.id_t);
        std::swap(e->a.ctmc, e->a.ctmc_t);
        std::swap(e->a.dye, e->a.dye_t);
        std::swap(e->a.bond, e->a.bond_t);
        std::swap(e->rep, e->rep_t);
    }
    // particles of removed replicas now sit after every live one; their
    // slots may be reused, and every replica's particles are cell-sorted
    e->n_total -= e->n_retired;
    e->n_retired = 0;
    e->draining.clear();
    std::fill(e->fresh_beg.begin(), e->fresh_beg.end(), (int64_t)-1);
    if (ntiles == 0) {
        e->gstep += 1;
        return 0;
    }
    EnsPool<Real> Q;
    Q.x = (const Real*)e->a.x;
    Q.y = (const Real*)e->a.y;
    Q.xn = (Real*)e->a.xn;
    Q.yn = (Real*)e->a.yn;
    Q.sp = e->a.sp;
    Q.fl = e->a.fl;
    Q.id = e->a.id;
    Q.start = e->cell_start;
    Q.counts = e->counts;
    Q.delta = e->delta;
    Q.ns = p.n_species;
    Q.n_grids = (int)e->grids.size();
    Q.bond = e->a.bond;

    ForceParams<Real> F;
    F.sig = (const Real*)e->sig;
    F.eps = (const Real*)e->eps;
    F.kh = (const Real*)e->kh;
    F.Dbase = (const Real*)e->Dbase;
    F.ns = p.n_species;
    F.wca = p.use_wca != 0;
    F.harm = p.use_harmonic != 0;
    F.dfield = p.dfield;
    F.n_regions = 0;
    F.rcx = F.rcy = F.rr = F.rf = nullptr;
    F.smooth_w = (Real)p.smooth_w;
    F.sin_amp = (Real)p.sin_amp;
    F.sin_kx = F.sin_ky = (Real)0;
    F.dt = (Real)p.dt;
    F.max_disp = (Real)p.max_disp;
    F.rint2 = (Real)e->rint2;
    const dim3 tgrid(8 * ((ntiles + 7) / 8));
    const uint32_t gstep = e->gstep;
    bool any_sub = false;
    for (int r = 0; r < e->rmax; ++r) any_sub |= e->seg[r].active && e->seg[r].sub_rows;
    if (e->has_autocat) {
        // the reference's pair-list offsets of every replica (its order of
        // the pairs), over the pre-move positions
        hipLaunchKernelGGL(k_ens_pair_count<Real>, tgrid, dim3(256), 0, st,
                           (const EnsTile*)e->d_tiles, ntiles, (const EnsSeg*)e->d_seg, Q,
                           e->pcnt);
        size_t tb = e->cub_bytes64;
        EDM_HIP(hipcub::DeviceScan::ExclusiveSum(e->cub_tmp64, tb, e->pcnt, e->poff,
                                                 (int)e->n_total, st));
    }
    if (any_sub)
        hipLaunchKernelGGL((k_ens_force<Real, true>), tgrid, dim3(256),
                           force_sub_lds(e->tile_rows, e->tile_cols), st,
                           (const EnsTile*)e->d_tiles, ntiles,
                           (const EnsSeg*)e->d_seg, Q, F, e->sensor_mask_all, gstep);
    else
        hipLaunchKernelGGL((k_ens_force<Real, false>), tgrid, dim3(256), 0, st,
                           (const EnsTile*)e->d_tiles, ntiles, (const EnsSeg*)e->d_seg, Q, F,
                           e->sensor_mask_all, gstep);

    bool fold_pending = false;  // a catalysis run's deferred updates, not yet folded
    // a catalysis run that reads no counts (no feedback, not autocatalytic)
    // can fold the previous run's deferred updates in its own launch
    auto folds_in = [&](size_t j) {
        if (j >= e->rx.size()) return false;
        const edm_reaction& q = e->rx[j];
        return q.kind == EDM_RX_CAT && q.a != q.c && q.fb1.sensor < 0 && e->cat_run_of[j] >= 0;
    };
    for (size_t ri = 0; ri < e->rx.size(); ++ri) {
        const edm_reaction& r = e->rx[ri];
        if (r.kind == EDM_RX_BIND) {
            const Real c2 = (Real)r.contact * (Real)r.contact;
            hipLaunchKernelGGL(k_ens_bind_edges<Real>, tgrid, dim3(256), 0, st,
                               (const EnsTile*)e->d_tiles, ntiles, (const EnsSeg*)e->d_seg, Q,
                               r.a, r.b, c2, (int)ri, e->ekey, e->ea, e->eb, e->ne, gstep);
            hipLaunchKernelGGL(k_ens_bind_match<Real>, dim3(e->rmax), dim3(1024), 0, st,
                               (const EnsSeg*)e->d_seg, Q, r.a, r.b, r.c,
                               (const unsigned long long*)e->ekey, (const int32_t*)e->ea,
                               (const int32_t*)e->eb, e->ne, e->best, e->ctr);
            hipLaunchKernelGGL(k_ens_dissoc<Real>, tgrid, dim3(256), 0, st,
                               (const EnsTile*)e->d_tiles, ntiles, (const EnsSeg*)e->d_seg, Q,
                               r.c, r.a, r.b, (int)ri, e->ctr, gstep);
        } else if (r.kind == EDM_RX_ADAPT) {
            const int ag = p.adapt_grid;
            int mc = 1;
            for (int q = 0; q < e->rmax; ++q)
                if (e->seg[q].active)
                    mc = std::max(mc, e->seg[q].gdims[2 * ag] * e->seg[q].gdims[2 * ag + 1]);
            hipLaunchKernelGGL(k_ens_adapt<Real>, dim3(nblk(mc), e->rmax), dim3(kBlock), 0, st,
                               (const EnsSeg*)e->d_seg, e->counts, ag, p.adapt_species,
                               p.dt / p.adapt_tau, p.n_species);
        } else if (r.kind == EDM_RX_CAT) {
            const int ci = e->cat_run_of[ri];
            if (ci < 0) continue;  // member of a run launched before
            const int start = e->cat_runs[ci].first;
            if (e->rx[start].a == e->rx[start].c) {
                hipLaunchKernelGGL(k_ens_autocat_edges<Real>, tgrid, dim3(256), 0, st,
                                   (const EnsTile*)e->d_tiles, ntiles, (const EnsSeg*)e->d_seg, Q,
                                   (const CatArgs<Real>*)e->d_cats, (int)e->cat_runs.size(), ci,
                                   (const uint64_t*)e->poff, e->ekey, e->ea, e->eb, e->ne, gstep);
                hipLaunchKernelGGL(k_ens_autocat_resolve<Real>, dim3(e->rmax), dim3(1024), 0, st,
                                   (const EnsSeg*)e->d_seg, Q, (const CatArgs<Real>*)e->d_cats,
                                   (int)e->cat_runs.size(), ci,
                                   (const unsigned long long*)e->ekey, (const int32_t*)e->ea,
                                   (const int32_t*)e->eb, e->ne, e->tconv, e->ctr);
            } else {
                const int fold_n = fold_pending ? (int)e->n_counts : 0;
                fold_pending = false;
                if (e->cat_walk == 1 && e->cat_list_ok[ci] && e->wl_npass > 0)
                    hipLaunchKernelGGL((k_ens_cat_list<Real>), dim3(nblk(e->wl_npass)), dim3(256),
                                       0, st, (int)(e->wl_npass / 64), (const int32_t*)e->wl_pass,
                                       (const int32_t*)e->wl_wrep_pass,
                                       (const uint32_t*)e->wl_rcnt, (const EnsSeg*)e->d_seg, Q,
                                       (const CatArgs<Real>*)e->d_cats, (int)e->cat_runs.size(),
                                       ci, e->ctr, gstep, e->counts, e->delta, fold_n);
                else
                    hipLaunchKernelGGL((k_ens_cat<Real>), tgrid, dim3(256), 0, st,
                                       (const EnsTile*)e->d_tiles, ntiles,
                                       (const EnsSeg*)e->d_seg, Q,
                                       (const CatArgs<Real>*)e->d_cats, (int)e->cat_runs.size(),
                                       ci, e->ctr, gstep, e->counts, e->delta, fold_n);
            }
            const edm_reaction& r0 = e->rx[start];
            const bool defer =
                r0.fb1.sensor >= 0 && (r0.fb1.sensor == r0.b || r0.fb1.sensor == r0.c);
            if (defer && e->n_counts) {
                if (folds_in((size_t)start + (size_t)e->cat_runs[ci].second))
                    fold_pending = true;
                else
                    hipLaunchKernelGGL(k_fold, dim3(nblk(e->n_counts)), dim3(kBlock), 0, st,
                                       (int)e->n_counts, e->counts, e->delta);
            }
        } else if (r.kind == EDM_RX_RECRUIT || r.kind == EDM_RX_DETACH ||
                   r.kind == EDM_RX_CONVERT) {
            if (e->group_start[ri] < 0) continue;
            const int gi = e->group_start[ri];
            const std::vector<int>& grp_ix = e->pass_groups[gi];
            bool defer_any = false;
            for (int k : grp_ix) defer_any = defer_any || e->pass_deferred[k];
            const edm_reaction& last_rx = e->rx[grp_ix.back()];
            const bool local = last_rx.kind == EDM_RX_RECRUIT && last_rx.use_adapt;
            if (local)
                hipLaunchKernelGGL((k_ens_recruit_cdf<Real>), dim3(e->rmax), dim3(1024), 0, st,
                                   (const PassGroupDev*)e->d_pgrp, (int)e->pass_groups.size(), gi,
                                   (const int32_t*)e->counts);
            if (e->wl_npass > 0)
                hipLaunchKernelGGL((local ? k_ens_passes<Real, true> : k_ens_passes<Real, false>),
                                   dim3(nblk(e->wl_npass)), dim3(256), 0,
                                   st, (int)(e->wl_npass / 64), (const int32_t*)e->wl_pass,
                                   (const int32_t*)e->wl_wrep_pass, (const uint32_t*)e->wl_rcnt,
                                   (const EnsSeg*)e->d_seg, Q, (const PassGroupDev*)e->d_pgrp,
                                   (int)e->pass_groups.size(), gi, e->ctr, e->cand_id,
                                   e->cand_slot, e->cand_x, e->cand_y, e->n_cand, gstep);
            const edm_reaction& last = e->rx[grp_ix.back()];
            if (last.kind == EDM_RX_RECRUIT)
                hipLaunchKernelGGL((k_ens_recruit<Real>), dim3(e->rmax), dim3(1024), 0, st,
                                   (const EnsSeg*)e->d_seg, Q, grp_ix.back(), last.a,
                                   (const int32_t*)e->cand_id, (const int32_t*)e->cand_slot,
                                   (const double*)e->cand_x, (const double*)e->cand_y, e->n_cand,
                                   e->ctr);
            if (defer_any && e->n_counts)
                hipLaunchKernelGGL(k_fold, dim3(nblk(e->n_counts)), dim3(kBlock), 0, st,
                                   (int)e->n_counts, e->counts, e->delta);
        }
    }
    if (fold_pending && e->n_counts)
        hipLaunchKernelGGL(k_fold, dim3(nblk(e->n_counts)), dim3(kBlock), 0, st,
                           (int)e->n_counts, e->counts, e->delta);
    std::swap(e->a.x, e->a.xn);
    std::swap(e->a.y, e->a.yn);
    e->gstep += 1;
    EDM_HIP(hipGetLastError());
    return 0;
}

// Binding edges that did not fit a replica's region: the matching of that
// step would differ from the reference's sequential pass, so the run is
// reported as incomplete (EDM_ECAPACITY) and the record cleared.
int ens_check_edges(edm_ensemble* e) {
    std::vector<unsigned int> ne(2 * (size_t)e->rmax);
    EDM_HIP(hipMemcpyAsync(ne.data(), e->ne, ne.size() * sizeof(unsigned int),
                           hipMemcpyDeviceToHost, e->stream));
    EDM_HIP(hipStreamSynchronize(e->stream));
    for (int r = 0; r < e->rmax; ++r)
        if (ne[2 * r + 1] > 0) {
            EDM_HIP(hipMemset(e->ne, 0, ne.size() * sizeof(unsigned int)));
            set_error("edm_ens_run: replica slot " + std::to_string(r) + " accepted " +
                      std::to_string(ne[2 * r + 1]) +
                      " candidate binding edges in a step, more than its edge region holds; "
                      "the outcome of that step is incomplete");
            return EDM_ECAPACITY;
        }
    return 0;
}

int ens_step(edm_ensemble* e, bool image) {
    if (int rc = e->fp64 ? ens_step_t<double>(e) : ens_step_t<float>(e)) return rc;
    if (image && e->has_pp) return ens_photons(e);
    return 0;
}
