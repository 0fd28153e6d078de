// Tiled catalysis pass (dynamics.py:826-1090), included by edm_dynamics.hip
// after the reaction helpers.  Each workgroup stages its strip's
// neighbourhood (pre-move positions, species, flags; the post-move
// positions of the few pairs within the cutoff are read when those pairs
// are evaluated), lists
// the strip's own alive enzymes in LDS and tests every enzyme against the
// substrates of its 3x3 cells, one wavefront per enzyme.  When substrates
// are the rarer side the walk starts from them instead (own alive
// substrates, one wavefront each, against the enzymes of their 3x3 cells):
// the set of pairs, each pair's uniform and hence the outcome are the same
// either way, and every launch uses one side for all strips.  A pair qualifies
// as in the reference: in the (capped) pair list built on pre-move
// positions (d^2 < cutoff^2), within the effective contact radius after the
// move, and accepted by its uniform against P_cat x Hill(rho at the substrate) [x adaptation].  The
// substrate flips with a byte CAS on the global species array, so however
// many of its pairs succeed it converts once, which is the outcome of the
// reference's sequential loop; counts and density updates follow the CAS.

// Up to kCatFuse catalytic reactions that run back to back with the same
// substrate and product and without density feedback or adaptation act as
// one: a substrate converts if any pair of any of them accepts (each pair
// keeps its own uniform per reaction), which is the outcome of running them
// one after another, so they share one launch.  Member j has enzyme enz[j].
static constexpr int kCatFuse = 4;

template <typename Real>
struct CatArgs {
    int n_enz;
    int enz[kCatFuse];
    Real contact2[kCatFuse], prob[kCatFuse];
    uint32_t rxi[kCatFuse];
    int64_t doff[kCatFuse];  // offset of the member's pair uniforms in hook draws
    int sub, prod;
    FbDev fb;                // member 0 only (a fused run has none)
    const double* w;         // adaptation field, member 0 only
    int wnx, wny;
    Real Kw, nw;
    Draws dr;
    Rng rng;
    uint32_t step;
    const uint32_t* step_base;  // when set, the launch's steps count from step_base[0]
    bool deferred;
    bool by_substrate;  // walk the pairs from the substrates' side (the rarer one)
    int group_shift;    // log2 of the lanes that walk one entity's candidates (0..6)
    int32_t* app;       // a later run's work list: converted substrates are appended
    uint32_t* app_cnt;  // (region u / kWlSlots of slot u; nullptr: none)
    int32_t *fold_counts, *fold_delta;  // deferred count updates of the previous run,
    int fold_n;                         // folded by this launch (fold_n > 0)
};

// Appends slot u to its region of a work list (one atomic; the order of a
// region's entries does not matter to the list walks).
__device__ __forceinline__ void wl_append(int32_t* list, uint32_t* cnt, int u) {
    const int b = u / kWlSlots;
    const uint32_t k = atomicAdd(&cnt[b], 1u);
    list[(size_t)b * kWlSlots + k] = u;
}

// Folds the previous run's deferred count updates (counts += delta, delta
// = 0) with the threads of this launch; the launch reads no counts and
// adds its own updates atomically, so the order does not matter.
__device__ __forceinline__ void fold_in_launch(int32_t* counts, int32_t* delta, int n) {
    const int nb = (int)(gridDim.x * gridDim.y);
    const int b = (int)(blockIdx.y * gridDim.x + blockIdx.x);
    for (int i = b * (int)blockDim.x + (int)threadIdx.x; i < n; i += nb * (int)blockDim.x) {
        const int d = delta[i];
        if (d) {
            atomicAdd(&counts[i], d);
            delta[i] = 0;
        }
    }
}

template <typename T>
__device__ __forceinline__ T cat_pick(const T (&a)[kCatFuse], int j) {
    T v = a[0];
#pragma unroll
    for (int k = 1; k < kCatFuse; ++k)
        if (j == k) v = a[k];
    return v;
}

// Probability with which member j's pair converts a substrate at (xs, ys)
// (post-move): P_cat x Hill(rho at the substrate) [x adaptation].
template <typename Real>
__device__ __forceinline__ Real cat_prob(const CatArgs<Real>& A, const Grids<Real>& G, int j,
                                         Real xs, Real ys) {
    const Real prob = cat_pick(A.prob, j);
    if (A.fb.sensor < 0) return prob;
    const Real hf = fb_factor(G, A.fb, xs, ys);
    if (!A.w) return prob * hf;
    Real af = (Real)1;
    if (A.Kw > (Real)0) {
        // dynamics.py:1078 reads adaptation_field[gx, gy] with the
        // feedback grid's indices on the adaptation grid's array
        // (no bounds check): element gx * wny + gy where that lies in
        // the array; past it the reference has no defined value and
        // the field is read at the substrate's own adaptation cell
        const int gx = trunc_cell(xs, (Real)A.fb.gsx, A.fb.gnx);
        const int gy = trunc_cell(ys, (Real)A.fb.gsy, A.fb.gny);
        int wc = gx * A.wny + gy;
        if (wc >= A.wnx * A.wny) {
            const double asx = ((double)A.fb.gnx * A.fb.gsx) / A.wnx;
            const double asy = ((double)A.fb.gny * A.fb.gsy) / A.wny;
            wc = trunc_cell(xs, (Real)asx, A.wnx) * A.wny + trunc_cell(ys, (Real)asy, A.wny);
        }
        const Real wr = pow((Real)A.w[wc] / A.Kw, A.nw);
        af = (Real)1 / ((Real)1 + wr);
    }
    return prob * hf * af;
}

// Member j's uniform for the pair (sa, sb) (sa holds the lower particle
// index): the hook draw at the pair's index in the reference's pair list,
// or the Philox draw of the pair's two particle indices.
template <typename Real>
__device__ __forceinline__ double cat_uniform(const Cells<Real>& C, const CatArgs<Real>& A,
                                              uint32_t step, int j, int sa, int sb, int it,
                                              int iu) {
    if (A.dr.base) return A.dr.base[cat_pick(A.doff, j) + pair_index(C, sa, sb)];
    const uint32_t lo = (uint32_t)(min(it, iu) - C.kbase);
    const uint32_t hi = (uint32_t)(max(it, iu) - C.kbase);
    return (double)Uni<Real>::first(A.rng.draw(step, lo, hi, (kTagCat << 16) | cat_pick(A.rxi, j)));
}

// Member j's pair of enzyme t (particle index it) and substrate u (index
// iu), post-move positions given: converts u if the pair is within contact
// and its uniform accepts.
template <typename Real>
__device__ __forceinline__ int cat_try_ids(const Cells<Real>& C, const Post<Real>& P,
                                           const CatArgs<Real>& A, uint32_t step, int j,
                                           const Grids<Real>& G, int n_grids, int t, int it,
                                           int u, int iu, Real xnt, Real ynt, Real xnu,
                                           Real ynu) {
    Real dx, dy;
    if (min_image(xnt, ynt, xnu, ynu, C.Lx, C.Ly, C.px, C.py, dx, dy) > cat_pick(A.contact2, j))
        return 0;
    const bool tlow = it < iu;
    const int sa = tlow ? t : u, sb = tlow ? u : t;
    if (!pair_in_list(C, sa, sb)) return 0;
    const Real pe = cat_prob(A, G, j, xnu, ynu);
    const double uu = cat_uniform(C, A, step, j, sa, sb, it, iu);
    if (uu < (double)pe && cas_byte(P.sp, u, (int8_t)A.sub, (int8_t)A.prod)) {
        count_update(G, n_grids, A.sub, xnu, ynu, -1, A.deferred);
        count_update(G, n_grids, A.prod, xnu, ynu, +1, A.deferred);
        if (A.app) wl_append(A.app, A.app_cnt, u);
        return 1;
    }
    return 0;
}

template <typename Real>
__device__ __forceinline__ int cat_try(const Cells<Real>& C, const Post<Real>& P,
                                       const CatArgs<Real>& A, uint32_t step, int j,
                                       const Grids<Real>& G, int n_grids, int t, int it, int u,
                                       Real xnt, Real ynt, Real xnu, Real ynu) {
    return cat_try_ids(C, P, A, step, j, G, n_grids, t, it, u, (int)C.id[u], xnt, ynt, xnu, ynu);
}

// ----------------------------------------------------------------------
// Autocatalysis: enzyme == product, E + S -> 2E (dynamics.py:826-1090 with
// enzyme_id == product_id).  The reference's loop runs over the pair list
// in order, so a substrate converted at pair p is an enzyme for every later
// pair and conversions cascade along the pair order.  Substrate s converts
// at
//   t(s) = min { p : pair p holds s, is in the (capped) list and within
//                contact after the move, its uniform accepts with the Hill
//                factor read at s, and its partner is an enzyme at p: one
//                from the start of the pass, or a substrate with
//                t(partner) < p },
// or never.  The condition at p looks only at conversions before p, so t
// has exactly one solution, the loop's outcome.  k_autocat_edges lists
// every enzyme-substrate and substrate-substrate pair of the list with its
// index and the directions its uniform accepts; k_autocat_resolve lowers t
// from "never" in rounds over those edges until a round changes nothing.
// Each value it holds stays at or above the solution (a lowering needs a
// partner already converted before p), so the round without a change ends
// on the solution; then the substrates with a finite t convert.  Pair
// indices come from the pair-list offsets in both draw modes, so Philox
// runs cascade in the reference's pair order too.
// ----------------------------------------------------------------------
static constexpr unsigned long long kAcNever = ~0ull;
static constexpr unsigned long long kAcIndex = (1ull << 61) - 1ull;  // edge key: pair index
static constexpr int kAcFwd = 62, kAcBwd = 61;  // edge key bits: a converts b, b converts a

// The accepted pairs of particle t (of the enzyme or substrate species):
// each pair is listed once, from its lower particle index (ea = that slot).
template <typename Real>
__device__ __forceinline__ void autocat_edges_one(const Cells<Real>& C, const Post<Real>& P,
                                                  const CatArgs<Real>& A, uint32_t step,
                                                  const Grids<Real>& G, int t,
                                                  unsigned long long* ekey, int32_t* ea,
                                                  int32_t* eb, unsigned int* ne,
                                                  unsigned int cap) {
    const int E = A.enz[0], S = A.sub;
    const int st = P.sp[t];
    if ((st != E && st != S) || !(C.fl[t] & EDM_FLAG_ALIVE)) return;
    const Real xt = C.x[t], yt = C.y[t];
    const Real xnt = P.xn[t], ynt = P.yn[t];
    const int it = C.id[t];
    int cx, cy;
    cell_of(C, xt, yt, cx, cy);
    for (int k = 0; k < 9; ++k) {
        const int c = nb_cell(C, cx, cy, k);
        if (c < 0) continue;
        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
            const int su = P.sp[u];
            if (u == t || (su != E && su != S) || (st == E && su == E)) continue;
            const int iu = C.id[u];
            if (iu < it || !(C.fl[u] & EDM_FLAG_ALIVE)) continue;
            Real dx, dy;
            if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) >= C.cut2)
                continue;
            if (!pair_in_list(C, t, u)) continue;
            const Real xnu = P.xn[u], ynu = P.yn[u];
            if (min_image(xnt, ynt, xnu, ynu, C.Lx, C.Ly, C.px, C.py, dx, dy) > A.contact2[0])
                continue;
            const double uu = cat_uniform(C, A, step, 0, t, u, it, iu);
            const bool fwd = su == S && uu < (double)cat_prob(A, G, 0, xnu, ynu);
            const bool bwd = st == S && uu < (double)cat_prob(A, G, 0, xnt, ynt);
            if (!fwd && !bwd) continue;
            const unsigned int e = atomicAdd(ne, 1u);
            if (e < cap) {
                ekey[e] = ((unsigned long long)pair_index(C, t, u) & kAcIndex) |
                          ((unsigned long long)fwd << kAcFwd) |
                          ((unsigned long long)bwd << kAcBwd);
                ea[e] = t;
                eb[e] = u;
            }
        }
    }
}

template <typename Real>
__global__ void k_autocat_edges(int n, Cells<Real> C, Post<Real> P, CatArgs<Real> A,
                                Grids<Real> G, unsigned long long* ekey, int32_t* ea,
                                int32_t* eb, unsigned int* ne, unsigned int cap) {
    const uint32_t step = A.step + (A.step_base ? *A.step_base : 0u);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    autocat_edges_one(C, P, A, step, G, t, ekey, ea, eb, ne, cap);
}

// One workgroup over one membrane's accepted pairs: the rounds, the
// conversions, and t reset to "never" on the edges' particles (tconv holds
// "never" for every slot between passes).
template <typename Real>
__device__ __forceinline__ void autocat_resolve_block(const Post<Real>& P, const CatArgs<Real>& A,
                                                      const unsigned long long* ekey,
                                                      const int32_t* ea, const int32_t* eb,
                                                      unsigned int* ne, unsigned int cap,
                                                      unsigned long long* tconv,
                                                      unsigned long long* ctr,
                                                      const Grids<Real>& G, int n_grids) {
    const unsigned int m = min(*ne, cap);
    // more accepted pairs than the buffer holds: the outcome would differ
    // from the reference's loop, so keep the count in ne[1] for the host to
    // report (EDM_ECAPACITY)
    if (threadIdx.x == 0 && *ne > cap) atomicMax(&ne[1], *ne);
    const int E = A.enz[0], S = A.sub;
    const int8_t* sp = P.sp;  // the pass's starting species until the conversions
    for (;;) {
        int changed = 0;
        for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
            const unsigned long long k = ekey[e];
            const unsigned long long p = k & kAcIndex;
            const int a = ea[e], b = eb[e];
            // t is read with atomics: lowered by other threads of the round
            if ((k >> kAcFwd) & 1ull) {
                const bool enz = sp[a] == E || atomicAdd(&tconv[a], 0ull) < p;
                if (enz && p < atomicAdd(&tconv[b], 0ull) && atomicMin(&tconv[b], p) > p)
                    changed = 1;
            }
            if ((k >> kAcBwd) & 1ull) {
                const bool enz = sp[b] == E || atomicAdd(&tconv[b], 0ull) < p;
                if (enz && p < atomicAdd(&tconv[a], 0ull) && atomicMin(&tconv[a], p) > p)
                    changed = 1;
            }
        }
        __threadfence();
        if (!__syncthreads_or(changed)) break;
    }
    for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
        const int ends[2] = {ea[e], eb[e]};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int x = ends[q];
            if (atomicAdd(&tconv[x], 0ull) != kAcNever &&
                cas_byte(P.sp, x, (int8_t)S, (int8_t)E)) {
                atomicAdd(&ctr[2], 1ull);
                count_update(G, n_grids, S, P.xn[x], P.yn[x], -1, A.deferred);
                count_update(G, n_grids, E, P.xn[x], P.yn[x], +1, A.deferred);
            }
        }
    }
    __threadfence();
    __syncthreads();
    for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
        tconv[ea[e]] = kAcNever;
        tconv[eb[e]] = kAcNever;
    }
    __syncthreads();
    if (threadIdx.x == 0) *ne = 0u;  // ready for the next pass
}

template <typename Real>
__global__ __launch_bounds__(1024) void k_autocat_resolve(Post<Real> P, CatArgs<Real> A,
                                                          const unsigned long long* ekey,
                                                          const int32_t* ea, const int32_t* eb,
                                                          unsigned int* ne, unsigned int cap,
                                                          unsigned long long* tconv,
                                                          unsigned long long* ctr, Grids<Real> G,
                                                          int n_grids) {
    autocat_resolve_block(P, A, ekey, ea, eb, ne, cap, tconv, ctr, G, n_grids);
}

// Appends v to the block's LDS queue for the lanes with `want` (one atomic
// per wavefront); returns false for a lane whose entry did not fit.
__device__ __forceinline__ bool queue_push(int* q, int* nq, int cap, bool want, int v) {
    const uint64_t b = __ballot(want);
    if (!b) return true;
    const int leader = __ffsll((unsigned long long)b) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(nq, __popcll(b));
    base = __shfl(base, leader);
    if (!want) return true;
    const int k = base + __popcll(b & ((1ull << lane_id()) - 1ull));
    if (k >= cap) return false;
    q[k] = v;
    return true;
}

// Catalysis over one strip; every thread of the 256-wide block must call it.
// Staged strips run in two phases: the entity groups scan their candidates
// and queue the pairs within the cutoff in LDS, then every thread of the
// block evaluates queued pairs (contact, feedback, draw, CAS), so that the
// per-pair work runs on full wavefronts.  A pair's outcome does not depend
// on when it is evaluated (its substrate converts once, on the first
// accepted pair).
template <typename Real, int kCols>
__device__ __forceinline__ void cat_strip(const Cells<Real>& C, const Post<Real>& P,
                                          const CatArgs<Real>& A, uint32_t step,
                                          unsigned long long* ctr, const Grids<Real>& G,
                                          int n_grids, const StripGeo& S) {
    constexpr int kCap = 1024;
    __shared__ Real s_x[kCap], s_y[kCap];
    __shared__ int8_t s_sp[kCap];
    __shared__ uint8_t s_fl[kCap];
    __shared__ int s_list[kCap];
    __shared__ int s_off[strip_cells_max(kCols) + 1];
    __shared__ int s_gst[kStripGst];
    __shared__ int s_wsum[4];
    __shared__ int s_nlist;
    constexpr int kQCap = 512;
    __shared__ int s_q[kQCap];  // queued pairs: own index | candidate << 10 | member << 20
    __shared__ int s_nq;

    if (threadIdx.x == 0) {
        s_nlist = 0;
        s_nq = 0;
    }
    const int total = strip_cells<kCols>(C, S, s_gst, s_off, s_wsum);
    const bool staged = total <= kCap;
    if (staged) {
        // all loads first, then the LDS stores: one memory round trip
        constexpr int kPer = kCap / 256;
        Real vx[kPer], vy[kPer];
        int8_t vs[kPer];
        uint8_t vf[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int o = (int)threadIdx.x + 256 * j;
            vx[j] = vy[j] = (Real)0;
            vs[j] = 0;
            vf[j] = 0;
            if (o < total) {
                const int g = strip_src<kCols>(S, s_off, s_gst, o);
                vx[j] = C.x[g];
                vy[j] = C.y[g];
                vs[j] = P.sp[g];
                vf[j] = C.fl[g];
            }
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int o = (int)threadIdx.x + 256 * j;
            if (o < total) {
                s_x[o] = vx[j];
                s_y[o] = vy[j];
                s_sp[o] = vs[j];
                s_fl[o] = vf[j];
            }
        }
    }
    __syncthreads();
    // own particles: rows 1..rows of the own columns; an own entity is named
    // by its staged position (a position of the own ranges also when the
    // strip is not staged)
    const OwnCol O1 = own_col(S, s_off, s_gst, 1);
    const OwnCol O2 = kCols > 1 && S.ncol > 1 ? own_col(S, s_off, s_gst, 2) : OwnCol{0, 0, 0};
    const int n1 = O1.end - O1.begin, nown = n1 + (O2.end - O2.begin);
    auto own_at = [&](int q) { return q < n1 ? O1.begin + q : O2.begin + (q - n1); };
    auto own_c2 = [&](int i) { return kCols > 1 && S.ncol > 1 && i >= O2.begin; };
    auto own_slot = [&](int i) {
        return own_c2(i) ? O2.gbase + (i - O2.begin) : O1.gbase + (i - O1.begin);
    };
    // the strip's own alive enzymes as (member << 28 | own index) entries, or
    // its own alive substrates as (own index) entries
    const bool by_sub = A.by_substrate;
    for (int q = (int)threadIdx.x; q < nown; q += blockDim.x) {
        const int i = own_at(q);
        const int t = own_slot(i);
        const int s = staged ? s_sp[i] : P.sp[t];
        const uint8_t f = staged ? s_fl[i] : C.fl[t];
        if (!(f & EDM_FLAG_ALIVE)) continue;
        if (by_sub) {
            if (s == A.sub) {
                const int e = atomicAdd(&s_nlist, 1);
                if (e < kCap) s_list[e] = i;
            }
            continue;
        }
#pragma unroll
        for (int j = 0; j < kCatFuse; ++j)
            if (j < A.n_enz && s == A.enz[j]) {
                const int e = atomicAdd(&s_nlist, 1);
                if (e < kCap) s_list[e] = (j << 28) | i;
            }
    }
    __syncthreads();
    // more entries than list slots (only in pathologically dense strips):
    // walk the own range directly, once per member
    const bool use_list = s_nlist <= kCap;
    const int nl = use_list ? s_nlist : (by_sub ? nown : nown * A.n_enz);
    // a group of 2^group_shift lanes per entity (a wavefront at most), its
    // lanes over the entity's candidates: wide groups run the per-candidate
    // work (Hill lookups, draws, CAS) side by side, narrow ones keep more
    // lanes busy when few candidates qualify
    const int gs = A.group_shift, gsize = 1 << gs;
    const int wave = (int)(threadIdx.x >> gs), lane = (int)threadIdx.x & (gsize - 1);
    const int nwave = (int)(blockDim.x >> gs);
    int hits = 0;
    if (by_sub) {
        for (int e = wave; e < nl; e += nwave) {
            const int i = use_list ? s_list[e] : own_at(e);
            const int t = own_slot(i);  // the substrate
            const int c0 = own_c2(i) ? 1 : 0;  // staged column left of the own one
            if (!use_list) {
                const int s = staged ? s_sp[i] : P.sp[t];
                const uint8_t f = staged ? s_fl[i] : C.fl[t];
                if (s != A.sub || !(f & EDM_FLAG_ALIVE)) continue;  // group-uniform
            }
            if (staged) {
                const Real xt = s_x[i], yt = s_y[i];
                const int r = strip_row(C, S, yt);
                const int k0 = c0 * S.nrow + r - 1;
                const int b0 = s_off[k0], l0 = s_off[k0 + 3] - b0;
                const int b1 = s_off[k0 + S.nrow], l1 = s_off[k0 + S.nrow + 3] - b1;
                const int b2 = s_off[k0 + 2 * S.nrow], l2 = s_off[k0 + 2 * S.nrow + 3] - b2;
                for (int c = lane; c < l0 + l1 + l2; c += gsize) {
                    const int u =
                        c < l0 ? b0 + c : (c < l0 + l1 ? b1 + (c - l0) : b2 + (c - l0 - l1));
                    if (u == i || !(s_fl[u] & EDM_FLAG_ALIVE)) continue;
                    const int su = s_sp[u];
                    bool enz = false;
#pragma unroll
                    for (int j = 0; j < kCatFuse; ++j) enz |= j < A.n_enz && su == A.enz[j];
                    if (!enz) continue;
                    Real dx, dy;
                    if (min_image(s_x[u], s_y[u], xt, yt, C.Lx, C.Ly, C.px, C.py, dx, dy) >=
                        C.cut2)
                        continue;
                    for (int j = 0; j < A.n_enz; ++j) {
                        const bool want = su == cat_pick(A.enz, j);
                        if (!queue_push(s_q, &s_nq, kQCap, want, i | (u << 10) | (j << 20))) {
                            const int tu = strip_src<kCols>(S, s_off, s_gst, u);  // the enzyme
                            hits += cat_try(C, P, A, step, j, G, n_grids, tu, C.id[tu], t,
                                            P.xn[tu], P.yn[tu], P.xn[t], P.yn[t]);
                        }
                    }
                }
            } else {
                const Real xt = C.x[t], yt = C.y[t], xnt = P.xn[t], ynt = P.yn[t];
                int cx, cy;
                cell_of(C, xt, yt, cx, cy);
                for (int k = lane; k < 9; k += gsize) {
                    const int c = nb_cell(C, cx, cy, k);
                    if (c < 0) continue;
                    for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                        if (u == t || !(C.fl[u] & EDM_FLAG_ALIVE)) continue;
                        const int su = P.sp[u];
                        bool enz = false;
#pragma unroll
                        for (int j = 0; j < kCatFuse; ++j) enz |= j < A.n_enz && su == A.enz[j];
                        if (!enz) continue;
                        Real dx, dy;
                        if (min_image(C.x[u], C.y[u], xt, yt, C.Lx, C.Ly, C.px, C.py, dx, dy) >=
                            C.cut2)
                            continue;
                        const int iu = C.id[u];
                        for (int j = 0; j < A.n_enz; ++j)
                            if (su == cat_pick(A.enz, j))
                                hits += cat_try(C, P, A, step, j, G, n_grids, u, iu, t, P.xn[u],
                                                P.yn[u], xnt, ynt);
                    }
                }
            }
        }
    }
    for (int e = wave; e < (by_sub ? 0 : nl); e += nwave) {
        const int ent = use_list ? s_list[e] : ((e / nown) << 28) | own_at(e % nown);
        const int j = ent >> 28;
        const int i = ent & 0x0fffffff;
        const int t = own_slot(i);
        const int c0 = own_c2(i) ? 1 : 0;  // staged column left of the own one
        if (!use_list) {
            const int s = staged ? s_sp[i] : P.sp[t];
            const uint8_t f = staged ? s_fl[i] : C.fl[t];
            if (s != cat_pick(A.enz, j) || !(f & EDM_FLAG_ALIVE)) continue;  // group-uniform
        }
        const int it = C.id[t];
        if (staged) {
            const Real xt = s_x[i], yt = s_y[i];
            const int r = strip_row(C, S, yt);
            const int k0 = c0 * S.nrow + r - 1;
            const int b0 = s_off[k0], l0 = s_off[k0 + 3] - b0;
            const int b1 = s_off[k0 + S.nrow], l1 = s_off[k0 + S.nrow + 3] - b1;
            const int b2 = s_off[k0 + 2 * S.nrow], l2 = s_off[k0 + 2 * S.nrow + 3] - b2;
            for (int c = lane; c < l0 + l1 + l2; c += gsize) {
                const int u = c < l0 ? b0 + c : (c < l0 + l1 ? b1 + (c - l0) : b2 + (c - l0 - l1));
                if (u == i || s_sp[u] != A.sub || !(s_fl[u] & EDM_FLAG_ALIVE)) continue;
                Real dx, dy;
                if (min_image(xt, yt, s_x[u], s_y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) >= C.cut2)
                    continue;
                if (!queue_push(s_q, &s_nq, kQCap, true, i | (u << 10) | (j << 20))) {
                    const int tu = strip_src<kCols>(S, s_off, s_gst, u);  // the substrate
                    hits += cat_try(C, P, A, step, j, G, n_grids, t, it, tu, P.xn[t], P.yn[t],
                                    P.xn[tu], P.yn[tu]);
                }
            }
        } else {
            const Real xt = C.x[t], yt = C.y[t], xnt = P.xn[t], ynt = P.yn[t];
            int cx, cy;
            cell_of(C, xt, yt, cx, cy);
            for (int k = lane; k < 9; k += gsize) {
                const int c = nb_cell(C, cx, cy, k);
                if (c < 0) continue;
                for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                    if (u == t || P.sp[u] != A.sub || !(C.fl[u] & EDM_FLAG_ALIVE)) continue;
                    Real dx, dy;
                    if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) >=
                        C.cut2)
                        continue;
                    hits += cat_try(C, P, A, step, j, G, n_grids, t, it, u, xnt, ynt, P.xn[u],
                                    P.yn[u]);
                }
            }
        }
    }
    // phase 2: the queued pairs, one per thread (own entity i, candidate u)
    __syncthreads();
    const int nq = min(s_nq, kQCap);
    for (int k = (int)threadIdx.x; k < nq; k += (int)blockDim.x) {
        const int v = s_q[k];
        const int i = v & 1023, u = (v >> 10) & 1023, j = v >> 20;
        const int to = own_slot(i);                   // the own entity's slot
        const int tc = strip_src<kCols>(S, s_off, s_gst, u);  // the candidate's slot
        const Real xo = P.xn[to], yo = P.yn[to], xc = P.xn[tc], yc = P.yn[tc];
        if (by_sub)  // enzyme u, substrate i
            hits += cat_try(C, P, A, step, j, G, n_grids, tc, C.id[tc], to, xc, yc, xo, yo);
        else  // enzyme i, substrate u
            hits += cat_try(C, P, A, step, j, G, n_grids, to, C.id[to], tc, xo, yo, xc, yc);
    }
    unsigned long long v = hits;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane_id() == 0 && v) atomicAdd(&ctr[2], v);
}

template <typename Real>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 6>))) void k_cat_tile(Cells<Real> C, Post<Real> P, CatArgs<Real> A,
                                                  unsigned long long* ctr, Grids<Real> G,
                                                  int n_grids, int ntile_y, int ntiles) {
    // the parameter block is indexed with per-lane values: keep it in LDS
    // rather than in per-lane scratch
    __shared__ CatArgs<Real> sA;
    if (A.fold_n > 0) fold_in_launch(A.fold_counts, A.fold_delta, A.fold_n);
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // uniform across the block
    if (threadIdx.x == 0) {
        sA = A;
        if (A.step_base) sA.step += *A.step_base;
    }
    __syncthreads();
    cat_strip<Real, 1>(C, P, sA, sA.step, ctr, G, n_grids, strip_geo(C, tile, ntile_y));
}

// Position of the k-th set bit of m (k < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int k) {
    int base = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (k >= c) {
            k -= c;
            m >>= w;
            base += w;
        } else {
            m = low;
        }
    }
    return base;
}

// LDS of one wavefront of the flat catalysis walk: the batch's entities,
// their runs of cells, and the queue of candidate pairs within the cutoff.
template <typename Real>
struct CatFlatLds {
    int inc[64];            // inclusive candidate counts of the batch's entities
    int slot[64];           // the entities' slots
    Real x[64], y[64];      // their pre-move positions
    int pb[64 * 9];         // run p of entity e: slots [pb, pb + pe - pe_prev)
    int pe[64 * 9];         // cumulative run ends
    int qa[64], qb[64];     // queued (walk entity, candidate) slots within the cutoff
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Drains the queue: lane l < n evaluates queued pair l.
template <typename Real>
__device__ __forceinline__ int cat_flat_drain(const Cells<Real>& C, const Post<Real>& P,
                                              const CatArgs<Real>& A, uint32_t step, bool by_sub,
                                              const Grids<Real>& G, int n_grids,
                                              const CatFlatLds<Real>& L, int n) {
    const int lane = lane_id();
    if (lane >= n) return 0;
    // enzyme e, substrate s
    const int e = by_sub ? L.qb[lane] : L.qa[lane];
    const int s = by_sub ? L.qa[lane] : L.qb[lane];
    const int se = P.sp[e], ie = C.id[e], is = C.id[s];
    const Real xne = P.xn[e], yne = P.yn[e], xns = P.xn[s], yns = P.yn[s];
    int hits = 0;
    for (int j = 0; j < A.n_enz; ++j)
        if (se == cat_pick(A.enz, j))
            hits += cat_try_ids(C, P, A, step, j, G, n_grids, e, ie, s, is, xne, yne, xns, yns);
    return hits;
}

// Flat walk of one wavefront.  Lane l holds walk entity t (an alive enzyme
// of the run, or with by_sub an alive substrate) when bit l of `mask` is
// set.  The entities are taken `per` at a time; each entity's nine cells
// are merged into runs of consecutive cells (contiguous slot ranges), and
// the batch's candidates -- every slot of every entity's runs -- are
// spread over the wavefront's lanes (owner by binary search over the
// inclusive counts), so lanes do no redundant work whatever the cell
// populations.  Candidates of the right species, alive and within the
// pair-list cutoff (pre-move) are queued in LDS and evaluated 64 at a time
// by the full wavefront.  The candidates of an entity are the slots of its
// nine cells, as in every other catalysis walk, so the outcome is the same.
template <typename Real>
__device__ __forceinline__ int cat_flat_wave(const Cells<Real>& C, const Post<Real>& P,
                                             const CatArgs<Real>& A, uint32_t step, bool by_sub,
                                             const Grids<Real>& G, int n_grids, int t_lane,
                                             uint64_t mask, int per, CatFlatLds<Real>& L) {
    const int lane = lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const int n_ent = __popcll(mask);
    int hits = 0;
    int qn = 0;  // queued pairs (uniform)
    for (int b0 = 0; b0 < n_ent; b0 += per) {  // uniform
        const int ne = min(per, n_ent - b0);
        const bool has = lane < ne;
        const int t = __shfl(t_lane, has ? nth_set_bit(mask, b0 + lane) : 0);
        int tot = 0;
        if (has) {
            const Real xt = C.x[t], yt = C.y[t];
            L.slot[lane] = t;
            L.x[lane] = xt;
            L.y[lane] = yt;
            int cx, cy;
            cell_of(C, xt, yt, cx, cy);
            int np = 0, pbeg = 0, plen = 0, prevc = -2;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int c = nb_cell(C, cx, cy, k);
                if (c < 0) continue;
                const int b = (int)C.start[c], l = (int)C.start[c + 1] - b;
                if (c != prevc + 1 && plen > 0) {
                    L.pb[lane * 9 + np] = pbeg;
                    L.pe[lane * 9 + np] = tot;
                    ++np;
                    plen = 0;
                }
                if (plen == 0) pbeg = b;
                plen += l;
                tot += l;
                prevc = c;
            }
            if (plen > 0) {
                L.pb[lane * 9 + np] = pbeg;
                L.pe[lane * 9 + np] = tot;
            }
        }
        // inclusive scan of the candidate counts
        int inc = tot;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(inc, o);
            if (lane >= o) inc += v;
        }
        L.inc[lane] = inc;
        const int T = __shfl(inc, 63);
        wave_lds_sync();
        for (int c0 = 0; c0 < T; c0 += 64) {  // uniform
            const int c = c0 + lane;
            bool pass = false;
            int te = 0, u = 0;
            if (c < T) {
                // owner: the first entity whose inclusive count exceeds c
                int o = 0;
#pragma unroll
                for (int w = 32; w >= 1; w >>= 1)
                    if (o + w - 1 < ne && L.inc[o + w - 1] <= c) o += w;
                const int r = c - (o > 0 ? L.inc[o - 1] : 0);
                int k = 0;
                while (L.pe[o * 9 + k] <= r) ++k;
                u = L.pb[o * 9 + k] + r - (k > 0 ? L.pe[o * 9 + k - 1] : 0);
                te = L.slot[o];
                const Real xt = L.x[o], yt = L.y[o];
                const uint8_t fu = C.fl[u];
                const int su = P.sp[u];
                const Real xu = C.x[u], yu = C.y[u];
                if (u != te && (fu & EDM_FLAG_ALIVE)) {
                    bool sp_ok;
                    if (by_sub) {
                        sp_ok = false;
#pragma unroll
                        for (int j = 0; j < kCatFuse; ++j) sp_ok |= j < A.n_enz && su == A.enz[j];
                    } else {
                        sp_ok = su == A.sub;
                    }
                    Real dx, dy;
                    pass = sp_ok && min_image(xt, yt, xu, yu, C.Lx, C.Ly, C.px, C.py, dx, dy) <
                                        C.cut2;
                }
            }
            const uint64_t m = __ballot(pass);
            const int np = __popcll(m);
            if (qn + np > 64) {  // uniform
                wave_lds_sync();
                hits += cat_flat_drain(C, P, A, step, by_sub, G, n_grids, L, qn);
                qn = 0;
                wave_lds_sync();
            }
            if (pass) {
                const int pos = qn + __popcll(m & below);
                L.qa[pos] = te;
                L.qb[pos] = u;
            }
            qn += np;
        }
        // the next batch rewrites the entity tables
        wave_lds_sync();
    }
    if (qn > 0) {
        wave_lds_sync();
        hits += cat_flat_drain(C, P, A, step, by_sub, G, n_grids, L, qn);
    }
    return hits;
}

// Grid (list regions, kWlPer): each wavefront takes per = 64 >> group_shift
// consecutive entries of its region at a time (chunks y * 4 + wave,
// stepping by 4 * kWlPer) and walks them with cat_flat_wave.
template <typename Real>
__global__ __launch_bounds__(256) void k_cat_list(Cells<Real> C, Post<Real> P, CatArgs<Real> A,
                                                  WlView V, unsigned long long* ctr,
                                                  Grids<Real> G, int n_grids) {
    if (A.fold_n > 0) fold_in_launch(A.fold_counts, A.fold_delta, A.fold_n);
    const int cnt = (int)V.cnt[blockIdx.x];
    const int per = 64 >> A.group_shift;
    if ((int)blockIdx.y * 4 * per >= cnt) return;  // uniform: nothing listed for this block
    // the parameter block is indexed with per-lane values: keep it in LDS
    // rather than in per-lane scratch
    __shared__ CatArgs<Real> sA;
    __shared__ CatFlatLds<Real> sL[4];
    if (threadIdx.x == 0) {
        sA = A;
        if (A.step_base) sA.step += *A.step_base;
    }
    __syncthreads();
    const int lane = lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int hits = 0;
    for (int ch = (int)blockIdx.y * 4 + wave; ch * per < cnt; ch += 4 * (int)gridDim.y) {
        const int e = ch * per + lane;
        const bool has = lane < per && e < cnt;
        const int t = has ? V.list[(size_t)blockIdx.x * kWlSlots + e] : 0;
        hits += cat_flat_wave(C, P, sA, sA.step, sA.by_substrate, G, n_grids, t, __ballot(has),
                              64, sL[wave]);
    }
    unsigned long long v = hits;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if (lane == 0 && v) atomicAdd(&ctr[2], v);
}
