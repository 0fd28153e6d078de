// LDS strip staging shared by the tiled dynamics kernels (included by
// edm_dynamics.hip; relies on its Cells / ForceParams / Grids helpers).
//
// A workgroup of 256 threads owns a strip of C.tile_rows (at most
// kTileRows) cells of one cell column, or of two adjacent columns (sparse
// membranes whose columns are short: a workgroup then owns about as many
// particles as it has threads).  The strip's (ncol + 2) x (rows + 2)
// neighbourhood cells (columns cx-1..cx+ncol, rows cy0-1..cy0+rows, wrapped
// or clipped at the box edges) are staged in LDS, one particle per thread,
// so that consecutive threads read consecutive slots (strip_src).  Inside
// the strip every particle's 3x3 cell neighbourhood is then three
// contiguous LDS ranges, one per column.
// Blocks that the dispatcher sends to the same XCD (blockIdx % 8) take
// adjacent strips, so a strip's neighbour columns are usually in that XCD's
// L2.  A strip whose neighbourhood exceeds the LDS capacity is handled with
// global-memory neighbour walks instead.
#pragma once

static constexpr int kMaxOwnCols = 2;
// staged cells of a strip kernel built for at most kCols own columns
__host__ __device__ constexpr int strip_cells_max(int kCols) { return (kCols + 2) * (kTileRows + 2); }

static constexpr int kStripCells = strip_cells_max(kMaxOwnCols);
// global start of a staged column's first row, middle rows and last row:
// s_gst[3 * column + 0 / 1 / 2] (the middle rows are one slot range)
static constexpr int kStripGst = 3 * (kMaxOwnCols + 2);

// ncol own columns (1 or 2) = staged columns 1..ncol; nscol = ncol + 2
// staged columns of nrow = rows + 2 cells each, column-major.
struct StripGeo {
    int tile, cx, cy0, rows, nrow, ncells, ncol, nscol;
};

// Staged column of staged cell k or of a staged position o.
__device__ __forceinline__ int strip_col_of_cell(const StripGeo& S, int k) {
    return (k >= S.nrow) + (k >= 2 * S.nrow) + (k >= 3 * S.nrow);
}

// The own particles of own column c (1..ncol): staged positions [begin,
// end) and the global slot of begin (a column's middle rows are one
// contiguous slot range).
struct OwnCol {
    int begin, end, gbase;
};

__device__ __forceinline__ OwnCol own_col(const StripGeo& S, const int* s_off, const int* s_gst,
                                          int c) {
    const int k = c * S.nrow + 1;
    return OwnCol{s_off[k], s_off[k + S.rows], s_gst[3 * c + 1]};
}

__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + 7) / 8;
    return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

template <typename Real>
__device__ __forceinline__ StripGeo strip_geo(const Cells<Real>& C, int tile, int ntile_y) {
    StripGeo S;
    S.tile = tile;
    S.cx = tile / ntile_y;
    S.cy0 = (tile % ntile_y) * C.tile_rows;
    S.rows = min(C.tile_rows, C.ncy - S.cy0);
    S.nrow = S.rows + 2;
    S.ncol = 1;
    S.nscol = 3;
    S.ncells = 3 * S.nrow;
    return S;
}

// LDS offset of every staged cell, s_off[k], and the global starts
// s_gst[kStripGst]
// (exclusive scan of the cell counts, s_off[ncells] = total).  Every
// thread of the 256-wide block must call it.  Returns the total.
template <int kCols, typename Real>
__device__ __forceinline__ int strip_cells(const Cells<Real>& C, const StripGeo& S, int* s_gst,
                                           int* s_off, int* s_wsum) {
    // thread i takes the consecutive cells i*per .. i*per + per - 1 (one
    // cell per thread for strips of up to 256 cells)
    constexpr int kPer = (strip_cells_max(kCols) + 255) / 256;
    const int per = S.ncells <= 256 ? 1 : kPer;
    int cnt[kPer];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        cnt[j] = 0;
        if (j >= per) break;
        const int k = (int)threadIdx.x * per + j;
        if (k < S.ncells) {
            const int col = strip_col_of_cell(S, k), r = k - col * S.nrow;
            int gx = S.cx + col - 1, gy = S.cy0 + r - 1;
            bool ok = true;
            if (C.px) gx = gx < 0 ? gx + C.ncx : (gx >= C.ncx ? gx - C.ncx : gx);
            else if (gx < 0 || gx >= C.ncx) ok = false;
            if (C.py) gy = gy < 0 ? gy + C.ncy : (gy >= C.ncy ? gy - C.ncy : gy);
            else if (gy < 0 || gy >= C.ncy) ok = false;
            int gs = 0;
            if (ok) {
                const int c = gx * C.ncy + gy;
                gs = (int)C.start[c];
                cnt[j] = (int)C.start[c + 1] - gs;
            }
            if (r <= 1) s_gst[3 * col + r] = gs;
            if (r == S.nrow - 1) s_gst[3 * col + 2] = gs;
        }
        sum += cnt[j];
    }
    // block-wide exclusive scan of the threads' sums: wavefront scans, then
    // wavefront totals
    const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
    int v = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    if (lane == 63) s_wsum[wave] = v;
    __syncthreads();
    int run = v - sum;
    for (int w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        if (j >= per) break;
        const int k = (int)threadIdx.x * per + j;
        if (k < S.ncells) s_off[k] = run;
        if (k == S.ncells - 1) s_off[S.ncells] = run + cnt[j];
        run += cnt[j];
    }
    __syncthreads();
    return s_off[S.ncells];
}

// Global slot of staged position o (0 <= o < total).  Cells are stored in
// slot order by (column, row), so a column's middle rows cy0 .. cy0+rows-1
// form one contiguous slot range; only its first and last staged rows can
// be wrapped or clipped cells, and those are mapped on their own.  Staging
// therefore reads consecutive addresses with consecutive threads.
template <int kCols>
__device__ __forceinline__ int strip_src(const StripGeo& S, const int* s_off, const int* s_gst,
                                         int o) {
    // s_off[3 * nrow] is the total when there are three staged columns
    int col = (o >= s_off[S.nrow]) + (o >= s_off[2 * S.nrow]);
    if constexpr (kCols > 1) col += o >= s_off[3 * S.nrow];
    const int k0 = col * S.nrow;
    int k = k0 + 1, w = 1;
    if (o < s_off[k0 + 1]) {
        k = k0;
        w = 0;
    } else if (o >= s_off[k0 + S.nrow - 1]) {
        k = k0 + S.nrow - 1;
        w = 2;
    }
    return s_gst[3 * col + w] + (o - s_off[k]);
}

// Staged row (1..rows) of an own particle at pre-move position y.
template <typename Real>
__device__ __forceinline__ int strip_row(const Cells<Real>& C, const StripGeo& S, Real y) {
    return trunc_cell(y, C.csy, C.ncy) - S.cy0 + 1;
}

// Force sub-rows.  The potentials reach far less than a cell (about 11 nm
// against the 165 nm pair-list cells of the reaction contacts), so each
// cell row is cut into nsub sub-rows at least 1.01 x the longest potential
// range tall (at most kMaxSub): a pair within range is then at most one
// sub-row apart, and a particle's force candidates in a column are the
// particles of its own sub-row and the two next to it.
static constexpr int kMaxSub = 8;

// Dynamic LDS bytes of a force strip launch whose strips have at most
// `rows` cell rows and `ncol` own columns.
static inline size_t force_sub_lds(int rows, int ncol) {
    return (size_t)((ncol + 2) * (rows + 2) * kMaxSub + 1) * sizeof(int16_t);
}

template <typename Real>
__device__ __forceinline__ int force_nsub(const Cells<Real>& C, Real rl2) {
    if (!(rl2 > (Real)0)) return kMaxSub;
    const Real q = C.csy / ((Real)1.01 * sqrt(rl2));
    return q >= (Real)kMaxSub ? kMaxSub : max(1, (int)q);
}

// Sub-row (0..nsub-1) of a particle at y in global cell row gy.
template <typename Real>
__device__ __forceinline__ int sub_row(const Cells<Real>& C, Real y, int gy, int nsub,
                                       Real inv_h) {
    const int b = (int)((y - (Real)gy * C.csy) * inv_h);
    return min(max(b, 0), nsub - 1);
}

// Forces + Ermak-McCammon move over one strip (any pair summation order).
// Every thread of the 256-wide block must call it.  A particle's candidates
// are enumerated by column, then sub-row, then slot order, in both the LDS
// and the global-memory walk, so the sums do not depend on the strip shape.
// kSub = false: no membrane of the launch uses sub-rows (one per cell; no
// sort, no dynamic LDS); kSub = true: a membrane without them sorts with
// one sub-row per cell, which keeps its order.
template <typename Real, bool kSub, int kCols>
__device__ __forceinline__ void force_strip(const Cells<Real>& C, const ForceParams<Real>& F,
                                            Real* __restrict__ xn, Real* __restrict__ yn,
                                            const double* __restrict__ noise_hook, Rng rng,
                                            uint32_t step, const Grids<Real>& G, int n_grids,
                                            uint32_t keep_mask, const StripGeo& S) {
    constexpr int kCap = 1024;
    constexpr int kTabSpecies = 16;  // species-pair tables in LDS up to this many species
    constexpr int kTab = kTabSpecies * kTabSpecies;
    __shared__ Real s_x[kCap], s_y[kCap];
    __shared__ int8_t s_sp[kCap];
    __shared__ uint8_t s_fl[kCap];
    __shared__ int s_off[strip_cells_max(kCols) + 1];
    __shared__ int s_gst[kStripGst];
    __shared__ int s_wsum[4];
    __shared__ Real s_sig[kTab], s_eps[kTab], s_kh[kTab];
    // staged positions in (cell, sub-row, slot) order, each staged
    // particle's sub-row, and (dynamic LDS of force_sub_lds(rows) bytes) the
    // sub-row starts in s_ord: index cell * nsub + sub-row, then the total
    __shared__ int16_t s_ord[kCap];
    __shared__ uint8_t s_b[kCap];
    extern __shared__ int16_t s_sub[];

    // the species-pair tables in LDS (read per candidate pair; beyond
    // kTabSpecies species they stay in global memory); the barriers of
    // strip_cells and of the staging below order them
    const bool lds_tab = F.ns <= kTabSpecies;
    if (lds_tab)
        for (int k = (int)threadIdx.x; k < F.ns * F.ns; k += (int)blockDim.x) {
            s_sig[k] = F.sig[k];
            s_eps[k] = F.eps[k];
            s_kh[k] = F.kh[k];
        }
    ForceParams<Real> FL = F;
    FL.sig = s_sig;
    FL.eps = s_eps;
    FL.kh = s_kh;
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
                vs[j] = C.sp[g];
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
    // own particles: rows 1..rows of the own columns, each column's
    // contiguous in memory
    const OwnCol O1 = own_col(S, s_off, s_gst, 1);
    const OwnCol O2 = kCols > 1 && S.ncol > 1 ? own_col(S, s_off, s_gst, 2) : OwnCol{0, 0, 0};
    const int n1 = O1.end - O1.begin, n_own = n1 + (O2.end - O2.begin);
    // pairs beyond the pair-list cutoff or the longest potential range
    // contribute nothing
    const Real rl2 = min(C.cut2, F.rint2);
    const int nsub = kSub && C.sub_rows ? force_nsub(C, rl2) : 1;
    const Real inv_h = (Real)nsub / C.csy;
    if (kSub && staged) {
        // counting sort of every staged cell by sub-row (stable: slot order
        // inside a sub-row), one thread per cell
        for (int k = (int)threadIdx.x; k < S.ncells; k += (int)blockDim.x) {
            const int col = strip_col_of_cell(S, k), r = k - col * S.nrow;
            int gy = S.cy0 + r - 1;
            gy = gy < 0 ? gy + C.ncy : (gy >= C.ncy ? gy - C.ncy : gy);
            const int o0 = s_off[k], o1 = s_off[k + 1];
            int cnt[kMaxSub];
#pragma unroll
            for (int q = 0; q < kMaxSub; ++q) cnt[q] = 0;
            for (int o = o0; o < o1; ++o) {
                const int bq = sub_row(C, s_y[o], gy, nsub, inv_h);
                s_b[o] = (uint8_t)bq;
#pragma unroll
                for (int q = 0; q < kMaxSub; ++q) cnt[q] += bq == q;
            }
            int run = o0;
#pragma unroll
            for (int q = 0; q < kMaxSub; ++q) {
                if (q < nsub) s_sub[k * nsub + q] = (int16_t)run;
                const int c = cnt[q];
                cnt[q] = run;  // next position of sub-row q
                run += c;
            }
            for (int o = o0; o < o1; ++o) {
                const int bq = s_b[o];
                int pos = 0;
#pragma unroll
                for (int q = 0; q < kMaxSub; ++q)
                    if (bq == q) pos = cnt[q]++;
                s_ord[pos] = (int16_t)o;
            }
        }
        if (threadIdx.x == 0) s_sub[S.ncells * nsub] = (int16_t)total;
        __syncthreads();
    }
    for (int q0 = 0; q0 < n_own; q0 += blockDim.x) {
        const int q = q0 + (int)threadIdx.x;
        bool take_any = false;
        Real nxp = 0, nyp = 0;
        int si = 0;
        if (q < n_own) {
            const bool c1 = kCols == 1 || q < n1;
            const int oc = c1 ? 1 : 2;  // own column
            const int i = c1 ? O1.begin + q : O2.begin + (q - n1);
            const int t = c1 ? O1.gbase + q : O2.gbase + (q - n1);
            Real xt, yt;
            uint8_t ft;
            if (staged) {
                xt = s_x[i];
                yt = s_y[i];
                si = s_sp[i];
                ft = s_fl[i];
            } else {
                xt = C.x[t];
                yt = C.y[t];
                si = C.sp[t];
                ft = C.fl[t];
            }
            nxp = xt;
            nyp = yt;
            if (ft & EDM_FLAG_ALIVE) {
                Real Fx = 0, Fy = 0;
                if (staged) {
                    // own sub-row g of the strip's column (rows 1..rows): the
                    // candidates of a column are sub-rows g-1 .. g+1, one
                    // contiguous range of s_ord (the particle itself is among
                    // them: at d^2 = 0 no potential acts)
                    // (one sub-row per cell: the column's three cells)
                    const int r = strip_row(C, S, yt);
                    const int g = r * nsub + (kSub ? (int)s_b[i] : 0);
                    auto sweep = [&](const ForceParams<Real>& T) {
                        for (int col = oc - 1; col <= oc + 1; ++col) {
                            int v0, v1;
                            if constexpr (kSub) {
                                const int base = col * S.nrow * nsub;
                                v0 = s_sub[base + g - 1];
                                v1 = s_sub[base + g + 2];
                            } else {
                                const int k0 = col * S.nrow + r - 1;
                                v0 = s_off[k0];
                                v1 = s_off[k0 + 3];
                            }
                            for (int v = v0; v < v1; ++v) {
                                const int u = kSub ? (int)s_ord[v] : v;
                                Real dx, dy;
                                const Real dsq = min_image(xt, yt, s_x[u], s_y[u], C.Lx, C.Ly,
                                                           C.px, C.py, dx, dy);
                                if (dsq >= rl2 || !(s_fl[u] & EDM_FLAG_ALIVE)) continue;
                                const int su = s_sp[u];
                                Real f;
                                if (F.wca && pair_force(T, 0, si, su, dsq, f)) {
                                    Fx += f * dx;
                                    Fy += f * dy;
                                }
                                if (F.harm && pair_force(T, 1, si, su, dsq, f)) {
                                    Fx += f * dx;
                                    Fy += f * dy;
                                }
                            }
                        }
                    };
                    if (lds_tab)
                        sweep(FL);
                    else
                        sweep(F);
                } else {
                    int cxx, cyy;
                    cell_of(C, xt, yt, cxx, cyy);
                    const int bo = sub_row(C, yt, cyy, nsub, inv_h);
                    // columns, then sub-rows bo-1 .. bo+1 (possibly in the
                    // rows above / below), then slot order: the LDS order
                    for (int k = 0; k < 9; ++k) {
                        const int q = bo + (k % 3) - 1;  // sub-row, own-row relative
                        const int dr = q < 0 ? -1 : (q >= nsub ? 1 : 0);
                        const int qs = q - dr * nsub;
                        const int c = nb_cell(C, cxx, cyy, 3 * (k / 3) + dr + 1);
                        if (c < 0) continue;
                        const int gyc = c % C.ncy;
                        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                            if (u == t) continue;
                            if (sub_row(C, C.y[u], gyc, nsub, inv_h) != qs) continue;
                            Real dx, dy;
                            const Real dsq = min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px,
                                                       C.py, dx, dy);
                            if (dsq >= rl2 || !(C.fl[u] & EDM_FLAG_ALIVE)) continue;
                            const int su = C.sp[u];
                            Real f;
                            if (F.wca && pair_force(F, 0, si, su, dsq, f)) {
                                Fx += f * dx;
                                Fy += f * dy;
                            }
                            if (F.harm && pair_force(F, 1, si, su, dsq, f)) {
                                Fx += f * dx;
                                Fy += f * dy;
                            }
                        }
                    }
                }
                move_particle(C, F, si, C.id[t] - C.kbase, xt, yt, Fx, Fy, noise_hook, rng, step,
                              nxp, nyp);
                take_any = ((keep_mask >> si) & 1u) != 0;
            }
            xn[t] = nxp;
            yn[t] = nyp;
        }
        add_density(G, n_grids, take_any, si, nxp, nyp);
    }
}

template <typename Real, bool kSub>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 7>))) void k_force_tile(Cells<Real> C, ForceParams<Real> F,
                                                    Real* __restrict__ xn, Real* __restrict__ yn,
                                                    const double* __restrict__ noise_hook,
                                                    Rng rng, StepRef sref, Grids<Real> G,
                                                    int n_grids, uint32_t keep_mask, int ntile_y,
                                                    int ntiles) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // uniform across the block
    if (cap_gated_off(C)) return;  // uniform across the launch
    force_strip<Real, kSub, 1>(C, F, xn, yn, noise_hook, rng, sref.get(), G, n_grids, keep_mask,
                               strip_geo(C, tile, ntile_y));
}
