// LDS strip staging shared by the tiled dynamics kernels (included by
// edm_dynamics.hip; relies on its Cells / ForceParams / Grids helpers).
//
// A workgroup of 256 threads owns a strip of C.tile_rows (at most
// kTileRows) cells of one cell column (sparser membranes take taller strips,
// so that a workgroup owns about as many particles as it has threads).  The strip's 3 x (rows + 2) neighbourhood cells (columns cx-1..cx+1,
// rows cy0-1..cy0+rows, wrapped or clipped at the box edges) are staged in
// LDS, one particle per thread, so that consecutive threads read consecutive
// slots (strip_src).  Inside the strip every particle's 3x3
// cell neighbourhood is then three contiguous LDS ranges, one per column.
// Blocks that the dispatcher sends to the same XCD (blockIdx % 8) take
// adjacent strips, so a strip's neighbour columns are usually in that XCD's
// L2.  A strip whose neighbourhood exceeds the LDS capacity is handled with
// global-memory neighbour walks instead.
#pragma once

static constexpr int kStripCells = 3 * (kTileRows + 2);

struct StripGeo {
    int tile, cx, cy0, rows, nrow, ncells;
};

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
    S.ncells = 3 * S.nrow;
    return S;
}

// Global start and LDS offset of every staged cell: s_gst[k], s_off[k]
// (exclusive scan of the cell counts, s_off[ncells] = total).  Every
// thread of the 256-wide block must call it.  Returns the total.
template <typename Real>
__device__ __forceinline__ int strip_cells(const Cells<Real>& C, const StripGeo& S, int* s_gst,
                                           int* s_off, int* s_wsum) {
    // thread i takes the consecutive cells i*per .. i*per + per - 1 (one
    // cell per thread for strips of up to 256 cells)
    constexpr int kPer = (kStripCells + 255) / 256;
    const int per = S.ncells <= 256 ? 1 : kPer;
    int cnt[kPer];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        cnt[j] = 0;
        if (j >= per) break;
        const int k = (int)threadIdx.x * per + j;
        if (k < S.ncells) {
            const int col = (k >= S.nrow) + (k >= 2 * S.nrow), r = k - col * S.nrow;
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
            s_gst[k] = gs;
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
__device__ __forceinline__ int strip_src(const StripGeo& S, const int* s_off, const int* s_gst,
                                         int o) {
    const int col = (o >= s_off[S.nrow]) + (o >= s_off[2 * S.nrow]);
    const int k0 = col * S.nrow;
    int k = k0 + 1;
    if (o < s_off[k0 + 1]) k = k0;
    else if (o >= s_off[k0 + S.nrow - 1]) k = k0 + S.nrow - 1;
    return s_gst[k] + (o - s_off[k]);
}

// Staged row (1..rows) of an own particle at pre-move position y.
template <typename Real>
__device__ __forceinline__ int strip_row(const Cells<Real>& C, const StripGeo& S, Real y) {
    return trunc_cell(y, C.csy, C.ncy) - S.cy0 + 1;
}

// Forces + Ermak-McCammon move over one strip (any pair summation order).
// Every thread of the 256-wide block must call it.
template <typename Real>
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
    __shared__ int s_off[kStripCells + 1];
    __shared__ int s_gst[kStripCells];
    __shared__ int s_wsum[4];
    __shared__ Real s_sig[kTab], s_eps[kTab], s_kh[kTab];

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
    const int total = strip_cells(C, S, s_gst, s_off, s_wsum);
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
                const int g = strip_src(S, s_off, s_gst, o);
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
    // own particles: the middle column's rows 1..rows, contiguous in memory
    const int own_begin = s_off[S.nrow + 1], own_end = s_off[S.nrow + S.rows + 1];
    const int gbase = s_gst[S.nrow + 1];
    // pairs beyond the pair-list cutoff or the longest potential range
    // contribute nothing
    const Real rl2 = min(C.cut2, F.rint2);
    for (int i0 = own_begin; i0 < own_end; i0 += blockDim.x) {
        const int i = i0 + (int)threadIdx.x;
        bool take_any = false;
        Real nxp = 0, nyp = 0;
        int si = 0;
        if (i < own_end) {
            const int t = gbase + (i - own_begin);
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
                    const int r = strip_row(C, S, yt);
                    // (the particle itself is among the candidates: at d^2 = 0 no
                    // potential acts)
                    auto sweep = [&](const ForceParams<Real>& T) {
                        for (int col = 0; col < 3; ++col) {
                            const int k0 = col * S.nrow + r - 1;
                            const int ub = s_off[k0 + 3];
                            for (int u = s_off[k0]; u < ub; ++u) {
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
                    for (int k = 0; k < 9; ++k) {
                        const int c = nb_cell(C, cxx, cyy, k);
                        if (c < 0) continue;
                        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                            if (u == t) continue;
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

template <typename Real>
__global__ __launch_bounds__(256) void k_force_tile(Cells<Real> C, ForceParams<Real> F,
                                                    Real* __restrict__ xn, Real* __restrict__ yn,
                                                    const double* __restrict__ noise_hook,
                                                    Rng rng, uint32_t step, Grids<Real> G,
                                                    int n_grids, uint32_t keep_mask, int ntile_y,
                                                    int ntiles) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;  // uniform across the block
    force_strip(C, F, xn, yn, noise_hook, rng, step, G, n_grids, keep_mask,
                strip_geo(C, tile, ntile_y));
}
