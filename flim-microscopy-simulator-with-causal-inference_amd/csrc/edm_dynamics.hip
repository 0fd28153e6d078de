// Dynamics kernels of the ED-MRDT per-timestep path on gfx950:
// cell binning by counting sort (storage kept in cell order), the fused
// force + Ermak-McCammon kernel with incremental coarse-grid density counts,
// and the reaction pass (binding matching, dissociation, catalysis with Hill
// feedback, recruitment with the index-order cap, detachment, conversion).
//
// Reference: ed_mrdt/dynamics.py (see the per-kernel citations).  Numerics
// are restated so that FP64 runs with injected draws reproduce the
// reference's events exactly; see edm_device.h.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>

#include "edm_device.h"
#include "edm_internal.h"
#include "edm_ensemble.h"

namespace edm {

static constexpr int kBlock = 256;
static inline int nblk(int64_t n, int b = kBlock) { return (int)((n + b - 1) / b); }

// ----------------------------------------------------------------------
// Shared read-only view of the cell-ordered membrane
// ----------------------------------------------------------------------
template <typename Real>
struct Cells {
    const Real *x, *y;  // pre-move positions (slot order)
    const int8_t* sp;
    const uint8_t* fl;
    const int32_t* id;
    const uint32_t* start;  // [ncell + 1]
    int ncx, ncy;
    Real csx, csy, Lx, Ly, cut2;
    bool px, py;
    // reference pair bookkeeping (by particle index)
    const uint64_t* pcnt;
    const uint64_t* poff;
    const int64_t* npairs;  // device scalars: length of the reference pair list, pairs found
    bool cap_active;        // the host knows the cap truncates this step's list
    bool cap_dev;           // the device decides: the cap truncates when npairs[0] < npairs[1]
    int8_t cap_gate;        // +1: the launch runs only when the cap truncates, -1: only when not
    int kbase;  // first particle index of this membrane (0 unless batched); draw keys
                // use the index relative to it
    int poff_base;  // pair offsets of particle index i at poff[i + poff_base]
    int tile_rows;  // cell rows per strip of the tiled kernels (1..kTileRows)
    bool sub_rows;  // force strips enumerate candidates by force sub-row
};

template <typename Real>
__device__ __forceinline__ void cell_of(const Cells<Real>& C, Real x, Real y, int& cx, int& cy) {
    cx = trunc_cell(x, C.csx, C.ncx);
    cy = trunc_cell(y, C.csy, C.ncy);
}

// Neighbour cell k (0..8) of (cx, cy) in the reference's (dcx, dcy)
// lexicographic order (dynamics.py:244-258); -1 when outside a wall.
// (cx, cy) is a cell of the grid and the offsets are -1..1, so a periodic
// wrap is one conditional add or subtract: the same cell as the modulo.
template <typename Real>
__device__ __forceinline__ int nb_cell(const Cells<Real>& C, int cx, int cy, int k) {
    const int q = k >= 6 ? 2 : (k >= 3 ? 1 : 0);
    int nx = cx + q - 1;
    int ny = cy + (k - 3 * q) - 1;
    if (C.px) nx = nx < 0 ? nx + C.ncx : (nx >= C.ncx ? nx - C.ncx : nx);
    else if (nx < 0 || nx >= C.ncx) return -1;
    if (C.py) ny = ny < 0 ? ny + C.ncy : (ny >= C.ncy ? ny - C.ncy : ny);
    else if (ny < 0 || ny >= C.ncy) return -1;
    return nx * C.ncy + ny;
}

// Rank of slot `sb` in the pair enumeration of slot `sa` (id[sa] < id[sb]):
// reference order = neighbour cells in order, each chain in descending
// particle index (slots inside a cell ascend by index, so walk backwards).
template <typename Real>
__device__ int64_t pair_rank(const Cells<Real>& C, int sa, int sb) {
    const Real xa = C.x[sa], ya = C.y[sa];
    const int ia = C.id[sa];
    int cx, cy;
    cell_of(C, xa, ya, cx, cy);
    int64_t r = 0;
    for (int k = 0; k < 9; ++k) {
        const int c = nb_cell(C, cx, cy, k);
        if (c < 0) continue;
        for (int u = (int)C.start[c + 1] - 1; u >= (int)C.start[c]; --u) {
            if (C.id[u] <= ia) continue;
            Real dx, dy;
            const Real d2 = min_image(xa, ya, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy);
            if (d2 < C.cut2) {
                if (u == sb) return r;
                ++r;
            }
        }
    }
    return -1;
}

// Does the faithful pair cap truncate this step's pair list?
template <typename Real>
__device__ __forceinline__ bool cap_on(const Cells<Real>& C) {
    return C.cap_active || (C.cap_dev && C.npairs[0] < C.npairs[1]);
}

// A launch gated on the cap (cap_gate) that has nothing to do this step.
template <typename Real>
__device__ __forceinline__ bool cap_gated_off(const Cells<Real>& C) {
    return C.cap_gate != 0 && (C.cap_gate > 0) != cap_on(C);
}

// Does the pair (sa, sb) exist in the reference's capped pair list?
// sa must hold the lower particle index.
template <typename Real>
__device__ __forceinline__ bool pair_in_list(const Cells<Real>& C, int sa, int sb) {
    if (!cap_on(C)) return true;
    const int ia = C.id[sa];
    const uint64_t off = C.poff[ia], cnt = C.pcnt[ia];
    const uint64_t np = (uint64_t)*C.npairs;
    if (off + cnt <= np) return true;
    if (off >= np) return false;
    return off + (uint64_t)pair_rank(C, sa, sb) < np;
}

template <typename Real>
__device__ __forceinline__ int64_t pair_index(const Cells<Real>& C, int sa, int sb) {
    return (int64_t)C.poff[C.id[sa] + C.poff_base] + pair_rank(C, sa, sb);
}

// Density of a coarse-grid cell holding k sensor particles, as the
// reference's sequential `density += inv_area` leaves it (FP64), or k*inv.
template <typename Real>
__device__ __forceinline__ Real density_of(int k, Real inv) {
    if constexpr (sizeof(Real) == 8) {
        double d = 0.0;
        for (int i = 0; i < k; ++i) d += inv;
        return d;
    } else {
        return (Real)k * inv;
    }
}

// Coarse density grids (feedback rules) as seen by the reaction kernels.
template <typename Real>
struct Grids {
    const int32_t* off;   // [n_grids * ns]
    const int32_t* dims;  // [n_grids][2]
    const Real* s;        // [n_grids][2] cell sizes
    const Real* inv;      // [n_grids] 1/area
    int32_t* counts;      // read (and direct updates)
    int32_t* delta;       // deferred updates
    int ns;
};

template <typename Real>
__device__ __forceinline__ int grid_cell(const Grids<Real>& G, int g, Real x, Real y) {
    const int nx = G.dims[2 * g], ny = G.dims[2 * g + 1];
    const int gx = trunc_cell(x, G.s[2 * g], nx);
    const int gy = trunc_cell(y, G.s[2 * g + 1], ny);
    return gx * ny + gy;
}

// A feedback rule with its grid resolved on the host: the kept-count offset
// of (grid, sensor), the grid's shape, its cell size and 1 / cell area, so
// that a lookup is one count load.
struct FbDev {
    int sensor, grid, inhibition;
    double K, n, basal;
    int cbase, gnx, gny;
    double gsx, gsy, ginv;
};

static FbDev make_fb(const edm_membrane* m, const edm_feedback& f) {
    FbDev d{f.sensor, f.grid, f.inhibition, f.K_half, f.n_hill, f.basal, 0, 1, 1, 1.0, 1.0, 0.0};
    if (f.sensor >= 0 && f.grid >= 0) {
        const edm_grid& g = m->grids[f.grid];
        d.cbase = m->count_off[(size_t)f.grid * m->p.n_species + f.sensor];
        d.gnx = g.nx;
        d.gny = g.ny;
        d.gsx = g.sx;
        d.gsy = g.sy;
        const double area = g.sx * g.sy;
        d.ginv = area > 0.0 ? 1.0 / area : 0.0;
    }
    return d;
}

// Catalysis walks its pairs from the rarer side: the substrates when fewer of
// them than of the (fused) enzymes were alive at upload.  Either side gives
// the same pairs and outcome.
static bool cat_by_substrate(const std::vector<int64_t>& alive, const int* enz, int nf, int sub) {
    if (alive.empty()) return false;
    int64_t ne = 0;
    for (int j = 0; j < nf; ++j) ne += alive[enz[j]];
    return alive[sub] < ne;
}

// A rule without its grid resolved (the batched ensemble binds the grid of
// each replica on the device).
static FbDev to_fb_model(const edm_feedback& f) {
    return FbDev{f.sensor, f.grid, f.inhibition, f.K_half, f.n_hill, f.basal, 0, 1, 1, 1.0, 1.0,
                 0.0};
}

// Grid cell of (x, y) on the rule's grid (int() truncation, clamped).
template <typename Real>
__device__ __forceinline__ int fb_cell(const FbDev& f, Real x, Real y) {
    return trunc_cell(x, (Real)f.gsx, f.gnx) * f.gny + trunc_cell(y, (Real)f.gsy, f.gny);
}

// Hill factor of a grid cell holding k sensor particles.
template <typename Real>
__device__ __forceinline__ Real fb_hill(const FbDev& f, int k) {
    const Real rho = density_of<Real>(k, (Real)f.ginv);
    return hill_factor<Real>(rho, (Real)f.K, (Real)f.n, f.inhibition != 0, (Real)f.basal);
}

template <typename Real>
__device__ __forceinline__ Real fb_factor(const Grids<Real>& G, const FbDev& f, Real x, Real y) {
    if (f.sensor < 0) return (Real)1;
    return fb_hill<Real>(f, G.counts[f.cbase + fb_cell(f, x, y)]);
}

// Update the kept density counts after species s of a particle at (x, y)
// gains (+1) or loses (-1) one member.
template <typename Real>
__device__ __forceinline__ void count_update(const Grids<Real>& G, int n_grids, int s, Real x,
                                             Real y, int d, bool deferred) {
    for (int g = 0; g < n_grids; ++g) {
        const int o = G.off[g * G.ns + s];
        if (o < 0) continue;
        const int c = grid_cell(G, g, x, y);
        atomicAdd(&(deferred ? G.delta : G.counts)[o + c], d);
    }
}

// ----------------------------------------------------------------------
// Cell list: counting sort + per-cell index order + permutation
// ----------------------------------------------------------------------
template <typename Real>
__global__ void k_bin(int n, const Real* __restrict__ x, const Real* __restrict__ y, Real csx,
                      Real csy, int ncx, int ncy, uint32_t* __restrict__ key,
                      uint32_t* __restrict__ rank, uint32_t* __restrict__ cnt, int32_t* counts,
                      int n_counts, const int32_t* __restrict__ id, uint32_t* __restrict__ kid,
                      int nsub, Real inv_h, uint64_t* __restrict__ scan_state, int scan_tiles,
                      uint32_t* __restrict__ scan_ticket) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_counts) counts[s] = 0;  // density counts are rebuilt each step
    // the cell scan that follows starts from cleared tile states and ticket
    if (s < scan_tiles) scan_state[s] = 0ull;
    if (s < 2 && scan_ticket) scan_ticket[s] = 0u;
    const bool live = s < n;          // live lanes are a prefix of the wavefront
    uint32_t c = 0xffffffffu;
    if (live) {
        const Real ys = y[s];
        const int cx = trunc_cell(x[s], csx, ncx);
        const int cy = trunc_cell(ys, csy, ncy);
        c = (uint32_t)(cx * ncy + cy);
        key[s] = c;
        if (kid) {
            // the force sub-row lists' sort key: sub-row (as sub_row computes
            // it) above the particle index
            int b = (int)((ys - (Real)cy * csy) * inv_h);
            b = min(max(b, 0), nsub - 1);
            kid[s] = ((uint32_t)b << 28) | (uint32_t)id[s];
        }
    }
    // Slots are in the previous step's cell order, so equal keys mostly sit
    // in adjacent lanes: one atomic per run of equal keys, ranks inside the
    // run by lane.
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

// Exclusive scan of the n cell counts into the cell starts in one launch:
// tiles of kScanTile counts, each workgroup takes the next tile by ticket
// (so every earlier tile's workgroup has started), publishes its sum, then
// one wavefront looks back over the earlier tiles' published sums, 64 at a
// time, until one carries its inclusive prefix, and publishes its own.  A tile state is (flag << 32 |
// sum): 0 not yet, 1 the tile's own sum, 2 the inclusive prefix.  The
// states and the ticket are cleared by k_bin of the same step.  A wait
// that does not end within kScanSpin reads sets fail[0] and gives up
// (wrong starts rather than a hang; it does not happen while every
// started workgroup runs).
static constexpr int kScanTile = 2048;
static constexpr uint32_t kScanSpin = 1u << 20;

__global__ __launch_bounds__(256) void k_cell_scan(int n, const uint32_t* __restrict__ cnt,
                                                   uint32_t* __restrict__ start,
                                                   uint64_t* state, uint32_t* ticket) {
    __shared__ int s_tile;
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_prefix;
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(&ticket[0], 1u);
    __syncthreads();
    const int tile = s_tile;
    const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
    const int base = tile * kScanTile + (int)threadIdx.x * 8;
    uint32_t v[8];
    if (base + 8 <= n) {
        const uint4 a = *reinterpret_cast<const uint4*>(cnt + base);
        const uint4 b = *reinterpret_cast<const uint4*>(cnt + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = base + j < n ? cnt[base + j] : 0u;
    }
    uint32_t tsum = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) tsum += v[j];
    uint32_t incl = tsum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t wexcl = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint32_t ws = s_wsum[w];
        if (w < wave) wexcl += ws;
        total += ws;
    }
    if (wave == 0) {
        // look back 64 tiles at a time: lane l reads tile (hi - l); the
        // nearest tile with its inclusive prefix ends the walk
        uint32_t prefix = 0;
        if (tile > 0) {
            if (lane == 0)
                __hip_atomic_store(&state[tile], (1ull << 32) | total, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            bool gave_up = false;
            for (int hi = tile - 1; hi >= 0; hi -= 64) {
                const int p = hi - lane;
                uint64_t sv = 2ull << 32;  // before tile 0: an inclusive prefix of 0
                if (p >= 0) {
                    uint32_t spins = 0;
                    do {
                        sv = __hip_atomic_load(&state[p], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    } while ((sv >> 32) == 0u && ++spins <= kScanSpin);
                    if ((sv >> 32) == 0u) gave_up = true;
                }
                const uint64_t inc = __ballot((sv >> 32) == 2u || gave_up);
                const int stop = inc ? __builtin_ctzll(inc) : 63;
                uint32_t val = lane <= stop ? (uint32_t)sv : 0u;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o);
                prefix += val;
                if (inc) break;
            }
            if (__ballot(gave_up) && lane == 0) atomicOr(&ticket[1], 1u);
        }
        if (lane == 0) {
            __hip_atomic_store(&state[tile], (2ull << 32) | (prefix + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            s_prefix = prefix;
        }
    }
    __syncthreads();
    uint32_t run = s_prefix + wexcl + incl - tsum;
    if (base + 8 <= n) {
        uint32_t o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            o[j] = run;
            run += v[j];
        }
        *reinterpret_cast<uint4*>(start + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(start + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (base + j < n) start[base + j] = run;
            run += v[j];
        }
    }
}

__global__ void k_scatter(int n, const uint32_t* __restrict__ key, const uint32_t* __restrict__ rank,
                          const uint32_t* __restrict__ start, uint32_t* __restrict__ perm) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    perm[start[key[s]] + rank[s]] = (uint32_t)s;
}

// One thread per cell: order the cell's particles by particle index (the
// atomic ranks of k_bin are arbitrary) and clear the cell counter.  Cells of
// up to 8 particles are sorted in registers (all loads issued together);
// larger ones by the whole wavefront, one member per lane, each member's
// place its count of smaller indices.  Thread 0 also clears cnt[ncell] when
// clear_next is set (a bucket after the sorted cells).
__global__ void k_sort_cells(int ncell, const uint32_t* __restrict__ start, uint32_t* perm,
                             uint32_t* cnt, const int32_t* __restrict__ id, int clear_next) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool own = c < ncell;  // the whole wavefront stays for the large cells
    if (clear_next && c == 0) cnt[ncell] = 0;
    int lo = 0, hi = 0;
    if (own) {
        cnt[c] = 0;
        lo = (int)start[c];
        hi = (int)start[c + 1];
    }
    const int m = hi - lo;
    constexpr int K = 8;
    if (own && m >= 2 && m <= K) {
        uint32_t v[K];
        int32_t k[K];
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] = j < m ? perm[lo + j] : 0u;
#pragma unroll
        for (int j = 0; j < K; ++j) k[j] = j < m ? id[v[j]] : 0x7fffffff;
#pragma unroll
        for (int r = 0; r < K; ++r) {
#pragma unroll
            for (int j = r & 1; j + 1 < K; j += 2) {
                if (k[j] > k[j + 1]) {
                    const int32_t tk = k[j];
                    k[j] = k[j + 1];
                    k[j + 1] = tk;
                    const uint32_t tv = v[j];
                    v[j] = v[j + 1];
                    v[j + 1] = tv;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (j < m) perm[lo + j] = v[j];
    }
    const int lane = lane_id();
    for (uint64_t big = __ballot(own && m > K); big; big &= big - 1ull) {
        const int L = __builtin_ctzll(big);
        const int blo = __shfl(lo, L), bm = __shfl(m, L);
        if (bm <= 64) {
            const bool in = lane < bm;
            const uint32_t v = in ? perm[blo + lane] : 0u;
            const int32_t k = in ? id[v] : 0x7fffffff;
            int r = 0;
            for (int j = 0; j < bm; ++j) r += __shfl(k, j) < k;
            if (in) perm[blo + r] = v;  // every member was read above
        } else if (lane == L) {
            for (int i = lo + 1; i < hi; ++i) {
                const uint32_t v = perm[i];
                const int32_t vid = id[v];
                int j = i - 1;
                while (j >= lo && id[perm[j]] > vid) {
                    perm[j + 1] = perm[j];
                    --j;
                }
                perm[j + 1] = v;
            }
        }
    }
}

// One thread per new slot: gather every field from its old slot (the
// permutation is close to the identity, so reads stay nearly coalesced).
// Without bonds (kBond false) both bond arrays hold -1 everywhere and are
// left alone.
template <typename Real, bool kBond>
__global__ void k_gather(int n, const uint32_t* __restrict__ perm, DevArrays a,
                         bool write_slot_of) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t s = perm[t];
    ((Real*)a.xt)[t] = ((const Real*)a.x)[s];
    ((Real*)a.yt)[t] = ((const Real*)a.y)[s];
    a.sp_t[t] = a.sp[s];
    a.fl_t[t] = a.fl[s];
    const int32_t pid = a.id[s];
    a.id_t[t] = pid;
    if (kBond) a.bond_t[t] = a.bond[s];
    a.ctmc_t[t] = a.ctmc[s];
    a.dye_t[t] = a.dye[s];
    if (write_slot_of) a.slot_of[pid] = t;
}

__global__ void k_slot_of(int n, const int32_t* __restrict__ id, int32_t* __restrict__ slot_of) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) slot_of[id[t]] = t;
}

int dyn_ensure_slot_of(edm_membrane* m) {
    if (m->slot_of_valid || m->n == 0) return 0;
    hipLaunchKernelGGL(k_slot_of, dim3(nblk(m->n)), dim3(kBlock), 0, m->stream, m->n,
                       (const int32_t*)m->a.id, m->a.slot_of);
    EDM_HIP(hipGetLastError());
    m->slot_of_valid = true;
    return 0;
}

// Work lists, rebuilt after every sort: the slots of the particles the
// per-particle passes can touch, of the dye carriers and of the entities a
// catalysis walk starts from, so that those kernels run one listed slot per
// thread.  The ensemble builds its lists by replica (wl_build_block below,
// one atomic per chunk and list on the replica's counter); one membrane's
// lists need no atomics: block b of the build lists its kWlSlots slots into
// its own region [b * kWlSlots, ...) of every list, in slot order, and
// records how many (WlView).
struct WlBuild {
    WlSpec spec[kWlMaxLists];
    int32_t* list[kWlMaxLists];  // region b at b * kWlSlots
    uint32_t* cnt[kWlMaxLists];  // [blocks]
    int n_lists;
};

__device__ __forceinline__ bool wl_member(const WlSpec& S, int8_t sp, uint8_t fl) {
    return ((S.species >> sp) & 1u) && (fl & S.flags) == S.flags;
}

// Build block `blk` (256 threads; every thread of the block calls it).
__device__ __forceinline__ void wl_build_region(int n, const int8_t* __restrict__ sp,
                                               const uint8_t* __restrict__ fl, const WlBuild& W,
                                               int blk) {
    static_assert(kWlPer * kWlMaxLists <= 64, "membership bits of a thread");
    constexpr int kW = 4;  // wavefronts per block
    __shared__ int s_off[kWlMaxLists][kWlPer * kW];  // members per (row, wave), then offsets
    const int beg = blk * kWlSlots, end = min(n, beg + kWlSlots);
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    // row k of the block holds slots beg + k * 256 + thread (coalesced
    // loads); bit k * kWlMaxLists + L: that slot is a member of list L
    uint64_t memb = 0;
#pragma unroll
    for (int k = 0; k < kWlPer; ++k) {
        const int t = beg + k * 256 + (int)threadIdx.x;
        if (t < end) {
            const int8_t vs = sp[t];
            const uint8_t vf = fl[t];
            for (int L = 0; L < W.n_lists; ++L)
                if (wl_member(W.spec[L], vs, vf)) memb |= 1ull << (k * kWlMaxLists + L);
        }
    }
    for (int L = 0; L < W.n_lists; ++L) {
#pragma unroll
        for (int k = 0; k < kWlPer; ++k) {
            const uint64_t b = __ballot((memb >> (k * kWlMaxLists + L)) & 1ull);
            if (lane == 0) s_off[L][k * kW + w] = __popcll(b);
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < W.n_lists) {  // exclusive prefix, one thread per list
        const int L = (int)threadIdx.x;
        int run = 0;
        for (int i = 0; i < kWlPer * kW; ++i) {
            const int c = s_off[L][i];
            s_off[L][i] = run;
            run += c;
        }
        W.cnt[L][blk] = (uint32_t)run;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (int L = 0; L < W.n_lists; ++L) {
        int32_t* out = W.list[L] + (size_t)blk * kWlSlots;
#pragma unroll
        for (int k = 0; k < kWlPer; ++k) {
            const bool mem = (memb >> (k * kWlMaxLists + L)) & 1ull;
            const uint64_t b = __ballot(mem);
            if (mem) out[s_off[L][k * kW + w] + __popcll(b & below)] = beg + k * 256 + (int)threadIdx.x;
        }
    }
}

__global__ __launch_bounds__(256) void k_wl_build(int n, const int8_t* __restrict__ sp,
                                                  const uint8_t* __restrict__ fl, WlBuild W) {
    wl_build_region(n, sp, fl, W, (int)blockIdx.x);
}

// Ensemble list build (one replica's chunk): the chunk's members of both
// lists, appended to the replica's regions with one atomic per list.
struct WlRun {
    int32_t* list;
    uint32_t* cnt;  // entries appended this step
    uint32_t off;   // first entry of this list region
    uint32_t cap;   // entries the region holds
};

// Row k of the block holds slots beg + k * 256 + thread (coalesced loads); a
// member's place in the block's list is its (row, wave, lane) rank.  Every
// thread of the block (256) calls it.
template <typename PredP, typename PredD>
__device__ __forceinline__ void wl_build_block(int beg, int end, const PredP& in_pass,
                                               const PredD& in_dye, WlRun Rp, WlRun Rd) {
    constexpr int kW = 4;  // wavefronts per block
    __shared__ int32_t s_list[2][kWlSlots];
    __shared__ int s_cnt[2][kWlPer * kW];  // members per (row, wave), then their offsets
    __shared__ uint32_t s_base[2];
    const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
    uint64_t bal[2][kWlPer];
#pragma unroll
    for (int k = 0; k < kWlPer; ++k) {
        const int t = beg + k * 256 + (int)threadIdx.x;
        const bool live = t < end;
        bal[0][k] = __ballot(live && in_pass(t));
        bal[1][k] = __ballot(live && in_dye(t));
        if (lane == 0) {
            s_cnt[0][k * kW + w] = __popcll(bal[0][k]);
            s_cnt[1][k * kW + w] = __popcll(bal[1][k]);
        }
    }
    __syncthreads();
    if (threadIdx.x < 2) {  // exclusive prefix of the counts, list by list
        int run = 0;
        for (int i = 0; i < kWlPer * kW; ++i) {
            const int c = s_cnt[threadIdx.x][i];
            s_cnt[threadIdx.x][i] = run;
            run += c;
        }
        s_base[threadIdx.x] = run;  // the block's members, for now
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kWlPer; ++k) {
        const int t = beg + k * 256 + (int)threadIdx.x;
#pragma unroll
        for (int L = 0; L < 2; ++L)
            if ((bal[L][k] >> lane) & 1ull)
                s_list[L][s_cnt[L][k * kW + w] + __popcll(bal[L][k] & below)] = t;
    }
    __syncthreads();
    const int mp = (int)s_base[0], md = (int)s_base[1];
    __syncthreads();
    if (threadIdx.x == 0) {
        s_base[0] = mp ? atomicAdd(Rp.cnt, (uint32_t)mp) : 0u;
        s_base[1] = md ? atomicAdd(Rd.cnt, (uint32_t)md) : 0u;
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i < mp; i += 256) {
        const uint32_t k = s_base[0] + (uint32_t)i;
        if (k < Rp.cap) Rp.list[Rp.off + k] = s_list[0][i];
    }
    for (int i = (int)threadIdx.x; i < md; i += 256) {
        const uint32_t k = s_base[1] + (uint32_t)i;
        if (k < Rd.cap) Rd.list[Rd.off + k] = s_list[1][i];
    }
}


// Reference pair-list bookkeeping: per particle index, the number of
// partners j > i with d^2 < cutoff^2 (dynamics.py:237-276; dormant
// particles included, as in the reference).
template <typename Real>
__device__ __forceinline__ uint64_t pair_count_one(const Cells<Real>& C, int t) {
    const Real xt = C.x[t], yt = C.y[t];
    const int it = C.id[t];
    int cx, cy;
    cell_of(C, xt, yt, cx, cy);
    uint64_t cnt = 0;
    for (int k = 0; k < 9; ++k) {
        const int c = nb_cell(C, cx, cy, k);
        if (c < 0) continue;
        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
            if (C.id[u] <= it) continue;
            Real dx, dy;
            if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) < C.cut2) ++cnt;
        }
    }
    return cnt;
}

// Pairs (i < j within the cutoff) listed from each particle, by particle
// index: the reference's pair-list lengths per particle (dynamics.py:71-278).
template <typename Real>
__global__ void k_pair_count(int n, Cells<Real> C, uint64_t* __restrict__ pcnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    pcnt[C.id[t] + C.poff_base] = pair_count_one(C, t);
}

__global__ void k_pair_total(int n, const uint64_t* pcnt, const uint64_t* poff, int64_t cap,
                             int64_t* scal) {
    const uint64_t total = n > 0 ? poff[n - 1] + pcnt[n - 1] : 0;
    scal[1] = (int64_t)total;
    scal[0] = (cap > 0 && (int64_t)total > cap) ? cap : (int64_t)total;
}

// Materialise the reference pair list (validation / parity helper).
template <typename Real>
__global__ void k_pair_list(int n, Cells<Real> C, int32_t* __restrict__ out, int64_t cap) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Real xt = C.x[t], yt = C.y[t];
    const int it = C.id[t];
    int cx, cy;
    cell_of(C, xt, yt, cx, cy);
    int64_t p = (int64_t)C.poff[it];
    const int64_t np = *C.npairs;
    for (int k = 0; k < 9; ++k) {
        const int c = nb_cell(C, cx, cy, k);
        if (c < 0) continue;
        for (int u = (int)C.start[c + 1] - 1; u >= (int)C.start[c]; --u) {
            if (C.id[u] <= it) continue;
            Real dx, dy;
            if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) < C.cut2) {
                if (p < np && p < cap) {
                    out[2 * p] = it;
                    out[2 * p + 1] = C.id[u];
                }
                ++p;
            }
        }
    }
}

// ----------------------------------------------------------------------
// Fused forces + Ermak-McCammon step (dynamics.py:666-816, 323-406, 441-519)
// ----------------------------------------------------------------------
template <typename Real>
struct ForceParams {
    const Real *sig, *eps, *kh, *Dbase;
    int ns;
    bool wca, harm;
    int dfield;  // 0 regions, 1 sinusoidal
    int n_regions;
    const Real *rcx, *rcy, *rr, *rf;
    Real smooth_w, sin_amp, sin_kx, sin_ky;
    Real dt, max_disp;
    Real rint2;  // no species pair interacts at d^2 >= rint2
};

// Pair force over r on the particle the separation (dx, dy) points to;
// kind 0 = WCA, 1 = harmonic.  Returns false if the pair does not interact.
template <typename Real>
__device__ __forceinline__ bool pair_force(const ForceParams<Real>& F, int kind, int si, int su,
                                           Real dsq, Real& f) {
    const Real s = F.sig[si * F.ns + su];
    if (kind == 0) {
        const Real e = F.eps[si * F.ns + su];
        if (s <= (Real)0 || e <= (Real)0) return false;
        const Real rcut2 = s * s * (Real)1.2599210498948732;
        if (dsq >= rcut2 || dsq < (Real)1e-30) return false;
        // FP32: hardware reciprocals (about 1 ulp, inside the FP32 bound);
        // FP64: the reference's expressions, correctly rounded
        constexpr bool f32 = sizeof(Real) == 4;
        const Real inv_r2 = f32 ? (Real)__builtin_amdgcn_rcpf((float)dsq) : (Real)1 / dsq;
        const Real s2 = s * s * inv_r2;
        const Real s6 = s2 * s2 * s2;
        const Real s12 = s6 * s6;
        Real fr = (Real)24 * e * ((Real)2 * s12 - s6) * inv_r2;
        const Real fmax = f32 ? (Real)2400 * e * (Real)__builtin_amdgcn_rcpf((float)(s * s))
                              : (Real)2400 * e / (s * s);
        if (fr > fmax) fr = fmax;
        f = fr;
        return true;
    }
    const Real k = F.kh[si * F.ns + su];
    if (s <= (Real)0 || k <= (Real)0) return false;
    if (dsq >= s * s || dsq < (Real)1e-30) return false;
    if constexpr (sizeof(Real) == 4) {
        // k (s - r) / r = k (s / r - 1) with the hardware reciprocal square root
        f = k * (s * (Real)__builtin_amdgcn_rsqf((float)dsq) - (Real)1);
    } else {
        const Real r = sqrt(dsq);
        const Real fmag = k * (s - r);
        f = fmag / r;
    }
    return true;
}

template <typename Real>
__device__ __forceinline__ void dfield_eval(const ForceParams<Real>& F, int s, Real x, Real y,
                                            Real& D, Real& gx, Real& gy) {
    const Real base = F.Dbase[s];
    if (F.dfield == 1) {
        Real sx, cx, sy, cy;
        if constexpr (sizeof(Real) == 8) {
            det_sincos(F.sin_kx * x, sx, cx);
            det_sincos(F.sin_ky * y, sy, cy);
        } else {
            sx = sin(F.sin_kx * x);
            cx = cos(F.sin_kx * x);
            sy = sin(F.sin_ky * y);
            cy = cos(F.sin_ky * y);
        }
        D = base * ((Real)1 + F.sin_amp * sx * sy);
        gx = base * F.sin_amp * F.sin_kx * cx * sy;
        gy = base * F.sin_amp * F.sin_ky * sx * cy;
        return;
    }
    Real factor = (Real)1, gfx = (Real)0, gfy = (Real)0;
    for (int r = 0; r < F.n_regions; ++r) {
        const Real dxc = x - F.rcx[r], dyc = y - F.rcy[r];
        const Real dsq = dxc * dxc + dyc * dyc;
        const Real d = dsq > (Real)1e-30 ? sqrt(dsq) : (Real)1e-15;
        const Real arg = (d - F.rr[r]) / F.smooth_w;
        Real sig, dsig;
        if (arg > (Real)100) {
            sig = (Real)0;
            dsig = (Real)0;
        } else if (arg < -(Real)100) {
            sig = (Real)1;
            dsig = (Real)0;
        } else {
            const Real e = exp(arg);
            const Real den = (Real)1 + e;
            sig = (Real)1 / den;
            dsig = -e / (den * den * F.smooth_w);
        }
        const Real local = (Real)1 + (F.rf[r] - (Real)1) * sig;
        if (sig > (Real)0.01) {
            factor = local;
            const Real inv_d = (Real)1 / d;
            gfx = (F.rf[r] - (Real)1) * dsig * dxc * inv_d;
            gfy = (F.rf[r] - (Real)1) * dsig * dyc * inv_d;
            break;
        }
    }
    D = base * factor;
    gx = base * gfx;
    gy = base * gfy;
}

// Ermak-McCammon move of one alive particle (dynamics.py:359-406): Ito
// drift grad(D) dt, noise sqrt(2 D dt) xi, force drift D F dt, displacement
// cap on |dr|, then periodic wrap (Python modulo) or one reflection.
template <typename Real>
__device__ __forceinline__ void move_particle(const Cells<Real>& C, const ForceParams<Real>& F,
                                              int si, int it, Real xt, Real yt, Real Fx, Real Fy,
                                              const double* __restrict__ noise_hook, Rng rng,
                                              uint32_t step, Real& nxp, Real& nyp) {
    Real D, gx, gy;
    dfield_eval(F, si, xt, yt, D, gx, gy);
    Real n0, n1;
    if (noise_hook) {
        n0 = (Real)noise_hook[2 * (int64_t)it];
        n1 = (Real)noise_hook[2 * (int64_t)it + 1];
    } else {
        normal2<Real>(rng.draw(step, (uint32_t)it, 0u, kTagNoise), n0, n1);
    }
    const Real sn = sqrt((Real)2 * D * F.dt);
    Real dx = gx * F.dt;
    Real dy = gy * F.dt;
    dx += sn * n0;
    dy += sn * n1;
    dx += D * Fx * F.dt;
    dy += D * Fy * F.dt;
    const Real dr2 = dx * dx + dy * dy;
    if (dr2 > F.max_disp * F.max_disp) {
        const Real dr = sqrt(dr2);
        const Real sc = F.max_disp / dr;
        dx *= sc;
        dy *= sc;
    }
    Real px_ = xt + dx, py_ = yt + dy;
    if (C.px) px_ = py_mod(px_, C.Lx);
    else {
        if (px_ < (Real)0) px_ = -px_;
        else if (px_ >= C.Lx) px_ = (Real)2 * C.Lx - px_;
    }
    if (C.py) py_ = py_mod(py_, C.Ly);
    else {
        if (py_ < (Real)0) py_ = -py_;
        else if (py_ >= C.Ly) py_ = (Real)2 * C.Ly - py_;
    }
    nxp = px_;
    nyp = py_;
}

// Density counts of the kept (sensor) species at the post-move position;
// every lane of the wavefront must call this together.
template <typename Real>
__device__ __forceinline__ void add_density(const Grids<Real>& G, int n_grids, bool take_any,
                                            int si, Real x, Real y) {
    for (int g = 0; g < n_grids; ++g) {
        int o = -1;
        if (take_any) o = G.off[g * G.ns + si];
        const bool take = take_any && o >= 0;
        const int idx = take ? o + grid_cell(G, g, x, y) : 0;
        wave_aggregated_add(G.counts, idx, take, 1);
    }
}

// kOrdered: sum pair forces in the reference's pair order (all partners
// with a lower particle index ascending, then the particle's own
// enumeration order), WCA over the whole list first, then harmonic — the
// order compute_forces_wca / compute_forces_harmonic accumulate in.
template <typename Real, bool kOrdered>
__global__ __launch_bounds__(256) void k_force_bd(int n, Cells<Real> C, ForceParams<Real> F,
                                                  Real* __restrict__ xn, Real* __restrict__ yn,
                                                  const double* __restrict__ noise_hook, Rng rng,
                                                  StepRef sref, Grids<Real> G, int n_grids,
                                                  uint32_t keep_mask) {
    if (cap_gated_off(C)) return;  // uniform across the launch
    const uint32_t step = sref.get();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool take_any = false;
    Real nxp = 0, nyp = 0;
    int si = 0;
    if (t < n) {
        const Real xt = C.x[t], yt = C.y[t];
        si = C.sp[t];
        const bool alive = (C.fl[t] & EDM_FLAG_ALIVE) != 0;
        nxp = xt;
        nyp = yt;
        if (alive) {
            const int it = C.id[t];
            int cx, cy;
            cell_of(C, xt, yt, cx, cy);
            Real Fx = 0, Fy = 0;
            for (int kind = 0; kind < 2; ++kind) {
                if (kind == 0 && !F.wca) continue;
                if (kind == 1 && !F.harm) continue;
                if constexpr (kOrdered) {
                    // partners with lower index, ascending
                    int last = -1;
                    for (;;) {
                        int best = INT_MAX, bslot = -1;
                        for (int k = 0; k < 9; ++k) {
                            const int c = nb_cell(C, cx, cy, k);
                            if (c < 0) continue;
                            for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                                const int iu = C.id[u];
                                if (iu >= it || iu <= last || iu >= best) continue;
                                Real dx, dy;
                                if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx,
                                              dy) < C.cut2) {
                                    best = iu;
                                    bslot = u;
                                }
                            }
                        }
                        if (bslot < 0) break;
                        last = best;
                        if (!(C.fl[bslot] & EDM_FLAG_ALIVE)) continue;
                        if (!pair_in_list(C, bslot, t)) continue;
                        Real dx, dy;
                        const Real dsq = min_image(xt, yt, C.x[bslot], C.y[bslot], C.Lx, C.Ly,
                                                   C.px, C.py, dx, dy);
                        Real f;
                        if (pair_force(F, kind, C.sp[bslot], si, dsq, f)) {
                            Fx += f * dx;
                            Fy += f * dy;
                        }
                    }
                    // partners with higher index, reference enumeration order
                    for (int k = 0; k < 9; ++k) {
                        const int c = nb_cell(C, cx, cy, k);
                        if (c < 0) continue;
                        for (int u = (int)C.start[c + 1] - 1; u >= (int)C.start[c]; --u) {
                            if (C.id[u] <= it) continue;
                            Real dx, dy;
                            const Real dsq = min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px,
                                                       C.py, dx, dy);
                            if (dsq >= C.cut2) continue;
                            if (!(C.fl[u] & EDM_FLAG_ALIVE)) continue;
                            if (!pair_in_list(C, t, u)) continue;
                            Real f;
                            if (pair_force(F, kind, si, C.sp[u], dsq, f)) {
                                Fx += f * dx;
                                Fy += f * dy;
                            }
                        }
                    }
                } else {
                    for (int k = 0; k < 9; ++k) {
                        const int c = nb_cell(C, cx, cy, k);
                        if (c < 0) continue;
                        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
                            if (u == t) continue;
                            Real dx, dy;
                            const Real dsq = min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px,
                                                       C.py, dx, dy);
                            if (dsq >= C.cut2) continue;
                            if (!(C.fl[u] & EDM_FLAG_ALIVE)) continue;
                            if (cap_on(C)) {
                                const bool lower = C.id[u] < it;
                                if (!(lower ? pair_in_list(C, u, t) : pair_in_list(C, t, u)))
                                    continue;
                            }
                            Real f;
                            if (pair_force(F, kind, si, C.sp[u], dsq, f)) {
                                Fx += f * dx;
                                Fy += f * dy;
                            }
                        }
                    }
                }
            }
            move_particle(C, F, si, it, xt, yt, Fx, Fy, noise_hook, rng, step, nxp, nyp);
            take_any = ((keep_mask >> si) & 1u) != 0;
        }
        xn[t] = nxp;
        yn[t] = nyp;
    }
    add_density(G, n_grids, take_any, si, nxp, nyp);
}

#include "edm_tiles.h"

// ----------------------------------------------------------------------
// Force sub-row lists (dense membranes, steps without the reference pair
// order): the cell sort orders every cell by (force sub-row, particle
// index) and writes where each sub-row starts (fstart: kMaxSub entries per
// cell, those past the cell's nsub sub-rows holding the cell's end), so
// that a particle's force candidates in a column are one contiguous slot
// range — its own sub-row and the two next to it, which run on into the
// cells above and below — read straight from global memory with no strip
// staging.  Sub-rows are at least 1.01 x the longest potential range tall,
// so every pair within range is at most one sub-row apart (the 1 % margin
// is far above the rounding of the sub-row index, as for the strips).
// ----------------------------------------------------------------------
__global__ void k_sort_cells_fine(int ncell, const uint32_t* __restrict__ start, uint32_t* perm,
                                  uint32_t* cnt, const uint32_t* __restrict__ kid, int nsub,
                                  uint32_t* __restrict__ fstart) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool own = c < ncell;  // the whole wavefront stays for the large cells
    int lo = 0, hi = 0;
    if (own) {
        cnt[c] = 0;
        lo = (int)start[c];
        hi = (int)start[c + 1];
    }
    const int m = hi - lo;
    // (sub-row << 28 | particle index), written by k_bin
    auto key = [&](uint32_t v) { return (int32_t)kid[v]; };
    constexpr int K = 8;
    int nb[kMaxSub + 1];
#pragma unroll
    for (int q = 0; q <= kMaxSub; ++q) nb[q] = 0;  // members below sub-row q
    if (own && m <= K) {
        uint32_t v[K];
        int32_t k[K];
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] = j < m ? perm[lo + j] : 0u;
#pragma unroll
        for (int j = 0; j < K; ++j) k[j] = j < m ? key(v[j]) : 0x7fffffff;
        if (m >= 2) {
#pragma unroll
            for (int r = 0; r < K; ++r) {
#pragma unroll
                for (int j = r & 1; j + 1 < K; j += 2) {
                    if (k[j] > k[j + 1]) {
                        const int32_t tk = k[j];
                        k[j] = k[j + 1];
                        k[j + 1] = tk;
                        const uint32_t tv = v[j];
                        v[j] = v[j + 1];
                        v[j + 1] = tv;
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (j < m) perm[lo + j] = v[j];
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (j >= m) break;
            const int b = (int)((uint32_t)k[j] >> 28);
#pragma unroll
            for (int q = 1; q <= kMaxSub; ++q) nb[q] += b < q;
        }
    }
    // Cells of more than K (about 1 % at a mean of 3 per cell): the
    // wavefront sorts each in turn with one member per lane, each member's
    // place its count of smaller keys (keys are distinct), instead of one
    // lane inserting with a dependent pair of loads per step.
    const int lane = lane_id();
    for (uint64_t big = __ballot(own && m > K); big; big &= big - 1ull) {
        const int L = __builtin_ctzll(big);
        const int blo = __shfl(lo, L), bm = __shfl(m, L);
        if (bm <= 64) {
            const bool in = lane < bm;
            const uint32_t v = in ? perm[blo + lane] : 0u;
            const int32_t k = in ? key(v) : 0x7fffffff;
            int r = 0;
            for (int j = 0; j < bm; ++j) r += __shfl(k, j) < k;
            if (in) perm[blo + r] = v;  // every member was read above
            const int b = in ? (int)((uint32_t)k >> 28) : kMaxSub;
#pragma unroll
            for (int q = 1; q <= kMaxSub; ++q) {
                const int below = __popcll(__ballot(b < q));
                if (lane == L) nb[q] = below;
            }
        } else if (lane == L) {
            for (int i = lo + 1; i < hi; ++i) {
                const uint32_t v = perm[i];
                const int32_t vk = key(v);
                int j = i - 1;
                while (j >= lo && key(perm[j]) > vk) {
                    perm[j + 1] = perm[j];
                    --j;
                }
                perm[j + 1] = v;
            }
            for (int i = lo; i < hi; ++i) {
                const int b = (int)(kid[perm[i]] >> 28);
#pragma unroll
                for (int q = 1; q <= kMaxSub; ++q) nb[q] += b < q;
            }
        }
    }
    if (!own) return;
    // the cell's 8 entries as two 16-byte stores (rows of 32 bytes, so a
    // wavefront writes one contiguous 2 KiB run)
    static_assert(kMaxSub == 8, "fstart rows are two uint4");
    uint4* row = reinterpret_cast<uint4*>(fstart + (size_t)c * kMaxSub);
    row[0] = make_uint4((uint32_t)(lo + nb[0]), (uint32_t)(lo + nb[1]), (uint32_t)(lo + nb[2]),
                        (uint32_t)(lo + nb[3]));
    row[1] = make_uint4((uint32_t)(lo + nb[4]), (uint32_t)(lo + nb[5]), (uint32_t)(lo + nb[6]),
                        (uint32_t)(lo + nb[7]));
    if (c == ncell - 1) fstart[(size_t)ncell * kMaxSub] = (uint32_t)hi;
}

// Forces + Ermak-McCammon move over the sub-row lists, one particle per
// thread.  Candidates by column (left to right), then by (row, sub-row,
// particle index) with a periodic wrap's far row where the strip kernel
// stages it (below row 0, above the last row): the strip kernel's order, so
// both kernels sum every particle's forces in the same order and give the
// same bits.
template <typename Real>
__global__ __launch_bounds__(256) void k_force_fine(int n, Cells<Real> C, ForceParams<Real> F,
                                                    Real* __restrict__ xn, Real* __restrict__ yn,
                                                    Rng rng, StepRef sref, Grids<Real> G,
                                                    int n_grids, uint32_t keep_mask, int nsub,
                                                    Real inv_h,
                                                    const uint32_t* __restrict__ fstart,
                                                    WlBuild W, int wl_nb,
                                                    const int8_t* __restrict__ wl_sp,
                                                    const uint8_t* __restrict__ wl_fl) {
    // the first wl_nb blocks build the step's work lists beside the forces
    // (they read only species and flags, which this kernel does not write)
    if ((int)blockIdx.x < wl_nb) {
        wl_build_region(n, wl_sp, wl_fl, W, (int)blockIdx.x);
        return;
    }
    if (cap_gated_off(C)) return;  // uniform across the launch
    const uint32_t step = sref.get();
    const int t = ((int)blockIdx.x - wl_nb) * (int)blockDim.x + (int)threadIdx.x;
    bool take_any = false;
    Real nxp = 0, nyp = 0;
    int si = 0;
    if (t < n) {
        const Real xt = C.x[t], yt = C.y[t];
        si = C.sp[t];
        nxp = xt;
        nyp = yt;
        if (C.fl[t] & EDM_FLAG_ALIVE) {
            const Real rl2 = min(C.cut2, F.rint2);
            int cx, cy;
            cell_of(C, xt, yt, cx, cy);
            const int b = sub_row(C, yt, cy, nsub, inv_h);
            Real Fx = 0, Fy = 0;
            for (int dc = -1; dc <= 1; ++dc) {
                int gx = cx + dc;
                if (C.px) gx = gx < 0 ? gx + C.ncx : (gx >= C.ncx ? gx - C.ncx : gx);
                else if (gx < 0 || gx >= C.ncx) continue;
                const int col = gx * C.ncy;
                // sub-rows b-1 .. b+1 of row cy: one run of slots, from the
                // start of the first to the end of the last (they run on
                // into the rows above and below), plus the far row's end
                // sub-row when the run wraps round a periodic y edge.  Keys
                // are cell * S + sub-row
                constexpr int S = kMaxSub;
                const int kc = (col + cy) * S;
                int k0 = b > 0 ? kc + b - 1 : kc - S + nsub - 1;  // first sub-row's key
                int k1 = b < nsub - 1 ? kc + b + 1 : kc + S;      // last sub-row's key
                int below = -1, above = -1;  // wrapped far-row sub-row keys
                if (b == 0 && cy == 0) {
                    k0 = kc;
                    if (C.py) below = (col + C.ncy - 1) * S + nsub - 1;
                } else if (b == nsub - 1 && cy == C.ncy - 1) {
                    k1 = kc + nsub - 1;
                    if (C.py) above = col * S;
                }
                for (int r = 0; r < 3; ++r) {
                    int u0, u1;
                    const uint32_t* fs = fstart;
                    // the slot after a sub-row: the next key's start (past
                    // the cell's nsub sub-rows, entries hold the cell's end)
                    if (r == 1) {
                        u0 = (int)fs[k0];
                        u1 = (int)fs[k1 + 1];
                    } else {
                        const int e = r == 0 ? below : above;
                        if (e < 0) continue;
                        u0 = (int)fs[e];
                        u1 = (int)fs[e + 1];
                    }
                    for (int u = u0; u < u1; ++u) {
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
            move_particle(C, F, si, C.id[t] - C.kbase, xt, yt, Fx, Fy,
                          (const double*)nullptr, rng, step, nxp, nyp);
            take_any = ((keep_mask >> si) & 1u) != 0;
        }
        xn[t] = nxp;
        yn[t] = nyp;
    }
    add_density(G, n_grids, take_any, si, nxp, nyp);
}

// ----------------------------------------------------------------------
// Reaction kernels
// ----------------------------------------------------------------------
struct Draws {
    const double* base;  // the step's hook draws (nullptr = Philox)
    int64_t off;         // offset of this reaction's arrays
    int64_t n_pairs;     // length of the per-pair array (BIND, CAT)
    int n;               // particles
};

template <typename Real>
struct Post {
    Real *xn, *yn;  // post-move positions (writable)
    int8_t* sp;
    uint8_t* fl;
    int32_t* bond;
    const int32_t* slot_of;
};

__device__ __forceinline__ bool cas_byte(int8_t* base, int idx, int8_t expect, int8_t val) {
    const uintptr_t addr = (uintptr_t)(base + idx);
    uint32_t* w = (uint32_t*)(addr & ~(uintptr_t)3);
    const int sh = (int)(addr & 3) * 8;
    uint32_t old = *((volatile uint32_t*)w);
    for (;;) {
        if ((int8_t)((old >> sh) & 0xffu) != expect) return false;
        const uint32_t nw = (old & ~(0xffu << sh)) | ((uint32_t)(uint8_t)val << sh);
        const uint32_t prev = atomicCAS(w, old, nw);
        if (prev == old) return true;
        old = prev;
    }
}

// Candidate binding edges A(ra) - B(rb) of particle t (of species ra):
// both alive and free, in the (capped) pair list, within contact after the
// move, and accepted by their per-pair uniform (dynamics.py:529-606).
// Philox keys and edge keys use particle indices relative to C.kbase.
template <typename Real>
__device__ __forceinline__ void bind_edges_one(const Cells<Real>& C, const Post<Real>& P, int t,
                                               int ra, int rb, Real contact2, Real prob,
                                               const Draws& dr, const Rng& rng, uint32_t step,
                                               uint32_t rxi, unsigned long long* ekey,
                                               int32_t* ea, int32_t* eb, unsigned int* ne,
                                               unsigned int cap) {
    if (C.sp[t] != ra || !(C.fl[t] & EDM_FLAG_ALIVE) || P.bond[t] >= 0) return;
    const Real xt = C.x[t], yt = C.y[t];
    const int it = C.id[t];
    int cx, cy;
    cell_of(C, xt, yt, cx, cy);
    for (int k = 0; k < 9; ++k) {
        const int c = nb_cell(C, cx, cy, k);
        if (c < 0) continue;
        for (int u = (int)C.start[c]; u < (int)C.start[c + 1]; ++u) {
            if (u == t || C.sp[u] != rb) continue;
            const int iu = C.id[u];
            if (ra == rb && iu < it) continue;  // A is the lower index of an A+A pair
            if (!(C.fl[u] & EDM_FLAG_ALIVE) || P.bond[u] >= 0) continue;
            Real dx, dy;
            if (min_image(xt, yt, C.x[u], C.y[u], C.Lx, C.Ly, C.px, C.py, dx, dy) >= C.cut2)
                continue;
            const bool tlow = it < iu;
            const int sa = tlow ? t : u, sb = tlow ? u : t;
            if (!pair_in_list(C, sa, sb)) continue;
            if (min_image(P.xn[t], P.yn[t], P.xn[u], P.yn[u], C.Lx, C.Ly, C.px, C.py, dx, dy) >
                contact2)
                continue;
            unsigned long long key;
            double uu;
            if (dr.base) {
                const int64_t p = pair_index(C, sa, sb);
                key = (unsigned long long)p;
                uu = dr.base[dr.off + p];
            } else {
                const uint32_t lo = (uint32_t)(min(it, iu) - C.kbase);
                const uint32_t hi = (uint32_t)(max(it, iu) - C.kbase);
                key = ((unsigned long long)lo << 32) | hi;
                uu = (double)Uni<Real>::first(rng.draw(step, lo, hi, (kTagBind << 16) | rxi));
            }
            if (uu < (double)prob) {
                const unsigned int e = atomicAdd(ne, 1u);
                if (e < cap) {
                    ekey[e] = key;
                    ea[e] = t;
                    eb[e] = u;
                }
            }
        }
    }
}

template <typename Real>
__global__ void k_bind_edges(int n, Cells<Real> C, Post<Real> P, int ra, int rb, Real contact2,
                             Real prob, Draws dr, Rng rng, StepRef sref, uint32_t rxi,
                             unsigned long long* ekey, int32_t* ea, int32_t* eb, unsigned int* ne,
                             unsigned int cap) {
    const uint32_t step = sref.get();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    bind_edges_one(C, P, t, ra, rb, contact2, prob, dr, rng, step, rxi, ekey, ea, eb, ne, cap);
}

// Greedy matching in key order = the reference's sequential pass: an edge
// is accepted when it is the smallest live edge at both endpoints; repeat
// until no live edge is left.  One workgroup over one membrane's edges
// (best: per slot).  ne[1] keeps the largest count of accepted edges that
// exceeded the buffer, for the host to report.
template <typename Real>
__device__ __forceinline__ void bind_match_block(const Post<Real>& P,
                                                 const int32_t* __restrict__ id, int ra, int rb,
                                                 int prod, const unsigned long long* ekey,
                                                 const int32_t* ea, const int32_t* eb,
                                                 unsigned int* ne, unsigned int cap,
                                                 unsigned long long* best,
                                                 unsigned long long* ctr, const Grids<Real>& G,
                                                 int n_grids) {
    const unsigned int m = min(*ne, cap);
    // more accepted edges than the buffer holds: the matching below would
    // differ from the reference's sequential pass, so keep the largest such
    // count in ne[1] for the host to report (EDM_ECAPACITY)
    if (threadIdx.x == 0 && *ne > cap) atomicMax(&ne[1], *ne);
    volatile int8_t* vsp = P.sp;
    volatile uint8_t* vfl = P.fl;
    volatile int32_t* vbd = P.bond;
    for (unsigned int round = 0; round <= m; ++round) {
        for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
            const int a = ea[e], b = eb[e];
            const bool ok = vsp[a] == ra && vsp[b] == rb && (vfl[a] & EDM_FLAG_ALIVE) &&
                            (vfl[b] & EDM_FLAG_ALIVE) && vbd[a] < 0 && vbd[b] < 0;
            if (ok) {
                atomicExch(&best[a], ~0ull);
                atomicExch(&best[b], ~0ull);
            }
        }
        __threadfence();
        __syncthreads();
        for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
            const int a = ea[e], b = eb[e];
            const bool ok = vsp[a] == ra && vsp[b] == rb && (vfl[a] & EDM_FLAG_ALIVE) &&
                            (vfl[b] & EDM_FLAG_ALIVE) && vbd[a] < 0 && vbd[b] < 0;
            if (ok) {
                atomicMin(&best[a], ekey[e]);
                atomicMin(&best[b], ekey[e]);
            }
        }
        __threadfence();
        __syncthreads();
        int live = 0;
        for (unsigned int e = threadIdx.x; e < m; e += blockDim.x) {
            const int a = ea[e], b = eb[e];
            const bool ok = vsp[a] == ra && vsp[b] == rb && (vfl[a] & EDM_FLAG_ALIVE) &&
                            (vfl[b] & EDM_FLAG_ALIVE) && vbd[a] < 0 && vbd[b] < 0;
            if (!ok) continue;
            const unsigned long long k = ekey[e];
            if (atomicAdd(&best[a], 0ull) == k && atomicAdd(&best[b], 0ull) == k) {
                vsp[a] = (int8_t)prod;
                vfl[b] = vfl[b] & ~EDM_FLAG_ALIVE;
                vbd[a] = id[b];
                vbd[b] = id[a];
                atomicAdd(&ctr[0], 1ull);
                count_update(G, n_grids, ra, P.xn[a], P.yn[a], -1, false);
                count_update(G, n_grids, prod, P.xn[a], P.yn[a], +1, false);
                count_update(G, n_grids, rb, P.xn[b], P.yn[b], -1, false);
            } else {
                live = 1;
            }
        }
        __threadfence();
        if (!__syncthreads_or(live)) break;
    }
    __syncthreads();
    if (threadIdx.x == 0) *ne = 0u;  // ready for the next pass
}

template <typename Real>
__global__ __launch_bounds__(1024) void k_bind_match(Post<Real> P, const int32_t* __restrict__ id,
                                                     int ra, int rb, int prod,
                                                     const unsigned long long* ekey,
                                                     const int32_t* ea, const int32_t* eb,
                                                     unsigned int* ne, unsigned int cap,
                                                     unsigned long long* best,
                                                     unsigned long long* ctr, Grids<Real> G,
                                                     int n_grids) {
    bind_match_block(P, id, ra, rb, prod, ekey, ea, eb, ne, cap, best, ctr, G, n_grids);
}

// dynamics.py:610-654 apply_dissociation of particle t: a complex splits
// into its two species at the complex's position.  Bonds hold particle
// indices; the partner's slot is P.slot_of[partner - kbase], and the draw
// key is the particle index relative to kbase.
template <typename Real>
__device__ __forceinline__ bool dissoc_one(int t, const int32_t* __restrict__ id,
                                           const Post<Real>& P, int prod, int ra, int rb,
                                           Real prob, const Draws& dr, const Rng& rng,
                                           uint32_t step, uint32_t rxi, int kbase,
                                           const Grids<Real>& G, int n_grids) {
    if (!(P.fl[t] & EDM_FLAG_ALIVE) || P.sp[t] != prod || P.bond[t] < 0) return false;
    const int it = id[t];
    const double u = dr.base ? dr.base[dr.off + dr.n_pairs + it]
                             : (double)Uni<Real>::first(rng.draw(
                                   step, (uint32_t)(it - kbase), 0u, (kTagDissoc << 16) | rxi));
    if (!(u < (double)prob)) return false;
    const int s = P.slot_of[P.bond[t] - kbase];
    P.sp[t] = (int8_t)ra;
    P.sp[s] = (int8_t)rb;
    P.fl[s] = P.fl[s] | EDM_FLAG_ALIVE;
    P.xn[s] = P.xn[t];
    P.yn[s] = P.yn[t];
    P.bond[t] = -1;
    P.bond[s] = -1;
    count_update(G, n_grids, prod, P.xn[t], P.yn[t], -1, false);
    count_update(G, n_grids, ra, P.xn[t], P.yn[t], +1, false);
    count_update(G, n_grids, rb, P.xn[t], P.yn[t], +1, false);
    return true;
}

template <typename Real>
__global__ void k_dissoc(int n, const int32_t* __restrict__ id, Post<Real> P, int prod, int ra,
                         int rb, Real prob, Draws dr, Rng rng, StepRef sref, uint32_t rxi,
                         unsigned long long* ctr, Grids<Real> G, int n_grids) {
    const uint32_t step = sref.get();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool hit = t < n && dissoc_one(t, id, P, prod, ra, rb, prob, dr, rng, step, rxi, 0, G,
                                         n_grids);
    wave_count(&ctr[1], hit);
}

// dynamics.py:1094-1112 update_adaptation_field, reading the sensor density
template <typename Real>
__global__ void k_adapt(Grids<Real> G, int grid, int species, double* w, double rate, int ncells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const double rho = (double)density_of<Real>(G.counts[G.off[grid * G.ns + species] + c],
                                                G.inv[grid]);
    w[c] += rate * (rho - w[c]);
}

__global__ void k_zero_u32(int n, unsigned int* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

#include "edm_cat_tile.h"

// CDF of density-biased recruitment (dynamics.py:2355-2370): over the grid
// cells in index order, the running sum of max(density, 0.01) divided by the
// total.  The sum is sequential, in the reference's order, so the CDF is the
// reference's to the last bit: the floored densities are written in
// parallel, one wavefront accumulates them (the running total is uniform
// across its lanes; lane j keeps the value after cell j of each chunk of
// 64), then every thread divides.  (The reference's uniform fallback for a
// zero total cannot happen: every cell adds at least 0.01.)  One
// workgroup of 1024 threads per CDF.
template <typename Real>
__device__ __forceinline__ void recruit_cdf_block(const int32_t* __restrict__ counts, int nc,
                                                  Real inv, double* cdf) {
    __shared__ double s_total;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        const double d = (double)density_of<Real>(counts[k], inv);
        cdf[k] = d < 0.01 ? 0.01 : d;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double total = 0.0;
        for (int base = 0; base < nc; base += 64) {
            const int k = base + lane;
            const double v = k < nc ? cdf[k] : 0.0;
            const int m = min(64, nc - base);
            double mine = 0.0;
            for (int j = 0; j < m; ++j) {
                total += __shfl(v, j);
                if (lane == j) mine = total;
            }
            if (k < nc) cdf[k] = mine;
        }
        if (lane == 0) s_total = total;
    }
    __syncthreads();
    const double total = s_total;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) cdf[k] = cdf[k] / total;
}

template <typename Real>
__global__ __launch_bounds__(1024) void k_recruit_cdf(const int32_t* __restrict__ counts, int nc,
                                                      Real inv, double* cdf) {
    recruit_cdf_block<Real>(counts, nc, inv, cdf);
}

// Accept the max_recruits candidates with the lowest particle index (the
// reference stops its index-order scan after max_recruits), one workgroup.
// Index-order recruitment cap over one membrane's candidates (particle
// indices id_lo..id_hi): a binary search finds the smallest index threshold
// that admits max_recruits candidates, then the admitted ones come alive at
// their drawn sites.  Every thread of the block must call it.
template <typename Real>
__device__ __forceinline__ void recruit_apply_block(const Post<Real>& P, int target, int maxr,
                                                    int id_lo, int id_hi, const int32_t* cid,
                                                    const int32_t* cslot, const double* cxo,
                                                    const double* cyo, unsigned int* nc,
                                                    unsigned long long* ctr, const Grids<Real>& G,
                                                    int n_grids) {
    __shared__ int s_cnt;
    const int m = (int)*nc;
    int thr = INT_MAX;
    if (m > maxr) {
        int lo = id_lo, hi = id_hi;
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (threadIdx.x == 0) s_cnt = 0;
            __syncthreads();
            int local = 0;
            for (int k = threadIdx.x; k < m; k += blockDim.x) local += cid[k] <= mid;
            if (local) atomicAdd(&s_cnt, local);
            __syncthreads();
            const int cnt = s_cnt;
            __syncthreads();
            if (cnt >= maxr) hi = mid;
            else lo = mid + 1;
        }
        thr = lo;
    }
    int acc = 0;
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        if (cid[k] > thr || maxr <= 0) continue;
        const int t = cslot[k];
        P.fl[t] = P.fl[t] | EDM_FLAG_ALIVE;
        P.xn[t] = (Real)cxo[k];
        P.yn[t] = (Real)cyo[k];
        count_update(G, n_grids, target, P.xn[t], P.yn[t], +1, false);
        ++acc;
    }
    if (acc) atomicAdd(&ctr[3], (unsigned long long)acc);
    __syncthreads();
    if (threadIdx.x == 0) *nc = 0u;  // ready for the next pass
}

// The recruitments of one run of passes: block r applies recruitment r
// from its candidate buffer (entries r * stride .., counter nc[r]).
struct RecSet {
    int target[2], maxr[2];
};

template <typename Real>
__global__ __launch_bounds__(1024) void k_recruit_apply(Post<Real> P, RecSet R, int n,
                                                        const int32_t* cid, const int32_t* cslot,
                                                        const double* cxo, const double* cyo,
                                                        unsigned int* nc, int64_t stride,
                                                        unsigned long long* ctr, Grids<Real> G,
                                                        int n_grids) {
    const int r = (int)blockIdx.x;
    const int64_t o = r * stride;
    recruit_apply_block(P, R.target[r], R.maxr[r], 0, n - 1, cid + o, cslot + o, cxo + o,
                        cyo + o, nc + r, ctr, G, n_grids);
}

#include "edm_passes.h"

__global__ void k_fold(int n, int32_t* counts, int32_t* delta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    counts[i] += delta[i];
    delta[i] = 0;
}

template <typename Real>
__global__ void k_density_out(Grids<Real> G, int grid, int species, double* out, int ncells) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    out[c] = (double)density_of<Real>(G.counts[G.off[grid * G.ns + species] + c], G.inv[grid]);
}

// ----------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------
template <typename Real>
static Cells<Real> make_cells(edm_membrane* m) {
    Cells<Real> C;
    C.x = (const Real*)m->a.x;
    C.y = (const Real*)m->a.y;
    C.sp = m->a.sp;
    C.fl = m->a.fl;
    C.id = m->a.id;
    C.start = m->cell_start;
    C.ncx = m->p.ncx;
    C.ncy = m->p.ncy;
    C.csx = (Real)m->p.csx;
    C.csy = (Real)m->p.csy;
    C.Lx = (Real)m->p.Lx;
    C.Ly = (Real)m->p.Ly;
    C.cut2 = (Real)m->p.cutoff * (Real)m->p.cutoff;
    C.px = m->p.periodic_x != 0;
    C.py = m->p.periodic_y != 0;
    C.pcnt = m->pair_cnt;
    C.poff = m->pair_off;
    C.npairs = m->d_scal;
    C.cap_active = m->need_pair_offsets && m->pair_sync && m->n_pairs < m->n_pairs_found;
    C.cap_dev = m->need_pair_offsets && !m->pair_sync;
    C.cap_gate = 0;
    C.kbase = 0;
    C.poff_base = 0;
    C.tile_rows = m->tile_rows;
    C.sub_rows = m->sub_rows;
    return C;
}

template <typename Real>
static Grids<Real> make_grids(edm_membrane* m) {
    Grids<Real> G;
    G.off = m->d_count_off;
    G.dims = m->d_grid_dims;
    G.s = (const Real*)m->d_grid_s;
    G.inv = (const Real*)m->d_grid_s + 2 * m->grids.size();
    G.counts = m->counts;
    G.delta = m->delta;
    G.ns = m->p.n_species;
    return G;
}

template <typename Real>
static Post<Real> make_post(edm_membrane* m) {
    Post<Real> P;
    P.xn = (Real*)m->a.xn;
    P.yn = (Real*)m->a.yn;
    P.sp = m->a.sp;
    P.fl = m->a.fl;
    P.bond = m->a.bond;
    P.slot_of = m->a.slot_of;
    return P;
}

template <typename T>
static int dalloc(T** p, size_t count) {
    EDM_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T) + 16));
    return 0;
}

// Binding-edge buffers of at least `need` edges (hook draws: one per listed
// pair at most).  Called between steps; the stream is drained first.
static int ensure_edge_cap(edm_membrane* m, size_t need) {
    if (need <= m->edge_cap) return 0;
    graph_touch(m);
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipFree(m->edge_key));
    EDM_HIP(hipFree(m->edge_a));
    EDM_HIP(hipFree(m->edge_b));
    m->edge_key = nullptr;
    m->edge_a = m->edge_b = nullptr;
    m->edge_cap = need + need / 4;
    if (dalloc(&m->edge_key, m->edge_cap) || dalloc(&m->edge_a, m->edge_cap) ||
        dalloc(&m->edge_b, m->edge_cap))
        return EDM_EHIP;
    return 0;
}

template <typename Real>
static int upload_real(void** dst, const double* src, size_t n) {
    std::vector<Real> tmp(std::max<size_t>(n, 1));
    for (size_t i = 0; i < n; ++i) tmp[i] = (Real)src[i];
    EDM_HIP(hipMalloc(dst, tmp.size() * sizeof(Real)));
    EDM_HIP(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(Real), hipMemcpyHostToDevice));
    return 0;
}

// Fused per-particle passes.  Counts exist only for kept (sensor) species;
// a pass "writes" the kept species whose counts it changes (a recruitment
// when its cap is applied) and "reads" its feedback sensors, and it
// "touches" the particles of its species (a conversion: of both species).
// Two passes commute when neither reads what the other writes and they
// touch no common species; a pass may then run before an earlier one.
// Inside a run every thread applies the run's passes to its particle in
// order, so a run is closed before a pass that reads what an earlier pass
// of the run writes, after a pass that reads what it writes itself
// (deferred, folded after the run), and before a pass that touches, or
// reads the counts of, a species with a recruitment earlier in the run (a
// recruitment's index-order cap needs all its candidates, so its particles
// come alive only after the run).  Runs are formed greedily over each
// stretch of consecutive per-particle passes: a pass joins the open run
// unless one of those rules or a skipped earlier pass that it does not
// commute with stops it.  A run holds at most max_rec recruitments, each
// with its own candidate buffer, and a density-biased recruitment ends its
// run (one CDF per run).  Passes with zero probability change nothing and
// are skipped.  Every run starts at its lowest reaction index, and runs
// are launched in that order.
void plan_passes(const std::vector<edm_reaction>& rx, uint32_t sensor_mask,
                 std::vector<std::vector<int>>& groups, std::vector<int>& group_start,
                 std::vector<int>& deferred, int max_rec) {
    const size_t nr = rx.size();
    group_start.assign(nr, -1);
    deferred.assign(nr, 0);
    groups.clear();
    auto kept = [&](int s) { return s >= 0 && ((sensor_mask >> s) & 1u); };
    auto sb = [](int s) { return s >= 0 ? 1u << s : 0u; };
    auto reads = [&](const edm_reaction& r) { return sb(r.fb1.sensor) | sb(r.fb2.sensor); };
    auto writes = [&](const edm_reaction& r) {
        uint32_t w = kept(r.a) ? sb(r.a) : 0u;
        if (r.kind == EDM_RX_CONVERT && kept(r.b)) w |= sb(r.b);
        return w;
    };
    auto touches = [&](const edm_reaction& r) {
        return sb(r.a) | (r.kind == EDM_RX_CONVERT ? sb(r.b) : 0u);
    };
    auto commute = [&](const edm_reaction& p, const edm_reaction& q) {
        return !(reads(p) & writes(q)) && !(reads(q) & writes(p)) && !(touches(p) & touches(q));
    };
    for (size_t i = 0; i < nr;) {
        const bool pp = rx[i].kind == EDM_RX_DETACH || rx[i].kind == EDM_RX_CONVERT ||
                        rx[i].kind == EDM_RX_RECRUIT;
        if (!pp) {
            ++i;
            continue;
        }
        std::vector<int> rem;  // this stretch's passes, in order
        for (; i < nr; ++i) {
            const edm_reaction& r = rx[i];
            if (r.kind != EDM_RX_DETACH && r.kind != EDM_RX_CONVERT && r.kind != EDM_RX_RECRUIT)
                break;
            if (r.prob > 0.0) rem.push_back((int)i);
        }
        while (!rem.empty()) {
            std::vector<int> cur, skipped;
            uint32_t wmask = 0, pend = 0;
            int nrec = 0;
            bool closed = false;
            for (int k : rem) {
                const edm_reaction& r = rx[k];
                bool wait = closed;
                for (int q : skipped)
                    if (!wait && !commute(rx[q], r)) wait = true;
                const bool rec = r.kind == EDM_RX_RECRUIT;
                const uint32_t rmask = reads(r);
                const uint32_t own = r.kind == EDM_RX_RECRUIT ? 0u : writes(r);
                if (!wait && !cur.empty() &&
                    ((rmask & (wmask | pend)) || (touches(r) & pend) ||
                     (rec && nrec >= max_rec) || cur.size() >= (size_t)kMaxPasses))
                    wait = true;
                if (wait) {
                    skipped.push_back(k);
                    continue;
                }
                deferred[k] = (rmask & own) ? 1 : 0;
                cur.push_back(k);
                wmask |= own;
                if (rec) {
                    pend |= sb(r.a);
                    ++nrec;
                }
                if (deferred[k] || (rec && (r.use_adapt || nrec >= max_rec))) closed = true;
            }
            group_start[cur.front()] = (int)groups.size();
            groups.push_back(cur);
            rem.swap(skipped);
        }
    }
}

// Species the passes act on, closed under every reaction that turns one
// species into another: a particle outside the set is never touched by a
// pass and never enters the set, so the number of particles in the set is
// fixed (it sizes the batched ensemble's pass lists).
uint32_t pass_species_of(const std::vector<edm_reaction>& rx,
                         const std::vector<std::vector<int>>& groups) {
    uint32_t S = 0;
    for (auto& g : groups)
        for (int i : g) {
            const edm_reaction& r = rx[i];
            S |= 1u << r.a;
            if (r.kind == EDM_RX_CONVERT) S |= 1u << r.b;
        }
    for (bool grow = S != 0; grow;) {
        grow = false;
        for (auto& r : rx) {
            uint32_t T = 0;
            if (r.kind == EDM_RX_CAT) T = (1u << r.b) | (1u << r.c);
            else if (r.kind == EDM_RX_BIND) T = (1u << r.a) | (1u << r.b) | (1u << r.c);
            else if (r.kind == EDM_RX_CONVERT) T = (1u << r.a) | (1u << r.b);
            if ((T & S) && (T & ~S)) {
                S |= T;
                grow = true;
            }
        }
    }
    return S;
}

template <typename Real>
static int create_t(edm_membrane* m) {
    const edm_params& p = m->p;
    const int n = m->n, ns = p.n_species;
    const size_t rb = sizeof(Real);
    void** rbufs[6] = {&m->a.x, &m->a.y, &m->a.xn, &m->a.yn, &m->a.xt, &m->a.yt};
    for (auto b : rbufs) EDM_HIP(hipMalloc(b, (size_t)n * rb + 16));
    if (dalloc(&m->a.sp, n + 4) || dalloc(&m->a.sp_t, n + 4) || dalloc(&m->a.fl, n + 4) ||
        dalloc(&m->a.fl_t, n + 4) || dalloc(&m->a.id, n) || dalloc(&m->a.id_t, n) ||
        dalloc(&m->a.bond, n) || dalloc(&m->a.bond_t, n) || dalloc(&m->a.ctmc, n + 4) ||
        dalloc(&m->a.ctmc_t, n + 4) || dalloc(&m->a.dye, n + 4) || dalloc(&m->a.dye_t, n + 4) ||
        dalloc(&m->a.slot_of, n))
        return EDM_EHIP;
    m->ncell = p.ncx * p.ncy;
    if (dalloc(&m->key, n) || dalloc(&m->rank, n) || dalloc(&m->perm, n) ||
        dalloc(&m->cell_cnt, m->ncell + 1) || dalloc(&m->cell_start, m->ncell + 1))
        return EDM_EHIP;
    EDM_HIP(hipMemset(m->cell_cnt, 0, (m->ncell + 1) * sizeof(uint32_t)));
    // scan temp storage: the larger of the cell scan and the pair-count scan
    size_t b1 = 0, b2 = 0, b3 = 0;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b1, m->cell_cnt, m->cell_start,
                                             m->ncell + 1, m->stream));
    uint64_t* dummy64 = nullptr;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, dummy64, dummy64, n, m->stream));
    int64_t* dummyi = nullptr;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, dummyi, dummyi, n, m->stream));
    m->cub_tmp_bytes = std::max(b1, std::max(b2, b3)) + 256;
    EDM_HIP(hipMalloc(&m->cub_tmp, m->cub_tmp_bytes));
    m->scan_tiles = (m->ncell + 1 + kScanTile - 1) / kScanTile;
    EDM_HIP(hipMalloc(&m->scan_state, (size_t)m->scan_tiles * sizeof(uint64_t)));
    EDM_HIP(hipMalloc(&m->scan_ticket, 2 * sizeof(uint32_t)));
    EDM_HIP(hipMemset(m->scan_ticket, 0, 2 * sizeof(uint32_t)));
    if (dalloc(&m->pair_cnt, n) || dalloc(&m->pair_off, n) || dalloc(&m->d_scal, 16)) return EDM_EHIP;
    EDM_HIP(hipMemset(m->d_scal, 0, 16 * sizeof(int64_t)));
    // parameters
    int rc = 0;
    rc |= upload_real<Real>(&m->sig, p.sigma, (size_t)ns * ns);
    rc |= upload_real<Real>(&m->eps, p.epsilon, (size_t)ns * ns);
    rc |= upload_real<Real>(&m->kh, p.k_harm, (size_t)ns * ns);
    rc |= upload_real<Real>(&m->Dbase, p.D_base, ns);
    std::vector<double> z(1, 0.0);
    rc |= upload_real<Real>(&m->rcx, p.n_regions ? p.rcx : z.data(), p.n_regions);
    rc |= upload_real<Real>(&m->rcy, p.n_regions ? p.rcy : z.data(), p.n_regions);
    rc |= upload_real<Real>(&m->rr, p.n_regions ? p.rr : z.data(), p.n_regions);
    rc |= upload_real<Real>(&m->rf, p.n_regions ? p.rf : z.data(), p.n_regions);
    if (rc) return EDM_EHIP;
    // density grids: keep counts for every (grid, sensor) read by a feedback
    const int ng = (int)m->grids.size();
    m->count_off.assign((size_t)std::max(ng, 1) * ns, -1);
    auto keep = [&](int g, int s) {
        if (g < 0 || s < 0) return;
        m->count_off[(size_t)g * ns + s] = 0;
    };
    for (auto& r : m->rx) {
        keep(r.fb1.grid, r.fb1.sensor);
        keep(r.fb2.grid, r.fb2.sensor);
    }
    if (p.adapt_grid >= 0) keep(p.adapt_grid, p.adapt_species);
    int total = 0;
    for (int g = 0; g < ng; ++g)
        for (int s = 0; s < ns; ++s)
            if (m->count_off[(size_t)g * ns + s] == 0) {
                m->count_off[(size_t)g * ns + s] = total;
                total += m->grids[g].nx * m->grids[g].ny;
                m->sensor_mask_all |= 1u << s;
            }
    m->n_count = total;
    if (dalloc(&m->counts, total) || dalloc(&m->delta, total) ||
        dalloc(&m->d_count_off, m->count_off.size()) || dalloc(&m->d_grid_dims, 2 * std::max(ng, 1)))
        return EDM_EHIP;
    EDM_HIP(hipMemset(m->counts, 0, (size_t)std::max(total, 1) * 4));
    EDM_HIP(hipMemset(m->delta, 0, (size_t)std::max(total, 1) * 4));
    EDM_HIP(hipMemcpy(m->d_count_off, m->count_off.data(), m->count_off.size() * 4,
                      hipMemcpyHostToDevice));
    std::vector<int32_t> dims(2 * std::max(ng, 1), 1);
    std::vector<double> gs(3 * std::max(ng, 1), 1.0);
    for (int g = 0; g < ng; ++g) {
        dims[2 * g] = m->grids[g].nx;
        dims[2 * g + 1] = m->grids[g].ny;
        gs[2 * g] = m->grids[g].sx;
        gs[2 * g + 1] = m->grids[g].sy;
        const double area = m->grids[g].sx * m->grids[g].sy;
        gs[2 * ng + g] = area > 0.0 ? 1.0 / area : 0.0;
    }
    EDM_HIP(hipMemcpy(m->d_grid_dims, dims.data(), dims.size() * 4, hipMemcpyHostToDevice));
    if (upload_real<Real>(&m->d_grid_s, gs.data(), gs.size())) return EDM_EHIP;
    if (p.adapt_grid >= 0) {
        const auto& g = m->grids[p.adapt_grid];
        if (dalloc(&m->w_adapt, (size_t)g.nx * g.ny)) return EDM_EHIP;
        EDM_HIP(hipMemset(m->w_adapt, 0, (size_t)g.nx * g.ny * sizeof(double)));
    }
    plan_passes(m->rx, m->sensor_mask_all, m->pass_groups, m->group_start, m->pass_deferred,
                kMaxRecruits);
    // largest squared range of any pair potential (the tiled force kernel
    // skips candidates beyond it before looking up the species tables)
    {
        double r2 = 0.0;
        for (int k = 0; k < ns * ns; ++k) {
            const double sg = p.sigma[k];
            if (!(sg > 0.0)) continue;
            if (p.use_wca && p.epsilon[k] > 0.0) r2 = std::max(r2, sg * sg * 1.2599210498948732);
            if (p.use_harmonic && p.k_harm[k] > 0.0) r2 = std::max(r2, sg * sg);
        }
        m->rint2 = r2 * (1.0 + 1e-6);
    }
    // reaction scratch
    // accepted binding edges: with hook draws at most one per listed pair
    // (grown to the pair count before the step); with Philox draws at most
    // one per A-B pair within the cutoff, so 4 per particle covers a
    // membrane up to ~8x the 1M workload's density, and a step that accepts
    // more is reported (ne[1], EDM_ECAPACITY) rather than truncated silently
    // (autocatalysis lists its accepted enzyme-substrate and substrate-
    // substrate pairs in the same buffer)
    bool has_bind = m->has_autocat;
    for (const auto& r : m->rx) has_bind = has_bind || r.kind == EDM_RX_BIND;
    m->edge_cap = has_bind ? (size_t)std::max<int64_t>(4 * (int64_t)n, 4096) : 1024;
    if (m->has_autocat) {
        if (dalloc(&m->tconv, n)) return EDM_EHIP;
        EDM_HIP(hipMemset(m->tconv, 0xff, (size_t)n * sizeof(unsigned long long)));
    }
    if (dalloc(&m->edge_key, m->edge_cap) || dalloc(&m->edge_a, m->edge_cap) ||
        dalloc(&m->edge_b, m->edge_cap) || dalloc(&m->n_edges, 4) || dalloc(&m->best, n) ||
        dalloc(&m->cand_id, kMaxRecruits * n) || dalloc(&m->cand_slot, kMaxRecruits * n) ||
        dalloc(&m->cand_x, kMaxRecruits * n) || dalloc(&m->cand_y, kMaxRecruits * n) ||
        dalloc(&m->n_cand, 4) || dalloc(&m->d_ctr, 8))
        return EDM_EHIP;
    size_t ncdf = 0;
    for (const auto& r : m->rx)
        if (r.kind == EDM_RX_RECRUIT && r.use_adapt)
            ncdf = std::max(ncdf, (size_t)m->grids[r.fb1.grid].nx * m->grids[r.fb1.grid].ny);
    if (ncdf && dalloc(&m->rcdf, ncdf)) return EDM_EHIP;
    EDM_HIP(hipMemset(m->d_ctr, 0, 8 * sizeof(unsigned long long)));
    EDM_HIP(hipMemset(m->n_edges, 0, 16));
    EDM_HIP(hipMemset(m->n_cand, 0, 16));
    return 0;
}

int dyn_create(edm_membrane* m) { return m->fp64 ? create_t<double>(m) : create_t<float>(m); }

// Hook draws of one recruitment: u_acc[n] and u_pos[n,2]; the
// density-biased placement adds u_loc[n] and xi_off[n,2]
// (dynamics.py:2301-2302, 2366-2367).
static int64_t recruit_draws(const edm_reaction& r, int64_t n) {
    return (r.use_adapt ? 6 : 3) * n;
}

int64_t hook_draws_per_step(const edm_membrane* m, int64_t n_pairs) {
    int64_t total = 2 * (int64_t)m->n;
    for (const auto& r : m->rx) {
        switch (r.kind) {
            case EDM_RX_BIND: total += n_pairs + m->n; break;
            case EDM_RX_CAT: total += n_pairs; break;
            case EDM_RX_RECRUIT: total += recruit_draws(r, m->n); break;
            case EDM_RX_DETACH:
            case EDM_RX_CONVERT: total += m->n; break;
            default: break;
        }
    }
    return total;
}

// Reactions ri .. ri + len - 1 share one catalysis launch: up to kCatFuse
// back-to-back catalytic reactions with one substrate and product and
// without feedback or adaptation.
static int cat_run_len(const std::vector<edm_reaction>& rx, size_t ri) {
    const edm_reaction& r = rx[ri];
    // autocatalysis depends on the order of the reactions' pairs: never fused
    const bool plain = r.fb1.sensor < 0 && !r.use_adapt && r.a != r.c;
    int nf = 1;
    while (plain && nf < kCatFuse && ri + nf < rx.size()) {
        const edm_reaction& r2 = rx[ri + nf];
        if (r2.kind != EDM_RX_CAT || r2.b != r.b || r2.c != r.c || r2.fb1.sensor >= 0 ||
            r2.use_adapt || r2.a == r2.c)
            break;
        ++nf;
    }
    return nf;
}

static uint32_t sbit(int s) { return s >= 0 ? 1u << s : 0u; }

// Species whose particles reaction q can move into or out of a work list:
// species changes, and with `alive` also changes of the alive flag.
static uint32_t touched_species(const edm_reaction& q, bool alive) {
    switch (q.kind) {
        case EDM_RX_BIND: return sbit(q.a) | sbit(q.b) | sbit(q.c);
        case EDM_RX_CAT: return sbit(q.b) | sbit(q.c);
        case EDM_RX_CONVERT: return sbit(q.a) | sbit(q.b);
        case EDM_RX_RECRUIT:
        case EDM_RX_DETACH: return alive ? sbit(q.a) : 0u;
        default: return 0u;
    }
}

// The species a run of passes can touch: every pass's species a, and the
// species a particle may hold before the run and still reach a pass as its
// species a through the run's earlier conversions (reach).
static uint32_t pass_reach(const std::vector<edm_reaction>& rx, const std::vector<int>& g, int k) {
    uint32_t reach = sbit(rx[g[k]].a);
    for (int j = k - 1; j >= 0; --j) {
        const edm_reaction& qj = rx[g[j]];
        if (qj.kind == EDM_RX_CONVERT && ((reach >> qj.b) & 1u)) reach |= sbit(qj.a);
    }
    return reach;
}

// This step's runs of passes, each with a work list of its species (alive or
// not) -- built with the others after the sort, or just before the run when
// an earlier reaction of the step can change the species of its particles.
static void plan_pass_lists(edm_membrane* m, int& next) {
    m->pass_plan.clear();
    for (size_t gi = 0; gi < m->pass_groups.size(); ++gi) {
        const std::vector<int>& g = m->pass_groups[gi];
        PassPlan pp{};
        pp.ri = g.front();
        pp.list = -1;
        uint32_t S = 0;
        for (size_t k = 0; k < g.size(); ++k) S |= pass_reach(m->rx, g, (int)k);
        if (next < kWlMaxLists) {
            pp.list = next++;
            pp.spec = WlSpec{S, 0u};
            // reactions before the run: those before its first one, and
            // the members of earlier runs (a run may hold later reactions
            // that commute with the ones it skipped)
            for (int k = 0; k < pp.ri; ++k)
                if (touched_species(m->rx[k], false) & S) pp.late = true;
            for (size_t gj = 0; gj < gi; ++gj)
                for (int k : m->pass_groups[gj])
                    if (touched_species(m->rx[k], false) & S) pp.late = true;
        }
        m->pass_plan.push_back(pp);
    }
}

// This step's catalysis walks.  A run walks a work list of its entities
// (alive enzymes, or alive substrates) unless an enzyme species is also its
// substrate or product (an entity could then change identity during the
// run; the strip kernel keeps the species snapshot of the run's start).  The
// list is built with the others after the sort, or -- when an earlier
// reaction of the step can change the species or alive flag of its
// entities -- just before the run.
static void plan_cat_walks(edm_membrane* m, int& next) {
    m->cat_plan.clear();
    for (size_t ri = 0; ri < m->rx.size();) {
        const edm_reaction& r = m->rx[ri];
        if (r.kind != EDM_RX_CAT) {
            ++ri;
            continue;
        }
        CatPlan c{};
        c.ri = (int)ri;
        c.nf = cat_run_len(m->rx, ri);
        int enz[kCatFuse];
        uint32_t E = 0;
        for (int j = 0; j < c.nf; ++j) {
            enz[j] = m->rx[ri + j].a;
            E |= sbit(enz[j]);
        }
        c.by_sub = m->cat_side >= 0 ? m->cat_side == 1
                                    : cat_by_substrate(m->alive_at_upload, enz, c.nf, r.b);
        c.list = -1;
        if (m->cat_walk == 1 && !(E & (sbit(r.b) | sbit(r.c))) && next < kWlMaxLists) {
            c.list = next++;
            c.spec = WlSpec{c.by_sub ? sbit(r.b) : E, (uint32_t)EDM_FLAG_ALIVE};
            // An earlier catalysis run (not autocatalytic) that only converts
            // substrates into the list's species can append them to the list
            // built after the sort; any other earlier change to the entities
            // means a build just before the run.
            std::vector<size_t> app;
            for (size_t k = 0; k < ri; ++k) {
                const edm_reaction& q = m->rx[k];
                if (!(touched_species(q, true) & c.spec.species)) continue;
                CatPlan* src = nullptr;
                for (CatPlan& e : m->cat_plan)
                    if ((int)k >= e.ri && (int)k < e.ri + e.nf) src = &e;
                const bool adds = q.kind == EDM_RX_CAT && q.a != q.c &&
                                  (sbit(q.c) & c.spec.species) && !(sbit(q.b) & c.spec.species);
                if (adds && src && (src->app_list < 0 || src->app_list == c.list))
                    app.push_back((size_t)(src - m->cat_plan.data()));
                else
                    c.late = true;
            }
            if (!c.late)
                for (size_t e : app) m->cat_plan[e].app_list = c.list;
        }
        m->cat_plan.push_back(c);
        ri += c.nf;
    }
}


static WlBuild wl_make(edm_membrane* m, const int* lists, const WlSpec* specs, int nl) {
    WlBuild W{};
    W.n_lists = nl;
    for (int i = 0; i < nl; ++i) {
        W.spec[i] = specs[i];
        W.list[i] = m->wl + (size_t)lists[i] * m->wl_nblk * kWlSlots;
        W.cnt[i] = m->wl_cnt + (size_t)lists[i] * m->wl_nblk;
        m->wl_built |= 1u << lists[i];
    }
    return W;
}

// Launches the build staged by the step's sort (the force launch that
// follows takes it over when it can).
int dyn_wl_flush(edm_membrane* m) {
    if (!m->wl_pend_n) return 0;
    const WlBuild W = wl_make(m, m->wl_pend_list, m->wl_pend_spec, m->wl_pend_n);
    m->wl_pend_n = 0;
    hipLaunchKernelGGL(k_wl_build, dim3(m->wl_nblk), dim3(256), 0, m->stream, m->n,
                       m->wl_pend_sp, m->wl_pend_fl, W);
    EDM_HIP(hipGetLastError());
    return 0;
}

// Builds the listed work lists from the current arrays (sp, fl: the slot
// order of this step).
static void wl_build(edm_membrane* m, const int8_t* sp, const uint8_t* fl, const int* lists,
                     const WlSpec* specs, int nl) {
    if (nl == 0 || m->n == 0 || !m->wl) return;
    dyn_wl_flush(m);
    const WlBuild W = wl_make(m, lists, specs, nl);
    hipLaunchKernelGGL(k_wl_build, dim3(m->wl_nblk), dim3(256), 0, m->stream, m->n, sp, fl, W);
}

template <typename Real>
static int begin_t(edm_membrane* m) {
    const int n = m->n;
    hipStream_t st = m->stream;
    // bin on the current positions, counting sort, order cells by index
    // force sub-row lists: dense membranes, steps that do not need the
    // reference's in-cell order (parity draws, the pair cap)
    m->fine_active = m->fine_force == 1 && m->sub_rows && !m->need_pair_offsets &&
                     m->p.ncx >= 3 && m->p.ncy >= 3;
    if (m->fine_active) {
        const size_t need = (size_t)m->ncell * kMaxSub + 1;
        if (need > m->fstart_cap || !m->fkey) {
            graph_touch(m);
            EDM_HIP(hipStreamSynchronize(st));
            if (m->fstart) EDM_HIP(hipFree(m->fstart));
            if (m->fkey) EDM_HIP(hipFree(m->fkey));
            m->fstart = m->fkey = nullptr;
            m->fstart_cap = 0;
            EDM_HIP(hipMalloc((void**)&m->fstart, need * sizeof(uint32_t)));
            EDM_HIP(hipMalloc((void**)&m->fkey, (size_t)std::max(n, 1) * sizeof(uint32_t)));
            m->fstart_cap = need;
        }
        // sub-rows at least 1.01 x the longest potential range tall
        const double rl = std::sqrt(std::min(m->p.cutoff * m->p.cutoff, m->rint2));
        const double q = rl > 0.0 ? m->p.csy / (1.01 * rl) : (double)kMaxSub;
        m->fine_nsub = q >= kMaxSub ? kMaxSub : std::max(1, (int)q);
        m->fine_inv_h = (double)m->fine_nsub / m->p.csy;
    }
    hipLaunchKernelGGL(k_bin<Real>, dim3(nblk(std::max({n, m->n_count, m->scan_tiles}))),
                       dim3(kBlock), 0, st, n, (const Real*)m->a.x, (const Real*)m->a.y,
                       (Real)m->p.csx, (Real)m->p.csy, m->p.ncx, m->p.ncy, m->key, m->rank,
                       m->cell_cnt, m->counts, m->n_count, (const int32_t*)m->a.id,
                       m->fine_active ? m->fkey : nullptr, m->fine_nsub, (Real)m->fine_inv_h,
                       m->scan_state, m->scan_tiles, m->scan_ticket);
    hipLaunchKernelGGL(k_cell_scan, dim3(m->scan_tiles), dim3(256), 0, st, m->ncell + 1,
                       (const uint32_t*)m->cell_cnt, m->cell_start, m->scan_state,
                       m->scan_ticket);
    size_t tb = 0;
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(kBlock), 0, st, n, m->key, m->rank,
                       m->cell_start, m->perm);
    if (m->fine_active) {
        hipLaunchKernelGGL(k_sort_cells_fine, dim3(nblk(m->ncell)), dim3(kBlock), 0, st,
                           m->ncell, m->cell_start, m->perm, m->cell_cnt,
                           (const uint32_t*)m->fkey, m->fine_nsub, m->fstart);
    } else {
        hipLaunchKernelGGL(k_sort_cells, dim3(nblk(m->ncell)), dim3(kBlock), 0, st, m->ncell,
                           m->cell_start, m->perm, m->cell_cnt, (const int32_t*)m->a.id, 0);
    }
    // keep_slot_of: bonds can exist
    hipLaunchKernelGGL((m->keep_slot_of ? k_gather<Real, true> : k_gather<Real, false>),
                       dim3(nblk(n)), dim3(kBlock), 0, st, n, (const uint32_t*)m->perm, m->a,
                       m->keep_slot_of);
    m->slot_of_valid = m->keep_slot_of;
    if (m->wl) {
        // work lists of this step's slot order (the gather wrote it into the
        // _t arrays, swapped below): dye carriers, and the runs of passes and
        // catalysis walks that no earlier reaction of the step can change
        int next = kWlDye + 1;
        plan_pass_lists(m, next);
        plan_cat_walks(m, next);
        int lists[kWlMaxLists];
        WlSpec specs[kWlMaxLists];
        int nl = 0;
        if (m->n_dye > 0) {
            lists[nl] = kWlDye;
            specs[nl++] = WlSpec{0xffffffffu, (uint32_t)EDM_FLAG_DYE};
        }
        for (const PassPlan& pp : m->pass_plan)
            if (pp.list >= 0 && !pp.late) {
                lists[nl] = pp.list;
                specs[nl++] = pp.spec;
            }
        for (const CatPlan& c : m->cat_plan)
            if (c.list >= 0 && !c.late) {
                lists[nl] = c.list;
                specs[nl++] = c.spec;
            }
        m->wl_built = 0;
        m->wl_pend_n = 0;
        if (nl > 0 && n > 0) {
            // staged: built by the force launch of this step, or by
            // dyn_wl_flush before anything else reads a list
            std::copy(lists, lists + nl, m->wl_pend_list);
            std::copy(specs, specs + nl, m->wl_pend_spec);
            m->wl_pend_sp = (const int8_t*)m->a.sp_t;
            m->wl_pend_fl = (const uint8_t*)m->a.fl_t;
            m->wl_pend_n = nl;
            for (int i = 0; i < nl; ++i) m->wl_built |= 1u << lists[i];
        }
        m->wl_valid = true;
    }
    std::swap(m->a.x, m->a.xt);
    std::swap(m->a.y, m->a.yt);
    std::swap(m->a.sp, m->a.sp_t);
    std::swap(m->a.fl, m->a.fl_t);
    std::swap(m->a.id, m->a.id_t);
    std::swap(m->a.bond, m->a.bond_t);
    std::swap(m->a.ctmc, m->a.ctmc_t);
    std::swap(m->a.dye, m->a.dye_t);
    if (m->need_pair_offsets) {
        Cells<Real> C = make_cells<Real>(m);
        hipLaunchKernelGGL(k_pair_count<Real>, dim3(nblk(n)), dim3(kBlock), 0, st, n, C,
                           m->pair_cnt);
        tb = m->cub_tmp_bytes;
        EDM_HIP(hipcub::DeviceScan::ExclusiveSum(m->cub_tmp, tb, m->pair_cnt,
                                                 m->pair_off, n, st));
        hipLaunchKernelGGL(k_pair_total, dim3(1), dim3(1), 0, st, n, m->pair_cnt,
                           (const uint64_t*)m->pair_off, (int64_t)m->p.pair_cap, m->d_scal);
        if (m->pair_sync) {
            // parity steps hand the pair count to the caller (hook draws)
            int64_t sc[2];
            EDM_HIP(hipMemcpyAsync(sc, m->d_scal, 2 * sizeof(int64_t), hipMemcpyDeviceToHost,
                                   st));
            EDM_HIP(hipStreamSynchronize(st));
            m->n_pairs = sc[0];
            m->n_pairs_found = sc[1];
        } else {
            // the kernels read the counts from d_scal (Cells::cap_dev)
            m->n_pairs = m->n_pairs_found = -1;
        }
    }
    EDM_HIP(hipGetLastError());
    return 0;
}

int dyn_step_begin(edm_membrane* m) { return m->fp64 ? begin_t<double>(m) : begin_t<float>(m); }

int dyn_check_scan(edm_membrane* m) {
    if (!m->scan_ticket) return 0;
    unsigned int f = 0;
    EDM_HIP(hipMemcpyAsync(&f, m->scan_ticket + 1, sizeof(f), hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    if (f) {
        set_error("cell scan: a tile waited too long for an earlier tile's sum; the cell list "
                  "of that step is wrong");
        return EDM_EHIP;
    }
    return 0;
}

int dyn_check_overflow(edm_membrane* m) {
    if (!m->n_edges) return 0;
    if (int rc = dyn_check_scan(m)) return rc;
    unsigned int ne[2];
    EDM_HIP(hipMemcpyAsync(ne, m->n_edges, sizeof(ne), hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    if (ne[1] > 0) {
        set_error("binding / autocatalysis: a step accepted " + std::to_string(ne[1]) +
                  " candidate pairs, more than the edge buffer holds (" +
                  std::to_string(m->edge_cap) + "); the outcome of that step is incomplete");
        return EDM_ECAPACITY;
    }
    return 0;
}

template <typename Real>
static int finish_t(edm_membrane* m, const double* draws, bool hook) {
    const int n = m->n;
    hipStream_t st = m->stream;
    const edm_params& p = m->p;
    Cells<Real> C = make_cells<Real>(m);
    Grids<Real> G = make_grids<Real>(m);
    Post<Real> P = make_post<Real>(m);
    const int ng = (int)m->grids.size();
    Rng rng{(uint32_t)(p.seed & 0xffffffffu), (uint32_t)(p.seed >> 32)};
    const StepRef step = step_ref(m);

    ForceParams<Real> F;
    F.sig = (const Real*)m->sig;
    F.eps = (const Real*)m->eps;
    F.kh = (const Real*)m->kh;
    F.Dbase = (const Real*)m->Dbase;
    F.ns = p.n_species;
    F.wca = p.use_wca != 0;
    F.harm = p.use_harmonic != 0;
    F.dfield = p.dfield;
    F.n_regions = p.n_regions;
    F.rcx = (const Real*)m->rcx;
    F.rcy = (const Real*)m->rcy;
    F.rr = (const Real*)m->rr;
    F.rf = (const Real*)m->rf;
    F.smooth_w = (Real)p.smooth_w;
    F.sin_amp = (Real)p.sin_amp;
    F.sin_kx = (Real)p.sin_kx;
    F.sin_ky = (Real)p.sin_ky;
    F.dt = (Real)p.dt;
    F.max_disp = (Real)p.max_disp;
    F.rint2 = (Real)m->rint2;
    if (hook)
        for (const auto& r : m->rx)
            if (r.kind == EDM_RX_BIND || (r.kind == EDM_RX_CAT && r.a == r.c)) {
                if (int rc = ensure_edge_cap(m, (size_t)m->n_pairs)) return rc;
                break;
            }
    const double* noise = hook ? draws : nullptr;
    // the staged work lists go with k_force_fine; any other force kernel
    // has them built first
    const bool fine_launch = !(hook && m->fp64 && p.exact_force_order) && !C.cap_active &&
                             m->fine_active && !hook;
    WlBuild Wf{};
    int wl_nb = 0;
    const int8_t* wl_sp = nullptr;
    const uint8_t* wl_fl = nullptr;
    if (fine_launch && m->wl_pend_n) {
        Wf = wl_make(m, m->wl_pend_list, m->wl_pend_spec, m->wl_pend_n);
        wl_nb = m->wl_nblk;
        wl_sp = m->wl_pend_sp;
        wl_fl = m->wl_pend_fl;
        m->wl_pend_n = 0;
    } else if (int rc = dyn_wl_flush(m)) {
        return rc;
    }
    if (m->time_force) EDM_HIP(hipEventRecord(m->kev[0], st));
    if (hook && m->fp64 && p.exact_force_order)
        hipLaunchKernelGGL((k_force_bd<Real, true>), dim3(nblk(n)), dim3(kBlock), 0, st, n, C, F,
                           (Real*)m->a.xn, (Real*)m->a.yn, noise, rng, step, G, ng,
                           m->sensor_mask_all);
    else if (C.cap_active)
        hipLaunchKernelGGL((k_force_bd<Real, false>), dim3(nblk(n)), dim3(kBlock), 0, st, n, C, F,
                           (Real*)m->a.xn, (Real*)m->a.yn, noise, rng, step, G, ng,
                           m->sensor_mask_all);
    else {
        // with the cap decided on the device, both paths are launched and
        // exactly one of them runs (the other returns at once)
        Cells<Real> Ct = C;
        if (C.cap_dev) {
            Cells<Real> Cb = C;
            Cb.cap_gate = 1;
            Ct.cap_gate = -1;
            hipLaunchKernelGGL((k_force_bd<Real, false>), dim3(nblk(n)), dim3(kBlock), 0, st, n,
                               Cb, F, (Real*)m->a.xn, (Real*)m->a.yn, noise, rng, step, G, ng,
                               m->sensor_mask_all);
        }
        const int ntile_y = (p.ncy + m->tile_rows - 1) / m->tile_rows;
        const int ntiles = p.ncx * ntile_y;
        const dim3 grid(8 * ((ntiles + 7) / 8));
        if (fine_launch)
            hipLaunchKernelGGL((k_force_fine<Real>), dim3(nblk(n) + wl_nb), dim3(256), 0, st, n,
                               Ct, F, (Real*)m->a.xn, (Real*)m->a.yn, rng, step, G, ng,
                               m->sensor_mask_all, m->fine_nsub, (Real)m->fine_inv_h,
                               (const uint32_t*)m->fstart, Wf, wl_nb, wl_sp, wl_fl);
        else if (m->sub_rows)
            hipLaunchKernelGGL((k_force_tile<Real, true>), grid, dim3(256),
                               force_sub_lds(m->tile_rows, 1), st, Ct, F, (Real*)m->a.xn,
                               (Real*)m->a.yn, noise, rng, step, G, ng, m->sensor_mask_all,
                               ntile_y, ntiles);
        else
            hipLaunchKernelGGL((k_force_tile<Real, false>), grid, dim3(256), 0, st, Ct, F,
                               (Real*)m->a.xn, (Real*)m->a.yn, noise, rng, step, G, ng,
                               m->sensor_mask_all, ntile_y, ntiles);
    }
    if (m->time_force) EDM_HIP(hipEventRecord(m->kev[1], st));
    // the fused photon kernel of this step (ph_fused_image) may start once
    // the forces are done; it is released before the first run of
    // per-particle passes (those latency-bound launches leave the most room
    // beside them), or at the end of the reactions
    m->force_mark = false;
    bool mark_pending = false;
    if (!hook && m->ph_overlap_ok && !m->use_graphs && !m->graph_rec) {
        // ordering between two streams of this device only: no system-scope
        // fence (no write-back for the host) when the event is recorded
        if (!m->ev_force)
            EDM_HIP(hipEventCreateWithFlags(&m->ev_force,
                                            hipEventDisableTiming | hipEventDisableSystemFence));
        mark_pending = true;
    }
    auto release_photons = [&]() -> int {
        if (!mark_pending) return 0;
        EDM_HIP(hipEventRecord(m->ev_force, st));
        m->force_mark = true;
        mark_pending = false;
        return 0;
    };

    int64_t off = 2 * (int64_t)n;
    const int64_t npairs = m->n_pairs;
    // offsets of every reaction's arrays in the step's hook draws
    std::vector<int64_t> draw_off(m->rx.size());
    {
        int64_t o = 2 * (int64_t)n;
        for (size_t ri = 0; ri < m->rx.size(); ++ri) {
            draw_off[ri] = o;
            switch (m->rx[ri].kind) {
                case EDM_RX_BIND: o += npairs + n; break;
                case EDM_RX_CAT: o += npairs; break;
                case EDM_RX_RECRUIT: o += recruit_draws(m->rx[ri], n); break;
                case EDM_RX_DETACH:
                case EDM_RX_CONVERT: o += n; break;
                default: break;
            }
        }
    }
    int cat_skip = 0;
    bool fold_pending = false;  // the previous run's deferred updates, not yet folded
    // a catalysis run that reads no counts and defers none can fold the
    // previous run's deferred updates in its own launch
    auto folds_in = [&](size_t j) {
        if (j >= m->rx.size()) return false;
        const edm_reaction& q = m->rx[j];
        return q.kind == EDM_RX_CAT && q.a != q.c && q.fb1.sensor < 0;
    };
    for (size_t ri = 0; ri < m->rx.size(); ++ri) {
        const edm_reaction& r = m->rx[ri];
        Draws dr{hook ? draws : nullptr, off, npairs, n};
        const uint32_t rxi = (uint32_t)ri;
        const FbDev f1 = make_fb(m, r.fb1), f2 = make_fb(m, r.fb2);
        switch (r.kind) {
            case EDM_RX_BIND: {
                const Real c2 = (Real)r.contact * (Real)r.contact;
                hipLaunchKernelGGL(k_bind_edges<Real>, dim3(nblk(n)), dim3(kBlock), 0, st, n, C, P,
                                   r.a, r.b, c2, (Real)r.prob, dr, rng, step, rxi, m->edge_key,
                                   m->edge_a, m->edge_b, m->n_edges, (unsigned int)m->edge_cap);
                hipLaunchKernelGGL(k_bind_match<Real>, dim3(1), dim3(1024), 0, st, P,
                                   (const int32_t*)m->a.id, r.a, r.b, r.c,
                                   (const unsigned long long*)m->edge_key,
                                   (const int32_t*)m->edge_a, (const int32_t*)m->edge_b,
                                   m->n_edges, (unsigned int)m->edge_cap,
                                   m->best, m->d_ctr, G, ng);
                hipLaunchKernelGGL(k_dissoc<Real>, dim3(nblk(n)), dim3(kBlock), 0, st, n,
                                   (const int32_t*)m->a.id, P, r.c, r.a, r.b, (Real)r.prob2, dr,
                                   rng, step, rxi, m->d_ctr, G, ng);
                off += npairs + n;
                break;
            }
            case EDM_RX_ADAPT: {
                const auto& g = m->grids[p.adapt_grid];
                hipLaunchKernelGGL(k_adapt<Real>, dim3(nblk(g.nx * g.ny)), dim3(kBlock), 0, st, G,
                                   p.adapt_grid, p.adapt_species, m->w_adapt, p.dt / p.adapt_tau,
                                   g.nx * g.ny);
                break;
            }
            case EDM_RX_CAT: {
                if (cat_skip > 0) {  // member of the fused run launched before
                    --cat_skip;
                    off += npairs;
                    break;
                }
                const bool defer = f1.sensor >= 0 && (f1.sensor == r.b || f1.sensor == r.c);
                const CatPlan* cp = nullptr;
                for (const CatPlan& c : m->cat_plan)
                    if (c.ri == (int)ri) cp = &c;
                const int nf = cp ? cp->nf : cat_run_len(m->rx, ri);
                int wnx = 1, wny = 1;
                const double* w = nullptr;
                if (r.use_adapt && p.adapt_grid >= 0) {
                    w = m->w_adapt;
                    wnx = m->grids[p.adapt_grid].nx;
                    wny = m->grids[p.adapt_grid].ny;
                }
                CatArgs<Real> A{};
                A.n_enz = nf;
                for (int j = 0; j < nf; ++j) {
                    const edm_reaction& rj = m->rx[ri + j];
                    A.enz[j] = rj.a;
                    A.contact2[j] = (Real)rj.contact * (Real)rj.contact;
                    A.prob[j] = (Real)rj.prob;
                    A.rxi[j] = (uint32_t)(ri + j);
                    A.doff[j] = draw_off[ri + j];
                }
                A.sub = r.b;
                A.prod = r.c;
                A.fb = f1;
                A.w = w;
                A.wnx = wnx;
                A.wny = wny;
                A.Kw = (Real)p.adapt_K;
                A.nw = (Real)p.adapt_n;
                A.dr = dr;
                A.rng = rng;
                A.step = step.off;
                A.step_base = step.base;
                A.deferred = defer;
                A.group_shift = __builtin_ctz((unsigned)m->cat_lanes);
                A.app = nullptr;
                A.app_cnt = nullptr;
                if (cp && cp->app_list >= 0 && m->wl_valid &&
                    ((m->wl_built >> cp->app_list) & 1u)) {
                    A.app = m->wl + (size_t)cp->app_list * m->wl_nblk * kWlSlots;
                    A.app_cnt = m->wl_cnt + (size_t)cp->app_list * m->wl_nblk;
                }
                A.fold_counts = m->counts;
                A.fold_delta = m->delta;
                A.fold_n = 0;
                if (fold_pending) {
                    A.fold_n = m->n_count;
                    fold_pending = false;
                }
                A.by_substrate = cp ? cp->by_sub
                                    : m->cat_side >= 0
                                          ? m->cat_side == 1
                                          : cat_by_substrate(m->alive_at_upload, A.enz, nf, r.b);
                bool walked = false;
                if (r.a == r.c) {  // autocatalysis: conversions cascade in pair order
                    hipLaunchKernelGGL((k_autocat_edges<Real>), dim3(nblk(n)), dim3(kBlock), 0,
                                       st, n, C, P, A, G, m->edge_key, m->edge_a, m->edge_b,
                                       m->n_edges, (unsigned int)m->edge_cap);
                    hipLaunchKernelGGL((k_autocat_resolve<Real>), dim3(1), dim3(1024), 0, st, P,
                                       A, (const unsigned long long*)m->edge_key,
                                       (const int32_t*)m->edge_a, (const int32_t*)m->edge_b,
                                       m->n_edges, (unsigned int)m->edge_cap, m->tconv,
                                       m->d_ctr, G, ng);
                    walked = true;
                } else if (cp && cp->list >= 0 && m->wl_valid) {
                    if (cp->late)  // entities as they are now
                        wl_build(m, (const int8_t*)m->a.sp, (const uint8_t*)m->a.fl, &cp->list,
                                 &cp->spec, 1);
                    const WlView V = wl_view(m, cp->list);
                    if (V.list) {
                        hipLaunchKernelGGL((k_cat_list<Real>), dim3(m->wl_nblk, kWlPer),
                                           dim3(256), 0, st, C, P, A, V, m->d_ctr, G, ng);
                        walked = true;
                    }
                }
                if (!walked) {
                    const int ntile_y = (p.ncy + m->tile_rows - 1) / m->tile_rows;
                    const int ntiles = p.ncx * ntile_y;
                    hipLaunchKernelGGL((k_cat_tile<Real>), dim3(8 * ((ntiles + 7) / 8)),
                                       dim3(256), 0, st, C, P, A, m->d_ctr, G, ng, ntile_y,
                                       ntiles);
                }
                if (defer && m->n_count) {
                    if (folds_in(ri + nf))
                        fold_pending = true;
                    else
                        hipLaunchKernelGGL(k_fold, dim3(nblk(m->n_count)), dim3(kBlock), 0, st,
                                           m->n_count, m->counts, m->delta);
                }
                cat_skip = nf - 1;
                off += npairs;
                break;
            }
            case EDM_RX_RECRUIT:
            case EDM_RX_DETACH:
            case EDM_RX_CONVERT: {
                if (m->group_start[ri] >= 0) {
                    const std::vector<int>& g = m->pass_groups[m->group_start[ri]];
                    PassGroupDev grp{};
                    grp.n = (int)g.size();
                    bool defer_any = false, local = false;
                    RecSet R{};
                    int nrec = 0;
                    for (int k = 0; k < grp.n; ++k) {
                        const edm_reaction& q = m->rx[g[k]];
                        PassDev& d = grp.p[k];
                        d.rec = 0;
                        if (q.kind == EDM_RX_RECRUIT) {
                            d.rec = nrec;
                            R.target[nrec] = q.a;
                            R.maxr[nrec] = q.max_recruits;
                            ++nrec;
                        }
                        d.kind = q.kind;
                        d.a = q.a;
                        d.b = q.b;
                        d.rxi = g[k];
                        d.prob = q.prob;
                        d.f1 = make_fb(m, q.fb1);
                        d.f2 = make_fb(m, q.fb2);
                        d.draw_off = draw_off[g[k]];
                        d.deferred = m->pass_deferred[g[k]];
                        defer_any = defer_any || d.deferred;
                        d.reach = pass_reach(m->rx, g, k);
                        grp.species |= d.reach;
                        d.cdf = nullptr;
                        d.sigma = 0.0;
                        if (q.kind == EDM_RX_RECRUIT && q.use_adapt) {
                            // the CDF of this step's counts on the rule's grid;
                            // no earlier pass of the run changes them
                            d.cdf = m->rcdf;
                            d.sigma = q.contact;
                            local = true;
                            hipLaunchKernelGGL(k_recruit_cdf<Real>, dim3(1), dim3(1024), 0, st,
                                               (const int32_t*)m->counts + d.f1.cbase,
                                               d.f1.gnx * d.f1.gny, (Real)d.f1.ginv, m->rcdf);
                        }
                    }
                    WlView V{nullptr, nullptr};
                    for (const PassPlan& pp : m->pass_plan)
                        if (pp.ri == (int)ri && pp.list >= 0 && m->wl_valid) {
                            if (pp.late)  // the run's particles as they are now
                                wl_build(m, (const int8_t*)m->a.sp, (const uint8_t*)m->a.fl,
                                         &pp.list, &pp.spec, 1);
                            V = wl_view(m, pp.list);
                        }
                    if (int rc = release_photons()) return rc;
                    hipLaunchKernelGGL((local ? k_passes<Real, true> : k_passes<Real, false>),
                                       V.list ? dim3(m->wl_nblk, kWlPer) : dim3(nblk(n)),
                                       dim3(kBlock), 0, st, n, V,
                                       (const int32_t*)m->a.id, P, grp, hook ? draws : nullptr,
                                       rng, step, (Real)p.Lx, (Real)p.Ly, m->d_ctr, G, ng,
                                       m->cand_id, m->cand_slot, m->cand_x, m->cand_y, m->n_cand,
                                       (int64_t)n);
                    if (nrec > 0)
                        hipLaunchKernelGGL(k_recruit_apply<Real>, dim3(nrec), dim3(1024), 0, st, P,
                                           R, n, (const int32_t*)m->cand_id,
                                           (const int32_t*)m->cand_slot, (const double*)m->cand_x,
                                           (const double*)m->cand_y, m->n_cand, (int64_t)n,
                                           m->d_ctr, G, ng);
                    if (defer_any && m->n_count)
                        hipLaunchKernelGGL(k_fold, dim3(nblk(m->n_count)), dim3(kBlock), 0, st,
                                           m->n_count, m->counts, m->delta);
                }
                off += (r.kind == EDM_RX_RECRUIT) ? recruit_draws(r, n) : (int64_t)n;
                break;
            }
            default: break;
        }
    }
    if (int rc = release_photons()) return rc;
    // post-move positions become the current positions
    std::swap(m->a.x, m->a.xn);
    std::swap(m->a.y, m->a.yn);
    m->step += 1;
    EDM_HIP(hipGetLastError());
    return 0;
}

int dyn_step_finish(edm_membrane* m, const double* d_draws_dev, bool hook) {
    const int rc = m->fp64 ? finish_t<double>(m, d_draws_dev, hook)
                           : finish_t<float>(m, d_draws_dev, hook);
    if (rc || !m->st_acc) return rc;
    return stats_accumulate(m);
}

template <typename Real>
static int pairs_t(edm_membrane* m, int32_t* host_pairs, int64_t cap, int64_t* n_out) {
    // The pair list of the last begin: positions (x, y) were swapped to the
    // post-move buffer by finish, so this must run between begin and finish.
    if (!m->need_pair_offsets) return EDM_ESTATE;
    const int64_t np = std::min(m->n_pairs, cap);
    int32_t* d = nullptr;
    EDM_HIP(hipMalloc(&d, (size_t)std::max<int64_t>(np, 1) * 2 * sizeof(int32_t)));
    Cells<Real> C = make_cells<Real>(m);
    hipLaunchKernelGGL(k_pair_list<Real>, dim3(nblk(m->n)), dim3(kBlock), 0, m->stream, m->n, C, d,
                       np);
    EDM_HIP(hipMemcpyAsync(host_pairs, d, (size_t)np * 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                           m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipFree(d));
    *n_out = m->n_pairs;
    return 0;
}

int dyn_pairs(edm_membrane* m, int32_t* pairs, int64_t cap, int64_t* n_out) {
    if (!m->begun) return EDM_ESTATE;
    return m->fp64 ? pairs_t<double>(m, pairs, cap, n_out) : pairs_t<float>(m, pairs, cap, n_out);
}

template <typename Real>
static int density_t(edm_membrane* m, int32_t species, int32_t grid, double* out) {
    if (grid < 0 || grid >= (int)m->grids.size() || species < 0 || species >= m->p.n_species)
        return EDM_EINVAL;
    if (m->count_off[(size_t)grid * m->p.n_species + species] < 0) return EDM_EINVAL;
    const int nc = m->grids[grid].nx * m->grids[grid].ny;
    double* d = nullptr;
    EDM_HIP(hipMalloc(&d, (size_t)nc * sizeof(double)));
    Grids<Real> G = make_grids<Real>(m);
    hipLaunchKernelGGL(k_density_out<Real>, dim3(nblk(nc)), dim3(kBlock), 0, m->stream, G, grid,
                       species, d, nc);
    EDM_HIP(hipMemcpyAsync(out, d, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipFree(d));
    return 0;
}

int dyn_density(edm_membrane* m, int32_t species, int32_t grid, double* out) {
    return m->fp64 ? density_t<double>(m, species, grid, out)
                   : density_t<float>(m, species, grid, out);
}

#include "edm_ens_dyn.h"

}  // namespace edm
