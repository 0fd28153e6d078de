// The reference's free kernels (ed_mrdt/dynamics.py __all__, :38-60) as
// stateless entry points on device arrays in the reference's own layouts:
// positions / forces / grad_D / noise f64[N, 2] row-major, species and bond
// partners i32[N], is_alive u8[N], pair lists i32[M, 2] in the reference's
// order, density fields f64[nx, ny] row-major.  Each entry point keeps the
// reference kernel's sequential semantics (pair order, index-order caps,
// sequential FP64 sums) on a parallel decomposition:
//
//   * pair list: counting sort into cells, per-cell descending index (the
//     reference's head-inserted chains), per-particle pair counts, a scan,
//     then each particle writes its pairs in the reference's neighbour order
//   * pair forces: each particle gathers its pairs in pair order (a CSR of
//     (pair, side) entries sorted by pair index), so its force is the
//     reference's running sum to the last bit
//   * binding: greedy matching by rounds over the accepted candidate edges
//     (an edge is taken when it is the lowest live edge at both ends), which
//     is the reference's sequential pass
//   * catalysis: a substrate converts iff any of its pairs accepts (the
//     enzyme species never changes, so pair order does not matter)
//   * recruitment: accept flags, a scan, and the first max_recruits
//     accepted in index order
//   * density: FP64 atomic adds of the same 1/area value, whose partial sums
//     do not depend on the order, so each cell holds the reference's
//     sequential sum
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "edm_device.h"
#include "edm_internal.h"

using namespace edm;

namespace {

constexpr int kB = 256;

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kB - 1) / kB); }

int kinvalid(const std::string& m) {
    set_error(m);
    return EDM_EINVAL;
}

__device__ __forceinline__ int ref_cell(double x, double s, int n) {
    return clamp_cell((int)(x / s), n);
}

__device__ __forceinline__ double min_image_d(double x1, double y1, double x2, double y2,
                                              double Lx, double Ly, int px, int py, double& dx,
                                              double& dy) {
    dx = x1 - x2;
    dy = y1 - y2;
    if (px) {
        if (dx > 0.5 * Lx) dx -= Lx;
        else if (dx < -0.5 * Lx) dx += Lx;
    }
    if (py) {
        if (dy > 0.5 * Ly) dy -= Ly;
        else if (dy < -0.5 * Ly) dy += Ly;
    }
    return dx * dx + dy * dy;
}

__device__ __forceinline__ double hill_d(double rho, double K, double n, int inh, double basal) {
    if (K > 0.0) {
        const double rn = pow(rho / K, n);
        if (inh) return 1.0 / (1.0 + rn);
        return basal + (1.0 - basal) * (rn / (1.0 + rn));
    }
    return 1.0;
}

// ------------------------------------------------------------------ scans
// Exclusive scans in three launches (block sums, one block over the sums,
// block-local scans plus the block's offset); the partial sums live in a
// caller-provided array of scan_partials(n) elements.
constexpr int kScanBlock = 1024;

constexpr int64_t scan_partials(int64_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }

template <typename In, typename Out>
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(const In* in, int64_t n, Out* part) {
    __shared__ Out s[kScanBlock];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    s[threadIdx.x] = i < n ? (Out)in[i] : (Out)0;
    __syncthreads();
    for (int w = kScanBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// exclusive scan of part[0..nb) in place, one block; part[nb] = total
template <typename Out>
__global__ __launch_bounds__(kScanBlock) void k_scan_part(Out* part, int nb) {
    __shared__ Out s[kScanBlock];
    __shared__ Out carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += kScanBlock) {
        const int i = base + (int)threadIdx.x;
        const Out v = i < nb ? part[i] : (Out)0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int w = 1; w < kScanBlock; w <<= 1) {
            const Out t = (int)threadIdx.x >= w ? s[threadIdx.x - w] : (Out)0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) part[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanBlock - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

// out[i] = part[block] + sum(in[block start .. i)), out[n] = total
template <typename In, typename Out>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const In* in, int64_t n,
                                                           const Out* part, int nb, Out* out) {
    __shared__ Out s[kScanBlock];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const Out v = i < n ? (Out)in[i] : (Out)0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 1; w < kScanBlock; w <<= 1) {
        const Out t = (int)threadIdx.x >= w ? s[threadIdx.x - w] : (Out)0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = part[blockIdx.x] + s[threadIdx.x] - v;
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = part[nb];
}

// out[0..n] = exclusive scan of in[0..n) with out[n] the total.  out may
// alias in.
template <typename In, typename Out>
void exclusive_scan(const In* in, int64_t n, Out* out, Out* part, hipStream_t st) {
    const int nb = (int)((n + kScanBlock - 1) / kScanBlock);
    if (nb > 0) hipLaunchKernelGGL((k_scan_sums<In, Out>), dim3(nb), dim3(kScanBlock), 0, st, in, n, part);
    hipLaunchKernelGGL((k_scan_part<Out>), dim3(1), dim3(kScanBlock), 0, st, part, nb);
    hipLaunchKernelGGL((k_scan_apply<In, Out>), dim3(std::max(nb, 1)), dim3(kScanBlock), 0, st, in,
                       n, (const Out*)part, nb, out);
}

__global__ void k_fill_u32(uint32_t* a, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ------------------------------------------------------- cell list + pairs
// dynamics.py:177-196 _build_cell_list: cell counts (rank of each particle
// among its cell's arrivals) into cnt[c]
__global__ void k_cl_count(const double* __restrict__ pos, int n, int nx, int ny, double csx,
                           double csy, uint32_t* cnt, uint32_t* rank) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = ref_cell(pos[2 * i], csx, nx) * ny + ref_cell(pos[2 * i + 1], csy, ny);
    rank[i] = atomicAdd(&cnt[c], 1u);
}

__global__ void k_cl_scatter(const double* __restrict__ pos, int n, int nx, int ny, double csx,
                             double csy, const uint32_t* __restrict__ start,
                             const uint32_t* __restrict__ rank, int32_t* items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = ref_cell(pos[2 * i], csx, nx) * ny + ref_cell(pos[2 * i + 1], csy, ny);
    items[start[c] + rank[i]] = i;
}

// each cell's members in descending index: the reference's chain order
// (head insertion while i ascends)
__global__ void k_cl_order(int ncell, const uint32_t* __restrict__ start, int32_t* items) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = (int)start[c], e = (int)start[c + 1];
    for (int a = b + 1; a < e; ++a) {
        const int v = items[a];
        int k = a - 1;
        while (k >= b && items[k] < v) {
            items[k + 1] = items[k];
            --k;
        }
        items[k + 1] = v;
    }
}

struct PairGeom {
    int nx, ny, px, py;
    double csx, csy, Lx, Ly, cut2;
};

// dynamics.py:220-278 _cell_list_pairs for particle i: with emit == false
// counts its pairs, else writes them from out_base on (pairs beyond
// max_pairs are dropped, as the reference drops them)
template <bool emit>
__device__ __forceinline__ uint32_t pairs_of(int i, const double* __restrict__ pos,
                                             const PairGeom& g, const uint32_t* __restrict__ start,
                                             const int32_t* __restrict__ items, int64_t out_base,
                                             int64_t max_pairs, int32_t* pairs) {
    const double xi = pos[2 * i], yi = pos[2 * i + 1];
    const int cxi = ref_cell(xi, g.csx, g.nx), cyi = ref_cell(yi, g.csy, g.ny);
    uint32_t k = 0;
    for (int dcx = -1; dcx <= 1; ++dcx) {
        for (int dcy = -1; dcy <= 1; ++dcy) {
            int ncx = cxi + dcx, ncy = cyi + dcy;
            if (g.px) ncx = ((ncx % g.nx) + g.nx) % g.nx;
            else if (ncx < 0 || ncx >= g.nx) continue;
            if (g.py) ncy = ((ncy % g.ny) + g.ny) % g.ny;
            else if (ncy < 0 || ncy >= g.ny) continue;
            const int c = ncx * g.ny + ncy;
            for (uint32_t a = start[c]; a < start[c + 1]; ++a) {
                const int j = items[a];
                if (j <= i) continue;
                double dx, dy;
                if (min_image_d(xi, yi, pos[2 * j], pos[2 * j + 1], g.Lx, g.Ly, g.px, g.py, dx,
                                dy) < g.cut2) {
                    if (emit) {
                        const int64_t o = out_base + k;
                        if (o < max_pairs) {
                            pairs[2 * o] = i;
                            pairs[2 * o + 1] = j;
                        }
                    }
                    ++k;
                }
            }
        }
    }
    return k;
}

__global__ void k_pairs_count(const double* __restrict__ pos, int n, PairGeom g,
                              const uint32_t* __restrict__ start, const int32_t* __restrict__ items,
                              uint32_t* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cnt[i] = pairs_of<false>(i, pos, g, start, items, 0, 0, nullptr);
}

__global__ void k_pairs_emit(const double* __restrict__ pos, int n, PairGeom g,
                             const uint32_t* __restrict__ start, const int32_t* __restrict__ items,
                             const uint64_t* __restrict__ off, int64_t max_pairs, int32_t* pairs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if ((int64_t)off[i] >= max_pairs) return;
    pairs_of<true>(i, pos, g, start, items, (int64_t)off[i], max_pairs, pairs);
}

// Scratch of the pair-list build, in the order it is carved from one block.
struct PairScratch {
    uint32_t* start;   // ncell + 1 (cell counts, then cell starts)
    uint32_t* rank;    // n (arrival rank in the cell, then pair counts)
    int32_t* items;    // n
    uint64_t* off;     // n + 1
    uint32_t* part32;  // scan_partials(ncell)
    uint64_t* part64;  // scan_partials(n)
};

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

constexpr size_t pair_scratch_bytes(int64_t n, int64_t ncell) {
    return align256((ncell + 1) * 4) + align256(n * 4) + align256(n * 4) +
           align256((n + 1) * 8) + align256(scan_partials(ncell + 1) * 4) +
           align256(scan_partials(n) * 8);
}

PairScratch carve_pairs(void* base, int64_t n, int64_t ncell) {
    char* p = (char*)base;
    PairScratch s;
    s.start = (uint32_t*)p;
    p += align256((ncell + 1) * 4);
    s.rank = (uint32_t*)p;
    p += align256(n * 4);
    s.items = (int32_t*)p;
    p += align256(n * 4);
    s.off = (uint64_t*)p;
    p += align256((n + 1) * 8);
    s.part32 = (uint32_t*)p;
    p += align256(scan_partials(ncell + 1) * 4);
    s.part64 = (uint64_t*)p;
    return s;
}

int check_pair_args(const double* pos, int32_t n, int32_t nx, int32_t ny, double csx, double csy,
                    double cutoff, int64_t max_pairs, int32_t* pairs, int64_t* n_pairs) {
    if (!pos || !pairs || !n_pairs || n < 0) return kinvalid("edm_cell_pairs: null argument");
    if (nx <= 0 || ny <= 0 || !(csx > 0.0) || !(csy > 0.0))
        return kinvalid("edm_cell_pairs: bad cell grid");
    if (!(cutoff > 0.0)) return kinvalid("edm_cell_pairs: cutoff must be positive");
    if (max_pairs < 0) return kinvalid("edm_cell_pairs: negative pair capacity");
    return 0;
}

int run_pairs(const double* pos, int32_t n, int32_t nx, int32_t ny, double csx, double csy,
              double Lx, double Ly, int32_t px, int32_t py, double cutoff, int64_t max_pairs,
              int32_t* pairs, int64_t* n_pairs, int64_t* n_found, void* scratch, hipStream_t st) {
    const int64_t ncell = (int64_t)nx * ny;
    PairScratch s = carve_pairs(scratch, n, ncell);
    // a fill kernel rather than hipMemsetAsync, so the call can be captured
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ncell + 1)), dim3(kB), 0, st, s.start,
                       ncell + 1, 0u);
    PairGeom g{nx, ny, px ? 1 : 0, py ? 1 : 0, csx, csy, Lx, Ly, cutoff * cutoff};
    if (n > 0) {
        hipLaunchKernelGGL(k_cl_count, dim3(blocks_for(n)), dim3(kB), 0, st, pos, n, nx, ny, csx,
                           csy, s.start, s.rank);
        exclusive_scan<uint32_t, uint32_t>(s.start, ncell, s.start, s.part32, st);
        hipLaunchKernelGGL(k_cl_scatter, dim3(blocks_for(n)), dim3(kB), 0, st, pos, n, nx, ny, csx,
                           csy, (const uint32_t*)s.start, (const uint32_t*)s.rank, s.items);
        hipLaunchKernelGGL(k_cl_order, dim3(blocks_for(ncell)), dim3(kB), 0, st, (int)ncell,
                           (const uint32_t*)s.start, s.items);
        hipLaunchKernelGGL(k_pairs_count, dim3(blocks_for(n)), dim3(kB), 0, st, pos, n, g,
                           (const uint32_t*)s.start, (const int32_t*)s.items, s.rank);
    }
    exclusive_scan<uint32_t, uint64_t>(s.rank, n, s.off, s.part64, st);
    if (n > 0)
        hipLaunchKernelGGL(k_pairs_emit, dim3(blocks_for(n)), dim3(kB), 0, st, pos, n, g,
                           (const uint32_t*)s.start, (const int32_t*)s.items,
                           (const uint64_t*)s.off, max_pairs, pairs);
    uint64_t total = 0;
    EDM_HIP(hipMemcpyAsync(&total, s.off + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    EDM_HIP(hipStreamSynchronize(st));
    EDM_HIP(hipGetLastError());
    *n_pairs = std::min<int64_t>((int64_t)total, max_pairs);
    if (n_found) *n_found = (int64_t)total;
    return 0;
}

// ------------------------------------------- allocation-free variant's scratch
// edm_cell_pairs_noalloc builds its pair list in this array instead of a
// per-call allocation: cell starts, arrival ranks, cell members, pair
// offsets and scan partials for at most EDM_NOALLOC_MAX_PARTICLES particles
// and EDM_NOALLOC_MAX_CELLS cells (20 MiB; every GPU the library runs on
// holds its own copy until the process ends).  Calls share it: overlapping
// calls on one GPU race.
constexpr int64_t kNoAllocMaxN = EDM_NOALLOC_MAX_PARTICLES;
constexpr int64_t kNoAllocMaxCells = EDM_NOALLOC_MAX_CELLS;
constexpr size_t kNoAllocBytes = pair_scratch_bytes(kNoAllocMaxN, kNoAllocMaxCells);

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The device address of `scratch` on each device (each GPU holds its own
// copy at its own address).  The array's name in host code is a host-side
// stand-in, so host code only ever passes these addresses.
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int noalloc_scratch(void** out) {
    int dev = 0;
    EDM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return kinvalid("edm_cell_pairs_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        EDM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// ------------------------------------------------------------- bd_step
// dynamics.py:323-406
__global__ void k_bd(double* pos, const double* __restrict__ D, const double* __restrict__ F,
                     const double* __restrict__ gD, double dt, double Lx, double Ly, int px, int py,
                     int n, const double* __restrict__ noise, const uint8_t* __restrict__ alive,
                     double maxd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i]) return;
    const double Di = D[i];
    const double sn = sqrt(2.0 * Di * dt);
    double dx = gD[2 * i] * dt;
    double dy = gD[2 * i + 1] * dt;
    dx += sn * noise[2 * i];
    dy += sn * noise[2 * i + 1];
    dx += Di * F[2 * i] * dt;
    dy += Di * F[2 * i + 1] * dt;
    const double dr2 = dx * dx + dy * dy;
    if (dr2 > maxd * maxd) {
        const double s = maxd / sqrt(dr2);
        dx *= s;
        dy *= s;
    }
    double x = pos[2 * i] + dx, y = pos[2 * i + 1] + dy;
    if (px) x = py_mod(x, Lx);
    else if (x < 0.0) x = -x;
    else if (x >= Lx) x = 2.0 * Lx - x;
    if (py) y = py_mod(y, Ly);
    else if (y < 0.0) y = -y;
    else if (y >= Ly) y = 2.0 * Ly - y;
    pos[2 * i] = x;
    pos[2 * i + 1] = y;
}

// ----------------------------------------------------- D(x) regions
// dynamics.py:441-519 apply_diffusion_field_regions_smooth
__global__ void k_dfield_regions(double* D, double* gD, const double* __restrict__ Dbase,
                                 const int32_t* __restrict__ sp, const double* __restrict__ pos,
                                 int n, const double* __restrict__ rcx,
                                 const double* __restrict__ rcy, const double* __restrict__ rr,
                                 const double* __restrict__ rf, int nreg, double w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = pos[2 * i], y = pos[2 * i + 1];
    const double base = Dbase[sp[i]];
    double factor = 1.0, gfx = 0.0, gfy = 0.0;
    for (int r = 0; r < nreg; ++r) {
        const double dxc = x - rcx[r], dyc = y - rcy[r];
        const double dsq = dxc * dxc + dyc * dyc;
        const double d = dsq > 1e-30 ? sqrt(dsq) : 1e-15;
        const double arg = (d - rr[r]) / w;
        double sig, dsig;
        if (arg > 100.0) {
            sig = 0.0;
            dsig = 0.0;
        } else if (arg < -100.0) {
            sig = 1.0;
            dsig = 0.0;
        } else {
            const double e = exp(arg);
            const double den = 1.0 + e;
            sig = 1.0 / den;
            dsig = -e / (den * den * w);
        }
        const double local = 1.0 + (rf[r] - 1.0) * sig;
        if (sig > 0.01) {
            factor = local;
            const double inv_d = 1.0 / d;
            gfx = (rf[r] - 1.0) * dsig * dxc * inv_d;
            gfy = (rf[r] - 1.0) * dsig * dyc * inv_d;
            break;
        }
    }
    D[i] = base * factor;
    gD[2 * i] = base * gfx;
    gD[2 * i + 1] = base * gfy;
}

// -------------------------------------------------------------- forces
// Each particle sums its pairs' forces in pair order: entries (pair << 1 |
// side) are bucketed per particle and sorted.
__global__ void k_fc_count(const int32_t* __restrict__ pairs, int64_t m, uint32_t* cnt) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    atomicAdd(&cnt[pairs[2 * p]], 1u);
    atomicAdd(&cnt[pairs[2 * p + 1]], 1u);
}

__global__ void k_fc_fill(const int32_t* __restrict__ pairs, int64_t m,
                          const uint64_t* __restrict__ off, uint32_t* cur, uint64_t* ent) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    ent[off[i] + atomicAdd(&cur[i], 1u)] = (uint64_t)p << 1;
    ent[off[j] + atomicAdd(&cur[j], 1u)] = ((uint64_t)p << 1) | 1u;
}

// kind 0: WCA (dynamics.py:666-739), 1: harmonic (dynamics.py:753-816)
__global__ void k_fc_sum(int kind, int n, const double* __restrict__ pos, double* F,
                         const int32_t* __restrict__ sp, const uint8_t* __restrict__ alive,
                         const int32_t* __restrict__ pairs, const double* __restrict__ sig_m,
                         const double* __restrict__ par_m, int ns, double Lx, double Ly, int px,
                         int py, const uint64_t* __restrict__ off, uint64_t* ent) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint64_t b = off[q], e = off[q + 1];
    for (uint64_t a = b + 1; a < e; ++a) {
        const uint64_t v = ent[a];
        uint64_t k = a;
        while (k > b && ent[k - 1] > v) {
            ent[k] = ent[k - 1];
            --k;
        }
        ent[k] = v;
    }
    double fx_q = F[2 * q], fy_q = F[2 * q + 1];
    for (uint64_t a = b; a < e; ++a) {
        const uint64_t v = ent[a];
        const int64_t p = (int64_t)(v >> 1);
        const int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        const int si = sp[i], sj = sp[j];
        const double sigma = sig_m[si * ns + sj];
        const double par = par_m[si * ns + sj];
        if (sigma <= 0.0 || par <= 0.0) continue;
        double dx, dy;
        const double dsq = min_image_d(pos[2 * i], pos[2 * i + 1], pos[2 * j], pos[2 * j + 1], Lx,
                                       Ly, px, py, dx, dy);
        double f;
        if (kind == 0) {
            if (dsq >= sigma * sigma * 1.2599210498948732 || dsq < 1e-30) continue;
            const double inv_r2 = 1.0 / dsq;
            const double s2 = sigma * sigma * inv_r2;
            const double s6 = s2 * s2 * s2;
            const double s12 = s6 * s6;
            f = 24.0 * par * (2.0 * s12 - s6) * inv_r2;
            const double fmax = 2400.0 * par / (sigma * sigma);
            if (f > fmax) f = fmax;
        } else {
            if (dsq >= sigma * sigma || dsq < 1e-30) continue;
            const double r = sqrt(dsq);
            f = par * (sigma - r) / r;
        }
        const double fx = f * dx, fy = f * dy;
        if (v & 1u) {
            fx_q -= fx;
            fy_q -= fy;
        } else {
            fx_q += fx;
            fy_q += fy;
        }
    }
    F[2 * q] = fx_q;
    F[2 * q + 1] = fy_q;
}

// ------------------------------------------------------------- binding
// dynamics.py:529-606: the pair's accepted candidate edge, if any
__global__ void k_bd_cand(const double* __restrict__ pos, const int32_t* __restrict__ sp,
                          const int32_t* __restrict__ bond, const uint8_t* __restrict__ alive,
                          int ra, int rb, double c2, double prob, const int32_t* __restrict__ pairs,
                          int64_t m, const double* __restrict__ u, double Lx, double Ly, int px,
                          int py, uint64_t* cand, uint32_t* n_cand, int32_t* ea, int32_t* eb) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    if (!alive[i] || !alive[j] || bond[i] >= 0 || bond[j] >= 0) return;
    int a, b;
    if (sp[i] == ra && sp[j] == rb) {
        a = i;
        b = j;
    } else if (sp[i] == rb && sp[j] == ra) {
        a = j;
        b = i;
    } else {
        return;
    }
    double dx, dy;
    if (min_image_d(pos[2 * i], pos[2 * i + 1], pos[2 * j], pos[2 * j + 1], Lx, Ly, px, py, dx,
                    dy) > c2)
        return;
    if (!(u[p] < prob)) return;
    const uint32_t e = atomicAdd(n_cand, 1u);
    cand[e] = (uint64_t)p;
    ea[e] = a;
    eb[e] = b;
}

// greedy matching in pair order, one workgroup: rounds accept every live
// edge that is the lowest live edge at both of its ends
__global__ __launch_bounds__(1024) void k_bd_match(const uint32_t* n_cand,
                                                   const uint64_t* __restrict__ cand,
                                                   const int32_t* __restrict__ ea,
                                                   const int32_t* __restrict__ eb, int32_t* sp,
                                                   int32_t* bond, uint8_t* alive, int prod,
                                                   unsigned long long* best, uint32_t* count) {
    const uint32_t m = *n_cand;
    volatile int32_t* vbond = bond;
    for (uint32_t round = 0; round <= m; ++round) {
        for (uint32_t e = threadIdx.x; e < m; e += blockDim.x)
            if (vbond[ea[e]] < 0 && vbond[eb[e]] < 0) {
                atomicExch(&best[ea[e]], ~0ull);
                atomicExch(&best[eb[e]], ~0ull);
            }
        __threadfence();
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < m; e += blockDim.x)
            if (vbond[ea[e]] < 0 && vbond[eb[e]] < 0) {
                atomicMin(&best[ea[e]], (unsigned long long)cand[e]);
                atomicMin(&best[eb[e]], (unsigned long long)cand[e]);
            }
        __threadfence();
        __syncthreads();
        int live = 0;
        for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
            const int a = ea[e], b = eb[e];
            if (vbond[a] >= 0 || vbond[b] >= 0) continue;
            const unsigned long long k = cand[e];
            if (atomicAdd(&best[a], 0ull) == k && atomicAdd(&best[b], 0ull) == k) {
                sp[a] = prod;
                alive[b] = 0;
                vbond[a] = b;
                vbond[b] = a;
                atomicAdd(count, 1u);
            } else {
                live = 1;
            }
        }
        __threadfence();
        if (!__syncthreads_or(live)) break;
    }
}

// dynamics.py:610-654 apply_dissociation
__global__ void k_dissoc_free(int32_t* sp, int32_t* bond, uint8_t* alive, double* pos, int n,
                              int prod, int ra, int rb, double prob, const double* __restrict__ u,
                              uint32_t* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i] || sp[i] != prod || bond[i] < 0) return;
    if (u[i] < prob) {
        const int j = bond[i];
        sp[i] = ra;
        sp[j] = rb;
        alive[j] = 1;
        pos[2 * j] = pos[2 * i];
        pos[2 * j + 1] = pos[2 * i + 1];
        bond[i] = -1;
        bond[j] = -1;
        atomicAdd(count, 1u);
    }
}

// ----------------------------------------------------------- catalysis
struct FbArgs {
    const double* density;  // nullptr: no feedback factor
    int nx, ny;
    double sx, sy, K, n;
    int inh;
    double basal;
};

__device__ __forceinline__ double fb_at(const FbArgs& f, double x, double y, int* gx = nullptr,
                                        int* gy = nullptr) {
    const int a = ref_cell(x, f.sx, f.nx), b = ref_cell(y, f.sy, f.ny);
    if (gx) *gx = a;
    if (gy) *gy = b;
    return hill_d(f.density[a * f.ny + b], f.K, f.n, f.inh, f.basal);
}

struct AdaptArgs {
    const double* w;  // nullptr: no adaptation factor
    int nx, ny;
    double K, n;
};

// Conversion probability of a pair whose substrate sits at (xs, ys):
// prob x Hill(rho) [x adaptation] (dynamics.py:826-1090).
__device__ __forceinline__ double cat_free_prob(double prob, const FbArgs& fb,
                                                const AdaptArgs& ad, double xs, double ys) {
    if (!fb.density) return prob;
    int gx, gy;
    const double hf = fb_at(fb, xs, ys, &gx, &gy);
    if (!ad.w) return prob * hf;
    double af = 1.0;
    if (ad.K > 0.0) {
        // adaptation_field[gx, gy] with the feedback grid's indices
        // (dynamics.py:1078); past the array the reference has no defined
        // value and the substrate's own adaptation cell is read instead
        int wc = gx * ad.ny + gy;
        if (wc >= ad.nx * ad.ny) {
            const double asx = (fb.nx * fb.sx) / ad.nx, asy = (fb.ny * fb.sy) / ad.ny;
            wc = ref_cell(xs, asx, ad.nx) * ad.ny + ref_cell(ys, asy, ad.ny);
        }
        af = 1.0 / (1.0 + pow(ad.w[wc] / ad.K, ad.n));
    }
    return prob * hf * af;
}

// dynamics.py:826-1090: the plain, feedback and feedback + adaptation forms
__global__ void k_cat_free(const double* __restrict__ pos, int32_t* sp,
                           const uint8_t* __restrict__ alive, int enz, int sub, int prod, double c2,
                           double prob, const int32_t* __restrict__ pairs, int64_t m,
                           const double* __restrict__ u, double Lx, double Ly, int px, int py,
                           FbArgs fb, AdaptArgs ad, uint32_t* count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    if (!alive[i] || !alive[j]) return;
    int s;
    const int si = sp[i], sj = sp[j];
    if (si == enz && sj == sub) s = j;
    else if (sj == enz && si == sub) s = i;
    else return;
    double dx, dy;
    if (min_image_d(pos[2 * i], pos[2 * i + 1], pos[2 * j], pos[2 * j + 1], Lx, Ly, px, py, dx,
                    dy) > c2)
        return;
    const double pe = cat_free_prob(prob, fb, ad, pos[2 * s], pos[2 * s + 1]);
    if (u[p] < pe && atomicCAS(&sp[s], sub, prod) == sub) atomicAdd(count, 1u);
}

// Autocatalysis (enzyme == product) over a given pair list: the pair's
// index is its time in the reference's loop, so a substrate converted at
// pair p acts as an enzyme from pair p + 1 on.  Same fixpoint as the
// membrane's k_autocat_resolve: t(s) is the first pair that converts s,
// lowered from "never" in rounds until none changes.  Candidates: pairs of
// an enzyme or substrate with a substrate, within contact, with the
// directions their uniform accepts (bit 62: i converts j, bit 61: j
// converts i) above the pair index.
__global__ void k_ac_free_cand(const double* __restrict__ pos, const int32_t* __restrict__ sp,
                               const uint8_t* __restrict__ alive, int enz, int sub, double c2,
                               double prob, const int32_t* __restrict__ pairs, int64_t m,
                               const double* __restrict__ u, double Lx, double Ly, int px, int py,
                               FbArgs fb, AdaptArgs ad, uint64_t* cand, uint32_t* n_cand,
                               int32_t* ea, int32_t* eb) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    if (!alive[i] || !alive[j]) return;
    const int si = sp[i], sj = sp[j];
    if ((si != enz && si != sub) || (sj != enz && sj != sub) || (si == enz && sj == enz)) return;
    double dx, dy;
    if (min_image_d(pos[2 * i], pos[2 * i + 1], pos[2 * j], pos[2 * j + 1], Lx, Ly, px, py, dx,
                    dy) > c2)
        return;
    const bool fwd = sj == sub && u[p] < cat_free_prob(prob, fb, ad, pos[2 * j], pos[2 * j + 1]);
    const bool bwd = si == sub && u[p] < cat_free_prob(prob, fb, ad, pos[2 * i], pos[2 * i + 1]);
    if (!fwd && !bwd) return;
    const uint32_t e = atomicAdd(n_cand, 1u);
    cand[e] = (uint64_t)p | ((uint64_t)fwd << 62) | ((uint64_t)bwd << 61);
    ea[e] = i;
    eb[e] = j;
}

__global__ __launch_bounds__(1024) void k_ac_free_resolve(const uint32_t* __restrict__ n_cand,
                                                          const uint64_t* __restrict__ cand,
                                                          const int32_t* __restrict__ ea,
                                                          const int32_t* __restrict__ eb,
                                                          int32_t* sp, int enz, int sub,
                                                          unsigned long long* t,
                                                          uint32_t* count) {
    constexpr unsigned long long kNever = ~0ull, kIdx = (1ull << 61) - 1ull;
    const uint32_t m = *n_cand;
    for (;;) {
        int changed = 0;
        for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
            const uint64_t k = cand[e];
            const unsigned long long p = k & kIdx;
            const int a = ea[e], b = eb[e];
            if ((k >> 62) & 1ull) {
                const bool enzyme = sp[a] == enz || atomicAdd(&t[a], 0ull) < p;
                if (enzyme && p < atomicAdd(&t[b], 0ull) && atomicMin(&t[b], p) > p) changed = 1;
            }
            if ((k >> 61) & 1ull) {
                const bool enzyme = sp[b] == enz || atomicAdd(&t[b], 0ull) < p;
                if (enzyme && p < atomicAdd(&t[a], 0ull) && atomicMin(&t[a], p) > p) changed = 1;
            }
        }
        __threadfence();
        if (!__syncthreads_or(changed)) break;
    }
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) {
        const int ends[2] = {ea[e], eb[e]};
        for (int q = 0; q < 2; ++q)
            if (atomicAdd(&t[ends[q]], 0ull) != kNever && atomicCAS(&sp[ends[q]], sub, enz) == sub)
                atomicAdd(count, 1u);
    }
}

// ----------------------------------------------------------- recruitment
// dynamics.py:1116-1458: accept flags of the dormant target particles; the
// candidate site and Hill factor(s) as the reference computes them
struct RecArgs {
    int target;
    double prob, Lx, Ly;
    const double* u_acc;
    const double* u_pos;  // [n, 2]; plain / feedback / dual
    FbArgs f1, f2;        // f2.density: dual feedback
    const double* cdf;    // local: CDF over f1's grid
    const double* u_loc;
    const double* offs;   // [n, 2] standard normals
    double sigma;
};

__device__ __forceinline__ bool rec_site(const RecArgs& R, int i, double& cx, double& cy) {
    if (R.cdf) {
        const int ncell = R.f1.nx * R.f1.ny;
        const double r = R.u_loc[i];
        int lo = 0, hi = ncell - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (R.cdf[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        const int gx = lo / R.f1.ny, gy = lo % R.f1.ny;
        cx = py_mod(((double)gx + 0.5) * R.f1.sx + R.offs[2 * i] * R.sigma, R.Lx);
        cy = py_mod(((double)gy + 0.5) * R.f1.sy + R.offs[2 * i + 1] * R.sigma, R.Ly);
        return R.u_acc[i] < R.prob * fb_at(R.f1, cx, cy);
    }
    cx = R.u_pos[2 * i] * R.Lx;
    cy = R.u_pos[2 * i + 1] * R.Ly;
    double pe = R.prob;
    if (R.f1.density && R.f2.density) pe = R.prob * fb_at(R.f1, cx, cy) * fb_at(R.f2, cx, cy);
    else if (R.f1.density) pe = R.prob * fb_at(R.f1, cx, cy);
    return R.u_acc[i] < pe;
}

__global__ void k_rec_flag(const int32_t* __restrict__ sp, const uint8_t* __restrict__ alive,
                           int n, RecArgs R, uint32_t* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double cx, cy;
    flag[i] = (!alive[i] && sp[i] == R.target && rec_site(R, i, cx, cy)) ? 1u : 0u;
}

__global__ void k_rec_apply(const int32_t* __restrict__ sp, uint8_t* alive, double* pos, int n,
                            RecArgs R, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ rank, uint32_t max_recruits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i] || rank[i] >= max_recruits) return;
    double cx, cy;
    rec_site(R, i, cx, cy);
    alive[i] = 1;
    pos[2 * i] = cx;
    pos[2 * i + 1] = cy;
}

// dynamics.py:1462-1555 detachment and 1566-1657 conversion
__global__ void k_detach_free(const int32_t* __restrict__ sp, uint8_t* alive,
                              const double* __restrict__ pos, int n, int target, double prob,
                              const double* __restrict__ u, FbArgs fb, uint32_t* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i] || sp[i] != target) return;
    const double pe = fb.density ? prob * fb_at(fb, pos[2 * i], pos[2 * i + 1]) : prob;
    if (u[i] < pe) {
        alive[i] = 0;
        atomicAdd(count, 1u);
    }
}

__global__ void k_convert_free(int32_t* sp, const uint8_t* __restrict__ alive,
                               const double* __restrict__ pos, int n, int from, int to, double prob,
                               const double* __restrict__ u, FbArgs fb, uint32_t* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i] || sp[i] != from) return;
    const double pe = fb.density ? prob * fb_at(fb, pos[2 * i], pos[2 * i + 1]) : prob;
    if (u[i] < pe) {
        sp[i] = to;
        atomicAdd(count, 1u);
    }
}

// ------------------------------------------------- density + adaptation
// dynamics.py:1668-1703: every addend is the same 1/area, so the partial
// sums a cell passes through do not depend on the order of the adds
__global__ void k_density_free(const double* __restrict__ pos, const int32_t* __restrict__ sp,
                               const uint8_t* __restrict__ alive, int n, int target, int nx, int ny,
                               double sx, double sy, double inv, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i] || sp[i] != target) return;
    const int c = ref_cell(pos[2 * i], sx, nx) * ny + ref_cell(pos[2 * i + 1], sy, ny);
    atomicAdd(&out[c], inv);
}

// dynamics.py:1094-1112
__global__ void k_adapt_free(double* w, const double* __restrict__ rho, int ncell, double inv) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) w[c] += inv * (rho[c] - w[c]);
}

// stream-ordered scratch of one call, released on every return path
struct AsyncBuf {
    char* p = nullptr;
    hipStream_t st = nullptr;
    int alloc(size_t bytes, hipStream_t s) {
        st = s;
        EDM_HIP(hipMallocAsync((void**)&p, std::max<size_t>(bytes, 256), st));
        return 0;
    }
    ~AsyncBuf() {
        if (p) (void)hipFreeAsync(p, st);
    }
};

// a counter in device memory for the calls that return a count
struct Counter {
    uint32_t* d = nullptr;
    hipStream_t st;
    int init(hipStream_t s) {
        st = s;
        EDM_HIP(hipMallocAsync((void**)&d, sizeof(uint32_t), st));
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, st, d, 1, 0u);
        return 0;
    }
    int read(int64_t* out) {
        uint32_t v = 0;
        EDM_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, st));
        EDM_HIP(hipStreamSynchronize(st));
        EDM_HIP(hipGetLastError());
        if (out) *out = v;
        return 0;
    }
    ~Counter() {
        if (d) (void)hipFreeAsync(d, st);
    }
};

FbArgs make_fb(const double* d, int32_t nx, int32_t ny, double sx, double sy, double K,
               double n, int32_t inh, double basal) {
    FbArgs f{d, nx, ny, sx, sy, K, n, inh ? 1 : 0, basal};
    return f;
}

bool bad_grid(const double* d, int32_t nx, int32_t ny, double sx, double sy) {
    return d && (nx <= 0 || ny <= 0 || !(sx > 0.0) || !(sy > 0.0));
}

}  // namespace

extern "C" {

int edm_device_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return kinvalid("edm_device_alloc: bad argument");
    *out = nullptr;
    EDM_HIP(hipMalloc(out, (size_t)std::max<int64_t>(bytes, 1)));
    return 0;
}

int edm_device_free(void* p) {
    if (p) EDM_HIP(hipFree(p));
    return 0;
}

int edm_copy_to_device(void* dst, const void* src, int64_t bytes) {
    if ((!dst || !src) && bytes > 0) return kinvalid("edm_copy_to_device: null argument");
    if (bytes > 0) EDM_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return 0;
}

int edm_copy_to_host(void* dst, const void* src, int64_t bytes) {
    if ((!dst || !src) && bytes > 0) return kinvalid("edm_copy_to_host: null argument");
    if (bytes > 0) EDM_HIP(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return 0;
}

int edm_cell_pairs(const double* positions, int32_t n, int32_t nx, int32_t ny, double cell_sx,
                   double cell_sy, double Lx, double Ly, int32_t periodic_x, int32_t periodic_y,
                   double cutoff, int64_t max_pairs, int32_t* pairs, int64_t* n_pairs,
                   int64_t* n_found, void* stream) {
    if (int rc = check_pair_args(positions, n, nx, ny, cell_sx, cell_sy, cutoff, max_pairs, pairs,
                                 n_pairs))
        return rc;
    hipStream_t st = (hipStream_t)stream;
    AsyncBuf buf;
    if (int rc = buf.alloc(pair_scratch_bytes(n, (int64_t)nx * ny), st)) return rc;
    return run_pairs(positions, n, nx, ny, cell_sx, cell_sy, Lx, Ly, periodic_x, periodic_y,
                     cutoff, max_pairs, pairs, n_pairs, n_found, buf.p, st);
}

int edm_cell_pairs_noalloc(const double* positions, int32_t n, int32_t nx, int32_t ny,
                           double cell_sx, double cell_sy, double Lx, double Ly,
                           int32_t periodic_x, int32_t periodic_y, double cutoff, int64_t max_pairs,
                           int32_t* pairs, int64_t* n_pairs, int64_t* n_found, void* stream) {
    if (int rc = check_pair_args(positions, n, nx, ny, cell_sx, cell_sy, cutoff, max_pairs, pairs,
                                 n_pairs))
        return rc;
    if ((int64_t)n > kNoAllocMaxN || (int64_t)nx * ny > kNoAllocMaxCells)
        return kinvalid("edm_cell_pairs_noalloc: at most " + std::to_string(kNoAllocMaxN) +
                        " particles and " + std::to_string(kNoAllocMaxCells) +
                        " cells fit its static scratch; got " + std::to_string(n) +
                        " particles and " + std::to_string((int64_t)nx * ny) + " cells");
    void* buf = nullptr;
    if (int rc = noalloc_scratch(&buf)) return rc;
    return run_pairs(positions, n, nx, ny, cell_sx, cell_sy, Lx, Ly, periodic_x, periodic_y,
                     cutoff, max_pairs, pairs, n_pairs, n_found, buf, (hipStream_t)stream);
}

int edm_bd_step(double* positions, const double* D, const double* forces, const double* grad_D,
                double dt, double Lx, double Ly, int32_t periodic_x, int32_t periodic_y, int32_t n,
                const double* noise, const uint8_t* is_alive, double max_displacement,
                void* stream) {
    if (n < 0 || (n > 0 && (!positions || !D || !forces || !grad_D || !noise || !is_alive)))
        return kinvalid("edm_bd_step: null argument");
    hipStream_t st = (hipStream_t)stream;
    if (n > 0)
        hipLaunchKernelGGL(k_bd, dim3(blocks_for(n)), dim3(kB), 0, st, positions, D, forces, grad_D,
                           dt, Lx, Ly, periodic_x ? 1 : 0, periodic_y ? 1 : 0, n, noise, is_alive,
                           max_displacement);
    EDM_HIP(hipGetLastError());
    return 0;
}

int edm_diffusion_field_regions(double* D, double* grad_D, const double* D_base,
                                const int32_t* species_id, const double* positions, int32_t n,
                                const double* region_cx, const double* region_cy,
                                const double* region_r, const double* region_Dfactor,
                                int32_t n_regions, double smoothing_width, void* stream) {
    if (n < 0 || n_regions < 0 ||
        (n > 0 && (!D || !grad_D || !D_base || !species_id || !positions)) ||
        (n_regions > 0 && (!region_cx || !region_cy || !region_r || !region_Dfactor)))
        return kinvalid("edm_diffusion_field_regions: null argument");
    hipStream_t st = (hipStream_t)stream;
    if (n > 0)
        hipLaunchKernelGGL(k_dfield_regions, dim3(blocks_for(n)), dim3(kB), 0, st, D, grad_D,
                           D_base, species_id, positions, n, region_cx, region_cy, region_r,
                           region_Dfactor, n_regions, smoothing_width);
    EDM_HIP(hipGetLastError());
    return 0;
}

int edm_pair_forces(int32_t kind, const double* positions, double* forces,
                    const int32_t* species_id, const uint8_t* is_alive, int32_t n,
                    const int32_t* pairs, int64_t n_pairs, const double* sigma_matrix,
                    const double* param_matrix, int32_t n_species, double Lx, double Ly,
                    int32_t periodic_x, int32_t periodic_y, void* stream) {
    if (kind != 0 && kind != 1) return kinvalid("edm_pair_forces: kind must be 0 (WCA) or 1 (harmonic)");
    if (n < 0 || n_pairs < 0 || n_species <= 0 ||
        (n > 0 && (!positions || !forces || !species_id || !is_alive)) ||
        (n_pairs > 0 && (!pairs || !sigma_matrix || !param_matrix)))
        return kinvalid("edm_pair_forces: null argument");
    if (n == 0 || n_pairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const size_t b_cnt = align256((size_t)n * 4), b_off = align256(((size_t)n + 1) * 8),
                 b_part = align256((size_t)scan_partials(n) * 8),
                 b_ent = align256((size_t)n_pairs * 2 * 8);
    AsyncBuf ab;
    if (int rc = ab.alloc(2 * b_cnt + b_off + b_part + b_ent, st)) return rc;
    char* buf = ab.p;
    uint32_t* cnt = (uint32_t*)buf;
    uint32_t* cur = (uint32_t*)(buf + b_cnt);
    uint64_t* off = (uint64_t*)(buf + 2 * b_cnt);
    uint64_t* part = (uint64_t*)(buf + 2 * b_cnt + b_off);
    uint64_t* ent = (uint64_t*)(buf + 2 * b_cnt + b_off + b_part);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(2 * (int64_t)b_cnt / 4)), dim3(kB), 0, st, cnt,
                       2 * (int64_t)b_cnt / 4, 0u);
    hipLaunchKernelGGL(k_fc_count, dim3(blocks_for(n_pairs)), dim3(kB), 0, st, pairs, n_pairs, cnt);
    exclusive_scan<uint32_t, uint64_t>(cnt, n, off, part, st);
    hipLaunchKernelGGL(k_fc_fill, dim3(blocks_for(n_pairs)), dim3(kB), 0, st, pairs, n_pairs,
                       (const uint64_t*)off, cur, ent);
    hipLaunchKernelGGL(k_fc_sum, dim3(blocks_for(n)), dim3(kB), 0, st, (int)kind, n, positions,
                       forces, species_id, is_alive, pairs, sigma_matrix, param_matrix, n_species,
                       Lx, Ly, periodic_x ? 1 : 0, periodic_y ? 1 : 0, (const uint64_t*)off, ent);
    EDM_HIP(hipGetLastError());
    return 0;
}

int edm_binding(const double* positions, int32_t* species_id, int32_t* bond_partner,
                uint8_t* is_alive, int32_t n, int32_t reactant_a, int32_t reactant_b,
                int32_t product_id, double contact_radius_um, double prob_bind,
                const int32_t* pairs, int64_t n_pairs, const double* rng_uniform, double Lx,
                double Ly, int32_t periodic_x, int32_t periodic_y, int64_t* n_bound,
                void* stream) {
    if (n < 0 || n_pairs < 0 || n_pairs > 0xffffffffLL ||
        (n > 0 && (!positions || !species_id || !bond_partner || !is_alive)) ||
        (n_pairs > 0 && (!pairs || !rng_uniform)))
        return kinvalid("edm_binding: bad argument");
    if (n_bound) *n_bound = 0;
    if (n == 0 || n_pairs == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const size_t b_c = align256((size_t)n_pairs * 8), b_e = align256((size_t)n_pairs * 4),
                 b_best = align256((size_t)n * 8);
    AsyncBuf ab;
    if (int rc = ab.alloc(b_c + 2 * b_e + b_best + 256, st)) return rc;
    char* buf = ab.p;
    uint64_t* cand = (uint64_t*)buf;
    int32_t* ea = (int32_t*)(buf + b_c);
    int32_t* eb = (int32_t*)(buf + b_c + b_e);
    unsigned long long* best = (unsigned long long*)(buf + b_c + 2 * b_e);
    uint32_t* ctr = (uint32_t*)(buf + b_c + 2 * b_e + b_best);  // [0] candidates, [1] bound
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, ctr, 2, 0u);
    hipLaunchKernelGGL(k_bd_cand, dim3(blocks_for(n_pairs)), dim3(kB), 0, st, positions, species_id,
                       bond_partner, is_alive, reactant_a, reactant_b,
                       contact_radius_um * contact_radius_um, prob_bind, pairs, n_pairs,
                       rng_uniform, Lx, Ly, periodic_x ? 1 : 0, periodic_y ? 1 : 0, cand, ctr, ea,
                       eb);
    hipLaunchKernelGGL(k_bd_match, dim3(1), dim3(1024), 0, st, (const uint32_t*)ctr,
                       (const uint64_t*)cand, (const int32_t*)ea, (const int32_t*)eb, species_id,
                       bond_partner, is_alive, product_id, best, ctr + 1);
    uint32_t nb = 0;
    EDM_HIP(hipMemcpyAsync(&nb, ctr + 1, 4, hipMemcpyDeviceToHost, st));
    EDM_HIP(hipStreamSynchronize(st));
    EDM_HIP(hipGetLastError());
    if (n_bound) *n_bound = nb;
    return 0;
}

int edm_dissociation(int32_t* species_id, int32_t* bond_partner, uint8_t* is_alive,
                     double* positions, int32_t n, int32_t product_id, int32_t reactant_a,
                     int32_t reactant_b, double prob_dissoc, const double* rng_uniform,
                     int64_t* n_dissociated, void* stream) {
    if (n < 0 || (n > 0 && (!species_id || !bond_partner || !is_alive || !positions || !rng_uniform)))
        return kinvalid("edm_dissociation: null argument");
    if (product_id == reactant_a || product_id == reactant_b)
        return kinvalid("edm_dissociation: the product must differ from both reactants");
    hipStream_t st = (hipStream_t)stream;
    Counter c;
    if (int rc = c.init(st)) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_dissoc_free, dim3(blocks_for(n)), dim3(kB), 0, st, species_id,
                           bond_partner, is_alive, positions, n, product_id, reactant_a, reactant_b,
                           prob_dissoc, rng_uniform, c.d);
    return c.read(n_dissociated);
}

int edm_catalysis(const double* positions, int32_t* species_id, const uint8_t* is_alive,
                  int32_t n, int32_t enzyme_id, int32_t substrate_id, int32_t product_id,
                  double contact_radius_um, double prob, const int32_t* pairs, int64_t n_pairs,
                  const double* rng_uniform, double Lx, double Ly, int32_t periodic_x,
                  int32_t periodic_y, const double* density_field, int32_t nx_grid,
                  int32_t ny_grid, double cell_sx, double cell_sy, double K_half, double n_hill,
                  int32_t is_inhibition, double basal_fraction, const double* adaptation_field,
                  int32_t nx_adapt, int32_t ny_adapt, double K_adapt, double n_adapt,
                  int64_t* n_converted, void* stream) {
    if (n < 0 || n_pairs < 0 || (n_pairs > 0 && (!positions || !species_id || !is_alive || !pairs ||
                                                 !rng_uniform)))
        return kinvalid("edm_catalysis: null argument");
    if (enzyme_id == substrate_id)
        return kinvalid("edm_catalysis: the enzyme must differ from the substrate");
    if (bad_grid(density_field, nx_grid, ny_grid, cell_sx, cell_sy) ||
        (adaptation_field && (!density_field || nx_adapt <= 0 || ny_adapt <= 0)))
        return kinvalid("edm_catalysis: bad density or adaptation grid");
    hipStream_t st = (hipStream_t)stream;
    Counter c;
    if (int rc = c.init(st)) return rc;
    FbArgs fb = make_fb(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                        is_inhibition, basal_fraction);
    AdaptArgs ad{adaptation_field, nx_adapt, ny_adapt, K_adapt, n_adapt};
    if (n_pairs > 0 && enzyme_id == product_id) {
        if (n_pairs > 0xffffffffLL) return kinvalid("edm_catalysis: more than 2^32 pairs");
        const size_t b_c = align256((size_t)n_pairs * 8), b_e = align256((size_t)n_pairs * 4),
                     b_t = align256((size_t)std::max(n, 1) * 8);
        AsyncBuf ab;
        if (int rc = ab.alloc(b_c + 2 * b_e + b_t + 256, st)) return rc;
        char* buf = ab.p;
        uint64_t* cand = (uint64_t*)buf;
        int32_t* ea = (int32_t*)(buf + b_c);
        int32_t* eb = (int32_t*)(buf + b_c + b_e);
        unsigned long long* t = (unsigned long long*)(buf + b_c + 2 * b_e);
        uint32_t* nc = (uint32_t*)(buf + b_c + 2 * b_e + b_t);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, nc, 1, 0u);
        if (n > 0)
            hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(2 * (int64_t)n)), dim3(kB), 0, st,
                               (uint32_t*)t, 2 * (int64_t)n, 0xffffffffu);
        hipLaunchKernelGGL(k_ac_free_cand, dim3(blocks_for(n_pairs)), dim3(kB), 0, st, positions,
                           (const int32_t*)species_id, is_alive, enzyme_id, substrate_id,
                           contact_radius_um * contact_radius_um, prob, pairs, n_pairs,
                           rng_uniform, Lx, Ly, periodic_x ? 1 : 0, periodic_y ? 1 : 0, fb, ad,
                           cand, nc, ea, eb);
        hipLaunchKernelGGL(k_ac_free_resolve, dim3(1), dim3(1024), 0, st, (const uint32_t*)nc,
                           (const uint64_t*)cand, (const int32_t*)ea, (const int32_t*)eb,
                           species_id, enzyme_id, substrate_id, t, c.d);
        return c.read(n_converted);
    }
    if (n_pairs > 0)
        hipLaunchKernelGGL(k_cat_free, dim3(blocks_for(n_pairs)), dim3(kB), 0, st, positions,
                           species_id, is_alive, enzyme_id, substrate_id, product_id,
                           contact_radius_um * contact_radius_um, prob, pairs, n_pairs,
                           rng_uniform, Lx, Ly, periodic_x ? 1 : 0, periodic_y ? 1 : 0, fb, ad,
                           c.d);
    return c.read(n_converted);
}

int edm_recruitment(const int32_t* species_id, uint8_t* is_alive, double* positions, int32_t n,
                    int32_t target_species_id, double prob, double Lx, double Ly,
                    const double* rng_accept, const double* rng_positions, int32_t max_recruits,
                    const double* density_1, int32_t nx_1, int32_t ny_1, double sx_1, double sy_1,
                    double K_1, double n_1, int32_t inh_1, double basal_1, const double* density_2,
                    int32_t nx_2, int32_t ny_2, double sx_2, double sy_2, double K_2, double n_2,
                    int32_t inh_2, double basal_2, const double* cdf, const double* rng_location,
                    const double* rng_offsets, double recruit_sigma, int64_t* n_recruited,
                    void* stream) {
    const bool local = cdf != nullptr;
    if (n < 0 || (n > 0 && (!species_id || !is_alive || !positions || !rng_accept)) ||
        (n > 0 && !local && !rng_positions) ||
        (local && (!density_1 || density_2 || !rng_location || !rng_offsets)))
        return kinvalid("edm_recruitment: bad argument");
    if (density_2 && !density_1) return kinvalid("edm_recruitment: second feedback without a first");
    if (bad_grid(density_1, nx_1, ny_1, sx_1, sy_1) || bad_grid(density_2, nx_2, ny_2, sx_2, sy_2))
        return kinvalid("edm_recruitment: bad density grid");
    if (n_recruited) *n_recruited = 0;
    if (n == 0 || max_recruits <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    RecArgs R;
    R.target = target_species_id;
    R.prob = prob;
    R.Lx = Lx;
    R.Ly = Ly;
    R.u_acc = rng_accept;
    R.u_pos = rng_positions;
    R.f1 = make_fb(density_1, nx_1, ny_1, sx_1, sy_1, K_1, n_1, inh_1, basal_1);
    R.f2 = make_fb(density_2, nx_2, ny_2, sx_2, sy_2, K_2, n_2, inh_2, basal_2);
    R.cdf = cdf;
    R.u_loc = rng_location;
    R.offs = rng_offsets;
    R.sigma = recruit_sigma;
    const size_t b_f = align256((size_t)n * 4), b_r = align256(((size_t)n + 1) * 4),
                 b_p = align256((size_t)scan_partials(n) * 4);
    AsyncBuf ab;
    if (int rc = ab.alloc(b_f + b_r + b_p, st)) return rc;
    char* buf = ab.p;
    uint32_t* flag = (uint32_t*)buf;
    uint32_t* rank = (uint32_t*)(buf + b_f);
    uint32_t* part = (uint32_t*)(buf + b_f + b_r);
    hipLaunchKernelGGL(k_rec_flag, dim3(blocks_for(n)), dim3(kB), 0, st, species_id,
                       (const uint8_t*)is_alive, n, R, flag);
    exclusive_scan<uint32_t, uint32_t>(flag, n, rank, part, st);
    hipLaunchKernelGGL(k_rec_apply, dim3(blocks_for(n)), dim3(kB), 0, st, species_id, is_alive,
                       positions, n, R, (const uint32_t*)flag, (const uint32_t*)rank,
                       (uint32_t)max_recruits);
    uint32_t tot = 0;
    EDM_HIP(hipMemcpyAsync(&tot, rank + n, 4, hipMemcpyDeviceToHost, st));
    EDM_HIP(hipStreamSynchronize(st));
    EDM_HIP(hipGetLastError());
    if (n_recruited) *n_recruited = std::min<int64_t>(tot, max_recruits);
    return 0;
}

int edm_detachment(const int32_t* species_id, uint8_t* is_alive, const double* positions,
                   int32_t n, int32_t target_species_id, double prob, const double* rng_uniform,
                   const double* density_field, int32_t nx_grid, int32_t ny_grid, double cell_sx,
                   double cell_sy, double K_half, double n_hill, int32_t is_inhibition,
                   double basal_fraction, int64_t* n_detached, void* stream) {
    if (n < 0 || (n > 0 && (!species_id || !is_alive || !rng_uniform)) ||
        (density_field && n > 0 && !positions))
        return kinvalid("edm_detachment: null argument");
    if (bad_grid(density_field, nx_grid, ny_grid, cell_sx, cell_sy))
        return kinvalid("edm_detachment: bad density grid");
    hipStream_t st = (hipStream_t)stream;
    Counter c;
    if (int rc = c.init(st)) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_detach_free, dim3(blocks_for(n)), dim3(kB), 0, st, species_id,
                           is_alive, positions, n, target_species_id, prob, rng_uniform,
                           make_fb(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                   n_hill, is_inhibition, basal_fraction),
                           c.d);
    return c.read(n_detached);
}

int edm_conversion(int32_t* species_id, const uint8_t* is_alive, const double* positions,
                   int32_t n, int32_t from_id, int32_t to_id, double prob,
                   const double* rng_uniform, const double* density_field, int32_t nx_grid,
                   int32_t ny_grid, double cell_sx, double cell_sy, double K_half, double n_hill,
                   int32_t is_inhibition, double basal_fraction, int64_t* n_converted,
                   void* stream) {
    if (n < 0 || (n > 0 && (!species_id || !is_alive || !rng_uniform)) ||
        (density_field && n > 0 && !positions))
        return kinvalid("edm_conversion: null argument");
    if (bad_grid(density_field, nx_grid, ny_grid, cell_sx, cell_sy))
        return kinvalid("edm_conversion: bad density grid");
    if (from_id == to_id) return kinvalid("edm_conversion: from and to species must differ");
    hipStream_t st = (hipStream_t)stream;
    Counter c;
    if (int rc = c.init(st)) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_convert_free, dim3(blocks_for(n)), dim3(kB), 0, st, species_id,
                           is_alive, positions, n, from_id, to_id, prob, rng_uniform,
                           make_fb(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                   n_hill, is_inhibition, basal_fraction),
                           c.d);
    return c.read(n_converted);
}

int edm_density_field(const double* positions, const int32_t* species_id, const uint8_t* is_alive,
                      int32_t n, int32_t target_species_id, int32_t nx_grid, int32_t ny_grid,
                      double cell_sx, double cell_sy, double* density_out, void* stream) {
    if (n < 0 || !density_out || (n > 0 && (!positions || !species_id || !is_alive)))
        return kinvalid("edm_density_field: null argument");
    if (nx_grid <= 0 || ny_grid <= 0) return kinvalid("edm_density_field: bad grid");
    const double area = cell_sx * cell_sy;
    const double inv = area > 0.0 ? 1.0 / area : 0.0;
    hipStream_t st = (hipStream_t)stream;
    if (n > 0)
        hipLaunchKernelGGL(k_density_free, dim3(blocks_for(n)), dim3(kB), 0, st, positions,
                           species_id, is_alive, n, target_species_id, nx_grid, ny_grid, cell_sx,
                           cell_sy, inv, density_out);
    EDM_HIP(hipGetLastError());
    return 0;
}

int edm_adaptation_update(double* adaptation_field, const double* density_field, int32_t nx,
                          int32_t ny, double dt, double tau_w, void* stream) {
    if (!adaptation_field || !density_field || nx < 0 || ny < 0)
        return kinvalid("edm_adaptation_update: bad argument");
    const int ncell = nx * ny;
    hipStream_t st = (hipStream_t)stream;
    if (ncell > 0)
        hipLaunchKernelGGL(k_adapt_free, dim3(blocks_for(ncell)), dim3(kB), 0, st,
                           adaptation_field, density_field, ncell, dt / tau_w);
    EDM_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
