// Observables of a membrane computed on the device (the statistics the
// reference computes on the host from copied state):
//
//   * mean squared displacement with periodic jumps unwrapped, the
//     reference's B1 benchmark (validation.py:256-335): after every step each
//     particle's minimum-image displacement since the previous step is added
//     to its accumulator (by particle index, so the per-step cell sort does
//     not matter); a reduction gives sum |r(t) - r(0)|^2 over the alive
//     particles of chosen species
//   * pair-distance histogram behind the RDF g(r) of
//     spatial_correlation_ground_truth (analysis.py:733-790): all pairs of
//     alive particles of chosen species, tiled through LDS, one histogram
//     per workgroup in LDS
//   * per-pixel TCSPC moments of the current frame behind mean_arrival_time
//     (analysis.py:87-127) and phasor_transform (analysis.py:497-566): one
//     wavefront per pixel reduces its 256 bins to sum I, sum I cos(phi_k),
//     sum I sin(phi_k), sum I t_k
//
// Reductions run in a fixed order (block partials summed by one block), so
// repeated runs give the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "edm_device.h"
#include "edm_internal.h"

namespace edm {

namespace {

constexpr int kSB = 256;
inline int sblocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kSB - 1) / kSB); }

template <typename Real>
__global__ void k_st_reset(int n, const int32_t* __restrict__ id, const Real* __restrict__ x,
                           const Real* __restrict__ y, Real* prev, double* acc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = id[t];
    prev[2 * i] = x[t];
    prev[2 * i + 1] = y[t];
    acc[2 * i] = 0.0;
    acc[2 * i + 1] = 0.0;
}

// validation.py:276-290: delta = after - before, each component wrapped by
// the nearest multiple of the box
template <typename Real>
__global__ void k_st_accum(int n, const int32_t* __restrict__ id, const Real* __restrict__ x,
                           const Real* __restrict__ y, Real* prev, double* acc, double Lx,
                           double Ly) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = id[t];
    double dx = (double)x[t] - (double)prev[2 * i];
    double dy = (double)y[t] - (double)prev[2 * i + 1];
    dx -= Lx * rint(dx / Lx);
    dy -= Ly * rint(dy / Ly);
    acc[2 * i] += dx;
    acc[2 * i + 1] += dy;
    prev[2 * i] = x[t];
    prev[2 * i + 1] = y[t];
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    return v;
}

// block partials: [b][0] = sum |acc|^2, [b][1] = count
__global__ __launch_bounds__(256) void k_st_msd(int n, const int32_t* __restrict__ id,
                                                const int8_t* __restrict__ sp,
                                                const uint8_t* __restrict__ fl, uint32_t mask,
                                                const double* __restrict__ acc, double* part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0, c = 0.0;
    if (t < n && (fl[t] & EDM_FLAG_ALIVE) && ((mask >> sp[t]) & 1u)) {
        const int i = id[t];
        r2 = acc[2 * i] * acc[2 * i] + acc[2 * i + 1] * acc[2 * i + 1];
        c = 1.0;
    }
    r2 = wave_sum_d(r2);
    c = wave_sum_d(c);
    __shared__ double s[2][4];
    if (lane_id() == 0) {
        s[0][threadIdx.x >> 6] = r2;
        s[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ((s[0][0] + s[0][1]) + s[0][2]) + s[0][3];
        part[2 * blockIdx.x + 1] = ((s[1][0] + s[1][1]) + s[1][2]) + s[1][3];
    }
}

// sums the block partials in block order (one block, 256 threads striding,
// then a fixed tree), into part[2 * nb], part[2 * nb + 1]
__global__ __launch_bounds__(256) void k_st_final(int nb, double* part) {
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        a += part[2 * k];
        b += part[2 * k + 1];
    }
    __shared__ double s[2][256];
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * nb] = s[0][0];
        part[2 * nb + 1] = s[1][0];
    }
}

// unordered pairs of alive particles of the species in mask: blockIdx.x
// takes a tile of 256 rows, loops over the column tiles from its own, and
// counts each pair once (column slot > row slot)
template <typename Real>
__global__ __launch_bounds__(256) void k_st_pairs(int n, const Real* __restrict__ x,
                                                  const Real* __restrict__ y,
                                                  const int8_t* __restrict__ sp,
                                                  const uint8_t* __restrict__ fl, uint32_t mask,
                                                  double Lx, double Ly, int px, int py,
                                                  double dr, int n_bins,
                                                  unsigned long long* hist) {
    extern __shared__ unsigned int lh[];  // n_bins
    __shared__ double tx[256], ty[256];
    __shared__ int tok[256];
    for (int b = threadIdx.x; b < n_bins; b += blockDim.x) lh[b] = 0u;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool mine = i < n && (fl[i] & EDM_FLAG_ALIVE) && ((mask >> sp[i]) & 1u);
    const double xi = mine ? (double)x[i] : 0.0, yi = mine ? (double)y[i] : 0.0;
    for (int j0 = blockIdx.x * blockDim.x; j0 < n; j0 += blockDim.x) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        const bool ok = j < n && (fl[j] & EDM_FLAG_ALIVE) && ((mask >> sp[j]) & 1u);
        tx[threadIdx.x] = ok ? (double)x[j] : 0.0;
        ty[threadIdx.x] = ok ? (double)y[j] : 0.0;
        tok[threadIdx.x] = ok;
        __syncthreads();
        if (!mine) continue;
        const int kmax = min((int)blockDim.x, n - j0);
        for (int k = 0; k < kmax; ++k) {
            if (!tok[k] || j0 + k <= i) continue;
            double dx = tx[k] - xi, dy = ty[k] - yi;
            if (px) {
                if (dx > 0.5 * Lx) dx -= Lx;
                else if (dx < -0.5 * Lx) dx += Lx;
            }
            if (py) {
                if (dy > 0.5 * Ly) dy -= Ly;
                else if (dy < -0.5 * Ly) dy += Ly;
            }
            const int b = (int)(sqrt(dx * dx + dy * dy) / dr);
            if (b < n_bins) atomicAdd(&lh[b], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < n_bins; b += blockDim.x)
        if (lh[b]) atomicAdd(&hist[b], (unsigned long long)lh[b]);
}

// analysis.py:87-127, 497-566: one wavefront per pixel of the channel;
// out[pixel] = {sum I, sum I cos(phi_k), sum I sin(phi_k), sum I t_k} with
// phi_k = 2 pi h (k + 0.5) / n_bins and t_k = (k + 0.5) bin width (ns)
__global__ __launch_bounds__(256) void k_st_flim(const uint32_t* __restrict__ flim, int64_t npix,
                                                 int n_bins, int harmonic, double bin_ns,
                                                 double* out) {
    extern __shared__ double tab[];  // [2][n_bins]: cos, sin
    for (int k = threadIdx.x; k < n_bins; k += blockDim.x) {
        double s, c;
        det_sincos(2.0 * 3.141592653589793 * harmonic * (k + 0.5) / n_bins, s, c);
        tab[k] = c;
        tab[n_bins + k] = s;
    }
    __syncthreads();
    const int64_t p = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (p >= npix) return;  // whole wavefronts
    const uint32_t* h = flim + p * n_bins;
    double I = 0.0, gc = 0.0, gs = 0.0, tt = 0.0;
    for (int k = lane_id(); k < n_bins; k += 64) {
        const double v = (double)h[k];
        I += v;
        gc += v * tab[k];
        gs += v * tab[n_bins + k];
        tt += v * ((k + 0.5) * bin_ns);
    }
    I = wave_sum_d(I);
    gc = wave_sum_d(gc);
    gs = wave_sum_d(gs);
    tt = wave_sum_d(tt);
    if (lane_id() == 0) {
        out[4 * p] = I;
        out[4 * p + 1] = gc;
        out[4 * p + 2] = gs;
        out[4 * p + 3] = tt;
    }
}

template <typename Real>
int open_t(edm_membrane* m) {
    const int n = m->n;
    if (!m->st_acc) {
        EDM_HIP(hipMalloc(&m->st_acc, (size_t)n * 2 * sizeof(double)));
        EDM_HIP(hipMalloc(&m->st_prev, (size_t)n * 2 * sizeof(Real)));
        EDM_HIP(hipMalloc(&m->st_part, ((size_t)sblocks(n) + 1) * 2 * sizeof(double)));
    }
    hipLaunchKernelGGL(k_st_reset<Real>, dim3(sblocks(n)), dim3(kSB), 0, m->stream, n,
                       (const int32_t*)m->a.id, (const Real*)m->a.x, (const Real*)m->a.y,
                       (Real*)m->st_prev, m->st_acc);
    EDM_HIP(hipGetLastError());
    return 0;
}

template <typename Real>
int accum_t(edm_membrane* m) {
    hipLaunchKernelGGL(k_st_accum<Real>, dim3(sblocks(m->n)), dim3(kSB), 0, m->stream, m->n,
                       (const int32_t*)m->a.id, (const Real*)m->a.x, (const Real*)m->a.y,
                       (Real*)m->st_prev, m->st_acc, m->p.Lx, m->p.Ly);
    EDM_HIP(hipGetLastError());
    return 0;
}

template <typename Real>
int pairs_hist_t(edm_membrane* m, uint32_t mask, int n_bins, double r_max) {
    const size_t lds = (size_t)n_bins * sizeof(unsigned int);
    hipLaunchKernelGGL(k_st_pairs<Real>, dim3(sblocks(m->n)), dim3(kSB), lds, m->stream, m->n,
                       (const Real*)m->a.x, (const Real*)m->a.y, (const int8_t*)m->a.sp,
                       (const uint8_t*)m->a.fl, mask, m->p.Lx, m->p.Ly, m->p.periodic_x,
                       m->p.periodic_y, r_max / n_bins, n_bins, m->st_hist);
    EDM_HIP(hipGetLastError());
    return 0;
}

int sinvalid(const std::string& msg) {
    set_error(msg);
    return EDM_EINVAL;
}

}  // namespace

int stats_accumulate(edm_membrane* m) {
    return m->fp64 ? accum_t<double>(m) : accum_t<float>(m);
}

void stats_free(edm_membrane* m) {
    void* b[] = {m->st_acc, m->st_prev, m->st_part, m->st_hist, m->st_pix};
    for (void* p : b)
        if (p) (void)hipFree(p);
    m->st_acc = nullptr;
    m->st_prev = nullptr;
    m->st_part = nullptr;
    m->st_hist = nullptr;
    m->st_pix = nullptr;
}

}  // namespace edm

using namespace edm;

extern "C" {

int edm_stats_open(edm_membrane* m) {
    if (!m) return sinvalid("edm_stats_open: null handle");
    graph_touch(m);
    if (m->begun) {
        set_error("edm_stats_open: a parity step is open");
        return EDM_ESTATE;
    }
    return m->fp64 ? open_t<double>(m) : open_t<float>(m);
}

int edm_stats_close(edm_membrane* m) {
    if (!m) return sinvalid("edm_stats_close: null handle");
    graph_touch(m);
    EDM_HIP(hipStreamSynchronize(m->stream));
    stats_free(m);
    return 0;
}

int edm_stats_msd(edm_membrane* m, uint32_t species_mask, double* sum_r2, int64_t* count) {
    if (!m || !sum_r2 || !count) return sinvalid("edm_stats_msd: null argument");
    if (!m->st_acc) {
        set_error("edm_stats_msd: edm_stats_open was not called");
        return EDM_ESTATE;
    }
    const int nb = sblocks(m->n);
    hipLaunchKernelGGL(k_st_msd, dim3(nb), dim3(kSB), 0, m->stream, m->n,
                       (const int32_t*)m->a.id, (const int8_t*)m->a.sp, (const uint8_t*)m->a.fl,
                       species_mask, (const double*)m->st_acc, m->st_part);
    hipLaunchKernelGGL(k_st_final, dim3(1), dim3(kSB), 0, m->stream, nb, m->st_part);
    double r[2];
    EDM_HIP(hipMemcpyAsync(r, m->st_part + 2 * (size_t)nb, sizeof(r), hipMemcpyDeviceToHost,
                           m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    *sum_r2 = r[0];
    *count = (int64_t)r[1];
    return 0;
}

int edm_stats_pair_hist(edm_membrane* m, uint32_t species_mask, int32_t n_bins, double r_max,
                        uint64_t* counts) {
    if (!m || !counts || n_bins <= 0 || n_bins > 4096 || !(r_max > 0.0))
        return sinvalid("edm_stats_pair_hist: bad argument (1..4096 bins, r_max > 0)");
    if (n_bins > m->st_hist_cap) {
        if (m->st_hist) EDM_HIP(hipFree(m->st_hist));
        m->st_hist = nullptr;
        EDM_HIP(hipMalloc(&m->st_hist, (size_t)n_bins * sizeof(unsigned long long)));
        m->st_hist_cap = n_bins;
    }
    EDM_HIP(hipMemsetAsync(m->st_hist, 0, (size_t)n_bins * 8, m->stream));
    const int rc = m->fp64 ? pairs_hist_t<double>(m, species_mask, n_bins, r_max)
                           : pairs_hist_t<float>(m, species_mask, n_bins, r_max);
    if (rc) return rc;
    EDM_HIP(hipMemcpyAsync(counts, m->st_hist, (size_t)n_bins * 8, hipMemcpyDeviceToHost,
                           m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int edm_stats_flim(edm_membrane* m, int32_t channel, int32_t harmonic, double* out) {
    if (!m || !out || !m->has_ip) return sinvalid("edm_stats_flim: imaging not set");
    const edm_image_params& ip = m->ip;
    if (channel < 0 || channel >= ip.n_ch || harmonic < 1)
        return sinvalid("edm_stats_flim: bad channel or harmonic");
    const int64_t npix = (int64_t)ip.nx * ip.ny;
    if ((size_t)npix * 4 > m->st_pix_cap) {
        if (m->st_pix) EDM_HIP(hipFree(m->st_pix));
        m->st_pix = nullptr;
        EDM_HIP(hipMalloc(&m->st_pix, (size_t)npix * 4 * sizeof(double)));
        m->st_pix_cap = (size_t)npix * 4;
    }
    const int64_t threads = npix * 64;
    const int blocks = (int)((threads + kSB - 1) / kSB);
    hipLaunchKernelGGL(k_st_flim, dim3(blocks), dim3(kSB), 2 * ip.n_bins * sizeof(double),
                       m->stream, (const uint32_t*)(m->flim + (size_t)channel * npix * ip.n_bins),
                       npix, ip.n_bins, harmonic, ip.bin_w * 1e9, m->st_pix);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipMemcpyAsync(out, m->st_pix, (size_t)npix * 4 * sizeof(double),
                           hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

}  // extern "C"
