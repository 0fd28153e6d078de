// Census of a membrane on gfx950: every count the reference's
// SimulationPipeline takes with numpy after a BD step (pipeline.py:567-642:
// complexes, bleached dyes, donors, alive particles per species, the mean
// proximity-mode FRET efficiency) and at a frame end (_extract_ground_truth,
// pipeline.py:357-455), as ONE pass over the slot arrays instead of
// (1 + n_species) numpy reductions, written into a device ring of rows so
// that a stretch of steps records with no host synchronisation.  The host
// turns rows into the reference's SubFrameData / GroundTruth.
//
// Layout of the reduction: k_census runs 1024-thread blocks; each lane reads
// 4 consecutive slots with one 32-bit load of each byte array (flags,
// species, dye type) and one 128-bit load of the bond partners.  Fixed
// counts are per-lane registers reduced by wave shuffles; species counts are
// wave ballots.  Each block writes its partial row; k_census_final (one
// block, next in stream order, so no device-scope fences are needed) sums
// the partials in block order and writes the census row, so the double FRET
// sum is deterministic run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

#include "edm_internal.h"

namespace edm {

static constexpr int kCenThreads = 1024;
static constexpr int kCenWaves = kCenThreads / 64;
static constexpr int kCenFixed = 12;               // n_bond_alive .. n_free_dye[7]
static constexpr int kCenCols = kCenFixed + 32;    // + species_alive[32]
static constexpr int kCenMaxBlocks = 256;
static constexpr int kCenRows = 16;                // partial rows summed per column in parallel

static_assert(offsetof(edm_census, species_alive) - offsetof(edm_census, n_bond_alive) ==
                  kCenFixed * sizeof(int64_t),
              "census counts must be consecutive");

struct CensusArgs {
    int n, n_species;
    const uint8_t* fl;
    const int8_t* sp;
    const int32_t* bond;
    const int8_t* dye;
    // FRET sum over the type-0 donors (k_fret == nullptr: none)
    int n_dye;
    const int32_t* dye_ids;
    const int32_t* slot_of;
    const double* k_fret;
    double tau0;
    unsigned long long* part;  // [gridDim.x][kCenCols]
    double* part_e;            // [gridDim.x]
};

struct CensusFinal {
    int nblk;  // blocks of the k_census launch
    const unsigned long long* part;
    const double* part_e;
    const unsigned long long* ctr;
    edm_census* row;
    int64_t step;
    bool fret_sum;
};

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Four consecutive slots from quad g (slots past n read as flags 0 and no
// partner: counted nowhere).
struct Quad {
    uint32_t f4, s4, d4;
    int b[4];
};

__device__ __forceinline__ Quad load_quad(const CensusArgs& A, int64_t g) {
    Quad Q{0u, 0u, 0u, {-1, -1, -1, -1}};
    const int64_t i0 = g * 4;
    if (i0 + 3 < A.n) {
        Q.f4 = *(const uint32_t*)(A.fl + i0);
        Q.s4 = *(const uint32_t*)(A.sp + i0);
        Q.d4 = *(const uint32_t*)(A.dye + i0);
        const int4 bb = *(const int4*)(A.bond + i0);
        Q.b[0] = bb.x;
        Q.b[1] = bb.y;
        Q.b[2] = bb.z;
        Q.b[3] = bb.w;
    } else if (i0 < A.n) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < A.n) {
                Q.f4 |= (uint32_t)A.fl[i0 + k] << (8 * k);
                Q.s4 |= (uint32_t)(uint8_t)A.sp[i0 + k] << (8 * k);
                Q.d4 |= (uint32_t)(uint8_t)A.dye[i0 + k] << (8 * k);
                Q.b[k] = A.bond[i0 + k];
            }
        }
    }
    return Q;
}

// Counts of one quad: fixed columns into the lane's registers, alive per
// species by wave ballot into lane q's register (every lane must call it).
__device__ __forceinline__ void count_quad(const Quad& Q, int n_species, int lane,
                                           unsigned (&c)[kCenFixed], unsigned& my_sc) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const unsigned f = (Q.f4 >> (8 * k)) & 0xffu;
        const int s = (int)(int8_t)((Q.s4 >> (8 * k)) & 0xffu);
        const int d = (int)(int8_t)((Q.d4 >> (8 * k)) & 0xffu);
        const bool alive = f & EDM_FLAG_ALIVE, dyed = f & EDM_FLAG_DYE, bl = f & EDM_FLAG_BLEACHED;
        const bool bound = Q.b[k] >= 0;
        c[0] += alive && bound;
        c[1] += bound;
        c[2] += bl;
        c[3] += dyed && alive && !bl && d == 0;
        const bool free_dye = dyed && !bl && !bound;
#pragma unroll
        for (int t = 0; t < 8; ++t) c[4 + t] += free_dye && d == t;
        for (int q = 0; q < n_species; ++q) {
            const unsigned cnt = (unsigned)__popcll(__ballot(alive && s == q));
            my_sc += lane == q ? cnt : 0u;
        }
    }
}

__global__ __launch_bounds__(kCenThreads) void k_census(CensusArgs A) {
    __shared__ unsigned s_w[kCenWaves][kCenCols];
    __shared__ double s_e[kCenWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;

    unsigned c[kCenFixed];
#pragma unroll
    for (int k = 0; k < kCenFixed; ++k) c[k] = 0u;
    unsigned my_sc = 0u;  // lane q < n_species: alive particles of species q
    const int64_t quads = ((int64_t)A.n + 3) / 4;
    const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
    // two quads per lane per pass, both loaded before either is counted;
    // block-uniform loop bound: every lane reaches the ballots
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < quads; base += 2 * gstride) {
        const Quad Q0 = load_quad(A, base + threadIdx.x);
        const Quad Q1 = load_quad(A, base + gstride + threadIdx.x);
        count_quad(Q0, A.n_species, lane, c, my_sc);
        count_quad(Q1, A.n_species, lane, c, my_sc);
    }
#pragma unroll
    for (int k = 0; k < kCenFixed; ++k) {
        const unsigned v = wave_sum_u32(c[k]);
        if (lane == 0) s_w[w][k] = v;
    }
    if (lane < 32) s_w[w][kCenFixed + lane] = my_sc;

    // proximity-mode FRET efficiency of the type-0 donors (pipeline.py:594-605)
    double e = 0.0;
    if (A.k_fret) {
        for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.n_dye; q += gstride) {
            const int t = A.slot_of[A.dye_ids[q]];
            const unsigned f = A.fl[t];
            if ((f & EDM_FLAG_DYE) && (f & EDM_FLAG_ALIVE) && !(f & EDM_FLAG_BLEACHED) &&
                A.dye[t] == 0) {
                const double kt = A.k_fret[q] * A.tau0;
                e += kt / (1.0 + kt);
            }
        }
    }
    e = wave_sum_f64(e);
    if (lane == 0) s_e[w] = e;
    __syncthreads();

    if (threadIdx.x < kCenCols) {
        unsigned long long v = 0ull;
        for (int ww = 0; ww < kCenWaves; ++ww) v += s_w[ww][threadIdx.x];
        A.part[(size_t)blockIdx.x * kCenCols + threadIdx.x] = v;
    }
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int ww = 0; ww < kCenWaves; ++ww) v += s_e[ww];
        A.part_e[blockIdx.x] = v;
    }
}

// Sum of the block partials in block order, the event and photon counters,
// into one census row.
__global__ __launch_bounds__(kCenThreads) void k_census_final(CensusFinal F) {
    __shared__ unsigned long long s_fin[kCenRows][kCenCols];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x < kCenRows * kCenCols) {
        const int col = threadIdx.x % kCenCols, r = threadIdx.x / kCenCols;
        unsigned long long v = 0ull;
#pragma unroll 8
        for (int bk = r; bk < F.nblk; bk += kCenRows) v += F.part[(size_t)bk * kCenCols + col];
        s_fin[r][col] = v;
    }
    double ev = 0.0;
    if (w == 0) {
        for (int bk = lane; bk < F.nblk; bk += 64) ev += F.part_e[bk];
        ev = wave_sum_f64(ev);
    }
    __syncthreads();
    int64_t* counts = &F.row->n_bond_alive;
    if (threadIdx.x < kCenCols) {
        unsigned long long v = 0ull;
        for (int r = 0; r < kCenRows; ++r) v += s_fin[r][threadIdx.x];
        counts[threadIdx.x] = (int64_t)v;
    }
    if (threadIdx.x == 64) {
        F.row->step = F.step;
        for (int k = 0; k < 6; ++k) F.row->events[k] = (int64_t)F.ctr[k];
        F.row->photons = (int64_t)F.ctr[6];
    }
    if (threadIdx.x == 0) F.row->e_sum = F.fret_sum ? ev : 0.0;
}

// Frame-end FRET ground truth of the dye carriers (pipeline.py:384-411).
struct GtMaps {
    int n;
    int donor[8];
    double tau[8];
};

__global__ void k_ground_truth(int n_dye, const int32_t* dye_ids, const int32_t* slot_of,
                               const uint8_t* fl, const int8_t* dye, const double* k_fret,
                               GtMaps M, double* fret_e, double* tau_da) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int t = slot_of[dye_ids[q]];
    const unsigned f = fl[t];
    if (!(f & EDM_FLAG_DYE) || (f & EDM_FLAG_BLEACHED)) return;
    const int d = dye[t];
    const double k = k_fret[q];
    bool hit = false;
    double e = 0.0, ta = 0.0;
    for (int j = 0; j < M.n; ++j) {
        if (d != M.donor[j]) continue;
        const double tD = M.tau[j];
        if (k > 0.0) {
            const double kt = k * tD;
            e = kt / (1.0 + kt);
            ta = tD / (1.0 + kt);
        } else {
            e = 0.0;
            ta = tD;
        }
        hit = true;
    }
    if (hit) {
        fret_e[dye_ids[q]] = e;
        tau_da[dye_ids[q]] = ta;
    }
}

static int cen_scratch(edm_membrane* m) {
    if (m->cen_part) return 0;
    EDM_HIP(hipMalloc((void**)&m->cen_part, sizeof(unsigned long long) * kCenMaxBlocks * kCenCols));
    EDM_HIP(hipMalloc((void**)&m->cen_part_e, sizeof(double) * kCenMaxBlocks));
    EDM_HIP(hipMalloc((void**)&m->cen_row, sizeof(edm_census)));
    return 0;
}

// One census of the current state into `row` (device, non-null), FRET sum
// when `fret_sum`.  The census scratch must be allocated (cen_scratch).
static int census(edm_membrane* m, edm_census* row, bool fret_sum) {
    if (!m->cen_part || !row) {
        set_error("census: scratch or row not allocated");
        return EDM_ESTATE;
    }
    CensusArgs A{};
    A.n = m->n;
    A.n_species = m->p.n_species;
    A.fl = m->a.fl;
    A.sp = m->a.sp;
    A.bond = m->a.bond;
    A.dye = m->a.dye;
    A.n_dye = m->n_dye;
    A.dye_ids = m->dye_ids;
    A.slot_of = m->a.slot_of;
    A.k_fret = nullptr;
    if (fret_sum && m->has_pp && m->n_dye > 0) {
        if (int rc = dyn_ensure_slot_of(m)) return rc;
        A.k_fret = m->k_fret;
        A.tau0 = m->pp.tau[0];
    }
    A.part = m->cen_part;
    A.part_e = m->cen_part_e;
    const int64_t quads = ((int64_t)m->n + 3) / 4;
    const int64_t want = std::max<int64_t>((quads + 2 * kCenThreads - 1) / (2 * kCenThreads),
                                           ((int64_t)m->n_dye + kCenThreads - 1) / kCenThreads);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(kCenMaxBlocks, want));
    hipLaunchKernelGGL(k_census, dim3(grid), dim3(kCenThreads), 0, m->stream, A);
    CensusFinal F{grid, m->cen_part, m->cen_part_e, m->d_ctr, row, (int64_t)m->step,
                  A.k_fret != nullptr};
    hipLaunchKernelGGL(k_census_final, dim3(1), dim3(kCenThreads), 0, m->stream, F);
    EDM_HIP(hipGetLastError());
    return 0;
}

int rec_step(edm_membrane* m) {
    if (!m->rec) {
        set_error("edm_record_step: recording is not open");
        return EDM_ESTATE;
    }
    if (m->rec_n >= m->rec_cap) {
        set_error("edm_record_step: recording ring full (read it out)");
        return EDM_ECAPACITY;
    }
    if (int rc = census(m, m->rec + m->rec_n, (m->rec_flags & EDM_REC_FRET_SUM) != 0)) return rc;
    ++m->rec_n;
    return 0;
}

void rec_free(edm_membrane* m) {
    void* bufs[] = {m->rec, m->cen_part, m->cen_part_e, m->cen_row, m->gt_e, m->gt_tau};
    for (void* b : bufs)
        if (b) hipFree(b);
    m->rec = nullptr;
    m->cen_part = nullptr;
    m->cen_part_e = nullptr;
    m->cen_row = nullptr;
    m->gt_e = m->gt_tau = nullptr;
    m->rec_cap = m->rec_n = m->rec_flags = 0;
}

}  // namespace edm

using namespace edm;

static int invalid(const std::string& msg) {
    set_error(msg);
    return EDM_EINVAL;
}

extern "C" {

int edm_run_frame(edm_membrane* m, int32_t n_steps) {
    if (!m || n_steps < 0) return invalid("edm_run_frame: bad argument");
    if (m->begun) {
        set_error("edm_run_frame: a parity step is open");
        return EDM_ESTATE;
    }
    return steps_run(m, n_steps, m->has_pp && m->has_ip);
}

int edm_record_open(edm_membrane* m, int32_t capacity, int32_t flags) {
    if (!m || capacity <= 0 || capacity > (1 << 24))
        return invalid("edm_record_open: capacity must be 1 .. 2^24 rows");
    if (flags & ~(EDM_REC_FRET_SUM | EDM_REC_AUTO)) return invalid("edm_record_open: unknown flag");
    graph_touch(m);
    if ((flags & EDM_REC_FRET_SUM) && !m->has_pp) {
        set_error("edm_record_open: the FRET sum needs the photophysics tables");
        return EDM_ESTATE;
    }
    if (m->rec && capacity != m->rec_cap) {
        EDM_HIP(hipStreamSynchronize(m->stream));
        EDM_HIP(hipFree(m->rec));
        m->rec = nullptr;
    }
    if (!m->rec) EDM_HIP(hipMalloc((void**)&m->rec, sizeof(edm_census) * (size_t)capacity));
    if (int rc = cen_scratch(m)) return rc;
    m->rec_cap = capacity;
    m->rec_n = 0;
    m->rec_flags = flags;
    return 0;
}

int edm_record_step(edm_membrane* m) {
    if (!m) return invalid("edm_record_step: null handle");
    return rec_step(m);
}

int edm_record_rows(edm_membrane* m, int32_t* n_rows) {
    if (!m || !n_rows) return invalid("edm_record_rows: null argument");
    *n_rows = m->rec_n;
    return 0;
}

int edm_record_read(edm_membrane* m, edm_census* rows, int32_t cap, int32_t* n_rows) {
    if (!m || !n_rows || (cap > 0 && !rows)) return invalid("edm_record_read: null argument");
    if (!m->rec) {
        set_error("edm_record_read: recording is not open");
        return EDM_ESTATE;
    }
    if (cap < m->rec_n) {
        set_error("edm_record_read: " + std::to_string(m->rec_n) + " rows pending, room for " +
                  std::to_string(cap));
        return EDM_ECAPACITY;
    }
    EDM_HIP(hipStreamSynchronize(m->stream));
    if (m->rec_n > 0)
        EDM_HIP(hipMemcpy(rows, m->rec, sizeof(edm_census) * (size_t)m->rec_n,
                          hipMemcpyDeviceToHost));
    *n_rows = m->rec_n;
    m->rec_n = 0;
    return 0;
}

int edm_record_close(edm_membrane* m) {
    if (!m) return invalid("edm_record_close: null handle");
    graph_touch(m);
    if (m->rec) {
        EDM_HIP(hipStreamSynchronize(m->stream));
        EDM_HIP(hipFree(m->rec));
    }
    m->rec = nullptr;
    m->rec_cap = m->rec_n = m->rec_flags = 0;
    return 0;
}

int edm_ground_truth(edm_membrane* m, int32_t n_maps, const int32_t* donor, const double* tau_D,
                     double* fret_E, double* tau_DA, edm_census* out) {
    if (!m || n_maps < 0 || n_maps > 8 || (n_maps > 0 && (!donor || !tau_D)))
        return invalid("edm_ground_truth: bad argument");
    if (m->begun) {
        set_error("edm_ground_truth: a parity step is open");
        return EDM_ESTATE;
    }
    GtMaps M{};
    M.n = n_maps;
    for (int j = 0; j < n_maps; ++j) {
        M.donor[j] = donor[j];
        M.tau[j] = tau_D[j];
    }
    const int n = m->n;
    hipStream_t st = m->stream;
    const bool fret = (fret_E || tau_DA) && n > 0;
    if (fret) {
        if (!m->gt_e) {
            EDM_HIP(hipMalloc((void**)&m->gt_e, sizeof(double) * (size_t)n));
            EDM_HIP(hipMalloc((void**)&m->gt_tau, sizeof(double) * (size_t)n));
        }
        EDM_HIP(hipMemsetAsync(m->gt_e, 0, sizeof(double) * (size_t)n, st));
        EDM_HIP(hipMemsetAsync(m->gt_tau, 0, sizeof(double) * (size_t)n, st));
        if (m->has_pp && m->n_dye > 0 && n_maps > 0) {
            // update_fret_rates (pipeline.py:649), then the per-donor values
            if (int rc = ph_fret(m)) return rc;
            const int blocks = (m->n_dye + 255) / 256;
            hipLaunchKernelGGL(k_ground_truth, dim3(blocks), dim3(256), 0, st, m->n_dye,
                               (const int32_t*)m->dye_ids, (const int32_t*)m->a.slot_of, m->a.fl,
                               (const int8_t*)m->a.dye, (const double*)m->k_fret, M, m->gt_e,
                               m->gt_tau);
            EDM_HIP(hipGetLastError());
        }
    }
    if (out) {
        // the one-row result lives in the census scratch: allocate it before
        // taking its address
        if (int rc = cen_scratch(m)) return rc;
        if (int rc = census(m, m->cen_row, false)) return rc;
    }
    EDM_HIP(hipStreamSynchronize(st));
    if (fret_E && n > 0)
        EDM_HIP(hipMemcpy(fret_E, m->gt_e, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    if (tau_DA && n > 0)
        EDM_HIP(hipMemcpy(tau_DA, m->gt_tau, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    if (out) EDM_HIP(hipMemcpy(out, m->cen_row, sizeof(edm_census), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
