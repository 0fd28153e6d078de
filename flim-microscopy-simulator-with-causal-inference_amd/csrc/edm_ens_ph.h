// Batched ensemble: photon emission + imaging and frame end (included at
// the end of edm_photons.hip, inside namespace edm).  The dye carriers of
// the step's work list emit with their replica's draws, image geometry and
// TCSPC stack.

// Proximity FRET in the batch: every replica's alive targets of the map
// binned into its own grid (cells of the pooled target-cell space from its
// tg_off), as the single membrane's fret_prox_grid does.
template <typename Real>
__global__ void k_ens_tg_count(int n, const Real* x, const Real* y, const int8_t* sp,
                               const uint8_t* fl, const uint16_t* __restrict__ rep,
                               const EnsSeg* __restrict__ segs, int target, int32_t* cnt) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    if (!(fl[u] & EDM_FLAG_ALIVE) || sp[u] != target) return;
    const EnsSeg& g = segs[rep[u]];
    if (!g.active) return;
    const TgGrid tg{g.tg_ncx, g.tg_ncy, g.tg_ix, g.tg_iy};
    atomicAdd(&cnt[g.tg_off + tg.cell((double)x[u], (double)y[u])], 1);
}

template <typename Real>
__global__ void k_ens_tg_scatter(int n, const Real* x, const Real* y, const int8_t* sp,
                                 const uint8_t* fl, const uint16_t* __restrict__ rep,
                                 const EnsSeg* __restrict__ segs, int target,
                                 const int32_t* start, int32_t* fill, Real* tx, Real* ty,
                                 int32_t* tu) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    if (!(fl[u] & EDM_FLAG_ALIVE) || sp[u] != target) return;
    const EnsSeg& g = segs[rep[u]];
    if (!g.active) return;
    const TgGrid tg{g.tg_ncx, g.tg_ncy, g.tg_ix, g.tg_iy};
    const Real xu = x[u], yu = y[u];
    const int c = g.tg_off + tg.cell((double)xu, (double)yu);
    const int k = start[c] + atomicAdd(&fill[c], 1);
    tx[k] = xu;
    ty[k] = yu;
    tu[k] = u;
}

// The proximity map's binned targets (nullptr: no proximity map).
template <typename Real>
struct EnsTargets {
    int map;
    const int32_t* start;
    const Real *x, *y;
    const int32_t* u;
};

template <typename Real, int kWin>
__global__ __launch_bounds__(256) void k_ens_photons(int n_waves, const int32_t* __restrict__ list,
                                                     const int32_t* __restrict__ wrep,
                                                     const uint32_t* __restrict__ rcnt,
                                                     const EnsSeg* __restrict__ segs,
                                                     const edm_photo_params* pp, const Real* x,
                                                     const Real* y, uint8_t* fl, const int8_t* dye,
                                                     int8_t* ctmc, const int32_t* __restrict__ id,
                                                     const int32_t* __restrict__ bond,
                                                     EnsTargets<Real> TG,
                                                     unsigned long long* ctr, uint32_t gstep) {
    if ((int)blockIdx.x * (int)(blockDim.x >> 6) +
            __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) >= n_waves)
        return;  // whole waves
    const EnsWave W = ens_wave(wrep, segs, 1);
    const EnsSeg& g = segs[W.r];
    const int j = W.j + lane_id();
    int emitted = 0, detected = 0;
    const Rng rng{g.seed_lo ^ 0x5bd1e995u, g.seed_hi};
    Emit<Real> E{0, 0, 0u, 0, 0, 0, 0, (Real)0, (Real)0};
    if (j < min(g.wl_n[1], (int)rcnt[2 * W.r + 1])) {
        const int t = list[g.wl_off[1] + j];
        // FRET rates as the single membrane takes them (update_fret_rates,
        // photophysics.py:1021-1114): bond maps from the bonds (binding
        // models; an intermolecular acceptor through the replica's
        // slot_of), then the proximity map from the replica's target grid
        double kf = 0.0;
        int aid = -1;
        if (bond) fret_inline(pp, t, fl, dye, bond, g.slot_of, kf, aid, g.kbase);
        if (TG.map >= 0) {
            const int q = TG.map;
            const uint8_t f = fl[t];
            if ((f & EDM_FLAG_DYE) && (f & EDM_FLAG_ALIVE) && !(f & EDM_FLAG_BLEACHED) &&
                dye[t] == pp->fm_donor[q]) {
                const TgGrid tg{g.tg_ncx, g.tg_ncy, g.tg_ix, g.tg_iy};
                const bool found = prox_search<Real>(tg, TG.start + g.tg_off, TG.x, TG.y, TG.u, t,
                                                     x[t], y[t], (Real)pp->fm_radius[q],
                                                     (Real)g.Lx, (Real)g.Ly, g.px != 0,
                                                     g.py != 0);
                kf = found ? pp->fm_k_hi[q] : pp->fm_k_lo[q];
                aid = -1;
            }
        }
        E = emit_pss<Real>(pp, t, (uint32_t)(id[t] - g.kbase), fl, dye, ctmc, bond, g.slot_of,
                           true, kf, aid, rng, gstep - g.step0, emitted, g.kbase);
    }
    // the wave's emitters are all of replica W.r: its geometry, stack (and
    // its row flags) and draws
    wave_photons<Real, kWin>(E, &g.ip, x, y, g.flim,
                             row_flags(g.flim, (int64_t)g.ip.n_ch * g.ip.nx * g.ip.ny,
                                       g.ip.n_bins),
                             rng, gstep - g.step0, detected);
    photon_count(emitted, detected, ctr + 8 * W.r);
}

template <typename Real, int kWin>
static void launch_ens_photons(edm_ensemble* e, int nt) {
    EnsTargets<Real> TG{e->prox_map, (const int32_t*)e->tg_start, (const Real*)e->tg_x,
                        (const Real*)e->tg_y, (const int32_t*)e->tg_u};
    hipLaunchKernelGGL((k_ens_photons<Real, kWin>), dim3(nb(nt)), dim3(256), 0, e->stream, nt / 64,
                       (const int32_t*)e->wl_dye, (const int32_t*)e->wl_wrep_dye,
                       (const uint32_t*)e->wl_rcnt, (const EnsSeg*)e->d_seg,
                       (const edm_photo_params*)e->d_pp, (const Real*)e->a.x,
                       (const Real*)e->a.y, e->a.fl, (const int8_t*)e->a.dye, e->a.ctmc,
                       (const int32_t*)e->a.id, (const int32_t*)e->a.bond, TG, e->ctr, e->gstep);
}

template <typename Real>
static int ens_targets(edm_ensemble* e) {
    const int n = (int)e->n_total;
    const int nc = (int)e->n_tcells;
    hipStream_t st = e->stream;
    const int target = e->pp.fm_target[e->prox_map];
    hipLaunchKernelGGL(k_tg_zero, dim3(nb(nc + 1)), dim3(kB), 0, st, nc, e->tg_cnt, e->tg_fill);
    hipLaunchKernelGGL(k_ens_tg_count<Real>, dim3(nb(std::max(n, 1))), dim3(kB), 0, st, n,
                       (const Real*)e->a.x, (const Real*)e->a.y, (const int8_t*)e->a.sp,
                       (const uint8_t*)e->a.fl, (const uint16_t*)e->rep, (const EnsSeg*)e->d_seg,
                       target, e->tg_cnt);
    size_t tb = e->cub_bytes_tg;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(e->cub_tmp_tg, tb, e->tg_cnt, e->tg_start, nc + 1,
                                             st));
    hipLaunchKernelGGL(k_ens_tg_scatter<Real>, dim3(nb(std::max(n, 1))), dim3(kB), 0, st, n,
                       (const Real*)e->a.x, (const Real*)e->a.y, (const int8_t*)e->a.sp,
                       (const uint8_t*)e->a.fl, (const uint16_t*)e->rep, (const EnsSeg*)e->d_seg,
                       target, (const int32_t*)e->tg_start, e->tg_fill, (Real*)e->tg_x,
                       (Real*)e->tg_y, e->tg_u);
    return 0;
}

int ens_photons(edm_ensemble* e) {
    const int nt = (int)e->wl_ndye;
    if (nt == 0 || !e->has_pp) return 0;
    if (e->prox_map >= 0)
        if (int rc = e->fp64 ? ens_targets<double>(e) : ens_targets<float>(e)) return rc;
    // the PSF window capacity of the launch covers every replica's window
    bool small = true;
    for (int r = 0; r < e->rmax; ++r)
        if (e->seg[r].active) small = small && psf_small(e->seg[r].ip.R);
    if (e->fp64) {
        if (small) launch_ens_photons<double, 8>(e, nt);
        else launch_ens_photons<double, 31>(e, nt);
    } else {
        if (small) launch_ens_photons<float, 8>(e, nt);
        else launch_ens_photons<float, 31>(e, nt);
    }
    EDM_HIP(hipGetLastError());
    return 0;
}

int ens_frame_noise(edm_ensemble* e, int slot, int64_t* n_dark, int64_t* n_ap) {
    const EnsSeg& g = e->seg[slot];
    const edm_image_params& ip = g.ip;
    const int64_t npix = (int64_t)ip.n_ch * ip.nx * ip.ny;
    Rng rng{g.seed_lo ^ 0x2545F491u, g.seed_hi};
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, e->stream, (int64_t)4,
                       (uint32_t*)e->frame_out, 0u);
    hipLaunchKernelGGL(k_frame_noise, dim3(nb(npix)), dim3(kB), 0, e->stream, npix, ip.n_bins,
                       g.flim, row_flags(g.flim, npix, ip.n_bins), ip.dark_lambda, ip.p_ap, rng,
                       e->gstep - g.step0, e->frame_out);
    EDM_HIP(hipGetLastError());
    if (n_dark || n_ap) {
        unsigned long long c[2];
        EDM_HIP(hipMemcpyAsync(c, e->frame_out, 16, hipMemcpyDeviceToHost, e->stream));
        EDM_HIP(hipStreamSynchronize(e->stream));
        if (n_dark) *n_dark = (int64_t)c[0];
        if (n_ap) *n_ap = (int64_t)c[1];
    }
    return 0;
}

int ens_frame_reset(edm_ensemble* e, int slot) {
    const EnsSeg& g = e->seg[slot];
    if (!g.flim) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(nb(g.flim_words), 8192)), dim3(kB), 0, e->stream,
                       g.flim_words, g.flim, 0u);
    EDM_HIP(hipGetLastError());
    return 0;
}

// Frame end of several replicas at once, as the single membrane's streamed
// frame end does it (k_fr_noise_count, k_fr_emit) on each listed stack:
// blockIdx.y = position q in the list, wavefront w of the block's 4 = the
// replica's pixel rows 64 w .. 64 w + 63.  The first pass sums the flagged
// rows, draws every pixel's detector noise (frame_noise_pixel: the draws
// of the dense frame end) and counts the wavefront's nonzero bins into
// wcnt[wbase[q] + w]; after an exclusive scan of those counts the second
// pass writes the flagged rows' nonzero bins as (flat index, count) pairs
// in index order, replica after replica in list order (oidx nullptr: no
// output), zeroes the bins it read and clears the flags, which resets the
// stack.  Only rows that hold counts are read or written.
struct EnsFrameList {
    const int32_t* slots;
    const int64_t* wbase;
};

__device__ __forceinline__ bool ens_fr_wave(const EnsSeg& g, int64_t& q0, int64_t& npix, int& nb,
                                            int64_t& w) {
    if (!g.flim) return false;
    npix = (int64_t)g.ip.n_ch * g.ip.nx * g.ip.ny;
    nb = g.ip.n_bins;
    w = (int64_t)blockIdx.x * 4 + (int64_t)(threadIdx.x >> 6);
    q0 = w << 6;
    return q0 < npix;
}

__global__ __launch_bounds__(256) void k_ens_fr_noise_count(const EnsSeg* __restrict__ segs,
                                                            EnsFrameList L,
                                                            unsigned long long* noise_ctr,
                                                            uint32_t* wcnt, uint32_t gstep) {
    const int r = L.slots[blockIdx.y];
    const EnsSeg& g = segs[r];
    int64_t q0, npix, w;
    int nb;
    if (!ens_fr_wave(g, q0, npix, nb, w)) return;  // whole wavefronts
    uint32_t* flim = g.flim;
    uint8_t* flags = row_flags(flim, npix, nb);
    const int lane = lane_id();
    const int64_t q = q0 + lane;
    const bool in = q < npix;
    uint32_t tot0 = 0u, nz = 0u;
    for (unsigned long long act = __ballot(in && flags[q] != 0u); act; act &= act - 1ull) {
        const int j = __builtin_ctzll(act);
        uint32_t s, c;
        wave_row_sum_nz(flim + (q0 + j) * nb, nb, s, c);
        if (lane == j) {
            tot0 = s;
            nz = c;
        }
    }
    const Rng rng{g.seed_lo ^ 0x2545F491u, g.seed_hi};
    unsigned long long nd = 0, na = 0;
    if (in)
        nz += frame_noise_pixel(q, nb, flim, flags, tot0, g.ip.dark_lambda, g.ip.p_ap, rng,
                                gstep - g.step0, nd, na);
    noise_count(nd, na, noise_ctr + 2 * r);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nz += __shfl_xor(nz, o);
    if (wcnt && lane == 0) wcnt[L.wbase[blockIdx.y] + w] = nz;
}

__global__ __launch_bounds__(256) void k_ens_fr_emit(const EnsSeg* __restrict__ segs,
                                                     EnsFrameList L,
                                                     const uint32_t* __restrict__ off,
                                                     uint32_t* oidx, uint32_t* ocnt) {
    const EnsSeg& g = segs[L.slots[blockIdx.y]];
    int64_t q0, npix, w;
    int nb;
    if (!ens_fr_wave(g, q0, npix, nb, w)) return;  // whole wavefronts
    uint32_t* flim = g.flim;
    uint8_t* flags = row_flags(flim, npix, nb);
    const int lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int64_t q = q0 + lane;
    const bool mine = q < npix && flags[q] != 0u;
    uint32_t base = oidx ? off[L.wbase[blockIdx.y] + w] : 0u;
    for (unsigned long long act = __ballot(mine); act; act &= act - 1ull) {
        const int j = __builtin_ctzll(act);
        uint32_t* row = flim + (q0 + j) * nb;
        const uint32_t r0 = (uint32_t)((q0 + j) * nb);
        for (int b0 = 0; b0 < nb; b0 += 64) {
            const int b = b0 + lane;
            const uint32_t v = b < nb ? row[b] : 0u;
            const unsigned long long mk = __ballot(v != 0u);
            if (v) {
                if (oidx) {
                    const uint32_t pos = base + (uint32_t)__popcll(mk & lt);
                    oidx[pos] = r0 + (uint32_t)b;
                    ocnt[pos] = v;
                }
                row[b] = 0u;
            }
            base += (uint32_t)__popcll(mk);
        }
    }
    if (mine) flags[q] = 0u;
}

// The listed replicas' first wavefront entries (wbase[n] = total
// wavefronts); returns the grid's x extent (4 wavefronts per block).  On
// the device the list is n slots, then the n + 1 offsets (8-byte aligned).
static int64_t ens_frame_list(edm_ensemble* e, const int32_t* slots, int n,
                              std::vector<int64_t>& wbase) {
    wbase.assign((size_t)n + 1, 0);
    int64_t most = 0;
    for (int q = 0; q < n; ++q) {
        const EnsSeg& g = e->seg[slots[q]];
        const int64_t npix =
            g.flim ? (int64_t)g.ip.n_ch * g.ip.nx * g.ip.ny : 0;
        const int64_t nw = (npix + 63) / 64;
        wbase[q + 1] = wbase[q] + nw;
        most = std::max(most, nw);
    }
    return (most + 3) / 4;
}

static size_t ens_list_bytes(int n) {
    return ((size_t)n * sizeof(int32_t) + 7) / 8 * 8 + ((size_t)n + 1) * sizeof(int64_t);
}

int ens_frame_end(edm_ensemble* e, const int32_t* slots, int32_t n) {
    if (n <= 0) return 0;
    if (e->layout_dirty)
        if (int rc = ens_layout(e)) return rc;
    const size_t bytes = ens_list_bytes(n);
    if (bytes > e->slot_list_cap) {
        hipFree(e->d_slot_list);
        e->slot_list_cap = std::max(bytes, ens_list_bytes(e->rmax));
        EDM_HIP(hipMalloc(&e->d_slot_list, e->slot_list_cap));
    }
    std::vector<int64_t> wbase;
    const int64_t gx = ens_frame_list(e, slots, n, wbase);
    std::vector<char> host(bytes);
    std::memcpy(host.data(), slots, (size_t)n * sizeof(int32_t));
    const size_t wo = ((size_t)n * sizeof(int32_t) + 7) / 8 * 8;
    std::memcpy(host.data() + wo, wbase.data(), wbase.size() * sizeof(int64_t));
    EDM_HIP(hipMemcpyAsync(e->d_slot_list, host.data(), bytes, hipMemcpyHostToDevice, e->stream));
    const EnsFrameList L{(const int32_t*)e->d_slot_list,
                         (const int64_t*)((const char*)e->d_slot_list + wo)};
    if (gx > 0) {
        hipLaunchKernelGGL(k_ens_fr_noise_count, dim3((unsigned)gx, n), dim3(kB), 0, e->stream,
                           (const EnsSeg*)e->d_seg, L, e->noise_ctr, (uint32_t*)nullptr,
                           e->gstep);
        hipLaunchKernelGGL(k_ens_fr_emit, dim3((unsigned)gx, n), dim3(kB), 0, e->stream,
                           (const EnsSeg*)e->d_seg, L, (const uint32_t*)nullptr,
                           (uint32_t*)nullptr, (uint32_t*)nullptr);
    }
    EDM_HIP(hipGetLastError());
    // the list must stay valid until the kernels have read it
    EDM_HIP(hipStreamSynchronize(e->stream));
    return 0;
}

// ---------------------------------------------------------------- streamed frames
// A collect ends the frames of the listed replicas and returns every frame,
// sparse: after the detector noise, each stack's nonzero bins are written as
// (flat index, count) pairs, in index order, replica after replica in list
// order, and the stacks are reset (k_ens_fr_noise_count, k_sp_scan over
// the wavefront counts, k_ens_fr_emit).

// exclusive scan of blk[0 .. n) in place, blk[n] = total (one block)
__global__ __launch_bounds__(1024) void k_sp_scan(uint32_t* blk, int n) {
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const uint32_t v = i < n ? blk[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int w = 1; w < 1024; w <<= 1) {
            const uint32_t t = (int)threadIdx.x >= w ? s[threadIdx.x - w] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) blk[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk[n] = carry;
}

// per listed replica: cumulative dark counts, afterpulses and detected photons
__global__ void k_sp_meta(const int32_t* __restrict__ slots, int n,
                          const unsigned long long* __restrict__ noise_ctr,
                          const unsigned long long* __restrict__ ctr, unsigned long long* meta) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int r = slots[q];
    meta[3 * q] = noise_ctr[2 * r];
    meta[3 * q + 1] = noise_ctr[2 * r + 1];
    meta[3 * q + 2] = ctr[8 * r + 7];
}

void ens_frames_free(edm_ensemble* e) {
    for (auto& b : e->fbuf) {
        void* dev[] = {b.d_slots, b.d_blk, b.d_idx, b.d_cnt, b.d_meta};
        for (void* p : dev)
            if (p) (void)hipFree(p);
        void* host[] = {b.h_blk, b.h_idx, b.h_cnt, b.h_meta};
        for (void* p : host)
            if (p) (void)hipHostFree(p);
        if (b.copied) (void)hipEventDestroy(b.copied);
        if (b.written) (void)hipEventDestroy(b.written);
        b = edm_ensemble::FrameBuf{};
    }
    if (e->copy_stream) (void)hipStreamDestroy(e->copy_stream);
    e->copy_stream = nullptr;
}

// Buffers of a collect: the list (n slots and n + 1 wavefront offsets), the
// per-replica metadata, and nw + 1 wavefront counts / offsets.
static int sp_grow_slots(edm_ensemble::FrameBuf& b, int64_t n, int64_t nw) {
    if (n > b.slot_cap) {
        const int64_t c = std::max<int64_t>(n, 64);
        if (b.d_slots) EDM_HIP(hipFree(b.d_slots));
        if (b.d_meta) EDM_HIP(hipFree(b.d_meta));
        if (b.h_meta) EDM_HIP(hipHostFree(b.h_meta));
        b.d_slots = nullptr;
        b.d_meta = nullptr;
        b.h_meta = nullptr;
        EDM_HIP(hipMalloc(&b.d_slots, ens_list_bytes((int)c)));
        EDM_HIP(hipMalloc(&b.d_meta, c * 3 * sizeof(unsigned long long)));
        EDM_HIP(hipHostMalloc((void**)&b.h_meta, c * 3 * sizeof(unsigned long long)));
        b.slot_cap = c;
    }
    if (nw + 1 > b.blk_cap) {
        const int64_t c = std::max<int64_t>(nw + 1, 1024);
        if (b.d_blk) EDM_HIP(hipFree(b.d_blk));
        if (b.h_blk) EDM_HIP(hipHostFree(b.h_blk));
        b.d_blk = nullptr;
        b.h_blk = nullptr;
        EDM_HIP(hipMalloc(&b.d_blk, c * sizeof(uint32_t)));
        EDM_HIP(hipHostMalloc((void**)&b.h_blk, c * sizeof(uint32_t)));
        b.blk_cap = c;
    }
    return 0;
}

static int sp_grow_entries(edm_ensemble::FrameBuf& b, int64_t n) {
    if (n <= b.cap) return 0;
    const int64_t c = std::max<int64_t>(n + n / 2, 1 << 16);
    if (b.d_idx) EDM_HIP(hipFree(b.d_idx));
    if (b.d_cnt) EDM_HIP(hipFree(b.d_cnt));
    if (b.h_idx) EDM_HIP(hipHostFree(b.h_idx));
    if (b.h_cnt) EDM_HIP(hipHostFree(b.h_cnt));
    b.d_idx = b.d_cnt = b.h_idx = b.h_cnt = nullptr;
    EDM_HIP(hipMalloc(&b.d_idx, c * sizeof(uint32_t)));
    EDM_HIP(hipMalloc(&b.d_cnt, c * sizeof(uint32_t)));
    EDM_HIP(hipHostMalloc((void**)&b.h_idx, c * sizeof(uint32_t)));
    EDM_HIP(hipHostMalloc((void**)&b.h_cnt, c * sizeof(uint32_t)));
    b.cap = c;
    return 0;
}

int ens_frames_collect(edm_ensemble* e, const int32_t* slots, int32_t n, int32_t* ticket) {
    edm_ensemble::FrameBuf& b = e->fbuf[e->next_ticket & 1];
    if (b.ticket >= 0) {
        set_error("edm_ens_frames_collect: two collects are pending; fetch the older one first");
        return EDM_ESTATE;
    }
    if (e->layout_dirty)
        if (int rc = ens_layout(e)) return rc;
    if (!e->copy_stream) EDM_HIP(hipStreamCreate(&e->copy_stream));
    if (!b.copied) EDM_HIP(hipEventCreateWithFlags(&b.copied, hipEventDisableTiming));
    if (!b.written) EDM_HIP(hipEventCreateWithFlags(&b.written, hipEventDisableTiming));
    std::vector<int64_t> wbase;
    const int64_t gx = ens_frame_list(e, slots, n, wbase);
    const int64_t nw = wbase[(size_t)n];
    if (nw >= (int64_t)INT32_MAX) {
        set_error("edm_ens_frames_collect: too many pixel rows in one collect");
        return EDM_EINVAL;
    }
    if (int rc = sp_grow_slots(b, n, nw)) return rc;
    hipStream_t st = e->stream;
    const size_t wo = ((size_t)n * sizeof(int32_t) + 7) / 8 * 8;
    // the host copy of the list outlives the transfer: it is freed after the
    // synchronisation below
    std::vector<char> host(ens_list_bytes(n));
    std::memcpy(host.data(), slots, (size_t)n * sizeof(int32_t));
    std::memcpy(host.data() + wo, wbase.data(), wbase.size() * sizeof(int64_t));
    EDM_HIP(hipMemcpyAsync(b.d_slots, host.data(), host.size(), hipMemcpyHostToDevice, st));
    const EnsFrameList L{(const int32_t*)b.d_slots, (const int64_t*)((const char*)b.d_slots + wo)};
    if (n > 0 && gx > 0)
        hipLaunchKernelGGL(k_ens_fr_noise_count, dim3((unsigned)gx, n), dim3(kB), 0, st,
                           (const EnsSeg*)e->d_seg, L, e->noise_ctr, b.d_blk, e->gstep);
    hipLaunchKernelGGL(k_sp_scan, dim3(1), dim3(1024), 0, st, b.d_blk, (int)nw);
    EDM_HIP(hipMemcpyAsync(b.h_blk, b.d_blk, ((size_t)nw + 1) * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, st));
    // the offsets size the entry buffers: wait for the noise and the counts
    EDM_HIP(hipStreamSynchronize(st));
    const int64_t total = b.h_blk[nw];
    if (int rc = sp_grow_entries(b, total)) return rc;
    if (n > 0) {
        if (gx > 0)
            hipLaunchKernelGGL(k_ens_fr_emit, dim3((unsigned)gx, n), dim3(kB), 0, st,
                               (const EnsSeg*)e->d_seg, L, (const uint32_t*)b.d_blk, b.d_idx,
                               b.d_cnt);
        hipLaunchKernelGGL(k_sp_meta, dim3(nb(n)), dim3(kB), 0, st, (const int32_t*)b.d_slots, n,
                           (const unsigned long long*)e->noise_ctr,
                           (const unsigned long long*)e->ctr, b.d_meta);
    }
    EDM_HIP(hipGetLastError());
    b.wbase = wbase;
    // the copy stream moves the entries once the compute stream has written
    // them; the compute stream goes on with the next steps meanwhile
    EDM_HIP(hipEventRecord(b.written, st));
    EDM_HIP(hipStreamWaitEvent(e->copy_stream, b.written, 0));
    if (total > 0) {
        EDM_HIP(hipMemcpyAsync(b.h_idx, b.d_idx, (size_t)total * 4, hipMemcpyDeviceToHost,
                               e->copy_stream));
        EDM_HIP(hipMemcpyAsync(b.h_cnt, b.d_cnt, (size_t)total * 4, hipMemcpyDeviceToHost,
                               e->copy_stream));
    }
    if (n > 0)
        EDM_HIP(hipMemcpyAsync(b.h_meta, b.d_meta, (size_t)n * 3 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, e->copy_stream));
    EDM_HIP(hipEventRecord(b.copied, e->copy_stream));
    b.slots.assign(slots, slots + n);
    b.ticket = e->next_ticket++;
    *ticket = b.ticket;
    return 0;
}

int ens_frames_fetch(edm_ensemble* e, int32_t ticket, int64_t* counts, int64_t* meta,
                     uint32_t* idx, uint32_t* cnt, int64_t cap, int64_t* total) {
    edm_ensemble::FrameBuf& b = e->fbuf[ticket & 1];
    if (ticket < 0 || b.ticket != ticket) {
        set_error("edm_ens_frames_fetch: no pending collect with that ticket");
        return EDM_ESTATE;
    }
    EDM_HIP(hipEventSynchronize(b.copied));
    const int n = (int)b.slots.size();
    const int64_t tot = b.h_blk[b.wbase[(size_t)n]];
    if (total) *total = tot;
    if (idx || cnt) {
        if (cap < tot) {
            set_error("edm_ens_frames_fetch: output holds " + std::to_string(cap) +
                      " entries, the frames have " + std::to_string(tot));
            return EDM_EINVAL;
        }
        if (idx) std::memcpy(idx, b.h_idx, (size_t)tot * 4);
        if (cnt) std::memcpy(cnt, b.h_cnt, (size_t)tot * 4);
    }
    for (int q = 0; q < n; ++q) {
        if (counts)
            counts[q] = (int64_t)b.h_blk[b.wbase[(size_t)q + 1]] -
                        (int64_t)b.h_blk[b.wbase[(size_t)q]];
        if (meta)
            for (int k = 0; k < 3; ++k) meta[3 * q + k] = (int64_t)b.h_meta[3 * q + k];
    }
    // a call without output arrays only reports the sizes; one with them
    // releases the buffer for a later collect
    if (idx || cnt) b.ticket = -1;
    return 0;
}
