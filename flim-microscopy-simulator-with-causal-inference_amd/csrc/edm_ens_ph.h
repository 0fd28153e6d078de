// Batched ensemble: photon emission + imaging and frame end (included at
// the end of edm_photons.hip, inside namespace edm).  The dye carriers of
// the step's work list emit with their replica's draws, image geometry and
// TCSPC stack.

template <typename Real, int kWin>
__global__ __launch_bounds__(256) void k_ens_photons(int n_waves, const int32_t* __restrict__ list,
                                                     const int32_t* __restrict__ wrep,
                                                     const uint32_t* __restrict__ rcnt,
                                                     const EnsSeg* __restrict__ segs,
                                                     const edm_photo_params* pp, const Real* x,
                                                     const Real* y, uint8_t* fl, const int8_t* dye,
                                                     int8_t* ctmc, const int32_t* __restrict__ id,
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
        E = emit_pss<Real>(pp, t, (uint32_t)(id[t] - g.kbase), fl, dye, ctmc, nullptr, nullptr,
                           true, 0.0, -1, rng, gstep - g.step0, emitted);
    }
    // the wave's emitters are all of replica W.r: its geometry, stack and draws
    wave_photons<Real, kWin>(E, &g.ip, x, y, g.flim, rng, gstep - g.step0, detected);
    photon_count(emitted, detected, ctr + 8 * W.r);
}

template <typename Real, int kWin>
static void launch_ens_photons(edm_ensemble* e, int nt) {
    hipLaunchKernelGGL((k_ens_photons<Real, kWin>), dim3(nb(nt)), dim3(256), 0, e->stream, nt / 64,
                       (const int32_t*)e->wl_dye, (const int32_t*)e->wl_wrep_dye,
                       (const uint32_t*)e->wl_rcnt, (const EnsSeg*)e->d_seg,
                       (const edm_photo_params*)e->d_pp, (const Real*)e->a.x,
                       (const Real*)e->a.y, e->a.fl, (const int8_t*)e->a.dye, e->a.ctmc,
                       (const int32_t*)e->a.id, e->ctr, e->gstep);
}

int ens_photons(edm_ensemble* e) {
    const int nt = (int)e->wl_ndye;
    if (nt == 0 || !e->has_pp) return 0;
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
                       g.flim, ip.dark_lambda, ip.p_ap, rng, e->gstep - g.step0, e->frame_out);
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
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(nb(g.flim_elems), 8192)), dim3(kB), 0, e->stream,
                       g.flim_elems, g.flim, 0u);
    EDM_HIP(hipGetLastError());
    return 0;
}

// Frame end of several replicas at once (blockIdx.y = position in the slot
// list): detector noise into each listed stack, counted per replica.
// Compiled for 5 waves per SIMD: latency bound, it runs faster at 5 than at
// the 3 its registers would allow, despite the spills (1051 against 1135 us
// for 512 stacks of 64 x 64 x 256).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_ens_frame_noise(const EnsSeg* __restrict__ segs,
                                                         const int32_t* __restrict__ slots,
                                                         unsigned long long* noise_ctr,
                                                         uint32_t gstep) {
    const int r = slots[blockIdx.y];
    const EnsSeg& g = segs[r];
    // the few replica fields the frame end needs, once
    uint32_t* const flim = g.flim;
    const int nb = g.ip.n_bins;
    const int64_t npix = (int64_t)g.ip.n_ch * g.ip.nx * g.ip.ny;
    const double lam = g.ip.dark_lambda, p_ap = g.ip.p_ap;
    const Rng rng{g.seed_lo ^ 0x2545F491u, g.seed_hi};
    const uint32_t frame = gstep - g.step0;
    if (!flim) return;  // block-uniform
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t end = (npix + 63) / 64 * 64;  // whole wavefronts take part in the counts
#pragma unroll 1
    for (int64_t q0 = (int64_t)blockIdx.x * blockDim.x; q0 < end; q0 += stride) {
        // this wavefront's 64 pixels: totals read coalesced, then one pixel per lane
        const int64_t qw = q0 + (threadIdx.x & ~63u);
        const int64_t q = q0 + threadIdx.x;
        unsigned long long nd = 0, na = 0;
        if (qw < npix) {
            const uint32_t tot0 = p_ap > 0.0 ? wave_pixel_totals(flim, qw, npix, nb) : 0u;
            if (q < npix) frame_noise_pixel(q, nb, flim, tot0, lam, p_ap, rng, frame, nd, na);
        }
        noise_count(nd, na, noise_ctr + 2 * r);
    }
}

__global__ void k_ens_frame_reset(const EnsSeg* __restrict__ segs,
                                  const int32_t* __restrict__ slots) {
    const EnsSeg& g = segs[slots[blockIdx.y]];
    if (!g.flim) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.flim_elems;
         i += (int64_t)gridDim.x * blockDim.x)
        g.flim[i] = 0u;
}

int ens_frame_end(edm_ensemble* e, const int32_t* slots, int32_t n) {
    if (n <= 0) return 0;
    if (e->layout_dirty)
        if (int rc = ens_layout(e)) return rc;
    if ((size_t)n > e->slot_list_cap) {
        hipFree(e->d_slot_list);
        e->slot_list_cap = (size_t)std::max(n, e->rmax);
        EDM_HIP(hipMalloc(&e->d_slot_list, e->slot_list_cap * sizeof(int32_t)));
    }
    EDM_HIP(hipMemcpyAsync(e->d_slot_list, slots, (size_t)n * sizeof(int32_t),
                           hipMemcpyHostToDevice, e->stream));
    hipLaunchKernelGGL(k_ens_frame_noise, dim3(32, n), dim3(kB), 0, e->stream,
                       (const EnsSeg*)e->d_seg, (const int32_t*)e->d_slot_list, e->noise_ctr,
                       e->gstep);
    hipLaunchKernelGGL(k_ens_frame_reset, dim3(64, n), dim3(kB), 0, e->stream,
                       (const EnsSeg*)e->d_seg, (const int32_t*)e->d_slot_list);
    EDM_HIP(hipGetLastError());
    // the slot list must stay valid until the kernels have read it
    EDM_HIP(hipStreamSynchronize(e->stream));
    return 0;
}
