// C-ABI entry points of the batched ensemble (edm_ens_* in edm_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "edm_ensemble.h"

using namespace edm;

static int ens_invalid(const std::string& msg) {
    set_error(msg);
    return EDM_EINVAL;
}

static void ens_free(edm_ensemble* e) {
    void* bufs[] = {e->a.x, e->a.y, e->a.xn, e->a.yn, e->a.xt, e->a.yt, e->a.sp, e->a.sp_t,
                    e->a.fl, e->a.fl_t, e->a.id, e->a.id_t, e->a.ctmc, e->a.ctmc_t, e->a.dye,
                    e->a.dye_t, e->rep, e->rep_t, e->key, e->rank, e->perm, e->cell_cnt,
                    e->cell_start, e->cub_tmp, e->counts, e->delta, e->cand_id, e->cand_slot,
                    e->cand_x, e->cand_y, e->n_cand, e->ctr, e->frame_out, e->sig, e->eps, e->kh,
                    e->Dbase, e->d_pp, e->d_seg, e->d_tiles, e->noise_ctr, e->d_slot_list,
                    e->d_cats, e->d_pgrp, e->wl_pass, e->wl_dye, e->wl_rcnt,
                    e->wl_wrep_pass, e->wl_wrep_dye, e->d_wl_chunks, e->a.bond, e->a.bond_t,
                    e->ekey, e->ea, e->eb, e->ne, e->best, e->pcnt, e->poff, e->tconv,
                    e->cub_tmp64, e->tg_cnt, e->tg_fill, e->tg_start, e->tg_u, e->tg_x,
                    e->tg_y, e->cub_tmp_tg};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (auto& g : e->seg) {
        if (g.flim) hipFree(g.flim);
        if (g.rcdf) hipFree(g.rcdf);
        if (g.reg) hipFree(g.reg);
        if (g.w_adapt) hipFree(g.w_adapt);
        if (g.slot_of) hipFree(g.slot_of);
    }
    ens_frames_free(e);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

extern "C" {

int edm_ens_create(const edm_params* p, const edm_photo_params* pp, const edm_image_params* ip,
                   int64_t capacity, int32_t max_replicas, edm_ensemble** out) {
    if (!p || !out) return ens_invalid("edm_ens_create: null argument");
    *out = nullptr;
    if (capacity <= 0 || capacity >= INT32_MAX / 2)
        return ens_invalid("edm_ens_create: particle capacity out of range");
    if (max_replicas <= 0 || max_replicas > 65535)
        return ens_invalid("edm_ens_create: 1..65535 replicas");
    if (p->n_species <= 0 || p->n_species > kEnsMaxSpecies)
        return ens_invalid("edm_ens_create: 1..32 species supported");
    if (p->n_grids < 0 || p->n_grids > kEnsMaxGrids || (p->n_grids > 0 && !p->grids))
        return ens_invalid("edm_ens_create: at most 4 density grids");
    if (p->n_reactions < 0 || p->n_reactions > kEnsMaxRx || (p->n_reactions > 0 && !p->reactions))
        return ens_invalid("edm_ens_create: at most 16 reactions");
    if (!p->sigma || !p->epsilon || !p->k_harm || !p->D_base)
        return ens_invalid("edm_ens_create: missing parameter arrays");
    if (p->adapt_grid >= p->n_grids ||
        (p->adapt_grid >= 0 && (p->adapt_species < 0 || p->adapt_species >= p->n_species)))
        return ens_invalid("edm_ens_create: bad adaptation grid or species");
    for (int i = 0; i < p->n_reactions; ++i) {
        const edm_reaction& r = p->reactions[i];
        if (r.kind == EDM_RX_BIND && (r.a < 0 || r.b < 0 || r.c < 0 || r.a >= p->n_species ||
                                      r.b >= p->n_species || r.c >= p->n_species))
            return ens_invalid("edm_ens_create: binding refers to a missing species");
        if (r.kind == EDM_RX_ADAPT && p->adapt_grid < 0)
            return ens_invalid("edm_ens_create: adaptation step without an adaptation grid");
        if (r.kind == EDM_RX_RECRUIT && r.use_adapt &&
            (r.fb1.sensor < 0 || r.fb2.sensor >= 0 || !(r.contact >= 0.0)))
            return ens_invalid("edm_ens_create: local recruitment needs one feedback rule and "
                               "sigma >= 0");
        if (r.kind == EDM_RX_CAT && (r.a == r.b || r.b == r.c))
            return ens_invalid("edm_ens_create: catalysis needs enzyme and product distinct "
                               "from the substrate");
        if (r.kind < EDM_RX_BIND || r.kind > EDM_RX_ADAPT)
            return ens_invalid("edm_ens_create: unknown reaction kind");
        const edm_feedback* fbs[2] = {&r.fb1, &r.fb2};
        for (auto f : fbs)
            if (f->sensor >= 0 && (f->grid < 0 || f->grid >= p->n_grids || f->sensor >= p->n_species))
                return ens_invalid("edm_ens_create: feedback refers to a missing grid or species");
    }
    int prox = -1;
    if (pp) {
        for (int j = 0; j < pp->n_fret_maps; ++j)
            if (!pp->fm_skip[j] && pp->fm_proximity[j]) {
                if (prox >= 0)
                    return ens_invalid("edm_ens_create: at most one proximity FRET map is "
                                       "batched");
                if (!(pp->fm_radius[j] > 0.0) || pp->fm_target[j] < 0 ||
                    pp->fm_target[j] >= p->n_species)
                    return ens_invalid("edm_ens_create: bad proximity FRET map");
                prox = j;
            }
    }
    edm_ensemble* e = new edm_ensemble();
    e->p = *p;
    e->fp64 = p->fp64 != 0;
    e->rs = e->fp64 ? 8 : 4;
    e->cap = capacity;
    e->rmax = max_replicas;
    e->grids.assign(p->grids, p->grids + p->n_grids);
    e->rx.assign(p->reactions, p->reactions + p->n_reactions);
    hipError_t he = hipGetDevice(&e->device);
    if (he == hipSuccess) he = hipStreamCreate(&e->stream);
    if (he != hipSuccess) {
        delete e;
        return hip_fail(he, "edm_ens_create: stream");
    }
    int rc = ens_create(e);
    e->p.grids = nullptr;
    e->p.reactions = nullptr;
    e->p.sigma = e->p.epsilon = e->p.k_harm = e->p.D_base = nullptr;
    e->p.rcx = e->p.rcy = e->p.rr = e->p.rf = nullptr;
    if (!rc && pp && prox >= 0) {
        e->prox_map = prox;
        const size_t rb = e->rs;
        if (hipMalloc(&e->tg_x, (size_t)capacity * rb + 16) != hipSuccess ||
            hipMalloc(&e->tg_y, (size_t)capacity * rb + 16) != hipSuccess ||
            hipMalloc(&e->tg_u, (size_t)capacity * sizeof(int32_t) + 16) != hipSuccess)
            rc = hip_fail(hipErrorOutOfMemory, "edm_ens_create: proximity FRET buffers");
    }
    if (!rc && pp) {
        e->pp = *pp;
        e->has_pp = true;
        he = hipMalloc(&e->d_pp, sizeof(edm_photo_params));
        if (he == hipSuccess)
            he = hipMemcpy(e->d_pp, pp, sizeof(edm_photo_params), hipMemcpyHostToDevice);
        if (he != hipSuccess) rc = hip_fail(he, "edm_ens_create: photophysics");
    }
    if (!rc && ip) {
        if (ip->n_bins <= 0 || ip->n_ch <= 0 || ip->n_sigma <= 0 || ip->n_sigma > 8)
            rc = ens_invalid("edm_ens_create: bad imaging constants");
        e->ip = *ip;
    }
    if (rc) {
        ens_free(e);
        return rc;
    }
    *out = e;
    return 0;
}

int edm_ens_destroy(edm_ensemble* e) {
    if (!e) return 0;
    hipStreamSynchronize(e->stream);
    ens_free(e);
    return 0;
}

int edm_ens_add(edm_ensemble* e, const edm_replica* r, const edm_host_state* s, int32_t* slot) {
    if (!e || !r || !s || !slot || !s->positions || !s->species_id || !s->is_alive)
        return ens_invalid("edm_ens_add: null argument");
    *slot = -1;
    if (r->n <= 0 || r->ncx < 3 || r->ncy < 3 || !(r->Lx > 0.0) || !(r->Ly > 0.0))
        return ens_invalid("edm_ens_add: bad box or cell grid");
    if (r->n_grids != (int)e->grids.size() || r->n_reactions != (int)e->rx.size())
        return ens_invalid("edm_ens_add: the replica does not match the model's grids/reactions");
    if (r->n_regions < 0 || r->n_regions > kEnsMaxRegions ||
        (r->n_regions > 0 && e->p.dfield != 0))
        return ens_invalid("edm_ens_add: 0..8 circular D(x) regions, and none with the "
                           "sinusoidal field");
    for (int q = 0; q < r->n_regions; ++q)
        if (!(r->reg_r[q] >= 0.0) || !(r->reg_f[q] >= 0.0))
            return ens_invalid("edm_ens_add: a D(x) region needs radius >= 0 and factor >= 0");
    if (e->n_total + r->n > e->cap) {
        set_error("edm_ens_add: particle pool full");
        return EDM_ECAPACITY;
    }
    if (e->next_id + r->n >= INT32_MAX) {
        set_error("edm_ens_add: particle index space exhausted");
        return EDM_ECAPACITY;
    }
    const bool imaging = e->has_pp && e->ip.n_bins > 0;
    if (imaging && (r->img_nx <= 0 || r->img_ny <= 0 || r->img_R < 0 || 2 * r->img_R + 1 > 31))
        return ens_invalid("edm_ens_add: bad image geometry (PSF window at most 31 pixels)");
    int k = -1;
    for (int q = 0; q < e->rmax && k < 0; ++q)
        if (!e->seg[q].active &&
            std::find(e->draining.begin(), e->draining.end(), q) == e->draining.end())
            k = q;
    if (k < 0) {
        set_error("edm_ens_add: no free replica slot");
        return EDM_ECAPACITY;
    }
    const int n = r->n;
    const int ns = e->p.n_species;
    for (int i = 0; i < n; ++i)
        if (s->species_id[i] < 0 || s->species_id[i] >= ns)
            return ens_invalid("edm_ens_add: species id out of range");
    EnsSeg g{};
    // g owns its CDF and TCSPC stack until it is stored in e->seg[k]: any
    // error return before that frees them
    struct SegGuard {
        EnsSeg* g;
        bool armed;
        ~SegGuard() {
            if (!armed) return;
            if (g->rcdf) hipFree(g->rcdf);
            if (g->flim) hipFree(g->flim);
            if (g->reg) hipFree(g->reg);
            if (g->w_adapt) hipFree(g->w_adapt);
            if (g->slot_of) hipFree(g->slot_of);
        }
    } guard{&g, true};
    g.active = 1;
    g.kbase = (int32_t)e->next_id;
    g.n = n;
    g.ncx = r->ncx;
    g.ncy = r->ncy;
    g.px = e->p.periodic_x;
    g.py = e->p.periodic_y;
    g.Lx = r->Lx;
    g.Ly = r->Ly;
    g.csx = r->csx;
    g.csy = r->csy;
    g.cutoff = r->cutoff;
    g.sin_kx = r->sin_kx;
    g.sin_ky = r->sin_ky;
    g.seed_lo = (uint32_t)(r->seed & 0xffffffffu);
    g.seed_hi = (uint32_t)(r->seed >> 32);
    g.step0 = e->gstep;
    g.n_grids = r->n_grids;
    for (int q = 0; q < r->n_grids; ++q) {
        g.gdims[2 * q] = r->grids[q].nx;
        g.gdims[2 * q + 1] = r->grids[q].ny;
        g.gs_d[2 * q] = r->grids[q].sx;
        g.gs_d[2 * q + 1] = r->grids[q].sy;
        g.gs_f[2 * q] = (float)r->grids[q].sx;
        g.gs_f[2 * q + 1] = (float)r->grids[q].sy;
        const double area = r->grids[q].sx * r->grids[q].sy;
        g.ginv_d[q] = area > 0.0 ? 1.0 / area : 0.0;
        g.ginv_f[q] = (float)g.ginv_d[q];
    }
    for (int q = 0; q < r->n_reactions; ++q) {
        g.prob[q] = r->prob[q];
        g.prob2[q] = r->prob2[q];
        g.maxr[q] = r->max_recruits[q];
    }
    // proximity FRET: this replica's target grid (as the single membrane's,
    // fret_prox_grid: cells at least the radius plus a rounding margin, at
    // most max(n, 64) of them)
    if (e->prox_map >= 0) {
        const double Lx = r->Lx, Ly = r->Ly, rad = e->pp.fm_radius[e->prox_map];
        const double ulp = e->fp64 ? 0x1p-52 : 0x1p-23;
        const double cmin = rad * 1.001 + 64.0 * ulp * std::max(Lx, Ly);
        long ncx = std::max(1L, std::min(4096L, (long)std::floor(Lx / cmin)));
        long ncy = std::max(1L, std::min(4096L, (long)std::floor(Ly / cmin)));
        const long maxc = std::max<long>(n, 64);
        while (ncx * ncy > maxc) {
            if (ncx >= ncy) ncx = (ncx + 1) / 2;
            else ncy = (ncy + 1) / 2;
        }
        g.tg_ncx = (int32_t)ncx;
        g.tg_ncy = (int32_t)ncy;
        g.tg_ix = (double)ncx / Lx;
        g.tg_iy = (double)ncy / Ly;
    }
    // binding models: the slot of each particle, written by every sort
    if (e->has_bind) EDM_HIP(hipMalloc(&g.slot_of, (size_t)std::max(n, 1) * sizeof(int32_t)));
    // the CDF of density-biased recruitment, sized by this replica's grid
    g.rcdf = nullptr;
    {
        size_t ncdf = 0;
        for (const auto& rx : e->rx)
            if (rx.kind == EDM_RX_RECRUIT && rx.use_adapt)
                ncdf = std::max(ncdf, (size_t)g.gdims[2 * rx.fb1.grid] * g.gdims[2 * rx.fb1.grid + 1]);
        if (ncdf) EDM_HIP(hipMalloc(&g.rcdf, ncdf * sizeof(double)));
    }
    // this replica's D(x) regions in the position type (as the single
    // membrane's, upload_real) and its adaptation field, zero at admission
    g.n_regions = r->n_regions;
    if (r->n_regions > 0) {
        const int R = kEnsMaxRegions;
        std::vector<double> v(4 * R, 0.0);
        for (int q = 0; q < r->n_regions; ++q) {
            v[q] = r->reg_cx[q];
            v[R + q] = r->reg_cy[q];
            v[2 * R + q] = r->reg_r[q];
            v[3 * R + q] = r->reg_f[q];
        }
        EDM_HIP(hipMalloc(&g.reg, v.size() * e->rs));
        if (e->fp64) {
            EDM_HIP(hipMemcpy(g.reg, v.data(), v.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(v.begin(), v.end());
            EDM_HIP(hipMemcpy(g.reg, f.data(), f.size() * 4, hipMemcpyHostToDevice));
        }
    }
    if (e->p.adapt_grid >= 0) {
        const size_t nw = (size_t)g.gdims[2 * e->p.adapt_grid] * g.gdims[2 * e->p.adapt_grid + 1];
        EDM_HIP(hipMalloc(&g.w_adapt, std::max<size_t>(nw, 1) * sizeof(double)));
        EDM_HIP(hipMemset(g.w_adapt, 0, std::max<size_t>(nw, 1) * sizeof(double)));
    }
    if (imaging) {
        g.ip = e->ip;
        g.ip.nx = r->img_nx;
        g.ip.ny = r->img_ny;
        g.ip.R = r->img_R;
        g.ip.pixel = r->img_pixel;
        g.ip.offx = r->img_offx;
        g.ip.offy = r->img_offy;
        const int64_t npix = (int64_t)g.ip.n_ch * g.ip.nx * g.ip.ny;
        g.flim_elems = npix * g.ip.n_bins;
        g.flim_words = g.flim_elems + (npix + 255) / 256 * 64;  // counts, row flags
        EDM_HIP(hipMalloc(&g.flim, (size_t)g.flim_words * sizeof(uint32_t)));
        EDM_HIP(hipMemsetAsync(g.flim, 0, (size_t)g.flim_words * sizeof(uint32_t), e->stream));
    }
    // particles, appended after every pooled one (particle-index order)
    const int64_t beg = e->n_total;
    std::vector<double> x(n), y(n);
    std::vector<int8_t> sp(n), ctmc(n, 0), dye(n, -1);
    std::vector<uint8_t> fl(n);
    std::vector<int32_t> id(n);
    std::vector<uint16_t> rep(n, (uint16_t)k);
    std::vector<int32_t> bond(n, -1);
    for (int i = 0; i < n; ++i) {
        x[i] = s->positions[2 * i];
        y[i] = s->positions[2 * i + 1];
        sp[i] = (int8_t)s->species_id[i];
        uint8_t f = s->is_alive[i] ? EDM_FLAG_ALIVE : 0;
        if (s->has_dye && s->has_dye[i]) f |= EDM_FLAG_DYE;
        if (s->is_bleached && s->is_bleached[i]) f |= EDM_FLAG_BLEACHED;
        fl[i] = f;
        id[i] = g.kbase + i;
        if (s->ctmc_state) ctmc[i] = s->ctmc_state[i];
        if (s->dye_type_id) dye[i] = (int8_t)s->dye_type_id[i];
        if (s->bond_partner && s->bond_partner[i] >= 0) {
            if (!e->has_bind || s->bond_partner[i] >= n)
                return ens_invalid("edm_ens_add: bond partner without binding, or out of range");
            bond[i] = g.kbase + s->bond_partner[i];
        }
    }
    hipStream_t st = e->stream;
    EDM_HIP(hipStreamSynchronize(st));
    if (e->fp64) {
        EDM_HIP(hipMemcpy((double*)e->a.x + beg, x.data(), n * 8, hipMemcpyHostToDevice));
        EDM_HIP(hipMemcpy((double*)e->a.y + beg, y.data(), n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> xf(x.begin(), x.end()), yf(y.begin(), y.end());
        EDM_HIP(hipMemcpy((float*)e->a.x + beg, xf.data(), n * 4, hipMemcpyHostToDevice));
        EDM_HIP(hipMemcpy((float*)e->a.y + beg, yf.data(), n * 4, hipMemcpyHostToDevice));
    }
    EDM_HIP(hipMemcpy(e->a.sp + beg, sp.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(e->a.fl + beg, fl.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(e->a.id + beg, id.data(), n * 4, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(e->a.ctmc + beg, ctmc.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(e->a.dye + beg, dye.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(e->rep + beg, rep.data(), n * 2, hipMemcpyHostToDevice));
    if (e->has_bind)
        EDM_HIP(hipMemcpy(e->a.bond + beg, bond.data(), n * 4, hipMemcpyHostToDevice));
    EDM_HIP(hipMemset(e->ctr + 8 * (size_t)k, 0, 8 * sizeof(unsigned long long)));
    EDM_HIP(hipMemset(e->noise_ctr + 2 * (size_t)k, 0, 2 * sizeof(unsigned long long)));
    int64_t n_alive = 0;
    for (int i = 0; i < n; ++i) n_alive += (fl[i] & EDM_FLAG_ALIVE) ? 1 : 0;
    g.sub_rows = force_sub_rows(n_alive, (int64_t)g.ncx * g.ncy) ? 1 : 0;
    e->seg[k] = g;
    guard.armed = false;
    e->fresh_beg[k] = beg;
    int64_t npass = 0, ndye = 0;
    for (int i = 0; i < n; ++i) {
        npass += (e->pass_species >> sp[i]) & 1u;
        ndye += (fl[i] & EDM_FLAG_DYE) ? 1 : 0;
    }
    e->seg_npass[k] = npass;
    e->seg_ndye[k] = ndye;
    e->seg_alive[k].assign(ns, 0);
    for (int i = 0; i < n; ++i)
        if (fl[i] & EDM_FLAG_ALIVE) ++e->seg_alive[k][sp[i]];
    e->n_total += n;
    e->next_id += n;
    e->layout_dirty = true;
    *slot = k;
    return 0;
}

int edm_ens_remove(edm_ensemble* e, int32_t slot) {
    if (!e || slot < 0 || slot >= e->rmax || !e->seg[slot].active)
        return ens_invalid("edm_ens_remove: no such replica");
    EDM_HIP(hipStreamSynchronize(e->stream));
    EnsSeg& g = e->seg[slot];
    if (g.flim) EDM_HIP(hipFree(g.flim));
    g.flim = nullptr;
    // the next step's layout uploads this slot's tables without the pointer
    if (g.rcdf) EDM_HIP(hipFree(g.rcdf));
    g.rcdf = nullptr;
    if (g.reg) EDM_HIP(hipFree(g.reg));
    g.reg = nullptr;
    if (g.w_adapt) EDM_HIP(hipFree(g.w_adapt));
    g.w_adapt = nullptr;
    if (g.slot_of) EDM_HIP(hipFree(g.slot_of));
    g.slot_of = nullptr;
    g.active = 0;
    e->n_retired += g.n;
    e->draining.push_back(slot);
    e->layout_dirty = true;
    return 0;
}

int edm_ens_run(edm_ensemble* e, int32_t n_steps, int32_t image) {
    if (!e || n_steps < 0) return ens_invalid("edm_ens_run: bad argument");
    for (int k = 0; k < n_steps; ++k)
        if (int rc = ens_step(e, image != 0)) return rc;
    return (e->has_bind || e->has_autocat) && n_steps > 0 ? ens_check_edges(e) : 0;
}

int edm_ens_frame(edm_ensemble* e, int32_t slot, uint32_t* flim, int64_t n_elems, int64_t* n_dark,
                  int64_t* n_ap) {
    if (!e || slot < 0 || slot >= e->rmax || !e->seg[slot].active || !e->seg[slot].flim)
        return ens_invalid("edm_ens_frame: no imaged replica in that slot");
    EnsSeg& g = e->seg[slot];
    if (int rc = ens_frame_noise(e, slot, n_dark, n_ap)) return rc;
    if (flim) {
        if (n_elems < g.flim_elems) return ens_invalid("edm_ens_frame: output too small");
        EDM_HIP(hipMemcpyAsync(flim, g.flim, (size_t)g.flim_elems * 4, hipMemcpyDeviceToHost,
                               e->stream));
    }
    if (int rc = ens_frame_reset(e, slot)) return rc;
    EDM_HIP(hipStreamSynchronize(e->stream));
    return 0;
}

int edm_ens_frames_collect(edm_ensemble* e, const int32_t* slots, int32_t n_slots,
                           int32_t* ticket) {
    if (!e || !ticket || n_slots < 0 || (n_slots > 0 && !slots))
        return ens_invalid("edm_ens_frames_collect: null argument");
    for (int i = 0; i < n_slots; ++i) {
        if (slots[i] < 0 || slots[i] >= e->rmax || !e->seg[slots[i]].active)
            return ens_invalid("edm_ens_frames_collect: no such replica");
        for (int j = 0; j < i; ++j)
            if (slots[j] == slots[i]) return ens_invalid("edm_ens_frames_collect: slot listed twice");
    }
    return ens_frames_collect(e, slots, n_slots, ticket);
}

int edm_ens_frames_fetch(edm_ensemble* e, int32_t ticket, int64_t* counts, int64_t* meta,
                         uint32_t* flat_index, uint32_t* count, int64_t cap, int64_t* total) {
    if (!e) return ens_invalid("edm_ens_frames_fetch: null handle");
    return ens_frames_fetch(e, ticket, counts, meta, flat_index, count, cap, total);
}

int edm_ens_frame_end(edm_ensemble* e, const int32_t* slots, int32_t n_slots) {
    if (!e || (n_slots > 0 && !slots)) return ens_invalid("edm_ens_frame_end: null argument");
    for (int i = 0; i < n_slots; ++i)
        if (slots[i] < 0 || slots[i] >= e->rmax || !e->seg[slots[i]].active)
            return ens_invalid("edm_ens_frame_end: no such replica");
    return ens_frame_end(e, slots, n_slots);
}

int edm_ens_noise_counters(edm_ensemble* e, int32_t slot, int64_t out[2]) {
    if (!e || !out || slot < 0 || slot >= e->rmax || !e->seg[slot].active)
        return ens_invalid("edm_ens_noise_counters: no such replica");
    unsigned long long c[2];
    EDM_HIP(hipMemcpyAsync(c, e->noise_ctr + 2 * (size_t)slot, sizeof(c), hipMemcpyDeviceToHost,
                           e->stream));
    EDM_HIP(hipStreamSynchronize(e->stream));
    out[0] = (int64_t)c[0];
    out[1] = (int64_t)c[1];
    return 0;
}

int edm_ens_download(edm_ensemble* e, int32_t slot, edm_host_state* s) {
    if (!e || !s || slot < 0 || slot >= e->rmax || !e->seg[slot].active)
        return ens_invalid("edm_ens_download: no such replica");
    const EnsSeg& g = e->seg[slot];
    hipStream_t st = e->stream;
    EDM_HIP(hipStreamSynchronize(st));
    int64_t beg = e->fresh_beg[slot];
    if (beg < 0) {
        uint32_t se[2];
        EDM_HIP(hipMemcpy(&se[0], e->cell_start + g.coff, 4, hipMemcpyDeviceToHost));
        EDM_HIP(hipMemcpy(&se[1], e->cell_start + g.coff + (int64_t)g.ncx * g.ncy, 4,
                          hipMemcpyDeviceToHost));
        if ((int64_t)(se[1] - se[0]) != g.n) {
            set_error("edm_ens_download: replica range inconsistent with its size");
            return EDM_ESTATE;
        }
        beg = se[0];
    }
    const int n = g.n;
    std::vector<double> x(n), y(n);
    std::vector<int8_t> sp(n), ctmc(n), dye(n);
    std::vector<uint8_t> fl(n);
    std::vector<int32_t> id(n);
    if (e->fp64) {
        EDM_HIP(hipMemcpy(x.data(), (const double*)e->a.x + beg, n * 8, hipMemcpyDeviceToHost));
        EDM_HIP(hipMemcpy(y.data(), (const double*)e->a.y + beg, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> xf(n), yf(n);
        EDM_HIP(hipMemcpy(xf.data(), (const float*)e->a.x + beg, n * 4, hipMemcpyDeviceToHost));
        EDM_HIP(hipMemcpy(yf.data(), (const float*)e->a.y + beg, n * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            x[i] = xf[i];
            y[i] = yf[i];
        }
    }
    EDM_HIP(hipMemcpy(sp.data(), e->a.sp + beg, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(fl.data(), e->a.fl + beg, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(id.data(), e->a.id + beg, n * 4, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(ctmc.data(), e->a.ctmc + beg, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(dye.data(), e->a.dye + beg, n, hipMemcpyDeviceToHost));
    std::vector<int32_t> bond(n, -1);
    if (e->has_bind)
        EDM_HIP(hipMemcpy(bond.data(), e->a.bond + beg, n * 4, hipMemcpyDeviceToHost));
    for (int t = 0; t < n; ++t) {
        const int i = id[t] - g.kbase;
        if (i < 0 || i >= n) {
            set_error("edm_ens_download: particle index outside the replica");
            return EDM_ESTATE;
        }
        s->positions[2 * i] = x[t];
        s->positions[2 * i + 1] = y[t];
        s->species_id[i] = sp[t];
        s->is_alive[i] = (fl[t] & EDM_FLAG_ALIVE) ? 1 : 0;
        if (s->has_dye) s->has_dye[i] = (fl[t] & EDM_FLAG_DYE) ? 1 : 0;
        if (s->is_bleached) s->is_bleached[i] = (fl[t] & EDM_FLAG_BLEACHED) ? 1 : 0;
        if (s->ctmc_state) s->ctmc_state[i] = ctmc[t];
        if (s->dye_type_id) s->dye_type_id[i] = dye[t];
        if (s->bond_partner) s->bond_partner[i] = bond[t] >= 0 ? bond[t] - g.kbase : -1;
    }
    return 0;
}

int edm_ens_counters(edm_ensemble* e, int32_t slot, int64_t out[8]) {
    if (!e || !out || slot < 0 || slot >= e->rmax || !e->seg[slot].active)
        return ens_invalid("edm_ens_counters: no such replica");
    unsigned long long c[8];
    EDM_HIP(hipMemcpyAsync(c, e->ctr + 8 * (size_t)slot, sizeof(c), hipMemcpyDeviceToHost,
                           e->stream));
    EDM_HIP(hipStreamSynchronize(e->stream));
    for (int i = 0; i < 8; ++i) out[i] = (int64_t)c[i];
    return 0;
}

int edm_ens_set_tuning(edm_ensemble* e, int32_t key, int64_t value) {
    if (!e) return ens_invalid("edm_ens_set_tuning: null handle");
    switch (key) {
        case EDM_ENS_TUNE_STRIP_PARTICLES:
            if (value < 32 || value > 512)
                return ens_invalid("edm_ens_set_tuning: strip particles must be in 32..512");
            e->strip_parts = (int)value;
            break;
        case EDM_ENS_TUNE_PAIR_BELOW:
            if (value < 0 || value > 1024)
                return ens_invalid("edm_ens_set_tuning: pair-below must be in 0..1024");
            e->pair_below = (int)value;
            break;
        case EDM_ENS_TUNE_CAT_LANES:
            if (value < 1 || value > 64 || (value & (value - 1)))
                return ens_invalid("edm_ens_set_tuning: catalysis lanes must be a power of two "
                                   "in 1..64");
            e->cat_lanes = (int)value;
            break;
        case EDM_ENS_TUNE_CAT_WALK:
            if (value != 0 && value != 1)
                return ens_invalid("edm_ens_set_tuning: catalysis walk must be 0 or 1");
            e->cat_walk = (int)value;
            break;
        default:
            return ens_invalid("edm_ens_set_tuning: unknown key");
    }
    // strips and catalysis tables are rebuilt before the next step
    e->layout_dirty = true;
    return 0;
}

int edm_ens_get_tuning(edm_ensemble* e, int32_t key, int64_t* value) {
    if (!e || !value) return ens_invalid("edm_ens_get_tuning: null argument");
    switch (key) {
        case EDM_ENS_TUNE_STRIP_PARTICLES: *value = e->strip_parts; return 0;
        case EDM_ENS_TUNE_PAIR_BELOW: *value = e->pair_below; return 0;
        case EDM_ENS_TUNE_CAT_LANES: *value = e->cat_lanes; return 0;
        case EDM_ENS_TUNE_CAT_WALK: *value = e->cat_walk; return 0;
        default: return ens_invalid("edm_ens_get_tuning: unknown key");
    }
}

int edm_ens_info(edm_ensemble* e, int64_t out[4]) {
    if (!e || !out) return ens_invalid("edm_ens_info: null argument");
    int64_t live = 0;
    for (auto& g : e->seg) live += g.active ? 1 : 0;
    out[0] = e->n_total;
    out[1] = e->n_cells;
    out[2] = live;
    out[3] = e->cap;
    return 0;
}

int edm_ens_synchronize(edm_ensemble* e) {
    if (!e) return ens_invalid("edm_ens_synchronize: null handle");
    EDM_HIP(hipStreamSynchronize(e->stream));
    return 0;
}

}  // extern "C"
