// C-ABI entry points of libedm_hip.so (declared in include/edm_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "edm_internal.h"

namespace edm {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return EDM_EHIP;
}

int stage_draws(edm_membrane* m, const double* host, int64_t n) {
    if ((size_t)n > m->draws_cap) {
        EDM_HIP(hipStreamSynchronize(m->stream));
        if (m->d_draws) EDM_HIP(hipFree(m->d_draws));
        m->draws_cap = (size_t)std::max<int64_t>(n, 1024);
        m->draws_cap += m->draws_cap / 4;
        EDM_HIP(hipMalloc(&m->d_draws, m->draws_cap * sizeof(double)));
    }
    if (host && n > 0) {
        EDM_HIP(hipMemcpyAsync(m->d_draws, host, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                               m->stream));
        EDM_HIP(hipStreamSynchronize(m->stream));
    }
    return 0;
}

}  // namespace edm

using namespace edm;

static int invalid(const std::string& msg) {
    set_error(msg);
    return EDM_EINVAL;
}

extern "C" {

const char* edm_last_error(void) { return g_err.c_str(); }

int edm_device_count(int* count) {
    int c = 0;
    EDM_HIP(hipGetDeviceCount(&c));
    *count = c;
    return 0;
}

int edm_create(const edm_params* p, edm_membrane** out) {
    if (!p || !out) return invalid("edm_create: null argument");
    *out = nullptr;
    if (p->n <= 0) return invalid("edm_create: particle count must be positive");
    if (p->n_species <= 0 || p->n_species > 32) return invalid("edm_create: 1..32 species supported");
    if (p->ncx < 3 || p->ncy < 3) return invalid("edm_create: cell grid must be at least 3x3");
    if (p->n_grids < 0 || (p->n_grids > 0 && !p->grids)) return invalid("edm_create: bad grids");
    if (p->n_reactions < 0 || (p->n_reactions > 0 && !p->reactions))
        return invalid("edm_create: bad reactions");
    if (!p->sigma || !p->epsilon || !p->k_harm || !p->D_base)
        return invalid("edm_create: missing parameter arrays");
    for (int i = 0; i < p->n_reactions; ++i) {
        const edm_reaction& r = p->reactions[i];
        const edm_feedback* fbs[2] = {&r.fb1, &r.fb2};
        for (auto f : fbs)
            if (f->sensor >= 0 && (f->grid < 0 || f->grid >= p->n_grids || f->sensor >= p->n_species))
                return invalid("edm_create: feedback refers to a missing grid or species");
        if (r.kind == EDM_RX_CAT && r.a == r.b)
            return invalid("edm_create: catalysis needs an enzyme distinct from its substrate");
        if (r.kind < EDM_RX_BIND || r.kind > EDM_RX_ADAPT)
            return invalid("edm_create: unknown reaction kind");
        if (r.kind == EDM_RX_RECRUIT && r.use_adapt &&
            (r.fb1.sensor < 0 || r.fb2.sensor >= 0 || !(r.contact >= 0.0)))
            return invalid("edm_create: local recruitment needs one feedback rule and sigma >= 0");
    }
    if (p->adapt_grid >= p->n_grids) return invalid("edm_create: bad adaptation grid");
    edm_membrane* m = new edm_membrane();
    m->p = *p;
    m->n = p->n;
    m->fp64 = p->fp64 != 0;
    m->rs = m->fp64 ? 8 : 4;
    m->grids.assign(p->grids, p->grids + p->n_grids);
    m->rx.assign(p->reactions, p->reactions + p->n_reactions);
    for (const auto& r : m->rx)
        if (r.kind == EDM_RX_CAT && r.a == r.c) m->has_autocat = true;
    m->p.grids = nullptr;
    m->p.reactions = nullptr;
    m->p.sigma = m->p.epsilon = m->p.k_harm = m->p.D_base = nullptr;
    m->p.rcx = m->p.rcy = m->p.rr = m->p.rf = nullptr;
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = hipStreamCreate(&m->stream);
    if (e != hipSuccess) {
        delete m;
        return hip_fail(e, "edm_create: stream");
    }
    edm_params tmp = *p;  // keep caller's arrays valid during create
    m->p = tmp;
    int rc = dyn_create(m);
    m->p.grids = nullptr;
    m->p.reactions = nullptr;
    m->p.sigma = m->p.epsilon = m->p.k_harm = m->p.D_base = nullptr;
    m->p.rcx = m->p.rcy = m->p.rr = m->p.rf = nullptr;
    if (rc) {
        edm_destroy(m);
        return rc;
    }
    EDM_HIP(hipEventCreate(&m->ev0));
    EDM_HIP(hipEventCreate(&m->ev1));
    *out = m;
    return 0;
}

int edm_destroy(edm_membrane* m) {
    if (!m) return 0;
    if (m->stream) hipStreamSynchronize(m->stream);
    graphs_free(m);
    frames_free(m);
    if (m->fr_stream) hipStreamDestroy(m->fr_stream);
    if (m->fr_photons) hipEventDestroy(m->fr_photons);
    if (m->ph_stream) {
        hipStreamSynchronize(m->ph_stream);
        hipStreamDestroy(m->ph_stream);
    }
    if (m->ev_force) hipEventDestroy(m->ev_force);
    if (m->ev_photons) hipEventDestroy(m->ev_photons);
    if (m->fr_done) hipEventDestroy(m->fr_done);
    if (m->d_step) hipFree(m->d_step);
    rec_free(m);
    stats_free(m);
    void* bufs[] = {m->a.x, m->a.y, m->a.xn, m->a.yn, m->a.xt, m->a.yt, m->a.sp, m->a.sp_t, m->a.fl,
                    m->a.fl_t, m->a.id, m->a.id_t, m->a.bond, m->a.bond_t, m->a.ctmc, m->a.ctmc_t,
                    m->a.dye, m->a.dye_t, m->a.slot_of, m->key, m->rank, m->cell_cnt,
                    m->cell_start, m->perm, m->cub_tmp, m->sig, m->eps, m->kh, m->Dbase, m->rcx,
                    m->rcy, m->rr, m->rf, m->d_count_off, m->d_grid_dims, m->d_grid_s, m->counts,
                    m->delta, m->pair_cnt, m->pair_off, m->d_ctr, m->d_draws, m->edge_key,
                    m->edge_a, m->edge_b, m->n_edges, m->best, m->tconv, m->cand_id, m->cand_slot, m->rcdf,
                    m->cand_x, m->cand_y, m->n_cand, m->w_adapt, m->d_pp, m->k_fret, m->acc,
                    m->ph_nd, m->ph_nf, m->ph_cnt, m->ph_off, m->ph_pid, m->ph_t, m->ph_ch,
                    m->ph_fret, m->flim, m->pcx, m->det_flag, m->det_pos, m->sv_key, m->sv_key2,
                    m->sv_val, m->sv_val2, m->asg_px, m->asg_py, m->asg_flag, m->asg_pos,
                    m->asg_t, m->asg_ch, m->sv_px, m->sv_py, m->pix_tot, m->d_scal,
                    m->dye_ids, m->wl, m->wl_cnt, m->fstart, m->fkey, m->tg_cnt, m->tg_fill, m->tg_start,
                    m->tg_x, m->tg_y, m->tg_u, m->scan_state, m->scan_ticket};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (auto e : m->kev) hipEventDestroy(e);
    if (m->ev0) hipEventDestroy(m->ev0);
    if (m->ev1) hipEventDestroy(m->ev1);
    if (m->mark0) hipEventDestroy(m->mark0);
    if (m->mark1) hipEventDestroy(m->mark1);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

int edm_upload_state(edm_membrane* m, const edm_host_state* s) {
    if (!m || !s || !s->positions || !s->species_id || !s->is_alive)
        return invalid("edm_upload_state: null argument");
    graph_touch(m);
    const int n = m->n;
    std::vector<double> x(n), y(n);
    std::vector<int8_t> sp(n), ctmc(n, 0), dye(n, -1);
    std::vector<uint8_t> fl(n);
    std::vector<int32_t> id(n), bond(n, -1);
    for (int i = 0; i < n; ++i) {
        x[i] = s->positions[2 * i];
        y[i] = s->positions[2 * i + 1];
        if (s->species_id[i] < 0 || s->species_id[i] >= m->p.n_species)
            return invalid("edm_upload_state: species id out of range");
        sp[i] = (int8_t)s->species_id[i];
        uint8_t f = s->is_alive[i] ? EDM_FLAG_ALIVE : 0;
        if (s->has_dye && s->has_dye[i]) f |= EDM_FLAG_DYE;
        if (s->is_bleached && s->is_bleached[i]) f |= EDM_FLAG_BLEACHED;
        fl[i] = f;
        id[i] = i;
        if (s->bond_partner) bond[i] = s->bond_partner[i];
        if (s->ctmc_state) ctmc[i] = s->ctmc_state[i];
        if (s->dye_type_id) dye[i] = (int8_t)s->dye_type_id[i];
    }
    hipStream_t st = m->stream;
    EDM_HIP(hipStreamSynchronize(st));
    if (m->fp64) {
        EDM_HIP(hipMemcpy(m->a.x, x.data(), n * 8, hipMemcpyHostToDevice));
        EDM_HIP(hipMemcpy(m->a.y, y.data(), n * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> xf(x.begin(), x.end()), yf(y.begin(), y.end());
        EDM_HIP(hipMemcpy(m->a.x, xf.data(), n * 4, hipMemcpyHostToDevice));
        EDM_HIP(hipMemcpy(m->a.y, yf.data(), n * 4, hipMemcpyHostToDevice));
    }
    EDM_HIP(hipMemcpy(m->a.sp, sp.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.fl, fl.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.id, id.data(), n * 4, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.slot_of, id.data(), n * 4, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.bond, bond.data(), n * 4, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.ctmc, ctmc.data(), n, hipMemcpyHostToDevice));
    EDM_HIP(hipMemcpy(m->a.dye, dye.data(), n, hipMemcpyHostToDevice));
    // dye carriers never change: the photon kernels run over this list
    std::vector<int32_t> dyes;
    for (int i = 0; i < n; ++i)
        if (fl[i] & EDM_FLAG_DYE) dyes.push_back(i);
    if (m->dye_ids) EDM_HIP(hipFree(m->dye_ids));
    m->dye_ids = nullptr;
    EDM_HIP(hipMalloc(&m->dye_ids, std::max<size_t>(dyes.size(), 1) * sizeof(int32_t)));
    if (!dyes.empty())
        EDM_HIP(hipMemcpy(m->dye_ids, dyes.data(), dyes.size() * 4, hipMemcpyHostToDevice));
    m->n_dye = (int)dyes.size();
    // bonds can exist: given at upload, or made by a binding reaction
    bool bonds = false;
    for (auto& r : m->rx) bonds = bonds || r.kind == EDM_RX_BIND;
    for (int i = 0; i < n && !bonds; ++i) bonds = bond[i] >= 0;
    // photons beside the reactions: the species a dye carrier can hold
    // (its species at upload, closed under catalysis and conversion) are
    // touched by no recruitment, detachment or binding, and no FRET map
    // reads positions of other particles (proximity)
    {
        uint32_t C = 0;
        for (int i = 0; i < n; ++i)
            if (fl[i] & EDM_FLAG_DYE) C |= 1u << sp[i];
        for (bool grow = true; grow;) {
            grow = false;
            for (auto& r : m->rx) {
                uint32_t T = 0;
                if (r.kind == EDM_RX_CAT) T = (1u << r.b) | (1u << r.c);
                else if (r.kind == EDM_RX_CONVERT) T = (1u << r.a) | (1u << r.b);
                if ((T & C) && (T & ~C)) {
                    C |= T;
                    grow = true;
                }
            }
        }
        bool ok = !bonds;
        for (auto& r : m->rx)
            if ((r.kind == EDM_RX_RECRUIT || r.kind == EDM_RX_DETACH) && ((C >> r.a) & 1u))
                ok = false;
        if (m->has_pp)
            for (int j = 0; j < m->pp.n_fret_maps; ++j)
                if (m->pp.fm_proximity[j] && !m->pp.fm_skip[j]) ok = false;
        m->ph_overlap_ok = ok;
    }
    m->keep_slot_of = bonds;
    // without bonds the step does not gather the bond arrays: both hold -1
    if (!bonds && n > 0)
        EDM_HIP(hipMemcpy(m->a.bond_t, bond.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    m->slot_of_valid = true;  // the identity, copied above
    m->alive_at_upload.assign(m->p.n_species, 0);
    int64_t n_alive = 0;
    for (int i = 0; i < n; ++i)
        if (fl[i] & EDM_FLAG_ALIVE) {
            ++m->alive_at_upload[sp[i]];
            ++n_alive;
        }
    m->sub_rows = force_sub_rows(n_alive, (int64_t)m->p.ncx * m->p.ncy);
    const int nb = std::max(1, (n + kWlSlots - 1) / kWlSlots);
    if (nb != m->wl_nblk) {
        if (m->wl) EDM_HIP(hipFree(m->wl));
        if (m->wl_cnt) EDM_HIP(hipFree(m->wl_cnt));
        m->wl = nullptr;
        m->wl_cnt = nullptr;
        EDM_HIP(hipMalloc(&m->wl, (size_t)kWlMaxLists * nb * kWlSlots * sizeof(int32_t)));
        EDM_HIP(hipMalloc(&m->wl_cnt, (size_t)kWlMaxLists * nb * sizeof(uint32_t)));
        m->wl_nblk = nb;
    }
    EDM_HIP(hipMemset(m->wl_cnt, 0, (size_t)kWlMaxLists * nb * sizeof(uint32_t)));
    m->wl_valid = false;
    m->wl_pend_n = 0;  // a pending build would write the old lists
    return 0;
}

int edm_set_tuning(edm_membrane* m, int32_t key, int64_t value) {
    if (!m) return invalid("edm_set_tuning: null membrane");
    graph_touch(m);
    switch (key) {
        case EDM_TUNE_TILE_ROWS:
            if (value < 1 || value > kTileRows)
                return invalid("edm_set_tuning: tile rows must be in 1..256");
            m->tile_rows = (int)value;
            return 0;
        case EDM_TUNE_CAT_SIDE:
            if (value < -1 || value > 1)
                return invalid("edm_set_tuning: catalysis side must be -1, 0 or 1");
            m->cat_side = (int)value;
            return 0;
        case EDM_TUNE_CAT_WALK:
            if (value < 0 || value > 1) return invalid("edm_set_tuning: catalysis walk must be 0 or 1");
            m->cat_walk = (int)value;
            return 0;
        case EDM_TUNE_CAT_LANES:
            if (value < 1 || value > 64 || (value & (value - 1)))
                return invalid("edm_set_tuning: catalysis lanes must be a power of two in 1..64");
            m->cat_lanes = (int)value;
            return 0;
        case EDM_TUNE_FRET_SEARCH:
            if (value < 0 || value > 1)
                return invalid("edm_set_tuning: FRET search must be 0 (grid) or 1 (scan)");
            m->fret_search = (int)value;
            return 0;
        case EDM_TUNE_GRAPHS:
            if (value < 0 || value > 1) return invalid("edm_set_tuning: graphs must be 0 or 1");
            m->use_graphs = (int)value;
            return 0;
        case EDM_TUNE_FORCE_LISTS:
            if (value < 0 || value > 1) return invalid("edm_set_tuning: force lists must be 0 or 1");
            m->fine_force = (int)value;
            return 0;
        case EDM_TUNE_GRAPH_REPLAYS:
        case EDM_TUNE_GRAPH_STATUS:
            return invalid("edm_set_tuning: read-only key");
        default:
            return invalid("edm_set_tuning: unknown key");
    }
}

int edm_get_tuning(edm_membrane* m, int32_t key, int64_t* value) {
    if (!m || !value) return invalid("edm_get_tuning: null argument");
    switch (key) {
        case EDM_TUNE_TILE_ROWS: *value = m->tile_rows; return 0;
        case EDM_TUNE_CAT_SIDE: *value = m->cat_side; return 0;
        case EDM_TUNE_CAT_WALK: *value = m->cat_walk; return 0;
        case EDM_TUNE_CAT_LANES: *value = m->cat_lanes; return 0;
        case EDM_TUNE_FRET_SEARCH: *value = m->fret_search; return 0;
        case EDM_TUNE_GRAPHS: *value = m->use_graphs; return 0;
        case EDM_TUNE_GRAPH_REPLAYS: *value = (int64_t)m->graph_replays; return 0;
        case EDM_TUNE_GRAPH_STATUS: *value = m->graph_status; return 0;
        case EDM_TUNE_FORCE_LISTS: *value = m->fine_force; return 0;
        default: return invalid("edm_get_tuning: unknown key");
    }
}

int edm_download_state(edm_membrane* m, edm_host_state* s) {
    if (!m || !s) return invalid("edm_download_state: null argument");
    const int n = m->n;
    if (int rc = dyn_check_overflow(m)) return rc;
    if (int rc = dyn_check_scan(m)) return rc;
    std::vector<double> x(n), y(n);
    if (m->fp64) {
        EDM_HIP(hipMemcpy(x.data(), m->a.x, n * 8, hipMemcpyDeviceToHost));
        EDM_HIP(hipMemcpy(y.data(), m->a.y, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> xf(n), yf(n);
        EDM_HIP(hipMemcpy(xf.data(), m->a.x, n * 4, hipMemcpyDeviceToHost));
        EDM_HIP(hipMemcpy(yf.data(), m->a.y, n * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            x[i] = xf[i];
            y[i] = yf[i];
        }
    }
    std::vector<int8_t> sp(n), ctmc(n), dye(n);
    std::vector<uint8_t> fl(n);
    std::vector<int32_t> id(n), bond(n);
    EDM_HIP(hipMemcpy(sp.data(), m->a.sp, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(fl.data(), m->a.fl, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(id.data(), m->a.id, n * 4, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(bond.data(), m->a.bond, n * 4, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(ctmc.data(), m->a.ctmc, n, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(dye.data(), m->a.dye, n, hipMemcpyDeviceToHost));
    for (int t = 0; t < n; ++t) {
        const int i = id[t];
        if (i < 0 || i >= n) return invalid("edm_download_state: corrupt slot map");
        if (s->positions) {
            s->positions[2 * i] = x[t];
            s->positions[2 * i + 1] = y[t];
        }
        if (s->species_id) s->species_id[i] = sp[t];
        if (s->is_alive) s->is_alive[i] = (fl[t] & EDM_FLAG_ALIVE) ? 1 : 0;
        if (s->has_dye) s->has_dye[i] = (fl[t] & EDM_FLAG_DYE) ? 1 : 0;
        if (s->is_bleached) s->is_bleached[i] = (fl[t] & EDM_FLAG_BLEACHED) ? 1 : 0;
        if (s->bond_partner) s->bond_partner[i] = bond[t];
        if (s->ctmc_state) s->ctmc_state[i] = ctmc[t];
        if (s->dye_type_id) s->dye_type_id[i] = dye[t];
    }
    return 0;
}

int edm_set_photophysics(edm_membrane* m, const edm_photo_params* pp) {
    if (!m || !pp) return invalid("edm_set_photophysics: null argument");
    graph_touch(m);
    if (pp->n_fluoro < 0 || pp->n_fluoro > 8 || pp->n_fret_maps < 0 || pp->n_fret_maps > 8)
        return invalid("edm_set_photophysics: at most 8 fluorophores and 8 FRET pairs");
    m->pp = *pp;
    m->has_pp = true;
    return ph_alloc(m);
}

int edm_set_imaging(edm_membrane* m, const edm_image_params* ip) {
    if (!m || !ip) return invalid("edm_set_imaging: null argument");
    graph_touch(m);
    if (ip->nx <= 0 || ip->ny <= 0 || ip->n_bins <= 0 || ip->n_ch <= 0 || ip->n_sigma <= 0 ||
        ip->n_sigma > 8 || ip->R < 0)
        return invalid("edm_set_imaging: bad image geometry");
    m->ip = *ip;
    m->has_ip = true;
    return img_alloc(m);
}

int edm_step_begin(edm_membrane* m, int64_t* n_pairs) {
    if (!m) return invalid("edm_step_begin: null handle");
    graph_touch(m);
    m->need_pair_offsets = true;
    m->pair_sync = true;
    if (int rc = dyn_check_overflow(m)) return rc;
    if (int rc = dyn_step_begin(m)) return rc;
    m->begun = true;
    if (n_pairs) *n_pairs = m->n_pairs;
    return 0;
}

int edm_step_finish(edm_membrane* m, const double* draws, int64_t n_draws) {
    if (!m) return invalid("edm_step_finish: null handle");
    graph_touch(m);
    if (!m->begun) {
        set_error("edm_step_finish: no edm_step_begin for this step");
        return EDM_ESTATE;
    }
    if (draws) {
        const int64_t need = hook_draws_per_step(m, m->n_pairs);
        if (n_draws != need)
            return invalid("edm_step_finish: expected " + std::to_string(need) + " draws, got " +
                           std::to_string(n_draws));
        if (int rc = stage_draws(m, draws, n_draws)) return rc;
    }
    m->begun = false;
    return dyn_step_finish(m, draws ? m->d_draws : nullptr, draws != nullptr);
}

int edm_run(edm_membrane* m, int32_t n_steps) {
    if (!m) return invalid("edm_run: null handle");
    graph_touch(m);
    if (m->begun) {
        set_error("edm_run: a parity step is open");
        return EDM_ESTATE;
    }
    return steps_run(m, n_steps, false);
}

int edm_counters(edm_membrane* m, int64_t out[6]) {
    if (!m || !out) return invalid("edm_counters: null argument");
    unsigned long long c[8];
    EDM_HIP(hipMemcpyAsync(c, m->d_ctr, sizeof(c), hipMemcpyDeviceToHost, m->stream));
    if (int rc = dyn_check_overflow(m)) return rc;
    for (int i = 0; i < 6; ++i) out[i] = (int64_t)c[i];
    return 0;
}

int edm_photon_counters(edm_membrane* m, int64_t out[2]) {
    if (!m || !out) return invalid("edm_photon_counters: null argument");
    unsigned long long c[8];
    EDM_HIP(hipMemcpyAsync(c, m->d_ctr, sizeof(c), hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    out[0] = (int64_t)c[6];
    out[1] = (int64_t)c[7];
    return 0;
}

int edm_pairs(edm_membrane* m, int32_t* pairs, int64_t cap, int64_t* n_pairs) {
    if (!m || !pairs || !n_pairs) return invalid("edm_pairs: null argument");
    return dyn_pairs(m, pairs, cap, n_pairs);
}

int edm_density(edm_membrane* m, int32_t species, int32_t grid, double* out) {
    if (!m || !out) return invalid("edm_density: null argument");
    int rc = dyn_density(m, species, grid, out);
    if (rc == EDM_EINVAL) set_error("edm_density: density of that species/grid is not kept");
    return rc;
}

int edm_photons(edm_membrane* m, const double* u_pois, const double* u_trans, const double* u_arr,
                int64_t n_arr, int64_t* n_photons) {
    if (!m || !u_pois || !u_trans || !n_photons || (n_arr > 0 && !u_arr))
        return invalid("edm_photons: null argument");
    if (!m->has_pp) {
        set_error("edm_photons: photophysics not set");
        return EDM_ESTATE;
    }
    return ph_hook(m, u_pois, u_trans, u_arr, n_arr, n_photons);
}

int edm_photon_batch(edm_membrane* m, int32_t* particle_idx, double* arrival_time, int8_t* channel,
                     uint8_t* is_fret, int64_t cap) {
    if (!m) return invalid("edm_photon_batch: null handle");
    const int64_t n = std::min(cap, m->n_ph);
    if (n <= 0) return 0;
    EDM_HIP(hipStreamSynchronize(m->stream));
    if (particle_idx) EDM_HIP(hipMemcpy(particle_idx, m->ph_pid, n * 4, hipMemcpyDeviceToHost));
    if (arrival_time) EDM_HIP(hipMemcpy(arrival_time, m->ph_t, n * 8, hipMemcpyDeviceToHost));
    if (channel) EDM_HIP(hipMemcpy(channel, m->ph_ch, n, hipMemcpyDeviceToHost));
    if (is_fret) EDM_HIP(hipMemcpy(is_fret, m->ph_fret, n, hipMemcpyDeviceToHost));
    return 0;
}

int edm_fret_rates(edm_membrane* m, double* k_fret, int32_t* acceptor_idx) {
    if (!m) return invalid("edm_fret_rates: null handle");
    if (!m->has_pp) {
        set_error("edm_fret_rates: photophysics not set");
        return EDM_ESTATE;
    }
    if (int rc = ph_fret(m)) return rc;
    const int n = m->n, nd = m->n_dye;
    if (k_fret)
        for (int i = 0; i < n; ++i) k_fret[i] = 0.0;
    if (acceptor_idx)
        for (int i = 0; i < n; ++i) acceptor_idx[i] = -1;
    if (nd == 0) return 0;
    std::vector<double> k(nd);
    std::vector<int32_t> a(nd), id(nd);
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipMemcpy(k.data(), m->k_fret, nd * 8, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(a.data(), m->acc, nd * 4, hipMemcpyDeviceToHost));
    EDM_HIP(hipMemcpy(id.data(), m->dye_ids, nd * 4, hipMemcpyDeviceToHost));
    for (int q = 0; q < nd; ++q) {
        if (k_fret) k_fret[id[q]] = k[q];
        if (acceptor_idx) acceptor_idx[id[q]] = a[q];
    }
    return 0;
}

int edm_image_detect(edm_membrane* m, const double* u_qe, int64_t n, int64_t* n_surviving) {
    if (!m || (n > 0 && !u_qe) || !n_surviving) return invalid("edm_image_detect: null argument");
    if (!m->has_ip) {
        set_error("edm_image_detect: imaging not set");
        return EDM_ESTATE;
    }
    int rc = img_detect(m, u_qe, n, n_surviving);
    if (rc == EDM_EINVAL) set_error("edm_image_detect: draw count must equal the photon count");
    return rc;
}

int edm_image_assign(edm_membrane* m, const double* u_pix, int64_t n, int64_t* n_assigned) {
    if (!m || (n > 0 && !u_pix) || !n_assigned) return invalid("edm_image_assign: null argument");
    int rc = img_assign(m, u_pix, n, n_assigned);
    if (rc == EDM_EINVAL) set_error("edm_image_assign: draw count must equal the surviving count");
    return rc;
}

int edm_image_bin(edm_membrane* m, const double* jitter, int64_t n) {
    if (!m || (n > 0 && !jitter)) return invalid("edm_image_bin: null argument");
    int rc = img_bin(m, jitter, n);
    if (rc == EDM_EINVAL) set_error("edm_image_bin: jitter count must equal the assigned count");
    return rc;
}

int edm_photons_image(edm_membrane* m, int64_t* n_emitted, int64_t* n_detected) {
    if (!m) return invalid("edm_photons_image: null handle");
    if (!m->has_pp || !m->has_ip) {
        set_error("edm_photons_image: photophysics and imaging must be set");
        return EDM_ESTATE;
    }
    return ph_fused_image(m, n_emitted, n_detected);
}

int edm_frame_pixel_totals(edm_membrane* m, uint32_t* totals) {
    if (!m || !totals || !m->has_ip) return invalid("edm_frame_pixel_totals: bad argument");
    return frame_totals(m, totals);
}

int edm_frame_add_counts(edm_membrane* m, const int64_t* flat_idx, int64_t n) {
    if (!m || (n > 0 && !flat_idx) || !m->has_ip) return invalid("edm_frame_add_counts: bad argument");
    return frame_add(m, flat_idx, n);
}

int edm_frame_noise(edm_membrane* m, int64_t* n_dark, int64_t* n_afterpulse) {
    if (!m || !m->has_ip) return invalid("edm_frame_noise: imaging not set");
    return frame_noise(m, n_dark, n_afterpulse);
}

int edm_frame_download(edm_membrane* m, uint32_t* flim, int64_t n_elems) {
    if (!m || !flim || !m->has_ip) return invalid("edm_frame_download: bad argument");
    if ((size_t)n_elems != m->flim_elems) return invalid("edm_frame_download: size mismatch");
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipMemcpy(flim, m->flim, m->flim_elems * 4, hipMemcpyDeviceToHost));
    return 0;
}

int edm_frame_collect(edm_membrane* m, int32_t* ticket) {
    if (!m || !ticket || !m->has_ip) return invalid("edm_frame_collect: imaging not set");
    return frame_collect(m, ticket);
}

int edm_frame_fetch(edm_membrane* m, int32_t ticket, int64_t* meta, uint32_t* flat_index,
                    uint32_t* count, int64_t cap, int64_t* total) {
    if (!m) return invalid("edm_frame_fetch: null handle");
    return frame_fetch(m, ticket, meta, flat_index, count, cap, total);
}

int edm_frame_reset(edm_membrane* m) {
    if (!m || !m->has_ip) return invalid("edm_frame_reset: imaging not set");
    return frame_reset(m);
}

int edm_synchronize(edm_membrane* m) {
    if (!m) return invalid("edm_synchronize: null handle");
    return dyn_check_overflow(m);
}

int edm_time_steps(edm_membrane* m, int32_t n_steps, int32_t image, double* total_ms,
                   double* force_ms, double* breakdown_ms) {
    if (!m || n_steps <= 0) return invalid("edm_time_steps: bad argument");
    if (image && (!m->has_pp || !m->has_ip)) return invalid("edm_time_steps: imaging not set");
    // two events per step bracket the force/BD kernel launch
    const size_t need = 2 * (size_t)n_steps;
    while (m->kev.size() < need) {
        hipEvent_t e;
        EDM_HIP(hipEventCreate(&e));
        m->kev.push_back(e);
    }
    std::vector<hipEvent_t> saved = m->kev;
    EDM_HIP(hipStreamSynchronize(m->stream));
    if (graphs_eligible(m)) {
        // replayed graphs: the force launches are not timed one by one
        EDM_HIP(hipEventRecord(m->ev0, m->stream));
        if (int rc = steps_run(m, n_steps, image != 0)) return rc;
        EDM_HIP(hipEventRecord(m->ev1, m->stream));
        EDM_HIP(hipEventSynchronize(m->ev1));
        float ms = 0.f;
        EDM_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        if (total_ms) *total_ms = ms;
        if (force_ms) *force_ms = -1.0;
        if (breakdown_ms)
            for (int i = 0; i < 8; ++i) breakdown_ms[i] = 0.0;
        return 0;
    }
    EDM_HIP(hipEventRecord(m->ev0, m->stream));
    for (int k = 0; k < n_steps; ++k) {
        m->need_pair_offsets = loop_pair_offsets(m);
        m->pair_sync = false;
        if (int rc = dyn_step_begin(m)) return rc;
        m->time_force = true;
        // finish records kev[0] / kev[1] around the force launch: point them
        // at this step's pair
        std::swap(m->kev[0], m->kev[2 * k]);
        std::swap(m->kev[1], m->kev[2 * k + 1]);
        int rc = dyn_step_finish(m, nullptr, false);
        std::swap(m->kev[0], m->kev[2 * k]);
        std::swap(m->kev[1], m->kev[2 * k + 1]);
        m->time_force = false;
        if (rc) return rc;
        if (image) {
            if (int rc2 = ph_fused_image(m, nullptr, nullptr)) return rc2;
        }
        if (m->rec_flags & EDM_REC_AUTO)
            if (int rc2 = rec_step(m)) return rc2;
    }
    EDM_HIP(hipEventRecord(m->ev1, m->stream));
    EDM_HIP(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    EDM_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    double fsum = 0.0;
    for (int k = 0; k < n_steps; ++k) {
        float f = 0.f;
        EDM_HIP(hipEventElapsedTime(&f, m->kev[2 * k], m->kev[2 * k + 1]));
        fsum += f;
    }
    if (total_ms) *total_ms = ms;
    if (force_ms) *force_ms = fsum;
    if (breakdown_ms)
        for (int i = 0; i < 8; ++i) breakdown_ms[i] = 0.0;
    return 0;
}

int edm_time_mark(edm_membrane* m, int32_t end, double* elapsed_ms) {
    if (!m || (end != 0 && end != 1)) return invalid("edm_time_mark: bad argument");
    if (!m->mark0) EDM_HIP(hipEventCreate(&m->mark0));
    if (!m->mark1) EDM_HIP(hipEventCreate(&m->mark1));
    if (!end) {
        EDM_HIP(hipEventRecord(m->mark0, m->stream));
        m->mark_open = true;
        return 0;
    }
    if (!m->mark_open) {
        set_error("edm_time_mark: no window was opened");
        return EDM_ESTATE;
    }
    EDM_HIP(hipEventRecord(m->mark1, m->stream));
    EDM_HIP(hipEventSynchronize(m->mark1));
    m->mark_open = false;
    float ms = 0.f;
    EDM_HIP(hipEventElapsedTime(&ms, m->mark0, m->mark1));
    if (elapsed_ms) *elapsed_ms = ms;
    return 0;
}

}  // extern "C"
