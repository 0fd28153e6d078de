// Photophysics and virtual-microscope kernels of the ED-MRDT path on gfx950:
// bond / proximity FRET rates, the PSS photon kernel as count -> scan ->
// emit (so each photon's arrival draw index is the reference's sequential
// index), the emitter-centred PSF pixel draw, IRF jitter + TCSPC binning,
// and the frame-end detector noise.  A fused per-emitter kernel runs the
// whole photon path with Philox draws and no photon list.
//
// Reference: ed_mrdt/photophysics.py and ed_mrdt/microscope.py.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "edm_device.h"
#include "edm_internal.h"
#include "edm_ensemble.h"

namespace edm {

static constexpr int kB = 256;
static inline int nb(int64_t n, int b = kB) { return (int)std::max<int64_t>(1, (n + b - 1) / b); }

template <typename T>
static int dalloc(T** p, size_t count) {
    EDM_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T) + 16));
    return 0;
}

// ----------------------------------------------------------------------
// FRET rates (photophysics.py:1021-1114)
// ----------------------------------------------------------------------
__global__ void k_fret_zero(int n, double* k_fret, int32_t* acc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    k_fret[t] = 0.0;
    acc[t] = -1;
}

// photophysics.py:529-588, one bond-mode FRET map
__global__ void k_fret_bond(int n_dye, const int32_t* dye_ids, const edm_photo_params* pp, int j,
                            const uint8_t* fl, const int8_t* dye, const int32_t* bond,
                            const int32_t* slot_of, double* k_fret, int32_t* acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int t = slot_of[dye_ids[q]];
    if (!(fl[t] & EDM_FLAG_DYE) || (fl[t] & EDM_FLAG_BLEACHED)) return;
    if (dye[t] != pp->fm_donor[j]) return;
    const int partner = bond[t];
    if (partner < 0) return;
    if (pp->fm_intra[j]) {
        k_fret[q] = pp->fm_kfret[j];
        acc[q] = -1;
    } else {
        const int s = slot_of[partner];
        if ((fl[s] & EDM_FLAG_DYE) && dye[s] == pp->fm_acceptor[j]) {
            k_fret[q] = pp->fm_kfret[j];
            acc[q] = partner;
        }
    }
}

// photophysics.py:592-683, brute-force proximity search (as the reference)
template <typename Real>
__global__ void k_fret_prox(int n_dye, const int32_t* dye_ids, const int32_t* slot_of, int n,
                            const edm_photo_params* pp, int j, const Real* x, const Real* y,
                            const int8_t* sp, const uint8_t* fl, const int8_t* dye, Real Lx,
                            Real Ly, bool px, bool py, double* k_fret, int32_t* acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int t = slot_of[dye_ids[q]];
    const uint8_t f = fl[t];
    if (!(f & EDM_FLAG_DYE) || !(f & EDM_FLAG_ALIVE) || (f & EDM_FLAG_BLEACHED)) return;
    if (dye[t] != pp->fm_donor[j]) return;
    const Real r2 = (Real)pp->fm_radius[j] * (Real)pp->fm_radius[j];
    const int target = pp->fm_target[j];
    const Real xi = x[t], yi = y[t];
    bool found = false;
    for (int u = 0; u < n && !found; ++u) {
        if (u == t) continue;
        if (!(fl[u] & EDM_FLAG_ALIVE) || sp[u] != target) continue;
        Real dx = x[u] - xi, dy = y[u] - yi;
        if (px) {
            if (dx > Lx / (Real)2) dx -= Lx;
            else if (dx < -Lx / (Real)2) dx += Lx;
        }
        if (py) {
            if (dy > Ly / (Real)2) dy -= Ly;
            else if (dy < -Ly / (Real)2) dy += Ly;
        }
        if (dx * dx + dy * dy < r2) found = true;
    }
    k_fret[q] = found ? pp->fm_k_hi[j] : pp->fm_k_lo[j];
    acc[q] = -1;
}

// The same search through a grid of the map's alive targets.  Cells are no
// narrower than the proximity radius (with a margin), so every target within
// the radius of a donor, minimum image included, lies in the donor's cell or
// one of its 8 neighbours; the cell index is taken in FP64 so rounding cannot
// move a particle by a whole cell.  "Some target within the radius" does not
// depend on the order in which targets are met, and the distance test is the
// brute-force kernel's expression on the same values, so both kernels give
// the same answer.
struct TgGrid {
    int ncx, ncy;
    double ix, iy;  // cells per um
    __device__ __forceinline__ int cell(double x, double y) const {
        int cx = (int)floor(x * ix), cy = (int)floor(y * iy);
        cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
        cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
        return cx * ncy + cy;
    }
};

__global__ void k_tg_zero(int ncell, int32_t* cnt, int32_t* fill) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncell) return;
    cnt[c] = 0;
    if (c < ncell) fill[c] = 0;
}

template <typename Real>
__global__ void k_tg_count(int n, const Real* x, const Real* y, const int8_t* sp,
                           const uint8_t* fl, int target, TgGrid g, int32_t* cnt) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    if (!(fl[u] & EDM_FLAG_ALIVE) || sp[u] != target) return;
    atomicAdd(&cnt[g.cell((double)x[u], (double)y[u])], 1);
}

template <typename Real>
__global__ void k_tg_scatter(int n, const Real* x, const Real* y, const int8_t* sp,
                             const uint8_t* fl, int target, TgGrid g, const int32_t* start,
                             int32_t* fill, Real* tx, Real* ty, int32_t* tu) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    if (!(fl[u] & EDM_FLAG_ALIVE) || sp[u] != target) return;
    const Real xu = x[u], yu = y[u];
    const int c = g.cell((double)xu, (double)yu);
    const int k = start[c] + atomicAdd(&fill[c], 1);
    tx[k] = xu;
    ty[k] = yu;
    tu[k] = u;
}

// neighbour cells of c along one axis: distinct indices, wrapped on a
// periodic axis, clipped on a walled one
__device__ __forceinline__ int tg_axis(int c, int nc, bool periodic, int* out) {
    if (periodic && nc < 3) {
        for (int k = 0; k < nc; ++k) out[k] = k;
        return nc;
    }
    int m = 0;
    for (int o = -1; o <= 1; ++o) {
        int v = c + o;
        if (periodic) v = (v + nc) % nc;
        else if (v < 0 || v >= nc) continue;
        out[m++] = v;
    }
    return m;
}

// Some alive target other than slot t within radius r of (xi, yi), through
// the target grid: the donor's cell and its neighbours (start: the grid's
// cell starts).
template <typename Real>
__device__ __forceinline__ bool prox_search(const TgGrid& g, const int32_t* start, const Real* tx,
                                            const Real* ty, const int32_t* tu, int t, Real xi,
                                            Real yi, Real r, Real Lx, Real Ly, bool px, bool py) {
    const Real r2 = r * r;
    const int c = g.cell((double)xi, (double)yi);
    int cxs[3], cys[3];
    const int nx = tg_axis(c / g.ncy, g.ncx, px, cxs);
    const int ny = tg_axis(c % g.ncy, g.ncy, py, cys);
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
            const int cc = cxs[a] * g.ncy + cys[b];
            for (int k = start[cc], e = start[cc + 1]; k < e; ++k) {
                if (tu[k] == t) continue;
                Real dx = tx[k] - xi, dy = ty[k] - yi;
                if (px) {
                    if (dx > Lx / (Real)2) dx -= Lx;
                    else if (dx < -Lx / (Real)2) dx += Lx;
                }
                if (py) {
                    if (dy > Ly / (Real)2) dy -= Ly;
                    else if (dy < -Ly / (Real)2) dy += Ly;
                }
                if (dx * dx + dy * dy < r2) return true;
            }
        }
    return false;
}

template <typename Real>
__global__ void k_fret_prox_grid(int n_dye, const int32_t* dye_ids, const int32_t* slot_of,
                                 const edm_photo_params* pp, int j, const Real* x, const Real* y,
                                 const uint8_t* fl, const int8_t* dye, Real Lx, Real Ly, bool px,
                                 bool py, TgGrid g, const int32_t* start, const Real* tx,
                                 const Real* ty, const int32_t* tu, double* k_fret, int32_t* acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int t = slot_of[dye_ids[q]];
    const uint8_t f = fl[t];
    if (!(f & EDM_FLAG_DYE) || !(f & EDM_FLAG_ALIVE) || (f & EDM_FLAG_BLEACHED)) return;
    if (dye[t] != pp->fm_donor[j]) return;
    const bool found = prox_search<Real>(g, start, tx, ty, tu, t, x[t], y[t],
                                         (Real)pp->fm_radius[j], Lx, Ly, px, py);
    k_fret[q] = found ? pp->fm_k_hi[j] : pp->fm_k_lo[j];
    acc[q] = -1;
}

// ----------------------------------------------------------------------
// PSS kernel (photophysics.py:339-483)
// ----------------------------------------------------------------------
struct PssOut {
    int nd, nf;      // donor photons, acceptor photons actually emitted
    double tau;      // donor arrival-time scale
    int ty;          // dye type
};

// Evaluate one emitter: photon counts and the CTMC transition.  Returns
// false when the particle does not take part (no dye, dormant, bleached).
__device__ __forceinline__ bool pss_eval(const edm_photo_params* pp, uint8_t& fl, int8_t& ctmc,
                                         int8_t dye, double kf, int32_t acc_id, double u_p,
                                         double u_t, PssOut& o) {
    if (!(fl & EDM_FLAG_DYE) || !(fl & EDM_FLAG_ALIVE) || (fl & EDM_FLAG_BLEACHED)) return false;
    const int ty = dye;
    if (ty < 0 || ty >= pp->n_fluoro) return false;
    const int s0 = pp->gs[ty], s1 = pp->es[ty], t1 = pp->t1[ty], bl = pp->bl[ty];
    const int cur = ctmc;
    int pss;
    if (cur == s0 || cur == s1) pss = 0;
    else if (t1 >= 0 && cur == t1) pss = 1;
    else if (bl >= 0 && cur == bl) {
        fl = fl | EDM_FLAG_BLEACHED;
        return false;
    } else pss = 0;
    const bool has_fret = kf > 0.0;
    double md, mf, tau, T00, T01;
    if (has_fret) {
        const double tD = pp->tau[ty];
        tau = tD / (1.0 + kf * tD);
        const double kb = pp->k_bound[ty];
        double frac = 0.0;
        if (kb > 0.0) {
            frac = kf / kb;
            if (frac > 1.0) frac = 1.0;
        }
        md = pp->mean_ph[ty][pss] * (1.0 - frac) + pp->mean_ph_f[ty][pss] * frac;
        mf = pp->mean_fr[ty][pss] * (1.0 - frac) + pp->mean_fr_f[ty][pss] * frac;
        T00 = pp->T_step[ty][pss * 3 + 0] * (1.0 - frac) + pp->T_step_f[ty][pss * 3 + 0] * frac;
        T01 = pp->T_step[ty][pss * 3 + 1] * (1.0 - frac) + pp->T_step_f[ty][pss * 3 + 1] * frac;
    } else {
        md = pp->mean_ph[ty][pss];
        mf = pp->mean_fr[ty][pss];
        tau = pp->tau[ty];
        T00 = pp->T_step[ty][pss * 3 + 0];
        T01 = pp->T_step[ty][pss * 3 + 1];
    }
    const int nd = poisson_lcg(md, u_p);
    int nfr = 0;
    if (mf > 0.0 && has_fret) nfr = poisson_lcg(mf, u_p * 0.5 + 0.25);
    o.nd = nd;
    o.nf = (nfr > 0 && acc_id >= 0) ? nfr : 0;
    o.tau = tau;
    o.ty = ty;
    if (u_t < T00) ctmc = (int8_t)s0;
    else if (u_t < T00 + T01) ctmc = (int8_t)(t1 >= 0 ? t1 : s0);
    else {
        ctmc = (int8_t)(bl >= 0 ? bl : s0);
        fl = fl | EDM_FLAG_BLEACHED;
    }
    return true;
}

__global__ void k_pss_count(int n_dye, const int32_t* dye_ids, const int32_t* slot_of,
                            const edm_photo_params* pp, uint8_t* fl, const int8_t* dye,
                            int8_t* ctmc, const double* k_fret, const int32_t* acc,
                            const double* u_pois, const double* u_trans, int32_t* nd_out,
                            int32_t* nf_out, int64_t* cnt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int it = dye_ids[q];
    const int t = slot_of[it];
    uint8_t f = fl[t];
    int8_t c = ctmc[t];
    PssOut o{0, 0, 0.0, 0};
    const bool on = pss_eval(pp, f, c, dye[t], k_fret[q], acc[q], u_pois[it], u_trans[it], o);
    fl[t] = f;
    ctmc[t] = c;
    nd_out[it] = on ? o.nd : 0;
    nf_out[it] = on ? o.nf : 0;
    cnt[it] = on ? (int64_t)(o.nd + o.nf) : 0;
}

__global__ void k_pss_emit(int n_dye, const int32_t* dye_ids, const edm_photo_params* pp,
                           const int8_t* dye, const int32_t* slot_of, const double* k_fret,
                           const int32_t* acc, const int32_t* nd_in, const int32_t* nf_in,
                           const int64_t* off, const double* u_arr, int64_t n_arr,
                           int64_t max_photons, int32_t* o_pid, double* o_t, int8_t* o_ch,
                           uint8_t* o_fret) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n_dye) return;
    const int it = dye_ids[q];
    const int t = slot_of[it];
    const int nd = nd_in[it], nf = nf_in[it];
    if (nd + nf == 0) return;
    const int ty = dye[t];
    const double kf = k_fret[q];
    const double tau = kf > 0.0 ? pp->tau[ty] / (1.0 + kf * pp->tau[ty]) : pp->tau[ty];
    const int64_t base = off[it];
    for (int k = 0; k < nd; ++k) {
        const int64_t pos = base + k;
        if (pos >= max_photons) break;
        double ta;
        if (pos < n_arr) {
            double u = u_arr[pos];
            if (u < 1e-30) u = 1e-30;
            ta = -tau * log(u);
        } else {
            ta = tau;
        }
        o_pid[pos] = it;
        o_t[pos] = ta;
        o_ch[pos] = (int8_t)ty;
        o_fret[pos] = 0;
    }
    if (nf > 0) {
        const int a = acc[q];
        const int at = dye[slot_of[a]];
        const double tau_a = (at >= 0 && at < pp->n_fluoro) ? pp->tau[at] : tau;
        for (int k = 0; k < nf; ++k) {
            const int64_t pos = base + nd + k;
            if (pos >= max_photons) break;
            double ta;
            if (pos < n_arr) {
                double u = u_arr[pos];
                if (u < 1e-30) u = 1e-30;
                ta = -tau_a * log(u);
            } else {
                ta = tau_a;
            }
            o_pid[pos] = a;
            o_t[pos] = ta;
            o_ch[pos] = (int8_t)at;
            o_fret[pos] = 1;
        }
    }
}

// ----------------------------------------------------------------------
// Emitter-centred PSF pixel draw (microscope.py:523-631)
// ----------------------------------------------------------------------
__device__ __forceinline__ bool psf_pick(const edm_image_params* ip, const double* pcx,
                                         const double* pcy, double x, double y, int ch, double u,
                                         int& opx, int& opy) {
    if (ch < 0) ch = 0;
    if (ch >= ip->n_sigma) ch = 0;
    double sigma = ip->sigma_ch[ch];
    if (sigma <= 0.0) sigma = ip->sigma_ch[0];
    const int R = ip->R, nx = ip->nx, ny = ip->ny;
    const double fx = (x - ip->offx) / ip->pixel - 0.5;
    const double fy = (y - ip->offy) / ip->pixel - 0.5;
    const int i0 = (int)floor(fx + 0.5);
    const int j0 = (int)floor(fy + 0.5);
    if (i0 < -R || i0 >= nx + R) return false;
    if (j0 < -R || j0 >= ny + R) return false;
    const double inv2 = 0.5 / (sigma * sigma);
    double total = 0.0;
    int first_i = -1, first_j = -1;
    for (int di = -R; di <= R; ++di) {
        const int ii = i0 + di;
        if (ii < 0 || ii >= nx) continue;
        const double dx = x - pcx[ii];
        for (int dj = -R; dj <= R; ++dj) {
            const int jj = j0 + dj;
            if (jj < 0 || jj >= ny) continue;
            const double dy = y - pcy[jj];
            total += exp(-(dx * dx + dy * dy) * inv2);
            if (first_i < 0) {
                first_i = ii;
                first_j = jj;
            }
        }
    }
    if (first_i < 0 || total <= 0.0) return false;
    const double uu = u * total;
    double cum = 0.0;
    for (int di = -R; di <= R; ++di) {
        const int ii = i0 + di;
        if (ii < 0 || ii >= nx) continue;
        const double dx = x - pcx[ii];
        for (int dj = -R; dj <= R; ++dj) {
            const int jj = j0 + dj;
            if (jj < 0 || jj >= ny) continue;
            const double dy = y - pcy[jj];
            cum += exp(-(dx * dx + dy * dy) * inv2);
            if (cum >= uu) {
                opx = ii;
                opy = jj;
                return true;
            }
        }
    }
    opx = first_i;
    opy = first_j;
    return true;
}

// Row flags of a single membrane's stacks: one byte per pixel histogram,
// stored right after the stack's n_pix x n_bins counts and set by every
// kernel that adds counts to a row, so that a clear flag means an empty row
// and a frame end reads only the flagged rows.  Concurrent writers all
// store 1.  Every stack, a single membrane's and each replica's, has them.
__host__ __device__ __forceinline__ uint8_t* row_flags(uint32_t* flim, int64_t npix, int nb) {
    return (uint8_t*)(flim + npix * (int64_t)nb);
}

__device__ __forceinline__ void tcspc_add(const edm_image_params* ip, uint32_t* flim,
                                          uint8_t* flags, int ch, int px, int py, double t) {
    if (ch < 0 || ch >= ip->n_ch) return;
    if (px < 0 || px >= ip->nx || py < 0 || py >= ip->ny) return;
    if (t < 0.0) t = 0.0;
    const long long b = (long long)(t / ip->bin_w);
    if (b >= ip->n_bins) return;
    const int64_t q = ((int64_t)ch * ip->nx + px) * ip->ny + py;
    atomicAdd(&flim[q * ip->n_bins + b], 1u);
    if (flags) flags[q] = 1u;
}

// ----------------------------------------------------------------------
// Fused per-emitter photon path with Philox draws
// ----------------------------------------------------------------------
// Bond-mode FRET rate of one donor, all maps in config order (later maps
// overwrite earlier ones, as update_fret_rates does; photophysics.py:529-588).
// Bonds hold particle indices; the acceptor's slot is slot_of[index - kbase].
__device__ __forceinline__ void fret_inline(const edm_photo_params* pp, int t, const uint8_t* fl,
                                            const int8_t* dye, const int32_t* bond,
                                            const int32_t* slot_of, double& kf, int& acc,
                                            int kbase = 0) {
    kf = 0.0;
    acc = -1;
    if (!(fl[t] & EDM_FLAG_DYE) || (fl[t] & EDM_FLAG_BLEACHED)) return;
    const int partner = bond[t];
    if (partner < 0) return;
    for (int j = 0; j < pp->n_fret_maps; ++j) {
        if (pp->fm_proximity[j] || pp->fm_skip[j] || dye[t] != pp->fm_donor[j]) continue;
        if (pp->fm_intra[j]) {
            kf = pp->fm_kfret[j];
            acc = -1;
        } else if (slot_of) {
            const int s = slot_of[partner - kbase];
            if ((fl[s] & EDM_FLAG_DYE) && dye[s] == pp->fm_acceptor[j]) {
                kf = pp->fm_kfret[j];
                acc = partner;
            }
        }
    }
}

// Pixel draw from the emitter's (2R+1)^2 Gaussian window factorised as
// exp(-dx^2 k) exp(-dy^2 k): the same discrete distribution as the
// reference's joint CDF walk (di-major), evaluated with 2(2R+1) exps in
// the working precision.  Windows wider than 31 pixels are rejected when
// the imaging parameters are set; kMax (8 or 31) is the window the launch was
// built for (the weights live in registers).
template <typename Real, int kMax>
__device__ __forceinline__ bool psf_pick_sep(const edm_image_params* ip, Real x, Real y, int ch,
                                             Real u, int& opx, int& opy) {
    const int R = ip->R, nx = ip->nx, ny = ip->ny;
    if (ch < 0) ch = 0;
    if (ch >= ip->n_sigma) ch = 0;
    Real sigma = (Real)ip->sigma_ch[ch];
    if (sigma <= (Real)0) sigma = (Real)ip->sigma_ch[0];
    const Real pix = (Real)ip->pixel;
    const int i0 = (int)floor((x - (Real)ip->offx) / pix - (Real)0.5 + (Real)0.5);
    const int j0 = (int)floor((y - (Real)ip->offy) / pix - (Real)0.5 + (Real)0.5);
    if (i0 < -R || i0 >= nx + R || j0 < -R || j0 >= ny + R) return false;
    if (2 * R + 1 > kMax) return false;  // rejected on the host
    const Real inv2 = (Real)0.5 / (sigma * sigma);
    Real wx[kMax], wy[kMax];
    Real Wx = 0, Wy = 0;
    for (int d = -R; d <= R; ++d) {
        const int ii = i0 + d, jj = j0 + d;
        Real a = 0, b = 0;
        // pixel centres (i + 0.5) * pixel + offset, as the scanner computes them
        if (ii >= 0 && ii < nx) {
            const Real dx = x - (Real)(((double)ii + 0.5) * ip->pixel + ip->offx);
            a = exp(-dx * dx * inv2);
        }
        if (jj >= 0 && jj < ny) {
            const Real dy = y - (Real)(((double)jj + 0.5) * ip->pixel + ip->offy);
            b = exp(-dy * dy * inv2);
        }
        wx[d + R] = a;
        wy[d + R] = b;
        Wx += a;
        Wy += b;
    }
    const Real total = Wx * Wy;
    if (!(total > (Real)0)) return false;
    Real uu = u * total;
    int si = -1;
    for (int d = 0; d < 2 * R + 1; ++d) {
        if (wx[d] <= (Real)0) continue;
        const Real row = wx[d] * Wy;
        si = d;
        if (uu <= row) break;
        uu -= row;
    }
    int sj = -1;
    for (int d = 0; d < 2 * R + 1; ++d) {
        if (wy[d] <= (Real)0) continue;
        const Real cell = wx[si] * wy[d];
        sj = d;
        if (uu <= cell) break;
        uu -= cell;
    }
    if (si < 0 || sj < 0) return false;
    opx = i0 + si - R;
    opy = j0 + sj - R;
    return true;
}

// The photons of one emitter after its PSS step: n = nd donor photons then
// n - nd FRET-sensitised acceptor photons, drawn with the emitter's key and
// the photon's index.
template <typename Real>
struct Emit {
    int n, nd;
    uint32_t key;
    int es_d, ch_d, es_a, ch_a;  // emitting slot and channel of donor / acceptor photons
    Real tau_d, tau_a;
};

// One emitter of the fused photon path with Philox draws: FRET rate, PSS
// counts and CTMC transition.  `key` is the particle's draw key; FRET comes
// from (kf_in, acc_in) when have_kf, else from the bond maps (acceptor
// lookups need slot_of).
template <typename Real>
__device__ __forceinline__ Emit<Real> emit_pss(const edm_photo_params* pp, int t, uint32_t key,
                                               uint8_t* fl, const int8_t* dye, int8_t* ctmc,
                                               const int32_t* bond, const int32_t* slot_of,
                                               bool have_kf, double kf_in, int acc_in, Rng rng,
                                               uint32_t step, int& emitted, int kbase = 0) {
    Emit<Real> E{0, 0, key, t, 0, t, 0, (Real)0, (Real)0};
    uint8_t f = fl[t];
    int8_t c = ctmc[t];
    double kf;
    int aid;
    if (have_kf) {
        kf = kf_in;
        aid = acc_in;
    } else {
        fret_inline(pp, t, fl, dye, bond, slot_of, kf, aid, kbase);
    }
    const U4 r0 = rng.draw(step, key, 0u, kTagPhotonPois);
    const double up = u01d(r0.x, r0.y), ut = u01d(r0.z, r0.w);
    PssOut o{0, 0, 0.0, 0};
    if (pss_eval(pp, f, c, dye[t], kf, aid, up, ut, o)) {
        emitted += o.nd + o.nf;
        E.n = o.nd + o.nf;
        E.nd = o.nd;
        E.ch_d = o.ty;
        E.tau_d = (Real)o.tau;
        if (o.nf > 0) {
            E.es_a = slot_of[aid - kbase];
            E.ch_a = dye[E.es_a];
            E.tau_a = (E.ch_a >= 0 && E.ch_a < pp->n_fluoro) ? (Real)pp->tau[E.ch_a] : (Real)o.tau;
        }
    }
    fl[t] = f;
    ctmc[t] = c;
    return E;
}

// Arrival time, detection, PSF pixel, IRF jitter and TCSPC bin of every
// photon of the wavefront's emitters, spread over its lanes (few emitters
// emit in a step, so walking each emitter's photons in its own lane leaves
// most lanes idle).  Photon k of an emitter uses the draws keyed by
// (step, emitter key, k) whichever lane takes it.  Every lane must call it;
// ip, flim, flags (row flags or nullptr), rng and step are the same for
// the whole wavefront.
template <typename Real, int kWin>
__device__ __forceinline__ void wave_photons(const Emit<Real>& E, const edm_image_params* ip,
                                             const Real* x, const Real* y, uint32_t* flim,
                                             uint8_t* flags, Rng rng, uint32_t step,
                                             int& detected) {
    const int lane = lane_id();
    int incl = E.n;  // photons of lanes 0 .. lane
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    const Real qe = (Real)ip->qe;
    const Real sirf = (Real)ip->sigma_irf;
    for (int p0 = 0; p0 < total; p0 += 64) {
        const int p = p0 + lane;
        // the emitter lane of photon p: lanes whose photons all come before p
        int lo = 0;
#pragma unroll
        for (int s = 32; s > 0; s >>= 1) {
            const int v = __shfl(incl, lo + s - 1);
            if (v <= p) lo += s;
        }
        const int ow = min(lo, 63);
        const int k = p - (__shfl(incl, ow) - __shfl(E.n, ow));
        const uint32_t key = (uint32_t)__shfl((int)E.key, ow);
        const int nd = __shfl(E.nd, ow);
        const int es_d = __shfl(E.es_d, ow), ch_d = __shfl(E.ch_d, ow);
        const int es_a = __shfl(E.es_a, ow), ch_a = __shfl(E.ch_a, ow);
        const Real tau_d = __shfl(E.tau_d, ow), tau_a = __shfl(E.tau_a, ow);
        if (p < total) {
            const bool donor = k < nd;
            const int es = donor ? es_d : es_a, ch = donor ? ch_d : ch_a;
            const Real tau = donor ? tau_d : tau_a;
            const U4 ra = rng.draw(step, key, (uint32_t)k, kTagArrival);
            if (Uni<Real>::second(ra) < qe) {  // detection efficiency
                Real ua = Uni<Real>::first(ra);
                if (ua < (Real)1e-30) ua = (Real)1e-30;
                const Real ta = -tau * log(ua);
                const U4 rp = rng.draw(step, key, (uint32_t)k, kTagPixel);
                int px_, py_;
                if (psf_pick_sep<Real, kWin>(ip, x[es], y[es], ch, Uni<Real>::first(rp), px_,
                                             py_)) {
                    Real z0, z1;
                    normal2<Real>(rng.draw(step, key, (uint32_t)k, kTagJitter), z0, z1);
                    tcspc_add(ip, flim, flags, ch, px_, py_, (double)(ta + sirf * z0));
                    ++detected;
                }
            }
        }
    }
}

// PSF window capacity of a launch: 8 pixels when 2R + 1 <= 8, else 31.
static inline bool psf_small(int R) { return 2 * R + 1 <= 8; }

__device__ __forceinline__ void photon_count(int emitted, int detected, unsigned long long* ctr) {
    unsigned long long e = emitted, d = detected;
    for (int o = 32; o > 0; o >>= 1) {
        e += __shfl_down(e, o);
        d += __shfl_down(d, o);
    }
    if (lane_id() == 0) {
        if (e) atomicAdd(&ctr[6], e);
        if (d) atomicAdd(&ctr[7], d);
    }
}

// Single membrane: one thread per dye carrier.
template <typename Real, int kWin>
__global__ __launch_bounds__(256) void k_pss_image(int n_dye, const int32_t* dye_ids, WlView V, const int32_t* id,
                            const edm_photo_params* pp, const edm_image_params* ip, const Real* x,
                            const Real* y, uint8_t* fl, const int8_t* dye, int8_t* ctmc,
                            const int32_t* bond, const double* k_fret, const int32_t* acc,
                            const int32_t* slot_of, uint32_t* flim, Rng rng, StepRef sref,
                            unsigned long long* ctr) {
    const uint32_t step = sref.get();
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    int emitted = 0, detected = 0;
    Emit<Real> E{0, 0, 0u, 0, 0, 0, 0, (Real)0, (Real)0};
    if (V.list) {
        // this step's dye work list (slots); FRET from the bond maps
        const int m = wl_slice(V);
        if (m == 0) return;  // uniform: an empty slice
        if ((int)threadIdx.x < m) {
            const int t = wl_slot(V);
            E = emit_pss<Real>(pp, t, (uint32_t)id[t], fl, dye, ctmc, bond, slot_of, false, 0.0,
                               -1, rng, step, emitted);
        }
    } else if (q < n_dye) {
        const int it = dye_ids[q];
        E = emit_pss<Real>(pp, slot_of[it], (uint32_t)it, fl, dye, ctmc, bond, slot_of,
                           k_fret != nullptr, k_fret ? k_fret[q] : 0.0, acc ? acc[q] : -1, rng,
                           step, emitted);
    }
    wave_photons<Real, kWin>(E, ip, x, y, flim,
                             row_flags(flim, (int64_t)ip->n_ch * ip->nx * ip->ny, ip->n_bins),
                             rng, step, detected);
    photon_count(emitted, detected, ctr);
}

// ----------------------------------------------------------------------
// Hook-run imaging kernels (microscope.py:1069-1157)
// ----------------------------------------------------------------------
__global__ void k_det_flag(int64_t n, const double* u_qe, double qe, int64_t* flag) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    flag[p] = u_qe[p] < qe ? 1 : 0;
}

__global__ void k_det_compact(int64_t n, const int64_t* flag, const int64_t* pos,
                              const int32_t* pid, uint32_t* key, uint32_t* val) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !flag[p]) return;
    key[pos[p]] = (uint32_t)pid[p];
    val[pos[p]] = (uint32_t)p;
}

template <typename Real>
__global__ void k_assign(int64_t n, const edm_image_params* ip, const double* pcx,
                         const double* pcy, const Real* x, const Real* y, const int32_t* slot_of,
                         const uint32_t* sv_val, const int32_t* pid, const int8_t* ch,
                         const double* u_pix, int64_t* aflag, int32_t* spx, int32_t* spy) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t p = sv_val[q];
    const int s = slot_of[pid[p]];
    int a, b;
    const bool ok = psf_pick(ip, pcx, pcy, (double)x[s], (double)y[s], ch[p], u_pix[q], a, b);
    aflag[q] = ok ? 1 : 0;
    spx[q] = a;
    spy[q] = b;
}

__global__ void k_assign_compact(int64_t n, const int64_t* aflag, const int64_t* apos,
                                 const uint32_t* sv_val, const int32_t* spx, const int32_t* spy,
                                 const double* ph_t, const int8_t* ph_ch, int32_t* apx,
                                 int32_t* apy, double* at, int8_t* ach) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || !aflag[q]) return;
    const int64_t o = apos[q];
    const uint32_t p = sv_val[q];
    apx[o] = spx[q];
    apy[o] = spy[q];
    at[o] = ph_t[p];
    ach[o] = ph_ch[p];
}

__global__ void k_bin_tcspc(int64_t n, const edm_image_params* ip, uint32_t* flim,
                            const int32_t* apx, const int32_t* apy, const double* at,
                            const int8_t* ach, const double* jitter) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    tcspc_add(ip, flim, row_flags(flim, (int64_t)ip->n_ch * ip->nx * ip->ny, ip->n_bins), ach[a],
              apx[a], apy[a], at[a] + jitter[a]);
}

// ----------------------------------------------------------------------
// Frame end (microscope.py:859-920)
// ----------------------------------------------------------------------
// Totals of the pixel histograms q0 .. q0 + 63 (nb bins each), every row
// read coalesced by the whole wavefront (16 B per lane when nb % 4 == 0);
// lane j returns the total of pixel q0 + j (0 past npix).  Wave-uniform
// arguments; every lane must call it.
__device__ __forceinline__ uint32_t wave_pixel_totals(const uint32_t* flim, int64_t q0,
                                                      int64_t npix, int nb) {
    const int lane = lane_id();
    uint32_t mine = 0u;
    const int jn = (int)min<int64_t>(64, npix - q0);
#pragma unroll 4
    for (int j = 0; j < jn; ++j) {
        const uint32_t* row = flim + (q0 + j) * nb;
        uint32_t s = 0u;
        if ((nb & 3) == 0) {
            const uint4* r4 = (const uint4*)row;
            for (int b = lane; b < (nb >> 2); b += 64) {
                const uint4 v = r4[b];
                s += v.x + v.y + v.z + v.w;
            }
        } else {
            for (int b = lane; b < nb; b += 64) s += row[b];
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        mine = lane == j ? s : mine;
    }
    return mine;
}

__global__ void k_pixel_totals(int64_t npix, int nb, const uint32_t* flim, uint32_t* tot) {
    const int64_t q0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~(int64_t)63;
    if (q0 >= npix) return;  // whole wavefronts
    const uint32_t s = wave_pixel_totals(flim, q0, npix, nb);
    const int64_t q = q0 + lane_id();
    if (q < npix) tot[q] = s;
}

__global__ void k_add_counts(int64_t n, const int64_t* idx, uint32_t* flim, int64_t nelem,
                             int nb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t k = idx[i];
    if (k >= 0 && k < nelem) {
        atomicAdd(&flim[k], 1u);
        ((uint8_t*)(flim + nelem))[k / nb] = 1u;
    }
}

// Poisson by sequential inversion (small means: dark counts per pixel).
__device__ __forceinline__ int poisson_inv(double lam, double u) {
    double p = exp(-lam), F = p;
    int k = 0;
    while (u > F && k < 1000) {
        ++k;
        p *= lam / k;
        F += p;
    }
    return k;
}

// Binomial by inversion (Kachitvichyanukul & Schmeiser 1988, BINV) for
// small n*p, normal approximation otherwise.
__device__ __forceinline__ int binomial_draw(int nn, double p, double u, double z) {
    if (nn <= 0 || p <= 0.0) return 0;
    const double mean = nn * p;
    if (mean > 30.0) {
        const double v = mean + sqrt(mean * (1.0 - p)) * z + 0.5;
        long long iv = (long long)v;
        if (iv < 0) iv = 0;
        if (iv > nn) iv = nn;
        return (int)iv;
    }
    const double q = 1.0 - p, s = p / q, a = (nn + 1) * s;
    double r = pow(q, (double)nn);
    int k = 0;
    while (u > r && k < nn) {
        u -= r;
        ++k;
        r *= (a / k - s);
    }
    return k;
}

// One count added to a bin by the thread that owns its row; returns 1 if
// the bin was empty.
__device__ __forceinline__ uint32_t bump(uint32_t* row, int b) {
    const uint32_t o = row[b];
    row[b] = o + 1u;
    return o == 0u;
}

// Dark counts (Poisson, uniform bins) and afterpulses (binomial of the
// pixel's counts after the dark counts, uniform bins) of one pixel
// histogram whose counts before the dark counts total `tot0`, by the one
// thread that owns the row.  Sets the row's flag (if the stack keeps flags)
// when dark counts land and returns how many counts landed on empty bins.
__device__ __forceinline__ uint32_t frame_noise_pixel(int64_t q, int nb, uint32_t* flim,
                                                      uint8_t* flags, uint32_t tot0, double lam,
                                                      double p_ap, Rng rng, uint32_t frame,
                                                      unsigned long long& nd,
                                                      unsigned long long& na) {
    uint32_t* row = flim + q * nb;
    const uint32_t lo = (uint32_t)q, hi = (uint32_t)(q >> 32);
    uint32_t tot = tot0, fresh = 0u;
    if (lam >= 1e-10) {
        const U4 r = rng.draw(frame, lo, hi, kTagDark);
        const int k = poisson_inv(lam, u01d(r.x, r.y));
        for (int i = 0; i < k; ++i) {
            const U4 rb = rng.draw(frame, lo, hi + 0x10000u * (uint32_t)(i + 1), kTagDark);
            fresh += bump(row, (int)(u01d(rb.x, rb.y) * nb) % nb);
        }
        if (k > 0 && flags) flags[q] = 1u;
        nd += (unsigned long long)k;
        tot += (uint32_t)k;
    }
    if (p_ap > 0.0) {
        if (tot) {
            const U4 r = rng.draw(frame, lo, hi, kTagAfterpulse);
            double z0, z1;
            normal2<double>(r, z0, z1);
            const int k = binomial_draw((int)tot, p_ap, u01d(r.x, r.z), z0);
            for (int i = 0; i < k; ++i) {
                const U4 rb =
                    rng.draw(frame, lo, hi + 0x10000u * (uint32_t)(i + 1), kTagAfterpulse);
                fresh += bump(row, (int)(u01d(rb.x, rb.y) * nb) % nb);
            }
            na += (unsigned long long)k;
        }
    }
    return fresh;
}

__device__ __forceinline__ void noise_count(unsigned long long nd, unsigned long long na,
                                            unsigned long long* out) {
    for (int o = 32; o > 0; o >>= 1) {
        nd += __shfl_down(nd, o);
        na += __shfl_down(na, o);
    }
    if (lane_id() == 0) {
        if (nd) atomicAdd(&out[0], nd);
        if (na) atomicAdd(&out[1], na);
    }
}

// One wavefront per 64 pixels: their totals read coalesced, then each
// lane draws its own pixel's noise.  flags: the stack's row flags, or
// nullptr for a stack that keeps none.
__global__ void k_frame_noise(int64_t npix, int nb, uint32_t* flim, uint8_t* flags, double lam,
                              double p_ap, Rng rng, uint32_t frame, unsigned long long* out) {
    const int64_t q0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) & ~(int64_t)63;
    if (q0 >= npix) return;  // whole wavefronts
    const uint32_t tot0 = p_ap > 0.0 ? wave_pixel_totals(flim, q0, npix, nb) : 0u;
    const int64_t q = q0 + lane_id();
    unsigned long long nd = 0, na = 0;
    if (q < npix)
        (void)frame_noise_pixel(q, nb, flim, flags, tot0, lam, p_ap, rng, frame, nd, na);
    noise_count(nd, na, out);
}

// cumulative photon counter of the hook path (the fused kernel counts its own)
__global__ void k_ctr_add(unsigned long long* c, unsigned long long v) {
    if (threadIdx.x == 0) *c += v;
}

__global__ void k_fill_u32(int64_t n, uint32_t* p, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ----------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------
int ph_alloc(edm_membrane* m) {
    const int n = m->n;
    graph_touch(m);
    if (!m->d_pp) EDM_HIP(hipMalloc((void**)&m->d_pp, sizeof(edm_photo_params)));
    EDM_HIP(hipMemcpy(m->d_pp, &m->pp, sizeof(edm_photo_params), hipMemcpyHostToDevice));
    if (!m->k_fret) {
        if (dalloc(&m->k_fret, n) || dalloc(&m->acc, n) || dalloc(&m->ph_nd, n) ||
            dalloc(&m->ph_nf, n) || dalloc(&m->ph_cnt, n) || dalloc(&m->ph_off, n))
            return EDM_EHIP;
        // only dye carriers are ever written: everything else stays zero
        EDM_HIP(hipMemset(m->ph_nd, 0, (size_t)n * 4));
        EDM_HIP(hipMemset(m->ph_nf, 0, (size_t)n * 4));
        EDM_HIP(hipMemset(m->ph_cnt, 0, (size_t)n * 8));
        hipLaunchKernelGGL(k_fret_zero, dim3(nb(n)), dim3(kB), 0, m->stream, n, m->k_fret, m->acc);
    }
    return 0;
}

static int ensure_photon_cap(edm_membrane* m, int64_t need) {
    if (need <= m->ph_cap) return 0;
    graph_touch(m);
    int64_t cap = std::max<int64_t>(need, 1024);
    cap = cap + cap / 2;
    hipFree(m->ph_pid);
    hipFree(m->ph_t);
    hipFree(m->ph_ch);
    hipFree(m->ph_fret);
    hipFree(m->det_flag);
    hipFree(m->det_pos);
    hipFree(m->sv_key);
    hipFree(m->sv_key2);
    hipFree(m->sv_val);
    hipFree(m->sv_val2);
    hipFree(m->asg_flag);
    hipFree(m->asg_pos);
    hipFree(m->asg_px);
    hipFree(m->asg_py);
    hipFree(m->asg_t);
    hipFree(m->asg_ch);
    hipFree(m->sv_px);
    hipFree(m->sv_py);
    if (dalloc(&m->ph_pid, cap) || dalloc(&m->ph_t, cap) || dalloc(&m->ph_ch, cap) ||
        dalloc(&m->ph_fret, cap) || dalloc(&m->det_flag, cap) || dalloc(&m->det_pos, cap) ||
        dalloc(&m->sv_key, cap) || dalloc(&m->sv_key2, cap) || dalloc(&m->sv_val, cap) ||
        dalloc(&m->sv_val2, cap) || dalloc(&m->asg_flag, cap) || dalloc(&m->asg_pos, cap) ||
        dalloc(&m->asg_px, cap) || dalloc(&m->asg_py, cap) || dalloc(&m->asg_t, cap) ||
        dalloc(&m->asg_ch, cap) || dalloc(&m->sv_px, cap) || dalloc(&m->sv_py, cap))
        return EDM_EHIP;
    m->ph_cap = cap;
    return 0;
}

static int scan_i64(edm_membrane* m, const int64_t* in, int64_t* out, int64_t n) {
    size_t need = 0;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)n, m->stream));
    if (need > m->cub_tmp_bytes) {
        graph_touch(m);
        EDM_HIP(hipStreamSynchronize(m->stream));
        EDM_HIP(hipFree(m->cub_tmp));
        m->cub_tmp_bytes = need + 256;
        EDM_HIP(hipMalloc(&m->cub_tmp, m->cub_tmp_bytes));
    }
    size_t tb = m->cub_tmp_bytes;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(m->cub_tmp, tb, in, out, (int)n, m->stream));
    return 0;
}

// total = out[n-1] + in[n-1] read back to the host
static int scan_total(edm_membrane* m, const int64_t* in, const int64_t* out, int64_t n,
                      int64_t* total) {
    if (n <= 0) {
        *total = 0;
        return 0;
    }
    int64_t a = 0, b = 0;
    EDM_HIP(hipMemcpyAsync(&a, in + n - 1, 8, hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipMemcpyAsync(&b, out + n - 1, 8, hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    *total = a + b;
    return 0;
}

// Grid of map j's alive targets and the donors' search through it.  Cell
// sides are at least the radius plus a margin that covers the rounding of
// the distance test (a few ulps of the box side in the position type), and
// the grid holds at most max(n, 64) cells.
template <typename Real>
static int fret_prox_grid(edm_membrane* m, int j) {
    const int n = m->n, nd = m->n_dye;
    hipStream_t st = m->stream;
    const double Lx = m->p.Lx, Ly = m->p.Ly, r = m->pp.fm_radius[j];
    const double ulp = sizeof(Real) == 4 ? 0x1p-23 : 0x1p-52;
    const double cmin = r * 1.001 + 64.0 * ulp * std::max(Lx, Ly);
    long ncx = std::max(1L, std::min(4096L, (long)std::floor(Lx / cmin)));
    long ncy = std::max(1L, std::min(4096L, (long)std::floor(Ly / cmin)));
    const long maxc = std::max<long>(n, 64);
    while (ncx * ncy > maxc) {
        if (ncx >= ncy) ncx = (ncx + 1) / 2;
        else ncy = (ncy + 1) / 2;
    }
    TgGrid g{(int)ncx, (int)ncy, (double)ncx / Lx, (double)ncy / Ly};
    const int ncell = (int)(ncx * ncy);
    if (ncell + 1 > m->tg_cell_cap) {
        graph_touch(m);
        EDM_HIP(hipStreamSynchronize(st));
        hipFree(m->tg_cnt);
        hipFree(m->tg_fill);
        hipFree(m->tg_start);
        m->tg_cnt = m->tg_fill = m->tg_start = nullptr;
        m->tg_cell_cap = 0;
        if (dalloc(&m->tg_cnt, ncell + 1) || dalloc(&m->tg_fill, ncell + 1) ||
            dalloc(&m->tg_start, ncell + 1))
            return EDM_EHIP;
        m->tg_cell_cap = ncell + 1;
    }
    if (n > m->tg_cap) {
        graph_touch(m);
        EDM_HIP(hipStreamSynchronize(st));
        hipFree(m->tg_x);
        hipFree(m->tg_y);
        hipFree(m->tg_u);
        m->tg_x = m->tg_y = nullptr;
        m->tg_u = nullptr;
        m->tg_cap = 0;
        if (dalloc((double**)&m->tg_x, n) || dalloc((double**)&m->tg_y, n) ||
            dalloc(&m->tg_u, n))
            return EDM_EHIP;
        m->tg_cap = n;
    }
    const int target = m->pp.fm_target[j];
    const Real* x = (const Real*)m->a.x;
    const Real* y = (const Real*)m->a.y;
    hipLaunchKernelGGL(k_tg_zero, dim3(nb(ncell + 1)), dim3(kB), 0, st, ncell, m->tg_cnt,
                       m->tg_fill);
    hipLaunchKernelGGL(k_tg_count<Real>, dim3(nb(n)), dim3(kB), 0, st, n, x, y, m->a.sp, m->a.fl,
                       target, g, m->tg_cnt);
    size_t need = 0;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, m->tg_cnt, m->tg_start, ncell + 1,
                                             st));
    if (need > m->cub_tmp_bytes) {
        graph_touch(m);
        EDM_HIP(hipStreamSynchronize(st));
        EDM_HIP(hipFree(m->cub_tmp));
        m->cub_tmp_bytes = need + 256;
        EDM_HIP(hipMalloc(&m->cub_tmp, m->cub_tmp_bytes));
    }
    size_t tb = m->cub_tmp_bytes;
    EDM_HIP(hipcub::DeviceScan::ExclusiveSum(m->cub_tmp, tb, m->tg_cnt, m->tg_start, ncell + 1,
                                             st));
    hipLaunchKernelGGL(k_tg_scatter<Real>, dim3(nb(n)), dim3(kB), 0, st, n, x, y, m->a.sp,
                       m->a.fl, target, g, (const int32_t*)m->tg_start, m->tg_fill,
                       (Real*)m->tg_x, (Real*)m->tg_y, m->tg_u);
    hipLaunchKernelGGL(k_fret_prox_grid<Real>, dim3(nb(nd)), dim3(kB), 0, st, nd,
                       (const int32_t*)m->dye_ids, (const int32_t*)m->a.slot_of,
                       (const edm_photo_params*)m->d_pp, j, x, y, m->a.fl, m->a.dye, (Real)Lx,
                       (Real)Ly, m->p.periodic_x != 0, m->p.periodic_y != 0, g,
                       (const int32_t*)m->tg_start, (const Real*)m->tg_x, (const Real*)m->tg_y,
                       (const int32_t*)m->tg_u, m->k_fret, m->acc);
    return 0;
}

int ph_fret(edm_membrane* m) {
    if (int rc = dyn_ensure_slot_of(m)) return rc;
    const int n = m->n, nd = m->n_dye;
    hipStream_t st = m->stream;
    if (nd == 0) return 0;
    hipLaunchKernelGGL(k_fret_zero, dim3(nb(nd)), dim3(kB), 0, st, nd, m->k_fret, m->acc);
    const edm_photo_params* dpp = (const edm_photo_params*)m->d_pp;
    for (int j = 0; j < m->pp.n_fret_maps; ++j) {
        if (m->pp.fm_proximity[j] || m->pp.fm_skip[j]) continue;
        hipLaunchKernelGGL(k_fret_bond, dim3(nb(nd)), dim3(kB), 0, st, nd,
                           (const int32_t*)m->dye_ids, dpp, j, m->a.fl, m->a.dye, m->a.bond,
                           m->a.slot_of, m->k_fret, m->acc);
    }
    for (int j = 0; j < m->pp.n_fret_maps; ++j) {
        if (!m->pp.fm_proximity[j] || m->pp.fm_skip[j]) continue;
        if (m->fret_search == 0) {
            if (int rc = m->fp64 ? fret_prox_grid<double>(m, j) : fret_prox_grid<float>(m, j))
                return rc;
            continue;
        }
        if (m->fp64)
            hipLaunchKernelGGL(k_fret_prox<double>, dim3(nb(nd)), dim3(kB), 0, st, nd,
                               (const int32_t*)m->dye_ids, (const int32_t*)m->a.slot_of, n, dpp, j,
                               (const double*)m->a.x, (const double*)m->a.y, m->a.sp, m->a.fl,
                               m->a.dye, m->p.Lx, m->p.Ly, m->p.periodic_x != 0,
                               m->p.periodic_y != 0, m->k_fret, m->acc);
        else
            hipLaunchKernelGGL(k_fret_prox<float>, dim3(nb(nd)), dim3(kB), 0, st, nd,
                               (const int32_t*)m->dye_ids, (const int32_t*)m->a.slot_of, n, dpp, j,
                               (const float*)m->a.x, (const float*)m->a.y, m->a.sp, m->a.fl,
                               m->a.dye, (float)m->p.Lx, (float)m->p.Ly, m->p.periodic_x != 0,
                               m->p.periodic_y != 0, m->k_fret, m->acc);
    }
    EDM_HIP(hipGetLastError());
    return 0;
}

int ph_hook(edm_membrane* m, const double* u_pois, const double* u_trans, const double* u_arr,
            int64_t n_arr, int64_t* n_photons) {
    if (int rc = dyn_ensure_slot_of(m)) return rc;
    const int n = m->n;
    hipStream_t st = m->stream;
    if (int rc = ph_fret(m)) return rc;
    // stage the three draw arrays contiguously
    const int64_t tot = 2 * (int64_t)n + n_arr;
    if (int rc = stage_draws(m, nullptr, tot)) return rc;
    EDM_HIP(hipMemcpyAsync(m->d_draws, u_pois, (size_t)n * 8, hipMemcpyHostToDevice, st));
    EDM_HIP(hipMemcpyAsync(m->d_draws + n, u_trans, (size_t)n * 8, hipMemcpyHostToDevice, st));
    if (n_arr > 0)
        EDM_HIP(hipMemcpyAsync(m->d_draws + 2 * (int64_t)n, u_arr, (size_t)n_arr * 8,
                               hipMemcpyHostToDevice, st));
    const edm_photo_params* dpp = (const edm_photo_params*)m->d_pp;
    hipLaunchKernelGGL(k_pss_count, dim3(nb(m->n_dye)), dim3(kB), 0, st, m->n_dye,
                       (const int32_t*)m->dye_ids, (const int32_t*)m->a.slot_of, dpp, m->a.fl,
                       (const int8_t*)m->a.dye, m->a.ctmc,
                       (const double*)m->k_fret, (const int32_t*)m->acc,
                       (const double*)m->d_draws, (const double*)(m->d_draws + n), m->ph_nd,
                       m->ph_nf, m->ph_cnt);
    if (int rc = scan_i64(m, m->ph_cnt, m->ph_off, n)) return rc;
    int64_t total = 0;
    if (int rc = scan_total(m, m->ph_cnt, m->ph_off, n, &total)) return rc;
    const int64_t np = std::min<int64_t>(total, m->pp.max_photons);
    if (int rc = ensure_photon_cap(m, np)) return rc;
    hipLaunchKernelGGL(k_pss_emit, dim3(nb(m->n_dye)), dim3(kB), 0, st, m->n_dye,
                       (const int32_t*)m->dye_ids, dpp, (const int8_t*)m->a.dye, (const int32_t*)m->a.slot_of,
                       (const double*)m->k_fret, (const int32_t*)m->acc,
                       (const int32_t*)m->ph_nd, (const int32_t*)m->ph_nf,
                       (const int64_t*)m->ph_off, (const double*)(m->d_draws + 2 * (int64_t)n),
                       n_arr, (int64_t)m->pp.max_photons, m->ph_pid, m->ph_t, m->ph_ch,
                       m->ph_fret);
    if (np > 0)
        hipLaunchKernelGGL(k_ctr_add, dim3(1), dim3(64), 0, st, m->d_ctr + 6,
                           (unsigned long long)np);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipStreamSynchronize(st));
    m->n_ph = np;
    *n_photons = np;
    return 0;
}

static const edm_image_params* dev_ip(edm_membrane* m) {
    return (const edm_image_params*)(m->pcx + m->ip.nx + m->ip.ny);
}

template <typename Real, int kWin>
static void launch_pss_image(edm_membrane* m, hipStream_t st, dim3 grid, WlView V,
                             const edm_photo_params* dpp, const edm_image_params* d_ip,
                             const double* kf, const int32_t* ac, Rng rng, StepRef step) {
    hipLaunchKernelGGL((k_pss_image<Real, kWin>), grid, dim3(kB), 0, st, m->n_dye,
                       (const int32_t*)m->dye_ids, V, (const int32_t*)m->a.id, dpp, d_ip,
                       (const Real*)m->a.x, (const Real*)m->a.y, m->a.fl, (const int8_t*)m->a.dye,
                       m->a.ctmc, (const int32_t*)m->a.bond, kf, ac, (const int32_t*)m->a.slot_of,
                       m->flim, rng, step, m->d_ctr);
}

int ph_fused_image(edm_membrane* m, int64_t* n_emitted, int64_t* n_detected) {
    if (!m->has_ip) return EDM_ESTATE;
    if (2 * m->ip.R + 1 > 31) {
        set_error("edm_photons_image: PSF window wider than 31 pixels (use the hook path)");
        return EDM_EINVAL;
    }
    hipStream_t st = m->stream;
    if (m->n_dye == 0) {
        if (n_emitted) *n_emitted = 0;
        if (n_detected) *n_detected = 0;
        return 0;
    }
    // proximity FRET needs the global search kernel; bond FRET is computed
    // inside the fused kernel
    bool prox = false;
    for (int j = 0; j < m->pp.n_fret_maps; ++j)
        if (m->pp.fm_proximity[j] && !m->pp.fm_skip[j]) prox = true;
    if (prox)
        if (int rc = ph_fret(m)) return rc;
    const double* kf = prox ? (const double*)m->k_fret : nullptr;
    const int32_t* ac = prox ? (const int32_t*)m->acc : nullptr;
    const edm_photo_params* dpp = (const edm_photo_params*)m->d_pp;
    Rng rng{(uint32_t)(m->p.seed & 0xffffffffu) ^ 0x5bd1e995u, (uint32_t)(m->p.seed >> 32)};
    // the photon step uses the count of completed dynamics steps as its counter
    const StepRef step = step_ref(m);
    const edm_image_params* d_ip = dev_ip(m);
    // after a step the dye carriers' slots are listed in cell order
    const WlView V = prox ? WlView{nullptr, nullptr} : wl_view(m, kWlDye);
    if (!V.list)  // the dye-index path looks carriers up through slot_of
        if (int rc = dyn_ensure_slot_of(m)) return rc;
    const dim3 grid = V.list ? dim3(m->wl_nblk, kWlPer) : dim3(nb(m->n_dye));
    const bool small = psf_small(m->ip.R);
    // beside the step's reactions when none of them can change what the
    // photons read (the plan's ph_overlap_ok): after the step's forces only
    hipStream_t ps = st;
    const bool beside =
        m->force_mark && m->ph_overlap_ok && !prox && !m->use_graphs && !m->graph_rec;
    if (beside) {
        if (!m->ph_stream) EDM_HIP(hipStreamCreate(&m->ph_stream));
        if (!m->ev_photons)
            EDM_HIP(hipEventCreateWithFlags(&m->ev_photons,
                                            hipEventDisableTiming | hipEventDisableSystemFence));
        EDM_HIP(hipStreamWaitEvent(m->ph_stream, m->ev_force, 0));
        ps = m->ph_stream;
    }
    m->force_mark = false;
    if (m->fp64) {
        if (small) launch_pss_image<double, 8>(m, ps, grid, V, dpp, d_ip, kf, ac, rng, step);
        else launch_pss_image<double, 31>(m, ps, grid, V, dpp, d_ip, kf, ac, rng, step);
    } else {
        if (small) launch_pss_image<float, 8>(m, ps, grid, V, dpp, d_ip, kf, ac, rng, step);
        else launch_pss_image<float, 31>(m, ps, grid, V, dpp, d_ip, kf, ac, rng, step);
    }
    if (beside) {
        // everything queued on the main stream from here on follows the photons
        EDM_HIP(hipEventRecord(m->ev_photons, m->ph_stream));
        EDM_HIP(hipStreamWaitEvent(st, m->ev_photons, 0));
    }
    EDM_HIP(hipGetLastError());
    if (n_emitted || n_detected) {
        unsigned long long c[2];
        EDM_HIP(hipMemcpyAsync(c, m->d_ctr + 6, 16, hipMemcpyDeviceToHost, st));
        EDM_HIP(hipStreamSynchronize(st));
        if (n_emitted) *n_emitted = (int64_t)c[0];
        if (n_detected) *n_detected = (int64_t)c[1];
    }
    return 0;
}

// Imaging buffers: flim stack, pixel centres and the parameter block live
// in one allocation after the pixel-centre arrays.
// 32-bit words of a stack allocation: the n_pix x n_bins counts, then one
// flag byte per pixel row (row_flags), padded to 256 bytes
static size_t stack_words(const edm_image_params& ip) {
    const size_t npix = (size_t)ip.n_ch * ip.nx * ip.ny;
    return npix * ip.n_bins + (npix + 255) / 256 * 64;
}

int img_alloc(edm_membrane* m) {
    graph_touch(m);
    const edm_image_params& ip = m->ip;
    const size_t elems = (size_t)ip.n_ch * ip.nx * ip.ny * ip.n_bins;
    if (elems != m->flim_elems || stack_words(ip) != m->flim_words) {
        frames_free(m);  // a pending streamed frame of the old geometry is dropped
        hipFree(m->flim);
        hipFree(m->pix_tot);
        m->flim = nullptr;
        if (dalloc(&m->flim, stack_words(ip)) ||
            dalloc(&m->pix_tot, (size_t)ip.n_ch * ip.nx * ip.ny))
            return EDM_EHIP;
        m->flim_elems = elems;
        m->flim_words = stack_words(ip);
        if (int rc = frame_reset(m)) return rc;
    }
    hipFree(m->pcx);
    const size_t ip_words = (sizeof(edm_image_params) + 7) / 8;
    if (dalloc(&m->pcx, (size_t)ip.nx + ip.ny + ip_words)) return EDM_EHIP;
    m->pcy = m->pcx + ip.nx;
    std::vector<double> host((size_t)ip.nx + ip.ny + ip_words, 0.0);
    for (int i = 0; i < ip.nx; ++i) host[i] = ((double)i + 0.5) * ip.pixel + ip.offx;
    for (int j = 0; j < ip.ny; ++j) host[ip.nx + j] = ((double)j + 0.5) * ip.pixel + ip.offy;
    memcpy(host.data() + ip.nx + ip.ny, &ip, sizeof(edm_image_params));
    EDM_HIP(hipMemcpy(m->pcx, host.data(), host.size() * 8, hipMemcpyHostToDevice));
    return 0;
}

int img_detect(edm_membrane* m, const double* u_qe, int64_t n, int64_t* n_surv) {
    hipStream_t st = m->stream;
    if (n != m->n_ph) return EDM_EINVAL;
    if (n == 0) {
        m->n_surv = 0;
        *n_surv = 0;
        return 0;
    }
    if (int rc = stage_draws(m, u_qe, n)) return rc;
    hipLaunchKernelGGL(k_det_flag, dim3(nb(n)), dim3(kB), 0, st, n, (const double*)m->d_draws,
                       m->ip.qe, m->det_flag);
    if (int rc = scan_i64(m, m->det_flag, m->det_pos, n)) return rc;
    int64_t ns = 0;
    if (int rc = scan_total(m, m->det_flag, m->det_pos, n, &ns)) return rc;
    hipLaunchKernelGGL(k_det_compact, dim3(nb(n)), dim3(kB), 0, st, n,
                       (const int64_t*)m->det_flag, (const int64_t*)m->det_pos,
                       (const int32_t*)m->ph_pid, m->sv_key, m->sv_val);
    if (ns > 0) {
        size_t need = 0;
        EDM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, m->sv_key, m->sv_key2, m->sv_val,
                                                   m->sv_val2, (int)ns, 0, 32, st));
        if (need > m->cub_tmp_bytes) {
            graph_touch(m);
            EDM_HIP(hipStreamSynchronize(st));
            EDM_HIP(hipFree(m->cub_tmp));
            m->cub_tmp_bytes = need + 256;
            EDM_HIP(hipMalloc(&m->cub_tmp, m->cub_tmp_bytes));
        }
        size_t tb = m->cub_tmp_bytes;
        EDM_HIP(hipcub::DeviceRadixSort::SortPairs(m->cub_tmp, tb, m->sv_key, m->sv_key2,
                                                   m->sv_val, m->sv_val2, (int)ns, 0, 32, st));
    }
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipStreamSynchronize(st));
    m->n_surv = ns;
    *n_surv = ns;
    return 0;
}

int img_assign(edm_membrane* m, const double* u_pix, int64_t n, int64_t* n_asg) {
    if (int rc = dyn_ensure_slot_of(m)) return rc;
    hipStream_t st = m->stream;
    if (n != m->n_surv) return EDM_EINVAL;
    if (n == 0) {
        m->n_asg = 0;
        *n_asg = 0;
        return 0;
    }
    if (int rc = stage_draws(m, u_pix, n)) return rc;
    if (m->fp64)
        hipLaunchKernelGGL(k_assign<double>, dim3(nb(n)), dim3(kB), 0, st, n, dev_ip(m),
                           (const double*)m->pcx, (const double*)m->pcy, (const double*)m->a.x,
                           (const double*)m->a.y, (const int32_t*)m->a.slot_of,
                           (const uint32_t*)m->sv_val2, (const int32_t*)m->ph_pid,
                           (const int8_t*)m->ph_ch, (const double*)m->d_draws, m->asg_flag,
                           m->sv_px, m->sv_py);
    else
        hipLaunchKernelGGL(k_assign<float>, dim3(nb(n)), dim3(kB), 0, st, n, dev_ip(m),
                           (const double*)m->pcx, (const double*)m->pcy, (const float*)m->a.x,
                           (const float*)m->a.y, (const int32_t*)m->a.slot_of,
                           (const uint32_t*)m->sv_val2, (const int32_t*)m->ph_pid,
                           (const int8_t*)m->ph_ch, (const double*)m->d_draws, m->asg_flag,
                           m->sv_px, m->sv_py);
    if (int rc = scan_i64(m, m->asg_flag, m->asg_pos, n)) return rc;
    int64_t na = 0;
    if (int rc = scan_total(m, m->asg_flag, m->asg_pos, n, &na)) return rc;
    hipLaunchKernelGGL(k_assign_compact, dim3(nb(n)), dim3(kB), 0, st, n,
                       (const int64_t*)m->asg_flag, (const int64_t*)m->asg_pos,
                       (const uint32_t*)m->sv_val2, (const int32_t*)m->sv_px,
                       (const int32_t*)m->sv_py, (const double*)m->ph_t, (const int8_t*)m->ph_ch,
                       m->asg_px, m->asg_py, m->asg_t, m->asg_ch);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipStreamSynchronize(st));
    m->n_asg = na;
    *n_asg = na;
    return 0;
}

int img_bin(edm_membrane* m, const double* jitter, int64_t n) {
    hipStream_t st = m->stream;
    if (n != m->n_asg) return EDM_EINVAL;
    if (n == 0) return 0;
    if (int rc = stage_draws(m, jitter, n)) return rc;
    hipLaunchKernelGGL(k_bin_tcspc, dim3(nb(n)), dim3(kB), 0, st, n, dev_ip(m), m->flim,
                       (const int32_t*)m->asg_px, (const int32_t*)m->asg_py,
                       (const double*)m->asg_t, (const int8_t*)m->asg_ch,
                       (const double*)m->d_draws);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipStreamSynchronize(st));
    return 0;
}

int frame_totals(edm_membrane* m, uint32_t* totals) {
    const int64_t npix = (int64_t)m->ip.n_ch * m->ip.nx * m->ip.ny;
    hipLaunchKernelGGL(k_pixel_totals, dim3(nb(npix)), dim3(kB), 0, m->stream, npix, m->ip.n_bins,
                       (const uint32_t*)m->flim, m->pix_tot);
    EDM_HIP(hipMemcpyAsync(totals, m->pix_tot, (size_t)npix * 4, hipMemcpyDeviceToHost, m->stream));
    EDM_HIP(hipStreamSynchronize(m->stream));
    return 0;
}

int frame_add(edm_membrane* m, const int64_t* flat, int64_t n) {
    if (n <= 0) return 0;
    int64_t* d = nullptr;
    EDM_HIP(hipMalloc(&d, (size_t)n * 8));
    EDM_HIP(hipMemcpyAsync(d, flat, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_add_counts, dim3(nb(n)), dim3(kB), 0, m->stream, n, (const int64_t*)d,
                       m->flim, (int64_t)m->flim_elems, m->ip.n_bins);
    EDM_HIP(hipStreamSynchronize(m->stream));
    EDM_HIP(hipFree(d));
    return 0;
}

int frame_noise(edm_membrane* m, int64_t* n_dark, int64_t* n_ap) {
    const int64_t npix = (int64_t)m->ip.n_ch * m->ip.nx * m->ip.ny;
    Rng rng{(uint32_t)(m->p.seed & 0xffffffffu) ^ 0x2545F491u, (uint32_t)(m->p.seed >> 32)};
    unsigned long long* out = (unsigned long long*)(m->d_scal + 8);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, m->stream, (int64_t)4, (uint32_t*)out, 0u);
    hipLaunchKernelGGL(k_frame_noise, dim3(nb(npix)), dim3(kB), 0, m->stream, npix, m->ip.n_bins,
                       m->flim, (uint8_t*)(m->flim + m->flim_elems), m->ip.dark_lambda,
                       m->ip.p_ap, rng, m->step, out);
    EDM_HIP(hipGetLastError());
    if (n_dark || n_ap) {
        unsigned long long c[2];
        EDM_HIP(hipMemcpyAsync(c, out, 16, hipMemcpyDeviceToHost, m->stream));
        EDM_HIP(hipStreamSynchronize(m->stream));
        if (n_dark) *n_dark = (int64_t)c[0];
        if (n_ap) *n_ap = (int64_t)c[1];
    }
    return 0;
}

int frame_reset(edm_membrane* m) {
    if (!m->flim) return 0;
    const int64_t n = (int64_t)m->flim_words;  // counts and row flags
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min(nb(n), 8192)), dim3(kB), 0, m->stream, n, m->flim,
                       0u);
    EDM_HIP(hipGetLastError());
    return 0;
}

// Sum and nonzero-bin count of one row, read coalesced by the wavefront
// (16 B per lane when nb % 4 == 0); wave-uniform arguments, every lane calls.
__device__ __forceinline__ void wave_row_sum_nz(const uint32_t* row, int nb, uint32_t& sum,
                                                uint32_t& nz) {
    const int lane = lane_id();
    uint32_t s = 0u, c = 0u;
    if ((nb & 3) == 0) {
        const uint4* r4 = (const uint4*)row;
        for (int b = lane; b < (nb >> 2); b += 64) {
            const uint4 v = r4[b];
            s += v.x + v.y + v.z + v.w;
            c += (v.x != 0u) + (v.y != 0u) + (v.z != 0u) + (v.w != 0u);
        }
    } else {
        for (int b = lane; b < nb; b += 64) {
            const uint32_t v = row[b];
            s += v;
            c += v != 0u;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o);
        c += __shfl_xor(c, o);
    }
    sum = s;
    nz = c;
}

#include "edm_ens_ph.h"

// ----------------------------------------------------------------------
// Streamed frames of one membrane (microscope.py:1159-1203 for every frame,
// the reference keeping them all, pipeline.py:645).  edm_frame_collect ends
// the current frame: the steps that follow write their photons into a
// second stack, while a second stream gives the ended stack its detector
// noise (the same draws as edm_frame_noise at this step), compacts it to
// (flat index, count) pairs in index order, and resets it.  edm_frame_fetch
// waits for that work and copies the pairs out.  One frame may be pending.
// ----------------------------------------------------------------------
// Both passes run one wavefront per 64 pixel rows and touch only the rows
// whose flag is set (at 1024 x 1024 px and ~3.6e5 counts a frame, ~30 % of
// them), so a frame end moves about a fifth of the stack instead of reading
// it whole three times and writing it once.

// First pass: the flagged rows' totals (for the afterpulses) and nonzero
// bins, then every pixel's detector noise drawn as edm_frame_noise draws it
// (frame_noise_pixel), and the wavefront's count of nonzero bins after it.
__global__ __launch_bounds__(256) void k_fr_noise_count(int64_t npix, int nb, uint32_t* flim,
                                                        double lam, double p_ap, Rng rng,
                                                        uint32_t frame,
                                                        unsigned long long* noise,
                                                        uint32_t* wcnt) {
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t q0 = w << 6;
    if (q0 >= npix) return;  // whole wavefronts
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
    unsigned long long nd = 0, na = 0;
    if (in) nz += frame_noise_pixel(q, nb, flim, flags, tot0, lam, p_ap, rng, frame, nd, na);
    noise_count(nd, na, noise);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nz += __shfl_xor(nz, o);
    if (lane == 0) wcnt[w] = nz;
}

// Second pass: the flagged rows' nonzero bins written as (flat index,
// count) pairs from the wavefront's scanned offset, in index order; the
// bins written are zeroed and the flags cleared, which resets the stack.
__global__ __launch_bounds__(256) void k_fr_emit(int64_t npix, int nb, uint32_t* flim,
                                                 const uint32_t* __restrict__ off, uint32_t* oidx,
                                                 uint32_t* ocnt) {
    const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t q0 = w << 6;
    if (q0 >= npix) return;  // whole wavefronts
    uint8_t* flags = row_flags(flim, npix, nb);
    const int lane = lane_id();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int64_t q = q0 + lane;
    const bool in = q < npix;
    const bool mine = in && flags[q] != 0u;
    uint32_t base = off[w];
    for (unsigned long long act = __ballot(mine); act; act &= act - 1ull) {
        const int j = __builtin_ctzll(act);
        uint32_t* row = flim + (q0 + j) * nb;
        const uint32_t r0 = (uint32_t)((q0 + j) * nb);
        if ((nb & 3) == 0) {
            uint4* r4 = (uint4*)row;
            const int n4 = nb >> 2;
            for (int b0 = 0; b0 < n4; b0 += 64) {
                const int b = b0 + lane;
                const uint4 v = b < n4 ? r4[b] : make_uint4(0u, 0u, 0u, 0u);
                const unsigned long long m0 = __ballot(v.x != 0u), m1 = __ballot(v.y != 0u),
                                         m2 = __ballot(v.z != 0u), m3 = __ballot(v.w != 0u);
                uint32_t pos = base + (uint32_t)(__popcll(m0 & lt) + __popcll(m1 & lt) +
                                                 __popcll(m2 & lt) + __popcll(m3 & lt));
                const uint32_t i0 = r0 + 4u * (uint32_t)b;
                const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (vv[k]) {
                        oidx[pos] = i0 + k;
                        ocnt[pos] = vv[k];
                        ++pos;
                    }
                if ((v.x | v.y | v.z | v.w) != 0u) r4[b] = make_uint4(0u, 0u, 0u, 0u);
                base += (uint32_t)(__popcll(m0) + __popcll(m1) + __popcll(m2) + __popcll(m3));
            }
        } else {
            for (int b0 = 0; b0 < nb; b0 += 64) {
                const int b = b0 + lane;
                const uint32_t v = b < nb ? row[b] : 0u;
                const unsigned long long mk = __ballot(v != 0u);
                if (v) {
                    const uint32_t pos = base + (uint32_t)__popcll(mk & lt);
                    oidx[pos] = r0 + (uint32_t)b;
                    ocnt[pos] = v;
                    row[b] = 0u;
                }
                base += (uint32_t)__popcll(mk);
            }
        }
    }
    if (mine) flags[q] = 0u;
}

// the detected-photon counter at the frame's end, beside its noise counts
__global__ void k_fr_meta(const unsigned long long* __restrict__ ctr, unsigned long long* meta) {
    if (blockIdx.x == 0 && threadIdx.x == 0) meta[2] = ctr[7];
}

void frames_free(edm_membrane* m) {
    if (m->fr_stream) (void)hipStreamSynchronize(m->fr_stream);
    void* dev[] = {m->flim_alt, m->fr_blk, m->fr_idx, m->fr_cnt, m->fr_meta};
    for (void* p : dev)
        if (p) (void)hipFree(p);
    m->flim_alt = nullptr;
    m->fr_blk = m->fr_idx = m->fr_cnt = nullptr;
    m->fr_meta = nullptr;
    m->fr_ticket = -1;
    m->fr_nwave = 0;
}

int frame_collect(edm_membrane* m, int32_t* ticket) {
    if (m->fr_ticket >= 0) {
        set_error("edm_frame_collect: a collected frame is pending; fetch it first");
        return EDM_ESTATE;
    }
    const int64_t n = (int64_t)m->flim_elems;
    if (n >= (int64_t)UINT32_MAX) {
        set_error("edm_frame_collect: the stack has more than 2^32 bins");
        return EDM_EINVAL;
    }
    const int64_t npix = (int64_t)m->ip.n_ch * m->ip.nx * m->ip.ny;
    const int64_t nwave = (npix + 63) / 64;
    hipStream_t st = m->stream;
    if (!m->fr_stream) EDM_HIP(hipStreamCreate(&m->fr_stream));
    if (!m->fr_photons) EDM_HIP(hipEventCreateWithFlags(&m->fr_photons, hipEventDisableTiming));
    if (!m->fr_done) EDM_HIP(hipEventCreateWithFlags(&m->fr_done, hipEventDisableTiming));
    if (!m->flim_alt) {
        graph_touch(m);
        m->fr_nwave = (int)nwave;
        if (dalloc(&m->flim_alt, m->flim_words) || dalloc(&m->fr_blk, (size_t)nwave + 1) ||
            dalloc(&m->fr_idx, (size_t)n) || dalloc(&m->fr_cnt, (size_t)n) ||
            dalloc(&m->fr_meta, 4))
            return EDM_EHIP;
        hipLaunchKernelGGL(k_fill_u32, dim3(std::min(nb((int64_t)m->flim_words), 8192)), dim3(kB),
                           0, st, (int64_t)m->flim_words, m->flim_alt, 0u);
    }
    // kernel arguments of recorded step graphs name the stack being swapped
    graph_touch(m);
    uint32_t* ended = m->flim;
    EDM_HIP(hipEventRecord(m->fr_photons, st));
    EDM_HIP(hipStreamWaitEvent(m->fr_stream, m->fr_photons, 0));
    std::swap(m->flim, m->flim_alt);
    hipStream_t fs = m->fr_stream;
    Rng rng{(uint32_t)(m->p.seed & 0xffffffffu) ^ 0x2545F491u, (uint32_t)(m->p.seed >> 32)};
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, fs, (int64_t)8, (uint32_t*)m->fr_meta,
                       0u);
    hipLaunchKernelGGL(k_fr_meta, dim3(1), dim3(1), 0, fs, (const unsigned long long*)m->d_ctr,
                       m->fr_meta);
    const int nblk = (int)((nwave + 3) / 4);
    hipLaunchKernelGGL(k_fr_noise_count, dim3(nblk), dim3(256), 0, fs, npix, m->ip.n_bins, ended,
                       m->ip.dark_lambda, m->ip.p_ap, rng, m->step, m->fr_meta, m->fr_blk);
    hipLaunchKernelGGL(k_sp_scan, dim3(1), dim3(1024), 0, fs, m->fr_blk, (int)nwave);
    hipLaunchKernelGGL(k_fr_emit, dim3(nblk), dim3(256), 0, fs, npix, m->ip.n_bins, ended,
                       (const uint32_t*)m->fr_blk, m->fr_idx, m->fr_cnt);
    EDM_HIP(hipGetLastError());
    EDM_HIP(hipEventRecord(m->fr_done, fs));
    // the steps after this one write into the other stack: zeroed when it was
    // allocated, or reset by the previous collect, whose fetch waited for it
    m->fr_ticket = m->fr_next++;
    *ticket = m->fr_ticket;
    return 0;
}

int frame_fetch(edm_membrane* m, int32_t ticket, int64_t* meta, uint32_t* idx, uint32_t* cnt,
                int64_t cap, int64_t* total) {
    if (ticket < 0 || ticket != m->fr_ticket) {
        set_error("edm_frame_fetch: no pending frame with that ticket");
        return EDM_ESTATE;
    }
    EDM_HIP(hipEventSynchronize(m->fr_done));
    uint32_t tot32 = 0;
    unsigned long long mt[3] = {0, 0, 0};
    EDM_HIP(hipMemcpyAsync(&tot32, m->fr_blk + m->fr_nwave, 4, hipMemcpyDeviceToHost,
                           m->fr_stream));
    EDM_HIP(hipMemcpyAsync(mt, m->fr_meta, sizeof(mt), hipMemcpyDeviceToHost, m->fr_stream));
    EDM_HIP(hipStreamSynchronize(m->fr_stream));
    const int64_t tot = tot32;
    if (total) *total = tot;
    if (meta)
        for (int k = 0; k < 3; ++k) meta[k] = (int64_t)mt[k];
    if (idx || cnt) {
        if (cap < tot) {
            set_error("edm_frame_fetch: output holds " + std::to_string(cap) +
                      " entries, the frame has " + std::to_string(tot));
            return EDM_EINVAL;
        }
        if (tot > 0) {
            if (idx)
                EDM_HIP(hipMemcpyAsync(idx, m->fr_idx, (size_t)tot * 4, hipMemcpyDeviceToHost,
                                       m->fr_stream));
            if (cnt)
                EDM_HIP(hipMemcpyAsync(cnt, m->fr_cnt, (size_t)tot * 4, hipMemcpyDeviceToHost,
                                       m->fr_stream));
            EDM_HIP(hipStreamSynchronize(m->fr_stream));
        }
        // a call without output arrays only reports the sizes; one with them
        // releases the frame for the next collect
        m->fr_ticket = -1;
    }
    return 0;
}

}  // namespace edm
