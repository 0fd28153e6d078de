// Fused per-particle reaction passes (included by edm_dynamics.hip).
//
// Detachment (dynamics.py:1462-1555), conversion (dynamics.py:1566-1657)
// and recruitment-candidate flagging (dynamics.py:1230-1458) each look at
// one particle at a time.  A run of such passes can share one kernel in
// which every thread applies the passes to its own particle in the
// reference's order, provided no pass reads a density count that an
// earlier pass of the run changes (the run is split there on the host).
// A recruitment's index-order cap needs all its candidates before any is
// accepted (k_recruit_apply), so nothing after it in its run may depend on
// it; a run may hold a few recruitments of different species, each with its
// own candidate buffer.  With the step's
// work list of the species the passes can touch a thread takes one listed
// slot instead of one slot of the whole state.

struct PassDev {
    int kind, a, b, rxi;
    double prob;
    FbDev f1, f2;
    int64_t draw_off;
    int deferred;
    uint32_t reach;  // species a particle may hold before the run and still
                     // reach this pass as species a (through earlier conversions)
    double* cdf;  // RECRUIT, density-biased placement: the CDF over f1's grid
                  // cells, written by the CDF kernel before the run (nullptr:
                  // uniform placement)
    double sigma;       // its Gaussian offset around the chosen cell centre
    int rec;            // RECRUIT: its candidate buffer (0 .. kMaxRecruits - 1)
};

static constexpr int kMaxPasses = 8;

struct PassGroupDev {
    int n;
    uint32_t species;  // union of the passes' reach masks
    PassDev p[kMaxPasses];
};

// Recruitments of one run (each with its own candidate buffer).
static constexpr int kMaxRecruits = 2;

// A recruitment candidate of this thread's particle: at most one per run
// (the run's recruitments are of different species, and a particle
// recruited or not keeps its species through the run).
template <typename Real>
struct Cand {
    bool on;
    int it;
    Real x, y;
    int rec;  // candidate buffer
};

// Appends the wavefront's candidates with one atomic per buffer (the order
// of candidates in a buffer does not matter: the cap is applied by particle
// index).  Buffer r starts at r * stride entries and counts in nc[r *
// nc_stride].  Every lane of the wavefront must call it; the buffers must
// be the same for the whole wavefront.
template <typename Real>
__device__ __forceinline__ void append_candidates(const Cand<Real>& c, int t, int32_t* cid,
                                                  int32_t* cslot, double* cxo, double* cyo,
                                                  unsigned int* nc, int64_t stride = 0,
                                                  int nc_stride = 1) {
    if (!__ballot(c.on)) return;  // uniform
    const int lane = lane_id();
#pragma unroll
    for (int r = 0; r < kMaxRecruits; ++r) {
        const bool mine = c.on && c.rec == r;
        const unsigned long long b = __ballot(mine);
        if (!b) continue;  // uniform
        const int leader = __ffsll((long long)b) - 1;
        unsigned int base = 0u;
        if (lane == leader) base = atomicAdd(nc + r * nc_stride, (unsigned int)__popcll(b));
        base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);
        if (mine) {
            const int64_t k =
                r * stride + base + (unsigned int)__popcll(b & ((1ull << lane) - 1ull));
            cid[k] = c.it;
            cslot[k] = t;
            cxo[k] = (double)c.x;
            cyo[k] = (double)c.y;
        }
    }
}

// Cell chosen by a uniform r on a CDF of nc cells: the first cell whose
// CDF value is not below r, found by bisection (dynamics.py:1163-1172).
__device__ __forceinline__ int cdf_cell(const double* __restrict__ cdf, int nc, double r) {
    int lo = 0, hi = nc - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Density-biased recruitment site (dynamics.py:1162-1184): the centre of the
// cell the CDF draw picks, plus sigma times two standard normals, wrapped
// into the box.
template <typename Real>
__device__ __forceinline__ void local_site(const PassDev& d, double r, Real o0, Real o1, Real Lx,
                                           Real Ly, Real& px, Real& py) {
    const int c = cdf_cell(d.cdf, d.f1.gnx * d.f1.gny, r);
    const int gx = c / d.f1.gny, gy = c - gx * d.f1.gny;
    const Real cx = ((Real)gx + (Real)0.5) * (Real)d.f1.gsx + o0 * (Real)d.sigma;
    const Real cy = ((Real)gy + (Real)0.5) * (Real)d.f1.gsy + o1 * (Real)d.sigma;
    px = py_mod(cx, Lx);
    py = py_mod(cy, Ly);
}

// Applies a run of passes to the particle in slot t (species s0, known to
// be in the run's species set).  Draw keys use the particle index minus
// kbase; hook draws are indexed by the particle index.  kLocal: the run may
// end with a density-biased recruitment (its code is compiled only then).
template <typename Real, bool kLocal = false>
__device__ __forceinline__ void passes_particle(int t, int s0, int n, const int32_t* __restrict__ id,
                                                int kbase, const Post<Real>& P,
                                                const PassGroupDev& grp,
                                                const double* __restrict__ draws, Rng rng,
                                                uint32_t step, Real Lx, Real Ly,
                                                const Grids<Real>& G, int n_grids,
                                                Cand<Real>& cand, int& det_hits, int& conv_hits) {
    {
        uint8_t f = P.fl[t];
        const uint8_t f0 = f;
        int8_t s = (int8_t)s0;
        const int it = id[t];
        const uint32_t key = (uint32_t)(it - kbase);
        const Real x = P.xn[t], y = P.yn[t];
        // Phase 1: every pass this particle can reach gets its uniform, its
        // position (the particle's, or a recruitment's drawn site) and its
        // density counts up front, so the loads of all passes are in flight
        // together.  No pass of a run reads a count that another pass of the
        // run changes, so reading them early sees the same values.
        double u[kMaxPasses];
        Real px[kMaxPasses], py[kMaxPasses];
        int k1[kMaxPasses], k2[kMaxPasses];
#pragma unroll
        for (int k = 0; k < kMaxPasses; ++k) {
            u[k] = 1.0;
            px[k] = x;
            py[k] = y;
            k1[k] = 0;
            k2[k] = 0;
            if (k >= grp.n) continue;
            const PassDev& d = grp.p[k];
            if (!((d.reach >> s0) & 1u)) continue;
            if (d.kind == EDM_RX_RECRUIT) {
                Real ux, uy;
                if (draws) {
                    u[k] = draws[d.draw_off + it];
                    ux = (Real)draws[d.draw_off + n + 2 * (int64_t)it];
                    uy = (Real)draws[d.draw_off + n + 2 * (int64_t)it + 1];
                } else {
                    const U4 r = rng.draw(step, key, 0u, (kTagRecruitAcc << 16) | d.rxi);
                    u[k] = (double)Uni<Real>::first(r);
                    const U4 w = rng.draw(step, key, 1u, (kTagRecruitPos << 16) | d.rxi);
                    ux = Uni<Real>::first(w);
                    uy = Uni<Real>::second(w);
                }
                if (kLocal && d.cdf) {
                    // the uniform positions are drawn but unused here, as
                    // the reference draws them in this mode too
                    double rl;
                    Real o0, o1;
                    if (draws) {
                        rl = draws[d.draw_off + 3 * (int64_t)n + it];
                        o0 = (Real)draws[d.draw_off + 4 * (int64_t)n + 2 * (int64_t)it];
                        o1 = (Real)draws[d.draw_off + 4 * (int64_t)n + 2 * (int64_t)it + 1];
                    } else {
                        const U4 w = rng.draw(step, key, 0u, (kTagRecruitLoc << 16) | d.rxi);
                        rl = u01d(w.x, w.y);
                        normal2<Real>(rng.draw(step, key, 0u, (kTagRecruitOff << 16) | d.rxi),
                                      o0, o1);
                    }
                    local_site<Real>(d, rl, o0, o1, Lx, Ly, px[k], py[k]);
                } else {
                    px[k] = ux * Lx;
                    py[k] = uy * Ly;
                }
                if (d.f2.sensor >= 0) k2[k] = G.counts[d.f2.cbase + fb_cell(d.f2, px[k], py[k])];
            } else {
                const uint32_t tag = d.kind == EDM_RX_DETACH ? kTagDetach : kTagConvert;
                u[k] = draws ? draws[d.draw_off + it]
                             : (double)Uni<Real>::first(
                                   rng.draw(step, key, 0u, (tag << 16) | d.rxi));
            }
            if (d.f1.sensor >= 0) k1[k] = G.counts[d.f1.cbase + fb_cell(d.f1, px[k], py[k])];
        }
        // Phase 2: the passes in the reference's order.
#pragma unroll
        for (int k = 0; k < kMaxPasses; ++k) {
            if (k >= grp.n) continue;
            const PassDev& d = grp.p[k];
            if (d.kind == EDM_RX_DETACH) {
                if (!(f & EDM_FLAG_ALIVE) || s != d.a) continue;
                const Real pe =
                    d.f1.sensor >= 0 ? (Real)d.prob * fb_hill<Real>(d.f1, k1[k]) : (Real)d.prob;
                if (u[k] < (double)pe) {
                    f = f & ~EDM_FLAG_ALIVE;
                    count_update(G, n_grids, d.a, x, y, -1, d.deferred != 0);
                    ++det_hits;
                }
            } else if (d.kind == EDM_RX_CONVERT) {
                if (!(f & EDM_FLAG_ALIVE) || s != d.a) continue;
                const Real pe =
                    d.f1.sensor >= 0 ? (Real)d.prob * fb_hill<Real>(d.f1, k1[k]) : (Real)d.prob;
                if (u[k] < (double)pe) {
                    s = (int8_t)d.b;
                    count_update(G, n_grids, d.a, x, y, -1, d.deferred != 0);
                    count_update(G, n_grids, d.b, x, y, +1, d.deferred != 0);
                    ++conv_hits;
                }
            } else if (d.kind == EDM_RX_RECRUIT) {
                if ((f & EDM_FLAG_ALIVE) || s != d.a) continue;
                Real pe = (Real)d.prob;
                if (d.f1.sensor >= 0 && d.f2.sensor >= 0) {
                    const Real h1 = fb_hill<Real>(d.f1, k1[k]);
                    const Real h2 = fb_hill<Real>(d.f2, k2[k]);
                    pe = (Real)d.prob * h1 * h2;
                } else if (d.f1.sensor >= 0) {
                    pe = (Real)d.prob * fb_hill<Real>(d.f1, k1[k]);
                }
                if (u[k] < (double)pe) {
                    cand.on = true;
                    cand.it = it;
                    cand.x = px[k];
                    cand.y = py[k];
                    cand.rec = d.rec;
                }
            }
        }
        if (f != f0) P.fl[t] = f;
        if (s != s0) P.sp[t] = s;
    }
}

template <typename Real>
__device__ __forceinline__ void passes_count(int det_hits, int conv_hits, unsigned long long* ctr) {
    unsigned long long dh = det_hits, ch = conv_hits;
    for (int o = 32; o > 0; o >>= 1) {
        dh += __shfl_down(dh, o);
        ch += __shfl_down(ch, o);
    }
    if (lane_id() == 0) {
        if (dh) atomicAdd(&ctr[4], dh);
        if (ch) atomicAdd(&ctr[5], ch);
    }
}

// One listed slot per thread (the step's pass-species work list), or one
// slot per thread without a list.
template <typename Real, bool kLocal>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFp32Waves<Real, 6>))) void k_passes(int n, WlView V, const int32_t* __restrict__ id,
                                                Post<Real> P, PassGroupDev grp,
                                                const double* __restrict__ draws, Rng rng,
                                                StepRef sref, Real Lx, Real Ly,
                                                unsigned long long* ctr, Grids<Real> G,
                                                int n_grids, int32_t* cid, int32_t* cslot,
                                                double* cxo, double* cyo, unsigned int* nc,
                                                int64_t cand_stride) {
    const uint32_t step = sref.get();
    int t = -1;
    if (V.list) {
        const int m = wl_slice(V);
        if (m == 0) return;  // uniform: an empty slice
        if ((int)threadIdx.x < m) t = wl_slot(V);
    } else {
        const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        if (q < n) t = q;
    }
    int det_hits = 0, conv_hits = 0;
    Cand<Real> cand{false, 0, (Real)0, (Real)0, 0};
    if (t >= 0) {
        const int s0 = P.sp[t];
        if ((grp.species >> s0) & 1u)
            passes_particle<Real, kLocal>(t, s0, n, id, 0, P, grp, draws, rng, step, Lx, Ly, G,
                                          n_grids, cand, det_hits, conv_hits);
    }
    append_candidates(cand, t, cid, cslot, cxo, cyo, nc, cand_stride);
    passes_count<Real>(det_hits, conv_hits, ctr);
}
