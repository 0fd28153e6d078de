/*
 * ORACLE — test infrastructure only.
 *
 * CPU restatement, in plain C and FP64, of the ED-MRDT per-timestep hot path:
 * the numba kernels of ed_mrdt/dynamics.py, ed_mrdt/photophysics.py and
 * ed_mrdt/microscope.py.  Every function keeps the reference's loop order,
 * its sequential (greedy, first-come) reaction semantics, Python float-modulo
 * boundaries, int() truncation and its silent caps, so that with the same
 * injected random arrays it produces the same results as the reference.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the timed CPU baseline.  The
 * product path (the HIP library) never calls it.
 *
 * Pinning (which check fixes each function) is listed in oracle/README.md;
 * orc_pss_kernel's photon order, arrival-array consumption, max_photons cut,
 * exhaustion and transitions are fixed by a hand-built known answer and by
 * the CTMC statistics in tests/test_oracle_pins.py.
 *
 * Build: oracle/build_oracle.py (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Python float modulo for a positive divisor (numba mirrors CPython's
 * float_rem): r = fmod(x, L); a negative non-zero remainder gets +L, so the
 * result lies in [0, L] and can equal L for tiny negative x. */
static inline double py_mod(double x, double L) {
    double r = fmod(x, L);
    if (r != 0.0) {
        if ((L < 0.0) != (r < 0.0)) r += L;
    } else {
        r = copysign(0.0, L);
    }
    return r;
}

static inline int clampi(int v, int lo, int hi) {
    if (v >= hi) v = hi - 1;
    if (v < lo) v = lo;
    return v;
}

/* dynamics.py:200-216 _min_image_dist_sq */
static inline double min_image(double x1, double y1, double x2, double y2,
                               double Lx, double Ly, int px, int py,
                               double *odx, double *ody) {
    double dx = x1 - x2, dy = y1 - y2;
    if (px) {
        if (dx > 0.5 * Lx) dx -= Lx;
        else if (dx < -0.5 * Lx) dx += Lx;
    }
    if (py) {
        if (dy > 0.5 * Ly) dy -= Ly;
        else if (dy < -0.5 * Ly) dy += Ly;
    }
    *odx = dx;
    *ody = dy;
    return dx * dx + dy * dy;
}

double orc_min_image_dsq(double x1, double y1, double x2, double y2,
                         double Lx, double Ly, int px, int py,
                         double *dx, double *dy) {
    return min_image(x1, y1, x2, y2, Lx, Ly, px, py, dx, dy);
}

double orc_py_mod(double x, double L) { return py_mod(x, L); }

/* dynamics.py:177-196 _build_cell_list: head insertion, so each chain runs
 * in descending particle index. positions are [N,2] row-major. */
void orc_build_cell_list(const double *pos, int N, int *head, int *next,
                         int nx, int ny, double csx, double csy) {
    for (int c = 0; c < nx * ny; ++c) head[c] = -1;
    for (int i = 0; i < N; ++i) {
        int cx = (int)(pos[2 * i] / csx);
        int cy = (int)(pos[2 * i + 1] / csy);
        cx = clampi(cx, 0, nx);
        cy = clampi(cy, 0, ny);
        int c = cx * ny + cy;
        next[i] = head[c];
        head[c] = i;
    }
}

/* dynamics.py:220-278 _cell_list_pairs.  Pairs (i<j, d^2 < cutoff^2) in the
 * reference order: i ascending, 3x3 neighbour cells in (dcx, dcy)
 * lexicographic order, then chain order.  Pairs beyond max_pairs are counted
 * in *n_found but not stored (the reference drops them silently). */
int orc_cell_list_pairs(const double *pos, int N, const int *head,
                        const int *next, int nx, int ny, double csx,
                        double csy, double Lx, double Ly, int px, int py,
                        double cutoff, long long max_pairs, int *pairs,
                        long long *n_found) {
    double cut2 = cutoff * cutoff;
    long long n = 0, found = 0;
    for (int i = 0; i < N; ++i) {
        double xi = pos[2 * i], yi = pos[2 * i + 1];
        int cxi = clampi((int)(xi / csx), 0, nx);
        int cyi = clampi((int)(yi / csy), 0, ny);
        for (int dcx = -1; dcx <= 1; ++dcx) {
            for (int dcy = -1; dcy <= 1; ++dcy) {
                int ncx = cxi + dcx, ncy = cyi + dcy;
                if (px) ncx = ((ncx % nx) + nx) % nx;
                else if (ncx < 0 || ncx >= nx) continue;
                if (py) ncy = ((ncy % ny) + ny) % ny;
                else if (ncy < 0 || ncy >= ny) continue;
                int j = head[ncx * ny + ncy];
                while (j >= 0) {
                    if (j > i) {
                        double dx, dy;
                        double dsq = min_image(xi, yi, pos[2 * j], pos[2 * j + 1],
                                               Lx, Ly, px, py, &dx, &dy);
                        if (dsq < cut2) {
                            if (n < max_pairs) {
                                pairs[2 * n] = i;
                                pairs[2 * n + 1] = j;
                                ++n;
                            }
                            ++found;
                        }
                    }
                    j = next[j];
                }
            }
        }
    }
    if (n_found) *n_found = found;
    return (int)n;
}

/* orc_cell_list_pairs on n_threads threads (the timing variant with every
 * loop threaded; the serial routine above is the restatement): the pairs of
 * each i are counted in parallel, an exclusive scan over i gives every i its
 * offset in the list, and each i's pairs are written there, so the list, its
 * order and the cap are the serial routine's. */
static long long pairs_of(const double *pos, const int *head, const int *next,
                          int nx, int ny, double csx, double csy, double Lx,
                          double Ly, int px, int py, double cut2, int i,
                          int *out, long long room) {
    double xi = pos[2 * i], yi = pos[2 * i + 1];
    int cxi = clampi((int)(xi / csx), 0, nx);
    int cyi = clampi((int)(yi / csy), 0, ny);
    long long k = 0;
    for (int dcx = -1; dcx <= 1; ++dcx) {
        for (int dcy = -1; dcy <= 1; ++dcy) {
            int ncx = cxi + dcx, ncy = cyi + dcy;
            if (px) ncx = ((ncx % nx) + nx) % nx;
            else if (ncx < 0 || ncx >= nx) continue;
            if (py) ncy = ((ncy % ny) + ny) % ny;
            else if (ncy < 0 || ncy >= ny) continue;
            for (int j = head[ncx * ny + ncy]; j >= 0; j = next[j]) {
                if (j <= i) continue;
                double dx, dy;
                if (min_image(xi, yi, pos[2 * j], pos[2 * j + 1], Lx, Ly, px, py,
                              &dx, &dy) < cut2) {
                    if (out && k < room) {
                        out[2 * k] = i;
                        out[2 * k + 1] = j;
                    }
                    ++k;
                }
            }
        }
    }
    return k;
}

int orc_cell_list_pairs_par(const double *pos, int N, const int *head,
                            const int *next, int nx, int ny, double csx,
                            double csy, double Lx, double Ly, int px, int py,
                            double cutoff, long long max_pairs, int *pairs,
                            long long *n_found, int n_threads) {
    const double cut2 = cutoff * cutoff;
    const int nt = n_threads > 0 ? n_threads : 1;
    long long *off = (long long *)malloc(((size_t)N + 1) * sizeof(long long));
    if (!off) return -1;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < N; ++i)
        off[i] = pairs_of(pos, head, next, nx, ny, csx, csy, Lx, Ly, px, py, cut2, i, NULL, 0);
    long long run = 0;
    for (int i = 0; i < N; ++i) {
        const long long c = off[i];
        off[i] = run;
        run += c;
    }
    off[N] = run;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < N; ++i) {
        if (off[i] >= max_pairs || off[i + 1] == off[i]) continue;
        pairs_of(pos, head, next, nx, ny, csx, csy, Lx, Ly, px, py, cut2, i,
                 pairs + 2 * off[i], max_pairs - off[i]);
    }
    if (n_found) *n_found = run;
    free(off);
    return (int)(run < max_pairs ? run : max_pairs);
}

/* dynamics.py:291-307 _brute_force_pairs_numba (validation helper). */
int orc_brute_force_pairs(const double *pos, int N, double cutoff, double Lx,
                          double Ly, int px, int py, long long max_pairs,
                          int *pairs) {
    double cut2 = cutoff * cutoff;
    long long n = 0;
    for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) {
            double dx, dy;
            double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                                   pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
            if (dsq < cut2 && n < max_pairs) {
                pairs[2 * n] = i;
                pairs[2 * n + 1] = j;
                ++n;
            }
        }
    return (int)n;
}

/* dynamics.py:441-519 apply_diffusion_field_regions_smooth */
void orc_diffusion_field_smooth(double *D, double *gradD, const double *D_base,
                                const int *species, const double *pos, int N,
                                const double *rcx, const double *rcy,
                                const double *rr, const double *rf,
                                int n_regions, double w) {
    for (int i = 0; i < N; ++i) {
        double x = pos[2 * i], y = pos[2 * i + 1];
        double base = D_base[species[i]];
        double factor = 1.0, gfx = 0.0, gfy = 0.0;
        for (int r = 0; r < n_regions; ++r) {
            double dxc = x - rcx[r], dyc = y - rcy[r];
            double dsq = dxc * dxc + dyc * dyc;
            double d = dsq > 1e-30 ? sqrt(dsq) : 1e-15;
            double arg = (d - rr[r]) / w;
            double sig, dsig;
            if (arg > 100.0) {
                sig = 0.0;
                dsig = 0.0;
            } else if (arg < -100.0) {
                sig = 1.0;
                dsig = 0.0;
            } else {
                double e = exp(arg);
                double den = 1.0 + e;
                sig = 1.0 / den;
                dsig = -e / (den * den * w);
            }
            double local = 1.0 + (rf[r] - 1.0) * sig;
            if (sig > 0.01) {
                factor = local;
                double inv_d = 1.0 / d;
                gfx = (rf[r] - 1.0) * dsig * dxc * inv_d;
                gfy = (rf[r] - 1.0) * dsig * dyc * inv_d;
                break;
            }
        }
        D[i] = base * factor;
        gradD[2 * i] = base * gfx;
        gradD[2 * i + 1] = base * gfy;
    }
}

/* Sinusoidal D(x) field of the wave-system workload (not in the reference):
 * D = D_s (1 + a sin(kx x) sin(ky y)), grad D analytic. */
/* sin and cos by Cody-Waite reduction modulo pi/2 (two 33-bit parts of
 * pi/2 and a 53-bit tail) and Taylor polynomials through r^17 / r^18 on
 * [-pi/4, pi/4], Horner form; IEEE add, multiply and rint only.  The
 * sinusoidal field is not in the reference, so the field's definition is
 * this build's: this method, checked against libm sin/cos to 2e-16 and by
 * the closed-form gradient (tests/test_oracle_kernels.py). */
void orc_sincos(double x, double *s, double *c) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    const double r = ((x - k * 0x1.921fb54400000p+0) - k * 0x1.0b4611a600000p-34) -
                     k * 0x1.3198a2e037073p-69;
    const double r2 = r * r;
    static const double S[8] = {0x1.952c77030ad4ap-49, -0x1.ae7f3e733b81fp-41,
                                0x1.6124613a86d09p-33, -0x1.ae64567f544e4p-26,
                                0x1.71de3a556c734p-19, -0x1.a01a01a01a01ap-13,
                                0x1.1111111111111p-7, -0x1.5555555555555p-3};
    static const double C[9] = {-0x1.6827863b97d97p-53, 0x1.ae7f3e733b81fp-45,
                                -0x1.93974a8c07c9dp-37, 0x1.1eed8eff8d898p-29,
                                -0x1.27e4fb7789f5cp-22, 0x1.a01a01a01a01ap-16,
                                -0x1.6c16c16c16c17p-10, 0x1.5555555555555p-5,
                                -0x1.0000000000000p-1};
    double ps = S[0];
    for (int q = 1; q < 8; ++q) ps = S[q] + r2 * ps;
    double pc = C[0];
    for (int q = 1; q < 9; ++q) pc = C[q] + r2 * pc;
    const double sr = r + r * (r2 * ps);
    const double cr = 1.0 + r2 * pc;
    switch (((long long)k) & 3) {
        case 0: *s = sr; *c = cr; break;
        case 1: *s = cr; *c = -sr; break;
        case 2: *s = -sr; *c = -cr; break;
        default: *s = -cr; *c = sr; break;
    }
}

void orc_diffusion_field_sinusoidal(double *D, double *gradD,
                                    const double *D_base, const int *species,
                                    const double *pos, int N, double amp,
                                    double kx, double ky) {
    for (int i = 0; i < N; ++i) {
        double x = pos[2 * i], y = pos[2 * i + 1];
        double base = D_base[species[i]];
        double sx, cx, sy, cy;
        orc_sincos(kx * x, &sx, &cx);
        orc_sincos(ky * y, &sy, &cy);
        D[i] = base * (1.0 + amp * sx * sy);
        gradD[2 * i] = base * amp * kx * cx * sy;
        gradD[2 * i + 1] = base * amp * ky * sx * cy;
    }
}

/* dynamics.py:666-739 compute_forces_wca */
void orc_forces_wca(const double *pos, double *forces, const int *species,
                    const uint8_t *alive, const int *pairs, int n_pairs,
                    const double *sig_m, const double *eps_m, int n_sp,
                    double Lx, double Ly, int px, int py) {
    for (int p = 0; p < n_pairs; ++p) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        double sigma = sig_m[species[i] * n_sp + species[j]];
        double eps = eps_m[species[i] * n_sp + species[j]];
        if (sigma <= 0.0 || eps <= 0.0) continue;
        double dx, dy;
        double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                               pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
        double rcut2 = sigma * sigma * 1.2599210498948732;
        if (dsq >= rcut2 || dsq < 1e-30) continue;
        double inv_r2 = 1.0 / dsq;
        double s2 = sigma * sigma * inv_r2;
        double s6 = s2 * s2 * s2;
        double s12 = s6 * s6;
        double f = 24.0 * eps * (2.0 * s12 - s6) * inv_r2;
        double fmax = 2400.0 * eps / (sigma * sigma);
        if (f > fmax) f = fmax;
        double fx = f * dx, fy = f * dy;
        forces[2 * i] += fx;
        forces[2 * i + 1] += fy;
        forces[2 * j] -= fx;
        forces[2 * j + 1] -= fy;
    }
}

/* dynamics.py:753-816 compute_forces_harmonic */
void orc_forces_harmonic(const double *pos, double *forces, const int *species,
                         const uint8_t *alive, const int *pairs, int n_pairs,
                         const double *sig_m, const double *k_m, int n_sp,
                         double Lx, double Ly, int px, int py) {
    for (int p = 0; p < n_pairs; ++p) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        double sigma = sig_m[species[i] * n_sp + species[j]];
        double k = k_m[species[i] * n_sp + species[j]];
        if (sigma <= 0.0 || k <= 0.0) continue;
        double dx, dy;
        double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                               pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
        if (dsq >= sigma * sigma || dsq < 1e-30) continue;
        double r = sqrt(dsq);
        double fmag = k * (sigma - r);
        double f = fmag / r;
        double fx = f * dx, fy = f * dy;
        forces[2 * i] += fx;
        forces[2 * i + 1] += fy;
        forces[2 * j] -= fx;
        forces[2 * j + 1] -= fy;
    }
}

/* orc_forces_wca / orc_forces_harmonic over the pair list on n_threads
 * threads (the timing variant with every loop threaded): the same pair
 * terms, added with atomic updates, so the sums may differ from the serial
 * routines' in the last bits. */
void orc_forces_par(int harmonic, const double *pos, double *forces,
                    const int *species, const uint8_t *alive, const int *pairs,
                    int n_pairs, const double *sig_m, const double *p_m, int n_sp,
                    double Lx, double Ly, int px, int py, int n_threads) {
    const int nt = n_threads > 0 ? n_threads : 1;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int p = 0; p < n_pairs; ++p) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        double sigma = sig_m[species[i] * n_sp + species[j]];
        double a = p_m[species[i] * n_sp + species[j]];
        if (sigma <= 0.0 || a <= 0.0) continue;
        double dx, dy;
        double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                               pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
        double f;
        if (harmonic) {
            if (dsq >= sigma * sigma || dsq < 1e-30) continue;
            double r = sqrt(dsq);
            f = a * (sigma - r) / r;
        } else {
            double rcut2 = sigma * sigma * 1.2599210498948732;
            if (dsq >= rcut2 || dsq < 1e-30) continue;
            double inv_r2 = 1.0 / dsq;
            double s2 = sigma * sigma * inv_r2;
            double s6 = s2 * s2 * s2;
            f = 24.0 * a * (2.0 * s6 * s6 - s6) * inv_r2;
            double fmax = 2400.0 * a / (sigma * sigma);
            if (f > fmax) f = fmax;
        }
        double fx = f * dx, fy = f * dy;
#pragma omp atomic
        forces[2 * i] += fx;
#pragma omp atomic
        forces[2 * i + 1] += fy;
#pragma omp atomic
        forces[2 * j] -= fx;
#pragma omp atomic
        forces[2 * j + 1] -= fy;
    }
}

/* dynamics.py:323-406 bd_step (prange in the reference; the per-particle
 * updates are independent, so threads change nothing but the timing). */
void orc_bd_step(double *pos, const double *D, const double *forces,
                 const double *gradD, double dt, double Lx, double Ly, int px,
                 int py, int N, const double *noise, const uint8_t *alive,
                 double maxd, int n_threads) {
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : 1;
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (int i = 0; i < N; ++i) {
        if (!alive[i]) continue;
        double Di = D[i];
        double sn = sqrt(2.0 * Di * dt);
        double dx = gradD[2 * i] * dt;
        double dy = gradD[2 * i + 1] * dt;
        dx += sn * noise[2 * i];
        dy += sn * noise[2 * i + 1];
        dx += Di * forces[2 * i] * dt;
        dy += Di * forces[2 * i + 1] * dt;
        double dr2 = dx * dx + dy * dy;
        if (dr2 > maxd * maxd) {
            double dr = sqrt(dr2);
            double s = maxd / dr;
            dx *= s;
            dy *= s;
        }
        double x = pos[2 * i] + dx;
        double y = pos[2 * i + 1] + dy;
        if (px) x = py_mod(x, Lx);
        else {
            if (x < 0.0) x = -x;
            else if (x >= Lx) x = 2.0 * Lx - x;
        }
        if (py) y = py_mod(y, Ly);
        else {
            if (y < 0.0) y = -y;
            else if (y >= Ly) y = 2.0 * Ly - y;
        }
        pos[2 * i] = x;
        pos[2 * i + 1] = y;
    }
    (void)n_threads;
}

/* Hill factor shared by every feedback kernel (dynamics.py:966-976 and the
 * identical blocks in the recruitment/detachment/conversion kernels). */
static inline double hill(double rho, double K, double n, int inhibition,
                          double basal) {
    if (K > 0) {
        double ratio = rho / K;
        double rn = pow(ratio, n);
        if (inhibition) return 1.0 / (1.0 + rn);
        double h = rn / (1.0 + rn);
        return basal + (1.0 - basal) * h;
    }
    return 1.0;
}

double orc_hill(double rho, double K, double n, int inhibition, double basal) {
    return hill(rho, K, n, inhibition, basal);
}

static inline int grid_cell(double x, double y, int nxg, int nyg, double sx,
                            double sy, int *gx, int *gy) {
    int a = clampi((int)(x / sx), 0, nxg);
    int b = clampi((int)(y / sy), 0, nyg);
    *gx = a;
    *gy = b;
    return a * nyg + b;
}

/* A feedback rule bound to a density field ([nxg, nyg] row-major). */
typedef struct {
    const double *density; /* NULL = no feedback */
    int nxg, nyg;
    double sx, sy;
    double K, n;
    int inhibition;
    double basal;
} orc_fb;

static inline double fb_factor(const orc_fb *fb, double x, double y) {
    if (!fb || !fb->density) return 1.0;
    int gx, gy;
    int c = grid_cell(x, y, fb->nxg, fb->nyg, fb->sx, fb->sy, &gx, &gy);
    return hill(fb->density[c], fb->K, fb->n, fb->inhibition, fb->basal);
}

/* dynamics.py:529-606 apply_binding_reactions */
int orc_binding(const double *pos, int *species, int *bond, uint8_t *alive,
                int ra, int rb, int prod, double contact, double prob,
                const int *pairs, int n_pairs, const double *u, double Lx,
                double Ly, int px, int py) {
    int nb = 0;
    double c2 = contact * contact;
    for (int p = 0; p < n_pairs; ++p) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        if (bond[i] >= 0 || bond[j] >= 0) continue;
        int a, b;
        if (species[i] == ra && species[j] == rb) {
            a = i;
            b = j;
        } else if (species[i] == rb && species[j] == ra) {
            a = j;
            b = i;
        } else
            continue;
        double dx, dy;
        double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                               pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
        if (dsq > c2) continue;
        if (u[p] < prob) {
            species[a] = prod;
            alive[b] = 0;
            bond[a] = b;
            bond[b] = a;
            ++nb;
        }
    }
    return nb;
}

/* dynamics.py:610-654 apply_dissociation */
int orc_dissociation(int *species, int *bond, uint8_t *alive, double *pos,
                     int N, int prod, int ra, int rb, double prob,
                     const double *u) {
    int nd = 0;
    for (int i = 0; i < N; ++i) {
        if (!alive[i]) continue;
        if (species[i] != prod) continue;
        if (bond[i] < 0) continue;
        if (u[i] < prob) {
            int j = bond[i];
            species[i] = ra;
            species[j] = rb;
            alive[j] = 1;
            pos[2 * j] = pos[2 * i];
            pos[2 * j + 1] = pos[2 * i + 1];
            bond[i] = -1;
            bond[j] = -1;
            ++nd;
        }
    }
    return nd;
}

/* Cell of the adaptation field read for a substrate in feedback-grid cell
 * (gx, gy).  dynamics.py:1078 reads adaptation_field[gx, gy] with the
 * FEEDBACK grid's indices on the adaptation grid's [anx, any] array, and
 * numba does not bounds-check: the element read is the flat one,
 * gx * any + gy, which is another cell when the two grids differ and lies
 * past the array when gx * any + gy >= anx * any.  Where it lies inside the
 * array this follows the reference exactly; past it the reference reads
 * memory it does not own (no defined value), and this restatement reads
 * the adaptation field at the substrate's own cell instead (the adaptation
 * grid spans the feedback grid's extent nxg * sx by nyg * sy). */
static inline int adapt_cell(double x, double y, int gx, int gy, int nxg, int nyg,
                             double sx, double sy, int anx, int any) {
    long flat = (long)gx * any + gy;
    if (flat < (long)anx * any) return (int)flat;
    double asx = (nxg * sx) / anx, asy = (nyg * sy) / any;
    int ax = clampi((int)(x / asx), 0, anx);
    int ay = clampi((int)(y / asy), 0, any);
    return ax * any + ay;
}

int orc_adapt_cell(double x, double y, int gx, int gy, int nxg, int nyg, double sx,
                   double sy, int anx, int any) {
    return adapt_cell(x, y, gx, gy, nxg, nyg, sx, sy, anx, any);
}

/* dynamics.py:826-1090: apply_catalytic_reaction, _with_feedback and
 * _with_feedback_and_adaptation in one body.  fb->density == NULL selects
 * the plain kernel; adapt != NULL ([anx, any]) multiplies the adaptation
 * factor read as adapt_cell gives it (dynamics.py:1078-1084). */
int orc_catalytic(const double *pos, int *species, const uint8_t *alive,
                  int enz, int sub, int prod, double contact, double prob,
                  const int *pairs, int n_pairs, const double *u, double Lx,
                  double Ly, int px, int py, const double *density, int nxg,
                  int nyg, double gsx, double gsy, double K, double n,
                  int inhibition, double basal, const double *adapt,
                  int anx, int any, double K_adapt, double n_adapt) {
    int nc = 0;
    double c2 = contact * contact;
    for (int p = 0; p < n_pairs; ++p) {
        int i = pairs[2 * p], j = pairs[2 * p + 1];
        if (!alive[i] || !alive[j]) continue;
        int s;
        if (species[i] == enz && species[j] == sub) s = j;
        else if (species[j] == enz && species[i] == sub) s = i;
        else continue;
        double dx, dy;
        double dsq = min_image(pos[2 * i], pos[2 * i + 1], pos[2 * j],
                               pos[2 * j + 1], Lx, Ly, px, py, &dx, &dy);
        if (dsq > c2) continue;
        double pe = prob;
        if (density) {
            int gx, gy;
            int c = grid_cell(pos[2 * s], pos[2 * s + 1], nxg, nyg, gsx, gsy,
                              &gx, &gy);
            double hf = hill(density[c], K, n, inhibition, basal);
            if (adapt) {
                double af = 1.0;
                if (K_adapt > 0) {
                    int wc = adapt_cell(pos[2 * s], pos[2 * s + 1], gx, gy, nxg, nyg, gsx,
                                        gsy, anx, any);
                    double wr = pow(adapt[wc] / K_adapt, n_adapt);
                    af = 1.0 / (1.0 + wr);
                }
                pe = prob * hf * af;
            } else {
                pe = prob * hf;
            }
        }
        if (u[p] < pe) {
            species[s] = prod;
            ++nc;
        }
    }
    return nc;
}

/* dynamics.py:1094-1112 update_adaptation_field */
void orc_adaptation_update(double *w, const double *rho, int nx, int ny,
                           double dt, double tau) {
    double inv = dt / tau;
    for (int k = 0; k < nx * ny; ++k) w[k] += inv * (rho[k] - w[k]);
}

/* dynamics.py:1230-1458 apply_membrane_recruitment, _with_feedback and
 * _with_dual_feedback (fb1/fb2 density NULL = that factor is absent; the
 * reference multiplies prob * hill_1 * hill_2 in that order). */
int orc_recruit(const int *species, uint8_t *alive,
                double *pos, int N, int target, double prob, double Lx,
                double Ly, const double *u_acc, const double *u_pos,
                int max_recruits, const double *d1, int nx1, int ny1,
                double sx1, double sy1, double K1, double n1, int inh1,
                double b1, const double *d2, int nx2, int ny2, double sx2,
                double sy2, double K2, double n2, int inh2, double b2) {
    int nr = 0;
    for (int i = 0; i < N; ++i) {
        if (nr >= max_recruits) break;
        if (alive[i]) continue;
        if (species[i] != target) continue;
        double cx = u_pos[2 * i] * Lx;
        double cy = u_pos[2 * i + 1] * Ly;
        double pe = prob;
        if (d1 && d2) {
            int gx, gy;
            int c1 = grid_cell(cx, cy, nx1, ny1, sx1, sy1, &gx, &gy);
            double h1 = hill(d1[c1], K1, n1, inh1, b1);
            int c2 = grid_cell(cx, cy, nx2, ny2, sx2, sy2, &gx, &gy);
            double h2 = hill(d2[c2], K2, n2, inh2, b2);
            pe = prob * h1 * h2;
        } else if (d1) {
            int gx, gy;
            int c1 = grid_cell(cx, cy, nx1, ny1, sx1, sy1, &gx, &gy);
            pe = prob * hill(d1[c1], K1, n1, inh1, b1);
        }
        if (u_acc[i] < pe) {
            alive[i] = 1;
            pos[2 * i] = cx;
            pos[2 * i + 1] = cy;
            ++nr;
        }
    }
    return nr;
}

/* dynamics.py:1116-1217 apply_membrane_recruitment_local (CDF-biased). */
int orc_recruit_local(const int *species, uint8_t *alive, double *pos, int N,
                      int target, double prob, double Lx, double Ly,
                      const double *u_acc, const double *offsets,
                      int max_recruits, const double *density, int nxg,
                      int nyg, double sx, double sy, double K, double n,
                      int inh, double basal, const double *cdf,
                      const double *u_loc, double sigma_r) {
    int nr = 0, ncell = nxg * nyg;
    for (int i = 0; i < N; ++i) {
        if (nr >= max_recruits) break;
        if (alive[i]) continue;
        if (species[i] != target) continue;
        double r = u_loc[i];
        int lo = 0, hi = ncell - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (cdf[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        int gx = lo / nyg, gy = lo % nyg;
        double cx = (gx + 0.5) * sx + offsets[2 * i] * sigma_r;
        double cy = (gy + 0.5) * sy + offsets[2 * i + 1] * sigma_r;
        cx = py_mod(cx, Lx);
        cy = py_mod(cy, Ly);
        int g2x, g2y;
        int c = grid_cell(cx, cy, nxg, nyg, sx, sy, &g2x, &g2y);
        double pe = prob * hill(density[c], K, n, inh, basal);
        if (u_acc[i] < pe) {
            alive[i] = 1;
            pos[2 * i] = cx;
            pos[2 * i + 1] = cy;
            ++nr;
        }
    }
    return nr;
}

/* dynamics.py:1462-1555 apply_membrane_detachment(_with_feedback) */
int orc_detach(const int *species, uint8_t *alive, const double *pos, int N,
               int target, double prob, const double *u,
               const double *density, int nxg, int nyg, double sx, double sy,
               double K, double n, int inh, double basal) {
    int nd = 0;
    for (int i = 0; i < N; ++i) {
        if (!alive[i]) continue;
        if (species[i] != target) continue;
        double pe = prob;
        if (density) {
            int gx, gy;
            int c = grid_cell(pos[2 * i], pos[2 * i + 1], nxg, nyg, sx, sy, &gx,
                              &gy);
            pe = prob * hill(density[c], K, n, inh, basal);
        }
        if (u[i] < pe) {
            alive[i] = 0;
            ++nd;
        }
    }
    return nd;
}

/* dynamics.py:1566-1657 apply_species_conversion(_with_feedback) */
int orc_convert(int *species, const uint8_t *alive, const double *pos, int N,
                int from, int to, double prob, const double *u,
                const double *density, int nxg, int nyg, double sx, double sy,
                double K, double n, int inh, double basal) {
    int nc = 0;
    for (int i = 0; i < N; ++i) {
        if (!alive[i]) continue;
        if (species[i] != from) continue;
        double pe = prob;
        if (density) {
            int gx, gy;
            int c = grid_cell(pos[2 * i], pos[2 * i + 1], nxg, nyg, sx, sy, &gx,
                              &gy);
            pe = prob * hill(density[c], K, n, inh, basal);
        }
        if (u[i] < pe) {
            species[i] = to;
            ++nc;
        }
    }
    return nc;
}

/* dynamics.py:1668-1703 compute_density_field: sequential FP64 sums of
 * inv_area, so a cell with k particles holds fl(...fl(inv+inv)...+inv). */
void orc_density(const double *pos, const int *species, const uint8_t *alive,
                 int N, int target, int nxg, int nyg, double sx, double sy,
                 double *out) {
    double area = sx * sy;
    double inv = area > 0.0 ? 1.0 / area : 0.0;
    for (int i = 0; i < N; ++i) {
        if (!alive[i]) continue;
        if (species[i] != target) continue;
        int gx, gy;
        int c = grid_cell(pos[2 * i], pos[2 * i + 1], nxg, nyg, sx, sy, &gx, &gy);
        out[c] += inv;
    }
}

/* ------------------------------------------------------------------ */
/* Photophysics                                                        */
/* ------------------------------------------------------------------ */

/* photophysics.py:487-521 _poisson_draw: one U(0,1) seeds a 64-bit LCG;
 * Knuth's product method for mean < 30, else one Box-Muller normal. */
int orc_poisson_draw(double mean, double u_seed) {
    if (mean <= 0.0) return 0;
    uint64_t s = (uint64_t)(int64_t)(u_seed * 9.007199254740992e15);
    const uint64_t A = 6364136223846793005ULL, C = 1442695040888963407ULL;
    if (mean < 30.0) {
        double L = exp(-mean);
        int k = 0;
        double p = 1.0;
        for (;;) {
            s = s * A + C;
            double u = (double)(s >> 11) / 9.007199254740992e15;
            p *= u;
            if (p <= L) break;
            ++k;
            if (k > 10000) break;
        }
        return k;
    }
    s = s * A + C;
    double u1 = (double)(s >> 11) / 9.007199254740992e15;
    s = s * A + C;
    double u2 = (double)(s >> 11) / 9.007199254740992e15;
    if (u1 < 1e-30) u1 = 1e-30;
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double v = mean + sqrt(mean) * z + 0.5;
    long long iv = (long long)v;
    return iv > 0 ? (int)iv : 0;
}

/* photophysics.py:529-588 _compute_fret_rates_numba (bonded donors). */
void orc_fret_rates_bond(double *k_fret, int *acc_idx, const uint8_t *has_dye,
                         const uint8_t *bleached, const int *bond,
                         const int *dye_type, int N, double R0, double tau_D,
                         int donor_fid, int acceptor_fid, double r_ic,
                         int intramolecular) {
    for (int i = 0; i < N; ++i) {
        if (!has_dye[i] || bleached[i]) continue;
        if (dye_type[i] != donor_fid) continue;
        int partner = bond[i];
        if (partner >= 0) {
            if (intramolecular) {
                double ratio = R0 / r_ic;
                double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
                k_fret[i] = r6 / tau_D;
                acc_idx[i] = -1;
            } else if (has_dye[partner] && dye_type[partner] == acceptor_fid) {
                double ratio = R0 / r_ic;
                double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
                k_fret[i] = r6 / tau_D;
                acc_idx[i] = partner;
            }
        }
    }
}

/* photophysics.py:592-683 _compute_fret_rates_proximity_numba (O(N^2)). */
void orc_fret_rates_proximity(double *k_fret, int *acc_idx, const double *pos,
                              const int *species, const uint8_t *has_dye,
                              const uint8_t *alive, const uint8_t *bleached,
                              const int *dye_type, int N, int donor_fid,
                              int target, double radius, double k_hi,
                              double k_lo, double Lx, double Ly, int px,
                              int py) {
    double r2 = radius * radius;
    for (int i = 0; i < N; ++i) {
        if (!has_dye[i] || !alive[i] || bleached[i]) continue;
        if (dye_type[i] != donor_fid) continue;
        int found = 0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            if (!alive[j] || species[j] != target) continue;
            double dx = pos[2 * j] - pos[2 * i];
            double dy = pos[2 * j + 1] - pos[2 * i + 1];
            if (px) {
                if (dx > Lx / 2.0) dx -= Lx;
                else if (dx < -Lx / 2.0) dx += Lx;
            }
            if (py) {
                if (dy > Ly / 2.0) dy -= Ly;
                else if (dy < -Ly / 2.0) dy += Ly;
            }
            if (dx * dx + dy * dy < r2) {
                found = 1;
                break;
            }
        }
        k_fret[i] = found ? k_hi : k_lo;
        acc_idx[i] = -1;
    }
}

/* photophysics.py:339-483 _evolve_pss_kernel.  Tables are per fluorophore
 * type: T_step[nf,3,3], mean_*[nf,2], tau[nf], k_bound[nf], states[nf]. */
int orc_pss_kernel(int8_t *ctmc, uint8_t *bleached, const uint8_t *has_dye,
                   const uint8_t *alive, const int *dye_type,
                   const double *k_fret, const int *acc_idx, int nf,
                   const double *T_step, const double *T_step_f,
                   const double *mean_ph, const double *mean_ph_f,
                   const double *mean_fr, const double *mean_fr_f,
                   const double *tau_all, const double *k_bound,
                   const int *gs, const int *es, const int *t1s,
                   const int *bls, int *out_particle, double *out_time,
                   int8_t *out_channel, int *out_pulse, uint8_t *out_is_fret,
                   int max_photons, int N, const double *u_pois,
                   const double *u_trans, const double *u_arr, int n_arr) {
    int n = 0, ai = 0;
    for (int i = 0; i < N; ++i) {
        if (!has_dye[i] || !alive[i] || bleached[i]) continue;
        int t = dye_type[i];
        if (t < 0 || t >= nf) continue;
        int s0 = gs[t], s1 = es[t], t1 = t1s[t], bl = bls[t];
        int cur = ctmc[i], pss;
        if (cur == s0 || cur == s1) pss = 0;
        else if (t1 >= 0 && cur == t1) pss = 1;
        else if (bl >= 0 && cur == bl) {
            bleached[i] = 1;
            continue;
        } else pss = 0;
        int has_fret = k_fret[i] > 0.0;
        double md, mf, tau, T00, T01;
        if (has_fret) {
            double tD = tau_all[t];
            double kf = k_fret[i];
            tau = tD / (1.0 + kf * tD);
            double kb = k_bound[t];
            double frac = 0.0;
            if (kb > 0.0) {
                frac = kf / kb;
                if (frac > 1.0) frac = 1.0;
            }
            md = mean_ph[t * 2 + pss] * (1.0 - frac) + mean_ph_f[t * 2 + pss] * frac;
            mf = mean_fr[t * 2 + pss] * (1.0 - frac) + mean_fr_f[t * 2 + pss] * frac;
            T00 = T_step[t * 9 + pss * 3 + 0] * (1.0 - frac) +
                  T_step_f[t * 9 + pss * 3 + 0] * frac;
            T01 = T_step[t * 9 + pss * 3 + 1] * (1.0 - frac) +
                  T_step_f[t * 9 + pss * 3 + 1] * frac;
        } else {
            md = mean_ph[t * 2 + pss];
            mf = mean_fr[t * 2 + pss];
            tau = tau_all[t];
            T00 = T_step[t * 9 + pss * 3 + 0];
            T01 = T_step[t * 9 + pss * 3 + 1];
        }
        int nd = orc_poisson_draw(md, u_pois[i]);
        int nfr = 0;
        if (mf > 0.0 && has_fret) nfr = orc_poisson_draw(mf, u_pois[i] * 0.5 + 0.25);
        for (int k = 0; k < nd; ++k) {
            if (n >= max_photons) break;
            double ta;
            if (ai < n_arr) {
                double u = u_arr[ai++];
                if (u < 1e-30) u = 1e-30;
                ta = -tau * log(u);
            } else
                ta = tau;
            out_particle[n] = i;
            out_time[n] = ta;
            out_channel[n] = (int8_t)t;
            out_pulse[n] = 0;
            out_is_fret[n] = 0;
            ++n;
        }
        int acc = acc_idx[i];
        if (nfr > 0 && acc >= 0) {
            int at = dye_type[acc];
            double ta_tau = (at >= 0 && at < nf) ? tau_all[at] : tau;
            for (int k = 0; k < nfr; ++k) {
                if (n >= max_photons) break;
                double ta;
                if (ai < n_arr) {
                    double u = u_arr[ai++];
                    if (u < 1e-30) u = 1e-30;
                    ta = -ta_tau * log(u);
                } else
                    ta = ta_tau;
                out_particle[n] = acc;
                out_time[n] = ta;
                out_channel[n] = (int8_t)at;
                out_pulse[n] = 0;
                out_is_fret[n] = 1;
                ++n;
            }
        }
        double ut = u_trans[i];
        if (ut < T00) ctmc[i] = (int8_t)s0;
        else if (ut < T00 + T01) ctmc[i] = (int8_t)(t1 >= 0 ? t1 : s0);
        else {
            ctmc[i] = (int8_t)(bl >= 0 ? bl : s0);
            bleached[i] = 1;
        }
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* Microscope                                                          */
/* ------------------------------------------------------------------ */

/* microscope.py:523-631 _assign_photons_emitter_centric.  Photons must be
 * sorted by particle index.  Returns the number assigned; detected[p] marks
 * them.  The cache of the last emitter's CDF is kept as in the reference. */
int orc_assign_emitter_centric(const int *pidx, const int8_t *channel,
                               const double *posx, const double *posy,
                               int n_ph, const double *pcx, const double *pcy,
                               int nx, int ny, double pixel, double offx,
                               double offy, int R, const double *sig_ch,
                               int n_sig, const double *u, int *out_px,
                               int *out_py, uint8_t *detected) {
    int n_det = 0;
    int last_pid = -999;
    double last_sigma = -1.0;
    int maxn = (2 * R + 1) * (2 * R + 1);
    double *cw = (double *)malloc(sizeof(double) * maxn);
    int *ci = (int *)malloc(sizeof(int) * maxn);
    int *cj = (int *)malloc(sizeof(int) * maxn);
    int cn = 0;
    double ct = 0.0;
    for (int p = 0; p < n_ph; ++p) {
        int pid = pidx[p];
        int ch = channel[p];
        if (ch < 0) ch = 0;
        if (ch >= n_sig) ch = 0;
        double sigma = sig_ch[ch];
        if (sigma <= 0.0) sigma = sig_ch[0];
        if (pid != last_pid || sigma != last_sigma) {
            double x = posx[pid], y = posy[pid];
            double fx = (x - offx) / pixel - 0.5;
            double fy = (y - offy) / pixel - 0.5;
            int i0 = (int)floor(fx + 0.5);
            int j0 = (int)floor(fy + 0.5);
            if (i0 < -R || i0 >= nx + R) continue;
            if (j0 < -R || j0 >= ny + R) continue;
            double inv2 = 0.5 / (sigma * sigma);
            cn = 0;
            ct = 0.0;
            for (int di = -R; di <= R; ++di) {
                int ii = i0 + di;
                if (ii < 0 || ii >= nx) continue;
                double dx = x - pcx[ii];
                for (int dj = -R; dj <= R; ++dj) {
                    int jj = j0 + dj;
                    if (jj < 0 || jj >= ny) continue;
                    double dy = y - pcy[jj];
                    double w = exp(-(dx * dx + dy * dy) * inv2);
                    ct += w;
                    cw[cn] = ct;
                    ci[cn] = ii;
                    cj[cn] = jj;
                    ++cn;
                }
            }
            last_pid = pid;
            last_sigma = sigma;
        }
        if (cn == 0 || ct <= 0.0) continue;
        double uu = u[p] * ct;
        int si = ci[0], sj = cj[0];
        for (int k = 0; k < cn; ++k)
            if (cw[k] >= uu) {
                si = ci[k];
                sj = cj[k];
                break;
            }
        out_px[n_det] = si;
        out_py[n_det] = sj;
        detected[p] = 1;
        ++n_det;
    }
    free(cw);
    free(ci);
    free(cj);
    return n_det;
}

/* microscope.py:778-803 _accumulate_tcspc_numba (times already jittered). */
void orc_accumulate_tcspc(uint32_t *flim, const int *px, const int *py,
                          const double *t, const int8_t *channel, int n,
                          double bin_w, int n_bins, int n_ch, int nx, int ny) {
    for (int p = 0; p < n; ++p) {
        int ch = channel[p];
        if (ch < 0 || ch >= n_ch) continue;
        int x = px[p], y = py[p];
        if (x < 0 || x >= nx || y < 0 || y >= ny) continue;
        double tt = t[p];
        if (tt < 0.0) tt = 0.0;
        long long b = (long long)(tt / bin_w);
        if (b >= n_bins) continue;
        flim[(((size_t)ch * nx + x) * ny + y) * n_bins + b] += 1;
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
