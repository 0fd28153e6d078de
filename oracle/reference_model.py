"""ORACLE driver — test infrastructure only.

Replays the reference's per-step composition on the CPU in FP64:

* ``OracleState``       — particles.py:80-243 (init RNG order, SoA arrays)
* ``OracleDynamics``    — dynamics.py:1724-2148 (derived constants) and
                          dynamics.py:2150-2465 (step order and numpy draws)
* ``OraclePhotons``     — photophysics.py:59-331, 690-1197
* ``OracleMicroscope``  — microscope.py:73-391, 650-1203

Every numba kernel is replaced by its C restatement in oracle/edm_oracle.c;
numpy Generators are used exactly where the reference uses them, in the same
order and with the same sizes, so that the oracle consumes the same random
streams as the reference does for the same seed.  Derived constants are
restated here from the reference source, independently of the product
package (the product's own host code is compared against them in tests).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module.

parity unpinned: the compositions (OracleDynamics.step, the photon step,
process_photons / finalize_frame) follow the reference source text; no
reference test or fixture fixes their multi-step output.  They are frozen by
the golden vectors of tests/golden/ (oracle/README.md lists the pinning of
every kernel they call).
"""

from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libedm_oracle.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i8p = ctypes.POINTER(ctypes.c_int8)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build_recruit_cdf(density_field, nx_grid, ny_grid) -> np.ndarray:
    """dynamics.py:2355-2370: running sum of max(density, 0.01) over the
    cells in index order (gx major), divided by the total; the uniform CDF
    when the total is zero."""
    dens = np.asarray(density_field, dtype=np.float64).reshape(-1)
    nc = nx_grid * ny_grid
    cdf = np.empty(nc)
    total = 0.0
    for k in range(nc):
        total += max(float(dens[k]), 0.01)
        cdf[k] = total
    if total > 0:
        for k in range(nc):
            cdf[k] /= total
    else:
        for k in range(nc):
            cdf[k] = (k + 1.0) / nc
    return cdf



def lib():
    """Load (building first if needed) the oracle shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from build_oracle import build  # type: ignore
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_cell_list_pairs.restype = ctypes.c_int
        _lib.orc_cell_list_pairs_par.restype = ctypes.c_int
        _lib.orc_poisson_draw.restype = ctypes.c_int
        _lib.orc_poisson_draw.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_hill.restype = ctypes.c_double
        _lib.orc_hill.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, ctypes.c_double]
        _lib.orc_py_mod.restype = ctypes.c_double
        _lib.orc_py_mod.argtypes = [ctypes.c_double, ctypes.c_double]
        _lib.orc_min_image_dsq.restype = ctypes.c_double
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def P(a: Optional[np.ndarray]):
    """ctypes pointer for a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


D = ctypes.c_double
I = ctypes.c_int
LL = ctypes.c_longlong


# ----------------------------------------------------------------------
# Particle state (particles.py:80-243)
# ----------------------------------------------------------------------

class OracleState:
    def __init__(self, config, rng: Optional[np.random.Generator] = None):
        if rng is None:
            rng = np.random.default_rng(config.random_seed)
        self.species_names = [s.name for s in config.species]
        name_to_id = {n: i for i, n in enumerate(self.species_names)}
        counts = []
        N = 0
        for sp in config.species:
            c = config.initial_populations.get(sp.name, 0)
            counts.append((name_to_id[sp.name], c))
            N += c
        pools = []
        for ex in config.compartment_exchanges:
            if ex.cytosol_pool > 0:
                sid = name_to_id.get(ex.species, -1)
                if sid >= 0:
                    pools.append((sid, N, ex.cytosol_pool))
                    N += ex.cytosol_pool
        self.N = N
        self.positions = np.zeros((N, 2))
        self.species_id = np.zeros(N, dtype=np.int32)
        self.bond_partner = np.full(N, -1, dtype=np.int32)
        self.has_dye = np.zeros(N, dtype=np.uint8)
        self.dye_type_id = np.full(N, -1, dtype=np.int32)
        self.ctmc_state = np.zeros(N, dtype=np.int8)
        self.is_bleached = np.zeros(N, dtype=np.uint8)
        self.is_alive = np.ones(N, dtype=np.uint8)
        Lx, Ly = config.membrane.size
        off = 0
        for sid, c in counts:
            self.species_id[off:off + c] = sid
            self.positions[off:off + c, 0] = rng.uniform(0.0, Lx, size=c)
            self.positions[off:off + c, 1] = rng.uniform(0.0, Ly, size=c)
            off += c
        for sid, start, c in pools:
            self.species_id[start:start + c] = sid
            self.positions[start:start + c, 0] = rng.uniform(0.0, Lx, size=c)
            self.positions[start:start + c, 1] = rng.uniform(0.0, Ly, size=c)
            self.is_alive[start:start + c] = 0
        fl = {f.name: i for i, f in enumerate(config.fluorophores)}
        for att in config.dye_attachments:
            sid = name_to_id.get(att.species, -1)
            if sid < 0:
                continue
            fid = fl[att.fluorophore]
            idx = np.where(self.species_id == sid)[0]
            if len(idx) == 0:
                continue
            lab = rng.random(len(idx)) < att.labeling_efficiency
            sel = idx[lab]
            self.has_dye[sel] = 1
            self.dye_type_id[sel] = fid
            self.ctmc_state[sel] = 0


# ----------------------------------------------------------------------
# Dynamics (dynamics.py:1724-2465)
# ----------------------------------------------------------------------

def _fb_info(fb, Lx, Ly, name_to_id):
    nxg = max(3, int(Lx / fb.grid_resolution))
    nyg = max(3, int(Ly / fb.grid_resolution))
    return dict(sensor=name_to_id[fb.sensor_species], nxg=nxg, nyg=nyg,
                sx=Lx / nxg, sy=Ly / nyg, K=fb.K_half, n=fb.n_hill,
                inh=1 if fb.function_type == "hill_inhibition" else 0,
                basal=fb.basal_fraction, density=np.zeros(nxg * nyg))


class OracleDynamics:
    """Faithful FP64 restatement of DynamicsEngine (host constants + step)."""

    def __init__(self, config, state: OracleState, pair_cap: bool = True,
                 n_threads: int = 1, all_loops: bool = False):
        """n_threads: threads of bd_step (the reference's prange); with
        all_loops the pair list and the pair forces run on them too (a
        timing variant: the same pairs in the same order, forces summed
        with atomic adds)."""
        self.config = config
        self.state = state
        self.pair_cap = pair_cap
        self.n_threads = n_threads
        self.all_loops = all_loops
        Lx, Ly = config.membrane.size
        self.Lx, self.Ly = float(Lx), float(Ly)
        self.px, self.py = (1 if p else 0 for p in config.membrane.periodic)
        self.dt = config.bd_timestep
        sp = config.species
        name_to_id = {s.name: i for i, s in enumerate(sp)}
        self.name_to_id = name_to_id

        # cutoff (dynamics.py:1738-1782)
        cut_nm = config.max_potential_cutoff
        for r in config.reactions:
            cut_nm = max(cut_nm, r.contact_radius)
        for c in config.catalytic_reactions:
            cut_nm = max(cut_nm, c.contact_radius)
        cut_um = cut_nm / 1000.0
        for r in config.reactions:
            if r.is_bimolecular and r.kon > 0:
                Da = sp[name_to_id[r.reactants[0]]].diffusion_coefficient
                Db = sp[name_to_id[r.reactants[1]]].diffusion_coefficient
                se = max(r.contact_radius / 1000.0, math.sqrt(4.0 * (Da + Db) * self.dt))
                cut_um = max(cut_um, se)
        for c in config.catalytic_reactions:
            De = sp[name_to_id[c.enzyme]].diffusion_coefficient
            Ds = sp[name_to_id[c.substrate]].diffusion_coefficient
            se = max(c.contact_radius / 1000.0, math.sqrt(4.0 * (De + Ds) * self.dt))
            cut_um = max(cut_um, se)
        self.cutoff = max(cut_um, cut_nm / 1000.0) * 1.1
        if self.cutoff < 0.001:
            self.cutoff = 0.05
        self.nx = max(3, int(math.floor(self.Lx / self.cutoff)))
        self.ny = max(3, int(math.floor(self.Ly / self.cutoff)))
        self.csx = self.Lx / self.nx
        self.csy = self.Ly / self.ny

        # potentials (dynamics.py:1784-1824)
        ns = len(sp)
        self.n_sp = ns
        self.sig = np.zeros((ns, ns))
        self.eps = np.zeros((ns, ns))
        self.kh = np.zeros((ns, ns))
        self.use_wca = self.use_harm = False
        for pot in config.potentials:
            a, b = name_to_id[pot.species_a], name_to_id[pot.species_b]
            if pot.potential_type in ("wca", "lennard_jones"):
                self.use_wca = True
                s = pot.parameters.get("sigma", 5.0) / 1000.0
                e = pot.parameters.get("epsilon", 2.5)
                self.sig[a, b] = self.sig[b, a] = s
                self.eps[a, b] = self.eps[b, a] = e
            elif pot.potential_type == "harmonic_repulsion":
                self.use_harm = True
                s = pot.parameters.get("sigma", 5.0) / 1000.0
                k = pot.parameters.get("spring_constant", pot.parameters.get("epsilon", 50.0))
                self.sig[a, b] = self.sig[b, a] = s
                self.kh[a, b] = self.kh[b, a] = k

        # displacement cap (dynamics.py:1826-1864)
        pos_sig = self.sig[self.sig > 0]
        sig_min = float(pos_sig.min()) if pos_sig.size else np.inf
        D_max = max(s.diffusion_coefficient for s in sp)
        contact_max = 0.0
        for c in config.catalytic_reactions:
            contact_max = max(contact_max, c.contact_radius / 1000.0)
        if self.use_wca:
            self.max_disp = max(config.displacement_cap_fraction * sig_min, contact_max)
        elif self.use_harm:
            rms = np.sqrt(2.0 * D_max * config.bd_timestep)
            self.max_disp = max(5.0 * rms, contact_max, sig_min if sig_min < np.inf else 0.1)
        elif sig_min < np.inf:
            self.max_disp = config.displacement_cap_fraction * sig_min
        else:
            self.max_disp = 0.1

        # binding reactions (dynamics.py:1866-1963)
        self.bind = []
        for r in config.reactions:
            if not r.is_bimolecular:
                continue
            ra, rb = name_to_id[r.reactants[0]], name_to_id[r.reactants[1]]
            prod = name_to_id[r.products[0]]
            contact = r.contact_radius / 1000.0
            Drel = sp[ra].diffusion_coefficient + sp[rb].diffusion_coefficient
            s_bd = math.sqrt(4.0 * Drel * self.dt)
            se = max(contact, s_bd)
            A = math.pi * (se ** 2)
            if A > 0 and r.kon > 0:
                Nb = config.initial_populations.get(r.reactants[1], 0)
                area = config.membrane.size[0] * config.membrane.size[1]
                cb = Nb / area if area > 0 else 0
                Lh = min(config.membrane.size) / 2.0
                if cb > 0 and Drel > 0 and contact > 0:
                    R = min(1.0 / math.sqrt(math.pi * cb), Lh)
                    lr = math.log(R / contact)
                    if lr > 0:
                        inv = 1.0 / r.kon - lr / (2.0 * math.pi * Drel)
                        kint = 1.0 / inv if inv > 0 else 1e12
                        pb = min(1.0, kint * self.dt / A)
                    else:
                        pb = min(1.0, r.kon * self.dt / A)
                else:
                    pb = min(1.0, r.kon * self.dt / A)
            else:
                pb = 0.0
            self.bind.append(dict(ra=ra, rb=rb, prod=prod, contact=se, prob=pb,
                                  pdis=min(1.0, r.koff * self.dt)))

        fbmap = {fb.name: fb for fb in config.feedback_rules}

        def fbi(name):
            if name and name in fbmap:
                return _fb_info(fbmap[name], self.Lx, self.Ly, name_to_id)
            return None

        # catalytic (dynamics.py:1965-2004)
        self.cat = []
        for c in config.catalytic_reactions:
            e, s, p = name_to_id[c.enzyme], name_to_id[c.substrate], name_to_id[c.product]
            Drel = sp[e].diffusion_coefficient + sp[s].diffusion_coefficient
            se = max(c.contact_radius / 1000.0, math.sqrt(4.0 * Drel * self.dt))
            A = math.pi * se ** 2
            pc = min(1.0, c.k_cat * self.dt / max(A, 1e-30))
            self.cat.append(dict(e=e, s=s, p=p, contact=se, prob=pc, fb=fbi(c.feedback_rule),
                                 adapt='PI3K' in c.enzyme))
        # exchanges (dynamics.py:2006-2041)
        self.exch = []
        for ex in config.compartment_exchanges:
            self.exch.append(dict(
                sid=name_to_id[ex.species], pon=min(1.0, ex.k_on * self.dt),
                poff=min(1.0, ex.k_off * self.dt), maxr=ex.max_recruits_per_step,
                fb=fbi(ex.feedback_rule), fb2=fbi(ex.secondary_feedback_rule),
                fbd=fbi(ex.detachment_feedback_rule),
                local=bool(ex.feedback_rule and 'pip3_activation' in ex.feedback_rule)))
        # conversions (dynamics.py:2043-2060)
        self.conv = []
        for cv in config.unimolecular_conversions:
            self.conv.append(dict(f=name_to_id[cv.from_species], t=name_to_id[cv.to_species],
                                  prob=min(1.0, cv.rate * self.dt), fb=fbi(cv.feedback_rule)))
        # adaptation (dynamics.py:2068-2096) and local recruitment (:2098-2119)
        self.use_adapt = getattr(config, 'use_adaptation', False)
        self.tau_w = getattr(config, 'adaptation_tau_w', 60.0)
        self.K_w = getattr(config, 'adaptation_K_w', 10.0)
        self.n_w = getattr(config, 'adaptation_n_w', 2.0)
        self.anx = max(3, int(self.Lx / 1.0))
        self.any = max(3, int(self.Ly / 1.0))
        self.asx, self.asy = self.Lx / self.anx, self.Ly / self.any
        self.w = np.zeros(self.anx * self.any)
        self.pip3_rho = np.zeros(self.anx * self.any)
        self.pip3_id = name_to_id.get('PIP3', -1)
        self.use_local = getattr(config, 'use_local_pi3k_recruitment', False)
        self.recruit_sigma = getattr(config, 'recruit_sigma', 0.5)

        # D field (dynamics.py:2121-2134)
        df = config.diffusion_field
        self.D_base = np.array([s.diffusion_coefficient for s in sp], dtype=np.float64)
        circ = [r for r in df.regions if r.get("shape") == "circle"]
        self.nreg = len(circ)
        self.rcx = np.array([r["center"][0] for r in circ], dtype=np.float64)
        self.rcy = np.array([r["center"][1] for r in circ], dtype=np.float64)
        self.rr = np.array([r["radius"] for r in circ], dtype=np.float64)
        self.rf = np.array([r.get("D_factor", 1.0) for r in circ], dtype=np.float64)
        self.smooth_w = config.region_smoothing_width
        self.sinus = df.type == "sinusoidal"
        if self.sinus:
            lx, ly = df.wavelength
            self.sin_amp = float(df.amplitude)
            self.sin_kx = 2.0 * math.pi / (lx if lx > 0 else self.Lx)
            self.sin_ky = 2.0 * math.pi / (ly if ly > 0 else self.Ly)
        N = state.N
        self.Dp = np.empty(N)
        self.gradD = np.zeros((N, 2))
        self.forces = np.zeros((N, 2))
        self.head = np.full(self.nx * self.ny, -1, dtype=np.int32)
        self.next = np.empty(N, dtype=np.int32)
        self.rng = np.random.default_rng(config.random_seed)
        self.n_bindings = self.n_dissociations = 0
        self.n_catalytic = self.n_recruited = self.n_detached = self.n_converted = 0
        self.last_pairs = None
        self.last_n_found = 0

    # -- helpers -------------------------------------------------------
    def _density(self, fb):
        st = self.state
        fb['density'][:] = 0.0
        lib().orc_density(P(st.positions), P(st.species_id), P(st.is_alive), I(st.N),
                          I(fb['sensor']), I(fb['nxg']), I(fb['nyg']), D(fb['sx']),
                          D(fb['sy']), P(fb['density']))

    def max_pairs(self):
        N = self.state.N
        return min(N * 30, 500000) if self.pair_cap else max(1, N * 64)

    def build_pairs(self):
        st = self.state
        L = lib()
        L.orc_build_cell_list(P(st.positions), I(st.N), P(self.head), P(self.next),
                              I(self.nx), I(self.ny), D(self.csx), D(self.csy))
        mp = self.max_pairs()
        pairs = np.empty((mp, 2), dtype=np.int32)
        found = ctypes.c_longlong(0)
        args = (P(st.positions), I(st.N), P(self.head), P(self.next), I(self.nx), I(self.ny),
                D(self.csx), D(self.csy), D(self.Lx), D(self.Ly), I(self.px), I(self.py),
                D(self.cutoff), LL(mp), P(pairs), ctypes.byref(found))
        if self.all_loops:
            n = L.orc_cell_list_pairs_par(*args, I(self.n_threads))
            if n < 0:
                raise MemoryError("orc_cell_list_pairs_par: no memory for the pair offsets")
        else:
            n = L.orc_cell_list_pairs(*args)
        self.last_n_found = found.value
        return np.ascontiguousarray(pairs[:n])

    # -- one step (dynamics.py:2150-2465) --------------------------------
    def step(self):
        st = self.state
        L = lib()
        N = st.N
        pos, sp, alive, bp = st.positions, st.species_id, st.is_alive, st.bond_partner
        Lx, Ly, px, py = D(self.Lx), D(self.Ly), I(self.px), I(self.py)
        if self.sinus:
            L.orc_diffusion_field_sinusoidal(P(self.Dp), P(self.gradD), P(self.D_base), P(sp),
                                             P(pos), I(N), D(self.sin_amp), D(self.sin_kx),
                                             D(self.sin_ky))
        else:
            L.orc_diffusion_field_smooth(P(self.Dp), P(self.gradD), P(self.D_base), P(sp), P(pos),
                                         I(N), P(self.rcx), P(self.rcy), P(self.rr), P(self.rf),
                                         I(self.nreg), D(self.smooth_w))
        pairs = self.build_pairs()
        npairs = pairs.shape[0]
        self.last_pairs = pairs
        self.forces[:] = 0.0
        if self.all_loops:
            if self.use_wca:
                L.orc_forces_par(I(0), P(pos), P(self.forces), P(sp), P(alive), P(pairs),
                                 I(npairs), P(self.sig), P(self.eps), I(self.n_sp), Lx, Ly, px,
                                 py, I(self.n_threads))
            if self.use_harm:
                L.orc_forces_par(I(1), P(pos), P(self.forces), P(sp), P(alive), P(pairs),
                                 I(npairs), P(self.sig), P(self.kh), I(self.n_sp), Lx, Ly, px,
                                 py, I(self.n_threads))
        else:
            if self.use_wca:
                L.orc_forces_wca(P(pos), P(self.forces), P(sp), P(alive), P(pairs), I(npairs),
                                 P(self.sig), P(self.eps), I(self.n_sp), Lx, Ly, px, py)
            if self.use_harm:
                L.orc_forces_harmonic(P(pos), P(self.forces), P(sp), P(alive), P(pairs),
                                      I(npairs), P(self.sig), P(self.kh), I(self.n_sp), Lx, Ly,
                                      px, py)
        noise = self.rng.standard_normal((N, 2))
        L.orc_bd_step(P(pos), P(self.Dp), P(self.forces), P(self.gradD), D(self.dt), Lx, Ly,
                      px, py, I(N), P(noise), P(alive), D(self.max_disp), I(self.n_threads))
        for r in self.bind:
            u = self.rng.random(npairs)
            self.n_bindings += L.orc_binding(P(pos), P(sp), P(bp), P(alive), I(r['ra']),
                                             I(r['rb']), I(r['prod']), D(r['contact']),
                                             D(r['prob']), P(pairs), I(npairs), P(u), Lx, Ly,
                                             px, py)
            u = self.rng.random(N)
            self.n_dissociations += L.orc_dissociation(P(sp), P(bp), P(alive), P(pos), I(N),
                                                       I(r['prod']), I(r['ra']), I(r['rb']),
                                                       D(r['pdis']), P(u))
        if self.use_adapt and self.pip3_id >= 0:
            self.pip3_rho[:] = 0.0
            L.orc_density(P(pos), P(sp), P(alive), I(N), I(self.pip3_id), I(self.anx),
                          I(self.any), D(self.asx), D(self.asy), P(self.pip3_rho))
            L.orc_adaptation_update(P(self.w), P(self.pip3_rho), I(self.anx), I(self.any),
                                    D(self.dt), D(self.tau_w))
        for c in self.cat:
            u = self.rng.random(npairs)
            fb = c['fb']
            if fb is not None:
                self._density(fb)
                adapt = self.w if (self.use_adapt and c['adapt']) else None
                self.n_catalytic += L.orc_catalytic(
                    P(pos), P(sp), P(alive), I(c['e']), I(c['s']), I(c['p']),
                    D(c['contact']), D(c['prob']), P(pairs), I(npairs), P(u), Lx, Ly, px, py,
                    P(fb['density']), I(fb['nxg']), I(fb['nyg']), D(fb['sx']), D(fb['sy']),
                    D(fb['K']), D(fb['n']), I(fb['inh']), D(fb['basal']), P(adapt),
                    I(self.anx), I(self.any), D(self.K_w), D(self.n_w))
            else:
                self.n_catalytic += L.orc_catalytic(
                    P(pos), P(sp), P(alive), I(c['e']), I(c['s']), I(c['p']),
                    D(c['contact']), D(c['prob']), P(pairs), I(npairs), P(u), Lx, Ly, px, py,
                    None, I(0), I(0), D(1.0), D(1.0), D(0.0), D(0.0), I(0), D(0.0), None,
                    I(1), I(1), D(0.0), D(0.0))
        for ex in self.exch:
            u_acc = self.rng.random(N)
            u_pos = self.rng.random((N, 2))
            fb, fb2 = ex['fb'], ex['fb2']
            if fb is not None and fb2 is not None:
                self._density(fb)
                self._density(fb2)
                nr = L.orc_recruit(P(sp), P(alive), P(pos), I(N), I(ex['sid']), D(ex['pon']),
                                   Lx, Ly, P(u_acc), P(u_pos), I(ex['maxr']),
                                   P(fb['density']), I(fb['nxg']), I(fb['nyg']), D(fb['sx']),
                                   D(fb['sy']), D(fb['K']), D(fb['n']), I(fb['inh']),
                                   D(fb['basal']),
                                   P(fb2['density']), I(fb2['nxg']), I(fb2['nyg']), D(fb2['sx']),
                                   D(fb2['sy']), D(fb2['K']), D(fb2['n']), I(fb2['inh']),
                                   D(fb2['basal']))
            elif fb is not None:
                self._density(fb)
                if self.use_local and ex['local']:
                    cdf = build_recruit_cdf(fb['density'], fb['nxg'], fb['nyg'])
                    u_loc = self.rng.random(N)
                    off = self.rng.standard_normal((N, 2))
                    nr = L.orc_recruit_local(P(sp), P(alive), P(pos), I(N), I(ex['sid']),
                                             D(ex['pon']), Lx, Ly, P(u_acc), P(off),
                                             I(ex['maxr']), P(fb['density']), I(fb['nxg']),
                                             I(fb['nyg']), D(fb['sx']), D(fb['sy']), D(fb['K']),
                                             D(fb['n']), I(fb['inh']), D(fb['basal']), P(cdf),
                                             P(u_loc), D(self.recruit_sigma))
                else:
                    nr = L.orc_recruit(P(sp), P(alive), P(pos), I(N), I(ex['sid']), D(ex['pon']),
                                       Lx, Ly, P(u_acc), P(u_pos), I(ex['maxr']),
                                       P(fb['density']), I(fb['nxg']), I(fb['nyg']), D(fb['sx']),
                                       D(fb['sy']), D(fb['K']), D(fb['n']), I(fb['inh']),
                                       D(fb['basal']),
                                       None, I(0), I(0), D(1.0), D(1.0), D(0.0), D(0.0), I(0),
                                       D(0.0))
            else:
                nr = L.orc_recruit(P(sp), P(alive), P(pos), I(N), I(ex['sid']), D(ex['pon']),
                                   Lx, Ly, P(u_acc), P(u_pos), I(ex['maxr']),
                                   None, I(0), I(0), D(1.0), D(1.0), D(0.0), D(0.0), I(0), D(0.0),
                                   None, I(0), I(0), D(1.0), D(1.0), D(0.0), D(0.0), I(0), D(0.0))
            self.n_recruited += nr
            u_det = self.rng.random(N)
            fbd = ex['fbd']
            if fbd is not None:
                self._density(fbd)
                nd = L.orc_detach(P(sp), P(alive), P(pos), I(N), I(ex['sid']), D(ex['poff']),
                                  P(u_det), P(fbd['density']), I(fbd['nxg']), I(fbd['nyg']),
                                  D(fbd['sx']), D(fbd['sy']), D(fbd['K']), D(fbd['n']),
                                  I(fbd['inh']), D(fbd['basal']))
            else:
                nd = L.orc_detach(P(sp), P(alive), P(pos), I(N), I(ex['sid']), D(ex['poff']),
                                  P(u_det), None, I(0), I(0), D(1.0), D(1.0), D(0.0), D(0.0),
                                  I(0), D(0.0))
            self.n_detached += nd
        for cv in self.conv:
            u = self.rng.random(N)
            fb = cv['fb']
            if fb is not None:
                self._density(fb)
                nc = L.orc_convert(P(sp), P(alive), P(pos), I(N), I(cv['f']), I(cv['t']),
                                   D(cv['prob']), P(u), P(fb['density']), I(fb['nxg']),
                                   I(fb['nyg']), D(fb['sx']), D(fb['sy']), D(fb['K']),
                                   D(fb['n']), I(fb['inh']), D(fb['basal']))
            else:
                nc = L.orc_convert(P(sp), P(alive), P(pos), I(N), I(cv['f']), I(cv['t']),
                                   D(cv['prob']), P(u), None, I(0), I(0), D(1.0), D(1.0),
                                   D(0.0), D(0.0), I(0), D(0.0))
            self.n_converted += nc

    def counters(self) -> Dict[str, int]:
        return dict(n_bindings=self.n_bindings, n_dissociations=self.n_dissociations,
                    n_catalytic=self.n_catalytic, n_recruited=self.n_recruited,
                    n_detached=self.n_detached, n_converted=self.n_converted)


# ----------------------------------------------------------------------
# Photophysics (photophysics.py:59-331, 690-1197)
# ----------------------------------------------------------------------

class _Rates:
    """RateMatrix restatement (photophysics.py:59-151)."""

    def __init__(self, fl):
        self.states = fl.states
        n = len(fl.states)
        idx = {s.name: i for i, s in enumerate(fl.states)}
        Q = np.zeros((n, n))
        for a, tr in fl.thermal_rates.items():
            for b, r in tr.items():
                Q[idx[a], idx[b]] = r
        for i in range(n):
            Q[i, i] = 0.0
            Q[i, i] = -np.sum(Q[i, :])
        self.Q = Q
        self.exit = np.array([-Q[i, i] for i in range(n)])
        self.branch = np.zeros((n, n))
        for i in range(n):
            if self.exit[i] > 0:
                for j in range(n):
                    if j != i:
                        self.branch[i, j] = Q[i, j] / self.exit[i]
        self.gs = next((i for i, s in enumerate(fl.states) if s.is_absorbing and not s.is_fluorescent), 0)
        self.es = next((i for i, s in enumerate(fl.states) if s.is_fluorescent), -1)
        self.t1 = -1
        self.bl = -1
        for i, s in enumerate(fl.states):
            if not s.is_absorbing and not s.is_fluorescent and not s.is_terminal and self.t1 < 0:
                self.t1 = i
            if s.is_terminal and self.bl < 0:
                self.bl = i

    def lifetime(self, st):
        k = self.exit[st]
        return float('inf') if k <= 0 else 1.0 / k


def _pss(rm: _Rates, p_exc, T_laser, n_pulses, k_fret=0.0, QYd=1.0, QYa=1.0):
    """_build_pss_matrices restatement (photophysics.py:218-331)."""
    s0, s1, t1, bl = rm.gs, rm.es, rm.t1, rm.bl
    kS1 = rm.exit[s1] + k_fret if (k_fret > 0 and s1 >= 0) else rm.exit[s1]
    if kS1 > 0 and s1 >= 0:
        p_rad = rm.Q[s1, s0] / kS1
        p_fr = k_fret / kS1 if k_fret > 0 else 0.0
        p_isc = rm.Q[s1, t1] / kS1 if t1 >= 0 else 0.0
        p_bl1 = rm.Q[s1, bl] / kS1 if bl >= 0 else 0.0
    else:
        p_rad = p_fr = p_isc = p_bl1 = 0.0
    p_ret = p_rad + p_fr
    if t1 >= 0:
        kT1 = rm.exit[t1]
        pTx = 1.0 - math.exp(-kT1 * T_laser) if kT1 > 0 else 0.0
        pT0 = rm.branch[t1, s0]
        pTb = rm.branch[t1, bl] if bl >= 0 else 0.0
    else:
        pTx = pT0 = pTb = 0.0
    T = np.zeros((3, 3))
    T[0, 0] = (1.0 - p_exc) + p_exc * p_ret
    T[0, 1] = p_exc * p_isc
    T[0, 2] = p_exc * p_bl1
    T[1, 0] = pTx * pT0
    T[1, 1] = 1.0 - pTx
    T[1, 2] = pTx * pTb
    T[2, 2] = 1.0
    for row in range(3):
        rs = T[row].sum()
        if abs(rs - 1.0) > 1e-12:
            T[row] /= rs
    T_step = np.linalg.matrix_power(T, n_pulses)
    Ts = T[:2, :2].copy()
    I2 = np.eye(2)
    Tsn = np.linalg.matrix_power(Ts, n_pulses)
    try:
        S = (I2 - Tsn) @ np.linalg.inv(I2 - Ts)
    except np.linalg.LinAlgError:
        S = np.zeros((2, 2))
        Tk = I2.copy()
        for _ in range(n_pulses):
            S += Tk
            Tk = Tk @ Ts
    pd = p_exc * p_rad * QYd
    mean_ph = np.zeros(2) if pd <= 0.0 else np.array([S[0, 0] * pd, S[1, 0] * pd])
    pf = p_exc * p_fr * QYa
    mean_fr = np.zeros(2) if pf <= 0.0 else np.array([S[0, 0] * pf, S[1, 0] * pf])
    tau = 1.0 / kS1 if kS1 > 0 else 4.1e-9
    return T_step, mean_ph, mean_fr, tau


class OraclePhotons:
    def __init__(self, config, state: OracleState, scan_duty_fraction: float = 1.0):
        self.config = config
        self.state = state
        self.rms = [_Rates(f) for f in config.fluorophores]
        nf = len(self.rms)
        self.nf = nf
        self.active = nf > 0
        if not self.active:
            return
        self.duty = max(0.0, min(1.0, scan_duty_fraction))
        self.T_laser = config.laser.period
        self.n_pulses = max(1, int(round(config.bd_timestep / self.T_laser)))
        self.p_exc = np.array([self._pexc(f) for f in config.fluorophores])
        fl = {f.name: i for i, f in enumerate(config.fluorophores)}
        self.fl = fl
        sp_id = {s.name: i for i, s in enumerate(config.species)}
        self.sp_id = sp_id
        self.fmaps = []
        for fp in config.fret_pairs:
            self.fmaps.append(dict(R0=fp.R0, r=fp.intra_complex_distance,
                                   d=fl.get(fp.donor_type, -1), a=fl.get(fp.acceptor_type, -1),
                                   cfg=fp))
        self.QY = np.ones(nf)
        for t, rm in enumerate(self.rms):
            if rm.es >= 0:
                self.QY[t] = config.fluorophores[t].states[rm.es].quantum_yield
        self.T_step = np.zeros((nf, 3, 3))
        self.T_step_f = np.zeros((nf, 3, 3))
        self.mean_ph = np.zeros((nf, 2))
        self.mean_ph_f = np.zeros((nf, 2))
        self.mean_fr = np.zeros((nf, 2))
        self.mean_fr_f = np.zeros((nf, 2))
        self.tau = np.zeros(nf)
        self.tau_f = np.zeros(nf)
        self.k_bound = np.zeros(nf)
        for t, rm in enumerate(self.rms):
            pe = self.p_exc[t] * self.duty
            QYa = 1.0
            for fm in self.fmaps:
                if fm['d'] == t:
                    if 0 <= fm['a'] < nf:
                        QYa = float(self.QY[fm['a']])
                    break
            T, m, mf, tau = _pss(rm, pe, self.T_laser, self.n_pulses, 0.0, float(self.QY[t]), QYa)
            self.T_step[t], self.mean_ph[t], self.mean_fr[t], self.tau[t] = T, m, mf, tau
            kb = self._kbound(t)
            self.k_bound[t] = kb
            if kb > 0:
                T, m, mf, tau = _pss(rm, pe, self.T_laser, self.n_pulses, kb, float(self.QY[t]), QYa)
            self.T_step_f[t], self.mean_ph_f[t], self.mean_fr_f[t], self.tau_f[t] = T, m, mf, tau
        self.gs = np.array([r.gs for r in self.rms], dtype=np.int32)
        self.es = np.array([r.es for r in self.rms], dtype=np.int32)
        self.t1 = np.array([r.t1 for r in self.rms], dtype=np.int32)
        self.bl = np.array([r.bl for r in self.rms], dtype=np.int32)
        n_lab = int(np.sum(state.has_dye))
        mm = max(float(self.mean_ph[t, 0]) for t in range(nf))
        self.max_photons = max(1000, int(n_lab * mm * 3))
        N = state.N
        self.k_fret = np.zeros(N)
        self.acc = np.full(N, -1, dtype=np.int32)
        self.rng = np.random.default_rng(config.random_seed + 1000)

    def _pexc(self, fl):
        """_compute_excitation_prob restatement (photophysics.py:953-991)."""
        h, c = 6.626e-34, 3.0e8
        las = self.config.laser
        Eph = h * c / (las.wavelength * 1e-9)
        A = math.pi * (las.beam_waist * 1e-9) ** 2
        nph = (las.power / las.repetition_rate) / Eph
        flux = nph / A
        sig = fl.absorption_cross_section * 1e-4
        if fl.excitation_wavelength > 0:
            dl = abs(las.wavelength - fl.excitation_wavelength)
            if dl > 1.0:
                fwhm = getattr(fl, 'absorption_fwhm', 45.0)
                sig *= math.exp(-0.5 * (dl / (fwhm / 2.355)) ** 2)
        return min(sig * flux, 0.9)

    def _kbound(self, t):
        """_get_bound_fret_rate restatement (photophysics.py:914-951)."""
        for fm in self.fmaps:
            if fm['d'] == t:
                fp = fm['cfg']
                rm = self.rms[t]
                tD = rm.lifetime(rm.es)
                if tD <= 0 or math.isinf(tD):
                    return 0.0
                if getattr(fp, 'fret_mode', 'bond') == 'proximity':
                    if fp.proximity_E_high >= 1.0:
                        return 1e15
                    return fp.proximity_E_high / ((1.0 - fp.proximity_E_high) * tD)
                ratio = fm['R0'] / fm['r']
                return ratio ** 6 / tD
        return 0.0

    def update_fret_rates(self):
        """photophysics.py:1021-1114"""
        st = self.state
        L = lib()
        self.k_fret[:] = 0.0
        self.acc[:] = -1
        Lx, Ly = self.config.membrane.size
        px, py = (1 if p else 0 for p in self.config.membrane.periodic)
        for fm in self.fmaps:
            fp = fm['cfg']
            if getattr(fp, 'fret_mode', 'bond') == 'proximity':
                continue
            d = fm['d']
            if d < 0 or d >= self.nf:
                continue
            rm = self.rms[d]
            tD = rm.lifetime(rm.es)
            if tD <= 0 or math.isinf(tD):
                continue
            L.orc_fret_rates_bond(P(self.k_fret), P(self.acc), P(st.has_dye), P(st.is_bleached),
                                  P(st.bond_partner), P(st.dye_type_id), I(st.N), D(fm['R0']),
                                  D(tD), I(d), I(fm['a']), D(fm['r']),
                                  I(1 if getattr(fp, 'intramolecular', False) else 0))
        for fp in self.config.fret_pairs:
            if getattr(fp, 'fret_mode', 'bond') != 'proximity':
                continue
            d = self.fl.get(fp.donor_type, -1)
            if d < 0 or d >= self.nf:
                continue
            rm = self.rms[d]
            tD = rm.lifetime(rm.es)
            if tD <= 0 or math.isinf(tD):
                continue
            tgt = self.sp_id.get(fp.proximity_target_species, -1)
            if tgt < 0:
                continue
            Eh, El = fp.proximity_E_high, fp.proximity_E_low
            kh = Eh / ((1.0 - Eh) * tD) if Eh < 1.0 else 1e15
            kl = El / ((1.0 - El) * tD) if El < 1.0 else 1e15
            L.orc_fret_rates_proximity(P(self.k_fret), P(self.acc), P(st.positions),
                                       P(st.species_id), P(st.has_dye), P(st.is_alive),
                                       P(st.is_bleached), P(st.dye_type_id), I(st.N), I(d),
                                       I(tgt), D(fp.proximity_radius * 1e-3), D(kh), D(kl),
                                       D(Lx), D(Ly), I(px), I(py))

    def evolve_one_bd_step(self):
        """photophysics.py:1116-1197; returns a dict PhotonBatch."""
        if not self.active:
            return dict(particle_idx=np.empty(0, np.int32), arrival_time=np.empty(0),
                        channel=np.empty(0, np.int8), is_fret=np.empty(0, np.uint8), n_photons=0)
        st = self.state
        N = st.N
        self.update_fret_rates()
        u_p = self.rng.random(N)
        u_t = self.rng.random(N)
        mm = 0.0
        for t in range(self.nf):
            mm = max(mm, max(float(self.mean_ph[t, 0]), float(self.mean_ph_f[t, 0])))
        est = max(1000, int(N * mm * 2.5))
        u_a = self.rng.random(est)
        mp = self.max_photons
        op = np.empty(mp, np.int32)
        ot = np.empty(mp)
        oc = np.empty(mp, np.int8)
        opl = np.empty(mp, np.int32)
        of = np.empty(mp, np.uint8)
        n = lib().orc_pss_kernel(
            P(st.ctmc_state), P(st.is_bleached), P(st.has_dye), P(st.is_alive), P(st.dye_type_id),
            P(self.k_fret), P(self.acc), I(self.nf), P(self.T_step), P(self.T_step_f),
            P(self.mean_ph), P(self.mean_ph_f), P(self.mean_fr), P(self.mean_fr_f), P(self.tau),
            P(self.k_bound), P(self.gs), P(self.es), P(self.t1), P(self.bl), P(op), P(ot), P(oc),
            P(opl), P(of), I(mp), I(N), P(u_p), P(u_t), P(u_a), I(est))
        return dict(particle_idx=op[:n].copy(), arrival_time=ot[:n].copy(),
                    channel=oc[:n].copy(), is_fret=of[:n].copy(), n_photons=n)


# ----------------------------------------------------------------------
# Microscope (microscope.py:73-391, 650-1203)
# ----------------------------------------------------------------------

class OracleMicroscope:
    def __init__(self, config, state: OracleState):
        self.config = config
        self.state = state
        NA = config.objective.NA
        self.sigma = {}
        for ch, fl in enumerate(config.fluorophores):
            lam_em = config.psf.emission_wavelength
            for s in fl.states:
                if s.is_fluorescent and s.emission_wavelength > 0:
                    lam_em = s.emission_wavelength
                    break
            si = 0.21 * config.laser.wavelength / NA
            sd = 0.21 * lam_em / NA
            self.sigma[ch] = 1.0 / math.sqrt(1.0 / si ** 2 + 1.0 / sd ** 2) / 1000.0
        self.pixel = config.scan.pixel_size / 1000.0
        self.nx, self.ny = config.scan.pixels_x, config.scan.pixels_y
        Lx, Ly = config.membrane.size
        self.offx = (Lx - self.nx * self.pixel) / 2.0
        self.offy = (Ly - self.ny * self.pixel) / 2.0
        ms = max(self.sigma.values()) if self.sigma else 0.1
        self.R = int(math.ceil(4.0 * ms / self.pixel))
        self.pcx = (np.arange(self.nx) + 0.5) * self.pixel + self.offx
        self.pcy = (np.arange(self.ny) + 0.5) * self.pixel + self.offy
        nch = max(self.sigma.keys()) + 1 if self.sigma else 1
        self.sig_ch = np.zeros(max(nch, 2))
        for ch, s in self.sigma.items():
            self.sig_ch[ch] = s
        self.n_bins = config.tcspc.n_bins
        self.bin_w = config.tcspc.bin_width
        self.n_ch = max(2, len(config.fluorophores))
        f2s = 1.0 / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        self.sigma_irf = math.sqrt((config.laser.pulse_width * f2s) ** 2
                                   + config.detector.timing_jitter ** 2)
        self.flim = np.zeros((self.n_ch, self.nx, self.ny, self.n_bins), dtype=np.uint32)
        self.qe = config.detector.detection_efficiency
        self.dark_rate = config.detector.dark_count_rate
        self.p_ap = config.detector.afterpulsing_prob
        self.dwell = config.scan.dwell_time
        self.rng = np.random.default_rng(config.random_seed + 2000)
        self.n_signal = 0

    def process_photons(self, batch) -> int:
        n = batch['n_photons']
        if n == 0:
            return 0
        det = self.rng.random(n) < self.qe
        di = np.where(det)[0]
        if len(di) == 0:
            return 0
        pidx = batch['particle_idx'][di]
        pos = self.state.positions
        ok = (pidx >= 0) & (pidx < pos.shape[0])
        if not np.all(ok):
            di = di[ok]
            pidx = pidx[ok]
        ns = len(di)
        if ns == 0:
            return 0
        ch = batch['channel'][di]
        arr = batch['arrival_time'][di]
        order = np.argsort(pidx, kind='mergesort')
        pidx_s = np.ascontiguousarray(pidx[order]).astype(np.int32)
        ch_s = np.ascontiguousarray(ch[order]).astype(np.int8)
        arr_s = arr[order]
        u = self.rng.random(ns)
        opx = np.empty(ns, np.int32)
        opy = np.empty(ns, np.int32)
        dmask = np.zeros(ns, np.uint8)
        posx = np.ascontiguousarray(pos[:, 0])
        posy = np.ascontiguousarray(pos[:, 1])
        na = lib().orc_assign_emitter_centric(
            P(pidx_s), P(ch_s), P(posx), P(posy), I(ns), P(self.pcx), P(self.pcy), I(self.nx),
            I(self.ny), D(self.pixel), D(self.offx), D(self.offy), I(self.R), P(self.sig_ch),
            I(self.sig_ch.size), P(u), P(opx), P(opy), P(dmask))
        if na == 0:
            return 0
        ai = np.where(dmask.astype(bool))[0]
        t = arr_s[ai] + self.rng.normal(0.0, self.sigma_irf, size=na)
        t = np.ascontiguousarray(t)
        cha = np.ascontiguousarray(ch_s[ai])
        lib().orc_accumulate_tcspc(P(self.flim), P(opx), P(opy), P(t), P(cha), I(na),
                                   D(self.bin_w), I(self.n_bins), I(self.n_ch), I(self.nx),
                                   I(self.ny))
        self.n_signal += na
        return na

    def finalize_frame(self):
        """Dark counts then afterpulsing (microscope.py:859-920, 1159-1203)."""
        flim = self.flim
        n_ch, nx, ny, nb = flim.shape
        n_dark = 0
        lam = self.dark_rate * self.dwell * 1
        if lam >= 1e-10:
            for c in range(n_ch):
                for i in range(nx):
                    for j in range(ny):
                        k = self.rng.poisson(lam)
                        if k > 0:
                            for b in self.rng.integers(0, nb, size=k):
                                flim[c, i, j, b] += 1
                            n_dark += k
        n_ap = 0
        if self.p_ap > 0:
            for c in range(n_ch):
                for i in range(nx):
                    for j in range(ny):
                        tot = int(flim[c, i, j].sum())
                        if tot == 0:
                            continue
                        k = self.rng.binomial(tot, self.p_ap)
                        if k > 0:
                            for b in self.rng.integers(0, nb, size=k):
                                flim[c, i, j, b] += 1
                            n_ap += k
        out = dict(flim_stack=flim.copy(), n_signal=self.n_signal, n_dark=n_dark, n_ap=n_ap)
        flim[:] = 0
        self.n_signal = 0
        return out


class OraclePipeline:
    """Frame loop of pipeline.py:489-738 (no sub-frame recording)."""

    def __init__(self, config, pair_cap: bool = True, n_threads: int = 1):
        self.config = config
        self.state = OracleState(config)
        self.dyn = OracleDynamics(config, self.state, pair_cap=pair_cap, n_threads=n_threads)
        duty = min(1.0, config.scan.dwell_time / config.scan.frame_interval)
        self.ph = OraclePhotons(config, self.state, duty)
        self.mic = OracleMicroscope(config, self.state)
        self.n_bd_per_frame = int(round(config.scan.frame_interval / config.bd_timestep))

    def run_steps(self, n_steps: int, image: bool = True):
        emitted = detected = 0
        for _ in range(n_steps):
            self.dyn.step()
            b = self.ph.evolve_one_bd_step()
            emitted += b['n_photons']
            if image:
                detected += self.mic.process_photons(b)
        return emitted, detected
