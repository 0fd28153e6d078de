"""Brownian dynamics + reactions on the GPU (reference dynamics.py).

``DynamicsEngine(config, state)`` keeps the reference's constructor and
``step()`` / ``run()`` interface and its derived constants
(dynamics.py:1724-2148): cell-list cutoff and grid, species potential
matrices, displacement cap, Collins-Kimball binding probability, catalytic
probability normalised by the effective contact area, exchange/conversion
probabilities and the feedback grids.  One step runs entirely on the device
in the order of dynamics.py:2150-2465.

Random numbers:
  rng="numpy"  (default) the reference's own numpy stream — the same
               Generator(seed) drawn in the same order and sizes — is
               injected each step, so events and positions follow the
               reference's for the same seed (host sync once per step);
  rng="philox" counter-based draws generated on the device (no host sync);
               statistically equivalent, used for throughput runs.
"""

from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, List, Optional

import numpy as np

from . import _hip
from ._hip import check, lib, ptr
from .config import ConfigError, SimulationConfig
from .particles import ParticleState

from .free_kernels import (CellList, DeviceArray, apply_binding_reactions,  # noqa: F401
                           apply_catalytic_reaction, apply_catalytic_with_feedback,
                           apply_catalytic_with_feedback_and_adaptation,
                           apply_diffusion_field_regions_smooth, apply_dissociation,
                           apply_membrane_detachment, apply_membrane_detachment_with_feedback,
                           apply_membrane_recruitment, apply_membrane_recruitment_local,
                           apply_membrane_recruitment_with_dual_feedback,
                           apply_membrane_recruitment_with_feedback, apply_species_conversion,
                           apply_species_conversion_with_feedback, bd_step,
                           compute_density_field, compute_forces_harmonic, compute_forces_wca,
                           update_adaptation_field)

# the reference's export list (dynamics.py:38-60) plus this module's helper
__all__ = [
    "CellList",
    "bd_step",
    "apply_binding_reactions",
    "apply_dissociation",
    "compute_forces_wca",
    "DynamicsEngine",
    "apply_catalytic_reaction",
    "apply_catalytic_with_feedback",
    "apply_membrane_recruitment",
    "apply_membrane_recruitment_with_feedback",
    "apply_membrane_recruitment_with_dual_feedback",
    "apply_membrane_detachment",
    "apply_membrane_detachment_with_feedback",
    "apply_species_conversion",
    "apply_species_conversion_with_feedback",
    "compute_density_field",
    "compute_forces_harmonic",
    "apply_diffusion_field_regions_smooth",
    "apply_catalytic_with_feedback_and_adaptation",
    "update_adaptation_field",
    "apply_membrane_recruitment_local",
    "derive_constants",
]


def _fb_key(fb, Lx, Ly):
    nxg = max(3, int(Lx / fb.grid_resolution))
    nyg = max(3, int(Ly / fb.grid_resolution))
    return (nxg, nyg, Lx / nxg, Ly / nyg)


def derive_constants(config: SimulationConfig) -> Dict:
    """Host-side constants of DynamicsEngine.__init__ (reference
    dynamics.py:1724-2148), returned as a plain dict."""
    sp = config.species
    name_to_id = {s.name: i for i, s in enumerate(sp)}
    Lx, Ly = (float(v) for v in config.membrane.size)
    dt = float(config.bd_timestep)

    def Dof(name):
        return sp[name_to_id[name]].diffusion_coefficient

    # cell-list cutoff (dynamics.py:1738-1782)
    cut_nm = config.max_potential_cutoff
    for r in config.reactions:
        cut_nm = max(cut_nm, r.contact_radius)
    for c in config.catalytic_reactions:
        cut_nm = max(cut_nm, c.contact_radius)
    cut_um = cut_nm / 1000.0
    for r in config.reactions:
        if r.is_bimolecular and r.kon > 0:
            s_bd = math.sqrt(4.0 * (Dof(r.reactants[0]) + Dof(r.reactants[1])) * dt)
            cut_um = max(cut_um, max(r.contact_radius / 1000.0, s_bd))
    for c in config.catalytic_reactions:
        s_bd = math.sqrt(4.0 * (Dof(c.enzyme) + Dof(c.substrate)) * dt)
        cut_um = max(cut_um, max(c.contact_radius / 1000.0, s_bd))
    cutoff = max(cut_um, cut_nm / 1000.0) * 1.1
    if cutoff < 0.001:
        cutoff = 0.05
    nx = max(3, int(math.floor(Lx / cutoff)))
    ny = max(3, int(math.floor(Ly / cutoff)))

    # potentials (dynamics.py:1784-1824)
    ns = len(sp)
    sig = np.zeros((ns, ns))
    eps = np.zeros((ns, ns))
    kh = np.zeros((ns, ns))
    use_wca = use_harm = False
    for pot in config.potentials:
        a, b = name_to_id[pot.species_a], name_to_id[pot.species_b]
        s = pot.parameters.get("sigma", 5.0) / 1000.0
        if pot.potential_type in ("wca", "lennard_jones"):
            use_wca = True
            sig[a, b] = sig[b, a] = s
            eps[a, b] = eps[b, a] = pot.parameters.get("epsilon", 2.5)
        elif pot.potential_type == "harmonic_repulsion":
            use_harm = True
            sig[a, b] = sig[b, a] = s
            k = pot.parameters.get("spring_constant", pot.parameters.get("epsilon", 50.0))
            kh[a, b] = kh[b, a] = k

    # displacement cap (dynamics.py:1826-1864)
    sig_min = float(sig[sig > 0].min()) if np.any(sig > 0) else math.inf
    D_max = max(s.diffusion_coefficient for s in sp)
    contact_max = max([c.contact_radius / 1000.0 for c in config.catalytic_reactions] + [0.0])
    if use_wca:
        max_disp = max(config.displacement_cap_fraction * sig_min, contact_max)
    elif use_harm:
        max_disp = max(5.0 * math.sqrt(2.0 * D_max * dt), contact_max,
                       sig_min if sig_min < math.inf else 0.1)
    elif sig_min < math.inf:
        max_disp = config.displacement_cap_fraction * sig_min
    else:
        max_disp = 0.1

    return dict(Lx=Lx, Ly=Ly, dt=dt, cutoff=cutoff, nx=nx, ny=ny, csx=Lx / nx, csy=Ly / ny,
                sigma=sig, epsilon=eps, k_harm=kh, use_wca=use_wca, use_harmonic=use_harm,
                max_disp=max_disp, name_to_id=name_to_id)


def binding_probability(config: SimulationConfig, rxn, dt: float):
    """Collins-Kimball 2D binding probability and detection radius
    (reference dynamics.py:1866-1963). Returns (sigma_eff, P_bind)."""
    sp = {s.name: s for s in config.species}
    Da = sp[rxn.reactants[0]].diffusion_coefficient
    Db = sp[rxn.reactants[1]].diffusion_coefficient
    Drel = Da + Db
    contact = rxn.contact_radius / 1000.0
    s_eff = max(contact, math.sqrt(4.0 * Drel * dt))
    A = math.pi * s_eff ** 2
    if not (A > 0 and rxn.kon > 0):
        return s_eff, 0.0
    Nb = config.initial_populations.get(rxn.reactants[1], 0)
    area = config.membrane.size[0] * config.membrane.size[1]
    cb = Nb / area if area > 0 else 0
    if cb > 0 and Drel > 0 and contact > 0:
        R = min(1.0 / math.sqrt(math.pi * cb), min(config.membrane.size) / 2.0)
        lr = math.log(R / contact)
        if lr > 0:
            inv = 1.0 / rxn.kon - lr / (2.0 * math.pi * Drel)
            k_int = 1.0 / inv if inv > 0 else 1e12
            return s_eff, min(1.0, k_int * dt / A)
    return s_eff, min(1.0, rxn.kon * dt / A)


class _Plan:
    """Reactions in the reference's execution order + the grids they read."""

    def __init__(self, config: SimulationConfig, consts: Dict):
        self.grids: List[tuple] = []
        self.reactions: List[_hip.Reaction] = []
        self.kinds: List[str] = []
        name_to_id = consts["name_to_id"]
        Lx, Ly, dt = consts["Lx"], consts["Ly"], consts["dt"]
        fbmap = {fb.name: fb for fb in config.feedback_rules}

        def grid_index(key):
            if key not in self.grids:
                self.grids.append(key)
            return self.grids.index(key)

        def feedback(name):
            if not name or name not in fbmap:
                return _hip.no_feedback()
            fb = fbmap[name]
            return _hip.Feedback(sensor=name_to_id[fb.sensor_species],
                                 grid=grid_index(_fb_key(fb, Lx, Ly)),
                                 inhibition=1 if fb.function_type == "hill_inhibition" else 0,
                                 pad=0, K_half=fb.K_half, n_hill=fb.n_hill,
                                 basal=fb.basal_fraction)

        def rx(kind, a=-1, b=-1, c=-1, contact=0.0, prob=0.0, prob2=0.0, maxr=0, adapt=0,
               fb1=None, fb2=None, label=""):
            self.reactions.append(_hip.Reaction(
                kind=kind, a=a, b=b, c=c, contact=contact, prob=prob, prob2=prob2,
                max_recruits=maxr, use_adapt=adapt, fb1=fb1 or _hip.no_feedback(),
                fb2=fb2 or _hip.no_feedback()))
            self.kinds.append(label)

        for r in config.reactions:
            if not r.is_bimolecular:
                continue
            s_eff, pb = binding_probability(config, r, dt)
            rx(_hip.EDM_RX_BIND, name_to_id[r.reactants[0]], name_to_id[r.reactants[1]],
               name_to_id[r.products[0]], contact=s_eff, prob=pb,
               prob2=min(1.0, r.koff * dt), label="bind")
        # adaptation (dynamics.py:2068-2096, ad-hoc config attributes)
        self.use_adapt = bool(getattr(config, "use_adaptation", False))
        pip3 = name_to_id.get("PIP3", -1)
        self.adapt_grid = -1
        self.adapt_species = -1
        self.adapt = (float(getattr(config, "adaptation_tau_w", 60.0)),
                      float(getattr(config, "adaptation_K_w", 10.0)),
                      float(getattr(config, "adaptation_n_w", 2.0)))
        if self.use_adapt and pip3 >= 0:
            anx, any_ = max(3, int(Lx / 1.0)), max(3, int(Ly / 1.0))
            self.adapt_grid = grid_index((anx, any_, Lx / anx, Ly / any_))
            self.adapt_species = pip3
            rx(_hip.EDM_RX_ADAPT, label="adapt")
        for c in config.catalytic_reactions:
            e, s, p = name_to_id[c.enzyme], name_to_id[c.substrate], name_to_id[c.product]
            Drel = config.species[e].diffusion_coefficient + config.species[s].diffusion_coefficient
            s_eff = max(c.contact_radius / 1000.0, math.sqrt(4.0 * Drel * dt))
            pc = min(1.0, c.k_cat * dt / max(math.pi * s_eff ** 2, 1e-30))
            fb = feedback(c.feedback_rule)
            adapt = 1 if (self.adapt_grid >= 0 and "PI3K" in c.enzyme and fb.sensor >= 0) else 0
            rx(_hip.EDM_RX_CAT, e, s, p, contact=s_eff, prob=pc, adapt=adapt, fb1=fb, label="cat")
        # density-biased recruitment (dynamics.py:2034-2041, 2116-2119, 2354):
        # an exchange whose rule name holds 'pip3_activation' and that has
        # exactly one feedback rule, when the ad-hoc config flag is set
        use_local = bool(getattr(config, "use_local_pi3k_recruitment", False))
        sigma = float(getattr(config, "recruit_sigma", 0.5))
        for ex in config.compartment_exchanges:
            sid = name_to_id[ex.species]
            fb1 = feedback(ex.feedback_rule)
            fb2 = feedback(ex.secondary_feedback_rule) if fb1.sensor >= 0 else _hip.no_feedback()
            local = (use_local and bool(ex.feedback_rule) and "pip3_activation" in ex.feedback_rule
                     and fb1.sensor >= 0 and fb2.sensor < 0)
            rx(_hip.EDM_RX_RECRUIT, sid, prob=min(1.0, ex.k_on * dt), maxr=ex.max_recruits_per_step,
               fb1=fb1, fb2=fb2, adapt=1 if local else 0, contact=sigma if local else 0.0,
               label="recruit_local" if local else "recruit")
            rx(_hip.EDM_RX_DETACH, sid, prob=min(1.0, ex.k_off * dt),
               fb1=feedback(ex.detachment_feedback_rule), label="detach")
        for cv in config.unimolecular_conversions:
            rx(_hip.EDM_RX_CONVERT, name_to_id[cv.from_species], name_to_id[cv.to_species],
               prob=min(1.0, cv.rate * dt), fb1=feedback(cv.feedback_rule), label="convert")


class BuiltParams:
    """An edm_params block with the host arrays its pointers refer to."""

    def __init__(self, params, keep):
        self.params = params
        self.keep = keep


def build_params(config: SimulationConfig, c: Dict, plan: "_Plan", n: int,
                 precision: str = "fp64", pair_cap: bool = True,
                 exact_force_order: bool = True) -> BuiltParams:
    """The C parameter block of one membrane (edm_params): geometry, the
    species tables, the D field, the reaction plan and the seed."""
    df = config.diffusion_field
    circles = [r for r in df.regions if r.get("shape") == "circle"]
    keep: List = []

    def arr(v):
        a = np.ascontiguousarray(np.asarray(v, dtype=np.float64))
        keep.append(a)
        return a.ctypes.data

    Lx, Ly = c["Lx"], c["Ly"]
    grids = (_hip.Grid * max(1, len(plan.grids)))()
    for i, (gx, gy, sx, sy) in enumerate(plan.grids):
        grids[i] = _hip.Grid(nx=gx, ny=gy, sx=sx, sy=sy)
    rxs = (_hip.Reaction * max(1, len(plan.reactions)))(*plan.reactions)
    keep += [grids, rxs]
    p = _hip.Params()
    p.n = n
    p.n_species = len(config.species)
    p.fp64 = 1 if precision == "fp64" else 0
    periodic_x, periodic_y = config.membrane.periodic
    p.periodic_x = 1 if periodic_x else 0
    p.periodic_y = 1 if periodic_y else 0
    p.ncx, p.ncy = c["nx"], c["ny"]
    p.use_wca = 1 if c["use_wca"] else 0
    p.use_harmonic = 1 if c["use_harmonic"] else 0
    p.dfield = 1 if df.type == "sinusoidal" else 0
    p.n_regions = 0 if p.dfield else len(circles)
    p.n_grids = len(plan.grids)
    p.n_reactions = len(plan.reactions)
    p.pair_cap = min(n * 30, 500000) if pair_cap else 0
    p.exact_force_order = 1 if exact_force_order else 0
    p.Lx, p.Ly, p.dt = Lx, Ly, c["dt"]
    p.cutoff, p.csx, p.csy = c["cutoff"], c["csx"], c["csy"]
    p.max_disp = c["max_disp"]
    p.smooth_w = config.region_smoothing_width
    if p.dfield:
        lx, ly = df.wavelength
        p.sin_amp = float(df.amplitude)
        p.sin_kx = 2.0 * math.pi / (lx if lx > 0 else Lx)
        p.sin_ky = 2.0 * math.pi / (ly if ly > 0 else Ly)
    tau_w, K_w, n_w = plan.adapt
    p.adapt_tau, p.adapt_K, p.adapt_n = tau_w, K_w, n_w
    p.adapt_grid, p.adapt_species = plan.adapt_grid, plan.adapt_species
    p.seed = int(config.random_seed) & 0xFFFFFFFFFFFFFFFF
    p.sigma, p.epsilon, p.k_harm = arr(c["sigma"]), arr(c["epsilon"]), arr(c["k_harm"])
    p.D_base = arr([s.diffusion_coefficient for s in config.species])
    p.rcx = arr([r["center"][0] for r in circles] or [0.0])
    p.rcy = arr([r["center"][1] for r in circles] or [0.0])
    p.rr = arr([r["radius"] for r in circles] or [0.0])
    p.rf = arr([r.get("D_factor", 1.0) for r in circles] or [0.0])
    p.grids = ctypes.cast(grids, ctypes.c_void_p)
    p.reactions = ctypes.cast(rxs, ctypes.c_void_p)
    return BuiltParams(p, keep)


class DynamicsEngine:
    """GPU DynamicsEngine with the reference's interface.

    Parameters beyond the reference's (config, state):
      precision  "fp64" (default) or "fp32"
      rng        "numpy" (default, reference streams) or "philox"
      pair_cap   True: keep the reference's min(30 N, 500000) pair-list cap
                 (default); False: no cap (every pair within the cutoff)
      exact_force_order
                 True (default): with numpy draws in FP64, sum pair forces in
                 the reference's pair order (bit-identical positions); False:
                 the LDS-tiled force kernel (positions agree to ~1e-15)
    """

    def __init__(self, config: SimulationConfig, state: ParticleState, precision: str = None,
                 rng: str = None, pair_cap: bool = True, exact_force_order: bool = True):
        self.config = config
        self.state = state
        precision = precision or os.environ.get("EDM_PRECISION", "fp64")
        rng = rng or os.environ.get("EDM_RNG", "numpy")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if rng not in ("numpy", "philox"):
            raise ValueError("rng must be 'numpy' or 'philox'")
        self.precision = precision
        self.rng_mode = rng
        self.pair_cap = pair_cap
        c = derive_constants(config)
        self._c = c
        self.Lx, self.Ly, self.dt = c["Lx"], c["Ly"], c["dt"]
        self.periodic_x, self.periodic_y = config.membrane.periodic
        self.cutoff_um = c["cutoff"]
        self.max_displacement = c["max_disp"]
        self.sigma_matrix, self.epsilon_matrix, self.harmonic_matrix = (
            c["sigma"], c["epsilon"], c["k_harm"])
        self._use_wca, self._use_harmonic = c["use_wca"], c["use_harmonic"]
        self.plan = _Plan(config, c)
        N = state.N
        built = build_params(config, c, self.plan, N, precision, pair_cap, exact_force_order)
        p = built.params
        self._keep = built.keep
        h = ctypes.c_void_p()
        check(lib().edm_create(ctypes.byref(p), ctypes.byref(h)), "edm_create")
        self._h = h
        self._params = p
        self._keep = []
        self._upload_from(state._a)
        state._attach(self)
        self._rng = np.random.default_rng(config.random_seed)
        self._ctr_cache = None
        self.n_steps_done = 0

    # ---- device state transfer ------------------------------------------
    def _host_struct(self, a) -> _hip.HostState:
        return _hip.HostState(
            positions=a["positions"].ctypes.data, species_id=a["species_id"].ctypes.data,
            bond_partner=a["bond_partner"].ctypes.data,
            is_alive=a["is_alive"].view(np.uint8).ctypes.data,
            has_dye=a["has_dye"].view(np.uint8).ctypes.data,
            is_bleached=a["is_bleached"].view(np.uint8).ctypes.data,
            dye_type_id=a["dye_type_id"].ctypes.data, ctmc_state=a["ctmc_state"].ctypes.data)

    def _upload_from(self, a) -> None:
        s = self._host_struct(a)
        check(lib().edm_upload_state(self._h, ctypes.byref(s)), "edm_upload_state")

    def _download_into(self, a) -> None:
        s = self._host_struct(a)
        check(lib().edm_download_state(self._h, ctypes.byref(s)), "edm_download_state")

    @property
    def handle(self):
        return self._h

    # ---- stepping (dynamics.py:2150-2465) --------------------------------
    def _numpy_draws(self, n_pairs: int) -> np.ndarray:
        """The reference's per-step draws in its order (dynamics.py:2194-2438)."""
        N = self.state.N
        g = self._rng
        parts = [g.standard_normal((N, 2)).ravel()]
        for r, kind in zip(self.plan.reactions, self.plan.kinds):
            if kind == "bind":
                parts.append(g.random(n_pairs))
                parts.append(g.random(N))
            elif kind == "cat":
                parts.append(g.random(n_pairs))
            elif kind in ("recruit", "recruit_local"):
                parts.append(g.random(N))
                parts.append(g.random((N, 2)).ravel())
                if kind == "recruit_local":
                    parts.append(g.random(N))
                    parts.append(g.standard_normal((N, 2)).ravel())
            elif kind in ("detach", "convert"):
                parts.append(g.random(N))
        return np.ascontiguousarray(np.concatenate(parts))

    def step(self) -> None:
        L = lib()
        if self.rng_mode == "numpy":
            npairs = ctypes.c_int64(0)
            check(L.edm_step_begin(self._h, ctypes.byref(npairs)), "edm_step_begin")
            d = self._numpy_draws(int(npairs.value))
            check(L.edm_step_finish(self._h, ptr(d), d.size), "edm_step_finish")
        else:
            check(L.edm_run(self._h, 1), "edm_run")
        self.n_steps_done += 1
        self._ctr_cache = None
        self.state.mark_stale()

    def run(self, n_steps: int, report_every: int = 0) -> None:
        if self.rng_mode == "philox" and report_every <= 0:
            check(lib().edm_run(self._h, int(n_steps)), "edm_run")
            self.n_steps_done += n_steps
            self._ctr_cache = None
            self.state.mark_stale()
            return
        for k in range(n_steps):
            self.step()
            if report_every > 0 and (k + 1) % report_every == 0:
                print(f"  t={(k + 1) * self.dt:.4f}s | sum bind={self.n_bindings} "
                      f"sum dissoc={self.n_dissociations}")

    def synchronize(self) -> None:
        """Wait for the queued steps; raises HipError if one of them
        overflowed a device buffer (EDM_ECAPACITY)."""
        check(lib().edm_synchronize(self._h), "edm_synchronize")

    def step_begin(self) -> int:
        """First half of a parity step: cell list on pre-move positions;
        returns the reference pair-list length."""
        npairs = ctypes.c_int64(0)
        check(lib().edm_step_begin(self._h, ctypes.byref(npairs)), "edm_step_begin")
        return int(npairs.value)

    def step_finish(self, draws: Optional[np.ndarray]) -> None:
        if draws is None:
            check(lib().edm_step_finish(self._h, None, 0), "edm_step_finish")
        else:
            d = np.ascontiguousarray(draws, dtype=np.float64)
            check(lib().edm_step_finish(self._h, ptr(d), d.size), "edm_step_finish")
        self.n_steps_done += 1
        self._ctr_cache = None
        self.state.mark_stale()

    def pairs(self) -> np.ndarray:
        """Reference-ordered pair list of the open parity step."""
        cap = self.state.N * 64
        out = np.empty((cap, 2), dtype=np.int32)
        n = ctypes.c_int64(0)
        check(lib().edm_pairs(self._h, ptr(out), cap, ctypes.byref(n)), "edm_pairs")
        return out[: min(int(n.value), cap)].copy()

    def density(self, species: str, grid_resolution: float) -> np.ndarray:
        sid = self._c["name_to_id"][species]
        nxg = max(3, int(self.Lx / grid_resolution))
        nyg = max(3, int(self.Ly / grid_resolution))
        key = (nxg, nyg, self.Lx / nxg, self.Ly / nyg)
        if key not in self.plan.grids:
            raise ValueError("no feedback grid with that resolution")
        g = self.plan.grids.index(key)
        out = np.empty(nxg * nyg)
        check(lib().edm_density(self._h, sid, g, ptr(out)), "edm_density")
        return out.reshape(nxg, nyg)

    # ---- tuning (no reference counterpart) ----------------------------------
    _TUNE = (("tile_rows", _hip.TUNE_TILE_ROWS), ("cat_side", _hip.TUNE_CAT_SIDE),
             ("cat_walk", _hip.TUNE_CAT_WALK), ("cat_lanes", _hip.TUNE_CAT_LANES),
             ("fret_search", _hip.TUNE_FRET_SEARCH), ("graphs", _hip.TUNE_GRAPHS),
             ("force_lists", _hip.TUNE_FORCE_LISTS))

    def set_tuning(self, tile_rows: Optional[int] = None, cat_side: Optional[int] = None,
                   cat_walk: Optional[int] = None, cat_lanes: Optional[int] = None,
                   fret_search: Optional[int] = None, graphs: Optional[int] = None,
                   force_lists: Optional[int] = None) -> None:
        """Work-decomposition knobs that leave every result unchanged: cell
        rows per strip of the tiled kernels (1..256), the side catalysis walks
        its pairs from (0 enzymes, 1 substrates, -1 the rarer one), how
        (1 work lists where possible, 0 LDS strips), the lanes per entity
        of the strip walk (a power of two in 1..64) and how proximity FRET
        finds its targets (0 a grid of the targets, 1 a scan of every
        particle per donor, the reference's search), and whether runs of
        steps are recorded into HIP graphs and replayed (1) or launched one
        by one (0)."""
        vals = dict(tile_rows=tile_rows, cat_side=cat_side, cat_walk=cat_walk,
                    cat_lanes=cat_lanes, fret_search=fret_search, graphs=graphs,
                    force_lists=force_lists)
        for name, key in self._TUNE:
            if vals[name] is not None:
                check(lib().edm_set_tuning(self._h, key, int(vals[name])), "edm_set_tuning")

    def tuning(self) -> Dict[str, int]:
        out = {}
        for name, key in self._TUNE:
            v = ctypes.c_int64(0)
            check(lib().edm_get_tuning(self._h, key, ctypes.byref(v)), "edm_get_tuning")
            out[name] = int(v.value)
        return out

    @property
    def graph_replays(self) -> int:
        """Step graphs replayed so far (see set_tuning(graphs=...))."""
        v = ctypes.c_int64(0)
        check(lib().edm_get_tuning(self._h, _hip.TUNE_GRAPH_REPLAYS, ctypes.byref(v)),
              "edm_get_tuning")
        return int(v.value)

    @property
    def graph_status(self) -> int:
        """0, or the code of the last failed graph recording (the library's
        header lists them); the membrane then steps launch by launch."""
        v = ctypes.c_int64(0)
        check(lib().edm_get_tuning(self._h, _hip.TUNE_GRAPH_STATUS, ctypes.byref(v)),
              "edm_get_tuning")
        return int(v.value)

    # ---- counters (dynamics.py:2068-2071, 2147-2148) ----------------------
    def _counters(self):
        if self._ctr_cache is None:
            out = np.zeros(6, dtype=np.int64)
            check(lib().edm_counters(self._h, ptr(out)), "edm_counters")
            self._ctr_cache = out
        return self._ctr_cache

    n_bindings = property(lambda self: int(self._counters()[0]))
    n_dissociations = property(lambda self: int(self._counters()[1]))
    n_catalytic = property(lambda self: int(self._counters()[2]))
    n_recruited = property(lambda self: int(self._counters()[3]))
    n_detached = property(lambda self: int(self._counters()[4]))
    n_converted = property(lambda self: int(self._counters()[5]))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.state.pull()
            self.state._device = None
            lib().edm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _hip._lib is not None:
                _hip._lib.edm_destroy(self._h)
                self._h = None
        except Exception:
            pass
