"""Batched ensemble: many independent membranes of one model on one GPU.

The reference simulates replicas one after another, one SimulationPipeline
per replica (pipeline.py:489-738).  ``EnsembleEngine`` keeps every resident
replica in one device pool and advances all of them with the same launches
(edm_ens_* in include/edm_hip.h).  A replica brings its own configuration
(box, particle counts, seed, rates, image size) as long as the reaction
network, species, potentials, fluorophores and time step are the model's.

Draws are Philox, keyed by the replica's seed, its own step count and its
own particle indices, so a replica's trajectory, events and FLIM stacks do
not depend on which other replicas are resident (tests/test_gpu_ensemble.py).
"""

from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Optional

import numpy as np

from . import _hip
from ._hip import check, lib, ptr
from .config import ConfigError, SimulationConfig
from .dynamics import _Plan, build_params, derive_constants
from .microscope import FLIMFrame, SparseFrame, image_params
from .particles import ParticleState
from .photophysics import photo_params


def _duty(config: SimulationConfig) -> float:
    return min(1.0, config.scan.dwell_time / config.scan.frame_interval)


def _model_signature(config: SimulationConfig, c: Dict, plan: _Plan):
    """Everything that must be shared by the replicas of one ensemble."""
    rx = [(r.kind, r.a, r.b, r.c, round(r.contact, 15), r.use_adapt,
           (r.fb1.sensor, r.fb1.grid, r.fb1.inhibition, r.fb1.K_half, r.fb1.n_hill, r.fb1.basal),
           (r.fb2.sensor, r.fb2.grid, r.fb2.inhibition, r.fb2.K_half, r.fb2.n_hill, r.fb2.basal))
          for r in plan.reactions]
    return (tuple(s.name for s in config.species),
            tuple(s.diffusion_coefficient for s in config.species),
            tuple(rx), len(plan.grids), c["cutoff"], c["max_disp"], c["dt"],
            c["sigma"].tobytes(), c["epsilon"].tobytes(), c["k_harm"].tobytes(),
            tuple(config.membrane.periodic), config.diffusion_field.type,
            float(config.region_smoothing_width),
            (plan.adapt_grid >= 0, plan.adapt_species, tuple(plan.adapt)),
            tuple(f.name for f in config.fluorophores), _duty(config))


class EnsembleEngine:
    """Resident replicas of one model, stepped together (FP32 by default).

    capacity      particles of the shared device pool
    max_replicas  resident replicas at most
    """

    def __init__(self, model: SimulationConfig, capacity: int, max_replicas: int = 1024,
                 precision: str = "fp32", image: bool = True):
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.image = image
        c = derive_constants(model)
        plan = _Plan(model, c)
        self._sig = _model_signature(model, c, plan)
        built = build_params(model, c, plan, 0, precision, pair_cap=False,
                             exact_force_order=False)
        pp = photo_params(model, _duty(model), 0) if image else None
        ip = image_params(model) if image else None
        h = ctypes.c_void_p()
        check(lib().edm_ens_create(ctypes.byref(built.params),
                                   ctypes.byref(pp) if pp is not None else None,
                                   ctypes.byref(ip) if ip is not None else None,
                                   int(capacity), int(max_replicas), ctypes.byref(h)),
              "edm_ens_create")
        self._h = h
        self.max_replicas = int(max_replicas)
        self.capacity = int(capacity)
        self._draining = 0  # removed replicas whose slots free up at the next step
        self._pooled = 0    # particles in the pool, removed replicas' included
        self._n_ch = ip.n_ch if ip is not None else 0
        self._n_bins = ip.n_bins if ip is not None else 0
        self.replicas: Dict[int, Dict] = {}
        self._tickets: Dict[int, List] = {}   # pending collects: (slot, uid, shape, frame)
        self._frame_meta: Dict[int, tuple] = {}  # per replica uid: cumulative dark, AP, detected
        self._uid = 0

    # ---- membership --------------------------------------------------------
    def add(self, config: SimulationConfig, state: Optional[ParticleState] = None) -> int:
        """Admit a replica (its initial state as ParticleState draws it);
        returns its slot."""
        c = derive_constants(config)
        plan = _Plan(config, c)
        if _model_signature(config, c, plan) != self._sig:
            raise ConfigError("replica does not share the ensemble's model (species, "
                              "reaction network, potentials, fluorophores, time step)")
        st = state if state is not None else ParticleState(config)
        a = st._a
        r = _hip.Replica()
        r.n = st.N
        r.ncx, r.ncy = c["nx"], c["ny"]
        r.n_grids = len(plan.grids)
        r.Lx, r.Ly = c["Lx"], c["Ly"]
        r.csx, r.csy, r.cutoff = c["csx"], c["csy"], c["cutoff"]
        df = config.diffusion_field
        if df.type == "sinusoidal":
            lx, ly = df.wavelength
            r.sin_kx = 2.0 * math.pi / (lx if lx > 0 else c["Lx"])
            r.sin_ky = 2.0 * math.pi / (ly if ly > 0 else c["Ly"])
        else:
            # the replica's own circular regions (dynamics.py:441-519)
            circles = [g for g in df.regions if g.get("shape") == "circle"]
            if len(circles) > 8:
                raise ConfigError("a batched replica holds at most 8 circular D(x) regions")
            r.n_regions = len(circles)
            for i, g in enumerate(circles):
                r.reg_cx[i], r.reg_cy[i] = float(g["center"][0]), float(g["center"][1])
                r.reg_r[i] = float(g["radius"])
                r.reg_f[i] = float(g.get("D_factor", 1.0))
        r.seed = int(config.random_seed) & 0xFFFFFFFFFFFFFFFF
        for i, (gx, gy, sx, sy) in enumerate(plan.grids):
            r.grids[i] = _hip.Grid(nx=gx, ny=gy, sx=sx, sy=sy)
        for i, rx in enumerate(plan.reactions):
            r.prob[i] = rx.prob
            r.prob2[i] = rx.prob2
            r.max_recruits[i] = rx.max_recruits
        r.n_reactions = len(plan.reactions)
        shape = None
        if self.image:
            ip = image_params(config)
            r.img_nx, r.img_ny, r.img_R = ip.nx, ip.ny, ip.R
            r.img_pixel, r.img_offx, r.img_offy = ip.pixel, ip.offx, ip.offy
            shape = (ip.n_ch, ip.nx, ip.ny, ip.n_bins)
        hs = _hip.HostState(
            positions=a["positions"].ctypes.data, species_id=a["species_id"].ctypes.data,
            bond_partner=a["bond_partner"].ctypes.data,
            is_alive=a["is_alive"].view(np.uint8).ctypes.data,
            has_dye=a["has_dye"].view(np.uint8).ctypes.data,
            is_bleached=a["is_bleached"].view(np.uint8).ctypes.data,
            dye_type_id=a["dye_type_id"].ctypes.data, ctmc_state=a["ctmc_state"].ctypes.data)
        slot = ctypes.c_int32(-1)
        check(lib().edm_ens_add(self._h, ctypes.byref(r), ctypes.byref(hs), ctypes.byref(slot)),
              "edm_ens_add")
        k = int(slot.value)
        self._pooled += st.N
        self._uid += 1
        self.replicas[k] = dict(config=config, N=st.N, shape=shape, frame_index=0, uid=self._uid,
                                frame_every=max(1, int(round(config.scan.frame_interval /
                                                             config.bd_timestep))))
        return k

    def remove(self, slot: int) -> None:
        """Remove a replica; its slot and pool space are reusable after the
        next step."""
        check(lib().edm_ens_remove(self._h, int(slot)), "edm_ens_remove")
        info = self.replicas.pop(slot, None)
        self._draining += 1
        self._retired = getattr(self, "_retired", 0) + (info["N"] if info else 0)

    def free_slots(self) -> int:
        return self.max_replicas - len(self.replicas) - self._draining

    def free_particles(self) -> int:
        return self.capacity - self._pooled

    # ---- stepping ------------------------------------------------------------
    def run(self, n_steps: int) -> None:
        """Advance every resident replica by n_steps (asynchronous)."""
        check(lib().edm_ens_run(self._h, int(n_steps), 1 if self.image else 0), "edm_ens_run")
        if n_steps > 0:
            self._draining = 0
            self._pooled -= getattr(self, "_retired", 0)
            self._retired = 0

    def synchronize(self) -> None:
        check(lib().edm_ens_synchronize(self._h), "edm_ens_synchronize")

    # ---- results -------------------------------------------------------------
    def frame(self, slot: int, download: bool = True) -> Optional[FLIMFrame]:
        """End the replica's current frame: dark counts and afterpulses, then
        (optionally) copy its TCSPC stack out, and reset it."""
        info = self.replicas[slot]
        nd, na = ctypes.c_int64(0), ctypes.c_int64(0)
        out = None
        n_el = 0
        if download:
            out = np.zeros(info["shape"], dtype=np.uint32)
            n_el = out.size
        check(lib().edm_ens_frame(self._h, int(slot), ptr(out) if out is not None else None,
                                  n_el, ctypes.byref(nd), ctypes.byref(na)), "edm_ens_frame")
        info["frame_index"] += 1
        if out is None:
            return None
        c = self.counters(slot)
        return FLIMFrame.from_stack(out, frame_index=info["frame_index"] - 1,
                                    n_signal=c["photons_detected"], n_dark=int(nd.value),
                                    n_ap=int(na.value))

    def end_frames(self, slots) -> None:
        """End the current frame of several replicas at once (detector noise
        and reset on the device, nothing copied out)."""
        sl = np.ascontiguousarray(np.asarray(list(slots), dtype=np.int32))
        if sl.size == 0:
            return
        check(lib().edm_ens_frame_end(self._h, ptr(sl), int(sl.size)), "edm_ens_frame_end")
        for k in sl.tolist():
            self.replicas[k]["frame_index"] += 1

    def collect_frames(self, slots) -> int:
        """End the current frame of several replicas and stream every one
        back (edm_ens_frames_collect): detector noise, compaction of each
        stack's nonzero bins on the device, reset; the copy to host memory
        runs while later steps do.  Returns a ticket for fetch_frames; at
        most two collects may be pending."""
        sl = np.ascontiguousarray(np.asarray(list(slots), dtype=np.int32))
        t = ctypes.c_int32(-1)
        check(lib().edm_ens_frames_collect(self._h, ptr(sl), int(sl.size), ctypes.byref(t)),
              "edm_ens_frames_collect")
        rec = []
        for k in sl.tolist():
            info = self.replicas[k]
            rec.append((k, info["uid"], info["shape"], info["frame_index"]))
            info["frame_index"] += 1
        self._tickets[int(t.value)] = rec
        return int(t.value)

    def fetch_frames(self, ticket: int) -> Dict[int, SparseFrame]:
        """The frames of a collect, by slot (waits for their copy)."""
        rec = self._tickets.pop(ticket)
        n = len(rec)
        counts = np.zeros(max(n, 1), dtype=np.int64)
        meta = np.zeros((max(n, 1), 3), dtype=np.int64)
        total = ctypes.c_int64(0)
        check(lib().edm_ens_frames_fetch(self._h, int(ticket), ptr(counts), ptr(meta), None,
                                         None, 0, ctypes.byref(total)), "edm_ens_frames_fetch")
        tot = int(total.value)
        idx = np.empty(max(tot, 1), dtype=np.uint32)
        cnt = np.empty(max(tot, 1), dtype=np.uint32)
        check(lib().edm_ens_frames_fetch(self._h, int(ticket), None, None, ptr(idx), ptr(cnt),
                                         idx.size, ctypes.byref(total)), "edm_ens_frames_fetch")
        out: Dict[int, SparseFrame] = {}
        off = 0
        for q, (k, uid, shape, fi) in enumerate(rec):
            c = int(counts[q])
            last = self._frame_meta.get(uid, (0, 0, 0))
            m = tuple(int(v) for v in meta[q])
            self._frame_meta[uid] = m
            out[k] = SparseFrame(shape, idx[off:off + c].copy(), cnt[off:off + c].copy(), fi,
                                 n_signal_photons=m[2] - last[2], n_dark_counts=m[0] - last[0],
                                 n_afterpulses=m[1] - last[1])
            off += c
        return out

    def noise_counters(self, slot: int) -> Dict[str, int]:
        """Dark counts and afterpulses added by end_frames so far."""
        out = np.zeros(2, dtype=np.int64)
        check(lib().edm_ens_noise_counters(self._h, int(slot), ptr(out)), "edm_ens_noise_counters")
        return dict(n_dark=int(out[0]), n_afterpulses=int(out[1]))

    def state(self, slot: int) -> Dict[str, np.ndarray]:
        """The replica's particle arrays in particle-index order."""
        n = self.replicas[slot]["N"]
        a = dict(positions=np.zeros((n, 2)), species_id=np.zeros(n, np.int32),
                 bond_partner=np.full(n, -1, np.int32), is_alive=np.zeros(n, np.bool_),
                 has_dye=np.zeros(n, np.bool_), is_bleached=np.zeros(n, np.bool_),
                 dye_type_id=np.full(n, -1, np.int32), ctmc_state=np.zeros(n, np.int8))
        hs = _hip.HostState(
            positions=a["positions"].ctypes.data, species_id=a["species_id"].ctypes.data,
            bond_partner=a["bond_partner"].ctypes.data,
            is_alive=a["is_alive"].view(np.uint8).ctypes.data,
            has_dye=a["has_dye"].view(np.uint8).ctypes.data,
            is_bleached=a["is_bleached"].view(np.uint8).ctypes.data,
            dye_type_id=a["dye_type_id"].ctypes.data, ctmc_state=a["ctmc_state"].ctypes.data)
        check(lib().edm_ens_download(self._h, int(slot), ctypes.byref(hs)), "edm_ens_download")
        return a

    _CTR = ("n_bindings", "n_dissociations", "n_catalytic", "n_recruited", "n_detached",
            "n_converted", "photons_emitted", "photons_detected")

    def counters(self, slot: int) -> Dict[str, int]:
        out = np.zeros(8, dtype=np.int64)
        check(lib().edm_ens_counters(self._h, int(slot), ptr(out)), "edm_ens_counters")
        return dict(zip(self._CTR, (int(v) for v in out)))

    # ---- tuning (no reference counterpart) ------------------------------------
    _TUNE = (("strip_particles", _hip.ENS_TUNE_STRIP_PARTICLES),
             ("pair_below", _hip.ENS_TUNE_PAIR_BELOW), ("cat_lanes", _hip.ENS_TUNE_CAT_LANES),
             ("cat_walk", _hip.ENS_TUNE_CAT_WALK))

    def set_tuning(self, strip_particles: Optional[int] = None, pair_below: Optional[int] = None,
                   cat_lanes: Optional[int] = None, cat_walk: Optional[int] = None) -> None:
        """Work decomposition of the batched step, leaving every result
        unchanged: own particles a strip aims at (32..512), the column
        population below which two columns share a strip (0..1024, 0 never),
        the lanes per catalysis entity (a power of two in 1..64) and whether
        catalysis walks strips of cells (0) or its enzymes from the pass
        work list (1).  Applies from the next step."""
        vals = dict(strip_particles=strip_particles, pair_below=pair_below, cat_lanes=cat_lanes,
                    cat_walk=cat_walk)
        for name, key in self._TUNE:
            if vals[name] is not None:
                check(lib().edm_ens_set_tuning(self._h, key, int(vals[name])),
                      "edm_ens_set_tuning")

    def tuning(self) -> Dict[str, int]:
        out = {}
        for name, key in self._TUNE:
            v = ctypes.c_int64(0)
            check(lib().edm_ens_get_tuning(self._h, key, ctypes.byref(v)), "edm_ens_get_tuning")
            out[name] = int(v.value)
        return out

    def info(self) -> Dict[str, int]:
        out = np.zeros(4, dtype=np.int64)
        check(lib().edm_ens_info(self._h, ptr(out)), "edm_ens_info")
        return dict(particles=int(out[0]), cells=int(out[1]), replicas=int(out[2]),
                    capacity=int(out[3]))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().edm_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None) and _hip._lib is not None:
                _hip._lib.edm_ens_destroy(self._h)
                self._h = None
        except Exception:
            pass


def replica_bytes(config: SimulationConfig, precision: str = "fp32", image: bool = True) -> int:
    """Device memory a replica occupies in an ensemble: particle pool entries
    (positions, post-move and sort copies, species/flags/index/CTMC/dye and
    their copies, sort keys, recruitment candidates), cell keys, density
    counts and its TCSPC stack."""
    c = derive_constants(config)
    n = sum(config.initial_populations.values()) + sum(
        ex.cytosol_pool for ex in config.compartment_exchanges if ex.cytosol_pool > 0)
    rb = 8 if precision == "fp64" else 4
    per_particle = 6 * rb + 2 * (1 + 1 + 4 + 1 + 1 + 2) + 3 * 4 + (4 + 4 + 8 + 8)
    cells = c["nx"] * c["ny"] * 2 * 4
    grids = 0
    plan = _Plan(config, c)
    for (gx, gy, _, _) in plan.grids:
        grids += gx * gy * 4 * 2 * max(1, len(config.species))
    flim = 0
    if image:
        ip = image_params(config)
        flim = ip.n_ch * ip.nx * ip.ny * ip.n_bins * 4
    return int(n * per_particle + cells + grids + flim)
