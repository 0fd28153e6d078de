"""Named workloads of the per-timestep path, built from the plain membrane.

* ``plain_membrane()``   10 x 10 um, 3500 alive particles, WCA, FP64 parity case
* ``wave_system()``      10 x 10 um, 13,000 particles, sinusoidal D(x), 128 px
* ``large_membrane()``   100 x 100 um, 1,000,000 particles, harmonic k = 1e7 kT/um^2
* ``ensemble_replica()`` one replica of the 512-replica batch (box side U[5,30] um)
* ``served_job()``       one job of the served stream (box side logU[5,50] um)
* ``binding_fret()``     small test system with PHAkt + PIP3 <-> PHAkt_PIP3 binding,
                         a circular slow-diffusion region and bond-mode FRET
                         (or proximity-mode FRET against PIP2)

Everything is derived from ``examples/pip3_pten_plain.yaml`` by rescaling the
box, the particle counts (species proportions kept), the cytosolic pools and
the recruitment caps, so the reaction network is the same in every workload.
"""

from __future__ import annotations

import copy
import math
import os
from typing import Any, Dict, Optional

import numpy as np
import yaml

from .config import SimulationConfig, config_from_dict

EXAMPLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples")
PLAIN_YAML = os.path.join(EXAMPLES, "pip3_pten_plain.yaml")

_SPECIES_ORDER = ("PIP2", "PIP3", "PI3K", "PTEN_T1", "PTEN_T2")


def plain_raw() -> Dict[str, Any]:
    with open(PLAIN_YAML, "r", encoding="utf-8") as f:
        return yaml.safe_load(f)


def _split_counts(total: int, fractions) -> Dict[str, int]:
    """Integer counts with the given proportions that sum to ``total``
    (largest-remainder rounding)."""
    raw = [total * f for f in fractions]
    base = [int(math.floor(r)) for r in raw]
    rem = total - sum(base)
    order = sorted(range(len(raw)), key=lambda k: raw[k] - base[k], reverse=True)
    for k in order[:rem]:
        base[k] += 1
    return dict(zip(_SPECIES_ORDER, base))


def rescale(raw: Dict[str, Any], side: float, n_alive: int,
            pools_scale: Optional[float] = None) -> Dict[str, Any]:
    """Return a copy of ``raw`` on a ``side`` x ``side`` box with ``n_alive``
    alive particles in the plain membrane's species proportions; cytosolic
    pools and per-step recruitment caps scale with the particle count."""
    out = copy.deepcopy(raw)
    base_pop = raw["initial_populations"]
    base_total = sum(base_pop[s] for s in _SPECIES_ORDER)
    fr = [base_pop[s] / base_total for s in _SPECIES_ORDER]
    out["initial_populations"] = _split_counts(int(n_alive), fr)
    scale = (n_alive / base_total) if pools_scale is None else pools_scale
    for ex in out.get("compartment_exchanges", []):
        ex["cytosol_pool"] = max(1, int(round(ex.get("cytosol_pool", 100) * scale)))
        ex["max_recruits_per_step"] = max(1, int(round(ex.get("max_recruits_per_step", 10) * scale)))
    out["membrane"]["size"] = [float(side), float(side)]
    return out


def plain_membrane(seed: int = 0) -> SimulationConfig:
    raw = plain_raw()
    raw["random_seed"] = seed
    return config_from_dict(raw)


def wave_system(seed: int = 0, n_steps: int = 20000) -> SimulationConfig:
    raw = rescale(plain_raw(), 10.0, 13000)
    raw["name"] = "PIP3/PTEN wave system (13k)"
    raw["diffusion_field"] = {"type": "sinusoidal", "amplitude": 0.3, "wavelength": [10.0, 10.0]}
    raw["scan"].update(pixels_x=128, pixels_y=128, pixel_size=78.125, frame_interval=0.1,
                       n_frames=max(1, n_steps // 100))
    raw["simulation_time"] = n_steps * raw["bd_timestep"]
    raw["random_seed"] = seed
    return config_from_dict(raw)


def _harmonic_potentials(k_kT_per_um2: float, sigma_nm: float):
    pots = []
    for a in range(len(_SPECIES_ORDER)):
        for b in range(a, len(_SPECIES_ORDER)):
            pots.append({"species_a": _SPECIES_ORDER[a], "species_b": _SPECIES_ORDER[b],
                         "potential_type": "harmonic_repulsion",
                         "parameters": {"sigma": sigma_nm, "spring_constant": k_kT_per_um2,
                                        "interaction_distance": sigma_nm}})
    return pots


def large_membrane(seed: int = 0, side: float = 100.0, n_alive: int = 1_000_000,
                   pixels: int = 1024, proximity_target: Optional[str] = None,
                   proximity_radius_nm: float = 50.0) -> SimulationConfig:
    """1M particles on 100 x 100 um, harmonic k = 10 kT/nm^2 = 1e7 kT/um^2,
    r_c = sigma = 10 nm, density grid 200 x 200 (0.5 um cells).  With
    ``proximity_target`` the PIP3-borne mTFP1 is a proximity biosensor of
    that species (E = 0.5 within the radius, else 0.02)."""
    raw = rescale(plain_raw(), side, n_alive)
    if proximity_target:
        raw["fret_pairs"][0].update(fret_mode="proximity",
                                    proximity_target_species=proximity_target,
                                    proximity_radius=float(proximity_radius_nm),
                                    proximity_E_high=0.5, proximity_E_low=0.02)
    raw["name"] = f"PIP3/PTEN large membrane ({n_alive})"
    raw["potentials"] = _harmonic_potentials(1.0e7, 10.0)
    raw["scan"].update(pixels_x=pixels, pixels_y=pixels, pixel_size=side * 1000.0 / pixels,
                       frame_interval=0.1)
    raw["random_seed"] = seed
    raw["ram_limit_gb"] = 1.0e6
    return config_from_dict(raw)


def ensemble_replica(r: int, rng_seed: int = 12345) -> SimulationConfig:
    """Replica ``r`` of the batched ensemble: side ~ U[5,30] um, count scaled to
    the plain membrane's density (35 / um^2), k_cat and k_on jittered
    log-uniformly by +/-50 %, seed r, 64 x 64 px image."""
    g = np.random.default_rng([rng_seed, r])
    side = float(g.uniform(5.0, 30.0))
    n = int(round(35.0 * side * side))
    raw = rescale(plain_raw(), side, n)
    for c in raw["catalytic_reactions"]:
        c["k_cat"] = float(c["k_cat"] * math.exp(g.uniform(math.log(0.5), math.log(1.5))))
    for e in raw["compartment_exchanges"]:
        e["k_on"] = float(e["k_on"] * math.exp(g.uniform(math.log(0.5), math.log(1.5))))
    raw["scan"].update(pixels_x=64, pixels_y=64, pixel_size=side * 1000.0 / 64,
                       frame_interval=0.05)
    raw["random_seed"] = int(r)
    raw["name"] = f"ensemble replica {r}"
    return config_from_dict(raw)


def served_job(j: int, rng_seed: int = 777, max_density: float = 100.0):
    """Job ``j`` of the served stream: side logU[5,50] um, particle count
    U[1k, 250k] clamped to ``max_density`` per um^2 (independent draws would
    reach 10^4 / um^2), steps U[500, 5000], image 64..256 px.  Returns
    (config, n_steps)."""
    g = np.random.default_rng([rng_seed, j])
    side = float(math.exp(g.uniform(math.log(5.0), math.log(50.0))))
    n = int(g.integers(1000, 250001))
    n = int(min(n, max_density * side * side))
    n = max(n, 200)
    steps = int(g.integers(500, 5001))
    px = int(g.choice([64, 128, 256]))
    raw = rescale(plain_raw(), side, n)
    raw["scan"].update(pixels_x=px, pixels_y=px, pixel_size=side * 1000.0 / px,
                       frame_interval=0.1)
    raw["random_seed"] = int(j)
    raw["name"] = f"served job {j}"
    return config_from_dict(raw), steps


def binding_fret(seed: int = 3, intramolecular: bool = True, side: float = 4.0,
                 reflective: bool = False, regions: bool = True,
                 harmonic: bool = False, proximity: bool = False) -> SimulationConfig:
    """Small test system exercising binding/dissociation, bond FRET, the
    smooth circular D(x) regions and (optionally) reflective walls.  With
    ``proximity`` the FRET pair is a proximity biosensor: a PHAkt donor is
    quenched (E = 0.5) while a PIP2 lies within 50 nm, else E = 0.02."""
    raw = rescale(plain_raw(), side, int(round(35 * side * side)))
    raw["species"] += [
        {"name": "PHAkt", "diffusion_coefficient": 0.08, "radius": 3.0},
        {"name": "PHAkt_PIP3", "diffusion_coefficient": 0.06, "radius": 3.5},
    ]
    raw["initial_populations"]["PHAkt"] = int(6 * side * side)
    raw["initial_populations"]["PHAkt_PIP3"] = 0
    raw["reactions"] = [{"name": "PHAkt_PIP3_binding", "reactants": ["PHAkt", "PIP3"],
                         "products": ["PHAkt_PIP3"], "kon": 0.5, "koff": 2.0,
                         "contact_radius": 30.0}]
    raw["compartment_exchanges"].append(
        {"name": "PHAkt_membrane_cytosol", "species": "PHAkt", "k_on": 0.5, "k_off": 0.5,
         "cytosol_pool": int(10 * side * side), "max_recruits_per_step": 3,
         "feedback_rule": "pip3_activation"})
    if intramolecular:
        raw["dye_attachments"] = [{"species": "PHAkt", "fluorophore": "mTFP1"}]
    else:
        raw["fret_pairs"][0]["intramolecular"] = False
        raw["dye_attachments"] = [{"species": "PHAkt", "fluorophore": "mTFP1"},
                                  {"species": "PIP3", "fluorophore": "Venus",
                                   "labeling_efficiency": 0.7}]
    if proximity:
        raw["fret_pairs"][0].update(fret_mode="proximity", proximity_target_species="PIP2",
                                    proximity_radius=50.0, proximity_E_high=0.5,
                                    proximity_E_low=0.02)
    if regions:
        raw["diffusion_field"] = {"type": "regions", "regions": [
            {"shape": "circle", "center": [side / 2, side / 2], "radius": side / 5,
             "D_factor": 0.3}]}
    if reflective:
        raw["membrane"]["periodic"] = [False, False]
    if harmonic:
        raw["potentials"] = _harmonic_potentials(50.0, 8.0)
    # bright enough that every step carries photons
    raw["scan"].update(pixels_x=32, pixels_y=32, pixel_size=side * 1000.0 / 32,
                       dwell_time=2.0e-4, frame_interval=0.02)
    raw["random_seed"] = seed
    return config_from_dict(raw)
