"""Simulation configuration: dataclasses and the YAML loader.

Keeps the reference's schema (ed_mrdt/config.py:27-50 names, :61-490 fields
and defaults, :947-1094 YAML keys) so that existing YAML files and scripts
that set attributes on a loaded config keep working.  Two quirks of the
reference loader are kept on purpose (config.py:1061-1089):
``displacement_cap_fraction``, ``region_smoothing_width`` and a top-level
``n_frames`` are not read from YAML, so their dataclass defaults apply.

Additions for the GPU path (all optional, default = reference behaviour):
``diffusion_field.type == "sinusoidal"`` with ``amplitude`` / ``wavelength``
(a smooth D(x) = D_s (1 + a sin(2 pi x/lx) sin(2 pi y/ly)) field), read by
the dynamics engine only.
"""

from __future__ import annotations

import math
import warnings as _warnings
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Dict, List, Optional, Tuple, Union

import yaml

__all__ = [
    "Species", "Reaction", "PairPotential", "DiffusionField",
    "CatalyticReaction", "CompartmentExchange", "UnimolecularConversion",
    "FeedbackRule", "MembraneGeometry", "PhotoState", "FluorophoreType",
    "FRETpair", "DyeAttachment", "ObjectiveLens", "PSFmodel", "Detector",
    "LaserSource", "ScanParameters", "TCSPCsettings", "SimulationConfig",
    "load_config", "config_from_dict", "ConfigError",
]


class ConfigError(Exception):
    """Raised for an invalid configuration (reference config.py:57)."""


@dataclass
class Species:
    name: str
    diffusion_coefficient: float
    radius: float
    color: str = "#FFFFFF"
    states: List[str] = field(default_factory=lambda: ["default"])
    state_transitions: Dict[str, Dict[str, float]] = field(default_factory=dict)
    D_per_state: Optional[Dict[str, float]] = None

    def get_D(self, state: str = "default") -> float:
        if self.D_per_state and state in self.D_per_state:
            return self.D_per_state[state]
        return self.diffusion_coefficient


@dataclass
class Reaction:
    name: str
    reactants: List[str]
    products: List[str]
    kon: float = 0.0
    koff: float = 0.0
    contact_radius: float = 5.0
    Keq: Optional[float] = None
    f_bound: Optional[float] = None
    required_state: Optional[Dict[str, str]] = None
    product_state: Optional[Dict[str, str]] = None

    @property
    def is_bimolecular(self) -> bool:
        return len(self.reactants) == 2

    @property
    def is_reversible(self) -> bool:
        return self.koff > 0.0


@dataclass
class PairPotential:
    species_a: str
    species_b: str
    potential_type: str
    parameters: Dict[str, float] = field(default_factory=dict)

    def get_cutoff(self) -> float:
        """Interaction cutoff in nm (reference config.py:130-139)."""
        p = self.parameters
        if self.potential_type == "wca":
            return p.get("sigma", 5.0) * 2.0 ** (1.0 / 6.0)
        if self.potential_type == "lennard_jones":
            return p.get("sigma", 5.0) * 2.5
        if self.potential_type == "harmonic_repulsion":
            return p.get("interaction_distance", 5.0)
        return 10.0


@dataclass
class DiffusionField:
    type: str = "uniform"
    default_D_factor: float = 1.0
    regions: List[Dict[str, Any]] = field(default_factory=list)
    # GPU-path addition: sinusoidal field parameters
    amplitude: float = 0.0
    wavelength: Tuple[float, float] = (0.0, 0.0)

    def get_D_factor(self, x: float, y: float) -> float:
        if self.type == "uniform":
            return self.default_D_factor
        for region in self.regions:
            shape = region.get("shape")
            if shape == "circle":
                cx, cy = region["center"]
                if (x - cx) ** 2 + (y - cy) ** 2 <= region["radius"] ** 2:
                    return region.get("D_factor", self.default_D_factor)
            elif shape == "rectangle":
                x0, y0 = region["origin"]
                w, h = region["size"]
                if x0 <= x <= x0 + w and y0 <= y <= y0 + h:
                    return region.get("D_factor", self.default_D_factor)
        return self.default_D_factor


@dataclass
class CatalyticReaction:
    name: str
    enzyme: str
    substrate: str
    product: str
    k_cat: float = 1.0
    contact_radius: float = 5.0
    feedback_rule: Optional[str] = None


@dataclass
class CompartmentExchange:
    name: str
    species: str
    k_on: float = 0.01
    k_off: float = 0.01
    cytosol_pool: int = 100
    max_recruits_per_step: int = 10
    feedback_rule: Optional[str] = None
    secondary_feedback_rule: Optional[str] = None
    detachment_feedback_rule: Optional[str] = None


@dataclass
class UnimolecularConversion:
    name: str
    from_species: str
    to_species: str
    rate: float = 0.1
    feedback_rule: Optional[str] = None


@dataclass
class FeedbackRule:
    name: str
    sensor_species: str
    function_type: str = "hill_inhibition"
    K_half: float = 100.0
    n_hill: float = 3.0
    grid_resolution: float = 0.2
    basal_fraction: float = 0.0


@dataclass
class MembraneGeometry:
    size: Tuple[float, float] = (2.0, 2.0)
    periodic: Tuple[bool, bool] = (True, True)
    geometry_type: str = "planar"

    @property
    def area(self) -> float:
        return self.size[0] * self.size[1]


@dataclass
class PhotoState:
    name: str
    is_fluorescent: bool = False
    is_absorbing: bool = True
    is_terminal: bool = False
    quantum_yield: float = 1.0
    emission_wavelength: float = 0.0


@dataclass
class FluorophoreType:
    name: str
    states: List[PhotoState] = field(default_factory=list)
    thermal_rates: Dict[str, Dict[str, float]] = field(default_factory=dict)
    photo_yields: Dict[str, Dict[str, float]] = field(default_factory=dict)
    absorption_cross_section: float = 0.0
    excitation_wavelength: float = 0.0
    extinction_coeff: float = 0.0
    tau_rad: float = 4.0e-9
    tau_nr_base: float = 1.0e-8
    environment_sensitivity: str = "none"

    def get_state_index(self, state_name: str) -> int:
        for i, s in enumerate(self.states):
            if s.name == state_name:
                return i
        raise ConfigError(f"State '{state_name}' not found in fluorophore '{self.name}'")

    @property
    def n_states(self) -> int:
        return len(self.states)

    @property
    def fluorescent_states(self) -> List[int]:
        return [i for i, s in enumerate(self.states) if s.is_fluorescent]

    @property
    def terminal_states(self) -> List[int]:
        return [i for i, s in enumerate(self.states) if s.is_terminal]


@dataclass
class FRETpair:
    donor_type: str
    acceptor_type: str
    R0: float
    kappa2: float = 2.0 / 3.0
    intra_complex_distance: float = 5.0
    fret_mode: str = "bond"
    intramolecular: bool = False
    proximity_radius: float = 10.0
    proximity_E_high: float = 0.5
    proximity_E_low: float = 0.02
    proximity_target_species: str = ""

    def efficiency(self, r_nm: float) -> float:
        if r_nm <= 0:
            return 1.0
        return 1.0 / (1.0 + (r_nm / self.R0) ** 6)


@dataclass
class DyeAttachment:
    species: str
    fluorophore: str
    labeling_efficiency: float = 1.0
    linker_length: float = 1.0
    position_on_species: str = "N-term"


@dataclass
class ObjectiveLens:
    NA: float = 1.4
    magnification: float = 100.0
    immersion: str = "oil"
    n_immersion: float = 1.515

    def airy_radius(self, wavelength_nm: float) -> float:
        return 0.61 * wavelength_nm / self.NA


@dataclass
class PSFmodel:
    type: str = "gibson_lanni"
    emission_wavelength: float = 520.0
    sigma_xy: float = 0.0
    sigma_z: float = 0.0

    def get_sigma_xy(self, NA: float) -> float:
        if self.sigma_xy > 0:
            return self.sigma_xy
        return 0.21 * self.emission_wavelength / NA


@dataclass
class Detector:
    type: str = "SPAD"
    dark_count_rate: float = 100.0
    detection_efficiency: float = 0.2
    timing_jitter: float = 50e-12
    dead_time: float = 100e-9
    afterpulsing_prob: float = 0.005


@dataclass
class LaserSource:
    wavelength: float = 488.0
    repetition_rate: float = 40e6
    pulse_width: float = 100e-12
    power: float = 1e-6
    beam_profile: str = "gaussian"
    beam_waist: float = 250.0

    @property
    def period(self) -> float:
        return 1.0 / self.repetition_rate


@dataclass
class ScanParameters:
    mode: str = "confocal"
    pixel_size: float = 50.0
    pixels_x: int = 256
    pixels_y: int = 256
    dwell_time: float = 10e-6
    n_frames: int = 1
    frame_interval: float = 1.0
    pinhole_diameter: float = 1.0

    @property
    def fov_x(self) -> float:
        return self.pixels_x * self.pixel_size / 1000.0

    @property
    def fov_y(self) -> float:
        return self.pixels_y * self.pixel_size / 1000.0

    @property
    def frame_time(self) -> float:
        return self.pixels_x * self.pixels_y * self.dwell_time


@dataclass
class TCSPCsettings:
    n_bins: int = 256
    time_range: float = 25e-9

    @property
    def bin_width(self) -> float:
        return self.time_range / self.n_bins


@dataclass
class SimulationConfig:
    name: str = "Unnamed simulation"
    version: str = "1.1.0"
    membrane: MembraneGeometry = field(default_factory=lambda: MembraneGeometry(size=(2.0, 2.0)))
    diffusion_field: DiffusionField = field(default_factory=DiffusionField)
    species: List[Species] = field(default_factory=list)
    reactions: List[Reaction] = field(default_factory=list)
    potentials: List[PairPotential] = field(default_factory=list)
    initial_populations: Dict[str, int] = field(default_factory=dict)
    catalytic_reactions: List[CatalyticReaction] = field(default_factory=list)
    compartment_exchanges: List[CompartmentExchange] = field(default_factory=list)
    unimolecular_conversions: List[UnimolecularConversion] = field(default_factory=list)
    feedback_rules: List[FeedbackRule] = field(default_factory=list)
    fluorophores: List[FluorophoreType] = field(default_factory=list)
    fret_pairs: List[FRETpair] = field(default_factory=list)
    dye_attachments: List[DyeAttachment] = field(default_factory=list)
    objective: ObjectiveLens = field(default_factory=ObjectiveLens)
    laser: LaserSource = field(default_factory=LaserSource)
    detector: Detector = field(default_factory=Detector)
    scan: ScanParameters = field(default_factory=ScanParameters)
    tcspc: TCSPCsettings = field(default_factory=TCSPCsettings)
    psf: PSFmodel = field(default_factory=PSFmodel)
    simulation_time: float = 60.0
    bd_timestep: float = 1e-5
    random_seed: int = 42
    pss_mode: bool = True
    n_threads: int = 8
    ram_limit_gb: float = 12.0
    displacement_cap_fraction: float = 0.5
    region_smoothing_width: float = 0.05
    _species_map: Dict[str, int] = field(default_factory=dict, repr=False)
    _fluoro_map: Dict[str, int] = field(default_factory=dict, repr=False)

    def species_by_name(self, name: str) -> Species:
        idx = self._species_map.get(name)
        if idx is None:
            raise ConfigError(f"Species '{name}' not found")
        return self.species[idx]

    def fluorophore_by_name(self, name: str) -> FluorophoreType:
        idx = self._fluoro_map.get(name)
        if idx is None:
            raise ConfigError(f"Fluorophore '{name}' not found")
        return self.fluorophores[idx]

    @property
    def total_particles(self) -> int:
        return sum(self.initial_populations.values())

    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def max_potential_cutoff(self) -> float:
        if not self.potentials:
            return 0.0
        return max(p.get_cutoff() for p in self.potentials)

    def validate(self) -> List[str]:
        """Consistency checks (reference config.py:579-766, the same 18
        rules); returns warnings, raises ConfigError listing every error."""
        warnings: List[str] = []
        errors: List[str] = []
        self._species_map = {s.name: i for i, s in enumerate(self.species)}
        self._fluoro_map = {f.name: i for i, f in enumerate(self.fluorophores)}
        sp = self._species_map
        fl = self._fluoro_map
        fb_names = {fb.name for fb in self.feedback_rules}

        names = [s.name for s in self.species]
        if len(names) != len(set(names)):
            errors.append("Duplicate species names detected")
        for rxn in self.reactions:
            for r in rxn.reactants:
                if r not in sp:
                    errors.append(f"Reaction '{rxn.name}': reactant '{r}' not in species list")
            for p in rxn.products:
                if p not in sp:
                    errors.append(f"Reaction '{rxn.name}': product '{p}' not in species list")
        for pot in self.potentials:
            if pot.species_a not in sp:
                errors.append(f"Potential: species_a '{pot.species_a}' not in species list")
            if pot.species_b not in sp:
                errors.append(f"Potential: species_b '{pot.species_b}' not in species list")
        for name in self.initial_populations:
            if name not in sp:
                errors.append(f"initial_populations: species '{name}' not in species list")
        for att in self.dye_attachments:
            if att.species not in sp:
                errors.append(f"DyeAttachment: species '{att.species}' not in species list")
            if att.fluorophore not in fl:
                errors.append(f"DyeAttachment: fluorophore '{att.fluorophore}' not in fluorophores")
        for fp in self.fret_pairs:
            if fp.donor_type not in fl:
                errors.append(f"FRETpair: donor '{fp.donor_type}' not in fluorophores")
            if fp.acceptor_type not in fl:
                errors.append(f"FRETpair: acceptor '{fp.acceptor_type}' not in fluorophores")
        for fluo in self.fluorophores:
            state_names = {s.name for s in fluo.states}
            for src, trans in fluo.thermal_rates.items():
                if src not in state_names:
                    errors.append(f"Fluorophore '{fluo.name}': thermal_rate from state '{src}' not in states")
                for dst in trans:
                    if dst not in state_names:
                        errors.append(f"Fluorophore '{fluo.name}': thermal_rate to state '{dst}' not in states")
        if self.tcspc.time_range > self.laser.period * 1.01:
            errors.append(f"TCSPC time_range ({self.tcspc.time_range:.2e}s) > laser period ({self.laser.period:.2e}s)")
        fov_x, fov_y = self.scan.fov_x, self.scan.fov_y
        mem_x, mem_y = self.membrane.size
        if fov_x > mem_x * 1.01 or fov_y > mem_y * 1.01:
            warnings.append(f"Scan FOV ({fov_x:.1f}x{fov_y:.1f} um) exceeds membrane ({mem_x:.1f}x{mem_y:.1f} um)")
        for s in self.species:
            if s.diffusion_coefficient < 0:
                errors.append(f"Species '{s.name}': negative D")
            if s.radius <= 0:
                errors.append(f"Species '{s.name}': non-positive radius")
        if self.bd_timestep <= 0:
            errors.append("bd_timestep must be positive")
        if self.simulation_time <= 0:
            errors.append("simulation_time must be positive")
        # RAM estimate of the reference's host layout (config.py:676-689)
        N = self.total_particles
        Px, Py, Tb = self.scan.pixels_x, self.scan.pixels_y, self.tcspc.n_bins
        ram_gb = (N * 50 + N * 2 * 8 * 2 + Px * Py * Tb * 4 + Px * Py * 4 + N * 50) / 1e9
        if ram_gb > self.ram_limit_gb:
            errors.append(f"Estimated RAM ({ram_gb:.1f} GB) exceeds limit ({self.ram_limit_gb} GB)")
        elif ram_gb > self.ram_limit_gb * 0.8:
            warnings.append(f"Estimated RAM ({ram_gb:.1f} GB) is close to limit ({self.ram_limit_gb} GB)")
        for cat in self.catalytic_reactions:
            for role, nm in (("enzyme", cat.enzyme), ("substrate", cat.substrate), ("product", cat.product)):
                if nm not in sp:
                    errors.append(f"CatalyticReaction '{cat.name}': {role} '{nm}' not in species")
            if cat.feedback_rule and cat.feedback_rule not in fb_names:
                errors.append(f"CatalyticReaction '{cat.name}': feedback_rule '{cat.feedback_rule}' not found")
        for ex in self.compartment_exchanges:
            if ex.species not in sp:
                errors.append(f"CompartmentExchange '{ex.name}': species '{ex.species}' not in species")
            for key in ("feedback_rule", "secondary_feedback_rule", "detachment_feedback_rule"):
                v = getattr(ex, key)
                if v and v not in fb_names:
                    errors.append(f"CompartmentExchange '{ex.name}': {key} '{v}' not found")
        for cv in self.unimolecular_conversions:
            if cv.from_species not in sp:
                errors.append(f"UnimolecularConversion '{cv.name}': from '{cv.from_species}' not in species")
            if cv.to_species not in sp:
                errors.append(f"UnimolecularConversion '{cv.name}': to '{cv.to_species}' not in species")
            if cv.feedback_rule and cv.feedback_rule not in fb_names:
                errors.append(f"UnimolecularConversion '{cv.name}': feedback_rule '{cv.feedback_rule}' not found")
        for fb in self.feedback_rules:
            if fb.sensor_species not in sp:
                errors.append(f"FeedbackRule '{fb.name}': sensor_species '{fb.sensor_species}' not in species")
            if fb.function_type not in ("hill_inhibition", "hill_activation"):
                errors.append(f"FeedbackRule '{fb.name}': function_type must be 'hill_inhibition' or 'hill_activation'")
            if not (0.0 <= fb.basal_fraction <= 1.0):
                errors.append(f"FeedbackRule '{fb.name}': basal_fraction must be in [0, 1], got {fb.basal_fraction}")
        for fp in self.fret_pairs:
            if fp.fret_mode == "proximity":
                if not fp.proximity_target_species:
                    errors.append(f"FRETpair {fp.donor_type}/{fp.acceptor_type}: proximity mode requires proximity_target_species")
                elif fp.proximity_target_species not in sp:
                    errors.append(f"FRETpair: proximity_target_species '{fp.proximity_target_species}' not in species")
        D_max = max((s.diffusion_coefficient for s in self.species), default=0.0)
        sigma_min = min((p.parameters.get("sigma", 5.0) for p in self.potentials
                         if p.potential_type in ("wca", "lennard_jones")), default=float("inf"))
        if sigma_min < float("inf") and D_max > 0:
            dt_stable = (sigma_min / 1000.0) ** 2 / (4.0 * D_max)
            if self.bd_timestep > dt_stable:
                _warnings.warn(
                    f"BD timestep ({self.bd_timestep:.1e}s) exceeds WCA stability limit "
                    f"sigma^2/(4D) = {dt_stable:.1e}s", stacklevel=2)
        if not (0.001 <= self.region_smoothing_width <= 0.5):
            errors.append(f"region_smoothing_width ({self.region_smoothing_width}) must be in [0.001, 0.5] um")
        if errors:
            raise ConfigError(f"Configuration has {len(errors)} error(s):\n"
                              + "\n".join(f"  - {e}" for e in errors))
        return warnings


# ----------------------------------------------------------------------
# YAML loading (reference config.py:778-1094)
# ----------------------------------------------------------------------

def _photo_state(d: Dict[str, Any]) -> PhotoState:
    return PhotoState(
        name=d["name"], is_fluorescent=d.get("is_fluorescent", False),
        is_absorbing=d.get("is_absorbing", True), is_terminal=d.get("is_terminal", False),
        quantum_yield=d.get("quantum_yield", 1.0),
        emission_wavelength=d.get("emission_wavelength", 0.0))


def _fluorophore(d: Dict[str, Any]) -> FluorophoreType:
    return FluorophoreType(
        name=d["name"], states=[_photo_state(s) for s in d.get("states", [])],
        thermal_rates=d.get("thermal_rates", {}), photo_yields=d.get("photo_yields", {}),
        absorption_cross_section=d.get("absorption_cross_section", 0.0),
        excitation_wavelength=d.get("excitation_wavelength", 0.0),
        extinction_coeff=d.get("extinction_coeff", 0.0),
        tau_rad=d.get("tau_rad", 4.0e-9), tau_nr_base=d.get("tau_nr_base", 1.0e-8),
        environment_sensitivity=d.get("environment_sensitivity", "none"))


def _reaction(d: Dict[str, Any]) -> Reaction:
    """Reaction with the three kon modes (direct, Keq, f_bound) of
    reference config.py:833-879."""
    kon = d.get("kon", 0.0)
    koff = d.get("koff", 0.0)
    Keq = d.get("Keq", None)
    f_bound = d.get("f_bound", None)
    if f_bound is not None and Keq is None and kon == 0.0:
        if not (0.0 < f_bound < 1.0):
            raise ConfigError(f"Reaction '{d['name']}': f_bound must be in (0,1), got {f_bound}")
        lig = d.get("ligand_density", None)
        if lig is None:
            raise ConfigError(f"Reaction '{d['name']}': f_bound mode requires 'ligand_density'")
        Keq = 1.0 / (lig * (1.0 - f_bound) / f_bound)
    if Keq is not None and kon == 0.0:
        if koff <= 0.0:
            raise ConfigError(f"Reaction '{d['name']}': Keq mode requires koff > 0")
        kon = Keq * koff
    return Reaction(
        name=d["name"], reactants=d["reactants"], products=d["products"], kon=kon,
        koff=koff, contact_radius=d.get("contact_radius", 5.0), Keq=Keq, f_bound=f_bound,
        required_state=d.get("required_state", None), product_state=d.get("product_state", None))


def config_from_dict(raw: Dict[str, Any]) -> SimulationConfig:
    """Build and validate a SimulationConfig from a parsed YAML mapping."""
    if not isinstance(raw, dict):
        raise ConfigError(f"YAML root must be a dict, got {type(raw).__name__}")
    m = raw.get("membrane", {})
    membrane = MembraneGeometry(size=tuple(m.get("size", [2.0, 2.0])),
                                periodic=tuple(m.get("periodic", [True, True])),
                                geometry_type=m.get("geometry_type", "planar"))
    df = raw.get("diffusion_field", {})
    diffusion_field = DiffusionField(
        type=df.get("type", "uniform"), default_D_factor=df.get("default_D_factor", 1.0),
        regions=df.get("regions", []), amplitude=df.get("amplitude", 0.0),
        wavelength=tuple(df.get("wavelength", [0.0, 0.0])))
    species = [Species(name=s["name"], diffusion_coefficient=s["diffusion_coefficient"],
                       radius=s["radius"], color=s.get("color", "#FFFFFF"),
                       states=s.get("states", ["default"]),
                       state_transitions=s.get("state_transitions", {}),
                       D_per_state=s.get("D_per_state", None))
               for s in raw.get("species", [])]
    reactions = [_reaction(r) for r in raw.get("reactions", [])]
    potentials = [PairPotential(species_a=p["species_a"], species_b=p["species_b"],
                                potential_type=p["potential_type"],
                                parameters=p.get("parameters", {}))
                  for p in raw.get("potentials", [])]
    catalytic = [CatalyticReaction(name=c["name"], enzyme=c["enzyme"], substrate=c["substrate"],
                                   product=c["product"], k_cat=c.get("k_cat", 1.0),
                                   contact_radius=c.get("contact_radius", 5.0),
                                   feedback_rule=c.get("feedback_rule", None))
                 for c in raw.get("catalytic_reactions", [])]
    exchanges = [CompartmentExchange(
        name=e["name"], species=e["species"], k_on=e.get("k_on", 0.01), k_off=e.get("k_off", 0.01),
        cytosol_pool=e.get("cytosol_pool", 100),
        max_recruits_per_step=e.get("max_recruits_per_step", 10),
        feedback_rule=e.get("feedback_rule", None),
        secondary_feedback_rule=e.get("secondary_feedback_rule", None),
        detachment_feedback_rule=e.get("detachment_feedback_rule", None))
        for e in raw.get("compartment_exchanges", [])]
    conversions = [UnimolecularConversion(name=u["name"], from_species=u["from_species"],
                                          to_species=u["to_species"], rate=u.get("rate", 0.1),
                                          feedback_rule=u.get("feedback_rule", None))
                   for u in raw.get("unimolecular_conversions", [])]
    feedback = [FeedbackRule(name=f["name"], sensor_species=f["sensor_species"],
                             function_type=f.get("function_type", "hill_inhibition"),
                             K_half=f.get("K_half", 100.0), n_hill=f.get("n_hill", 3.0),
                             grid_resolution=f.get("grid_resolution", 0.2),
                             basal_fraction=f.get("basal_fraction", 0.0))
                for f in raw.get("feedback_rules", [])]
    fluorophores = [_fluorophore(f) for f in raw.get("fluorophores", [])]
    fret_pairs = [FRETpair(
        donor_type=f["donor_type"], acceptor_type=f["acceptor_type"], R0=f["R0"],
        kappa2=f.get("kappa2", 2.0 / 3.0), intra_complex_distance=f.get("intra_complex_distance", 5.0),
        fret_mode=f.get("fret_mode", "bond"), intramolecular=f.get("intramolecular", False),
        proximity_radius=f.get("proximity_radius", 10.0),
        proximity_E_high=f.get("proximity_E_high", 0.5), proximity_E_low=f.get("proximity_E_low", 0.02),
        proximity_target_species=f.get("proximity_target_species", ""))
        for f in raw.get("fret_pairs", [])]
    dyes = [DyeAttachment(species=a["species"], fluorophore=a["fluorophore"],
                          labeling_efficiency=a.get("labeling_efficiency", 1.0),
                          linker_length=a.get("linker_length", 1.0),
                          position_on_species=a.get("position_on_species", "N-term"))
            for a in raw.get("dye_attachments", [])]
    o = raw.get("objective", {})
    objective = ObjectiveLens(NA=o.get("NA", 1.4), magnification=o.get("magnification", 100.0),
                              immersion=o.get("immersion", "oil"), n_immersion=o.get("n_immersion", 1.515))
    la = raw.get("laser", {})
    laser = LaserSource(wavelength=la.get("wavelength", 488.0), repetition_rate=la.get("repetition_rate", 40e6),
                        pulse_width=la.get("pulse_width", 100e-12), power=la.get("power", 1e-6),
                        beam_profile=la.get("beam_profile", "gaussian"), beam_waist=la.get("beam_waist", 250.0))
    de = raw.get("detector", {})
    detector = Detector(type=de.get("type", "SPAD"), dark_count_rate=de.get("dark_count_rate", 100.0),
                        detection_efficiency=de.get("detection_efficiency", 0.2),
                        timing_jitter=de.get("timing_jitter", 50e-12), dead_time=de.get("dead_time", 100e-9),
                        afterpulsing_prob=de.get("afterpulsing_prob", 0.005))
    sc = raw.get("scan", {})
    scan = ScanParameters(mode=sc.get("mode", "confocal"), pixel_size=sc.get("pixel_size", 50.0),
                          pixels_x=sc.get("pixels_x", 256), pixels_y=sc.get("pixels_y", 256),
                          dwell_time=sc.get("dwell_time", 10e-6), n_frames=sc.get("n_frames", 1),
                          frame_interval=sc.get("frame_interval", 1.0),
                          pinhole_diameter=sc.get("pinhole_diameter", 1.0))
    tc = raw.get("tcspc", {})
    tcspc = TCSPCsettings(n_bins=tc.get("n_bins", 256), time_range=tc.get("time_range", 25e-9))
    ps = raw.get("psf", {})
    psf = PSFmodel(type=ps.get("type", "gibson_lanni"), emission_wavelength=ps.get("emission_wavelength", 520.0),
                   sigma_xy=ps.get("sigma_xy", 0.0), sigma_z=ps.get("sigma_z", 0.0))
    cfg = SimulationConfig(
        name=raw.get("name", "Unnamed simulation"), version=raw.get("version", "1.1.0"),
        membrane=membrane, diffusion_field=diffusion_field, species=species, reactions=reactions,
        potentials=potentials, initial_populations=raw.get("initial_populations", {}),
        catalytic_reactions=catalytic, compartment_exchanges=exchanges,
        unimolecular_conversions=conversions, feedback_rules=feedback, fluorophores=fluorophores,
        fret_pairs=fret_pairs, dye_attachments=dyes, objective=objective, laser=laser,
        detector=detector, scan=scan, tcspc=tcspc, psf=psf,
        simulation_time=raw.get("simulation_time", 60.0), bd_timestep=raw.get("bd_timestep", 1e-5),
        random_seed=raw.get("random_seed", 42), pss_mode=raw.get("pss_mode", True),
        n_threads=raw.get("n_threads", 8), ram_limit_gb=raw.get("ram_limit_gb", 12.0))
    cfg.validate()
    return cfg


def load_config(path: Union[str, Path]) -> SimulationConfig:
    """Load and validate a YAML configuration (reference config.py:947)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"Config file not found: {path}")
    with open(path, "r", encoding="utf-8") as f:
        raw = yaml.safe_load(f)
    return config_from_dict(raw)
