"""Particle state: structure of arrays (reference particles.py:36-368).

Initialisation draws in the reference's order (particles.py:94-243), so the
same seed gives the same initial state: for each species in config order
uniform x then y, then each cytosolic pool (x then y, dormant), then one
Bernoulli labelling draw per particle of each dye-carrying species.

Once a DynamicsEngine has taken the state to the GPU the device copy is
authoritative: the array attributes are host mirrors refreshed on access
after device work (``pull()``); after editing them on the host call
``push()`` to write them back.
"""

from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

from .config import ConfigError, SimulationConfig

__all__ = ["ParticleState"]

_ARRAYS = ("positions", "species_id", "state_id", "bond_partner", "has_dye", "dye_type_id",
           "ctmc_state", "is_bleached", "is_alive")


class ParticleState:
    def __init__(self, config: SimulationConfig, rng: Optional[np.random.Generator] = None):
        if rng is None:
            rng = np.random.default_rng(config.random_seed)
        self._rng = rng
        self._config = config
        self._device = None      # membrane handle owner (DynamicsEngine)
        self._stale = False      # host mirror older than the device copy

        self.species_names: List[str] = [s.name for s in config.species]
        self._species_name_to_id: Dict[str, int] = {n: i for i, n in enumerate(self.species_names)}
        counts: List[Tuple[int, int]] = []
        N = 0
        for sp in config.species:
            c = config.initial_populations.get(sp.name, 0)
            counts.append((self._species_name_to_id[sp.name], c))
            N += c
        self._dormant_pools: List[Tuple[int, int, int]] = []
        for ex in getattr(config, "compartment_exchanges", []):
            if ex.cytosol_pool > 0:
                sid = self._species_name_to_id.get(ex.species, -1)
                if sid >= 0:
                    self._dormant_pools.append((sid, N, ex.cytosol_pool))
                    N += ex.cytosol_pool
        if N == 0:
            raise ConfigError("Total particle count is 0. Check initial_populations in config.")
        self.N: int = N
        a = {}
        a["positions"] = np.zeros((N, 2), dtype=np.float64)
        a["species_id"] = np.zeros(N, dtype=np.int32)
        a["state_id"] = np.zeros(N, dtype=np.int32)
        a["bond_partner"] = np.full(N, -1, dtype=np.int32)
        a["has_dye"] = np.zeros(N, dtype=np.bool_)
        a["dye_type_id"] = np.full(N, -1, dtype=np.int32)
        a["ctmc_state"] = np.zeros(N, dtype=np.int8)
        a["is_bleached"] = np.zeros(N, dtype=np.bool_)
        a["is_alive"] = np.ones(N, dtype=np.bool_)
        self._a = a

        self.species_ranges: Dict[str, Tuple[int, int]] = {}
        Lx, Ly = config.membrane.size
        off = 0
        for sid, c in counts:
            a["species_id"][off:off + c] = sid
            a["positions"][off:off + c, 0] = rng.uniform(0.0, Lx, size=c)
            a["positions"][off:off + c, 1] = rng.uniform(0.0, Ly, size=c)
            self.species_ranges[self.species_names[sid]] = (off, off + c)
            off += c
        for sid, start, c in self._dormant_pools:
            a["species_id"][start:start + c] = sid
            a["positions"][start:start + c, 0] = rng.uniform(0.0, Lx, size=c)
            a["positions"][start:start + c, 1] = rng.uniform(0.0, Ly, size=c)
            a["is_alive"][start:start + c] = False
        self._assign_dyes(config)
        self._D_array = np.array([s.diffusion_coefficient for s in config.species], dtype=np.float64)

    def _assign_dyes(self, config: SimulationConfig) -> None:
        """Bernoulli labelling of every particle of each attached species,
        dormant ones included (reference particles.py:191-243)."""
        fl = {f.name: i for i, f in enumerate(config.fluorophores)}
        a = self._a
        for att in config.dye_attachments:
            sid = self._species_name_to_id.get(att.species, -1)
            if sid < 0:
                continue
            fid = fl.get(att.fluorophore)
            if fid is None:
                raise ConfigError(f"DyeAttachment references fluorophore '{att.fluorophore}' "
                                  f"not found in config.fluorophores")
            idx = np.where(a["species_id"] == sid)[0]
            if len(idx) == 0:
                continue
            lab = self._rng.random(len(idx)) < att.labeling_efficiency
            sel = idx[lab]
            a["has_dye"][sel] = True
            a["dye_type_id"][sel] = fid
            a["ctmc_state"][sel] = 0

    # ---- device mirror ------------------------------------------------
    def _attach(self, device) -> None:
        self._device = device
        self._stale = False

    def mark_stale(self) -> None:
        if self._device is not None:
            self._stale = True

    def pull(self) -> None:
        """Refresh the host arrays from the device copy."""
        if self._device is not None and self._stale:
            self._stale = False
            self._device._download_into(self._a)

    def push(self) -> None:
        """Write host edits back to the device copy."""
        if self._device is not None:
            self._stale = False
            self._device._upload_from(self._a)

    def __getattr__(self, name):
        if name in _ARRAYS:
            a = self.__dict__.get("_a")
            if a is None:
                raise AttributeError(name)
            if self.__dict__.get("_stale"):
                self.pull()
            return a[name]
        raise AttributeError(name)

    # ---- queries (reference particles.py:247-339) ----------------------
    def get_positions_of(self, species_name: str) -> np.ndarray:
        sid = self._species_name_to_id.get(species_name)
        if sid is None:
            raise ConfigError(f"Species '{species_name}' not found")
        return self.positions[self.species_id == sid]

    def get_alive_mask(self) -> np.ndarray:
        return self.is_alive

    def get_labeled_mask(self, fluorophore_name: Optional[str] = None) -> np.ndarray:
        if fluorophore_name is None:
            return self.has_dye.copy()
        fid = None
        for i, f in enumerate(self._config.fluorophores):
            if f.name == fluorophore_name:
                fid = i
                break
        if fid is None:
            raise ConfigError(f"Fluorophore '{fluorophore_name}' not found")
        return self.has_dye & (self.dye_type_id == fid)

    def get_free_mask(self) -> np.ndarray:
        return self.bond_partner == -1

    def get_bound_mask(self) -> np.ndarray:
        return self.bond_partner >= 0

    def count_species(self, species_name: str) -> int:
        sid = self._species_name_to_id.get(species_name)
        if sid is None:
            raise ConfigError(f"Species '{species_name}' not found")
        return int(np.sum((self.species_id == sid) & self.is_alive))

    def get_D(self, idx: int) -> float:
        return float(self._D_array[self.species_id[idx]])

    def get_D_array(self) -> np.ndarray:
        return self._D_array[self.species_id]

    def get_numba_arrays(self) -> dict:
        """Raw arrays under the reference's key names (host mirrors)."""
        d = {k: getattr(self, k) for k in _ARRAYS if k != "state_id"}
        d["state_id"] = self.state_id
        d["D_per_species"] = self._D_array
        return d

    @property
    def ram_bytes(self) -> int:
        return int(sum(self._a[k].nbytes for k in _ARRAYS))

    def __repr__(self) -> str:
        lines = [f"ParticleState(N={self.N})"]
        for name, (s, e) in self.species_ranges.items():
            lines.append(f"  {name}: {e - s} particles, {int(np.sum(self.has_dye[s:e]))} labeled")
        for sid, start, c in self._dormant_pools:
            lines.append(f"  {self.species_names[sid]} (dormant cytosol): {c} particles")
        lines.append(f"  RAM: {self.ram_bytes / 1e6:.2f} MB")
        return "\n".join(lines)
