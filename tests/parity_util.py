"""Helpers shared by the parity tests: run the GPU engines and the oracle
side by side from the same configuration and seed, with the reference's
numpy draws on both sides."""

from __future__ import annotations

import numpy as np

import reference_model as rm
from ed_mrdt.particles import ParticleState


def host_arrays(state):
    return dict(positions=state.positions.copy(), species_id=state.species_id.copy(),
                is_alive=state.is_alive.astype(np.uint8), bond_partner=state.bond_partner.copy(),
                ctmc_state=state.ctmc_state.copy(), is_bleached=state.is_bleached.astype(np.uint8),
                has_dye=state.has_dye.astype(np.uint8), dye_type_id=state.dye_type_id.copy())


def oracle_arrays(st: rm.OracleState):
    return dict(positions=st.positions.copy(), species_id=st.species_id.copy(),
                is_alive=st.is_alive.copy(), bond_partner=st.bond_partner.copy(),
                ctmc_state=st.ctmc_state.copy(), is_bleached=st.is_bleached.copy(),
                has_dye=st.has_dye.copy(), dye_type_id=st.dye_type_id.copy())


def compare_states(g, o, pos_rtol=1e-9, label=""):
    """Integer/flag arrays must match exactly; positions within pos_rtol
    relative to the box (FP64 parity bound)."""
    for k in ("species_id", "is_alive", "bond_partner", "ctmc_state", "is_bleached"):
        bad = np.flatnonzero(g[k] != o[k])
        assert bad.size == 0, f"{label}{k} differs at {bad[:10]} (gpu {g[k][bad[:5]]}, oracle {o[k][bad[:5]]})"
    d = np.abs(g["positions"] - o["positions"])
    scale = np.maximum(np.abs(o["positions"]), 1.0)
    worst = float(np.max(d / scale)) if d.size else 0.0
    assert worst <= pos_rtol, f"{label}positions differ: max rel {worst:.3e} > {pos_rtol}"
    return worst


def autocatalytic_config(feedback: bool, seed: int = 5, k_cat: float = 8.0):
    """The plain PIP3/PTEN membrane with an autocatalytic reaction added
    after PI3K's: PIP3 + PIP2 -> 2 PIP3 within 150 nm (per-pair probability
    k_cat dt / (pi 0.15^2) = 0.11 at k_cat = 8), optionally under the
    pip3_activation feedback (its sensor is the enzyme, so density updates
    are deferred)."""
    from ed_mrdt import workloads
    raw = workloads.plain_raw()
    rx = {"name": "PIP3_autocatalysis", "enzyme": "PIP3", "substrate": "PIP2",
          "product": "PIP3", "k_cat": float(k_cat), "contact_radius": 150.0}
    if feedback:
        rx["feedback_rule"] = "pip3_activation"
    raw["catalytic_reactions"].insert(1, rx)
    raw["random_seed"] = seed
    return workloads.config_from_dict(raw)
