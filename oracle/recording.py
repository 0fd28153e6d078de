"""ORACLE — test infrastructure only.

Caller-side recording of the reference's SimulationPipeline, restated with
numpy on the oracle's arrays (OracleState / OracleDynamics / OraclePhotons):

* ``OracleSubframe``  — the per-BD-step recording of SimulationPipeline.run
                        (pipeline.py:536-566 set-up, 567-642 one step,
                        700-717 SubFrameData)
* ``ground_truth``    — _extract_ground_truth (pipeline.py:357-455), called
                        after update_fret_rates (pipeline.py:649)
* ``census_row``      — the counts one device census row holds, from the
                        same masks

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module.

Pinning (tests/test_recording.py): the reference's own ground-truth tests
(tests/test_pipeline.py:307-380 of the reference: array shapes, E in [0, 1],
tau_DA >= 0, donor-only tau_DA = tau_D within 5 % of 4.1 ns, complex count
bounded by the receptor count) run on the oracle; the FRET values are tied
to closed forms (Förster E = 1 / (1 + (r/R0)^6) for a bound donor with the
bond-mode rate, E = E_high / E_low for the proximity biosensor rates,
E + tau_DA / tau_D = 1 for every donor); counts are tied to brute-force
loops (mutual bond pairs, per-particle tallies) on small states; the
per-step event deltas sum to the dynamics' cumulative counters.
"""

from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np


def _b(a) -> np.ndarray:
    return np.asarray(a).astype(bool)


def _donor_index(config, donor_name: str) -> Optional[int]:
    """First fluorophore of that name (pipeline.py:386-390)."""
    for fi, f in enumerate(config.fluorophores):
        if f.name == donor_name:
            return fi
    return None


def ground_truth(config, state, photons) -> Dict:
    """pipeline.py:357-455 on the oracle state; ``photons.k_fret`` must hold
    the rates of a fresh update_fret_rates()."""
    N = state.N
    fret_E = np.zeros(N)
    tau_DA = np.zeros(N)
    has_dye = _b(state.has_dye)
    bleached = _b(state.is_bleached)
    alive = _b(state.is_alive)
    bond = np.asarray(state.bond_partner)
    dye = np.asarray(state.dye_type_id)
    k = np.asarray(photons.k_fret) if getattr(photons, "active", True) else np.zeros(N)
    for fp in config.fret_pairs:
        d = _donor_index(config, fp.donor_type)
        if d is None:
            continue
        rm = photons.rms[d]
        tD = rm.lifetime(rm.es)
        if tD <= 0 or np.isinf(tD):
            continue
        sel = has_dye & ~bleached & (dye == d)
        on = sel & (k > 0)
        off = sel & ~(k > 0)
        kt = k[on] * tD
        fret_E[on] = kt / (1.0 + kt)
        tau_DA[on] = tD / (1.0 + kt)
        fret_E[off] = 0.0
        tau_DA[off] = tD
    n_complexes = int(np.sum(bond >= 0)) // 2
    n_bleached = int(np.sum(bleached))
    n_free = 0
    for fp in config.fret_pairs:
        for fi, f in enumerate(config.fluorophores):
            if f.name == fp.donor_type:
                n_free += int(np.sum(has_dye & (dye == fi) & ~bleached & (bond < 0)))
    sp = np.asarray(state.species_id)
    counts = {s.name: int(np.sum((sp == i) & alive)) for i, s in enumerate(config.species)}
    return dict(fret_efficiency_exact=fret_E, donor_lifetime_exact=tau_DA,
                n_complexes=n_complexes, n_bleached=n_bleached, n_free_donors=n_free,
                species_counts=counts)


def census_row(config, state, photons, counters: Dict[str, int], photons_emitted: int,
               step: int, fret_sum: bool) -> Dict:
    """The counts of one census (the fields of edm_census, restated from the
    reference's masks at pipeline.py:570, 582, 594-597, 622-625 per step and
    418-439 per frame) on the oracle state: what one device census row must
    hold.  ``counters`` are the dynamics' cumulative event counters."""
    alive = _b(state.is_alive)
    bleached = _b(state.is_bleached)
    has_dye = _b(state.has_dye)
    bond = np.asarray(state.bond_partner)
    dye = np.asarray(state.dye_type_id)
    sp = np.asarray(state.species_id)
    donor0 = has_dye & alive & ~bleached & (dye == 0)
    free = has_dye & ~bleached & (bond < 0)
    e_sum = 0.0
    if fret_sum and int(np.sum(donor0)):
        k = np.asarray(photons.k_fret)[donor0]
        kt = k * photons.tau[0]
        e_sum = float(np.sum(kt / (1.0 + kt)))
    return dict(step=int(step),
                n_bond_alive=int(np.sum(alive & (bond >= 0))),
                n_bond=int(np.sum(bond >= 0)),
                n_bleached=int(np.sum(bleached)),
                n_donor0=int(np.sum(donor0)),
                n_free_dye=[int(np.sum(free & (dye == t))) for t in range(8)],
                species_alive=[int(np.sum(alive & (sp == q))) for q in range(32)],
                events=[counters[k] for k in ("n_bindings", "n_dissociations", "n_catalytic",
                                              "n_recruited", "n_detached", "n_converted")],
                photons=int(photons_emitted), e_sum=e_sum)


class OracleSubframe:
    """Per-step recording (pipeline.py:536-566, 567-642).  Call ``record``
    after each dynamics step + photon step with that step's photon count."""

    def __init__(self, config, dynamics, photons):
        self.config = config
        self.dyn = dynamics
        self.ph = photons
        self.prev = dict(dynamics.counters())
        self.rows: List[Dict] = []

    def record(self, state, n_photons: int) -> Dict:
        cfg = self.config
        alive = _b(state.is_alive)
        bleached = _b(state.is_bleached)
        has_dye = _b(state.has_dye)
        bond = np.asarray(state.bond_partner)
        n_bound = int(np.sum(bond[alive] >= 0)) // 2
        cur = dict(self.dyn.counters())
        row = dict(n_complexes=n_bound,
                   n_bindings=cur["n_bindings"] - self.prev["n_bindings"],
                   n_dissociations=cur["n_dissociations"] - self.prev["n_dissociations"],
                   n_bleached=int(np.sum(bleached)),
                   total_photons=int(n_photons))
        mean_E = 0.0
        if len(cfg.fret_pairs) > 0:
            fp = cfg.fret_pairs[0]
            mode = getattr(fp, "fret_mode", "bond")
            donor_mask = has_dye & alive & ~bleached & (np.asarray(state.dye_type_id) == 0)
            n_total = int(np.sum(donor_mask))
            if mode == "proximity" and n_total > 0:
                k = np.asarray(self.ph.k_fret)[donor_mask]
                tau0 = self.ph.tau[0]
                mean_E = float(np.mean(k * tau0 / (1.0 + k * tau0)))
            elif mode != "proximity" and n_total > 0:
                E_const = 1.0 / (1.0 + (fp.intra_complex_distance / fp.R0) ** 6)
                n_free = n_total - n_bound
                mean_E = E_const * n_bound / (n_bound + n_free)
        row["mean_fret_E"] = mean_E
        sp = np.asarray(state.species_id)
        row["species_counts"] = {s.name: int(np.sum((sp == i) & alive))
                                 for i, s in enumerate(cfg.species)}
        for key in ("n_catalytic", "n_recruited", "n_detached", "n_converted"):
            row[key] = cur[key] - self.prev[key]
        self.prev = cur
        self.rows.append(row)
        return row

    def arrays(self) -> Dict[str, np.ndarray]:
        """SubFrameData fields (pipeline.py:700-717) as arrays."""
        out = {}
        for key in ("n_complexes", "n_bindings", "n_dissociations", "n_bleached", "total_photons",
                    "n_catalytic", "n_recruited", "n_detached", "n_converted"):
            out[key] = np.array([r[key] for r in self.rows], dtype=np.int32)
        out["mean_fret_E"] = np.array([r["mean_fret_E"] for r in self.rows], dtype=np.float64)
        out["species_counts"] = {s.name: np.array([r["species_counts"][s.name] for r in self.rows],
                                                  dtype=np.int32)
                                 for s in self.config.species}
        return out
