"""Photophysics: CTMC rate matrices, PSS precompute and the per-step photon
kernel on the GPU (reference photophysics.py).

The PSS tables are per-configuration constants computed once on the host
in FP64 (the 3x3 per-pulse transition matrix raised to n_pulses with
numpy.linalg.matrix_power and the geometric-series photon means,
photophysics.py:218-331); only they are FP64 on purpose.  Each BD step the
device updates FRET rates and runs the PSS kernel for every labelled,
alive, unbleached particle.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _hip
from ._hip import check, lib, ptr
from .config import ConfigError, FluorophoreType, FRETpair, SimulationConfig
from .particles import ParticleState

__all__ = ["RateMatrix", "PhotonBatch", "PhotonSimulator", "fret_efficiency", "fret_rate",
           "build_pss_matrices"]


class RateMatrix:
    """CTMC generator of one fluorophore (reference photophysics.py:59-151)."""

    def __init__(self, fluorophore: FluorophoreType):
        self.fluorophore = fluorophore
        self.n_states = fluorophore.n_states
        self.state_names = [s.name for s in fluorophore.states]
        idx = {n: i for i, n in enumerate(self.state_names)}
        n = self.n_states
        Q = np.zeros((n, n))
        for a, trans in fluorophore.thermal_rates.items():
            for b, rate in trans.items():
                Q[idx[a], idx[b]] = rate
        for i in range(n):
            Q[i, i] = 0.0
            Q[i, i] = -np.sum(Q[i, :])
        self.Q = Q
        st = fluorophore.states
        self.fluorescent_mask = np.array([s.is_fluorescent for s in st], dtype=bool)
        self.terminal_mask = np.array([s.is_terminal for s in st], dtype=bool)
        self.absorbing_mask = np.array([s.is_absorbing for s in st], dtype=bool)
        self.ground_state = next((i for i, s in enumerate(st) if s.is_absorbing and not s.is_fluorescent), 0)
        self.excited_state = next((i for i, s in enumerate(st) if s.is_fluorescent), -1)
        self.exit_rates = np.array([-Q[i, i] for i in range(n)])
        self.branch_probs = np.zeros((n, n))
        for i in range(n):
            if self.exit_rates[i] > 0:
                for j in range(n):
                    if j != i:
                        self.branch_probs[i, j] = Q[i, j] / self.exit_rates[i]
        self.t1_state = next((i for i, s in enumerate(st)
                              if not s.is_absorbing and not s.is_fluorescent and not s.is_terminal), -1)
        self.bleached_state = next((i for i, s in enumerate(st) if s.is_terminal), -1)

    def get_exit_rate(self, state: int) -> float:
        return self.exit_rates[state]

    def get_lifetime(self, state: int) -> float:
        k = self.exit_rates[state]
        return float("inf") if k <= 0 else 1.0 / k

    def copy_with_fret(self, k_fret: float):
        exit_rates = self.exit_rates.copy()
        branch = self.branch_probs.copy()
        fret_branch = np.zeros(self.n_states)
        if self.excited_state < 0 or k_fret <= 0:
            return exit_rates, branch, fret_branch
        s1 = self.excited_state
        new_total = exit_rates[s1] + k_fret
        exit_rates[s1] = new_total
        if new_total > 0:
            fret_branch[s1] = k_fret / new_total
            for j in range(self.n_states):
                if j != s1:
                    branch[s1, j] = self.Q[s1, j] / new_total
        return exit_rates, branch, fret_branch


def fret_efficiency(r_nm: float, R0_nm: float) -> float:
    """E(r) = 1 / (1 + (r/R0)^6) (reference photophysics.py:159-166)."""
    if r_nm <= 0.0:
        return 1.0
    q = r_nm / R0_nm
    return 1.0 / (1.0 + q * q * q * q * q * q)


def fret_rate(r_nm: float, R0_nm: float, tau_D: float) -> float:
    """k_FRET = (R0/r)^6 / tau_D (reference photophysics.py:169-178)."""
    if r_nm <= 0.0:
        return 1.0e15
    if tau_D <= 0.0:
        return 0.0
    q = R0_nm / r_nm
    return q * q * q * q * q * q / tau_D


class PhotonBatch:
    """Photons emitted during one BD step.  On the GPU path the arrays stay
    on the device until an attribute is read."""

    def __init__(self, n_photons: int = 0, source=None, arrays: Optional[Dict] = None):
        self.n_photons = int(n_photons)
        self._src = source
        self._arr = arrays

    @staticmethod
    def empty() -> "PhotonBatch":
        return PhotonBatch(0, arrays=dict(particle_idx=np.empty(0, np.int32),
                                          arrival_time=np.empty(0),
                                          channel=np.empty(0, np.int8),
                                          pulse_index=np.empty(0, np.int32),
                                          is_fret=np.empty(0, np.bool_)))

    def _materialize(self):
        if self._arr is None:
            n = self.n_photons
            pid = np.empty(n, np.int32)
            t = np.empty(n)
            ch = np.empty(n, np.int8)
            fr = np.empty(n, np.uint8)
            if n:
                check(lib().edm_photon_batch(self._src, ptr(pid), ptr(t), ptr(ch), ptr(fr), n),
                      "edm_photon_batch")
            self._arr = dict(particle_idx=pid, arrival_time=t, channel=ch,
                             pulse_index=np.zeros(n, np.int32), is_fret=fr.astype(bool))
        return self._arr

    particle_idx = property(lambda self: self._materialize()["particle_idx"])
    arrival_time = property(lambda self: self._materialize()["arrival_time"])
    channel = property(lambda self: self._materialize()["channel"])
    pulse_index = property(lambda self: self._materialize()["pulse_index"])
    is_fret = property(lambda self: self._materialize()["is_fret"])

    def __repr__(self) -> str:
        return f"PhotonBatch(n={self.n_photons})"


def build_pss_matrices(rm: RateMatrix, p_exc: float, T_laser: float, n_pulses: int,
                       k_fret: float = 0.0, QY_donor: float = 1.0, QY_acceptor: float = 1.0) -> Dict:
    """Per-step PSS tables of one fluorophore (reference photophysics.py:218-331):
    per-pulse 3x3 chain {S0, T1, bleached}, its n_pulses-th power, and the
    geometric-series photon means from S0 / T1."""
    s0, s1, t1, bl = rm.ground_state, rm.excited_state, rm.t1_state, rm.bleached_state
    kS1 = rm.exit_rates[s1] + k_fret if (k_fret > 0 and s1 >= 0) else rm.exit_rates[s1]
    if kS1 > 0 and s1 >= 0:
        p_rad = rm.Q[s1, s0] / kS1
        p_fr = k_fret / kS1 if k_fret > 0 else 0.0
        p_isc = rm.Q[s1, t1] / kS1 if t1 >= 0 else 0.0
        p_bl = rm.Q[s1, bl] / kS1 if bl >= 0 else 0.0
    else:
        p_rad = p_fr = p_isc = p_bl = 0.0
    if t1 >= 0:
        kT1 = rm.exit_rates[t1]
        pT = 1.0 - math.exp(-kT1 * T_laser) if kT1 > 0 else 0.0
        pT0 = rm.branch_probs[t1, s0]
        pTb = rm.branch_probs[t1, bl] if bl >= 0 else 0.0
    else:
        pT = pT0 = pTb = 0.0
    T = np.zeros((3, 3))
    T[0] = [(1.0 - p_exc) + p_exc * (p_rad + p_fr), p_exc * p_isc, p_exc * p_bl]
    T[1] = [pT * pT0, 1.0 - pT, pT * pTb]
    T[2, 2] = 1.0
    for row in range(3):
        rs = T[row].sum()
        if abs(rs - 1.0) > 1e-12:
            T[row] /= rs
    T_step = np.linalg.matrix_power(T, n_pulses)
    Ts = T[:2, :2].copy()
    I2 = np.eye(2)
    try:
        S = (I2 - np.linalg.matrix_power(Ts, n_pulses)) @ np.linalg.inv(I2 - Ts)
    except np.linalg.LinAlgError:
        S = np.zeros((2, 2))
        Tk = I2.copy()
        for _ in range(n_pulses):
            S += Tk
            Tk = Tk @ Ts
    pd = p_exc * p_rad * QY_donor
    pf = p_exc * p_fr * QY_acceptor
    mean_ph = np.zeros(2) if pd <= 0.0 else np.array([S[0, 0] * pd, S[1, 0] * pd])
    mean_fret = np.zeros(2) if pf <= 0.0 else np.array([S[0, 0] * pf, S[1, 0] * pf])
    return {"T_step": T_step, "mean_ph": mean_ph, "mean_fret": mean_fret,
            "tau_eff": 1.0 / kS1 if kS1 > 0 else 4.1e-9}


def photo_params(config: SimulationConfig, scan_duty_fraction: float = 1.0,
                 n_labeled: int = 0) -> Optional[_hip.PhotoParams]:
    """edm_photo_params of a configuration, computed on the host (None when
    the configuration has no fluorophores)."""
    sim = PhotonSimulator.__new__(PhotonSimulator)
    sim._init_tables(config, scan_duty_fraction)
    return sim._build_pp(n_labeled) if sim._active else None


class PhotonSimulator:
    """GPU PhotonSimulator with the reference's interface (photophysics.py:690)."""

    def __init__(self, config: SimulationConfig, state: ParticleState, scan_duty_fraction: float = 1.0):
        self.state = state
        self._init_tables(config, scan_duty_fraction)
        if not self._active:
            return
        dev = state._device
        if dev is None:
            raise RuntimeError("PhotonSimulator needs the state on the GPU: build the "
                               "DynamicsEngine first (as SimulationPipeline does)")
        self._dev = dev
        self._upload(int(np.sum(state._a["has_dye"])))
        self._rng = np.random.default_rng(config.random_seed + 1000)
        self.total_donor_photons = 0
        self.total_acceptor_photons = 0
        self.total_fret_events = 0
        self.total_bleached = 0

    def _init_tables(self, config: SimulationConfig, scan_duty_fraction: float) -> None:
        """Host-side tables (photophysics.py:690-887): rate matrices,
        excitation probabilities, FRET maps and the PSS matrices."""
        self.config = config
        self.rate_matrices = [RateMatrix(f) for f in config.fluorophores]
        self.n_fluoro_types = len(self.rate_matrices)
        self._active = self.n_fluoro_types > 0
        if not self._active:
            return
        if self.n_fluoro_types > 8 or len(config.fret_pairs) > 8:
            raise ConfigError("the GPU path supports at most 8 fluorophores and 8 FRET pairs")
        nf = self.n_fluoro_types
        self.scan_duty_fraction = max(0.0, min(1.0, scan_duty_fraction))
        self.laser_period = config.laser.period
        self.n_pulses_per_step = max(1, int(round(config.bd_timestep / self.laser_period)))
        self._p_excitation = np.array([self._compute_excitation_prob(f) for f in config.fluorophores])
        self._fret_pairs = config.fret_pairs
        self._species_name_to_id = {s.name: i for i, s in enumerate(config.species)}
        self._fluoro_name_to_id = {f.name: i for i, f in enumerate(config.fluorophores)}
        self._fret_maps = [dict(R0=fp.R0, kappa2=fp.kappa2, intra_complex_dist=fp.intra_complex_distance,
                                donor_fluoro_id=self._fluoro_name_to_id.get(fp.donor_type, -1),
                                acceptor_fluoro_id=self._fluoro_name_to_id.get(fp.acceptor_type, -1),
                                cfg=fp) for fp in config.fret_pairs]
        if self._fret_maps:
            self.donor_channel = self._fret_maps[0]["donor_fluoro_id"]
            self.acceptor_channel = self._fret_maps[0]["acceptor_fluoro_id"]
        else:
            self.donor_channel = 0
            self.acceptor_channel = 1 if nf > 1 else 0
        self._precompute_pss_matrices()

    # ---- host precompute (photophysics.py:796-991) ------------------------
    def _compute_excitation_prob(self, fluo: FluorophoreType) -> float:
        h, c = 6.626e-34, 3.0e8
        las = self.config.laser
        E_photon = h * c / (las.wavelength * 1e-9)
        A_focus = math.pi * (las.beam_waist * 1e-9) ** 2
        flux = (las.power / las.repetition_rate) / E_photon / A_focus
        sigma = fluo.absorption_cross_section * 1e-4
        if fluo.excitation_wavelength > 0:
            dl = abs(las.wavelength - fluo.excitation_wavelength)
            if dl > 1.0:
                sigma *= math.exp(-0.5 * (dl / (getattr(fluo, "absorption_fwhm", 45.0) / 2.355)) ** 2)
        return min(sigma * flux, 0.9)

    def _get_bound_fret_rate(self, t: int) -> float:
        for fm in self._fret_maps:
            if fm["donor_fluoro_id"] != t:
                continue
            fp = fm["cfg"]
            rm = self.rate_matrices[t]
            tD = rm.get_lifetime(rm.excited_state)
            if tD <= 0 or math.isinf(tD):
                return 0.0
            if getattr(fp, "fret_mode", "bond") == "proximity":
                E = fp.proximity_E_high
                return 1e15 if E >= 1.0 else E / ((1.0 - E) * tD)
            return (fm["R0"] / fm["intra_complex_dist"]) ** 6 / tD
        return 0.0

    def _precompute_pss_matrices(self) -> None:
        nf = self.n_fluoro_types
        self._QY_per_type = np.ones(nf)
        for t, rm in enumerate(self.rate_matrices):
            if rm.excited_state >= 0:
                self._QY_per_type[t] = rm.fluorophore.states[rm.excited_state].quantum_yield
        self._T_step_all = np.zeros((nf, 3, 3))
        self._T_step_fret_all = np.zeros((nf, 3, 3))
        self._mean_ph_all = np.zeros((nf, 2))
        self._mean_ph_fret_all = np.zeros((nf, 2))
        self._mean_fret_all = np.zeros((nf, 2))
        self._mean_fret_fret_all = np.zeros((nf, 2))
        self._tau_all = np.zeros(nf)
        self._tau_fret_all = np.zeros(nf)
        self._k_fret_bound_all = np.zeros(nf)
        for t, rm in enumerate(self.rate_matrices):
            p_exc = self._p_excitation[t] * self.scan_duty_fraction
            QYa = 1.0
            for fm in self._fret_maps:
                if fm["donor_fluoro_id"] == t:
                    if 0 <= fm["acceptor_fluoro_id"] < nf:
                        QYa = float(self._QY_per_type[fm["acceptor_fluoro_id"]])
                    break
            a = build_pss_matrices(rm, p_exc, self.laser_period, self.n_pulses_per_step, 0.0,
                                   float(self._QY_per_type[t]), QYa)
            self._T_step_all[t], self._mean_ph_all[t] = a["T_step"], a["mean_ph"]
            self._mean_fret_all[t], self._tau_all[t] = a["mean_fret"], a["tau_eff"]
            kb = self._get_bound_fret_rate(t)
            self._k_fret_bound_all[t] = kb
            b = (build_pss_matrices(rm, p_exc, self.laser_period, self.n_pulses_per_step, kb,
                                    float(self._QY_per_type[t]), QYa) if kb > 0 else a)
            self._T_step_fret_all[t], self._mean_ph_fret_all[t] = b["T_step"], b["mean_ph"]
            self._mean_fret_fret_all[t], self._tau_fret_all[t] = b["mean_fret"], b["tau_eff"]

    def _upload(self, n_labeled: int) -> None:
        pp = self._build_pp(n_labeled)
        self._pp = pp
        check(lib().edm_set_photophysics(self._dev.handle, ctypes.byref(pp)), "edm_set_photophysics")

    def _build_pp(self, n_labeled: int) -> _hip.PhotoParams:
        """The edm_photo_params block (photon-buffer bound from the number of
        labelled particles, photophysics.py:879-887)."""
        mm = max(float(self._mean_ph_all[t, 0]) for t in range(self.n_fluoro_types))
        self._max_photons = max(1000, int(n_labeled * mm * 3))
        pp = _hip.PhotoParams()
        nf = self.n_fluoro_types
        pp.n_fluoro = nf
        pp.n_fret_maps = len(self._fret_maps)
        for t, rm in enumerate(self.rate_matrices):
            for k in range(9):
                pp.T_step[t][k] = float(self._T_step_all[t].ravel()[k])
                pp.T_step_f[t][k] = float(self._T_step_fret_all[t].ravel()[k])
            for s in range(2):
                pp.mean_ph[t][s] = float(self._mean_ph_all[t, s])
                pp.mean_ph_f[t][s] = float(self._mean_ph_fret_all[t, s])
                pp.mean_fr[t][s] = float(self._mean_fret_all[t, s])
                pp.mean_fr_f[t][s] = float(self._mean_fret_fret_all[t, s])
            pp.tau[t] = float(self._tau_all[t])
            pp.k_bound[t] = float(self._k_fret_bound_all[t])
            pp.gs[t], pp.es[t] = rm.ground_state, rm.excited_state
            pp.t1[t], pp.bl[t] = rm.t1_state, rm.bleached_state
        for j, fm in enumerate(self._fret_maps):
            fp = fm["cfg"]
            d, a = fm["donor_fluoro_id"], fm["acceptor_fluoro_id"]
            pp.fm_donor[j], pp.fm_acceptor[j] = d, a
            pp.fm_intra[j] = 1 if getattr(fp, "intramolecular", False) else 0
            prox = getattr(fp, "fret_mode", "bond") == "proximity"
            pp.fm_proximity[j] = 1 if prox else 0
            skip = 0
            if d < 0 or d >= nf:
                skip = 1
            else:
                rm = self.rate_matrices[d]
                tD = rm.get_lifetime(rm.excited_state)
                if tD <= 0 or math.isinf(tD):
                    skip = 1
                elif prox:
                    tgt = self._species_name_to_id.get(fp.proximity_target_species, -1)
                    if tgt < 0:
                        skip = 1
                    pp.fm_target[j] = tgt
                    Eh, El = fp.proximity_E_high, fp.proximity_E_low
                    pp.fm_k_hi[j] = Eh / ((1.0 - Eh) * tD) if Eh < 1.0 else 1e15
                    pp.fm_k_lo[j] = El / ((1.0 - El) * tD) if El < 1.0 else 1e15
                    pp.fm_radius[j] = fp.proximity_radius * 1e-3
                else:
                    q = fm["R0"] / fm["intra_complex_dist"]
                    pp.fm_kfret[j] = q * q * q * q * q * q / tD
            pp.fm_skip[j] = skip
        pp.max_photons = int(self._max_photons)
        return pp

    # ---- per step (photophysics.py:1021-1197) --------------------------------
    def update_fret_rates(self) -> None:
        if self._active:
            check(lib().edm_fret_rates(self._dev.handle, None, None), "edm_fret_rates")

    @property
    def _k_fret(self) -> np.ndarray:
        N = self.state.N
        k = np.zeros(N)
        a = np.full(N, -1, np.int32)
        check(lib().edm_fret_rates(self._dev.handle, ptr(k), ptr(a)), "edm_fret_rates")
        return k

    @property
    def _fret_acceptor_idx(self) -> np.ndarray:
        N = self.state.N
        k = np.zeros(N)
        a = np.full(N, -1, np.int32)
        check(lib().edm_fret_rates(self._dev.handle, ptr(k), ptr(a)), "edm_fret_rates")
        return a

    def evolve_one_bd_step(self) -> PhotonBatch:
        if not self._active:
            return PhotonBatch.empty()
        N = self.state.N
        if self._dev.rng_mode == "philox":
            # emission, detection, PSF and TCSPC run fused per emitter; the
            # batch carries only its counts (process_photons returns them)
            ne, nd = ctypes.c_int64(0), ctypes.c_int64(0)
            check(lib().edm_photons_image(self._dev.handle, ctypes.byref(ne), ctypes.byref(nd)),
                  "edm_photons_image (the VirtualMicroscope must exist)")
            prev_e, prev_d = getattr(self, "_fused_prev", (0, 0))
            self._fused_prev = (int(ne.value), int(nd.value))
            self.state.mark_stale()
            b = PhotonBatch(int(ne.value) - prev_e, arrays=None)
            b._n_detected = int(nd.value) - prev_d
            return b
        u_p = self._rng.random(N)
        u_t = self._rng.random(N)
        mm = 0.0
        for t in range(self.n_fluoro_types):
            mm = max(mm, float(self._mean_ph_all[t, 0]), float(self._mean_ph_fret_all[t, 0]))
        est = max(1000, int(N * mm * 2.5))
        u_a = self._rng.random(est)
        n = ctypes.c_int64(0)
        check(lib().edm_photons(self._dev.handle, ptr(u_p), ptr(u_t), ptr(u_a), est,
                                ctypes.byref(n)), "edm_photons")
        self.state.mark_stale()
        batch = PhotonBatch(int(n.value), source=self._dev.handle)
        if batch.n_photons > 0:
            ch = batch.channel
            self.total_donor_photons += int(np.sum(ch == self.donor_channel))
            self.total_acceptor_photons += int(np.sum(ch == self.acceptor_channel))
            self.total_fret_events += int(np.sum(batch.is_fret))
        self.total_bleached = int(np.sum(self.state.is_bleached))
        return batch

    def get_donor_lifetime_histogram(self, photon_batch: PhotonBatch, n_bins: int = 256,
                                     time_range: float = 25e-9):
        edges = np.linspace(0, time_range, n_bins + 1)
        mask = photon_batch.channel == self.donor_channel
        counts, _ = np.histogram(photon_batch.arrival_time[mask], bins=edges)
        return counts, edges
