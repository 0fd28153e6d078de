"""Observables of a running membrane, reduced on the device.

The reference computes its validation and analysis statistics on the host
from copies of the state and the FLIM stacks: the MSD of benchmark B1
(validation.py:256-335), the RDF (analysis.py:733-790), the mean arrival
time (analysis.py:87-127) and the phasor (analysis.py:497-566).  Here the
reductions run on the GPU (``edm_stats_*``) and only their results cross to
the host; the final ratios and normalisations are the reference's.

    msd = MSDTracker(engine)          # accumulates from now on, every step
    engine.run(1000)
    msd.msd(["PIP2", "PIP3"])          # MSD in um^2
    r, g = rdf(engine, ["PIP2"], n_bins=50, r_max_um=1.0)
    tau, I = mean_arrival_time(microscope, channel=0)
    ph = phasor_transform(microscope, channel=0)

Species-count trajectories come from the per-step census of
``pipeline.SubFrameRecorder`` (also a device reduction).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Iterable, Optional, Tuple

import numpy as np

from ._hip import check, lib, ptr


def _mask(engine, species: Optional[Iterable[str]]) -> int:
    names = [s.name for s in engine.config.species]
    if species is None:
        return (1 << len(names)) - 1
    m = 0
    for s in species:
        m |= 1 << names.index(s)
    return m


class MSDTracker:
    """Mean squared displacement since construction, periodic jumps
    unwrapped (validation.py:276-290), over the alive particles of the
    chosen species."""

    def __init__(self, engine):
        self.engine = engine
        check(lib().edm_stats_open(engine.handle), "edm_stats_open")

    def sums(self, species: Optional[Iterable[str]] = None) -> Tuple[float, int]:
        s, c = ctypes.c_double(0.0), ctypes.c_int64(0)
        check(lib().edm_stats_msd(self.engine.handle, _mask(self.engine, species),
                                  ctypes.byref(s), ctypes.byref(c)), "edm_stats_msd")
        return float(s.value), int(c.value)

    def msd(self, species: Optional[Iterable[str]] = None) -> float:
        s, c = self.sums(species)
        return s / c if c else 0.0

    def close(self) -> None:
        check(lib().edm_stats_close(self.engine.handle), "edm_stats_close")


def pair_distance_histogram(engine, species: Optional[Iterable[str]] = None, n_bins: int = 50,
                            r_max_um: float = 1.0) -> np.ndarray:
    """Unordered pairs of alive particles of the species by distance
    (bins of r_max / n_bins)."""
    out = np.zeros(n_bins, dtype=np.uint64)
    check(lib().edm_stats_pair_hist(engine.handle, _mask(engine, species), int(n_bins),
                                    float(r_max_um), ptr(out)), "edm_stats_pair_hist")
    return out


def rdf(engine, species: Optional[Iterable[str]] = None, n_bins: int = 50,
        r_max_um: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """g(r) of the alive particles of the species, normalised as
    spatial_correlation_ground_truth (analysis.py:733-790): each pair
    counts twice, g = counts / (N rho 2 pi r dr)."""
    counts = 2.0 * pair_distance_histogram(engine, species, n_bins, r_max_um).astype(np.float64)
    st = engine.state
    names = [s.name for s in engine.config.species]
    sel = st.is_alive.copy()
    if species is not None:
        sel &= np.isin(st.species_id, [names.index(s) for s in species])
    N = int(sel.sum())
    Lx, Ly = engine.config.membrane.size
    rho = N / (Lx * Ly)
    dr = r_max_um / n_bins
    r = (np.arange(n_bins) + 0.5) * dr
    g = np.zeros(n_bins)
    shell = 2.0 * np.pi * r * dr
    if N > 1:
        g = counts / (N * rho * shell)
    return r, g


def flim_moments(microscope, channel: int = 0, harmonic: int = 1) -> np.ndarray:
    """Per pixel of the current frame: [sum I, sum I cos(phi_k),
    sum I sin(phi_k), sum I t_k (ns)], shape [nx, ny, 4]."""
    dev = microscope._dev
    nx, ny = microscope.frame_shape[1], microscope.frame_shape[2]
    out = np.zeros((nx, ny, 4), dtype=np.float64)
    check(lib().edm_stats_flim(dev.handle, int(channel), int(harmonic), ptr(out)),
          "edm_stats_flim")
    return out


def mean_arrival_time(microscope, channel: int = 0, min_counts: int = 10):
    """analysis.py:87-127 on the current frame: (tau_map in ns, NaN below
    min_counts; intensity)."""
    m = flim_moments(microscope, channel)
    I = m[..., 0]
    tau = np.full_like(I, np.nan)
    ok = I >= min_counts
    tau[ok] = m[..., 3][ok] / I[ok]
    return tau, I


@dataclass
class PhasorResult:
    g: np.ndarray
    s: np.ndarray
    intensity: np.ndarray
    tau_phi: np.ndarray
    tau_mod: np.ndarray
    mask: np.ndarray
    omega: float
    harmonic: int


def phasor_transform(microscope, channel: int = 0, harmonic: int = 1,
                     min_counts: int = 20) -> PhasorResult:
    """analysis.py:497-566 on the current frame."""
    m = flim_moments(microscope, channel, harmonic)
    time_range = microscope.config.tcspc.time_range
    omega = 2.0 * np.pi * harmonic / time_range
    I = m[..., 0]
    mask = I >= min_counts
    g = np.zeros_like(I)
    s = np.zeros_like(I)
    g[mask] = m[..., 1][mask] / I[mask]
    s[mask] = m[..., 2][mask] / I[mask]
    omega_ns = omega * 1e-9
    tau_phi = np.full_like(g, np.nan)
    tau_mod = np.full_like(g, np.nan)
    valid = mask & (g > 1e-10)
    tau_phi[valid] = (s[valid] / g[valid]) / omega_ns
    m2 = g ** 2 + s ** 2
    mv = valid & (m2 > 1e-10) & (m2 < 1.0 - 1e-10)
    tau_mod[mv] = np.sqrt(1.0 / m2[mv] - 1.0) / omega_ns
    return PhasorResult(g, s, I, tau_phi, tau_mod, mask, omega, harmonic)
