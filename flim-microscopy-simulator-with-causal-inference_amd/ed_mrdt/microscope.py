"""Virtual confocal microscope on the GPU (reference microscope.py).

PSF sigma from NA and wavelengths, pixel grid centred on the membrane, the
emitter-centred PSF pixel draw, IRF jitter and TCSPC binning into the
device-resident ``uint32[n_ch, nx, ny, n_bins]`` stack, and frame-end dark
counts / afterpulses.  With the reference's numpy streams (rng="numpy") the
host draws exactly what VirtualMicroscope draws (microscope.py:1095-1148,
859-920) and the device consumes them; with rng="philox" detection, pixel
draw, jitter and binning run fused with photon emission and the frame noise
is drawn on the device.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np

from . import _hip
from ._hip import check, lib, ptr
from .config import ConfigError, SimulationConfig
from .particles import ParticleState

__all__ = ["PSF", "ConfocalScanner", "TCSPCAccumulator", "DetectorModel", "VirtualMicroscope",
           "FLIMFrame"]


class PSF:
    """Gaussian confocal PSF per channel (reference microscope.py:73-226):
    sigma_conf = 1 / sqrt(sigma_ill^-2 + sigma_det^-2), sigma = 0.21 lambda / NA."""

    def __init__(self, config: SimulationConfig):
        self.config = config
        self.NA = config.objective.NA
        self.psf_type = config.psf.type
        self.sigma_um: Dict[int, float] = {}
        self.fwhm_um: Dict[int, float] = {}
        for ch, fl in enumerate(config.fluorophores):
            lam_em = config.psf.emission_wavelength
            for st in fl.states:
                if st.is_fluorescent and st.emission_wavelength > 0:
                    lam_em = st.emission_wavelength
                    break
            s_ill = 0.21 * config.laser.wavelength / self.NA
            s_det = 0.21 * lam_em / self.NA
            s = 1.0 / math.sqrt(1.0 / (s_ill ** 2) + 1.0 / (s_det ** 2)) / 1000.0
            self.sigma_um[ch] = s
            self.fwhm_um[ch] = s * 2.0 * math.sqrt(2.0 * math.log(2.0))
        if 0 in self.sigma_um:
            self._default_sigma = self.sigma_um[0]
        else:
            self._default_sigma = 0.21 * config.psf.emission_wavelength / self.NA / 1000.0 / math.sqrt(2.0)

    def get_sigma(self, channel: int = 0) -> float:
        return self.sigma_um.get(channel, self._default_sigma)

    def evaluate(self, r_um: float, channel: int = 0) -> float:
        s = self.get_sigma(channel)
        return math.exp(-0.5 * (r_um / s) ** 2) / (2.0 * math.pi * s ** 2)

    def evaluate_array(self, r_um: np.ndarray, channel: int = 0) -> np.ndarray:
        s = self.get_sigma(channel)
        return np.exp(-0.5 * (r_um / s) ** 2) / (2.0 * math.pi * s ** 2)


class ConfocalScanner:
    """Pixel grid centred on the membrane (reference microscope.py:241-296)."""

    def __init__(self, config: SimulationConfig, psf: PSF):
        self.config = config
        self.psf = psf
        self.pixel_size_um = config.scan.pixel_size / 1000.0
        self.nx, self.ny = config.scan.pixels_x, config.scan.pixels_y
        self.fov_x = self.nx * self.pixel_size_um
        self.fov_y = self.ny * self.pixel_size_um
        Lx, Ly = config.membrane.size
        self.offset_x = (Lx - self.fov_x) / 2.0
        self.offset_y = (Ly - self.fov_y) / 2.0
        max_sigma = max(psf.sigma_um.values()) if psf.sigma_um else 0.1
        self.psf_range_um = 4.0 * max_sigma
        self.psf_range_px = int(math.ceil(self.psf_range_um / self.pixel_size_um))
        self.pixel_cx = (np.arange(self.nx) + 0.5) * self.pixel_size_um + self.offset_x
        self.pixel_cy = (np.arange(self.ny) + 0.5) * self.pixel_size_um + self.offset_y


class TCSPCAccumulator:
    """TCSPC settings and IRF width (reference microscope.py:650-712)."""

    def __init__(self, config: SimulationConfig):
        self.n_bins = config.tcspc.n_bins
        self.time_range = config.tcspc.time_range
        self.bin_width = config.tcspc.bin_width
        self.nx, self.ny = config.scan.pixels_x, config.scan.pixels_y
        self.n_channels = max(2, len(config.fluorophores))
        f2s = 1.0 / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        self.sigma_irf = math.sqrt((config.laser.pulse_width * f2s) ** 2
                                   + config.detector.timing_jitter ** 2)


class DetectorModel:
    """Detector constants (reference microscope.py:819-838)."""

    def __init__(self, config: SimulationConfig):
        d = config.detector
        self.det_efficiency = d.detection_efficiency
        self.dark_count_rate = d.dark_count_rate
        self.afterpulsing_prob = d.afterpulsing_prob
        self.dead_time = d.dead_time
        self.dwell_time = config.scan.dwell_time


@dataclass
class FLIMFrame:
    flim_stack: np.ndarray
    intensity: np.ndarray
    frame_index: int = 0
    n_signal_photons: int = 0
    n_dark_counts: int = 0
    n_afterpulses: int = 0

    @staticmethod
    def from_stack(flim_stack, frame_index=0, n_signal=0, n_dark=0, n_ap=0) -> "FLIMFrame":
        return FLIMFrame(flim_stack=flim_stack, intensity=flim_stack.sum(axis=-1).astype(np.float64),
                         frame_index=frame_index, n_signal_photons=n_signal, n_dark_counts=n_dark,
                         n_afterpulses=n_ap)

    @property
    def total_counts(self) -> int:
        return int(self.flim_stack.sum())

    @property
    def shape(self) -> Tuple[int, ...]:
        return self.flim_stack.shape


@dataclass
class SparseFrame:
    """One FLIM frame as its nonzero TCSPC bins: ``flat_index`` into the
    stack ``[n_ch, nx, ny, n_bins]`` (C order, ascending) and ``count``.  A
    served or batched job streams its frames back in this form;
    ``to_frame()`` gives the reference's dense ``FLIMFrame``."""
    shape: Tuple[int, int, int, int]
    flat_index: np.ndarray
    count: np.ndarray
    frame_index: int = 0
    n_signal_photons: int = 0
    n_dark_counts: int = 0
    n_afterpulses: int = 0

    @staticmethod
    def from_stack(flim_stack, frame_index=0, n_signal=0, n_dark=0, n_ap=0) -> "SparseFrame":
        flat = np.ascontiguousarray(flim_stack).reshape(-1)
        nz = np.flatnonzero(flat)
        return SparseFrame(tuple(int(v) for v in flim_stack.shape), nz.astype(np.uint32),
                           flat[nz].astype(np.uint32), frame_index, n_signal, n_dark, n_ap)

    def to_frame(self) -> FLIMFrame:
        stack = np.zeros(self.shape, dtype=np.uint32)
        stack.reshape(-1)[self.flat_index.astype(np.int64)] = self.count
        return FLIMFrame.from_stack(stack, frame_index=self.frame_index,
                                    n_signal=self.n_signal_photons, n_dark=self.n_dark_counts,
                                    n_ap=self.n_afterpulses)

    @property
    def total_counts(self) -> int:
        return int(self.count.sum(dtype=np.int64))

    def same_counts(self, other: "SparseFrame") -> bool:
        return (self.shape == other.shape and np.array_equal(self.flat_index, other.flat_index)
                and np.array_equal(self.count, other.count))


def image_params(config: SimulationConfig, psf: "PSF" = None, scanner: "ConfocalScanner" = None,
                 tcspc: "TCSPCAccumulator" = None,
                 detector: "DetectorModel" = None) -> _hip.ImageParams:
    """edm_image_params of a configuration (microscope.py:73-391, 650-838)."""
    psf = psf or PSF(config)
    sc = scanner or ConfocalScanner(config, psf)
    tc = tcspc or TCSPCAccumulator(config)
    de = detector or DetectorModel(config)
    ip = _hip.ImageParams()
    ip.nx, ip.ny, ip.n_bins, ip.n_ch = sc.nx, sc.ny, tc.n_bins, tc.n_channels
    ip.R = sc.psf_range_px
    n_sig = max((max(psf.sigma_um.keys()) + 1) if psf.sigma_um else 1, 2)
    if n_sig > 8:
        raise ConfigError("the GPU path supports at most 8 channels")
    ip.n_sigma = n_sig
    for ch, s in psf.sigma_um.items():
        ip.sigma_ch[ch] = s
    ip.pixel, ip.offx, ip.offy = sc.pixel_size_um, sc.offset_x, sc.offset_y
    ip.bin_w, ip.sigma_irf = tc.bin_width, tc.sigma_irf
    ip.qe = de.det_efficiency
    ip.dark_lambda = de.dark_count_rate * de.dwell_time * 1
    ip.p_ap = de.afterpulsing_prob
    return ip


class VirtualMicroscope:
    """GPU VirtualMicroscope with the reference's interface (microscope.py:1021)."""

    def __init__(self, config: SimulationConfig, state: ParticleState):
        self.config = config
        self.state = state
        self.psf = PSF(config)
        self.scanner = ConfocalScanner(config, self.psf)
        self.tcspc = TCSPCAccumulator(config)
        self.detector = DetectorModel(config)
        dev = state._device
        if dev is None:
            raise RuntimeError("VirtualMicroscope needs the state on the GPU: build the "
                               "DynamicsEngine first (as SimulationPipeline does)")
        self._dev = dev
        ip = image_params(config, self.psf, self.scanner, self.tcspc, self.detector)
        self._ip = ip
        check(lib().edm_set_imaging(dev.handle, ctypes.byref(ip)), "edm_set_imaging")
        self._rng = np.random.default_rng(config.random_seed + 2000)
        self._frame_index = 0
        self._n_signal = 0
        self._n_bd_steps_in_frame = 0
        self.total_photons_received = 0
        self.total_photons_detected = 0
        self.total_dark_counts = 0
        self.total_afterpulses = 0

    @property
    def frame_shape(self) -> Tuple[int, int, int, int]:
        return (self.tcspc.n_channels, self.scanner.nx, self.scanner.ny, self.tcspc.n_bins)

    def process_photons(self, photon_batch) -> int:
        """microscope.py:1069-1157."""
        n_ph = photon_batch.n_photons
        if self._dev.rng_mode == "philox":
            nd = int(getattr(photon_batch, "_n_detected", 0))
            self.total_photons_received += max(n_ph, 0)
            self.total_photons_detected += nd
            self._n_signal += nd
            self._n_bd_steps_in_frame += 1
            return nd
        if n_ph == 0:
            return 0
        self.total_photons_received += n_ph
        L, h = lib(), self._dev.handle
        u_qe = self._rng.random(n_ph)
        ns = ctypes.c_int64(0)
        check(L.edm_image_detect(h, ptr(u_qe), n_ph, ctypes.byref(ns)), "edm_image_detect")
        if ns.value == 0:
            self._n_bd_steps_in_frame += 1
            return 0
        u_pix = self._rng.random(int(ns.value))
        na = ctypes.c_int64(0)
        check(L.edm_image_assign(h, ptr(u_pix), ns.value, ctypes.byref(na)), "edm_image_assign")
        if na.value == 0:
            self._n_bd_steps_in_frame += 1
            return 0
        jit = np.ascontiguousarray(self._rng.normal(0.0, self.tcspc.sigma_irf, size=int(na.value)))
        check(L.edm_image_bin(h, ptr(jit), na.value), "edm_image_bin")
        self.total_photons_detected += int(na.value)
        self._n_signal += int(na.value)
        self._n_bd_steps_in_frame += 1
        return int(na.value)

    def _frame_noise_numpy(self) -> Tuple[int, int]:
        """Dark counts and afterpulses drawn exactly as DetectorModel does
        (microscope.py:859-920); the counts are added on the device."""
        L, h = lib(), self._dev.handle
        n_ch, nx, ny, nb = self.frame_shape
        lam = self.detector.dark_count_rate * self.detector.dwell_time * 1
        g = self._rng
        idx = []
        n_dark = 0
        if lam >= 1e-10:
            for c in range(n_ch):
                for i in range(nx):
                    for j in range(ny):
                        k = g.poisson(lam)
                        if k > 0:
                            base = ((c * nx + i) * ny + j) * nb
                            idx.extend(int(base + b) for b in g.integers(0, nb, size=k))
                            n_dark += k
        if idx:
            a = np.asarray(idx, dtype=np.int64)
            check(L.edm_frame_add_counts(h, ptr(a), a.size), "edm_frame_add_counts")
        n_ap = 0
        p = self.detector.afterpulsing_prob
        if p > 0:
            tot = np.empty(n_ch * nx * ny, dtype=np.uint32)
            check(L.edm_frame_pixel_totals(h, ptr(tot)), "edm_frame_pixel_totals")
            idx = []
            for q in np.flatnonzero(tot):
                k = g.binomial(int(tot[q]), p)
                if k > 0:
                    idx.extend(int(q * nb + b) for b in g.integers(0, nb, size=k))
                    n_ap += k
            if idx:
                a = np.asarray(idx, dtype=np.int64)
                check(L.edm_frame_add_counts(h, ptr(a), a.size), "edm_frame_add_counts")
        return n_dark, n_ap

    def finalize_frame(self) -> FLIMFrame:
        """microscope.py:1159-1203."""
        L, h = lib(), self._dev.handle
        if self._dev.rng_mode == "philox":
            nd, na = ctypes.c_int64(0), ctypes.c_int64(0)
            check(L.edm_frame_noise(h, ctypes.byref(nd), ctypes.byref(na)), "edm_frame_noise")
            n_dark, n_ap = int(nd.value), int(na.value)
        else:
            n_dark, n_ap = self._frame_noise_numpy()
        self.total_dark_counts += n_dark
        self.total_afterpulses += n_ap
        stack = np.empty(self.frame_shape, dtype=np.uint32)
        check(L.edm_frame_download(h, ptr(stack), stack.size), "edm_frame_download")
        frame = FLIMFrame.from_stack(stack, frame_index=self._frame_index, n_signal=self._n_signal,
                                     n_dark=n_dark, n_ap=n_ap)
        check(L.edm_frame_reset(h), "edm_frame_reset")
        self._frame_index += 1
        self._n_signal = 0
        self._n_bd_steps_in_frame = 0
        if self._dev.rng_mode == "philox":
            self._last_detected = self._detected_counter()
        return frame

    # ---- streamed frames (no reference counterpart: the reference keeps every
    # frame of a run, pipeline.py:645, by copying each dense stack) ------------
    def _detected_counter(self) -> int:
        c = np.zeros(2, dtype=np.int64)
        check(lib().edm_photon_counters(self._dev.handle, ptr(c)), "edm_photon_counters")
        return int(c[1])

    def collect_frame(self) -> int:
        """End the current frame as finalize_frame does (Philox draws), but
        leave its detector noise, compaction to nonzero bins and reset to a
        second stream while the next steps run into a second stack; returns
        a ticket for fetch_frame.  One frame may be pending."""
        if self._dev.rng_mode != "philox":
            raise ValueError("collect_frame needs the Philox draws; use finalize_frame")
        t = ctypes.c_int32(-1)
        check(lib().edm_frame_collect(self._dev.handle, ctypes.byref(t)), "edm_frame_collect")
        self._pending = (int(t.value), self._frame_index)
        self._frame_index += 1
        self._n_signal = 0
        self._n_bd_steps_in_frame = 0
        return int(t.value)

    def fetch_frame(self, ticket: int) -> SparseFrame:
        """The collected frame (waits for its noise and compaction)."""
        pend = getattr(self, "_pending", None)
        if pend is None or pend[0] != ticket:
            raise ValueError("no pending frame with that ticket")
        L, h = lib(), self._dev.handle
        meta = np.zeros(3, dtype=np.int64)
        total = ctypes.c_int64(0)
        check(L.edm_frame_fetch(h, int(ticket), ptr(meta), None, None, 0, ctypes.byref(total)),
              "edm_frame_fetch")
        n = int(total.value)
        idx = np.empty(max(n, 1), dtype=np.uint32)
        cnt = np.empty(max(n, 1), dtype=np.uint32)
        check(L.edm_frame_fetch(h, int(ticket), None, ptr(idx), ptr(cnt), idx.size,
                                ctypes.byref(total)), "edm_frame_fetch")
        self._pending = None
        n_dark, n_ap = int(meta[0]), int(meta[1])
        n_signal = int(meta[2]) - getattr(self, "_last_detected", 0)
        self._last_detected = int(meta[2])
        self.total_dark_counts += n_dark
        self.total_afterpulses += n_ap
        return SparseFrame(self.frame_shape, idx[:n].copy(), cnt[:n].copy(), pend[1],
                           n_signal_photons=n_signal, n_dark_counts=n_dark, n_afterpulses=n_ap)
