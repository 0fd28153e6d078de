"""ctypes binding of libedm_hip.so (C ABI in include/edm_hip.h).

The library is loaded from this package directory.  There is no fallback:
if the library is missing or no GPU is visible, ``lib()`` raises.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libedm_hip.so")

c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p

EDM_RX_BIND, EDM_RX_CAT, EDM_RX_RECRUIT, EDM_RX_DETACH, EDM_RX_CONVERT, EDM_RX_ADAPT = 1, 2, 3, 4, 5, 6


class Feedback(ctypes.Structure):
    _fields_ = [("sensor", c_int32), ("grid", c_int32), ("inhibition", c_int32), ("pad", c_int32),
                ("K_half", c_double), ("n_hill", c_double), ("basal", c_double)]


def no_feedback() -> Feedback:
    return Feedback(sensor=-1, grid=-1, inhibition=0, pad=0, K_half=0.0, n_hill=0.0, basal=0.0)


class Grid(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("sx", c_double), ("sy", c_double)]


class Reaction(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("a", c_int32), ("b", c_int32), ("c", c_int32),
                ("contact", c_double), ("prob", c_double), ("prob2", c_double),
                ("max_recruits", c_int32), ("use_adapt", c_int32),
                ("fb1", Feedback), ("fb2", Feedback)]


class Params(ctypes.Structure):
    _fields_ = [("n", c_int32), ("n_species", c_int32), ("fp64", c_int32),
                ("periodic_x", c_int32), ("periodic_y", c_int32), ("ncx", c_int32),
                ("ncy", c_int32), ("use_wca", c_int32), ("use_harmonic", c_int32),
                ("dfield", c_int32), ("n_regions", c_int32), ("n_grids", c_int32),
                ("n_reactions", c_int32), ("pair_cap", c_int32), ("exact_force_order", c_int32),
                ("Lx", c_double), ("Ly", c_double), ("dt", c_double), ("cutoff", c_double),
                ("csx", c_double), ("csy", c_double), ("max_disp", c_double),
                ("smooth_w", c_double), ("sin_amp", c_double), ("sin_kx", c_double),
                ("sin_ky", c_double), ("adapt_tau", c_double), ("adapt_K", c_double),
                ("adapt_n", c_double), ("adapt_grid", c_int32), ("adapt_species", c_int32),
                ("seed", ctypes.c_uint64),
                ("sigma", c_void_p), ("epsilon", c_void_p), ("k_harm", c_void_p),
                ("D_base", c_void_p), ("rcx", c_void_p), ("rcy", c_void_p), ("rr", c_void_p),
                ("rf", c_void_p), ("grids", c_void_p), ("reactions", c_void_p)]


class HostState(ctypes.Structure):
    _fields_ = [("positions", c_void_p), ("species_id", c_void_p), ("bond_partner", c_void_p),
                ("is_alive", c_void_p), ("has_dye", c_void_p), ("is_bleached", c_void_p),
                ("dye_type_id", c_void_p), ("ctmc_state", c_void_p)]


D8 = c_double * 8
I8 = c_int32 * 8


class PhotoParams(ctypes.Structure):
    _fields_ = [("n_fluoro", c_int32), ("n_fret_maps", c_int32),
                ("T_step", (c_double * 9) * 8), ("T_step_f", (c_double * 9) * 8),
                ("mean_ph", (c_double * 2) * 8), ("mean_ph_f", (c_double * 2) * 8),
                ("mean_fr", (c_double * 2) * 8), ("mean_fr_f", (c_double * 2) * 8),
                ("tau", D8), ("k_bound", D8),
                ("gs", I8), ("es", I8), ("t1", I8), ("bl", I8),
                ("fm_donor", I8), ("fm_acceptor", I8), ("fm_intra", I8), ("fm_proximity", I8),
                ("fm_target", I8), ("fm_kfret", D8), ("fm_k_hi", D8), ("fm_k_lo", D8),
                ("fm_radius", D8), ("fm_skip", I8), ("max_photons", c_int32)]


class ImageParams(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("n_bins", c_int32), ("n_ch", c_int32),
                ("R", c_int32), ("n_sigma", c_int32), ("pixel", c_double), ("offx", c_double),
                ("offy", c_double), ("bin_w", c_double), ("sigma_irf", c_double),
                ("qe", c_double), ("dark_lambda", c_double), ("p_ap", c_double),
                ("sigma_ch", D8)]


class Replica(ctypes.Structure):
    _fields_ = [("n", c_int32), ("ncx", c_int32), ("ncy", c_int32), ("n_grids", c_int32),
                ("Lx", c_double), ("Ly", c_double), ("csx", c_double), ("csy", c_double),
                ("cutoff", c_double), ("sin_kx", c_double), ("sin_ky", c_double),
                ("seed", ctypes.c_uint64), ("grids", Grid * 4), ("prob", c_double * 16),
                ("max_recruits", c_int32 * 16), ("img_nx", c_int32), ("img_ny", c_int32),
                ("img_R", c_int32), ("n_reactions", c_int32), ("img_pixel", c_double),
                ("img_offx", c_double), ("img_offy", c_double), ("n_regions", c_int32),
                ("pad_", c_int32), ("reg_cx", c_double * 8), ("reg_cy", c_double * 8),
                ("reg_r", c_double * 8), ("reg_f", c_double * 8), ("prob2", c_double * 16)]


class Census(ctypes.Structure):
    """edm_census: one reduction over a membrane (pipeline.py:567-642, 357-455)."""
    _fields_ = [("step", c_int64), ("n_bond_alive", c_int64), ("n_bond", c_int64),
                ("n_bleached", c_int64), ("n_donor0", c_int64), ("n_free_dye", c_int64 * 8),
                ("species_alive", c_int64 * 32), ("events", c_int64 * 6), ("photons", c_int64),
                ("e_sum", c_double)]


# numpy view of an array of edm_census rows
CENSUS_DTYPE = np.dtype([("step", "<i8"), ("n_bond_alive", "<i8"), ("n_bond", "<i8"),
                         ("n_bleached", "<i8"), ("n_donor0", "<i8"), ("n_free_dye", "<i8", (8,)),
                         ("species_alive", "<i8", (32,)), ("events", "<i8", (6,)),
                         ("photons", "<i8"), ("e_sum", "<f8")])

# recording flags (edm_record_open)
REC_FRET_SUM = 1
REC_AUTO = 2

# tuning keys (edm_set_tuning)
TUNE_TILE_ROWS = 1
TUNE_CAT_SIDE = 2
TUNE_CAT_WALK = 3
TUNE_CAT_LANES = 4
TUNE_FRET_SEARCH = 5
TUNE_GRAPHS = 6
TUNE_GRAPH_REPLAYS = 7
TUNE_GRAPH_STATUS = 8
TUNE_FORCE_LISTS = 9
ENS_TUNE_STRIP_PARTICLES = 1
ENS_TUNE_PAIR_BELOW = 2
ENS_TUNE_CAT_LANES = 3
ENS_TUNE_CAT_WALK = 4

# function name -> (argtypes) ; every function returns int32 status
_SIGS = {
    "edm_device_count": [ctypes.POINTER(ctypes.c_int)],
    "edm_create": [ctypes.POINTER(Params), ctypes.POINTER(c_void_p)],
    "edm_destroy": [c_void_p],
    "edm_upload_state": [c_void_p, ctypes.POINTER(HostState)],
    "edm_download_state": [c_void_p, ctypes.POINTER(HostState)],
    "edm_set_photophysics": [c_void_p, ctypes.POINTER(PhotoParams)],
    "edm_set_imaging": [c_void_p, ctypes.POINTER(ImageParams)],
    "edm_step_begin": [c_void_p, ctypes.POINTER(c_int64)],
    "edm_step_finish": [c_void_p, c_void_p, c_int64],
    "edm_run": [c_void_p, c_int32],
    "edm_counters": [c_void_p, c_void_p],
    "edm_photon_counters": [c_void_p, c_void_p],
    "edm_set_tuning": [c_void_p, c_int32, c_int64],
    "edm_get_tuning": [c_void_p, c_int32, ctypes.POINTER(c_int64)],
    "edm_pairs": [c_void_p, c_void_p, c_int64, ctypes.POINTER(c_int64)],
    "edm_density": [c_void_p, c_int32, c_int32, c_void_p],
    "edm_photons": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, ctypes.POINTER(c_int64)],
    "edm_photon_batch": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64],
    "edm_fret_rates": [c_void_p, c_void_p, c_void_p],
    "edm_image_detect": [c_void_p, c_void_p, c_int64, ctypes.POINTER(c_int64)],
    "edm_image_assign": [c_void_p, c_void_p, c_int64, ctypes.POINTER(c_int64)],
    "edm_image_bin": [c_void_p, c_void_p, c_int64],
    "edm_photons_image": [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)],
    "edm_frame_pixel_totals": [c_void_p, c_void_p],
    "edm_frame_add_counts": [c_void_p, c_void_p, c_int64],
    "edm_frame_noise": [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)],
    "edm_frame_download": [c_void_p, c_void_p, c_int64],
    "edm_frame_reset": [c_void_p],
    "edm_synchronize": [c_void_p],
    "edm_time_steps": [c_void_p, c_int32, c_int32, ctypes.POINTER(c_double),
                       ctypes.POINTER(c_double), c_void_p],
    "edm_time_mark": [c_void_p, c_int32, ctypes.POINTER(c_double)],
    "edm_run_frame": [c_void_p, c_int32],
    "edm_frame_collect": [c_void_p, c_void_p],
    "edm_frame_fetch": [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "edm_record_open": [c_void_p, c_int32, c_int32],
    "edm_record_step": [c_void_p],
    "edm_record_rows": [c_void_p, ctypes.POINTER(c_int32)],
    "edm_record_read": [c_void_p, c_void_p, c_int32, ctypes.POINTER(c_int32)],
    "edm_record_close": [c_void_p],
    "edm_ground_truth": [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                         ctypes.POINTER(Census)],
    "edm_ens_create": [ctypes.POINTER(Params), ctypes.POINTER(PhotoParams),
                       ctypes.POINTER(ImageParams), c_int64, c_int32, ctypes.POINTER(c_void_p)],
    "edm_ens_destroy": [c_void_p],
    "edm_ens_add": [c_void_p, ctypes.POINTER(Replica), ctypes.POINTER(HostState),
                    ctypes.POINTER(c_int32)],
    "edm_ens_remove": [c_void_p, c_int32],
    "edm_ens_run": [c_void_p, c_int32, c_int32],
    "edm_ens_frame": [c_void_p, c_int32, c_void_p, c_int64, ctypes.POINTER(c_int64),
                      ctypes.POINTER(c_int64)],
    "edm_ens_frame_end": [c_void_p, c_void_p, c_int32],
    "edm_ens_noise_counters": [c_void_p, c_int32, c_void_p],
    "edm_ens_frames_collect": [c_void_p, c_void_p, c_int32, ctypes.POINTER(c_int32)],
    "edm_ens_frames_fetch": [c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                             ctypes.POINTER(c_int64)],
    "edm_ens_download": [c_void_p, c_int32, ctypes.POINTER(HostState)],
    "edm_ens_counters": [c_void_p, c_int32, c_void_p],
    "edm_ens_info": [c_void_p, c_void_p],
    "edm_ens_set_tuning": [c_void_p, c_int32, c_int64],
    "edm_ens_get_tuning": [c_void_p, c_int32, c_void_p],
    "edm_ens_synchronize": [c_void_p],
    # observables
    "edm_stats_open": [c_void_p],
    "edm_stats_close": [c_void_p],
    "edm_stats_msd": [c_void_p, ctypes.c_uint32, ctypes.POINTER(c_double),
                      ctypes.POINTER(c_int64)],
    "edm_stats_pair_hist": [c_void_p, ctypes.c_uint32, c_int32, c_double, c_void_p],
    "edm_stats_flim": [c_void_p, c_int32, c_int32, c_void_p],
    # device memory for the free kernels
    "edm_device_alloc": [c_int64, ctypes.POINTER(c_void_p)],
    "edm_device_free": [c_void_p],
    "edm_copy_to_device": [c_void_p, c_void_p, c_int64],
    "edm_copy_to_host": [c_void_p, c_void_p, c_int64],
    # the reference's free kernels on device arrays (dynamics.py:38-60)
    "edm_cell_pairs": [c_void_p, c_int32, c_int32, c_int32, c_double, c_double, c_double,
                       c_double, c_int32, c_int32, c_double, c_int64, c_void_p,
                       ctypes.POINTER(c_int64), ctypes.POINTER(c_int64), c_void_p],
    "edm_cell_pairs_noalloc": [c_void_p, c_int32, c_int32, c_int32, c_double, c_double,
                               c_double, c_double, c_int32, c_int32, c_double, c_int64,
                               c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64),
                               c_void_p],
    "edm_bd_step": [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double,
                    c_int32, c_int32, c_int32, c_void_p, c_void_p, c_double, c_void_p],
    "edm_diffusion_field_regions": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_double,
                                    c_void_p],
    "edm_pair_forces": [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                        c_int64, c_void_p, c_void_p, c_int32, c_double, c_double, c_int32,
                        c_int32, c_void_p],
    "edm_binding": [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                    c_double, c_double, c_void_p, c_int64, c_void_p, c_double, c_double, c_int32,
                    c_int32, ctypes.POINTER(c_int64), c_void_p],
    "edm_dissociation": [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                         c_int32, c_double, c_void_p, ctypes.POINTER(c_int64), c_void_p],
    "edm_catalysis": [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_double,
                      c_double, c_void_p, c_int64, c_void_p, c_double, c_double, c_int32,
                      c_int32, c_void_p, c_int32, c_int32, c_double, c_double, c_double,
                      c_double, c_int32, c_double, c_void_p, c_int32, c_int32, c_double,
                      c_double, ctypes.POINTER(c_int64), c_void_p],
    "edm_recruitment": [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_double, c_double,
                        c_double, c_void_p, c_void_p, c_int32,
                        c_void_p, c_int32, c_int32, c_double, c_double, c_double, c_double,
                        c_int32, c_double,
                        c_void_p, c_int32, c_int32, c_double, c_double, c_double, c_double,
                        c_int32, c_double,
                        c_void_p, c_void_p, c_void_p, c_double, ctypes.POINTER(c_int64),
                        c_void_p],
    "edm_detachment": [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_double, c_void_p,
                       c_void_p, c_int32, c_int32, c_double, c_double, c_double, c_double,
                       c_int32, c_double, ctypes.POINTER(c_int64), c_void_p],
    "edm_conversion": [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_double,
                       c_void_p, c_void_p, c_int32, c_int32, c_double, c_double, c_double,
                       c_double, c_int32, c_double, ctypes.POINTER(c_int64), c_void_p],
    "edm_density_field": [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                          c_double, c_double, c_void_p, c_void_p],
    "edm_adaptation_update": [c_void_p, c_void_p, c_int32, c_int32, c_double, c_double,
                              c_void_p],
}

EXPORTS = tuple(_SIGS) + ("edm_last_error",)

_lib: Optional[ctypes.CDLL] = None


class HipError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libedm_hip.so and declare every entry point (no GPU call)."""
    if not os.path.exists(path):
        raise HipError(f"{path} is missing: build it with build_hip.py (no CPU fallback exists)")
    L = ctypes.CDLL(path)
    for name, argt in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = argt
        f.restype = c_int32
    L.edm_last_error.argtypes = []
    L.edm_last_error.restype = ctypes.c_char_p
    return L


def lib() -> ctypes.CDLL:
    """The loaded library, after checking that a GPU is present."""
    global _lib
    if _lib is None:
        L = load_library()
        n = ctypes.c_int(0)
        rc = L.edm_device_count(ctypes.byref(n))
        if rc != 0 or n.value < 1:
            raise HipError("no HIP device visible: the ED-MRDT GPU path needs an MI355X "
                           "(there is no CPU fallback)")
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().edm_last_error().decode(errors="replace")
        raise HipError(f"{what or 'libedm_hip'} failed ({rc}): {msg}")


def ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)
