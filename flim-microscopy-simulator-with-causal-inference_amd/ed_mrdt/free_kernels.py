"""The reference's free dynamics kernels on the GPU, with its signatures.

``ed_mrdt.dynamics`` re-exports every name of the reference's
``dynamics.__all__`` (dynamics.py:38-60).  Each function takes the
reference's arguments, runs the matching ``libedm_hip.so`` entry point
(``edm_cell_pairs``, ``edm_bd_step``, ``edm_pair_forces``, ``edm_binding``,
...) on device copies of the arrays and writes the arrays the reference
mutates back in place; the ones that return a count return it.

Arrays may be numpy arrays (copied to the device and back per call) or
``DeviceArray`` objects, which live in device memory between calls and are
used in place: a caller that keeps its state on the device calls the
kernels with no host transfer.  Device memory comes from the library itself
(``edm_device_alloc``), so these functions need nothing but
``libedm_hip.so``.  There is no CPU path: without the library or a GPU
every call raises ``HipError``.
"""

from __future__ import annotations

import ctypes
import math
from typing import Optional, Tuple

import numpy as np

from ._hip import check, lib

__all__ = [
    "CellList",
    "bd_step",
    "apply_binding_reactions",
    "apply_dissociation",
    "compute_forces_wca",
    "apply_catalytic_reaction",
    "apply_catalytic_with_feedback",
    "apply_membrane_recruitment",
    "apply_membrane_recruitment_with_feedback",
    "apply_membrane_recruitment_with_dual_feedback",
    "apply_membrane_detachment",
    "apply_membrane_detachment_with_feedback",
    "apply_species_conversion",
    "apply_species_conversion_with_feedback",
    "compute_density_field",
    "compute_forces_harmonic",
    "apply_diffusion_field_regions_smooth",
    "apply_catalytic_with_feedback_and_adaptation",
    "update_adaptation_field",
    "apply_membrane_recruitment_local",
    "DeviceArray",
]


class DeviceArray:
    """A C-contiguous array in device memory (float64, int32 or uint8; a
    bool array is held as uint8).  ``DeviceArray(host_array)`` copies it to
    the device; ``DeviceArray.empty(shape, dtype)`` allocates; ``numpy()``
    copies it back; ``upload(a)`` overwrites it from the host."""

    def __init__(self, host=None, *, shape=None, dtype=None):
        if host is not None:
            arr = np.asarray(host)
            if arr.dtype == np.bool_:
                arr = arr.astype(np.uint8)
            shape, dtype = arr.shape, arr.dtype
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().edm_device_alloc(self.nbytes, ctypes.byref(p)), "edm_device_alloc")
        self._p = p
        if host is not None:
            self.upload(arr)

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        return cls(shape=shape if isinstance(shape, tuple) else (int(shape),), dtype=dtype)

    @property
    def ptr(self):
        return self._p

    def __len__(self):
        return self.shape[0]

    def upload(self, a) -> None:
        arr = np.ascontiguousarray(a, dtype=self.dtype)
        if arr.nbytes != self.nbytes:
            raise ValueError("upload of a differently sized array")
        check(lib().edm_copy_to_device(self._p, arr.ctypes.data, self.nbytes),
              "edm_copy_to_device")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().edm_copy_to_host(out.ctypes.data, self._p, self.nbytes), "edm_copy_to_host")
        return out

    def __del__(self):
        try:
            if getattr(self, "_p", None) and self._p.value:
                lib().edm_device_free(self._p)
                self._p = ctypes.c_void_p()
        except Exception:
            pass


class _Arg:
    """One array argument on the device: a DeviceArray is used as it is; a
    numpy array is copied to a temporary DeviceArray and, when ``inout``,
    copied back into the caller's array after the call."""

    def __init__(self, a, dtype, inout: bool = False):
        dtype = np.dtype(dtype)
        self.src = a
        self.inout = inout
        if isinstance(a, DeviceArray):
            if a.dtype != dtype:
                raise TypeError(f"expected a {dtype} DeviceArray, got {a.dtype}")
            self.dev = a
            self.host = False
        else:
            self.dev = DeviceArray(np.ascontiguousarray(a, dtype=dtype))
            self.host = True

    @property
    def ptr(self):
        return self.dev.ptr

    def back(self):
        if self.inout and self.host:
            out = self.dev.numpy()
            self.src[...] = out.reshape(np.shape(self.src)).astype(np.asarray(self.src).dtype,
                                                                   copy=False)


def _opt(a, dtype):
    return None if a is None else _Arg(a, dtype)


def _p(arg: Optional[_Arg]):
    return None if arg is None else arg.ptr


def _stream():
    return None  # the null stream; the copies back are synchronous


def _finish(*args: _Arg):
    for a in args:
        a.back()


# ----------------------------------------------------------------------
# Cell list and pair list (dynamics.py:71-278)
# ----------------------------------------------------------------------

class CellList:
    """dynamics.py:71-175 CellList with the pair search on the GPU.

    ``nx``, ``ny``, ``cell_sx``, ``cell_sy`` as the reference derives them
    (cells at least the cutoff, at least 3 per side).  ``get_neighbor_pairs``
    returns the reference's pair list: i ascending, neighbour cells in
    (dcx, dcy) order, each cell's members in descending index, capped at
    min(30 N, 500000) pairs (``max_pairs`` overrides the cap).  ``head`` and
    ``next_particle`` are the reference's linked-list arrays of the last
    ``build``."""

    def __init__(self, Lx: float, Ly: float, cutoff: float,
                 periodic: Tuple[bool, bool] = (True, True)):
        if cutoff <= 0:
            raise ValueError(f"cutoff must be positive, got {cutoff}")
        self.Lx = Lx
        self.Ly = Ly
        self.cutoff = cutoff
        self.periodic_x, self.periodic_y = periodic
        self.nx = max(3, int(math.floor(Lx / cutoff)))
        self.ny = max(3, int(math.floor(Ly / cutoff)))
        self.cell_sx = Lx / self.nx
        self.cell_sy = Ly / self.ny
        self.head = np.full(self.nx * self.ny, -1, dtype=np.int32)
        self.next_particle = np.empty(0, dtype=np.int32)
        self.n_found = 0

    def _cells(self, positions, N):
        pos = np.asarray(positions, dtype=np.float64)[:N]
        cx = np.clip(np.trunc(pos[:, 0] / self.cell_sx), -1, self.nx).astype(np.int64)
        cy = np.clip(np.trunc(pos[:, 1] / self.cell_sy), -1, self.ny).astype(np.int64)
        return np.clip(cx, 0, self.nx - 1) * self.ny + np.clip(cy, 0, self.ny - 1)

    def build(self, positions, N: int) -> None:
        """The reference's head / next chains (dynamics.py:177-196): each
        cell's chain runs from its highest index down."""
        if len(self.next_particle) < N:
            self.next_particle = np.empty(N, dtype=np.int32)
        self.head[:] = -1
        if N == 0:
            return
        if isinstance(positions, DeviceArray):
            positions = positions.numpy()
        c = self._cells(positions, N)
        order = np.lexsort((np.arange(N), c))  # by cell, then index ascending
        cs = c[order]
        nxt = np.full(N, -1, dtype=np.int32)
        same = cs[1:] == cs[:-1]
        nxt[order[1:][same]] = order[:-1][same]
        self.next_particle[:N] = nxt
        last = np.r_[cs[1:] != cs[:-1], True]
        self.head[cs[last]] = order[last]

    def get_neighbors_brute(self, positions, N: int, cutoff: float) -> np.ndarray:
        """dynamics.py:140-151 / 281-320: the pairs (i < j, d < cutoff) in
        (i, j) ascending order, at most 20 N of them, by a search over the
        whole box on the GPU (one cell per non-periodic side, three, all
        neighbours of each other, per periodic side)."""
        nx = 3 if self.periodic_x else 1
        ny = 3 if self.periodic_y else 1
        cap = max(1, N * (N - 1) // 2)
        pairs = self._pairs(positions, N, nx, ny, self.Lx / nx, self.Ly / ny, cutoff, cap)
        order = np.lexsort((pairs[:, 1], pairs[:, 0]))
        return np.ascontiguousarray(pairs[order][: N * 20])

    def get_neighbor_pairs(self, positions, N: int,
                           max_pairs: Optional[int] = None) -> np.ndarray:
        cap = min(N * 30, 500000) if max_pairs is None else int(max_pairs)
        return self._pairs(positions, N, self.nx, self.ny, self.cell_sx, self.cell_sy,
                           self.cutoff, cap)

    def _pairs(self, positions, N, nx, ny, sx, sy, cutoff, cap):
        pos = _Arg(positions, np.float64)
        out = DeviceArray.empty((max(cap, 1), 2), np.int32)
        n = ctypes.c_int64(0)
        found = ctypes.c_int64(0)
        check(lib().edm_cell_pairs(pos.ptr, int(N), int(nx), int(ny), float(sx), float(sy),
                                   float(self.Lx), float(self.Ly), int(bool(self.periodic_x)),
                                   int(bool(self.periodic_y)), float(cutoff), int(cap), out.ptr,
                                   ctypes.byref(n), ctypes.byref(found), _stream()),
              "edm_cell_pairs")
        self.n_found = int(found.value)
        k = int(n.value)
        res = np.empty((k, 2), dtype=np.int32)
        if k:
            check(lib().edm_copy_to_host(res.ctypes.data, out.ptr, res.nbytes),
                  "edm_copy_to_host")
        return res


# ----------------------------------------------------------------------
# Brownian step, D(x) and forces (dynamics.py:323-519, 666-816)
# ----------------------------------------------------------------------

def bd_step(positions, D_per_particle, forces, grad_D, dt, Lx, Ly, periodic_x, periodic_y,
            N, noise, is_alive, max_displacement) -> None:
    """dynamics.py:323-406: Ermak-McCammon step of the alive particles,
    displacement cap on |dr|, periodic (Python modulo) or reflective walls;
    positions updated in place."""
    pos = _Arg(positions, np.float64, inout=True)
    a = [_Arg(D_per_particle, np.float64), _Arg(forces, np.float64), _Arg(grad_D, np.float64),
         _Arg(noise, np.float64), _Arg(is_alive, np.uint8)]
    check(lib().edm_bd_step(pos.ptr, a[0].ptr, a[1].ptr, a[2].ptr, float(dt), float(Lx),
                            float(Ly), int(bool(periodic_x)), int(bool(periodic_y)), int(N),
                            a[3].ptr, a[4].ptr, float(max_displacement), _stream()),
          "edm_bd_step")
    _finish(pos)


def apply_diffusion_field_regions_smooth(D_per_particle, grad_D, D_base, species_id, positions,
                                         N, region_cx, region_cy, region_r, region_Dfactor,
                                         n_regions, smoothing_width) -> None:
    """dynamics.py:441-519: D and grad D from the first circular region
    whose sigmoid exceeds 0.01; D and grad_D written in place."""
    Dp = _Arg(D_per_particle, np.float64, inout=True)
    gD = _Arg(grad_D, np.float64, inout=True)
    a = [_Arg(D_base, np.float64), _Arg(species_id, np.int32), _Arg(positions, np.float64)]
    nr = int(n_regions)
    r = [_Arg(np.asarray(v, dtype=np.float64)[:max(nr, 1)] if nr else np.zeros(1), np.float64)
         for v in (region_cx, region_cy, region_r, region_Dfactor)]
    check(lib().edm_diffusion_field_regions(Dp.ptr, gD.ptr, a[0].ptr, a[1].ptr, a[2].ptr, int(N),
                                            r[0].ptr, r[1].ptr, r[2].ptr, r[3].ptr, nr,
                                            float(smoothing_width), _stream()),
          "edm_diffusion_field_regions")
    _finish(Dp, gD)


def _pair_forces(kind, positions, forces, species_id, is_alive, pairs, n_pairs, sigma_matrix,
                 param_matrix, Lx, Ly, periodic_x, periodic_y):
    f = _Arg(forces, np.float64, inout=True)
    sig = np.ascontiguousarray(sigma_matrix, dtype=np.float64)
    ns = int(sig.shape[0])
    pos, sp, al = _Arg(positions, np.float64), _Arg(species_id, np.int32), _Arg(is_alive, np.uint8)
    m = int(n_pairs)
    pr = _Arg(np.asarray(pairs)[:m] if isinstance(pairs, np.ndarray) else pairs, np.int32)
    s, q = _Arg(sig, np.float64), _Arg(param_matrix, np.float64)
    n = int(species_id.shape[0])
    check(lib().edm_pair_forces(int(kind), pos.ptr, f.ptr, sp.ptr, al.ptr, n, pr.ptr, m, s.ptr,
                                q.ptr, ns, float(Lx), float(Ly), int(bool(periodic_x)),
                                int(bool(periodic_y)), _stream()), "edm_pair_forces")
    _finish(f)


def compute_forces_wca(positions, forces, species_id, is_alive, pairs, n_pairs, sigma_matrix,
                       epsilon_matrix, Lx, Ly, periodic_x, periodic_y) -> None:
    """dynamics.py:666-739: WCA pair forces (capped at 2400 eps / sigma^2)
    added to forces in place, Newton's third law, pair order."""
    _pair_forces(0, positions, forces, species_id, is_alive, pairs, n_pairs, sigma_matrix,
                 epsilon_matrix, Lx, Ly, periodic_x, periodic_y)


def compute_forces_harmonic(positions, forces, species_id, is_alive, pairs, n_pairs,
                            sigma_matrix, k_matrix, Lx, Ly, periodic_x, periodic_y) -> None:
    """dynamics.py:753-816: harmonic repulsion k (sigma - r) for r < sigma."""
    _pair_forces(1, positions, forces, species_id, is_alive, pairs, n_pairs, sigma_matrix,
                 k_matrix, Lx, Ly, periodic_x, periodic_y)


# ----------------------------------------------------------------------
# Reactions (dynamics.py:529-1657)
# ----------------------------------------------------------------------

def _pairs_arg(pairs, n_pairs):
    m = int(n_pairs)
    if isinstance(pairs, np.ndarray):
        return _Arg(np.asarray(pairs)[:m] if m else np.zeros((1, 2), np.int32), np.int32), m
    return _Arg(pairs, np.int32), m


def _count(fn, name, *args) -> int:
    c = ctypes.c_int64(0)
    check(fn(*args, ctypes.byref(c), _stream()), name)
    return int(c.value)


def apply_binding_reactions(positions, species_id, bond_partner, is_alive, N, reactant_a,
                            reactant_b, product_id, contact_radius_um, prob_bind, pairs, n_pairs,
                            rng_uniform, Lx, Ly, periodic_x, periodic_y) -> int:
    """dynamics.py:529-606: A + B -> A-B over the pair list in pair order
    (both alive and unbound, within contact, u[p] < P): A becomes the
    product, B is hidden, both record the bond.  Returns the bindings."""
    pos = _Arg(positions, np.float64)
    sp = _Arg(species_id, np.int32, inout=True)
    bd = _Arg(bond_partner, np.int32, inout=True)
    al = _Arg(is_alive, np.uint8, inout=True)
    pr, m = _pairs_arg(pairs, n_pairs)
    u = _Arg(rng_uniform if m else np.zeros(1), np.float64)
    n = _count(lib().edm_binding, "edm_binding", pos.ptr, sp.ptr, bd.ptr, al.ptr, int(N),
               int(reactant_a), int(reactant_b), int(product_id), float(contact_radius_um),
               float(prob_bind), pr.ptr, m, u.ptr, float(Lx), float(Ly), int(bool(periodic_x)),
               int(bool(periodic_y)))
    _finish(sp, bd, al)
    return n


def apply_dissociation(species_id, bond_partner, is_alive, positions, N, product_id, reactant_a,
                       reactant_b, prob_dissoc, rng_uniform) -> int:
    """dynamics.py:610-654: a complex splits with u[i] < P_off; B reappears
    at A's position."""
    sp = _Arg(species_id, np.int32, inout=True)
    bd = _Arg(bond_partner, np.int32, inout=True)
    al = _Arg(is_alive, np.uint8, inout=True)
    pos = _Arg(positions, np.float64, inout=True)
    u = _Arg(rng_uniform, np.float64)
    n = _count(lib().edm_dissociation, "edm_dissociation", sp.ptr, bd.ptr, al.ptr, pos.ptr,
               int(N), int(product_id), int(reactant_a), int(reactant_b), float(prob_dissoc),
               u.ptr)
    _finish(sp, bd, al, pos)
    return n


def _fb_args(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill, is_inhibition,
             basal_fraction):
    d = _opt(density_field, np.float64)
    return d, [_p(d), int(nx_grid), int(ny_grid), float(cell_sx), float(cell_sy), float(K_half),
               float(n_hill), int(bool(is_inhibition)), float(basal_fraction)]


_NO_FB = [None, 0, 0, 1.0, 1.0, 0.0, 1.0, 0, 0.0]


def _catalysis(positions, species_id, is_alive, N, enzyme_id, substrate_id, product_id,
               contact_radius_um, prob, pairs, n_pairs, rng_uniform, Lx, Ly, periodic_x,
               periodic_y, fb=None, adapt=None, K_adapt=0.0, n_adapt=1.0) -> int:
    pos = _Arg(positions, np.float64)
    sp = _Arg(species_id, np.int32, inout=True)
    al = _Arg(is_alive, np.uint8)
    pr, m = _pairs_arg(pairs, n_pairs)
    u = _Arg(rng_uniform if m else np.zeros(1), np.float64)
    keep, fbv = (None, _NO_FB) if fb is None else _fb_args(*fb)
    w = _opt(adapt, np.float64)
    anx, any_ = (0, 0) if adapt is None else tuple(int(v) for v in adapt.shape[:2])
    n = _count(lib().edm_catalysis, "edm_catalysis", pos.ptr, sp.ptr, al.ptr, int(N),
               int(enzyme_id), int(substrate_id), int(product_id), float(contact_radius_um),
               float(prob), pr.ptr, m, u.ptr, float(Lx), float(Ly), int(bool(periodic_x)),
               int(bool(periodic_y)), *fbv, _p(w), anx, any_, float(K_adapt), float(n_adapt))
    del keep
    _finish(sp)
    return n


def apply_catalytic_reaction(positions, species_id, is_alive, N, enzyme_id, substrate_id,
                             product_id, contact_radius_um, prob_catalysis, pairs, n_pairs,
                             rng_uniform, Lx, Ly, periodic_x, periodic_y) -> int:
    """dynamics.py:826-884: E + S -> E + P over the pair list; a substrate
    converts once if any of its pairs accepts."""
    return _catalysis(positions, species_id, is_alive, N, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis, pairs, n_pairs, rng_uniform, Lx, Ly,
                      periodic_x, periodic_y)


def apply_catalytic_with_feedback(positions, species_id, is_alive, N, enzyme_id, substrate_id,
                                  product_id, contact_radius_um, prob_catalysis_base, pairs,
                                  n_pairs, rng_uniform, Lx, Ly, periodic_x, periodic_y,
                                  density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                  n_hill, is_inhibition, basal_fraction=0.0) -> int:
    """dynamics.py:888-983: catalysis x Hill(rho at the substrate's cell)."""
    return _catalysis(positions, species_id, is_alive, N, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis_base, pairs, n_pairs, rng_uniform, Lx,
                      Ly, periodic_x, periodic_y,
                      fb=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                          is_inhibition, basal_fraction))


def apply_catalytic_with_feedback_and_adaptation(positions, species_id, is_alive, N, enzyme_id,
                                                 substrate_id, product_id, contact_radius_um,
                                                 prob_catalysis_base, pairs, n_pairs,
                                                 rng_uniform, Lx, Ly, periodic_x, periodic_y,
                                                 density_field, nx_grid, ny_grid, cell_sx,
                                                 cell_sy, K_half, n_hill, is_inhibition,
                                                 basal_fraction, adaptation_field, K_adapt,
                                                 n_adapt) -> int:
    """dynamics.py:987-1090: catalysis x Hill x 1 / (1 + (w / K_adapt)^n),
    w read at the feedback grid's indices (dynamics.py:1078)."""
    return _catalysis(positions, species_id, is_alive, N, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis_base, pairs, n_pairs, rng_uniform, Lx,
                      Ly, periodic_x, periodic_y,
                      fb=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                          is_inhibition, basal_fraction),
                      adapt=adaptation_field, K_adapt=K_adapt, n_adapt=n_adapt)


def update_adaptation_field(adaptation_field, density_field, nx, ny, dt, tau_w) -> None:
    """dynamics.py:1094-1112: w += dt / tau_w (rho - w)."""
    w = _Arg(adaptation_field, np.float64, inout=True)
    rho = _Arg(density_field, np.float64)
    check(lib().edm_adaptation_update(w.ptr, rho.ptr, int(nx), int(ny), float(dt), float(tau_w),
                                      _stream()), "edm_adaptation_update")
    _finish(w)


def _recruit(species_id, is_alive, positions, N, target_species_id, prob, Lx, Ly, rng_accept,
             rng_positions, max_recruits, fb1=None, fb2=None, cdf=None, rng_location=None,
             rng_offsets=None, recruit_sigma=0.0) -> int:
    sp = _Arg(species_id, np.int32)
    al = _Arg(is_alive, np.uint8, inout=True)
    pos = _Arg(positions, np.float64, inout=True)
    ua = _Arg(rng_accept, np.float64)
    up = _opt(rng_positions, np.float64)
    k1, f1 = (None, _NO_FB) if fb1 is None else _fb_args(*fb1)
    k2, f2 = (None, _NO_FB) if fb2 is None else _fb_args(*fb2)
    c, ul, uo = _opt(cdf, np.float64), _opt(rng_location, np.float64), _opt(rng_offsets, np.float64)
    n = _count(lib().edm_recruitment, "edm_recruitment", sp.ptr, al.ptr, pos.ptr, int(N),
               int(target_species_id), float(prob), float(Lx), float(Ly), ua.ptr, _p(up),
               int(max_recruits), *f1, *f2, _p(c), _p(ul), _p(uo), float(recruit_sigma))
    del k1, k2
    _finish(al, pos)
    return n


def apply_membrane_recruitment(species_id, is_alive, positions, N, target_species_id,
                               prob_recruit, Lx, Ly, rng_accept, rng_positions,
                               max_recruits) -> int:
    """dynamics.py:1230-1264: dormant particles come alive at u * L with
    u_acc < P_on, the first max_recruits in index order."""
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit, Lx, Ly,
                    rng_accept, rng_positions, max_recruits)


def apply_membrane_recruitment_with_feedback(species_id, is_alive, positions, N,
                                             target_species_id, prob_recruit_base, Lx, Ly,
                                             rng_accept, rng_positions, max_recruits,
                                             density_field, nx_grid, ny_grid, cell_sx, cell_sy,
                                             K_half, n_hill, is_inhibition,
                                             basal_fraction=0.0) -> int:
    """dynamics.py:1268-1343: recruitment x Hill at the candidate site."""
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit_base, Lx,
                    Ly, rng_accept, rng_positions, max_recruits,
                    fb1=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                         is_inhibition, basal_fraction))


def apply_membrane_recruitment_with_dual_feedback(
        species_id, is_alive, positions, N, target_species_id, prob_recruit_base, Lx, Ly,
        rng_accept, rng_positions, max_recruits, density_field_1, nx_grid_1, ny_grid_1,
        cell_sx_1, cell_sy_1, K_half_1, n_hill_1, is_inhibition_1, basal_fraction_1,
        density_field_2, nx_grid_2, ny_grid_2, cell_sx_2, cell_sy_2, K_half_2, n_hill_2,
        is_inhibition_2, basal_fraction_2) -> int:
    """dynamics.py:1347-1458: recruitment x Hill_1 x Hill_2 at the site."""
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit_base, Lx,
                    Ly, rng_accept, rng_positions, max_recruits,
                    fb1=(density_field_1, nx_grid_1, ny_grid_1, cell_sx_1, cell_sy_1, K_half_1,
                         n_hill_1, is_inhibition_1, basal_fraction_1),
                    fb2=(density_field_2, nx_grid_2, ny_grid_2, cell_sx_2, cell_sy_2, K_half_2,
                         n_hill_2, is_inhibition_2, basal_fraction_2))


def apply_membrane_recruitment_local(species_id, is_alive, positions, N, target_species_id,
                                     prob_recruit_base, Lx, Ly, rng_accept, rng_offsets,
                                     max_recruits, density_field, nx_grid, ny_grid, cell_sx,
                                     cell_sy, K_half, n_hill, is_inhibition, basal_fraction, cdf,
                                     rng_location, recruit_sigma) -> int:
    """dynamics.py:1116-1217: the site is the centre of a grid cell drawn
    from the CDF plus recruit_sigma times a normal pair, wrapped; Hill at
    the site."""
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit_base, Lx,
                    Ly, rng_accept, None, max_recruits,
                    fb1=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                         is_inhibition, basal_fraction),
                    cdf=cdf, rng_location=rng_location, rng_offsets=rng_offsets,
                    recruit_sigma=recruit_sigma)


def _detach(species_id, is_alive, positions, N, target, prob, rng_uniform, fb=None) -> int:
    sp = _Arg(species_id, np.int32)
    al = _Arg(is_alive, np.uint8, inout=True)
    pos = _opt(positions, np.float64)
    u = _Arg(rng_uniform, np.float64)
    keep, fbv = (None, _NO_FB) if fb is None else _fb_args(*fb)
    n = _count(lib().edm_detachment, "edm_detachment", sp.ptr, al.ptr, _p(pos), int(N),
               int(target), float(prob), u.ptr, *fbv)
    del keep
    _finish(al)
    return n


def apply_membrane_detachment(species_id, is_alive, N, target_species_id, prob_detach,
                              rng_uniform) -> int:
    """dynamics.py:1462-1486: alive particles go dormant with u < P_off."""
    return _detach(species_id, is_alive, None, N, target_species_id, prob_detach, rng_uniform)


def apply_membrane_detachment_with_feedback(species_id, is_alive, positions, N,
                                            target_species_id, prob_detach_base, rng_uniform,
                                            density_field, nx_grid, ny_grid, cell_sx, cell_sy,
                                            K_half, n_hill, is_inhibition,
                                            basal_fraction=0.0) -> int:
    """dynamics.py:1490-1555: detachment x Hill at the particle's cell."""
    return _detach(species_id, is_alive, positions, N, target_species_id, prob_detach_base,
                   rng_uniform, fb=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                    n_hill, is_inhibition, basal_fraction))


def _convert(species_id, is_alive, positions, N, from_id, to_id, prob, rng_uniform,
             fb=None) -> int:
    sp = _Arg(species_id, np.int32, inout=True)
    al = _Arg(is_alive, np.uint8)
    pos = _opt(positions, np.float64)
    u = _Arg(rng_uniform, np.float64)
    keep, fbv = (None, _NO_FB) if fb is None else _fb_args(*fb)
    n = _count(lib().edm_conversion, "edm_conversion", sp.ptr, al.ptr, _p(pos), int(N),
               int(from_id), int(to_id), float(prob), u.ptr, *fbv)
    del keep
    _finish(sp)
    return n


def apply_species_conversion(species_id, is_alive, N, from_id, to_id, prob_convert,
                             rng_uniform) -> int:
    """dynamics.py:1566-1591: A -> B with u < P."""
    return _convert(species_id, is_alive, None, N, from_id, to_id, prob_convert, rng_uniform)


def apply_species_conversion_with_feedback(species_id, is_alive, positions, N, from_id, to_id,
                                           prob_convert_base, rng_uniform, density_field,
                                           nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                                           is_inhibition, basal_fraction=0.0) -> int:
    """dynamics.py:1595-1657: conversion x Hill at the particle's cell."""
    return _convert(species_id, is_alive, positions, N, from_id, to_id, prob_convert_base,
                    rng_uniform, fb=(density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                     n_hill, is_inhibition, basal_fraction))


def compute_density_field(positions, species_id, is_alive, N, target_species_id, nx_grid,
                          ny_grid, cell_sx, cell_sy, density_out) -> None:
    """dynamics.py:1668-1703: density_out[gx, gy] += 1 / cell area per alive
    particle of the species (the caller zeroes it, as in the reference)."""
    out = _Arg(density_out, np.float64, inout=True)
    pos, sp, al = _Arg(positions, np.float64), _Arg(species_id, np.int32), _Arg(is_alive, np.uint8)
    check(lib().edm_density_field(pos.ptr, sp.ptr, al.ptr, int(N), int(target_species_id),
                                  int(nx_grid), int(ny_grid), float(cell_sx), float(cell_sy),
                                  out.ptr, _stream()), "edm_density_field")
    _finish(out)
