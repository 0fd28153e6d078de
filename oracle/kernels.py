"""ORACLE kernels with the reference's Python signatures — test infrastructure only.

numpy front-ends for the C restatement in oracle/edm_oracle.c, named and
parameterised like the reference's numba kernels so that restated
known-answer tests read like the reference's own.  Arrays are modified in
place exactly where the reference modifies them (boolean arrays are passed
to C as bytes and copied back).  Each function cites the reference kernel
it stands for.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from reference_model import D, I, LL, P, build_recruit_cdf, lib

__all__ = [
    "_min_image_dist_sq", "py_mod", "CellList", "bd_step",
    "apply_diffusion_field_regions_smooth", "apply_diffusion_field_sinusoidal",
    "compute_forces_wca", "compute_forces_harmonic", "hill_factor",
    "apply_binding_reactions", "apply_dissociation", "apply_catalytic_reaction",
    "apply_catalytic_with_feedback", "apply_catalytic_with_feedback_and_adaptation",
    "update_adaptation_field", "apply_membrane_recruitment",
    "apply_membrane_recruitment_with_feedback",
    "apply_membrane_recruitment_with_dual_feedback", "apply_membrane_detachment",
    "apply_membrane_detachment_with_feedback", "apply_species_conversion",
    "apply_species_conversion_with_feedback", "compute_density_field", "_poisson_draw",
    "_compute_fret_rates_numba", "_compute_fret_rates_proximity_numba",
    "assign_emitter_centric", "accumulate_tcspc",
]


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


class _Bytes:
    """A bool/uint8 array seen by C as bytes; copied back on exit."""

    def __init__(self, a: np.ndarray):
        self.a = a
        self.b = np.ascontiguousarray(a, dtype=np.uint8).copy()

    def back(self):
        self.a[...] = self.b.astype(self.a.dtype)


class _InPlace:
    """A float64/int32 array passed by pointer and written back if it had
    to be converted (arrays of the right type are passed directly)."""

    def __init__(self, a: np.ndarray, dtype):
        self.a = a
        self.direct = a.dtype == dtype and a.flags["C_CONTIGUOUS"]
        self.b = a if self.direct else np.ascontiguousarray(a, dtype=dtype).copy()

    def back(self):
        if not self.direct:
            self.a[...] = self.b


# ----------------------------------------------------------------------
# Geometry and neighbour search (dynamics.py:71-320)
# ----------------------------------------------------------------------

def _min_image_dist_sq(x1, y1, x2, y2, Lx, Ly, periodic_x, periodic_y):
    """dynamics.py:200-217: squared minimum-image distance, dx, dy."""
    import ctypes
    dx = ctypes.c_double()
    dy = ctypes.c_double()
    f = lib().orc_min_image_dsq
    f.argtypes = [D, D, D, D, D, D, I, I, ctypes.POINTER(D), ctypes.POINTER(D)]
    dsq = f(x1, y1, x2, y2, Lx, Ly, int(bool(periodic_x)), int(bool(periodic_y)),
            ctypes.byref(dx), ctypes.byref(dy))
    return dsq, dx.value, dy.value


def py_mod(x: float, L: float) -> float:
    """Python's float modulo (result has the sign of L)."""
    return lib().orc_py_mod(x, L)


class CellList:
    """dynamics.py:71-175: linked-list cell list; cells >= cutoff, at least 3."""

    def __init__(self, Lx: float, Ly: float, cutoff: float,
                 periodic: Tuple[bool, bool] = (True, True)):
        if cutoff <= 0:
            raise ValueError(f"cutoff must be positive, got {cutoff}")
        self.Lx, self.Ly, self.cutoff = Lx, Ly, cutoff
        self.periodic_x, self.periodic_y = periodic
        self.nx = max(3, int(math.floor(Lx / cutoff)))
        self.ny = max(3, int(math.floor(Ly / cutoff)))
        self.cell_sx = Lx / self.nx
        self.cell_sy = Ly / self.ny
        self.head = np.full(self.nx * self.ny, -1, dtype=np.int32)
        self.next_particle = np.empty(0, dtype=np.int32)

    def build(self, positions: np.ndarray, N: int) -> None:
        if len(self.next_particle) < N:
            self.next_particle = np.empty(N, dtype=np.int32)
        pos = _f64(positions)
        lib().orc_build_cell_list(P(pos), I(N), P(self.head), P(self.next_particle),
                                  I(self.nx), I(self.ny), D(self.cell_sx), D(self.cell_sy))

    def get_neighbor_pairs(self, positions: np.ndarray, N: int,
                           max_pairs: Optional[int] = None) -> np.ndarray:
        pos = _f64(positions)
        cap = max_pairs if max_pairs is not None else max(1, N * (N - 1) // 2)
        out = np.empty((cap, 2), dtype=np.int32)
        found = LL(0)
        import ctypes
        n = lib().orc_cell_list_pairs(
            P(pos), I(N), P(self.head), P(self.next_particle), I(self.nx), I(self.ny),
            D(self.cell_sx), D(self.cell_sy), D(self.Lx), D(self.Ly),
            I(int(self.periodic_x)), I(int(self.periodic_y)), D(self.cutoff), LL(cap),
            P(out), ctypes.byref(found))
        return out[:n].copy()

    def get_neighbors_brute(self, positions: np.ndarray, N: int, cutoff: float) -> np.ndarray:
        pos = _f64(positions)
        cap = max(1, N * (N - 1) // 2)
        out = np.empty((cap, 2), dtype=np.int32)
        f = lib().orc_brute_force_pairs
        f.restype = I
        n = f(P(pos), I(N), D(cutoff), D(self.Lx), D(self.Ly), I(int(self.periodic_x)),
              I(int(self.periodic_y)), LL(cap), P(out))
        return out[:n].copy()


# ----------------------------------------------------------------------
# Diffusion field, forces, integrator (dynamics.py:323-816)
# ----------------------------------------------------------------------

def bd_step(positions, D_arr, forces, grad_D, dt, Lx, Ly, periodic_x, periodic_y, N,
            noise, is_alive, max_displacement) -> None:
    """dynamics.py:324-406: Ermak-McCammon step in place."""
    pos = _InPlace(positions, np.float64)
    alive = np.ascontiguousarray(is_alive, dtype=np.uint8)
    lib().orc_bd_step(P(pos.b), P(_f64(D_arr)), P(_f64(forces)), P(_f64(grad_D)), D(dt),
                      D(Lx), D(Ly), I(int(bool(periodic_x))), I(int(bool(periodic_y))), I(N),
                      P(_f64(noise)), P(alive), D(max_displacement), I(1))
    pos.back()


def apply_diffusion_field_regions_smooth(D_out, grad_D, D_base, species_id, positions, N,
                                         region_cx, region_cy, region_r, region_D_factor,
                                         n_regions, smoothing_width) -> None:
    """dynamics.py:441-519: D and grad(D) from smooth circular regions."""
    d = _InPlace(D_out, np.float64)
    g = _InPlace(grad_D, np.float64)
    lib().orc_diffusion_field_smooth(
        P(d.b), P(g.b), P(_f64(D_base)), P(_i32(species_id)), P(_f64(positions)), I(N),
        P(_f64(region_cx)), P(_f64(region_cy)), P(_f64(region_r)),
        P(_f64(region_D_factor)), I(n_regions), D(smoothing_width))
    d.back()
    g.back()


def apply_diffusion_field_sinusoidal(D_out, grad_D, D_base, species_id, positions, N,
                                     amplitude, kx, ky) -> None:
    """D = D_s (1 + a sin(kx x) sin(ky y)) with its analytic gradient (no
    counterpart in the reference; pinned by finite differences)."""
    d = _InPlace(D_out, np.float64)
    g = _InPlace(grad_D, np.float64)
    lib().orc_diffusion_field_sinusoidal(P(d.b), P(g.b), P(_f64(D_base)), P(_i32(species_id)),
                                         P(_f64(positions)), I(N), D(amplitude), D(kx), D(ky))
    d.back()
    g.back()


def _forces(fn, positions, forces, species_id, is_alive, pairs, n_pairs, sig_m, other_m,
            Lx, Ly, px, py):
    f = _InPlace(forces, np.float64)
    sig = _f64(sig_m)
    ns = sig.shape[0]
    fn(P(_f64(positions)), P(f.b), P(_i32(species_id)),
       P(np.ascontiguousarray(is_alive, dtype=np.uint8)), P(_i32(pairs)), I(n_pairs), P(sig),
       P(_f64(other_m)), I(ns), D(Lx), D(Ly), I(int(bool(px))), I(int(bool(py))))
    f.back()


def compute_forces_wca(positions, forces, species_id, is_alive, pairs, n_pairs,
                       sigma_matrix, epsilon_matrix, Lx, Ly, periodic_x, periodic_y) -> None:
    """dynamics.py:666-739: WCA pair forces (force cap 2400 eps/sigma^2), Newton 3."""
    _forces(lib().orc_forces_wca, positions, forces, species_id, is_alive, pairs, n_pairs,
            sigma_matrix, epsilon_matrix, Lx, Ly, periodic_x, periodic_y)


def compute_forces_harmonic(positions, forces, species_id, is_alive, pairs, n_pairs,
                            sigma_matrix, k_matrix, Lx, Ly, periodic_x, periodic_y) -> None:
    """dynamics.py:753-816: soft harmonic repulsion k (sigma - r) for r < sigma."""
    _forces(lib().orc_forces_harmonic, positions, forces, species_id, is_alive, pairs,
            n_pairs, sigma_matrix, k_matrix, Lx, Ly, periodic_x, periodic_y)


# ----------------------------------------------------------------------
# Reactions (dynamics.py:529-1703)
# ----------------------------------------------------------------------

def hill_factor(rho, K_half, n_hill, is_inhibition, basal_fraction=0.0) -> float:
    """dynamics.py:960-980 (shared by every feedback kernel)."""
    return lib().orc_hill(rho, K_half, n_hill, int(bool(is_inhibition)), basal_fraction)


def apply_binding_reactions(positions, species_id, bond_partner, is_alive, N, reactant_a,
                            reactant_b, product, contact_radius, prob_bind, pairs, n_pairs,
                            rng_uniform, Lx, Ly, periodic_x, periodic_y) -> int:
    """dynamics.py:529-606: greedy sequential binding over the pair list."""
    sp = _InPlace(species_id, np.int32)
    bd = _InPlace(bond_partner, np.int32)
    al = _Bytes(is_alive)
    f = lib().orc_binding
    f.restype = I
    n = f(P(_f64(positions)), P(sp.b), P(bd.b), P(al.b), I(reactant_a), I(reactant_b),
          I(product), D(contact_radius), D(prob_bind), P(_i32(pairs)), I(n_pairs),
          P(_f64(rng_uniform)), D(Lx), D(Ly), I(int(bool(periodic_x))),
          I(int(bool(periodic_y))))
    sp.back()
    bd.back()
    al.back()
    return n


def apply_dissociation(species_id, bond_partner, is_alive, positions, N, product,
                       reactant_a, reactant_b, prob_dissoc, rng_uniform) -> int:
    """dynamics.py:610-654: per-complex dissociation; partner reappears on A."""
    sp = _InPlace(species_id, np.int32)
    bd = _InPlace(bond_partner, np.int32)
    al = _Bytes(is_alive)
    pos = _InPlace(positions, np.float64)
    f = lib().orc_dissociation
    f.restype = I
    n = f(P(sp.b), P(bd.b), P(al.b), P(pos.b), I(N), I(product), I(reactant_a),
          I(reactant_b), D(prob_dissoc), P(_f64(rng_uniform)))
    sp.back()
    bd.back()
    al.back()
    pos.back()
    return n


def _catalytic(positions, species_id, is_alive, enzyme_id, substrate_id, product_id,
               contact_radius_um, prob, pairs, n_pairs, rng_uniform, Lx, Ly, px, py,
               density=None, nxg=1, nyg=1, sx=1.0, sy=1.0, K=0.0, n=1.0, inh=False,
               basal=0.0, adapt=None, K_adapt=0.0, n_adapt=1.0) -> int:
    sp = _InPlace(species_id, np.int32)
    f = lib().orc_catalytic
    f.restype = I
    dens = None if density is None else _f64(density)
    ad = None if adapt is None else _f64(adapt)
    anx, any_ = (1, 1) if adapt is None else np.shape(adapt)
    res = f(P(_f64(positions)), P(sp.b), P(np.ascontiguousarray(is_alive, dtype=np.uint8)),
            I(enzyme_id), I(substrate_id), I(product_id), D(contact_radius_um), D(prob),
            P(_i32(pairs)), I(n_pairs), P(_f64(rng_uniform)), D(Lx), D(Ly),
            I(int(bool(px))), I(int(bool(py))), P(dens), I(nxg), I(nyg), D(sx), D(sy), D(K),
            D(n), I(int(bool(inh))), D(basal), P(ad), I(anx), I(any_), D(K_adapt),
            D(n_adapt))
    sp.back()
    return res


def apply_catalytic_reaction(positions, species_id, is_alive, N, enzyme_id, substrate_id,
                             product_id, contact_radius_um, prob_catalysis, pairs, n_pairs,
                             rng_uniform, Lx, Ly, periodic_x, periodic_y) -> int:
    """dynamics.py:826-884: E + S -> E + P over the pair list."""
    return _catalytic(positions, species_id, is_alive, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis, pairs, n_pairs, rng_uniform, Lx, Ly,
                      periodic_x, periodic_y)


def apply_catalytic_with_feedback(positions, species_id, is_alive, N, enzyme_id,
                                  substrate_id, product_id, contact_radius_um,
                                  prob_catalysis_base, pairs, n_pairs, rng_uniform, Lx, Ly,
                                  periodic_x, periodic_y, density_field, nx_grid, ny_grid,
                                  cell_sx, cell_sy, K_half, n_hill, is_inhibition,
                                  basal_fraction=0.0) -> int:
    """dynamics.py:888-984: catalysis x Hill(rho at the substrate)."""
    return _catalytic(positions, species_id, is_alive, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis_base, pairs, n_pairs, rng_uniform, Lx,
                      Ly, periodic_x, periodic_y, density_field, nx_grid, ny_grid, cell_sx,
                      cell_sy, K_half, n_hill, is_inhibition, basal_fraction)


def apply_catalytic_with_feedback_and_adaptation(positions, species_id, is_alive, N,
                                                 enzyme_id, substrate_id, product_id,
                                                 contact_radius_um, prob_catalysis_base,
                                                 pairs, n_pairs, rng_uniform, Lx, Ly,
                                                 periodic_x, periodic_y, density_field,
                                                 nx_grid, ny_grid, cell_sx, cell_sy, K_half,
                                                 n_hill, is_inhibition, basal_fraction,
                                                 adaptation_field, K_adapt,
                                                 n_adapt) -> int:
    """dynamics.py:987-1090: catalysis x Hill x 1/(1 + (w/K_adapt)^n_adapt)."""
    return _catalytic(positions, species_id, is_alive, enzyme_id, substrate_id, product_id,
                      contact_radius_um, prob_catalysis_base, pairs, n_pairs, rng_uniform, Lx,
                      Ly, periodic_x, periodic_y, density_field, nx_grid, ny_grid, cell_sx,
                      cell_sy, K_half, n_hill, is_inhibition, basal_fraction,
                      adaptation_field, K_adapt, n_adapt)


def update_adaptation_field(w, rho, nx, ny, dt, tau) -> None:
    """dynamics.py:1094-1112: w += dt/tau (rho - w)."""
    ww = _InPlace(w, np.float64)
    lib().orc_adaptation_update(P(ww.b), P(_f64(rho)), I(nx), I(ny), D(dt), D(tau))
    ww.back()


_NOFB = (None, 1, 1, 1.0, 1.0, 0.0, 1.0, 0, 0.0)


def _recruit(species_id, is_alive, positions, N, target, prob, Lx, Ly, rng_accept,
             rng_positions, max_recruits, fb1=_NOFB, fb2=_NOFB) -> int:
    al = _Bytes(is_alive)
    pos = _InPlace(positions, np.float64)
    f = lib().orc_recruit
    f.restype = I

    def fbargs(fb):
        d, nx, ny, sx, sy, K, n, inh, b = fb
        return [P(None if d is None else _f64(d)), I(nx), I(ny), D(sx), D(sy), D(K), D(n),
                I(int(bool(inh))), D(b)]

    res = f(P(_i32(species_id)), P(al.b), P(pos.b), I(N), I(target), D(prob), D(Lx), D(Ly),
            P(_f64(rng_accept)), P(_f64(rng_positions)), I(max_recruits), *fbargs(fb1),
            *fbargs(fb2))
    al.back()
    pos.back()
    return res


def apply_membrane_recruitment(species_id, is_alive, positions, N, target_species_id,
                               prob_recruit, Lx, Ly, rng_accept, rng_positions,
                               max_recruits) -> int:
    """dynamics.py:1230-1264: dormant -> alive in index order up to the cap."""
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit, Lx,
                    Ly, rng_accept, rng_positions, max_recruits)


def apply_membrane_recruitment_with_feedback(species_id, is_alive, positions, N,
                                             target_species_id, prob_recruit_base, Lx, Ly,
                                             rng_accept, rng_positions, max_recruits,
                                             density_field, nx_grid, ny_grid, cell_sx,
                                             cell_sy, K_half, n_hill, is_inhibition,
                                             basal_fraction=0.0) -> int:
    """dynamics.py:1268-1343: recruitment x Hill(rho at the candidate site)."""
    fb = (density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill, is_inhibition,
          basal_fraction)
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit_base,
                    Lx, Ly, rng_accept, rng_positions, max_recruits, fb)


def apply_membrane_recruitment_with_dual_feedback(
        species_id, is_alive, positions, N, target_species_id, prob_recruit_base, Lx, Ly,
        rng_accept, rng_positions, max_recruits, density_field_1, nx_grid_1, ny_grid_1,
        cell_sx_1, cell_sy_1, K_half_1, n_hill_1, is_inhibition_1, basal_fraction_1,
        density_field_2, nx_grid_2, ny_grid_2, cell_sx_2, cell_sy_2, K_half_2, n_hill_2,
        is_inhibition_2, basal_fraction_2) -> int:
    """dynamics.py:1347-1458: recruitment x Hill_1 x Hill_2 at the site."""
    fb1 = (density_field_1, nx_grid_1, ny_grid_1, cell_sx_1, cell_sy_1, K_half_1, n_hill_1,
           is_inhibition_1, basal_fraction_1)
    fb2 = (density_field_2, nx_grid_2, ny_grid_2, cell_sx_2, cell_sy_2, K_half_2, n_hill_2,
           is_inhibition_2, basal_fraction_2)
    return _recruit(species_id, is_alive, positions, N, target_species_id, prob_recruit_base,
                    Lx, Ly, rng_accept, rng_positions, max_recruits, fb1, fb2)


def apply_membrane_recruitment_local(species_id, is_alive, positions, N, target_species_id,
                                     prob_recruit_base, Lx, Ly, rng_accept, rng_offsets,
                                     max_recruits, density_field, nx_grid, ny_grid, cell_sx,
                                     cell_sy, K_half, n_hill, is_inhibition, basal_fraction,
                                     cdf, rng_location, recruit_sigma) -> int:
    """dynamics.py:1116-1217: the candidate site is the centre of the cell
    the CDF bisection picks plus sigma * offsets, wrapped; Hill feedback is
    read at that site; index-order cap."""
    al = _Bytes(is_alive)
    pos = _InPlace(positions, np.float64)
    f = lib().orc_recruit_local
    f.restype = I
    res = f(P(_i32(species_id)), P(al.b), P(pos.b), I(N), I(target_species_id),
            D(prob_recruit_base), D(Lx), D(Ly), P(_f64(rng_accept)), P(_f64(rng_offsets)),
            I(max_recruits), P(_f64(density_field)), I(nx_grid), I(ny_grid), D(cell_sx),
            D(cell_sy), D(K_half), D(n_hill), I(int(bool(is_inhibition))), D(basal_fraction),
            P(_f64(cdf)), P(_f64(rng_location)), D(recruit_sigma))
    al.back()
    pos.back()
    return res


def _detach(species_id, is_alive, positions, N, target, prob, rng_uniform, fb=_NOFB) -> int:
    al = _Bytes(is_alive)
    d, nx, ny, sx, sy, K, n, inh, b = fb
    f = lib().orc_detach
    f.restype = I
    pos = np.zeros((N, 2)) if positions is None else _f64(positions)
    res = f(P(_i32(species_id)), P(al.b), P(pos), I(N), I(target), D(prob),
            P(_f64(rng_uniform)), P(None if d is None else _f64(d)), I(nx), I(ny), D(sx),
            D(sy), D(K), D(n), I(int(bool(inh))), D(b))
    al.back()
    return res


def apply_membrane_detachment(species_id, is_alive, N, target_species_id, prob_detach,
                              rng_uniform) -> int:
    """dynamics.py:1462-1486: alive -> dormant."""
    return _detach(species_id, is_alive, None, N, target_species_id, prob_detach,
                   rng_uniform)


def apply_membrane_detachment_with_feedback(species_id, is_alive, positions, N,
                                            target_species_id, prob_detach_base,
                                            rng_uniform, density_field, nx_grid, ny_grid,
                                            cell_sx, cell_sy, K_half, n_hill, is_inhibition,
                                            basal_fraction=0.0) -> int:
    """dynamics.py:1490-1555: detachment x Hill(rho at the particle)."""
    fb = (density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill, is_inhibition,
          basal_fraction)
    return _detach(species_id, is_alive, positions, N, target_species_id, prob_detach_base,
                   rng_uniform, fb)


def _convert(species_id, is_alive, positions, N, from_id, to_id, prob, rng_uniform,
             fb=_NOFB) -> int:
    sp = _InPlace(species_id, np.int32)
    d, nx, ny, sx, sy, K, n, inh, b = fb
    f = lib().orc_convert
    f.restype = I
    pos = np.zeros((N, 2)) if positions is None else _f64(positions)
    res = f(P(sp.b), P(np.ascontiguousarray(is_alive, dtype=np.uint8)), P(pos), I(N),
            I(from_id), I(to_id), D(prob), P(_f64(rng_uniform)),
            P(None if d is None else _f64(d)), I(nx), I(ny), D(sx), D(sy), D(K), D(n),
            I(int(bool(inh))), D(b))
    sp.back()
    return res


def apply_species_conversion(species_id, is_alive, N, from_id, to_id, prob_convert,
                             rng_uniform) -> int:
    """dynamics.py:1566-1591: A -> B."""
    return _convert(species_id, is_alive, None, N, from_id, to_id, prob_convert, rng_uniform)


def apply_species_conversion_with_feedback(species_id, is_alive, positions, N, from_id, to_id,
                                           prob_convert_base, rng_uniform, density_field,
                                           nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill,
                                           is_inhibition, basal_fraction=0.0) -> int:
    """dynamics.py:1595-1657: conversion x Hill(rho at the particle)."""
    fb = (density_field, nx_grid, ny_grid, cell_sx, cell_sy, K_half, n_hill, is_inhibition,
          basal_fraction)
    return _convert(species_id, is_alive, positions, N, from_id, to_id, prob_convert_base,
                    rng_uniform, fb)


def compute_density_field(positions, species_id, is_alive, N, target_species_id, nx_grid,
                          ny_grid, cell_sx, cell_sy, density_out) -> None:
    """dynamics.py:1668-1703: per-cell sequential += 1/area."""
    out = _InPlace(density_out, np.float64)
    lib().orc_density(P(_f64(positions)), P(_i32(species_id)),
                      P(np.ascontiguousarray(is_alive, dtype=np.uint8)), I(N),
                      I(target_species_id), I(nx_grid), I(ny_grid), D(cell_sx), D(cell_sy),
                      P(out.b))
    out.back()


# ----------------------------------------------------------------------
# Photophysics and imaging (photophysics.py:487-683, microscope.py:523-803)
# ----------------------------------------------------------------------

def _poisson_draw(mean: float, u_seed: float) -> int:
    """photophysics.py:487-521: LCG-seeded Knuth (< 30) or Box-Muller draw."""
    return lib().orc_poisson_draw(mean, u_seed)


def _compute_fret_rates_numba(k_fret, fret_acceptor_idx, has_dye, is_bleached, bond_partner,
                              dye_type_id, N, R0_nm, tau_D, donor_fid, acceptor_fid,
                              r_intra_nm, intramolecular) -> None:
    """photophysics.py:529-588: bond-mode FRET rates."""
    k = _InPlace(k_fret, np.float64)
    a = _InPlace(fret_acceptor_idx, np.int32)
    lib().orc_fret_rates_bond(P(k.b), P(a.b), P(np.ascontiguousarray(has_dye, dtype=np.uint8)),
                              P(np.ascontiguousarray(is_bleached, dtype=np.uint8)),
                              P(_i32(bond_partner)), P(_i32(dye_type_id)), I(N), D(R0_nm),
                              D(tau_D), I(donor_fid), I(acceptor_fid), D(r_intra_nm),
                              I(int(bool(intramolecular))))
    k.back()
    a.back()


def _compute_fret_rates_proximity_numba(k_fret, fret_acceptor_idx, positions, species_id,
                                        has_dye, is_alive, is_bleached, dye_type_id, N, tau_D,
                                        donor_fid, target_species_id, proximity_radius_um,
                                        k_fret_high, k_fret_low, Lx, Ly, periodic_x,
                                        periodic_y) -> None:
    """photophysics.py:592-683: high rate if a live target is within the radius."""
    k = _InPlace(k_fret, np.float64)
    a = _InPlace(fret_acceptor_idx, np.int32)
    u8 = lambda v: P(np.ascontiguousarray(v, dtype=np.uint8))  # noqa: E731
    lib().orc_fret_rates_proximity(P(k.b), P(a.b), P(_f64(positions)), P(_i32(species_id)),
                                   u8(has_dye), u8(is_alive), u8(is_bleached),
                                   P(_i32(dye_type_id)), I(N), I(donor_fid),
                                   I(target_species_id), D(proximity_radius_um),
                                   D(k_fret_high), D(k_fret_low), D(Lx), D(Ly),
                                   I(int(bool(periodic_x))), I(int(bool(periodic_y))))
    k.back()
    a.back()


def assign_emitter_centric(particle_idx, channel, positions, pixel_cx, pixel_cy, pixel_size,
                           offset_x, offset_y, R, sigma_by_channel, u_pixel):
    """microscope.py:523-631: pixel of every photon by the emitter-centred
    (di-major) CDF walk; returns (px, py, detected) per photon."""
    pidx = _i32(particle_idx)
    n = len(pidx)
    pos = _f64(positions)
    posx = np.ascontiguousarray(pos[:, 0])
    posy = np.ascontiguousarray(pos[:, 1])
    ch = np.ascontiguousarray(channel, dtype=np.int8)
    pcx, pcy = _f64(pixel_cx), _f64(pixel_cy)
    sig = _f64(sigma_by_channel)
    px = np.zeros(n, dtype=np.int32)
    py = np.zeros(n, dtype=np.int32)
    det = np.zeros(n, dtype=np.uint8)
    f = lib().orc_assign_emitter_centric
    f.restype = I
    f(P(pidx), P(ch), P(posx), P(posy), I(n), P(pcx), P(pcy), I(len(pcx)), I(len(pcy)),
      D(pixel_size), D(offset_x), D(offset_y), I(R), P(sig), I(len(sig)), P(_f64(u_pixel)),
      P(px), P(py), P(det))
    return px, py, det.astype(bool)


def accumulate_tcspc(flim_stack, px, py, arrival_times, channel, bin_width, n_bins) -> None:
    """microscope.py:778-803: bin = int(t / bin_width), dropped if >= n_bins."""
    assert flim_stack.dtype == np.uint32 and flim_stack.flags["C_CONTIGUOUS"]
    n_ch, nx, ny, nb = flim_stack.shape
    assert nb == n_bins
    t = _f64(arrival_times)
    lib().orc_accumulate_tcspc(P(flim_stack), P(_i32(px)), P(_i32(py)), P(t),
                               P(np.ascontiguousarray(channel, dtype=np.int8)), I(len(t)),
                               D(bin_width), I(n_bins), I(n_ch), I(nx), I(ny))
