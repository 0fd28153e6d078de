"""Known-answer and closed-form checks that pin the oracle's kernels.

The oracle (oracle/edm_oracle.c through oracle/kernels.py) is the CPU
restatement the GPU path is compared against, so it is checked here first,
without a GPU:

* the reference's own known-answer tests, restated with the same inputs and
  expected outcomes (tests/test_generic_dynamics.py T1-T6,
  tests/test_expert_audit_50.py 07-16 and 21-23, tests/test_bd_fixes.py 1-3,
  tests/test_dynamics.py step 1.3);
* closed forms (force laws, the Ermak-McCammon update, Hill factors, Foerster
  rates, grad D against finite differences);
* brute force on small inputs (pair lists, the emitter-centred pixel CDF).

Expected values are hand-derived numbers or independent numpy computations,
chosen so that a transposed index, a dropped term or a flipped comparison
fails at least one test.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from kernels import (CellList, _compute_fret_rates_numba, _compute_fret_rates_proximity_numba,
                     _min_image_dist_sq, _poisson_draw, accumulate_tcspc,
                     apply_binding_reactions, apply_catalytic_reaction,
                     apply_catalytic_with_feedback, apply_catalytic_with_feedback_and_adaptation,
                     apply_diffusion_field_regions_smooth, apply_diffusion_field_sinusoidal,
                     apply_dissociation, apply_membrane_detachment,
                     apply_membrane_detachment_with_feedback, apply_membrane_recruitment,
                     apply_membrane_recruitment_local,
                     apply_membrane_recruitment_with_dual_feedback,
                     apply_membrane_recruitment_with_feedback, apply_species_conversion,
                     apply_species_conversion_with_feedback, assign_emitter_centric, bd_step,
                     build_recruit_cdf, compute_density_field, compute_forces_harmonic,
                     compute_forces_wca,
                     hill_factor, py_mod, update_adaptation_field)


@pytest.fixture(autouse=True, scope="module")
def _built(oracle_lib):
    return oracle_lib


# ----------------------------------------------------------------------
# Catalysis (reference tests/test_generic_dynamics.py:213-355, 689-756)
# ----------------------------------------------------------------------

def test_catalytic_reaction_simple():
    pos = np.array([[0.1, 0.1], [0.11, 0.11], [0.5, 0.5]])
    sp = np.array([0, 1, 1], dtype=np.int32)
    alive = np.ones(3, dtype=bool)
    n = apply_catalytic_reaction(pos, sp, alive, 3, 0, 1, 2, 0.02, 1.0,
                                 np.array([[0, 1]], dtype=np.int32), 1, np.array([0.0]),
                                 1.0, 1.0, True, True)
    assert n == 1
    assert sp.tolist() == [0, 2, 1]
    assert alive.all()


def test_catalytic_conservation_multiple_pairs():
    pos = np.array([[0.0, 0.0], [0.001, 0.001], [0.2, 0.2], [0.201, 0.201], [0.3, 0.3],
                    [0.4, 0.4], [0.5, 0.5], [0.9, 0.9]])
    sp = np.array([0, 1, 0, 1, 1, 1, 1, 1], dtype=np.int32)
    before = int(np.sum(sp == 1))
    n = apply_catalytic_reaction(pos, sp, np.ones(8, bool), 8, 0, 1, 2, 0.02, 1.0,
                                 np.array([[0, 1], [2, 3]], dtype=np.int32), 2,
                                 np.array([0.1, 0.1]), 1.0, 1.0, True, True)
    assert n == 2
    assert int(np.sum(sp == 1) + np.sum(sp == 2)) == before
    assert sp.tolist() == [0, 2, 0, 2, 1, 1, 1, 1]


def test_catalytic_enzyme_unchanged_and_either_pair_order():
    # the enzyme may be the second member of a pair; the substrate flips once
    pos = np.array([[0.5, 0.5], [0.51, 0.5], [0.5, 0.51]])
    sp = np.array([1, 0, 0], dtype=np.int32)  # substrate 0, enzymes 1 and 2
    pairs = np.array([[0, 1], [0, 2]], dtype=np.int32)
    n = apply_catalytic_reaction(pos, sp, np.ones(3, bool), 3, 0, 1, 2, 0.05, 1.0, pairs, 2,
                                 np.array([0.0, 0.0]), 1.0, 1.0, True, True)
    assert n == 1, "a substrate converts once however many pairs accept"
    assert sp.tolist() == [2, 0, 0]


def test_autocatalysis_cascades_in_pair_order():
    """enzyme == product (E + S -> 2E): the reference's loop visits the pairs
    in order, so a substrate converted at pair p acts as an enzyme for the
    pairs after p only.  A row E, S1, S2, S3 at 10 nm spacing (contact 15 nm:
    only neighbours touch), prob 0.5."""
    pos = np.array([[0.50, 0.5], [0.51, 0.5], [0.52, 0.5], [0.53, 0.5]])

    def run(pairs, u=None):
        sp = np.array([1, 0, 0, 0], dtype=np.int32)  # 1: enzyme and product, 0: substrate
        pr = np.array(pairs, dtype=np.int32)
        uu = np.zeros(len(pr)) if u is None else np.array(u, dtype=np.float64)
        n = apply_catalytic_reaction(pos, sp, np.ones(4, bool), 4, 1, 0, 1, 0.015, 0.5, pr,
                                     len(pr), uu, 1.0, 1.0, True, True)
        return n, sp.tolist()

    # in chain order every link fires: E -> S1 -> S2 -> S3
    assert run([[0, 1], [1, 2], [2, 3]]) == (3, [1, 1, 1, 1])
    # reversed, the later links ran while S1 and S2 were still substrates
    assert run([[2, 3], [1, 2], [0, 1]]) == (1, [1, 1, 0, 0])
    # a pair's orientation does not matter
    assert run([[1, 0], [2, 1], [3, 2]]) == (3, [1, 1, 1, 1])
    # two substrates alone do nothing; S3's only pair came before S2 converted
    assert run([[1, 2], [2, 3]]) == (0, [1, 0, 0, 0])
    assert run([[0, 1], [2, 3], [1, 2]]) == (2, [1, 1, 1, 0])
    # a rejected link (u >= prob) stops the cascade there
    assert run([[0, 1], [1, 2], [2, 3]], u=[0.0, 0.5, 0.0]) == (1, [1, 1, 0, 0])
    # the pair (E, S2) is beyond contact
    assert run([[0, 2]]) == (0, [1, 0, 0, 0])


def test_catalytic_strict_acceptance_and_contact_boundary():
    pos = np.array([[0.0, 0.0], [0.03, 0.04]])  # distance exactly 0.05
    pairs = np.array([[0, 1]], dtype=np.int32)
    for u, contact, expect in [(0.5, 0.05, 0), (0.4999, 0.05, 1), (0.0, 0.0499, 0),
                               (0.0, 0.0501, 1)]:
        sp = np.array([0, 1], dtype=np.int32)
        n = apply_catalytic_reaction(pos, sp, np.ones(2, bool), 2, 0, 1, 2, contact, 0.5,
                                     pairs, 1, np.array([u]), 1.0, 1.0, False, False)
        assert n == expect, (u, contact)
    # a dead member blocks the pair
    sp = np.array([0, 1], dtype=np.int32)
    assert apply_catalytic_reaction(pos, sp, np.array([True, False]), 2, 0, 1, 2, 0.1, 1.0,
                                    pairs, 1, np.array([0.0]), 1.0, 1.0, False, False) == 0


def _feedback_case():
    pos = np.array([[0.1, 0.1], [0.11, 0.11], [0.3, 0.3], [0.31, 0.31], [0.9, 0.9],
                    [0.91, 0.91], [0.5, 0.1], [0.51, 0.11], [0.7, 0.7], [0.71, 0.71]])
    sp = np.array([0, 1] * 5, dtype=np.int32)
    pairs = np.array([[0, 1], [2, 3], [4, 5], [6, 7], [8, 9]], dtype=np.int32)
    density = np.array([[500.0, 500.0], [10.0, 10.0]])  # [gx, gy]
    return pos, sp, pairs, density


def test_catalytic_with_inhibitory_feedback_known_answer():
    pos, sp, pairs, density = _feedback_case()
    # substrates at gx = int(x / 0.5): S0, S1 in row gx=0 (rho 500), S2..S4 in gx=1 (rho 10)
    n = apply_catalytic_with_feedback(pos, sp, np.ones(10, bool), 10, 0, 1, 2, 0.02, 1.0, pairs,
                                      5, np.full(5, 0.5), 1.0, 1.0, True, True, density, 2, 2,
                                      0.5, 0.5, 200.0, 3.0, True)
    # Hill inhibition: 1/(1+2.5^3) = 0.0602 < 0.5 (blocked); 1/(1+0.05^3) ~ 1 (accepted)
    assert n == 3
    assert sp.tolist() == [0, 1, 0, 1, 0, 2, 0, 2, 0, 2]


def test_catalytic_with_activation_feedback_and_basal():
    pos, sp, pairs, density = _feedback_case()
    h_hi = hill_factor(500.0, 200.0, 3.0, False, 0.1)
    h_lo = hill_factor(10.0, 200.0, 3.0, False, 0.1)
    assert h_hi == pytest.approx(0.1 + 0.9 * 2.5 ** 3 / (1 + 2.5 ** 3), rel=1e-14)
    assert h_lo == pytest.approx(0.1 + 0.9 * 0.05 ** 3 / (1 + 0.05 ** 3), rel=1e-14)
    u = np.full(5, 0.5)
    n = apply_catalytic_with_feedback(pos, sp, np.ones(10, bool), 10, 0, 1, 2, 0.02, 1.0, pairs,
                                      5, u, 1.0, 1.0, True, True, density, 2, 2, 0.5, 0.5,
                                      200.0, 3.0, False, 0.1)
    assert n == 2 and sp.tolist() == [0, 2, 0, 2, 0, 1, 0, 1, 0, 1]


def test_catalytic_adaptation_factor():
    pos, sp, pairs, density = _feedback_case()
    w = np.array([[0.0, 0.0], [30.0, 30.0]])  # adaptation field on the same grid
    # activation on all cells ~1 (K tiny); adaptation 1/(1+(30/10)^2) = 0.1 on gx=1
    u = np.full(5, 0.2)
    n = apply_catalytic_with_feedback_and_adaptation(
        pos, sp, np.ones(10, bool), 10, 0, 1, 2, 0.02, 1.0, pairs, 5, u, 1.0, 1.0, True, True,
        density, 2, 2, 0.5, 0.5, 1e-6, 1.0, False, 0.0, w, 10.0, 2.0)
    assert n == 2 and sp.tolist() == [0, 2, 0, 2, 0, 1, 0, 1, 0, 1]


def test_catalytic_adaptation_reads_feedback_indices_on_the_adaptation_array():
    """dynamics.py:1059-1084 takes gx, gy on the FEEDBACK grid (here 4 x 4,
    0.5 um cells) and reads adaptation_field[gx, gy] from the adaptation
    grid's own array (here 2 x 2): numba's unchecked 2-D index is the flat
    element gx * 2 + gy.  Substrate 1 sits in feedback cell (0, 3): flat 3 =
    w[1, 1] = 100 blocks it, although its own adaptation cell (0, 1) holds 0.
    Substrate 3 sits in feedback cell (2, 1): flat 5 lies past the array,
    where the reference has no defined value and the restatement reads the
    substrate's own cell (1, 0) = 0, so it converts (a clamp to the last row
    would read w[1, 1] and block it).  Substrate 5, feedback cell (0, 1),
    flat 1 = 0: converts."""
    pos = np.array([[0.3, 1.69], [0.3, 1.70], [1.2, 0.69], [1.2, 0.70], [0.3, 0.69],
                    [0.3, 0.70]])
    sp = np.array([0, 1] * 3, dtype=np.int32)
    pairs = np.array([[0, 1], [2, 3], [4, 5]], dtype=np.int32)
    density = np.zeros((4, 4))
    w = np.array([[0.0, 0.0], [0.0, 100.0]])
    n = apply_catalytic_with_feedback_and_adaptation(
        pos, sp, np.ones(6, bool), 6, 0, 1, 2, 0.05, 1.0, pairs, 3, np.full(3, 0.2), 2.0, 2.0,
        True, True, density, 4, 4, 0.5, 0.5, 1.0, 1.0, True, 0.0, w, 10.0, 2.0)
    assert n == 2 and sp.tolist() == [0, 1, 0, 2, 0, 2]
    from reference_model import lib
    L = lib()
    L.orc_adapt_cell.restype = ctypes.c_int
    L.orc_adapt_cell.argtypes = [ctypes.c_double] * 2 + [ctypes.c_int] * 4 + \
        [ctypes.c_double] * 2 + [ctypes.c_int] * 2
    # coinciding grids (the reference's production setting, 1 um feedback
    # cells): the read cell is the substrate's own cell
    for gx in range(10):
        for gy in range(10):
            assert L.orc_adapt_cell(gx + 0.5, gy + 0.5, gx, gy, 10, 10, 1.0, 1.0, 10, 10) == \
                gx * 10 + gy


def test_hill_limits():
    assert hill_factor(7.0, 7.0, 3.0, True) == pytest.approx(0.5, abs=1e-15)
    assert hill_factor(7.0, 7.0, 3.0, False, 0.2) == pytest.approx(0.6, abs=1e-15)
    assert hill_factor(0.0, 7.0, 2.0, False, 0.25) == pytest.approx(0.25, abs=1e-15)
    assert hill_factor(1e9, 7.0, 2.0, True) < 1e-15
    assert hill_factor(3.0, 0.0, 2.0, True) == 1.0  # K <= 0: no modulation


# ----------------------------------------------------------------------
# Exchange and conversion (tests/test_generic_dynamics.py:365-564)
# ----------------------------------------------------------------------

def test_membrane_recruitment_index_order_cap():
    N = 50
    sp = np.zeros(N, dtype=np.int32)
    alive = np.concatenate([np.ones(20, bool), np.zeros(30, bool)])
    pos = np.zeros((N, 2))
    acc = np.concatenate([np.ones(20), np.linspace(0.05, 0.45, 30)])
    rpos = np.random.default_rng(42).uniform(0, 1, (N, 2))
    n = apply_membrane_recruitment(sp, alive, pos, N, 0, 0.5, 2.0, 3.0, acc, rpos, 15)
    # every dormant particle accepts (all u < 0.5); the cap takes the first 15 by index
    assert n == 15
    assert alive[:35].all() and not alive[35:].any()
    np.testing.assert_array_equal(pos[20:35], rpos[20:35] * np.array([2.0, 3.0]))
    assert (pos[35:] == 0).all()


def test_membrane_recruitment_acceptance_threshold():
    N = 60
    sp = np.zeros(N, dtype=np.int32)
    alive = np.concatenate([np.ones(30, bool), np.zeros(30, bool)])
    acc = np.concatenate([np.ones(30), np.linspace(0.1, 0.4, 30)])
    rpos = np.random.default_rng(43).uniform(0, 1, (N, 2))
    n = apply_membrane_recruitment(sp, alive, np.zeros((N, 2)), N, 0, 0.3, 1.0, 1.0, acc, rpos,
                                   20)
    expect = int(np.sum(np.linspace(0.1, 0.4, 30) < 0.3))
    assert n == expect == 20
    assert int(alive.sum()) == 30 + n


def test_recruitment_feedback_at_candidate_site():
    N = 4
    sp = np.zeros(N, dtype=np.int32)
    alive = np.zeros(N, bool)
    # candidate sites: (0.25, 0.25) -> cell [0,0]; (0.75, 0.25) -> cell [1,0]
    rpos = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]])
    d1 = np.array([[0.0, 0.0], [100.0, 100.0]])  # high in gx=1
    d2 = np.array([[100.0, 0.0], [100.0, 0.0]])  # high in gy=0
    fb1 = (d1, 2, 2, 0.5, 0.5, 10.0, 2.0, False, 0.0)  # activation by d1
    fb2 = (d2, 2, 2, 0.5, 0.5, 10.0, 2.0, True, 0.0)   # inhibition by d2
    n = apply_membrane_recruitment_with_feedback(sp, alive, np.zeros((N, 2)), N, 0, 1.0, 1.0,
                                                 1.0, np.full(N, 0.5), rpos, 10, *fb1[:8])
    assert n == 2 and alive.tolist() == [False, True, False, True]
    alive[:] = False
    n = apply_membrane_recruitment_with_dual_feedback(sp, alive, np.zeros((N, 2)), N, 0, 1.0,
                                                      1.0, 1.0, np.full(N, 0.5), rpos, 10, *fb1,
                                                      *fb2)
    assert n == 1 and alive.tolist() == [False, False, False, True]


def test_recruit_cdf_floor_and_cell_order():
    """dynamics.py:2355-2370 on a 3 x 5 grid (gx major): a zero field gives
    the uniform CDF; one hot cell puts the whole jump at its flat index
    gx * 5 + gy; densities below 0.01 count as 0.01 and only those."""
    cdf = build_recruit_cdf(np.zeros((3, 5)), 3, 5)
    np.testing.assert_allclose(cdf, np.arange(1, 16) / 15.0, rtol=1e-14)
    assert cdf[-1] == 1.0
    d = np.zeros((3, 5))
    d[2, 1] = 100.0
    cdf = build_recruit_cdf(d, 3, 5)
    total = 100.0 + 14 * 0.01
    np.testing.assert_allclose(cdf[10], 11 * 0.01 / total, rtol=1e-12)
    np.testing.assert_allclose(cdf[11], (11 * 0.01 + 100.0) / total, rtol=1e-12)
    assert np.searchsorted(cdf, 0.5) == 11
    d = np.array([[0.005, 0.02, 0.0], [3.0, 0.0099, 0.5]])
    cdf = build_recruit_cdf(d, 2, 3)
    w = np.array([0.01, 0.02, 0.01, 3.0, 0.01, 0.5])
    np.testing.assert_allclose(cdf, np.cumsum(w) / w.sum(), rtol=1e-13)


def test_local_recruitment_site_offsets_wrap_and_feedback():
    """dynamics.py:1116-1217 on a 3 x 5 grid of 1 um cells over a 3 x 5 um
    box, one hot cell (gx=2, gy=1), activation Hill with no basal rate, so a
    candidate is accepted (u = 0.5) only on the hot cell.  The chosen cell is
    the first with CDF >= u_loc (numpy's searchsorted); the site is its
    centre + sigma * (xi_x, xi_y), wrapped; Hill is read at the site."""
    d = np.zeros((3, 5))
    d[2, 1] = 100.0
    cdf = build_recruit_cdf(d, 3, 5)
    hot = 0.5 * (cdf[10] + cdf[11])
    # slot 0 alive, slot 1 another species, slots 2.. dormant targets
    sp = np.array([0, 1, 0, 0, 0, 0, 0], dtype=np.int32)
    alive = np.array([1, 0, 0, 0, 0, 0, 0], dtype=bool)
    u_loc = np.array([hot, hot, hot, hot, 0.0, 1e-9, cdf[12] + 1e-9])
    # sigma = 0.4: xi = 2.5 moves the site 1 um
    off = np.array([[0, 0], [0, 0], [0.0, 0.0], [2.5, 0.0], [-2.5, 2.5], [2.5, -2.5],
                    [-2.5, 0.0]])
    pos = np.full((7, 2), -1.0)
    n = apply_membrane_recruitment_local(sp, alive, pos, 7, 0, 1.0, 3.0, 5.0, np.full(7, 0.5),
                                         off, 10, d, 3, 5, 1.0, 1.0, 10.0, 2.0, False, 0.0, cdf,
                                         u_loc, 0.4)
    # slot 2: hot cell, no offset -> (2.5, 1.5)
    # slot 3: hot cell, +1 um in x -> 3.5 wraps to 0.5: cold, refused
    # slot 4: cell 0 (0.5, 0.5) moved (-1, +1) -> (2.5, 1.5) after the wrap: hot
    # slot 5: cell 0 moved (+1, -1) -> (1.5, 4.5): cold (the transposed offsets
    #         would have landed on the hot cell)
    # slot 6: cell 13 = (2, 3) (2.5, 3.5) moved -1 um in x -> (1.5, 3.5): cold
    assert n == 2
    assert alive.tolist() == [True, False, True, False, True, False, False]
    np.testing.assert_allclose(pos[2], [2.5, 1.5], rtol=0, atol=1e-15)
    np.testing.assert_allclose(pos[4], [2.5, 1.5], rtol=0, atol=1e-12)
    assert (pos[[1, 3, 5, 6]] == -1.0).all()
    assert np.searchsorted(cdf, u_loc[6]) == 13
    # the index-order cap: with room for one, only slot 2 comes alive
    alive[:] = [True, False, False, False, False, False, False]
    n = apply_membrane_recruitment_local(sp, alive, pos, 7, 0, 1.0, 3.0, 5.0, np.full(7, 0.5),
                                         off, 1, d, 3, 5, 1.0, 1.0, 10.0, 2.0, False, 0.0, cdf,
                                         u_loc, 0.4)
    assert n == 1 and alive.tolist() == [True, False, True, False, False, False, False]


def test_membrane_detachment_counts():
    N = 50
    sp = np.zeros(N, dtype=np.int32)
    alive = np.ones(N, bool)
    u = np.linspace(0.1, 0.9, N)
    n = apply_membrane_detachment(sp, alive, N, 0, 0.5, u)
    assert n == int(np.sum(u < 0.5))
    np.testing.assert_array_equal(alive, u >= 0.5)


def test_detachment_and_conversion_feedback_at_particle():
    N = 4
    pos = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]])
    dens = np.array([[0.0, 100.0], [0.0, 100.0]])  # high in gy=1
    sp = np.zeros(N, dtype=np.int32)
    alive = np.ones(N, bool)
    n = apply_membrane_detachment_with_feedback(sp, alive, pos, N, 0, 1.0, np.full(N, 0.5), dens,
                                                2, 2, 0.5, 0.5, 10.0, 2.0, True)
    assert n == 2 and alive.tolist() == [False, False, True, True]
    alive[:] = True
    n = apply_species_conversion_with_feedback(sp, alive, pos, N, 0, 3, 1.0, np.full(N, 0.5),
                                               dens, 2, 2, 0.5, 0.5, 10.0, 2.0, False)
    assert n == 2 and sp.tolist() == [0, 0, 3, 3]


def test_species_conversion_known_answers():
    sp = np.concatenate([np.zeros(60, np.int32), np.ones(40, np.int32)])
    u = np.random.default_rng(42).uniform(0, 1, 100)
    n = apply_species_conversion(sp, np.ones(100, bool), 100, 0, 1, 0.1, u)
    assert n == int(np.sum(u[:60] < 0.1))
    assert int(np.sum(sp == 0)) == 60 - n and int(np.sum(sp == 1)) == 40 + n
    sp = np.zeros(20, np.int32)
    assert apply_species_conversion(sp, np.ones(20, bool), 20, 0, 2, 1.0, np.zeros(20)) == 20
    assert (sp == 2).all()
    sp = np.zeros(50, np.int32)
    alive = np.concatenate([np.ones(30, bool), np.zeros(20, bool)])
    assert apply_species_conversion(sp, alive, 50, 0, 1, 1.0, np.zeros(50)) == 30
    assert sp[0] == 1 and sp[35] == 0


# ----------------------------------------------------------------------
# Density field (tests/test_generic_dynamics.py:574-679) and adaptation
# ----------------------------------------------------------------------

def test_density_field_uniform_exact():
    pos = np.array([[0.5, 0.5], [1.5, 0.5], [0.5, 1.5], [1.5, 1.5]] * 25)
    out = np.zeros((2, 2))
    compute_density_field(pos, np.zeros(100, np.int32), np.ones(100, bool), 100, 0, 2, 2, 1.0,
                          1.0, out)
    np.testing.assert_array_equal(out, np.full((2, 2), 25.0))


def test_density_field_clustered_and_dormant():
    pos = np.concatenate([np.random.default_rng(42).uniform(0.0, 0.5, (80, 2)),
                          np.random.default_rng(43).uniform(1.5, 2.0, (20, 2))])
    out = np.zeros((2, 2))
    compute_density_field(pos, np.zeros(100, np.int32), np.ones(100, bool), 100, 0, 2, 2, 1.0,
                          1.0, out)
    assert out[0, 0] == 80.0 and out[1, 1] == 20.0 and out[0, 1] == 0.0
    pos = np.random.default_rng(42).uniform(0, 2, (100, 2))
    alive = np.concatenate([np.ones(60, bool), np.zeros(40, bool)])
    out[:] = 0.0
    compute_density_field(pos, np.zeros(100, np.int32), alive, 100, 0, 2, 2, 1.0, 1.0, out)
    assert out.sum() == 60.0


def test_density_field_is_a_sequential_sum():
    # 0.3 x 0.7 cells: inv_area is not a power of two, so k * inv and the
    # reference's repeated += differ in the last bits for some k
    sx, sy = 0.3, 0.7
    inv = 1.0 / (sx * sy)
    rng = np.random.default_rng(5)
    pos = np.column_stack([rng.uniform(0, 0.3, 97), rng.uniform(0, 0.7, 97)])
    out = np.zeros((1, 1))
    compute_density_field(pos, np.zeros(97, np.int32), np.ones(97, bool), 97, 0, 1, 1, sx, sy,
                          out)
    acc = 0.0
    for _ in range(97):
        acc += inv
    assert out[0, 0] == acc
    assert acc != 97 * inv  # the two summations really differ here


def test_grid_cell_truncates_and_clamps():
    # x = Lx falls into the last cell (clamped), negative x into the first
    pos = np.array([[1.0, 0.2], [-0.01, 0.2], [0.49999, 0.2]])
    out = np.zeros((2, 1))
    compute_density_field(pos, np.zeros(3, np.int32), np.ones(3, bool), 3, 0, 2, 1, 0.5, 1.0,
                          out)
    np.testing.assert_array_equal(out, [[4.0], [2.0]])


def test_adaptation_update():
    w = np.array([[1.0, 2.0], [3.0, 4.0]])
    rho = np.array([[5.0, 2.0], [0.0, 8.0]])
    update_adaptation_field(w, rho, 2, 2, 0.01, 0.5)
    np.testing.assert_allclose(w, [[1.08, 2.0], [2.94, 4.08]], rtol=0, atol=1e-15)


# ----------------------------------------------------------------------
# Binding and dissociation (tests/test_expert_audit_50.py:364-433)
# ----------------------------------------------------------------------

def test_binding_conserves_and_sets_bonds():
    N = 100
    pos = np.random.default_rng(1).uniform(0, 1, (N, 2))
    sp = np.zeros(N, np.int32)
    sp[60:] = 1
    bond = np.full(N, -1, np.int32)
    alive = np.ones(N, bool)
    pairs = np.array([[i, 60 + i] for i in range(40)], dtype=np.int32)
    for i in range(40):
        pos[60 + i] = pos[i] + 0.001
    n = apply_binding_reactions(pos, sp, bond, alive, N, 0, 1, 2, 0.01, 1.0, pairs, 40,
                                np.zeros(40), 2.0, 2.0, True, True)
    assert n == 40
    assert int(np.sum((sp == 0) & alive) + np.sum((sp == 2) & alive)) == 60
    assert not alive[60:].any()
    np.testing.assert_array_equal(bond[:40], np.arange(60, 100))
    np.testing.assert_array_equal(bond[60:], np.arange(40))


def test_binding_requires_both_free_and_is_greedy():
    pos = np.array([[0.5, 0.5], [0.501, 0.5], [0.502, 0.5]])
    sp = np.array([0, 1, 1], np.int32)
    bond = np.array([1, 0, -1], np.int32)
    assert apply_binding_reactions(pos, sp, bond, np.ones(3, bool), 3, 0, 1, 2, 1.0, 1.0,
                                   np.array([[0, 2]], np.int32), 1, np.zeros(1), 5.0, 5.0, True,
                                   True) == 0
    # greedy: the first accepted pair in list order wins; (B, A) order swaps roles
    sp = np.array([1, 0, 1], np.int32)
    bond = np.full(3, -1, np.int32)
    alive = np.ones(3, bool)
    pairs = np.array([[0, 1], [1, 2]], np.int32)
    n = apply_binding_reactions(pos, sp, bond, alive, 3, 0, 1, 2, 1.0, 1.0, pairs, 2,
                                np.zeros(2), 5.0, 5.0, True, True)
    assert n == 1
    assert sp.tolist() == [1, 2, 1] and alive.tolist() == [False, True, True]
    assert bond.tolist() == [1, 0, -1]


def test_dissociation_restores_species():
    sp = np.array([2, 1, 2, 1], np.int32)
    bond = np.array([1, 0, 3, 2], np.int32)
    alive = np.array([True, False, True, False])
    pos = np.array([[1.0, 1.0], [7.0, 7.0], [2.0, 2.0], [9.0, 9.0]])
    u = np.array([0.2, 0.0, 0.8, 0.0])
    n = apply_dissociation(sp, bond, alive, pos, 4, 2, 0, 1, 0.5, u)
    assert n == 1  # u is read at the complex's index: 0.2 < 0.5, 0.8 >= 0.5
    assert sp.tolist() == [0, 1, 2, 1]
    assert alive.tolist() == [True, True, True, False]
    assert bond.tolist() == [-1, -1, 3, 2]
    np.testing.assert_array_equal(pos[1], [1.0, 1.0])  # partner reappears on A


# ----------------------------------------------------------------------
# Forces and the integrator (tests/test_bd_fixes.py:41-151,
# tests/test_expert_audit_50.py:197-309)
# ----------------------------------------------------------------------

def test_wca_force_closed_form_and_newton3():
    sigma, eps = 0.008, 2.5
    r = 0.9 * sigma
    pos = np.array([[1.0, 1.0], [1.0 + r, 1.0]])
    f = np.zeros((2, 2))
    compute_forces_wca(pos, f, np.zeros(2, np.int32), np.ones(2, bool),
                       np.array([[0, 1]], np.int32), 1, [[sigma]], [[eps]], 10.0, 10.0, True,
                       True)
    s6 = (sigma / r) ** 6
    f_over_r = 24 * eps * (2 * s6 * s6 - s6) / r ** 2
    assert f_over_r < 2400 * eps / sigma ** 2  # uncapped here
    assert f[0, 0] == pytest.approx(-f_over_r * r, rel=1e-12)
    assert f[1, 0] == pytest.approx(f_over_r * r, rel=1e-12)
    assert f[0, 1] == 0.0 and f[1, 1] == 0.0


def test_wca_force_cap_and_range():
    sigma, eps = 0.008, 2.5
    for r, expect in [(0.1 * sigma, 2400 * eps / sigma ** 2 * 0.1 * sigma),
                      (2 ** (1 / 6) * sigma * 1.0001, 0.0)]:
        pos = np.array([[1.0, 1.0], [1.0, 1.0 + r]])
        f = np.zeros((2, 2))
        compute_forces_wca(pos, f, np.zeros(2, np.int32), np.ones(2, bool),
                           np.array([[0, 1]], np.int32), 1, [[sigma]], [[eps]], 10.0, 10.0,
                           True, True)
        assert f[1, 1] == pytest.approx(expect, rel=1e-12, abs=0.0)


def test_wca_across_periodic_boundary_and_species_matrix():
    # particles 0.004 apart across x = 0; the (1, 0) matrix entry applies
    pos = np.array([[0.002, 0.5], [9.998, 0.5]])
    sig = np.array([[0.0, 0.008], [0.008, 0.0]])
    eps = np.array([[0.0, 1.0], [1.0, 0.0]])
    f = np.zeros((2, 2))
    compute_forces_wca(pos, f, np.array([1, 0], np.int32), np.ones(2, bool),
                       np.array([[0, 1]], np.int32), 1, sig, eps, 10.0, 10.0, True, True)
    assert f[0, 0] > 0 and f[1, 0] < 0  # pushed apart through the boundary
    f2 = np.zeros((2, 2))
    compute_forces_wca(pos, f2, np.array([1, 1], np.int32), np.ones(2, bool),
                       np.array([[0, 1]], np.int32), 1, sig, eps, 10.0, 10.0, True, True)
    assert (f2 == 0).all()  # sigma(1, 1) = 0: no interaction


def test_harmonic_force_value_and_bound():
    sigma, k = 0.008, 50.0
    pos = np.array([[0.0, 0.0], [0.006, 0.0]])
    f = np.zeros((2, 2))
    args = (np.zeros(2, np.int32), np.ones(2, bool), np.array([[0, 1]], np.int32), 1,
            [[sigma]], [[k]], 100.0, 100.0, True, True)
    compute_forces_harmonic(pos, f, *args)
    assert f[1, 0] == pytest.approx(k * (sigma - 0.006), rel=1e-12)
    assert f[0, 0] == -f[1, 0] and f[0, 1] == 0.0 and f[1, 1] == 0.0
    pos = np.array([[0.0, 0.0], [0.1 * sigma, 0.0]])
    f[:] = 0
    compute_forces_harmonic(pos, f, *args)
    assert f[1, 0] == pytest.approx(k * 0.9 * sigma, rel=1e-12)
    # dead particles feel nothing
    f[:] = 0
    compute_forces_harmonic(pos, f, np.zeros(2, np.int32), np.array([True, False]),
                            *args[2:])
    assert (f == 0).all()


def test_bd_step_closed_form():
    pos = np.array([[1.0, 2.0]])
    D, g, F, xi, dt = 0.5, (0.2, -0.1), (3.0, 4.0), (0.7, -1.2), 1e-3
    bd_step(pos, np.array([D]), np.array([F]), np.array([g]), dt, 10.0, 10.0, True, True, 1,
            np.array([xi]), np.ones(1, bool), 1.0)
    sn = math.sqrt(2 * D * dt)
    ex = 1.0 + (g[0] * dt + sn * xi[0] + D * F[0] * dt)
    ey = 2.0 + (g[1] * dt + sn * xi[1] + D * F[1] * dt)
    assert pos[0, 0] == pytest.approx(ex, rel=0, abs=1e-15)
    assert pos[0, 1] == pytest.approx(ey, rel=0, abs=1e-15)


def test_bd_step_displacement_cap_keeps_direction():
    pos = np.array([[50.0, 50.0]])
    bd_step(pos, np.array([1.0]), np.array([[3e8, 4e8]]), np.zeros((1, 2)), 1e-3, 100.0, 100.0,
            True, True, 1, np.zeros((1, 2)), np.ones(1, bool), 0.004)
    d = pos[0] - 50.0
    assert math.hypot(*d) == pytest.approx(0.004, rel=1e-12)
    assert d[0] == pytest.approx(0.0024, rel=1e-12) and d[1] == pytest.approx(0.0032, rel=1e-12)


def test_bd_step_boundaries_and_dead():
    # periodic: Python modulo; reflective: one mirror image
    pos = np.array([[0.001, 9.999], [5.0, 5.0]])
    bd_step(pos, np.array([1.0, 1.0]), np.array([[-1.0, 1.0], [0.0, 0.0]]), np.zeros((2, 2)),
            1.0, 10.0, 10.0, True, True, 2, np.array([[0.0, 0.0], [9.0, 9.0]]),
            np.array([True, False]), 0.003)
    assert pos[0, 0] == pytest.approx(py_mod(0.001 - 0.003 / math.sqrt(2), 10.0), abs=1e-15)
    assert 0.0 <= pos[0, 0] < 10.0 and pos[0, 0] > 9.99
    assert pos[1].tolist() == [5.0, 5.0]  # dead: immobile
    pos = np.array([[0.001, 9.999]])
    bd_step(pos, np.array([1.0]), np.array([[-1.0, 1.0]]), np.zeros((1, 2)), 1.0, 10.0, 10.0,
            False, False, 1, np.zeros((1, 2)), np.ones(1, bool), 0.003)
    step = 0.003 / math.sqrt(2)
    assert pos[0, 0] == pytest.approx(-(0.001 - step), abs=1e-15)
    assert pos[0, 1] == pytest.approx(20.0 - (9.999 + step), abs=1e-12)


def test_bd_positions_stay_in_box_and_msd():
    rng = np.random.default_rng(99)
    N, L = 400, 50.0
    pos = rng.uniform(10.0, 40.0, (N, 2))
    p0 = pos.copy()
    D, dt, steps = 0.1, 1e-3, 200
    for _ in range(steps):
        bd_step(pos, np.full(N, D), np.zeros((N, 2)), np.zeros((N, 2)), dt, L, L, True, True, N,
                rng.standard_normal((N, 2)), np.ones(N, bool), 100.0)
        assert (pos >= 0).all() and (pos < L).all()
    dr = pos - p0
    dr -= np.round(dr / L) * L
    msd = float(np.mean(np.sum(dr ** 2, axis=1)))
    assert abs(msd - 4 * D * dt * steps) / (4 * D * dt * steps) < 0.12


# ----------------------------------------------------------------------
# Geometry and neighbour search (tests/test_dynamics.py:47-99,
# tests/test_expert_audit_50.py:316-357)
# ----------------------------------------------------------------------

def test_min_image_symmetry_and_bounds():
    rng = np.random.default_rng(3)
    Lx, Ly = 2.0, 3.0
    for _ in range(200):
        x1, x2 = rng.uniform(0, Lx, 2)
        y1, y2 = rng.uniform(0, Ly, 2)
        d12, dx, dy = _min_image_dist_sq(x1, y1, x2, y2, Lx, Ly, True, True)
        d21, _, _ = _min_image_dist_sq(x2, y2, x1, y1, Lx, Ly, True, True)
        assert d12 == d21
        assert abs(dx) <= Lx / 2 and abs(dy) <= Ly / 2
        assert d12 == pytest.approx(dx * dx + dy * dy, rel=1e-15)
    d, dx, dy = _min_image_dist_sq(0.1, 0.1, 1.9, 2.9, Lx, Ly, True, True)
    assert dx == pytest.approx(0.2) and dy == pytest.approx(0.2)
    d, dx, dy = _min_image_dist_sq(0.1, 0.1, 1.9, 2.9, Lx, Ly, False, False)
    assert dx == pytest.approx(-1.8) and dy == pytest.approx(-2.8)


@pytest.mark.parametrize("periodic", [(True, True), (False, False), (True, False)])
def test_cell_list_equals_brute_force(periodic):
    rng = np.random.default_rng(42)
    N, L, cut = 500, 5.0, 0.5
    pos = rng.uniform(0, L, (N, 2))
    cl = CellList(L, L, cut, periodic=periodic)
    cl.build(pos, N)
    pc = cl.get_neighbor_pairs(pos, N)
    pb = cl.get_neighbors_brute(pos, N, cut)
    sc = set(map(tuple, pc.tolist()))
    sb = set(map(tuple, pb.tolist()))
    assert len(sc) == len(pc), "duplicate pairs"
    assert all(i < j for i, j in sc)
    assert sc == sb
    # every listed pair is within the cutoff, by an independent distance
    d = pos[pc[:, 0]] - pos[pc[:, 1]]
    for ax, per in enumerate(periodic):
        if per:
            d[:, ax] -= L * np.round(d[:, ax] / L)
    assert (np.sum(d * d, axis=1) < cut * cut).all()


def test_cell_list_small_box_and_cap():
    rng = np.random.default_rng(7)
    pos = rng.uniform(0, 1.0, (60, 2))
    cl = CellList(1.0, 1.0, 0.4, periodic=(True, True))  # nx = 3 (minimum)
    assert cl.nx == 3
    cl.build(pos, 60)
    full = cl.get_neighbor_pairs(pos, 60)
    assert set(map(tuple, full.tolist())) == set(map(tuple, cl.get_neighbors_brute(
        pos, 60, 0.4).tolist()))
    capped = cl.get_neighbor_pairs(pos, 60, max_pairs=25)
    np.testing.assert_array_equal(capped, full[:25])  # the cap keeps the enumeration prefix


def test_cell_list_rejects_nonpositive_cutoff():
    with pytest.raises(ValueError):
        CellList(1.0, 1.0, 0.0)


# ----------------------------------------------------------------------
# Diffusion fields (tests/test_expert_audit_50.py:873-921)
# ----------------------------------------------------------------------

def _smooth(x, y):
    D = np.zeros(1)
    g = np.zeros((1, 2))
    apply_diffusion_field_regions_smooth(D, g, np.array([0.1]), np.zeros(1, np.int32),
                                         np.array([[x, y]]), 1, [1.0], [1.0], [0.5], [0.3], 1,
                                         0.05)
    return D[0], g[0]


def test_smooth_field_gradient_matches_finite_difference():
    eps = 1e-6
    for x0, y0 in [(1.49, 1.0), (1.0, 1.52), (1.3, 1.3)]:
        D0, g = _smooth(x0, y0)
        gx = (_smooth(x0 + eps, y0)[0] - _smooth(x0 - eps, y0)[0]) / (2 * eps)
        gy = (_smooth(x0, y0 + eps)[0] - _smooth(x0, y0 - eps)[0]) / (2 * eps)
        assert g[0] == pytest.approx(gx, rel=1e-4, abs=1e-9)
        assert g[1] == pytest.approx(gy, rel=1e-4, abs=1e-9)
    # deep inside the region D = D_base * factor; far outside D = D_base, no gradient
    D_in, g_in = _smooth(1.0, 1.05)
    assert D_in == pytest.approx(0.03, rel=1e-3)
    D_out, g_out = _smooth(5.0, 5.0)
    assert D_out == 0.1 and (g_out == 0).all()


def test_sinusoidal_field_gradient_matches_finite_difference():
    L = 10.0
    k = 2 * math.pi / L

    def ev(x, y):
        D = np.zeros(1)
        g = np.zeros((1, 2))
        apply_diffusion_field_sinusoidal(D, g, np.array([0.15]), np.zeros(1, np.int32),
                                         np.array([[x, y]]), 1, 0.3, k, k)
        return D[0], g[0]

    for x0, y0 in [(1.0, 2.0), (7.3, 4.4), (2.5, 2.5)]:
        D0, g = ev(x0, y0)
        assert D0 == pytest.approx(0.15 * (1 + 0.3 * math.sin(k * x0) * math.sin(k * y0)),
                                   rel=1e-14)
        e = 1e-6
        assert g[0] == pytest.approx((ev(x0 + e, y0)[0] - ev(x0 - e, y0)[0]) / (2 * e),
                                     rel=1e-5, abs=1e-10)
        assert g[1] == pytest.approx((ev(x0, y0 + e)[0] - ev(x0, y0 - e)[0]) / (2 * e),
                                     rel=1e-5, abs=1e-10)


def test_sincos_of_the_sinusoidal_field_matches_libm():
    """The field's sin/cos (Cody-Waite reduction + Taylor polynomials, IEEE
    operations only, so the GPU's FP64 build computes the same bits) agree
    with libm within one unit in the last place of 1 (2.3e-16 absolute;
    measured 1.1e-16) over the arguments the field takes (k x for x in the
    box, many wavelengths), in every quadrant; sin^2 + cos^2 = 1 to 5e-16 and
    the odd/even symmetries hold exactly."""
    from reference_model import lib
    L = lib()
    f = L.orc_sincos
    f.restype = None
    f.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                  ctypes.POINTER(ctypes.c_double)]
    s, c = ctypes.c_double(), ctypes.c_double()
    xs = np.concatenate([np.linspace(-40.0, 40.0, 20001),
                         np.random.default_rng(0).uniform(0.0, 2 * math.pi * 50, 20000),
                         np.arange(-8, 9) * (math.pi / 4), [0.0, 1e-300, -1e-12]])
    err_s = err_c = err_1 = 0.0
    for x in xs:
        f(float(x), ctypes.byref(s), ctypes.byref(c))
        err_s = max(err_s, abs(s.value - math.sin(x)))
        err_c = max(err_c, abs(c.value - math.cos(x)))
        err_1 = max(err_1, abs(s.value * s.value + c.value * c.value - 1.0))
        f(float(-x), ctypes.byref(s2 := ctypes.c_double()), ctypes.byref(c2 := ctypes.c_double()))
        if abs(x) < 2.0:  # exact symmetry where the reduction has k in {-1, 0, 1}
            assert s2.value == -s.value and c2.value == c.value
    assert err_s < 2.3e-16 and err_c < 2.3e-16, (err_s, err_c)
    assert err_1 < 5e-16


# ----------------------------------------------------------------------
# Photophysics (tests/test_expert_audit_50.py:503-523,
# tests/test_generic_dynamics.py:792-935)
# ----------------------------------------------------------------------

def test_poisson_draw_guards_and_means():
    assert all(_poisson_draw(0.0, i / 100) == 0 for i in range(100))
    assert _poisson_draw(-5.0, 0.5) == 0
    for mean in (0.7, 10.0, 55.0):  # Knuth branch and Box-Muller branch
        n = 20000
        draws = np.array([_poisson_draw(mean, (i + 0.5) / n) for i in range(n)])
        assert abs(draws.mean() - mean) < 4 * math.sqrt(mean / n) + 0.02
        assert abs(draws.var() - mean) < 0.1 * mean + 0.05
        assert (draws >= 0).all()


def test_poisson_draw_lcg_known_answer():
    # The draw is a pure function of (mean, u): a 64-bit LCG seeded with
    # int(u * 2^53); Knuth's product method below mean 30.  Replayed here with
    # Python integers (no wrap-around arithmetic shared with the C code).
    A, C, M = 6364136223846793005, 1442695040888963407, 1 << 64

    def knuth(mean, u):
        s = int(u * 9007199254740992.0) % M
        L, k, p = math.exp(-mean), 0, 1.0
        while True:
            s = (s * A + C) % M
            p *= (s >> 11) / 9007199254740992.0
            if p <= L:
                return k
            k += 1

    rng = np.random.default_rng(11)
    for mean, u in zip(rng.uniform(0.01, 29.9, 300), rng.random(300)):
        assert _poisson_draw(float(mean), float(u)) == knuth(float(mean), float(u))


def test_bond_fret_rates_closed_form():
    N = 6
    has_dye = np.array([1, 1, 1, 1, 1, 0], bool)
    bleached = np.array([0, 0, 0, 1, 0, 0], bool)
    bond = np.array([1, 0, 5, 4, 3, 2], np.int32)
    dye = np.array([0, 1, 0, 0, 1, -1], np.int32)
    tau_D, R0, r = 2.793e-9, 5.4, 5.0
    k = np.zeros(N)
    acc = np.full(N, -1, np.int32)
    _compute_fret_rates_numba(k, acc, has_dye, bleached, bond, dye, N, R0, tau_D, 0, 1, r, False)
    kf = (R0 / r) ** 6 / tau_D
    # donor 0 bonded to acceptor-dyed 1; donor 2's partner has no dye; donor 3 bleached
    assert k[0] == pytest.approx(kf, rel=1e-14) and acc[0] == 1
    assert k[2] == 0.0 and acc[2] == -1
    assert k[3] == 0.0 and k[1] == 0.0 and k[4] == 0.0
    E = kf * tau_D / (1 + kf * tau_D)
    assert E == pytest.approx(1 / (1 + (r / R0) ** 6), rel=1e-12)
    assert tau_D * (1 - E) == pytest.approx(1.0 / (1.0 / tau_D + kf), rel=1e-12)
    k[:] = 0
    acc[:] = -1
    _compute_fret_rates_numba(k, acc, has_dye, bleached, bond, dye, N, R0, tau_D, 0, 1, r, True)
    # intramolecular: any bonded unbleached donor, acceptor index -1
    assert k[0] == pytest.approx(kf) and k[2] == pytest.approx(kf) and (acc == -1).all()


def test_proximity_fret_rates():
    pos = np.array([[0.0, 0.0], [0.001, 0.001], [0.1, 0.1], [0.5, 0.5], [0.2, 0.2],
                    [0.8, 0.8], [0.99, 0.99]])
    sp = np.array([1, 2, 1, 2, 1, 2, 1], np.int32)
    alive = np.array([1, 1, 1, 0, 1, 1, 1], bool)
    bleached = np.array([0, 0, 0, 0, 0, 1, 0], bool)
    k = np.zeros(7)
    acc = np.full(7, 5, np.int32)
    _compute_fret_rates_proximity_numba(k, acc, pos, sp, np.ones(7, bool), alive, bleached,
                                        np.zeros(7, np.int32), 7, 4.1e-9, 0, 2, 0.02, 1e8, 1e6,
                                        1.0, 1.0, True, True)
    assert k[0] == 1e8  # live target 1.4 nm away
    assert k[2] == 1e6 and k[4] == 1e6
    assert k[6] == 1e8  # the periodic image of target 1 is 15.6 nm away
    assert k[5] == 0.0  # bleached donor untouched
    assert k[3] == 0.0  # dead donor untouched
    assert acc[0] == -1 and acc[5] == 5


# ----------------------------------------------------------------------
# Imaging (microscope.py:523-631, 778-803; tests/test_expert_audit_50.py:557-565)
# ----------------------------------------------------------------------

def test_emitter_centric_pixel_matches_brute_force_cdf():
    rng = np.random.default_rng(21)
    nx = ny = 16
    pix = 0.1
    offx = offy = 0.0
    pcx = offx + (np.arange(nx) + 0.5) * pix
    pcy = offy + (np.arange(ny) + 0.5) * pix
    R, sig = 3, np.array([0.06, 0.08])
    npart = 40
    pos = rng.uniform(-0.2, 1.8, (npart, 2))
    pidx = np.sort(rng.integers(0, npart, 300)).astype(np.int32)
    ch = rng.integers(0, 2, 300).astype(np.int8)
    u = rng.random(300)
    px, py, det = assign_emitter_centric(pidx, ch, pos, pcx, pcy, pix, offx, offy, R, sig, u)
    k = 0
    for p in range(300):
        x, y = pos[pidx[p]]
        s = sig[ch[p]]
        i0 = math.floor((x - offx) / pix - 0.5 + 0.5)
        j0 = math.floor((y - offy) / pix - 0.5 + 0.5)
        if not (-R <= i0 < nx + R and -R <= j0 < ny + R):
            assert not det[p]
            continue
        cells = [(i, j) for i in range(i0 - R, i0 + R + 1) if 0 <= i < nx
                 for j in range(j0 - R, j0 + R + 1) if 0 <= j < ny]
        if not cells:
            assert not det[p]
            continue
        w = np.array([math.exp(-((x - pcx[i]) ** 2 + (y - pcy[j]) ** 2) / (2 * s * s))
                      for i, j in cells])
        c = np.cumsum(w)
        q = int(np.searchsorted(c, u[p] * c[-1], side="left"))
        q = min(q, len(cells) - 1)
        assert det[p]
        assert (px[k], py[k]) == cells[q], p
        k += 1
    assert k == int(det.sum())


def test_tcspc_binning_and_range():
    flim = np.zeros((2, 3, 3, 256), np.uint32)
    bw = 25e-9 / 256
    t = np.array([0.0, bw * 0.999, bw * 1.0001, -1e-12, 26e-9, 255.5 * bw, 3 * bw])
    px = np.array([0, 0, 0, 1, 1, 2, 5], np.int32)  # the last photon is off the image
    py = np.array([0, 0, 0, 1, 1, 2, 0], np.int32)
    ch = np.array([0, 0, 0, 1, 1, 1, 0], np.int8)
    accumulate_tcspc(flim, px, py, t, ch, bw, 256)
    assert flim[0, 0, 0, 0] == 2 and flim[0, 0, 0, 1] == 1
    assert flim[1, 1, 1, 0] == 1  # negative times clamp to bin 0; 26 ns (bin 266) dropped
    assert flim[1, 2, 2, 255] == 1
    assert int(flim.sum()) == 5
    assert int(26e-9 / bw) >= 256
