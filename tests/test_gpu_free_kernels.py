"""The reference's free dynamics kernels (dynamics.py:38-60) on the GPU.

* The restated known-answer tests of the reference (tests/test_oracle_kernels.py,
  which pin the oracle) are run again with every kernel they call replaced by
  the GPU function of the same name from ``ed_mrdt.dynamics``: the same
  inputs must give the same hand-derived answers.
* Each kernel is compared with the oracle on larger seeded inputs (a few
  thousand particles, the oracle's reference-ordered pair lists, numpy
  uniforms): integer and flag arrays and event counts bit-exact, positions
  and forces bit-exact where the arithmetic is IEEE-exact (pair order kept),
  D(x) fields within 1e-15 relative (exp differs by an ulp).
* The same kernels on DeviceArray arguments act in place with no host copy.
"""

import numpy as np
import pytest

import kernels as ok
import test_oracle_kernels as T
from ed_mrdt import dynamics as gd

pytestmark = pytest.mark.gpu

GPU_NAMES = [n for n in gd.__all__ if n not in ("DynamicsEngine", "derive_constants")]


def _uses_gpu_names(fn):
    names = set(fn.__code__.co_names)
    for c in fn.__code__.co_consts:
        if hasattr(c, "co_names"):
            names |= set(c.co_names)
    return bool(names & set(GPU_NAMES))


def _cases(name):
    """(name, kwargs) for each case of a known-answer test, expanding its
    parametrize marks."""
    fn = getattr(T, name)
    cases = [{}]
    for mark in getattr(fn, "pytestmark", []):
        if mark.name != "parametrize":
            continue
        argnames, values = mark.args[0], mark.args[1]
        names = [a.strip() for a in argnames.split(",")] if isinstance(argnames, str) \
            else list(argnames)
        expanded = []
        for c in cases:
            for v in values:
                v = v if len(names) > 1 else (v,)
                expanded.append({**c, **dict(zip(names, v))})
        cases = expanded
    return [(name, c) for c in cases]


KNOWN = [c for n in sorted(dir(T)) if n.startswith("test_") and _uses_gpu_names(getattr(T, n))
         for c in _cases(n)]


def test_every_reference_export_is_a_gpu_function():
    ref_all = ["CellList", "bd_step", "apply_binding_reactions", "apply_dissociation",
               "compute_forces_wca", "DynamicsEngine", "apply_catalytic_reaction",
               "apply_catalytic_with_feedback", "apply_membrane_recruitment",
               "apply_membrane_recruitment_with_feedback",
               "apply_membrane_recruitment_with_dual_feedback", "apply_membrane_detachment",
               "apply_membrane_detachment_with_feedback", "apply_species_conversion",
               "apply_species_conversion_with_feedback", "compute_density_field",
               "compute_forces_harmonic", "apply_diffusion_field_regions_smooth",
               "apply_catalytic_with_feedback_and_adaptation", "update_adaptation_field",
               "apply_membrane_recruitment_local"]
    for n in ref_all:
        assert n in gd.__all__ and getattr(gd, n).__module__.startswith("ed_mrdt."), n
    assert len(KNOWN) >= 20, KNOWN


@pytest.mark.parametrize("name,kwargs", KNOWN,
                         ids=[n + "".join(f"-{v}" for v in k.values()) for n, k in KNOWN])
def test_reference_known_answers_on_gpu(oracle_lib, monkeypatch, name, kwargs):
    for n in GPU_NAMES:
        if hasattr(T, n):
            monkeypatch.setattr(T, n, getattr(gd, n))
    getattr(T, name)(**kwargs)


# ---------------------------------------------------------------- seeded
def _system(seed=0, n=4000, L=6.0, ns=5):
    g = np.random.default_rng(seed)
    pos = g.uniform(0.0, L, (n, 2))
    sp = g.integers(0, ns, n).astype(np.int32)
    alive = g.random(n) < 0.9
    return g, pos, sp, alive


def _pairs(pos, n, L, cutoff=0.165, periodic=(True, True)):
    cl = ok.CellList(L, L, cutoff, periodic)
    cl.build(pos, n)
    return cl, cl.get_neighbor_pairs(pos, n, max_pairs=min(30 * n, 500000))


@pytest.mark.parametrize("periodic", [(True, True), (False, False), (True, False)])
def test_pair_list_matches_oracle_order(oracle_lib, periodic):
    g, pos, sp, alive = _system(1, n=6000, L=5.0)
    pos[:10] = [[0, 0], [5.0, 5.0], [4.9999999, 0.1], [0.0, 2.5], [2.5, 0.0], [1e-12, 1e-12],
                [5.0, 0.0], [0.165, 0.165], [0.33, 0.33], [2.0, 2.0]]
    cl, want = _pairs(pos, 6000, 5.0, periodic=periodic)
    gcl = gd.CellList(5.0, 5.0, 0.165, periodic)
    gcl.build(pos, 6000)
    np.testing.assert_array_equal(gcl.head, cl.head)
    np.testing.assert_array_equal(gcl.next_particle[:6000], cl.next_particle[:6000])
    got = gcl.get_neighbor_pairs(pos, 6000)
    np.testing.assert_array_equal(got, want)
    # the cap: fewer stored than found, the first ones in the same order
    few = gcl.get_neighbor_pairs(pos, 6000, max_pairs=1000)
    np.testing.assert_array_equal(few, want[:1000])
    assert gcl.n_found == len(want) > 1000
    bf = gcl.get_neighbors_brute(pos[:800], 800, 0.165)
    obf = cl.get_neighbors_brute(pos[:800], 800, 0.165)
    np.testing.assert_array_equal(np.asarray(bf), np.asarray(obf))


def test_pair_list_reference_cap_crossed(oracle_lib):
    """A dense box (about 45 pairs per particle): the reference's
    min(30 N, 500000) cap truncates the list at the same pair."""
    g = np.random.default_rng(3)
    n = 3000
    pos = g.uniform(0.0, 1.6, (n, 2))
    cl = ok.CellList(1.6, 1.6, 0.165)
    cl.build(pos, n)
    full = cl.get_neighbor_pairs(pos, n)
    assert len(full) > 30 * n
    got = gd.CellList(1.6, 1.6, 0.165).get_neighbor_pairs(pos, n)
    np.testing.assert_array_equal(got, full[: 30 * n])


@pytest.mark.parametrize("potential", ["wca", "harmonic"])
def test_pair_forces_match_oracle_bit_for_bit(oracle_lib, potential):
    g, pos, sp, alive = _system(2, n=5000, L=3.0)
    cl, pairs = _pairs(pos, 5000, 3.0, cutoff=0.03)
    sig = g.uniform(0.005, 0.012, (5, 5))
    sig = 0.5 * (sig + sig.T)
    par = g.uniform(0.5, 2.0, (5, 5)) * (1.0 if potential == "wca" else 1e5)
    par[0, 1] = par[1, 0] = 0.0  # a switched-off species pair
    f0 = g.normal(size=(5000, 2))
    fo, fg = f0.copy(), f0.copy()
    fn_o = ok.compute_forces_wca if potential == "wca" else ok.compute_forces_harmonic
    fn_g = gd.compute_forces_wca if potential == "wca" else gd.compute_forces_harmonic
    fn_o(pos, fo, sp, alive, pairs, len(pairs), sig, par, 3.0, 3.0, True, True)
    fn_g(pos, fg, sp, alive, pairs, len(pairs), sig, par, 3.0, 3.0, True, True)
    assert np.count_nonzero(fo != f0) > 100
    np.testing.assert_array_equal(fg, fo)


@pytest.mark.parametrize("periodic", [True, False])
def test_bd_step_matches_oracle_bit_for_bit(oracle_lib, periodic):
    g, pos, sp, alive = _system(4, n=5000, L=2.0)
    D = g.uniform(0.05, 0.2, 5000)
    F = g.normal(scale=50.0, size=(5000, 2))
    F[:50] *= 1e4  # displacement cap
    gD = g.normal(scale=0.1, size=(5000, 2))
    noise = g.standard_normal((5000, 2))
    po, pg = pos.copy(), pos.copy()
    ok.bd_step(po, D, F, gD, 1e-3, 2.0, 2.0, periodic, periodic, 5000, noise, alive, 0.15)
    gd.bd_step(pg, D, F, gD, 1e-3, 2.0, 2.0, periodic, periodic, 5000, noise, alive, 0.15)
    np.testing.assert_array_equal(pg, po)
    assert np.all(pg[~alive] == pos[~alive])


def test_diffusion_regions_match_oracle(oracle_lib):
    g, pos, sp, alive = _system(5, n=5000, L=4.0)
    Db = np.array([0.15, 0.15, 0.1, 0.08, 0.08])
    rc = (np.array([1.0, 3.0]), np.array([1.0, 2.5]), np.array([0.6, 0.9]), np.array([0.3, 2.0]))
    Do, go = np.zeros(5000), np.zeros((5000, 2))
    Dg, gg = np.zeros(5000), np.zeros((5000, 2))
    ok.apply_diffusion_field_regions_smooth(Do, go, Db, sp, pos, 5000, *rc, 2, 0.05)
    gd.apply_diffusion_field_regions_smooth(Dg, gg, Db, sp, pos, 5000, *rc, 2, 0.05)
    np.testing.assert_allclose(Dg, Do, rtol=1e-15, atol=0)
    np.testing.assert_allclose(gg, go, rtol=1e-13, atol=1e-300)
    assert np.count_nonzero(Do != Db[sp]) > 500


def _reaction_inputs(seed):
    g, pos, sp, alive = _system(seed, n=5000, L=4.0)
    cl, pairs = _pairs(pos, 5000, 4.0)
    dens = g.uniform(0.0, 60.0, (8, 8))
    return g, pos, sp, alive, pairs, dens


def test_binding_and_dissociation_match_oracle(oracle_lib):
    g, pos, sp, alive, pairs, _ = _reaction_inputs(6)
    bond = np.full(5000, -1, np.int32)
    u = g.random(len(pairs))
    a = [x.copy() for x in (sp, bond, alive)]
    b = [x.copy() for x in (sp, bond, alive)]
    no = ok.apply_binding_reactions(pos, a[0], a[1], a[2], 5000, 1, 2, 3, 0.1, 0.6, pairs,
                                    len(pairs), u, 4.0, 4.0, True, True)
    ng = gd.apply_binding_reactions(pos, b[0], b[1], b[2], 5000, 1, 2, 3, 0.1, 0.6, pairs,
                                    len(pairs), u, 4.0, 4.0, True, True)
    assert ng == no > 50
    for x, y in zip(a, b):
        np.testing.assert_array_equal(y, x)
    ud = g.random(5000)
    po, pg = pos.copy(), pos.copy()
    no = ok.apply_dissociation(a[0], a[1], a[2], po, 5000, 3, 1, 2, 0.4, ud)
    ng = gd.apply_dissociation(b[0], b[1], b[2], pg, 5000, 3, 1, 2, 0.4, ud)
    assert ng == no > 10
    for x, y in zip(a, b):
        np.testing.assert_array_equal(y, x)
    np.testing.assert_array_equal(pg, po)


@pytest.mark.parametrize("form", ["plain", "feedback", "adaptation"])
def test_catalysis_matches_oracle(oracle_lib, form):
    g, pos, sp, alive, pairs, dens = _reaction_inputs(7)
    u = g.random(len(pairs))
    so, sg = sp.copy(), sp.copy()
    args = (pos, None, alive, 5000, 2, 0, 1, 0.15, 0.4, pairs, len(pairs), u, 4.0, 4.0, True,
            True)
    fb = (dens, 8, 8, 0.5, 0.5, 20.0, 2.0, False, 0.1)
    w = g.uniform(0.0, 30.0, (3, 5))  # another shape than the feedback grid
    if form == "plain":
        no = ok.apply_catalytic_reaction(*args[:1], so, *args[2:])
        ng = gd.apply_catalytic_reaction(*args[:1], sg, *args[2:])
    elif form == "feedback":
        no = ok.apply_catalytic_with_feedback(*args[:1], so, *args[2:], *fb)
        ng = gd.apply_catalytic_with_feedback(*args[:1], sg, *args[2:], *fb)
    else:
        no = ok.apply_catalytic_with_feedback_and_adaptation(*args[:1], so, *args[2:], *fb, w,
                                                             10.0, 2.0)
        ng = gd.apply_catalytic_with_feedback_and_adaptation(*args[:1], sg, *args[2:], *fb, w,
                                                             10.0, 2.0)
    assert ng == no > 20
    np.testing.assert_array_equal(sg, so)


@pytest.mark.parametrize("form", ["plain", "feedback", "adaptation"])
def test_autocatalysis_matches_oracle(oracle_lib, form):
    """enzyme == product: conversions cascade in the order of the given pair
    list (shuffled here), as the oracle's sequential loop leaves them."""
    g, pos, sp, alive, pairs, dens = _reaction_inputs(11)
    pairs = pairs[g.permutation(len(pairs))]
    u = g.random(len(pairs))
    fb = (dens, 8, 8, 0.5, 0.5, 20.0, 2.0, False, 0.1)
    w = g.uniform(0.0, 30.0, (3, 5))
    out = []
    for prod in (1, 3):  # 1: autocatalytic; 3: a product that is no enzyme
        so, sg = sp.copy(), sp.copy()
        args = (alive, 5000, 1, 0, prod, 0.15, 0.6, pairs, len(pairs), u, 4.0, 4.0, True, True)
        if form == "plain":
            no = ok.apply_catalytic_reaction(pos, so, *args)
            ng = gd.apply_catalytic_reaction(pos, sg, *args)
        elif form == "feedback":
            no = ok.apply_catalytic_with_feedback(pos, so, *args, *fb)
            ng = gd.apply_catalytic_with_feedback(pos, sg, *args, *fb)
        else:
            no = ok.apply_catalytic_with_feedback_and_adaptation(pos, so, *args, *fb, w, 10.0,
                                                                 2.0)
            ng = gd.apply_catalytic_with_feedback_and_adaptation(pos, sg, *args, *fb, w, 10.0,
                                                                 2.0)
        assert ng == no
        np.testing.assert_array_equal(sg, so)
        out.append(no)
    auto, direct = out
    assert auto > direct > 5, "the autocatalytic run must convert through cascades"


@pytest.mark.parametrize("form", ["plain", "feedback", "dual", "local"])
def test_recruitment_matches_oracle(oracle_lib, form):
    g, pos, sp, alive, pairs, dens = _reaction_inputs(8)
    dens2 = g.uniform(0.0, 40.0, (10, 10))
    ua, up = g.random(5000), g.random((5000, 2))
    ul, uo = g.random(5000), g.standard_normal((5000, 2))
    cdf = np.cumsum(np.maximum(dens.ravel(), 0.01))
    cdf /= cdf[-1]
    fb1 = (dens, 8, 8, 0.5, 0.5, 20.0, 2.0, False, 0.3)
    fb2 = (dens2, 10, 10, 0.4, 0.4, 15.0, 4.0, True, 0.0)
    out = []
    for K in (ok, gd):
        a, p = alive.copy(), pos.copy()
        if form == "plain":
            n = K.apply_membrane_recruitment(sp, a, p, 5000, 3, 0.7, 4.0, 4.0, ua, up, 60)
        elif form == "feedback":
            n = K.apply_membrane_recruitment_with_feedback(sp, a, p, 5000, 3, 0.7, 4.0, 4.0, ua,
                                                           up, 60, *fb1)
        elif form == "dual":
            n = K.apply_membrane_recruitment_with_dual_feedback(sp, a, p, 5000, 3, 0.9, 4.0, 4.0,
                                                                ua, up, 60, *fb1, *fb2)
        else:
            n = K.apply_membrane_recruitment_local(sp, a, p, 5000, 3, 0.7, 4.0, 4.0, ua, uo, 60,
                                                   *fb1, cdf, ul, 0.3)
        out.append((n, a, p))
    (no, ao, po), (ng, ag, pg) = out
    assert ng == no and no > 5
    np.testing.assert_array_equal(ag, ao)
    np.testing.assert_array_equal(pg, po)


@pytest.mark.parametrize("kind", ["detach", "detach_fb", "convert", "convert_fb"])
def test_detachment_and_conversion_match_oracle(oracle_lib, kind):
    g, pos, sp, alive, pairs, dens = _reaction_inputs(9)
    u = g.random(5000)
    fb = (dens, 8, 8, 0.5, 0.5, 20.0, 2.0, True, 0.0)
    out = []
    for K in (ok, gd):
        s, a = sp.copy(), alive.copy()
        if kind == "detach":
            n = K.apply_membrane_detachment(s, a, 5000, 2, 0.3, u)
        elif kind == "detach_fb":
            n = K.apply_membrane_detachment_with_feedback(s, a, pos, 5000, 2, 0.3, u, *fb)
        elif kind == "convert":
            n = K.apply_species_conversion(s, a, 5000, 3, 4, 0.3, u)
        else:
            n = K.apply_species_conversion_with_feedback(s, a, pos, 5000, 3, 4, 0.3, u, *fb)
        out.append((n, s, a))
    assert out[0][0] == out[1][0] > 20
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])


def test_density_and_adaptation_match_oracle(oracle_lib):
    g, pos, sp, alive = _system(10, n=20000, L=10.0)
    for nxg, s in ((20, 0.5), (7, 10.0 / 7)):  # power-of-two and other cell areas
        do, dg = np.zeros((nxg, nxg)), np.zeros((nxg, nxg))
        ok.compute_density_field(pos, sp, alive, 20000, 1, nxg, nxg, s, s, do)
        gd.compute_density_field(pos, sp, alive, 20000, 1, nxg, nxg, s, s, dg)
        np.testing.assert_array_equal(dg, do)
    wo, wg = g.uniform(0, 5, (7, 7)), None
    wg = wo.copy()
    ok.update_adaptation_field(wo, do, 7, 7, 1e-3, 0.5)
    gd.update_adaptation_field(wg, do, 7, 7, 1e-3, 0.5)
    np.testing.assert_array_equal(wg, wo)


def test_device_resident_arrays_in_place(oracle_lib):
    """The kernels on DeviceArray arguments: mutated in place on the
    device, no host copy, over several calls."""
    DA = gd.DeviceArray
    g, pos, sp, alive = _system(11, n=3000, L=3.0)
    D = np.full(3000, 0.1)
    F = np.zeros((3000, 2))
    gD = np.zeros((3000, 2))
    dp, da = DA(pos), DA(alive)
    dD, dF, dg, dsp = DA(D), DA(F), DA(gD), DA(sp)
    want = pos.copy()
    for k in range(3):
        noise = g.standard_normal((3000, 2))
        gd.bd_step(dp, dD, dF, dg, 1e-3, 3.0, 3.0, True, True, 3000, DA(noise), da, 0.15)
        ok.bd_step(want, D, F, gD, 1e-3, 3.0, 3.0, True, True, 3000, noise, alive, 0.15)
    np.testing.assert_array_equal(dp.numpy(), want)
    dens = DA(np.zeros((6, 6)))
    gd.compute_density_field(dp, dsp, da, 3000, 2, 6, 6, 0.5, 0.5, dens)
    dw = np.zeros((6, 6))
    ok.compute_density_field(want, sp, alive, 3000, 2, 6, 6, 0.5, 0.5, dw)
    np.testing.assert_array_equal(dens.numpy(), dw)
    assert da.dtype == np.uint8 and dsp.dtype == np.int32
