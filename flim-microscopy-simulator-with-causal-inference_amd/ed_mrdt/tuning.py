"""Autotuning of the kernels' work decomposition.

The knobs -- how the forces of dense membranes find their candidates
(sub-row lists or LDS strips), how catalysis walks its pairs (work lists of
its enzymes or substrates, or LDS strips of cells), from which side
(enzymes, substrates, or per run the side with fewer alive particles), with
how many lanes per entity, and the cell rows per strip of the tiled force
and catalysis kernels -- leave
every result bit-identical, so the fastest setting can be found by timing.
`autotune` builds a scratch engine from the configuration (its own initial
state, Philox draws), times each candidate over a few steps and returns the
fastest settings; `DynamicsEngine.set_tuning(**settings)` applies them.

`autotune_ensemble` does the same for the batched step of a set of
replicas: the own particles per strip workgroup, the column population
below which two columns share a strip, and the lanes per catalysis entity
(`EnsembleEngine.set_tuning(**settings)`).  It times the candidates on the
live engine it is given (the settings leave results unchanged, so the
timed steps are ordinary steps of the run), or on a scratch engine built
from the replicas' configurations.
"""

from __future__ import annotations

import ctypes
import time
from typing import Dict, Iterable

from . import _hip
from .config import SimulationConfig

TILE_ROWS = (16, 32, 64, 128, 256)
CAT_SIDES = (-1, 0, 1)
CAT_WALKS = (1, 0)
CAT_LANES = (64, 16, 4, 1)
FORCE_LISTS = (1, 0)


def autotune(config: SimulationConfig, precision: str = "fp32", steps: int = 20,
             warmup: int = 3, tile_rows: Iterable[int] = TILE_ROWS,
             cat_sides: Iterable[int] = CAT_SIDES,
             cat_walks: Iterable[int] = CAT_WALKS,
             cat_lanes: Iterable[int] = CAT_LANES,
             force_lists: Iterable[int] = FORCE_LISTS) -> Dict:
    """Fastest force_lists, (cat_walk, cat_lanes), cat_side and tile_rows
    for this configuration (tuned one after another in that order, each with
    the best of the earlier ones); `times_ms` holds the measured
    milliseconds per step of every candidate."""
    from .dynamics import DynamicsEngine
    from .particles import ParticleState

    L = _hip.lib()
    st = ParticleState(config)
    eng = DynamicsEngine(config, st, precision=precision, rng="philox", pair_cap=False)
    tot, fms = ctypes.c_double(0), ctypes.c_double(0)

    def per_step() -> float:
        _hip.check(L.edm_time_steps(eng.handle, int(warmup), 0, ctypes.byref(tot),
                                    ctypes.byref(fms), None), "edm_time_steps")
        _hip.check(L.edm_time_steps(eng.handle, int(steps), 0, ctypes.byref(tot),
                                    ctypes.byref(fms), None), "edm_time_steps")
        return tot.value / steps

    times: Dict[str, float] = {}
    try:
        lists = list(force_lists)
        for v in lists:
            eng.set_tuning(force_lists=v)
            times[f"force_lists={v}"] = per_step()
        best_lists = min(lists, key=lambda v: times[f"force_lists={v}"])
        eng.set_tuning(force_lists=best_lists)
        # the lanes mean different things to the two walks: tuned together
        walks, lanes = list(cat_walks), list(cat_lanes)
        for w in walks:
            for g in lanes:
                eng.set_tuning(cat_walk=w, cat_lanes=g)
                times[f"cat_walk={w},cat_lanes={g}"] = per_step()
        best_walk, best_lanes = min(((w, g) for w in walks for g in lanes),
                                    key=lambda wg: times[f"cat_walk={wg[0]},cat_lanes={wg[1]}"])
        eng.set_tuning(cat_walk=best_walk, cat_lanes=best_lanes)
        sides = list(cat_sides)
        for s in sides:
            eng.set_tuning(cat_side=s)
            times[f"cat_side={s}"] = per_step()
        best_side = min(sides, key=lambda s: times[f"cat_side={s}"])
        eng.set_tuning(cat_side=best_side)
        rows = list(tile_rows)
        for r in rows:
            eng.set_tuning(tile_rows=r)
            times[f"tile_rows={r}"] = per_step()
        best_rows = min(rows, key=lambda r: times[f"tile_rows={r}"])
    finally:
        st.mark_stale()
        eng.close()
    return dict(force_lists=best_lists, tile_rows=best_rows, cat_side=best_side,
                cat_walk=best_walk, cat_lanes=best_lanes, times_ms=times)


ENS_STRIP_PARTICLES = (96, 128, 192, 256, 384)
ENS_PAIR_BELOW = (0, 64, 128, 256)
ENS_CAT_LANES = (64, 16, 4, 1)
ENS_CAT_WALK = (0, 1)


def autotune_ensemble(engine=None, configs=None, steps: int = 10, warmup: int = 2,
                      strip_particles: Iterable[int] = ENS_STRIP_PARTICLES,
                      pair_below: Iterable[int] = ENS_PAIR_BELOW,
                      cat_lanes: Iterable[int] = ENS_CAT_LANES,
                      cat_walk: Iterable[int] = ENS_CAT_WALK) -> Dict:
    """Fastest (cat_walk, cat_lanes), pair_below and strip_particles of the
    batched step (tuned one after another in that order, each with the best
    of the earlier ones).  With `engine` the candidates are timed on it (its replicas
    advance by the timed steps, with results independent of the settings)
    and the best settings are left applied; with `configs` a scratch
    ensemble of those replicas is built and closed.  `times_ms` holds the
    milliseconds per step of every candidate."""
    from .ensemble import EnsembleEngine
    from .particles import ParticleState

    own = engine is None
    if own:
        if not configs:
            raise ValueError("autotune_ensemble needs an engine or replica configurations")
        states = [ParticleState(c) for c in configs]
        cap = sum(s.N for s in states) + 1024
        engine = EnsembleEngine(configs[0], cap, max_replicas=len(configs))
        for c, s in zip(configs, states):
            engine.add(c, s)

    def per_step() -> float:
        engine.run(int(warmup))
        engine.synchronize()
        t0 = time.perf_counter()
        engine.run(int(steps))
        engine.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    times: Dict[str, float] = {}
    try:
        best = {}
        # the lanes mean different things to the two walks: tuned together
        walks, lanes = list(cat_walk), list(cat_lanes)
        for w in walks:
            for g in lanes:
                engine.set_tuning(cat_walk=w, cat_lanes=g)
                times[f"cat_walk={w},cat_lanes={g}"] = per_step()
        best["cat_walk"], best["cat_lanes"] = min(
            ((w, g) for w in walks for g in lanes),
            key=lambda wg: times[f"cat_walk={wg[0]},cat_lanes={wg[1]}"])
        engine.set_tuning(cat_walk=best["cat_walk"], cat_lanes=best["cat_lanes"])
        for name, cands in (("pair_below", list(pair_below)),
                            ("strip_particles", list(strip_particles))):
            for v in cands:
                engine.set_tuning(**{name: v})
                times[f"{name}={v}"] = per_step()
            best[name] = min(cands, key=lambda v: times[f"{name}={v}"])
            engine.set_tuning(**{name: best[name]})
    finally:
        if own:
            engine.close()
    return dict(best, times_ms=times)
