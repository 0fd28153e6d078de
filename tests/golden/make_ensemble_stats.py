"""Writes tests/golden/oracle_ensemble_stats.npz: the oracle's long-run
ensemble observables (tests/ensemble_stats.py) for every case, one row per
replica.  The oracle (FP64, the reference's numpy draws) is too slow to run
32 x 1000 steps of these systems inside a GPU test, so its ensembles are
computed once here and the GPU test compares against them.

    python tests/golden/make_ensemble_stats.py [--jobs 8]
"""

from __future__ import annotations

import argparse
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "flim-microscopy-simulator-with-causal-inference_amd"),
          os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import ensemble_stats as es  # noqa: E402

OUT = os.path.join(HERE, "oracle_ensemble_stats.npz")


def _one(args):
    case, seed = args
    return case, seed, es.oracle_replica(case, seed)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    import build_oracle
    build_oracle.build()
    tasks = [(c, s) for c in es.CASES for s in range(es.REPLICAS)]
    rows = {c: [None] * es.REPLICAS for c in es.CASES}
    t0 = time.time()
    with ProcessPoolExecutor(max_workers=a.jobs) as ex:
        for case, seed, r in ex.map(_one, tasks):
            rows[case][seed] = r
            print(f"{case} seed {seed} done at {time.time() - t0:.0f} s", flush=True)
    data = {c: np.stack(v) for c, v in rows.items()}
    for c in es.CASES:
        assert data[c].shape == (es.REPLICAS, len(es.observable_names(c)))
    np.savez_compressed(OUT, steps=es.STEPS, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
