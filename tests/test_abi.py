"""The C ABI of libedm_hip.so, checked without a GPU.

* every entry point that include/edm_hip.h declares is exported by the
  built library and declared in the ctypes binding;
* the ctypes structures have the C layout (sizes and field offsets from a C
  program compiled against the header with gcc);
* without a GPU the product path refuses to run instead of falling back.
"""

from __future__ import annotations

import ctypes
import os
import re
import subprocess
import sys

import pytest

from conftest import PKG, ROOT

HEADER = os.path.join(ROOT, "include", "edm_hip.h")


def _declared():
    text = open(HEADER, encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def hip():
    from ed_mrdt import _hip
    if not os.path.exists(_hip.LIB_PATH):
        sys.path.insert(0, PKG)
        import build_hip
        build_hip.build(jobs=3)
    return _hip


def test_header_declares_the_boundary():
    names = _declared()
    for n in ("edm_create", "edm_step_begin", "edm_step_finish", "edm_run", "edm_photons",
              "edm_photons_image", "edm_image_assign", "edm_frame_noise", "edm_counters"):
        assert n in names


def test_library_exports_every_declared_symbol(hip):
    L = hip.load_library()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, f"declared in edm_hip.h but not exported: {missing}"
    unbound = [n for n in _declared() if n not in hip.EXPORTS]
    assert not unbound, f"declared in edm_hip.h but not bound in _hip.py: {unbound}"


def _prototypes():
    """name -> parameter count of every function the header declares."""
    text = open(HEADER, encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(edm_[a-z0-9_]+)\s*\(([^()]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_ctypes_argument_counts_match_the_header(hip):
    protos = _prototypes()
    assert len(protos) >= 60
    bad = {n: (k, len(hip._SIGS[n])) for n, k in protos.items()
           if n in hip._SIGS and len(hip._SIGS[n]) != k}
    assert not bad, f"header vs ctypes parameter counts: {bad}"


_STRUCTS = {
    "edm_feedback": ("Feedback", ["sensor", "grid", "K_half", "basal"]),
    "edm_grid": ("Grid", ["nx", "sx", "sy"]),
    "edm_reaction": ("Reaction", ["kind", "contact", "prob2", "max_recruits", "use_adapt",
                                  "fb1", "fb2"]),
    "edm_params": ("Params", ["pair_cap", "exact_force_order", "Lx", "adapt_grid", "seed",
                              "sigma", "reactions"]),
    "edm_host_state": ("HostState", []),
    "edm_photo_params": ("PhotoParams", []),
    "edm_image_params": ("ImageParams", []),
    "edm_census": ("Census", ["n_bond_alive", "n_free_dye", "species_alive", "events",
                              "photons", "e_sum"]),
}


def test_ctypes_layout_matches_c(hip, tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "edm_hip.h"',
             "int main(void) {"]
    for cname, (_, fields) in _STRUCTS.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    c = dict(line.split() for line in out.strip().splitlines())
    for cname, (pyname, fields) in _STRUCTS.items():
        S = getattr(hip, pyname)
        assert ctypes.sizeof(S) == int(c[cname]), cname
        for f in fields:
            assert getattr(S, f).offset == int(c[f"{cname}.{f}"]), f"{cname}.{f}"


def test_census_numpy_view_matches_the_struct(hip):
    """The recorder reads rows through a numpy dtype: same size and offsets
    as the ctypes (and so the C) layout."""
    dt = hip.CENSUS_DTYPE
    assert dt.itemsize == ctypes.sizeof(hip.Census)
    for name, _ in hip.Census._fields_:
        assert dt.fields[name][1] == getattr(hip.Census, name).offset, name


def test_no_cpu_fallback_without_a_gpu(hip):
    L = hip.load_library()
    n = ctypes.c_int(-1)
    rc = L.edm_device_count(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is visible here")
    with pytest.raises(hip.HipError, match="no HIP device"):
        hip.lib()
    from ed_mrdt import workloads
    from ed_mrdt.dynamics import DynamicsEngine
    from ed_mrdt.particles import ParticleState
    cfg = workloads.binding_fret()
    with pytest.raises(hip.HipError):
        DynamicsEngine(cfg, ParticleState(cfg))
