"""CPU checks of the C-ABI library: it loads, exports every function that
include/*.h declares, its structure layouts equal the C header's (measured
by a gcc-compiled probe), and the host-side helpers agree with the oracle.
No device compute is called."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_functions():
    names = []
    for h in ("fullmag_fdm.h", "fullmag_fdm_ext.h"):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(fullmag_fdm_\w+)\s*\(", text)
    return sorted(set(n for n in names if not n.endswith("_fn")))


def test_library_exports_every_declared_symbol():
    from fullmag_amd import _abi
    lib = _abi.load()
    declared = _declared_functions()
    assert len(declared) >= 22 + 5
    for name in declared:
        assert hasattr(lib, name), name
    # the 22 reference entry points are all declared in fullmag_fdm.h
    assert set(_abi.REFERENCE_SYMBOLS) <= set(declared)
    assert set(_abi.EXT_SYMBOLS) <= set(declared)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fullmag_fdm_\w+)", out))
    assert set(declared) == exported


def test_soname():
    from fullmag_amd import _abi
    out = subprocess.run(["readelf", "-d", _abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "libfullmag_fdm.so.0" in out


_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "fullmag_fdm.h"
#include "fullmag_fdm_ext.h"
#define F(s, f) printf("%s.%s %zu\n", #s, #f, offsetof(s, f));
int main(void) {
  printf("sizeof.plan %zu\nsizeof.stats %zu\nsizeof.info %zu\nsizeof.snap %zu\n",
         sizeof(fullmag_fdm_plan_desc), sizeof(fullmag_fdm_step_stats),
         sizeof(fullmag_fdm_device_info), sizeof(fullmag_fdm_snapshot_desc));
  printf("sizeof.ext %zu\nsizeof.adj %zu\n", sizeof(fullmag_fdm_plan_ext), sizeof(fullmag_fdm_adjoint_desc));
  FIELDS
  return 0;
}
"""


def test_struct_layout_matches_c_header(tmp_path):
    from fullmag_amd import _abi
    lines = []
    structs = (("fullmag_fdm_plan_desc", _abi.PlanDesc), ("fullmag_fdm_step_stats", _abi.StepStats),
               ("fullmag_fdm_device_info", _abi.DeviceInfo), ("fullmag_fdm_plan_ext", _abi.PlanExt),
               ("fullmag_fdm_adjoint_desc", _abi.AdjointDesc))
    for st, cls in structs:
        for f, _ in cls._fields_:
            lines.append(f"F({st}, {f})")
    src = tmp_path / "probe.c"
    src.write_text(_PROBE.replace("FIELDS", "\n  ".join(lines)))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.split("\n") if l)
    assert int(out["sizeof.plan"]) == ctypes.sizeof(_abi.PlanDesc)
    assert int(out["sizeof.stats"]) == ctypes.sizeof(_abi.StepStats)
    assert int(out["sizeof.info"]) == ctypes.sizeof(_abi.DeviceInfo)
    assert int(out["sizeof.snap"]) == ctypes.sizeof(_abi.SnapshotDesc)
    assert int(out["sizeof.ext"]) == ctypes.sizeof(_abi.PlanExt)
    assert int(out["sizeof.adj"]) == ctypes.sizeof(_abi.AdjointDesc)
    for st, cls in structs:
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_no_device_reports_unavailable_or_available():
    from fullmag_amd import _abi
    assert _abi.load().fullmag_fdm_is_available() in (0, 1)


def test_random_unit_vectors_match_oracle_and_reject_seed_zero():
    from fullmag_amd.backend import random_unit_vectors
    from oracle import oracle as O
    a = random_unit_vectors(42, 5000)
    assert np.array_equal(a, O.random_unit_vectors(42, 5000))
    assert np.abs(np.linalg.norm(a, axis=1) - 1).max() < 1e-15
    with pytest.raises(ValueError):
        random_unit_vectors(0, 3)


def test_hole_mask_matches_planner_rule():
    from fullmag_amd.backend import hole_mask
    from oracle import oracle as O
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")))["planner"]
    g = gold["hole_cells_512"]
    m = hole_mask(512, 512, 1, g["cell"][0], g["cell"][1], g["radius"])
    assert int((m == 0).sum()) == g["inactive"]
    assert np.array_equal(m, O.hole_mask(512, 512, 1, g["cell"][0], g["cell"][1], g["radius"]))
