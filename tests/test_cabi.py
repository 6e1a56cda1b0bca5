"""The C-ABI library loads and exports every entry point include/suli.h
declares; host-side argument checks fail cleanly (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

import pysuli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "suli.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|const char \*)\s*\*?\s*(suli_\w+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = pysuli.lib()
    names = _declared()
    assert len(names) >= 20, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_uninitialised_errors():
    L = pysuli.lib()
    assert b"gfx950" in L.suli_version()
    # every state call before suli_init reports an error instead of crashing
    assert L.suli_visco() != 0
    buf = C.create_string_buffer(256)
    L.suli_last_error(buf, 256)
    assert b"not initialised" in buf.value


def test_params_layout_matches_header():
    # the ctypes mirror of suli_params must have the header's field order
    txt = open(os.path.join(ROOT, "include", "suli.h")).read()
    body = txt[txt.index("typedef struct {"):txt.index("} suli_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(?:int32_t|double)\s+([^;]+);", body)
    names = [n.strip() for grp in fields for n in grp.split(",")]
    assert names == [f for f, _ in pysuli.Params._fields_]


def test_bad_grid_rejected():
    p = pysuli.Params()
    p.nx, p.ny, p.nz = 2, 8, 8
    assert pysuli.lib().suli_init(C.byref(p)) != 0


def test_shipped_learned_weights():
    """Every shipped network has the parameter count of its architecture, one
    of them is the default and comes first, and its grid constraint fits the
    benchmark grid."""
    import numpy as np
    import unet_ref as U

    shipped = pysuli.shipped_learned()
    assert shipped and shipped[0]["default"] and sum(m["default"] for m in shipped) == 1
    for m in shipped:
        w = np.load(m["path"])
        assert w.dtype == np.float32 and w.size == U.nparams(m["levels"]), m["name"]
        assert np.all(np.isfinite(w))
        assert 0 <= m["entry"] <= 4 and 2 <= m["levels"] <= 6
    d = shipped[0]
    assert 128 % (1 << (d["entry"] + d["levels"] - 1)) == 0
