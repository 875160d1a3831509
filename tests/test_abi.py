"""The C-ABI library loads and exports every entry point include/sdfx.h
declares, with the declared status values. No compute call (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "sdfx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sdfx_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("sdfx_generate_sdf", "sdfx_generate_grid_sdf", "sdfx_build_sdf_acceleration", "sdfx_query_sdf",
                     "sdfx_query_vector_to_closest_point", "sdfx_free_sdf_acceleration", "sdfx_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from mesh_to_sdf_amd import _lib
    L = _lib.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert set(declared_functions()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with sdfx.h"


def test_version_and_error_state():
    from mesh_to_sdf_amd import _lib
    L = _lib.lib()
    assert L.sdfx_version() == 1
    assert L.sdfx_last_error() == b""


def test_status_values_match_header_and_oracle():
    import oracle as O
    from mesh_to_sdf_amd import _lib
    text = open(HEADER).read()
    values = dict((k, int(v)) for k, v in re.findall(r"(SDFX_(?:OK|EMPTY_RESULT|ERR_[A-Z_]+))\s*=\s*(-?\d+)", text))
    for name, v in values.items():
        short = name[len("SDFX_"):]
        assert getattr(_lib, short) == v, name
        if hasattr(O, short):
            assert getattr(O, short) == v, name


def test_no_oracle_in_product():
    """The product package never imports or loads the oracle."""
    pkg = os.path.join(ROOT, "mesh_to_sdf_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")) or f == "Makefile":
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "liboracle" not in text and "sdf_oracle" not in text and "import oracle" not in text, f


def test_tuning_knobs_and_ranges():
    """Every knob reads back what was set, out-of-range values and unknown
    keys are refused (no GPU needed: the knobs are host state)."""
    from mesh_to_sdf_amd import _lib
    L = _lib.lib()
    for key, good, bad in (("exhaustive_max_pairs", 12345, None), ("grid_prop_blocks", 512, 4),
                           ("lbvh_visit_budget", 64, 15), ("grid_batch_prop_blocks", 1024, 65537),
                           ("grid_item_major_min", 4096, -1)):
        old = L.sdfx_get_tuning(key.encode())
        assert old >= 0 or key == "exhaustive_max_pairs"
        assert L.sdfx_set_tuning(key.encode(), good) == 0 and L.sdfx_get_tuning(key.encode()) == good
        if bad is not None:
            assert L.sdfx_set_tuning(key.encode(), bad) == -1 and L.sdfx_get_tuning(key.encode()) == good
        assert L.sdfx_set_tuning(key.encode(), old) == 0
    assert L.sdfx_set_tuning(b"lbvh_visit_budget", 1 << 31) == -1
    old = L.sdfx_get_tuning(b"lbvh_visit_budget")
    assert L.sdfx_set_tuning(b"lbvh_visit_budget", 0) == 0  # 0: chosen per mesh from its size
    assert L.sdfx_set_tuning(b"lbvh_visit_budget", old) == 0
    assert L.sdfx_set_tuning(b"no_such_knob", 1) == -1 and L.sdfx_get_tuning(b"no_such_knob") == -1
