"""Rust-style messages: the reference's panic and Err strings print f32 values
with `{:?}`; the library reproduces that formatting (checked here through the
non-finite query message, lib.rs:374-376, which is raised before any device
work)."""
import numpy as np
import pytest

import refcases as R


@pytest.mark.parametrize("value,text", R.RUST_DEBUG_F32)
def test_python_statement_matches_known_rust_output(value, text):
    assert R.rust_debug_f32(value) == text


@pytest.mark.parametrize("seed", range(4))
def test_library_formatting_matches(seed):
    """The C++ formatter (via the lib.rs:375 message) against the numpy
    statement on random f32 bit patterns of every magnitude."""
    from mesh_to_sdf_amd import _lib, generate_sdf, Topology, AccelerationMethod
    g = np.random.default_rng(seed)
    bits = g.integers(0, 2 ** 32, size=300, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = np.concatenate([vals, np.float32([v for v, _ in R.RUST_DEBUG_F32])])
    for k in range(0, len(vals) - 1, 2):
        q = np.array([[vals[k], np.nan, vals[k + 1]]], np.float32)
        with pytest.raises(_lib.SdfxError) as e:
            generate_sdf(R.TETRA_VERTICES, Topology.TriangleList(R.TETRA_INDICES), q, AccelerationMethod.RtreeBvh)
        assert e.value.status == _lib.ERR_NON_FINITE_QUERY
        want = f"Query point [{R.rust_debug_f32(vals[k])}, NaN, {R.rust_debug_f32(vals[k + 1])}] " \
               "contains non-finite values"
        assert e.value.message == want
