"""Inputs and expected values transcribed from the reference crate's own tests
(paths relative to mesh_to_sdf/src/). Shared by the oracle and GPU suites."""
import math

import numpy as np

# lib.rs:491-492
TETRA_INDICES = np.array([0, 1, 2, 0, 2, 3, 0, 3, 1, 1, 3, 2], dtype=np.uint32)
TETRA_VERTICES = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, -1.0]], dtype=np.float32)

_f = np.float32
# lib.rs:500: (0.5 / FRAC_1_SQRT_2.hypot(1.0)) * FRAC_1_SQRT_2, evaluated in f32
_point_five = _f(_f(0.5) / _f(math.hypot(_f(1 / math.sqrt(2)), 1.0))) * _f(1 / math.sqrt(2))

# lib.rs:501-516: (query, expected sdf)
TETRA_POINTS = [
    ([0.0, 0.0, 0.0], 0.0),
    ([0.5, 0.0, 0.0], 0.0),
    ([0.2, 0.2, 0.0], 0.0),
    ([0.2, 0.2, 0.05], 0.05),
    ([0.2, 0.2, -0.05], -0.05),
    ([1.0, 1.0, 0.05], float(np.sqrt(_f(0.05 * 0.05 + 0.5 * 0.5 + 0.5 * 0.5)))),
    ([1.0, 1.0, -0.05], float(np.sqrt(_f(0.05 * 0.05 + 0.5 * 0.5 + 0.5 * 0.5)))),
    ([1.0, 1.0, 0.0], math.hypot(0.5, 0.5)),
    ([0.0, 0.0, -1.05], 0.05),
    ([0.0, 0.0, -0.95], 0.0),
    ([0.5, 0.5, -0.5], float(_point_five)),
    ([0.5 - 1e-6, 0.5, -0.5], float(_point_five)),
    ([0.5, 0.5 - 1e-6, -0.5], float(_point_five)),
    ([0.5, 0.5, -0.5 - 1e-6], float(_point_five)),
]

# lib.rs:594-617: unit cube, outward faces
CUBE_VERTICES = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 0.0],
                          [1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [1.0, 1.0, 1.0]], dtype=np.float32)
CUBE_INDICES = np.array([0, 2, 4, 0, 4, 1, 3, 5, 7, 3, 7, 6, 0, 1, 5, 0, 5, 3, 2, 6, 7, 2, 7, 4, 0, 3, 6, 0, 6, 2,
                         1, 4, 7, 1, 7, 5], dtype=np.uint32)
# lib.rs:620-629 (expected -0.2) and lib.rs:652-662 (expected sqrt(0.12))
CUBE_INSIDE = np.array([[0.2, 0.2, 0.2], [0.8, 0.2, 0.2], [0.2, 0.8, 0.2], [0.2, 0.2, 0.8], [0.8, 0.8, 0.2],
                        [0.8, 0.2, 0.8], [0.2, 0.8, 0.8], [0.8, 0.8, 0.8]], dtype=np.float32)
CUBE_OUTSIDE = np.array([[-0.2, -0.2, -0.2], [1.2, -0.2, -0.2], [-0.2, 1.2, -0.2], [-0.2, -0.2, 1.2],
                         [1.2, 1.2, -0.2], [1.2, -0.2, 1.2], [-0.2, 1.2, 1.2], [1.2, 1.2, 1.2]], dtype=np.float32)
CUBE_OUTSIDE_EXPECTED = float(np.float32(math.sqrt((0.2 * 0.2) * 3.0)))

# generic/rtree_bvh.rs:385-393: (query, expected closest - query, expected inside)
CLOSEST_VECTOR_POINTS = [
    ([0.0, 0.0, 0.0], [0.0, 0.0, 0.0], False),
    ([0.0, 0.5, 0.0], [0.0, 0.0, 0.0], False),
    ([0.2, 0.2, 0.0], [0.0, 0.0, 0.0], False),
    ([-0.1, -0.1, 0.2], [0.1, 0.1, -0.2], False),
    ([0.1, -0.1, 0.2], [0.0, 0.1, -0.2], False),
    ([0.2, 0.2, 0.2], [0.0, 0.0, -0.2], False),
    ([0.2, 0.2, -0.1], [0.0, 0.0, 0.1], True),
]

# lib.rs:13-31 crate doctest: RtreeBvh gives [1.0]
DOC_VERTICES_1 = np.array([[0.5, 1.5, 0.5], [1.0, 2.0, 3.0], [1.0, 3.0, 7.0]], dtype=np.float32)
DOC_QUERY_1 = np.array([[0.5, 0.5, 0.5]], dtype=np.float32)
# lib.rs:342-362 generate_sdf doctest and lib.rs:441-461 query_sdf doctest: [1.0]
DOC_VERTICES_2 = np.array([[0.0, 1.0, 0.0], [1.0, 2.0, 3.0], [1.0, 3.0, 4.0]], dtype=np.float32)
DOC_QUERY_2 = np.array([[0.0, 0.0, 0.0]], dtype=np.float32)
# generate/grid.rs:207-231 doctest: grid 0..10, 10 cells per axis, sdf[0] == 1.0
DOC_VERTICES_GRID = np.array([[0.5, 1.5, 0.5], [1.0, 2.0, 3.0], [1.0, 3.0, 4.0]], dtype=np.float32)

# generate/grid.rs:696-698 (two triangles, 5^3 grid over 0..5)
GRID2_VERTICES = np.array([[0.0, 1.0, 0.0], [1.0, 2.0, 3.0], [1.0, 3.0, 4.0], [2.0, 0.0, 0.0]], dtype=np.float32)
GRID2_INDICES = np.array([0, 1, 2, 1, 2, 3], dtype=np.uint32)

# generic/bvh.rs:346-352 / generic/rtree.rs:170-176 query points on suzanne
SUZANNE_POINTS = np.array([[0.01, 0.01, 0.5], [1.0, 1.0, 1.0], [0.1, 0.2, 0.2], [1.1, 2.2, 5.2], [-0.1, 0.2, -0.2]],
                          dtype=np.float32)
# default.rs:120-124 with the pysdf / mesh_to_sdf baseline of default.rs:136-138 (+-0.1)
SUZANNE_BASELINE_POINTS = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.1, 0.2, 0.2]], dtype=np.float32)
SUZANNE_BASELINE = [-0.42, 0.69, -0.46]

# proptest-regressions/geo.txt:7-8
PROPTEST_REGRESSIONS = [
    ([0.0, -8.055119, 1.1846914], [0.0, 0.0, 0.0], [0.0, 0.0, 8.367966], [-7.806354, 9.330519, 0.0]),
    ([0.0, -5.8359632, 4.405388], [0.0, 0.9572999, 9.758267], [6.9999175, -4.739112, 7.5462694],
     [0.0, -9.673183, 0.52112055]),
]


def rust_debug_f32(x) -> str:
    """Rust's `{:?}` of an f32 (core::fmt::float: shortest round-trip digits;
    exponential when the value is nonzero and below 1e-4 or at least 1e16),
    built from numpy's shortest float32 digits: an independent statement of
    the formatting the library does in C++."""
    x = np.float32(x)
    if np.isnan(x):
        return "NaN"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    sign = "-" if np.signbit(x) else ""
    a = np.float32(abs(x))
    if a == 0:
        return sign + "0.0"
    mant, exp = np.format_float_scientific(a, unique=True, trim="-").split("e")
    digits, e = mant.replace(".", ""), int(exp)
    if a < np.float32(1e-4) or a >= np.float32(1e16):
        return sign + digits[0] + ("." + digits[1:] if len(digits) > 1 else "") + "e" + str(e)
    if e >= 0:
        ip = digits[: e + 1].ljust(e + 1, "0")
        fp = digits[e + 1:] or "0"
        return sign + ip + "." + fp
    return sign + "0." + "0" * (-e - 1) + digits


# Known outputs of Rust's f32 Debug formatting.
RUST_DEBUG_F32 = [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.0, "-1.0"), (0.1, "0.1"), (0.5, "0.5"), (123456.7, "123456.7"),
    (16777217.0, "16777216.0"), (1e-4, "0.0001"), (1e-5, "1e-5"), (1.5e-7, "1.5e-7"), (1e15, "1000000000000000.0"),
    (1e16, "1e16"), (3.4028235e38, "3.4028235e38"), (1e-45, "1e-45"), (float("nan"), "NaN"), (float("inf"), "inf"),
    (float("-inf"), "-inf"), (2.5, "2.5"), (100.0, "100.0"), (0.001, "0.001"), (-0.00012345, "-0.00012345"),
]
