"""The on-disk format of serde.rs:77-221, pinned by the reference's own V1
fixtures (tests/sdf_generic_v1.bin, tests/sdf_grid_v1.bin, read back in
serde.rs:315-374) and cross-checked with an independent MessagePack decoder."""
import os

import numpy as np
import pytest

import mesh_to_sdf_amd as m
from mesh_to_sdf_amd import serde as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _fixture(name):
    with open(os.path.join(GOLDEN, name), "rb") as fh:
        return fh.read()


def test_read_generic_v1_fixture():
    """serde.rs:315-342: queries (1,2,3), (6,5,4) and distances 1, 3."""
    de = S.read_from_file(os.path.join(GOLDEN, "sdf_generic_v1.bin"))
    assert isinstance(de, S.DeserializeSdf.Generic)
    np.testing.assert_array_equal(de.query_points, np.array([[1, 2, 3], [6, 5, 4]], np.float32))
    np.testing.assert_array_equal(de.distances, np.array([1, 3], np.float32))


def test_read_grid_v1_fixture():
    """serde.rs:344-374: Grid::new([1,2,3], [4,5,6], [7,8,9]), distances 0..504."""
    de = S.read_from_file(os.path.join(GOLDEN, "sdf_grid_v1.bin"))
    assert isinstance(de, S.DeserializeSdf.Grid)
    assert de.grid == m.Grid([1, 2, 3], [4, 5, 6], [7, 8, 9])
    np.testing.assert_array_equal(de.distances, np.arange(7 * 8 * 9, dtype=np.float32))


def test_write_reproduces_fixture_bytes():
    gen = S.SerializeSdf.Generic(np.array([[1, 2, 3], [6, 5, 4]], np.float32), np.array([1, 3], np.float32))
    assert S.serialize(gen) == _fixture("sdf_generic_v1.bin")
    grid = m.Grid([1, 2, 3], [4, 5, 6], [7, 8, 9])
    ser = S.SerializeSdf.Grid(grid, np.arange(grid.get_total_cell_count(), dtype=np.float32))
    assert S.serialize(ser) == _fixture("sdf_grid_v1.bin")


@pytest.mark.parametrize("n", [0, 1, 15, 16, 65535, 65536, 100003])
def test_round_trip_and_msgpack_agree(n, tmp_path):
    g = np.random.default_rng(n)
    q = g.normal(size=(n, 3)).astype(np.float32)
    d = g.normal(size=n).astype(np.float32)
    d[: min(n, 3)] = [np.float32(-0.0), np.float32(np.inf), np.float32(np.finfo(np.float32).max)][: min(n, 3)]
    path = tmp_path / "sdf.bin"
    S.save_to_file(S.SerializeSdf.Generic(q, d), path)
    de = S.read_from_file(path)
    assert de.query_points.view(np.uint32).tobytes() == q.view(np.uint32).tobytes()
    assert de.distances.view(np.uint32).tobytes() == d.view(np.uint32).tobytes()
    msgpack = pytest.importorskip("msgpack")
    obj = msgpack.unpackb(path.read_bytes(), raw=False)
    (payload,) = obj["V1"]["Generic"][1:]
    assert len(obj["V1"]["Generic"][0]) == n and len(payload) == n
    # the reference's encoder (rmp_serde): the same bytes as msgpack's float32 packing
    ref = msgpack.packb({"V1": {"Generic": [q.tolist(), d.tolist()]}}, use_single_float=True)
    assert ref == path.read_bytes()


@pytest.mark.parametrize("count", [(7, 8, 9), (1, 200, 3), (300, 70000, 1)])
def test_grid_counts_encode_like_rmp(count):
    msgpack = pytest.importorskip("msgpack")
    grid = m.Grid([0.5, -1, 2], [0.25, 0.125, 3], count)
    d = np.zeros(5, np.float32)
    data = S.serialize(S.SerializeSdf.Grid(grid, d))
    ref = msgpack.packb({"V1": {"Grid": [[[0.5, -1.0, 2.0], [0.25, 0.125, 3.0], list(count)], d.tolist()]}},
                        use_single_float=True)
    assert data == ref
    de = S.deserialize(data)
    assert de.grid == grid


def test_reader_accepts_other_number_widths():
    msgpack = pytest.importorskip("msgpack")
    data = msgpack.packb({"V1": {"Grid": [[[1.0, 2.0, 3.0], [4, 5, 6], [7, 8, 9]], [1.5, 2, -3]]}})  # f64 and ints
    de = S.deserialize(data)
    assert de.grid == m.Grid([1, 2, 3], [4, 5, 6], [7, 8, 9])
    np.testing.assert_array_equal(de.distances, np.array([1.5, 2, -3], np.float32))


def test_errors():
    with pytest.raises(S.SerdeError) as e:
        S.deserialize(_fixture("sdf_grid_v1.bin")[:100])
    assert e.value.kind == S.SerdeError.DESERIALIZATION_FAILED
    with pytest.raises(S.SerdeError) as e:
        S.deserialize(b"\x81\xa2V2\x81\xa4Grid\x92")
    assert e.value.kind == S.SerdeError.DESERIALIZATION_FAILED
    with pytest.raises(S.SerdeError) as e:
        S.read_from_file(os.path.join(GOLDEN, "does_not_exist.bin"))
    assert e.value.kind == S.SerdeError.IO_ERROR
    with pytest.raises(S.SerdeError) as e:
        S.serialize("not an sdf")
    assert e.value.kind == S.SerdeError.SERIALIZATION_FAILED
