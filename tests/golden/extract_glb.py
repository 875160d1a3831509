"""Extracts the POSITION / index arrays of a reference test mesh (.glb) into
a small .npz fixture. Data only: the glTF container is parsed as JSON plus a
binary chunk; nothing from the file is executed.

Run in the build container (the reference tree is not on the GPU machines):
    python tests/golden/extract_glb.py /root/reference/mesh_to_sdf/assets/suzanne.glb tests/golden/suzanne.npz
"""
import json
import struct
import sys

import numpy as np

_COMPONENT = {5121: np.uint8, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def read_accessor(js, binchunk, idx):
    acc = js["accessors"][idx]
    view = js["bufferViews"][acc["bufferView"]]
    dtype = np.dtype(_COMPONENT[acc["componentType"]])
    ncomp = _NCOMP[acc["type"]]
    offset = view.get("byteOffset", 0) + acc.get("byteOffset", 0)
    stride = view.get("byteStride", ncomp * dtype.itemsize)
    count = acc["count"]
    rows = [np.frombuffer(binchunk, dtype=dtype, count=ncomp, offset=offset + i * stride) for i in range(count)]
    return np.stack(rows).reshape(count, ncomp)


def extract(path, mesh_index=0, primitive=0):
    data = open(path, "rb").read()
    magic, _, _ = struct.unpack("<III", data[:12])
    assert magic == 0x46546C67, "not a glb file"
    jlen, _ = struct.unpack("<II", data[12:20])
    js = json.loads(data[20:20 + jlen])
    off = 20 + jlen
    blen, _ = struct.unpack("<II", data[off:off + 8])
    binchunk = data[off + 8:off + 8 + blen]
    prim = js["meshes"][mesh_index]["primitives"][primitive]
    pos = read_accessor(js, binchunk, prim["attributes"]["POSITION"]).astype(np.float32)
    idx = read_accessor(js, binchunk, prim["indices"]).reshape(-1).astype(np.uint32)
    return pos, idx


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    pos, idx = extract(src)
    np.savez_compressed(dst, vertices=pos, indices=idx)
    print(dst, pos.shape, idx.shape)
