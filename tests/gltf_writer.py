"""A small glTF 2.0 writer for the loader tests: known arrays in, a GLB or a
JSON document with a base64 data URI out, with the accessor layouts glTF
allows (separate or interleaved views, strides with padding, byte offsets,
every component type)."""
import base64
import json
import struct

import numpy as np

CT = {"i8": 5120, "u8": 5121, "i16": 5122, "u16": 5123, "u32": 5125, "f32": 5126}
NP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
TYPES = {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4", 16: "MAT4"}


class Builder:
    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": []}

    def _align(self, n=4):
        while len(self.bin) % n:
            self.bin.append(0)

    def accessor(self, array, ct, normalized=False, pad=0, offset=0):
        """One accessor in a view of its own; pad adds bytes after each
        element (a byteStride larger than the element), offset bytes before
        the first element (the accessor's byteOffset)."""
        ctype = CT[ct] if isinstance(ct, str) else ct
        a = np.ascontiguousarray(array, dtype=NP[ctype])
        if a.ndim == 1:
            a = a.reshape(-1, 1)
        n, k = a.shape
        esize = a.dtype.itemsize * k
        stride = esize + pad
        if pad and stride % 4:
            stride += 4 - stride % 4
        self._align()
        start = len(self.bin)
        body = bytearray(offset + (n - 1) * stride + esize if n else offset)
        for i in range(n):
            body[offset + i * stride: offset + i * stride + esize] = a[i].tobytes()
        self.bin += body
        view = {"buffer": 0, "byteOffset": start, "byteLength": len(body)}
        if pad:
            view["byteStride"] = stride
        self.doc["bufferViews"].append(view)
        acc = {"bufferView": len(self.doc["bufferViews"]) - 1, "componentType": ctype, "count": n, "type": TYPES[k]}
        if offset:
            acc["byteOffset"] = offset
        if normalized:
            acc["normalized"] = True
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def interleaved(self, arrays):
        """Several (array, ct, normalized) in one view, element by element."""
        parts = []
        for array, ct, norm in arrays:
            ctype = CT[ct]
            a = np.ascontiguousarray(array, dtype=NP[ctype])
            if a.ndim == 1:
                a = a.reshape(-1, 1)
            parts.append((a, ctype, norm))
        n = len(parts[0][0])
        offs, off = [], 0
        for a, _, _ in parts:
            offs.append(off)
            off += a.dtype.itemsize * a.shape[1]
            off += (-off) % 4
        stride = off
        self._align()
        start = len(self.bin)
        body = bytearray(n * stride)
        for i in range(n):
            for (a, _, _), o in zip(parts, offs):
                row = a[i].tobytes()
                body[i * stride + o: i * stride + o + len(row)] = row
        self.bin += body
        self.doc["bufferViews"].append({"buffer": 0, "byteOffset": start, "byteLength": len(body), "byteStride": stride})
        v = len(self.doc["bufferViews"]) - 1
        out = []
        for (a, ctype, norm), o in zip(parts, offs):
            acc = {"bufferView": v, "byteOffset": o, "componentType": ctype, "count": n, "type": TYPES[a.shape[1]]}
            if norm:
                acc["normalized"] = True
            self.doc["accessors"].append(acc)
            out.append(len(self.doc["accessors"]) - 1)
        return out

    def mesh(self, name, primitives, extras=None, weights=None):
        m = {"name": name, "primitives": primitives}
        if extras is not None:
            m["extras"] = extras
        if weights is not None:
            m["weights"] = weights
        self.doc["meshes"].append(m)
        return len(self.doc["meshes"]) - 1

    def skin(self, joints, inverse_bind=None, name="skin"):
        self.doc.setdefault("nodes", [])
        while len(self.doc["nodes"]) <= max(joints):
            self.doc["nodes"].append({"name": "node%d" % len(self.doc["nodes"])})
        s = {"name": name, "joints": list(joints)}
        if inverse_bind is not None:
            s["inverseBindMatrices"] = self.accessor(np.asarray(inverse_bind, dtype=np.float32).reshape(-1, 16), "f32")
        self.doc.setdefault("skins", []).append(s)

    def glb(self):
        self._align()
        doc = dict(self.doc)
        doc["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(doc).encode()
        js += b" " * ((-len(js)) % 4)
        out = bytearray()
        out += struct.pack("<III", 0x46546C67, 2, 12 + 8 + len(js) + 8 + len(self.bin))
        out += struct.pack("<II", len(js), 0x4E4F534A) + js
        out += struct.pack("<II", len(self.bin), 0x004E4942) + bytes(self.bin)
        return bytes(out)

    def gltf(self):
        doc = dict(self.doc)
        uri = "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()
        doc["buffers"] = [{"byteLength": len(self.bin), "uri": uri}]
        return json.dumps(doc).encode()
