"""Minimal reader of netCDF-3 files (CDF-1, CDF-2 and CDF-5) written from the
published "NetCDF Classic and 64-bit Offset File Format" grammar and its
CDF-5 extension: big-endian header of dimensions, global attributes and
variables (name, dimension ids, attributes, type, vsize, begin), then each
non-record variable's data at its begin offset.  Test infrastructure: it
checks the library's writer; scipy.io.netcdf_file (CDF-1/2 only) checks this
reader."""
import struct

import numpy as np

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
TYPES = {1: ("b", 1), 2: ("S1", 1), 3: (">i2", 2), 4: (">i4", 4), 5: (">f4", 4), 6: (">f8", 8),
         7: ("u1", 1), 8: (">u2", 2), 9: (">u4", 4), 10: (">i8", 8), 11: (">u8", 8)}


class _Buf:
    def __init__(self, data: bytes, version: int):
        self.d, self.p, self.v = data, 4, version

    def u32(self):
        x = struct.unpack_from(">I", self.d, self.p)[0]
        self.p += 4
        return x

    def u64(self):
        x = struct.unpack_from(">Q", self.d, self.p)[0]
        self.p += 8
        return x

    def nonneg(self):
        return self.u64() if self.v == 5 else self.u32()

    def offset(self):
        return self.u32() if self.v == 1 else self.u64()

    def name(self):
        n = self.nonneg()
        s = self.d[self.p:self.p + n].decode()
        self.p += (n + 3) // 4 * 4
        return s

    def values(self, t, n):
        dt, size = TYPES[t]
        raw = self.d[self.p:self.p + n * size]
        self.p += (n * size + 3) // 4 * 4
        if t == 2:
            return raw.decode()
        return np.frombuffer(raw, dt).copy()

    def atts(self):
        tag, n = self.u32(), self.nonneg()
        if tag == 0:
            assert n == 0
            return {}
        assert tag == NC_ATTRIBUTE, tag
        out = {}
        for _ in range(n):
            name = self.name()
            t = self.u32()
            out[name] = (t, self.values(t, self.nonneg()))
        return out


def read(path: str) -> dict:
    """{'version', 'dims': {name: len}, 'attrs': {name: (type, value)},
    'vars': {name: {'dims', 'type', 'attrs', 'data', 'begin', 'vsize'}}}"""
    data = open(path, "rb").read()
    assert data[:3] == b"CDF", "not a netCDF-3 file"
    version = data[3]
    assert version in (1, 2, 5), version
    b = _Buf(data, version)
    numrecs = b.nonneg()
    assert numrecs == 0, "record variables are not read here"
    tag, n = b.u32(), b.nonneg()
    dims = []
    if tag:
        assert tag == NC_DIMENSION
        for _ in range(n):
            dims.append((b.name(), b.nonneg()))
    gatts = b.atts()
    tag, n = b.u32(), b.nonneg()
    variables = {}
    if tag:
        assert tag == NC_VARIABLE
        for _ in range(n):
            name = b.name()
            ids = [b.nonneg() for _ in range(b.nonneg())]
            vatts = b.atts()
            t, vsize, begin = b.u32(), b.nonneg(), b.offset()
            shape = tuple(dims[i][1] for i in ids)
            dt, size = TYPES[t]
            count = int(np.prod(shape)) if shape else 1
            raw = data[begin:begin + count * size]
            arr = np.frombuffer(raw, dt).reshape(shape).copy()
            variables[name] = dict(dims=[dims[i][0] for i in ids], type=t, attrs=vatts, data=arr, begin=begin,
                                   vsize=vsize)
    return dict(version=version, dims=dict(dims), attrs=gatts, vars=variables, size=len(data))
