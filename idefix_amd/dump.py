"""Restart dumps in the reference's .dmp format (src/output/dump.cpp:812-979 writer,
:628-810 reader; pytools/dump_io.py:16-117 documents the layout).

Layout: a 128-byte header "Idefix <version> Dump Data <little|big> endian" (C string,
zero padded), then fields.  A field is a 16-byte zero-padded name, an int32 type
(0 double, 1 float, 2 int32, 3 bool), an int32 ndim, ndim int32 sizes (x size first) and
the raw values with x varying fastest.  The grid comes first (x1, xl1, xr1, x2, ...: the
interior cell centres and edges of each direction), then the registered fields in name
order (the writer iterates a std::map), then "eof" (one double).  Cell fields hold the
interior (Vc-RHO, Vc-VX1.., Vc-PRS; Vc-BX2/Vc-BX3 only when the dimension is below 2/3),
face fields the interior faces (Vs-BX1s .. with one extra face along their direction),
scalars are 1-element arrays (time, dt, ...).
"""
import struct

import numpy as np

HEADER_SIZE = 128
NAME_SIZE = 16
TYPES = {0: ("d", np.float64), 1: ("f", np.float32), 2: ("i", np.int32), 3: ("?", np.bool_)}
VERSION = "v2.2.02"


def _cvar_names(dim, comps, mhd, ntracer=0, isothermal=False):
    """Vc variables saved in dumps (fluid.hpp:587-638): name -> variable index; passive
    tracers follow the fluid variables as Vc-TR0, Vc-TR1, ..."""
    names = {"Vc-RHO": 0, "Vc-VX1": 1}
    if comps >= 2:
        names["Vc-VX2"] = 2
    if comps >= 3:
        names["Vc-VX3"] = 3
    he = 0 if isothermal else 1  # no PRS in isothermal builds
    if he:
        names["Vc-PRS"] = comps + 1
    if mhd:
        if comps >= 2 and dim < 2:
            names["Vc-BX2"] = comps + 2 + he
        if comps >= 3 and dim < 3:
            names["Vc-BX3"] = comps + 3 + he
    nfluid = 2 * comps + 1 + he if mhd else comps + 1 + he
    for t in range(ntracer):
        names["Vc-TR%d" % t] = nfluid + t
    return names


def _field(fh, name, arr, typ=0):
    arr = np.asarray(arr)
    raw = name.encode()[:NAME_SIZE - 1]
    fh.write(raw + b"\0" * (NAME_SIZE - len(raw)))
    fh.write(struct.pack("<ii", typ, arr.ndim))
    fh.write(struct.pack("<%di" % arr.ndim, *arr.shape[::-1]))
    code, dt = TYPES[typ]
    fh.write(np.ascontiguousarray(arr, dtype=dt).astype("<" + np.dtype(dt).str[1:], copy=False).tobytes())


def write_dump(filename, g, extra=None):
    """Writes the state of a DataBlock-like object g (numpy/torch Vc [nvar][nz][ny][nx],
    Vs [dim][..], coordinates x/xl/xr, beg/end, t, dt) as a reference-format dump.
    `extra` adds scalar entries {name: (value, type)}."""
    Vc, Vs = g.download()
    b, e = g.beg, g.end
    with open(filename, "wb") as fh:
        head = ("Idefix %s Dump Data little endian" % VERSION).encode()
        fh.write(head + b"\0" * (HEADER_SIZE - len(head)))
        xr = getattr(g, "xr", None)
        for d in range(3):
            x = np.asarray(g.x[d])[b[d]:e[d]]
            xl = np.asarray(g.xl[d])[b[d]:e[d]]
            xrr = (np.asarray(xr[d])[b[d]:e[d]] if xr is not None
                   else xl + np.asarray(g.dx[d])[b[d]:e[d]])
            _field(fh, "x%d" % (d + 1), x)
            _field(fh, "xl%d" % (d + 1), xl)
            _field(fh, "xr%d" % (d + 1), xrr)
        entries = {}
        for name, n in _cvar_names(g.dim, g.comps, g.mhd, getattr(g, "ntracer", 0),
                                     getattr(g, "isothermal", False)).items():
            entries[name] = ("cell", n)
        if g.mhd:
            for d in range(g.dim):
                entries["Vs-BX%ds" % (d + 1)] = ("face", d)
        scalars = {"time": (g.t, 0), "dt": (g.dt, 0)}
        if extra:
            scalars.update(extra)
        for name, (v, typ) in scalars.items():
            entries[name] = ("scalar", (v, typ))
        for name in sorted(entries):
            kind, what = entries[name]
            if kind == "cell":
                _field(fh, name, Vc[what][b[2]:e[2], b[1]:e[1], b[0]:e[0]])
            elif kind == "face":
                d = what
                sl = [slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0])]
                sl[2 - d] = slice(b[d], e[d] + 1)
                _field(fh, name, Vs[d][tuple(sl)])
            else:
                v, typ = what
                _field(fh, name, np.atleast_1d(np.asarray(v, dtype=TYPES[typ][1])), typ)
        _field(fh, "eof", np.zeros(1))


def read_dump(filename):
    """Returns (header, {name: array}); arrays are indexed [k][j][i] (x fastest in the
    file, as numpy C order with the sizes reversed)."""
    out = {}
    with open(filename, "rb") as fh:
        head = fh.read(HEADER_SIZE)
        header = head[:head.index(b"\0")].decode()
        order = "<" if "little" in header or "endian" not in header else ">"
        while True:
            nm = fh.read(NAME_SIZE)
            if len(nm) < NAME_SIZE:
                raise ValueError("truncated dump: no eof field")
            name = nm[:nm.index(b"\0")].decode()
            typ, ndim = struct.unpack(order + "ii", fh.read(8))
            dims = struct.unpack(order + "%di" % ndim, fh.read(4 * ndim))
            code, dt = TYPES[typ]
            n = int(np.prod(dims))
            a = np.frombuffer(fh.read(n * np.dtype(dt).itemsize), dtype=np.dtype(dt).newbyteorder(order))
            arr = a.reshape(dims[::-1]).astype(dt)
            if name == "eof":
                break
            out[name] = arr
    return header, out


def restart(g, filename):
    """Loads a dump into a DataBlock (interior cells, interior faces, t, dt), the
    counterpart of Dump::Read (dump.cpp:628-810); ghost zones are filled by the next
    boundary step and the cell-centred field along the face directions from the faces."""
    _, F = read_dump(filename)
    Vc, Vs = g.download()
    b, e = g.beg, g.end
    for name, n in _cvar_names(g.dim, g.comps, g.mhd, getattr(g, "ntracer", 0),
                                     getattr(g, "isothermal", False)).items():
        Vc[n][b[2]:e[2], b[1]:e[1], b[0]:e[0]] = F[name]
    if g.mhd:
        for d in range(g.dim):
            sl = [slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0])]
            sl[2 - d] = slice(b[d], e[d] + 1)
            Vs[d][tuple(sl)] = F["Vs-BX%ds" % (d + 1)]
    g.upload(Vc, Vs if g.mhd else None)
    g.t = float(F["time"][0])
    g.dt = float(F["dt"][0])
    if g.mhd:
        g.set_boundaries()
    return F
