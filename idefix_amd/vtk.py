"""Legacy-VTK snapshots in the reference's layout (src/output/vtk.cpp:288-600, Cartesian:
RECTILINEAR_GRID, binary, big endian, float32).

Header lines "# vtk DataFile Version 2.0", "Idefix <version> VTK Data", "BINARY",
"DATASET RECTILINEAR_GRID", "FIELD FieldData 9" then the fields GEOMETRY (int),
PERIODICITY (3 int), TIME (float), X1L/X2L/X3L_NATIVE_COORDINATES (nodes) and
X1C/X2C/X3C_NATIVE_COORDINATES (centres), "DIMENSIONS" (nodes per direction, one more than
cells along active directions), X/Y/Z_COORDINATES (nodes), "CELL_DATA n" and one
"SCALARS <name> float / LOOKUP_TABLE default" block per cell variable in name order (the
writer iterates a std::map), values x fastest.  Along an inactive direction the single
node is the cell centre.
"""
import numpy as np

VERSION = "v2.2.02"
GEOMETRY_CARTESIAN = 0  # idefix.hpp geometry codes


def _names(comps, mhd, ntracer=0, isothermal=False):
    n = ["RHO", "VX1", "VX2", "VX3"][:1 + comps] + ([] if isothermal else ["PRS"])
    if mhd:
        n += ["BX1", "BX2", "BX3"][:comps]
    return n + ["TR%d" % t for t in range(ntracer)]  # fluid.hpp:633-638


def write_vtk(filename, g, periodic=(True, True, True)):
    Vc, _ = g.download()
    b, e = g.beg, g.end
    nx = [e[d] - b[d] for d in range(3)]
    active = [g.np_tot[d] > 1 for d in range(3)]
    off = [1 if a else 0 for a in active]
    be = lambda a: np.asarray(a, dtype=">f4").tobytes()
    nodes, centres = [], []
    for d in range(3):
        x = np.asarray(g.x[d])
        xl = np.asarray(g.xl[d])
        if active[d]:
            nd = xl[b[d]:b[d] + nx[d] + 1] if len(xl) > b[d] + nx[d] else np.append(
                xl[b[d]:e[d]], np.asarray(g.xr[d])[e[d] - 1])
            ce = x[b[d]:e[d]]
        else:
            nd = x[:nx[d] + off[d]]
            ce = nd[:nx[d]]
        nodes.append(nd.astype(np.float64))
        centres.append(ce.astype(np.float64))
    with open(filename, "wb") as fh:
        w = lambda s: fh.write(s.encode())
        w("# vtk DataFile Version 2.0\n")
        w("Idefix %s VTK Data\n" % VERSION)
        w("BINARY\n")
        w("DATASET RECTILINEAR_GRID\n")
        w("FIELD FieldData 9\n")
        w("GEOMETRY 1 1 int\n")
        fh.write(np.asarray([GEOMETRY_CARTESIAN], dtype=">i4").tobytes())
        w("\nPERIODICITY 1 3 int\n")
        fh.write(np.asarray([1 if p else 0 for p in periodic], dtype=">i4").tobytes())
        w("\nTIME 1 1 float\n")
        fh.write(be([g.t]))
        for d, lab in enumerate(("X1L", "X2L", "X3L")):
            w("\n%s_NATIVE_COORDINATES 1 %d float\n" % (lab, nx[d] + off[d]))
            fh.write(be(nodes[d]))
        for d, lab in enumerate(("X1C", "X2C", "X3C")):
            w("\n%s_NATIVE_COORDINATES 1 %d float\n" % (lab, nx[d]))
            fh.write(be(centres[d]))
        w("\nDIMENSIONS %d %d %d\n" % tuple(nx[d] + off[d] for d in range(3)))
        for d, lab in enumerate("XYZ"):
            w(("" if d == 0 else "\n") + "%s_COORDINATES %d float\n" % (lab, nx[d] + off[d]))
            fh.write(be(nodes[d]))
        w("\nCELL_DATA %d\n" % (nx[0] * nx[1] * nx[2]))
        names = _names(g.comps, g.mhd, getattr(g, "ntracer", 0), getattr(g, "isothermal", False))
        for name in sorted(names):
            n = names.index(name)
            w("\nSCALARS %s float\n" % name)
            w("LOOKUP_TABLE default\n")
            fh.write(be(Vc[n][b[2]:e[2], b[1]:e[1], b[0]:e[0]].ravel()))


def read_vtk(filename):
    """Minimal reader for the files write_vtk produces: returns (fields, coords) with
    fields[name] shaped [nz][ny][nx] and coords = (x nodes, y nodes, z nodes)."""
    raw = open(filename, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        s = raw[pos:end].decode()
        pos = end + 1
        return s

    def floats(n, dt=">f4"):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + 4 * n], dtype=dt).astype(np.float64)
        pos += 4 * n
        return a

    coords, fields, dims, meta = [], {}, None, {}
    while pos < len(raw):
        s = line()
        if not s:
            continue
        tok = s.split()
        if tok[0] in ("GEOMETRY", "PERIODICITY"):
            meta[tok[0]] = floats(int(tok[2]), ">i4")
        elif tok[0] == "TIME":
            meta["TIME"] = floats(1)[0]
        elif tok[0].endswith("_NATIVE_COORDINATES"):
            meta[tok[0]] = floats(int(tok[2]))
        elif tok[0] == "DIMENSIONS":
            dims = [int(v) for v in tok[1:4]]
        elif tok[0] in ("X_COORDINATES", "Y_COORDINATES", "Z_COORDINATES"):
            coords.append(floats(int(tok[1])))
        elif tok[0] == "SCALARS":
            name = tok[1]
            line()  # LOOKUP_TABLE
            ncell = [max(d - 1, 1) if d > 1 else 1 for d in dims]
            # a direction with one node holds one cell
            fields[name] = floats(int(np.prod(ncell))).reshape(ncell[::-1])
    return fields, coords, meta
