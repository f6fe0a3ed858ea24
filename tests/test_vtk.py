"""Legacy-VTK snapshots in the reference's layout (src/output/vtk.cpp:365-600): header
fields, coordinates and big-endian float cell data, read back."""
import types

import numpy as np

from idefix_amd import vtk


def _fake(dim=2, comps=2, nx=(5, 4, 1), ng=2, seed=1):
    rng = np.random.default_rng(seed)
    nt = [nx[d] + 2 * ng if d < dim else 1 for d in range(3)]
    Vc = rng.normal(size=(2 * comps + 2, nt[2], nt[1], nt[0]))
    beg = [ng if d < dim else 0 for d in range(3)]
    end = [beg[d] + nx[d] for d in range(3)]
    dxs = [0.1, 0.2, 1.0]
    xl = [(np.arange(nt[d]) - beg[d]) * dxs[d] for d in range(3)]
    x = [xl[d] + 0.5 * dxs[d] for d in range(3)]
    xr = [xl[d] + dxs[d] for d in range(3)]
    g = types.SimpleNamespace(dim=dim, comps=comps, mhd=True, beg=beg, end=end, np_tot=nt,
                              x=x, xl=xl, xr=xr, t=0.125)
    g.download = lambda: (Vc.copy(), None)
    return g, Vc


def test_vtk_layout_and_round_trip(tmp_path):
    g, Vc = _fake()
    fn = tmp_path / "data.0000.vtk"
    vtk.write_vtk(fn, g, periodic=(True, False, False))
    raw = fn.read_bytes()
    assert raw.startswith(b"# vtk DataFile Version 2.0\nIdefix v2.2.02 VTK Data\nBINARY\n"
                          b"DATASET RECTILINEAR_GRID\nFIELD FieldData 9\nGEOMETRY 1 1 int\n")
    assert b"\nDIMENSIONS 6 5 1\n" in raw
    assert b"\nCELL_DATA 20\n" in raw
    fields, coords, meta = vtk.read_vtk(fn)
    assert list(fields) == ["BX1", "BX2", "PRS", "RHO", "VX1", "VX2"]   # name order
    b, e = g.beg, g.end
    for name, n in (("RHO", 0), ("VX1", 1), ("VX2", 2), ("PRS", 3), ("BX1", 4), ("BX2", 5)):
        assert np.array_equal(fields[name], Vc[n][b[2]:e[2], b[1]:e[1], b[0]:e[0]].astype(np.float32))
    assert np.allclose(coords[0], np.arange(6) * 0.1)
    assert np.allclose(coords[1], np.arange(5) * 0.2)
    assert np.allclose(coords[2], [0.5])           # inactive direction: the cell centre
    assert meta["TIME"] == np.float32(0.125)
    assert meta["PERIODICITY"].tolist() == [1, 0, 0]
    assert np.allclose(meta["X1C_NATIVE_COORDINATES"], 0.05 + 0.1 * np.arange(5))
