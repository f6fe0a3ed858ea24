"""Restart dumps in the reference's .dmp layout (src/output/dump.cpp:812-979): byte-level
layout of a small dump, read/write round trip, and (GPU) restart bit-exactness as in the
reference's test/MHD/OrszagTang3D restart check."""
import struct
import types

import numpy as np
import pytest

from idefix_amd import dump


def _fake(dim=2, comps=2, nx=(5, 4, 1), ng=2, seed=0):
    rng = np.random.default_rng(seed)
    nt = [nx[d] + 2 * ng if d < dim else 1 for d in range(3)]
    nvar = 2 * comps + 2
    Vc = rng.normal(size=(nvar, nt[2], nt[1], nt[0]))
    vs = [nt[d] + (1 if d < dim else 0) for d in range(3)]
    Vs = rng.normal(size=(dim, vs[2], vs[1], vs[0]))
    beg = [ng if d < dim else 0 for d in range(3)]
    end = [beg[d] + nx[d] for d in range(3)]
    x = [np.arange(nt[d]) + 0.5 for d in range(3)]
    xl = [np.arange(nt[d]) + 0.0 for d in range(3)]
    xr = [np.arange(nt[d]) + 1.0 for d in range(3)]
    g = types.SimpleNamespace(dim=dim, comps=comps, mhd=True, beg=beg, end=end, x=x, xl=xl,
                              xr=xr, dx=[np.ones(nt[d]) for d in range(3)], t=0.25, dt=1e-3)
    g.download = lambda: (Vc.copy(), Vs.copy())
    return g, Vc, Vs


def test_layout_bytes(tmp_path):
    g, Vc, Vs = _fake()
    fn = tmp_path / "dump.0000.dmp"
    dump.write_dump(fn, g)
    raw = fn.read_bytes()
    assert raw[:128].rstrip(b"\0") == b"Idefix v2.2.02 Dump Data little endian"
    assert len(raw) >= 128
    # first field: x1, double, 1-D, 5 interior centres 2.5 .. 6.5
    assert raw[128:144].rstrip(b"\0") == b"x1"
    typ, ndim, n = struct.unpack("<iii", raw[144:156])
    assert (typ, ndim, n) == (0, 1, 5)
    assert np.array_equal(np.frombuffer(raw[156:196], "<f8"), np.arange(2, 7) + 0.5)
    _, F = dump.read_dump(fn)
    names = list(F)
    # grid, then the registered fields in name order
    assert names[:9] == ["x1", "xl1", "xr1", "x2", "xl2", "xr2", "x3", "xl3", "xr3"]
    assert names[9:] == sorted(names[9:])
    assert names[9:] == ["Vc-PRS", "Vc-RHO", "Vc-VX1", "Vc-VX2", "Vs-BX1s", "Vs-BX2s", "dt", "time"]
    # face field: one extra face along its direction, x size first in the file
    assert F["Vs-BX1s"].shape == (1, 4, 6)
    assert F["Vs-BX2s"].shape == (1, 5, 5)


def test_round_trip(tmp_path):
    g, Vc, Vs = _fake(dim=3, comps=3, nx=(4, 3, 2), seed=2)
    fn = tmp_path / "d.dmp"
    dump.write_dump(fn, g, extra={"dumpFileNumber": (7, 2), "periodicity": ([1, 1, 0], 3)})
    h, F = dump.read_dump(fn)
    b, e = g.beg, g.end
    assert np.array_equal(F["Vc-RHO"], Vc[0][b[2]:e[2], b[1]:e[1], b[0]:e[0]])
    assert np.array_equal(F["Vc-PRS"], Vc[4][b[2]:e[2], b[1]:e[1], b[0]:e[0]])
    assert np.array_equal(F["Vs-BX3s"], Vs[2][b[2]:e[2] + 1, b[1]:e[1], b[0]:e[0]])
    assert "Vc-BX3" not in F and "Vc-BX1" not in F
    assert F["time"][0] == 0.25 and F["dt"][0] == 1e-3
    assert F["dumpFileNumber"].dtype == np.int32 and F["dumpFileNumber"][0] == 7
    assert F["periodicity"].tolist() == [True, True, False]


@pytest.mark.gpu
def test_restart_is_bit_exact(tmp_path):
    """Dump after 5 cycles; 5 more cycles from the live state and from the restarted state
    give identical interiors (test/MHD/OrszagTang3D/setup.cpp:27-160 checks the same)."""
    from idefix_amd import problems
    from idefix_amd.datablock import DataBlock

    def block():
        return DataBlock(3, (16, 12, 10), mhd=True, recon="plm", solver="hlld",
                         emf="uct_contact", cfl=0.9, nstages=2)

    a = block()
    Vc, Vs = problems.orszag_tang_3d(a)
    a.upload(Vc, Vs)
    for _ in range(5):
        a.cycle()
    fn = tmp_path / "dump.0001.dmp"
    dump.write_dump(fn, a)
    for _ in range(5):
        a.cycle()
    b = block()
    dump.restart(b, fn)
    for _ in range(5):
        b.cycle()
    Va, Sa = a.download()
    Vb, Sb = b.download()
    lo, hi = a.beg, a.end
    assert np.array_equal(Va[:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]],
                          Vb[:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]])
    for d in range(3):
        sl = [slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0])]
        sl[2 - d] = slice(lo[d], hi[d] + 1)
        assert np.array_equal(Sa[d][tuple(sl)], Sb[d][tuple(sl)])
    assert a.t == b.t and a.dt == b.dt


def test_reads_a_dump_written_by_the_reference(tmp_path):
    """tests/golden/mti_init.dmp is the reference's own test/MHD/MTI/init.dmp (written by
    Idefix 2.0.00, 2-D, two components, 32 x 32 on [0, 0.1]^2, first_dt 1e-8 in its input).
    The reader recovers the header, every field with its type and shape, the grid of the
    input file, a face field whose discrete divergence is at round-off, and the
    restart scalars; writing the fields back with this writer and reading them again
    gives the same arrays."""
    import os
    fn = os.path.join(os.path.dirname(__file__), "golden", "mti_init.dmp")
    head, F = dump.read_dump(fn)
    assert head == "Idefix 2.0.00-ba41b38 Dump Data little endian"
    assert list(F)[:9] == ["x1", "xl1", "xr1", "x2", "xl2", "xr2", "x3", "xl3", "xr3"]
    names = list(F)[9:]
    assert names == sorted(names)
    assert {"Vc-RHO", "Vc-VX1", "Vc-VX2", "Vc-PRS", "Vs-BX1s", "Vs-BX2s", "time", "dt",
            "geometry", "periodicity", "dumpFileNumber", "vtkFileNumber"} <= set(names)
    for name in ("Vc-RHO", "Vc-VX1", "Vc-VX2", "Vc-PRS"):
        assert F[name].shape == (1, 32, 32) and F[name].dtype == np.float64
    assert F["Vs-BX1s"].shape == (1, 32, 33) and F["Vs-BX2s"].shape == (1, 33, 32)
    assert F["periodicity"].dtype == np.int32 and F["dumpFileNumber"].dtype == np.int32
    h = 0.1 / 32
    assert np.allclose(F["xl1"], np.arange(32) * h, rtol=0, atol=1e-15)
    assert np.allclose(F["x2"], (np.arange(32) + 0.5) * h, rtol=0, atol=1e-15)
    assert np.allclose(F["xr2"] - F["xl2"], h, rtol=1e-12)
    assert F["time"][0] == 0.0 and F["dt"][0] == 1e-8
    assert list(F["periodicity"]) [:2] == [1, 0]   # X1 periodic, X2 userdef (idefix.ini)
    assert np.all(F["Vc-RHO"] > 0) and np.all(F["Vc-PRS"] > 0)
    bx, by = F["Vs-BX1s"][0], F["Vs-BX2s"][0]
    div = (bx[:, 1:] - bx[:, :-1]) / h + (by[1:, :] - by[:-1, :]) / h
    assert np.max(np.abs(div)) * h <= 1e-12 * max(np.max(np.abs(bx)), np.max(np.abs(by)))
    # write the same interior back through this writer, read it again
    nt = [36, 36, 1]
    Vc = np.zeros((6, 1, 36, 36))
    Vs = np.zeros((2, 1, 37, 37))
    for n, name in enumerate(["Vc-RHO", "Vc-VX1", "Vc-VX2", "Vc-PRS"]):
        Vc[n, :, 2:34, 2:34] = F[name]
    Vs[0, :, 2:34, 2:35] = F["Vs-BX1s"]
    Vs[1, :, 2:35, 2:34] = F["Vs-BX2s"]
    g = types.SimpleNamespace(dim=2, comps=2, mhd=True, beg=[2, 2, 0], end=[34, 34, 1],
                              x=[np.concatenate(([0, 0], F["x1"], [0, 0])),
                                 np.concatenate(([0, 0], F["x2"], [0, 0])), F["x3"]],
                              xl=[np.concatenate(([0, 0], F["xl1"], [0, 0])),
                                  np.concatenate(([0, 0], F["xl2"], [0, 0])), F["xl3"]],
                              xr=[np.concatenate(([0, 0], F["xr1"], [0, 0])),
                                  np.concatenate(([0, 0], F["xr2"], [0, 0])), F["xr3"]],
                              dx=[np.ones(nt[d]) for d in range(3)], t=0.0, dt=1e-8)
    g.download = lambda: (Vc.copy(), Vs.copy())
    out = tmp_path / "again.dmp"
    dump.write_dump(out, g)
    _, G = dump.read_dump(out)
    for name in ("x1", "xl2", "Vc-RHO", "Vc-VX1", "Vc-VX2", "Vc-PRS", "Vs-BX1s", "Vs-BX2s", "dt"):
        assert np.array_equal(G[name], F[name]), name
