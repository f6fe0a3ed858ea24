"""The command-line driver on the GPU: an input file in the reference's format, the output
cadence, and a restart from one of its dumps that continues bit for bit."""
import os

import numpy as np
import pytest

from idefix_amd import dump, vtk
from idefix_amd.run import simulate

pytestmark = pytest.mark.gpu

OT_INI = """
[Grid]
X1-grid    1  0.0  64  u  1.0
X2-grid    1  0.0  64  u  1.0

[TimeIntegrator]
CFL         0.6
tstop       0.05
first_dt    1.e-4
nstages     2

[Hydro]
solver    %s

[Boundary]
X1-beg    periodic
X1-end    periodic
X2-beg    periodic
X2-end    periodic

[Output]
vtk    0.02
dmp    0.02
"""


SOD_INI = """
[Grid]
X1-grid    1  0.0  500  u  1.0
[TimeIntegrator]
CFL         0.8
tstop       0.2
first_dt    1.e-4
nstages     2
[Hydro]
solver    hllc
gamma     1.4
[Boundary]
X1-beg    outflow
X1-end    outflow
"""


@pytest.mark.parametrize("solver", ["hll", "hlld"])
def test_driver_outputs_and_restart(tmp_path, solver):
    d = str(tmp_path)
    path = os.path.join(d, "idefix.ini")
    with open(path, "w") as fh:
        fh.write(OT_INI % solver)
    rc, g, out = simulate(["-i", path, "--setup", "orszag_tang", "--dir", d])
    assert rc == 0 and g.t == pytest.approx(0.05, abs=1e-15)
    names = sorted(f for f in os.listdir(d) if f.endswith((".vtk", ".dmp")))
    # outputs at t = 0, 0.02, 0.04 (the first cycle beyond each period boundary)
    assert names == ["data.0000.vtk", "data.0001.vtk", "data.0002.vtk",
                     "dump.0000.dmp", "dump.0001.dmp", "dump.0002.dmp"]
    _, F = dump.read_dump(os.path.join(d, "dump.0001.dmp"))
    t1 = float(F["time"][0])
    assert 0.02 <= t1 < 0.03
    fields, _, meta = vtk.read_vtk(os.path.join(d, "data.0001.vtk"))
    assert meta["TIME"] == pytest.approx(t1, rel=1e-6)
    Vc, Vs = g.download()
    b, e = g.beg, g.end
    # restart from dump 1 and run to the same tstop: identical final state
    rc2, h, out2 = simulate(["-i", path, "--setup", "orszag_tang", "--dir", d,
                             "-restart", "1"])
    assert rc2 == 0 and h.t == g.t
    Vc2, Vs2 = h.download()
    assert np.array_equal(Vc2[:, :, b[1]:e[1], b[0]:e[0]], Vc[:, :, b[1]:e[1], b[0]:e[0]])
    assert np.array_equal(Vs2[0, :, b[1]:e[1], b[0]:e[0] + 1], Vs[0, :, b[1]:e[1], b[0]:e[0] + 1])
    assert np.array_equal(Vs2[1, :, b[1]:e[1] + 1, b[0]:e[0]], Vs[1, :, b[1]:e[1] + 1, b[0]:e[0]])
    # the restarted run continued the numbering: its next dump overwrote dump.0002
    assert out2.dump_number == 3 and out2.vtk_number == 3


def test_driver_sod_hllc(tmp_path):
    d = str(tmp_path)
    path = os.path.join(d, "idefix.ini")
    with open(path, "w") as fh:
        fh.write(SOD_INI)
    rc, g, out = simulate(["-i", path, "--setup", "sod", "--dir", d, "-nowrite"])
    assert rc == 0 and g.t == pytest.approx(0.2, abs=1e-15)
    assert not [f for f in os.listdir(d) if f.endswith((".vtk", ".dmp"))]
    Vc, _ = g.download()
    rho = Vc[0, 0, 0, g.beg[0]:g.end[0]]
    assert rho.min() > 0.12 and rho.max() < 1.0 + 1e-12
