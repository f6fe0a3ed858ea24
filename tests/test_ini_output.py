"""Input-file parsing, the output cadence of Output::CheckForWrites and the restart
counters, on the CPU (fake blocks; the GPU run of the driver is in test_gpu_driver.py)."""
import numpy as np
import pytest

from idefix_amd import dump, ini, output, vtk
from idefix_amd.output import Output

# test/MHD/OrszagTang/idefix-hll.ini, as the reference ships it (plus a comment line)
OT_HLL = """
[Grid]
X1-grid    1  0.0  128  u  1.0
X2-grid    1  0.0  128  u  1.0

[TimeIntegrator]
CFL         0.6
tstop       0.5
first_dt    1.e-4
nstages     2

[Hydro]
solver    hll    # the HLL solver

[Boundary]
X1-beg    periodic
X1-end    periodic
X2-beg    periodic
X2-end    periodic

[Output]
vtk    0.5
dmp    0.5
log    100
"""

SOD = """
[Grid]
X1-grid    1  0.0  500  u  1.0
[TimeIntegrator]
CFL            0.8
CFL_max_var    1.1      # not used
tstop          0.2
first_dt       1.e-4
nstages        2
[Hydro]
solver    hllc
gamma     1.4
[Boundary]
X1-beg    outflow
X1-end    outflow
"""


def test_parse_blocks_and_comments():
    b = ini.parse_ini(OT_HLL)
    assert set(b) == {"Grid", "TimeIntegrator", "Hydro", "Boundary", "Output"}
    assert b["Grid"]["X1-grid"] == ["1", "0.0", "128", "u", "1.0"]
    assert b["Hydro"]["solver"] == ["hll"]
    assert b["Output"]["log"] == ["100"]
    with pytest.raises(ValueError):
        ini.parse_ini("CFL 0.5\n")
    with pytest.raises(ValueError):
        ini.parse_ini("[Grid\nX1-grid 1 0 8 u 1\n")


def test_block_config_with_reference_defaults():
    kw, run = ini.block_config(ini.parse_ini(OT_HLL))
    assert kw["dim"] == 2 and kw["nx"] == (128, 128) and kw["comps"] == 2
    assert kw["solver"] == "hll" and kw["emf"] == "uct_contact"  # default average
    assert kw["gamma"] == pytest.approx(5.0 / 3.0) and kw["cfl_max_var"] == 1.1
    assert kw["cfl"] == 0.6 and kw["nstages"] == 2 and kw["first_dt"] == 1e-4
    assert kw["bc"] == [("periodic", "periodic")] * 2
    assert run["tstop"] == 0.5 and run["output"]["vtk"] == ["0.5"]
    kw, run = ini.block_config(ini.parse_ini(SOD), mhd=False)
    assert kw["dim"] == 1 and kw["nx"] == (500,) and kw["gamma"] == 1.4
    assert kw["bc"] == [("outflow", "outflow")] and "emf" not in kw


def test_block_config_refusals():
    with pytest.raises(ValueError):
        ini.block_config(ini.parse_ini(SOD), mhd=True)  # hllc needs a HD fluid
    with pytest.raises(NotImplementedError):
        ini.block_config(ini.parse_ini(OT_HLL + "[Hydro]\nhall constant 0.1\n"))
    kw, _ = ini.block_config(ini.parse_ini(OT_HLL.replace("solver    hll", "solver hlld\ntracer 2")))
    assert kw["ntracer"] == 2 and kw["solver"] == "hlld"
    with pytest.raises(NotImplementedError):
        ini.block_config(ini.parse_ini(OT_HLL.replace("128  u", "128  l")))
    # isothermal sound speed: constant, userdef (eos_isothermal.hpp:23-35), else refused
    hydro = lambda line: ini.parse_ini(SOD.replace("gamma     1.4", "gamma     1.4\n" + line))
    kw, run = ini.block_config(hydro("csiso constant 0.5"), mhd=False)
    assert kw["csiso"] == 0.5 and not run["csiso_userdef"]
    kw, run = ini.block_config(hydro("csiso userdef"), mhd=False)
    assert run["csiso_userdef"] and kw["csiso"] > 0
    with pytest.raises(ValueError):
        ini.block_config(hydro("csiso table"), mhd=False)


class FakeBlock:
    """The attributes the writers read, with a small 2-D MHD state."""

    def __init__(self, t=0.0):
        self.t, self.dt = t, 1e-3
        self.dim, self.comps, self.mhd = 2, 2, True
        ng = 2
        self.nx = [6, 4, 1]
        self.beg = [ng, ng, 0]
        self.end = [ng + 6, ng + 4, 1]
        self.np_tot = [10, 8, 1]
        self.x, self.xl, self.xr, self.dx = [], [], [], []
        for d in range(3):
            n, h = self.np_tot[d], (1.0 / self.nx[d] if d < 2 else 1.0)
            i = np.arange(n) - (ng if d < 2 else 0)
            self.xl.append(i * h)
            self.xr.append((i + 1) * h)
            self.x.append((i + 0.5) * h)
            self.dx.append(np.full(n, h))
        rng = np.random.default_rng(1)
        self.Vc = rng.uniform(0.5, 1.5, (6, 1, 8, 10))
        self.Vs = rng.uniform(-1, 1, (2, 1, 9, 11))
        self.boundary_calls = 0

    def download(self):
        return self.Vc.copy(), self.Vs.copy()

    def upload(self, Vc, Vs=None):
        self.Vc = np.array(Vc)
        if Vs is not None:
            self.Vs = np.array(Vs)

    def set_boundaries(self):
        self.boundary_calls += 1


def test_check_for_writes_cadence(tmp_path):
    """vtk every 0.1 and dmp every 0.25 from t = 0: first files at t = 0, one file per
    crossing of a period boundary, and after a jump over several periods the clock moves
    to the last period boundary not beyond t (output.cpp:173-190, :279-316)."""
    g = FakeBlock()
    out = Output({"vtk": ["0.1"], "dmp": ["0.25"]}, g, directory=str(tmp_path))
    written = {}
    for t in (0.0, 0.05, 0.1, 0.12, 0.25, 0.35, 0.39, 0.41, 1.02, 1.05, 1.12):
        g.t = t
        written[t] = out.check_for_writes(g)
    assert {t: n for t, n in written.items() if n} == {0.0: 2, 0.1: 1, 0.25: 2, 0.35: 1,
                                                      0.41: 1, 1.02: 2, 1.12: 1}
    names = [p.split("/")[-1] for p in out.files]
    assert names == ["data.0000.vtk", "dump.0000.dmp", "data.0001.vtk", "data.0002.vtk",
                     "dump.0001.dmp", "data.0003.vtk", "data.0004.vtk", "data.0005.vtk",
                     "dump.0002.dmp", "data.0006.vtk"]
    assert out.vtk_last == pytest.approx(1.1)
    assert out.dump_last == pytest.approx(1.0)


def test_negative_period_and_nowrite(tmp_path):
    g = FakeBlock()
    out = Output({"vtk": ["-1"], "dmp": ["-1"]}, g, directory=str(tmp_path))
    assert not out.vtk_enabled and not out.dump_enabled
    assert out.check_for_writes(g) == 0
    out = Output({"vtk": ["0.1"]}, g, directory=str(tmp_path), force_no_write=True)
    assert out.check_for_writes(g) == 0 and not list(tmp_path.iterdir())


def test_dump_carries_output_counters_and_restart_restores_them(tmp_path):
    g = FakeBlock()
    out = Output({"vtk": ["0.1"], "dmp": ["0.2"]}, g, directory=str(tmp_path))
    for t in (0.0, 0.1, 0.2, 0.3):
        g.t = t
        out.check_for_writes(g)
    _, F = dump.read_dump(str(tmp_path / "dump.0001.dmp"))
    assert F["dumpFileNumber"].dtype == np.int32 and int(F["dumpFileNumber"][0]) == 2
    assert int(F["vtkFileNumber"][0]) == 3 and float(F["vtkLast"][0]) == pytest.approx(0.2)
    assert float(F["time"][0]) == 0.2
    assert list(F["periodicity"]) == [1, 1, 1] and int(F["geometry"][0]) == 0
    # restart from the latest dump into a fresh block
    h = FakeBlock()
    h.Vc[:] = 0
    h.Vs[:] = 0
    o2 = Output({"vtk": ["0.1"], "dmp": ["0.2"]}, h, directory=str(tmp_path))
    assert o2.restart_from_dump(h, -1)
    assert h.t == 0.2 and o2.dump_number == 2 and o2.vtk_number == 3
    assert o2.vtk_last == pytest.approx(0.2) and o2.dump_last == pytest.approx(0.2)
    b, e = h.beg, h.end
    # cell fields other than the face-held components, and the interior faces
    assert np.array_equal(h.Vc[:4, :, b[1]:e[1], b[0]:e[0]], g.Vc[:4, :, b[1]:e[1], b[0]:e[0]])
    assert np.array_equal(h.Vs[0, :, b[1]:e[1], b[0]:e[0] + 1],
                          g.Vs[0, :, b[1]:e[1], b[0]:e[0] + 1])
    assert np.array_equal(h.Vs[1, :, b[1]:e[1] + 1, b[0]:e[0]],
                          g.Vs[1, :, b[1]:e[1] + 1, b[0]:e[0]])
    assert h.boundary_calls == 1  # the face-held cell components come from the faces
    assert not o2.restart_from_dump(h, 7)  # no such dump


def test_vtk_snapshot_numbering_and_time(tmp_path):
    g = FakeBlock(t=0.0)
    out = Output({"vtk": ["0.5"]}, g, directory=str(tmp_path))
    out.check_for_writes(g)
    g.t = 0.5
    out.check_for_writes(g)
    fields, coords, meta = vtk.read_vtk(str(tmp_path / "data.0001.vtk"))
    assert meta["TIME"] == pytest.approx(0.5)
    assert sorted(fields) == ["BX1", "BX2", "PRS", "RHO", "VX1", "VX2"]
    assert output.INT == 2


def test_setup_loader(tmp_path):
    """--setup takes a built-in setup or a user's setup file: init_flow() is required,
    setup() optional, MHD and COMPONENTS mirror definitions.hpp."""
    from idefix_amd.run import SETUPS, load_setup
    fn, mhd, comps, enroll = load_setup("brio_wu")
    assert fn is SETUPS["brio_wu"][0] and mhd and comps == 2 and enroll is None
    good = tmp_path / "good.py"
    good.write_text("MHD = False\nCOMPONENTS = 1\n"
                    "def setup(data, inp):\n    pass\n"
                    "def init_flow(data):\n    return None, None\n")
    fn, mhd, comps, enroll = load_setup(str(good))
    assert fn.__name__ == "init_flow" and mhd is False and comps == 1
    assert enroll.__name__ == "setup"
    bare = tmp_path / "bare.py"
    bare.write_text("def init_flow(data):\n    return None, None\n")
    fn, mhd, comps, enroll = load_setup(str(bare))
    assert mhd is True and comps is None and enroll is None
    bad = tmp_path / "bad.py"
    bad.write_text("X = 1\n")
    with pytest.raises(ValueError, match="init_flow"):
        load_setup(str(bad))
    with pytest.raises(FileNotFoundError):
        load_setup(str(tmp_path / "missing.py"))
    with pytest.raises(ValueError, match="unknown setup"):
        load_setup("no_such_setup")
