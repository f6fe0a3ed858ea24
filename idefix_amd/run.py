"""Command-line driver with the reference's main loop (src/main.cpp:120-260):

    python -m idefix_amd.run -i idefix.ini --setup orszag_tang|path/setup.py [-restart [n]]
                             [-maxcycles N] [-nowrite] [--recon plm|ppm] [--dir DIR]

Reads the input file (Input, src/input.cpp), builds the DataBlock, creates the initial
state with the setup or restarts from a dump, then cycles to tstop with the dt clipping
of main.cpp:187 and calls the output cadence after every cycle.  Options keep the
reference's spelling (-i, -restart, -maxcycles, -nowrite).

The setup is a built-in one (functions of idefix_amd.problems reproducing reference
tests) or a user's setup file, the counterpart of the setup.cpp the reference compiles
with each problem (src/setup.hpp:23-30):

    MHD = True          # definitions.hpp: MHD or hydro (default True)
    COMPONENTS = 3      # definitions.hpp: COMPONENTS (default: the grid's dimensions)

    def setup(data, inp):
        # Setup::Setup: enroll functions, e.g. data.set_userdef_boundary(fn) for
        # "userdef" sides (fn(dir, side, t, Vc, Vs) fills their ghost zones) and
        # data.set_iso_sound_speed(cs) for "csiso userdef"; inp holds the input file's
        # blocks ({block: {entry: [values]}})
        ...

    def init_flow(data):
        # Setup::InitFlow: return (Vc, Vs) as numpy arrays of the block's shapes
        ...

setup() is optional; init_flow() is required.  It runs once per run, before the
initial state or the restart, as the reference constructs Setup before both.
"""
import argparse
import importlib.util
import os
import sys
import time

from . import capi, ini, problems
from .output import Output

SETUPS = {
    # name -> (function, mhd, COMPONENTS or None for DIMENSIONS): the reference's tests
    # these setups reproduce
    "orszag_tang": (problems.orszag_tang, True, None),        # test/MHD/OrszagTang
    "orszag_tang_3d": (problems.orszag_tang_3d, True, None),  # test/MHD/OrszagTang3D
    "sod": (problems.sod, False, None),                       # test/HD/sod
    "brio_wu": (problems.brio_wu, True, 2),                   # test/MHD/sod
    "sod_iso": (problems.sod_iso, False, None),               # test/HD/sod-iso
    "brio_wu_iso": (problems.brio_wu_iso, True, 2),           # test/MHD/sod-iso
    "kelvin_helmholtz": (problems.kelvin_helmholtz, True, 3),
    "turbulence": (problems.turbulence, True, 3),
}


def load_setup(name):
    """(init_flow, mhd, COMPONENTS or None, setup or None) of a built-in setup name or of
    a user's setup file (*.py)."""
    if name in SETUPS:
        fn, mhd, comps = SETUPS[name]
        return fn, mhd, comps, None
    if name.endswith(".py"):
        if not os.path.isfile(name):
            raise FileNotFoundError("setup file %s not found" % name)
        spec = importlib.util.spec_from_file_location("idefix_user_setup", name)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        init = getattr(mod, "init_flow", None)
        if not callable(init):
            raise ValueError("setup file %s defines no init_flow(data)" % name)
        enroll = getattr(mod, "setup", None)
        if enroll is not None and not callable(enroll):
            raise ValueError("setup file %s: setup must be a function setup(data, inp)" % name)
        comps = getattr(mod, "COMPONENTS", None)
        return init, bool(getattr(mod, "MHD", True)), None if comps is None else int(comps), enroll
    raise ValueError("unknown setup %r: one of %s, or a setup file *.py"
                     % (name, ", ".join(sorted(SETUPS))))


def build(blocks, setup, recon="plm", precision="fp64", comps=None):
    from .datablock import DataBlock
    fn, mhd, comps0, enroll = load_setup(setup)
    comps = comps if comps is not None else comps0
    kw, run = ini.block_config(blocks, recon=recon, mhd=mhd, precision=precision, comps=comps)
    g = DataBlock(**kw)
    if enroll is not None:
        enroll(g, blocks)
    if run.get("csiso_userdef") and getattr(g, "_cs", None) is None:
        # eos_isothermal.hpp:41-45: "userdef" without an enrolled sound speed
        raise ValueError("csiso userdef: no user-defined isothermal sound speed has been "
                         "enrolled (the setup calls data.set_iso_sound_speed)")
    periodic = [all(b == "periodic" for b in kw["bc"][d]) if d < kw["dim"] else True
                for d in range(3)]
    return g, fn, run, periodic


def main(argv=None):
    return simulate(argv)[0]


def simulate(argv=None):
    """The main loop; returns (return code, DataBlock, Output)."""
    ap = argparse.ArgumentParser(prog="idefix_amd.run")
    ap.add_argument("-i", dest="ini", default="idefix.ini")
    ap.add_argument("--setup", required=True,
                    help="a built-in setup (%s) or a setup file *.py" % ", ".join(sorted(SETUPS)))
    ap.add_argument("--recon", default="plm", choices=["constant", "plm", "ppm"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--comps", type=int, default=None)
    ap.add_argument("--dir", default=".")
    ap.add_argument("-restart", nargs="?", type=int, const=-1, default=None)
    ap.add_argument("-maxcycles", type=int, default=-1)
    ap.add_argument("-nowrite", action="store_true")
    a = ap.parse_args(argv)

    blocks = ini.read_ini(a.ini)
    g, fn, run, periodic = build(blocks, a.setup, a.recon, a.precision, a.comps)
    out = Output(run["output"], g, directory=a.dir, force_no_write=a.nowrite,
                 periodic=periodic)
    restarted = False
    if a.restart is not None:
        restarted = out.restart_from_dump(g, a.restart)
        print("Main: restart %s." % ("done" if restarted else "aborted"))
        if restarted:
            g.set_boundaries()
    if not restarted:
        Vc, Vs = fn(g)
        g.upload(Vc, Vs)
        g.set_boundaries()
        out.check_for_writes(g)
    tstop = run["tstop"]
    rc = 0
    t0 = time.perf_counter()
    n0 = g.ncycles
    while g.t < tstop:
        if tstop - g.t < g.dt:
            g.dt = tstop - g.t
        try:
            g.cycle()
        except capi.IdfxError as e:  # the TimeIntegrator's exceptions (NaN, small dt)
            print("Main: WARNING! Caught an exception in TimeIntegrator.\n%s" % e)
            out.force_write_vtk(g)
            rc = 1
            break
        out.check_for_writes(g)
        if 0 <= a.maxcycles <= g.ncycles:
            print("Main: Reached maximum number of integration cycles.")
            break
    g.torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = g.nx[0] * g.nx[1] * g.nx[2]
    ncyc = g.ncycles - n0
    print("Main: Reached t=%.10g after %d cycles" % (g.t, ncyc))
    if ncyc > 0:
        print("Main: Perfs are %.4g cell updates/second" % (cells * ncyc / wall))
    return rc, g, out


if __name__ == "__main__":
    sys.exit(main())
