"""Input files in the reference's .ini format and the DataBlock they describe.

Parsing follows Input::Input (src/input.cpp:60-112): text after '#' is a comment, a line
"[Name]" opens a block, every other line is "parameter value value ...", values split on
white space.  `block_config` turns the blocks the hot path reads into DataBlock keyword
arguments, with the reference's defaults:
  [Grid]           X1-grid .. X3-grid: nPatches, then (start, size, type) per patch and
                   the end (gridHost.cpp:55-75); one uniform patch per direction here
  [TimeIntegrator] CFL, CFL_max_var (1.1), first_dt (1e-10), nstages, tstop
                   (timeIntegrator.cpp:30-45, main.cpp:184)
  [Hydro]          solver (riemannSolver.hpp:98-138), gamma (5/3, eos_adiabatic.hpp:21),
                   emf (uct_contact, constrainedTransport.hpp:156-177), tracer (0,
                   fluid.hpp:315-325)
  [Boundary]       X1-beg .. X3-end
The reconstruction is a build option of the reference (ORDER), so it is an argument here.
"""

UNSUPPORTED = {
    # [Hydro] entries that switch on physics this library does not implement
    "Hydro": ("hall", "resistivity", "ambipolar", "viscosity", "bragTDiffusion",
              "bragViscosity", "rotation", "shearingBox", "fargo",
              "gravPotential", "bodyForce", "drag"),
}


def parse_ini(text):
    """Returns {block: {parameter: [values as strings]}}."""
    blocks, cur = {}, None
    for raw in text.splitlines():
        line = raw.split("#", 1)[0]
        s = line.strip()
        if not s:
            continue
        if s.startswith("["):
            end = s.find("]")
            if end < 0:
                raise ValueError("block name %r not properly ended" % s)
            cur = s[1:end]
            blocks.setdefault(cur, {})
            continue
        if cur is None:
            raise ValueError("a block name must come before the first parameter")
        tok = s.split()
        blocks[cur][tok[0]] = tok[1:]
    return blocks


def read_ini(path):
    with open(path) as fh:
        return parse_ini(fh.read())


def _get(blocks, block, name, i=0, cast=str, default=None):
    vals = blocks.get(block, {}).get(name)
    if vals is None or len(vals) <= i:
        if default is None:
            raise KeyError("[%s] %s is required" % (block, name))
        return default
    return cast(vals[i])


def block_config(blocks, recon="plm", mhd=True, precision="fp64", comps=None):
    """DataBlock keyword arguments plus the run control values (tstop, outputs)."""
    for block, names in UNSUPPORTED.items():
        for n in names:
            if n in blocks.get(block, {}):
                raise NotImplementedError("[%s] %s is not supported" % (block, n))
    grid = blocks.get("Grid", {})
    nx, xbeg, xend = [], [], []
    for d in range(3):
        key = "X%d-grid" % (d + 1)
        if key not in grid:
            break
        v = grid[key]
        npatch = int(v[0])
        if npatch != 1 or v[3] != "u":
            raise NotImplementedError("%s: only one uniform patch per direction" % key)
        nx.append(int(v[2]))
        xbeg.append(float(v[1]))
        xend.append(float(v[4]))
    dim = sum(1 for n in nx if n > 1) if nx else 0
    if dim == 0 or any(n <= 1 for n in nx[:dim]):
        raise ValueError("[Grid] must describe 1 to 3 active directions, the first ones")
    while len(xbeg) < 3:
        xbeg.append(0.0)
        xend.append(1.0)
    solver = _get(blocks, "Hydro", "solver")
    if mhd and solver == "hllc":
        raise ValueError("hllc Riemann solver requires a HD fluid")
    if not mhd and solver == "hlld":
        raise ValueError("hlld Riemann solver requires a MHD fluid")
    bc = []
    for d in range(dim):
        bc.append((_get(blocks, "Boundary", "X%d-beg" % (d + 1)),
                   _get(blocks, "Boundary", "X%d-end" % (d + 1))))
    kw = dict(dim=dim, nx=tuple(nx[:dim]), xbeg=tuple(xbeg), xend=tuple(xend), mhd=mhd,
              comps=dim if comps is None else comps,
              recon=recon, solver=solver, precision=precision, bc=bc,
              gamma=_get(blocks, "Hydro", "gamma", cast=float, default=5.0 / 3.0),
              nstages=_get(blocks, "TimeIntegrator", "nstages", cast=int),
              cfl=_get(blocks, "TimeIntegrator", "CFL", cast=float),
              cfl_max_var=_get(blocks, "TimeIntegrator", "CFL_max_var", cast=float, default=1.1),
              first_dt=_get(blocks, "TimeIntegrator", "first_dt", cast=float, default=1.0e-10),
              ntracer=_get(blocks, "Hydro", "tracer", cast=int, default=0),
              shock_flattening=(_get(blocks, "Hydro", "shockFlattening", cast=float)
                                if "shockFlattening" in blocks.get("Hydro", {}) else None))
    user_cs = False
    if "csiso" in blocks.get("Hydro", {}):
        # ISOTHERMAL builds (eos_isothermal.hpp:23-35): "csiso constant <value>", or
        # "csiso userdef", whose per-cell speeds the setup supplies
        # (DataBlock.set_iso_sound_speed; EnrollIsoSoundSpeed in the reference)
        kind = _get(blocks, "Hydro", "csiso")
        if kind == "constant":
            kw["csiso"] = _get(blocks, "Hydro", "csiso", i=1, cast=float)
        elif kind == "userdef":
            kw["csiso"] = 1.0  # placeholder until the setup hands over its array
            user_cs = True
        else:
            raise ValueError("csiso admits only constant or userdef entries")
    if mhd:
        kw["emf"] = _get(blocks, "Hydro", "emf", default="uct_contact")
    run = dict(tstop=_get(blocks, "TimeIntegrator", "tstop", cast=float),
               output=blocks.get("Output", {}), csiso_userdef=user_cs)
    return kw, run
