"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
as the results the GPU path is compared with.  It wraps oracle/liboracle.so (built by
oracle/Makefile) and exposes the solver arrays as numpy views in the reference's
LayoutRight shapes: Vc/Uc [nvar][nz][ny][nx], Vs [dim][nz+KOFF][ny+JOFF][nx+IOFF].
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

RECON = {"constant": 1, "plm": 2, "linear": 2, "limo3": 3, "ppm": 4, "parabolic": 4}
SOLVER = {"hllc": 1, "hlld": 2, "roe": 3, "hll": 4, "tvdlf": 5}
EMF = {"arithmetic": 0, "uct0": 1, "uct_contact": 2, "uct_hll": 3, "uct_hlld": 4}
BC = {"periodic": 0, "outflow": 1, "reflective": 2, "userdef": 3}
USERDEF_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double)

ARRAYS = ["Vc", "Vs", "Uc", "InvDt", "cMax", "Flux", "ex", "ey", "ez", "ezi", "ezj",
          "exj", "exk", "eyi", "eyk", "svx", "svy", "svz", "axL", "axR", "ayL", "ayR",
          "azL", "azR", "dxL", "dxR", "dyL", "dyR", "dzL", "dzR", "Uc0", "Vs0"]


class OrcCfg(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("comps", ctypes.c_int), ("mhd", ctypes.c_int),
                ("order", ctypes.c_int), ("solver", ctypes.c_int), ("emf", ctypes.c_int),
                ("nstages", ctypes.c_int), ("nx", ctypes.c_int * 3),
                ("xbeg", ctypes.c_double * 3), ("xend", ctypes.c_double * 3),
                ("lbound", ctypes.c_int * 3), ("rbound", ctypes.c_int * 3),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cfl_max_var", ctypes.c_double), ("first_dt", ctypes.c_double),
                ("ntracer", ctypes.c_int), ("shock_flattening", ctypes.c_int),
                ("shock_smoothing", ctypes.c_double), ("isothermal", ctypes.c_int),
                ("csiso", ctypes.c_double)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, dp, ip, d, i, lp = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int), ctypes.c_double, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_long))
        cp = ctypes.POINTER(OrcCfg)
        sig = {
            "orc_new": (vp, [cp]), "orc_free": (None, [vp]), "orc_nthreads": (i, [i]),
            "orc_layout": (None, [vp, ip, ip, ip, ip]),
            "orc_array": (dp, [vp, i, lp]), "orc_coord": (dp, [vp, i, i]),
            "orc_set_boundaries": (None, [vp]), "orc_prim_to_cons": (None, [vp]),
            "orc_set_userdef": (None, [vp, USERDEF_FN, vp]),
            "orc_set_iso_sound_speed": (i, [vp, dp]),
            "orc_cons_to_prim": (None, [vp]), "orc_reset_stage": (None, [vp]),
            "orc_calc_flux": (None, [vp, i]), "orc_calc_rhs": (None, [vp, i, d]),
            "orc_find_shock": (None, [vp, vp]),
            "orc_calc_corner_emf": (None, [vp]), "orc_evolve_mag_field": (None, [vp, d]),
            "orc_reconstruct_vc_field_uc": (None, [vp]),
            "orc_evolve_stage": (None, [vp, d, d]), "orc_compute_timestep": (d, [vp]),
            "orc_check_divb": (d, [vp]), "orc_add_and_store": (None, [vp, d, d]),
            "orc_copy_begin": (None, [vp]), "orc_cycle": (i, [vp]),
            "orc_run": (i, [vp, d, ctypes.c_long]), "orc_get_t": (d, [vp]),
            "orc_get_dt": (d, [vp]), "orc_set_t": (None, [vp, d]),
            "orc_set_dt": (None, [vp, d]), "orc_get_ncycles": (ctypes.c_long, [vp]),
            "orc_face_flux": (i, [cp, i, dp, dp, dp, dp, dp]),
            "orc_plant": (i, [i]),
            "orc_vanleer": (d, [d, d]), "orc_mclim": (d, [d, d]), "orc_limo3lim": (d, [d, d, d]),
            "orc_ppm_states": (None, [d, d, d, d, d, dp, dp]),
            "orc_prim2cons_pt": (None, [cp, dp, dp]), "orc_cons2prim_pt": (None, [cp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def make_cfg(dim, comps=None, mhd=True, recon="plm", solver="hlld", emf="uct_contact",
             nstages=2, nx=(64, 1, 1), xbeg=(0.0, 0.0, 0.0), xend=(1.0, 1.0, 1.0),
             bc="periodic", gamma=5.0 / 3.0, cfl=0.6, cfl_max_var=1.1, first_dt=1e-4,
             ntracer=0, shock_flattening=None, csiso=None):
    """shock_flattening: None (off) or the smoothing value of [Hydro] shockFlattening;
    csiso: None (adiabatic) or the constant isothermal sound speed."""
    c = OrcCfg()
    c.dim = dim
    c.comps = dim if comps is None else comps
    c.mhd = 1 if mhd else 0
    c.order = RECON[recon] if isinstance(recon, str) else recon
    c.solver = SOLVER[solver] if isinstance(solver, str) else solver
    c.emf = EMF[emf] if isinstance(emf, str) else emf
    c.nstages = nstages
    nx = list(nx) + [1] * (3 - len(nx))
    for d in range(3):
        c.nx[d] = nx[d] if d < dim else 1
        c.xbeg[d] = xbeg[d]
        c.xend[d] = xend[d]
        if isinstance(bc, str):
            lb = rb = bc
        else:
            bd = bc[d] if d < len(bc) else "periodic"
            lb, rb = bd if isinstance(bd, (tuple, list)) else (bd, bd)
        c.lbound[d] = BC[lb]
        c.rbound[d] = BC[rb]
    c.gamma = gamma
    c.cfl = cfl
    c.cfl_max_var = cfl_max_var
    c.first_dt = first_dt
    c.ntracer = ntracer
    c.shock_flattening = 0 if shock_flattening is None else 1
    c.shock_smoothing = 0.0 if shock_flattening is None else float(shock_flattening)
    c.isothermal = 0 if csiso is None else 1
    c.csiso = 0.0 if csiso is None else float(csiso)
    return c


class Oracle:
    """One oracle solver state; the arrays are numpy views into the C++ storage."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._h = lib().orc_new(ctypes.byref(cfg))
        if not self._h:
            raise ValueError("oracle refused this configuration")
        np_tot = (ctypes.c_int * 3)()
        ng = (ctypes.c_int * 3)()
        nv = ctypes.c_int()
        vs = (ctypes.c_int * 3)()
        lib().orc_layout(self._h, np_tot, ng, ctypes.byref(nv), vs)
        self.np_tot = tuple(np_tot)
        self.nghost = tuple(ng)
        self.nvar = nv.value
        self.vs_shape = tuple(vs)
        self.dim = cfg.dim
        self.beg = self.nghost
        self.end = tuple(self.nghost[d] + (cfg.nx[d] if d < cfg.dim else 1) for d in range(3))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def array(self, name):
        n = ctypes.c_long()
        p = lib().orc_array(self._h, ARRAYS.index(name), ctypes.byref(n))
        if n.value == 0:
            return None
        flat = np.ctypeslib.as_array(p, shape=(n.value,))
        nx, ny, nz = self.np_tot
        if name in ("Vc", "Uc", "Flux", "Uc0"):
            return flat.reshape(self.nvar, nz, ny, nx)
        if name in ("Vs", "Vs0"):
            return flat.reshape(self.dim, self.vs_shape[2], self.vs_shape[1], self.vs_shape[0])
        return flat.reshape(nz, ny, nx)

    def coord(self, d, kind="x"):
        p = lib().orc_coord(self._h, d, {"x": 0, "xl": 1, "xr": 2, "dx": 3}[kind])
        return np.ctypeslib.as_array(p, shape=(self.np_tot[d],)).copy()

    @property
    def Vc(self):
        return self.array("Vc")

    @property
    def Vs(self):
        return self.array("Vs")

    @property
    def Uc(self):
        return self.array("Uc")

    def set_boundaries(self):
        lib().orc_set_boundaries(self._h)

    def prim_to_cons(self):
        lib().orc_prim_to_cons(self._h)

    def cons_to_prim(self):
        lib().orc_cons_to_prim(self._h)

    def reset_stage(self):
        lib().orc_reset_stage(self._h)

    def calc_flux(self, d):
        lib().orc_calc_flux(self._h, d)

    def find_shock(self):
        """Shock flags (uint8, [nz][ny][nx]) of the current Vc."""
        nx, ny, nz = self.np_tot
        out = np.zeros((nz, ny, nx), dtype=np.uint8)
        lib().orc_find_shock(self._h, out.ctypes.data_as(ctypes.c_void_p))
        return out

    def calc_rhs(self, d, dt):
        lib().orc_calc_rhs(self._h, d, dt)

    def calc_corner_emf(self):
        lib().orc_calc_corner_emf(self._h)

    def evolve_mag_field(self, dt):
        lib().orc_evolve_mag_field(self._h, dt)

    def evolve_stage(self, t, dt):
        lib().orc_evolve_stage(self._h, t, dt)

    def compute_timestep(self):
        return lib().orc_compute_timestep(self._h)

    def check_divb(self):
        return lib().orc_check_divb(self._h)

    def add_and_store(self, wl, wr):
        lib().orc_add_and_store(self._h, wl, wr)

    def copy_begin(self):
        lib().orc_copy_begin(self._h)

    def set_userdef(self, fn):
        """Enrolls fn(dir, side, t) for the user-defined sides: it writes the ghost zones of
        that side into self.Vc / self.Vs (numpy views of the oracle's arrays)."""
        if fn is None:
            self._ufn = None
            lib().orc_set_userdef(self._h, USERDEF_FN(0), None)
            return
        self._ufn = USERDEF_FN(lambda arg, d, side, t: fn(d, side, t))
        lib().orc_set_userdef(self._h, self._ufn, None)

    def set_iso_sound_speed(self, cs):
        """Per-cell isothermal sound speed ([k][j][i] over the whole array, ghosts
        included), or None for the constant csiso of the configuration."""
        if cs is None:
            ptr = None
        else:
            self._cs = np.ascontiguousarray(cs, dtype=np.float64)
            nx, ny, nz = self.np_tot
            assert self._cs.size == nx * ny * nz
            ptr = self._cs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if lib().orc_set_iso_sound_speed(self._h, ptr) != 0:
            raise ValueError("a per-cell sound speed needs an isothermal configuration")

    def cycle(self):
        return lib().orc_cycle(self._h)

    def run(self, tstop, maxcycles=-1):
        return lib().orc_run(self._h, tstop, maxcycles)

    @property
    def t(self):
        return lib().orc_get_t(self._h)

    @t.setter
    def t(self, v):
        lib().orc_set_t(self._h, v)

    @property
    def dt(self):
        return lib().orc_get_dt(self._h)

    @dt.setter
    def dt(self, v):
        lib().orc_set_dt(self._h, v)

    @property
    def ncycles(self):
        return lib().orc_get_ncycles(self._h)

    def interior(self, a):
        """Interior cells of a cell-centred array ([nv][k][j][i] or [k][j][i])."""
        b, e = self.beg, self.end
        return a[..., b[2]:e[2], b[1]:e[1], b[0]:e[0]]


class OracleGrid:
    """Grid view of an oracle state (coordinates, extents) for the initial-condition
    functions of idefix_amd.problems."""

    def __init__(self, o):
        self.o = o
        self.np_tot = o.np_tot
        self.vs_shape = o.vs_shape if o.cfg.mhd else None
        self.nvar = o.nvar
        self.dim = o.dim
        self.comps = o.cfg.comps
        self.mhd = bool(o.cfg.mhd)
        self.isothermal = bool(o.cfg.isothermal)
        self.x = [o.coord(d, "x") for d in range(3)]
        self.xl = [o.coord(d, "xl") for d in range(3)]
        self.dx = [o.coord(d, "dx") for d in range(3)]
        self.beg = list(o.beg)
        self.end = list(o.end)


def set_threads(n):
    return lib().orc_nthreads(n)


def face_flux(cfg, d, vL, vR):
    vL = np.ascontiguousarray(vL, dtype=np.float64)
    vR = np.ascontiguousarray(vR, dtype=np.float64)
    F = np.zeros(8)
    cmax = np.zeros(1)
    emf = np.zeros(7)
    dp = ctypes.POINTER(ctypes.c_double)
    wrote = lib().orc_face_flux(ctypes.byref(cfg), d, vL.ctypes.data_as(dp),
                                vR.ctypes.data_as(dp), F.ctypes.data_as(dp),
                                cmax.ctypes.data_as(dp), emf.ctypes.data_as(dp))
    return F, cmax[0], emf, bool(wrote)


PLANTS = {"hlld_sbx": 1, "contact_sv": 2, "uct0_index": 3, "riemann_phi": 4,
          "ppm_extremum": 5, "roe_alfven": 6}


class planted:
    """Context manager switching on one planted mistake of the oracle (a negative
    control: a pin must fail while it is on)."""

    def __init__(self, name):
        self.which = PLANTS[name]

    def __enter__(self):
        self.prev = lib().orc_plant(self.which)
        return self

    def __exit__(self, *exc):
        lib().orc_plant(self.prev)
        return False


def vanleer(dvp, dvm):
    return lib().orc_vanleer(dvp, dvm)


def mclim(dvp, dvm):
    return lib().orc_mclim(dvp, dvm)


def limo3lim(dvp, dvm, dx):
    return lib().orc_limo3lim(dvp, dvm, dx)


def ppm_states(vm2, vm1, v0, vp1, vp2):
    vl = ctypes.c_double()
    vr = ctypes.c_double()
    lib().orc_ppm_states(vm2, vm1, v0, vp1, vp2, ctypes.byref(vl), ctypes.byref(vr))
    return vl.value, vr.value


def prim2cons(cfg, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    u = np.zeros(8)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().orc_prim2cons_pt(ctypes.byref(cfg), v.ctypes.data_as(dp), u.ctypes.data_as(dp))
    return u


def cons2prim(cfg, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.zeros(8)
    dp = ctypes.POINTER(ctypes.c_double)
    lib().orc_cons2prim_pt(ctypes.byref(cfg), u.ctypes.data_as(dp), v.ctypes.data_as(dp))
    return v
