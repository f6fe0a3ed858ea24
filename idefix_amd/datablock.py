"""Host-side DataBlock/Fluid/TimeIntegrator for Python callers.

Mirrors the reference's host objects around the hot path:
  - uniform grid construction of GridHost::MakeGrid (src/gridHost.cpp:79-86, :200-221),
  - the Fluid arrays Vc, Vs, Uc, InvDt in their LayoutRight shapes (src/fluid/fluid.hpp:529-560),
  - TimeIntegrator::Cycle and the main loop's tstop clipping (src/timeIntegrator.cpp:246-414,
    src/main.cpp:190-225).
Device memory comes from torch (plumbing only); every computation runs in
libidefix_amd.so through the C ABI.  Construction fails when no GPU is present.
"""
import ctypes

import numpy as np

from . import capi


def make_grid(dim, nx, xbeg, xend, nghost):
    """Cell centres, left/right faces and widths per direction over np_tot cells, as
    GridHost::MakeGrid builds them for uniform patches (gridHost.cpp:79-86) and for the
    unused directions (gridHost.cpp:200-221)."""
    x, xl, xr, dx = [], [], [], []
    for d in range(3):
        n = nx[d] + 2 * nghost[d]
        if d < dim:
            ps, pe = float(xbeg[d]), float(xend[d])
            dxv = (pe - ps) / nx[d]
            ii = np.arange(n) - nghost[d]
            dxa = np.full(n, dxv)
            dx.append(dxa)
            x.append(ps + (ii + 0.5) * dxa)
            xl.append(ps + ii * dxa)
            xr.append(ps + (ii + 1) * dxa)
        else:
            dx.append(np.full(n, 0.5 - (-0.5)))
            x.append(np.full(n, 0.5 * (-0.5 + 0.5)))
            xl.append(np.full(n, -0.5))
            xr.append(np.full(n, 0.5))
    return x, xl, xr, dx


class DataBlock:
    def __init__(self, dim, nx, xbeg=(0.0, 0.0, 0.0), xend=(1.0, 1.0, 1.0), comps=None,
                 mhd=True, recon="plm", solver="hlld", emf="uct_contact", precision="fp64",
                 bc="periodic", gamma=5.0 / 3.0, nstages=2, cfl=0.6, cfl_max_var=1.1,
                 first_dt=1e-4, device="cuda:0", ntracer=0, shock_flattening=None,
                 csiso=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("DataBlock needs a GPU: the hot path has no CPU fallback")
        self.torch = torch
        self.device = torch.device(device)
        self.dim = dim
        self.comps = dim if comps is None else comps
        self.mhd = bool(mhd)
        self.nstages = nstages
        self.cfl = cfl
        self.cfl_max_var = cfl_max_var
        self.gamma = gamma
        nx = list(nx) + [1] * (3 - len(nx))
        self.nx = [nx[d] if d < dim else 1 for d in range(3)]
        self.order = capi.RECON[recon]
        ng = 2 if self.order < 4 else 3
        self.nghost = [ng if d < dim else 0 for d in range(3)]
        self.np_tot = [self.nx[d] + 2 * self.nghost[d] for d in range(3)]
        self.beg = list(self.nghost)
        self.end = [self.nghost[d] + self.nx[d] for d in range(3)]
        self.x, self.xl, self.xr, self.dx = make_grid(dim, self.nx, xbeg, xend, self.nghost)
        desc = capi.IdfxDesc()
        desc.dim, desc.comps, desc.mhd = dim, self.comps, int(self.mhd)
        for d in range(3):
            desc.nx[d] = self.nx[d]
            if isinstance(bc, str):
                lb = rb = bc
            else:
                bd = bc[d] if d < len(bc) else "periodic"
                lb, rb = bd if isinstance(bd, (tuple, list)) else (bd, bd)
            desc.lbound[d] = capi.BC[lb]
            desc.rbound[d] = capi.BC[rb]
        desc.recon = self.order
        desc.solver = capi.SOLVER[solver]
        desc.emf = capi.EMF[emf]
        desc.precision = capi.PREC[precision]
        desc.gamma = gamma
        desc.ntracer = ntracer
        # None: off; else the smoothing value of [Hydro] shockFlattening
        desc.shock_flattening = 0 if shock_flattening is None else 1
        desc.shock_smoothing = 0.0 if shock_flattening is None else float(shock_flattening)
        # None: adiabatic; else the constant isothermal sound speed
        desc.isothermal = 0 if csiso is None else 1
        desc.csiso = 0.0 if csiso is None else float(csiso)
        self.isothermal = csiso is not None
        self.ntracer = ntracer
        L = capi.load()
        self._L = L
        h = ctypes.c_void_p()
        dps = [np.ascontiguousarray(a, dtype=np.float64) for a in self.dx]
        self._dx_host = dps
        dp = ctypes.POINTER(ctypes.c_double)
        capi.check(L.idfx_ctx_create(ctypes.byref(desc), dps[0].ctypes.data_as(dp),
                                     dps[1].ctypes.data_as(dp), dps[2].ctypes.data_as(dp),
                                     ctypes.byref(h)), "idfx_ctx_create")
        self.ctx = h
        np_tot = (ctypes.c_int * 3)()
        ngh = (ctypes.c_int * 3)()
        nv = ctypes.c_int()
        vs = (ctypes.c_int * 3)()
        capi.check(L.idfx_ctx_layout(h, np_tot, ngh, ctypes.byref(nv), vs), "layout")
        self.nvar = nv.value
        self.vs_shape = list(vs)
        kw = dict(dtype=torch.float64, device=self.device)
        nxt, nyt, nzt = self.np_tot
        self.Vc = torch.zeros((self.nvar, nzt, nyt, nxt), **kw)
        self.Uc = torch.zeros((self.nvar, nzt, nyt, nxt), **kw)
        self.InvDt = torch.zeros((nzt, nyt, nxt), **kw)
        if self.mhd:
            self.Vs = torch.zeros((dim, vs[2], vs[1], vs[0]), **kw)
        else:
            self.Vs = None
        self.t = 0.0
        self.dt = first_dt
        self.ncycles = 0

    def __del__(self):
        ctx = getattr(self, "ctx", None)
        if ctx is not None and ctx.value:
            self._L.idfx_ctx_destroy(ctx)
            self.ctx = None

    # ---- helpers -------------------------------------------------------------------
    def _p(self, t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def interior(self, a):
        b, e = self.beg, self.end
        return a[..., b[2]:e[2], b[1]:e[1], b[0]:e[0]]

    # ---- reference operations ------------------------------------------------------
    def set_boundaries(self):
        capi.check(self._L.idfx_set_boundaries(self.ctx, self._p(self.Vc), self._p(self.Vs),
                                               self._stream()), "SetBoundaries")

    def prim_to_cons(self):
        capi.check(self._L.idfx_prim_to_cons(self.ctx, self._p(self.Vc), self._p(self.Uc),
                                             self._stream()), "PrimToCons")

    def cons_to_prim(self):
        capi.check(self._L.idfx_cons_to_prim(self.ctx, self._p(self.Uc), self._p(self.Vc),
                                             self._p(self.Vs), self._stream()), "ConsToPrim")

    def reset_stage(self):
        capi.check(self._L.idfx_reset_stage(self.ctx, self._p(self.InvDt), self._stream()),
                   "ResetStage")

    def calc_flux(self, d, flux):
        capi.check(self._L.idfx_calc_flux(self.ctx, d, self._p(self.Vc), self._p(self.Vs),
                                          self._p(flux), self._stream()), "CalcFlux")

    def calc_rhs(self, d, dt, flux):
        capi.check(self._L.idfx_calc_rhs(self.ctx, d, dt, self._p(flux), self._p(self.Uc),
                                         self._p(self.InvDt), self._stream()), "CalcRHS")

    def sweep(self, d, dt, flags):
        capi.check(self._L.idfx_sweep(self.ctx, d, dt, self._p(self.Vc), self._p(self.Vs),
                                      self._p(self.Uc), self._p(self.InvDt), flags,
                                      self._stream()), "sweep")

    def calc_corner_emf(self):
        capi.check(self._L.idfx_corner_emf(self.ctx, self._p(self.Vc), self._p(self.Vs),
                                           self._stream()), "CalcCornerEMF")

    def evolve_mag_field(self, dt):
        capi.check(self._L.idfx_evolve_mag_field(self.ctx, dt, self._p(self.Vs),
                                                 self._stream()), "EvolveMagField")

    def evolve_stage(self, t, dt):
        capi.check(self._L.idfx_evolve_stage(self.ctx, t, dt, self._p(self.Vc), self._p(self.Vs),
                                             self._p(self.Uc), self._p(self.InvDt),
                                             self._stream()), "EvolveStage")

    def compute_timestep(self):
        out = ctypes.c_double()
        capi.check(self._L.idfx_compute_timestep(self.ctx, self._p(self.InvDt),
                                                 ctypes.byref(out), self._stream()),
                   "ComputeTimestep")
        return out.value

    def check_divb(self):
        out = ctypes.c_double()
        capi.check(self._L.idfx_check_divb(self.ctx, self._p(self.Vs), ctypes.byref(out),
                                           self._stream()), "CheckDivB")
        return out.value

    def ctx_array(self, which):
        """Context-owned scratch array (face/corner EMFs, cMax) as a torch view; the
        uct_contact upwind weights (9..11, stored as byte codes 2 w) as a float64 copy."""
        ptr = ctypes.c_void_p()
        capi.check(self._L.idfx_ctx_array(self.ctx, which, ctypes.byref(ptr)), "ctx_array")
        shape = (self.np_tot[2], self.np_tot[1], self.np_tot[0])
        code = 9 <= which <= 11

        class _View:  # zero-copy wrapper through the CUDA array interface
            __cuda_array_interface__ = {"shape": shape, "typestr": "|u1" if code else "<f8",
                                        "data": (ptr.value, False), "version": 2,
                                        "strides": None}
        t = self.torch.as_tensor(_View(), device=self.device)
        return t.double() * 0.5 if code else t

    # ---- driver ---------------------------------------------------------------------
    def stage(self, stage, dt, skip_boundaries=False, write_uc=True, want_dt=False):
        newdt = ctypes.c_double(0.0)
        capi.check(self._L.idfx_stage(self.ctx, stage, self.nstages, dt, self.cfl,
                                      self._p(self.Vc), self._p(self.Vs), self._p(self.Uc),
                                      self._p(self.InvDt), int(skip_boundaries), int(write_uc),
                                      ctypes.byref(newdt) if want_dt else None, self._stream()),
                   "stage")
        return newdt.value

    def set_userdef_boundary(self, fn):
        """Boundary::EnrollUserDefBoundary: fn(dir, side, t, Vc, Vs) fills the ghost zones
        of a user-defined side (side 0 left, 1 right) of direction dir; Vc and Vs are the
        torch tensors the calling entry point works on (the block's own arrays on the
        per-kernel schedule that user-defined sides select), and its work goes on the
        current stream.  None unenrolls."""
        if fn is None:
            self._ufn = None
            capi.check(self._L.idfx_ctx_set_userdef_boundary(self.ctx, capi.USERDEF_FN(0), None),
                       "set_userdef_boundary")
            return

        def trampoline(arg, d, side, t, vc, vs, stream):
            assert vc == self.Vc.data_ptr() and (vs or 0) == (self.Vs.data_ptr() if self.mhd else 0)
            fn(d, side, t, self.Vc, self.Vs)
        self._ufn = capi.USERDEF_FN(trampoline)
        capi.check(self._L.idfx_ctx_set_userdef_boundary(self.ctx, self._ufn, None),
                   "set_userdef_boundary")

    def set_time(self, t):
        """Time passed to the user boundary by set_boundaries() and stage()."""
        capi.check(self._L.idfx_ctx_set_time(self.ctx, t), "set_time")

    def set_iso_sound_speed(self, cs):
        """Per-cell isothermal sound speed ("csiso userdef"): an array of shape
        (nz, ny, nx) over all cells, ghosts included (numpy or torch), kept on the device
        and read by every sweep; None returns to the constant csiso."""
        if cs is None:
            self._cs = None
            capi.check(self._L.idfx_ctx_set_iso_sound_speed(self.ctx, None),
                       "set_iso_sound_speed")
            return
        t = self.torch.as_tensor(np.asarray(cs) if not self.torch.is_tensor(cs) else cs,
                                 dtype=self.torch.float64, device=self.device).contiguous()
        nx, ny, nz = self.np_tot
        if tuple(t.shape) != (nz, ny, nx):
            raise ValueError("sound speed array must have shape %s" % ((nz, ny, nx),))
        self._cs = t
        capi.check(self._L.idfx_ctx_set_iso_sound_speed(self.ctx, ctypes.c_void_p(t.data_ptr())),
                   "set_iso_sound_speed")

    def set_mlp_params(self, params):
        """Loads packed FP32 learned-limiter parameters (numpy, idfx_mlp_param_count()
        floats) for BF16 inference, or unloads them with None (PPM weights)."""
        if params is None:
            capi.check(self._L.idfx_ctx_set_mlp_params(self.ctx, None), "set_mlp_params")
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        capi.check(self._L.idfx_ctx_set_mlp_params(
            self.ctx, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "set_mlp_params")

    def mlp_faces(self, d):
        """Interface weights of direction d from the current Vc: torch float32
        [16, nz, ny, nx] (face index = cell right of the face)."""
        nxt, nyt, nzt = self.np_tot
        W = self.torch.zeros((16, nzt, nyt, nxt), dtype=self.torch.float32, device=self.device)
        capi.check(self._L.idfx_mlp_faces(self.ctx, d, self._p(self.Vc), ctypes.c_void_p(W.data_ptr()),
                                          self._stream()), "mlp_faces")
        return W

    def set_fusion(self, level):
        """Kernel schedule of cycle(): 0 = one kernel per reference step, 1 = the fused
        schedules (2-D MHD second half; 3-D ping-pong stage, with the fused CT kernel on
        blocks of at least 2^26 interior cells), the default where supported, 2 = as 1 with
        the fused 3-D CT kernel at every size.  Results are identical."""
        capi.check(self._L.idfx_ctx_set_fusion(self.ctx, int(level)), "set_fusion")

    def set_profiling(self, on):
        """Region profiler (idfx::pushRegion/popRegion with profiling on, global.cpp:66-95):
        HIP event pairs around each region of cycle() and stage()."""
        capi.check(self._L.idfx_ctx_set_profiling(self.ctx, int(bool(on))), "set_profiling")

    def region_times(self):
        """{region: (total ms, launches)} since the last call; resets the totals."""
        n = len(capi.REGIONS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        capi.check(self._L.idfx_ctx_region_times(self.ctx, ms, cnt, n), "region_times")
        return {name: (ms[q], cnt[q]) for q, name in enumerate(capi.REGIONS) if cnt[q] > 0}

    def cycle(self):
        """TimeIntegrator::Cycle on the fused path (timeIntegrator.cpp:246-414)."""
        t = ctypes.c_double(self.t)
        dt = ctypes.c_double(self.dt)
        st = self._L.idfx_cycle(self.ctx, self.nstages, self.cfl, self.cfl_max_var,
                                self.ncycles, ctypes.byref(t), ctypes.byref(dt),
                                self._p(self.Vc), self._p(self.Vs), self._p(self.Uc),
                                self._p(self.InvDt), self._stream())
        self.t, self.dt = t.value, dt.value
        self.ncycles += 1
        capi.check(st, "Cycle")

    def run(self, tstop, maxcycles=-1):
        """Main loop of src/main.cpp:190-225 (without outputs)."""
        while self.t < tstop:
            if tstop - self.t < self.dt:
                self.dt = tstop - self.t
            self.cycle()
            if 0 <= maxcycles <= self.ncycles:
                break

    # ---- state transfer (DataBlockHost::SyncToDevice / SyncFromDevice) ---------------
    def upload(self, Vc, Vs=None):
        self.Vc.copy_(self.torch.from_numpy(np.ascontiguousarray(Vc)))
        if self.mhd and Vs is not None:
            self.Vs.copy_(self.torch.from_numpy(np.ascontiguousarray(Vs)))

    def download(self):
        self.torch.cuda.synchronize(self.device)
        Vc = self.Vc.cpu().numpy()
        Vs = self.Vs.cpu().numpy() if self.mhd else None
        return Vc, Vs

