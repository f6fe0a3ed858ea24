"""ctypes binding of the C ABI in include/idefix_amd.h (libidefix_amd.so).

This is the binding a Python caller (tests, bench) uses; the C++ host code of the
package calls the same entry points directly.  Loading fails loudly when the library
is missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libidefix_amd.so")

OK = 0
E_ARGS, E_LAUNCH, E_NAN, E_DT_SMALL, E_FIRST_DT, E_ALLOC = 1, 2, 3, 4, 5, 6

SOLVER = {"hllc": 1, "hlld": 2, "roe": 3, "hll": 4, "tvdlf": 5}
RECON = {"constant": 1, "plm": 2, "linear": 2, "limo3": 3, "ppm": 4, "parabolic": 4, "mlp": 8}
EMF = {"arithmetic": 0, "uct0": 1, "uct_contact": 2, "uct_hll": 3, "uct_hlld": 4}
BC = {"periodic": 0, "outflow": 1, "reflective": 2, "userdef": 3}
PREC = {"fp64": 0, "fp32": 1}

SWEEP_INIT_UC, SWEEP_SAVE_UC0, SWEEP_INVDT_SET, SWEEP_INVDT_ADD = 1, 2, 4, 8

# exported symbols (one per declaration of include/idefix_amd.h)
EXPORTS = [
    "idfx_ctx_create", "idfx_ctx_destroy", "idfx_ctx_layout", "idfx_strerror",
    "idfx_ctx_array", "idfx_set_boundaries", "idfx_prim_to_cons", "idfx_cons_to_prim",
    "idfx_reconstruct_vc_field", "idfx_reset_stage", "idfx_calc_flux", "idfx_calc_rhs",
    "idfx_sweep", "idfx_corner_emf", "idfx_evolve_mag_field", "idfx_evolve_stage",
    "idfx_compute_timestep", "idfx_add_and_store", "idfx_check_divb", "idfx_stage",
    "idfx_cycle", "idfx_ctx_set_fusion", "idfx_ctx_set_profiling", "idfx_ctx_region_times",
    "idfx_mlp_param_count", "idfx_mlp_create", "idfx_mlp_destroy", "idfx_mlp_set_params",
    "idfx_mlp_get_params", "idfx_mlp_loss_grad", "idfx_mlp_train_step", "idfx_mlp_loss",
    "idfx_ctx_set_mlp_params", "idfx_mlp_faces", "idfx_ctx_set_iso_sound_speed",
    "idfx_enforce_boundary_dir", "idfx_reconstruct_normal_field",
    "idfx_ctx_set_userdef_boundary", "idfx_ctx_set_time", "idfx_compute_timestep_noalloc",
]

# idfx_userdef_boundary_fn: (arg, dir, side, t, Vc, Vs, stream)
USERDEF_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)

REGIONS = ["boundaries", "sweep_x", "sweep_y", "sweep_z", "corner_emf", "evolve_b",
           "cons_to_prim", "fused2d_y", "ct_fused", "mlp"]


class IdfxDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("comps", ctypes.c_int), ("mhd", ctypes.c_int),
                ("nx", ctypes.c_int * 3), ("recon", ctypes.c_int), ("solver", ctypes.c_int),
                ("emf", ctypes.c_int), ("precision", ctypes.c_int),
                ("lbound", ctypes.c_int * 3), ("rbound", ctypes.c_int * 3),
                ("gamma", ctypes.c_double), ("ntracer", ctypes.c_int),
                ("shock_flattening", ctypes.c_int), ("shock_smoothing", ctypes.c_double),
                ("isothermal", ctypes.c_int), ("csiso", ctypes.c_double)]


class IdfxError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__("%s: %s (status %d)" % (what, strerror(status), status))


_lib = None


def load():
    """Load libidefix_amd.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libidefix_amd.so not built: run `python -m idefix_amd.build` "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    i, d, vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sigs = {
        "idfx_ctx_create": [ctypes.POINTER(IdfxDesc), dp, dp, dp, ctypes.POINTER(vp)],
        "idfx_ctx_destroy": [vp],
        "idfx_ctx_layout": [vp, ip, ip, ip, ip],
        "idfx_ctx_array": [vp, i, ctypes.POINTER(vp)],
        "idfx_set_boundaries": [vp, vp, vp, vp],
        "idfx_prim_to_cons": [vp, vp, vp, vp],
        "idfx_cons_to_prim": [vp, vp, vp, vp, vp],
        "idfx_reconstruct_vc_field": [vp, vp, vp, vp],
        "idfx_reset_stage": [vp, vp, vp],
        "idfx_calc_flux": [vp, i, vp, vp, vp, vp],
        "idfx_calc_rhs": [vp, i, d, vp, vp, vp, vp],
        "idfx_sweep": [vp, i, d, vp, vp, vp, vp, i, vp],
        "idfx_corner_emf": [vp, vp, vp, vp],
        "idfx_evolve_mag_field": [vp, d, vp, vp],
        "idfx_evolve_stage": [vp, d, d, vp, vp, vp, vp, vp],
        "idfx_compute_timestep": [vp, vp, dp, vp],
        "idfx_add_and_store": [vp, d, d, vp, vp, ctypes.c_longlong, vp],
        "idfx_check_divb": [vp, vp, dp, vp],
        "idfx_stage": [vp, i, i, d, d, vp, vp, vp, vp, i, i, dp, vp],
        "idfx_cycle": [vp, i, d, d, ctypes.c_long, dp, dp, vp, vp, vp, vp, vp],
        "idfx_ctx_set_fusion": [vp, i],
        "idfx_ctx_set_profiling": [vp, i],
        "idfx_ctx_region_times": [vp, dp, ctypes.POINTER(ctypes.c_long), i],
        "idfx_mlp_param_count": [],
        "idfx_mlp_create": [i, ctypes.POINTER(vp)],
        "idfx_mlp_destroy": [vp],
        "idfx_mlp_set_params": [vp, ctypes.POINTER(ctypes.c_float)],
        "idfx_mlp_get_params": [vp, ctypes.POINTER(ctypes.c_float)],
        "idfx_mlp_loss_grad": [vp, vp, vp, i, dp, ctypes.POINTER(ctypes.c_float), vp],
        "idfx_mlp_train_step": [vp, vp, vp, i, ctypes.c_float, vp],
        "idfx_mlp_loss": [vp, dp, vp],
        "idfx_ctx_set_mlp_params": [vp, ctypes.POINTER(ctypes.c_float)],
        "idfx_ctx_set_iso_sound_speed": [vp, vp],
        "idfx_mlp_faces": [vp, i, vp, vp, vp],
        "idfx_enforce_boundary_dir": [vp, i, vp, vp, vp],
        "idfx_reconstruct_normal_field": [vp, i, vp, vp],
        "idfx_ctx_set_userdef_boundary": [vp, USERDEF_FN, vp],
        "idfx_ctx_set_time": [vp, d],
        "idfx_compute_timestep_noalloc": [vp, vp, dp, vp],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.idfx_strerror.argtypes = [ctypes.c_int]
    L.idfx_strerror.restype = ctypes.c_char_p
    _lib = L
    return L


def strerror(status):
    return load().idfx_strerror(status).decode()


def check(status, what):
    if status != OK:
        raise IdfxError(status, what)
