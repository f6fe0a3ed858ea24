"""CPU checks of the product's host side: the C-ABI library loads and exports every
entry point that include/idefix_amd.h declares (no compute calls without a GPU), the
grid matches the oracle's, and the parity comparison detects a planted mistake."""
import ctypes
import os
import re

import numpy as np
import pytest

from idefix_amd import capi
from idefix_amd.datablock import make_grid
from oracle import oracle as O
from tests import pairs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    txt = open(os.path.join(ROOT, "include", "idefix_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(idfx_\w+)\s*\(", txt, re.M)))


def test_header_declares_the_binding_list():
    assert declared_entry_points() == sorted(capi.EXPORTS)


def test_region_names_follow_the_header_enum():
    """capi.REGIONS[q] names region q of the header's IDFX_REGION_* enum, and there are
    IDFX_NREGIONS of them (DataBlock.region_times reads them by index)."""
    txt = open(os.path.join(ROOT, "include", "idefix_amd.h")).read()
    enum = {name.lower(): int(v) for name, v in re.findall(r"IDFX_REGION_(\w+)\s*=\s*(\d+)", txt)}
    n = int(re.search(r"IDFX_NREGIONS\s*=\s*(\d+)", txt).group(1))
    assert len(capi.REGIONS) == n == len(enum)
    for q, name in enumerate(capi.REGIONS):
        assert enum[name] == q, (name, q)


def test_library_exports_every_declared_symbol():
    L = capi.load()
    missing = [n for n in declared_entry_points() if not hasattr(L, n)]
    assert not missing, missing
    assert capi.strerror(capi.E_NAN) == "Nan found after integration cycle"


def test_ctx_create_rejects_bad_configs_without_gpu_work():
    """Argument validation happens before any device call."""
    L = capi.load()
    d = capi.IdfxDesc()
    d.dim, d.comps, d.mhd = 2, 2, 1
    d.nx[0], d.nx[1], d.nx[2] = 8, 8, 1
    d.recon, d.solver, d.emf = 2, capi.SOLVER["hllc"], 0  # HLLC with MHD: refused
    h = ctypes.c_void_p()
    assert L.idfx_ctx_create(ctypes.byref(d), None, None, None, ctypes.byref(h)) == capi.E_ARGS
    d.solver = capi.SOLVER["hlld"]
    d.recon = 5  # no such reconstruction
    assert L.idfx_ctx_create(ctypes.byref(d), None, None, None, ctypes.byref(h)) == capi.E_ARGS
    d.recon, d.solver = 3, capi.SOLVER["roe"]  # LimO3 with Roe MHD is not built
    assert L.idfx_ctx_create(ctypes.byref(d), None, None, None, ctypes.byref(h)) == capi.E_ARGS
    # HllMHD has no uct_hlld face store (hllMHD.hpp:449-452)
    for solver in ("hll", "tvdlf"):
        d.recon, d.solver, d.emf = 2, capi.SOLVER[solver], capi.EMF["uct_hlld"]
        assert L.idfx_ctx_create(ctypes.byref(d), None, None, None, ctypes.byref(h)) == capi.E_ARGS


def _desc(**kw):
    d = capi.IdfxDesc()
    d.dim, d.comps, d.mhd = 2, 2, 1
    d.nx[0], d.nx[1], d.nx[2] = 8, 8, 1
    d.recon, d.solver, d.emf = 2, capi.SOLVER["hlld"], capi.EMF["uct_contact"]
    d.gamma = 5.0 / 3.0
    for k, v in kw.items():
        if k == "nx":
            for q in range(3):
                d.nx[q] = v[q]
        else:
            setattr(d, k, v)
    return d


@pytest.mark.parametrize("kw", [
    dict(nx=(0, 8, 1)),                                   # empty grid
    dict(nx=(8, -3, 1)),
    dict(dim=3, comps=3, nx=(1024, 1024, 1024)),          # beyond 32-bit byte offsets
    dict(dim=0), dict(dim=4), dict(comps=1),              # COMPONENTS < DIMENSIONS
    dict(ntracer=-1), dict(ntracer=65),
    dict(isothermal=1, csiso=0.0), dict(isothermal=1, csiso=-1.0),
    dict(shock_flattening=1, shock_smoothing=0.05, recon=4),  # PPM has no flattening
    dict(emf=5), dict(emf=-1),
    dict(mhd=0, solver=capi.SOLVER["hlld"]),              # hlld needs a MHD fluid
    dict(dim=1, comps=1, mhd=0, solver=capi.SOLVER["roe"], ntracer=1, nx=(64, 1, 1)),
])
def test_ctx_create_rejects_invalid_descriptions(kw):
    """Empty or oversized grids and inconsistent options are refused with IDFX_E_ARGS
    before any device call (this runs without a GPU)."""
    L = capi.load()
    h = ctypes.c_void_p()
    d = _desc(**kw)
    assert L.idfx_ctx_create(ctypes.byref(d), None, None, None, ctypes.byref(h)) == capi.E_ARGS
    assert not h.value


@pytest.mark.parametrize("dim,nx,recon", [(1, (100,), "plm"), (2, (48, 40), "ppm"),
                                          (3, (8, 12, 10), "plm")])
def test_grid_matches_oracle(dim, nx, recon):
    ng = 2 if recon == "plm" else 3
    nghost = [ng if d < dim else 0 for d in range(3)]
    nxx = [nx[d] if d < dim else 1 for d in range(3)]
    x, xl, xr, dx = make_grid(dim, nxx, (0.0, -1.0, 0.25), (1.0, 2.0, 0.5), nghost)
    o = O.Oracle(O.make_cfg(dim, mhd=False, solver="hllc", recon=recon, nx=nx,
                            xbeg=(0.0, -1.0, 0.25), xend=(1.0, 2.0, 0.5)))
    for d in range(3):
        assert np.array_equal(x[d], o.coord(d, "x"))
        assert np.array_equal(xl[d], o.coord(d, "xl"))
        assert np.array_equal(xr[d], o.coord(d, "xr"))
        assert np.array_equal(dx[d], o.coord(d, "dx"))


class _FakeGpu:
    def __init__(self, Vc, Vs):
        self._Vc, self._Vs = Vc, Vs

    def download(self):
        return self._Vc, self._Vs


def test_parity_check_catches_a_planted_mistake():
    """Negative control of the GPU parity harness: an oracle state with one cell nudged
    by 1 ulp-scale amounts passes, a 1e-9 relative change fails."""
    kw = pairs.cfg_kwargs(2, (16, 16), emf="uct_contact")
    o = pairs.make_oracle(kw)
    from idefix_amd import problems
    Vc, Vs = problems.orszag_tang(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    o.cycle()
    good = _FakeGpu(o.Vc.copy(), o.Vs.copy())
    pairs.compare_state(o, good)
    bad_vc = o.Vc.copy()
    bad_vc[3, 0, 7, 9] *= 1 + 1e-9
    with pytest.raises(AssertionError):
        pairs.compare_state(o, _FakeGpu(bad_vc, o.Vs.copy()))
    bad_vs = o.Vs.copy()
    bad_vs[1, 0, 5, 6] += 1e-9
    with pytest.raises(AssertionError):
        pairs.compare_state(o, _FakeGpu(o.Vc.copy(), bad_vs))
