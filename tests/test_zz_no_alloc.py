"""idfx_compute_timestep_noalloc (the allocation-free ComputeTimestep) through the C ABI:
on two calls in a row, the second on a smaller block, it returns exactly what
idfx_compute_timestep returns; blocks beyond its static scratch are refused."""
import ctypes

import numpy as np
import pytest

from idefix_amd import capi, problems

pytestmark = pytest.mark.gpu


def _noalloc(g):
    out = ctypes.c_double()
    st = g._L.idfx_compute_timestep_noalloc(g.ctx, g._p(g.InvDt), ctypes.byref(out), g._stream())
    return st, out.value


def _block(dim, nx, seed):
    from idefix_amd.datablock import DataBlock
    g = DataBlock(dim, nx, comps=3, emf="uct_contact", recon="plm", cfl=0.5)
    Vc, Vs = problems.random_mhd(g, seed=seed, amp=0.2)
    g.upload(Vc, Vs)
    # InvDt of a real stage (stage 0 of the fused path accumulates it)
    g.set_boundaries()
    g.prim_to_cons()
    g.reset_stage()
    g.evolve_stage(0.0, 1e-4)
    return g


def test_noalloc_timestep_matches_the_allocating_entry_point():
    import torch
    for dim, nx, seed in ((3, (160, 130, 150), 1), (2, (70, 33), 2)):
        g = _block(dim, nx, seed)
        ref = g.compute_timestep()
        st, got = _noalloc(g)
        assert st == capi.OK
        assert got == ref, (dim, got, ref)
        # the InvDt maximum placed in the last interior cell is found
        b, e = g.beg, g.end
        g.InvDt[e[2] - 1, e[1] - 1, e[0] - 1] = 10.0 * float(torch.max(g.InvDt))
        st, got = _noalloc(g)
        assert st == capi.OK and got == g.compute_timestep()
        del g


def test_noalloc_timestep_refuses_blocks_beyond_its_scratch():
    import torch
    from idefix_amd.datablock import DataBlock
    # 2^28 + a row of interior cells on a thin block (the arrays themselves stay small)
    g = DataBlock(1, (2 ** 28 + 64,), mhd=False, solver="hllc", recon="plm", bc="outflow")
    st, _ = _noalloc(g)
    assert st == capi.E_ARGS
    torch.cuda.synchronize()
