"""Pins of the oracle's passive tracers (tracer.hpp:58-111, tracer.cpp): consistency with
the continuity equation, exact conservation in periodic boxes, and second-order
advection by a uniform flow."""
import numpy as np
import pytest

from oracle import oracle as O
from idefix_amd import problems


def _tracer_total(o, n):
    o.prim_to_cons()
    return o.interior(o.Uc)[n].sum()


@pytest.mark.parametrize("solver,recon", [("hlld", "plm"), ("hll", "ppm"), ("roe", "limo3")])
def test_uniform_tracer_stays_exactly_one(solver, recon):
    """A tracer equal to 1 everywhere has rho*tr = rho and flux F_rho * 1 * A, the mass
    flux: it evolves exactly like the density and stays 1 to the last bit."""
    cfg = O.make_cfg(2, mhd=True, recon=recon, solver=solver, emf="uct_contact",
                     nx=(32, 24), cfl=0.6, ntracer=1)
    o = O.Oracle(cfg)
    Vc, Vs = problems.orszag_tang(O.OracleGrid(o))
    assert o.nvar == 7  # RHO VX1 VX2 PRS BX1 BX2 TR0
    Vc[6] = 1.0
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    for _ in range(8):
        assert o.cycle() == 0
    assert np.all(o.interior(o.Vc)[6] == 1.0)


def test_tracers_conserved_in_periodic_3d():
    """test/MHD/OrszagTang3D/idefix.ini (hlld, CFL 0.9, 2 tracers), at 16 x 32 x 16."""
    cfg = O.make_cfg(3, mhd=True, recon="plm", solver="hlld", emf="uct_contact",
                     nx=(16, 32, 16), cfl=0.9, ntracer=2)
    o = O.Oracle(cfg)
    Vc, Vs = problems.orszag_tang_3d(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    t0 = [_tracer_total(o, 8), _tracer_total(o, 9)]
    assert t0[0] > 0 and t0[1] > 0
    for _ in range(4):
        assert o.cycle() == 0
    t1 = [_tracer_total(o, 8), _tracer_total(o, 9)]
    assert t1 == pytest.approx(t0, rel=1e-13)
    tr = o.interior(o.Vc)[8:10]
    assert np.any((tr > 1e-6) & (tr < 1 - 1e-6))  # the fronts have moved


def _advect(n):
    cfg = O.make_cfg(1, mhd=False, recon="plm", solver="hllc", nx=(n,), cfl=0.5,
                     gamma=1.4, ntracer=1)
    o = O.Oracle(cfg)
    x = o.coord(0, "x")
    Vc = np.zeros_like(o.Vc)
    Vc[0] = 1.0
    Vc[1] = 1.0
    Vc[2] = 1.0
    Vc[3] = 0.5 + 0.25 * np.sin(2 * np.pi * x)[None, None, :]
    o.Vc[...] = Vc
    o.set_boundaries()
    tr0 = o.interior(o.Vc)[3].copy()
    assert o.run(1.0) == 0
    # the carrier flow is uniform and stays so exactly
    assert np.all(o.interior(o.Vc)[0] == 1.0)
    return np.mean(np.abs(o.interior(o.Vc)[3] - tr0))


def test_tracer_advection_second_order():
    """One period of a sine wave advected by rho = v = p = 1: the L1 error falls by more
    than 3x per halving of the cell size."""
    e1, e2 = _advect(64), _advect(128)
    assert e2 < 2e-3
    assert e1 / e2 > 3.0, (e1, e2)
