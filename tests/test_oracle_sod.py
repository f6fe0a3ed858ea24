"""Pins the oracle's hydro path (PLM + HLLC/Roe + RK2 + outflow boundaries) to the exact
Riemann solution, and the exact solver itself to Toro's published star states."""
import numpy as np
import pytest

from oracle import oracle as O
from idefix_amd import problems
from tests.exact_riemann import sample, star_state

# Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, 3rd ed., Table 4.2:
# (rhoL, uL, pL, rhoR, uR, pR) -> (p*, u*), five significant digits as published
TORO_TABLE_4_2 = [
    ((1.0, 0.0, 1.0, 0.125, 0.0, 0.1), (0.30313, 0.92745)),
    ((1.0, -2.0, 0.4, 1.0, 2.0, 0.4), (0.00189, 0.00000)),
    ((1.0, 0.0, 1000.0, 1.0, 0.0, 0.01), (460.894, 19.5975)),
    ((1.0, 0.0, 0.01, 1.0, 0.0, 100.0), (46.0950, -6.19633)),
    ((5.99924, 19.5975, 460.894, 5.99242, -6.19633, 46.0950), (1691.64, 8.68975)),
]


@pytest.mark.parametrize("case", range(5))
def test_exact_solver_matches_toro_table(case):
    st, (p_ref, u_ref) = TORO_TABLE_4_2[case]
    p, u = star_state(*st, gamma=1.4)
    assert p == pytest.approx(p_ref, rel=2e-5, abs=1e-5)
    assert u == pytest.approx(u_ref, rel=2e-5, abs=1e-5)


def _sod(solver, recon="plm", n=1000):
    cfg = O.make_cfg(1, mhd=False, recon=recon, solver=solver, nx=(n,), bc="outflow",
                     gamma=1.4, cfl=0.8, nstages=2)
    o = O.Oracle(cfg)
    Vc, _ = problems.sod(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.set_boundaries()
    assert o.run(0.2) == 0
    return o


@pytest.mark.parametrize("solver,recon", [("hllc", "plm"), ("hllc", "ppm"), ("roe", "plm"),
                                          ("hll", "plm"), ("tvdlf", "plm"),
                                          ("hllc", "limo3"), ("roe", "limo3"),
                                          ("hll", "limo3"), ("tvdlf", "limo3")])
def test_sod_pressure_error(solver, recon):
    """Mean |p - p_exact| < 2e-3: the reference's own acceptance bound for its Sod test
    (test/HD/sod/python/testidefix.py:71-77), here at the 1000-cell configuration."""
    o = _sod(solver, recon)
    x = o.coord(0)[o.beg[0]:o.end[0]]
    Vc = o.interior(o.Vc)
    exact = sample(x, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), gamma=1.4)
    err_p = np.mean(np.abs(Vc[2, 0, 0] - exact[2]))
    err_rho = np.mean(np.abs(Vc[0, 0, 0] - exact[0]))
    assert o.t == pytest.approx(0.2, abs=1e-14)
    assert err_p < 2e-3, err_p
    assert err_rho < 5e-3, err_rho


def test_sod_converges_first_order_at_discontinuities():
    """L1 errors of a shock-capturing scheme fall roughly like 1/N: halving the cell size
    must reduce the error by a clear factor (a wrong flux or wave speed stalls it)."""
    errs = []
    for n in (200, 800):
        o = _sod("hllc", n=n)
        x = o.coord(0)[o.beg[0]:o.end[0]]
        exact = sample(x, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), gamma=1.4)
        errs.append(np.mean(np.abs(o.interior(o.Vc)[0, 0, 0] - exact[0])))
    assert errs[0] / errs[1] > 2.5, errs


def test_sod_conservation_before_waves_reach_boundaries():
    """Mass, momentum minus boundary pressure impulse and energy are conserved by the
    flux form while the outflow boundaries see a uniform state."""
    cfg = O.make_cfg(1, mhd=False, recon="plm", solver="hllc", nx=(400,), bc="outflow",
                     gamma=1.4, cfl=0.8)
    o = O.Oracle(cfg)
    Vc, _ = problems.sod(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.set_boundaries()
    o.prim_to_cons()
    dx = 1.0 / 400
    m0 = o.interior(o.Uc)[0].sum() * dx
    e0 = o.interior(o.Uc)[2].sum() * dx
    o.run(0.15)
    o.prim_to_cons()
    m1 = o.interior(o.Uc)[0].sum() * dx
    e1 = o.interior(o.Uc)[2].sum() * dx
    assert m1 == pytest.approx(m0, rel=1e-13)
    assert e1 == pytest.approx(e0, rel=1e-13)
