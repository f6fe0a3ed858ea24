"""Pins of the oracle's shock flattening (shockFlattening.hpp:62-168 and the flagged
branches of extrapolateToFaces.hpp): flags sit on the shock of a Sod tube and nowhere
else, follow a mirror of the state, an infinite threshold flags nothing and reproduces
the unflattened run bit for bit, and the flattened Sod tube still meets the reference's
acceptance bound.  No input of the reference's test suite enables shock flattening, so no
fixture covers it."""
import numpy as np
import pytest

from oracle import oracle as O
from idefix_amd import problems
from tests.exact_riemann import sample


def _sod(n=400, recon="plm", solver="hllc", smoothing=None, tend=0.1):
    cfg = O.make_cfg(1, mhd=False, recon=recon, solver=solver, nx=(n,), bc="outflow",
                     gamma=1.4, cfl=0.8, shock_flattening=smoothing)
    o = O.Oracle(cfg)
    Vc, _ = problems.sod(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.set_boundaries()
    assert o.run(tend) == 0
    return o


def test_flags_sit_on_the_shock_only():
    o = _sod(smoothing=0.5)
    f = o.find_shock()[0, 0]
    x = o.coord(0, "x")
    flagged = x[f != 0]
    assert len(flagged) >= 1
    # exact shock position at t = 0.1: x0 + S t with S = 1.75216 (Toro, Sod test)
    xs = 0.5 + 1.75216 * 0.1
    assert np.all(np.abs(flagged - xs) < 4 * (x[1] - x[0]))


def test_flags_follow_a_mirror_image():
    o = _sod(smoothing=0.5)
    f = o.find_shock()[0, 0]
    V = o.Vc.copy()
    o.Vc[...] = V[:, :, :, ::-1]
    o.Vc[1] = -o.Vc[1]
    g = o.find_shock()[0, 0]
    assert np.array_equal(g[1:-1], f[::-1][1:-1])


@pytest.mark.parametrize("recon", ["plm", "limo3"])
def test_infinite_threshold_is_no_flattening(recon):
    a = _sod(recon=recon, smoothing=1e300)
    b = _sod(recon=recon)
    assert np.array_equal(a.interior(a.Vc), b.interior(b.Vc))


@pytest.mark.parametrize("recon,solver", [("plm", "hllc"), ("limo3", "hllc"), ("plm", "hll"),
                                          ("limo3", "tvdlf")])
def test_flattened_sod_meets_the_reference_bound(recon, solver):
    o = _sod(n=1000, recon=recon, solver=solver, smoothing=1.0, tend=0.2)
    x = o.coord(0, "x")[o.beg[0]:o.end[0]]
    Vc = o.interior(o.Vc)
    exact = sample(x, 0.2, 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), gamma=1.4)
    assert np.mean(np.abs(Vc[2, 0, 0] - exact[2])) < 2e-3
    # the flags were active: the solution differs from the unflattened run
    plain = _sod(n=1000, recon=recon, solver=solver, tend=0.2)
    assert np.abs(Vc[0, 0, 0] - plain.interior(plain.Vc)[0, 0, 0]).max() > 1e-6

