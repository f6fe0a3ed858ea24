"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bar: FP64 results within 1e-12 relative per cell of the oracle (absolute floor
1e-12 x field scale), interior cells and interior faces, after each checked step.
"""
import numpy as np
import pytest

from idefix_amd import problems
from tests.pairs import cfg_kwargs, compare_state, make_pair, rel_err

pytestmark = pytest.mark.gpu


def _cycle_both(o, g, ncycles):
    for n in range(ncycles):
        so = o.cycle()
        g.cycle()
        assert so == 0
        assert g.t == pytest.approx(o.t, rel=1e-14, abs=0)
        assert g.dt == pytest.approx(o.dt, rel=1e-12, abs=0), (n, g.dt, o.dt)


MHD2D_EMFS = ["uct_hlld", "uct_contact", "arithmetic", "uct0", "uct_hll"]


@pytest.mark.parametrize("emf", MHD2D_EMFS)
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_orszag_tang_cycles(emf, recon):
    kw = cfg_kwargs(2, (48, 40), emf=emf, recon=recon, cfl=0.6)
    o, g = make_pair(kw, problems.orszag_tang)
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT %s %s" % (emf, recon))


@pytest.mark.parametrize("emf", ["uct_hlld", "uct_contact", "arithmetic"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_orszag_tang_roe_cycles(emf, recon):
    kw = cfg_kwargs(2, (48, 40), emf=emf, recon=recon, solver="roe", cfl=0.5)
    o, g = make_pair(kw, problems.orszag_tang)
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT roe %s %s" % (emf, recon))


@pytest.mark.parametrize("solver", ["hlld", "roe"])
@pytest.mark.parametrize("emf", ["uct_hlld", "uct_contact"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_mhd3d_cycles(emf, recon, solver):
    kw = cfg_kwargs(3, (16, 12, 10), emf=emf, recon=recon, cfl=0.4, nstages=3, solver=solver)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=3, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="3D %s %s %s" % (solver, emf, recon))


@pytest.mark.parametrize("emf", ["arithmetic", "uct0"])
@pytest.mark.parametrize("nstages", [2, 3])
def test_mhd3d_arithmetic_uct0_cycles(emf, nstages):
    """The two corner-EMF averages the 3-D cases above leave out (calcCornerEmf.hpp:14-126),
    RK2 and RK3."""
    kw = cfg_kwargs(3, (16, 12, 10), emf=emf, recon="plm", cfl=0.4, nstages=nstages,
                    solver="hlld")
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=5, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="3D hlld %s RK%d" % (emf, nstages))


@pytest.mark.parametrize("solver", ["hllc", "roe"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_sod_cycles(solver, recon):
    kw = cfg_kwargs(1, (1000,), mhd=False, solver=solver, recon=recon, bc="outflow",
                    gamma=1.4, cfl=0.8)
    o, g = make_pair(kw, problems.sod)
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="sod %s %s" % (solver, recon))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("solver", ["hllc", "roe"])
def test_hd_multid_cycles(dim, solver):
    nx = (40, 24) if dim == 2 else (16, 12, 10)
    bc = ["outflow", "periodic"] if dim == 2 else "periodic"
    kw = cfg_kwargs(dim, nx, mhd=False, solver=solver, recon="plm", bc=bc, gamma=1.4, cfl=0.5)
    o, g = make_pair(kw, lambda gg: problems.random_hd(gg, seed=5, amp=0.3))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="HD%dD %s" % (dim, solver))


@pytest.mark.parametrize("solver", ["hll", "tvdlf"])
@pytest.mark.parametrize("emf", ["uct_contact", "uct_hll", "arithmetic", "uct0"])
def test_orszag_tang_hll_tvdlf_cycles(emf, solver):
    """HllMHD and TvdlfMHD (test/MHD/OrszagTang/idefix-hll.ini, idefix-tvdlf.ini: PLM, RK2,
    CFL 0.6)."""
    kw = cfg_kwargs(2, (48, 40), emf=emf, solver=solver, cfl=0.6)
    o, g = make_pair(kw, problems.orszag_tang)
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT %s %s" % (solver, emf))


@pytest.mark.parametrize("solver", ["hll", "tvdlf"])
@pytest.mark.parametrize("emf", ["uct_hll", "uct_contact"])
def test_mhd3d_hll_tvdlf_cycles(emf, solver):
    kw = cfg_kwargs(3, (16, 12, 10), emf=emf, solver=solver, cfl=0.4, nstages=3)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=4, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="3D %s %s" % (solver, emf))


@pytest.mark.parametrize("solver", ["hll", "tvdlf"])
def test_sod_hll_tvdlf_cycles(solver):
    """HllHD and TvdlfHD on the Sod tube (test/HD/sod/idefix-hll.ini, idefix-tvdlf.ini)."""
    kw = cfg_kwargs(1, (1000,), mhd=False, solver=solver, recon="plm", bc="outflow",
                    gamma=1.4, cfl=0.8)
    o, g = make_pair(kw, problems.sod)
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="sod %s" % solver)


@pytest.mark.parametrize("solver", ["hll", "tvdlf"])
@pytest.mark.parametrize("dim", [2, 3])
def test_hd_multid_hll_tvdlf_cycles(dim, solver):
    nx = (40, 24) if dim == 2 else (16, 12, 10)
    bc = ["outflow", "periodic"] if dim == 2 else "periodic"
    kw = cfg_kwargs(dim, nx, mhd=False, solver=solver, recon="plm", bc=bc, gamma=1.4, cfl=0.5)
    o, g = make_pair(kw, lambda gg: problems.random_hd(gg, seed=6, amp=0.3))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="HD%dD %s" % (dim, solver))


@pytest.mark.parametrize("solver,emf,recon", [("hlld", "uct_contact", "limo3"),
                                              ("hlld", "uct_hlld", "limo3"),
                                              ("hll", "uct_hll", "limo3"),
                                              ("tvdlf", "uct_contact", "limo3"),
                                              ("hll", "uct_contact", "ppm"),
                                              ("tvdlf", "uct_hll", "ppm")])
def test_orszag_tang_reconstruction_matrix(solver, emf, recon):
    """The reference's OrszagTang test runs every solver input with reconstructions 2, 3
    and 4 (test/MHD/OrszagTang/testme.py:46-53); the LimO3 and PPM members here."""
    kw = cfg_kwargs(2, (48, 40), emf=emf, solver=solver, recon=recon, cfl=0.6)
    o, g = make_pair(kw, problems.orszag_tang)
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT %s %s %s" % (solver, emf, recon))


def test_mhd3d_limo3_cycles():
    kw = cfg_kwargs(3, (16, 12, 10), emf="uct_contact", solver="hlld", recon="limo3",
                    cfl=0.4, nstages=3)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=8, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="3D hlld limo3")


@pytest.mark.parametrize("solver,recon", [("hllc", "limo3"), ("roe", "limo3"),
                                          ("hll", "limo3"), ("tvdlf", "limo3"),
                                          ("hll", "ppm"), ("tvdlf", "ppm")])
def test_sod_reconstruction_matrix(solver, recon):
    """test/HD/sod/testme.py runs its solver inputs with reconstructions 2-4."""
    kw = cfg_kwargs(1, (1000,), mhd=False, solver=solver, recon=recon, bc="outflow",
                    gamma=1.4, cfl=0.8)
    o, g = make_pair(kw, problems.sod)
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="sod %s %s" % (solver, recon))


def test_hd2d_limo3_cycles():
    kw = cfg_kwargs(2, (40, 24), mhd=False, solver="hllc", recon="limo3",
                    bc=["outflow", "periodic"], gamma=1.4, cfl=0.5)
    o, g = make_pair(kw, lambda gg: problems.random_hd(gg, seed=9, amp=0.3))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="HD2D limo3")


def _with_tracers(ic, nfluid, ntr):
    """Initial state of `ic` plus ntr tracers: a step in x and a smooth profile in y."""
    def f(gg):
        Vc, Vs = ic(gg)
        x = np.asarray(gg.x[0])[None, None, :]
        y = np.asarray(gg.x[1])[None, :, None]
        for t in range(ntr):
            Vc[nfluid + t] = (np.where(x > 0.5, 1.0, 0.0) + 0 * y) if t % 2 == 0 else \
                (0.5 + 0.4 * np.sin(2 * np.pi * (y + 0.3 * t)) + 0 * x)
        return Vc, Vs
    return f


def test_orszag_tang_3d_tracers_reference_config():
    """test/MHD/OrszagTang3D/idefix.ini: 32 x 64 x 32, hlld, CFL 0.9, 2 tracers
    (TRG = [x > 1/2], TRG+1 = [z > 1/2] as its setup)."""
    kw = cfg_kwargs(3, (32, 64, 32), emf="uct_contact", solver="hlld", cfl=0.9, ntracer=2)
    o, g = make_pair(kw, problems.orszag_tang_3d)
    assert g.nvar == 10 and o.nvar == 10
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="OT3D tracers")


@pytest.mark.parametrize("emf,recon", [("uct_hlld", "plm"), ("uct_contact", "ppm"),
                                       ("uct_hll", "limo3")])
def test_orszag_tang_2d_tracers(emf, recon):
    kw = cfg_kwargs(2, (48, 40), emf=emf, recon=recon, cfl=0.6, ntracer=2)
    o, g = make_pair(kw, _with_tracers(problems.orszag_tang, 6, 2))
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT tracers %s %s" % (emf, recon))


def test_uniform_tracer_stays_one_on_gpu():
    kw = cfg_kwargs(2, (48, 40), emf="uct_contact", cfl=0.6, ntracer=1)
    def ic(gg):
        Vc, Vs = problems.orszag_tang(gg)
        Vc[6] = 1.0
        return Vc, Vs
    o, g = make_pair(kw, ic)
    g.set_boundaries()
    for _ in range(10):
        g.cycle()
    Vc, _ = g.download()
    b, e = g.beg, g.end
    assert np.all(Vc[6, 0, b[1]:e[1], b[0]:e[0]] == 1.0)


@pytest.mark.parametrize("bc", [["outflow", "periodic"], ["reflective", "periodic"]])
def test_tracers_with_outflow_and_reflective_sides(bc):
    """Tracers in the ghost fill of every side type; with MHD the reflective fill flips
    their sign like the transverse field (boundary.hpp:697)."""
    kw = cfg_kwargs(2, (32, 24), emf="uct_contact", bc=bc, cfl=0.4, ntracer=2)
    o, g = make_pair(kw, _with_tracers(lambda gg: problems.random_mhd(gg, seed=12, amp=0.1), 6, 2))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="tracers bc %s" % bc)


@pytest.mark.parametrize("dim,solver", [(1, "hllc"), (2, "hll"), (3, "tvdlf")])
def test_hd_tracers(dim, solver):
    nx = {1: (400,), 2: (40, 24), 3: (16, 12, 10)}[dim]
    bc = {1: "outflow", 2: ["outflow", "periodic"], 3: "periodic"}[dim]
    ic = problems.sod if dim == 1 else (lambda gg: problems.random_hd(gg, seed=5, amp=0.3))
    kw = cfg_kwargs(dim, nx, mhd=False, solver=solver, recon="plm", bc=bc, gamma=1.4,
                    cfl=0.5, ntracer=2)
    nfluid = dim + 2

    def f(gg):
        Vc, Vs = ic(gg)
        x = np.asarray(gg.x[0])[None, None, :]
        Vc[nfluid] = np.where(x > 0.3, 1.0, 0.0) + 0 * Vc[0]
        Vc[nfluid + 1] = 0.5 + 0.4 * np.cos(6 * np.pi * x) + 0 * Vc[0]
        return Vc, Vs
    o, g = make_pair(kw, f)
    o.set_boundaries()
    _cycle_both(o, g, 8)
    compare_state(o, g, what="HD%dD tracers" % dim)


@pytest.mark.parametrize("solver,recon,nstages", [
    (s, r, 2) for s in ("roe", "hll", "hlld", "tvdlf") for r in ("plm", "limo3")] +
    [("roe", "ppm", 3), ("hlld", "ppm", 3)])
def test_brio_wu_reference_matrix(solver, recon, nstages):
    """test/MHD/sod (Brio-Wu, 800 cells on [0, 100], CFL 0.8, gamma 5/3): its inputs
    idefix.ini / -hll / -hlld / -tvdlf with reconstructions 2 and 3, and the RK3 inputs
    with reconstruction 4 (test/MHD/sod/testme.py:18-20)."""
    kw = cfg_kwargs(1, (800,), comps=2, solver=solver, recon=recon, bc="outflow",
                    gamma=1.66666666666667, cfl=0.8, nstages=nstages, xend=(100, 1, 1))
    o, g = make_pair(kw, problems.brio_wu)
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="Brio-Wu %s %s" % (solver, recon))


def test_fast_wave_1d_periodic():
    kw = cfg_kwargs(1, (128,), comps=2, solver="hlld", recon="plm", cfl=0.8)
    o, g = make_pair(kw, lambda gg: problems.linear_fast_wave_1d(gg, eps=1e-3)[:2])
    o.set_boundaries()
    _cycle_both(o, g, 30)
    compare_state(o, g, what="1-D fast wave")


@pytest.mark.parametrize("solver,recon", [("hllc", "plm"), ("hll", "limo3"), ("tvdlf", "plm")])
def test_sod_shock_flattening(solver, recon):
    kw = cfg_kwargs(1, (1000,), mhd=False, solver=solver, recon=recon, bc="outflow",
                    gamma=1.4, cfl=0.8, shock_flattening=0.5)
    o, g = make_pair(kw, problems.sod)
    o.set_boundaries()
    _cycle_both(o, g, 60)
    compare_state(o, g, what="sod flattening %s %s" % (solver, recon))


@pytest.mark.parametrize("emf,recon", [("uct_hlld", "plm"), ("uct_contact", "limo3"),
                                       ("uct_hll", "plm")])
def test_orszag_tang_shock_flattening(emf, recon):
    """Flags on the fused 2-D path (both sweeps read them), from a perturbed state that
    has compressive cells from the start."""
    kw = cfg_kwargs(2, (48, 40), emf=emf, recon=recon, cfl=0.6, shock_flattening=0.05)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=23, amp=0.3))
    o.set_boundaries()
    _cycle_both(o, g, 12)
    compare_state(o, g, what="OT flattening %s %s" % (emf, recon))


def test_shock_flattening_3d_and_hd2d():
    kw = cfg_kwargs(3, (16, 12, 10), emf="uct_contact", cfl=0.4, shock_flattening=0.05)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=21, amp=0.3))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="3D flattening")
    kw = cfg_kwargs(2, (40, 24), mhd=False, solver="hllc", bc=["outflow", "periodic"],
                    gamma=1.4, cfl=0.5, shock_flattening=0.05)
    o, g = make_pair(kw, lambda gg: problems.random_hd(gg, seed=22, amp=0.4))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="HD2D flattening")


@pytest.mark.parametrize("solver,recon,nstages", [
    ("roe", "plm", 2), ("hll", "plm", 2), ("hllc", "plm", 2), ("tvdlf", "plm", 2),
    ("hllc", "limo3", 2), ("roe", "ppm", 3), ("hllc", "ppm", 3)])
def test_sod_iso_reference_matrix(solver, recon, nstages):
    """test/HD/sod-iso (500 cells, CFL 0.8, csiso 1): idefix.ini / -hll / -hllc / -tvdlf and
    the RK3 inputs with PPM (test/HD/sod-iso/testme.py:18-20)."""
    kw = cfg_kwargs(1, (500,), mhd=False, solver=solver, recon=recon, bc="outflow",
                    cfl=0.8, nstages=nstages, csiso=1.0)
    o, g = make_pair(kw, problems.sod_iso)
    assert g.nvar == 2
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="sod-iso %s %s" % (solver, recon))


@pytest.mark.parametrize("solver,recon,nstages", [
    ("hlld", "plm", 2), ("hll", "plm", 2), ("tvdlf", "plm", 2), ("hlld", "limo3", 2),
    ("hlld", "ppm", 3), ("roe", "plm", 2), ("roe", "limo3", 2), ("roe", "ppm", 3)])
def test_brio_wu_iso_reference_matrix(solver, recon, nstages):
    """test/MHD/sod-iso (800 cells on [0, 100], CFL 0.8, csiso 1): idefix.ini (roe),
    -hll, -hlld, -tvdlf and the RK3 inputs (roe, hlld) with PPM
    (test/MHD/sod-iso/testme.py:17-20)."""
    kw = cfg_kwargs(1, (800,), comps=2, solver=solver, recon=recon, bc="outflow", cfl=0.8,
                    nstages=nstages, xend=(100, 1, 1), csiso=1.0)
    o, g = make_pair(kw, problems.brio_wu_iso)
    o.set_boundaries()
    _cycle_both(o, g, 40)
    compare_state(o, g, what="Brio-Wu iso %s %s" % (solver, recon))


@pytest.mark.parametrize("solver,emf", [("hlld", "uct_hlld"), ("hlld", "uct_contact"),
                                        ("roe", "uct_contact"), ("roe", "uct_hlld")])
def test_mhd2d_iso_fused(solver, emf):
    """Isothermal 2-D MHD through the fused 2-D kernel."""
    kw = cfg_kwargs(2, (48, 40), solver=solver, emf=emf, cfl=0.6, csiso=0.8)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=31, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="2-D MHD iso %s %s" % (solver, emf))


def test_hd2d_iso_with_flattening():
    kw = cfg_kwargs(2, (40, 24), mhd=False, solver="hllc", bc=["outflow", "periodic"],
                    cfl=0.5, csiso=1.0, shock_flattening=0.05)
    o, g = make_pair(kw, lambda gg: problems.random_hd(gg, seed=32, amp=0.4))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="HD2D iso flattening")


def test_outflow_mhd2d():
    kw = cfg_kwargs(2, (32, 24), emf="uct_contact", bc=["outflow", "periodic"], cfl=0.4)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=11, amp=0.1))
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="outflow MHD")


@pytest.mark.parametrize("emf", ["uct_hlld", "uct_contact"])
def test_unfused_kernels_match_oracle(emf):
    """CalcFlux (FluxRiemann, cMax, face EMFs), CalcRightHandSide, CalcCornerEMF and
    EvolveMagField one at a time, as the reference sequences them."""
    import torch
    kw = cfg_kwargs(2, (40, 36), emf=emf)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=7, amp=0.25))
    o.set_boundaries()
    g.set_boundaries()
    o.prim_to_cons()
    g.prim_to_cons()
    o.reset_stage()
    g.reset_stage()
    dt = 1e-3
    flux = torch.zeros_like(g.Uc)
    names = {"uct_hlld": [(3, "ezi"), (12, "axL"), (13, "axR"), (18, "dxL"), (19, "dxR")],
             "uct_contact": [(3, "ezi"), (9, "svx")]}[emf]
    ynames = {"ezi": "ezj", "axL": "ayL", "axR": "ayR", "dxL": "dyL", "dxR": "dyR",
              "svx": "svy"}
    yidx = {"ezj": 4, "ayL": 14, "ayR": 15, "dyL": 20, "dyR": 21, "svy": 10}
    ext = 2 if emf == "uct_hlld" else 1
    b, e = o.beg, o.end
    for d in range(2):
        o.calc_flux(d)
        g.calc_flux(d, flux)
        torch.cuda.synchronize()
        # faces [beg, end+1) along d, [beg-ext, end+ext) across
        sl = [slice(None), slice(0, 1), None, None]
        if d == 0:
            sl[2], sl[3] = slice(b[1] - ext, e[1] + ext), slice(b[0], e[0] + 1)
        else:
            sl[2], sl[3] = slice(b[1], e[1] + 1), slice(b[0] - ext, e[0] + ext)
        assert rel_err(flux.cpu().numpy()[tuple(sl)], o.array("Flux")[tuple(sl)]) <= 1e-12
        for which, name in names:
            if d == 1:
                name = ynames[name]
                which = yidx[name]
            got = g.ctx_array(which).cpu().numpy()[tuple(sl[1:])]
            ref = o.array(name)[tuple(sl[1:])]
            assert rel_err(got, ref) <= 1e-12, (d, name)
        o.calc_rhs(d, dt)
        g.calc_rhs(d, dt, flux)
    torch.cuda.synchronize()
    assert rel_err(o.interior(g.Uc.cpu().numpy()), o.interior(o.Uc)) <= 1e-12
    assert rel_err(o.interior(g.InvDt.cpu().numpy()), o.interior(o.array("InvDt"))) <= 1e-12
    o.calc_corner_emf()
    g.calc_corner_emf()
    torch.cuda.synchronize()
    ez = g.ctx_array(2).cpu().numpy()[0, b[1]:e[1] + 1, b[0]:e[0] + 1]
    assert rel_err(ez, o.array("ez")[0, b[1]:e[1] + 1, b[0]:e[0] + 1]) <= 1e-12
    o.evolve_mag_field(dt)
    g.evolve_mag_field(dt)
    torch.cuda.synchronize()
    compare_state(o, g, what="unfused stage pieces")


def test_compute_timestep_and_divb():
    kw = cfg_kwargs(2, (64, 48), emf="uct_hlld")
    o, g = make_pair(kw, problems.orszag_tang)
    o.set_boundaries()
    for _ in range(3):
        o.cycle()
        g.cycle()
    dg = g.check_divb()
    do = o.check_divb()
    assert dg == pytest.approx(do, rel=1e-6, abs=1e-15)
    assert dg < 1e-12
    # ComputeTimestep through the reference-mode entry points
    g.set_boundaries()
    g.prim_to_cons()
    g.reset_stage()
    g.evolve_stage(0.0, g.dt)
    o.set_boundaries()
    o.prim_to_cons()
    o.reset_stage()
    o.evolve_stage(0.0, o.dt)
    assert g.compute_timestep() == o.compute_timestep()


def test_fp32_stencil_close_to_fp64():
    """FP32 stencil/flux/EMF arithmetic with FP64 accumulation and ConsToPrim stays
    within single-precision error of the FP64 oracle over a few cycles."""
    kw = cfg_kwargs(2, (64, 64), emf="uct_hlld")
    o, g = make_pair(kw, problems.orszag_tang, precision="fp32")
    o.set_boundaries()
    for _ in range(5):
        o.cycle()
        g.cycle()
    Vc_g, _ = g.download()
    for n in range(o.nvar):
        a = o.interior(Vc_g)[n]
        b = o.interior(o.Vc)[n]
        assert np.max(np.abs(a - b)) <= 1e-4 * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("emf", ["uct_hlld", "uct_contact", "uct0"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_orszag_tang_per_kernel_schedule(emf, recon):
    """idfx_cycle with one kernel per reference step (fusion level 0)."""
    kw = cfg_kwargs(2, (48, 40), emf=emf, recon=recon, cfl=0.6)
    o, g = make_pair(kw, problems.orszag_tang)
    g.set_fusion(0)
    o.set_boundaries()
    _cycle_both(o, g, 6)
    compare_state(o, g, what="OT per-kernel %s %s" % (emf, recon))


@pytest.mark.parametrize("nstages", [1, 3])
@pytest.mark.parametrize("emf", ["uct_hlld", "arithmetic"])
def test_mhd2d_stage_counts(nstages, emf):
    """Odd stage counts: RK3 runs its first stage in place, then two fused stages; the
    single-stage integrator runs per kernel."""
    kw = cfg_kwargs(2, (40, 36), emf=emf, recon="ppm", cfl=0.4, nstages=nstages)
    o, g = make_pair(kw, lambda gg: problems.random_mhd(gg, seed=13, amp=0.2))
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="2D MHD %d stages %s" % (nstages, emf))


@pytest.mark.parametrize("emf,recon,bc", [
    ("uct_hlld", "plm", "periodic"),
    ("uct_hll", "ppm", ["outflow", "periodic"]),
    ("uct_contact", "ppm", ["periodic", "reflective"]),
])
def test_fused_schedule_bitwise_equals_per_kernel(emf, recon, bc):
    """At a size the oracle does not reach in seconds: the fused 2-D schedule and the
    per-kernel schedule give bit-identical interiors and the same dt sequence.  The x
    extent leaves a partial last column block for every halo width."""
    import torch
    from idefix_amd.datablock import DataBlock
    nx = (301, 263)
    blocks = []
    for level in (1, 0):
        g = DataBlock(2, nx, emf=emf, recon=recon, bc=bc, cfl=0.5)
        g.set_fusion(level)
        Vc, Vs = problems.orszag_tang(g)
        g.upload(Vc, Vs)
        blocks.append(g)
    for _ in range(5):
        for g in blocks:
            g.cycle()
        assert blocks[0].dt == blocks[1].dt
    (a_c, a_s), (b_c, b_s) = blocks[0].download(), blocks[1].download()
    b, e = blocks[0].beg, blocks[0].end
    assert np.array_equal(a_c[:, :, b[1]:e[1], b[0]:e[0]], b_c[:, :, b[1]:e[1], b[0]:e[0]])
    assert np.array_equal(a_s[0][:, b[1]:e[1], b[0]:e[0] + 1], b_s[0][:, b[1]:e[1], b[0]:e[0] + 1])
    assert np.array_equal(a_s[1][:, b[1]:e[1] + 1, b[0]:e[0]], b_s[1][:, b[1]:e[1] + 1, b[0]:e[0]])
    torch.cuda.synchronize()


def test_region_profiler_counts():
    """Region profiler: one event pair per region per stage, positive durations."""
    kw = cfg_kwargs(2, (64, 64), emf="uct_hlld")
    _, g = make_pair(kw, problems.orszag_tang)
    g.set_profiling(True)
    for _ in range(3):
        g.cycle()
    rt = g.region_times()
    g.set_profiling(False)
    assert set(rt) == {"boundaries", "sweep_x", "fused2d_y"}
    for name, (ms, n) in rt.items():
        assert n == 6 and ms > 0.0, (name, ms, n)
    g.set_fusion(0)
    g.set_profiling(True)
    g.cycle()
    rt = g.region_times()
    assert set(rt) == {"boundaries", "sweep_x", "sweep_y", "corner_emf", "evolve_b",
                       "cons_to_prim"}
    assert all(n == 2 for _, n in rt.values())


def test_orszag_tang_3d_reference_benchmark_config():
    """The reference's published benchmark setup (3-D Orszag-Tang, PLM, HLLD, uct_contact,
    RK2, CFL 0.9) at a small size."""
    kw = cfg_kwargs(3, (16, 24, 12), emf="uct_contact", recon="plm", cfl=0.9, nstages=2)
    o, g = make_pair(kw, problems.orszag_tang_3d)
    o.set_boundaries()
    _cycle_both(o, g, 4)
    compare_state(o, g, what="OT3D")


@pytest.mark.parametrize("emf", ["uct_hlld", "uct_contact"])
@pytest.mark.parametrize("recon", ["plm", "ppm"])
def test_turbulence_2d_three_components(emf, recon):
    """2-D MHD with three velocity/field components (BX3 cell-centred), seeded turbulence,
    fused 2-D schedule."""
    kw = cfg_kwargs(2, (40, 36), emf=emf, recon=recon, cfl=0.5, comps=3)
    o, g = make_pair(kw, problems.turbulence)
    o.set_boundaries()
    _cycle_both(o, g, 5)
    compare_state(o, g, what="turb2d c3 %s %s" % (emf, recon))


def test_turbulence_3d_ppm_rk3():
    """3-D decaying turbulence, PPM, HLLD, RK3 (the 3-D turbulence configuration)."""
    kw = cfg_kwargs(3, (16, 12, 14), emf="uct_contact", recon="ppm", cfl=0.5, nstages=3)
    o, g = make_pair(kw, problems.turbulence)
    o.set_boundaries()
    _cycle_both(o, g, 3)
    compare_state(o, g, what="turb3d ppm rk3")


def _cs_field(o, amp=0.3):
    """A smooth positive per-cell sound speed over every cell of the array."""
    nx, ny, nz = o.np_tot
    x = o.coord(0, "x")[None, None, :]
    y = o.coord(1, "x")[None, :, None] if o.dim >= 2 else 0.0
    return (1.0 + amp * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)) * np.ones((nz, ny, nx))


@pytest.mark.parametrize("dim,nx,mhd,solver,recon,emf,flat", [
    (1, (200,), False, "hllc", "plm", "uct_contact", None),
    (1, (200,), False, "roe", "ppm", "uct_contact", None),
    (1, (200,), True, "hll", "plm", "uct_contact", None),
    (1, (200,), True, "roe", "limo3", "uct_contact", None),
    (2, (40, 24), False, "hllc", "plm", "uct_contact", 0.05),
    (2, (48, 40), True, "hlld", "plm", "uct_hlld", None),
    (2, (48, 40), True, "roe", "plm", "uct_contact", None),
])
def test_iso_cell_sound_speed(dim, nx, mhd, solver, recon, emf, flat):
    """"csiso userdef": a spatially varying sound speed held per cell, the same array in
    the oracle and on the GPU (face value = average of the two cells; shock flags with
    each cell's own)."""
    kw = cfg_kwargs(dim, nx, mhd=mhd, solver=solver, recon=recon, emf=emf, cfl=0.5,
                    comps=2 if mhd else dim, csiso=1.0, shock_flattening=flat,
                    bc="periodic" if dim == 2 else "outflow")
    ic = (lambda gg: problems.random_mhd(gg, seed=41, amp=0.2)) if mhd else \
        (lambda gg: problems.random_hd(gg, seed=42, amp=0.3))
    o, g = make_pair(kw, ic)
    cs = _cs_field(o)
    o.set_iso_sound_speed(cs)
    g.set_iso_sound_speed(cs)
    o.set_boundaries()
    _cycle_both(o, g, 5)
    compare_state(o, g, what="iso cs(x) %dD %s %s" % (dim, solver, recon))
    # back to the constant sound speed: both follow
    o.set_iso_sound_speed(None)
    g.set_iso_sound_speed(None)
    _cycle_both(o, g, 2)
    compare_state(o, g, what="iso cs constant again %dD %s" % (dim, solver))


@pytest.mark.parametrize("dim,nx,recon,emf,nstages,fusion", [
    (3, (16, 12, 10), "ppm", "uct_contact", 3, None),  # ping-pong: k_prs_from_faces
    (3, (16, 12, 10), "plm", "uct_contact", 3, 0),     # k_cons_to_prim
    (2, (24, 20), "plm", "uct_contact", 3, None),      # fused y kernel
    (2, (24, 20), "ppm", "uct_hlld", 2, None),
    (1, (64,), "plm", "uct_contact", 2, None),
])
def test_pressure_floor(dim, nx, recon, emf, nstages, fusion):
    """ConsToPrim's pressure floor (convertConsToPrim.hpp:40-51: p <= 0 -> SMALL_PRESSURE_FIX
    = 1e-5, the re-synced energy local to the kernel) on every path that converts: a
    magnetised random state with its pressure scaled by 1e-4 drives (E - kin - mag) below
    zero in some cells of every cycle.  The oracle must floor cells, the GPU must floor the
    same cells, and the whole state must match."""
    comps = 2 if dim == 1 else 3
    kw = cfg_kwargs(dim, nx, emf=emf, recon=recon, cfl=0.4, nstages=nstages, comps=comps,
                    solver="hlld")

    def ic(gg):
        Vc, Vs = problems.random_mhd(gg, seed=5, amp=0.3)
        Vc[comps + 1] *= 1e-4
        return Vc, Vs

    o, g = make_pair(kw, ic)
    if fusion is not None:
        g.set_fusion(fusion)
    o.set_boundaries()
    _cycle_both(o, g, 3)
    p_o = o.interior(o.Vc)[comps + 1]
    p_g = o.interior(g.download()[0])[comps + 1]
    assert np.sum(p_o == 1e-5) > 0, "the state does not reach the floor"
    assert np.array_equal(p_o == 1e-5, p_g == 1e-5)
    compare_state(o, g, what="pressure floor %dD %s %s" % (dim, recon, emf))
