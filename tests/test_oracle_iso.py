"""Pins of the oracle's isothermal equation of state (ISOTHERMAL builds, constant csiso:
eos_isothermal.hpp and the HAVE_ENERGY == 0 branches of the solvers): the exact
isothermal Riemann solution checked by its jump conditions, flux consistency with the
textbook isothermal fluxes, the Roe property of the Roe MHD solver, the isothermal Sod tube
against its exact solution with every solver, and the isothermal Brio-Wu tube: constant
normal field, conservation, and every solver against a 4x finer solution."""
import numpy as np
import pytest

from oracle import oracle as O
from idefix_amd import problems
from tests.exact_riemann import sample_iso, star_state_iso

rng = np.random.default_rng(7)
CS = 1.0


def test_exact_isothermal_solver_jump_conditions():
    for _ in range(200):
        rL, rR = rng.uniform(0.1, 3.0, 2)
        uL, uR = rng.normal(size=2) * 0.5
        rs, us = star_state_iso(rL, uL, rR, uR, CS)
        for rk, uk, side in ((rL, uL, -1), (rR, uR, 1)):
            if rs > rk:  # shock: mass and momentum fluxes continuous in its frame
                S = uk + side * CS * np.sqrt(rs / rk)
                assert S * (rs - rk) == pytest.approx(rs * us - rk * uk, rel=1e-9, abs=1e-12)
                lhs = S * (rs * us - rk * uk)
                rhs = (rs * us * us + CS ** 2 * rs) - (rk * uk * uk + CS ** 2 * rk)
                assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-12)
            else:  # rarefaction: Riemann invariant u -+ cs ln rho conserved
                assert us - side * CS * np.log(rs) == pytest.approx(
                    uk - side * CS * np.log(rk), rel=1e-12, abs=1e-12)


def _iso_flux_hd(v, d):
    rho, vel = v[0], v[1:4]
    F = np.zeros(4)
    F[0] = rho * vel[d]
    F[1:4] = rho * vel * vel[d]
    F[1 + d] += CS ** 2 * rho
    return F


def _iso_flux_mhd(v, d):
    rho, vel, B = v[0], v[1:4], v[4:7]
    F = np.zeros(7)
    F[0] = rho * vel[d]
    F[1:4] = rho * vel * vel[d] - B * B[d]
    F[1 + d] += CS ** 2 * rho + 0.5 * B @ B
    F[4:7] = vel[d] * B - B[d] * vel
    return F


@pytest.mark.parametrize("solver", ["hllc", "hll", "tvdlf", "roe"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_iso_hd_flux_consistency(solver, d):
    cfg = O.make_cfg(3, mhd=False, solver=solver, nx=(8, 8, 8), comps=3, csiso=CS)
    for _ in range(100):
        v = np.zeros(8)
        v[0] = rng.uniform(0.2, 3)
        v[1:4] = rng.normal(size=3)
        F, cmax, _, wrote = O.face_flux(cfg, d, v, v)
        assert wrote
        assert np.allclose(F[:4], _iso_flux_hd(v, d), rtol=1e-12, atol=1e-12)
        if solver != "tvdlf":
            assert cmax == pytest.approx(abs(v[1 + d]) + CS, rel=1e-12)


@pytest.mark.parametrize("solver,emf", [("hlld", "uct_contact"), ("hlld", "uct_hlld"),
                                        ("hll", "uct_hll"), ("tvdlf", "uct_contact"),
                                        ("roe", "uct_contact")])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_iso_mhd_flux_consistency(solver, emf, d):
    cfg = O.make_cfg(3, mhd=True, solver=solver, emf=emf, nx=(8, 8, 8), comps=3, csiso=CS)
    for _ in range(100):
        v = np.zeros(8)
        v[0] = rng.uniform(0.2, 3)
        v[1:4] = rng.normal(size=3)
        v[4:7] = rng.normal(size=3)
        F, _, _, _ = O.face_flux(cfg, d, v, v)
        assert np.allclose(F[:7], _iso_flux_mhd(v, d), rtol=1e-11, atol=1e-11)


def _flux_mhd_energy(v, d, gamma=5.0 / 3.0):
    rho, vel, p, B = v[0], v[1:4], v[4], v[5:8]
    pt = p + 0.5 * B @ B
    E = p / (gamma - 1.0) + 0.5 * rho * vel @ vel + 0.5 * B @ B
    F = np.zeros(8)
    F[0] = rho * vel[d]
    F[1:4] = rho * vel * vel[d] - B * B[d]
    F[1 + d] += pt
    F[4] = (E + pt) * vel[d] - B[d] * (vel @ B)
    F[5:8] = vel[d] * B - B[d] * vel
    return F


@pytest.mark.parametrize("iso", [True, False])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_roe_mhd_is_a_roe_linearisation(iso, d):
    """Roe's property F(U_R) - F(U_L) = A(U_R - U_L), summed over the waves, for finite
    jumps: when every wave speed has one sign the dissipation sum_k |lambda_k| eta_k R_k
    equals +-(F_R - F_L), so the Roe flux is exactly the upwind physical flux.  This pins
    the eigenvectors, the wave strengths and the X term of the sound speed (the exact
    linearisation of roeMHD.hpp; without X the deviation is ~5e-3)."""
    cfg = O.make_cfg(3, mhd=True, solver="roe", emf="uct_contact", nx=(8, 8, 8), comps=3,
                     csiso=CS if iso else None)
    nv, b0 = (7, 4) if iso else (8, 5)
    flux = (lambda v: _iso_flux_mhd(v, d)) if iso else (lambda v: _flux_mhd_energy(v, d))
    for _ in range(100):
        vL, vR = np.zeros(8), np.zeros(8)
        for v in (vL, vR):
            v[0] = rng.uniform(0.5, 2.0)
            v[1:4] = rng.normal(size=3) * 0.3
            if iso:
                v[4:7] = rng.normal(size=3) * 0.5
            else:
                v[4] = rng.uniform(0.5, 2.0)
                v[5:8] = rng.normal(size=3) * 0.5
        vR[b0 + d] = vL[b0 + d]  # the normal field is continuous (it comes from Vs)
        sgn = rng.choice([-1.0, 1.0])
        vL[1 + d] += 8.0 * sgn  # faster than every wave of both states
        vR[1 + d] += 8.0 * sgn
        F, _, _, _ = O.face_flux(cfg, d, vL, vR)
        ref = flux(vL if sgn > 0 else vR)
        assert np.max(np.abs(F[:nv] - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("solver,recon", [("roe", "plm"), ("hll", "plm"), ("hllc", "plm"),
                                          ("tvdlf", "plm"), ("hllc", "limo3")])
def test_iso_sod_against_exact_solution(solver, recon):
    """test/HD/sod-iso (500 cells, CFL 0.8, t = 0.2, csiso 1)."""
    cfg = O.make_cfg(1, mhd=False, recon=recon, solver=solver, nx=(500,), bc="outflow",
                     cfl=0.8, csiso=CS)
    o = O.Oracle(cfg)
    assert o.nvar == 2
    Vc, _ = problems.sod_iso(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.set_boundaries()
    assert o.run(0.2) == 0
    x = o.coord(0, "x")[o.beg[0]:o.end[0]]
    v = o.interior(o.Vc)[:, 0, 0]
    ex = sample_iso(x, 0.2, 0.5, (1.0, 0.0), (0.125, 0.0), CS)
    assert np.mean(np.abs(v[0] - ex[0])) < 3e-3
    assert np.mean(np.abs(v[1] - ex[1])) < 6e-3


def _bw_iso(solver, n=800, recon="plm", nstages=2, bc="outflow"):
    cfg = O.make_cfg(1, comps=2, mhd=True, recon=recon, solver=solver, emf="uct_contact",
                     nx=(n,), xend=(100, 1, 1), bc=bc, cfl=0.8, nstages=nstages, csiso=CS)
    o = O.Oracle(cfg)
    Vc, Vs = problems.brio_wu_iso(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    return o


def test_iso_brio_wu_solvers_agree_with_fine_solution():
    fine = _bw_iso("hlld", n=3200)
    assert fine.run(10.0) == 0
    ref = fine.interior(fine.Vc)[:, 0, 0].reshape(5, 800, 4).mean(axis=2)
    for solver, recon in [("hlld", "plm"), ("hll", "plm"), ("tvdlf", "plm"), ("hlld", "limo3"),
                          ("roe", "plm"), ("roe", "limo3")]:
        o = _bw_iso(solver, recon=recon)
        assert o.run(10.0) == 0
        v = o.interior(o.Vc)[:, 0, 0]
        err = np.mean(np.abs(v - ref), axis=1)
        assert err[0] < 4e-3 and err[4] < 6e-3, (solver, recon, err)
        assert np.all(v[3] == 0.75)


def test_iso_brio_wu_periodic_conservation():
    o = _bw_iso("hlld", n=400, bc="periodic")
    o.prim_to_cons()
    t0 = o.interior(o.Uc).sum(axis=(1, 2, 3)).copy()
    for _ in range(60):
        assert o.cycle() == 0
    o.prim_to_cons()
    t1 = o.interior(o.Uc).sum(axis=(1, 2, 3))
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.maximum(np.abs(t0), 400.0))


# ---- per-cell sound speed ("csiso userdef", eos_isothermal.hpp:29-33, :70-75) ---------

def _cs_field(o, amp=0.3):
    """A smooth positive sound speed over every cell of the array."""
    nx, ny, nz = o.np_tot
    x = o.coord(0, "x")[None, None, :]
    y = o.coord(1, "x")[None, :, None] if o.dim >= 2 else 0.0
    return (1.0 + amp * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)) * np.ones((nz, ny, nx))


@pytest.mark.parametrize("mhd,solver", [(False, "hllc"), (False, "roe"), (False, "tvdlf"),
                                        (True, "hlld"), (True, "roe"), (True, "hll")])
def test_uniform_cell_sound_speed_is_the_constant_one(mhd, solver):
    """An array holding csiso everywhere reproduces the constant path bit for bit."""
    kw = dict(mhd=mhd, solver=solver, nx=(24, 20), comps=2, cfl=0.5, csiso=0.9,
              emf="uct_contact")
    runs = []
    for use_array in (False, True):
        o = O.Oracle(O.make_cfg(2, **kw))
        Vc, Vs = (problems.random_mhd if mhd else problems.random_hd)(O.OracleGrid(o), seed=4)
        o.Vc[...] = Vc
        if mhd:
            o.Vs[...] = Vs
        if use_array:
            nx, ny, nz = o.np_tot
            o.set_iso_sound_speed(np.full((nz, ny, nx), 0.9))
        o.set_boundaries()
        for _ in range(3):
            assert o.cycle() == 0
        runs.append(o.Vc.copy())
    assert np.array_equal(runs[0], runs[1])


@pytest.mark.parametrize("mhd,solver", [(False, "hllc"), (False, "hll"), (False, "tvdlf"),
                                        (False, "roe"), (True, "hlld"), (True, "hll"),
                                        (True, "tvdlf"), (True, "roe")])
def test_face_sound_speed_is_the_average_of_the_two_cells(mhd, solver):
    """Gas at rest with uniform density (and field): every solver returns the physical
    flux, whose normal momentum is rho c^2 (+ B^2/2) with c the face average
    (cs_left + cs_right)/2 -- not either cell's speed, nor the mean of the squares."""
    o = O.Oracle(O.make_cfg(2, mhd=mhd, solver=solver, nx=(16, 12), comps=2, csiso=1.0,
                            emf="uct_contact"))
    cs = _cs_field(o)
    o.set_iso_sound_speed(cs)
    Vc = np.zeros_like(o.Vc)
    Vc[0] = 1.3
    b2 = 0.0
    if mhd:
        Vc[3], Vc[4] = 0.2, -0.1  # BX1, BX2
        o.Vs[0] = 0.2
        o.Vs[1] = -0.1
        b2 = 0.2 ** 2 + 0.1 ** 2
    o.Vc[...] = Vc
    o.set_boundaries()
    for d in (0, 1):
        o.calc_flux(d)
        F = o.array("Flux")
        b, e = o.beg, o.end
        sl = (slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0]))
        shift = [0, 0, 0]
        shift[2 - d] = 1
        left = cs[tuple(slice(s.start - sh, s.stop - sh) for s, sh in zip(sl, shift))]
        cface = 0.5 * (cs[sl] + left)
        bn2 = (0.2 if d == 0 else -0.1) ** 2
        expect = 1.3 * cface * cface + 0.5 * b2 - (bn2 if mhd else 0.0)
        assert np.allclose(F[1 + d][sl], expect, rtol=1e-14, atol=1e-14), (solver, d)


def test_shock_flags_use_each_cells_sound_speed():
    """A converging flow of uniform density and a sound speed that jumps by 1.5x inside
    the converging region: only the two cells beside the jump, whose neighbours' pressures
    rho cs^2 differ by 125 %, exceed the smoothing threshold (shockFlattening.hpp:96-128);
    with a uniform sound speed no cell is flagged."""
    cfg = O.make_cfg(1, mhd=False, solver="hllc", nx=(64,), bc="periodic", cfl=0.5,
                     csiso=1.0, shock_flattening=0.1)
    o = O.Oracle(cfg)
    x = o.coord(0, "x")
    Vc = np.zeros_like(o.Vc)
    Vc[0] = 1.0
    Vc[1] = -0.5 * np.sin(2 * np.pi * x)[None, None, :]  # div v < 0 for x < 1/4
    o.Vc[...] = Vc
    o.set_boundaries()
    assert not o.find_shock().any()
    o.set_iso_sound_speed(np.where(x > 0.125, 1.5, 1.0)[None, None, :])
    flagged = np.nonzero(o.find_shock()[0, 0])[0]
    jump = np.nonzero(np.diff(x > 0.125))[0][0]  # cells jump, jump + 1 straddle it
    assert list(flagged) == [jump, jump + 1]
