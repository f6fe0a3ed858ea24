"""Pins of the oracle's pointwise algebra by properties the mathematics fixes:
limiter identities, PPM exactness on polynomials, conversion round trips, flux
consistency with the textbook fluxes, upwinding, mirror symmetry and exact resolution of
stationary contact discontinuities."""
import numpy as np
import pytest

from oracle import oracle as O

rng = np.random.default_rng(1234)


# ---------------- limiters ----------------
def test_vanleer_properties():
    for _ in range(2000):
        a, b = rng.normal(size=2) * 10 ** rng.uniform(-3, 3)
        v = O.vanleer(a, b)
        if a * b <= 0:
            assert v == 0.0
            continue
        # harmonic mean of the two slopes: 1/v = (1/a + 1/b)/2
        assert 1.0 / v == pytest.approx(0.5 * (1.0 / a + 1.0 / b), rel=1e-13)
        assert O.vanleer(b, a) == v
        assert min(abs(a), abs(b)) * (1 - 1e-15) <= abs(v) <= 2 * min(abs(a), abs(b)) * (1 + 1e-15)
    assert O.vanleer(0.3, 0.3) == 0.3  # linear data is reconstructed exactly


def test_mc_limiter_is_minmod_of_three():
    def minmod3(x, y, z):
        if x > 0 and y > 0 and z > 0:
            return min(x, y, z)
        if x < 0 and y < 0 and z < 0:
            return max(x, y, z)
        return 0.0
    for _ in range(2000):
        a, b = rng.normal(size=2)
        assert O.mclim(a, b) == pytest.approx(minmod3(2 * a, 2 * b, 0.5 * (a + b)), rel=1e-15, abs=0)


def _limo3_face(vm, v0, vp, dx):
    """Left state of the face between cells 0 and +1 from the LimO3 slope of cell 0."""
    dvm, dvp = v0 - vm, vp - v0
    return v0 + 0.5 * dvp * O.limo3lim(dvp, dvm, dx)


def test_limo3_third_order_on_smooth_quadratics():
    """Where the differences are small against r dx (r = 0.1) LimO3 is the unlimited
    third-order interface value (-V[-1] + 5 V[0] + 2 V[+1]) / 6, exact for the cell
    averages of any quadratic (Cada & Torrilhon 2009, eq. 3.6)."""
    h = 0.01
    used = 0
    for _ in range(500):
        c = rng.normal(size=3) * np.array([1.0, 0.01, 0.01])
        x0 = rng.uniform(0.1, 0.9)
        avg = [_cell_avg_poly(c, x0 + (s - 0.5) * h, x0 + (s + 0.5) * h) for s in (-1, 0, 1)]
        dvm, dvp = avg[1] - avg[0], avg[2] - avg[1]
        if (dvm ** 2 + dvp ** 2) / (0.1 * h) ** 2 > 1.0 - 1e-9:
            continue  # outside the smooth regime
        used += 1
        exact = c[0] + c[1] * (x0 + h / 2) + c[2] * (x0 + h / 2) ** 2
        assert _limo3_face(*avg, h) == pytest.approx(exact, rel=0, abs=1e-13)
    assert used > 400


def test_limo3_limited_regime_properties():
    """Above the smoothness threshold the limiter lies in [0, 1.5], never exceeds twice the
    slope ratio on monotone data (no new extremum at the face), and reconstructs linear
    data exactly in both regimes."""
    r = np.random.default_rng(17)  # own stream: the result must not depend on test order
    for _ in range(3000):
        dvp, dvm = r.normal(size=2)
        dx = 10 ** r.uniform(-3, -1)  # differences of order 1: the limited regime
        if (dvm ** 2 + dvp ** 2) / (0.1 * dx) ** 2 <= 1.0:
            continue  # a rare draw inside the smooth regime (unlimited third order)
        lim = O.limo3lim(dvp, dvm, dx)
        th = dvm / dvp
        assert 0.0 <= lim <= 1.5
        if th >= 0:
            assert lim <= 2 * th * (1 + 1e-12) + 1e-300
        else:
            assert lim <= -0.5 * th * (1 + 1e-12)
    for dx in (1e-3, 1.0, 1e3):
        assert O.limo3lim(0.7, 0.7, dx) == pytest.approx(1.0, rel=1e-15)


def _cell_avg_poly(c, xl, xr):
    # average of sum c_k x^k over [xl, xr]
    return sum(ck * (xr ** (k + 1) - xl ** (k + 1)) / (k + 1) for k, ck in enumerate(c)) / (xr - xl)


def test_ppm_exact_on_monotone_cubic():
    """4th-order face interpolation is exact for cubics; on monotone data away from
    extrema no limiter fires, so PPM returns the exact face values."""
    for _ in range(200):
        c = [rng.uniform(-1, 1), rng.uniform(1, 2), rng.uniform(-0.05, 0.05), rng.uniform(-0.01, 0.01)]
        h = 0.01
        x0 = rng.uniform(-0.2, 0.2)
        avg = [_cell_avg_poly(c, x0 + (k - 0.5) * h, x0 + (k + 0.5) * h) for k in range(-2, 3)]
        vl, vr = O.ppm_states(*avg)
        exact_l = sum(ck * (x0 - 0.5 * h) ** k for k, ck in enumerate(c))
        exact_r = sum(ck * (x0 + 0.5 * h) ** k for k, ck in enumerate(c))
        assert vl == pytest.approx(exact_l, abs=1e-12)
        assert vr == pytest.approx(exact_r, abs=1e-12)


def test_ppm_no_new_extrema_at_step():
    for _ in range(500):
        lo, hi = sorted(rng.uniform(0.1, 2.0, 2))
        k = rng.integers(0, 4)
        v = [lo] * (k + 1) + [hi] * (4 - k)
        vl, vr = O.ppm_states(*v)
        assert lo - 1e-14 <= vl <= hi + 1e-14
        assert lo - 1e-14 <= vr <= hi + 1e-14


# ---------------- conversions ----------------
def _cfg(dim, mhd, solver, comps=None, emf="uct_contact", gamma=5.0 / 3.0):
    return O.make_cfg(dim, comps=comps, mhd=mhd, solver=solver, emf=emf, nx=(8, 8, 8), gamma=gamma)


@pytest.mark.parametrize("mhd,comps", [(False, 1), (False, 3), (True, 2), (True, 3)])
def test_prim_cons_round_trip(mhd, comps):
    cfg = _cfg(comps, mhd, "hlld" if mhd else "hllc", comps=comps)
    nv = 2 * comps + 2 if mhd else comps + 2
    for _ in range(500):
        v = np.zeros(8)
        v[0] = rng.uniform(0.1, 10)
        v[1:1 + comps] = rng.normal(size=comps)
        v[comps + 1] = rng.uniform(0.1, 10)
        if mhd:
            v[comps + 2:comps + 2 + comps] = rng.normal(size=comps)
        u = O.prim2cons(cfg, v)
        w = O.cons2prim(cfg, u)
        assert np.allclose(w[:nv], v[:nv], rtol=1e-13, atol=1e-14)
        # E = p/(g-1) + rho v^2/2 + B^2/2
        e = v[comps + 1] / (cfg.gamma - 1) + 0.5 * v[0] * np.sum(v[1:1 + comps] ** 2)
        if mhd:
            e += 0.5 * np.sum(v[comps + 2:comps + 2 + comps] ** 2)
        assert u[comps + 1] == pytest.approx(e, rel=1e-14)


def test_cons2prim_pressure_floor():
    cfg = _cfg(3, True, "hlld", comps=3)
    u = np.array([1.0, 1.0, 0.0, 0.0, 0.4, 0.1, 0.0, 0.0])  # E below kinetic energy
    v = O.cons2prim(cfg, u)
    assert v[4] == 1e-5


# ---------------- fluxes ----------------
def euler_flux(v, d, gamma, comps):
    rho, vel, p = v[0], v[1:1 + comps], v[comps + 1]
    E = p / (gamma - 1) + 0.5 * rho * np.sum(vel ** 2)
    F = np.zeros(comps + 2)
    F[0] = rho * vel[d]
    F[1:1 + comps] = rho * vel * vel[d]
    F[1 + d] += p
    F[comps + 1] = (E + p) * vel[d]
    return F


def mhd_flux(v, d, gamma, comps=3):
    rho, vel, p, B = v[0], v[1:4], v[4], v[5:8]
    E = p / (gamma - 1) + 0.5 * rho * vel @ vel + 0.5 * B @ B
    pt = p + 0.5 * B @ B
    F = np.zeros(8)
    F[0] = rho * vel[d]
    F[1:4] = rho * vel * vel[d] - B * B[d]
    F[1 + d] += pt
    F[4] = (E + pt) * vel[d] - B[d] * (vel @ B)
    F[5:8] = vel[d] * B - B[d] * vel
    return F


def _rand_hd(comps):
    v = np.zeros(8)
    v[0] = rng.uniform(0.2, 3)
    v[1:1 + comps] = rng.normal(size=comps)
    v[comps + 1] = rng.uniform(0.2, 3)
    return v


def _rand_mhd():
    v = np.zeros(8)
    v[0] = rng.uniform(0.2, 3)
    v[1:4] = rng.normal(size=3)
    v[4] = rng.uniform(0.2, 3)
    v[5:8] = rng.normal(size=3)
    return v


@pytest.mark.parametrize("solver", ["hllc", "roe", "hll", "tvdlf"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hd_flux_consistency(solver, d):
    cfg = _cfg(3, False, solver, comps=3, gamma=1.4)
    for _ in range(200):
        v = _rand_hd(3)
        F, cmax, _, wrote = O.face_flux(cfg, d, v, v)
        assert wrote
        assert np.allclose(F[:5], euler_flux(v, d, 1.4, 3), rtol=1e-12, atol=1e-12)
        a = np.sqrt(1.4 * v[4] / v[0])
        assert cmax == pytest.approx(abs(v[1 + d]) + a, rel=1e-12)


@pytest.mark.parametrize("solver,emf", [("hlld", "uct_contact"), ("hlld", "uct_hlld"),
                                        ("roe", "uct_contact"), ("roe", "uct_hlld"),
                                        ("hll", "uct_contact"), ("hll", "uct_hll"),
                                        ("tvdlf", "uct_contact"), ("tvdlf", "uct_hll")])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_mhd_flux_consistency(solver, emf, d):
    cfg = _cfg(3, True, solver, comps=3, emf=emf)
    for _ in range(200):
        v = _rand_mhd()
        F, cmax, _, _ = O.face_flux(cfg, d, v, v)
        assert np.allclose(F, mhd_flux(v, d, 5.0 / 3.0), rtol=1e-11, atol=1e-11)


def _mirror(v, d, mhd):
    w = v.copy()
    w[1 + d] = -w[1 + d]
    if mhd:
        # a reflection x_d -> -x_d flips the polar vector v_d and the axial B components
        # other than B_d; the normal field is kept (it is shared by both sides)
        for c in range(3):
            if c != d:
                w[5 + c] = -w[5 + c]
    return w


@pytest.mark.parametrize("solver", ["hllc", "hll", "tvdlf"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hllc_mirror_symmetry(solver, d):
    """Swapping and reflecting the two states reflects the flux: mass and energy fluxes
    change sign, the normal momentum flux does not."""
    cfg = _cfg(3, False, solver, comps=3, gamma=1.4)
    for _ in range(300):
        vL, vR = _rand_hd(3), _rand_hd(3)
        F, _, _, _ = O.face_flux(cfg, d, vL, vR)
        G, _, _, _ = O.face_flux(cfg, d, _mirror(vR, d, False), _mirror(vL, d, False))
        assert G[0] == pytest.approx(-F[0], rel=1e-12, abs=1e-13)
        assert G[1 + d] == pytest.approx(F[1 + d], rel=1e-12, abs=1e-13)
        assert G[4] == pytest.approx(-F[4], rel=1e-12, abs=1e-13)


@pytest.mark.parametrize("solver", ["hlld", "roe", "hll", "tvdlf"])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_mhd_mirror_symmetry(solver, d):
    cfg = _cfg(3, True, solver, comps=3)
    for _ in range(300):
        vL, vR = _rand_mhd(), _rand_mhd()
        vR[5 + d] = vL[5 + d]  # one normal field on the face
        F, _, _, _ = O.face_flux(cfg, d, vL, vR)
        G, _, _, _ = O.face_flux(cfg, d, _mirror(vR, d, True), _mirror(vL, d, True))
        assert G[0] == pytest.approx(-F[0], rel=1e-11, abs=1e-12)
        assert G[1 + d] == pytest.approx(F[1 + d], rel=1e-11, abs=1e-12)
        assert G[4] == pytest.approx(-F[4], rel=1e-11, abs=1e-12)


@pytest.mark.parametrize("solver", ["hllc", "roe", "hll"])
def test_supersonic_upwinding(solver):
    cfg = _cfg(3, False, solver, comps=3, gamma=1.4)
    for _ in range(100):
        vL, vR = _rand_hd(3), _rand_hd(3)
        vL[1] = vR[1] = 10.0  # flow to the right faster than any wave
        F, _, _, _ = O.face_flux(cfg, 0, vL, vR)
        if solver != "roe":
            assert np.allclose(F[:5], euler_flux(vL, 0, 1.4, 3), rtol=1e-13)


def _fast_speed(v, d, gamma):
    """Textbook fast magnetosonic speed: c_f^2 = (a^2 + b^2 + sqrt((a^2 + b^2)^2 -
    4 a^2 b_n^2)) / 2 with a^2 = gamma p / rho, b^2 = |B|^2 / rho, b_n^2 = B_d^2 / rho."""
    a2 = gamma * v[4] / v[0]
    b2 = (v[5:8] @ v[5:8]) / v[0]
    bn2 = v[5 + d] ** 2 / v[0]
    return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bn2)))


def _cons(v, gamma, mhd):
    u = v.copy()
    u[1:4] = v[0] * v[1:4]
    e = v[4] / (gamma - 1) + 0.5 * v[0] * v[1:4] @ v[1:4]
    u[4] = e + (0.5 * v[5:8] @ v[5:8] if mhd else 0.0)
    return u


@pytest.mark.parametrize("mhd", [False, True])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_hll_satisfies_both_outer_jump_conditions(mhd, d):
    """The HLL flux is the one whose single intermediate state U* meets the
    Rankine-Hugoniot conditions across both outer waves (Harten, Lax & van Leer 1983):
    F = F_L + S_L (U* - U_L) = F_R + S_R (U* - U_R), with S_L, S_R the Davis estimates
    min/max(v_n -+ c) over the two states and U* their integral average.  The wave speeds
    come from the textbook formulas, not from the oracle's expressions."""
    gamma = 5.0 / 3.0 if mhd else 1.4
    cfg = _cfg(3, mhd, "hll", comps=3, gamma=gamma, **({"emf": "uct_contact"} if mhd else {}))
    nv = 8 if mhd else 5
    nsub = 0
    for _ in range(300):
        vL, vR = (_rand_mhd(), _rand_mhd()) if mhd else (_rand_hd(3), _rand_hd(3))
        if mhd:
            vR[5 + d] = vL[5 + d]
            c = (_fast_speed(vL, d, gamma), _fast_speed(vR, d, gamma))
            fL, fR = mhd_flux(vL, d, gamma), mhd_flux(vR, d, gamma)
        else:
            c = tuple(np.sqrt(gamma * v[4] / v[0]) for v in (vL, vR))
            fL, fR = euler_flux(vL, d, gamma, 3), euler_flux(vR, d, gamma, 3)
        SL = min(vL[1 + d] - c[0], vR[1 + d] - c[1])
        SR = max(vL[1 + d] + c[0], vR[1 + d] + c[1])
        F, cmax, _, _ = O.face_flux(cfg, d, vL, vR)
        F = F[:nv]
        assert cmax == pytest.approx(max(abs(SL), abs(SR)), rel=1e-12)
        uL, uR = _cons(vL, gamma, mhd)[:nv], _cons(vR, gamma, mhd)[:nv]
        if SL > 0:
            assert np.allclose(F, fL, rtol=1e-12, atol=1e-12)
        elif SR < 0:
            assert np.allclose(F, fR, rtol=1e-12, atol=1e-12)
        else:
            nsub += 1
            us = (SR * uR - SL * uL - (fR - fL)) / (SR - SL)
            scale = 1e-11 * (1 + np.abs(fL) + np.abs(fR) + abs(SL) * np.abs(uL) + abs(SR) * np.abs(uR))
            assert np.all(np.abs(F - (fL + SL * (us - uL))) < scale)
            assert np.all(np.abs(F - (fR + SR * (us - uR))) < scale)
    assert nsub > 100


@pytest.mark.parametrize("solver", ["hllc", "hlld"])
def test_stationary_contact_is_exact(solver):
    """HLLC and HLLD resolve an isolated stationary contact exactly: zero mass flux and
    the pressure (plus magnetic pressure) as the only momentum flux."""
    mhd = solver == "hlld"
    cfg = _cfg(3, mhd, solver, comps=3)
    for _ in range(100):
        p = rng.uniform(0.2, 2)
        vL, vR = np.zeros(8), np.zeros(8)
        vL[0], vR[0] = rng.uniform(0.2, 3, 2)
        vL[4] = vR[4] = p
        if mhd:
            B = rng.normal(size=3)
            B[0] = 0.0  # tangential field only: a pure contact / tangential discontinuity
            vL[5:8] = vR[5:8] = B
        F, _, _, _ = O.face_flux(cfg, 0, vL, vR)
        assert abs(F[0]) < 1e-14
        pt = p + (0.5 * vL[5:8] @ vL[5:8] if mhd else 0.0)
        assert F[1] == pytest.approx(pt, rel=1e-13)
        assert abs(F[4]) < 1e-13


def test_hlld_emf_store_signs():
    """Face EMF of the x sweep is -F(By) (= E_z = v_y B_x - v_x B_y), of the y sweep
    +F(Bx) (storeFlux.hpp:14-29 with the signs of hlldMHD.hpp:79-142)."""
    cfg = _cfg(3, True, "hlld", comps=3, emf="uct_contact")
    for _ in range(100):
        v = _rand_mhd()
        F0, _, E0, _ = O.face_flux(cfg, 0, v, v)
        F1, _, E1, _ = O.face_flux(cfg, 1, v, v)
        ez = v[2] * v[5] - v[1] * v[6]
        assert E0[0] == pytest.approx(ez, rel=1e-12, abs=1e-13)
        assert E1[0] == pytest.approx(ez, rel=1e-12, abs=1e-13)


def test_compute_timestep_is_brute_force_min():
    cfg = O.make_cfg(2, mhd=True, solver="hlld", nx=(16, 12))
    o = O.Oracle(cfg)
    inv = o.array("InvDt")
    inv[...] = rng.uniform(0.5, 5.0, inv.shape)
    assert o.compute_timestep() == np.min(1.0 / o.interior(inv))


def test_roe_mhd_2d_consistency():
    """COMPONENTS = 2: the 6-wave Roe flux of a uniform state is the physical flux."""
    cfg = _cfg(2, True, "roe", comps=2)
    for _ in range(200):
        v3 = _rand_mhd()
        v3[3] = v3[7] = 0.0  # no third components
        v = np.zeros(8)
        v[0:3], v[3], v[4:6] = v3[0:3], v3[4], v3[5:7]
        for d in range(2):
            F, _, _, _ = O.face_flux(cfg, d, v, v)
            ref = mhd_flux(v3, d, 5.0 / 3.0)
            assert np.allclose(F[:6], [ref[0], ref[1], ref[2], ref[4], ref[5], ref[6]],
                               rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("mhd", [False, True])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_tvdlf_is_rusanov_flux_of_mean_state(mhd, d):
    """The local Lax-Friedrichs (Rusanov) flux, Toro 3rd ed. sec. 10.5.1:
    F = (F_L + F_R)/2 - a (U_R - U_L)/2 with a = |v_n| + c of the arithmetic mean of the
    two primitive states.  c is the sound speed in hydro.  In MHD the reference's speed
    (tvdlfMHD.hpp:192-194: cRL = gpr - B2; cRL = cRL + B2 + sqrt(..)) lacks the B^2 term of
    the fast speed: c^2 = c_f^2 - |B|^2 / (2 rho), with c_f the textbook fast speed."""
    gamma = 5.0 / 3.0 if mhd else 1.4
    cfg = _cfg(3, mhd, "tvdlf", comps=3, gamma=gamma)
    nv = 8 if mhd else 5
    for _ in range(300):
        vL, vR = (_rand_mhd(), _rand_mhd()) if mhd else (_rand_hd(3), _rand_hd(3))
        if mhd:
            vR[5 + d] = vL[5 + d]
        vm = 0.5 * (vL + vR)
        if mhd:
            c = np.sqrt(_fast_speed(vm, d, gamma) ** 2 - 0.5 * (vm[5:8] @ vm[5:8]) / vm[0])
        else:
            c = np.sqrt(gamma * vm[4] / vm[0])
        a = abs(vm[1 + d]) + c
        if mhd:
            fL, fR = mhd_flux(vL, d, gamma), mhd_flux(vR, d, gamma)
        else:
            fL, fR = euler_flux(vL, d, gamma, 3), euler_flux(vR, d, gamma, 3)
        uL, uR = _cons(vL, gamma, mhd)[:nv], _cons(vR, gamma, mhd)[:nv]
        F, cmax, E, _ = O.face_flux(cfg, d, vL, vR)
        assert cmax == pytest.approx(a, rel=1e-12)
        want = 0.5 * (fL + fR) - 0.5 * a * (uR - uL)
        scale = 1e-11 * (1 + np.abs(fL) + np.abs(fR) + a * (np.abs(uL) + np.abs(uR)))
        assert np.all(np.abs(F[:nv] - want) < scale)
