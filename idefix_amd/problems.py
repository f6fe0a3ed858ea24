"""Initial conditions of the workloads (the Setup::InitFlow of each problem).

Each function fills host arrays Vc [nvar][nz][ny][nx] and Vs [dim][nz+1][ny+1][nx+1]
from the cell coordinates of a grid object exposing `x`, `xl` (lists of per-direction
arrays over np_tot), `np_tot`, `vs_shape`, `nvar`, `dim`, `comps`; both the GPU
DataBlock and the CPU oracle provide them.  Variable order: RHO, VX1..VXc, PRS, BX1..BXc.
"""
import numpy as np


def _coords(g, kind="x"):
    if hasattr(g, "coord"):
        return [g.coord(d, kind) for d in range(3)]
    return list(getattr(g, kind))


def _empty(g):
    nx, ny, nz = g.np_tot
    Vc = np.zeros((g.nvar, nz, ny, nx))
    Vs = None
    if getattr(g, "vs_shape", None) is not None and g.dim >= 1 and _is_mhd(g):
        vx, vy, vz = g.vs_shape
        Vs = np.zeros((g.dim, vz, vy, vx))
    return Vc, Vs


def _is_mhd(g):
    if hasattr(g, "mhd"):
        return bool(g.mhd)
    return bool(g.cfg.mhd)


def _is_iso(g):
    if hasattr(g, "isothermal"):
        return bool(g.isothermal)
    return bool(getattr(getattr(g, "cfg", None), "isothermal", 0))


def _nfluid(g):
    """Number of fluid variables (no energy when isothermal); passive tracers follow."""
    he = 0 if _is_iso(g) else 1
    return 2 * _comps(g) + 1 + he if _is_mhd(g) else _comps(g) + 1 + he


def _comps(g):
    if hasattr(g, "comps"):
        return g.comps
    return g.cfg.comps


def orszag_tang(g):
    """2-D Orszag-Tang vortex, test/MHD/OrszagTang/setup.cpp:31-51 (gamma 5/3)."""
    Vc, Vs = _empty(g)
    x, y = _coords(g)[0], _coords(g)[1]
    nx, ny, nz = g.np_tot
    X = x[None, None, :]
    Y = y[None, :, None]
    B0 = 1.0 / np.sqrt(4.0 * np.pi)
    Vc[0] = 25.0 / (36.0 * np.pi)
    Vc[3] = 5.0 / (12.0 * np.pi)
    Vc[1] = -np.sin(2.0 * np.pi * Y) + 0 * X
    Vc[2] = np.sin(2.0 * np.pi * X) + 0 * Y
    Vs[0, :nz, :ny, :nx] = -B0 * np.sin(2.0 * np.pi * Y) + 0 * X
    Vs[1, :nz, :ny, :nx] = B0 * np.sin(4.0 * np.pi * X) + 0 * Y
    return Vc, Vs


def orszag_tang_3d(g):
    """3-D Orszag-Tang vortex, test/MHD/OrszagTang3D/setup.cpp:213-263: the velocity field
    on cell centres and B = curl A from the edge vector potential
    A = B0/(2 pi) (cos 2 pi y, sin 2 pi x, cos 2 pi y + cos(4 pi x)/2) taken on the
    interior faces as DataBlockHost::MakeVsFromAmag does (dataBlockHost.cpp:186-214)."""
    Vc, Vs = _empty(g)
    x, y, z = _coords(g)
    xl, yl, _ = _coords(g, "xl")
    nx, ny, nz = g.np_tot
    X = x[None, None, :]
    Y = y[None, :, None]
    Z = z[:, None, None]
    B0 = 1.0 / np.sqrt(4.0 * np.pi)
    Vc[0] = 25.0 / (36.0 * np.pi)
    Vc[4] = 5.0 / (12.0 * np.pi)
    Vc[1] = -np.sin(2.0 * np.pi * Y) + 0 * X + 0 * Z
    Vc[2] = np.sin(2.0 * np.pi * X) + np.cos(2.0 * np.pi * Z) + 0 * Y
    Vc[3] = np.cos(2.0 * np.pi * X) + 0 * Y + 0 * Z
    # vector potential on cell corners (xl, yl, zl) of cells 0 .. np_tot-1
    XL = xl[None, None, :]
    YL = yl[None, :, None]
    c = B0 / (2.0 * np.pi)
    A = np.zeros((3, nz, ny, nx))
    A[0] = c * np.cos(2.0 * np.pi * YL) + 0 * XL
    A[1] = c * np.sin(2.0 * np.pi * XL) + 0 * YL
    A[2] = c * (np.cos(2.0 * np.pi * YL) + np.cos(4.0 * np.pi * XL) / 2.0)
    b, e = g.beg, g.end
    k0, k1 = b[2], e[2] + 1
    j0, j1 = b[1], e[1] + 1
    i0, i1 = b[0], e[0] + 1
    d1 = _widths(g, 0)[None, None, i0:i1]
    d2 = _widths(g, 1)[None, j0:j1, None]
    d3 = _widths(g, 2)[k0:k1, None, None]
    Ai = A[:, k0:k1, j0:j1, i0:i1]
    Aip = A[:, k0:k1, j0:j1, i0 + 1:i1 + 1]
    Ajp = A[:, k0:k1, j0 + 1:j1 + 1, i0:i1]
    Akp = A[:, k0 + 1:k1 + 1, j0:j1, i0:i1]
    Vs[0, k0:k1, j0:j1, i0:i1] = (0.0 + 1 / d2 * (Ajp[2] - Ai[2])) - 1 / d3 * (Akp[1] - Ai[1])
    Vs[1, k0:k1, j0:j1, i0:i1] = -1 / d1 * (Aip[2] - Ai[2]) + 1 / d3 * (Akp[0] - Ai[0])
    Vs[2, k0:k1, j0:j1, i0:i1] = 1 / d1 * (Aip[1] - Ai[1]) - 1 / d2 * (Ajp[0] - Ai[0])
    # passive tracers (setup.cpp:250-253): TRG = [x > 1/2], TRG+1 = [z > 1/2]
    trg = _nfluid(g)
    if g.nvar > trg:
        Vc[trg] = np.where(X > 0.5, 1.0, 0.0) + 0 * Y + 0 * Z
    if g.nvar > trg + 1:
        Vc[trg + 1] = np.where(Z > 0.5, 1.0, 0.0) + 0 * X + 0 * Y
    return Vc, Vs


class ParkMiller:
    """The reference's portable generator idfx::randm (global.cpp:126-137): state <-
    16807 * state mod (2^31 - 1) (exact in double), returns state / (2^31 - 1); rank 0
    seeds it with 13763."""

    def __init__(self, seed=13763):
        self.state = int(seed)

    def __call__(self):
        self.state = (16807 * self.state) % 2147483647
        return self.state / 2147483647.0

    def draws(self, n, block=4096):
        """The next n values of the sequence at once (the same numbers as n calls).
        Block jump-ahead: with s_q the state after q steps, s_{q+B} = a^B s_q mod M, so
        each block is the previous one times a^B mod M (products below 2^62 in int64)."""
        M, a = 2147483647, 16807
        n = int(n)
        out = np.empty(n, dtype=np.float64)
        if n == 0:
            return out
        b = min(block, n)
        first = np.empty(b, dtype=np.int64)
        s = self.state
        for q in range(b):
            s = (a * s) % M
            first[q] = s
        aB = pow(a, b, M)
        cur = first
        for o in range(0, n, b):
            m = min(b, n - o)
            out[o:o + m] = cur[:m] / float(M)
            last = int(cur[m - 1])
            cur = (cur * aB) % M
        self.state = last
        return out


def _half_space_modes(dim, kmin, kmax):
    """Integer wave vectors n with kmin <= |n| <= kmax, one of each +-n pair, in a fixed
    order (nz outermost, then ny, then nx)."""
    out = []
    r = range(-kmax, kmax + 1)
    for nz in (range(0, kmax + 1) if dim == 3 else [0]):
        for ny in (r if dim >= 2 else [0]):
            for nx in r:
                if nz < 0 or (nz == 0 and (ny < 0 or (ny == 0 and nx <= 0))):
                    continue
                k2 = nx * nx + ny * ny + nz * nz
                if kmin * kmin <= k2 <= kmax * kmax:
                    out.append((nx, ny, nz))
    return out


def _synth(n_int, dim, modes, coef, shift):
    """Real field sum_n Re(coef_n exp(2 pi i n.(x + shift))) on the interior grid
    x_j = j / N_d (per direction), through one inverse FFT; coef is complex per mode."""
    shape = tuple(n_int[d] for d in range(dim))[::-1]  # (nz, ny, nx) order of the arrays
    F = np.zeros(shape, dtype=np.complex128)
    total = float(np.prod(shape))
    for (nx, ny, nz), c in zip(modes, coef):
        ph = np.exp(2j * np.pi * (nx * shift[0] + ny * shift[1] + nz * shift[2]))
        idx = [nx % n_int[0]]
        if dim >= 2:
            idx.insert(0, ny % n_int[1])
        if dim >= 3:
            idx.insert(0, nz % n_int[2])
        F[tuple(idx)] += total * c * ph
    return np.real(np.fft.ifftn(F))


def _wrap(g, field, dim):
    """Periodic extension of an interior field ((nz,) ny, nx) to the index box
    np_tot + 1 per direction (cells and faces): out[k, j, i] = field[(k - beg_z) mod nz,
    (j - beg_y) mod ny, (i - beg_x) mod nx]."""
    f3 = field.reshape((1,) * (3 - field.ndim) + field.shape)
    idx = []
    for d in (2, 1, 0):
        n = f3.shape[2 - d]
        idx.append((np.arange(g.np_tot[d] + 1) - g.beg[d]) % n)
    return f3[np.ix_(idx[0], idx[1], idx[2])]


def turbulence(g, seed=13763, kmin=1, kmax=8, v_rms=0.5, db_rms=0.1, b0=0.1, gamma=5.0 / 3.0):
    """Decaying MHD turbulence on the periodic unit box (2-D or 3-D, three components):
    rho = 1, p = 1/gamma, uniform Bx = b0, plus random-phase solenoidal velocity and
    vector-potential modes with kmin <= |n| <= kmax and a shell spectrum E(k) ~ k^-2
    (per-mode amplitude |n|^-(dim+1)/2 for v, one power more for A so that B = curl A
    follows the same spectrum).  Draws from ParkMiller(seed), per mode in the order of
    _half_space_modes: velocity direction (3 draws, centred), velocity phase, potential
    direction (3 draws), potential phase.  The velocity is projected normal to n
    (solenoidal); v is sampled at cell centres and A at cell corners, B = discrete curl
    of A on the faces (MakeVsFromAmag, dataBlockHost.cpp:186-214), so div B = 0 to
    round-off.  v and delta-B are scaled to the requested rms over the interior."""
    dim = g.dim
    assert _comps(g) == 3 and _is_mhd(g), "turbulence needs three-component MHD"
    Vc, Vs = _empty(g)
    n_int = [g.end[d] - g.beg[d] if d < dim else 1 for d in range(3)]
    rng = ParkMiller(seed)
    modes = _half_space_modes(dim, kmin, kmax)
    cv = [[], [], []]
    ca = [[], [], []]
    for (nx, ny, nz) in modes:
        k = np.array([nx, ny, nz], dtype=np.float64)
        kk = np.sqrt(k @ k)
        u = np.array([rng() - 0.5, rng() - 0.5, rng() - 0.5])
        phi = 2.0 * np.pi * rng()
        u = u - (u @ k) / (kk * kk) * k
        amp = kk ** (-(dim + 1) / 2.0)
        w = np.array([rng() - 0.5, rng() - 0.5, rng() - 0.5])
        psi = 2.0 * np.pi * rng()
        ampa = kk ** (-(dim + 1) / 2.0 - 1.0)
        for c in range(3):
            cv[c].append(amp * u[c] * np.exp(1j * phi))
            ca[c].append(ampa * w[c] * np.exp(1j * psi))
    h = [0.5 / n_int[d] if d < dim else 0.0 for d in range(3)]
    v = [_synth(n_int, dim, modes, cv[c], h) for c in range(3)]
    A = [_synth(n_int, dim, modes, ca[c], [0.0, 0.0, 0.0]) for c in range(3)]
    vscale = v_rms / np.sqrt(np.mean(v[0] ** 2 + v[1] ** 2 + v[2] ** 2))
    # discrete curl of corner A on the interior faces (periodic differences)
    dxs = [1.0 / n_int[d] if d < dim else 1.0 for d in range(3)]
    ax = {0: -1, 1: -2, 2: -3}  # numpy axis of direction d in (nz, ny, nx) / (ny, nx)

    def dA(f, d):  # forward difference along direction d / width
        if d >= dim:
            return np.zeros_like(f)
        return (np.roll(f, -1, axis=ax[d]) - f) / dxs[d]

    bx = dA(A[2], 1) - dA(A[1], 2)
    by = -dA(A[2], 0) + dA(A[0], 2)
    if dim == 3:
        bz = dA(A[1], 0) - dA(A[0], 1)
    else:
        # cell-centred Bz from the corner potential, averaged over the two edges per side
        ay_x = dA(A[1], 0)
        ax_y = dA(A[0], 1)
        bz = 0.5 * (ay_x + np.roll(ay_x, -1, axis=-2)) - 0.5 * (ax_y + np.roll(ax_y, -1, axis=-1))
    bscale = db_rms / np.sqrt(np.mean(bx ** 2 + by ** 2 + bz ** 2))
    full = lambda f: _wrap(g, f, dim)[:g.np_tot[2], :g.np_tot[1], :g.np_tot[0]]
    fullf = lambda f: _wrap(g, f, dim)
    Vc[0] = 1.0
    Vc[4] = 1.0 / gamma
    for c in range(3):
        Vc[1 + c] = vscale * full(v[c])
    vz, vy, vx = Vs.shape[1:]
    Vs[0] = b0 + bscale * fullf(bx)[:vz, :vy, :vx]
    Vs[1] = bscale * fullf(by)[:vz, :vy, :vx]
    if dim == 3:
        Vs[2] = bscale * fullf(bz)[:vz, :vy, :vx]
        Vc[7] = 0.0
    else:
        Vc[7] = bscale * full(bz)
    # cell-centred B consistent with the faces (ReconstructVcField)
    Vc[5] = 0.5 * (Vs[0][:g.np_tot[2], :g.np_tot[1], :g.np_tot[0]] +
                   Vs[0][:g.np_tot[2], :g.np_tot[1], 1:g.np_tot[0] + 1])
    Vc[6] = 0.5 * (Vs[1][:g.np_tot[2], :g.np_tot[1], :g.np_tot[0]] +
                   Vs[1][:g.np_tot[2], 1:g.np_tot[1] + 1, :g.np_tot[0]])
    if dim == 3:
        Vc[7] = 0.5 * (Vs[2][:g.np_tot[2], :g.np_tot[1], :g.np_tot[0]] +
                       Vs[2][1:g.np_tot[2] + 1, :g.np_tot[1], :g.np_tot[0]])
    return Vc, Vs


def kelvin_helmholtz(g, seed=13763, v0=0.5, a=0.05, rho=1.0, p=2.5, b0=0.05, noise=0.01):
    """3-D MHD Kelvin-Helmholtz shear layer (three components): vx = v0 tanh((y - 1/2)/a),
    rho, p uniform, uniform Bx = b0, and white-noise vy = noise (2 u - 1)
    exp(-((y - 1/2)/a)^2) with u drawn from ParkMiller(seed) once per interior cell in
    (k, j, i) order."""
    Vc, Vs = _empty(g)
    x, y, z = _coords(g)
    comps = _comps(g)
    assert comps == 3 and _is_mhd(g)
    nx, ny, nz = g.np_tot
    Y = y[None, :, None] + 0 * x[None, None, :] + 0 * z[:, None, None]
    Vc[0] = rho
    Vc[4] = p
    Vc[1] = v0 * np.tanh((Y - 0.5) / a)
    rng = ParkMiller(seed)
    b, e = g.beg, g.end
    shape = (e[2] - b[2], e[1] - b[1], e[0] - b[0])
    u = rng.draws(int(np.prod(shape))).reshape(shape)
    env = np.exp(-((y[b[1]:e[1]] - 0.5) / a) ** 2)[None, :, None]
    Vc[2][b[2]:e[2], b[1]:e[1], b[0]:e[0]] = noise * (2.0 * u - 1.0) * env
    Vs[0] = b0
    Vc[5] = b0
    return Vc, Vs


def _widths(g, d):
    """Cell widths along direction d over np_tot cells."""
    if hasattr(g, "dx"):
        return np.asarray(g.dx[d], dtype=np.float64)
    xl = np.asarray(_coords(g, "xl")[d], dtype=np.float64)
    xr = np.asarray(_coords(g, "xr")[d], dtype=np.float64)
    return xr - xl


def sod(g, x0=0.5):
    """1-D Sod shock tube, test/HD/sod/setup.cpp:35-38 (interface at x0)."""
    Vc, _ = _empty(g)
    x = _coords(g)[0][None, None, :]
    Vc[0] = np.where(x > x0, 0.125, 1.0)
    Vc[1] = 0.0
    Vc[_comps(g) + 1] = np.where(x > x0, 0.1, 1.0)
    return Vc, None


def random_mhd(g, seed=0, amp=0.3):
    """Smooth random perturbation of a uniform magnetised state (parity tests)."""
    rng = np.random.default_rng(seed)
    Vc, Vs = _empty(g)
    comps = _comps(g)
    shape = Vc.shape[1:]
    Vc[0] = 1.0 + amp * rng.uniform(-1, 1, shape)
    for c in range(comps):
        Vc[1 + c] = amp * rng.uniform(-1, 1, shape)
    he = 0 if _is_iso(g) else 1
    prs = 0.6 + amp * rng.uniform(-1, 1, shape)  # drawn either way: same sequence
    if he:
        Vc[comps + 1] = prs
    if Vs is not None:
        for c in range(Vs.shape[0]):
            Vs[c] = (0.4 - 0.1 * c) + amp * rng.uniform(-1, 1, Vs.shape[1:])
        for c in range(Vs.shape[0], comps):
            Vc[comps + 1 + he + c] = 0.2 + amp * rng.uniform(-1, 1, shape)
    return Vc, Vs


def random_hd(g, seed=0, amp=0.3):
    rng = np.random.default_rng(seed)
    Vc, _ = _empty(g)
    comps = _comps(g)
    shape = Vc.shape[1:]
    Vc[0] = 1.0 + amp * rng.uniform(-1, 1, shape)
    for c in range(comps):
        Vc[1 + c] = amp * rng.uniform(-1, 1, shape)
    prs = 0.6 + amp * rng.uniform(-1, 1, shape)
    if not _is_iso(g):
        Vc[comps + 1] = prs
    return Vc, None


def linear_fast_wave_2d(g, eps=1e-6, gamma=5.0 / 3.0, Lx=2.0, Ly=1.0):
    """Oblique fast magnetosonic eigenmode on a periodic [0,Lx]x[0,Ly] box.

    The background and the right eigenvector are those of the reference's linear-wave
    test (test/MHD/LinearWaveTest/setup.cpp:27-70, Gardiner & Stone 2008): rho=1, p=1/gamma,
    B=(1, 3/2, 0) in the frame of the wave, fast speed 2.  The wave vector
    k = 2 pi (1/Lx, 1/Ly) makes the box hold exactly one wavelength along each axis.
    Face fields come from a corner vector potential, so div B = 0 to round-off.
    Returns (Vc, Vs, period)."""
    Vc, Vs = _empty(g)
    x, y = _coords(g)[0], _coords(g)[1]
    xl, yl = _coords(g, "xl")[0], _coords(g, "xl")[1]
    kx, ky = 2 * np.pi / Lx, 2 * np.pi / Ly
    kk = np.hypot(kx, ky)
    ct, st = kx / kk, ky / kk
    R0 = 1.0 / (2.0 * np.sqrt(5.0))
    Rrho, Rvpar, Rvperp, RBperp, RE = 2.0, 4.0, -2.0, 4.0, 9.0
    Bpar, Bperp = 1.0, 1.5
    Rp = (gamma - 1.0) * (RE - Bperp * RBperp)
    X = x[None, None, :]
    Y = y[None, :, None]
    ph = np.cos(kx * X + ky * Y)
    a = eps * R0
    Vc[0] = 1.0 + a * Rrho * ph
    vpar, vperp = a * Rvpar * ph, a * Rvperp * ph
    Vc[1] = vpar * ct - vperp * st
    Vc[2] = vpar * st + vperp * ct
    Vc[3] = 1.0 / gamma + a * Rp * ph
    # B = curl(Az z): Az = B0x y - B0y x - (a RBperp / |k|) sin(k.x)
    B0x = Bpar * ct - Bperp * st
    B0y = Bpar * st + Bperp * ct
    amp = a * RBperp / kk

    def Az(xx, yy):
        return B0x * yy - B0y * xx - amp * np.sin(kx * xx + ky * yy)
    nx, ny = len(x), len(y)
    dxv = np.diff(xl)[0] if nx > 1 else 1.0
    dyv = np.diff(yl)[0] if ny > 1 else 1.0
    XL = xl[None, :]
    YC = y[:, None]
    Vs[0, 0, :ny, :nx] = (Az(XL, YC + 0.5 * dyv) - Az(XL, YC - 0.5 * dyv)) / dyv
    XC = x[None, :]
    YL = yl[:, None]
    Vs[1, 0, :ny, :nx] = -(Az(XC + 0.5 * dxv, YL) - Az(XC - 0.5 * dxv, YL)) / dxv
    cf = 2.0
    period = (1.0 / np.hypot(1.0 / Lx, 1.0 / Ly)) / cf
    return Vc, Vs, period


def linear_fast_wave_1d(g, eps=1e-6, gamma=5.0 / 3.0, L=1.0):
    """The fast magnetosonic eigenmode of linear_fast_wave_2d propagating along x on a
    periodic [0, L] line (two-component MHD, B = (1, 3/2) in the background, fast speed 2):
    Bx is the constant face field, By a cell variable.  Returns (Vc, Vs, period)."""
    Vc, Vs = _empty(g)
    x = _coords(g)[0]
    k = 2 * np.pi / L
    R0 = 1.0 / (2.0 * np.sqrt(5.0))
    Rrho, Rvpar, Rvperp, RBperp, RE = 2.0, 4.0, -2.0, 4.0, 9.0
    Bpar, Bperp = 1.0, 1.5
    Rp = (gamma - 1.0) * (RE - Bperp * RBperp)
    ph = np.cos(k * x)[None, None, :]
    a = eps * R0
    Vc[0] = 1.0 + a * Rrho * ph
    Vc[1] = a * Rvpar * ph
    Vc[2] = a * Rvperp * ph
    Vc[3] = 1.0 / gamma + a * Rp * ph
    Vc[4] = Bpar
    Vc[5] = Bperp + a * RBperp * ph
    Vs[0] = Bpar
    return Vc, Vs, L / 2.0


# Gardiner & Stone (2008) eigenvectors of the reference's linear-wave test
# (test/MHD/LinearWaveTest/setup.cpp:48-98): (R0, right eigenvector in the wave frame
# (rho, v1, v2, v3, B1, B2, B3, E)) of each |mode|; the sign of `mode` picks the
# propagation direction of the fast, Alfven and slow waves.  The entropy mode rides on a
# mean flow v1 = 1.
LINEAR_WAVE_MODES = {
    1: ("fast", 1.0 / (2.0 * np.sqrt(5.0)), (2.0, 4.0, -2.0, 0.0, 0.0, 4.0, 0.0, 9.0), 2.0),
    2: ("alfven", 1.0, (0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 1.0, 0.0), 1.0),
    3: ("slow", 1.0 / (2.0 * np.sqrt(5.0)), (4.0, 2.0, 4.0, 0.0, 0.0, -2.0, 0.0, 3.0), 0.5),
    4: ("entropy", 0.5, (2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0), 1.0),
}


def linear_wave_3d(g, mode, eps=1e-6, gamma=5.0 / 3.0):
    """The reference's oblique linear MHD wave (test/MHD/LinearWaveTest/setup.cpp:29-120):
    background rho = 1, p = 1/gamma, B = (1, 3/2, 0) in the wave frame (plus v1 = 1 for the
    entropy mode), perturbed by eps * R0 * R * cos(2 pi x1) along
    x1 = cos(a) cos(b) x + cos(a) sin(b) y + sin(a) z with sin(a) = 2/3, sin(b) = 2/sqrt(5).
    The face field is the discrete curl of the edge vector potential
    A = (0, B3 x1 + R0 R_B3 phi, -(B2 x1 + R0 R_B2 phi) + B1 x2) (wave frame), phi =
    eps sin(2 pi x1) / (2 pi), as MakeVsFromAmag builds it (dataBlockHost.cpp:186-214).
    For the fast and slow modes the pressure perturbation is (gamma - 1) (R_E - B.R_B)
    (plus - v1 R_rho / 2 for the entropy mode).  Returns (Vc, Vs, period): the wavelength
    along x1 is 1, so the period is 1 / wave speed."""
    name, R0, R, speed = LINEAR_WAVE_MODES[abs(mode)]
    sign = 1.0 if mode > 0 else -1.0
    Vc, Vs = _empty(g)
    Bg = np.array([1.0, 1.5, 0.0])  # background B in the wave frame
    vg = np.array([1.0 if abs(mode) == 4 else 0.0, 0.0, 0.0])
    Rr, R1, R2, R3, RB1, RB2, RB3, RE = R
    if abs(mode) in (1, 3):  # fast and slow modes propagate along sign
        R1, R2 = sign * R1, sign * R2
    if abs(mode) == 2:
        R3 = sign * R3
    Rp = (gamma - 1.0) * (RE - Bg @ np.array([RB1, RB2, RB3]) - vg[0] * Rr / 2.0)
    sa, sb = 2.0 / 3.0, 2.0 / np.sqrt(5.0)
    ca, cb = np.sqrt(1.0 - sa * sa), np.sqrt(1.0 - sb * sb)
    e1 = np.array([ca * cb, ca * sb, sa])
    e2 = np.array([-sb, cb, 0.0])
    e3 = np.array([-sa * cb, -sa * sb, ca])
    x, y, z = _coords(g)
    xl, yl, zl = _coords(g, "xl")
    X, Y, Z = x[None, None, :], y[None, :, None], z[:, None, None]
    x1 = e1[0] * X + e1[1] * Y + e1[2] * Z
    ph = eps * np.cos(2.0 * np.pi * x1)
    Vc[0] = 1.0 + R0 * Rr * ph
    Vc[4] = 1.0 / gamma + R0 * Rp * ph
    for c in range(3):
        Vc[1 + c] = (vg[0] * e1[c] + vg[1] * e2[c] + vg[2] * e3[c] +
                     R0 * (R1 * e1[c] + R2 * e2[c] + R3 * e3[c]) * ph)

    def wave_frame_A(xx, yy, zz):
        q1 = e1[0] * xx + e1[1] * yy + e1[2] * zz
        q2 = e2[0] * xx + e2[1] * yy
        phase = eps * np.sin(2.0 * np.pi * q1) / (2.0 * np.pi)
        return Bg[2] * q1 + R0 * RB3 * phase, -(Bg[1] * q1 + R0 * RB2 * phase) + Bg[0] * q2
    nx, ny, nz = g.np_tot
    A = np.zeros((3, nz, ny, nx))
    # component c of A lives on the edge along c: (x, yl, zl), (xl, y, zl), (xl, yl, z)
    for c, (xx, yy, zz) in enumerate(((X, yl[None, :, None], zl[:, None, None]),
                                      (xl[None, None, :], Y, zl[:, None, None]),
                                      (xl[None, None, :], yl[None, :, None], Z))):
        A2, A3 = wave_frame_A(xx, yy, zz)
        A[c] = A2 * e2[c] + A3 * e3[c] + 0 * X * Y * Z
    b, e = g.beg, g.end
    k0, k1 = b[2], e[2] + 1
    j0, j1 = b[1], e[1] + 1
    i0, i1 = b[0], e[0] + 1
    d1 = _widths(g, 0)[None, None, i0:i1]
    d2 = _widths(g, 1)[None, j0:j1, None]
    d3 = _widths(g, 2)[k0:k1, None, None]
    Ai = A[:, k0:k1, j0:j1, i0:i1]
    Aip = A[:, k0:k1, j0:j1, i0 + 1:i1 + 1]
    Ajp = A[:, k0:k1, j0 + 1:j1 + 1, i0:i1]
    Akp = A[:, k0 + 1:k1 + 1, j0:j1, i0:i1]
    Vs[0, k0:k1, j0:j1, i0:i1] = (0.0 + 1 / d2 * (Ajp[2] - Ai[2])) - 1 / d3 * (Akp[1] - Ai[1])
    Vs[1, k0:k1, j0:j1, i0:i1] = -1 / d1 * (Aip[2] - Ai[2]) + 1 / d3 * (Akp[0] - Ai[0])
    Vs[2, k0:k1, j0:j1, i0:i1] = 1 / d1 * (Aip[1] - Ai[1]) - 1 / d2 * (Ajp[0] - Ai[0])
    return Vc, Vs, 1.0 / speed


def linear_wave_error(V0, V1, beg, end):
    """Error measure of the reference's linear-wave test
    (test/MHD/LinearWaveTest/python/testidefix.py:26-49): for RHO, VX1..3, PRS and the
    three face fields, the mean absolute difference of the mean-subtracted initial and
    final fields; the result is the rms of the eight.  V0, V1 = (Vc, Vs) pairs."""
    b, e = beg, end
    errs = []
    for n in (0, 1, 2, 3, 4):
        q0 = V0[0][n, b[2]:e[2], b[1]:e[1], b[0]:e[0]]
        q1 = V1[0][n, b[2]:e[2], b[1]:e[1], b[0]:e[0]]
        errs.append(np.mean(np.abs((q0 - q0.mean()) - (q1 - q1.mean()))))
    for d in range(3):
        sl = [slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0])]
        sl[2 - d] = slice(b[d], e[d] + 1)
        q0 = V0[1][d][tuple(sl)]
        q1 = V1[1][d][tuple(sl)]
        errs.append(np.mean(np.abs((q0 - q0.mean()) - (q1 - q1.mean()))))
    return float(np.sqrt(np.mean(np.square(errs))))


def brio_wu(g, x0=50.0):
    """MHD shock tube of test/MHD/sod/setup.cpp:20-40 (Brio & Wu 1988 states, the
    reference's gamma 5/3 and [0, 100] domain): rho, p = 1, 1 | 0.125, 0.1; Bx = 0.75;
    By = 1 | -1; at rest."""
    Vc, Vs = _empty(g)
    x = _coords(g)[0][None, None, :]
    comps = _comps(g)
    Vc[0] = np.where(x < x0, 1.0, 0.125) + 0 * Vc[0]
    Vc[comps + 1] = np.where(x < x0, 1.0, 0.1) + 0 * Vc[0]
    Vc[comps + 2] = 0.75
    Vc[comps + 3] = np.where(x < x0, 1.0, -1.0) + 0 * Vc[0]
    Vs[0] = 0.75
    return Vc, Vs


def sod_iso(g, x0=0.5):
    """Isothermal shock tube of test/HD/sod-iso/setup.cpp:20-30: rho = 1 | 0.125 at rest
    (csiso 1 in its input)."""
    Vc, Vs = _empty(g)
    x = _coords(g)[0][None, None, :]
    Vc[0] = np.where(x > x0, 0.125, 1.0) + 0 * Vc[0]
    return Vc, Vs


def brio_wu_iso(g, x0=50.0):
    """Isothermal MHD shock tube of test/MHD/sod-iso/setup.cpp:20-40: rho = 1 | 0.125,
    Bx = 0.75, By = 1 | -1, at rest, on [0, 100] (two components)."""
    Vc, Vs = _empty(g)
    x = _coords(g)[0][None, None, :]
    comps = _comps(g)
    Vc[0] = np.where(x < x0, 1.0, 0.125) + 0 * Vc[0]
    Vc[comps + 1] = 0.75
    Vc[comps + 2] = np.where(x < x0, 1.0, -1.0) + 0 * Vc[0]
    Vs[0] = 0.75
    return Vc, Vs
