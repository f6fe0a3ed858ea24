"""Learned limiter: the training handle over the C ABI (idfx_mlp_*) and the data
pipeline that builds training samples from block-averaged high-resolution runs.

No counterpart in the reference.  The network, its features and the interface states
are specified in include/idefix_amd.h; the HIP kernels are in csrc/mlp.hpp (BF16 MFMA
inference) and csrc/mlp_train.hip (FP32 training).
"""
import ctypes

import numpy as np

from . import capi

# face-frame variable order of the features: rho, v_n, v_t, v_b, p, B_n, B_t, B_b, with
# (n, t, b) = (x, y, z) for x faces, (y, x, z) for y faces, (z, x, y) for z faces
# (the Xn/Xt/Xb convention of the sweeps, hlldMHD.hpp:153-159)
FRAME = {0: [0, 1, 2, 3, 4, 5, 6, 7], 1: [0, 2, 1, 3, 4, 6, 5, 7], 2: [0, 3, 1, 2, 4, 7, 5, 6]}


def param_count():
    return capi.load().idfx_mlp_param_count()


def init_params(seed=0):
    """Glorot-uniform weights and zero biases, packed FP32 (W1, b1, W2, b2, W3, b3)."""
    rng = np.random.default_rng(seed)
    parts = []
    for n_in, n_out in ((48, 64), (64, 64), (64, 16)):
        lim = np.sqrt(6.0 / (n_in + n_out))
        parts.append(rng.uniform(-lim, lim, size=(n_in, n_out)).ravel())
        parts.append(np.zeros(n_out))
    return np.concatenate(parts).astype(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (48, 64)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)),
                        ("W3", (64, 16)), ("b3", (16,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(P):
    return np.concatenate([np.asarray(P[k], dtype=np.float64).ravel()
                           for k in ("W1", "b1", "W2", "b2", "W3", "b3")]).astype(np.float32)


class Trainer:
    """FP32 training state on the GPU (parameters and Adam moments)."""

    def __init__(self, max_batch, params=None):
        import torch
        self.torch = torch
        self._L = capi.load()
        h = ctypes.c_void_p()
        capi.check(self._L.idfx_mlp_create(int(max_batch), ctypes.byref(h)), "idfx_mlp_create")
        self.h = h
        self.max_batch = max_batch
        self.set_params(init_params() if params is None else params)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self._L.idfx_mlp_destroy(h)
            self.h = None

    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        capi.check(self._L.idfx_mlp_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "idfx_mlp_set_params")

    def params(self):
        p = np.zeros(param_count(), dtype=np.float32)
        capi.check(self._L.idfx_mlp_get_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "idfx_mlp_get_params")
        return p

    def loss_grad(self, X, tau):
        """X, tau: float32 cuda tensors [B, 72] (48 features, the 16 PPM weights, 8 loss
        weights) and [B, 8] (normalised targets).  Returns (loss, grad numpy)."""
        g = np.zeros(param_count(), dtype=np.float32)
        loss = ctypes.c_double()
        capi.check(self._L.idfx_mlp_loss_grad(self.h, ctypes.c_void_p(X.data_ptr()),
                                              ctypes.c_void_p(tau.data_ptr()), X.shape[0],
                                              ctypes.byref(loss),
                                              g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                              self._stream()), "idfx_mlp_loss_grad")
        return loss.value, g

    def step(self, X, tau, lr):
        capi.check(self._L.idfx_mlp_train_step(self.h, ctypes.c_void_p(X.data_ptr()),
                                               ctypes.c_void_p(tau.data_ptr()), X.shape[0],
                                               ctypes.c_float(lr), self._stream()),
                   "idfx_mlp_train_step")

    def loss(self):
        v = ctypes.c_double()
        capi.check(self._L.idfx_mlp_loss(self.h, ctypes.byref(v), self._stream()), "idfx_mlp_loss")
        return v.value


def stencils(V, d, lo_face, hi_face):
    """Host helper: the 6-cell stencils [faces, 8, 6] of the interfaces lo..hi-1 along
    direction d of a primitive array V [8][nz][ny][nx] (numpy), face-frame variable
    order, every transverse index whose stencil is inside the array."""
    Vf = V[FRAME[d]]
    ax = 3 - d
    n = V.shape[ax]
    faces = np.arange(lo_face, hi_face)
    assert faces.min() >= 3 and faces.max() + 2 < n
    idx = faces[:, None] + np.arange(-3, 3)[None, :]       # [F, 6]
    S = np.take(Vf, idx, axis=ax)                          # inserts [F, 6] at ax
    # move the face and stencil axes to the front: [F, 6, 8, rest...]
    S = np.moveaxis(S, (ax, ax + 1), (0, 1))
    S = np.moveaxis(S, 2, -1)                              # [F, 6, rest..., 8]
    S = np.moveaxis(S, 1, -1)                              # [F, rest..., 8, 6]
    return S


# ---------------------------------------------------------------------------------------
# Training data: block-averaged high-resolution runs (2-D, three components, periodic)
# ---------------------------------------------------------------------------------------
def interior(g):
    """Interior primitives [8, ny, nx] and face fields (bx [ny, nx+1], by [ny+1, nx]) of a
    2-D DataBlock, as torch views."""
    b, e = g.beg, g.end
    V = g.Vc[:, 0, b[1]:e[1], b[0]:e[0]]
    bx = g.Vs[0, 0, b[1]:e[1], b[0]:e[0] + 1]
    by = g.Vs[1, 0, b[1]:e[1] + 1, b[0]:e[0]]
    return V, bx, by


def coarse_state(V, bx, by, f, gamma):
    """Block averages over f x f cells of a fine state (torch, float64): conserved
    variables averaged over the block, face fields over the coarse face; returns coarse
    primitives [8, N, N] (cell B from the coarse faces, Bz averaged) and coarse faces."""
    import torch
    F = torch.nn.functional
    rho = V[0]
    v = V[1:4]
    B = V[5:8]
    E = V[4] / (gamma - 1.0) + 0.5 * rho * (v * v).sum(0) + 0.5 * (B * B).sum(0)
    pool = lambda a: F.avg_pool2d(a[None, None], f)[0, 0]
    rc = pool(rho)
    mc = torch.stack([pool(rho * v[c]) for c in range(3)])
    Ec = pool(E)
    bzc = pool(B[2])
    # coarse faces: every f-th fine face, averaged over the f fine faces along it
    bxc = bx[:, ::f].reshape(-1, f, bx[:, ::f].shape[1]).mean(1)
    byc = by[::f, :].reshape(by[::f, :].shape[0], -1, f).mean(2)
    Bc = torch.stack([0.5 * (bxc[:, :-1] + bxc[:, 1:]), 0.5 * (byc[:-1] + byc[1:]), bzc])
    vc = mc / rc
    pc = (gamma - 1.0) * (Ec - 0.5 * rc * (vc * vc).sum(0) - 0.5 * (Bc * Bc).sum(0))
    Vc = torch.cat([rc[None], vc, pc[None], Bc])
    return Vc, bxc, byc


def face_targets(V, bx, by, f):
    """Interface values of the fine solution on the coarse faces (face-frame order):
    x faces [8, N, N] at coarse face I (between coarse cells I-1 and I, periodic), y faces
    likewise; the mean over the coarse face of (V(left) + V(right)) / 2, and the face
    field itself for the normal component."""
    import torch
    N = V.shape[-1] // f
    left = torch.roll(V, 1, dims=2)[:, :, ::f]            # cell i-1 for fine faces i = f I
    tx = 0.5 * (left + V[:, :, ::f])                     # [8, n_fine, N]
    tx = tx.reshape(8, N, f, N).mean(2)
    tx[5] = bx[:, ::f][:, :N].reshape(N, f, N).mean(1)
    down = torch.roll(V, 1, dims=1)[:, ::f, :]
    ty = 0.5 * (down + V[:, ::f, :])                     # [8, N, n_fine]
    ty = ty.reshape(8, N, N, f).mean(3)
    ty[6] = by[::f, :][:N].reshape(N, N, f).mean(2)
    return tx[FRAME[0]], ty[FRAME[1]]


def samples(Vc, targets, d, wp):
    """Training rows of every coarse interface along d (periodic stencils), from coarse
    primitives Vc [8, N, N], the fine interface values `targets` [8, N, N] and the PPM
    weights of the same interfaces wp [16, N, N] (face index = cell right of the face):
    rows [F, 72] = features (48) | PPM weights (16) | loss weights (8, set to 1 here),
    normalised targets [F, 8] and the stencil scales s [F, 8] (float32, torch)."""
    import torch
    Vf = Vc[FRAME[d]]
    dim = 2 if d == 0 else 1
    S = torch.stack([torch.roll(Vf, 3 - s, dims=dim) for s in range(6)], dim=-1)  # [8,N,N,6]
    S = S.permute(1, 2, 0, 3).reshape(-1, 8, 6)          # [F, 8, 6]
    t = targets.permute(1, 2, 0).reshape(-1, 8)
    c = 0.5 * (S[..., 2] + S[..., 3])
    s = (S - c[..., None]).abs().amax(-1)
    safe = torch.where(s > 0, s, torch.ones_like(s))
    x = torch.where(s[..., None] > 0, (S - c[..., None]) / safe[..., None], torch.zeros_like(S))
    # targets outside [-1, 1] (beyond the stencil's range) are unreachable by any weight
    # and would dominate the squared loss; clip them
    tau = torch.where(s > 0, (t - c) / safe, torch.zeros_like(s)).clamp(-1.0, 1.0)
    w = wp.permute(1, 2, 0).reshape(-1, 16).to(x.dtype)
    rows = torch.cat([x.reshape(-1, 48), w, torch.ones_like(s)], dim=1)
    return rows.float(), tau.float(), s.float()


def ppm_face_weights(Vc, bxc, byc):
    """PPM weights [16, N, N] of the x and y interfaces of a coarse periodic state, from the
    inference kernel itself with no network loaded (the fallback path)."""
    g = coarse_block(Vc, bxc, byc, "mlp", "fp32")
    g.set_boundaries()
    g.set_mlp_params(None)
    b, e = g.beg, g.end
    out = []
    for d in (0, 1):
        W = g.mlp_faces(d)
        out.append(W[:, 0, b[1]:e[1], b[0]:e[0]].double().clone())
    del g
    return out


def coarse_block(Vc, bxc, byc, recon, precision, nstages=3, cfl=0.4, emf="uct_contact",
                 gamma=5.0 / 3.0):
    """A 2-D periodic DataBlock holding a coarse state (interior given, ghosts filled by
    the first boundary step)."""
    import torch
    from .datablock import DataBlock
    N = Vc.shape[-1]
    g = DataBlock(2, (N, N), comps=3, mhd=True, recon=recon, solver="hlld", emf=emf,
                  precision=precision, nstages=nstages, cfl=cfl, gamma=gamma)
    b, e = g.beg, g.end
    with torch.no_grad():
        g.Vc.zero_()
        g.Vs.zero_()
        g.Vc[:, 0, b[1]:e[1], b[0]:e[0]] = Vc
        g.Vs[0, 0, b[1]:e[1], b[0]:e[0] + 1] = bxc
        g.Vs[1, 0, b[1]:e[1] + 1, b[0]:e[0]] = byc
    return g


def l1_error(V, bx, by, Vref, bxref, byref):
    """Mean over variables (rho, v1..3, p, face Bx, face By, Bz) of the mean absolute
    difference over the cells."""
    errs = [float((V[n] - Vref[n]).abs().mean()) for n in (0, 1, 2, 3, 4, 7)]
    errs.append(float((bx - bxref).abs().mean()))
    errs.append(float((by - byref).abs().mean()))
    return float(np.mean(errs)), errs


def fine_run(n, seed, times, recon="ppm", nstages=3, cfl=0.4, emf="uct_contact", log=None):
    """Plain FP64 run of 2-D decaying turbulence at n^2, yielding (t, V, bx, by) torch
    copies of the interior at each requested time."""
    from .datablock import DataBlock
    from . import problems
    g = DataBlock(2, (n, n), comps=3, mhd=True, recon=recon, solver="hlld", emf=emf,
                  precision="fp64", nstages=nstages, cfl=cfl)
    Vc, Vs = problems.turbulence(g, seed=seed)
    g.upload(Vc, Vs)
    g.set_boundaries()
    out = []
    V, bx, by = interior(g)
    out.append((0.0, V.clone(), bx.clone(), by.clone()))
    for t in times:
        g.run(t)
        V, bx, by = interior(g)
        out.append((g.t, V.clone(), bx.clone(), by.clone()))
        if log:
            log("  fine n=%d seed=%d t=%.3f cycles=%d" % (n, seed, g.t, g.ncycles))
    return out


def run_coarse(V, bx, by, tend, recon, precision, params=None, nstages=3, cfl=0.4):
    """Runs a coarse 2-D block from the given interior state to tend; returns
    (V, bx, by, wall seconds, cycles)."""
    import time
    import torch
    g = coarse_block(V, bx, by, recon, precision, nstages=nstages, cfl=cfl)
    if recon == "mlp":
        g.set_mlp_params(params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(tend)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    Vo, bxo, byo = interior(g)
    return Vo.clone(), bxo.clone(), byo.clone(), wall, g.ncycles


def main(argv=None):
    """Training pipeline: fine plain runs -> block-averaged samples -> FP32 training on the
    GPU -> comparison of the learned coarse run with plain coarse runs against the
    block-averaged fine solution.  Writes params.npy and report.json into --out."""
    import argparse
    import json
    import os
    import time
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=2048)
    ap.add_argument("--factor", type=int, default=8)
    ap.add_argument("--seeds", type=int, nargs="+", default=[1, 2, 3])
    ap.add_argument("--test-seed", type=int, default=4)
    ap.add_argument("--tend", type=float, default=1.0)
    ap.add_argument("--snapshots", type=int, default=10)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="learned_out")
    ap.add_argument("--no-calibration", action="store_true",
                    help="keep the a-priori network (no a-posteriori output scales)")
    ap.add_argument("--loss-weights", choices=("unit", "abs"), default="unit",
                    help="per-sample loss weights: 1 (scale-free interface errors) or the "
                         "stencil scale over the variable's mean scale (absolute errors)")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    gamma = 5.0 / 3.0
    t_start = time.perf_counter()
    log = lambda s: print("[%.0fs] %s" % (time.perf_counter() - t_start, s), flush=True)
    times = [a.tend * (k + 1) / a.snapshots for k in range(a.snapshots)]
    f = a.factor
    N = a.fine // f
    # ---- training samples -----------------------------------------------------------
    Xs, Ts, Ss = [], [], []
    calib = []  # (coarse initial state, coarse reference at tend) of each training seed
    for seed in a.seeds:
        snaps = fine_run(a.fine, seed, times, log=log)
        calib.append((coarse_state(*snaps[0][1:], f, gamma), coarse_state(*snaps[-1][1:], f, gamma)))
        for (t, V, bx, by) in snaps:
            Vc, bxc, byc = coarse_state(V, bx, by, f, gamma)
            tx, ty = face_targets(V, bx, by, f)
            wps = ppm_face_weights(Vc, bxc, byc)
            for d, tgt in ((0, tx), (1, ty)):
                X, T, S = samples(Vc, tgt, d, wps[d])
                Xs.append(X)
                Ts.append(T)
                Ss.append(S)
        torch.cuda.empty_cache()
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    S = torch.cat(Ss)
    del Xs, Ts, Ss
    if a.loss_weights == "abs":
        # absolute interface errors, each variable on the scale of its mean local variation
        X[:, 64:72] = S / S.mean(0, keepdim=True).clamp_min(1e-30)
    log("samples: %d" % X.shape[0])
    # ---- training -------------------------------------------------------------------
    p0 = unpack(init_params(0))
    p0["W3"][:] = 0.0  # the network starts as PPM (zero correction)
    tr = Trainer(a.batch, pack(p0))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    hist = []
    for step in range(a.steps):
        idx = torch.randint(0, X.shape[0], (a.batch,), device="cuda", generator=gen)
        xb = X[idx].contiguous()
        tb = T[idx].contiguous()
        lr = a.lr * (0.5 * (1.0 + np.cos(np.pi * step / a.steps)))  # cosine decay
        tr.step(xb, tb, max(lr, 1e-5))
        if step % 500 == 0 or step == a.steps - 1:
            hist.append((step, tr.loss()))
            log("step %d loss %.5f" % hist[-1])
    params = tr.params()
    np.save(os.path.join(a.out, "params_apriori.npy"), params)
    # ---- a-posteriori calibration on the training seeds --------------------------------
    alpha = np.zeros(8) if a.no_calibration else None
    if not a.no_calibration:
        alpha, cal_log = calibrate(params, calib, a.tend, log)
        params = scale_outputs(params, alpha)
    np.save(os.path.join(a.out, "params.npy"), params)
    # the van Leer member's loss on the same data, for reference
    vl = van_leer_loss(X[:200000], T[:200000])
    ppm = ppm_loss(X[:200000], T[:200000])
    # ---- evaluation on the held-out seed ---------------------------------------------
    fine = fine_run(a.fine, a.test_seed, [a.tend], log=log)
    (_, V0, bx0, by0), (_, V1, bx1, by1) = fine[0], fine[-1]
    ref = coarse_state(V1, bx1, by1, f, gamma)
    rep = {"fine": a.fine, "coarse": N, "factor": f, "tend": a.tend, "train_seeds": a.seeds,
           "test_seed": a.test_seed, "samples": int(X.shape[0]), "loss_history": hist,
           "van_leer_loss": vl, "ppm_loss": ppm, "loss_weights": a.loss_weights,
           "output_scales": [float(v) for v in alpha] if alpha is not None else None,
           "calibration": None if a.no_calibration else cal_log, "runs": {}}
    init = coarse_state(V0, bx0, by0, f, gamma)
    cases = [("learned_fp32_%d" % N, N, "mlp", "fp32", params),
             ("plain_ppm_fp64_%d" % N, N, "ppm", "fp64", None),
             ("plain_plm_fp64_%d" % N, N, "plm", "fp64", None),
             ("ppm_via_mlp_fp32_%d" % N, N, "mlp", "fp32", None),
             ("plain_ppm_fp64_%d" % (2 * N), 2 * N, "ppm", "fp64", None)]
    for name, n, recon, prec, p in cases:
        st = init if n == N else coarse_state(V0, bx0, by0, f // 2, gamma)
        Vo, bxo, byo, wall, cyc = run_coarse(*st, a.tend, recon, prec, p)
        if n != N:
            Vo, bxo, byo = coarse_state(Vo, bxo, byo, 2, gamma)
        l1, per = l1_error(Vo, bxo, byo, *ref)
        rep["runs"][name] = {"n": n, "recon": recon, "precision": prec, "L1": l1,
                             "L1_per_variable": per, "wall_s": wall, "cycles": cyc}
        log("%s: L1 %.4e wall %.2fs cycles %d" % (name, l1, wall, cyc))
    with open(os.path.join(a.out, "report.json"), "w") as fh:
        json.dump(rep, fh, indent=1)
    log("done")


def scale_outputs(params, alpha):
    """The network with the correction of variable v (output 2v) scaled by alpha[v]."""
    P = unpack(params)
    for v in range(8):
        P["W3"][:, 2 * v] *= alpha[v]
        P["b3"][2 * v] *= alpha[v]
    return pack(P)


def calibrate(params, calib, tend, log):
    """A-posteriori calibration of the a-priori network: per-variable scales alpha of its
    corrections that minimise the mean L1 error of coarse learned runs against the
    block-averaged fine solutions of the training seeds (alpha = 0 is PPM, so the result
    is never worse than PPM on these seeds).  A line search of a common scale, then one
    coordinate sweep per variable."""
    def err(alpha):
        p = scale_outputs(params, alpha)
        e = []
        for init, ref in calib:
            Vo, bxo, byo, _, _ = run_coarse(*init, tend, "mlp", "fp32", p)
            e.append(l1_error(Vo, bxo, byo, *ref)[0])
        return float(np.mean(e))
    hist = []
    best_a, best_e = None, None
    for c in (0.0, 0.15, 0.3, 0.5, 0.75, 1.0):
        e = err(np.full(8, c))
        hist.append(([c] * 8, e))
        log("calibration: common scale %.2f -> L1 %.4e" % (c, e))
        if best_e is None or e < best_e:
            best_a, best_e = np.full(8, c), e
    for v in range(8):
        for m in (0.0, 0.5, 1.5, 2.0):
            trial = best_a.copy()
            trial[v] = best_a[v] * m if best_a[v] > 0 else 0.3 * m
            e = err(trial)
            hist.append((list(trial), e))
            if e < best_e:
                best_a, best_e = trial, e
        log("calibration: variable %d scale %.3f -> L1 %.4e" % (v, best_a[v], best_e))
    return best_a, hist


def ppm_loss(X, T):
    """Training loss of the PPM member (zero correction) on a batch."""
    x = X[:, :48].double().reshape(-1, 8, 6)
    wp = X[:, 48:64].double()
    om2 = X[:, 64:72].double() ** 2
    jump = x[..., 3] - x[..., 2]
    eL = x[..., 2] + wp[:, 0::2] * jump - T.double()
    eR = x[..., 3] - wp[:, 1::2] * jump - T.double()
    return float((om2 * (eL ** 2 + eR ** 2)).sum(1).mean())


def van_leer_loss(X, T):
    """Training loss of the van Leer member on a batch (features are shift/scale
    normalised, so the member's weights follow from the features themselves)."""
    om2 = X[:, 64:72].double() ** 2
    x = X[:, :48].double().reshape(-1, 8, 6)
    a, b, c, d = x[..., 1], x[..., 2], x[..., 3], x[..., 4]

    def vl(dp, dm):
        den = dp + dm
        return torch_where(dp * dm > 0, 2 * dp * dm / torch_where(den != 0, den, 1.0), 0.0)

    import torch

    def torch_where(cond, u, v):
        return torch.where(cond, u, torch.as_tensor(v, dtype=x.dtype, device=x.device))

    jump = c - b
    safe = torch_where(jump != 0, jump, 1.0)
    wl = torch_where(jump != 0, 0.5 * vl(c - b, b - a) / safe, 0.0)
    wr = torch_where(jump != 0, 0.5 * vl(d - c, c - b) / safe, 0.0)
    tau = T.double()
    eL = b + wl * jump - tau
    eR = c - wr * jump - tau
    return float((om2 * (eL ** 2 + eR ** 2)).sum(1).mean())


if __name__ == "__main__":
    main()
