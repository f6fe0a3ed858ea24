"""3-D evaluation of the learned limiter on decaying MHD turbulence.

The network sees one interface at a time through its 6-cell face-frame stencil, so the
parameters trained on 2-D turbulence (learned_turb2d.npy) apply unchanged to x, y and z
interfaces of a 3-D block.  This module runs the 3-D comparison that one GPU holds:

* reference: a plain FP64 run (PPM, HLLD, uct_contact, RK3, CFL 0.4) at --fine^3 to
  --tend, block-averaged to --coarse^3;
* coarse runs from the fine initial state block-averaged to their resolution: the
  learned limiter (FP32 stencil, BF16 MFMA inference), PPM through the same path without
  a network (FP32), plain PPM and PLM in FP64, and plain PPM at twice the coarse
  resolution (block-averaged 2x before the comparison);
* L1 = mean over (rho, v1, v2, v3, p, face Bx, face By, face Bz) of the mean absolute
  difference from the reference; wall time of each coarse run (device synchronised).

Run (repository root, one GPU):
    python -u -m idefix_amd.learned3d --fine 512 --coarse 128 --tend 0.5 --seed 4 --out DIR

A run at 512^3 is the largest a 288 GB device holds with the per-kernel arrays; a 1024^3
reference for a 256^3 / 512^3 pair does not fit (about 45 words of FP64 state, fluxes and
EMF scratch per cell: 1030^3 x 45 x 8 B = 393 GB).

`--train` trains a network on 3-D data instead (the pipeline of idefix_amd.learned with
x, y and z interfaces): plain FP64 runs at --fine^3 of the training seeds, block-averaged
to --coarse^3 at --snapshots times, interface targets from the fine solution, FP32
training on the GPU, a-posteriori calibration of the output scales on the training seeds,
then the evaluation above on the held-out seed with the trained network.  Everything
runs in one process, since a fine state cannot leave the GPU between calls:
    python -u -m idefix_amd.learned3d --train --fine 256 --coarse 64 --seeds 1 2 \
        --seed 4 --tend 0.5 --out DIR
"""
import numpy as np

GAMMA = 5.0 / 3.0


def interior3(g):
    """Interior primitives [8, nz, ny, nx] and face fields bx [nz, ny, nx+1],
    by [nz, ny+1, nx], bz [nz+1, ny, nx] of a 3-D DataBlock, as torch views."""
    b, e = g.beg, g.end
    V = g.Vc[:, b[2]:e[2], b[1]:e[1], b[0]:e[0]]
    bx = g.Vs[0, b[2]:e[2], b[1]:e[1], b[0]:e[0] + 1]
    by = g.Vs[1, b[2]:e[2], b[1]:e[1] + 1, b[0]:e[0]]
    bz = g.Vs[2, b[2]:e[2] + 1, b[1]:e[1], b[0]:e[0]]
    return V, bx, by, bz


def coarse_state3(V, bx, by, bz, f, gamma=GAMMA):
    """Block averages over f^3 cells of a fine periodic 3-D state (torch, float64):
    density, momentum and total energy averaged over the block, each face field averaged
    over the f^2 fine faces of its coarse face (so the coarse faces keep div B = 0);
    returns coarse primitives [8, N, N, N] (cell B from the coarse faces) and the coarse
    faces."""
    import torch
    F = torch.nn.functional
    rho, v, B = V[0], V[1:4], V[5:8]
    E = V[4] / (gamma - 1.0) + 0.5 * rho * (v * v).sum(0) + 0.5 * (B * B).sum(0)
    pool = lambda a: F.avg_pool3d(a[None, None], f)[0, 0]
    rc = pool(rho)
    mc = torch.stack([pool(rho * v[c]) for c in range(3)])
    Ec = pool(E)
    del E
    nz, ny, nx = rho.shape
    Nz, Ny, Nx = nz // f, ny // f, nx // f
    bxc = bx[:, :, ::f].reshape(Nz, f, Ny, f, Nx + 1).mean((1, 3))
    byc = by[:, ::f, :].reshape(Nz, f, Ny + 1, Nx, f).mean((1, 4))
    bzc = bz[::f].reshape(Nz + 1, Ny, f, Nx, f).mean((2, 4))
    Bc = torch.stack([0.5 * (bxc[..., :-1] + bxc[..., 1:]), 0.5 * (byc[:, :-1] + byc[:, 1:]),
                      0.5 * (bzc[:-1] + bzc[1:])])
    vc = mc / rc
    pc = (gamma - 1.0) * (Ec - 0.5 * rc * (vc * vc).sum(0) - 0.5 * (Bc * Bc).sum(0))
    return torch.cat([rc[None], vc, pc[None], Bc]), bxc, byc, bzc


def l1_error3(V, bx, by, bz, Vref, bxref, byref, bzref):
    """Mean over (rho, v1, v2, v3, p, face Bx, face By, face Bz) of the mean absolute
    difference; returns (mean, per-quantity list)."""
    errs = [float((V[n] - Vref[n]).abs().mean()) for n in range(5)]
    errs += [float((a - b).abs().mean()) for a, b in ((bx, bxref), (by, byref), (bz, bzref))]
    return float(np.mean(errs)), errs


def block3(n, recon, precision, nstages=3, cfl=0.4):
    from .datablock import DataBlock
    return DataBlock(3, (n, n, n), comps=3, mhd=True, recon=recon, solver="hlld",
                     emf="uct_contact", precision=precision, bc="periodic", gamma=GAMMA,
                     nstages=nstages, cfl=cfl)


def load_state3(g, V, bx, by, bz):
    """Writes an interior state into a DataBlock (ghosts filled by the first boundary
    step of the run)."""
    import torch
    b, e = g.beg, g.end
    with torch.no_grad():
        g.Vc.zero_()
        g.Vs.zero_()
        g.Vc[:, b[2]:e[2], b[1]:e[1], b[0]:e[0]] = V
        g.Vs[0, b[2]:e[2], b[1]:e[1], b[0]:e[0] + 1] = bx
        g.Vs[1, b[2]:e[2], b[1]:e[1] + 1, b[0]:e[0]] = by
        g.Vs[2, b[2]:e[2] + 1, b[1]:e[1], b[0]:e[0]] = bz


def run_coarse3(state, tend, recon, precision, params=None):
    """Runs a coarse periodic 3-D block from `state` to tend; returns the interior state
    (torch copies), the wall seconds of the run and its cycle count."""
    import time
    import torch
    n = state[0].shape[-1]
    g = block3(n, recon, precision)
    load_state3(g, *state)
    if recon == "mlp":
        g.set_mlp_params(params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.run(tend)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = tuple(a.clone() for a in interior3(g))
    ncyc = g.ncycles
    del g
    torch.cuda.empty_cache()
    return out, wall, ncyc


def face_targets3(V, bx, by, bz, f):
    """Interface values of the fine solution on the coarse faces of a periodic 3-D block,
    per direction in face-frame order [8, N, N, N] (face index = coarse cell right of the
    face): the mean over the coarse face of (V(left) + V(right)) / 2 across the fine faces
    it covers, and the face field itself for the normal component."""
    import torch
    from .learned import FRAME
    nz, ny, nx = V.shape[1:]
    Nz, Ny, Nx = nz // f, ny // f, nx // f
    left = torch.roll(V, 1, dims=3)[..., ::f]
    tx = (0.5 * (left + V[..., ::f])).reshape(8, Nz, f, Ny, f, Nx).mean((2, 4))
    tx[5] = bx[:, :, ::f][..., :Nx].reshape(Nz, f, Ny, f, Nx).mean((1, 3))
    down = torch.roll(V, 1, dims=2)[:, :, ::f, :]
    ty = (0.5 * (down + V[:, :, ::f, :])).reshape(8, Nz, f, Ny, Nx, f).mean((2, 5))
    ty[6] = by[:, ::f, :][:, :Ny].reshape(Nz, f, Ny, Nx, f).mean((1, 4))
    back = torch.roll(V, 1, dims=1)[:, ::f]
    tz = (0.5 * (back + V[:, ::f])).reshape(8, Nz, Ny, f, Nx, f).mean((3, 5))
    tz[7] = bz[::f][:Nz].reshape(Nz, Ny, f, Nx, f).mean((2, 4))
    return tx[FRAME[0]], ty[FRAME[1]], tz[FRAME[2]]


def samples3(Vc, targets, d, wp):
    """Training rows of every coarse interface along d of a periodic 3-D state, as
    idefix_amd.learned.samples: features (48) | PPM weights (16) | loss weights (8),
    normalised targets and stencil scales (float32)."""
    import torch
    from .learned import FRAME
    Vf = Vc[FRAME[d]]
    ax = 3 - d
    S = torch.stack([torch.roll(Vf, 3 - s, dims=ax) for s in range(6)], dim=-1)  # [8,N,N,N,6]
    S = S.permute(1, 2, 3, 0, 4).reshape(-1, 8, 6)
    t = targets.permute(1, 2, 3, 0).reshape(-1, 8)
    c = 0.5 * (S[..., 2] + S[..., 3])
    sc = (S - c[..., None]).abs().amax(-1)
    safe = torch.where(sc > 0, sc, torch.ones_like(sc))
    x = torch.where(sc[..., None] > 0, (S - c[..., None]) / safe[..., None], torch.zeros_like(S))
    tau = torch.where(sc > 0, (t - c) / safe, torch.zeros_like(sc)).clamp(-1.0, 1.0)
    w = wp.permute(1, 2, 3, 0).reshape(-1, 16).to(x.dtype)
    rows = torch.cat([x.reshape(-1, 48), w, torch.ones_like(sc)], dim=1)
    return rows.float(), tau.float()


def ppm_face_weights3(state):
    """PPM weights [16, N, N, N] of the x, y, z interfaces of a coarse periodic 3-D state,
    from the inference kernel with no network loaded."""
    g = block3(state[0].shape[-1], "mlp", "fp32")
    load_state3(g, *state)
    g.set_boundaries()
    g.set_mlp_params(None)
    b, e = g.beg, g.end
    out = [g.mlp_faces(d)[:, b[2]:e[2], b[1]:e[1], b[0]:e[0]].double().clone() for d in range(3)]
    del g
    return out


def fine_run3(n, seed, times, log):
    """Plain FP64 run of the 3-D turbulence at n^3; interior states (torch copies) at t = 0
    and at each of `times`."""
    g = block3(n, "ppm", "fp64")
    from . import problems
    Vc, Vs = problems.turbulence(g, seed=seed)
    g.upload(Vc, Vs)
    g.set_boundaries()
    out = [tuple(x.clone() for x in interior3(g))]
    for t in times:
        g.run(t)
        out.append(tuple(x.clone() for x in interior3(g)))
        log("  fine %d^3 seed %d t=%.3f cycles=%d" % (n, seed, g.t, g.ncycles))
    del g
    return out


def calibrate3(params, calib, tend, log):
    """idefix_amd.learned.calibrate with 3-D coarse runs: per-variable output scales chosen
    by the L1 error of coarse learned runs against the training seeds' references."""
    from .learned import scale_outputs

    def err(alpha):
        p = scale_outputs(params, alpha)
        e = []
        for init, ref in calib:
            st, _, _ = run_coarse3(init, tend, "mlp", "fp32", p)
            e.append(l1_error3(*st, *ref)[0])
        return float(np.mean(e))
    best_a, best_e = None, None
    for c in (0.0, 0.15, 0.3, 0.5, 0.75, 1.0):
        e = err(np.full(8, c))
        log("calibration: common scale %.2f -> L1 %.5e" % (c, e))
        if best_e is None or e < best_e:
            best_a, best_e = np.full(8, c), e
    for v in range(8):
        for m in (0.0, 0.5, 1.5, 2.0):
            trial = best_a.copy()
            trial[v] = best_a[v] * m if best_a[v] > 0 else 0.3 * m
            e = err(trial)
            if e < best_e:
                best_a, best_e = trial, e
        log("calibration: variable %d scale %.3f -> L1 %.5e" % (v, best_a[v], best_e))
    return best_a


def train3(a, log):
    """Data, training and calibration on 3-D turbulence; returns the calibrated parameters
    and a report of the training."""
    import torch
    from .learned import Trainer, init_params, pack, unpack, scale_outputs
    f = a.fine // a.coarse
    times = [a.tend * (q + 1) / a.snapshots for q in range(a.snapshots)]
    Xs, Ts, calib = [], [], []
    for seed in a.seeds:
        snaps = fine_run3(a.fine, seed, times, log)
        calib.append((coarse_state3(*snaps[0], f), coarse_state3(*snaps[-1], f)))
        for st in snaps:
            cs = coarse_state3(*st, f)
            tg = face_targets3(*st, f)
            wps = ppm_face_weights3(cs)
            for d in range(3):
                X, T = samples3(cs[0], tg[d], d, wps[d])
                Xs.append(X)
                Ts.append(T)
        del snaps
        torch.cuda.empty_cache()
    X, T = torch.cat(Xs), torch.cat(Ts)
    del Xs, Ts
    log("samples: %d" % X.shape[0])
    p0 = unpack(init_params(0))
    p0["W3"][:] = 0.0  # starts as PPM
    tr = Trainer(a.batch, pack(p0))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    hist = []
    for step in range(a.steps):
        idx = torch.randint(0, X.shape[0], (a.batch,), device="cuda", generator=gen)
        lr = a.lr * (0.5 * (1.0 + np.cos(np.pi * step / a.steps)))
        tr.step(X[idx].contiguous(), T[idx].contiguous(), max(lr, 1e-5))
        if step % 1000 == 0 or step == a.steps - 1:
            hist.append((step, tr.loss()))
            log("step %d loss %.5f" % hist[-1])
    params = tr.params()
    del X, T
    torch.cuda.empty_cache()
    alpha = calibrate3(params, calib, a.tend, log)
    return scale_outputs(params, alpha), {"loss_history": hist, "output_scales":
                                           [float(v) for v in alpha], "train_seeds": a.seeds}


def main(argv=None):
    import argparse
    import json
    import os
    import time
    import torch
    from . import problems
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=512)
    ap.add_argument("--coarse", type=int, default=128)
    ap.add_argument("--tend", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--params", default=os.path.join(os.path.dirname(__file__),
                                                     "learned_turb2d.npy"))
    ap.add_argument("--log-every", type=int, default=200,
                    help="progress line every this many cycles of the fine run")
    ap.add_argument("--out", default="learned_out/turb3d")
    ap.add_argument("--train", action="store_true",
                    help="train and calibrate a network on 3-D data first (see above)")
    ap.add_argument("--seeds", type=int, nargs="+", default=[1, 2])
    ap.add_argument("--snapshots", type=int, default=5)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    t_start = time.perf_counter()
    log = lambda s: print("[%.0fs] %s" % (time.perf_counter() - t_start, s), flush=True)
    train_rep = None
    if a.train:
        assert a.seed not in a.seeds, "the evaluation seed must be held out"
        params, train_rep = train3(a, log)
        np.save(os.path.join(a.out, "params3d.npy"), params)
    else:
        params = np.load(a.params)
    N, f = a.coarse, a.fine // a.coarse
    assert N * f == a.fine and f >= 2
    # ---- reference --------------------------------------------------------------------
    g = block3(a.fine, "ppm", "fp64")
    log("initial state %d^3, seed %d" % (a.fine, a.seed))
    Vc, Vs = problems.turbulence(g, seed=a.seed)
    g.upload(Vc, Vs)
    del Vc, Vs
    g.set_boundaries()
    st0 = tuple(x.clone() for x in interior3(g))
    init = {N: coarse_state3(*st0, f), 2 * N: coarse_state3(*st0, f // 2)}
    del st0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while g.t < a.tend:
        g.run(a.tend, g.ncycles + a.log_every)
        log("fine %d^3 t=%.4f cycles=%d" % (a.fine, g.t, g.ncycles))
    torch.cuda.synchronize()
    fine_wall, fine_cyc = time.perf_counter() - t0, g.ncycles
    ref = coarse_state3(*interior3(g), f)
    del g
    torch.cuda.empty_cache()
    rep = {"fine": a.fine, "coarse": N, "tend": a.tend, "seed": a.seed,
           "params": "trained here (3-D)" if a.train else os.path.basename(a.params),
           "training": train_rep,
           "fine_run": {"wall_s": fine_wall, "cycles": fine_cyc}, "runs": {}}
    cases = [("learned_fp32_%d" % N, N, "mlp", "fp32", params),
             ("ppm_via_mlp_fp32_%d" % N, N, "mlp", "fp32", None),
             ("plain_ppm_fp64_%d" % N, N, "ppm", "fp64", None),
             ("plain_plm_fp64_%d" % N, N, "plm", "fp64", None),
             ("plain_ppm_fp64_%d" % (2 * N), 2 * N, "ppm", "fp64", None)]
    if 2 * N < a.fine:
        cases.insert(1, ("learned_fp32_%d" % (2 * N), 2 * N, "mlp", "fp32", params))
    for name, n, recon, prec, p in cases:
        st, wall, cyc = run_coarse3(init[n], a.tend, recon, prec, p)
        if n != N:
            st = coarse_state3(*st, n // N)
        l1, per = l1_error3(*st, *ref)
        rep["runs"][name] = {"n": n, "recon": recon, "precision": prec, "L1": l1,
                             "L1_per_quantity": per, "wall_s": wall, "cycles": cyc}
        log("%s: L1 %.5e wall %.2fs cycles %d" % (name, l1, wall, cyc))
    with open(os.path.join(a.out, "report.json"), "w") as fh:
        json.dump(rep, fh, indent=1)
    log("done")


if __name__ == "__main__":
    main()
