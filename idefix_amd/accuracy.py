"""FP32-stencil mode against the plain FP64 method: runs one workload in both modes to the
same time and reports the L1 difference (per variable, mean over the interior cells of
|q_fp32 - q_fp64|, averaged over rho, velocity, pressure and magnetic components, face
fields for the components held on faces) and the throughput of each run.

    python -m idefix_amd.accuracy --config ot --n 512 --tend 0.5
    python -m idefix_amd.accuracy --config turb3d --n 256 --tend 1.0
    python -m idefix_amd.accuracy --config kh --n 128 --tend 2.0

Runs too long for one process pair (512^3 to t = 1) run one precision at a time and keep
full-resolution planes of the final state (`--precision fp64 --save-planes a.npz`, then
fp32 into b.npz); `--compare a.npz b.npz` gives the L1 difference over those planes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

CONFIGS = {
    # 2-D Orszag-Tang: PLM, HLLD, UCT-HLLD, RK2, CFL 0.6
    "ot": dict(dim=2, comps=2, recon="plm", emf="uct_hlld", nstages=2, cfl=0.6,
               shape=lambda n: (n, n), xend=(1, 1, 1), ic="orszag_tang"),
    # 3-D decaying turbulence: PPM, HLLD, CT, RK3
    "turb3d": dict(dim=3, comps=3, recon="ppm", emf="uct_contact", nstages=3, cfl=0.4,
                   shape=lambda n: (n, n, n), xend=(1, 1, 1), ic="turbulence"),
    # 2-D decaying turbulence (three components): PPM, HLLD, CT, RK3
    "turb2d": dict(dim=2, comps=3, recon="ppm", emf="uct_contact", nstages=3, cfl=0.4,
                   shape=lambda n: (n, n), xend=(1, 1, 1), ic="turbulence"),
    # 3-D MHD Kelvin-Helmholtz on [0,1] x [0,1] x [0,1/2]: PPM, HLLD, CT, RK2
    "kh": dict(dim=3, comps=3, recon="ppm", emf="uct_contact", nstages=2, cfl=0.4,
               shape=lambda n: (2 * n, 2 * n, n), xend=(1, 1, 0.5), ic="kelvin_helmholtz"),
}


def run(config, n, tend, precision, recon=None, params=None, perturb=0.0):
    """perturb > 0 multiplies the initial cell state by 1 + perturb x (uniform in [-1, 1],
    fixed seed): a control run that shows how a perturbation of that size grows."""
    import torch
    from .datablock import DataBlock
    from . import problems
    c = CONFIGS[config]
    g = DataBlock(c["dim"], c["shape"](n), comps=c["comps"], mhd=True,
                  recon=recon or c["recon"], solver="hlld", emf=c["emf"], precision=precision,
                  nstages=c["nstages"], cfl=c["cfl"], xend=c["xend"])
    Vc, Vs = getattr(problems, c["ic"])(g)
    if perturb > 0:
        rng = np.random.default_rng(12345)
        Vc = Vc * (1.0 + perturb * rng.uniform(-1.0, 1.0, Vc.shape))
    g.upload(Vc, Vs)
    if params is not None:
        g.set_mlp_params(params)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while g.t < tend:  # the main loop in chunks of cycles, with a progress line each
        g.run(tend, maxcycles=g.ncycles + 500)
        print("%s %s: t = %.4g after %d cycles" % (config, precision, g.t, g.ncycles),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = int(np.prod(c["shape"](n)))
    return g, {"wall_s": wall, "cycles": g.ncycles, "t": g.t,
               "zone_updates_per_s": cells * g.ncycles / wall}


def l1(ga, gb):
    """Mean over variables of the mean absolute difference over interior cells/faces."""
    b, e = ga.beg, ga.end
    sl = (slice(b[2], e[2]), slice(b[1], e[1]), slice(b[0], e[0]))
    errs = []
    dim = ga.dim
    for n in range(ga.nvar):
        if n >= ga.comps + 2 and n - (ga.comps + 2) < dim:
            continue  # face-held components compared on the faces below
        errs.append(float((ga.Vc[n][sl] - gb.Vc[n][sl]).abs().mean()))
    for d in range(dim):
        fs = list(sl)
        fs[2 - d] = slice(b[d], e[d] + 1)
        errs.append(float((ga.Vs[d][tuple(fs)] - gb.Vs[d][tuple(fs)]).abs().mean()))
    return float(np.mean(errs)), errs


def sample_planes(g, nplanes=2):
    """Full-resolution planes of the final interior cell state of a 3-D block (all
    variables, the cell field being the face average): k at (q + 1/2) / nplanes of the
    interior, q = 0 .. nplanes - 1.  Returns (k indices, array [nplanes, nvar, ny, nx])."""
    b, e = g.beg, g.end
    nz = e[2] - b[2]
    ks = [b[2] + int((q + 0.5) * nz / nplanes) for q in range(nplanes)]
    planes = np.stack([g.Vc[:, k, b[1]:e[1], b[0]:e[0]].cpu().numpy() for k in ks])
    return ks, planes


def compare_planes(fa, fb):
    """L1 difference between two saved plane files: mean over variables of the mean
    absolute difference over the sampled cells."""
    a, b = np.load(fa), np.load(fb)
    assert np.array_equal(a["k"], b["k"]) and a["planes"].shape == b["planes"].shape
    assert abs(float(a["t"]) - float(b["t"])) < 1e-12
    pa, pb = a["planes"], b["planes"]
    per = [float(np.abs(pa[:, n] - pb[:, n]).mean()) for n in range(pa.shape[1])]
    return {"L1": float(np.mean(per)), "L1_per_variable": per, "t": float(a["t"]),
            "planes_k": [int(k) for k in a["k"]],
            "a": {k: float(a[k]) for k in ("cycles", "wall_s", "zone_updates_per_s", "divB")},
            "b": {k: float(b[k]) for k in ("cycles", "wall_s", "zone_updates_per_s", "divB")}}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=sorted(CONFIGS), default="ot")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tend", type=float, default=0.5)
    ap.add_argument("--out", default=None)
    ap.add_argument("--precision", choices=("both", "fp64", "fp32"), default="both",
                    help="one precision only: with --save-planes, keep planes of the result")
    ap.add_argument("--save-planes", default=None,
                    help="npz file for full-resolution planes of the final state (3-D)")
    ap.add_argument("--compare", nargs=2, default=None, metavar=("A", "B"),
                    help="L1 difference between two --save-planes files (no GPU)")
    ap.add_argument("--perturb", type=float, default=0.0,
                    help="relative random perturbation of the initial cell state (control)")
    a = ap.parse_args(argv)
    if a.compare:
        rep = compare_planes(*a.compare)
        print(json.dumps(rep))
        return
    if a.precision != "both":
        g, r = run(a.config, a.n, a.tend, a.precision, perturb=a.perturb)
        r["divB"] = g.check_divb()
        print(json.dumps({"config": a.config, "n": a.n, "tend": a.tend,
                          "precision": a.precision, "perturb": a.perturb, **r}), flush=True)
        if a.save_planes:
            ks, planes = sample_planes(g)
            os.makedirs(os.path.dirname(a.save_planes) or ".", exist_ok=True)
            np.savez(a.save_planes, k=np.array(ks), planes=planes, t=r["t"],
                     cycles=r["cycles"], wall_s=r["wall_s"],
                     zone_updates_per_s=r["zone_updates_per_s"], divB=r["divB"])
        return
    g64, r64 = run(a.config, a.n, a.tend, "fp64")
    g32, r32 = run(a.config, a.n, a.tend, "fp32")
    err, per = l1(g32, g64)
    rep = {"config": a.config, "n": a.n, "shape": CONFIGS[a.config]["shape"](a.n),
           "tend": a.tend, "fp64": r64, "fp32": r32, "L1_fp32_vs_fp64": err,
           "L1_per_variable": per, "divB_fp64": g64.check_divb(), "divB_fp32": g32.check_divb()}
    print(json.dumps(rep))
    if a.out:
        os.makedirs(os.path.dirname(a.out) or ".", exist_ok=True)
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)


if __name__ == "__main__":
    main()
