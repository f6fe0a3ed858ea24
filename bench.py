#!/usr/bin/env python3
"""Throughput of the per-stage finite-volume update on one MI355X.

Workload: 2-D ideal-MHD Orszag-Tang vortex at 2048^2 (the largest single-GPU size of
that configuration): PLM van Leer, HLLD, constrained transport with UCT-HLLD corner
EMFs, RK2, CFL 0.6, periodic boundaries, FP64.  A step is one full RK2 cycle of the
whole grid (TimeIntegrator::Cycle); a zone-update is one cell advanced by one cycle.
Inputs are synthetic (the analytic Orszag-Tang initial state), resident in HBM before
the timed region starts.

Prints ONE JSON line.  `value` is zone-updates/s of the plain FP64 method; the FP32
stencil mode is reported beside it.  `roofline` is computed for the kernel with the
largest share of the cycle, timed with HIP events on the stream it runs on;
`cpu_baseline` times the CPU oracle (an OpenMP restatement of the same algorithm) on a
bounded sample of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nx", type=int, default=2048)
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--kernel-reps", type=int, default=10)
    p.add_argument("--no-headline", action="store_true",
                   help="skip the reference's own 3-D benchmark configuration")
    p.add_argument("--headline-only", action="store_true",
                   help="only time the reference's own 3-D benchmark configuration")
    p.add_argument("--matrix", action="store_true",
                   help="only time the Orszag-Tang workload across solvers, reconstructions, "
                        "EMF averages and tracers (one JSON line per configuration)")
    p.add_argument("--configs", action="store_true",
                   help="only time every configured workload at its full size (one JSON "
                        "line per workload and precision)")
    p.add_argument("--only", default=None,
                   help="with --configs: comma-separated workload names to run")
    p.add_argument("--copy-probe", action="store_true",
                   help="only run a plain 1 GiB device copy 10 times (counter calibration)")
    return p.parse_args()


def copy_probe():
    """A plain device-to-device copy of known size: reads 1 GiB and writes 1 GiB per
    launch, the yardstick for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read_per_launch": n * 8, "launches": 10}))


def make_block(nx, precision, solver="hlld", recon="plm", emf="uct_hlld", ntracer=0):
    from idefix_amd.datablock import DataBlock
    from idefix_amd import problems
    g = DataBlock(2, (nx, nx), mhd=True, recon=recon, solver=solver, emf=emf,
                  precision=precision, bc="periodic", gamma=5.0 / 3.0, nstages=2, cfl=0.6,
                  first_dt=1e-4, ntracer=ntracer)
    Vc, Vs = problems.orszag_tang(g)
    if ntracer:
        x = g.x[0][None, None, :]
        for t in range(ntracer):
            Vc[6 + t] = (x > 0.5) + 0.0 * Vc[0]
    g.upload(Vc, Vs)
    return g


# Orszag-Tang configurations of the reference's test inputs (test/MHD/OrszagTang/*.ini
# with reconstructions 2-4) and of its OrszagTang3D input (2 tracers), timed by --matrix
MATRIX = [("hlld", "plm", "uct_hlld", 0), ("hlld", "plm", "uct_contact", 0),
          ("hlld", "plm", "uct0", 0), ("hlld", "plm", "arithmetic", 0),
          ("hlld", "plm", "uct_hll", 0), ("hlld", "ppm", "uct_contact", 0),
          ("hlld", "limo3", "uct_contact", 0), ("hll", "plm", "uct_contact", 0),
          ("hll", "plm", "uct_hll", 0), ("tvdlf", "plm", "uct_contact", 0),
          ("roe", "plm", "uct_contact", 0), ("hlld", "plm", "uct_contact", 2)]


def matrix(nx, steps, warmup):
    import torch
    for solver, recon, emf, ntr in MATRIX:
        g = make_block(nx, "fp64", solver, recon, emf, ntr)
        el = time_cycles(g, steps, warmup)
        print(json.dumps({"workload": "2D Orszag-Tang %dx%d, RK2, CFL 0.6, FP64" % (nx, nx),
                          "solver": solver, "recon": recon, "emf": emf, "tracers": ntr,
                          "zone_updates_per_s": nx * nx * steps / el,
                          "ms_per_step": 1e3 * el / steps}), flush=True)
        del g
        torch.cuda.empty_cache()


# The configured workloads at their full sizes.  S = state words per cell (conserved
# hydro words + face-held B components), k = state-sized HBM transfers per cycle of the
# two-register RK scheme (RK2 5, RK3 7): B_zu = k * S * 8 bytes is the compulsory HBM
# traffic of one zone-update.
WORKLOADS = [
    dict(name="sod1000", label="1-D Sod 1000, PLM, HLLC, RK2, CFL 0.8, outflow",
         dim=1, shape=(1000,), mhd=False, recon="plm", solver="hllc", nstages=2, cfl=0.8,
         gamma=1.4, bc="outflow", ic="sod", S=3, k=5, modes=("fp64",)),
    dict(name="ot512", label="2-D Orszag-Tang 512^2, PLM, HLLD, UCT-HLLD, RK2, CFL 0.6",
         dim=2, shape=(512, 512), comps=2, recon="plm", emf="uct_hlld", nstages=2, cfl=0.6,
         ic="orszag_tang", S=6, k=5, modes=("fp64", "fp32")),
    dict(name="ot2048", label="2-D Orszag-Tang 2048^2, PLM, HLLD, UCT-HLLD, RK2, CFL 0.6",
         dim=2, shape=(2048, 2048), comps=2, recon="plm", emf="uct_hlld", nstages=2,
         cfl=0.6, ic="orszag_tang", S=6, k=5, modes=("fp64", "fp32")),
    dict(name="turb3d256", label="3-D decaying turbulence 256^3, PPM, HLLD, CT, RK3, CFL 0.4",
         dim=3, shape=(256, 256, 256), comps=3, recon="ppm", emf="uct_contact", nstages=3,
         cfl=0.4, ic="turbulence", S=8, k=7, modes=("fp64", "fp32", "fp32+learned")),
    dict(name="turb3d512", label="3-D decaying turbulence 512^3, PPM, HLLD, CT, RK3, CFL 0.4",
         dim=3, shape=(512, 512, 512), comps=3, recon="ppm", emf="uct_contact", nstages=3,
         cfl=0.4, ic="turbulence", S=8, k=7, modes=("fp64", "fp32")),
    dict(name="turb2d256", label="2-D decaying turbulence 256^2 (3 components), PPM, HLLD, "
         "CT, RK3, CFL 0.4", dim=2, shape=(256, 256), comps=3, recon="ppm",
         emf="uct_contact", nstages=3, cfl=0.4, ic="turbulence", S=7, k=7,
         modes=("fp64", "fp32", "fp32+learned")),
    dict(name="turb2d2048", label="2-D decaying turbulence 2048^2 (3 components), PPM, HLLD, "
         "CT, RK3, CFL 0.4", dim=2, shape=(2048, 2048), comps=3, recon="ppm",
         emf="uct_contact", nstages=3, cfl=0.4, ic="turbulence", S=7, k=7,
         modes=("fp64", "fp32", "fp32+learned")),
    dict(name="kh512", label="3-D MHD Kelvin-Helmholtz 512x512x256 on [0,1]^2 x [0,1/2], PPM, "
         "HLLD, CT, RK2, CFL 0.4", dim=3, shape=(512, 512, 256), xend=(1.0, 1.0, 0.5),
         comps=3, recon="ppm", emf="uct_contact", nstages=2, cfl=0.4, ic="kelvin_helmholtz",
         S=8, k=5, modes=("fp64", "fp32", "fp32+learned")),
]


def workloads(only, steps, warmup):
    """Throughput of each configured workload at its full size: `warmup` untimed cycles,
    then `steps` timed cycles from the workload's initial state.  The learned mode uses
    the network trained on the 2-D turbulence (idefix_amd/learned_turb2d.npy)."""
    import numpy as np
    import torch
    from idefix_amd.datablock import DataBlock
    from idefix_amd import problems
    names = set(only.split(",")) if only else None
    for w in WORKLOADS:
        if names is not None and w["name"] not in names:
            continue
        cells = int(np.prod(w["shape"]))
        print("setting up %s" % w["name"], file=sys.stderr, flush=True)
        ic = None
        for mode in w["modes"]:
            precision = "fp32" if mode.startswith("fp32") else "fp64"
            recon = "mlp" if mode.endswith("learned") else w["recon"]
            kw = dict(mhd=w.get("mhd", True), recon=recon, solver=w["solver"] if "solver" in w
                      else "hlld", precision=precision, bc=w.get("bc", "periodic"),
                      gamma=w.get("gamma", 5.0 / 3.0), nstages=w["nstages"], cfl=w["cfl"],
                      first_dt=1e-4)
            if "comps" in w:
                kw["comps"] = w["comps"]
            if "emf" in w:
                kw["emf"] = w["emf"]
            if "xend" in w:
                kw["xend"] = w["xend"]
            t0 = time.perf_counter()
            g = DataBlock(w["dim"], w["shape"], **kw)
            if ic is None:
                ic = getattr(problems, w["ic"])(g)
            g.upload(*ic)
            if recon == "mlp":
                g.set_mlp_params(np.load(os.path.join(ROOT, "idefix_amd", "learned_turb2d.npy")))
            setup_s = time.perf_counter() - t0
            el = time_cycles(g, steps, warmup)
            zu = cells * steps / el
            bzu = w["k"] * w["S"] * 8
            print(json.dumps({"workload": w["name"], "label": w["label"], "mode": mode,
                              "cells": cells, "steps": steps, "warmup": warmup,
                              "zone_updates_per_s": zu, "ms_per_step": 1e3 * el / steps,
                              "bytes_per_zone_update": bzu,
                              "cycle_hbm_frac": bzu * zu / (HBM_PEAK_GBS * 1e9),
                              "setup_s": round(setup_s, 1)}), flush=True)
            del g
            torch.cuda.empty_cache()
        del ic


def time_cycles(g, steps, warmup):
    import torch
    for _ in range(warmup):
        g.cycle()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.cycle()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_times(g, cycles):
    """Average duration per launch of each region of the cycle, from the library's region
    profiler: HIP event pairs recorded on the stream the kernels run on."""
    g.set_profiling(True)
    for _ in range(cycles):
        g.cycle()
    rt = g.region_times()
    g.set_profiling(False)
    return {k: ms / n for k, (ms, n) in rt.items()}, {k: ms / cycles for k, (ms, n) in rt.items()}


def algorithmic_bytes(nx):
    """Compulsory HBM bytes per launch of the single-kernel regions of a 2-D MHD RK2 stage
    (each input read once, each output written once, 8-byte words per interior cell),
    averaged over the two stages of the cycle.
      sweep_x   stage 0: Vc 6 + BX1s 1 -> Uc 4 + Uc0 4 + InvDt 1 + face data 5 = 21
                stage 1: Vc 6 + BX1s 1 -> Uc 4 + face data 5                  = 16
      fused2d_y stage 0: Vc 6 + Vs 2 + Uc 4 + x-face data 5 + InvDt 1
                         -> Vc 6 + Vs 2 + Vs0 2 + InvDt 1                    = 29
                stage 1: Vc 6 + Vs 2 + Uc 4 + Uc0 4 + Vs0 2 + x-face data 5
                         -> Vc 6 + Vs 2 + Uc 6                                = 37
    """
    n = nx * nx
    w = 8
    return {"sweep_x": n * w * (21 + 16) / 2, "fused2d_y": n * w * (29 + 37) / 2}


# HBM traffic per launch of each kernel, measured with rocprofv3 --pmc FETCH_SIZE and
# --pmc WRITE_SIZE on this workload and scaled by the plain-copy calibration (reads x2.0,
# writes x1.0); the summary and the commands are in
# profiles/ot2048_hlld_uct_hlld_fp64_kernels.md.  None where not measured.
MEASURED_TRAFFIC = {2048: {"fused2d_y": (859.6 + 473.9) * 1e6, "sweep_x": (276.9 + 393.0) * 1e6}}


# The reference's published benchmark (doc/source/performances.rst:5-10, bench.json): 3-D
# Orszag-Tang, 128^3 cells per GPU, PLM, HLLD, uct_contact, RK2; cell-updates/s.
REFERENCE_PUBLISHED = {"NVIDIA H100": 3.079643e8, "NVIDIA A100": 2.044728e8,
                       "NVIDIA V100": 1.193720e8, "AMD MI250X (one GCD)": 1.436580e8}


def reference_headline(steps, warmup, n=128):
    """The reference's own benchmark configuration timed on this GPU (FP64)."""
    import torch
    from idefix_amd.datablock import DataBlock
    from idefix_amd import problems
    g = DataBlock(3, (n, n, n), mhd=True, recon="plm", solver="hlld", emf="uct_contact",
                  precision="fp64", bc="periodic", gamma=5.0 / 3.0, nstages=2, cfl=0.9,
                  first_dt=1e-4)
    Vc, Vs = problems.orszag_tang_3d(g)
    g.upload(Vc, Vs)
    el = time_cycles(g, steps, warmup)
    zu = n ** 3 * steps / el
    kt, _ = kernel_times(g, 3)
    del g
    torch.cuda.empty_cache()
    return {"workload": "3D Orszag-Tang %d^3, PLM, HLLD, CT uct_contact, RK2, CFL 0.9, "
                        "periodic, FP64 (the reference's published benchmark)" % n,
            "value": zu, "unit": "cell-updates/s", "ms_per_step": 1e3 * el / steps,
            "kernel_ms": kt,
            "published": REFERENCE_PUBLISHED,
            "ratio_vs_published": {k: zu / v for k, v in REFERENCE_PUBLISHED.items()}}


def cpu_baseline(nx, seconds):
    """CPU oracle on the host cores: Orszag-Tang at the same resolution, as many cycles as
    fit in about `seconds`."""
    import numpy as np
    from oracle import oracle as O
    from idefix_amd import problems
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or threads))
    O.set_threads(threads)
    cfg = O.make_cfg(2, mhd=True, recon="plm", solver="hlld", emf="uct_hlld", nstages=2,
                     nx=(nx, nx), bc="periodic", gamma=5.0 / 3.0, cfl=0.6)
    o = O.Oracle(cfg)
    Vc, Vs = problems.orszag_tang(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    o.cycle()  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        o.cycle()
        n += 1
        el = time.perf_counter() - t0
        if el > seconds or n >= 1000:
            break
    rate = nx * nx * n / el
    return {"value": rate, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": "Orszag-Tang %dx%d, PLM/HLLD/UCT-HLLD/RK2 FP64, %d cycles after 1 "
                      "warm-up cycle (%.1f s)" % (nx, nx, n, el)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box "
              "here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    import torch
    if args.headline_only:
        print(json.dumps(reference_headline(args.steps, args.warmup)))
        return
    if args.matrix:
        matrix(args.nx, args.steps, args.warmup)
        return
    if args.configs:
        workloads(args.only, args.steps, args.warmup)
        return
    nx = args.nx
    g = make_block(nx, "fp64")
    el = time_cycles(g, args.steps, args.warmup)
    zu = nx * nx * args.steps / el
    ms = 1e3 * el / args.steps
    kt, per_cycle = kernel_times(g, args.kernel_reps)
    ab = algorithmic_bytes(nx)
    # the single-kernel region with the largest share of the cycle
    dom = max((k for k in per_cycle if k in ab), key=lambda k: per_cycle[k])
    achieved = ab[dom] / (kt[dom] * 1e-3) / 1e9
    traffic = MEASURED_TRAFFIC.get(nx, {}).get(dom)
    del g
    torch.cuda.empty_cache()
    fp32 = None
    if not args.no_fp32:
        g32 = make_block(nx, "fp32")
        el32 = time_cycles(g32, args.steps, args.warmup)
        fp32 = nx * nx * args.steps / el32
        del g32
        torch.cuda.empty_cache()
    headline = None if args.no_headline else reference_headline(args.steps, args.warmup)
    cpu = None if args.no_cpu_baseline else cpu_baseline(nx, args.cpu_seconds)
    cycle_bytes = 5 * 6 * 8  # RK2: 5 state transfers of S = 6 words per zone-update
    line = {
        "metric": "zone-updates/s on 1 MI355X: plain FP64 vs FP32+learned at equal L1 error",
        "value": zu,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D Orszag-Tang MHD %dx%d, PLM van Leer, HLLD, CT UCT-HLLD, "
                               "RK2, CFL 0.6, periodic" % (nx, nx),
                   "cells": nx * nx, "step": "one RK2 cycle",
                   "parallelism": "1 GPU"},
        "fp32_stencil_value": fp32,
        "kernel_ms": kt,
        "kernel_ms_per_cycle": per_cycle,
        "cycle_hbm_frac": cycle_bytes * zu / (HBM_PEAK_GBS * 1e9),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes": ab[dom], "ms_per_launch": kt[dom],
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "reference_benchmark": headline,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
