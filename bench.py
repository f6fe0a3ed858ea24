#!/usr/bin/env python3
"""Throughput of the per-stage finite-volume update on one MI355X.

Workload (the configuration the metric's target is quoted on, and the largest
single-GPU configuration): 3-D decaying ideal-MHD turbulence at 512^3 on the periodic
unit cube, PPM, HLLD, constrained transport (uct_contact), RK3, CFL 0.4, gamma 5/3, the
seeded solenoidal initial state (rms v 0.5, rms dB 0.1, Bx 0.1).  A step is one full RK3
cycle of the whole grid (TimeIntegrator::Cycle); a zone-update is one cell advanced by one
cycle.  Inputs are synthetic, resident in HBM before the timed region starts.

Prints ONE JSON line.  `value` is zone-updates/s of the plain FP64 method at 512^3; beside
it the FP32 stencil mode at 512^3 and the FP32 + learned-limiter mode at 256^3 (the two
legs of the metric), with the equal-L1 comparison of the committed accuracy study.
`roofline` is computed for the kernel with the largest share of the cycle: the whole
stage's compulsory HBM bytes (448 B per zone-update per RK3 cycle, one third per stage)
over its average launch duration from HIP events on the stream it runs on; `traffic` is
that kernel's HBM bytes per launch measured in this run by two rocprofv3 --pmc child runs
(FETCH_SIZE, WRITE_SIZE) when rocprofv3 is available.  `cpu_baseline` times the CPU oracle
(an OpenMP restatement of the same algorithm) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor peak
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, vendor peak
# one interface through the 48 -> 64 -> 64 -> 16 network: 8192 multiply-adds
MLP_FLOP_PER_FACE = 2 * (48 * 64 + 64 * 64 + 64 * 16)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nx", type=int, default=2048,
                   help="Orszag-Tang size of --matrix")
    p.add_argument("--n", type=int, default=512, help="3-D turbulence size of the headline")
    p.add_argument("--no-learned", action="store_true")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 --pmc child runs that measure HBM traffic")
    p.add_argument("--cpu-n", type=int, default=128,
                   help="size of the CPU oracle sample of the headline workload")
    p.add_argument("--traffic-child", default=None, help=argparse.SUPPRESS)
    p.add_argument("--ic-cache", default=None, help=argparse.SUPPRESS)
    p.add_argument("--no-fp32", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
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


def copy_bandwidth():
    """Measured device copy bandwidth (GB/s): a 1 GiB float64 copy (1 GiB read + 1 GiB
    written) 10 times after one warm-up, timed with events on torch's stream, where the
    copies run.  The STREAM-like practical ceiling beside the vendor peak."""
    import torch
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 10 * 2 * n * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    torch.cuda.empty_cache()
    return gbs


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
    the network trained on data of the workload's dimension (idefix_amd/learned_turb2d.npy
    for 1-D and 2-D, idefix_amd/learned_turb3d.npy for 3-D)."""
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
                net = "learned_turb3d.npy" if w["dim"] == 3 else "learned_turb2d.npy"
                g.set_mlp_params(np.load(os.path.join(ROOT, "idefix_amd", net)))
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


# The headline workload: 3-D decaying turbulence (the metric's configuration), its
# compulsory HBM bytes per zone-update (7 state-sized transfers per RK3 cycle of S = 8
# words: rho, 3 momenta, E, 3 face fields) and per stage launch (one third of a cycle).
B_ZU_TURB3D = 7 * 8 * 8
STAGES_TURB3D = 3
# regions of the per-kernel 3-D schedule that are one kernel each, and the kernel names the
# profiler reports for them (FP64 instantiations)
# region -> (kernel name prefix, further substring of the demangled name): the y and z
# sweeps are one template, told apart by its direction argument
SINGLE_KERNEL_REGIONS = {"sweep_x": ("k_sweep_x<", ""),
                         "sweep_y": ("k_sweep_march<", ", double, 1, "),
                         "sweep_z": ("k_sweep_march<", ", double, 2, "),
                         "corner_emf": ("k_corner_emf<", ""), "evolve_b": ("k_evolve_b<", ""),
                         "cons_to_prim": ("k_cons_to_prim<", ""), "ct_fused": ("k_ct3d<", "")}
# FP64 VALU issue rate of the device: a wave64 FP64 instruction occupies a SIMD (16 lanes)
# for 4 cycles; 256 CUs x 4 SIMDs at 2.4 GHz
VALU_SIMDS, VALU_CLOCK_HZ, VALU_CYCLES_PER_INST = 1024, 2.4e9, 4
# plain-copy calibration of the traffic counters (1 GiB read + 1 GiB written per launch,
# bench.py --copy-probe): FETCH_SIZE reads back half the bytes, WRITE_SIZE all of them
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0

# The equal-L1 comparison of the two legs: the learned-limiter accuracy studies on one
# MI355X.  2-D (python -m idefix_amd.learned --fine 2048 --factor 8 ...): decaying
# turbulence, held-out seed, L1 at t = 1 against the 2048^2 solution block-averaged to
# 256^2.  3-D (python -m idefix_amd.learned3d --fine 512 --coarse 128 --tend 0.5 --seed 4):
# the 2-D network on 3-D turbulence, L1 at t = 0.5 against 512^3 block-averaged to 128^3.
# A network trained on 3-D data (256^3 -> 64^3, python -m idefix_amd.learned3d --train ...)
# gives the same picture.  In none does the learned run reach the error of the plain run
# at twice the resolution, so no equal-L1 wall-time ratio exists.
EQUAL_L1 = {"study_2d": {"case": "2-D decaying MHD turbulence, seeds 1-3 trained, seed 4 "
                                 "held out, L1 at t = 1 vs block-averaged 2048^2",
                         "L1": {"learned_fp32_256": 2.1035e-3, "plain_ppm_fp64_256": 2.1046e-3,
                                "plain_ppm_fp64_512": 4.1035e-4},
                         "wall_s": {"learned_fp32_256": 0.58, "plain_ppm_fp64_256": 0.48,
                                    "plain_ppm_fp64_512": 1.29}},
            "study_3d": {"case": "3-D decaying MHD turbulence, seed 4, 2-D network, L1 at "
                                 "t = 0.5 vs 512^3 block-averaged to 128^3",
                         "L1": {"learned_fp32_128": 6.5057e-3, "plain_ppm_fp64_128": 6.5146e-3,
                                "learned_fp32_256": 1.7048e-3, "plain_ppm_fp64_256": 1.7099e-3},
                         "wall_s": {"learned_fp32_128": 5.50, "plain_ppm_fp64_128": 2.96,
                                    "learned_fp32_256": 71.58, "plain_ppm_fp64_256": 35.96},
                         "note": "the learned 128^3 run was the first and includes module load"},
            "study_3d_trained": {"case": "3-D decaying MHD turbulence, network trained on 3-D "
                                         "data (seeds 1-2, 256^3 -> 64^3), seed 4 held out, L1 "
                                         "at t = 0.5 vs 256^3 block-averaged to 64^3",
                                 "L1": {"learned_fp32_64": 1.27584e-2,
                                        "plain_ppm_fp64_64": 1.27860e-2,
                                        "learned_fp32_128": 3.75875e-3,
                                        "plain_ppm_fp64_128": 3.75529e-3},
                                 "wall_s": {"learned_fp32_64": 0.55, "plain_ppm_fp64_64": 0.28,
                                            "learned_fp32_128": 5.33,
                                            "plain_ppm_fp64_128": 2.65}},
            "equal_l1_reached": False, "wall_time_ratio_at_equal_l1": None,
            "source": "profiles/learned_turb2d.md, profiles/learned_turb3d.md"}


def turb3d_block(n, precision, recon="ppm", ic=None):
    from idefix_amd.datablock import DataBlock
    from idefix_amd import problems
    g = DataBlock(3, (n, n, n), comps=3, mhd=True, recon=recon, solver="hlld",
                  emf="uct_contact", precision=precision, bc="periodic", gamma=5.0 / 3.0,
                  nstages=3, cfl=0.4, first_dt=1e-4)
    if ic is None:
        ic = problems.turbulence(g)
    g.upload(*ic)
    return g, ic


def learned_params():
    """The trained network of the learned limiter: the 3-D one when it exists, else the
    2-D one (the inference cost does not depend on the weights)."""
    import numpy as np
    for name in ("learned_turb3d.npy", "learned_turb2d.npy"):
        path = os.path.join(ROOT, "idefix_amd", name)
        if os.path.exists(path):
            return np.load(path), name
    return None, None


def traffic_child(cache, n, region, cycles):
    """Run under rocprofv3 --pmc: the FP64 headline block from the cached initial state,
    `cycles` cycles of the same schedule as the timed run."""
    import numpy as np
    import torch
    ic = (np.load(cache + "_Vc.npy", mmap_mode="r"), np.load(cache + "_Vs.npy", mmap_mode="r"))
    g, _ = turb3d_block(n, "fp64", ic=ic)
    for _ in range(cycles):
        g.cycle()
    torch.cuda.synchronize()
    print("traffic child done", region, flush=True)


def kernel_counter_average(path, counter, region):
    """Average per dispatch of `counter` over the FP64 dispatches of the kernel of
    `region` in a rocprofv3 counter_collection.csv (rows of one dispatch are summed:
    rocprofv3 may report a counter in several rows per dispatch); None if none."""
    import csv
    prefix, sub = SINGLE_KERNEL_REGIONS[region]
    per = {}
    with open(path) as fh:
        for row in csv.DictReader(fh):
            name = row["Kernel_Name"]
            if row["Counter_Name"] != counter or prefix not in name or sub not in name:
                continue
            if "double" not in name:
                continue
            per[row["Dispatch_Id"]] = per.get(row["Dispatch_Id"], 0.0) + float(row["Counter_Value"])
    return sum(per.values()) / len(per) if per else None


def measure_traffic(cache, n, region, timeout=240):
    """HBM bytes per launch of the kernel of `region`, from two rocprofv3 --pmc runs of
    traffic_child (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch, averaged over the
    kernel's dispatches), scaled by the plain-copy calibration, and its VALU instructions
    per launch from a third run (SQ_INSTS_VALU).  (None, reason, None) if rocprofv3 is
    missing or a run fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found", None
    env = dict(os.environ, TMPDIR="/tmp")
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU"):
        d = tempfile.mkdtemp(prefix="idfx_pmc_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", str(timeout), prof, "--pmc", counter,
               "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable, "-u",
               os.path.abspath(__file__), "--traffic-child", region, "--n", str(n),
               "--ic-cache", cache]
        print("traffic: rocprofv3 --pmc %s (%s)" % (counter, region), file=sys.stderr, flush=True)
        r = subprocess.run(cmd, env=env, cwd="/tmp", capture_output=True, text=True)
        files = [os.path.join(root, f) for root, _, fs in os.walk(d) for f in fs
                 if f.endswith("counter_collection.csv")]
        if not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, "rocprofv3 --pmc %s failed (rc %d)" % (counter, r.returncode), None
        avg = kernel_counter_average(files[0], counter, region)
        shutil.rmtree(d, ignore_errors=True)
        if avg is None:
            prefix, sub = SINGLE_KERNEL_REGIONS[region]
            return None, "no %s dispatch of %s in the profile" % (counter, prefix + sub), None
        out[counter] = avg
    return (out["FETCH_SIZE"] * 1024.0 * FETCH_SCALE + out["WRITE_SIZE"] * 1024.0 * WRITE_SCALE,
            {"fetch_bytes": out["FETCH_SIZE"] * 1024.0 * FETCH_SCALE,
             "write_bytes": out["WRITE_SIZE"] * 1024.0 * WRITE_SCALE,
             "scale": "FETCH_SIZE x %.1f, WRITE_SIZE x %.1f (1 GiB copy calibration)"
                      % (FETCH_SCALE, WRITE_SCALE)},
            out["SQ_INSTS_VALU"])


def cpu_baseline_turb3d(n, seconds):
    """CPU oracle on the host cores: the same 3-D turbulence configuration at n^3, as many
    RK3 cycles as fit in about `seconds` after one warm-up cycle."""
    from oracle import oracle as O
    from idefix_amd import problems
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    threads = max(1, min(threads, os.cpu_count() or threads))
    O.set_threads(threads)
    cfg = O.make_cfg(3, comps=3, mhd=True, recon="ppm", solver="hlld", emf="uct_contact",
                     nstages=3, nx=(n, n, n), bc="periodic", gamma=5.0 / 3.0, cfl=0.4)
    o = O.Oracle(cfg)
    Vc, Vs = problems.turbulence(O.OracleGrid(o))
    o.Vc[...] = Vc
    o.Vs[...] = Vs
    o.set_boundaries()
    o.cycle()
    k = 0
    t0 = time.perf_counter()
    while True:
        o.cycle()
        k += 1
        el = time.perf_counter() - t0
        if el > seconds or k >= 1000:
            break
    return {"value": n ** 3 * k / el, "unit": "zone-updates/s", "cores": threads,
            "kind": "port",
            "sample": "3-D turbulence %d^3 (1/%d of the headline cells), PPM/HLLD/uct_contact/"
                      "RK3 FP64, %d cycles after 1 warm-up cycle (%.1f s)"
                      % (n, (512 // n) ** 3, k, el)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; run with "
              "--gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.traffic_child:
        traffic_child(args.ic_cache, args.n, args.traffic_child, 2)
        return
    import numpy as np
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
    n = args.n
    print("bench: 3-D turbulence %d^3, initial state" % n, file=sys.stderr, flush=True)
    g, ic = turb3d_block(n, "fp64")
    el = time_cycles(g, args.steps, args.warmup)
    cells = n ** 3
    zu = cells * args.steps / el
    ms = 1e3 * el / args.steps
    print("bench: FP64 %.3e zone-updates/s" % zu, file=sys.stderr, flush=True)
    kt, per_cycle = kernel_times(g, args.kernel_reps)
    dom = max((k for k in per_cycle if k in SINGLE_KERNEL_REGIONS), key=lambda k: per_cycle[k])
    stage_bytes = B_ZU_TURB3D / STAGES_TURB3D * cells
    achieved = stage_bytes / (kt[dom] * 1e-3) / 1e9
    divb = g.check_divb()
    del g
    torch.cuda.empty_cache()
    copy_gbs = copy_bandwidth()
    print("bench: device copy %.0f GB/s" % copy_gbs, file=sys.stderr, flush=True)
    traffic, traffic_note, valu_insts = None, "not measured (--no-traffic)", None
    if not args.no_traffic:
        # the children start from this run's initial state, handed over in memory-backed
        # files (removed afterwards)
        shm = "/dev/shm" if os.path.isdir("/dev/shm") else "/tmp"
        cache = os.path.join(shm, "idfx_bench_ic_%d_%d" % (n, os.getpid()))
        try:
            np.save(cache + "_Vc.npy", ic[0])
            np.save(cache + "_Vs.npy", ic[1])
            traffic, traffic_note, valu_insts = measure_traffic(cache, n, dom)
        finally:
            for suffix in ("_Vc.npy", "_Vs.npy"):
                if os.path.exists(cache + suffix):
                    os.remove(cache + suffix)
    fp32 = None
    if not args.no_fp32:
        g32, _ = turb3d_block(n, "fp32", ic=ic)
        el32 = time_cycles(g32, args.steps, args.warmup)
        fp32 = cells * args.steps / el32
        del g32
        torch.cuda.empty_cache()
    del ic
    learned = None
    if not args.no_learned:
        nl = n // 2
        params, pname = learned_params()
        gl, _ = turb3d_block(nl, "fp32", recon="mlp")
        if params is not None:
            gl.set_mlp_params(params)
        ell = time_cycles(gl, args.steps, args.warmup)
        learned = {"zone_updates_per_s": nl ** 3 * args.steps / ell,
                   "ms_per_step": 1e3 * ell / args.steps,
                   "workload": "3-D turbulence %d^3, FP32 stencil + learned limiter (BF16 MFMA "
                               "inference, PPM fallback), HLLD, uct_contact, RK3, CFL 0.4" % nl,
                   "network": pname}
        # MFMA rate of the inference kernel (region timer, HIP events on its stream): one
        # launch per direction covers the faces beg..end along it and every transverse
        # index, ghosts included
        kl, _ = kernel_times(gl, 3)
        if "mlp" in kl:
            faces = (nl + 1) * gl.np_tot[1] * gl.np_tot[2]
            flop = MLP_FLOP_PER_FACE * faces
            tf = flop / (kl["mlp"] * 1e-3) / 1e12
            learned["mlp_roofline"] = {
                "bound": "mfma", "kernel": "k_mlp_infer", "achieved": tf,
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / BF16_PEAK_TFLOPS,
                "flop_per_launch": flop, "interfaces_per_launch": faces,
                "ms_per_launch": kl["mlp"]}
            learned["kernel_ms"] = kl
        del gl
        torch.cuda.empty_cache()
    headline = None if args.no_headline else reference_headline(args.steps, args.warmup)
    cpu = None if args.no_cpu_baseline else cpu_baseline_turb3d(args.cpu_n, args.cpu_seconds)
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
        "config": {"workload": "3-D decaying MHD turbulence %d^3, periodic unit cube, PPM, HLLD, "
                               "CT uct_contact, RK3, CFL 0.4, plain FP64" % n,
                   "cells": cells, "step": "one RK3 cycle", "parallelism": "1 GPU"},
        "fp32_stencil_value": fp32,
        "fp32_learned_half_resolution": learned,
        "equal_l1": EQUAL_L1,
        "kernel_ms": kt,
        "kernel_ms_per_cycle": per_cycle,
        "cycle_hbm_frac": B_ZU_TURB3D * zu / (HBM_PEAK_GBS * 1e9),
        "div_b": divb,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "algorithmic_bytes": stage_bytes, "ms_per_launch": kt[dom],
                     "traffic": traffic, "traffic_note": traffic_note,
                     "copy_gbs": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs if copy_gbs else None,
                     "valu": None if valu_insts is None else {
                         "insts_per_launch": valu_insts,
                         "busy": valu_insts * VALU_CYCLES_PER_INST
                                 / (VALU_SIMDS * VALU_CLOCK_HZ * kt[dom] * 1e-3),
                         "note": "SQ_INSTS_VALU per launch (rocprofv3 --pmc child run) x 4 "
                                 "cycles / (1024 SIMDs x 2.4 GHz x launch time): the FP64 "
                                 "VALU issue fraction, the kernel's actual bound"}},
        "cpu_baseline": cpu,
        "reference_benchmark": headline,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
