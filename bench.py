#!/usr/bin/env python3
"""Throughput of the implicit film-height step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024]

Workload: the 2-D well-mixed all-physics film on 1024 x 1024 cells
(``yuzen_amd.cases.allphys_2d``: capillary, disjoining, vertical and
in-plane gravity, spin, Marangoni on a frozen concentration, slip; Symmetry
walls; fixed dt; Newton tolerance 1e-8). A step is one implicit time step
(``C_NewtonKantorovich::TimeAdvance``): every Newton pass assembles the
Jacobian on the device, solves with BiCGSTAB preconditioned by the
structured smoothed-aggregation multigrid, and updates the film. The film,
substrate and concentration are resident in HBM before the timed region;
W untimed steps run first, then exactly K steps are timed between two
device synchronisations.

Printed: one JSON line with the steps/s, the roofline of the dominant
kernel (the fine-level stencil product of the Krylov iteration, timed with
HIP events on its own stream over the timed region) and a CPU baseline:
the FP64 CPU oracle on one host core over a bounded sample of the same
workload.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Newton steps/s, 1024² 2D film: learned-precond FP32 vs plain FP64 AMG @1e-8"
HBM_PEAK_GBS = 8000.0
DT = 2e-6


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n, passes_per_step):
    """The oracle (CPU FP64 restatement of the reference method, one
    core) on a bounded sample of the same workload: the first Newton pass
    of the first step of the same case (the explicit half A(h^n)h^n, the
    assembly of A and J, the multigrid set-up and the BiCGSTAB solve),
    scaled to steps/s with the Newton passes per step of the device run
    (the two take the same passes: the parity tests pin the counts)."""
    from yuzen_amd import cases
    from oracle import binding as ob
    c, h, s, cc, dt, _ = cases.allphys_2d(n=n, dt=DT)
    c.nk_max = 1          # stop after one pass
    o = ob.OracleStep(c, h, s, cc, dt)
    t0 = time.perf_counter()
    rc, nk, lin, r = o.advance()
    el = time.perf_counter() - t0
    per_step = el * passes_per_step
    return {
        "value": round(1.0 / per_step, 6),
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one Newton pass ({lin} BiCGSTAB iterations) of the first step of the same {n}x{n} "
                  f"case, CPU FP64 oracle (scalar, -O2), {el:.1f} s on {cpu_model()}; "
                  f"x {passes_per_step:.2f} passes/step",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the step is single-GPU and no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    from yuzen_amd import cases
    from yuzen_amd.step import HipStep

    c, h, s, cc, dt, _ = cases.allphys_2d(n=args.n, dt=DT)
    step = HipStep(c, h, s, cc, dt)
    for _ in range(args.warmup):
        rc, nk, lin, r = step.advance()
        if rc != 0:
            raise SystemExit(f"warm-up step failed (rc={rc})")
    step.kernel_time(1)
    passes = lins = 0
    phases = np.zeros(6)
    step.h()  # synchronise
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rc, nk, lin, r = step.advance()
        if rc != 0:
            raise SystemExit(f"timed step failed (rc={rc})")
        passes += nk
        lins += lin
        phases += step.timing()
    hfinal = step.h()  # synchronise
    el = time.perf_counter() - t0
    kcount, kms = step.kernel_time(0)
    steps_per_s = args.steps / el

    N = c.nx * c.ny
    # algorithmic bytes of one launch of the fine-level product
    # v = A p, (rh, v): 13 planes of FP64 values, p once, rh once, v written
    bytes_per_launch = N * (13 * 8 + 8 + 8 + 8)
    avg_s = (kms / 1e3) / max(kcount, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if kcount else 0.0
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "kernel": "yz::k_spmv_dot<13> (fine-level stencil SpMV + fused dot)",
        "launches": kcount,
        "avg_us": round(avg_s * 1e6, 2),
        "bytes_per_launch": bytes_per_launch,
    }
    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"2-D WM all-physics film {args.n}x{args.n}, plain FP64 AMG-BiCGSTAB "
                        "(the learned FP32 path is not built yet)",
            "cells": N, "dt": DT, "newton_tol": c.nk_eps, "linear_eps": c.lin_eps,
            "precond": "structured SA-AMG + SPAI0", "domain": "[0,10]^2, Symmetry walls",
        },
        "newton_passes_per_step": round(passes / args.steps, 3),
        "krylov_iters_per_pass": round(lins / max(passes, 1), 2),
        "phases_ms_per_step": {k: round(v / args.steps, 3) for k, v in
                               zip(["assembly", "residual", "amg_setup", "krylov", "update", "step"], phases)},
        "roofline": roof,
        "cpu_baseline": None,
        "h_checksum": float(np.sum(hfinal)),
    }
    step.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n, passes / args.steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
