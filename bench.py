#!/usr/bin/env python3
"""Benchmark of the isothermal-MHD cell-update path on one MI355X.

Workload (one step = one pass of the driver loop body over the box): 256^3 periodic,
rho = 1, B0 = (0,0,1), seeded solenoidal k^-2 velocity on 1 <= |k| <= 8 at rms Mach 2,
one passive scalar of seeded uniform [0,1) noise; VL2 + CT, HLLD, primitive PLM
(MC-beta 1.01); step = copy_velocity, VL2 step, passive step, ghost fill, CFL reduction and
device-side dt.  Both the FP32 path (FP32 bulk, FP64 face B / corner EMF / CFL, Kahan-
compensated rho, rho v) and the FP64 path run the same K steps from the same initial
state; `value` is the FP32 path, the FP64 path and the L1 difference are reported beside.

Timing: W untimed warm-up steps on the initial state, then the state is reset and exactly
K steps are timed (device synchronised on both sides).  Per-kernel durations come from HIP
events recorded on the solver's stream around every launch of the timed region.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s + HBM GB/s % peak, FP32 vs FP64 path at stated L1 error"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def bytes_per_item(name, esz):
    """Algorithmic (compulsory) HBM bytes per work item of each kernel: every input field
    read once and every output written once per cell or face."""
    d = 8
    if name.startswith("flux_"):          # rho, 3 m, 2 transverse B (real) + normal face B (f64) in; 4 fluxes + 2 face EMFs out
        return 12 * esz + d
    if name.startswith("corner"):         # 7 state + 6 face EMFs in (real); 3 corner EMFs out (f64)
        return 13 * esz + 3 * d
    if name.startswith("update"):         # 4 state + 12 fluxes in (real), 3 face B + 3 corner EMF in (f64); 7 state (real) + 3 face B (f64) out
        return 23 * esz + 9 * d
    return None


def run_path(ics_state, n, prec, steps, warmup, cour):
    from isothermal_mhd_amd import VanLeerSolver
    q, bf, qp = ics_state
    S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", bc="periodic", precision=prec, cour=cour)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    if warmup:
        S.run(warmup, graph=False)
    S.sync()
    # reset to the initial state: the timed steps start from t = 0, dt = 1e-15 (driver.f90:249)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    S.set_time(t=0.0, dt=1.0e-15)
    S.set_profiling(True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=False)
    S.sync()
    t1 = time.perf_counter()
    prof = S.profile()
    state = S.get_state()
    diag = S.diagnostics()
    t, dt = S.get_time()
    floors = S.floor_count()
    S.close()
    return {"sec": t1 - t0, "prof": prof, "state": state, "diag": diag, "t": float(t[0]),
            "dt": float(dt[0]), "floors": floors}


def roofline(prof, esz):
    # dominant kernel family: largest total event time
    fam = {}
    for name, (ms, nl, items) in prof.items():
        key = name[:-2] if name[-2:] in ("_x", "_y", "_z") else name
        f = fam.setdefault(key, [0.0, 0, 0, []])
        f[0] += ms
        f[1] += nl
        f[2] += items
        f[3].append(name)
    key = max(fam, key=lambda k: fam[k][0])
    ms, nl, items, names = fam[key]
    bpi = bytes_per_item(key, esz)
    avg_ms = ms / nl
    items_per_launch = items / nl
    out = {"bound": "hbm", "kernel": key, "launches": nl, "avg_launch_ms": avg_ms,
           "items_per_launch": items_per_launch, "bytes_per_item": bpi,
           "peak": PEAK_HBM_GBS, "unit": "GB/s", "traffic": None}
    if bpi is not None:
        ach = bpi * items_per_launch / (avg_ms * 1e-3) / 1e9
        out.update({"achieved": ach, "frac": ach / PEAK_HBM_GBS})
    pmc = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            tr = json.load(f)
        if key in tr:
            out["traffic"] = tr[key]
    return out


def l1_errors(sa, sb, nb=6):
    import numpy as np
    qa, ba, pa = sa
    qb, bb, pb = sb
    I = (slice(nb, -nb),) * 3
    out = {}
    for f, nm in enumerate(["rho", "mx", "my", "mz"]):
        a, b = qa[I + (f,)], qb[I + (f,)]
        out[nm] = float(np.abs(a - b).sum() / np.abs(b).sum())
        out[nm + "_linf"] = float(np.abs(a - b).max() / np.abs(b).max())
    for d, nm in enumerate(["bx", "by", "bz"]):
        a, b = ba[I + (d,)], bb[I + (d,)]
        out[nm] = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))
    Ip = (slice(nb - 2, -(nb - 2)),) * 3
    a, b = pa[Ip + (0,)], pb[Ip + (0,)]
    out["passive"] = float(np.abs(a - b).sum() / np.abs(b).sum())
    return out


def cpu_baseline(n_cpu, steps, cour):
    """The CPU restatement (oracle/imhd_oracle.c, FP64, one thread) on a bounded sample of
    the same workload: the same IC generator at n_cpu^3, `steps` driver steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from isothermal_mhd_amd import ics
    from refutil import Oracle
    q, bf, qp = ics.config_c2(n=n_cpu)
    O = Oracle(n_cpu, n_cpu, n_cpu, npass=1, mhd=2, riemann=2, recon=0, bc=2)
    O.set_state(q, bf, qp)
    O.bc()
    dt = 1.0e-15
    t0 = time.perf_counter()
    for _ in range(steps):
        O.copy_velocity(0)
        O.step(dt)
        O.copy_velocity(1)
        O.passive_step(dt)
        O.bc()
        dt = cour * (1.0 / n_cpu) / O.max_speed()
    sec = time.perf_counter() - t0
    return {"value": n_cpu ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} driver steps of the same workload at {n_cpu}^3 (same IC generator), "
                      f"oracle/imhd_oracle.c FP64 single-threaded, {sec:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cour", type=float, default=0.35)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the box would shard by 3-D blocks with an RCCL halo exchange, not implemented)",
              file=sys.stderr)
        return 2
    from isothermal_mhd_amd import ics
    n, K, W = args.n, args.steps, args.warmup
    state0 = ics.config_c2(n=n)
    r32 = run_path(state0, n, 32, K, W, args.cour)
    r64 = None if args.no_fp64 else run_path(state0, n, 64, K, W, args.cour)
    cups32 = n ** 3 * K / r32["sec"] / 1e6
    res = {
        "metric": METRIC,
        "value": cups32,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r32["sec"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{n}^3 periodic isothermal MHD, VL2+CT, HLLD, primitive PLM (MC-beta 1.01), "
                        "1 passive scalar; rho=1, B0=(0,0,1), solenoidal k^-2 velocity 1<=|k|<=8 at "
                        "rms Mach 2 (seed 2002), passive U[0,1) (seed 2003)",
            "cells": n ** 3, "cfl": args.cour,
            "precision": "FP32 bulk; FP64 face B, corner-EMF sum, CFL reduction, diagnostics; "
                         "Kahan-compensated rho and rho v",
            "step": "copy_velocity + VL2 predictor/corrector + CT + passive + ghost fill + CFL dt",
        },
        "fp32": {"ms_per_step": r32["sec"] / K * 1e3, "t_end": r32["t"],
                 "max_divb": r32["diag"]["divb_max"], "floor_resets": r32["floors"]},
        "roofline": roofline(r32["prof"], 4),
        "kernels_fp32_ms": {k: round(v[0] / v[1], 5) for k, v in r32["prof"].items()},
    }
    if r64 is not None:
        cups64 = n ** 3 * K / r64["sec"] / 1e6
        res["fp64"] = {"value": cups64, "ms_per_step": r64["sec"] / K * 1e3, "t_end": r64["t"],
                       "max_divb": r64["diag"]["divb_max"], "floor_resets": r64["floors"],
                       "roofline": roofline(r64["prof"], 8)}
        res["fp32_over_fp64"] = cups32 / cups64
        res["l1_fp32_vs_fp64"] = l1_errors(r32["state"], r64["state"])
        res["kernels_fp64_ms"] = {k: round(v[0] / v[1], 5) for k, v in r64["prof"].items()}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps, args.cour)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
