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
events recorded on the solver's stream around every launch of the timed region.  The same
K steps replayed from one captured hipGraph (no events) are timed beside it ("graph").

Prints ONE JSON line.

HBM traffic (roofline.traffic) comes from two rocprofv3 counter runs of `--pmc-run`
(FETCH_SIZE, then WRITE_SIZE), scaled by what the same counters read for a plain copy of
known size (k_convert_in<double> of an FP64 set_state), and reduced by `--pmc-summarize`
into profiles/pmc_traffic.json:
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d gpurun_out/pmc_fetch -- python3 bench.py --pmc-run
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d gpurun_out/pmc_write -- python3 bench.py --pmc-run
  python3 bench.py --pmc-summarize gpurun_out/pmc_fetch gpurun_out/pmc_write
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s + HBM GB/s % peak, FP32 vs FP64 path at stated L1 error"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def bytes_per_item(name, esz, npass=1):
    """Algorithmic (compulsory) HBM bytes per work item of each kernel: every input field
    read once and every output written once per cell or face (npass: passive scalars of
    the workload, whose velocities the fused stages write)."""
    d = 8
    pv = 3 * esz if npass > 0 else 0     # passive velocities written by the fused stages
    if name.startswith("stage_pred"):     # fused predictor: 7 state + 3 face B in (t^n), 7 state + 3 face B out (half step)
        return 14 * esz + 6 * d + pv
    if name.startswith("stage_corr"):     # fused corrector: half step 7 + 3 face B, t^n rho, m + 3 face B in; 7 + 3 out
        b = 18 * esz + 9 * d + pv
        if esz == 4:
            b += 8 * esz                  # Kahan terms of rho, m: read and written (FP32 path)
        return b
    if name.startswith("flux_"):          # fused sweeps, per cell: 7 state (real) + 3 face B (f64) in; 3 x (4 fluxes + 2 face EMFs) out
        return 25 * esz + 3 * d
    if name.startswith("corner"):         # 7 state + 6 face EMFs in (real); 3 corner EMFs out (f64)
        return 13 * esz + 3 * d
    if name.startswith("update"):         # 4 state + 12 fluxes in (real), 3 face B + 3 corner EMF in (f64); 7 state (real) + 3 face B (f64) out
        b = 23 * esz + 9 * d
        if name.startswith("update_corr") and esz == 4:
            b += 8 * esz                  # Kahan terms of rho, m: read and written (FP32 path)
        return b
    return None


def run_path(ics_state, n, prec, steps, warmup, cour):
    from isothermal_mhd_amd import VanLeerSolver
    q, bf, qp = ics_state
    S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", bc="periodic", precision=prec, cour=cour)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    tune = S.autotune(steps=5)      # launch shapes (set-up, outside the timed region)
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
    # the same K steps replayed from a captured hipGraph, without per-kernel events
    S.set_profiling(False)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    S.set_time(t=0.0, dt=1.0e-15)
    S.run(2, graph=True)          # capture both buffer parities outside the timed region
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    S.set_time(t=0.0, dt=1.0e-15)
    S.sync()
    g0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    g1 = time.perf_counter()
    S.close()
    return {"sec": t1 - t0, "graph_sec": g1 - g0, "prof": prof, "state": state, "diag": diag,
            "t": float(t[0]), "dt": float(dt[0]), "floors": floors, "autotune": tune.split("; ")[-1]}


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
    # the fused stage kernels are VALU-bound: their VALU issue rate against the MI355X issue
    # peak, from the counted instructions per launch and this run's launch time
    sq = os.path.join(ROOT, "profiles", "sq_valu.json")
    if esz == 4 and os.path.exists(sq):
        with open(sq) as f:
            vq = json.load(f)
        if key in vq:
            ach = vq[key]["valu_wave_instr"] / (avg_ms * 1e-3) / 1e9
            out["valu_issue"] = {"achieved": ach, "peak": VALU_PEAK_GWI, "unit": "G wave-instr/s",
                                 "frac": ach / VALU_PEAK_GWI, "wave_instr_per_launch": vq[key]["valu_wave_instr"]}
    return out


# family (bench profiler name) -> kernel symbol it launches (fused three-direction sweeps,
# z-marching column updates)
FAMILY_SYMBOL = {"stage_pred": "k_vl_stage", "stage_corr_cfl": "k_vl_stage",
                 "flux_pred": "k_flux3", "flux_corr": "k_flux3", "corner_pred": "k_corner",
                 "corner_corr": "k_corner", "update_pred": "k_update_z", "update_corr_cfl": "k_update_z",
                 "pass_stage_pred": "k_pass_stage_z", "pass_stage_corr": "k_pass_stage_z",
                 "fill_ghosts": "k_fill_ghost_band", "fill_ghosts_pass": "k_fill_ghost_band"}


def pmc_run(n=256, steps=3):
    """Short FP32 run of the workload for counter collection.  Its set_state converts each
    FP64 state field to FP32 with a k_convert_in<float> dispatch over the padded box (8 bytes
    read and 4 written per cell, no reuse): the calibration of the counters.  Writes the
    family -> (kernel, grid size) map for --pmc-summarize."""
    from isothermal_mhd_amd import VanLeerSolver, ics
    q, bf, qp = ics.config_c2(n=n)
    S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", bc="periodic", precision=32)
    S.set_state(q, bf, qp)
    S.boundary_conditions()
    S.set_profiling(True)
    S.run(steps, graph=False)
    prof = S.profile()
    S.close()
    fam = {}
    for name, (ms, nl, items) in prof.items():
        key = name[:-2] if name[-2:] in ("_x", "_y", "_z") else name
        if key in FAMILY_SYMBOL:
            grids = fam.setdefault(key, {"symbol": FAMILY_SYMBOL[key], "grids": [], "order": len(fam)})
            grids["items"] = items // nl
    nc = (n + 12) ** 3
    out = {"n": n, "families": fam, "copy": {"symbol": "k_convert_in<float>", "grid": (nc + 255) // 256 * 256,
                                              "bytes_read": nc * 8, "bytes_written": nc * 4}}
    os.makedirs(os.path.join(ROOT, "gpurun_out"), exist_ok=True)
    with open(os.path.join(ROOT, "gpurun_out", "pmc_map.json"), "w") as f:
        json.dump(out, f)
    print(json.dumps({"pmc_run": "done", "families": len(fam)}))
    return 0


def _read_counter(dirname, counter, grid=None):
    """per-dispatch (kernel name, counter value KiB) in dispatch order (only the dispatches
    of that grid size when grid is given)"""
    rows = []
    files = glob.glob(os.path.join(dirname, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {dirname}")
    if len(files) > 1:
        raise SystemExit(f"{len(files)} counter_collection.csv files under {dirname}: one run per directory")
    for fn in files:
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                if grid is not None and int(r.get("Grid_Size", -1)) != grid:
                    continue
                rows.append((int(r.get("Dispatch_Id", 0)), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(n, v) for _, n, v in rows]


def _per_step_sequence(rows, symbols):
    """the kernels of one step, in dispatch order, from the middle of the timed steps"""
    seq = [(n, v) for n, v in rows if any(s in n for s in symbols)]
    return seq


def pmc_summarize(fetch_dir, write_dir):
    with open(os.path.join(ROOT, "gpurun_out", "pmc_map.json")) as f:
        mp = json.load(f)
    fr, wr = _read_counter(fetch_dir, "FETCH_SIZE"), _read_counter(write_dir, "WRITE_SIZE")
    cp = mp["copy"]
    sym = cp["symbol"]
    cf = [v for n, v in _read_counter(fetch_dir, "FETCH_SIZE", cp.get("grid")) if sym in n]
    cw = [v for n, v in _read_counter(write_dir, "WRITE_SIZE", cp.get("grid")) if sym in n]
    if not cf or not cw:
        raise SystemExit(f"no {sym} dispatches (the calibration copy) in the counter files")
    cf, cw = sum(cf) / len(cf), sum(cw) / len(cw)
    f_fac = cf * 1024.0 / cp["bytes_read"]     # counter bytes per true byte (copy kernel)
    w_fac = cw * 1024.0 / cp["bytes_written"]
    out = {"_calibration": {"kernel": sym, "bytes_read": cp["bytes_read"], "bytes_written": cp["bytes_written"],
                            "FETCH_SIZE_KiB": cf, "WRITE_SIZE_KiB": cw, "fetch_factor": f_fac, "write_factor": w_fac}}
    # a native step launches each family's kernel in a fixed order; families sharing a kernel
    # symbol are told apart by their position in that order
    fams = sorted(mp["families"].items(), key=lambda kv: kv[1]["order"])
    lines = ["| family | kernel | cells or faces / launch | FETCH KiB/dispatch | WRITE KiB/dispatch | HBM bytes/launch (calibrated) |",
             "|---|---|---|---|---|---|"]
    for key, v in fams:
        same = [k for k, w in fams if w["symbol"] == v["symbol"]]
        pos = same.index(key)
        fv = [val for n, val in fr if v["symbol"] + "<float" in n]
        wv = [val for n, val in wr if v["symbol"] + "<float" in n]
        fv = fv[pos::len(same)]
        wv = wv[pos::len(same)]
        if not fv or not wv:
            continue
        a, b = sum(fv) / len(fv), sum(wv) / len(wv)
        t = a * 1024.0 / f_fac + b * 1024.0 / w_fac
        lines.append(f"| {key} | {v['symbol']}<float> | {v['items']} | {a:.0f} | {b:.0f} | {t:.4g} |")
        out[key] = t
    with open(os.path.join(ROOT, "profiles", "pmc_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))
    print(json.dumps(out["_calibration"]))
    return 0


# Nominal VALU issue rate of MI355X: 256 CUs x 4 SIMDs, a wave64 VALU instruction every 4
# cycles per SIMD (16 lanes), 2.4 GHz engine clock.  Some simple instructions issue faster
# (a passive-stage kernel measures above it), so a fraction of it is an approximate
# utilisation, not a bound.
VALU_PEAK_GWI = 256 * 4 / 4 * 2.4


def sq_summarize(sq_dir):
    """VALU wave-instructions per launch of each family (SQ_INSTS_VALU from a rocprofv3 run
    of --pmc-run), matched to families by dispatch order as in pmc_summarize; written to
    profiles/sq_valu.json, from which bench.py reports each kernel's VALU issue rate."""
    with open(os.path.join(ROOT, "gpurun_out", "pmc_map.json")) as f:
        mp = json.load(f)
    counters = {}
    for c in ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "SQ_WAVES"):
        counters[c] = _read_counter(sq_dir, c)
    fams = sorted(mp["families"].items(), key=lambda kv: kv[1]["order"])
    out = {"_peak_G_wave_instr_per_s": VALU_PEAK_GWI}
    lines = ["| family | kernel | VALU wave-instr / launch | VALU instr / wave | active VALU / wave-cycles |",
             "|---|---|---|---|---|"]
    for key, v in fams:
        same = [k for k, w in fams if w["symbol"] == v["symbol"]]
        pos = same.index(key)
        per = {}
        for c, rows in counters.items():
            vals = [val for n, val in rows if v["symbol"] + "<float" in n][pos::len(same)]
            per[c] = sum(vals) / len(vals) if vals else 0.0
        if not per["SQ_WAVES"]:
            continue
        out[key] = {"valu_wave_instr": per["SQ_INSTS_VALU"], "waves": per["SQ_WAVES"],
                    "active_valu_per_wave_cycle": per["SQ_ACTIVE_INST_VALU"] / per["SQ_WAVE_CYCLES"]}
        lines.append(f"| {key} | {v['symbol']}<float> | {per['SQ_INSTS_VALU']:.4g} | "
                     f"{per['SQ_INSTS_VALU'] / per['SQ_WAVES']:.0f} | {out[key]['active_valu_per_wave_cycle']:.3f} |")
    with open(os.path.join(ROOT, "profiles", "sq_valu.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))
    return 0


def run_big(n, steps, cour, warmup=2):
    """Throughput of one large FP32 box (HLLD + VL2 + CT, PLM, fused stages, whole step in
    a hipGraph) on the separable solenoidal IC of ics.separable_large, with the context's
    device memory: the single-GPU capacity line (512^3, 1024^3)."""
    from isothermal_mhd_amd import VanLeerSolver, ics
    t0 = time.perf_counter()
    q, bf = ics.separable_large(n)
    t_ic = time.perf_counter() - t0
    S = VanLeerSolver((n, n, n), npass=0, riemann="hlld", bc="periodic", precision=32, cour=cour)
    t0 = time.perf_counter()
    S.set_state(q, bf)
    S.sync()
    t_up = time.perf_counter() - t0
    del q, bf
    S.boundary_conditions()
    S.run(warmup, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    sec = time.perf_counter() - t0
    d = S.diagnostics()
    mem = S.device_bytes()
    S.close()
    return {"n": n, "value": n ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "steps": steps,
            "ms_per_step": sec / steps * 1e3, "device_GB": mem / 1e9, "upload_s": t_up, "ic_s": t_ic,
            "max_divb": float(d["divb_max"]), "dtype": "f32",
            "workload": f"{n}^3 periodic, rho=1, B0=(0.05,0.05,0), separable solenoidal velocity (6 modes per "
                        "component, |k|<=16, rms Mach 2, seed 5005); HLLD + VL2 + CT, PLM, FP32, hipGraph"}


def run_ensemble(steps, cour, warmup=2):
    """Batched-ensemble configuration: 48 periodic boxes of uneven sizes drawn from
    {32..192}^3 (seed 4004), FP32, HLLE + VL2 + CT, every kernel one launch for all boxes,
    the whole step in a hipGraph.  Returns throughput over all boxes."""
    from isothermal_mhd_amd import VanLeerSolver, ics
    dims, states, _ = ics.ensemble_c4()
    S = VanLeerSolver(dims, npass=0, riemann="hlle", precision=32, cour=cour)
    for b, (q, bf, qp) in enumerate(states):
        S.set_state(q, bf, None, box=b)
    del states
    S.boundary_conditions()
    S.run(warmup, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    sec = time.perf_counter() - t0
    cells = sum(a * b * c for a, b, c in dims)
    S.close()
    return {"value": cells * steps / sec / 1e6, "unit": "Mcell-updates/s", "boxes": len(dims), "cells": cells,
            "steps": steps, "ms_per_step": sec / steps * 1e3, "dtype": "f32",
            "workload": "48 periodic boxes, sizes U{32..192}^3 (seed 4004), |B0| U[0.1,1] random direction, "
                        "rms Mach U[0.5,4], k^-2 velocity; HLLE + VL2 + CT, PLM; one launch per kernel for all "
                        "boxes, whole step in a hipGraph"}


def run_io(n, reps=2):
    """Restart and snapshot files of one n^3 FP32 box with one passive scalar (the 256^3
    turbulence IC): imhd_write_cube (REAL*4 and REAL*8 xyz cube, packed on the device by
    k_pack_cube), imhd_write_restart and imhd_read_restart, to files under $TMPDIR (rates
    include PCIe, the host page cache and the file system); then the device generation of
    the reference's random force (imhd_generate_force, k_max = 2) and the analysis
    program's velocity spectrum, density PDF and mode decomposition (imhd_spectrum,
    imhd_density_pdf, imhd_mode_spectra)."""
    import tempfile
    from isothermal_mhd_amd import VanLeerSolver, ics
    q, bf, qp = ics.config_c2(n=n)
    S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", bc="periodic", precision=32)
    S.set_state(q, bf, qp)
    del q, bf, qp
    out = {"n": n, "dtype": "f32", "reps": reps}
    with tempfile.TemporaryDirectory() as d:
        ops = [("cube_f32", lambda p: S.write_cube(p, double_prec=False)),
               ("cube_f64", lambda p: S.write_cube(p, double_prec=True)),
               ("write_restart", lambda p: S.write_restart(p, 1.0, 1, (1234, 5678, 9123))),
               ("read_restart", lambda p: S.read_restart(p))]
        for name, fn in ops:
            p = os.path.join(d, "restart.dat" if "restart" in name else name + ".dat")
            best = None
            for _ in range(reps):
                S.sync()
                t0 = time.perf_counter()
                fn(p)
                S.sync()
                sec = time.perf_counter() - t0
                best = sec if best is None else min(best, sec)
            size = os.path.getsize(p)
            out[name] = {"bytes": size, "s": best, "GB_per_s": size / best / 1e9}
            print(f"[io] {name}: {size / 1e6:.0f} MB in {best:.3f} s", flush=True)
    # generate_rand_force on the device (k_max = 2): host modes + x pass, device y/z passes
    seeds = (1234, 5678, 9123)
    best = None
    for _ in range(reps):
        S.sync()
        t0 = time.perf_counter()
        seeds = S.generate_rand_force(k_max=2, seeds=seeds)
        S.sync()
        sec = time.perf_counter() - t0
        best = sec if best is None else min(best, sec)
    out["generate_rand_force"] = {"s": best, "k_max": 2}
    print(f"[io] generate_rand_force: {best * 1e3:.1f} ms", flush=True)
    # in-situ diagnostics of the analysis program: velocity spectrum, density PDF
    for name, fn in (("spectrum_velocity", lambda: S.spectrum("velocity")), ("density_pdf", lambda: S.density_pdf()),
                     ("mode_spectra", lambda: S.mode_spectra())):
        best = None
        for _ in range(reps):
            S.sync()
            t0 = time.perf_counter()
            fn()
            sec = time.perf_counter() - t0
            best = sec if best is None else min(best, sec)
        out[name] = {"s": best}
        print(f"[io] {name}: {best * 1e3:.1f} ms", flush=True)
    S.close()
    return out


def run_tvd_big(n, steps, cour, warmup=2):
    """Forced-turbulence TVD workload at n^3: K98 Roe-TVD + CTU + CT with OU forcing
    (solenoidal, 1 <= |k| <= 2, t_corr 0.5, sigma 5, seed 3003) and two passive scalars, IC
    ics.config_c3; FP32 and FP64, whole step in a hipGraph."""
    from isothermal_mhd_amd import TvdSolver, ics
    q, bf, qp = ics.config_c3(n=n)
    out = {"n": n, "steps": steps}
    for prec in (32, 64):
        S = TvdSolver((n, n, n), npass=2, riemann="roe", bc="periodic", precision=prec, cour=cour)
        S.set_state(q, bf, qp)
        S.set_ou_forcing(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=3003, solenoidal=True)
        S.boundary_conditions()
        S.run(warmup, graph=True)
        S.sync()
        t0 = time.perf_counter()
        S.run(steps, graph=True)
        S.sync()
        sec = time.perf_counter() - t0
        out[f"fp{prec}"] = {"value": n ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s",
                            "ms_per_step": sec / steps * 1e3, "device_GB": S.device_bytes() / 1e9}
        print(f"[tvd] {n}^3 fp{prec}: {out[f'fp{prec}']['value']:.0f} Mcell/s", flush=True)
        S.close()
    out["workload"] = (f"{n}^3 periodic, rho=1, B0=(0.2,0,0), v=0, OU solenoidal forcing 1<=|k|<=2 "
                       "(t_corr 0.5, sigma 5, seed 3003), 2 passives (sin 2 pi x, jittered step); "
                       "K98 Roe-TVD + CTU + CT, hipGraph native loop")
    return out


def run_tvd_c1(steps, cour):
    """64^3 periodic, rho = 1, B0 = (0.5,0,0), k^-2 velocity 1<=|k|<=4 at rms Mach 1 (seed 1001),
    K98 Roe-TVD + CTU + CT, FP64, native loop in a hipGraph."""
    from isothermal_mhd_amd import TvdSolver, ics
    n = 64
    q, bf, qp = ics.config_c1(n=n)
    S = TvdSolver((n, n, n), npass=0, riemann="roe", precision=64, cour=cour)
    S.set_state(q, bf, None)
    S.boundary_conditions()
    S.run(2, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    sec = time.perf_counter() - t0
    S.close()
    return {"value": n ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "steps": steps,
            "ms_per_step": sec / steps * 1e3, "dtype": "f64",
            "workload": "64^3 periodic, rho=1, B0=(0.5,0,0), k^-2 velocity 1<=|k|<=4 rms Mach 1 (seed 1001); "
                        "K98 Roe-TVD + CTU + CT, FP64, hipGraph native loop"}


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


def l1_fields(pairs, nb=6, npass=0):
    """Relative L1 of the FP32 run against the FP64 run, sum|a - b| / sum|b| over the
    interiors of all boxes, per conserved field, cell-centred B and passive scalar; pairs is
    a list of ((q, bf, qp) FP32, (q, bf, qp) FP64) per box."""
    import numpy as np
    names = ["rho", "mx", "my", "mz", "bx", "by", "bz"] + [f"passive{p}" for p in range(npass)]
    num = dict.fromkeys(names, 0.0)
    den = dict.fromkeys(names, 0.0)
    I = (slice(nb, -nb),) * 3
    Ip = (slice(nb - 2, -(nb - 2)),) * 3
    for (qa, _, pa), (qb, _, pb) in pairs:
        for f, nm in enumerate(names[:7]):
            a, b = qa[I + (f,)], qb[I + (f,)]
            num[nm] += float(np.abs(a - b).sum())
            den[nm] += float(np.abs(b).sum())
        for p in range(npass):
            a, b = pa[Ip + (p,)], pb[Ip + (p,)]
            num[f"passive{p}"] += float(np.abs(a - b).sum())
            den[f"passive{p}"] += float(np.abs(b).sum())
    return {nm: (num[nm] / den[nm] if den[nm] > 0 else None) for nm in names}   # None: field identically 0


def run_l1_configs(cour):
    """FP32 against FP64 from the same IC, per configuration: the forced TVD workload at
    512^3 (500 steps), the 48-box ensemble (200 steps) and the 1024^3 physics (lognormal
    density, k^-5/3 velocity at Mach 3, HLLD + PLM + OU) at 512^3 (200 steps; an FP64 run
    of 1024^3 does not fit on one GPU).  Reports the relative L1 per field and max|div B|.
    For the forced workload it also measures how far two FP64 runs drift apart when one
    starts from rho perturbed by 1e-12, for the TVD and the VL integrator at 256^3: the
    FP32 error is only meaningful next to that sensitivity of the method itself."""
    import numpy as np
    from isothermal_mhd_amd import TvdSolver, VanLeerSolver, ics
    res = {}

    def advance(S, steps, tag):
        done = 0
        while done < steps:
            k = min(100, steps - done)
            S.run(k, graph=True)
            S.sync()
            done += k
            print(f"[l1] {tag}: {done}/{steps} steps", flush=True)

    n = 512
    q, bf, qp = ics.config_c3(n=n)
    runs = {}
    for prec in (32, 64):
        S = TvdSolver((n, n, n), npass=2, riemann="roe", bc="periodic", precision=prec, cour=cour)
        S.set_state(q, bf, qp)
        S.set_ou_forcing(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=3003, solenoidal=True)
        S.boundary_conditions()
        advance(S, 500, f"c3 fp{prec}")
        runs[prec] = (S.get_state(), S.diagnostics()["divb_max"], float(S.get_time()[0][0]))
        S.close()
    del q, bf, qp
    res["c3_forced_tvd_512"] = {
        "workload": "512^3, K98 Roe-TVD + CTU + CT, OU forcing (1<=|k|<=2, t_corr 0.5, sigma 5, seed 3003), "
                    "2 passives, 500 steps", "l1": l1_fields([(runs[32][0], runs[64][0])], npass=2),
        "max_divb_fp32": runs[32][1], "max_divb_fp64": runs[64][1], "t_fp32": runs[32][2], "t_fp64": runs[64][2]}
    del runs

    dims, states, _ = ics.ensemble_c4()
    runs = {}
    for prec in (32, 64):
        S = VanLeerSolver(dims, npass=0, riemann="hlle", precision=prec, cour=cour)
        for b, (q, bf, qp) in enumerate(states):
            S.set_state(q, bf, None, box=b)
        S.boundary_conditions()
        advance(S, 200, f"c4 fp{prec}")
        diag = S.diagnostics()
        runs[prec] = ([S.get_state(box=b) for b in range(len(dims))], max(d["divb_max"] for d in diag))
        S.close()
    del states
    res["c4_ensemble"] = {
        "workload": "48 boxes U{32..192}^3 (seed 4004), HLLE + VL2 + CT, PLM, 200 steps",
        "l1": l1_fields(list(zip(runs[32][0], runs[64][0]))),
        "max_divb_fp32": runs[32][1], "max_divb_fp64": runs[64][1]}
    del runs

    # sensitivity of the forced workload to its initial state, per integrator: FP64 against
    # FP64 started from rho perturbed by 1e-12 (relative, seeded), next to FP32 against FP64
    n = 256
    q, bf, qp = ics.config_c3(n=n)
    qpert = q.copy(order="F")
    I = (slice(6, -6),) * 3
    qpert[I + (0,)] *= 1.0 + 1e-12 * np.random.default_rng(7).standard_normal(qpert[I + (0,)].shape)
    sens = {}
    for integ in ("tvd", "vl"):
        st = {}
        for tag, prec, qq in (("fp64", 64, q), ("fp64_perturbed", 64, qpert), ("fp32", 32, q)):
            if integ == "tvd":
                S = TvdSolver((n, n, n), npass=2, riemann="roe", bc="periodic", precision=prec, cour=cour)
            else:
                S = VanLeerSolver((n, n, n), npass=2, riemann="hlld", recon="plm", bc="periodic", precision=prec,
                                  cour=cour)
            S.set_state(qq, bf, qp)
            S.set_ou_forcing(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=3003, solenoidal=True)
            S.boundary_conditions()
            st[tag] = []
            for k in (10, 90, 400):
                advance(S, k, f"sensitivity {integ} {tag}")
                st[tag].append(S.get_state())
            S.close()
        sens[integ] = {f"after_{s_}_steps": {
            "fp64_perturbed_vs_fp64": l1_fields([(st["fp64_perturbed"][m], st["fp64"][m])], npass=2),
            "fp32_vs_fp64": l1_fields([(st["fp32"][m], st["fp64"][m])], npass=2)}
            for m, s_ in enumerate((10, 100, 500))}
        del st
    del q, bf, qp, qpert
    res["c3_sensitivity_256"] = {
        "workload": "256^3 forced workload of c3 (OU 1<=|k|<=2, t_corr 0.5, sigma 5, 2 passives), K98 Roe-TVD "
                    "and, for comparison, HLLD + VL2 + PLM; FP64 from rho * (1 + 1e-12 g) against FP64",
        "l1": sens}

    n = 512
    q, bf, qp = ics.config_c5(n=n)
    runs = {}
    for prec in (32, 64):
        S = VanLeerSolver((n, n, n), npass=0, riemann="hlld", recon="plm", bc="periodic", precision=prec,
                          cour=cour)
        S.set_state(q, bf, None)
        S.set_ou_forcing(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=5005, solenoidal=True)
        S.boundary_conditions()
        advance(S, 200, f"c5 physics at 512^3 fp{prec}")
        runs[prec] = (S.get_state(), S.diagnostics()["divb_max"], float(S.get_time()[0][0]))
        S.close()
    del q, bf, qp
    res["c5_physics_512"] = {
        "workload": "512^3, rho = 1 +-10% lognormal, B0=(0.05,0.05,0), k^-5/3 velocity 1<=|k|<=16 rms Mach 3 "
                    "(seed 5005), HLLD + VL2 + CT, PLM, OU forcing (1<=|k|<=2, t_corr 0.5, sigma 5), 200 steps",
        "l1": l1_fields([(runs[32][0], runs[64][0])]),
        "max_divb_fp32": runs[32][1], "max_divb_fp64": runs[64][1], "t_fp32": runs[32][2], "t_fp64": runs[64][2]}
    return res


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


def _cpu_reference_run(n_cpu, steps, cour):
    """Body of cpu_reference, run in a child process (the reference's modules PRINT to
    standard output)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from isothermal_mhd_amd import ics
    from refutil import Reference
    q, bf, qp = ics.config_c2(n=n_cpu)
    R = Reference(f"vl_hlld_p{n_cpu}")
    R.set_state(q, bf, qp)
    R.bc()
    dt = 1.0e-15
    t0 = time.perf_counter()
    for _ in range(steps):
        R.copy_velocity(0)
        R.step(dt)
        R.copy_velocity(1)
        R.passive_step(dt)
        R.bc()
        dt = cour * (1.0 / n_cpu) / R.max_speed()
    return time.perf_counter() - t0


def cpu_reference(n_cpu, steps, cour):
    """The reference's own solver on the host, beside the port: its vanleer_mod /
    passive_mod / boundary_conditions_mod compiled as they lie with flang -O3 for an
    n_cpu^3 periodic HLLD box with one passive (oracle/ref/build_ref.py, config
    vl_hlld_p128; present only where that build was made), one thread, the driver's
    step order on the same IC generator.  As compiled the reference reconstructs with
    characteristic PPM, its default; the PLM branch of the benchmark line is unreachable
    in its build.  Runs in a child process; returns None without the build."""
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", f"libimhd_ref_vl_hlld_p{n_cpu}.so")):
        return None
    code = (f"import sys; sys.path.insert(0, {ROOT!r}); import bench; "
            f"print('SECONDS', bench._cpu_reference_run({n_cpu}, {steps}, {cour}))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=600)
    lines = [l for l in r.stdout.splitlines() if l.startswith("SECONDS ")]
    if r.returncode != 0 or not lines:
        return {"error": (r.stderr or r.stdout)[-300:]}
    sec = float(lines[-1].split()[1])
    return {"value": n_cpu ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"{steps} driver steps at {n_cpu}^3 (same IC generator) of the reference's solver modules "
                      f"built with flang -O3 (vanleer_solve_3d with HLLD and its default characteristic PPM, "
                      f"passive_solve_3d, boundary_conditions), FP64 single-threaded, {sec:.1f} s"}


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
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--pmc-run", action="store_true", help="short run for rocprofv3 counter collection")
    ap.add_argument("--no-extra", action="store_true", help="skip the ensemble and TVD side measurements")
    ap.add_argument("--l1", action="store_true",
                    help="only the FP32-vs-FP64 L1 errors of the forced TVD, ensemble and 1024^3-physics configurations")
    ap.add_argument("--extra-steps", type=int, default=50)
    ap.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--sq-summarize", metavar="SQ_DIR")
    ap.add_argument("--big", metavar="N[,N]", help="only the large-box capacity lines, e.g. 512,1024")
    ap.add_argument("--io", type=int, metavar="N", help="only the restart/snapshot file timing of an N^3 box")
    ap.add_argument("--tvd", type=int, metavar="N", help="only the forced K98 Roe-TVD workload at N^3 (FP32, FP64)")
    args = ap.parse_args()
    if args.pmc_summarize:
        return pmc_summarize(*args.pmc_summarize)
    if args.sq_summarize:
        return sq_summarize(args.sq_summarize)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the box would shard by 3-D blocks with an RCCL halo exchange, not implemented)",
              file=sys.stderr)
        return 2
    if args.pmc_run:
        return pmc_run(args.n)
    if args.tvd:
        print(json.dumps({"tvd": run_tvd_big(args.tvd, args.steps if args.steps != 200 else 20, args.cour)}),
              flush=True)
        return 0
    if args.io:
        print(json.dumps({"io": run_io(args.io)}), flush=True)
        return 0
    if args.l1:
        print(json.dumps({"l1": run_l1_configs(args.cour)}), flush=True)
        return 0
    if args.big:
        lines = []
        for nb_ in args.big.split(","):
            lines.append(run_big(int(nb_), args.steps, args.cour))
            print(json.dumps({"big": lines[-1]}), flush=True)
        return 0
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
        "graph": {"value": n ** 3 * K / r32["graph_sec"] / 1e6, "ms_per_step": r32["graph_sec"] / K * 1e3,
                  "note": "same K FP32 steps replayed from a captured hipGraph, no per-kernel events"},
        "fp32": {"ms_per_step": r32["sec"] / K * 1e3, "t_end": r32["t"], "autotune": r32["autotune"],
                 "max_divb": r32["diag"]["divb_max"], "floor_resets": r32["floors"]},
        "roofline": roofline(r32["prof"], 4),
        "kernels_fp32_ms": {k: round(v[0] / v[1], 5) for k, v in r32["prof"].items()},
    }
    if r64 is not None:
        cups64 = n ** 3 * K / r64["sec"] / 1e6
        res["fp64"] = {"value": cups64, "ms_per_step": r64["sec"] / K * 1e3, "t_end": r64["t"],
                       "graph_ms_per_step": r64["graph_sec"] / K * 1e3,
                       "max_divb": r64["diag"]["divb_max"], "floor_resets": r64["floors"],
                       "roofline": roofline(r64["prof"], 8)}
        res["fp32_over_fp64"] = cups32 / cups64
        res["l1_fp32_vs_fp64"] = l1_errors(r32["state"], r64["state"])
        res["kernels_fp64_ms"] = {k: round(v[0] / v[1], 5) for k, v in r64["prof"].items()}
    if not args.no_extra:
        for key, fn in (("ensemble_c4", run_ensemble), ("tvd_c1", run_tvd_c1),
                        ("big_512", lambda k, c: run_big(512, k // 2, c))):
            try:
                res[key] = fn(args.extra_steps, args.cour)
            except Exception as e:   # a side measurement never hides the main line
                res[key] = {"error": repr(e)}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps, args.cour)
        ref = cpu_reference(args.cpu_n, args.cpu_steps, args.cour)
        if ref is not None:
            res["cpu_reference"] = ref
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
