#!/usr/bin/env python3
"""Benchmark of the isothermal-MHD cell-update path on one MI355X.

Workload (the largest single-GPU configuration; one step = one pass of the driver loop body
over the box): 1024^3 periodic, rho = 1 with seeded +-10 % lognormal noise
(rho = exp(0.1 g - 0.005)), B0 = (0.05, 0.05, 0), seeded k^-5/3 velocity on 1 <= |k| <= 16
at rms Mach 3 (seed 5005), generated on the device (imhd_init_turbulence); VL2 + CT, HLLD,
primitive PLM (MC-beta 1.01), OU forcing (solenoidal, 1 <= |k| <= 2, t_corr 0.5, sigma 5,
seed 5005); step = VL2 predictor/corrector + CT, OU kick, ghost fill, CFL reduction and
device-side dt.  `value` is the FP32 path (FP32 bulk, FP64 face B / corner EMF / CFL /
diagnostics, Kahan-compensated rho, rho v); the FP64 path runs the same K steps from the
same start beside it, and the relative L1 of rho, rho v and the face field between the two
after K steps is reported with max|div B|.

Timing: W untimed warm-up steps, then the start is regenerated and exactly K steps are
timed (device synchronised on both sides).  Per-kernel durations come from HIP events
recorded on the solver's stream around every launch of the timed region.  The same K steps
replayed from one captured hipGraph (no events) are timed beside it ("graph").  Side lines:
the 256^3 configuration (HLLD, 1 passive; FP32 and FP64), the forced K98 Roe-TVD workload
at 512^3 (OU, 2 passives), the 48-box ensemble and the 64^3 K98 Roe-TVD box.

Prints ONE JSON line.

HBM traffic (roofline.traffic) comes from two rocprofv3 counter runs of `--pmc-run`
(FETCH_SIZE, then WRITE_SIZE), scaled by what the same counters read for a plain copy of
known size (k_convert_in<float> of an FP64 set_state) and checked against a read-only
reduction of known size (the diagnostics pass), reduced by `--pmc-summarize` into
profiles/pmc_traffic.json (`--pmc-map`: where the run records which launch is which
family, pmc_out/pmc_map.json by default):
  rocprofv3 --pmc FETCH_SIZE --output-format csv -d pmc_out/pmc_fetch -- python3 bench.py --pmc-run
  rocprofv3 --pmc WRITE_SIZE --output-format csv -d pmc_out/pmc_write -- python3 bench.py --pmc-run
  python3 bench.py --pmc-summarize pmc_out/pmc_fetch pmc_out/pmc_write
"""
import argparse
import csv
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PMC_MAP = os.path.join(ROOT, "pmc_out", "pmc_map.json")   # --pmc-run writes it, the summaries read it
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


def run_path(ics_state, n, prec, steps, warmup, cour, recon="plm"):
    from isothermal_mhd_amd import VanLeerSolver
    q, bf, qp = ics_state
    S = VanLeerSolver((n, n, n), npass=1, riemann="hlld", recon=recon, bc="periodic", precision=prec, cour=cour)
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


def run_c2(n, steps, warmup, cour, recon="plm"):
    """The 256^3 configuration (rho = 1, B0 = (0,0,1), solenoidal k^-2 velocity on
    1 <= |k| <= 8 at rms Mach 2, one passive scalar; HLLD, no forcing) with the given
    corrector reconstruction (plm; ppm_char: the reference's default), FP32 and FP64 from
    the same host-built start, with the L1 between them after K steps."""
    from isothermal_mhd_amd import ics
    state0 = ics.config_c2(n=n)
    r32 = run_path(state0, n, 32, steps, warmup, cour, recon)
    r64 = run_path(state0, n, 64, steps, warmup, cour, recon)
    v32, v64 = n ** 3 * steps / r32["sec"] / 1e6, n ** 3 * steps / r64["sec"] / 1e6
    return {"value": v32, "unit": "Mcell-updates/s", "steps": steps, "dtype": "f32",
            "ms_per_step": r32["sec"] / steps * 1e3, "graph_value": n ** 3 * steps / r32["graph_sec"] / 1e6,
            "fp64": {"value": v64, "ms_per_step": r64["sec"] / steps * 1e3,
                     "graph_value": n ** 3 * steps / r64["graph_sec"] / 1e6, "max_divb": r64["diag"]["divb_max"]},
            "fp32_over_fp64": v32 / v64, "max_divb": r32["diag"]["divb_max"], "autotune": r32["autotune"],
            "l1_fp32_vs_fp64": l1_errors(r32["state"], r64["state"]),
            "roofline": roofline(r32["prof"], 4, n ** 3, 1, r32["sec"] / steps * 1e3, f"c2_{n}"),
            "kernels_fp32_ms": {k: round(v[0] / v[1], 5) for k, v in r32["prof"].items()},
            "workload": f"{n}^3 periodic, VL2+CT, HLLD, "
                        + ("primitive PLM" if recon == "plm" else "characteristic PPM (the reference's default)")
                        + ", 1 passive scalar; rho=1, B0=(0,0,1), solenoidal k^-2 velocity 1<=|k|<=8 at rms Mach 2 "
                          "(seed 2002), passive U[0,1) (seed 2003)"}


# The 1024^3 workload: start generated on the device and OU forcing
C5_IC = dict(rho0=1.0, sigma_ln=0.1, b0=(0.05, 0.05, 0.0), kmin=1.0, kmax=16.0, slope=5.0 / 3.0, mach=3.0,
             seed=5005, solenoidal=False)
C5_OU = dict(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=5005, solenoidal=True)
C5_WORKLOAD = ("{n}^3 periodic isothermal MHD, VL2+CT, HLLD, primitive PLM (MC-beta 1.01), OU forcing "
               "(solenoidal, 1<=|k|<=2, t_corr 0.5, sigma 5, seed 5005); rho = exp(0.1 g - 0.005) per cell, "
               "B0=(0.05,0.05,0), k^-5/3 velocity 1<=|k|<=16 at rms Mach 3 (seed 5005), start generated on "
               "the device (imhd_init_turbulence)")
L1_FIELDS = ((0, "rho"), (1, "mx"), (2, "my"), (3, "mz"), (7, "bfx"), (8, "bfy"), (9, "bfz"))


def c5_start(S):
    """the workload's start: device-generated state, fresh OU amplitudes, t = 0, dt = 1e-15"""
    S.init_turbulence(**C5_IC)
    S.set_ou_forcing(**C5_OU)
    S.set_time(t=0.0, dt=1.0e-15)


def _slabs(nz, T=16):
    T = max(1, min(T, nz))
    return [(nz * t // T, nz * (t + 1) // T) for t in range(T)]


def _interior_f32(buf, n, nb=6):
    """the FP32 values of an FP32 context's padded field (exact), interior only, in threads
    over z-slabs (numpy releases the GIL)"""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    out = np.empty((n, n, n), dtype=np.float32, order="F")
    I = buf[nb:nb + n, nb:nb + n, nb:nb + n]

    def cp(sl):
        out[:, :, sl[0]:sl[1]] = I[:, :, sl[0]:sl[1]]
    with ThreadPoolExecutor(16) as ex:
        list(ex.map(cp, _slabs(n)))
    return out


def _l1_against(a32, buf, n, nb=6):
    """relative L1 and Linf of a (FP32 run, interior) against the padded FP64 field buf"""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    I = buf[nb:nb + n, nb:nb + n, nb:nb + n]

    def part(sl):
        a = a32[:, :, sl[0]:sl[1]].astype(np.float64)
        b = I[:, :, sl[0]:sl[1]]
        d = np.abs(a - b)
        return float(d.sum()), float(np.abs(b).sum()), float(d.max()), float(np.abs(b).max())
    with ThreadPoolExecutor(16) as ex:
        r = list(ex.map(part, _slabs(n)))
    num, den = sum(x[0] for x in r), sum(x[1] for x in r)
    dm, bm = max(x[2] for x in r), max(x[3] for x in r)
    return (num / den if den > 0 else None), (dm / bm if bm > 0 else None)


def run_c5(n, prec, steps, warmup, cour, ref=None, tune=True, graph=True):
    """One precision of the main workload: start on the device, autotune (FP32), W warm-up
    steps, then the start again and K timed steps with per-kernel events; the fields after
    K steps are kept (FP32, `ref` None) or compared with `ref` (FP64); then the same K steps
    from the start replayed from a captured hipGraph."""
    from isothermal_mhd_amd import VanLeerSolver
    t0 = time.perf_counter()
    S = VanLeerSolver((n, n, n), npass=0, riemann="hlld", recon="plm", bc="periodic", precision=prec, cour=cour)
    c5_start(S)
    S.sync()
    out = {"device_GB": S.device_bytes() / 1e9, "setup_s": time.perf_counter() - t0}
    print(f"[c5] {n}^3 fp{prec}: context {out['device_GB']:.1f} GB, start in {out['setup_s']:.1f} s", flush=True)
    if tune:
        t0 = time.perf_counter()
        rep = S.autotune(steps=2)
        out["autotune"] = rep.split("; ")[-1]
        print(f"[c5] autotune {time.perf_counter() - t0:.1f} s: {rep}", flush=True)
    if warmup:
        S.run(warmup, graph=False)
    c5_start(S)
    S.set_profiling(True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=False)
    S.sync()
    sec = time.perf_counter() - t0
    out["prof"] = S.profile()
    S.set_profiling(False)
    d = S.diagnostics()
    t, dt = S.get_time()
    out.update({"sec": sec, "value": n ** 3 * steps / sec / 1e6, "ms_per_step": sec / steps * 1e3,
                "max_divb": d["divb_max"], "t_end": float(t[0]), "floor_resets": S.floor_count(),
                "mass": d["mass"], "ke": d["ke"]})
    print(f"[c5] fp{prec}: {out['value']:.0f} Mcell-updates/s ({out['ms_per_step']:.2f} ms/step)", flush=True)
    buf = None
    fields, l1, linf = {}, {}, {}
    for f, nm in L1_FIELDS:
        buf = S.get_field(f, out=buf)
        if ref is None:
            fields[nm] = _interior_f32(buf, n)
        else:
            l1[nm], linf[nm] = _l1_against(ref[nm], buf, n)
    del buf
    if ref is None:
        out["fields"] = fields
    else:
        out["l1"], out["linf"] = l1, linf
    if graph:
        c5_start(S)
        S.run(2, graph=True)          # capture both buffer parities outside the timed region
        c5_start(S)
        S.sync()
        t0 = time.perf_counter()
        S.run(steps, graph=True)
        S.sync()
        out["graph_sec"] = time.perf_counter() - t0
    S.close()
    return out


def b_alg(esz, npass):
    """Algorithmic bytes per cell-update: one compulsory read and one write of rho, three
    rho v, the three face-field components (FP64 in both paths) and the passive scalars;
    the cell-centred field is derived from the face field.  FP32 cell fields: 80 + 8 npass;
    FP64: 112 + 16 npass."""
    return 2 * (4 * esz + 3 * 8 + npass * esz)


def roofline(prof, esz, cells, npass, step_ms, tag):
    """Roofline of the dominant kernel family (largest total event time): the algorithmic
    bytes of one cell-update (b_alg) times the cells one launch updates, over the average
    launch time, against MI355X's HBM peak; beside it the kernel's own compulsory bytes per
    work item (bytes_per_item), the whole step at b_alg, and the measured HBM traffic of the
    same workload (profiles/pmc_traffic.json, when it was measured on `tag`)."""
    fam = {}
    for name, (ms, nl, items) in prof.items():
        key = name[:-2] if name[-2:] in ("_x", "_y", "_z") else name
        f = fam.setdefault(key, [0.0, 0, 0])
        f[0] += ms
        f[1] += nl
        f[2] += items
    key = max(fam, key=lambda k: fam[k][0])
    ms, nl, items = fam[key]
    avg_ms = ms / nl
    ba = b_alg(esz, npass)
    ach = ba * cells / (avg_ms * 1e-3) / 1e9
    out = {"bound": "hbm", "kernel": key, "symbol": FAMILY_SYMBOL.get(key), "launches": nl,
           "avg_launch_ms": avg_ms, "cell_updates_per_launch": cells, "bytes_per_cell_update": ba,
           "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
           "whole_step": {"achieved": ba * cells / (step_ms * 1e-3) / 1e9,
                          "frac": ba * cells / (step_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}}
    bpi = bytes_per_item(key, esz, npass)
    if bpi is not None:
        a2 = bpi * (items / nl) / (avg_ms * 1e-3) / 1e9
        out["kernel_count"] = {"bytes_per_item": bpi, "items_per_launch": items / nl, "achieved": a2,
                               "frac": a2 / PEAK_HBM_GBS}
    pmc = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            tr = json.load(f)
        if tr.get("_workload") == tag and key in tr:
            out["traffic"] = tr[key]
            out["measured"] = {"achieved": tr[key] / (avg_ms * 1e-3) / 1e9,
                               "frac": tr[key] / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                               "traffic_over_algorithmic": tr[key] / (ba * cells)}
    # the fused stage kernels are VALU-bound: their VALU issue rate against the MI355X issue
    # peak, from the counted instructions per launch and this run's launch time
    sq = os.path.join(ROOT, "profiles", "sq_valu.json")
    if esz == 4 and os.path.exists(sq):
        with open(sq) as f:
            vq = json.load(f)
        if vq.get("_workload") == tag and key in vq:
            a3 = vq[key]["valu_wave_instr"] / (avg_ms * 1e-3) / 1e9
            out["valu_issue"] = {"achieved": a3, "peak": VALU_PEAK_GWI, "unit": "G wave-instr/s",
                                 "frac": a3 / VALU_PEAK_GWI, "wave_instr_per_launch": vq[key]["valu_wave_instr"]}
    return out


# family (bench profiler name) -> kernel symbol it launches (fused three-direction sweeps,
# z-marching column updates)
FAMILY_SYMBOL = {"stage_pred": "k_vl_stage", "stage_corr_cfl": "k_vl_stage",
                 "flux_pred": "k_flux3", "flux_corr": "k_flux3", "corner_pred": "k_corner",
                 "corner_corr": "k_corner", "update_pred": "k_update_z", "update_corr_cfl": "k_update_z",
                 "pass_stage_pred": "k_pass_stage_z", "pass_stage_corr": "k_pass_stage_z",
                 "fill_ghosts": "k_fill_ghost_band", "fill_ghosts_pass": "k_fill_ghost_band",
                 "stage_corr": "k_vl_stage", "ou_forcing": "k_ou_kick"}


# kernel dispatches of a family's symbol per launch of the family, where not one
FAMILY_DISPATCHES = {}


def pmc_run(n=1024, steps=3, n_cal=256, map_path=None):
    """Short FP32 run of the main workload for counter collection, then the kernels of known
    traffic that calibrate and check the counters, on an n_cal^3 box: the k_convert_in<float>
    dispatches of an FP64 set_state (8 bytes read and 4 written per padded cell, no reuse:
    the calibration) and the first diagnostics pass over its interior (read-only: rho and
    three rho v in FP32 and three face-field components in FP64, 40 bytes per interior cell:
    a check of the read calibration on a second kernel).  Writes the family -> kernel map
    for --pmc-summarize."""
    from isothermal_mhd_amd import VanLeerSolver, ics
    S = VanLeerSolver((n, n, n), npass=0, riemann="hlld", recon="plm", bc="periodic", precision=32)
    c5_start(S)
    S.set_profiling(True)
    S.run(steps, graph=False)
    prof = S.profile()
    S.close()
    fam = {}
    for name, (ms, nl, items) in prof.items():
        key = name[:-2] if name[-2:] in ("_x", "_y", "_z") else name
        if key in FAMILY_SYMBOL:
            fam[key] = {"symbol": FAMILY_SYMBOL[key], "order": len(fam), "launches": nl,
                        "per_launch": FAMILY_DISPATCHES.get(key, 1), "items": items // nl}
    q, bf, qp = ics.config_c2(n=n_cal)
    C = VanLeerSolver((n_cal,) * 3, npass=1, riemann="hlld", bc="periodic", precision=32)
    C.set_state(q, bf, qp)
    C.diagnostics()
    C.close()
    nc = (n_cal + 12) ** 3
    out = {"workload": f"c5_{n}", "n": n, "steps": steps, "families": fam,
           "copy": {"symbol": "k_convert_in<float>", "grid": (nc + 255) // 256 * 256,
                    "bytes_read": nc * 8, "bytes_written": nc * 4},
           "check": {"symbol": "k_diag_part<float>", "bytes_read": 40 * n_cal ** 3}}
    map_path = map_path or PMC_MAP
    os.makedirs(os.path.dirname(os.path.abspath(map_path)), exist_ok=True)
    with open(map_path, "w") as f:
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


def _family_values(rows, mp):
    """counter value per launch of each family: the last (launches x per_launch) dispatches
    of the family's kernel symbol (FP32 instantiation) are dealt to the families sharing
    that symbol in their launch order"""
    fams = sorted(mp["families"].items(), key=lambda kv: kv[1]["order"])
    out = {}
    for sym in {v["symbol"] for _, v in fams}:
        group = [(k, v) for k, v in fams if v["symbol"] == sym]
        cycle = []
        for k, v in group:
            cycle += [k] * v["per_launch"]
        launches = group[0][1]["launches"]
        vals = [val for n, val in rows if sym + "<float" in n][-launches * len(cycle):]
        acc = {k: 0.0 for k, _ in group}
        for m, val in enumerate(vals):
            acc[cycle[m % len(cycle)]] += val
        for k, v in group:
            out[k] = acc[k] / v["launches"] if vals else None
    return out


def pmc_summarize(fetch_dir, write_dir, map_path=None):
    with open(map_path or PMC_MAP) as f:
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
    ck = mp["check"]
    dv = [v for n, v in fr if ck["symbol"] in n]
    check = None
    if dv:
        meas = dv[-1] * 1024.0 / f_fac
        check = {"kernel": ck["symbol"], "bytes_read_known": ck["bytes_read"], "bytes_read_calibrated": meas,
                 "ratio": meas / ck["bytes_read"]}
    out = {"_workload": mp["workload"],
           "_calibration": {"kernel": sym, "bytes_read": cp["bytes_read"], "bytes_written": cp["bytes_written"],
                            "FETCH_SIZE_KiB": cf, "WRITE_SIZE_KiB": cw, "fetch_factor": f_fac, "write_factor": w_fac,
                            "check": check}}
    fv, wv = _family_values(fr, mp), _family_values(wr, mp)
    lines = ["| family | kernel | items / launch | FETCH KiB/launch | WRITE KiB/launch | HBM bytes/launch (calibrated) |",
             "|---|---|---|---|---|---|"]
    for key, v in sorted(mp["families"].items(), key=lambda kv: kv[1]["order"]):
        a, b = fv.get(key), wv.get(key)
        if a is None or b is None:
            continue
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


def sq_summarize(sq_dir, map_path=None):
    """VALU wave-instructions per launch of each family (SQ_INSTS_VALU from a rocprofv3 run
    of --pmc-run), matched to families as in pmc_summarize; written to
    profiles/sq_valu.json, from which bench.py reports each kernel's VALU issue rate."""
    with open(map_path or PMC_MAP) as f:
        mp = json.load(f)
    per = {c: _family_values(_read_counter(sq_dir, c), mp)
           for c in ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES", "SQ_WAVES")}
    out = {"_workload": mp["workload"], "_peak_G_wave_instr_per_s": VALU_PEAK_GWI}
    lines = ["| family | kernel | VALU wave-instr / launch | VALU instr / wave | active VALU / wave-cycles |",
             "|---|---|---|---|---|"]
    for key, v in sorted(mp["families"].items(), key=lambda kv: kv[1]["order"]):
        w = per["SQ_WAVES"].get(key)
        if not w:
            continue
        vi = per["SQ_INSTS_VALU"][key]
        out[key] = {"valu_wave_instr": vi, "waves": w,
                    "active_valu_per_wave_cycle": per["SQ_ACTIVE_INST_VALU"][key] / per["SQ_WAVE_CYCLES"][key]}
        lines.append(f"| {key} | {v['symbol']}<float> | {vi:.4g} | {vi / w:.0f} | "
                     f"{out[key]['active_valu_per_wave_cycle']:.3f} |")
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
    tune = S.autotune(steps=3)     # launch shapes (set-up, outside the timed region)
    S.run(warmup, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    sec = time.perf_counter() - t0
    cells = sum(a * b * c for a, b, c in dims)
    S.close()
    return {"value": cells * steps / sec / 1e6, "unit": "Mcell-updates/s", "boxes": len(dims), "cells": cells,
            "steps": steps, "ms_per_step": sec / steps * 1e3, "dtype": "f32", "autotune": tune.split("; ")[-1],
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


def run_tvd_big(n, steps, cour, warmup=2, precisions=(32, 64)):
    """Forced-turbulence TVD workload at n^3: K98 Roe-TVD + CTU + CT with OU forcing
    (solenoidal, 1 <= |k| <= 2, t_corr 0.5, sigma 5, seed 3003) and two passive scalars, IC
    ics.config_c3; FP32 and FP64, whole step in a hipGraph."""
    from isothermal_mhd_amd import TvdSolver, ics
    q, bf, qp = ics.config_c3(n=n)
    out = {"n": n, "steps": steps}
    for prec in precisions:
        S = TvdSolver((n, n, n), npass=2, riemann="roe", bc="periodic", precision=prec, cour=cour)
        S.set_state(q, bf, qp)
        S.set_ou_forcing(kmin=1.0, kmax=2.0, t_corr=0.5, sigma=5.0, seed=3003, solenoidal=True)
        S.boundary_conditions()
        tune = S.autotune(steps=3)     # launch shapes (set-up, outside the timed region)
        S.run(warmup, graph=True)
        S.sync()
        t0 = time.perf_counter()
        S.run(steps, graph=True)
        S.sync()
        sec = time.perf_counter() - t0
        out[f"fp{prec}"] = {"value": n ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s",
                            "ms_per_step": sec / steps * 1e3, "device_GB": S.device_bytes() / 1e9,
                            "autotune": tune.split("; ")[-1]}
        print(f"[tvd] {n}^3 fp{prec}: {out[f'fp{prec}']['value']:.0f} Mcell/s", flush=True)
        S.close()
    out["workload"] = (f"{n}^3 periodic, rho=1, B0=(0.2,0,0), v=0, OU solenoidal forcing 1<=|k|<=2 "
                       "(t_corr 0.5, sigma 5, seed 3003), 2 passives (sin 2 pi x, jittered step); "
                       "K98 Roe-TVD + CTU + CT, hipGraph native loop")
    return out


def run_tvd_c1(steps, cour):
    """64^3 periodic, rho = 1, B0 = (0.5,0,0), k^-2 velocity 1<=|k|<=4 at rms Mach 1 (seed 1001),
    K98 Roe-TVD + CTU + CT, FP64, native loop in a hipGraph, at the configuration's CFL 0.8."""
    from isothermal_mhd_amd import TvdSolver, ics
    n = 64
    q, bf, qp = ics.config_c1(n=n)
    S = TvdSolver((n, n, n), npass=0, riemann="roe", precision=64, cour=cour)
    S.set_state(q, bf, None)
    S.boundary_conditions()
    tune = S.autotune(steps=3)     # launch shapes (set-up, outside the timed region)
    S.run(2, graph=True)
    S.sync()
    t0 = time.perf_counter()
    S.run(steps, graph=True)
    S.sync()
    sec = time.perf_counter() - t0
    S.close()
    return {"value": n ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "steps": steps,
            "ms_per_step": sec / steps * 1e3, "dtype": "f64", "cfl": cour, "autotune": tune.split("; ")[-1],
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


def run_l1_configs(cour, which=("c3", "c4", "sens", "c5")):
    """FP32 against FP64 from the same start, per configuration: the forced TVD workload at
    512^3 (500 steps), the 48-box ensemble (200 steps) and the main workload at 1024^3
    (200 steps).  Reports the relative L1 per field and max|div B|.  For the forced
    workload it also measures how far two FP64 runs drift apart when one starts from rho
    perturbed by 1e-12, for the TVD and the VL integrator at 256^3: the FP32 error is only
    meaningful next to that sensitivity of the method itself."""
    res = {}

    def advance(S, steps, tag):
        done = 0
        while done < steps:
            k = min(100, steps - done)
            S.run(k, graph=True)
            S.sync()
            done += k
            print(f"[l1] {tag}: {done}/{steps} steps", flush=True)

    if "c3" in which:
        res.update(_l1_c3(cour, advance))
    if "c4" in which:
        res.update(_l1_c4(cour, advance))
    if "sens" in which:
        res.update(_l1_sensitivity(cour, advance))
    return _l1_c5(res, which, cour)


def _l1_c3(cour, advance):
    from isothermal_mhd_amd import TvdSolver, ics
    res = {}
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
    return res


def _l1_c4(cour, advance):
    from isothermal_mhd_amd import VanLeerSolver, ics
    res = {}
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
    return res


def _l1_sensitivity(cour, advance):
    import numpy as np
    from isothermal_mhd_amd import TvdSolver, VanLeerSolver, ics
    res = {}
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
    return res


def _l1_c5(res, which, cour):
    if "c5" in which:
        # the main workload itself at 1024^3: FP32 then FP64 from the same device-generated
        # start, 200 steps each (the FP64 context holds 267 GB, alone on the GPU)
        n = 1024
        r32 = run_c5(n, 32, 200, 0, cour, tune=False, graph=False)
        ref = r32.pop("fields")
        r64 = run_c5(n, 64, 200, 0, cour, ref=ref, tune=False, graph=False)
        del ref
        res["c5_1024"] = {
            "workload": C5_WORKLOAD.format(n=n) + ", 200 steps", "l1": r64["l1"], "linf": r64["linf"],
            "max_divb_fp32": r32["max_divb"], "max_divb_fp64": r64["max_divb"], "t_fp32": r32["t_end"],
            "t_fp64": r64["t_end"], "value_fp32": r32["value"], "value_fp64": r64["value"],
            "device_GB_fp32": r32["device_GB"], "device_GB_fp64": r64["device_GB"]}
    return res


def cpu_workers():
    """worker processes for the CPU legs: the host cores this process may run on, at most 16
    (the CPU share of one GPU on the benchmark hosts)"""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def _cpu_worker(job):
    """One worker of the CPU baseline: the oracle (FP64, one thread) on its own n^3 box of
    the main workload (start read from the parent's file), OU forcing, `steps` native-loop
    steps in the library's order (step, OU kick, CFL dt, ghost fill)."""
    path, n, steps, cour = job
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    from refutil import Oracle
    z = np.load(path)
    O = Oracle(n, n, n, npass=0, mhd=2, riemann=2, recon=0, bc=2)
    O.set_state(z["q"], z["bf"], None)
    O.bc()
    O.ou_setup(**C5_OU)
    dt = 1.0e-15
    t0 = time.perf_counter()
    for _ in range(steps):
        O.step(dt)
        O.ou_step(dt)
        dt = cour * (1.0 / n) / O.max_speed()
        O.bc()
    return time.perf_counter() - t0


def cpu_baseline(n_cpu, steps, cour):
    """The CPU restatement (oracle/imhd_oracle.c, FP64) on a bounded sample of the main
    workload, on every host core this process may use: one single-threaded process per core,
    each advancing its own n_cpu^3 box of the workload (the library's device-generated start
    at n_cpu^3, OU forcing) by `steps` steps, all at once.  Throughput = all processes'
    cell-updates over the slowest process's loop time."""
    import multiprocessing as mp
    import tempfile
    import numpy as np
    from isothermal_mhd_amd import VanLeerSolver
    S = VanLeerSolver((n_cpu,) * 3, npass=0, riemann="hlld", recon="plm", bc="periodic", precision=64, cour=cour)
    S.init_turbulence(**C5_IC)
    q, bf, _ = S.get_state()
    S.close()
    P = cpu_workers()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "start.npz")
        np.savez(path, q=q, bf=bf)
        with mp.get_context("spawn").Pool(P) as pool:
            secs = pool.map(_cpu_worker, [(path, n_cpu, steps, cour)] * P)
    sec = max(secs)
    return {"value": P * n_cpu ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "cores": P, "kind": "port",
            "sample": f"{P} concurrent single-threaded processes, one per host core, each {steps} steps of the "
                      f"main workload on its own {n_cpu}^3 box (same device-generated start, HLLD + PLM, OU "
                      f"forcing) with oracle/imhd_oracle.c FP64; slowest process {sec:.1f} s, fastest "
                      f"{min(secs):.1f} s"}


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
    vl_hlld_p128; present only where that build was made), one single-threaded process per
    host core, each on its own box, the driver's step order on the 256^3 configuration's
    start generator.  As compiled the reference reconstructs with characteristic PPM, its
    default (the PLM branch is unreachable in its build), and it has no OU forcing.  Its
    MPI driver cannot start in this image (MPI_SIZEOF, mpi_domain_mod.f90:515), so the
    cores run independent single-rank boxes instead of one decomposed box.  Returns None
    without the build."""
    import subprocess
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", f"libimhd_ref_vl_hlld_p{n_cpu}.so")):
        return None
    code = (f"import sys; sys.path.insert(0, {ROOT!r}); import bench; "
            f"print('SECONDS', bench._cpu_reference_run({n_cpu}, {steps}, {cour}))")
    P = cpu_workers()
    procs = [subprocess.Popen([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                              text=True, cwd=ROOT) for _ in range(P)]
    secs = []
    for pr in procs:
        out, err = pr.communicate(timeout=600)
        lines = [l for l in out.splitlines() if l.startswith("SECONDS ")]
        if pr.returncode != 0 or not lines:
            return {"error": (err or out)[-300:]}
        secs.append(float(lines[-1].split()[1]))
    sec = max(secs)
    return {"value": P * n_cpu ** 3 * steps / sec / 1e6, "unit": "Mcell-updates/s", "cores": P,
            "kind": "reference",
            "sample": f"{P} concurrent single-threaded processes, one per host core, each {steps} driver steps at "
                      f"{n_cpu}^3 of the reference's solver modules built with flang -O3 (vanleer_solve_3d with HLLD "
                      f"and its default characteristic PPM, passive_solve_3d, boundary_conditions; 256^3 "
                      f"configuration's start generator at {n_cpu}^3), FP64; slowest process {sec:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--cour", type=float, default=0.35)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner (profiling runs)")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--pmc-run", action="store_true", help="short run for rocprofv3 counter collection")
    ap.add_argument("--no-extra", action="store_true", help="skip the ensemble and TVD side measurements")
    ap.add_argument("--l1", nargs="?", const="c3,c4,sens,c5", metavar="c3,c4,sens,c5",
                    help="only the FP32-vs-FP64 L1 errors of the forced TVD (c3), ensemble (c4), "
                         "sensitivity (sens) and 1024^3 main-workload (c5) configurations")
    ap.add_argument("--extra-steps", type=int, default=50)
    ap.add_argument("--pmc-summarize", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"))
    ap.add_argument("--sq-summarize", metavar="SQ_DIR")
    ap.add_argument("--pmc-map", metavar="PATH", help="launch-to-family map of --pmc-run (default pmc_out/pmc_map.json)")
    ap.add_argument("--big", metavar="N[,N]", help="only the large-box capacity lines, e.g. 512,1024")
    ap.add_argument("--io", type=int, metavar="N", help="only the restart/snapshot file timing of an N^3 box")
    ap.add_argument("--tvd", type=int, metavar="N", help="only the forced K98 Roe-TVD workload at N^3 (FP32, FP64)")
    args = ap.parse_args()
    if args.pmc_summarize:
        return pmc_summarize(*args.pmc_summarize, map_path=args.pmc_map)
    if args.sq_summarize:
        return sq_summarize(args.sq_summarize, map_path=args.pmc_map)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the box would shard by 3-D blocks with an RCCL halo exchange, not implemented)",
              file=sys.stderr)
        return 2
    if args.pmc_run:
        return pmc_run(args.n, map_path=args.pmc_map)
    if args.tvd:
        print(json.dumps({"tvd": run_tvd_big(args.tvd, args.steps if args.steps != 200 else 20, args.cour)}),
              flush=True)
        return 0
    if args.io:
        print(json.dumps({"io": run_io(args.io)}), flush=True)
        return 0
    if args.l1:
        print(json.dumps({"l1": run_l1_configs(args.cour, tuple(args.l1.split(",")))}), flush=True)
        return 0
    if args.big:
        lines = []
        for nb_ in args.big.split(","):
            lines.append(run_big(int(nb_), args.steps, args.cour))
            print(json.dumps({"big": lines[-1]}), flush=True)
        return 0
    n, K, W = args.n, args.steps, args.warmup
    tag = f"c5_{n}"
    r32 = run_c5(n, 32, K, W, args.cour, tune=not args.no_tune)
    ref = r32.pop("fields")
    res = {
        "metric": METRIC,
        "value": r32["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": C5_WORKLOAD.format(n=n),
            "cells": n ** 3, "cfl": args.cour,
            "precision": "FP32 bulk; FP64 face B and its CT update, CFL reduction, diagnostics; "
                         "Kahan-compensated rho and rho v",
            "step": "VL2 predictor/corrector + CT, OU kick with the CFL speed of the kicked cells, "
                    "ghost fill, device-side dt",
        },
        "graph": {"value": n ** 3 * K / r32["graph_sec"] / 1e6, "ms_per_step": r32["graph_sec"] / K * 1e3,
                  "note": "same K FP32 steps from the same start replayed from a captured hipGraph, no per-kernel "
                          "events"},
        "fp32": {k: r32[k] for k in ("ms_per_step", "t_end", "max_divb", "floor_resets", "device_GB", "mass", "ke",
                                     "autotune", "setup_s") if k in r32},
        "roofline": roofline(r32["prof"], 4, n ** 3, 0, r32["ms_per_step"], tag),
        "kernels_fp32_ms": {k: round(v[0] / v[1], 5) for k, v in r32["prof"].items()},
    }
    if not args.no_fp64:
        try:
            r64 = run_c5(n, 64, K, W, args.cour, ref=ref, tune=n < 768, graph=False)
            res["fp64"] = {"value": r64["value"], "roofline": roofline(r64["prof"], 8, n ** 3, 0, r64["ms_per_step"],
                                                                       tag + "_fp64")}
            res["fp64"].update({k: r64[k] for k in ("ms_per_step", "t_end", "max_divb", "floor_resets", "device_GB",
                                                    "mass", "ke", "setup_s") if k in r64})
            res["fp32_over_fp64"] = r32["value"] / r64["value"]
            res["l1_fp32_vs_fp64"] = {"after_steps": K, "fields": "relative L1 over the interior, "
                                      "sum|fp32 - fp64| / sum|fp64|; face field components bfx, bfy, bfz",
                                      **r64["l1"], "linf": r64["linf"]}
            res["kernels_fp64_ms"] = {k: round(v[0] / v[1], 5) for k, v in r64["prof"].items()}
        except Exception as e:    # the FP64 context must fit beside nothing else; say so if not
            res["fp64"] = {"error": repr(e)}
    del ref
    if not args.no_extra:
        for key, fn in (("c2_256", lambda: run_c2(256, K, W, args.cour)),
                        ("c2_256_ppm_char", lambda: run_c2(256, K, W, args.cour, recon="ppm_char")),
                        ("tvd_c3_512", lambda: run_tvd_big(512, K, args.cour, precisions=(32,))),
                        ("ensemble_c4", lambda: run_ensemble(args.extra_steps, args.cour)),
                        ("tvd_c1", lambda: run_tvd_c1(args.extra_steps, 0.8))):
            try:
                res[key] = fn()
            except Exception as e:   # a side measurement never hides the main line
                res[key] = {"error": repr(e)}
            print(f"[side] {key} done", flush=True)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps, args.cour)
        ref_cpu = cpu_reference(args.cpu_n, args.cpu_steps, args.cour)
        if ref_cpu is not None:
            res["cpu_reference"] = ref_cpu
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
