"""Benchmark of the SMAUG per-step MHD update on one MI355X.

Workload: the 3-D gravitationally stratified atmosphere, 256^3 interior
cells over 6 x 4 x 4 Mm (vertical axis first), FP32-perturbation /
FP64-background mixed precision, classical RK4 with a fixed dt (the smaller
of CFL 0.3 and the hyperviscous limit of the initial state), 1000 timed steps
by default.  A step is one RK4 step of the whole grid (per stage: the aux/cmax,
hyperviscosity-ratio, hyperdiffusive-source and update passes, and the ghost
fill), replayed from a captured hipGraph; inputs are resident in HBM.

Printed JSON line (rank 0): the contract fields plus
  fp64        : the plain FP64 GPU path on the same workload and steps
  accuracy    : per-field relative L2 / Linf of the mixed result against the
                FP64 GPU result after the workload's 1000 RK4 steps (a leg
                of its own, whatever --steps says), and the FP64 path's
                change under a 1e-9 change of dt and under one rounding of
                its state to FP32 after the first step
  roofline    : the longest kernel of a stage against HBM peak, with the
                mean launch time and profiled HBM traffic of every pass
  cpu_baseline: the CPU oracle (OpenMP) on a bounded sample of the workload
  orszag_tang : the 1024^2 Orszag-Tang vortex, 500 RK4 steps at dt = 1e-4,
                mixed and FP64 paths and the mixed error against FP64
  ensemble    : the batched 2-D ensemble (64 Orszag-Tang-type periodic runs,
                sides drawn from {96,...,512}, 200 RK4 steps, mixed precision)
                advanced by the persistent queue-driven kernels, and the same
                runs stepped one grid at a time for comparison
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "smaug-all_amd"))

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per cell per RK4 step (low-storage RK4: 72 + 4*N_B
# reads and 56 writes of field values, N_B = 4 background fields):
#   FP64  (128 + 16) * 8        = 1152
#   mixed 128 * 4 + 16 * 8      = 640
BYTES_PER_CELL_STEP = {"mixed": 640.0, "fp64": 1152.0}
# 2-D: 6 evolved fields (16 accesses each per step) and N_B = 3 background
# fields read in each of the 4 stages: mixed 96 * 4 + 12 * 8 = 480
BYTES_PER_CELL_STEP_2D_MIXED = 480.0
STAGES_PER_STEP = 4
# forms of the 3-D update pass (smaug_set_zstream / smaug_get_zstream)
FORMS = {0: "per-cell", 1: "z-streamed update, z-streamed sources",
         2: "z-streamed update, per-cell sources", 3: "z-streamed update, LDS-plane sources"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="timed RK4 steps (1000: the stratified workload's step count, after "
                         "which the mixed-vs-FP64 error is reported)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="interior cells per axis")
    ap.add_argument("--large", action="store_true",
                    help="the large stratified atmosphere: 768 x 512 x 512 interior cells over "
                         "12 x 8 x 8 Mm (vertical first) instead of n^3 over 6 x 4 x 4 Mm")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-ot", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="fixed 3-D update form, z-stream depth and (optionally) ratio z-stream "
                         "depth and ratio form 'FORM,KZ[,KZR[,RFORM]]' instead of the autotuner "
                         "(FORM as in smaug_set_zstream: 0 per-cell, 1 z-streamed, 3 z-streamed "
                         "with LDS-plane sources; RFORM as in smaug_set_ratio_form: 1 every pair "
                         "in the ratio pass, 2 the in-plane pairs in the source kernel), e.g. for "
                         "profiling runs")
    ap.add_argument("--ens-steps", type=int, default=200)
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="RK4 steps of the mixed-vs-FP64 accuracy leg (the stratified "
                         "workload's horizon)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-ens-compare", action="store_true",
                    help="skip stepping the ensemble's runs one grid at a time for comparison")
    ap.add_argument("--ot-steps", type=int, default=500,
                    help="RK4 steps of the Orszag-Tang leg (500 is the workload's count)")
    return ap.parse_args()


def run_path(setup, precision, warmup, steps, dt_scale=1.0, tune=True, shape=None):
    """time `steps` RK4 steps after `warmup`; the launch shapes (XCD panels,
    z-streamed or per-cell update, z-stream depth) are picked by the
    engine's autotuner first"""
    from smaugamd.engine import Engine

    dt = setup.dt * dt_scale
    e = Engine(setup, precision)
    e.set_state()
    info = {}
    if shape is not None:
        vals = [int(x) for x in shape.split(",")]
        form, kz = vals[0], vals[1]
        e.set_zstream(form, kz)
        if len(vals) > 2:
            e.set_ratio_form(vals[3] if len(vals) > 3 else 1, vals[2])
        info["tuned"] = {"update_form": FORMS[form], "kz": kz,
                         "ratio_form": e.get_ratio_form()[0],
                         "ratio_kz": e.get_ratio_form()[1], "fixed": True}
    elif tune:
        best, ms = e.autotune(3)
        form, kz = e.get_zstream()
        info["tuned"] = {"update_form": FORMS[form], "kz": kz,
                         "ratio_form": e.get_ratio_form()[0],
                         "ratio_kz": e.get_ratio_form()[1],
                         "panel_ratio": list(best[0]), "panel_update": list(best[1])}
        e.set_state()  # the tuning launches leave the state untouched; re-fill anyway
    e.rk4(0.0, dt, warmup)
    t0 = warmup * dt
    tot_ms, stage_ms = e.time_rk4(t0, dt, steps)
    info["passes_ms"] = e.time_passes(t0 + (steps + 1) * dt, dt)
    info["form"] = e.get_zstream()[0]
    info["ratio_form"] = e.get_ratio_form()[0]
    info["bg_plane"] = e.get_bg_plane()
    w = e.get_state()
    e.close()
    return tot_ms, stage_ms, w, info


def _field_errors(a, b, m, names):
    import numpy as np

    l2, linf = {}, {}
    for f, name in enumerate(names):
        ref = b[f][m]
        d = a[f][m] - ref
        amp = np.abs(ref).max()
        rms = np.sqrt(np.mean(ref ** 2))
        l2[name] = float(np.sqrt(np.mean(d ** 2)) / rms) if rms > 0 else 0.0
        linf[name] = float(np.abs(d).max() / amp) if amp > 0 else 0.0
    return l2, linf


def accuracy_leg(setup, nsteps):
    """the mixed path against the FP64 path after nsteps RK4 steps from the
    workload's initial state (the stratified workload's horizon is 1000
    steps, whatever --steps says), and two measures of how far perturbations
    carry on this workload, both FP64 runs against the FP64 run: dt changed
    by 1e-9 (a smooth perturbation), and the state rounded to FP32 once,
    after the first step, the FP64 path continuing from the rounded state
    (the perturbation that storing the state in FP32 makes, once instead of
    at every stage); and mixed against FP64 once more with the hyperdiffusion
    switched off, the term whose ratio switch carries the amplification."""
    import numpy as np

    from smaugamd.engine import FP64, MIXED, Engine

    names = ["rho", "m1", "m2", "m3", "e", "b1", "b2", "b3"]
    n0, n1, n2 = setup.n
    m = np.zeros((n2, n1, n0), dtype=bool)  # interior cells
    m[2:-2, 2:-2, 2:-2] = True
    m = m.reshape(-1)
    out = {}
    nohd = dataclasses.replace(setup, hyperdif=0)
    for key, prec, scale in (("mixed", MIXED, 1.0), ("fp64", FP64, 1.0),
                             ("fp64_dt", FP64, 1.0 + 1e-9), ("fp64_round1", FP64, 1.0),
                             ("mixed_nohd", MIXED, 1.0), ("fp64_nohd", FP64, 1.0)):
        e = Engine(nohd if key.endswith("_nohd") else setup, prec)
        e.set_state()
        dt = setup.dt * scale
        if key == "fp64_round1":
            e.rk4(0.0, dt, 1)
            w = e.get_state().astype(np.float32).astype(np.float64)
            e.set_state(w, setup.bg)
            if nsteps > 1:
                e.rk4(dt, dt, nsteps - 1)
        else:
            e.rk4(0.0, dt, nsteps)
        out[key] = e.get_state()
        e.close()
    l2, linf = _field_errors(out["mixed"], out["fp64"], m, names)
    sens, _ = _field_errors(out["fp64_dt"], out["fp64"], m, names)
    rnd, _ = _field_errors(out["fp64_round1"], out["fp64"], m, names)
    nhd, _ = _field_errors(out["mixed_nohd"], out["fp64_nohd"], m, names)
    return {"reference": "FP64 GPU path (parity-tested against the FP64 oracle)",
            "steps": nsteps, "rel_l2": l2, "rel_linf": linf,
            "target_rel_l2": 1e-5,
            "meets_target": bool(all(v <= 1e-5 for v in l2.values())),
            "fp64_rel_l2_under_1e-9_dt_perturbation": sens,
            "fp64_rel_l2_after_one_fp32_rounding_of_the_state_at_step_1": rnd,
            "target_reachable_with_fp32_state": bool(all(v <= 1e-5 for v in rnd.values())),
            "rel_l2_without_hyperdiffusion": nhd}


def pass_kernels(form, bg_plane, fp32=True, rform=1):
    """the kernels of each pass of a 3-D RK4 step for the update form, ratio
    form and background layout in use: (name as rocprofv3 prints it,
    namespace and arguments dropped; launches per step).  With the k-plane
    background and the z-streamed update, the update of every stage also
    forms the next stage's aux quantities and cfast maxima (k_ghost_prep, the
    ghost fill, does the ghost cells'), stage 3 those of the next step's stage
    0: the step has no aux pass.  Ratio form 2 with the z-streamed sources:
    the ratio pass forms the pairs along k (k_ratio_kz) and the source kernel
    the in-plane pairs (k_hyper_zsr)."""
    t = "float" if fp32 else "double"
    s = "f" if fp32 else "d"
    bp = "true" if bg_plane else "false"
    fold = bool(form) and bg_plane
    rin = rform == 2 and form == 1
    src = {1: f"k_hyper_zsr_{s}<{bp}>" if rin else f"k_hyper_zs_{s}<{bp}>",
           2: f"k_hyper<3, {t}>", 3: f"k_hyper_lds_{s}"}
    if fold:
        upd = [(f"k_zstream_{s}<true, true>", 4)]
    elif form:
        upd = [(f"k_zstream_{s}<{bp}, false>", 4)]
    else:
        upd = [(f"k_update<3, {t}>", 4)]
    return {
        "aux+cmax": [] if fold else [(f"k_prep_col<{t}>", 4)] if bg_plane
        else [(f"k_prep<3, {t}>", 4)],
        "ratio": [(f"k_ratio_kz_{s}" if rin else f"k_ratio_zs_{s}", 4)],
        "hyperdiffusive sources": [(src.get(form, ""), 4)],
        "stage update": upd,
        "ghost fill": [(f"k_ghost_prep<{t}>", 4)] if fold
        else [(f"k_bc<3, {t}>", 1 if bg_plane else 4)],
    }


def profiled_traffic(kernel, cells):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by the copy-kernel
    calibration) when it was taken on a grid of the same size, else None"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("cells") != cells:
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return k["fetch_bytes"] * t["fetch_scale"] + k["write_bytes"] * t["write_scale"]


def pass_traffic(kernels, cells):
    """profiled HBM bytes of a pass per stage (its launches in one RK4 step /
    4), or None when a kernel of it is not in the summary"""
    tot = 0.0
    for name, n in kernels:
        b = profiled_traffic(name, cells)
        if b is None:
            return None
        tot += b * n
    return tot / STAGES_PER_STEP


def cpu_baseline(n):
    """oracle RK4 steps on a bounded sample of the workload: the same vertical
    column structure with a reduced horizontal extent"""
    from oracle.oracle import Oracle
    from smaugamd import configs

    nh = 192
    s = configs.stratified(nz=n, nx=nh, ny=nh, lx=4.0e6 * nh / 256, ly=4.0e6 * nh / 256)
    o = Oracle(s)
    w0 = o.boundary(s.w)
    o.run(w0, 0.0, s.dt, 1)  # warm caches and page in
    steps = 4
    t = time.perf_counter()
    o.run(w0, 0.0, s.dt, steps)
    el = time.perf_counter() - t
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": s.ninterior * steps / el / 1e6,
        "unit": "Mcell-updates/s",
        "cores": Oracle.threads(),
        "cpu_model": model,
        "cpus_available": len(os.sched_getaffinity(0)),
        "cpus_online": os.cpu_count(),
        "kind": "port",
        "sample": f"{steps} RK4 steps of the stratified workload on {n}x{nh}x{nh} interior "
                  f"cells (full vertical column, {nh}x{nh} horizontal), FP64, "
                  f"{el:.1f} s wall, {Oracle.threads()} OpenMP threads",
    }


def ensemble_leg(steps, warmup, seq_steps=20, compare=True):
    import numpy as np

    from smaugamd import configs
    from smaugamd.engine import MIXED, Engine, Ensemble

    sides, scales = configs.ensemble_sizes()
    setups = [configs.alfven_member(n, b) for n, b in zip(sides, scales)]
    dts = [s.dt for s in setups]
    ens = Ensemble(setups, MIXED)
    for r in range(len(setups)):
        ens.set_state(r)
    s0 = [ens.sums(r) for r in range(len(setups))]
    ens.rk4(0.0, dts, warmup)
    ms = ens.time_rk4([warmup * d for d in dts], dts, steps)
    passes = ens.time_passes([(warmup + steps) * d for d in dts], dts)
    finite = all(np.isfinite(ens.get_state(r)).all() for r in range(len(setups)))
    # drift of the compensated FP64 sums relative to the initial sum of
    # |field| (mass rho, energy e); the hyperviscosity windows are padded at
    # the array ends whatever the boundary type (as in the reference), so a
    # periodic run with hyperdiffusion does not conserve energy exactly
    drift = {"rho": 0.0, "e": 0.0}
    for r in range(len(setups)):
        s1 = ens.sums(r)
        w = setups[r].w
        for name, f in (("rho", 0), ("e", 3)):
            drift[name] = max(drift[name], abs(s1[f] - s0[r][f]) / np.abs(w[f]).sum())
    cells = ens.ninterior
    ens.close()
    # the same runs one grid at a time (single-grid engine, graph per grid)
    seq_ms = 0.0
    for s in (setups if compare else []):
        e = Engine(s, MIXED)
        e.set_state()
        e.rk4(0.0, s.dt, 2)
        tot, _ = e.time_rk4(2 * s.dt, s.dt, seq_steps)
        seq_ms += tot / seq_steps
        e.close()
    ms_step = ms / steps
    gbs = BYTES_PER_CELL_STEP_2D_MIXED * cells / (ms_step / 1e3) / 1e9
    return {
        "workload": f"{len(setups)} periodic 2-D Orszag-Tang-type runs, sides "
                    f"{sorted(set(sides))} (numpy default_rng(0)), b0 in U[0.5,1.5]/sqrt(4 pi), "
                    f"per-run dt = 0.3 CFL, mixed precision",
        "runs": len(setups),
        "cells": cells,
        "steps": steps,
        "ms_per_step": ms_step,
        "value": cells * steps / (ms / 1e3) / 1e6,
        "unit": "Mcell-updates/s",
        "hbm_gbs_algorithmic": gbs,
        "hbm_frac": gbs / HBM_PEAK_GBS,
        "finite": bool(finite),
        "passes_ms": passes,
        "conservation_max_rel_drift": drift,
        "one_grid_at_a_time_ms_per_step": seq_ms if compare else None,
        "speedup_over_one_grid_at_a_time": seq_ms / ms_step if compare else None,
    }


def orszag_tang_leg(nsteps=500):
    """the Orszag-Tang vortex, 1024^2 periodic, fixed dt = 1e-4, mixed and FP64
    paths: Mcell-updates/s of each and the per-field error of the mixed
    result against the FP64 one after nsteps"""
    import numpy as np

    from smaugamd import configs
    from smaugamd.engine import FP64, MIXED, Engine

    s = configs.orszag_tang(nx=1024)
    out, res = {}, {}
    for prec, name in ((MIXED, "mixed"), (FP64, "fp64")):
        e = Engine(s, prec)
        e.set_state()
        e.autotune(3)
        e.rk4(0.0, s.dt, 2)
        e.set_state()
        tot, _ = e.time_rk4(0.0, s.dt, nsteps)
        res[name] = e.get_state()
        rform = e.get_ratio_form()[0]
        e.close()
        out[name] = {"ms_per_step": tot / nsteps, "value": s.ninterior * nsteps / (tot / 1e3) / 1e6,
                     "ratio_form": rform}
    m = np.zeros((s.n[1], s.n[0]), dtype=bool)
    m[2:-2, 2:-2] = True
    m = m.reshape(-1)
    names = ["rho", "m1", "m2", "e", "b1", "b2"]
    l2 = {}
    for f in range(s.nve):
        ref = res["fp64"][f][m]
        rms = np.sqrt(np.mean(ref ** 2))
        l2[names[f]] = float(np.sqrt(np.mean((res["mixed"][f][m] - ref) ** 2)) / rms)
    return {"workload": "Orszag-Tang vortex 1024x1024 periodic on [0,2pi]^2, dt = 1e-4",
            "steps": nsteps + 1, "unit": "Mcell-updates/s", "mixed": out["mixed"],
            "fp64": out["fp64"], "speedup_mixed_over_fp64": out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"],
            "mixed_rel_l2_vs_fp64": l2}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np

    from smaugamd import configs
    from smaugamd.engine import FP64, MIXED

    n = a.n
    if a.large:
        setup = configs.stratified(nz=768, nx=512, ny=512, lz=12.0e6, lx=8.0e6, ly=8.0e6)
        shape_txt = "768x512x512 cells over 12x8x8 Mm"
    else:
        setup = configs.stratified(nz=n, nx=n, ny=n)
        shape_txt = f"{n}x{n}x{n} cells over 6x4x4 Mm"
    cells = setup.ninterior

    print(f"[bench] set-up done: {setup.n} dt={setup.dt:.4e}", file=sys.stderr, flush=True)
    tot, stage_ms, wm, info = run_path(setup, MIXED, a.warmup, a.steps, shape=a.shape)
    print(f"[bench] mixed path: {tot / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
    ms_step = tot / a.steps
    value = cells * a.steps / (tot / 1e3) / 1e6

    out_fp64 = None
    acc = None
    if not a.no_fp64:
        tot64, stage64, wd, info64 = run_path(setup, FP64, a.warmup, a.steps)
        print(f"[bench] fp64 path: {tot64 / a.steps:.3f} ms/step", file=sys.stderr, flush=True)
        out_fp64 = {
            "value": cells * a.steps / (tot64 / 1e3) / 1e6,
            "ms_per_step": tot64 / a.steps,
            "stage_kernel_ms": stage64,
            "tuned": info64.get("tuned"),
            "passes_ms": info64["passes_ms"],
            "speedup_mixed_over_fp64": (tot64 / tot),
        }
    if not a.no_accuracy:
        acc = accuracy_leg(setup, a.accuracy_steps)
        print(f"[bench] accuracy leg: {acc['steps']} steps", file=sys.stderr, flush=True)

    # the roofline is given for the dominant kernel of the step: of the
    # passes that run one kernel in every stage, the one with the longest
    # mean launch (HIP events on the engine stream; for the stage update the
    # events around each update launch of one more step).  Every pass of a
    # stage processes the stage of every cell, so its algorithmic bytes per
    # launch are the step figure / 4 stages x cells.
    bytes_launch = BYTES_PER_CELL_STEP["mixed"] / STAGES_PER_STEP * cells
    pms = info["passes_ms"]
    pk = pass_kernels(info["form"], info["bg_plane"], rform=info["ratio_form"])
    single = [p for p in pms if len(pk[p]) == 1 and pk[p][0][1] == STAGES_PER_STEP]
    dom = max(single or pms, key=pms.get)
    kname = " + ".join(k for k, _ in pk[dom])
    launch_ms = stage_ms if dom == "stage update" else pms[dom]
    achieved = bytes_launch / (launch_ms / 1e3) / 1e9
    stage_total_ms = sum(pms.values())
    roof = {
        "bound": "hbm",
        "kernel": f"{kname} ({dom} pass, the longest of the stage)",
        "algorithmic_bytes_per_launch": bytes_launch,
        "launch_ms": launch_ms,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": pass_traffic(pk[dom], cells),
        "passes_ms": pms,
        "passes_traffic_bytes": {k: pass_traffic(v, cells) for k, v in pk.items()},
        "stage_frac": bytes_launch / (stage_total_ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "step_level_frac": (BYTES_PER_CELL_STEP["mixed"] * cells / (ms_step / 1e3) / 1e9)
        / HBM_PEAK_GBS,
    }
    from smaugamd.engine import copy_probe, stream_probe, walk_probe

    roof["measured_copy_gbs"] = copy_probe()
    roof["frac_of_measured_copy"] = achieved / roof["measured_copy_gbs"]
    # the streams of a stage update (32 read, 20 written, one FP32 value per
    # cell each), for the measured traffic of the update pass to be read
    # against: what HBM delivers to them read linearly, and walked plane by
    # plane in the update's 64 x 4-column tiles at its tuned depth
    roof["measured_stream_gbs"] = stream_probe(cells)
    roof["measured_walk_gbs"] = walk_probe(setup.n, kz=(info.get("tuned") or {}).get("kz", 128))
    ens = None
    if not a.no_ensemble:
        ens = ensemble_leg(a.ens_steps, a.warmup, compare=not a.no_ens_compare)
        print(f"[bench] ensemble: {ens['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    ot = None
    if not a.no_ot:
        ot = orszag_tang_leg(a.ot_steps)
        print(f"[bench] orszag-tang: {ot['mixed']['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    cpu = None if a.no_cpu else cpu_baseline(n)
    rec = {
        "metric": "Mcell-updates/s per RK4 step, mixed vs FP64 at stated L2 error; % HBM peak",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"3-D stratified solar atmosphere {shape_txt}, RK4, "
                               f"dt={setup.dt:.4e} s (min of CFL 0.3 and the hyperviscous "
                               f"limit)",
                   "cells": cells,
                   "precision": "mixed: FP32 perturbation state and stencil arithmetic; FP64 "
                                "background, pressure recovery, cmax/dt and conservation sums",
                   "parallelism": "single GPU",
                   "launch_shapes": info.get("tuned")},
        "fp64": out_fp64,
        "accuracy": acc,
        "roofline": roof,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "orszag_tang": ot,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
