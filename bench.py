"""FoilCFD MI355X benchmark: GLUPS and HBM roofline fraction of the fused
D3Q19 TRT + Smagorinsky stream-collide on the 23.6 M-cell NACA 2412 wing
(512 x 192 x 240, chord 256, AoA 8 deg), 16-bit shifted storage (FP16C,
s1e4m11) vs FP32, in the full default-tunnel configuration: a factor-2
refinement patch hugging the airfoil (the reference app's patch margins
0.20 / 0.50 / 0.10 / 0.20 chords), the iMEM wall model on the wall cells of
both levels, and 1 M tracer particles advected once per call.

One "step" is one coarse time step with its two fine sub-steps (the
reference's LBMSolver::stepN unit); GLUPS counts coarse + 2 x fine cell
updates per step, the reference's perfStats rule. Steps run in calls of 100
(one macro store, one momentum-exchange sample, one wall-model update per
level and one particle advection per call, like the reference's frame).

Prints ONE JSON line. value = FP16C-storage GLUPS (whole job, one GPU).
--coarse-only measures the bare coarse wing instead (no patch, wall model or
particles).
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"f16": 76.0, "f16c": 76.0, "bf16": 76.0, "f32": 152.0, "f64": 304.0}
# restriction record per coarse cell: 19 fneq, valid mark, 4 moments; FP32 for FP32 / FP64 storage,
# fneq and the mark as IEEE halves beside FP32 moments for the 16-bit storages
RECORD_BYTES = {"f16": 64.0, "f16c": 64.0, "bf16": 64.0, "f32": 96.0, "f64": 96.0}
METRIC = "GLUPS & % HBM peak, 23.6M-cell wing, FP16-store vs FP32 at ≤1% Cl/Cd err"
REFERENCE_PUBLISHED_GLUPS_FP32 = 8.151  # RTX 5090, 23.6M cells, FP32 (notes/PERF.md:10-20)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_storage(case, storage, steps, warmup, per_call, kernel_steps, variant="auto"):
    from foil_cfd_amd import cases
    s = cases.make_solver(case, storage)
    s.set_step_variant(variant)
    n = case.dims[0] * case.dims[1] * case.dims[2]
    # warmup (not timed)
    done = 0
    while done < warmup:
        k = min(per_call, warmup - done)
        s.step_n(k)
        done += k
    s.synchronize()
    # timed region: synchronize on both sides, wall clock
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(per_call, steps - done)
        s.step_n(k)
        done += k
    s.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    # per-launch HIP-event timing of the stream-collide kernel on the solver's stream
    s.set_kernel_timing(True)
    done = 0
    while done < kernel_steps:
        k = min(per_call, kernel_steps - done)
        s.step_n(k)
        done += k
    launches, kms = s.kernel_timing()
    s.set_kernel_timing(False)
    nan = s.nan_detected()
    fr = s.forces()
    s.close()
    avg_kernel_ms = kms / launches if launches else float("nan")
    res = {
        "glups": n * steps / wall / 1e9,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms": avg_kernel_ms,
        "kernel_launches": launches,
        "kernel_gbs": BYTES_PER_CELL[storage] * n / (avg_kernel_ms * 1e-3) / 1e9,
        "nan": bool(nan),
        "cl_ema": fr.cl,
        "cd_ema": fr.cd,
    }
    log(f"[bench] {storage} ({variant}): {res['glups']:.2f} GLUPS wall, kernel {avg_kernel_ms:.4f} ms "
        f"({res['kernel_gbs']:.0f} GB/s algorithmic)")
    return res


def full_setup(case, storage, aoa, wall="newton"):
    """The default-tunnel configuration on top of the coarse wing
    (cases.default_tunnel): factor-2 patch with the app's margins, clean
    surface references, wall model on both levels, 1 M particles."""
    from foil_cfd_amd import cases
    t0 = time.perf_counter()
    s, box, info = cases.default_tunnel(case, storage, aoa, wall)
    log(f"[bench] {storage}: patch {box} fine grid {info['fine_dims']} "
        f"({info['device_bytes'] / 1e9:.2f} GB), wall cells coarse {s.wall_cell_count_level(0)} "
        f"fine {s.wall_cell_count_level(1)}, setup {time.perf_counter() - t0:.1f} s")
    return s, box, info


def train_wall_network(s, calls, per_call, seed=[1]):
    """The learned wall function for a solver: wall-cell samples (y_s, |u_t|,
    nu, Newton u_tau) harvested from its own wall cells over `calls` calls,
    a 3-32-32-1 network trained on them with the library's BF16 MFMA kernels
    and Adam, then installed (Newton stays the fallback outside the
    network's envelope). Untimed set-up; returns what was trained."""
    import numpy as np

    from foil_cfd_amd import mlp
    s.set_wall_harvest(True)
    got = []
    for _ in range(calls):
        s.step_n(per_call)
        h = s.wall_harvest()
        got.append(h[h[:, 3] > 0])
    s.set_wall_harvest(False)
    smp = np.concatenate(got)
    if len(smp) < 1024:
        raise RuntimeError(f"wall harvest gave {len(smp)} samples; the learned wall function needs a flow")
    tr = mlp.MlpTrainer(smp, seed=7)
    t0 = time.perf_counter()
    losses = tr.fit(2000, 1024, 3e-3, 1e-4)
    prm = tr.export()
    tr.close()
    s.set_wall_mlp(prm)
    s.set_wall_model("mlp")
    out = {"samples": int(len(smp)), "train_s": round(time.perf_counter() - t0, 3),
           "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:]))}
    log(f"[bench] learned wall function: {out['samples']} harvested samples, loss {out['loss_first']:.3e} -> "
        f"{out['loss_last']:.3e} in {out['train_s']} s")
    return out


def run_full(case, storage, steps, warmup, per_call, kernel_steps, aoa, variant="auto", wall_fn="newton"):
    n_c = case.dims[0] * case.dims[1] * case.dims[2]
    s, box, info = full_setup(case, storage, aoa)
    s.set_step_variant(variant)
    fd = info["fine_dims"]
    n_f = fd[0] * fd[1] * fd[2]
    updates = n_c + 2 * n_f
    seed = [1]

    def run(k_total):
        done = 0
        while done < k_total:
            k = min(per_call, k_total - done)
            s.step_n(k)
            s.particles_advect(float(k), seed[0])
            seed[0] += 1
            done += k

    run(warmup)
    # the learned wall function is trained (untimed) on samples harvested from
    # this solver's own wall cells after the startup ramps (2 nx steps; the
    # ramped viscosity would leave the later cells outside its envelope)
    network = None
    if wall_fn == "mlp":
        ramp_end = 2 * case.dims[0]
        while s.step_count() < ramp_end:
            s.step_n(min(100, ramp_end - s.step_count()))
        network = train_wall_network(s, 3, 100)
    s.synchronize()
    t0 = time.perf_counter()
    run(steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    s.set_kernel_timing(True)
    run(kernel_steps)
    lv = [s.kernel_timing_level(i) for i in (0, 1)]
    s.set_kernel_timing(False)
    nan = s.nan_detected()
    fr = s.forces()
    wr = s.wall_model_readout()
    s.close()
    bpc = BYTES_PER_CELL[storage]
    # With the default variant the second fine sub-step of each coarse step
    # also writes the restriction's child-average records (96 B per coarse
    # cell of the box shrunk by 2): counted as half a record set per launch.
    x0, y0, x1, y1 = box
    n_rec = (x1 - x0 - 4) * (y1 - y0 - 4) * case.dims[2]
    rec_bytes = RECORD_BYTES[storage] * n_rec / 2 if variant == "auto" else 0.0
    per_level = {}
    for name, (launches, ms), cells, extra in (("coarse", lv[0], n_c, 0.0), ("fine", lv[1], n_f, rec_bytes)):
        avg = ms / launches if launches else float("nan")
        per_level[name] = {"cells": cells, "launches": launches, "kernel_ms": avg,
                           "algorithmic_bytes_per_launch": bpc * cells + extra,
                           "kernel_gbs": (bpc * cells + extra) / (avg * 1e-3) / 1e9}
    tot_ms = lv[0][1] + lv[1][1]
    res = {
        "glups": updates * steps / wall / 1e9,
        "ms_per_step": wall * 1e3 / steps,
        "cell_updates_per_step": updates,
        "kernel_ms": per_level["fine"]["kernel_ms"],
        "kernel_gbs": per_level["fine"]["kernel_gbs"],
        "kernel_bytes": per_level["fine"]["algorithmic_bytes_per_launch"],
        "all_step_kernels_gbs": bpc * (n_c * lv[0][0] + n_f * lv[1][0]) / (tot_ms * 1e-3) / 1e9,
        "step_kernel_share_of_wall": tot_ms / (kernel_steps / steps * wall * 1e3) if wall > 0 else None,
        "per_level": per_level,
        "nan": bool(nan),
        "cl_ema": fr.cl,
        "cd_ema": fr.cd,
        "wall_cells_active_fine": wr.cells,
        "wall_function": wall_fn,
        "wall_network_cells": int(wr.mlp_cells),
        "wall_network": network,
        "patch_box": list(box),
        "fine_dims": list(fd),
    }
    log(f"[bench] {storage} full: {res['glups']:.2f} GLUPS wall ({res['ms_per_step']:.2f} ms/step), fine step "
        f"kernel {res['kernel_ms']:.3f} ms ({res['kernel_gbs']:.0f} GB/s), coarse step kernel "
        f"{per_level['coarse']['kernel_ms']:.3f} ms ({per_level['coarse']['kernel_gbs']:.0f} GB/s)")
    return res


def accuracy_run(args):
    """Time-averaged Cl / Cd of the full default-tunnel configuration for one
    storage format (the reference's protocol: one momentum-exchange sample per
    call of --per-call steps folded into the 256-sample ring, the readout
    valid after two flow-throughs, the average over the samples after the
    first one; lbm_solver.h:25, lbm_solver.cpp:633-654, :1414-1509). Every
    raw sample is kept as well, so other windows can be averaged afterwards.
    Writes the record to --out after every progress line."""
    import foil_cfd_amd as fc
    from foil_cfd_amd import cases
    fc.load()
    storage = args.storages.split(",")[0].strip()
    case = cases.config3_wing(args.aoa)
    if args.tau:  # same grid, patch and models at a lower Reynolds number
        case = cases.naca_tunnel(512, 192, 240, 256, aoa=args.aoa, u_lat=0.08, cs=0.1733, tau=args.tau)
    s, box, info = full_setup(case, storage, args.aoa, wall=args.wall)
    s.set_stochastic_rounding(args.rounding == "stochastic")
    ft = fc.flow_through_steps(case.scaling, case.dims[0])
    total = int(round(args.flow_throughs * ft / args.per_call)) * args.per_call
    span, fscale = info["fine_dims"][2], info["fine_scaling"]
    rec = {"workload": f"naca2412_wing_512x192x240_chord256_aoa{args.aoa:g}_trt_les_imem_patch2x_particles1M",
           "storage": storage, "aoa": args.aoa, "wall_model": args.wall, "per_call": args.per_call,
           "flow_through_steps": ft, "steps_planned": total, "patch_box": list(box),
           "fine_dims": list(info["fine_dims"]), "forces_from_fine": info["forces_from_fine"],
           "tau": case.scaling.tau, "fine_tau": fscale.tau, "re_chord": case.scaling.re_effective,
           "rounding": args.rounding if storage in ("f16c", "f16", "bf16") else "exact",
           "samples": []}
    t0 = time.perf_counter()
    seed = 1
    done = 0
    status = "ok"
    while done < total:
        s.step_n(args.per_call)
        s.particles_advect(float(args.per_call), seed)
        seed += 1
        done += args.per_call
        s.synchronize()
        F, at = s.force_raw()
        cl = fc.force_to_coefficient(fscale if info["forces_from_fine"] else case.scaling, float(F[1]), span)
        cd = fc.force_to_coefficient(fscale if info["forces_from_fine"] else case.scaling, float(F[0]), span)
        rec["samples"].append([int(at), float(F[0]), float(F[1]), cl, cd])
        if s.nan_detected():
            status = f"watchdog verdict {s.nan_kind()} at step {s.step_count()}"
        if done % (10 * args.per_call) == 0 or done >= total or status != "ok":
            fr = s.forces()
            rec.update({"steps_done": done, "status": status, "wall_s": round(time.perf_counter() - t0, 2),
                        "readout": dataclasses.asdict(fr)})
            with open(args.out, "w") as fh:
                json.dump(rec, fh)
            print(f"[accuracy] {storage} aoa {args.aoa:g}: step {done}/{total} ({done / ft:.2f} FT) "
                  f"Cl {cl:.4f} Cd {cd:.4f} avg Cl {fr.cl_avg:.4f} Cd {fr.cd_avg:.4f} valid {fr.valid} "
                  f"{time.perf_counter() - t0:.0f} s {status}", flush=True)
        if status != "ok":
            break
    s.close()


ACCURACY_PROFILE = os.path.join(ROOT, "profiles", "wing_cl_cd_accuracy.json")


def accuracy_summary(paths, out_path, batches=6):
    """Folds --accuracy records into the committed summary: per angle of
    attack and storage the time-averaged Cl / Cd over the samples taken from
    one flow-through on (the reference readout's window), the standard error
    of each average from `batches` batch means, the error against the FP64
    run of the same angle with its one-sigma uncertainty, and how long the
    sample traces of a storage and the FP64 run stay within 1 % of each other
    (the start-up transient is deterministic; the developed separated flow is
    chaotic, so traces then decorrelate and only averages compare)."""
    import numpy as np
    runs = {}
    for p in paths:
        with open(p) as fh:
            r = json.load(fh)
        # 16-bit records without a rounding field predate stochastic rounding
        rnd = r.get("rounding", "nearest" if r["storage"] in ("f16c", "f16", "bf16") else "exact")
        name = r["storage"] + ("_nearest" if rnd == "nearest" else "")
        runs.setdefault(f"{r['aoa']:g}", {})[name] = r
    out = {"workload": "naca2412_wing_512x192x240_chord256_trt_les_imem_patch2x_particles1M",
           "protocol": "one momentum-exchange sample per 100 coarse steps; average over samples at or after "
                       "1 flow-through (reference readout), 4 flow-throughs run; standard error from batch means",
           "batches": batches, "aoa": {}}
    try:  # angles not among the new records keep their committed entries
        with open(out_path) as fh:
            out["aoa"].update({k: v for k, v in json.load(fh).get("aoa", {}).items() if k not in runs})
    except (OSError, ValueError):
        pass
    for aoa, by in sorted(runs.items(), key=lambda kv: float(kv[0])):
        ent = {}
        for st, r in by.items():
            s = np.asarray(r["samples"], dtype=np.float64)
            ft = r["flow_through_steps"]
            m = s[:, 0] >= ft
            e = {"status": r.get("status"), "steps": int(r.get("steps_done", 0)), "wall_s": r.get("wall_s"),
                 "samples": int(m.sum()), "rounding": "nearest" if st.endswith("_nearest") else (
                     "stochastic" if st in ("f16c", "f16", "bf16") else "exact")}
            for name, col in (("cl", 3), ("cd", 4)):
                v = s[m, col]
                bm = np.array([b.mean() for b in np.array_split(v, batches)])
                e[name] = round(float(v.mean()), 6)
                e[name + "_se"] = round(float(bm.std(ddof=1) / np.sqrt(batches)), 6)
                e[name + "_batch_means"] = [round(float(x), 5) for x in bm]
            e["readout_cl_avg"] = r["readout"]["cl_avg"]
            e["readout_cd_avg"] = r["readout"]["cd_avg"]
            ent[st] = e
        if "f64" in by:
            ref = np.asarray(by["f64"]["samples"], dtype=np.float64)
            for st, e in ent.items():
                if st == "f64":
                    continue
                for name in ("cl", "cd"):
                    r64, s64 = ent["f64"][name], ent["f64"][name + "_se"]
                    e[name + "_err_vs_f64"] = round((e[name] - r64) / abs(r64), 5)
                    e[name + "_err_sigma"] = round(float(np.hypot(e[name + "_se"], s64)) / abs(r64), 5)
                # trace difference per sample, relative to the FP64 run's mean
                s = np.asarray(by[st]["samples"], dtype=np.float64)
                k = min(len(s), len(ref))
                rel = np.maximum(np.abs(s[:k, 3] - ref[:k, 3]) / abs(ent["f64"]["cl"]),
                                 np.abs(s[:k, 4] - ref[:k, 4]) / abs(ent["f64"]["cd"]))
                off = np.nonzero(rel > 0.01)[0]
                e["traces_within_1pct_until_step"] = int(ref[k - 1, 0]) if not len(off) else (
                    int(ref[off[0] - 1, 0]) if off[0] > 0 else 0)
                early = ref[:k, 0] <= 1100
                e["trace_max_rel_diff_to_step_1100"] = round(float(rel[early].max()), 6) if early.any() else None
        out["aoa"][aoa] = ent
    out["aoa"] = dict(sorted(out["aoa"].items(), key=lambda kv: float(kv[0])))
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps({k: {st: {kk: v for kk, v in e.items() if "batch" not in kk} for st, e in ent.items()}
                      for k, ent in out["aoa"].items()}), flush=True)


ACCURACY_LAMINAR_PROFILE = os.path.join(ROOT, "profiles", "wing_cl_cd_accuracy_re1000.json")


def accuracy_line():
    """The committed accuracy summaries (the LES case at the reference's
    Reynolds number and the laminar one at chord Re ~1000), reduced to the
    bench line's figures: per angle and storage the Cl / Cd error against
    FP64 and its one-sigma uncertainty from batch means."""
    out = {}
    for key, path in (("les", ACCURACY_PROFILE), ("laminar_re1000", ACCURACY_LAMINAR_PROFILE)):
        try:
            with open(path) as fh:
                rec = json.load(fh)
        except (OSError, ValueError):
            continue
        res = {}
        for aoa, ent in rec.get("aoa", {}).items():
            for st, e in ent.items():
                if "cl_err_vs_f64" in e:
                    res.setdefault(f"aoa{aoa}", {})[st] = {k: e[k] for k in ("cl_err_vs_f64", "cd_err_vs_f64",
                                                                            "cl_err_sigma", "cd_err_sigma")}
        if res:
            out[key] = {"per_aoa": res, "source": os.path.relpath(path, ROOT)}
    return out or None


def cpu_baseline(case, seconds=12.0):
    """The oracle's FP32 stream-collide (reference semantics, pull ping-pong)
    on the host cores, on the same grid, for a bounded number of steps."""
    import numpy as np

    from oracle import oracle as orc
    nx, ny, nz = case.dims
    n = nx * ny * nz
    threads = orc.threads()
    f = orc.init_equilibrium(case.flags, case.dims, 0.0)
    sc = case.scaling
    # one probe step to size the sample
    taus, uins = orc.ramp_schedule(sc.tau, sc.nu_lat, sc.u_lat, nx, 64)
    t0 = time.perf_counter()
    orc.run(f, case.flags, case.dims, taus[:1], uins[:1], cs=case.cs, macro=False)
    probe = time.perf_counter() - t0
    steps = int(max(1, min(60, seconds / max(probe, 1e-3))))
    t0 = time.perf_counter()
    orc.run(f, case.flags, case.dims, taus[1:1 + steps], uins[1:1 + steps], cs=case.cs,
            macro=False)
    dt = time.perf_counter() - t0
    del f
    glups = n * steps / dt / 1e9
    log(f"[bench] cpu oracle: {steps} steps in {dt:.1f} s on {threads} threads -> {glups:.4f} GLUPS")
    return {"value": glups, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 pull/ping-pong stream-collide, 512x192x240 wing, {steps} steps "
                      f"({dt:.1f} s wall)", "seconds": dt}


def measured_traffic(storage, level="coarse"):
    """HBM bytes per step-kernel launch of a level measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE on this workload (copy-kernel calibrated), from the
    committed profile summary; None when absent."""
    path = os.path.join(ROOT, "profiles", "stream_collide_wing_traffic.json")
    key = {"f16c": "StoreF16C", "f16": "StoreF16", "bf16": "StoreBF16", "f32": "StoreF32",
           "f64": "StoreF64"}[storage]
    if level != "coarse":
        key = f"{key}:{level}"
    try:
        with open(path) as fh:
            rec = json.load(fh).get(key)
    except (OSError, ValueError):
        return None
    return None if rec is None else round(rec["traffic_bytes_per_launch_mean"])


def copy_calibration(mib=1024, reps=5):
    """A plain device copy of known size (elementwise negate, N bytes read and
    N bytes written per launch) for scaling the HBM traffic counters."""
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "torch.neg float32", "bytes_read": n * 4,
                      "bytes_written": n * 4, "launches": reps}), flush=True)


def polar_bench(args):
    """Batched polar sweep: 16 NACA 2412 cases (AoA -4..16 deg, chords from a
    fixed-seed draw in [96, 192], grids 2c x 0.75c x 0.9c) stepped as one
    group (one batched launch per step, graph replay, autotuned launch shape).
    Aggregate GLUPS of the group, and each case's GLUPS when stepped alone."""
    import foil_cfd_amd as fc
    from foil_cfd_amd import cases, group
    fc.load()
    t_setup = time.perf_counter()
    cs = group.polar_sweep(16)
    log(f"[polar] 16 cases built in {time.perf_counter() - t_setup:.1f} s, "
        f"{sum(c.flags.size for c in cs) / 1e6:.1f} M cells")
    solvers = [cases.make_solver(c, args.polar_storage) for c in cs]
    if args.polar_wall != "off":
        for s in solvers:
            s.set_wall_model("newton")
    g = group.SolverGroup(solvers)
    g.step(args.warmup)
    trained = None
    if args.polar_wall == "mlp":
        import numpy as np
        from foil_cfd_amd import mlp
        for s in solvers:
            s.set_wall_harvest(True)
        got = []
        for _ in range(4):
            g.step(5)
            for s in solvers:
                h = s.wall_harvest()
                got.append(h[h[:, 3] > 0])
        for s in solvers:
            s.set_wall_harvest(False)
        smp = np.concatenate(got)
        tr = mlp.MlpTrainer(smp, seed=7)
        t0 = time.perf_counter()
        tr.fit(2000, 1024, 3e-3, 1e-4)
        trained = {"samples": int(len(smp)), "train_s": round(time.perf_counter() - t0, 3)}
        prm = tr.export()
        for s in solvers:
            s.set_wall_mlp(prm)
            s.set_wall_model("mlp")
        tr.close()
    best = g.autotune(steps_per_trial=args.polar_trial)
    log(f"[polar] autotuned launch shape: {best['waves_per_block']} waves/workgroup, "
        f"{'XCD slab' if best['slab_order'] else 'linear'} order, {best['glups']:.2f} GLUPS")
    for s in solvers:
        s.synchronize()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.per_call, args.steps - done)
        g.step(k)
        done += k
    for s in solvers:
        s.synchronize()
    wall = time.perf_counter() - t0
    ncells = sum(s.ncells for s in solvers)
    agg = ncells * args.steps / wall / 1e9
    g.close()
    per_case = []
    for c, s in zip(cs, solvers):
        s.step_n(args.polar_trial)
        s.synchronize()
        t0 = time.perf_counter()
        s.step_n(args.polar_solo_steps)
        s.synchronize()
        dt = time.perf_counter() - t0
        per_case.append({"dims": list(c.dims), "aoa": round(float(c.name.split("_a")[-1]), 3)
                         if "_a" in c.name else None,
                         "glups_alone": round(s.ncells * args.polar_solo_steps / dt / 1e9, 3)})
        s.close()
    seq_ms = sum(c["dims"][0] * c["dims"][1] * c["dims"][2] / (c["glups_alone"] * 1e6) for c in per_case)
    out = {
        "metric": "polar sweep aggregate GLUPS (16 uneven NACA 2412 cases, one batched launch per step)",
        "value": round(agg, 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "naca2412_polar16_aoa-4to16_chords96to192_grid2c_0.75c_0.9c",
                   "cells": ncells, "storage": args.polar_storage, "wall_model": args.polar_wall,
                   "launch": {k: best[k] for k in ("waves_per_block", "slab_order", "use_graph")},
                   "network": trained},
        "per_case": per_case,
        "sequential_ms_per_step": round(seq_ms, 4),
        "batched_vs_sequential": round(seq_ms / (wall / args.steps * 1e3), 3),
    }
    print(json.dumps(out), flush=True)


PROBE_PROFILE = os.path.join(ROOT, "profiles", "aa_traffic_probe.json")


def traffic_probe(reps_coarse=200, reps_fine=50):
    """Traffic-only ceiling of the in-place step's access pattern on this GPU
    (fc_probe_aa_traffic: the step kernels' decomposition moving 19 reads and
    19 writes per cell, no collision) for the coarse wing and the fine patch
    grid, 2- and 4-byte populations, the local (S->L) and shifted (L->S)
    step and the esoteric-twist pattern (every step reads and writes slot q
    at cell + max(e_q, 0)), with and without 20 FP32 FMAs per population in
    between; plus a plain device copy. One JSON line; rates in TB/s of
    algorithmic bytes."""
    import ctypes as C

    import foil_cfd_amd as fc
    from foil_cfd_amd._abi import FcDims
    lib = fc.load()
    rows = []
    for dims, reps in (((512, 192, 240), reps_coarse), ((864, 248, 480), reps_fine)):
        n = dims[0] * dims[1] * dims[2]
        for eb in (2, 4):
            for pattern, pname in ((0, "S->L"), (1, "L->S"), (2, "twist")):
                for work in (0, 20):
                    ms = C.c_double(0.0)
                    rc = lib.fc_probe_aa_traffic(FcDims(*dims), eb, pattern, work, reps, C.byref(ms))
                    if rc != 0:
                        raise RuntimeError(f"fc_probe_aa_traffic failed ({rc})")
                    tbs = 2.0 * 19 * eb * n / (ms.value * 1e-3) / 1e12
                    rows.append({"dims": list(dims), "bytes_per_pop": eb, "pattern": pname,
                                 "fma_per_pop": work, "ms": round(ms.value, 5), "tb_s": round(tbs, 3)})
                    log(f"[probe] {dims} {eb} B {pname} fma {work}: {ms.value * 1e3:.1f} us {tbs:.3f} TB/s")
    copies = []
    for gib in (1, 4):
        ms = C.c_double(0.0)
        if lib.fc_probe_copy(gib << 30, 50, C.byref(ms)) != 0:
            raise RuntimeError("fc_probe_copy failed")
        tbs = 2.0 * (gib << 30) / (ms.value * 1e-3) / 1e12
        copies.append({"gib": gib, "ms": round(ms.value, 5), "tb_s": round(tbs, 3)})
        log(f"[probe] copy {gib} GiB: {ms.value * 1e3:.1f} us {tbs:.3f} TB/s")
    print(json.dumps({"probe": "aa_traffic", "peak_tb_s": HBM_PEAK_GBS / 1e3, "aa": rows, "copy": copies}),
          flush=True)


def probe_line():
    """The committed traffic-only probe, reduced to the ceilings of the step
    patterns on the coarse and the fine grid with no arithmetic (TB/s)."""
    try:
        with open(PROBE_PROFILE) as fh:
            rec = json.load(fh)
    except (OSError, ValueError):
        return None
    out = {}
    for r in rec.get("aa", []):
        if r["fma_per_pop"] == 0:
            key = f"{'coarse' if r['dims'][0] == 512 else 'fine'}_{r['bytes_per_pop']}B_{r['pattern']}"
            out[key] = r["tb_s"]
    return {"tb_s": out, "source": "profiles/aa_traffic_probe.json"} if out else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--per-call", type=int, default=100)
    ap.add_argument("--kernel-steps", type=int, default=200)
    ap.add_argument("--aoa", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--storages", default="f16c,f32")
    ap.add_argument("--variant", default="auto", choices=["auto", "single", "buf", "wide"],
                    help="step kernel: auto (library default), single, buf or wide")
    ap.add_argument("--coarse-only", action="store_true",
                    help="bare coarse wing (no refinement patch, wall model or particles)")
    ap.add_argument("--periodic-box", action="store_true",
                    help="with --coarse-only: an all-fluid periodic box of the wing's size (bulk step cost)")
    ap.add_argument("--polar", action="store_true",
                    help="batched 16-case polar sweep instead of the wing (separate JSON line)")
    ap.add_argument("--polar-storage", default="f16c")
    ap.add_argument("--polar-wall", default="mlp", choices=["off", "newton", "mlp"])
    ap.add_argument("--polar-trial", type=int, default=10)
    ap.add_argument("--polar-solo-steps", type=int, default=50)
    ap.add_argument("--accuracy", action="store_true",
                    help="time-averaged Cl/Cd run of the full configuration for the first of --storages "
                         "(record written to --out)")
    ap.add_argument("--flow-throughs", type=float, default=4.0)
    ap.add_argument("--tau", type=float, default=None,
                    help="--accuracy at this relaxation time instead of the physical (clamped) one")
    ap.add_argument("--summary-out", default=None,
                    help="--accuracy-summary target (default profiles/wing_cl_cd_accuracy.json)")
    ap.add_argument("--wall", default="newton", choices=["off", "newton", "mlp"])
    ap.add_argument("--rounding", default="stochastic", choices=["stochastic", "nearest"],
                    help="--accuracy: rounding of 16-bit population stores")
    ap.add_argument("--out", default="accuracy.json")
    ap.add_argument("--accuracy-summary", nargs="+", metavar="RECORD",
                    help="fold --accuracy records into the committed accuracy summary (no GPU)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (counter calibration)")
    ap.add_argument("--headline-wall", default="mlp", choices=["newton", "mlp"],
                    help="wall model of the 16-bit storages in the full configuration (FP32 / FP64 use Newton)")
    ap.add_argument("--probe", action="store_true",
                    help="traffic-only ceiling of the step's access pattern (separate JSON line)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (the solver core is single-GPU; "
              "--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        copy_calibration()
        return
    if args.probe:
        traffic_probe()
        return
    if args.accuracy_summary:
        accuracy_summary(args.accuracy_summary, args.summary_out or ACCURACY_PROFILE)
        return
    if args.polar:
        polar_bench(args)
        return
    if args.accuracy:
        accuracy_run(args)
        return

    import foil_cfd_amd as fc
    from foil_cfd_amd import cases
    fc.load()
    t_setup = time.perf_counter()
    case = cases.config3_wing(args.aoa)
    if args.coarse_only and args.periodic_box:
        case = cases.periodic_box(case)
    log(f"[bench] case {case.name}: flags built in {time.perf_counter() - t_setup:.1f} s")

    storages = [s.strip() for s in args.storages.split(",") if s.strip()]
    results = {}
    for st in storages:
        if args.coarse_only:
            results[st] = run_storage(case, st, args.steps, args.warmup, args.per_call, args.kernel_steps,
                                      args.variant)
        else:
            wall = args.headline_wall if st in ("f16c", "f16", "bf16") else "newton"
            results[st] = run_full(case, st, args.steps, args.warmup, args.per_call, args.kernel_steps, args.aoa,
                                   args.variant, wall)
    head = next((h for h in ("f16c", "f16") if h in results), storages[0])
    r = results[head]
    cpu = None if args.no_cpu else cpu_baseline(case)

    ncells = case.dims[0] * case.dims[1] * case.dims[2]
    store = {"f16c": "fp16c_s1e4m11_shifted", "f16": "fp16_shifted", "bf16": "bf16_shifted", "f32": "fp32",
             "f64": "fp64"}[head]
    if args.coarse_only:
        workload = f"naca2412_wing_512x192x240_chord256_aoa{args.aoa:g}_trt_les_{head}store_coarse_only"
        if args.periodic_box:
            workload = f"periodic_box_512x192x240_trt_les_{head}store"
        config = {"workload": workload, "cells": ncells, "storage": store,
                  "not_included": "refinement patch, wall model, particles (coarse grid only)"}
        kernel = "stepKernel (fused stream-collide), coarse wing"
        traffic = measured_traffic(head)
        bytes_per_launch = BYTES_PER_CELL[head] * ncells
    else:
        fd = r["fine_dims"]
        workload = (f"naca2412_wing_512x192x240_chord256_aoa{args.aoa:g}_trt_les_imem_patch2x_"
                    f"particles1M_{head}store")
        config = {"workload": workload, "cells": ncells, "fine_cells": fd[0] * fd[1] * fd[2],
                  "fine_grid": fd, "patch_box_coarse": r["patch_box"], "refine_factor": 2,
                  "patch_margins_chords": dict(zip(("upstream", "wake", "below", "above"), cases.APP_PATCH_MARGINS_C)),
                  "cell_updates_per_step": r["cell_updates_per_step"], "storage": store,
                  "wall_model": ("iMEM Reichardt, coarse + fine wall cells; " +
                                 ("learned u_tau network (Newton fallback)" if r.get("wall_function") == "mlp"
                                  else "Newton u_tau") + " for the headline storage, Newton for FP32"),
                  "particles": cases.DEFAULT_PARTICLES}
        kernel = ("fine-level stream-collide launches (dominant; one of the two per coarse step also writes "
                  "the restriction's child-average records)")
        traffic = measured_traffic(head, "fine")
        bytes_per_launch = r["kernel_bytes"]
    config.update({"collision": "TRT(3/16) + Smagorinsky(0.1733)", "streaming": "in-place AA",
                   "rounding_16bit": "stochastic (hash of cell, direction, step)",
                   "step_variant": args.variant, "steps_per_call": min(args.per_call, args.steps),
                   "parallelism": "single_gpu"})
    out = {
        "metric": METRIC,
        "value": round(r["glups"], 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": {
            "bound": "hbm",
            "achieved": round(r["kernel_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["kernel_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
            "kernel": kernel,
            "kernel_ms": round(r["kernel_ms"], 5),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "algorithmic_bytes_per_cell": BYTES_PER_CELL[head],
        },
        "cpu_baseline": cpu,
        "per_storage": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv)
                            for kk, vv in v.items()} for k, v in results.items()},
        "reference_published_glups_fp32": REFERENCE_PUBLISHED_GLUPS_FP32,
    }
    out["cl_cd_err_vs_fp64"] = accuracy_line()
    out["traffic_probe_ceiling"] = probe_line()
    if head in ("f16c", "f16") and "f32" in results:
        out["fp16_vs_fp32_speedup"] = round(results[head]["glups"] / results["f32"]["glups"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
