#!/usr/bin/env python3
"""Benchmark: FP32 nonlinear forward modelling on one MI355X.

Workload (one "step" = one full shot): 400 x 400 x 250-cell seeded layered
model (padded 328 x 472 x 468), d = 12.5 m, explosive 10 Hz Ricker source,
2.5 s = 626 coarse samples x 8 fine steps (5000 fine steps), 200 x 200
surface receivers recording all 9 components. Inputs (medium, positions,
source traces) are resident in HBM before the timed region.

Prints one JSON line with the throughput in Gcell-updates/s, the roofline of
the stencil kernels (algorithmic bytes / HIP-event-timed launch duration vs
8 TB/s), the metric's second half -- FWI-gradient shots/s batched vs plain
on a bounded slice of the 64-shot windowed workload (`fwi_batch`: its first
4 shots, 0.5 s record; each path timed on the second of two gradient calls,
an FWI iteration's steady state, with the first call reported beside it) --
and the CPU baseline (the FP64 CPU restatement of
the scheme timed on a bounded slice of the 100^3 homogeneous workload on the
host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per FP32 cell-update of the two stencil kernels:
# reads of the other level (c), of the own level (o) and of the medium, own level written
BYTES_F32 = {
    # 8 scaled medium arrays: velocity 6 c + 3 o + 3 medium reads, 3 writes; stress 3 c + 6 o + 5 medium, 6
    "scaled": {"vel": 60, "stress": 80, "mix_vel": (12, 3), "mix_stress": (14, 6)},
    # medium formed in the steppers from rho (velocity) and lambda, mu (stress)
    "elastic": {"vel": 52, "stress": 68, "mix_vel": (10, 3), "mix_stress": (11, 6)},
}
METRIC = "Gcell-updates/s & %HBM peak; FWI-gradient shots/s batched vs plain, 1 GPU"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(shape, medium="scaled"):
    """HBM-side bytes per fine step of the stencil pair for launch shape
    (tile_x, y_chunk, order, x_group) and medium kind from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by the plain-copy calibration), or
    None if that shape was not profiled."""
    path = os.path.join(ROOT, "profiles", "stencil_layered_400_pmc.md")
    key = "traffic_bytes_per_step[{}{}x{}/{}/{}]:".format("" if medium == "scaled" else medium + " ", *shape)
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def gpu_setup(workload, precision="f32", medium="scaled"):
    import torch
    from elasticisolib_3d_amd import propagator as P
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    from elasticisolib_3d_amd.models import source_traces

    w = workload
    a = w.axes
    t0 = time.time()
    med = P.scale_medium(w.rho, w.lame, w.mu, w.dtw)
    src_geom = ShotGeometry(a, w.src_xyz[:, 2], w.src_xyz[:, 0], w.src_xyz[:, 1])
    rec_geom = ShotGeometry(a, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    src_reg = scale_regular_traces(med, src_geom, src_geom.to_regular(source_traces(w)), a.dx * a.dy * a.dz)
    log(f"[bench] host set-up {time.time() - t0:.1f}s; receiver points per grid {rec_geom.counts()}")
    prop = P.Propagator(a.nz, a.nx, a.ny, w.nts, w.sub, a.dz, a.dx, a.dy, w.dts, a.min_pad, w.alpha_cos, precision)
    dev = "cuda"
    dt = torch.float32 if precision == "f32" else torch.float64
    if medium == "elastic":   # (rho, lambda, mu): the steppers form the 8 scaled values
        prop.set_elastic(torch.tensor(np.stack([w.rho, w.lame, w.mu]), dtype=dt, device=dev), w.dtw)
    else:
        prop.set_medium(torch.tensor(med, dtype=dt, device=dev))
    del med
    spos = [torch.tensor(p, device=dev) for p in src_geom.positions()]
    rpos = [torch.tensor(p, device=dev) for p in rec_geom.positions()]
    src = [torch.tensor(s, dtype=dt, device=dev) for s in src_reg]
    rec = [torch.zeros((n, w.nts), dtype=dt, device=dev) for n in
           [rec_geom.counts()[g] for g in (1, 2, 3, 0, 0, 0, 4, 5, 6)]]
    torch.cuda.synchronize()
    return prop, spos, src, rpos, rec


def cpu_baseline(target_s: float = 15.0):
    """FP64 CPU restatement of the scheme on a bounded slice of the 100^3
    homogeneous workload (the first coarse samples of the 0.6 s record)."""
    from oracle import oracle as O
    from elasticisolib_3d_amd.models import homogeneous_100
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    from elasticisolib_3d_amd.models import source_traces

    O.build()
    w = homogeneous_100()
    a = w.axes

    def run(nts):
        g = O.Grid(a.nz, a.nx, a.ny, nts, w.sub, a.dz, a.dx, a.dy, w.dts, a.min_pad, w.alpha_cos)
        med = O.lib_scale_medium(w.rho, w.lame, w.mu, w.dtw)
        sg = ShotGeometry(a, w.src_xyz[:, 2], w.src_xyz[:, 0], w.src_xyz[:, 1])
        rg = ShotGeometry(a, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
        tr = [t[:, :nts] for t in source_traces(w)]
        src = scale_regular_traces(med, sg, sg.to_regular(tr), a.dx * a.dy * a.dz)
        t0 = time.perf_counter()
        O.lib_nl_forward(g, med, src, sg.positions(), rg.positions())
        return time.perf_counter() - t0

    probe_nts = 3
    t_probe = run(probe_nts)
    per_step = t_probe / ((probe_nts - 1) * w.sub)
    n_fine = max(8, int(target_s / max(per_step, 1e-6)))
    nts = min(w.nts, n_fine // w.sub + 1)
    t = run(nts)
    steps = (nts - 1) * w.sub
    cells = w.interior_cells * steps
    return {
        "value": round(cells / t / 1e9, 4),
        "unit": "Gcell-updates/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"FP64 CPU restatement, 100^3 homogeneous workload (padded {a.nz}x{a.nx}x{a.ny}), "
                   f"first {nts} of {w.nts} coarse samples = {steps} fine steps, {t:.1f} s"),
    }


def bench_born(args):
    """Born forward and adjoint (FP32) on the 400x400x250 model smoothed with a
    15-cell Gaussian, 1 % sparse reflectivity, 501 coarse samples, 150x150
    receivers; dot-product test of the shot-level operator (perturbation
    staggering + propagation) with FP64 inner products."""
    import torch
    from elasticisolib_3d_amd import propagator as P
    from elasticisolib_3d_amd.models import born_layered_400

    t0 = time.time()
    w = born_layered_400(nts=args.nts if args.nts != 626 else 501)
    log(f"[bench] model {w.name} padded {w.axes.nz}x{w.axes.nx}x{w.axes.ny} built in {time.time() - t0:.1f}s")
    prop, spos, src, rpos, rec = gpu_setup(w, "f32", args.medium)
    med = P.scale_medium(w.rho, w.lame, w.mu, w.dtw)
    dm3 = w.extra["dm3"]
    dm8 = P.born_stagger(dm3, med, w.dtw)
    dm8_d = torch.tensor(dm8, dtype=torch.float32, device="cuda")
    del dm8
    prop.set_born_storage(args.born_segment)
    tuned = None
    if not args.no_autotune:
        from elasticisolib_3d_amd.autotune import autotune
        tuned = autotune(prop, reps=5, target=args.tune_target)
        log(f"[bench] autotune ({args.tune_target}): tile {tuned['tile_x']}x{tuned['y_chunk']} order "
            f"{tuned['order']}/{tuned['x_group']}, transposed tile {prop.get_tiling_transposed()}")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.born_forward(spos, src, dm8_d, rpos, rec_out=rec)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    ms_f, steps_f, cells = prop.last_timing()
    log(f"[bench] Born forward {t_fwd:.1f} s ({steps_f} fine steps)")

    gen = torch.Generator(device="cuda").manual_seed(11)
    data = [torch.randn(r.shape, generator=gen, device="cuda", dtype=torch.float32) for r in rec]
    img = torch.zeros((8,) + tuple(w.rho.shape), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.born_adjoint(spos, src, data, rpos, img)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    ms_a, steps_a, _ = prop.last_timing()
    seg = prop.born_storage_used()
    log(f"[bench] Born adjoint {t_adj:.1f} s ({steps_a} fine steps, storage plan {seg})")

    lhs = float(sum(float((r.double() * d.double()).sum()) for r, d in zip(rec, data)))
    img_h = img.cpu().numpy().astype(np.float64)
    del img
    m_adj = P.born_unstagger(img_h, med, w.dtw)
    rhs = float(np.sum(dm3 * m_adj))
    dot_rel = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    log(f"[bench] dot test <B m, d> = {lhs:.9e}  <m, B* d> = {rhs:.9e}  rel {dot_rel:.2e}")
    gf = float(cells) * steps_f / t_fwd / 1e9
    ga = float(cells) * steps_a / t_adj / 1e9
    out = {
        "metric": "Born Gcell-updates/s (forward; adjoint in config)",
        "value": round(gf, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(t_fwd * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "FP32 Born forward + adjoint, one shot: 400x400x250 layered model smoothed sigma=15 cells "
                        "(padded 328x472x468), 1% sparse drho/dlambda/dmu N(0, 0.02 background), explosive 10 Hz "
                        f"Ricker, {w.nts} coarse samples x {w.sub} fine steps, 150x150 receivers",
            "interior_cells": cells,
            "forward_fine_steps": steps_f,
            "forward_s": round(t_fwd, 2),
            "adjoint_fine_steps": steps_a,
            "adjoint_s": round(t_adj, 2),
            "adjoint_gcell_updates_per_s": round(ga, 3),
            "adjoint_storage": "all coarse samples in HBM" if seg < 0 else f"exact checkpoints every {seg} coarse samples",
            "dot_product": {"lhs": lhs, "rhs": rhs, "rel_err": dot_rel, "reductions": "fp64"},
            "launch": None if tuned is None else {k: tuned[k] for k in ("tile_x", "y_chunk", "order", "x_group",
                                                                          "tile_x_transposed") if k in tuned},
        },
    }
    print(json.dumps(out), flush=True)


def bench_fp64_check(args):
    """FP32 vs FP64 GPU nonlinear forward on the 400x400x250 workload: relative
    L2 difference of all receiver traces (the FP64 run is the reference)."""
    import torch
    from elasticisolib_3d_amd import propagator as P
    from elasticisolib_3d_amd.models import layered_400

    w = layered_400(nts=args.nts)
    res = {}
    shapes = {}
    for prec in ("f64", "f32"):
        prop, spos, src, rpos, rec = gpu_setup(w, prec, args.medium)
        if not args.no_autotune:
            from elasticisolib_3d_amd.autotune import autotune
            t = autotune(prop, reps=5)
            shapes[prec] = f"{t['tile_x']}x{t['y_chunk']}/{t['order']}/{t['x_group']}"
        prop.set_graph(not args.no_graph)
        prop.nl_forward(spos, src, rpos, rec_out=rec)   # warm-up (code objects, graph capture)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prop.nl_forward(spos, src, rpos, rec_out=rec)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0   # the traces stay on the GPU until here
        res[prec] = ([r.double().cpu() for r in rec], elapsed)
        log(f"[bench] {prec} forward {res[prec][1]:.1f} s (launch shape {shapes.get(prec, 'default')})")
        del prop, spos, src, rpos, rec
        torch.cuda.empty_cache()
    num = sum(float(((a - b) ** 2).sum()) for a, b in zip(res["f32"][0], res["f64"][0]))
    den = sum(float((b ** 2).sum()) for b in res["f64"][0])
    err = (num / den) ** 0.5
    per = []
    for a, b in zip(res["f32"][0], res["f64"][0]):
        nb = float((b ** 2).sum()) ** 0.5
        per.append(float(((a - b) ** 2).sum()) ** 0.5 / nb if nb > 0 else 0.0)
    cells = float(w.interior_cells) * w.fine_steps
    out = {
        "metric": "FP32 vs FP64 receiver-trace relative L2 error",
        "value": err,
        "unit": "relative L2",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(res["f32"][1] * 1e3, 1),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 vs f64", "data": "synthetic",
        "config": {
            "workload": "nonlinear forward, 400x400x250 layered model (padded 328x472x468), "
                        f"{w.nts} coarse samples x {w.sub} fine steps, 200x200 receivers, 9 components",
            "medium": args.medium,
            "per_component_rel_err": [round(e, 9) for e in per],
            "f64_gcell_updates_per_s": round(cells / res["f64"][1] / 1e9, 3),
            "f32_gcell_updates_per_s": round(cells / res["f32"][1] / 1e9, 3),
            "f64_bytes_per_cell": 2 * (BYTES_F32[args.medium]["vel"] + BYTES_F32[args.medium]["stress"]),
            "f64_GBs_algorithmic": round(cells * 2 * (BYTES_F32[args.medium]["vel"] + BYTES_F32[args.medium]["stress"])
                                         / res["f64"][1] / 1e9, 1),
            "launch": shapes or None,
        },
    }
    print(json.dumps(out), flush=True)


def _fwi_shots(w, ax):
    from elasticisolib_3d_amd.acquisition import ShotGeometry
    from elasticisolib_3d_amd.fwi import FwiShot
    rec = ShotGeometry(ax, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    shots = []
    for xyz in w.extra["shots"]:
        src = ShotGeometry(ax, xyz[:, 2], xyz[:, 0], xyz[:, 1])
        sig = [np.zeros((1, w.nts)) for _ in range(9)]
        for c in (3, 4, 5):
            sig[c][0] = w.wavelet
        shots.append(FwiShot(src, rec, sig))
    return shots


def bench_fwi(args):
    """FWI misfit gradient (adjoint state) on the 256x256x160 workload: 4 shots,
    FP32 stencil, int16 block-scaled checkpoints (default) every 20 fine steps, FP64 gradient
    accumulation; one shot per launch sequence (the plain path). Accuracy of
    the FP32 gradient against the FP64 gradient (exact checkpoints) on shot 1,
    and central differences of the FP64 misfit on `--fd-cells` cells."""
    import torch
    from elasticisolib_3d_amd.fwi import FwiProblem
    from elasticisolib_3d_amd.models import fwi_256

    t0 = time.time()
    w = fwi_256(nts=args.nts if args.nts != 626 else 501)
    ax = w.axes
    start = w.extra["start"]
    log(f"[bench] model {w.name} padded {ax.nz}x{ax.nx}x{ax.ny} built in {time.time() - t0:.1f}s")
    shots = _fwi_shots(w, ax)
    prob = FwiProblem(ax, w.nts, w.dts, w.sub, w.alpha_cos, shots, "f32", interval=20, checkpoint=args.checkpoint,
                      derivatives=args.derivatives,
                      autotune=not args.no_autotune and not args.launch, medium=args.medium, graph=not args.no_graph)
    if args.launch:
        from elasticisolib_3d_amd.autotune import apply, _parse_shapes
        tx, yc, order, xg = _parse_shapes(args.launch)[0]
        given = {"tile_x": tx, "y_chunk": yc, "order": order, "x_group": xg}
        apply(prob.prop, given)
    t0 = time.perf_counter()
    prob.synthesize(w.rho, w.lame, w.mu)
    log(f"[bench] observed data of {len(shots)} shots synthesized in {time.perf_counter() - t0:.1f}s")
    times = []
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        J, g3 = prob.misfit_and_gradient(*start)
        times.append(time.perf_counter() - t0)
        log(f"[bench] gradient {i + 1}: J = {J:.6e}, {times[-1]:.1f}s ({prob.last_timing})")
    plan = prob.prop.gradient_plan_used()
    drv_fmt = prob.prop.gradient_derivatives_used()
    tuned = dict(given, given=True) if args.launch else prob.tuned
    del prob
    torch.cuda.empty_cache()
    t_plain = min(times)
    shots_per_s = len(shots) / t_plain
    cells = float(w.interior_cells)
    steps = w.fine_steps

    acc = {}
    if args.fd_cells >= 0:
        one = shots[:1]
        p64 = FwiProblem(ax, w.nts, w.dts, w.sub, w.alpha_cos, one, "f64", interval=20, checkpoint="native")
        p64.synthesize(w.rho, w.lame, w.mu)
        t0 = time.perf_counter()
        J64, g64 = p64.misfit_and_gradient(*start)
        log(f"[bench] FP64 gradient (shot 1): J = {J64:.9e}, {time.perf_counter() - t0:.1f}s, "
            f"plan {p64.prop.gradient_plan_used()}")
        rng = np.random.default_rng(3)
        score = np.abs(g64 * start).ravel()
        top = np.argsort(score)[-max(1, score.size // 10):]
        pick = rng.choice(top, size=args.fd_cells, replace=False) if args.fd_cells else []
        fd_rows = []
        h = 1e-3
        for q in pick:
            k, rest = divmod(int(q), start[0].size)
            idx = np.unravel_index(rest, start[0].shape)
            mp = [m.copy() for m in start]
            mm = [m.copy() for m in start]
            mp[k][idx] *= 1 + h
            mm[k][idx] *= 1 - h
            fd = (p64.misfit(*mp) - p64.misfit(*mm)) / (2 * h * start[k][idx])
            fd_rows.append((k, idx, fd, float(g64[k][idx])))
            log(f"[bench] FD cell param {k} {tuple(int(i) for i in idx)}: FD {fd:.6e} FP64 adjoint {g64[k][idx]:.6e}")
        del p64
        torch.cuda.empty_cache()
        res = {}
        # (checkpoints, derivative combinations): each checkpoint format with the derivative storage of
        # the run, both exact, and exact checkpoints with 16-bit derivatives alone
        d16 = "s16"
        for name, ck, drv in (("f32_bf16", "bf16", args.derivatives), ("f32_s16", "s16", args.derivatives),
                              ("f32_exact", "native", "native"), (f"f32_exact_ckpt_{d16}_drv", "native", d16)):
            p = FwiProblem(ax, w.nts, w.dts, w.sub, w.alpha_cos, one, "f32", interval=20, checkpoint=ck,
                           derivatives=drv)
            t0 = time.perf_counter()
            Jx, gx = p.misfit_and_gradient(*start)
            res[name] = (Jx, gx, time.perf_counter() - t0, p.prop.gradient_plan_used(),
                         p.prop.gradient_derivatives_used())
            log(f"[bench] {name} gradient (shot 1): J = {Jx:.9e}, {res[name][2]:.1f}s, plan {res[name][3]}")
            del p
            torch.cuda.empty_cache()

        def rel(a, b):
            return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))

        for name, (Jx, gx, tx, px, dx) in res.items():
            acc[name] = {
                "grad_rel_l2_vs_fp64": rel(gx, g64),
                "grad_rel_l2_per_param": [rel(gx[k], g64[k]) for k in range(3)],
                "misfit_rel_vs_fp64": abs(Jx - J64) / J64,
                "checkpoint_plan": {"interval": px[0], "outer_interval": px[1], "format": px[2],
                                    "checkpoint_GB": round(px[3] / 1e9, 2)},
                "derivatives": dx,
            }
            if fd_rows:
                errs = [abs(float(gx[k][idx]) - fd) / abs(fd) for (k, idx, fd, _) in fd_rows]
                acc[name]["fd_rel_err_max"] = max(errs)
                acc[name]["fd_rel_err_median"] = float(np.median(errs))
        if fd_rows:
            e64 = [abs(g - fd) / abs(fd) for (_, _, fd, g) in fd_rows]
            acc["f64_exact"] = {"fd_rel_err_max": max(e64), "fd_rel_err_median": float(np.median(e64)),
                                "fd_cells": len(fd_rows), "fd_h_relative": h}

    out = {
        "metric": "FWI-gradient shots/s (plain: one shot per launch sequence)",
        "value": round(shots_per_s, 4),
        "unit": "shots/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": 0,
        "ms_per_step": round(t_plain * 1e3, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f32 stencil, {args.checkpoint} checkpoints, f64 gradient accumulation",
        "data": "synthetic",
        "config": {
            "workload": "FWI gradient, 256x256x160 layered model (padded 232x328x324), d=15 m, start = true smoothed "
                        f"sigma=20 cells, {len(shots)} explosive shots (8 Hz Ricker), {w.nts} coarse samples x {w.sub} "
                        f"fine steps, {len(w.rec_xyz)} surface receivers, L2 misfit, checkpoints every 20 fine steps",
            "gradient_s_per_call": [round(t, 2) for t in times],
            "derivatives": drv_fmt,
            "checkpoint_plan": {"interval": plan[0], "outer_interval": plan[1], "format": plan[2],
                                "checkpoint_GB_per_shot": round(plan[3] / 1e9, 2)},
            "launch": None if tuned is None else {k: tuned[k] for k in ("tile_x", "y_chunk", "order", "x_group",
                                                                          "tile_x_transposed") if k in tuned},
            "interior_cells": w.interior_cells,
            "fine_steps_per_propagation": steps,
            "propagation_gcell_updates_per_s": round(3 * cells * steps * len(shots) / t_plain / 1e9, 2),
            "accuracy_shot1": acc,
        },
    }
    print(json.dumps(out), flush=True)


def fwi_batch_compare(n_shots, batch, nts, checkpoint, autotune, full_shots, smooth_device=None):
    """Batched vs plain FWI gradient with shot-local windows on the 400x400x250
    layered model: the first `n_shots` shots of the 8x8 jittered grid, plain =
    one shot per launch sequence, batched = `batch` windows stacked in one grid
    per launch sequence (same kernels, same checkpointing); optionally
    `full_shots` shots over the whole padded model, one per launch sequence
    (the reference's layout, whose shot window option is unused). Returns
    (result dict, log lines)."""
    import torch
    from elasticisolib_3d_amd.batch import BatchShot, BatchedFwi, Window
    from elasticisolib_3d_amd.models import batch_layered_400

    t0 = time.time()
    cfg = batch_layered_400(nts=nts, smooth_device=smooth_device)
    log(f"[bench] FWI models built in {time.time() - t0:.1f}s")
    nts = cfg["nts"]
    shots = []
    for src, rec, (x0, y0, wx, wy) in cfg["shots"][:n_shots]:
        sig = [np.zeros((1, nts)) for _ in range(9)]
        for c in (3, 4, 5):
            sig[c][0] = cfg["wavelet"]
        shots.append(BatchShot(src, rec, Window(x0, y0, wx, wy), sig))
    shape = cfg["true"][0].shape
    # progress every 8th batch on the log (long runs stay visibly alive)
    kw = dict(precision="f32", interval=20, checkpoint=checkpoint, autotune=autotune,
              progress=lambda m: log(f"[bench] {m}") if any(f"batch {k}/" in m for k in range(8, 1000, 8)) else None)
    ref = BatchedFwi(shape, cfg["d"], cfg["pad"], nts, cfg["dts"], cfg["sub"], cfg["alpha_cos"], shots, batch=1, **kw)
    t0 = time.perf_counter()
    ref.synthesize(cfg["true"])
    log(f"[bench] observed data of {len(shots)} shots in {time.perf_counter() - t0:.1f}s")
    del ref
    res = {}
    first = {}
    for name, b in (("plain", 1), ("batched", batch)):
        op = BatchedFwi(shape, cfg["d"], cfg["pad"], nts, cfg["dts"], cfg["sub"], cfg["alpha_cos"], shots, batch=b, **kw)
        # two calls, as two FWI iterations: the first includes the one-time set-up of the persistent
        # propagator (allocation, launch-shape tuning, graph capture), the second is the steady state
        J, g = op.misfit_and_gradient(cfg["start"])
        first[name] = dict(op.timing)
        log(f"[bench] {name} (batch {b}) first call: J = {J:.9e}, {op.timing}")
        J, g = op.misfit_and_gradient(cfg["start"])
        res[name] = (J, g, op.timing)
        log(f"[bench] {name} (batch {b}) second call: J = {J:.9e}, {op.timing}")
        del op
        torch.cuda.empty_cache()
    gp, gb = res["plain"][1], res["batched"][1]
    diff = float(np.linalg.norm(gb - gp) / np.linalg.norm(gp))
    tp, tb = res["plain"][2]["total_s"], res["batched"][2]["total_s"]
    full = {}
    if full_shots > 0:
        # the reference's layout: every shot propagated over the whole padded model, one shot per
        # launch sequence (timing only: the observed traces are the windowed ones)
        fshots = [BatchShot(s.src_xyz, s.rec_xyz, Window(0, 0, shape[1], shape[0]), s.signals, s.observed)
                  for s in shots[:full_shots]]
        op = BatchedFwi(shape, cfg["d"], cfg["pad"], nts, cfg["dts"], cfg["sub"], cfg["alpha_cos"], fshots,
                        batch=1, **kw)
        op.misfit_and_gradient(cfg["start"])
        tf1 = op.timing["total_s"]
        log(f"[bench] full grid (batch 1, {len(fshots)} shots) first call: {op.timing}")
        op.misfit_and_gradient(cfg["start"])
        tf = op.timing["total_s"]
        log(f"[bench] full grid (batch 1, {len(fshots)} shots) second call: {op.timing}")
        full = {"whole_model_shots": len(fshots),
                "whole_model_shots_per_s": round(len(fshots) / tf, 4),
                "whole_model_s": round(tf, 2),
                "whole_model_propagation_s": round(op.timing["propagation_s"], 2),
                "speedup_batched_over_whole_model": round((tf / len(fshots)) / (tb / len(shots)), 3),
                "whole_model_first_call_s": round(tf1, 2)}
        del op
        torch.cuda.empty_cache()
    wins = [s.window for s in shots]
    return {
        "workload": f"FWI gradient of {len(shots)} shots with shot-local windows on the 400x400x250 layered model "
                    f"(d=12.5 m, 30-cell sponge per window), explosive 10 Hz Ricker, {nts} coarse samples x "
                    f"{cfg['sub']} fine steps, FP32 stencil, {checkpoint} checkpoints every 20 fine steps, "
                    f"FP64 gradient accumulation",
        "windows": [[w.x0, w.y0, w.nx, w.ny] for w in wins],
        "receivers": [int(len(s.rec_xyz)) for s in shots],
        "timing": "second of two calls on the same operator (an FWI iteration's steady state); the first call "
                  "also allocates, tunes and captures the persistent propagator (first_call_*)",
        "plain_shots_per_s": round(len(shots) / tp, 4),
        "batched_shots_per_s": round(len(shots) / tb, 4),
        "first_call_plain_s": round(first["plain"]["total_s"], 2),
        "first_call_batched_s": round(first["batched"]["total_s"], 2),
        "batch": batch,
        "speedup_batched_over_plain": round(tp / tb, 3),
        "plain_s": round(tp, 2), "batched_s": round(tb, 2),
        "plain_propagation_s": round(res["plain"][2]["propagation_s"], 2),
        "batched_propagation_s": round(res["batched"][2]["propagation_s"], 2),
        "gradient_rel_diff_batched_vs_plain": diff,
        **full,
    }


def bench_fwi_batch(args):
    """Batched vs plain FWI gradient with shot-local windows on the 400x400x250
    layered model (fwi_batch_compare with the command-line sizes)."""
    r = fwi_batch_compare(args.shots, args.batch, args.nts, args.checkpoint, not args.no_autotune, args.full_shots,
                          "cuda")
    out = {
        "metric": "FWI-gradient shots/s, batched vs plain (shot-local windows)",
        "value": r["batched_shots_per_s"],
        "unit": "shots/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(r["batched_s"] * 1e3, 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": f"f32 stencil, {args.checkpoint} checkpoints, f64 gradient accumulation",
        "data": "synthetic",
        "config": r,
    }
    print(json.dumps(out), flush=True)


def bench_hbm(args):
    """HBM rates of plain streaming mixes (ela_stream_probe): reads only and
    read / in-place write mixes, on arrays of the 400x400x250 grid's size
    (290 MB, partly held by the 256 MB Infinity Cache) and of 1 GiB (HBM's
    own rate). The stencil's roofline context: what HBM delivers for a
    byte mix with writes."""
    import torch
    from elasticisolib_3d_amd.propagator import Propagator

    prop = Propagator(40, 40, 40, 4, 2, 10.0, 10.0, 10.0, 0.002, 10, 0.95, "f32")
    mixes = [(2, 0), (12, 0), (1, 1), (10, 3), (11, 6), (12, 3), (14, 6), (12, 9)]
    sizes = {"grid_400_290MB": 72453888, "1GiB": 1 << 28}
    table = {}
    for (r, wr) in mixes:
        row = {}
        for name, n in sizes.items():
            row[name] = round(prop.stream_probe(r, wr, n, reps=10), 1)
        table[f"{r}r_{wr}w"] = row
        log(f"[bench] {r:2d} reads {wr} writes: " + ", ".join(f"{k} {v:.0f} GB/s" for k, v in row.items()))
    prop.close()
    torch.cuda.synchronize()
    with_writes = [v["1GiB"] for k, v in table.items() if not k.endswith("_0w")]
    out = {
        "metric": "HBM GB/s of streaming read / write mixes",
        "value": round(min(with_writes), 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": 10,
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (zero-filled arrays)",
        "config": {"workload": "plain streaming kernels: R arrays read, the first W rewritten in place, 16-byte "
                               "accesses; value = the slowest mix with writes on 1 GiB arrays",
                   "peak": HBM_PEAK_GBS, "mixes": table,
                   "mix_with_writes_range_1GiB": [min(with_writes), max(with_writes)]},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed shots")
    ap.add_argument("--warmup", type=int, default=1, help="untimed shots")
    ap.add_argument("--nts", type=int, default=626, help="coarse samples (626 = 2.5 s)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--launch", default="", help="fixed launch shape tile_x x y_chunk / order / x_group, "
                                                 "e.g. 4x8/3/4 (no autotuning)")
    ap.add_argument("--no-graph", action="store_true", help="launch every fine step from the host")
    ap.add_argument("--tune-target", choices=("forward", "gradient"), default="gradient",
                    help="born: rank launch shapes by the nonlinear forward step, or (default) by the gradient "
                         "step's kernel pairs -- forward, forward with extra per-cell streams, transposed -- "
                         "the kernel mix of the Born passes (also picks the transposed steps' own tile)")
    ap.add_argument("--medium", choices=("elastic", "scaled"), default="elastic",
                    help="nl_forward, born, fwi, fp64_check: rho, lambda, mu with the medium formed in the steppers (120 B per cell-update), "
                         "or the 8 scaled arrays of the reference's interface (140 B)")
    ap.add_argument("--no-stream-ref", action="store_true",
                    help="skip the streaming-kernel reference of the roofline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (for counter calibration)")
    ap.add_argument("--workload", choices=("nl_forward", "born", "fp64_check", "fwi", "fwi_batch", "hbm"),
                    default="nl_forward",
                    help="nl_forward (default): the headline line; born: Born forward/adjoint and dot test; "
                         "fp64_check: FP32 vs FP64 traces on the forward workload; fwi: FWI gradient shots/s "
                         "and gradient accuracy; hbm: HBM rates of streaming read / write mixes")
    ap.add_argument("--shots", type=int, default=4, help="fwi_batch: number of shots")
    ap.add_argument("--fwi-slice-shots", type=int, default=4,
                    help="nl_forward: shots of the batched-vs-plain FWI slice added to the line (0: skip)")
    ap.add_argument("--fwi-slice-nts", type=int, default=126,
                    help="nl_forward: coarse samples of that slice (126 = 0.5 s, 1000 fine steps)")
    ap.add_argument("--fwi-slice-full", type=int, default=1,
                    help="nl_forward: shots of that slice also run over the whole padded model")
    ap.add_argument("--batch", type=int, default=2, help="fwi_batch: shots per batched launch sequence")
    ap.add_argument("--full-shots", type=int, default=1,
                    help="fwi_batch: shots also timed over the whole padded model, one per launch sequence "
                         "(the reference's layout); 0 skips them")
    ap.add_argument("--fd-cells", type=int, default=16,
                    help="fwi: cells checked by central differences (-1: skip the accuracy part)")
    ap.add_argument("--checkpoint", choices=("bf16", "s16", "native"), default="s16",
                    help="fwi, fwi_batch: storage of the gradient's outer checkpoints")
    ap.add_argument("--derivatives", choices=("s16", "native"), default=None,
                    help="fwi: storage of the backward's derivative combinations (default: the library's, exact)")
    ap.add_argument("--born-segment", type=int, default=0,
                    help="Born adjoint storage: 0 automatic, -1 all samples, K checkpoint segment")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "shots would be distributed one process per GPU, which this build does not launch.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "born":
        return bench_born(args)
    if args.workload == "fp64_check":
        return bench_fp64_check(args)
    if args.workload == "fwi":
        return bench_fwi(args)
    if args.workload == "hbm":
        return bench_hbm(args)
    if args.workload == "fwi_batch":
        return bench_fwi_batch(args)

    import torch
    from elasticisolib_3d_amd.models import layered_400

    t0 = time.time()
    w = layered_400(nts=args.nts)
    log(f"[bench] model {w.name} padded {w.axes.nz}x{w.axes.nx}x{w.axes.ny} built in {time.time() - t0:.1f}s")
    prop, spos, src, rpos, rec = gpu_setup(w, "f32", args.medium)
    by = BYTES_F32[args.medium]
    bpc = by["vel"] + by["stress"]
    tuned = None
    if args.launch:
        from elasticisolib_3d_amd.autotune import apply, _parse_shapes
        tx, yc, order, xg = _parse_shapes(args.launch)[0]
        apply(prop, {"tile_x": tx, "y_chunk": yc, "order": order, "x_group": xg})
        tuned = {"tile_x": tx, "y_chunk": yc, "order": order, "x_group": xg, "given": True}
        log(f"[bench] launch shape {tx}x{yc} order {order}/{xg} (given)")
    elif not args.no_autotune:
        from elasticisolib_3d_amd.autotune import autotune
        t0 = time.time()
        tuned = autotune(prop, reps=10)
        log(f"[bench] autotune: tile {tuned['tile_x']}x{tuned['y_chunk']} order {tuned['order']}/{tuned['x_group']} "
            f"({tuned['ms_step']:.4f} ms/step, {'cached' if tuned['cached'] else f'{time.time() - t0:.1f}s'})")
    prop.set_graph(not args.no_graph)

    for i in range(args.warmup):
        prop.nl_forward(spos, src, rpos, rec_out=rec)
        torch.cuda.synchronize()
        log(f"[bench] warmup shot {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_ms = 0.0
    for i in range(args.steps):
        prop.nl_forward(spos, src, rpos, rec_out=rec)
        ms, steps_done, cells = prop.last_timing()
        ev_ms += ms
        log(f"[bench] timed shot {i + 1}/{args.steps}: loop {ms:.1f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cell_updates = float(w.interior_cells) * w.fine_steps * args.steps
    gcells = cell_updates / wall / 1e9

    # per-kernel launch durations with HIP events on the propagator's stream
    kv, ks = prop.time_step_kernels(args.kernel_reps)
    step_ms = kv + ks
    shape = (*prop.get_tiling(), *((tuned["order"], tuned["x_group"]) if tuned else (0, 1)))
    ach = bpc * w.interior_cells / (step_ms * 1e-3) / 1e9
    # streaming reference: plain kernels with the same read / in-place write
    # mixes over arrays of the padded grid's size (no neighbour accesses)
    ceiling = None
    if not args.no_stream_ref:
        n_pad = w.axes.nz * w.axes.nx * w.axes.ny // 4 * 4
        (rv, wv), (rs, wsn) = by["mix_vel"], by["mix_stress"]
        c_v = prop.stream_probe(rv, wv, n_pad)
        c_s = prop.stream_probe(rs, wsn, n_pad)
        step_ref = bpc / (by["vel"] / c_v + by["stress"] / c_s)
        ceiling = {f"streaming_GBs_{rv}r_{wv}w": round(c_v, 1), f"streaming_GBs_{rs}r_{wsn}w": round(c_s, 1),
                   "streaming_GBs_step_mix": round(step_ref, 1), "frac_of_streaming": round(ach / step_ref, 4)}
        log(f"[bench] streaming reference: {rv}r/{wv}w {c_v:.0f} GB/s, {rs}r/{wsn}w {c_s:.0f} GB/s, step mix "
            f"{step_ref:.0f} GB/s; stencil at {ach / step_ref:.1%} of it")
    finite = all(bool(torch.isfinite(r).all()) for r in rec)
    norm = float(sum(float((r.double() ** 2).sum()) for r in rec) ** 0.5)

    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "FP32 nonlinear forward modelling, one shot per step: seeded 12-layer model 400x400x250 "
                        "cells (padded 328x472x468, 30-cell sponge), d=12.5 m, explosive 10 Hz Ricker, "
                        f"{w.nts} coarse samples x {w.sub} fine steps, 200x200 surface receivers, 9 components",
            "interior_cells": w.interior_cells,
            "medium": ("rho, lambda, mu; the 8 staggered, scaled values formed in the steppers" if args.medium == "elastic"
                       else "8 scaled medium arrays"),
            "fine_steps": w.fine_steps,
            "event_loop_ms_per_step": round(ev_ms / args.steps, 2),
            "data_finite": finite,
            "data_l2": norm,
            "launch": {"tile_x": prop.get_tiling()[0], "y_chunk": prop.get_tiling()[1],
                       "order": None if tuned is None else [tuned["order"], tuned["x_group"]],
                       "autotuned": tuned is not None and not tuned.get("given"), "graph": not args.no_graph},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_velocity + k_stress (one fine step)",
            "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(shape, args.medium),
            "bytes_per_cell": bpc,
            "medium": args.medium,
            "k_velocity_ms": round(kv, 4),
            "k_stress_ms": round(ks, 4),
            "k_velocity_GBs": round(by["vel"] * w.interior_cells / (kv * 1e-3) / 1e9, 1),
            "k_stress_GBs": round(by["stress"] * w.interior_cells / (ks * 1e-3) / 1e9, 1),
            "streaming_reference": ceiling,
        },
    }
    del prop, spos, src, rpos, rec
    torch.cuda.empty_cache()
    if args.fwi_slice_shots > 0:
        # the metric's second half on a bounded slice of the 64-shot workload
        t0 = time.time()
        out["fwi_batch"] = fwi_batch_compare(args.fwi_slice_shots, min(args.batch, args.fwi_slice_shots),
                                             args.fwi_slice_nts, args.checkpoint, not args.no_autotune,
                                             args.fwi_slice_full, "cuda")
        out["fwi_batch"]["wall_s"] = round(time.time() - t0, 1)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # reported, never silently replaced
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
