"""The library on the five benchmark configurations at their real sizes.

C1  100^3 homogeneous model, point force, the full 0.6 s record (151 coarse
    samples, 1200 fine steps, 81 receivers): FP64 and FP32 against the FP64
    CPU oracle.
C2  400x400x250 layered model, explosive source, the full 2.5 s record (5000
    fine steps, 200x200 receivers, 9 components): FP32 against FP64 on the GPU
    (the bar for FP32 forward data is 1e-5).
C3  the Born pair on the 400x400x250 model smoothed with a 15-cell Gaussian,
    1 % sparse perturbations, 150x150 receivers, record shortened to 0.5 s
    (126 coarse samples, 1000 fine steps): FP32 dot test of the shot-level
    operator (perturbation staggering + propagation) with FP64 inner
    products, bar 1e-6 (the FP64 pair at this size: tests/test_gpu_fullsize.py).
C4  256x256x160 FWI gradient, record shortened to 0.5 s (126 coarse samples,
    1000 fine steps): FP32 with the default int16 checkpoints against the FP64
    gradient with exact checkpoints (bar 1e-3), and the FP64 gradient against
    central differences of the FP64 misfit on 4 cells.
C5  the 64 shot-local windows of the 400x400x250 model with their receiver
    counts, record shortened to 6 coarse samples (40 fine steps): each shot's
    receiver traces and gradient from 2-window stacks (batched) equal its
    propagation alone (plain), FP32 with exact checkpoints.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    num = sum(float(((x - y) ** 2).sum()) for x, y in zip(a, b))
    den = sum(float((y ** 2).sum()) for y in b)
    return (num / den) ** 0.5


# ---- C1 -------------------------------------------------------------------

@pytest.fixture(scope="module")
def c1():
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    from elasticisolib_3d_amd.models import homogeneous_100, source_traces
    w = homogeneous_100()
    a = w.axes
    g = O.Grid(a.nz, a.nx, a.ny, w.nts, w.sub, a.dz, a.dx, a.dy, w.dts, a.min_pad, w.alpha_cos)
    med = O.lib_scale_medium(w.rho, w.lame, w.mu, w.dtw)
    sg = ShotGeometry(a, w.src_xyz[:, 2], w.src_xyz[:, 0], w.src_xyz[:, 1])
    rg = ShotGeometry(a, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    src = scale_regular_traces(med, sg, sg.to_regular(source_traces(w)), a.dx * a.dy * a.dz)
    print(f"C1: oracle, {a.nz}x{a.nx}x{a.ny}, {(w.nts - 1) * w.sub} fine steps, {O.num_threads()} threads",
          flush=True)
    ref = O.lib_nl_forward(g, med, src, sg.positions(), rg.positions())
    return w, g, med, src, sg, rg, ref


@pytest.mark.parametrize("precision,medium,tol", [("f64", "scaled", 1e-11), ("f32", "scaled", 5e-6),
                                                  ("f32", "elastic", 5e-6)])
def test_c1_full_record_against_the_oracle(c1, precision, medium, tol):
    from elasticisolib_3d_amd.propagator import Propagator
    w, g, med, src, sg, rg, ref = c1
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)
    if medium == "elastic":
        prop.set_elastic(np.stack([w.rho, w.lame, w.mu]), w.dtw)
    else:
        prop.set_medium(med)
    prop.set_graph(True)
    out = prop.nl_forward(sg.positions(), src, rg.positions())
    e = _rel(out, ref)
    print(f"C1 {precision} {medium}: rel L2 {e:.2e} over {len(rg.positions()[1])} vx points x {g.nts} samples",
          flush=True)
    assert e < tol
    assert float(np.abs(out[2]).max()) > 0


# ---- C2 -------------------------------------------------------------------

def test_c2_full_record_fp32_against_fp64():
    import torch
    from elasticisolib_3d_amd import propagator as P
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    from elasticisolib_3d_amd.models import layered_400, source_traces
    w = layered_400()
    a = w.axes
    med = P.scale_medium(w.rho, w.lame, w.mu, w.dtw)
    sg = ShotGeometry(a, w.src_xyz[:, 2], w.src_xyz[:, 0], w.src_xyz[:, 1])
    rg = ShotGeometry(a, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    src = scale_regular_traces(med, sg, sg.to_regular(source_traces(w)), a.dx * a.dy * a.dz)
    del med
    rlm = np.stack([w.rho, w.lame, w.mu])
    out = {}
    for precision in ("f64", "f32"):
        dt = torch.float64 if precision == "f64" else torch.float32
        print(f"C2 {precision}: {a.nz}x{a.nx}x{a.ny}, {w.fine_steps} fine steps", flush=True)
        prop = P.Propagator(a.nz, a.nx, a.ny, w.nts, w.sub, a.dz, a.dx, a.dy, w.dts, a.min_pad, w.alpha_cos, precision)
        prop.set_elastic(torch.tensor(rlm, dtype=dt, device="cuda"), w.dtw)
        prop.set_graph(precision == "f32")
        spos = [torch.tensor(p, device="cuda") for p in sg.positions()]
        rpos = [torch.tensor(p, device="cuda") for p in rg.positions()]
        rec = prop.nl_forward(spos, [torch.tensor(s, dtype=dt, device="cuda") for s in src], rpos)
        torch.cuda.synchronize()
        out[precision] = [r.double() for r in rec]
        del prop, rec
        torch.cuda.empty_cache()
    e = _rel(out["f32"], out["f64"])
    per = [_rel([x], [y]) for x, y in zip(out["f32"], out["f64"])]
    print(f"C2 FP32 vs FP64: rel L2 {e:.2e}, per component {np.round(per, 8)}", flush=True)
    assert e < 1e-5
    assert all(np.isfinite(per))


# ---- C3 -------------------------------------------------------------------

def test_c3_born_dot_test_fp32_shortened_record():
    import torch
    from elasticisolib_3d_amd import propagator as P
    from elasticisolib_3d_amd.acquisition import ShotGeometry, scale_regular_traces
    from elasticisolib_3d_amd.models import born_layered_400, source_traces
    w = born_layered_400(nts=126)
    a = w.axes
    med = P.scale_medium(w.rho, w.lame, w.mu, w.dtw)
    sg = ShotGeometry(a, w.src_xyz[:, 2], w.src_xyz[:, 0], w.src_xyz[:, 1])
    rg = ShotGeometry(a, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    src = scale_regular_traces(med, sg, sg.to_regular(source_traces(w)), a.dx * a.dy * a.dz)
    dm3 = w.extra["dm3"]
    dm8 = P.born_stagger(dm3, med, w.dtw)
    print(f"C3 FP32: {a.nz}x{a.nx}x{a.ny}, {w.fine_steps} fine steps, {len(rg.positions()[1])} vx points", flush=True)
    prop = P.Propagator(a.nz, a.nx, a.ny, w.nts, w.sub, a.dz, a.dx, a.dy, w.dts, a.min_pad, w.alpha_cos, "f32")
    prop.set_elastic(torch.tensor(np.stack([w.rho, w.lame, w.mu]), dtype=torch.float32, device="cuda"), w.dtw)
    spos = [torch.tensor(p, device="cuda") for p in sg.positions()]
    rpos = [torch.tensor(p, device="cuda") for p in rg.positions()]
    srcd = [torch.tensor(x, dtype=torch.float32, device="cuda") for x in src]
    rec = prop.born_forward(spos, srcd, torch.tensor(dm8, dtype=torch.float32, device="cuda"), rpos)
    gen = torch.Generator(device="cuda").manual_seed(11)
    data = [torch.randn(r.shape, generator=gen, device="cuda", dtype=torch.float32) for r in rec]
    img = torch.zeros((8,) + tuple(w.rho.shape), dtype=torch.float32, device="cuda")
    prop.born_adjoint(spos, srcd, data, rpos, img)
    torch.cuda.synchronize()
    lhs = float(sum(float((r.double() * d.double()).sum()) for r, d in zip(rec, data)))
    rhs = float(np.sum(dm3 * P.born_unstagger(img.cpu().numpy().astype(np.float64), med, w.dtw)))
    e = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"C3 FP32 dot test: <B m, d> = {lhs:.9e}, <m, B* d> = {rhs:.9e}, rel {e:.2e} "
          f"(storage plan {prop.born_storage_used()})", flush=True)
    del prop, rec, data, img
    torch.cuda.empty_cache()
    assert abs(lhs) > 0 and e < 1e-6


# ---- C4 -------------------------------------------------------------------

def test_c4_gradient_shortened_record():
    import torch
    from elasticisolib_3d_amd.acquisition import ShotGeometry
    from elasticisolib_3d_amd.fwi import FwiProblem, FwiShot
    from elasticisolib_3d_amd.models import fwi_256
    w = fwi_256(nts=126, smooth_device="cuda")
    ax = w.axes
    start = w.extra["start"]
    rec = ShotGeometry(ax, w.rec_xyz[:, 2], w.rec_xyz[:, 0], w.rec_xyz[:, 1])
    xyz = w.extra["shots"][0]
    src = ShotGeometry(ax, xyz[:, 2], xyz[:, 0], xyz[:, 1])
    sig = [np.zeros((1, w.nts)) for _ in range(9)]
    for c in (3, 4, 5):
        sig[c][0] = w.wavelet
    res = {}
    for name, prec, ck in (("f64_exact", "f64", "native"), ("f32_default", "f32", None)):
        kw = {} if ck is None else {"checkpoint": ck}
        p = FwiProblem(ax, w.nts, w.dts, w.sub, w.alpha_cos, [FwiShot(src, rec, sig)], prec, interval=20, **kw)
        p.synthesize(w.rho, w.lame, w.mu)
        print(f"C4 {name}: {ax.nz}x{ax.nx}x{ax.ny}, {w.fine_steps} fine steps", flush=True)
        J, g3 = p.misfit_and_gradient(*start)
        res[name] = (p, J, g3)
        if name == "f32_default":
            assert p.prop.gradient_plan_used()[2] == "s16"
            del p
            torch.cuda.empty_cache()
    g64, g32 = res["f64_exact"][2], res["f32_default"][2]
    e = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    per = [float(np.linalg.norm(g32[k] - g64[k]) / np.linalg.norm(g64[k])) for k in range(3)]
    print(f"C4 FP32 (int16 checkpoints) vs FP64 gradient: rel L2 {e:.2e} (rho, lambda, mu {np.round(per, 7)})",
          flush=True)
    # FP64 gradient against central differences of the FP64 misfit on 4 illuminated cells: the
    # shortened record illuminates only part of the grid, so each cell is drawn from the 200 cells
    # of largest |g m| of its own parameter (rho, lambda, mu, rho)
    p64 = res["f64_exact"][0]
    rng = np.random.default_rng(3)
    h = 1e-3
    checks = []
    for k in (0, 1, 2, 0):
        score = np.abs(g64[k] * start[k]).ravel()
        q = int(rng.choice(np.argsort(score)[-200:]))
        idx = np.unravel_index(q, start[k].shape)
        mp = [m.copy() for m in start]
        mm = [m.copy() for m in start]
        mp[k][idx] *= 1 + h
        mm[k][idx] *= 1 - h
        fd = (p64.misfit(*mp) - p64.misfit(*mm)) / (2 * h * start[k][idx])
        an = float(g64[k][idx])
        print(f"C4 FD cell param {k} {tuple(int(i) for i in idx)}: FD {fd:.8e} adjoint {an:.8e} "
              f"rel {abs(fd - an) / abs(an):.2e}", flush=True)
        checks.append((fd, an))
    # release the propagators (tens of GB of checkpoints) before any assertion can keep them alive
    del res, p64
    torch.cuda.empty_cache()
    assert e < 1e-3
    for fd, an in checks:
        assert abs(fd - an) <= 1e-4 * abs(an)


# ---- C5 -------------------------------------------------------------------

def _stack_on_device(ws, m):
    """ws.stacked_elastic on the GPU: windows cut from m ([3, ny, nx, nz]
    torch, device) and edge-replicated outwards, stacked along y, 1 beyond
    narrower windows and past the last window."""
    import torch
    import torch.nn.functional as F
    from elasticisolib_3d_amd.models import FAT
    out = torch.ones((3, ws.ny, ws.nx, ws.nz), dtype=m.dtype, device=m.device)
    zm, zp = ws.zpads
    for k, w in enumerate(ws.windows):
        xm, xp = ws.xpads[k]
        core = m[:, w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :][None]
        win = F.pad(core, (FAT + zm, FAT + zp, FAT + xm, FAT + xp, FAT + ws.pad, FAT + ws.pad), mode="replicate")[0]
        out[:, ws.y_off[k]:ws.y_off[k] + ws.ny_w[k], :ws.nx_w[k], :] = win
    return out


class _SharedStack:
    """One propagator for every stack of `per_stack` windows of the 64 shots
    (grid: the widest and tallest stack), so the 64 + 16 gradients allocate
    their buffers once."""

    def __init__(self, shots, model_dev, cfg, per_stack):
        from elasticisolib_3d_amd import propagator as P
        from elasticisolib_3d_amd.batch import Window, WindowStack
        self.cfg, self.model = cfg, model_dev
        ny_m, nx_m, nz_m = model_dev.shape[1:]
        self.shape = (nz_m, nx_m, ny_m)
        groups = [shots[i:i + per_stack] for i in range(0, len(shots), per_stack)]
        own = [WindowStack(self.shape, cfg["d"], cfg["pad"], [Window(*s[2]) for s in g]) for g in groups]
        self.grid = (max(w.nx for w in own), max(w.ny for w in own))
        w0 = own[0]
        self.prop = P.Propagator(w0.nz, self.grid[0], self.grid[1], cfg["nts"], cfg["sub"], cfg["d"], cfg["d"],
                                 cfg["d"], cfg["dts"], w0.min_pad, cfg["alpha_cos"], "f32")
        self.prop.set_gradient_checkpoints(8, 0, "native")
        self.prop.set_graph(True)

    def run(self, group):
        """Receiver traces (per shot, 9 device arrays) and FP64 gradients (per
        shot, [8, ny_w, nx_w, nz] device) of one stack of shots; source traces
        unscaled (the same inputs for both paths)."""
        import torch
        from elasticisolib_3d_amd.acquisition import COMP_GRID, ShotGeometry
        from elasticisolib_3d_amd.batch import Window, WindowStack
        cfg, prop = self.cfg, self.prop
        ws = WindowStack(self.shape, cfg["d"], cfg["pad"], [Window(*s[2]) for s in group], grid=self.grid)
        ws.set_windows(prop)
        prop.set_elastic(_stack_on_device(ws, self.model).float(), cfg["dts"] / cfg["sub"])
        geos = []
        for k, (src_xyz, rec_xyz, _) in enumerate(group):
            ax = ws.axes[k]
            geos.append((ShotGeometry(ax, src_xyz[:, 2], src_xyz[:, 0], src_xyz[:, 1]),
                         ShotGeometry(ax, rec_xyz[:, 2], rec_xyz[:, 0], rec_xyz[:, 1])))
        spos = [np.concatenate(p) for p in zip(*[ws.stacked_positions(k, g[0].positions()) for k, g in enumerate(geos)])]
        rpos = [np.concatenate(p) for p in zip(*[ws.stacked_positions(k, g[1].positions()) for k, g in enumerate(geos)])]
        sig = [np.zeros((1, cfg["nts"])) for _ in range(9)]
        for c in (3, 4, 5):
            sig[c][0] = cfg["wavelet"] * 1e9
        src = [torch.tensor(np.concatenate([g[0].to_regular(sig)[c] for g in geos]), dtype=torch.float32,
                            device="cuda") for c in range(9)]
        sp = [torch.tensor(p, device="cuda") for p in spos]
        rp = [torch.tensor(p, device="cuda") for p in rpos]
        rec = prop.gradient_forward(sp, src, rp)
        adj = [r * 1e-3 for r in rec]
        g8, _ = prop.gradient_backward(sp, src, rp, adj)
        torch.cuda.synchronize()
        traces, grads, off = [], [], [0] * 7
        for k, (sg, rg) in enumerate(geos):
            cnt = rg.counts()
            traces.append([rec[c][off[COMP_GRID[c]]:off[COMP_GRID[c]] + cnt[COMP_GRID[c]]].clone() for c in range(9)])
            for gi in range(7):
                off[gi] += cnt[gi]
            grads.append(g8[:, ws.y_off[k]:ws.y_off[k] + ws.ny_w[k], :ws.nx_w[k], :].clone())
        return traces, grads


def test_c5_64_windows_batched_equal_plain():
    """The 64 windows with their receiver counts, stack by stack: each
    2-window stack's traces and gradients against each of its shots
    propagated alone (each path on its own persistent propagator, exact
    checkpoints every 8 fine steps)."""
    import time
    import torch
    from elasticisolib_3d_amd.models import batch_layered_400
    t_start = time.time()
    cfg = batch_layered_400(nts=6, smooth_device="cuda")
    print(f"C5: models built in {time.time() - t_start:.1f} s", flush=True)
    shots = cfg["shots"]
    assert len(shots) == 64
    nrec = [len(s[1]) for s in shots]
    print(f"C5: 64 windows {min(s[2][2] for s in shots)}-{max(s[2][2] for s in shots)} x "
          f"{min(s[2][3] for s in shots)}-{max(s[2][3] for s in shots)} cells, {min(nrec)}-{max(nrec)} receivers",
          flush=True)
    model = torch.tensor(cfg["start"], dtype=torch.float64, device="cuda")
    plain, batched = _SharedStack(shots, model, cfg, 1), _SharedStack(shots, model, cfg, 2)
    print(f"C5: shared grids plain {plain.grid}, 2-window stacks {batched.grid} "
          f"({time.time() - t_start:.1f} s)", flush=True)
    worst_d = worst_g = 0.0
    same = True
    nonzero = True
    t0 = time.time()
    for i in range(0, 64, 2):
        group = shots[i:i + 2]
        btr, bg = batched.run(group)
        for k, shot in enumerate(group):
            ptr, pg = plain.run([shot])
            worst_d = max(worst_d, _rel([x.double() for x in btr[k]], [y.double() for y in ptr[0]]))
            worst_g = max(worst_g, float(torch.linalg.norm(bg[k] - pg[0]) / torch.linalg.norm(pg[0])))
            same = same and bool(torch.equal(bg[k], pg[0]))
            nonzero = nonzero and float(torch.linalg.norm(pg[0])) > 0
        print(f"C5: shots {i + 1}-{i + 2} done ({time.time() - t0:.1f} s): worst data rel {worst_d:.2e}, "
              f"worst gradient rel {worst_g:.2e}", flush=True)
    del plain, batched
    torch.cuda.empty_cache()
    print(f"C5 batched vs plain: worst data rel {worst_d:.2e}, worst gradient rel {worst_g:.2e}, "
          f"gradients bit-identical {same}", flush=True)
    assert worst_d < 1e-6 and worst_g < 1e-6
    assert nonzero
