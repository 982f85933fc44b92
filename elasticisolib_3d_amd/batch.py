"""Batched multi-shot propagation with shot-local windows.

Each shot is modelled on a window of the model (a horizontal sub-rectangle,
full depth) padded with its own sponge, as the reference's padding tool pads
a model (padElasticFileGpuMain_3D.cpp:47-86). A batch of B windows is stacked
along y in one grid (ela_set_windows): the stencil, injection and recording
launches of a fine step then advance all B shots at once, each window keeping
its own FAT layers and sponge, so every shot's result equals its propagation
alone. B = 1 is the plain path (one shot per launch sequence) on the same
code.

WindowStack maps between the full model and the stack: medium (staggered and
scaled per window, never across windows), point positions, and gradients back
onto the full model (the transpose of the cut-and-pad).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import medium_torch as MT
from . import propagator as P
from .acquisition import (COMP_GRID, PaddedAxes, ShotGeometry, add_source_scaling_gradient, point_medium,
                          scale_regular_traces_at)
from .models import FAT, pad_sizes


@dataclass
class Window:
    """Model cells x in [x0, x0 + nx), y in [y0, y0 + ny), all z."""
    x0: int
    y0: int
    nx: int
    ny: int


def _pad_index(n: int, lo: int, hi: int) -> np.ndarray:
    """Source index of every padded cell under edge replication."""
    return np.clip(np.arange(-lo, n + hi), 0, n - 1)


def _unpad(g: np.ndarray, axis: int, lo: int, n: int) -> np.ndarray:
    """Transpose of edge-replication padding along one axis: the n core
    planes, with the low padding summed into the first and the high padding
    into the last (slices along the axis itself, no transposed copies)."""
    axis %= g.ndim

    def along(a, b):
        idx = [slice(None)] * g.ndim
        idx[axis] = slice(a, b)
        return tuple(idx)

    core = g[along(lo, lo + n)].copy()
    core[along(0, 1)] += g[along(0, lo)].sum(axis=axis, keepdims=True)
    core[along(n - 1, n)] += g[along(lo + n, None)].sum(axis=axis, keepdims=True)
    return core


class WindowStack:
    """B windows of a model of (nz, nx, ny) cells, spacing d, each padded
    with `pad` sponge cells (plus FAT) and stacked along y. `grid` = (nx, ny)
    places the stack in a larger grid (a propagator shared by several
    stacks): planes past the last window and columns past a window's width
    are never stepped."""

    def __init__(self, model_shape: Tuple[int, int, int], d: float, pad: int, windows: Sequence[Window],
                 grid: Optional[Tuple[int, int]] = None):
        nz, nxm, nym = model_shape
        self.nz_model, self.nx_model, self.ny_model = nz, nxm, nym
        self.d, self.pad = float(d), int(pad)
        self.windows = list(windows)
        zm, zp = pad_sizes(nz, pad)
        self.zpads = (zm, zp)
        self.nz = nz + zm + zp + 2 * FAT
        self.axes: List[PaddedAxes] = []
        self.xpads, self.ny_w, self.nx_w, self.y_off = [], [], [], []
        y = 0
        for w in self.windows:
            if w.x0 < 0 or w.y0 < 0 or w.x0 + w.nx > nxm or w.y0 + w.ny > nym:
                raise ValueError(f"window {w} leaves the model")
            xm, xp = pad_sizes(w.nx, pad)
            nxp = w.nx + xm + xp + 2 * FAT
            nyp = w.ny + 2 * pad + 2 * FAT
            self.xpads.append((xm, xp))
            self.nx_w.append(nxp)
            self.ny_w.append(nyp)
            self.y_off.append(y)
            y += nyp
            self.axes.append(PaddedAxes(self.nz, nxp, nyp, -(FAT + zm) * d, (w.x0 - FAT - xm) * d,
                                        (w.y0 - FAT - pad) * d, d, d, d, FAT, zm, zp, xm, xp, pad))
        self.nx = max(self.nx_w)
        self.ny = y
        if grid is not None:
            if grid[0] < self.nx or grid[1] < self.ny:
                raise ValueError(f"grid {grid} smaller than the stack ({self.nx}, {self.ny})")
            self.nx, self.ny = int(grid[0]), int(grid[1])
        self.min_pad = min(min(zm, zp), min(min(x) for x in self.xpads))

    # -- model ------------------------------------------------------------
    def _cut_index(self, k: int):
        w = self.windows[k]
        zm, zp = self.zpads
        xm, xp = self.xpads[k]
        iz = _pad_index(self.nz_model, FAT + zm, FAT + zp)
        ix = w.x0 + _pad_index(w.nx, FAT + xm, FAT + xp)
        iy = w.y0 + _pad_index(w.ny, FAT + self.pad, FAT + self.pad)
        return iy, ix, iz

    def window_model(self, k: int, m: np.ndarray) -> np.ndarray:
        """Padded window k of a [..., ny, nx, nz] model array: the window's
        cells, edge-replicated outwards (the cells _cut_index names)."""
        w = self.windows[k]
        zm, zp = self.zpads
        xm, xp = self.xpads[k]
        core = m[..., w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :]
        widths = [(0, 0)] * (m.ndim - 3) + [(FAT + self.pad, FAT + self.pad), (FAT + xm, FAT + xp),
                                            (FAT + zm, FAT + zp)]
        return np.pad(core, widths, mode="edge")

    def window_adjoint(self, k: int, g: np.ndarray, out: np.ndarray) -> None:
        """out (full model [..., ny, nx, nz]) += the transpose of window_model
        applied to g (padded window k): the replicated cells of each padding
        slab add into the edge plane they copy."""
        w = self.windows[k]
        zm, zp = self.zpads
        xm, xp = self.xpads[k]
        g = _unpad(g, -1, FAT + zm, self.nz_model)
        g = _unpad(g, -2, FAT + xm, w.nx)
        g = _unpad(g, -3, FAT + self.pad, w.ny)
        out[..., w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :] += g

    # torch versions (any device): the same cut, pad, stack and transpose,
    # so that batched FWI keeps the model and its gradient on the GPU
    def window_model_t(self, k: int, m):
        """window_model on a torch [..., ny, nx, nz] tensor (3 or 4 dims)."""
        import torch.nn.functional as F
        w = self.windows[k]
        zm, zp = self.zpads
        xm, xp = self.xpads[k]
        core = m[..., w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :]
        pad = (FAT + zm, FAT + zp, FAT + xm, FAT + xp, FAT + self.pad, FAT + self.pad)
        if core.dim() == 3:
            return F.pad(core[None], pad, mode="replicate")[0]
        return F.pad(core, pad, mode="replicate")

    def window_adjoint_t(self, k: int, g, out) -> None:
        """window_adjoint on torch tensors: out += transpose of window_model_t(k) applied to g."""
        w = self.windows[k]
        zm, zp = self.zpads
        xm, xp = self.xpads[k]

        def unpad(a, dim, lo, n):
            core = a.narrow(dim, lo, n).clone()
            core.narrow(dim, 0, 1).add_(a.narrow(dim, 0, lo).sum(dim=dim, keepdim=True))
            core.narrow(dim, n - 1, 1).add_(a.narrow(dim, lo + n, a.shape[dim] - lo - n).sum(dim=dim, keepdim=True))
            return core

        g = unpad(g, -1, FAT + zm, self.nz_model)
        g = unpad(g, -2, FAT + xm, w.nx)
        g = unpad(g, -3, FAT + self.pad, w.ny)
        out[..., w.y0:w.y0 + w.ny, w.x0:w.x0 + w.nx, :] += g

    def stack_t(self, per_window, fill: float = 0.0):
        """stack on torch tensors [..., ny_w, nx_w, nz] -> [..., ny, nx, nz]."""
        import torch
        a0 = per_window[0]
        out = torch.full(tuple(a0.shape[:-3]) + (self.ny, self.nx, self.nz), fill, dtype=a0.dtype, device=a0.device)
        for k, a in enumerate(per_window):
            out[..., self.y_off[k]:self.y_off[k] + self.ny_w[k], :self.nx_w[k], :] = a
        return out

    def stack(self, per_window: Sequence[np.ndarray], dtype=None, fill: float = 0.0) -> np.ndarray:
        """[..., ny_w, nx_w, nz] arrays -> one stacked [..., ny, nx, nz] array
        (`fill` beyond a narrower window)."""
        lead = per_window[0].shape[:-3]
        out = np.full(lead + (self.ny, self.nx, self.nz), fill, dtype or per_window[0].dtype)
        for k, a in enumerate(per_window):
            out[..., self.y_off[k]:self.y_off[k] + self.ny_w[k], :self.nx_w[k], :] = a
        return out

    def unstack(self, k: int, a: np.ndarray) -> np.ndarray:
        return a[..., self.y_off[k]:self.y_off[k] + self.ny_w[k], :self.nx_w[k], :]

    def window_media(self, rho, lam, mu, dtw: float) -> List[np.ndarray]:
        """Per-window med8: staggering and scaling window by window."""
        return [P.scale_medium(self.window_model(k, rho), self.window_model(k, lam), self.window_model(k, mu), dtw)
                for k in range(len(self.windows))]

    def stacked_medium(self, rho, lam, mu, dtw: float):
        """(stacked med8, list of per-window med8)."""
        meds = self.window_media(rho, lam, mu, dtw)
        return self.stack(meds), meds

    def stacked_elastic(self, rho, lam, mu) -> np.ndarray:
        """Stacked [3, ny, nx, nz] (rho, lambda, mu) of the padded windows, for
        a propagator that forms the scaled medium itself. Cells beyond a
        narrower window (never within reach of an interior stencil) hold 1."""
        return self.stack([np.stack([self.window_model(k, rho), self.window_model(k, lam), self.window_model(k, mu)])
                           for k in range(len(self.windows))], fill=1.0)

    # -- points -----------------------------------------------------------
    def stacked_positions(self, k: int, pos: Sequence[np.ndarray]) -> List[np.ndarray]:
        """Flat positions on window k's padded grid -> flat stacked positions."""
        out = []
        plane_w = self.nx_w[k] * self.nz
        for p in pos:
            p = np.asarray(p, np.int64)
            iy, r = np.divmod(p, plane_w)
            ix, iz = np.divmod(r, self.nz)
            out.append(((iy + self.y_off[k]) * self.nx + ix) * self.nz + iz)
        return out

    def set_windows(self, prop: P.Propagator) -> None:
        prop.set_windows(self.y_off, self.ny_w, self.nx_w)


@dataclass
class BatchShot:
    src_xyz: np.ndarray          # [n, 3] (x, y, z) model coordinates
    rec_xyz: np.ndarray          # [m, 3]
    window: Window
    signals: List[np.ndarray]    # 9 irregular [n, nts]
    observed: Optional[List[np.ndarray]] = None   # 9 irregular [m, nts] (host FP64)


class BatchedFwi:
    """FWI misfit and gradient over shots with shot-local windows, `batch`
    shots per launch sequence on one GPU (FP32 stencil, int16 block-scaled
    (default), bfloat16 or exact checkpoints, FP64 gradient accumulation, as
    FwiProblem). `model_shape` is the [ny, nx, nz] shape of the (unpadded)
    model arrays; `checkpoint` is "s16", "bf16" or "native", and the older
    `bf16` flag, when given, means "bf16" / "native"; `autotune` times the launch shapes of the shared grid
    once unless `tiling` fixes them; `medium` ("elastic" or "scaled") chooses how
    the stacked model reaches the library, as in FwiProblem; `graph` replays
    the gradient's time steps as HIP graphs (as FwiProblem). Every batch runs
    on one persistent propagator whose grid holds the widest and tallest
    stack of windows (`shared_grid`); a batch sets its windows and stacked
    medium on it, so wavefields, checkpoints and graphs are allocated and
    captured once for all batches and calls (`release` frees them).
    `progress`, a callable taking a string, is told after every batch;
    `derivatives` is the storage of the backward's derivative combinations
    (as FwiProblem; None: the library's default)."""

    def __init__(self, model_shape, d: float, pad: int, nts: int, dts: float, sub: int, alpha_cos: float,
                 shots: Sequence[BatchShot], batch: int = 1, precision: str = "f32", interval: int = 20,
                 bf16=None, device: int = 0, tiling=None, checkpoint: str = "s16", autotune: bool = False,
                 medium: str = "elastic", graph: bool = True, progress=None, derivatives=None):
        self.model_shape, self.d, self.pad = tuple(model_shape), float(d), int(pad)
        self.nts, self.dts, self.sub, self.alpha = int(nts), float(dts), int(sub), float(alpha_cos)
        self.dtw = self.dts / self.sub
        self.shots = list(shots)
        self.batch = max(1, int(batch))
        self.precision, self.interval, self.device = precision, int(interval), int(device)
        self.checkpoint = checkpoint if bf16 is None else ("bf16" if bf16 else "native")
        self.tiling = tiling
        self.autotune = bool(autotune)   # time the launch shapes of every stacked grid (cached per grid shape)
        # "elastic": hand the stacked (rho, lambda, mu) to the library (its forward steps form the
        # scaled medium); "scaled": the stacked 8 scaled arrays (ela_set_medium)
        if medium not in ("elastic", "scaled"):
            raise ValueError("medium must be 'elastic' or 'scaled'")
        self.medium = medium
        self.graph = bool(graph)   # replay the gradient's step pairs and checkpoint segments as HIP graphs
        self.progress = progress   # callable(str) told after every batch, or None
        self.derivatives = derivatives
        self.timing = {}
        # one propagator for every batch and every call: its grid holds the widest and the tallest
        # stack, so its wavefields, checkpoints and graphs are allocated, tuned and captured once
        self._prop = None
        self._grid = None
        self._tune_pending = False

    def _groups(self):
        return [self.shots[i:i + self.batch] for i in range(0, len(self.shots), self.batch)]

    def _model_zxy(self):
        return (self.model_shape[2], self.model_shape[1], self.model_shape[0])

    def shared_grid(self) -> Tuple[int, int]:
        """(nx, ny) of the propagator grid the batches share."""
        if self._grid is None:
            stacks = [WindowStack(self._model_zxy(), self.d, self.pad, [s.window for s in g]) for g in self._groups()]
            self._grid = (max(w.nx for w in stacks), max(w.ny for w in stacks))
        return self._grid

    def _propagator(self, ws):
        """The persistent propagator, created and configured on first use."""
        if self._prop is None:
            prop = P.Propagator(ws.nz, ws.nx, ws.ny, self.nts, self.sub, self.d, self.d, self.d, self.dts,
                                ws.min_pad, self.alpha, self.precision, self.device)
            if self.tiling:
                prop.set_tiling(*self.tiling[:2])
                if len(self.tiling) > 2:
                    prop.set_tile_order(*self.tiling[2:])
            prop.set_gradient_checkpoints(self.interval, 0, self.checkpoint)
            if self.derivatives is not None:
                prop.set_gradient_derivatives(self.derivatives)
            prop.set_graph(self.graph)
            self._prop = prop
            self._tune_pending = self.autotune and not self.tiling
        return self._prop

    def release(self) -> None:
        """Free the persistent propagator's device memory (it is created again when needed)."""
        self._prop = None

    def _model_device(self, model):
        """The full (rho, lambda, mu) model as one FP64 tensor on the GPU."""
        import torch
        return torch.as_tensor(np.asarray(model, np.float64), device="cuda")

    def _setup(self, group, model_d):
        """Stack, propagator and per-shot geometry of one batch. The windows'
        media are cut, padded, staggered and scaled on the GPU (FP64, the
        operations of the host routines); `meds` are the per-window 8 scaled
        arrays (FP64, device)."""
        import torch
        ws = WindowStack(self._model_zxy(), self.d, self.pad, [s.window for s in group], grid=self.shared_grid())
        wins = [ws.window_model_t(k, model_d) for k in range(len(group))]
        meds = [MT.scale_medium(w[0], w[1], w[2], self.dtw) for w in wins]
        prop = self._propagator(ws)
        ws.set_windows(prop)
        td = torch.float32 if self.precision == "f32" else torch.float64
        if self.medium == "elastic":
            prop.set_elastic(ws.stack_t(wins, fill=1.0).to(td).contiguous(), self.dtw)
        else:
            prop.set_medium(ws.stack_t(meds).to(td).contiguous())
        del wins
        if self._tune_pending:
            from .autotune import autotune
            autotune(prop, reps=5, target="gradient")
            self._tune_pending = False
        geos = []
        for k, s in enumerate(group):
            ax = ws.axes[k]
            sg = ShotGeometry(ax, s.src_xyz[:, 2], s.src_xyz[:, 0], s.src_xyz[:, 1])
            rg = ShotGeometry(ax, s.rec_xyz[:, 2], s.rec_xyz[:, 0], s.rec_xyz[:, 1])
            geos.append((sg, rg))
        return ws, meds, prop, geos, td

    def _merged(self, ws, geos, which):
        """Per grid: stacked positions of all shots of the batch, concatenated."""
        pos = [[] for _ in range(7)]
        for k, g in enumerate(geos):
            sp = ws.stacked_positions(k, g[which].positions())
            for j in range(7):
                pos[j].append(sp[j])
        return [np.concatenate(p) if p else np.zeros(0, np.int64) for p in pos]

    def _src_traces(self, ws, meds, geos, group, td):
        """Scaled regular source traces of the batch (9 concatenated arrays)
        and the unscaled regular traces per shot (the scaled medium is read
        at the source points only)."""
        import torch
        raw, scaled = [], []
        for k, (s, (sg, _)) in enumerate(zip(group, geos)):
            r = sg.to_regular(s.signals)
            raw.append(r)
            scaled.append(scale_regular_traces_at(point_medium(meds[k], sg), r, self.d ** 3))
        cat = [np.concatenate([sc[c] for sc in scaled]) for c in range(9)]
        return [torch.tensor(c, dtype=td, device="cuda") for c in cat], raw

    def _split_rec(self, geos, rec):
        """Concatenated regular receiver traces -> per shot irregular FP64."""
        out, off = [], [0] * 7
        for sg, rg in geos:
            reg = []
            for c in range(9):
                g = COMP_GRID[c]
                n = rg.counts()[g]
                reg.append(rec[c][off[g]:off[g] + n].double().cpu().numpy())
            for g in range(7):
                off[g] += rg.counts()[g]
            out.append(rg.to_irregular(reg))
        return out

    def _tell(self, what: str, i: int, n: int, t0: float) -> None:
        if self.progress is not None:
            self.progress(f"{what}: batch {i + 1}/{n} done ({time.perf_counter() - t0:.1f} s)")

    def synthesize(self, model) -> None:
        import torch
        model_d = self._model_device(model)
        groups = self._groups()
        t0 = time.perf_counter()
        for gi, group in enumerate(groups):
            ws, meds, prop, geos, td = self._setup(group, model_d)
            src, _ = self._src_traces(ws, meds, geos, group, td)
            rec = prop.nl_forward(self._merged(ws, geos, 0), src, self._merged(ws, geos, 1))
            torch.cuda.synchronize()
            for s, d in zip(group, self._split_rec(geos, rec)):
                s.observed = d
            del prop
            self._tell("observed data", gi, len(groups), t0)

    def misfit_and_gradient(self, model):
        """(J, dJ/d(rho, lambda, mu) on the full model [3, ny, nx, nz]). The
        gradient stays on the GPU from the adjoint propagation to the full
        model (chain rule through the scaling, source-scaling term, transpose
        of the window cut) and is copied to the host once."""
        import torch
        t0 = time.perf_counter()
        model_d = self._model_device(model)
        g3 = torch.zeros_like(model_d)
        J = 0.0
        t_prop = 0.0
        groups = self._groups()
        for gi, group in enumerate(groups):
            ws, meds, prop, geos, td = self._setup(group, model_d)
            src, raw = self._src_traces(ws, meds, geos, group, td)
            spos, rpos = self._merged(ws, geos, 0), self._merged(ws, geos, 1)
            ta = time.perf_counter()
            rec = prop.gradient_forward(spos, src, rpos)
            torch.cuda.synchronize()
            t_prop += time.perf_counter() - ta
            pred = self._split_rec(geos, rec)
            adj_parts = []
            for s, p, (sg, rg) in zip(group, pred, geos):
                res = [p[c] - s.observed[c] for c in range(9)]
                J += 0.5 * sum(float(np.sum(r * r)) for r in res)
                adj_parts.append(rg.to_regular(res))
            adj = []
            for c in range(9):
                adj.append(torch.tensor(np.concatenate([a[c] for a in adj_parts]), dtype=td, device="cuda"))
            ta = time.perf_counter()
            g8, gsrc = prop.gradient_backward(spos, src, rpos, adj)
            torch.cuda.synchronize()
            t_prop += time.perf_counter() - ta
            off = [0] * 7
            for k, (sg, rg) in enumerate(geos):
                gs = []
                for c in range(9):
                    g = COMP_GRID[c]
                    n = sg.counts()[g]
                    gs.append(gsrc[c][off[g]:off[g] + n].double().cpu().numpy())
                for g in range(7):
                    off[g] += sg.counts()[g]
                g8k = ws.unstack(k, g8).contiguous()
                add_source_scaling_gradient(g8k, sg, raw[k], gs, self.d ** 3)
                ws.window_adjoint_t(k, MT.medium_gradient(g8k, meds[k], self.dtw), g3)
            del prop, g8, meds
            torch.cuda.empty_cache()
            self._tell("gradient", gi, len(groups), t0)
        g3h = g3.cpu().numpy()
        self.timing = {"total_s": time.perf_counter() - t0, "propagation_s": t_prop}
        return J, g3h
