"""Full-waveform-inversion misfit and its adjoint-state gradient, with a
torch.autograd.Function front end.

The misfit of a set of shots is J = 1/2 sum_shots ||S_r F(m; s) - d_obs||^2,
F being the nonlinear forward operator of the propagator (regular-grid
traces, medium-scaled sources), S_r the trilinear receiver interpolation and
m = (rho, lambda, mu) on the padded grid. The gradient is computed shot by
shot with the library's checkpointed adjoint propagation
(ela_gradient_forward / ela_gradient_backward: FP32 stencil, bfloat16 or
exact checkpoints, FP64 accumulation), then mapped to (rho, lambda, mu):

    dJ/dM8   = adjoint-state imaging + source-scaling term (FP64)
    dJ/dm    = medium_gradient(dJ/dM8)                      (FP64, host)

The residual, the misfit reduction and the receiver interpolation run in FP64
on the GPU (torch). The reference's inversion (python_float/efwiMain_3D.py)
uses its Born adjoint as the gradient of the same misfit; the Born operator's
coarse-time secondary sources make that an approximation of the exact
gradient computed here.

`ElasticFwiMisfit.apply(rho, lam, mu, problem)` returns J as a 0-d float64
tensor; backward() returns the stored gradient scaled by the incoming
gradient. Forward does all the propagation (the gradient is needed as soon
as J is, and keeping every shot's checkpoints alive until backward would not
fit in memory).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import propagator as P
from .acquisition import COMP_GRID, PaddedAxes, ShotGeometry, scale_regular_traces, source_scaling_gradient


def _torch():
    import torch
    return torch


class _DeviceInterp:
    """Trilinear receiver interpolation of one grid on the GPU, FP64: each
    device reads 8 regular points with fixed weights (gather); the adjoint
    scatters with index_add."""

    def __init__(self, interp, device):
        torch = _torch()
        self.nbr = torch.tensor(interp.nbr, dtype=torch.int64, device=device)       # [n_dev, 8]
        self.w = torch.tensor(interp.weight, dtype=torch.float64, device=device)    # [n_dev, 8]
        self.n_reg = interp.n_reg

    def forward(self, reg):
        """[n_reg, nt] -> [n_dev, nt]"""
        return (reg.double()[self.nbr] * self.w[:, :, None]).sum(dim=1)

    def adjoint(self, irr):
        """[n_dev, nt] -> [n_reg, nt]"""
        torch = _torch()
        out = torch.zeros((self.n_reg, irr.shape[1]), dtype=torch.float64, device=irr.device)
        out.index_add_(0, self.nbr.reshape(-1), (self.w[:, :, None] * irr[:, None, :]).reshape(-1, irr.shape[1]))
        return out


@dataclass
class FwiShot:
    """One shot: source and receiver geometries (7 grids each), the 9
    irregular source signals [n_src, nts] and the observed irregular data
    (9 arrays [n_rec, nts], filled by FwiProblem.synthesize)."""

    src: ShotGeometry
    rec: ShotGeometry
    signals: List[np.ndarray]
    observed: Optional[list] = None
    _cache: dict = field(default_factory=dict, repr=False)


class FwiProblem:
    """Shots on one padded grid, propagated on one GPU. `checkpoint` selects
    the storage of the gradient's outer checkpoints ("bf16", "s16" or
    "native"; None: the `bf16` flag), `autotune` times the stencil launch
    shapes once when the first medium is set, `medium` chooses between
    handing (rho, lambda, mu) to the library ("elastic") and the 8 host-scaled
    arrays ("scaled"); the two agree exactly in FP64."""

    def __init__(self, axes: PaddedAxes, nts: int, dts: float, sub: int, alpha_cos: float, shots: Sequence[FwiShot],
                 precision: str = "f32", interval: int = 20, outer_interval: int = 0, bf16: bool = True,
                 device: int = 0, checkpoint=None, autotune: bool = False, medium: str = "elastic"):
        self.axes = axes
        self.nts, self.dts, self.sub = int(nts), float(dts), int(sub)
        self.dtw = self.dts / self.sub
        self.shots = list(shots)
        self.precision = precision
        self.dev = f"cuda:{device}"
        self.prop = P.Propagator(axes.nz, axes.nx, axes.ny, nts, sub, axes.dz, axes.dx, axes.dy, dts, axes.min_pad,
                                 alpha_cos, precision, device)
        # `checkpoint` ("bf16", "s16", "native") overrides the older bf16 flag
        self.prop.set_gradient_checkpoints(interval, outer_interval, bf16 if checkpoint is None else checkpoint)
        self.dxyz = axes.dx * axes.dy * axes.dz
        self.last_timing = {}
        # autotune: time the stencil launch shapes once, when the first medium is set
        self._tune, self.tuned = bool(autotune), None
        # "elastic": hand (rho, lambda, mu) to the library, whose forward steps form the scaled medium
        # (fewer bytes per step); "scaled": the 8 host-scaled arrays
        if medium not in ("elastic", "scaled"):
            raise ValueError("medium must be 'elastic' or 'scaled'")
        self.medium = medium

    # -- helpers -----------------------------------------------------------
    def _tdtype(self):
        torch = _torch()
        return torch.float32 if self.precision == "f32" else torch.float64

    def _shot_device(self, shot: FwiShot):
        """Positions and interpolation matrices of a shot on the GPU (cached)."""
        c = shot._cache
        if "spos" not in c:
            torch = _torch()
            c["spos"] = [torch.tensor(p, device=self.dev) for p in shot.src.positions()]
            c["rpos"] = [torch.tensor(p, device=self.dev) for p in shot.rec.positions()]
            c["Mr"] = [_DeviceInterp(g, self.dev) for g in shot.rec.grids]
            c["raw_reg"] = [np.asarray(x, np.float64) for x in shot.src.to_regular(shot.signals)]
        return c

    def _scaled_sources(self, med8, shot):
        torch = _torch()
        c = self._shot_device(shot)
        reg = scale_regular_traces(med8, shot.src, c["raw_reg"], self.dxyz)
        return [torch.tensor(r, dtype=self._tdtype(), device=self.dev) for r in reg]

    def _set_model(self, rho, lam, mu):
        med8 = P.scale_medium(np.asarray(rho, np.float64), np.asarray(lam, np.float64), np.asarray(mu, np.float64),
                              self.dtw)
        torch = _torch()
        if self.medium == "elastic":
            rlm = np.stack([np.asarray(rho, np.float64), np.asarray(lam, np.float64), np.asarray(mu, np.float64)])
            self.prop.set_elastic(torch.tensor(rlm, dtype=self._tdtype(), device=self.dev), self.dtw)
        else:
            self.prop.set_medium(torch.tensor(med8, dtype=self._tdtype(), device=self.dev))
        if self._tune and self.tuned is None:
            from .autotune import autotune
            self.tuned = autotune(self.prop, reps=5)
        return med8

    def _to_irregular(self, shot, reg):
        """Regular traces (device) -> 9 irregular FP64 arrays (device)."""
        torch = _torch()
        c = self._shot_device(shot)
        return [c["Mr"][COMP_GRID[k]].forward(reg[k]) for k in range(9)]

    def _to_regular(self, shot, irr):
        c = self._shot_device(shot)
        return [c["Mr"][COMP_GRID[k]].adjoint(irr[k]).to(self._tdtype()).contiguous() for k in range(9)]

    # -- public ------------------------------------------------------------
    def synthesize(self, rho, lam, mu) -> None:
        """Observed data of every shot from a (true) model (FP64 device)."""
        torch = _torch()
        med8 = self._set_model(rho, lam, mu)
        for shot in self.shots:
            c = self._shot_device(shot)
            reg = self.prop.nl_forward(c["spos"], self._scaled_sources(med8, shot), c["rpos"])
            shot.observed = [x.clone() for x in self._to_irregular(shot, reg)]
        torch.cuda.synchronize()

    def misfit(self, rho, lam, mu) -> float:
        """J only (plain forward propagations; FP64 residual reduction)."""
        torch = _torch()
        med8 = self._set_model(rho, lam, mu)
        J = torch.zeros((), dtype=torch.float64, device=self.dev)
        for shot in self.shots:
            c = self._shot_device(shot)
            reg = self.prop.nl_forward(c["spos"], self._scaled_sources(med8, shot), c["rpos"])
            pred = self._to_irregular(shot, reg)
            for k in range(9):
                J += 0.5 * torch.sum((pred[k] - shot.observed[k]) ** 2)
        return float(J)

    def misfit_and_gradient(self, rho, lam, mu):
        """(J, dJ/d(rho, lambda, mu) as a float64 [3, ny, nx, nz] array)."""
        torch = _torch()
        med8 = self._set_model(rho, lam, mu)
        J = torch.zeros((), dtype=torch.float64, device=self.dev)
        g8 = torch.zeros((8,) + tuple(med8.shape[1:]), dtype=torch.float64, device=self.dev)
        g8_src = np.zeros(med8.shape, np.float64)
        ms_f = ms_b = 0.0
        for shot in self.shots:
            c = self._shot_device(shot)
            src = self._scaled_sources(med8, shot)
            reg = self.prop.gradient_forward(c["spos"], src, c["rpos"])
            ms_f += self.prop.last_timing()[0]
            pred = self._to_irregular(shot, reg)
            res = [pred[k] - shot.observed[k] for k in range(9)]
            for k in range(9):
                J += 0.5 * torch.sum(res[k] ** 2)
            adj = self._to_regular(shot, res)
            g8, gsrc = self.prop.gradient_backward(c["spos"], src, c["rpos"], adj, g8)
            ms_b += self.prop.last_timing()[0]
            g8_src += source_scaling_gradient(med8, shot.src, c["raw_reg"], [x.double().cpu().numpy() for x in gsrc],
                                              self.dxyz)
        torch.cuda.synchronize()
        self.last_timing = {"forward_ms": ms_f, "backward_ms": ms_b}
        g8h = g8.cpu().numpy() + g8_src
        return float(J), P.medium_gradient(g8h, med8, self.dtw)


def _autograd_function():
    torch = _torch()

    class ElasticFwiMisfit(torch.autograd.Function):
        """J(rho, lambda, mu) for an FwiProblem; gradient by the adjoint state."""

        @staticmethod
        def forward(ctx, rho, lam, mu, problem: FwiProblem):
            J, g3 = problem.misfit_and_gradient(rho.detach().double().cpu().numpy(), lam.detach().double().cpu().numpy(),
                                                mu.detach().double().cpu().numpy())
            g = torch.from_numpy(g3)
            ctx.save_for_backward(g)
            ctx.dtypes = (rho.dtype, lam.dtype, mu.dtype)
            ctx.devices = (rho.device, lam.device, mu.device)
            return torch.tensor(J, dtype=torch.float64, device=rho.device)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            s = gJ.detach().double().cpu()
            out = [(s * g[k]).to(dtype=ctx.dtypes[k], device=ctx.devices[k]) for k in range(3)]
            return out[0], out[1], out[2], None

    return ElasticFwiMisfit


_FN = None


def fwi_misfit(rho, lam, mu, problem: FwiProblem):
    """Differentiable misfit: torch tensors in, 0-d float64 tensor out."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(rho, lam, mu, problem)


def invert(problem, m0, n_iter: int = 5, mask=None, bounds=None, history_size: int = 10,
           line_search: str = "strong_wolfe", log=None):
    """FWI outer loop: L-BFGS (torch.optim.LBFGS) on x = m / m0 (each of
    rho, lambda, mu normalised by the starting model, which scales the three
    parameters alike), the gradient from `problem.misfit_and_gradient`
    (adjoint state). `mask` ([3, ny, nx, nz] or broadcastable) multiplies the
    gradient (the reference's maskGradient); `bounds` = ((min_rho, min_lambda,
    min_mu), (max_rho, max_lambda, max_mu)) are enforced by projection after
    every evaluation (scalars or arrays; None for no bound). Returns
    (model (3 arrays), history of (evaluation, J)). The counterpart of the
    reference's efwiMain_3D.py driver, whose pySolver solvers are not part of
    this project."""
    torch = _torch()
    m0 = [np.asarray(m, np.float64) for m in m0]
    ref = [torch.from_numpy(m) for m in m0]
    x = [torch.ones_like(r, requires_grad=True) for r in ref]
    msk = None if mask is None else torch.as_tensor(np.asarray(mask, np.float64))
    lo = hi = None
    if bounds is not None:
        lo = [None if b is None else torch.as_tensor(np.asarray(b, np.float64)) / r for b, r in zip(bounds[0], ref)]
        hi = [None if b is None else torch.as_tensor(np.asarray(b, np.float64)) / r for b, r in zip(bounds[1], ref)]
    opt = torch.optim.LBFGS(x, lr=1.0, max_iter=n_iter, max_eval=int(1.5 * n_iter) + 2, history_size=history_size,
                            line_search_fn=line_search, tolerance_grad=0.0, tolerance_change=0.0)
    history = []

    def project():
        if bounds is None:
            return
        with torch.no_grad():
            for k in range(3):
                if lo[k] is not None:
                    x[k].copy_(torch.maximum(x[k], lo[k]))
                if hi[k] is not None:
                    x[k].copy_(torch.minimum(x[k], hi[k]))

    def closure():
        opt.zero_grad()
        project()
        m = [xi.detach().numpy() * mi for xi, mi in zip(x, m0)]
        J, g3 = problem.misfit_and_gradient(*m)
        for k in range(3):
            g = torch.from_numpy(np.ascontiguousarray(g3[k] * m0[k]))   # d/dx = d/dm * m0
            if msk is not None:
                g = g * (msk[k] if msk.dim() == 4 else msk)
            x[k].grad = g
        history.append((len(history), float(J)))
        if log is not None:
            log(f"[fwi] evaluation {len(history)}: J = {J:.9e}")
        return torch.tensor(J, dtype=torch.float64)

    opt.step(closure)
    project()
    return [xi.detach().numpy() * mi for xi, mi in zip(x, m0)], history
