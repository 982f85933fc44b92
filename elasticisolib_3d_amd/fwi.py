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

Around the operator, as the reference's inversion chains them
(python_float/efwiMain_3D.py:199-237): component sampling of the data
(ElasticDatComp_3D, seis_utils/python/dataCompModule_3D.py:11-141, pressure
p = (sxx + syy + szz) / 3), a data mask applied to the predicted and the
observed data (its DiagonalOp), and optionally the (Vp, Vs, rho)
parameterisation through ElasticConv_3D and its Jacobian
(seis_utils/python/elasticParamConvertModule_3D.py:23-117):

    J = 1/2 sum_shots ||M (C S_r F(g(m)) - d_obs)||^2

with C the component sampling, M the data mask and g the conversion.

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


ALL_COMPONENTS = "vx,vy,vz,sxx,syy,szz,sxz,sxy,syz"
_NAMES = ALL_COMPONENTS.split(",")
P_FACTOR = 1.0 / 3.0   # pressure factor of dataCompModule_3D.py:8 (the code, not its docstring's 0.5)


def component_rows(components: str):
    """Sampling matrix rows of a component list (ElasticDatComp_3D,
    dataCompModule_3D.py:11-141): one list of (index among the 9 components,
    weight) per sampled component; "p" is (sxx + syy + szz) / 3. A component
    named twice, or a name that is not a component, raises ValueError."""
    comps = [c.strip().lower() for c in components.split(",")]
    if len(set(comps)) != len(comps):
        raise ValueError("ERROR! A component was provided multiple times! Check your input arguments")
    rows = []
    for c in comps:
        if c == "p":
            rows.append([(3, P_FACTOR), (4, P_FACTOR), (5, P_FACTOR)])
        elif c in _NAMES:
            rows.append([(_NAMES.index(c), 1.0)])
        else:
            raise ValueError(f"ERROR! Provided unknown data components: {components}")
    return rows


def sample_components(rows, d9):
    """9 component arrays -> the sampled components (numpy or torch)."""
    out = []
    for row in rows:
        acc = None
        for k, w in row:
            t = d9[k] if w == 1.0 else w * d9[k]
            acc = t if acc is None else acc + t
        out.append(acc)
    return out


def sample_components_adjoint(rows, ds, zeros_like):
    """Adjoint of sample_components: sampled arrays -> 9 component arrays
    (components not sampled are zero arrays made by zeros_like())."""
    out = [None] * 9
    for row, d in zip(rows, ds):
        for k, w in row:
            t = d if w == 1.0 else w * d
            out[k] = t if out[k] is None else out[k] + t
    return [zeros_like() if o is None else o for o in out]


def vp_vs_rho_to_elastic(vp, vs, rho):
    """ElasticConv_3D type 1 (elasticParamConvertModule_3D.py:23-47):
    (Vp, Vs, rho) -> (rho, lambda = rho (Vp^2 - 2 Vs^2), mu = rho Vs^2)."""
    return rho, rho * (vp * vp - 2.0 * vs * vs), rho * vs * vs


def vp_vs_rho_gradient(vp, vs, rho, g_rho, g_lam, g_mu):
    """Transpose of the conversion's Jacobian at (vp, vs, rho) applied to a
    gradient in (rho, lambda, mu) (ElasticConvJab_3D adjoint,
    elasticParamConvertModule_3D.py:85-117): the gradient in (Vp, Vs, rho)."""
    return (2.0 * vp * rho * g_lam,
            2.0 * vs * rho * g_mu - 4.0 * vs * rho * g_lam,
            g_rho + (vp * vp - 2.0 * vs * vs) * g_lam + vs * vs * g_mu)


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
    irregular source signals [n_src, nts], the observed irregular data (one
    array [n_rec, nts] per component of the problem's component list, filled
    by FwiProblem.synthesize) and an optional data mask (one array per
    sampled component, broadcastable to [n_rec, nts]; the reference's
    dataMaskFile, efwiMain_3D.py:224-237)."""

    src: ShotGeometry
    rec: ShotGeometry
    signals: List[np.ndarray]
    observed: Optional[list] = None
    data_mask: Optional[list] = None
    _cache: dict = field(default_factory=dict, repr=False)


class FwiProblem:
    """Shots on one padded grid, propagated on one GPU. `checkpoint` selects
    the storage of the gradient's outer checkpoints: "s16" (default: int16
    with one scale per 64 values, 5.4e-5 from the FP64 gradient on the
    256x256x160 workload, inside the 1e-3 target), "bf16" (bfloat16 with one
    power-of-two scale per array and snapshot: 1.8e-3 there, outside it,
    because 8 significant bits per value restart every segment) or "native"
    (exact); the older `bf16` flag, when given, means "bf16" (True) or
    "native" (False). `components` is the sampled component list of the
    misfit (the reference's par key comp, e.g. "vx,vy,vz,p"; all 9 by
    default). `autotune` times the stencil launch shapes once when the
    first medium is set, ranked by the gradient step's three kernel pairs
    (forward, recomputation with derivative store, transposed step with
    imaging), `medium` chooses between
    handing (rho, lambda, mu) to the library ("elastic") and the 8 host-scaled
    arrays ("scaled"); the two agree exactly in FP64. `graph` replays the
    gradient's pairs of forward steps and its checkpoint segments (forward
    with derivative store, transposed steps with imaging) as HIP graphs
    captured once and kept by the propagator; the results are bit-identical
    to launching every step from the host. `derivatives` ("native", the
    library's default, or "s16": half the memory, FP32 only) is the storage
    of the derivative combinations the backward recomputes and images with."""

    def __init__(self, axes: PaddedAxes, nts: int, dts: float, sub: int, alpha_cos: float, shots: Sequence[FwiShot],
                 precision: str = "f32", interval: int = 20, outer_interval: int = 0, bf16: Optional[bool] = None,
                 device: int = 0, checkpoint: str = "s16", autotune: bool = False, medium: str = "elastic",
                 components: str = ALL_COMPONENTS, graph: bool = True, derivatives: Optional[str] = None):
        self.axes = axes
        self.components = components
        self.rows = component_rows(components)
        self.nts, self.dts, self.sub = int(nts), float(dts), int(sub)
        self.dtw = self.dts / self.sub
        self.shots = list(shots)
        self.precision = precision
        self.dev = f"cuda:{device}"
        self.prop = P.Propagator(axes.nz, axes.nx, axes.ny, nts, sub, axes.dz, axes.dx, axes.dy, dts, axes.min_pad,
                                 alpha_cos, precision, device)
        if bf16 is not None:
            checkpoint = "bf16" if bf16 else "native"
        self.prop.set_gradient_checkpoints(interval, outer_interval, checkpoint)
        if derivatives is not None:
            self.prop.set_gradient_derivatives(derivatives)
        self.prop.set_graph(graph)
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
            self.tuned = autotune(self.prop, reps=5, target="gradient")
        return med8

    def _to_irregular(self, shot, reg):
        """Regular traces (device) -> 9 irregular FP64 arrays (device)."""
        torch = _torch()
        c = self._shot_device(shot)
        return [c["Mr"][COMP_GRID[k]].forward(reg[k]) for k in range(9)]

    def _to_regular(self, shot, irr):
        c = self._shot_device(shot)
        return [c["Mr"][COMP_GRID[k]].adjoint(irr[k]).to(self._tdtype()).contiguous() for k in range(9)]

    def _mask(self, shot):
        """The shot's data mask on the GPU (cached), or None."""
        if shot.data_mask is None:
            return None
        c = shot._cache
        if "mask" not in c:
            torch = _torch()
            if len(shot.data_mask) != len(self.rows):
                raise ValueError("a data mask needs one array per sampled component")
            c["mask"] = [torch.as_tensor(np.asarray(m, np.float64), device=self.dev) for m in shot.data_mask]
        return c["mask"]

    def _residual(self, shot, pred9):
        """Masked residual of the sampled components, M (C pred - d_obs)."""
        if shot.observed is None or len(shot.observed) != len(self.rows):
            raise ValueError("observed data must hold one array per sampled component (FwiProblem.synthesize)")
        pred = sample_components(self.rows, pred9)
        res = [p - o for p, o in zip(pred, shot.observed)]
        mask = self._mask(shot)
        if mask is not None:
            res = [m * r for m, r in zip(mask, res)]
        return res

    def _adjoint_source(self, shot, res, like):
        """dJ/d(pred9) = C^T M res (irregular, 9 components)."""
        torch = _torch()
        mask = self._mask(shot)
        if mask is not None:
            res = [m * r for m, r in zip(mask, res)]
        return sample_components_adjoint(self.rows, [torch.broadcast_to(r, like.shape) for r in res],
                                         lambda: torch.zeros_like(like))

    # -- public ------------------------------------------------------------
    def synthesize(self, rho, lam, mu) -> None:
        """Observed data of every shot from a (true) model (FP64 device)."""
        torch = _torch()
        med8 = self._set_model(rho, lam, mu)
        for shot in self.shots:
            c = self._shot_device(shot)
            reg = self.prop.nl_forward(c["spos"], self._scaled_sources(med8, shot), c["rpos"])
            shot.observed = [x.clone() for x in sample_components(self.rows, self._to_irregular(shot, reg))]
        torch.cuda.synchronize()

    def misfit(self, rho, lam, mu) -> float:
        """J only (plain forward propagations; FP64 residual reduction)."""
        torch = _torch()
        med8 = self._set_model(rho, lam, mu)
        J = torch.zeros((), dtype=torch.float64, device=self.dev)
        for shot in self.shots:
            c = self._shot_device(shot)
            reg = self.prop.nl_forward(c["spos"], self._scaled_sources(med8, shot), c["rpos"])
            for r in self._residual(shot, self._to_irregular(shot, reg)):
                J += 0.5 * torch.sum(r ** 2)
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
            res = self._residual(shot, pred)
            for r in res:
                J += 0.5 * torch.sum(r ** 2)
            adj = self._to_regular(shot, self._adjoint_source(shot, res, pred[0]))
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


def objective(problem, m, param: str = "rho_lambda_mu"):
    """(J, gradient) of the FWI misfit at a model m given in the inverted
    parameterisation: "rho_lambda_mu" (the propagator's own) or "vp_vs_rho"
    (the reference's mod_par=1 chain, ElasticConv_3D then the modelling,
    efwiMain_3D.py:199-208). Arrays are [ny, nx, nz]; the gradient is a
    float64 [3, ny, nx, nz] array in the same parameterisation."""
    m = [np.asarray(x, np.float64) for x in m]
    if param == "rho_lambda_mu":
        return problem.misfit_and_gradient(*m)
    if param == "vp_vs_rho":
        J, g3 = problem.misfit_and_gradient(*vp_vs_rho_to_elastic(*m))
        return J, np.stack(vp_vs_rho_gradient(m[0], m[1], m[2], g3[0], g3[1], g3[2]))
    raise ValueError("param must be 'rho_lambda_mu' or 'vp_vs_rho'")


def invert(problem, m0, n_iter: int = 5, mask=None, bounds=None, history_size: int = 10,
           line_search: str = "strong_wolfe", log=None, param: str = "rho_lambda_mu"):
    """FWI outer loop: L-BFGS (torch.optim.LBFGS) on x = m / m0 (each
    parameter normalised by the starting model, which scales the three alike),
    the gradient from the adjoint state (`objective`). `param` is the
    inverted parameterisation ("rho_lambda_mu" or "vp_vs_rho", m0 given in
    it). `mask` ([3, ny, nx, nz] or broadcastable) multiplies the gradient
    (the reference's maskGradient); `bounds` = ((min_0, min_1, min_2),
    (max_0, max_1, max_2)) in the inverted parameterisation (scalars or
    arrays; None for no bound). The misfit is evaluated at the projection
    P(x) of the iterate onto the bounds, and its gradient with respect to x
    is the gradient at P(x) with the components held at a bound (pointing
    outward) set to zero: the line search and the curvature pairs then see
    one function, J(P(x)), and its gradient. Returns (P(x) as 3 arrays,
    history of (evaluation, J)). The counterpart of the reference's
    efwiMain_3D.py driver, whose pySolver solvers are not part of this
    project."""
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

    def projected():
        out = []
        with torch.no_grad():
            for k in range(3):
                v = x[k].detach()
                if lo is not None and lo[k] is not None:
                    v = torch.maximum(v, lo[k])
                if hi is not None and hi[k] is not None:
                    v = torch.minimum(v, hi[k])
                out.append(v)
        return out

    def closure():
        opt.zero_grad()
        xp = projected()
        J, g3 = objective(problem, [xi.numpy() * mi for xi, mi in zip(xp, m0)], param)
        for k in range(3):
            g = torch.from_numpy(np.ascontiguousarray(g3[k] * m0[k]))   # d/dx = d/dm * m0
            if msk is not None:
                g = g * (msk[k] if msk.dim() == 4 else msk)
            if bounds is not None:
                xk = x[k].detach()
                if lo[k] is not None:   # at or below the lower bound: only an inward (increasing) step is free
                    g = torch.where((xk <= lo[k]) & (g > 0), torch.zeros_like(g), g)
                if hi[k] is not None:
                    g = torch.where((xk >= hi[k]) & (g < 0), torch.zeros_like(g), g)
            x[k].grad = g
        history.append((len(history), float(J)))
        if log is not None:
            log(f"[fwi] evaluation {len(history)}: J = {J:.9e}")
        return torch.tensor(J, dtype=torch.float64)

    opt.step(closure)
    return [xi.numpy() * mi for xi, mi in zip(projected(), m0)], history
