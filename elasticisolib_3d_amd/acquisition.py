"""Acquisition geometry on the 7 staggered grids.

Counterpart of the reference's ``spaceInterpGpu_3D`` objects and of the
Python geometry builders ``buildSourceGeometry_3D`` / ``buildReceiversGeometry_3D``
(python_float/Elastic_iso_float_3D.py:21-383). Weights and unique point
lists come from the library's C++ routine ``ela_interp_linear``; the trace
mapping between irregular devices and regular grid points is a sparse
matrix product.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse as sp

from . import _lib

GRID_NAMES = ("centre", "x", "y", "z", "xz", "xy", "yz")
# Half-cell origin shifts (z, x, y) of each grid (Elastic_iso_float_3D.py:97-110).
GRID_SHIFT = ((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), (1, 1, 0), (0, 1, 1), (1, 0, 1))
COMP_GRID = (1, 2, 3, 0, 0, 0, 4, 5, 6)


@dataclass
class PaddedAxes:
    """Padded model axes with the par-file padding parameters."""

    nz: int
    nx: int
    ny: int
    oz: float
    ox: float
    oy: float
    dz: float
    dx: float
    dy: float
    fat: int
    z_pad_minus: int
    z_pad_plus: int
    x_pad_minus: int
    x_pad_plus: int
    y_pad: int
    tolerance: float = 1e-6

    @property
    def min_pad(self) -> int:
        """Sponge width (fdParamElastic_3D.cpp:44-55, 88; freeSurface=0)."""
        return min(min(self.z_pad_minus, self.z_pad_plus), min(self.x_pad_minus, self.x_pad_plus))

    def grid(self, g: int) -> "PaddedAxes":
        sz, sx, sy = GRID_SHIFT[g]
        return PaddedAxes(self.nz, self.nx, self.ny, self.oz - 0.5 * self.dz * sz, self.ox - 0.5 * self.dx * sx,
                          self.oy - 0.5 * self.dy * sy, self.dz, self.dx, self.dy, self.fat, self.z_pad_minus,
                          self.z_pad_plus, self.x_pad_minus, self.x_pad_plus, self.y_pad, self.tolerance)

    def c(self) -> _lib.ElaInterpAxes:
        return _lib.ElaInterpAxes(self.nz, self.nx, self.ny, self.oz, self.ox, self.oy, self.dz, self.dx, self.dy,
                                  self.fat, self.z_pad_minus, self.z_pad_plus, self.x_pad_minus, self.x_pad_plus,
                                  self.y_pad, self.tolerance)


class SpaceInterp:
    """Interpolation of one device set onto one grid (spaceInterpGpu_3D):
    linear (8 corners) or sinc with half-length `half_length` per axis, each
    optionally with a dipole, a second negative pole at the device shifted by
    `dipole_shift` = (dz, dx, dy) in model units."""

    def __init__(self, axes: PaddedAxes, z, x, y, method: str = "linear", half_length: int = 1, dipole: int = 0,
                 dipole_shift=(0.0, 0.0, 0.0)):
        lib = _lib.load()
        z = np.ascontiguousarray(z, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        kinds = {"linear": 0, "sinc": 1}
        if method not in kinds:
            raise ValueError(f"spatial interpolation method {method!r} is not supported (linear or sinc)")
        if method == "linear" and half_length != 1:
            raise ValueError("Half-length of interpolation filter should be set to 1 for linear interpolation")
        n = z.shape[0]
        nf = (8 if method == "linear" else 8 * int(half_length) ** 3) * (2 if dipole else 1)
        nbr = np.empty(nf * max(n, 1), np.int64)
        w = np.empty(nf * max(n, 1), np.float64)
        uniq = np.empty(nf * max(n, 1), np.int64)
        nu = ctypes.c_int64()
        ax = axes.c()
        sz, sx, sy = (float(v) for v in dipole_shift)
        rc = lib.ela_interp_weights(ctypes.byref(ax), n, z.ctypes.data, x.ctypes.data, y.ctypes.data, kinds[method],
                                    int(half_length), int(dipole), sz, sx, sy, nbr.ctypes.data, w.ctypes.data,
                                    uniq.ctypes.data, ctypes.byref(nu))
        if rc != _lib.ELA_OK:
            raise ValueError("acquisition device out of bounds (or its interpolation stencil leaves the padded interior)")
        self.n_dev = n
        self.method, self.half_length, self.dipole = method, int(half_length), int(dipole)
        self.pos_unique = uniq[: nu.value].copy()
        self.nbr = nbr[: nf * n].reshape(n, nf)
        self.weight = w[: nf * n].reshape(n, nf)
        rows = np.repeat(np.arange(n), nf)
        # M: [n_dev, n_reg]; forward (reg -> irreg) is M @ reg, adjoint is M.T @ irreg (duplicate
        # entries of one device, e.g. the two poles of a dipole sharing a point, are summed)
        self.M = sp.csr_matrix((self.weight.ravel(), (rows, self.nbr.ravel())), shape=(n, nu.value))

    @property
    def n_reg(self) -> int:
        return len(self.pos_unique)

    def forward(self, reg: np.ndarray) -> np.ndarray:
        """regular [n_reg, nt] -> irregular [n_dev, nt] (spaceInterpGpu_3D.cpp:140-157)."""
        return np.asarray(self.M @ reg)

    def adjoint(self, irreg: np.ndarray) -> np.ndarray:
        """irregular [n_dev, nt] -> regular [n_reg, nt] (spaceInterpGpu_3D.cpp:159-175)."""
        return np.asarray(self.M.T @ irreg)


class ShotGeometry:
    """The 7 grid interpolators of one device set (one shot's sources or receivers)."""

    def __init__(self, axes: PaddedAxes, z, x, y, **interp):
        self.grids: List[SpaceInterp] = [SpaceInterp(axes.grid(g), z, x, y, **interp) for g in range(7)]
        self.n_dev = len(np.atleast_1d(z))

    @classmethod
    def from_grids(cls, grids: Sequence[SpaceInterp]) -> "ShotGeometry":
        """From 7 interpolators already built on the centre, x, y, z, xz, xy
        and yz grids (the per-grid lists of the geometry builders)."""
        if len(grids) != 7:
            raise ValueError("need 7 grid interpolators")
        n = {g.n_dev for g in grids}
        if len(n) != 1:
            raise ValueError("the 7 interpolators must hold the same devices")
        obj = cls.__new__(cls)
        obj.grids = list(grids)
        obj.n_dev = n.pop()
        return obj

    def positions(self) -> List[np.ndarray]:
        return [g.pos_unique for g in self.grids]

    def counts(self) -> List[int]:
        return [g.n_reg for g in self.grids]

    def to_regular(self, traces: Sequence[np.ndarray]) -> List[np.ndarray]:
        """9 irregular [n_dev, nt] -> 9 regular [n_reg(grid), nt]."""
        return [self.grids[COMP_GRID[c]].adjoint(traces[c]) for c in range(9)]

    def to_irregular(self, traces: Sequence[np.ndarray]) -> List[np.ndarray]:
        return [self.grids[COMP_GRID[c]].forward(traces[c]) for c in range(9)]


# Grid of the points at which each of the 8 scaled medium arrays scales a
# component (rhoxDtw at the X grid, ..., lamb2MuDtw and lambDtw at the centre).
_MEDIUM_GRID = (1, 2, 3, 0, 0, 4, 5, 6)


def point_medium(med8, geom: ShotGeometry) -> List[np.ndarray]:
    """The 8 scaled medium arrays at the regular points where the source
    scaling reads them (array k at the points of grid _MEDIUM_GRID[k]), as
    FP64 numpy arrays. med8 is a numpy or torch [8, ny, nx, nz] array (a
    device tensor is gathered on its device, only the point values travel)."""
    pos = geom.positions()
    if hasattr(med8, "device") and not isinstance(med8, np.ndarray):
        import torch
        flat = med8.reshape(8, -1)
        return [flat[k].index_select(0, torch.as_tensor(np.asarray(pos[g], np.int64), device=med8.device))
                .double().cpu().numpy() for k, g in enumerate(_MEDIUM_GRID)]
    m = np.asarray(med8).reshape(8, -1)
    return [np.asarray(m[k][pos[g]], np.float64) for k, g in enumerate(_MEDIUM_GRID)]


def scale_regular_traces_at(mp: Sequence[np.ndarray], reg: Sequence[np.ndarray], dxyz: float):
    """Medium scaling applied to injected traces: forces by 2dtw/rho, normal
    moments through the (lambda+2mu, lambda) block, shear moments by 2dtw mu,
    all divided by the cell volume (nonlinearPropElasticGpu_3D.cpp:200-277;
    the nonlinear adjoint applies it to the receiver traces, :481-558).
    `mp` are the medium values at the points (point_medium)."""
    out = [None] * 9
    out[0] = reg[0] * mp[0][:, None]
    out[1] = reg[1] * mp[1][:, None]
    out[2] = reg[2] * mp[2][:, None]
    l2m = mp[3][:, None]
    lam = mp[4][:, None]
    out[3] = reg[3] * l2m + (reg[4] + reg[5]) * lam
    out[4] = reg[4] * l2m + (reg[3] + reg[5]) * lam
    out[5] = reg[5] * l2m + (reg[3] + reg[4]) * lam
    out[6] = reg[6] * mp[5][:, None]
    out[7] = reg[7] * mp[6][:, None]
    out[8] = reg[8] * mp[7][:, None]
    inv = 1.0 / dxyz
    return [o * inv for o in out]


def scale_regular_traces(med8: np.ndarray, geom: ShotGeometry, reg: Sequence[np.ndarray], dxyz: float):
    """scale_regular_traces_at with the medium read from med8 (numpy or torch)."""
    return scale_regular_traces_at(point_medium(med8, geom), reg, dxyz)


def _source_scaling_terms(reg: Sequence[np.ndarray], gscaled: Sequence[np.ndarray], dxyz: float):
    """Per medium array k: the gradient values at the points of grid
    _MEDIUM_GRID[k] of a function of the scaled source traces, through
    scale_regular_traces_at (`reg` unscaled, `gscaled` the gradient with
    respect to the scaled traces)."""
    inv = 1.0 / dxyz
    r = [np.asarray(x, np.float64) for x in reg]
    s = [np.asarray(x, np.float64) for x in gscaled]
    return [
        np.sum(s[0] * r[0], axis=1) * inv,
        np.sum(s[1] * r[1], axis=1) * inv,
        np.sum(s[2] * r[2], axis=1) * inv,
        np.sum(s[3] * r[3] + s[4] * r[4] + s[5] * r[5], axis=1) * inv,
        np.sum(s[3] * (r[4] + r[5]) + s[4] * (r[3] + r[5]) + s[5] * (r[3] + r[4]), axis=1) * inv,
        np.sum(s[6] * r[6], axis=1) * inv,
        np.sum(s[7] * r[7], axis=1) * inv,
        np.sum(s[8] * r[8], axis=1) * inv,
    ]


def add_source_scaling_gradient(g8, geom: ShotGeometry, reg: Sequence[np.ndarray], gscaled: Sequence[np.ndarray],
                                dxyz: float) -> None:
    """source_scaling_gradient added into a torch FP64 [8, ny, nx, nz] tensor
    (any device; index_add_ accumulates points shared by components)."""
    import torch
    pos = geom.positions()
    flat = g8.view(8, -1)
    for k, v in enumerate(_source_scaling_terms(reg, gscaled, dxyz)):
        p = pos[_MEDIUM_GRID[k]]
        if len(p):
            flat[k].index_add_(0, torch.as_tensor(np.asarray(p, np.int64), device=g8.device),
                               torch.as_tensor(v, dtype=g8.dtype, device=g8.device))


def source_scaling_gradient(med8: np.ndarray, geom: ShotGeometry, reg: Sequence[np.ndarray],
                            gscaled: Sequence[np.ndarray], dxyz: float, out: Optional[np.ndarray] = None) -> np.ndarray:
    """Gradient, with respect to the 8 scaled medium arrays, of a function of
    the scaled source traces, through scale_regular_traces: `reg` are the
    unscaled regular traces, `gscaled` the gradient with respect to the scaled
    ones. Returns an [8, ny, nx, nz] array that is zero away from the source
    points (np.add.at accumulates points shared by components); with `out`
    (a C-contiguous float64 [8, ny, nx, nz] array) the gradient is added into
    it and `out` is returned, without a dense temporary."""
    shape = med8.shape
    if out is None:
        g = np.zeros((8, int(np.prod(shape[1:]))), np.float64)
    else:
        if out.shape != shape or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array of the medium's shape")
        g = out.reshape(8, -1)
    pos = geom.positions()
    for k, v in enumerate(_source_scaling_terms(reg, gscaled, dxyz)):
        p = pos[_MEDIUM_GRID[k]]
        if len(p):
            np.add.at(g[k], p, v)
    return out if out is not None else g.reshape(shape)
