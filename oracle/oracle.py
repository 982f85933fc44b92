"""CPU oracle for the elastic isotropic 3-D propagator (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker. The product
library never calls it.

Two layers:

* ``lib_*`` wrappers around ``oracle/lib/libela_oracle.so`` (built from
  ``oracle/ela_oracle.cpp``): the regular-grid time loops of the reference
  (nonlinear and Born, forward and adjoint), staggering and medium scaling.
* Pure-numpy host glue restating the reference's per-shot host code:
  trilinear weights and unique-point lists
  (elastic_iso_float/src/spaceInterpGpu_3D.cpp:113-175, 214-279), source and
  receiver medium scaling (elastic_iso_float/src/nonlinearPropElasticGpu_3D.cpp:200-277,
  481-558), Born perturbation staggering/scaling and un-staggering
  (elastic_iso_float/src/BornElasticShotsGpu_3D.cpp:167-235, 541-607).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libela_oracle.so")
_lib = None

FAT = 4
# Grid of each component (vx vy vz sxx syy szz sxz sxy syz); grids are
# 0=centre 1=X 2=Y 3=Z 4=XZ 5=XY 6=YZ.
COMP_GRID = (1, 2, 3, 0, 0, 0, 4, 5, 6)
# Half-cell origin shifts (z, x, y) of each grid
# (python_float/Elastic_iso_float_3D.py:97-110).
GRID_SHIFT = (
    (0, 0, 0),  # centre
    (0, 1, 0),  # X
    (0, 0, 1),  # Y
    (1, 0, 0),  # Z
    (1, 1, 0),  # XZ
    (0, 1, 1),  # XY
    (1, 0, 1),  # YZ
)
FD_COEFFS = (1.196289062541883, -0.079752604188901, 0.009570312506634, -6.975446437140719e-04)


class OrcGrid(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int64), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("nts", ctypes.c_int64), ("sub", ctypes.c_int64),
        ("dz", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("dts", ctypes.c_double), ("minPad", ctypes.c_int64), ("alphaCos", ctypes.c_double),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle library with its Makefile (CPU only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def num_threads() -> int:
    return int(lib().orc_num_threads())


@dataclass
class Grid:
    """Propagation grid: padded sizes and the reference's fine-time parameters.

    ``sub`` is the doubled sub-sampling the reference propagates with
    (elastic_iso_float/src/fdParamElastic_3D.cpp:29-36).
    """

    nz: int
    nx: int
    ny: int
    nts: int
    sub: int
    dz: float
    dx: float
    dy: float
    dts: float
    minPad: int
    alphaCos: float = 0.99

    @property
    def dtw(self) -> float:
        return self.dts / self.sub

    @property
    def shape(self):
        return (self.ny, self.nx, self.nz)

    def c(self) -> OrcGrid:
        return OrcGrid(self.nz, self.nx, self.ny, self.nts, self.sub, self.dz, self.dx,
                       self.dy, self.dts, self.minPad, self.alphaCos)


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _ct(dtype):
    return ctypes.c_double if np.dtype(dtype) == np.float64 else ctypes.c_float


def _pack_pos(pos_lists):
    counts = np.array([len(p) for p in pos_lists], dtype=np.int64)
    flat = np.concatenate([np.asarray(p, dtype=np.int64) for p in pos_lists]) if counts.sum() else np.zeros(1, np.int64)
    return np.ascontiguousarray(flat), counts


def _pack_traces(traces, counts, nts, dtype):
    parts = []
    for c in range(9):
        t = np.asarray(traces[c], dtype=dtype).reshape(counts[COMP_GRID[c]], nts)
        parts.append(t.ravel())
    return np.ascontiguousarray(np.concatenate(parts)) if parts else np.zeros(1, dtype)


def _unpack_traces(flat, counts, nts):
    out, off = [], 0
    for c in range(9):
        n = counts[COMP_GRID[c]] * nts
        out.append(flat[off:off + n].reshape(counts[COMP_GRID[c]], nts).copy())
        off += n
    return out


def lib_nl_forward(g: Grid, med8, src_traces, src_pos, rec_pos, dtype=np.float64):
    """Regular-grid nonlinear forward. ``med8`` is [8, ny, nx, nz] scaled medium;
    ``src_traces[c]`` is [nReg(grid of c), nts]; positions are 7 lists."""
    L = lib()
    sp, sn = _pack_pos(src_pos)
    rp, rn = _pack_pos(rec_pos)
    med = np.ascontiguousarray(med8, dtype=dtype)
    src = _pack_traces(src_traces, sn, g.nts, dtype)
    rec = np.zeros(max(1, int(sum(rn[COMP_GRID[c]] for c in range(9)) * g.nts)), dtype)
    ct = _ct(dtype)
    gc = g.c()
    getattr(L, "orc_nl_forward_" + _sfx(dtype))(
        ctypes.byref(gc), _ptr(med, ct), _ptr(src, ct), _ptr(sp, ctypes.c_int64), _ptr(sn, ctypes.c_int64),
        _ptr(rec, ct), _ptr(rp, ctypes.c_int64), _ptr(rn, ctypes.c_int64))
    return _unpack_traces(rec, rn, g.nts)


def lib_nl_adjoint(g: Grid, med8, rec_traces, src_pos, rec_pos, dtype=np.float64):
    L = lib()
    sp, sn = _pack_pos(src_pos)
    rp, rn = _pack_pos(rec_pos)
    med = np.ascontiguousarray(med8, dtype=dtype)
    rec = _pack_traces(rec_traces, rn, g.nts, dtype)
    src = np.zeros(max(1, int(sum(sn[COMP_GRID[c]] for c in range(9)) * g.nts)), dtype)
    ct = _ct(dtype)
    gc = g.c()
    getattr(L, "orc_nl_adjoint_" + _sfx(dtype))(
        ctypes.byref(gc), _ptr(med, ct), _ptr(src, ct), _ptr(sp, ctypes.c_int64), _ptr(sn, ctypes.c_int64),
        _ptr(rec, ct), _ptr(rp, ctypes.c_int64), _ptr(rn, ctypes.c_int64))
    return _unpack_traces(src, sn, g.nts)


def lib_born_forward(g: Grid, med8, src_traces, src_pos, dmod8, rec_pos, dtype=np.float64, physics=False):
    """Born forward (BornElasticGpuFunctions_3D.cu:1226-1461). ``physics=True``
    replaces the reference's xz derivative pattern of the yz secondary source
    (kernelsElaGpu_3D.cu:1419-1431) with the yz pattern of the stepper."""
    L = lib()
    sp, sn = _pack_pos(src_pos)
    rp, rn = _pack_pos(rec_pos)
    med = np.ascontiguousarray(med8, dtype=dtype)
    dm = np.ascontiguousarray(dmod8, dtype=dtype)
    src = _pack_traces(src_traces, sn, g.nts, dtype)
    rec = np.zeros(max(1, int(sum(rn[COMP_GRID[c]] for c in range(9)) * g.nts)), dtype)
    ct = _ct(dtype)
    gc = g.c()
    getattr(L, "orc_born_forward_" + _sfx(dtype))(
        ctypes.byref(gc), _ptr(med, ct), _ptr(src, ct), _ptr(sp, ctypes.c_int64), _ptr(sn, ctypes.c_int64),
        _ptr(dm, ct), _ptr(rec, ct), _ptr(rp, ctypes.c_int64), _ptr(rn, ctypes.c_int64), ctypes.c_int(int(physics)))
    return _unpack_traces(rec, rn, g.nts)


def lib_born_adjoint(g: Grid, med8, src_traces, src_pos, rec_traces, rec_pos, dtype=np.float64, physics=False):
    """Returns the 8 accumulated (scaled, staggered) perturbation images
    (BornElasticGpuFunctions_3D.cu:2103-2336); ``physics`` as for
    lib_born_forward (kernelsElaGpu_3D.cu:1622-1634)."""
    L = lib()
    sp, sn = _pack_pos(src_pos)
    rp, rn = _pack_pos(rec_pos)
    med = np.ascontiguousarray(med8, dtype=dtype)
    src = _pack_traces(src_traces, sn, g.nts, dtype)
    rec = _pack_traces(rec_traces, rn, g.nts, dtype)
    dm = np.zeros((8,) + g.shape, dtype)
    ct = _ct(dtype)
    gc = g.c()
    getattr(L, "orc_born_adjoint_" + _sfx(dtype))(
        ctypes.byref(gc), _ptr(med, ct), _ptr(src, ct), _ptr(sp, ctypes.c_int64), _ptr(sn, ctypes.c_int64),
        _ptr(dm, ct), _ptr(rec, ct), _ptr(rp, ctypes.c_int64), _ptr(rn, ctypes.c_int64), ctypes.c_int(int(physics)))
    return dm


def lib_gradient(g: Grid, med8, src_traces, src_pos, adj_traces, rec_pos, dtype=np.float64):
    """Adjoint-state gradient of <adj, nl_forward(med8)> with respect to the 8
    scaled medium arrays (fixed scaled source traces), and its gradient with
    respect to the scaled source traces. Returns (g8 [8, ny, nx, nz], gsrc
    [9 arrays [nReg, nts]])."""
    L = lib()
    sp, sn = _pack_pos(src_pos)
    rp, rn = _pack_pos(rec_pos)
    med = np.ascontiguousarray(med8, dtype=dtype)
    src = _pack_traces(src_traces, sn, g.nts, dtype)
    adj = _pack_traces(adj_traces, rn, g.nts, dtype)
    g8 = np.zeros((8,) + g.shape, dtype)
    gsrc = np.zeros_like(src)
    ct = _ct(dtype)
    gc = g.c()
    getattr(L, "orc_gradient_" + _sfx(dtype))(
        ctypes.byref(gc), _ptr(med, ct), _ptr(src, ct), _ptr(sp, ctypes.c_int64), _ptr(sn, ctypes.c_int64),
        _ptr(adj, ct), _ptr(rp, ctypes.c_int64), _ptr(rn, ctypes.c_int64), _ptr(g8, ct), _ptr(gsrc, ct))
    return g8, _unpack_traces(gsrc, sn, g.nts)


def lib_imaging(g: Grid, V_o, V_c, V_n, its, dm8=None, ss9=None, adjoint=False, physics=False):
    """The Born imaging kernels alone (FP64): forward (imagingElaFwdGpu_3D,
    kernelsElaGpu_3D.cu:1232-1438) maps dm8 to the 9 secondary sources;
    adjoint (imagingElaAdjGpu_3D, :1440-1641) maps ss9 to the 8 images. V_* are
    [3, ny, nx, nz] coarse velocity slices (None at the ends of the record)."""
    L = lib()
    gc = g.c()
    ct = ctypes.c_double
    def arr(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    vo, vc, vn = arr(V_o), arr(V_c), arr(V_n)
    po = _ptr(vo, ct) if vo is not None else None
    pn = _ptr(vn, ct) if vn is not None else None
    if not adjoint:
        dm = arr(dm8)
        ss = np.zeros((9,) + g.shape)
        L.orc_imaging_fwd_f64(ctypes.byref(gc), po, _ptr(vc, ct), pn, ctypes.c_int64(its), _ptr(dm, ct),
                              _ptr(ss, ct), ctypes.c_int(int(physics)))
        return ss
    ss = arr(ss9)
    dm = np.zeros((8,) + g.shape)
    L.orc_imaging_adj_f64(ctypes.byref(gc), po, _ptr(vc, ct), pn, ctypes.c_int64(its), _ptr(dm, ct),
                          _ptr(ss, ct), ctypes.c_int(int(physics)))
    return dm


def lib_stagger(axis: int, adjoint: bool, a, dtype=np.float64):
    """axis 0=z 1=x 2=y on a [ny, nx, nz] array."""
    a = np.ascontiguousarray(a, dtype=dtype)
    out = np.empty_like(a)
    ny, nx, nz = a.shape
    getattr(lib(), "orc_stagger_" + _sfx(dtype))(int(axis), int(adjoint), nz, nx, ny, _ptr(a, _ct(dtype)), _ptr(out, _ct(dtype)))
    return out


def lib_scale_medium(rho, lam, mu, dtw, dtype=np.float64):
    rho = np.ascontiguousarray(rho, dtype=dtype)
    lam = np.ascontiguousarray(lam, dtype=dtype)
    mu = np.ascontiguousarray(mu, dtype=dtype)
    ny, nx, nz = rho.shape
    out = np.zeros((8, ny, nx, nz), dtype)
    ct = _ct(dtype)
    getattr(lib(), "orc_scale_medium_" + _sfx(dtype))(nz, nx, ny, ctypes.c_double(dtw), _ptr(rho, ct), _ptr(lam, ct), _ptr(mu, ct), _ptr(out, ct))
    return out


def lib_step(g: Grid, med8, p0, p1, mode: int, dtype=np.float64):
    """One stepper application: mode 0 forward, 1 nonlinear adjoint, 2 Born adjoint.
    p0/p1 are [9, ny, nx, nz]; returns the new p0."""
    med = np.ascontiguousarray(med8, dtype=dtype)
    a = np.ascontiguousarray(p0, dtype=dtype).copy()
    b = np.ascontiguousarray(p1, dtype=dtype)
    ct = _ct(dtype)
    gc = g.c()
    getattr(lib(), "orc_step_" + _sfx(dtype))(ctypes.byref(gc), _ptr(med, ct), _ptr(a, ct), _ptr(b, ct), int(mode))
    return a


def lib_tables(g: Grid):
    filt = np.zeros(2 * (g.sub + 1))
    taper = np.zeros(max(1, g.minPad))
    gc = g.c()
    lib().orc_tables(ctypes.byref(gc), _ptr(filt, ctypes.c_double), _ptr(taper, ctypes.c_double))
    return filt, taper[: g.minPad]


# ---------------------------------------------------------------------------
# Host glue restated in numpy
# ---------------------------------------------------------------------------

@dataclass
class Interp:
    """One spaceInterpGpu_3D object (linear method, no dipole)."""

    pos_unique: np.ndarray  # [nReg] int64 flat indices (first-seen order)
    idx: np.ndarray         # [nDev, 8] index into pos_unique
    weight: np.ndarray      # [nDev, 8]

    @property
    def n_reg(self):
        return len(self.pos_unique)

    def adjoint(self, irreg):
        """irregular [nDev, nt] -> regular [nReg, nt]
        (spaceInterpGpu_3D.cpp:159-175)."""
        out = np.zeros((self.n_reg, irreg.shape[1]), dtype=np.float64)
        for f in range(8):
            np.add.at(out, self.idx[:, f], self.weight[:, f:f + 1] * irreg)
        return out

    def forward(self, reg):
        """regular [nReg, nt] -> irregular [nDev, nt] (spaceInterpGpu_3D.cpp:140-157)."""
        out = np.zeros((self.idx.shape[0], reg.shape[1]), dtype=np.float64)
        for f in range(8):
            out += self.weight[:, f:f + 1] * reg[self.idx[:, f]]
        return out


def linear_interp(zc, xc, yc, nz, nx, ny, oz, ox, oy, dz, dx, dy, fat, zpm, zpp, xpm, xpp, ypad):
    """calcLinearWeights + convertIrregToReg (spaceInterpGpu_3D.cpp:113-138, 214-279).

    Weight arithmetic follows the reference's single-precision steps: the
    fractional position is formed in float32.
    """
    zc = np.asarray(zc, np.float32)
    xc = np.asarray(xc, np.float32)
    yc = np.asarray(yc, np.float32)
    wz = (zc - np.float32(oz)) / np.float32(dz)
    wx = (xc - np.float32(ox)) / np.float32(dx)
    wy = (yc - np.float32(oy)) / np.float32(dy)
    zr = np.trunc(wz).astype(np.int64)
    xr = np.trunc(wx).astype(np.int64)
    yr = np.trunc(wy).astype(np.int64)
    wz = np.float32(1.0) - (wz - zr.astype(np.float32))
    wx = np.float32(1.0) - (wx - xr.astype(np.float32))
    wy = np.float32(1.0) - (wy - yr.astype(np.float32))
    if np.any(yr < fat + ypad) or np.any(yr + 1 >= ny - fat - ypad):
        raise ValueError("linear interpolation point out of bounds on the y-axis")
    if np.any(xr < fat + xpm) or np.any(xr + 1 >= nx - fat - xpp):
        raise ValueError("linear interpolation point out of bounds on the x-axis")
    if np.any(zr < fat + zpm) or np.any(zr + 1 >= nz - fat - zpp):
        raise ValueError("linear interpolation point out of bounds on the z-axis")
    wz = wz.astype(np.float64)
    wx = wx.astype(np.float64)
    wy = wy.astype(np.float64)
    base = yr * nz * nx + xr * nz + zr
    offs = [0, 1, nz, nz + 1, nz * nx, nz * nx + 1, nz * nx + nz, nz * nx + nz + 1]
    wts = [wz * wx * wy, (1 - wz) * wx * wy, wz * (1 - wx) * wy, (1 - wz) * (1 - wx) * wy,
           wz * wx * (1 - wy), (1 - wz) * wx * (1 - wy), wz * (1 - wx) * (1 - wy), (1 - wz) * (1 - wx) * (1 - wy)]
    gp = np.stack([base + o for o in offs], axis=1)
    wt = np.stack(wts, axis=1).astype(np.float32).astype(np.float64)
    flat = gp.ravel()
    uniq, first = np.unique(flat, return_index=True)
    order = np.argsort(first)            # first-seen order
    pos_unique = uniq[order]
    rank = np.empty(len(uniq), np.int64)
    rank[order] = np.arange(len(uniq))
    idx = rank[np.searchsorted(uniq, flat)].reshape(gp.shape)
    return Interp(pos_unique.astype(np.int64), idx, wt)


def scale_sources(med8, interps, reg, dxyz, adjoint_data=False):
    """Medium scaling of regular-grid traces (nonlinearPropElasticGpu_3D.cpp:200-277;
    the nonlinear adjoint applies the same to receiver traces, :481-558).
    ``reg[c]`` is [nReg(grid c), nts]; ``interps[g]`` the 7 grids."""
    m = [med8[i].ravel() for i in range(8)]
    out = [r.copy() for r in reg]
    p = lambda g: interps[g].pos_unique  # noqa: E731
    out[0] = reg[0] * m[0][p(1)][:, None]
    out[1] = reg[1] * m[1][p(2)][:, None]
    out[2] = reg[2] * m[2][p(3)][:, None]
    l2m = m[3][p(0)][:, None]
    lam = m[4][p(0)][:, None]
    out[3] = reg[3] * l2m + (reg[4] + reg[5]) * lam
    out[4] = reg[4] * l2m + (reg[3] + reg[5]) * lam
    out[5] = reg[5] * l2m + (reg[3] + reg[4]) * lam
    out[6] = reg[6] * m[5][p(4)][:, None]
    out[7] = reg[7] * m[6][p(5)][:, None]
    out[8] = reg[8] * m[7][p(6)][:, None]
    return [o / dxyz for o in out]


def born_stagger_perturbation(dm3, med8, dtw, stagger_fn):
    """(drho, dlambda, dmu) [3, ny, nx, nz] -> 8 staggered, scaled arrays
    (BornElasticShotsGpu_3D.cpp:167-235)."""
    drho, dlam, dmu = dm3
    out = np.zeros((8,) + drho.shape)
    out[0] = stagger_fn(1, True, drho) * med8[0]
    out[1] = stagger_fn(2, True, drho) * med8[1]
    out[2] = stagger_fn(0, True, drho) * med8[2]
    out[3] = dlam * 2.0 * dtw
    out[4] = dmu * 2.0 * dtw
    out[5] = stagger_fn(0, True, stagger_fn(1, True, dmu)) * 2.0 * dtw
    out[6] = stagger_fn(2, True, stagger_fn(1, True, dmu)) * 2.0 * dtw
    out[7] = stagger_fn(0, True, stagger_fn(2, True, dmu)) * 2.0 * dtw
    return out


def born_unstagger_image(img8, med8, dtw, stagger_fn):
    """Adjoint of born_stagger_perturbation (BornElasticShotsGpu_3D.cpp:541-607)."""
    a = [img8[0] * med8[0], img8[1] * med8[1], img8[2] * med8[2]] + [img8[k] * 2.0 * dtw for k in range(3, 8)]
    drho = stagger_fn(1, False, a[0]) + stagger_fn(2, False, a[1]) + stagger_fn(0, False, a[2])
    dlam = a[3]
    dmu = (a[4] + stagger_fn(1, False, stagger_fn(0, False, a[5]))
           + stagger_fn(1, False, stagger_fn(2, False, a[6]))
           + stagger_fn(2, False, stagger_fn(0, False, a[7])))
    return np.stack([drho, dlam, dmu])
