"""Synthetic models, wavelets and acquisition layouts of the benchmark workloads.

Recipes follow the workload definitions of this project's benchmark (a
homogeneous 100^3 model, a seeded 12-layer model on 400x400x250 cells, ...).
Padding follows the reference's padding tool
(elastic_iso_float/src/padElasticFileGpuMain_3D.cpp:31-155): edge replication,
with zPadPlus / xPadPlus chosen so that the padded interior is a multiple of
the block size; the y pads are symmetric. The Ricker wavelet is the
reference's (python_float/waveletFloatMain.py:88-111) and the (Vp, Vs, rho)
-> (rho, lambda, mu) conversion is ElasticConv_3D type 1
(seis_utils/python/elasticParamConvertModule_3D.py:23-47).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .acquisition import PaddedAxes

FAT = 4


def ricker(nts: int, dts: float, f_dom: float, delay: float) -> np.ndarray:
    alpha = (math.pi * f_dom) ** 2
    t = np.arange(nts) * dts - delay
    return (1.0 - 2.0 * alpha * t * t) * np.exp(-alpha * t * t)


def vpvsrho_to_rholamemu(vp, vs, rho):
    return rho, rho * (vp * vp - 2.0 * vs * vs), rho * vs * vs


def pad_sizes(n: int, pad: int, block: int = 16):
    """(padMinus, padPlus) so that pad + n + padPlus is a multiple of `block`."""
    total = math.ceil(float(2 * pad + n) / float(block)) * block
    return pad, total - n - pad


def pad_model(m: np.ndarray, z_pad: int, x_pad: int, y_pad: int, fat: int = FAT, block: int = 16):
    """Edge-replicating padding of a [..., ny, nx, nz] model.

    Returns the padded array and (zPadMinus, zPadPlus, xPadMinus, xPadPlus, yPad)."""
    ny, nx, nz = m.shape[-3:]
    zm, zp = pad_sizes(nz, z_pad, block)
    xm, xp = pad_sizes(nx, x_pad, block)
    widths = [(0, 0)] * (m.ndim - 3) + [(fat + y_pad, fat + y_pad), (fat + xm, fat + xp), (fat + zm, fat + zp)]
    return np.pad(m, widths, mode="edge"), (zm, zp, xm, xp, y_pad)


def gaussian_random_field(shape, sigma_cells: float, rng: np.random.Generator) -> np.ndarray:
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal(shape).astype(np.float32), sigma=sigma_cells, mode="reflect")
    return g / max(float(np.max(np.abs(g))), 1e-30)


def _smoothing_matrix(n: int, w: np.ndarray, r: int) -> np.ndarray:
    """The n x n matrix of a 1-D filter with weights w[0 .. 2r] centred on
    each sample and the array extended by its nearest value: row i sums
    w[k] into column clip(i + k - r, 0, n - 1)."""
    cols = np.clip(np.arange(n)[:, None] + np.arange(-r, r + 1)[None, :], 0, n - 1)
    m = np.zeros((n, n))
    np.add.at(m, (np.repeat(np.arange(n), 2 * r + 1), cols.ravel()), np.tile(w, n))
    return m


def gaussian_smooth(a: np.ndarray, sigma: float, device: Optional[str] = None) -> np.ndarray:
    """Gaussian smoothing of a 3-D array with scipy.ndimage.gaussian_filter's
    kernel (radius int(4 sigma + 0.5), weights exp(-x^2 / (2 sigma^2))
    normalised, edges extended by the nearest value), each separable pass
    rounded to float32 as scipy stores it. With `device` (e.g. "cuda") the
    passes run on that torch device as products with the n x n filter matrix
    of each axis (FP64 GEMMs: precompiled BLAS kernels, no convolution
    library to compile or tune on a fresh device), a fraction of a second
    instead of tens of seconds for a 400 x 400 x 250 model with sigma = 20."""
    if device is None:
        from scipy.ndimage import gaussian_filter
        return gaussian_filter(a.astype(np.float32), sigma, mode="nearest").astype(np.float64)
    import torch
    r = int(4.0 * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (x / sigma) ** 2)
    w /= w.sum()
    t = torch.tensor(np.asarray(a, np.float32), device=device)
    for ax in range(3):
        m = torch.tensor(_smoothing_matrix(t.shape[ax], w, r), dtype=torch.float64, device=device)
        u = t.movedim(ax, -1).double()
        t = torch.matmul(u, m.T).float().movedim(-1, ax).contiguous()
        del u, m
    return t.double().cpu().numpy()


def layered_model(nz: int, nx: int, ny: int, seed: int, n_layers: int = 12, vp_range=(1800.0, 4500.0),
                  pert: float = 0.05, sigma: float = 8.0):
    """Seeded 1-D layered (Vp, Vs, rho) with a smoothed random perturbation.

    Vp of equal-thickness layers: sorted U(vp_range), increasing with depth;
    Vs = Vp/sqrt(3) * U(0.9, 1.0) per layer; Vp and Vs scaled by
    (1 + pert * g / max|g|) with g a Gaussian-smoothed N(0,1) field; Gardner
    density rho = 310 Vp^0.25. Arrays are [ny, nx, nz] float64."""
    rng = np.random.default_rng(seed)
    vp_l = np.sort(rng.uniform(vp_range[0], vp_range[1], n_layers))
    vs_l = vp_l / math.sqrt(3.0) * rng.uniform(0.9, 1.0, n_layers)
    layer = np.minimum((np.arange(nz) * n_layers) // nz, n_layers - 1)
    vp1, vs1 = vp_l[layer], vs_l[layer]
    g = gaussian_random_field((ny, nx, nz), sigma, rng) if pert > 0 else np.zeros((ny, nx, nz), np.float32)
    scale = 1.0 + pert * g.astype(np.float64)
    vp = vp1[None, None, :] * scale
    vs = vs1[None, None, :] * scale
    rho = 310.0 * vp ** 0.25
    return vp, vs, rho


@dataclass
class Workload:
    """A padded model with its time axis and one shot's acquisition."""

    name: str
    rho: np.ndarray          # [ny, nx, nz] padded
    lame: np.ndarray
    mu: np.ndarray
    axes: PaddedAxes
    nts: int
    dts: float
    sub: int                 # fine steps per coarse sample (2 x par sub)
    alpha_cos: float
    src_xyz: np.ndarray      # [n, 3] (x, y, z) model coordinates [m]
    rec_xyz: np.ndarray      # [m, 3]
    wavelet: np.ndarray      # [nts]
    source_type: str         # "explosive" or "fz"
    extra: dict = field(default_factory=dict)

    @property
    def dtw(self) -> float:
        return self.dts / self.sub

    @property
    def interior_cells(self) -> int:
        a = self.axes
        return (a.nz - 2 * FAT) * (a.nx - 2 * FAT) * (a.ny - 2 * FAT)

    @property
    def fine_steps(self) -> int:
        return (self.nts - 1) * self.sub


def _axes(nz, nx, ny, d, pads, fat=FAT, o=(0.0, 0.0, 0.0)):
    zm, zp, xm, xp, yp = pads
    return PaddedAxes(nz=nz, nx=nx, ny=ny, oz=o[0] - (fat + zm) * d, ox=o[1] - (fat + xm) * d,
                      oy=o[2] - (fat + yp) * d, dz=d, dx=d, dy=d, fat=fat, z_pad_minus=zm, z_pad_plus=zp,
                      x_pad_minus=xm, x_pad_plus=xp, y_pad=yp)


def homogeneous_100(nts: int = 151) -> Workload:
    """100^3 cells, d = 10 m, 20-cell sponge, Vp 2500, Vs 1440, rho 2000; vertical
    point force at the centre, 15 Hz Ricker (delay 1/f), dts = 4 ms with 4
    fine sub-steps per coarse sample of the par file (8 after doubling),
    0.6 s; 81 receivers on a line at z = 5 cells."""
    n, d, pad = 100, 10.0, 20
    vp = np.full((n, n, n), 2500.0)
    vs = np.full((n, n, n), 1440.0)
    rho0 = np.full((n, n, n), 2000.0)
    rho, lam, mu = vpvsrho_to_rholamemu(vp, vs, rho0)
    m, pads = pad_model(np.stack([rho, lam, mu]), pad, pad, pad)
    ny, nx, nz = m.shape[1:]
    ax = _axes(nz, nx, ny, d, pads)
    dts = 0.004
    src = np.array([[50 * d, 50 * d, 50 * d]])
    rec = np.array([[x * d, 50 * d, 5 * d] for x in range(10, 91)])
    return Workload("homogeneous_100", m[0], m[1], m[2], ax, nts, dts, 8, 0.99, src, rec,
                    ricker(nts, dts, 15.0, 1.0 / 15.0), "fz")


def layered_400(nts: int = 626, seed: int = 2024, n_rec_axis: int = 200, sigma: float = 8.0) -> Workload:
    """400 x 400 x 250 cells (x, y, z), d = 12.5 m, 12-layer seeded model with a
    5% smoothed perturbation (sigma 8 cells), 30-cell sponge; explosive 10 Hz
    Ricker at the centre, z = 10 cells; 2.5 s (626 samples of 4 ms, 8 fine
    steps each); 200 x 200 surface receivers every 2 cells."""
    nxm, nym, nzm, d, pad = 400, 400, 250, 12.5, 30
    vp, vs, rho0 = layered_model(nzm, nxm, nym, seed, sigma=sigma)
    rho, lam, mu = vpvsrho_to_rholamemu(vp, vs, rho0)
    del vp, vs, rho0
    m, pads = pad_model(np.stack([rho, lam, mu]), pad, pad, pad)
    del rho, lam, mu
    ny, nx, nz = m.shape[1:]
    ax = _axes(nz, nx, ny, d, pads)
    dts = 0.004
    src = np.array([[200 * d, 200 * d, 10 * d]])
    step = 2
    xs = np.arange(n_rec_axis) * step * d
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    rec = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    return Workload("layered_400x400x250", m[0], m[1], m[2], ax, nts, dts, 8, 0.99, src, rec,
                    ricker(nts, dts, 10.0, 0.1), "explosive")


def born_layered_400(nts: int = 501, seed: int = 2024, n_rec_axis: int = 150, sigma_bg: float = 15.0,
                     density: float = 0.01, rel_amp: float = 0.02, seed_dm: int = 7) -> Workload:
    """Born workload on the 400 x 400 x 250 layered model: background = the
    model of layered_400 smoothed with a Gaussian of sigma_bg cells (Vp and
    Vs smoothed, Gardner density from the smoothed Vp); sparse reflectivity
    drho, dlambda, dmu nonzero on `density` of the model cells, N(0, rel_amp *
    background). One explosive shot at the centre (z = 10 cells), 2.0 s
    (501 samples of 4 ms, 8 fine steps each), 150 x 150 surface receivers
    every 2 cells. The perturbations are in extra["dm3"] on the padded grid
    (zero in the padding)."""
    from scipy.ndimage import gaussian_filter
    nxm, nym, nzm, d, pad = 400, 400, 250, 12.5, 30
    vp, vs, _ = layered_model(nzm, nxm, nym, seed, sigma=8.0)
    vp = gaussian_filter(vp.astype(np.float32), sigma_bg, mode="nearest").astype(np.float64)
    vs = gaussian_filter(vs.astype(np.float32), sigma_bg, mode="nearest").astype(np.float64)
    rho0 = 310.0 * vp ** 0.25
    rho, lam, mu = vpvsrho_to_rholamemu(vp, vs, rho0)
    del vp, vs, rho0
    rng = np.random.default_rng(seed_dm)
    ncell = rho.size
    k = max(1, int(round(density * ncell)))
    idx = rng.choice(ncell, size=k, replace=False)
    dm = np.zeros((3, ncell))
    for i, base in enumerate((rho, lam, mu)):
        dm[i, idx] = rng.standard_normal(k) * rel_amp * base.ravel()[idx]
    dm = dm.reshape((3,) + rho.shape)
    m, pads = pad_model(np.stack([rho, lam, mu]), pad, pad, pad)
    del rho, lam, mu
    zm, zp, xm, xp, yp = pads
    dmp = np.zeros_like(m)
    dmp[:, FAT + yp:FAT + yp + nym, FAT + xm:FAT + xm + nxm, FAT + zm:FAT + zm + nzm] = dm
    ny, nx, nz = m.shape[1:]
    ax = _axes(nz, nx, ny, d, pads)
    dts = 0.004
    src = np.array([[200 * d, 200 * d, 10 * d]])
    xs = np.arange(n_rec_axis) * 2 * d + (200 - n_rec_axis) * d
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    rec = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    return Workload("born_layered_400x400x250", m[0], m[1], m[2], ax, nts, dts, 8, 0.99, src, rec,
                    ricker(nts, dts, 10.0, 0.1), "explosive", extra={"dm3": dmp})


def fwi_256(nts: int = 501, seed: int = 2025, n_shot_axis: int = 2, rec_step: int = 4, sigma_start: float = 20.0,
            smooth_device: Optional[str] = None):
    """FWI workload: true model = the layered recipe on 256 x 256 x 160 cells
    (x, y, z), d = 15 m, 30-cell sponge; starting model = the true one
    smoothed with a sigma_start-cell Gaussian; 2 x 2 explosive shots at
    z = 10 cells (8 Hz Ricker, delay 0.125 s), 2.0 s = 501 samples of 4 ms
    with 8 fine steps each; a seeded surface receiver grid every `rec_step`
    cells (jittered by up to a quarter cell). Returns (true Workload, start
    (rho, lambda, mu) on the padded grid, list of source positions).
    `smooth_device` runs the starting model's smoothing on that torch device
    (gaussian_smooth)."""
    nxm, nym, nzm, d, pad = 256, 256, 160, 15.0, 30
    vp, vs, rho0 = layered_model(nzm, nxm, nym, seed, sigma=8.0)
    vp_s = gaussian_smooth(vp, sigma_start, smooth_device)
    vs_s = gaussian_smooth(vs, sigma_start, smooth_device)
    rho_s = gaussian_smooth(rho0, sigma_start, smooth_device)
    true = np.stack(vpvsrho_to_rholamemu(vp, vs, rho0))
    start = np.stack(vpvsrho_to_rholamemu(vp_s, vs_s, rho_s))
    del vp, vs, rho0, vp_s, vs_s, rho_s
    mt, pads = pad_model(true, pad, pad, pad)
    ms, _ = pad_model(start, pad, pad, pad)
    ny, nx, nz = mt.shape[1:]
    ax = _axes(nz, nx, ny, d, pads)
    dts = 0.004
    rng = np.random.default_rng(seed + 1)
    xs = np.arange(0, nxm - 1, rec_step) * d
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    jit = rng.uniform(-0.25, 0.25, (2,) + gx.shape) * d
    rx = np.clip(gx + jit[0], 0, (nxm - 2) * d)
    ry = np.clip(gy + jit[1], 0, (nym - 2) * d)
    rec = np.stack([rx.ravel(), ry.ravel(), np.zeros(rx.size)], axis=1)
    shots = []
    for i in range(n_shot_axis):
        for j in range(n_shot_axis):
            shots.append(np.array([[(i + 0.5) * nxm / n_shot_axis * d, (j + 0.5) * nym / n_shot_axis * d, 10 * d]]))
    w = Workload("fwi_256x256x160", mt[0], mt[1], mt[2], ax, nts, dts, 8, 0.99, shots[0], rec,
                 ricker(nts, dts, 8.0, 0.125), "explosive", extra={"start": ms, "shots": shots})
    return w


def batch_layered_400(nts: int = 626, seed: int = 2024, seed_geom: int = 11, sigma_start: float = 20.0,
                      n_axis: int = 8, pitch: int = 50, win_range=(180, 300), rec_range=(60, 200),
                      smooth_device: Optional[str] = None):
    """Multi-shot workload with shot-local windows on the 400 x 400 x 250
    layered model (d = 12.5 m, 30-cell sponge around every window): 8 x 8
    shots on a jittered 50-cell surface grid (explosive 10 Hz Ricker at
    z = 10 cells), each with a window of U(180, 300) cells per horizontal axis
    centred on the shot (clipped to the model) and U(60, 200) receivers
    uniformly placed at the surface inside its window; 2.5 s = 626 samples of
    4 ms, 8 fine steps each. Returns a dict with the true and starting
    (rho, lambda, mu) [ny, nx, nz] (start: the true model smoothed with a
    sigma_start-cell Gaussian) and the shots (src_xyz, rec_xyz, window
    (x0, y0, nx, ny)). `smooth_device` runs the starting model's smoothing
    on that torch device (gaussian_smooth)."""
    nxm, nym, nzm, d = 400, 400, 250, 12.5
    vp, vs, rho0 = layered_model(nzm, nxm, nym, seed, sigma=8.0)
    true = np.stack(vpvsrho_to_rholamemu(vp, vs, rho0))
    sm = [gaussian_smooth(a, sigma_start, smooth_device) for a in (vp, vs, rho0)]
    start = np.stack(vpvsrho_to_rholamemu(*sm))
    del vp, vs, rho0, sm
    rng = np.random.default_rng(seed_geom)
    shots = []
    for i in range(n_axis):
        for j in range(n_axis):
            sx = (i + 0.5) * (nxm / n_axis) + rng.uniform(-0.25, 0.25) * pitch
            sy = (j + 0.5) * (nym / n_axis) + rng.uniform(-0.25, 0.25) * pitch
            wx, wy = rng.integers(win_range[0], win_range[1] + 1, 2)
            x0 = int(np.clip(round(sx - wx / 2), 0, nxm - wx))
            y0 = int(np.clip(round(sy - wy / 2), 0, nym - wy))
            nrec = int(rng.integers(rec_range[0], rec_range[1] + 1))
            rx = rng.uniform(x0 + 1, x0 + wx - 3, nrec) * d
            ry = rng.uniform(y0 + 1, y0 + wy - 3, nrec) * d
            rec = np.stack([rx, ry, np.zeros(nrec)], axis=1)
            src = np.array([[sx * d, sy * d, 10 * d]])
            shots.append((src, rec, (x0, y0, int(wx), int(wy))))
    return {"true": true, "start": start, "shots": shots, "d": d, "pad": 30, "nts": nts, "dts": 0.004, "sub": 8,
            "alpha_cos": 0.99, "wavelet": ricker(nts, 0.004, 10.0, 0.1)}


def source_traces(w: Workload) -> List[np.ndarray]:
    """9 irregular source traces [n_src, nts] for the workload's source type."""
    n = w.src_xyz.shape[0]
    tr = [np.zeros((n, w.nts)) for _ in range(9)]
    if w.source_type == "explosive":
        for c in (3, 4, 5):
            tr[c][:] = w.wavelet
    elif w.source_type == "fz":
        tr[2][:] = w.wavelet
    else:
        raise ValueError(w.source_type)
    return tr
