"""Seeded random problems shared by the parity tests."""
from __future__ import annotations

import math

import numpy as np

from oracle import oracle as O

FAT = 4


def interior_points(g, n, rng, margin=0):
    """n unique random interior flat indices (at least `margin` cells inside the FAT layer)."""
    lo = FAT + margin
    iz = rng.integers(lo, g.nz - lo, 4 * n)
    ix = rng.integers(lo, g.nx - lo, 4 * n)
    iy = rng.integers(lo, g.ny - lo, 4 * n)
    flat = iy * g.nz * g.nx + ix * g.nz + iz
    _, first = np.unique(flat, return_index=True)
    return flat[np.sort(first)][:n].astype(np.int64)


def random_medium(g, rng, contrast=0.2):
    shape = g.shape
    rho = 2000.0 * (1 + contrast * rng.random(shape))
    vp = 2500.0 * (1 + contrast * rng.random(shape))
    vs = vp / np.sqrt(3.0) * (0.9 + 0.1 * rng.random(shape))
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    return rho, lam, mu


def random_problem(g, seed, nsrc=3, nrec=7, margin=0):
    rng = np.random.default_rng(seed)
    rho, lam, mu = random_medium(g, rng)
    med = O.lib_scale_medium(rho, lam, mu, g.dtw)
    spos = [interior_points(g, nsrc, rng, margin) for _ in range(7)]
    rpos = [interior_points(g, nrec, rng, margin) for _ in range(7)]
    src = [rng.standard_normal((len(spos[O.COMP_GRID[c]]), g.nts)) * 1e3 for c in range(9)]
    rec = [rng.standard_normal((len(rpos[O.COMP_GRID[c]]), g.nts)) for c in range(9)]
    dm = rng.standard_normal((8,) + g.shape) * 1e-3
    for k in range(8):
        dm[k][:FAT] = 0; dm[k][-FAT:] = 0
        dm[k][:, :FAT] = 0; dm[k][:, -FAT:] = 0
        dm[k][:, :, :FAT] = 0; dm[k][:, :, -FAT:] = 0
    return dict(rho=rho, lam=lam, mu=mu, med=med, spos=spos, rpos=rpos, src=src, rec=rec, dm=dm)


def rel(a, b):
    a = np.concatenate([np.ravel(x) for x in a]) if isinstance(a, (list, tuple)) else np.ravel(a)
    b = np.concatenate([np.ravel(x) for x in b]) if isinstance(b, (list, tuple)) else np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def born_blob_problem(param, blob_y=20):
    """Homogeneous 40^3 grid, explosive Ricker source at the centre, receivers
    on a line, and a smooth Gaussian perturbation blob of parameter `param`
    (0 density, 1 lambda, 2 mu) between them (e^-4.5 of its peak at the
    source), centred on plane y = blob_y (the source's plane by default).
    Returns the grid, the medium, the perturbation [3, ny, nx, nz], the
    positions and fwd(scale): the oracle data, scaled medium and
    medium-scaled source traces of m + scale * perturbation."""
    g = O.Grid(nz=40, nx=40, ny=40, nts=60, sub=8, dz=10., dx=10., dy=10., dts=0.004, minPad=6, alphaCos=0.99)
    shape = g.shape
    rho = np.full(shape, 2000.0)
    vp, vs = 2500.0, 1440.0
    mu = rho * vs ** 2
    lam = rho * vp ** 2 - 2 * mu
    # explosive source at the centre, receivers on a line
    nz, nx = g.nz, g.nx
    c = 20
    flat = lambda iz, ix, iy: iy * nz * nx + ix * nz + iz  # noqa: E731
    spos = [np.array([flat(c, c, c)])] + [np.zeros(0, np.int64)] * 6
    rpos = [np.array([flat(12, x, 20) for x in range(10, 31, 2)])] * 7
    t = np.arange(g.nts) * g.dts - 0.09
    a2 = (math.pi * 12.0) ** 2
    w = (1 - 2 * a2 * t * t) * np.exp(-a2 * t * t)
    src = [np.zeros((len(spos[O.COMP_GRID[k]]), g.nts)) for k in range(9)]
    # a smooth perturbation blob between source and receivers, zero at the source
    iy, ix, iz = np.meshgrid(np.arange(g.ny), np.arange(g.nx), np.arange(g.nz), indexing="ij")
    blob = np.exp(-((iz - 14) ** 2 + (ix - 20) ** 2 + (iy - blob_y) ** 2) / 8.0)
    d3 = np.zeros((3,) + shape)
    d3[param] = blob * (rho, lam, mu)[param] * 1.0

    def fwd(scale):
        r = rho + scale * d3[0]
        lm = lam + scale * d3[1]
        m = mu + scale * d3[2]
        med = O.lib_scale_medium(r, lm, m, g.dtw)
        m8 = med.reshape(8, -1)
        s = [np.zeros_like(x) for x in src]
        pc = spos[0][0]
        for k in (3, 4, 5):   # explosive: mxx = myy = mzz
            s[k][0] = w * (m8[3][pc] + 2 * m8[4][pc]) / (g.dx * g.dy * g.dz)
        return O.lib_nl_forward(g, med, s, spos, rpos), med, s

    return g, rho, lam, mu, d3, spos, rpos, fwd
