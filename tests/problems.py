"""Seeded random problems shared by the parity tests."""
from __future__ import annotations

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
