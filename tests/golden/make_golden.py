"""Writes tests/golden/small_fp64.npz: FP64 oracle outputs of the four
operators and of the adjoint-state gradient on one small seeded problem.

The reference ships no golden vectors (and cannot run here), so these
fixtures are produced by the oracle, whose parts are pinned against
mathematics in tests/test_oracle.py; they freeze its results so that the GPU
parity tests compare against committed data and a later change of the oracle
shows up as a fixture mismatch. Inputs are stored exactly (medium and traces
as float32-representable values, the perturbation as a sparse list);
3-D outputs are stored on the FD interior (they are zero on the FAT layer).
Run from the repository root:

    python -m tests.golden.make_golden
"""
from __future__ import annotations

import os

import numpy as np

from oracle import oracle as O
from tests.problems import interior_points, random_medium

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "small_fp64.npz")
GRID = dict(nz=20, nx=18, ny=16, nts=6, sub=8, dz=10.0, dx=12.0, dy=11.0, dts=0.004, minPad=4, alphaCos=0.93)
SEED = 77
FAT = 4


def grid() -> O.Grid:
    return O.Grid(**GRID)


def interior(a):
    return np.ascontiguousarray(a[..., FAT:-FAT, FAT:-FAT, FAT:-FAT])


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def inputs(z):
    """Medium (scaled by the oracle), positions, traces and perturbation from the fixture."""
    g = grid()
    med = O.lib_scale_medium(z["rho"].astype(np.float64), z["lam"].astype(np.float64), z["mu"].astype(np.float64),
                             g.dtw)
    spos, rpos = _split(z["spos"], z["spos_n"]), _split(z["rpos"], z["rpos_n"])
    src = [x.reshape(-1, g.nts) for x in _split(z["src"].astype(np.float64), z["src_n"])]
    rec = [x.reshape(-1, g.nts) for x in _split(z["rec"].astype(np.float64), z["rec_n"])]
    dm = np.zeros((8,) + g.shape)
    dm.reshape(8, -1)[z["dm_k"], z["dm_i"]] = z["dm_v"].astype(np.float64)
    return g, med, spos, rpos, src, rec, dm


def _split(flat, n):
    out, o = [], 0
    for k in n:
        out.append(flat[o:o + k])
        o += k
    return out


def _flat(lists):
    return np.concatenate([np.ravel(x) for x in lists]), np.array([np.size(x) for x in lists], np.int64)


def make() -> dict:
    g = grid()
    rng = np.random.default_rng(SEED)
    rho, lam, mu = (np.asarray(x, np.float32) for x in random_medium(g, rng))
    spos = [interior_points(g, 3, rng) for _ in range(7)]
    rpos = [interior_points(g, 9, rng) for _ in range(7)]
    src = [np.asarray(rng.standard_normal((len(spos[O.COMP_GRID[c]]), g.nts)) * 1e3, np.float32) for c in range(9)]
    rec = [np.asarray(rng.standard_normal((len(rpos[O.COMP_GRID[c]]), g.nts)), np.float32) for c in range(9)]
    nI = (g.nz - 2 * FAT) * (g.nx - 2 * FAT) * (g.ny - 2 * FAT)
    k = rng.integers(0, 8, 60)
    cells = interior_points(g, 60, rng)
    v = np.asarray(rng.standard_normal(60) * 1e-3, np.float32)
    z = dict(rho=rho, lam=lam, mu=mu, dm_k=k, dm_i=cells, dm_v=v)
    for name, lists in (("spos", spos), ("rpos", rpos), ("src", src), ("rec", rec)):
        z[name], z[name + "_n"] = _flat(lists)
    g, med, spos, rpos, src, rec, dm = inputs(z)
    assert nI > 0
    d = O.lib_nl_forward(g, med, src, spos, rpos)
    a = O.lib_nl_adjoint(g, med, rec, spos, rpos)
    b = O.lib_born_forward(g, med, src, spos, dm, rpos)
    bi = np.asarray(O.lib_born_adjoint(g, med, src, spos, rec, rpos))
    g8, gs = O.lib_gradient(g, med, src, spos, rec, rpos)
    for name, lists in (("nl_fwd", d), ("nl_adj", a), ("born_fwd", b), ("grad_src", gs)):
        z[name], z[name + "_n"] = _flat(lists)
    for name, arr in (("born_adj", bi), ("grad8", np.asarray(g8))):
        full = np.asarray(arr)
        assert not np.any(full - np.pad(interior(full), ((0, 0),) + ((FAT, FAT),) * 3)), name
        z[name] = interior(full)
    return z


def expected(z):
    """Expected outputs as lists / full arrays."""
    g = grid()
    out = {}
    for name in ("nl_fwd", "nl_adj", "born_fwd", "grad_src"):
        out[name] = [x.reshape(-1, g.nts) for x in _split(z[name], z[name + "_n"])]
    for name in ("born_adj", "grad8"):
        out[name] = np.pad(z[name], ((0, 0),) + ((FAT, FAT),) * 3)
    return out


def load():
    return dict(np.load(PATH, allow_pickle=False))


if __name__ == "__main__":
    np.savez_compressed(PATH, **make())
    print(PATH, os.path.getsize(PATH), "bytes")
