"""Full-size checks on the padded grid of the 400x400x250 workload
(328 x 472 x 468 cells, 72.5 M per array), where the oracle would take hours:
properties that hold at any size. A homogeneous medium with an explosive
source gives normal stresses that are mirror-symmetric about the source along
every axis and invariant under permuting the axes; FP32 stays within the FP32
tolerance of FP64; the operator is linear call after call (no state carried
from one propagation to the next). These runs cover the launch at its real
shape: every block of the grid, 32-bit offsets over 72.5 M cells, the tiles
and y chunks of the full grid."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NX, NY = 328, 472, 468
D = 12.5
NTS, SUB, DTS = 21, 8, 0.004          # 160 fine steps of 0.5 ms (2 vp dtw / d = 0.2)
RHO, VP, VS = 2000.0, 2500.0, 1440.0
CZ, CX, CY = 164, 236, 234            # source cell (centre grid), > 150 cells from every edge
KS = (4, 8, 12)                       # receiver offsets along each axis (the wave travels 16 cells)


def _flat(iz, ix, iy):
    return (iy * NX + ix) * NZ + iz


def _wavelet(t0, f0):
    t = np.arange(NTS) * DTS
    a2 = (math.pi * f0) ** 2
    return (1 - 2 * a2 * (t - t0) ** 2) * np.exp(-a2 * (t - t0) ** 2)


def _receivers():
    pts = []
    for k in KS:
        pts += [(CZ + k, CX, CY), (CZ - k, CX, CY), (CZ, CX + k, CY), (CZ, CX - k, CY), (CZ, CX, CY + k),
                (CZ, CX, CY - k)]
    return [np.array([_flat(*p) for p in pts], np.int64)] + [np.zeros(0, np.int64)] * 6


def _explosive(cells, wavelets):
    """Positions and traces of explosive sources (mxx = myy = mzz) at centre-grid cells."""
    from elasticisolib_3d_amd.acquisition import COMP_GRID
    spos = [np.array([_flat(*c) for c in cells], np.int64)] + [np.zeros(0, np.int64)] * 6
    src = [np.zeros((len(spos[COMP_GRID[k]]), NTS)) for k in range(9)]
    for comp in (3, 4, 5):
        src[comp][:] = np.stack(wavelets)
    return spos, src


def _propagator(precision):
    import torch
    from elasticisolib_3d_amd.propagator import Propagator
    prop = Propagator(NZ, NX, NY, NTS, SUB, D, D, D, DTS, 30, 0.95, precision, device=0)
    td = torch.float32 if precision == "f32" else torch.float64
    rlm = torch.empty((3, NY, NX, NZ), dtype=td, device="cuda")
    rlm[0].fill_(RHO)
    rlm[1].fill_(RHO * VP ** 2 - 2 * RHO * VS ** 2)
    rlm[2].fill_(RHO * VS ** 2)
    prop.set_elastic(rlm, DTS / SUB)
    del rlm
    return prop


def _check_symmetry(rec, tol_rel):
    sxx, syy, szz = rec[3], rec[4], rec[5]
    peak = max(float(np.abs(x).max()) for x in (sxx, syy, szz))
    assert peak > 0
    tol = tol_rel * peak
    worst = 0.0
    for j, k in enumerate(KS):
        zp, zm, xp, xm, yp, ym = range(6 * j, 6 * j + 6)
        pairs = [(f[a], f[b]) for f in (sxx, syy, szz) for a, b in ((zp, zm), (xp, xm), (yp, ym))]
        pairs += [(sxx[xp], syy[yp]), (sxx[xp], szz[zp])]
        pairs += [(a, syy[xp]) for a in (szz[xp], sxx[yp], szz[yp], sxx[zp], syy[zp])]
        for a, b in pairs:
            worst = max(worst, float(np.abs(a - b).max()))
        assert float(np.abs(sxx[xp] - syy[xp]).max()) > 1e3 * tol, k   # not vacuous
    assert worst <= tol, (worst / peak, tol_rel)
    return worst / peak


@pytest.fixture(scope="module")
def explosive_runs():
    """Data of the centred explosive source in FP64 and FP32 (one propagator
    at a time, freed before the next)."""
    import torch
    spos, src = _explosive([(CZ, CX, CY)], [_wavelet(0.03, 30.0)])
    rpos = _receivers()
    out = {}
    for precision in ("f64", "f32"):
        print(f"full size {NZ}x{NX}x{NY}: {precision} explosive source, {NTS} samples", flush=True)
        prop = _propagator(precision)
        out[precision] = [np.asarray(x, np.float64) for x in prop.nl_forward(spos, src, rpos)]
        del prop
        torch.cuda.empty_cache()
    return out


@pytest.mark.parametrize("precision,tol", [("f64", 1e-11), ("f32", 5e-6)])
def test_full_size_explosive_symmetry(explosive_runs, precision, tol):
    worst = _check_symmetry(explosive_runs[precision], tol)
    print(f"full size {precision}: worst symmetry defect {worst:.2e} of the peak", flush=True)


def test_full_size_fp32_matches_fp64(explosive_runs):
    a, b = explosive_runs["f32"], explosive_runs["f64"]
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y ** 2)) for y in b)
    e = math.sqrt(num / den)
    print(f"full size FP32 vs FP64: rel L2 {e:.2e}", flush=True)
    assert e < 5e-6


def test_full_size_superposition_fp64():
    """F(s1 + s2) = F(s1) + F(s2) in FP64 with two explosive sources at
    different cells and with different wavelets, propagated one after the
    other on the same handle."""
    import torch
    rpos = _receivers()
    c1, c2 = (CZ, CX, CY), (CZ + 5, CX - 7, CY + 3)
    w1, w2 = _wavelet(0.03, 30.0), -0.7 * _wavelet(0.035, 22.0)
    print("full size f64: superposition of two sources", flush=True)
    prop = _propagator("f64")
    d1 = prop.nl_forward(*_explosive([c1], [w1])[:2], rpos)
    d2 = prop.nl_forward(*_explosive([c2], [w2])[:2], rpos)
    d12 = prop.nl_forward(*_explosive([c1, c2], [w1, w2])[:2], rpos)
    del prop
    torch.cuda.empty_cache()
    num = sum(float(np.sum((z - x - y) ** 2)) for x, y, z in zip(d1, d2, d12))
    den = sum(float(np.sum(z ** 2)) for z in d12)
    e = math.sqrt(num / den)
    print(f"full size superposition: rel L2 {e:.2e}", flush=True)
    assert den > 0 and e < 1e-12


def _random_devices(rng, per_grid, spread):
    """`per_grid` distinct random interior cells on each of the 7 grids,
    within `spread` cells of the source cell."""
    pos = []
    for _ in range(7):
        cells = set()
        while len(cells) < per_grid:
            dz, dx, dy = rng.integers(-spread, spread + 1, 3)
            cells.add(_flat(CZ + int(dz), CX + int(dx), CY + int(dy)))
        pos.append(np.array(sorted(cells), np.int64))
    return pos


def _random_traces(rng, pos):
    from elasticisolib_3d_amd.acquisition import COMP_GRID
    return [rng.standard_normal((len(pos[COMP_GRID[c]]), NTS)) for c in range(9)]


class _Devices:
    """Positions of the 7 grids, as the trace scaling reads them."""

    def __init__(self, pos):
        self._pos = pos

    def positions(self):
        return self._pos


def test_full_size_nonlinear_dot_test_fp64():
    """<F s, r> = <s, F* r> for the reference's nonlinear operator pair on
    random sources and receivers on all 7 grids, FP64, full grid: the forward
    injects medium-scaled source traces, the adjoint injects medium-scaled
    receiver traces (nonlinearPropElasticGpu_3D.cpp:200-277, :481-558)."""
    import torch
    from elasticisolib_3d_amd.acquisition import scale_regular_traces
    from elasticisolib_3d_amd.propagator import scale_medium
    rng = np.random.default_rng(11)
    spos, rpos = _random_devices(rng, 2, 12), _random_devices(rng, 5, 14)
    s, r = _random_traces(rng, spos), _random_traces(rng, rpos)
    shape = (NY, NX, NZ)
    med8 = scale_medium(np.full(shape, RHO), np.full(shape, RHO * VP ** 2 - 2 * RHO * VS ** 2),
                        np.full(shape, RHO * VS ** 2), DTS / SUB)
    s_scaled = scale_regular_traces(med8, _Devices(spos), s, D ** 3)
    r_scaled = scale_regular_traces(med8, _Devices(rpos), r, D ** 3)
    del med8
    print("full size f64: nonlinear dot test", flush=True)
    prop = _propagator("f64")
    d = prop.nl_forward(spos, s_scaled, rpos)
    a = prop.nl_adjoint(spos, r_scaled, rpos)
    del prop
    torch.cuda.empty_cache()
    lhs = sum(float(np.sum(x * y)) for x, y in zip(d, r))
    rhs = sum(float(np.sum(x * y)) for x, y in zip(s, a))
    e = abs(lhs - rhs) / abs(lhs)
    print(f"full size nonlinear dot test: {lhs:.15e} {rhs:.15e} rel {e:.2e}", flush=True)
    assert e < 1e-12


def test_full_size_born_dot_test_fp64():
    """<B dm, r> = <dm, B* r> for the Born operator with a dense random
    perturbation of all 8 staggered arrays (zero in the FAT layers), FP64,
    full grid (every coarse background slice kept in HBM)."""
    import torch
    rng = np.random.default_rng(12)
    spos, rpos = _random_devices(rng, 2, 12), _random_devices(rng, 5, 14)
    s = [torch.tensor(x, device="cuda") for x in _random_traces(rng, spos)]
    r = [torch.tensor(x, device="cuda") for x in _random_traces(rng, rpos)]
    g = torch.Generator(device="cuda").manual_seed(13)
    dm = torch.randn((8, NY, NX, NZ), dtype=torch.float64, device="cuda", generator=g)
    for sl in ((slice(None), slice(0, 4)), (slice(None), slice(-4, None))):   # FAT layers stay 0
        dm[sl] = 0
        dm[:, :, sl[1]] = 0
        dm[:, :, :, sl[1]] = 0
    print("full size f64: Born dot test", flush=True)
    prop = _propagator("f64")
    b = prop.born_forward(spos, s, dm, rpos)
    img = prop.born_adjoint(spos, s, r, rpos)
    torch.cuda.synchronize()
    lhs = sum(float(torch.sum(x * y)) for x, y in zip(b, r))
    rhs = float(torch.dot(dm.reshape(-1), img.reshape(-1)))
    del prop, dm, img
    torch.cuda.empty_cache()
    e = abs(lhs - rhs) / abs(lhs)
    print(f"full size Born dot test: {lhs:.15e} {rhs:.15e} rel {e:.2e}", flush=True)
    assert e < 1e-11
