"""CPU tests of the product's host side: the C-ABI library loads and exports
every symbol of include/elastic_iso_3d.h, and its host-only routines (medium
staggering/scaling, trilinear interpolation) agree with the oracle. No GPU
compute call is made here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    from elasticisolib_3d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from elasticisolib_3d_amd import build
        build.build()
    return _lib.load()


def header_functions():
    text = open(os.path.join(ROOT, "include", "elastic_iso_3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ela_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol(lib):
    from elasticisolib_3d_amd import _lib
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert lib.ela_version().decode().startswith("elastic_iso_3d")


def test_scale_medium_matches_oracle(lib):
    from elasticisolib_3d_amd.propagator import scale_medium
    rng = np.random.default_rng(0)
    shape = (9, 11, 13)
    rho = 1800 + 400 * rng.random(shape)
    lam = 1e9 * (3 + rng.random(shape))
    mu = 1e9 * (2 + rng.random(shape))
    a = scale_medium(rho, lam, mu, 4e-4)
    b = O.lib_scale_medium(rho, lam, mu, 4e-4)
    assert np.max(np.abs(a - b) / np.abs(b)) < 1e-14


def test_interp_linear_matches_oracle(lib):
    from elasticisolib_3d_amd.acquisition import PaddedAxes, ShotGeometry, SpaceInterp
    rng = np.random.default_rng(1)
    d = 12.5
    ax = PaddedAxes(nz=40, nx=44, ny=38, oz=-(4 + 5) * d, ox=-(4 + 6) * d, oy=-(4 + 5) * d, dz=d, dx=d, dy=d,
                    fat=4, z_pad_minus=5, z_pad_plus=7, x_pad_minus=6, x_pad_plus=6, y_pad=5)
    n = 40
    z = rng.uniform(0, 18 * d, n)     # devices inside the unpadded model
    x = rng.uniform(0, 22 * d, n)
    y = rng.uniform(0, 18 * d, n)
    z[:5] = np.arange(5) * d          # on grid nodes
    for g in range(7):
        gax = ax.grid(g)
        a = SpaceInterp(gax, z, x, y)
        b = O.linear_interp(z, x, y, gax.nz, gax.nx, gax.ny, gax.oz, gax.ox, gax.oy, d, d, d, 4,
                            gax.z_pad_minus, gax.z_pad_plus, gax.x_pad_minus, gax.x_pad_plus, gax.y_pad)
        assert np.array_equal(a.pos_unique, b.pos_unique)
        assert np.array_equal(a.nbr, b.idx)
        assert np.allclose(a.weight, b.weight, rtol=0, atol=1e-7)
    # trace mapping: forward and adjoint are transposes
    geo = ShotGeometry(ax, z, x, y)
    reg = [rng.standard_normal((geo.grids[gi].n_reg, 5)) for gi in (1, 2, 3, 0, 0, 0, 4, 5, 6)]
    irr = [rng.standard_normal((n, 5)) for _ in range(9)]
    lhs = sum(np.sum(a * b) for a, b in zip(geo.to_irregular(reg), irr))
    rhs = sum(np.sum(a * b) for a, b in zip(reg, geo.to_regular(irr)))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_interp_rejects_out_of_bounds(lib):
    from elasticisolib_3d_amd.acquisition import PaddedAxes, SpaceInterp
    ax = PaddedAxes(nz=30, nx=30, ny=30, oz=-90.0, ox=-90.0, oy=-90.0, dz=10, dx=10, dy=10, fat=4,
                    z_pad_minus=5, z_pad_plus=5, x_pad_minus=5, x_pad_plus=5, y_pad=5)
    with pytest.raises(ValueError):
        SpaceInterp(ax, [-50.0], [10.0], [10.0])


def test_padding_rule_matches_reference_sizes():
    """padElasticFileGpuMain_3D.cpp:47-86: zPadPlus/xPadPlus complete the block multiple."""
    from elasticisolib_3d_amd.models import pad_sizes
    assert pad_sizes(250, 30) == (30, 40)
    assert pad_sizes(400, 30) == (30, 34)
    assert pad_sizes(100, 20) == (20, 24)
    for n, p in ((250, 30), (400, 30), (100, 20), (160, 30), (256, 30)):
        lo, hi = pad_sizes(n, p)
        assert (lo + n + hi) % 16 == 0 and hi >= p


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from elasticisolib_3d_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(_lib.ElaError):
        _lib.load()


@pytest.mark.parametrize("bad,match", [
    (dict(nz=8), "too small"), (dict(ny=7), "too small"), (dict(nts=1), "nts must be"), (dict(sub=0), "sub must be"),
    (dict(dx=0.0), "non-positive spacing"), (dict(dts=-1e-3), "non-positive spacing"),
    (dict(min_pad=-1), "negative min_pad"), (dict(min_pad=11), "larger than half the grid")])
def test_create_rejects_bad_configurations(bad, match):
    """ela_create validates its configuration before any HIP call, so these
    run without a GPU: each bad field is ELA_ERR_ARG with its message."""
    from elasticisolib_3d_amd import _lib
    from elasticisolib_3d_amd.propagator import Propagator
    cfg = dict(nz=21, nx=23, ny=25, nts=5, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, min_pad=5)
    cfg.update(bad)
    with pytest.raises(_lib.ElaError, match=match):
        Propagator(cfg["nz"], cfg["nx"], cfg["ny"], cfg["nts"], cfg["sub"], cfg["dz"], cfg["dx"], cfg["dy"],
                   cfg["dts"], cfg["min_pad"], 0.95, "f32")


def _rand_model(shape, rng):
    rho = 1800 + 400 * rng.random(shape)
    lam = 1e9 * (3 + rng.random(shape))
    mu = 1e9 * (2 + rng.random(shape))
    return rho, lam, mu


def test_born_stagger_matches_oracle(lib):
    from elasticisolib_3d_amd.propagator import scale_medium, born_stagger, born_unstagger
    rng = np.random.default_rng(3)
    shape = (7, 9, 11)
    rho, lam, mu = _rand_model(shape, rng)
    dtw = 5e-4
    med = scale_medium(rho, lam, mu, dtw)
    dm3 = rng.standard_normal((3,) + shape)
    a = born_stagger(dm3, med, dtw)
    b = O.born_stagger_perturbation(dm3, med, dtw, O.lib_stagger)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    img8 = rng.standard_normal((8,) + shape)
    a = born_unstagger(img8, med, dtw)
    b = O.born_unstagger_image(img8, med, dtw, O.lib_stagger)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # and they are transposes of each other
    lhs = np.sum(born_stagger(dm3, med, dtw) * img8)
    rhs = np.sum(dm3 * born_unstagger(img8, med, dtw))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_medium_gradient_is_transposed_jacobian(lib):
    """<J dm, g8> = <dm, medium_gradient(g8)> with J dm from a central
    difference of scale_medium (quadratic error ~ h^2, h = 1e-5 relative)."""
    from elasticisolib_3d_amd.propagator import scale_medium, medium_gradient
    rng = np.random.default_rng(4)
    shape = (6, 8, 7)
    rho, lam, mu = _rand_model(shape, rng)
    dtw = 5e-4
    med = scale_medium(rho, lam, mu, dtw)
    g8 = rng.standard_normal((8,) + shape)
    g3 = medium_gradient(g8, med, dtw)
    for k, base in enumerate((rho, lam, mu)):
        d = rng.standard_normal(shape) * base
        h = 1e-5
        args_p = [rho, lam, mu]
        args_m = [rho, lam, mu]
        args_p[k] = base + h * d
        args_m[k] = base - h * d
        jd = (scale_medium(*args_p, dtw) - scale_medium(*args_m, dtw)) / (2 * h)
        lhs = np.sum(jd * g8)
        rhs = np.sum(d * g3[k])
        assert abs(lhs - rhs) <= 1e-7 * abs(rhs), (k, lhs, rhs)


def test_autotuner_picks_fastest_and_caches(tmp_path, monkeypatch):
    """The autotuner applies the candidate with the smallest measured step
    time and reuses it from its cache for the same grid."""
    from elasticisolib_3d_amd import autotune as A
    monkeypatch.setenv("ELA_TUNE_CACHE", str(tmp_path / "t.json"))

    class FakeProp:
        precision = "f32"
        shape = (40, 50, 60)

        def __init__(self):
            self.calls = 0
            self.tiling, self.order = (16, 8), (0, 1)

        def set_tiling(self, tx, yc):
            self.tiling = (tx, yc)

        def set_tile_order(self, o, g):
            self.order = (o, g)

        def time_step_kernels(self, reps):
            self.calls += 1
            tx, yc = self.tiling
            o, g = self.order
            t = 1.0 + 0.1 * abs(yc - 16) + 0.05 * (16 - tx) + (0.0 if (o, g) == (3, 2) else 0.02)
            return t / 2, t / 2

    p = FakeProp()
    best = A.autotune(p)
    assert (best["tile_x"], best["y_chunk"], best["order"], best["x_group"]) == (16, 16, 3, 2)
    assert p.tiling == (16, 16) and p.order == (3, 2) and not best["cached"]
    q = FakeProp()
    again = A.autotune(q)
    assert again["cached"] and q.calls == 0 and q.tiling == (16, 16) and q.order == (3, 2)


def test_autotuner_gradient_target(tmp_path, monkeypatch):
    """With target="gradient" the autotuner ranks shapes by the sum of the
    gradient step's three kernel pairs, which can prefer another shape than
    the forward pair: the forward and recomputing pairs take the x tile best
    for them, the transposed pair its own (ela_set_tiling_transposed), all
    from the same measurements, and the choice is cached under its own key."""
    from elasticisolib_3d_amd import autotune as A
    monkeypatch.setenv("ELA_TUNE_CACHE", str(tmp_path / "t.json"))

    class FakeProp:
        precision = "f32"
        shape = (40, 50, 60)

        def __init__(self):
            self.tiling, self.order = (16, 8), (0, 1)
            self.grad_calls = 0
            self.tile_t = 0

        def set_tiling_transposed(self, tx):
            self.tile_t = tx

        def set_tiling(self, tx, yc):
            self.tiling = (tx, yc)

        def set_tile_order(self, o, g):
            self.order = (o, g)

        def gradient_derivatives_used(self):
            return "native"

        def time_step_kernels(self, reps):   # the forward pair prefers tile 8
            tx, yc = self.tiling
            return 0.5 + 0.01 * abs(tx - 8) + 0.01 * abs(yc - 8), 0.5

        def time_gradient_kernels(self, reps):   # the transposed pair prefers tile 4 more strongly
            self.grad_calls += 1
            tx, yc = self.tiling
            fw = sum(self.time_step_kernels(reps))
            return fw, 1.2 * fw, 2.0 + 0.05 * abs(tx - 4) + 0.01 * abs(yc - 8)

    p = FakeProp()
    fwd = A.autotune(p, candidates=[(4, 8, 3, 2), (8, 8, 3, 1), (16, 8, 3, 1)])
    assert (fwd["tile_x"], fwd["target"]) == (8, "forward") and p.tile_t == 0
    grad = A.autotune(p, target="gradient")
    assert (grad["tile_x"], grad["tile_x_transposed"], grad["y_chunk"]) == (8, 4, 8) and not grad["cached"]
    assert grad["ms_step"] == grad["ms_forward"] + grad["ms_recompute"] + grad["ms_transposed"]
    assert abs(grad["ms_step"] - (1.0 + 1.2 + 2.0)) < 1e-12    # split beats every common tile (4.288 at tile 4)
    assert p.tiling == (8, 8) and p.tile_t == 4
    q = FakeProp()
    again = A.autotune(q, target="gradient")
    assert again["cached"] and q.grad_calls == 0 and q.tiling == (8, 8) and q.tile_t == 4
    with pytest.raises(ValueError):
        A.autotune(q, target="adjoint")


def test_positions_follow_the_memory_kind_of_the_call():
    """Position lists are handed to the library in the memory kind of the
    call (host here: torch CPU tensors become numpy arrays), so that a
    device-memory call never receives host position pointers."""
    import torch
    from elasticisolib_3d_amd import _lib
    from elasticisolib_3d_amd.propagator import Propagator
    prop = Propagator.__new__(Propagator)
    prop._keep = []
    pos = [torch.arange(3, dtype=torch.int64), np.arange(2), [], [5], np.zeros(0, np.int64), [1, 2], [7]]
    p = prop._points(pos, _lib.ELA_MEM_HOST)
    assert [p.n[g] for g in range(7)] == [3, 2, 0, 1, 0, 2, 1]
    assert all(isinstance(a, np.ndarray) and a.dtype == np.int64 for a in prop._keep)
    first = np.ctypeslib.as_array(ctypes.cast(p.pos[0], ctypes.POINTER(ctypes.c_int64)), shape=(3,))
    assert list(first) == [0, 1, 2]


def _kernel_constant(name):
    src = open(os.path.join(os.path.dirname(__file__), "..", "elasticisolib_3d_amd", "csrc", "ela_kernels.hpp")).read()
    m = re.search(r"constexpr int %s = ([^;]+);" % name, src)
    expr = m.group(1).split("//")[0].strip()
    return eval(expr, {"TZ": 64})   # the expressions are integer literals or multiples of TZ


@pytest.mark.parametrize("esize", [4, 8])
def test_stencil_loads_stay_inside_the_guard(esize):
    """The stencils load their neighbours without bounds predicates; the
    farthest element any load reaches (last plane, last x tile plus its high x
    halo, last z tile plus its high z halo) must lie inside the array's
    stride, which carries GUARD_ROWS rows + GUARD_EXTRA elements past the
    grid (the stride rule of ela_create). Grids with partial tiles in x and z,
    every tile width, FP32 and FP64 row pitches."""
    FAT, TZ = 4, 64
    guard_rows, guard_extra = _kernel_constant("GUARD_ROWS"), _kernel_constant("GUARD_EXTRA")
    assert guard_rows >= 16   # the widest tile
    line = 128 // esize
    rup = lambda v, m: (v + m - 1) // m * m   # noqa: E731
    rng = np.random.default_rng(0)
    for nz, nx, ny in [(29, 31, 45), (81, 21, 19), (328, 472, 468), (232, 328, 324)] + \
            [tuple(rng.integers(9, 400, 3)) for _ in range(40)]:
        zoff = line - FAT
        nzi = rup(zoff + nz, line)
        n_int = nzi * nx * ny
        stride = rup(n_int + guard_rows * nzi + guard_extra, 64)
        for bx in (4, 8, 16):
            nbz = (nz - 2 * FAT + TZ - 1) // TZ
            nbx = (nx - 2 * FAT + bx - 1) // bx
            z0 = zoff + FAT + (nbz - 1) * TZ        # last z tile
            x0 = FAT + (nbx - 1) * bx               # last x tile
            iz_max = max(z0 + TZ - 1, z0 + TZ + FAT - 1)          # own column / high z halo
            ix_max = max(x0 + bx - 1, x0 + bx + FAT - 1)          # own column / high x halo
            last_plane = ny - 1                     # queue loads reach y1 + 3 <= ny - 1
            far = last_plane * nx * nzi + ix_max * nzi + iz_max
            assert far < stride, (nz, nx, ny, bx, esize, far - stride)


def test_host_staggering_threads_bit_identical(lib, monkeypatch):
    """The host staggering routines split their loops over threads above 64 Ki
    cells: on a 120 k-cell grid every output is bit-identical with one thread
    and with eight, and the medium matches the oracle."""
    from elasticisolib_3d_amd.propagator import born_stagger, born_unstagger, medium_gradient, scale_medium
    rng = np.random.default_rng(31)
    shape = (40, 50, 60)   # ny, nx, nz
    rho = 1000 + 2000 * rng.random(shape)
    lam = 1e9 * (1 + rng.random(shape))
    mu = 1e9 * (1 + rng.random(shape))
    dtw = 5e-4
    dm3 = rng.standard_normal((3,) + shape)
    a8 = rng.standard_normal((8,) + shape)
    out = {}
    for nt in ("1", "8"):
        monkeypatch.setenv("OMP_NUM_THREADS", nt)
        med = scale_medium(rho, lam, mu, dtw)
        out[nt] = (med, born_stagger(dm3, med, dtw), born_unstagger(a8, med, dtw), medium_gradient(a8, med, dtw))
    for x, y in zip(out["1"], out["8"]):
        assert np.array_equal(x, y)
    ref = O.lib_scale_medium(rho, lam, mu, dtw)
    assert np.allclose(out["8"][0], ref, rtol=1e-14, atol=0)
