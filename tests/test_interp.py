"""Sinc and dipole interpolation of acquisition devices (host library,
ela_interp_weights; spaceInterpGpu_3D.cpp:214-480 with its sinc bounds check,
dipole y weight and sinc dipole sign corrected).

The sinc path of the reference is unreachable (its support check is inverted
and rejects every device), so no reference output pins it: these tests pin
the weights by what the mathematics fixes -- a device on a grid node gets a
single unit weight, the truncated sinc sums tend to one as the half-length
grows, a device at a cell centre gets a mirror-symmetric stencil, the linear
method reproduces linear fields so a dipole returns an exact difference, and
a dipole whose shift is a whole number of cells has weights summing to zero.
"""
import numpy as np
import pytest

from elasticisolib_3d_amd.acquisition import PaddedAxes, SpaceInterp

pytestmark = pytest.mark.filterwarnings("ignore::DeprecationWarning")

# 40 x 36 x 30 grid, fat 4, pads 5; dz != dx != dy so that a mixed-up spacing shows
AX = PaddedAxes(40, 36, 30, 0.0, 0.0, 0.0, 10.0, 12.0, 15.0, 4, 5, 5, 5, 5, 5)


def _model_point(fz, fx, fy):
    """A point at fractional grid coordinates (z, x, y) in model units."""
    return np.array([fz * AX.dz]), np.array([fx * AX.dx]), np.array([fy * AX.dy])


def _dense(si: SpaceInterp, k: int = 0):
    """Weights of device k as a dense [ny, nx, nz] array."""
    out = np.zeros(AX.ny * AX.nx * AX.nz)
    np.add.at(out, si.pos_unique[si.nbr[k]], si.weight[k])
    return out.reshape(AX.ny, AX.nx, AX.nz)


@pytest.mark.parametrize("h", [1, 2, 3, 4])
def test_sinc_device_on_a_node_has_a_single_unit_weight(h):
    z, x, y = _model_point(20, 17, 14)
    si = SpaceInterp(AX, z, x, y, method="sinc", half_length=h)
    d = _dense(si)
    assert si.weight.shape == (1, 8 * h ** 3)
    assert d[14, 17, 20] == pytest.approx(1.0, abs=1e-7)
    d[14, 17, 20] = 0.0
    assert np.abs(d).max() < 1e-6


def test_sinc_weight_sums_converge_to_one_like_one_over_h():
    """sum_k sinc(t - k) over all integers is 1; cut to k = -h+1 .. h it
    misses by two alternating tails of size ~ sin(pi t) / (2 pi h) each, so
    the error halves when h doubles. The 3-D sum is the product of the three
    per-axis sums and converges at the same rate."""
    big = PaddedAxes(80, 80, 80, 0.0, 0.0, 0.0, 10.0, 12.0, 15.0, 4, 5, 5, 5, 5, 5)
    err = []
    for h in (4, 8, 16):
        si = SpaceInterp(big, np.array([40.3 * 10.0]), np.array([40.6 * 12.0]), np.array([40.45 * 15.0]),
                         method="sinc", half_length=h)
        err.append(abs(si.weight.sum() - 1.0))
    assert 0.4 < err[1] / err[0] < 0.6 and 0.4 < err[2] / err[1] < 0.6
    assert err[2] < 0.06


def test_sinc_half_length_one_stencil_is_separable_and_ordered():
    """h = 1: the 8 points are the cell corners, y slowest then x then z, and
    the weights factor into per-axis sinc values."""
    fz, fx, fy = 20.25, 17.5, 14.75
    si = SpaceInterp(AX, *_model_point(fz, fx, fy), method="sinc", half_length=1)
    nz, nx = AX.nz, AX.nx
    idx = si.pos_unique[si.nbr[0]]
    base = 14 * nz * nx + 17 * nz + 20
    want = [base + iy * nz * nx + ix * nz + iz for iy in (0, 1) for ix in (0, 1) for iz in (0, 1)]
    assert list(idx) == want
    w = si.weight[0].reshape(2, 2, 2)  # [iy, ix, iz]
    # a separable 2x2x2 tensor has rank one in every unfolding
    for ax in range(3):
        m = np.moveaxis(w, ax, 0).reshape(2, 4)
        assert abs(np.linalg.det(m[:, :2])) < 1e-7 and abs(np.linalg.det(m[:, 2:])) < 1e-7
    # t = 0.25 along z: sinc(0.25) / sinc(-0.75) = sin(pi/4)/(pi/4) / (sin(3pi/4)/(3pi/4)) = 3
    assert w[0, 0, 0] / w[0, 0, 1] == pytest.approx(3.0, rel=1e-6)


def test_sinc_cell_centre_stencil_is_mirror_symmetric():
    si = SpaceInterp(AX, *_model_point(20.5, 17.5, 14.5), method="sinc", half_length=3)
    w = si.weight[0].reshape(6, 6, 6)
    for ax in range(3):
        assert np.allclose(w, np.flip(w, axis=ax), rtol=0, atol=1e-7)


def test_sinc_interpolates_a_smooth_field_better_than_linear():
    fz, fx, fy = 19.37, 16.81, 14.22
    zz, xx, yy = np.meshgrid(np.arange(AX.nz) * AX.dz, np.arange(AX.nx) * AX.dx, np.arange(AX.ny) * AX.dy,
                             indexing="ij")
    k = 2 * np.pi / (6.0 * np.array([AX.dz, AX.dx, AX.dy]))  # six points per wavelength
    f = np.cos(k[0] * zz + 0.3) * np.cos(k[1] * xx - 0.2) * np.cos(k[2] * yy + 0.1)
    f = np.transpose(f, (2, 1, 0))  # [ny, nx, nz]
    exact = np.cos(k[0] * fz * AX.dz + 0.3) * np.cos(k[1] * fx * AX.dx - 0.2) * np.cos(k[2] * fy * AX.dy + 0.1)
    errs = {}
    for method, h in (("linear", 1), ("sinc", 5)):
        si = SpaceInterp(AX, *_model_point(fz, fx, fy), method=method, half_length=h)
        val = si.forward(f.reshape(-1)[si.pos_unique][:, None])[0, 0]
        errs[method] = abs(val - exact)
    assert errs["sinc"] < 0.5 * errs["linear"]


def test_linear_dipole_returns_the_exact_difference_of_a_linear_field():
    """Linear interpolation reproduces linear fields, so a dipole at p with
    shift s reads f(p) - f(p + s); a y position divided by dz instead of dy
    (the reference's slip) would miss it."""
    a, b, c = 0.7, -1.3, 2.1
    z, x, y = _model_point(19.4, 16.3, 13.8)
    shift = (2.5 * AX.dz, -1.25 * AX.dx, 1.5 * AX.dy)
    si = SpaceInterp(AX, z, x, y, dipole=1, dipole_shift=shift)
    assert si.weight.shape == (1, 16)
    Z, X, Y = np.meshgrid(np.arange(AX.nz) * AX.dz, np.arange(AX.nx) * AX.dx, np.arange(AX.ny) * AX.dy,
                          indexing="ij")
    f = np.transpose(a * Z + b * X + c * Y, (2, 1, 0)).reshape(-1)
    val = si.forward(f[si.pos_unique][:, None])[0, 0]
    assert val == pytest.approx(-(a * shift[0] + b * shift[1] + c * shift[2]), rel=1e-5)
    assert si.weight[0, :8].sum() == pytest.approx(1.0, abs=1e-6)
    assert si.weight[0, 8:].sum() == pytest.approx(-1.0, abs=1e-6)


@pytest.mark.parametrize("h", [1, 2, 4])
def test_sinc_dipole_with_whole_cell_shift_sums_to_zero(h):
    z, x, y = _model_point(19.3, 16.6, 13.45)
    shift = (2 * AX.dz, -1 * AX.dx, 1 * AX.dy)
    si = SpaceInterp(AX, z, x, y, method="sinc", half_length=h, dipole=1, dipole_shift=shift)
    n1 = 8 * h ** 3
    assert si.weight[0, :n1].sum() == pytest.approx(-si.weight[0, n1:].sum(), abs=1e-6)
    assert abs(si.weight.sum()) < 1e-6
    # the second pole is the first one translated by the shift
    one = SpaceInterp(AX, z, x, y, method="sinc", half_length=h)
    two = SpaceInterp(AX, z + shift[0], x + shift[1], y + shift[2], method="sinc", half_length=h)
    assert np.allclose(_dense(si), _dense(one) - _dense(two), atol=1e-7)


def test_dipole_poles_sharing_points_are_summed():
    """Poles one cell apart share grid points: the unique list holds each
    point once and forward/adjoint are transposes."""
    z, x, y = _model_point(19.5, 16.5, 13.5)
    si = SpaceInterp(AX, np.r_[z, z + 3], np.r_[x, x], np.r_[y, y], dipole=1, dipole_shift=(AX.dz, 0.0, 0.0))
    assert len(si.pos_unique) == len(set(si.pos_unique.tolist()))
    rng = np.random.default_rng(0)
    reg = rng.standard_normal((si.n_reg, 3))
    irr = rng.standard_normal((si.n_dev, 3))
    assert np.sum(si.forward(reg) * irr) == pytest.approx(np.sum(reg * si.adjoint(irr)), rel=1e-12)


def test_sinc_support_must_stay_in_the_stepped_interior():
    # a device on the model's top row (z = 9.2 cells): h = 6 reaches z = 4 = FAT, h = 7 reaches z = 3
    SpaceInterp(AX, *_model_point(9.2, 16.0, 14.0), method="sinc", half_length=6)
    with pytest.raises(ValueError):
        SpaceInterp(AX, *_model_point(9.2, 16.0, 14.0), method="sinc", half_length=7)
    with pytest.raises(ValueError):  # device itself outside the model
        SpaceInterp(AX, *_model_point(2.0, 16.0, 14.0), method="sinc", half_length=1)
    with pytest.raises(ValueError):
        SpaceInterp(AX, *_model_point(20.0, 16.0, 14.0), method="linear", half_length=2)
    with pytest.raises(ValueError):
        SpaceInterp(AX, *_model_point(20.0, 16.0, 14.0), method="spline")
