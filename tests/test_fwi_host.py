"""Host-side pieces of the FWI chain (no GPU): component sampling against the
ElasticDatComp_3D operator, the (Vp, Vs, rho) conversion's Jacobian
transpose against central differences and ElasticConvJab_3D, and the
bounded L-BFGS outer loop on a quadratic misfit with active bounds."""
import numpy as np
import pytest

from elasticisolib_3d_amd import fwi as F
from elasticisolib_3d_amd import operators as OP
from elasticisolib_3d_amd.sep import Axis, Hypercube, Vector


def test_component_sampling_matches_the_operator():
    rng = np.random.default_rng(1)
    hyp = Hypercube([Axis(6), Axis(4), Axis(9), Axis(1)])
    model = Vector(hyp, rng.standard_normal(hyp.shape))
    comps = "p,vz,sxy,vx"
    op = OP.ElasticDatComp_3D(comps, model)
    data = op.getRange().clone()
    op.forward(False, model, data)
    d9 = [model.getNdArray()[0, k] for k in range(9)]
    rows = F.component_rows(comps)
    ds = F.sample_components(rows, d9)
    assert all(np.allclose(a, b, rtol=0, atol=1e-15) for a, b in zip(ds, data.getNdArray()[0]))
    # adjoint: <C x, y> = <x, C^T y>
    ys = [rng.standard_normal(d9[0].shape) for _ in rows]
    back = F.sample_components_adjoint(rows, ys, lambda: np.zeros_like(d9[0]))
    lhs = sum(np.sum(a * b) for a, b in zip(ds, ys))
    rhs = sum(np.sum(a * b) for a, b in zip(d9, back))
    assert abs(lhs - rhs) <= 1e-13 * abs(rhs)
    # the operator's adjoint gives the same 9 arrays
    dv = op.getRange().clone()
    dv.getNdArray()[0] = np.stack(ys)
    mv = model.clone()
    op.adjoint(False, mv, dv)
    assert all(np.allclose(a, b, rtol=0, atol=1e-15) for a, b in zip(back, mv.getNdArray()[0]))
    with pytest.raises(ValueError):
        F.component_rows("vx,vx")
    with pytest.raises(ValueError):
        F.component_rows("vx,q")


def test_vp_vs_rho_jacobian_transpose():
    rng = np.random.default_rng(2)
    shape = (3, 4, 5)
    vp = 2500 + 300 * rng.random(shape)
    vs = 1400 + 200 * rng.random(shape)
    rho = 2000 + 200 * rng.random(shape)
    w = [rng.standard_normal(shape) for _ in range(3)]   # dJ/d(rho, lambda, mu) of J = <w, g(m)>
    g = F.vp_vs_rho_gradient(vp, vs, rho, *w)
    for k in range(3):
        d = [np.zeros(shape) for _ in range(3)]
        d[k] = rng.standard_normal(shape) * (vp, vs, rho)[k]
        h = 1e-6

        def J(s):
            m = [a + s * b for a, b in zip((vp, vs, rho), d)]
            return sum(np.sum(a * b) for a, b in zip(w, F.vp_vs_rho_to_elastic(*m)))
        fd = (J(h) - J(-h)) / (2 * h)
        an = float(np.sum(g[k] * d[k]))
        assert abs(fd - an) <= 1e-7 * abs(an), (k, fd, an)
    # the same numbers as the Jacobian operator's adjoint
    hyp = Hypercube([Axis(shape[2]), Axis(shape[1]), Axis(shape[0]), Axis(3)])
    b = Vector(hyp, np.stack([vp, vs, rho]))
    jab = OP.ElasticConvJab_3D(b, None, 1)
    out = b.clone()
    jab.adjoint(False, out, Vector(hyp, np.stack(w)))
    assert np.allclose(out.getNdArray(), np.stack(g), rtol=1e-14, atol=0)


class _Quadratic:
    """J = 1/2 sum_k ||(m_k - t_k) / s_k||^2 (a stand-in for the wave-equation
    misfit with the same misfit_and_gradient interface)."""

    def __init__(self, t, s):
        self.t, self.s, self.calls = t, s, 0

    def misfit_and_gradient(self, *m):
        self.calls += 1
        r = [(a - b) / c for a, b, c in zip(m, self.t, self.s)]
        return 0.5 * float(sum(np.sum(x * x) for x in r)), np.stack([x / c for x, c in zip(r, self.s)])


def test_bounded_lbfgs_reaches_the_projected_minimum():
    """With bounds active on part of the cells, the outer loop converges to
    the minimiser of the quadratic over the box (the clamped target), and
    every recorded misfit is the one of the projected model it evaluated."""
    rng = np.random.default_rng(3)
    shape = (4, 3, 5)
    m0 = [np.full(shape, 2000.0), np.full(shape, 4e9), np.full(shape, 3e9)]
    t = [m * (1 + 0.3 * rng.standard_normal(shape)) for m in m0]
    s = [m * 0.1 for m in m0]
    lo = [0.9 * m for m in m0]
    hi = [1.1 * m for m in m0]
    prob = _Quadratic(t, s)
    out, hist = F.invert(prob, m0, n_iter=30, bounds=(lo, hi))
    want = [np.clip(a, l, h) for a, l, h in zip(t, lo, hi)]
    active = sum(int(np.sum((a < l) | (a > h))) for a, l, h in zip(t, lo, hi))
    assert active > 10
    for o, w_, m in zip(out, want, m0):
        assert np.max(np.abs(o - w_)) < 1e-6 * np.max(np.abs(m))
    for o, l, h in zip(out, lo, hi):
        assert np.all(o >= l) and np.all(o <= h)
    assert hist[-1][1] <= min(j for _, j in hist) + 1e-12
    Jw = prob.misfit_and_gradient(*want)[0]
    assert abs(hist[-1][1] - Jw) <= 1e-9 * Jw


def test_objective_in_vp_vs_rho():
    """objective(..., "vp_vs_rho") chains the conversion: its gradient is the
    central difference of J(g(m))."""
    rng = np.random.default_rng(4)
    shape = (3, 3, 4)
    vp, vs, rho = 2500 + 100 * rng.random(shape), 1400 + 50 * rng.random(shape), 2000 + 100 * rng.random(shape)
    t = [x * (1 + 0.05 * rng.standard_normal(shape)) for x in F.vp_vs_rho_to_elastic(vp, vs, rho)]
    prob = _Quadratic(t, [np.abs(x) * 0.1 for x in t])
    J, g = F.objective(prob, (vp, vs, rho), "vp_vs_rho")
    d = [rng.standard_normal(shape) * x * 1e-2 for x in (vp, vs, rho)]
    h = 1e-4
    jp = F.objective(prob, [a + h * b for a, b in zip((vp, vs, rho), d)], "vp_vs_rho")[0]
    jm = F.objective(prob, [a - h * b for a, b in zip((vp, vs, rho), d)], "vp_vs_rho")[0]
    fd = (jp - jm) / (2 * h)
    an = float(sum(np.sum(a * b) for a, b in zip(g, d)))
    assert abs(fd - an) <= 1e-6 * abs(an)
    with pytest.raises(ValueError):
        F.objective(prob, (vp, vs, rho), "impedance")


def test_gaussian_smooth_torch_matches_scipy():
    """The torch separable smoother used for the starting models equals
    scipy.ndimage.gaussian_filter (float32, nearest edges) to FP32 rounding."""
    from elasticisolib_3d_amd.models import gaussian_smooth
    rng = np.random.default_rng(5)
    a = 10 + 3 * rng.standard_normal((30, 41, 37))
    for sigma in (1.5, 5.0):
        ref = gaussian_smooth(a, sigma)
        out = gaussian_smooth(a, sigma, "cpu")
        assert np.max(np.abs(out - ref)) <= 1e-6 * np.max(np.abs(ref))
