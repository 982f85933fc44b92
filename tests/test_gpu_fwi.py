"""FWI misfit gradient on the GPU: the FP64 adjoint-state gradient against
central differences of the FP64 misfit, FP32 with bfloat16 checkpoints
against FP64, and the torch.autograd front end."""
import numpy as np
import pytest

from tests.problems import random_medium

pytestmark = pytest.mark.gpu

D = 10.0


def _problem(precision, bf16, interval=4, outer=0, checkpoint=None, autotune=False):
    from oracle import oracle as O
    from elasticisolib_3d_amd.acquisition import PaddedAxes, ShotGeometry
    from elasticisolib_3d_amd.fwi import FwiProblem, FwiShot
    g = O.Grid(nz=30, nx=32, ny=28, nts=7, sub=8, dz=D, dx=D, dy=D, dts=0.004, minPad=5, alphaCos=0.95)
    pad = 5
    ax = PaddedAxes(nz=g.nz, nx=g.nx, ny=g.ny, oz=-(4 + pad) * D, ox=-(4 + pad) * D, oy=-(4 + pad) * D, dz=D, dx=D,
                    dy=D, fat=4, z_pad_minus=pad, z_pad_plus=pad, x_pad_minus=pad, x_pad_plus=pad, y_pad=pad)
    rng = np.random.default_rng(7)
    shots = []
    t = np.arange(g.nts) * g.dts
    w = np.exp(-((t - 0.008) / 0.004) ** 2)
    for xs in (40.0, 90.0):
        src = ShotGeometry(ax, [35.0], [xs], [50.0])
        rz = rng.uniform(2, 20, 6); rx = rng.uniform(10, 110, 6); ry = rng.uniform(10, 80, 6)
        rec = ShotGeometry(ax, rz, rx, ry)
        sig = [np.zeros((1, g.nts)) for _ in range(9)]
        for c in (3, 4, 5):
            sig[c][0] = w * 1e6
        sig[2][0] = 0.3e6 * w
        shots.append(FwiShot(src, rec, sig))
    prob = FwiProblem(ax, g.nts, g.dts, g.sub, g.alphaCos, shots, precision=precision, interval=interval,
                      outer_interval=outer, bf16=bf16, checkpoint=checkpoint, autotune=autotune)
    true = random_medium(g, np.random.default_rng(8), contrast=0.1)
    prob.synthesize(*true)
    start = [m * (1 + 0.03 * np.random.default_rng(9 + k).standard_normal(m.shape)) for k, m in enumerate(true)]
    return prob, start


def test_fp64_gradient_matches_central_differences():
    prob, m0 = _problem("f64", bf16=False, interval=4, outer=8)
    J, g3 = prob.misfit_and_gradient(*m0)
    assert J > 0
    rng = np.random.default_rng(10)
    h = 1e-5
    for trial in range(2):
        dm = [rng.standard_normal(m.shape) * m for m in m0]
        jp = prob.misfit(*[m + h * d for m, d in zip(m0, dm)])
        jm = prob.misfit(*[m - h * d for m, d in zip(m0, dm)])
        fd = (jp - jm) / (2 * h)
        an = float(sum(np.sum(g * d) for g, d in zip(g3, dm)))
        print(f"directional derivative: FD {fd:.10e} adjoint {an:.10e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an)
    # single cells: the largest |gradient| of each parameter
    for k in range(3):
        i = np.unravel_index(np.argmax(np.abs(g3[k])), g3[k].shape)
        dm = [np.zeros_like(m) for m in m0]
        dm[k][i] = m0[k][i]
        h = 1e-3
        jp = prob.misfit(*[m + h * d for m, d in zip(m0, dm)])
        jm = prob.misfit(*[m - h * d for m, d in zip(m0, dm)])
        fd = (jp - jm) / (2 * h)
        an = float(g3[k][i] * m0[k][i])
        print(f"cell {k} {i}: FD {fd:.8e} adjoint {an:.8e} rel {abs(fd - an) / abs(an):.2e}", flush=True)
        assert abs(fd - an) <= 1e-4 * abs(an)


def test_fp32_bf16_gradient_close_to_fp64():
    p64, m0 = _problem("f64", bf16=False)
    J64, g64 = p64.misfit_and_gradient(*m0)
    p32, _ = _problem("f32", bf16=True, interval=4, outer=8)
    J32, g32 = p32.misfit_and_gradient(*m0)
    p32e, _ = _problem("f32", bf16=False, interval=4)
    J32e, g32e = p32e.misfit_and_gradient(*m0)
    p32s, _ = _problem("f32", bf16=True, interval=4, outer=8, checkpoint="s16", autotune=True)
    J32s, g32s = p32s.misfit_and_gradient(*m0)
    e = float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64))
    ee = float(np.linalg.norm(g32e - g64) / np.linalg.norm(g64))
    es = float(np.linalg.norm(g32s - g64) / np.linalg.norm(g64))
    ej = abs(J32 - J64) / J64
    print(f"FP32 gradient vs FP64: bf16 checkpoints {e:.2e}, int16 block-scaled {es:.2e}, exact checkpoints "
          f"{ee:.2e}; misfit {ej:.2e}", flush=True)
    assert ee < 1e-4
    assert e < 1e-2
    assert es < 1e-4 and es < e / 10
    assert ej < 1e-4
    assert p32s.prop.gradient_plan_used()[2] == "s16"
    assert p32s.tuned is not None   # launch shape tuned on the first medium (results do not depend on it)


def test_autograd_front_end():
    import torch
    from elasticisolib_3d_amd.fwi import fwi_misfit
    prob, m0 = _problem("f64", bf16=False)
    ts = [torch.tensor(m, dtype=torch.float64, requires_grad=True) for m in m0]
    J = fwi_misfit(ts[0], ts[1], ts[2], prob)
    (2.0 * J).backward()
    Jr, g3 = prob.misfit_and_gradient(*m0)
    assert abs(float(J.detach()) - Jr) <= 1e-12 * Jr
    # repeated runs agree to rounding (atomic additions at the receivers make
    # the FP64 summation order run-dependent)
    for k in range(3):
        g = ts[k].grad.numpy()
        assert np.linalg.norm(g - 2.0 * g3[k]) <= 1e-10 * np.linalg.norm(2.0 * g3[k])


def test_outer_loop_reduces_the_misfit():
    """Three L-BFGS iterations with the adjoint-state gradient reduce the
    FWI misfit of the small problem."""
    from elasticisolib_3d_amd.fwi import invert
    prob, m0 = _problem("f32", bf16=True)
    out, hist = invert(prob, m0, n_iter=3)
    print("misfit history:", [f"{j:.4e}" for _, j in hist], flush=True)
    assert hist[-1][1] < 0.9 * hist[0][1]
    assert all(np.all(np.isfinite(o)) for o in out)
