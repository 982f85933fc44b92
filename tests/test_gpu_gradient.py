"""Adjoint-state gradient on the GPU (through the C ABI) against the FP64
oracle gradient, for exact and bfloat16 checkpoints, one and two levels."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import born_blob_problem, random_problem, rel

pytestmark = pytest.mark.gpu

G = O.Grid(nz=29, nx=27, ny=25, nts=6, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95)


@pytest.fixture(scope="module")
def problem():
    P = random_problem(G, seed=31)
    g8, gsrc = O.lib_gradient(G, P["med"], P["src"], P["spos"], P["rec"], P["rpos"])
    return P, g8, gsrc


def make_prop(precision, g=G):
    from elasticisolib_3d_amd.propagator import Propagator
    return Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)


# interior z extent 24, a multiple of 8: the int16 checkpoints take the vectorised pack and unpack kernels
G8 = O.Grid(nz=32, nx=23, ny=21, nts=6, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95)


@pytest.mark.parametrize("K,K1", [(2, 8), (3, 3)])
def test_vectorised_int16_checkpoints(K, K1):
    """On a grid whose interior z extent is a multiple of 8 the FP32 int16
    checkpoints are packed and restored 8 values per lane; the gradient
    against the FP64 oracle stays where the scalar kernels put it (same
    blocks, scales and values)."""
    P = random_problem(G8, seed=33)
    g8_ref, gsrc_ref = O.lib_gradient(G8, P["med"], P["src"], P["spos"], P["rec"], P["rpos"])
    prop = make_prop("f32", G8)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(K, K1, "s16")
    prop.gradient_forward(P["spos"], P["src"], P["rpos"])
    g8, gsrc = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
    e8, es = rel(g8, g8_ref), rel(gsrc, gsrc_ref)
    print(f"vectorised int16 checkpoints K={K} K1={K1}: g8 rel={e8:.3e} src rel={es:.3e}", flush=True)
    assert e8 < 5e-5 and es < 5e-5


CASES = [
    # precision, interval, outer, checkpoint format, derivative storage,
    # tolerance (relative L2 vs the FP64 oracle)
    ("f64", 3, 3, "native", "native", 1e-11),   # one level, exact checkpoints
    ("f64", 2, 6, "native", "native", 1e-11),   # two levels (outer segment recomputed)
    ("f64", 20, 0, "native", "native", 1e-11),  # one segment longer than the run
    ("f32", 4, 4, "native", "native", 1e-5),
    ("f32", 2, 8, "bf16", "native", 5e-3),      # bfloat16 outer checkpoints
    ("f32", 2, 8, "s16", "native", 5e-5),       # block-scaled int16 outer checkpoints
    ("f64", 3, 3, "s16", "native", 5e-5),
    ("f32", 4, 4, "native", "s16", 5e-5),       # block-scaled int16 derivative combinations
    ("f32", 2, 8, "s16", "s16", 5e-5),          # both, two levels
]


@pytest.mark.parametrize("precision,K,K1,ck,drv,tol", CASES)
def test_gradient_matches_oracle(problem, precision, K, K1, ck, drv, tol):
    P, g8_ref, gsrc_ref = problem
    prop = make_prop(precision)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(K, K1, ck)
    prop.set_gradient_derivatives(drv)
    rec = prop.gradient_forward(P["spos"], P["src"], P["rpos"])
    fwd = prop.nl_forward(P["spos"], P["src"], P["rpos"])
    assert all(np.array_equal(a, b) for a, b in zip(rec, fwd)), "gradient forward data differ from nl_forward"
    g8, gsrc = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
    e8, es = rel(g8, g8_ref), rel(gsrc, gsrc_ref)
    print(f"gradient {precision} K={K} K1={K1} {ck}/{drv}: g8 rel={e8:.3e} src rel={es:.3e} "
          f"plan={prop.gradient_plan_used()}", flush=True)
    assert e8 < tol and es < tol


def test_derivative_storage_formats(problem):
    """The derivative combinations the backward recomputes and images with
    are stored exactly by default; FP32 handles also take block-scaled int16
    (FP64 handles refuse it). The int16 gradient differs from the
    exact-storage one by what the format allows (each value within 2^-16 of
    its 32-value block's largest magnitude) and the source gradient, which
    does not image, not at all; switching back restores the exact result."""
    from elasticisolib_3d_amd._lib import ElaError
    P, _, _ = problem
    p64 = make_prop("f64")
    assert p64.gradient_derivatives_used() == "native"
    for fmt in ("s16", "bf16"):
        with pytest.raises(ElaError):
            p64.set_gradient_derivatives(fmt)
    p64.close()
    prop = make_prop("f32")
    with pytest.raises(ElaError):
        prop.set_gradient_derivatives("bf16")
    assert prop.gradient_derivatives_used() == "native"
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(4, 4, "native")
    out = {}
    for fmt in ("native", "s16", "native"):
        prop.set_gradient_derivatives(fmt)
        assert prop.gradient_derivatives_used() == fmt
        prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        out.setdefault(fmt, []).append(prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"]))
    (n1, ns1), (n2, ns2) = out["native"]
    assert np.array_equal(n1, n2) and all(np.array_equal(x, y) for x, y in zip(ns1, ns2))
    q, qs = out["s16"][0]
    d8 = rel(q, n1)
    print(f"int16 derivatives vs exact: g8 rel={d8:.3e}", flush=True)
    assert 0 < d8 < 5e-5
    assert all(np.array_equal(x, y) for x, y in zip(qs, ns1))


@pytest.mark.parametrize("precision,drv", [("f32", "native"), ("f32", "s16"), ("f64", "native")])
def test_gradient_kernel_timing_keeps_checkpoints(problem, precision, drv):
    """ela_time_gradient_kernels times the gradient step's three kernel
    pairs on the handle (for the launch-shape tuner) between a gradient
    forward and its backward without disturbing the stored checkpoints: the
    backward gives the result of an undisturbed run bit for bit."""
    P = problem[0]
    prop = make_prop(precision)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(4, 0, "native")
    prop.set_gradient_derivatives(drv)

    def grad(timed):
        prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        t = prop.time_gradient_kernels(2) if timed else None
        g8, gs = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
        return t, np.array(g8), [np.array(x) for x in gs]

    _, g_ref, s_ref = grad(False)
    t, g, s = grad(True)
    print(f"gradient kernel pairs {precision}/{drv}: " + ", ".join(f"{x:.4f} ms" for x in t), flush=True)
    assert all(x > 0 for x in t)
    assert np.array_equal(g, g_ref) and all(np.array_equal(x, y) for x, y in zip(s, s_ref))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_transposed_tile_does_not_change_results(problem, precision):
    """The premultiplied transposed steps (gradient backward, Born adjoint)
    may run with an x tile of their own (ela_set_tiling_transposed); the
    gradient, the source gradient and the Born image stay bit-identical."""
    P = problem[0]
    prop = make_prop(precision)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(4, 0, "native")
    prop.set_tiling(8, 8)
    out = []
    for tt in (0, 4, 16):
        prop.set_tiling_transposed(tt)
        assert prop.get_tiling_transposed() == (tt or 8)
        prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        g8, gs = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
        img = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
        out.append((np.array(g8), [np.array(x) for x in gs], np.array(img)))
    (ga, sa, ia) = out[0]
    for gb, sb, ib in out[1:]:
        assert np.array_equal(ga, gb) and np.array_equal(ia, ib)
        assert all(np.array_equal(x, y) for x, y in zip(sa, sb))


def test_gradient_accumulates_and_requires_forward(problem):
    from elasticisolib_3d_amd._lib import ElaError
    P, g8_ref, _ = problem
    prop = make_prop("f64")
    prop.set_medium(P["med"])
    with pytest.raises(ElaError):
        prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
    prop.set_gradient_checkpoints(3, 0, False)
    prop.gradient_forward(P["spos"], P["src"], P["rpos"])
    base = np.full((8,) + G.shape, 0.25)
    g8, _ = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"], base.copy())
    assert rel(g8 - base, g8_ref) < 1e-11


def test_gradient_device_tensors(problem):
    import torch
    P, g8_ref, gsrc_ref = problem
    prop = make_prop("f32")
    dev = "cuda"
    prop.set_medium(torch.tensor(P["med"], dtype=torch.float32, device=dev))
    spos = [torch.tensor(p, device=dev) for p in P["spos"]]
    rpos = [torch.tensor(p, device=dev) for p in P["rpos"]]
    src = [torch.tensor(s, dtype=torch.float32, device=dev) for s in P["src"]]
    adj = [torch.tensor(s, dtype=torch.float32, device=dev) for s in P["rec"]]
    prop.set_gradient_checkpoints(4, 0, False)
    prop.set_gradient_derivatives("native")  # exact storage: the bound is FP32 rounding
    prop.gradient_forward(spos, src, rpos)
    g8, gsrc = prop.gradient_backward(spos, src, rpos, adj)
    torch.cuda.synchronize()
    assert rel(g8.cpu().numpy(), g8_ref) < 1e-5
    assert rel([x.cpu().numpy() for x in gsrc], gsrc_ref) < 1e-5


@pytest.mark.parametrize("param", [0, 1])
def test_gradient_agrees_with_born_adjoint(param):
    """The product's adjoint-state gradient against the product's Born
    adjoint (the reference's Born-adjoint gradient path) for J = <r, F(m)>:
    <dJ/dM, dM> = <B^T r, dm> along a smooth density or lambda perturbation,
    to the Born operator's coarse-time linearisation (a few percent); the
    GPU gradient itself matches the FP64 oracle to rounding."""
    g, rho, lam, mu, d3, spos, rpos, fwd = born_blob_problem(param)
    r, med0, s0 = fwd(0.0)
    from elasticisolib_3d_amd.propagator import Propagator
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, "f64")
    prop.set_medium(med0)
    prop.set_gradient_checkpoints(20, 0, "native")
    prop.gradient_forward(spos, s0, rpos)
    g8, _ = prop.gradient_backward(spos, s0, rpos, r)
    g8_ref, _ = O.lib_gradient(g, med0, s0, spos, r, rpos)
    assert rel(g8, g8_ref) < 1e-11
    h = 1e-4
    dM = (O.lib_scale_medium(rho + h * d3[0], lam + h * d3[1], mu, g.dtw)
          - O.lib_scale_medium(rho - h * d3[0], lam - h * d3[1], mu, g.dtw)) / (2 * h)
    lhs = float(np.sum(g8 * dM))
    img = prop.born_adjoint(spos, s0, r, rpos)
    dm8 = O.born_stagger_perturbation(d3, med0, g.dtw, O.lib_stagger)
    rhs = float(np.sum(np.asarray(img) * dm8))
    print(f"GPU gradient vs GPU Born adjoint, parameter {param}: {lhs:.6e} {rhs:.6e} "
          f"rel {abs(lhs - rhs) / abs(rhs):.4f}", flush=True)
    assert abs(lhs - rhs) <= 0.1 * abs(rhs)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_gradient_and_born_adjoint_are_reproducible(problem, precision):
    """No atomics on the transposed paths: the residual injection updates the
    three normal stresses of a point from one thread, so the gradient and the
    Born adjoint image are bit-identical from run to run (receivers share
    their grid points across the normal-stress components)."""
    P, _, _ = problem
    prop = make_prop(precision)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(3, 0, "native")
    runs = []
    for _ in range(2):
        prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        g8, gs = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
        img = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
        runs.append((np.array(g8), [np.array(x) for x in gs], np.array(img)))
    (ga, sa, ia), (gb, sb, ib) = runs
    assert np.array_equal(ga, gb)
    assert all(np.array_equal(x, y) for x, y in zip(sa, sb))
    assert np.array_equal(ia, ib)


@pytest.mark.parametrize("precision,K,K1,ck", [("f64", 4, 0, "native"), ("f32", 2, 8, "s16"),
                                                ("f32", 6, 0, "bf16"), ("f64", 3, 3, "native")])
def test_gradient_graph_replay_matches_eager(problem, precision, K, K1, ck):
    """The gradient's graph replays (pairs of forward steps, whole checkpoint
    segments of forward-with-store and transposed steps, the fine-step index
    and filter weights read from device memory) give the host-launched
    traces, medium gradient and source gradient bit for bit, on two calls in
    a row (the second replays the graphs the handle kept). The cases cover
    one and two checkpoint levels, a last segment shorter than the others
    (K = 6 over 40 steps, launched from the host) and an odd interval (no
    segment graphs)."""
    P = problem[0]
    prop = make_prop(precision)
    prop.set_medium(P["med"])
    prop.set_gradient_checkpoints(K, K1, ck)

    def run():
        rec = prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        g8, gs = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
        return [np.array(r) for r in rec] + [np.array(g8)] + [np.array(x) for x in gs]

    prop.set_graph(False)
    ref = run()
    prop.set_graph(True)
    for call in range(2):
        out = run()
        bad = [i for i, (a, b) in enumerate(zip(out, ref)) if not np.array_equal(a, b)]
        print(f"gradient graph {precision} K={K} K1={K1} {ck} call {call}: differing outputs {bad}", flush=True)
        assert not bad
