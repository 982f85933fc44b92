"""Parity of the HIP propagator (through the C ABI) with the CPU oracle.

Bars: FP64 GPU vs FP64 oracle to 1e-11 relative L2 (same scheme, only the
summation order differs); FP32 GPU vs FP64 oracle to 5e-6 relative L2 on
these short runs, half the 1e-5 bar for FP32 forward data (measured <= 4e-7:
FP32 rounding of the scheme itself, see test_fp32_oracle_error_budget for the
matching CPU figure).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.problems import random_problem, rel

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-11, "f32": 5e-6}

GRIDS = {
    # odd sizes: partial z and x tiles, several y chunks
    "odd": O.Grid(nz=29, nx=31, ny=45, nts=5, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95),
    # more than one 64-wide z tile
    "wide": O.Grid(nz=81, nx=21, ny=19, nts=4, sub=8, dz=12.5, dx=10.0, dy=10.0, dts=0.004, minPad=4, alphaCos=0.9),
}


def make_prop(g, precision):
    from elasticisolib_3d_amd.propagator import Propagator
    return Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)


@pytest.fixture(scope="module")
def problems():
    return {k: random_problem(g, seed=11 + i) for i, (k, g) in enumerate(GRIDS.items())}


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_nl_forward(problems, grid, precision):
    g, P = GRIDS[grid], problems[grid]
    ref = O.lib_nl_forward(g, P["med"], P["src"], P["spos"], P["rpos"])
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    out = prop.nl_forward(P["spos"], P["src"], P["rpos"])
    print(f"nl_forward {grid} {precision}: rel={rel(out, ref):.3e}", flush=True)
    assert rel(out, ref) < TOL[precision]


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_nl_adjoint(problems, grid, precision):
    g, P = GRIDS[grid], problems[grid]
    ref = O.lib_nl_adjoint(g, P["med"], P["rec"], P["spos"], P["rpos"])
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    out = prop.nl_adjoint(P["spos"], P["rec"], P["rpos"])
    print(f"nl_adjoint {grid} {precision}: rel={rel(out, ref):.3e}", flush=True)
    assert rel(out, ref) < TOL[precision]


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("physics", [False, True])
def test_born_forward(problems, grid, precision, physics):
    """Reference mode (the yz quirk) and physics mode against the oracle in
    the same mode."""
    g, P = GRIDS[grid], problems[grid]
    ref = O.lib_born_forward(g, P["med"], P["src"], P["spos"], P["dm"], P["rpos"], physics=physics)
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    prop.set_born_mode(physics)
    out = prop.born_forward(P["spos"], P["src"], P["dm"], P["rpos"])
    print(f"born_forward {grid} {precision} physics={physics}: rel={rel(out, ref):.3e}", flush=True)
    assert rel(out, ref) < TOL[precision]


@pytest.mark.parametrize("grid", list(GRIDS))
@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("physics", [False, True])
def test_born_adjoint(problems, grid, precision, physics):
    g, P = GRIDS[grid], problems[grid]
    ref = O.lib_born_adjoint(g, P["med"], P["src"], P["spos"], P["rec"], P["rpos"], physics=physics)
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    prop.set_born_mode(physics)
    out = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
    print(f"born_adjoint {grid} {precision} physics={physics}: rel={rel(out, ref):.3e}", flush=True)
    assert rel(out, ref) < TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_born_modes_differ_only_in_the_yz_image(problems, precision):
    """On the GPU, reference and physics modes give the same 7 images and a
    different dmuyz image, each equal to the oracle's in its own mode (the
    image-wide error of the adjoint parity test is dominated by the other 7
    arrays, so this checks the yz array alone)."""
    g, P = GRIDS["odd"], problems["odd"]
    out = {}
    for physics in (False, True):
        prop = make_prop(g, precision)
        prop.set_medium(P["med"])
        prop.set_born_mode(physics)
        out[physics] = np.asarray(prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"]))
        ref = O.lib_born_adjoint(g, P["med"], P["src"], P["spos"], P["rec"], P["rpos"], physics=physics)
        e7 = rel(out[physics][7], ref[7])
        print(f"born_adjoint yz image {precision} physics={physics}: rel={e7:.3e}", flush=True)
        assert e7 < TOL[precision]
    assert np.array_equal(out[False][:7], out[True][:7])
    assert rel(out[True][7], out[False][7]) > 0.1


def test_born_adjoint_accumulates(problems):
    """The Born adjoint adds into the incoming perturbation arrays (reference
    BornElasticGpuFunctions_3D.cu:2113 copies them to the device first)."""
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, "f64")
    prop.set_medium(P["med"])
    base = np.full((8,) + g.shape, 0.5)
    out = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"], base.copy())
    ref = O.lib_born_adjoint(g, P["med"], P["src"], P["spos"], P["rec"], P["rpos"])
    assert rel(out - base, ref) < 1e-11


def test_device_memory_path(problems):
    """Device (torch) tensors through the same entry points give the host result."""
    import torch
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, "f32")
    prop.set_medium(torch.tensor(P["med"], dtype=torch.float32, device="cuda"))
    spos = [torch.tensor(p, device="cuda") for p in P["spos"]]
    rpos = [torch.tensor(p, device="cuda") for p in P["rpos"]]
    src = [torch.tensor(s, dtype=torch.float32, device="cuda") for s in P["src"]]
    out = prop.nl_forward(spos, src, rpos)
    torch.cuda.synchronize()
    out = [o.cpu().numpy() for o in out]
    host = make_prop(g, "f32")
    host.set_medium(P["med"])
    ref = host.nl_forward(P["spos"], P["src"], P["rpos"])
    assert rel(out, ref) < 1e-6


LONG = O.Grid(nz=29, nx=27, ny=25, nts=11, sub=8, dz=10.0, dx=10.0, dy=10.0, dts=0.004, minPad=5, alphaCos=0.95)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_born_adjoint_checkpointed_is_bit_identical(precision):
    """Checkpointed storage of the background coarse wavefield (exact state
    checkpoints + recomputation) gives the same bits as keeping every coarse
    sample, for segment lengths that do and do not divide nts - 1. The fine
    steps run: 80 forward + 80 adjoint, plus the recomputation of every
    segment but the last (segments of K counted back from coarse step 10,
    each recomputed through two steps past its end, at most to step 9),
    e.g. K = 3: [0, 1) -> steps 0-2, [1, 4) -> 1-5, [4, 7) -> 4-8, [7, 10)
    kept: 13 coarse steps = 104 fine steps."""
    steps_expected = {1: 368, 2: 288, 3: 264, 4: 240, 9: 184, 10: 160, 50: 160}
    g = LONG
    P = random_problem(g, seed=5)
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    prop.set_born_storage(-1)
    full = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
    assert prop.born_storage_used() == -1 and prop.last_timing()[1] == 160
    ref = O.lib_born_adjoint(g, P["med"], P["src"], P["spos"], P["rec"], P["rpos"])
    print(f"born_adjoint long {precision}: rel={rel(full, ref):.3e}", flush=True)
    assert rel(full, ref) < TOL[precision]
    for K in (1, 2, 3, 4, 9, 10, 50):
        prop.set_born_storage(K)
        out = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
        assert prop.born_storage_used() == min(K, g.nts - 1)
        same = np.array_equal(out, full)
        steps = prop.last_timing()[1]
        print(f"  segment {K}: bit-identical={same}, {steps} fine steps", flush=True)
        assert same and steps == steps_expected[K]
    prop.set_born_storage(0)
    out = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
    assert prop.born_storage_used() == -1 and np.array_equal(out, full)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_born_forward_streaming_long(precision):
    """The interleaved Born forward (background two coarse samples ahead of
    the scattered field, four coarse slices live) on a longer time axis."""
    g = LONG
    P = random_problem(g, seed=6)
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    out = prop.born_forward(P["spos"], P["src"], P["dm"], P["rpos"])
    ref = O.lib_born_forward(g, P["med"], P["src"], P["spos"], P["dm"], P["rpos"])
    print(f"born_forward long {precision}: rel={rel(out, ref):.3e}", flush=True)
    assert rel(out, ref) < TOL[precision]


@pytest.mark.parametrize("segment", [-1, 3])
def test_born_dot_product_fp32_fp64_reductions(segment):
    """Shot-level Born operator (perturbation staggering + FP32 propagation)
    against its adjoint, inner products in FP64: relative mismatch <= 1e-6."""
    from elasticisolib_3d_amd.propagator import born_stagger, born_unstagger
    g = LONG
    P = random_problem(g, seed=8)
    rng = np.random.default_rng(9)
    dm3 = rng.standard_normal((3,) + g.shape) * np.array([20.0, 1e8, 1e8])[:, None, None, None]
    prop = make_prop(g, "f32")
    prop.set_medium(P["med"])
    prop.set_born_storage(segment)
    d = prop.born_forward(P["spos"], P["src"], born_stagger(dm3, P["med"], g.dtw), P["rpos"])
    img = prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])
    lhs = sum(float(np.sum(np.asarray(a, np.float64) * b)) for a, b in zip(d, P["rec"]))
    rhs = float(np.sum(dm3 * born_unstagger(np.asarray(img, np.float64), P["med"], g.dtw)))
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"Born dot test FP32 (segment {segment}): {lhs:.9e} {rhs:.9e} rel {err:.2e}", flush=True)
    assert err < 1e-6


def test_device_outputs_ordered_with_default_stream(problems):
    """Outputs of a device-memory call are consumed by torch work on the
    default stream without an explicit synchronize (the library's stream is
    ordered with it), and repeated calls reuse fresh output tensors."""
    import torch
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, "f64")
    spos = [torch.tensor(p, device="cuda") for p in P["spos"]]
    rpos = [torch.tensor(p, device="cuda") for p in P["rpos"]]
    src = [torch.tensor(s, dtype=torch.float64, device="cuda") for s in P["src"]]
    host = make_prop(g, "f64")
    host.set_medium(P["med"])
    ref = host.nl_forward(P["spos"], P["src"], P["rpos"])
    ref_norm = sum(float(np.sum(r * r)) for r in ref)
    for scale in (1.0, 1.05, 1.0):
        prop.set_medium(torch.tensor(P["med"] * scale, dtype=torch.float64, device="cuda"))
        out = prop.nl_forward(spos, src, rpos)
        n = float(sum((o * o).sum() for o in out))      # no synchronize before the reduction
        if scale == 1.0:
            assert abs(n - ref_norm) <= 1e-10 * ref_norm, (n, ref_norm)
        else:
            assert abs(n - ref_norm) > 1e-6 * ref_norm


@pytest.mark.parametrize("order,xgroup", [(1, 1), (2, 1), (2, 3), (2, 64), (3, 1), (3, 2), (3, 64)])
def test_tile_orders_give_identical_results(problems, order, xgroup):
    """Every block-to-tile order covers every tile exactly once: the results
    are bit-identical to the hardware order, with and without partial x
    groups, for several tile shapes."""
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, "f32")
    prop.set_medium(P["med"])
    for tx, yc in ((16, 8), (8, 3), (4, 5)):
        prop.set_tiling(tx, yc)
        prop.set_tile_order(0, 1)
        ref = prop.nl_forward(P["spos"], P["src"], P["rpos"])
        prop.set_tile_order(order, xgroup)
        out = prop.nl_forward(P["spos"], P["src"], P["rpos"])
        assert all(np.array_equal(a, b) for a, b in zip(out, ref)), (tx, yc)


@pytest.mark.parametrize("medium", ["scaled", "elastic"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_tilings_give_identical_results(problems, precision, medium):
    """Every cell's update uses the same operands in the same order whatever
    the tile width and y chunk, so all launch shapes give bit-identical
    results: the tile rows and halo rows that FP32 loads straight into LDS
    are filled by whole waves, one wave per row, and a tile 16 wide has 12
    rows (16 waves) that take no x halo. Nonlinear forward (forward stepper)
    and Born adjoint (premultiplied transposed stepper), both medium kinds."""
    g, P = GRIDS["wide"], problems["wide"]
    prop = make_prop(g, precision)
    if medium == "scaled":
        prop.set_medium(P["med"])
    else:
        prop.set_elastic(np.stack([P["rho"], P["lam"], P["mu"]]), g.dtw)
    out = {}
    for tx, yc in ((4, 5), (8, 3), (16, 8), (16, 1)):
        prop.set_tiling(tx, yc)
        out[(tx, yc)] = (prop.nl_forward(P["spos"], P["src"], P["rpos"]),
                         [prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])])
    ref = out[(4, 5)]
    for k, v in out.items():
        for a, b in zip(v, ref):
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), (k, precision, medium)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_graph_replay_matches_eager(problems, precision):
    """The coarse step replayed as a HIP graph (nts - 1 replays per call, the
    coarse index advancing in device memory) gives the eager results bit for
    bit, call after call, with eager calls in between."""
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    ref_f = prop.nl_forward(P["spos"], P["src"], P["rpos"])
    ref_a = prop.nl_adjoint(P["spos"], P["rec"], P["rpos"])
    for _ in range(2):
        prop.set_graph(True)
        out_f = prop.nl_forward(P["spos"], P["src"], P["rpos"])
        out_a = prop.nl_adjoint(P["spos"], P["rec"], P["rpos"])
        prop.set_graph(False)
        eager = prop.nl_forward(P["spos"], P["src"], P["rpos"])
        assert all(np.array_equal(a, b) for a, b in zip(out_f, ref_f))
        assert all(np.array_equal(a, b) for a, b in zip(out_a, ref_a))
        assert all(np.array_equal(a, b) for a, b in zip(eager, ref_f))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_graph_recaptured_when_launches_change(problems, precision):
    """The handle keeps the graph of the last call and replays it only when
    the launches are the same: calls with other receiver lists, another
    launch shape and another medium in between give their eager results bit
    for bit, and a repeat of the first call replays the kept graph to its
    first result."""
    g, P = GRIDS["odd"], problems["odd"]
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    fewer = [p[:3] for p in P["rpos"]]
    med2 = P["med"] * 1.01

    def calls():
        out = [prop.nl_forward(P["spos"], P["src"], P["rpos"]),
               prop.nl_forward(P["spos"], P["src"], fewer)]
        prop.set_tiling(8, 3)
        out.append(prop.nl_forward(P["spos"], P["src"], P["rpos"]))
        prop.set_medium(med2)
        out.append(prop.nl_forward(P["spos"], P["src"], P["rpos"]))
        prop.set_medium(P["med"])
        prop.set_tiling(16, 8)
        out.append(prop.nl_forward(P["spos"], P["src"], P["rpos"]))
        return out

    prop.set_tiling(16, 8)
    prop.set_graph(False)
    ref = calls()
    prop.set_graph(True)
    out = calls()
    for i, (a, b) in enumerate(zip(out, ref)):
        assert all(np.array_equal(x, y) for x, y in zip(a, b)), (precision, i)
    assert all(np.array_equal(x, y) for x, y in zip(out[-1], ref[0]))


def test_stream_probe_reports_bandwidth():
    """The streaming reference of the roofline runs its five mixes and
    rejects others."""
    from elasticisolib_3d_amd._lib import ElaError
    from elasticisolib_3d_amd.propagator import Propagator
    g = GRIDS["odd"]
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, "f32")
    for r, w in ((1, 1), (12, 3), (14, 6), (10, 3), (11, 6)):
        gbs = prop.stream_probe(r, w, 1 << 20, reps=3)
        print(f"stream probe {r}r/{w}w over 4 MiB arrays: {gbs:.0f} GB/s", flush=True)
        assert gbs > 0
    with pytest.raises(ElaError):
        prop.stream_probe(2, 1, 1 << 20)
    with pytest.raises(ElaError):
        prop.stream_probe(12, 3, 1001)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("name", ["odd", "wide"])
def test_elastic_parameters_match_scaled_medium(problems, precision, name):
    """ela_set_elastic (rho, lambda, mu; medium formed in the stepper and the
    8 arrays derived on the GPU) against ela_set_medium with the host-scaled
    8 arrays: bit-identical in FP64 (same operations in the same order), FP32
    to rounding, for the nonlinear forward and adjoint, the Born forward and
    adjoint and the adjoint-state gradient."""
    from elasticisolib_3d_amd.propagator import Propagator, scale_medium
    g, P = GRIDS[name], problems[name]
    rlm = np.stack([P["rho"], P["lam"], P["mu"]])
    med8 = scale_medium(P["rho"], P["lam"], P["mu"], g.dtw)   # the library's host scaling
    out = {}
    for kind in ("med8", "elastic"):
        prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)
        if kind == "med8":
            prop.set_medium(med8)
        else:
            prop.set_elastic(rlm, g.dtw)
        prop.set_gradient_checkpoints(3, 0, "native")
        prop.gradient_forward(P["spos"], P["src"], P["rpos"])
        g8, gs = prop.gradient_backward(P["spos"], P["src"], P["rpos"], P["rec"])
        out[kind] = (prop.nl_forward(P["spos"], P["src"], P["rpos"]),
                     prop.nl_adjoint(P["spos"], P["rec"], P["rpos"]),
                     prop.born_forward(P["spos"], P["src"], P["dm"], P["rpos"]),
                     [prop.born_adjoint(P["spos"], P["src"], P["rec"], P["rpos"])],
                     [g8] + list(gs))
    for k, what in enumerate(("nl_forward", "nl_adjoint", "born_forward", "born_adjoint", "gradient")):
        a, b = out["elastic"][k], out["med8"][k]
        err = rel(a, b)
        print(f"elastic vs med8 {precision} {name} {what}: {err:.2e}", flush=True)
        if precision == "f64":
            assert all(np.array_equal(x, y) for x, y in zip(a, b)), what
        else:
            assert err < 2e-6, (what, err)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_derived_medium_equals_host_scaling(problems, precision):
    """The 8 arrays ela_set_elastic derives on the GPU against the host
    scaling (ela_scale_medium, FP64) of the same (rho, lambda, mu) values: in
    FP64 all 8 bit for bit; in FP32 (inputs rounded to FP32 first, the host
    scaling of those values rounded once to FP32) the buoyancies and normal
    moduli bit for bit -- the GPU forms them with a correctly rounded FP64
    division and products -- and the shear moduli, FP32 half-cell sums,
    within 2 units in the last place."""
    from elasticisolib_3d_amd.propagator import Propagator, scale_medium
    g = GRIDS["odd"]
    P = problems["odd"]
    dt = np.float64 if precision == "f64" else np.float32
    rlm = np.stack([P["rho"], P["lam"], P["mu"]]).astype(dt)
    host = scale_medium(*[x.astype(np.float64) for x in rlm], g.dtw).astype(dt)
    prop = Propagator(g.nz, g.nx, g.ny, g.nts, g.sub, g.dz, g.dx, g.dy, g.dts, g.minPad, g.alphaCos, precision)
    prop.set_elastic(rlm, g.dtw)
    dev = prop.get_medium()
    exact = range(8) if precision == "f64" else range(5)
    for k in exact:
        nbad = int(np.count_nonzero(dev[k] != host[k]))
        print(f"derived medium {precision} array {k}: {nbad} of {dev[k].size} values differ", flush=True)
        assert nbad == 0, k
    if precision == "f32":
        for k in range(5, 8):
            ulp = np.spacing(np.abs(host[k])).astype(np.float64)
            d = np.abs(dev[k].astype(np.float64) - host[k].astype(np.float64))
            assert np.all(d <= 2 * ulp), (k, float((d / np.maximum(ulp, 1e-300)).max()))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_empty_point_lists(problems, precision):
    """No receiver on any grid, or no source on any grid: the calls succeed,
    return zero-length traces or exact zeros (nothing is injected, so every
    field stays 0), and leave the handle as it was: the next full call still
    matches the oracle."""
    g, P = GRIDS["odd"], problems["odd"]
    none = [np.zeros(0, np.int64) for _ in range(7)]
    empty = [np.zeros((0, g.nts)) for _ in range(9)]
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    print(f"empty point lists {precision}", flush=True)
    out = prop.nl_forward(P["spos"], P["src"], none)
    assert len(out) == 9 and all(np.asarray(x).size == 0 for x in out)
    quiet = prop.nl_forward(none, empty, P["rpos"])
    assert sum(np.asarray(x).size for x in quiet) > 0
    assert all(np.count_nonzero(np.asarray(x)) == 0 for x in quiet)
    back = prop.nl_adjoint(P["spos"], empty, none)
    assert all(np.count_nonzero(np.asarray(x)) == 0 for x in back)
    img = prop.born_adjoint(P["spos"], P["src"], empty, none)
    assert np.count_nonzero(np.asarray(img)) == 0
    ref = O.lib_nl_forward(g, P["med"], P["src"], P["spos"], P["rpos"])
    full = prop.nl_forward(P["spos"], P["src"], P["rpos"])
    print(f"  after the empty calls: rel={rel(full, ref):.3e}", flush=True)
    assert rel(full, ref) < TOL[precision]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_positions_outside_the_interior_are_rejected(problems, precision):
    """A source or receiver in the FAT layer, at a negative index or past the
    grid is an ELA_ERR_ARG error (the conversion kernel flags it before any
    kernel uses a position); the handle stays usable."""
    from elasticisolib_3d_amd._lib import ElaError
    g, P = GRIDS["odd"], problems["odd"]
    c = (g.ny // 2) * g.nz * g.nx + (g.nx // 2) * g.nz
    prop = make_prop(g, precision)
    prop.set_medium(P["med"])
    print(f"positions outside the interior {precision}", flush=True)
    for bad in (c + 0, c + g.nz - 1, -5, g.nz * g.nx * g.ny):   # z = 0 and z = nz - 1 lie in the FAT layer
        spos = [p.copy() for p in P["spos"]]
        spos[0] = np.concatenate([spos[0], [bad]]).astype(np.int64)
        src = [np.vstack([s, np.ones((1, g.nts))]) if O.COMP_GRID[k] == 0 else s for k, s in enumerate(P["src"])]
        with pytest.raises(ElaError):
            prop.nl_forward(spos, src, P["rpos"])
        rpos = [p.copy() for p in P["rpos"]]
        rpos[3] = np.concatenate([rpos[3], [bad]]).astype(np.int64)
        with pytest.raises(ElaError):
            prop.nl_forward(P["spos"], P["src"], rpos)
    ref = O.lib_nl_forward(g, P["med"], P["src"], P["spos"], P["rpos"])
    full = prop.nl_forward(P["spos"], P["src"], P["rpos"])
    print(f"  after the rejected calls: rel={rel(full, ref):.3e}", flush=True)
    assert rel(full, ref) < TOL[precision]


@pytest.mark.parametrize("precision,shape", [("f32", (1100, 1000, 1000)), ("f64", (560, 1000, 1000))])
def test_grid_beyond_32_bit_offsets_is_rejected(precision, shape):
    """Kernels index an array with 32-bit offsets: a grid whose arrays would
    exceed 4 GiB is ELA_ERR_ARG at creation, before anything is allocated."""
    from elasticisolib_3d_amd._lib import ElaError
    from elasticisolib_3d_amd.propagator import Propagator
    nz, nx, ny = shape
    print(f"grid {nz}x{nx}x{ny} {precision}", flush=True)
    with pytest.raises(ElaError, match="4 GiB"):
        Propagator(nz, nx, ny, 5, 8, 10.0, 10.0, 10.0, 0.004, 5, 0.95, precision)
