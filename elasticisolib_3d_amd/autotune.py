"""Autotuner of the stencil launch shape.

The stencil kernels take three speed-only parameters (results are
bit-identical for every choice): the x extent of a 64-wide z tile (4, 8 or
16 rows, i.e. 256, 512 or 1024 threads), the number of y planes one block
marches through (the depth of its register queue pipeline before it hands the
column to the next block), and the block-to-tile order (hardware round robin
over the 8 XCDs, contiguous ranges per XCD, y-major ranges per XCD in groups
of x columns, or the round robin over groups of whole z columns). The best
choice depends on the grid (tile counts, partial tiles, L2 reuse of halo
planes), so it is measured: every candidate
is timed with HIP events on the propagator's own stream
(ela_time_step_kernels: a velocity and a stress launch of a forward step on
the handle's current wavefields, which it overwrites), and the fastest is
applied. Results are cached per device name, precision, medium kind and grid
shape in a JSON file (ELA_TUNE_CACHE, default ~/.cache/elasticisolib_3d_amd/tuning.json).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Sequence, Tuple

TILE_X = (16, 8, 4)
Y_CHUNK = (4, 8, 16, 32)
ORDERS = ((0, 1), (1, 1), (2, 1), (3, 1), (3, 2), (3, 4))

Candidate = Tuple[int, int, int, int]  # tile_x, y_chunk, order, x_group


def default_candidates(precision: str = "f32") -> Sequence[Candidate]:
    tiles = TILE_X if precision == "f32" else (8, 4)
    return [(tx, yc, o, g) for tx in tiles for yc in Y_CHUNK for (o, g) in ORDERS]


def _cache_path() -> str:
    return os.environ.get("ELA_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "elasticisolib_3d_amd",
                                                         "tuning.json"))


def _device_name() -> str:
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # the name only keys the cache
        return "unknown"


def _key(prop, target: str = "forward") -> str:
    ny, nx, nz = prop.shape
    kind = getattr(prop, "medium_kind", "scaled")   # the steppers differ with the medium kind
    key = f"{_device_name()}|{prop.precision}|{kind}|{nz}x{nx}x{ny}"
    if target == "gradient":   # the gradient's kernel mix, with its derivative storage
        key += f"|gradient|{prop.gradient_derivatives_used()}"
    return key


def _load() -> Dict[str, dict]:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: Dict[str, dict]) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def apply(prop, best: dict) -> None:
    prop.set_tiling(best["tile_x"], best["y_chunk"])
    prop.set_tile_order(best["order"], best["x_group"])
    if hasattr(prop, "set_tiling_transposed"):
        prop.set_tiling_transposed(best.get("tile_x_transposed", 0))


def autotune(prop, reps: int = 10, candidates: Optional[Iterable[Candidate]] = None, use_cache: bool = True,
             verbose: bool = False, target: str = "forward") -> dict:
    """Time every candidate launch shape on `prop` (its medium must be set),
    apply the fastest and return {tile_x, y_chunk, order, x_group, ms_step,
    table}. The wavefields of the handle are overwritten. `target`
    "forward" ranks the shapes by a forward step (velocity + stress
    launches); "gradient" by the three kernel pairs of an adjoint-state
    gradient step (forward, recomputation with derivative store, transposed
    step with imaging: ela_time_gradient_kernels), which do not share the
    forward's best shape: for every y chunk and block order it takes the
    x tile that is fastest for the forward and recomputing pairs and, for
    the transposed pair (ela_set_tiling_transposed), the x tile fastest for
    that pair, from the same measurements."""
    if target not in ("forward", "gradient"):
        raise ValueError("target must be 'forward' or 'gradient'")
    key = _key(prop, target)
    if use_cache and candidates is None:
        hit = _load().get(key)
        if hit:
            apply(prop, hit)
            return dict(hit, cached=True)
    cands = list(candidates) if candidates is not None else list(default_candidates(prop.precision))
    if hasattr(prop, "set_tiling_transposed"):
        prop.set_tiling_transposed(0)   # candidates are timed with one tile for every step
    prop.set_tiling(*cands[0][:2])
    prop.time_step_kernels(2)  # warm the code objects
    table = []
    for tx, yc, order, xg in cands:
        prop.set_tiling(tx, yc)
        prop.set_tile_order(order, xg)
        if target == "gradient":
            parts = prop.time_gradient_kernels(reps)
        else:
            parts = prop.time_step_kernels(reps)
        table.append((sum(parts),) + tuple(parts) + (tx, yc, order, xg))
        if verbose:
            print(f"[autotune] tile {tx}x{yc} order {order}/{xg}: " + " + ".join(f"{p:.4f}" for p in parts)
                  + " ms", flush=True)
    table.sort()
    t, tx, yc, order, xg = table[0][0], *table[0][-4:]
    best = {"tile_x": tx, "y_chunk": yc, "order": order, "x_group": xg, "ms_step": t, "target": target}
    if target == "gradient":
        best.update(ms_forward=table[0][1], ms_recompute=table[0][2], ms_transposed=table[0][3],
                    tile_x_transposed=tx)
        # per (y chunk, order): the best tile of the forward + recomputing pairs and of the transposed pair
        groups = {}
        for row in table:
            groups.setdefault(row[-3:], []).append(row)
        for (yc_, o_, g_), rows in groups.items():
            f = min(rows, key=lambda r: r[1] + r[2])
            tr = min(rows, key=lambda r: r[3])
            split = f[1] + f[2] + tr[3]
            if split < best["ms_step"]:
                best.update(tile_x=f[4], y_chunk=yc_, order=o_, x_group=g_, ms_step=split, ms_forward=f[1],
                            ms_recompute=f[2], ms_transposed=tr[3], tile_x_transposed=tr[4])
    else:
        best.update(ms_velocity=table[0][1], ms_stress=table[0][2])
    apply(prop, best)
    if use_cache and candidates is None:
        tab = _load()
        tab[key] = best
        _store(tab)
    return dict(best, cached=False, table=[list(r) for r in table])


def _parse_shapes(text: str):
    """"16x8/0/1,4x8/3/4" -> [(16, 8, 0, 1), (4, 8, 3, 4)]"""
    out = []
    for item in text.split(","):
        tile, order, group = item.strip().split("/")
        tx, yc = tile.split("x")
        out.append((int(tx), int(yc), int(order), int(group)))
    return out


def main(argv=None) -> int:
    """Time launch shapes of the stencil on a synthetic medium of the given
    padded grid: every candidate (as the autotuner does, without its cache),
    or the listed shapes in order (--shapes), printing one line per shape."""
    import argparse
    import torch
    from .propagator import Propagator
    ap = argparse.ArgumentParser(prog="python -m elasticisolib_3d_amd.autotune")
    ap.add_argument("--grid", default="328,472,468", help="padded nz,nx,ny")
    ap.add_argument("--precision", default="f32", choices=("f32", "f64"))
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--shapes", default="", help="tile_x x y_chunk / order / x_group, comma separated")
    ap.add_argument("--medium", default="scaled", choices=("scaled", "elastic"),
                    help="8 scaled medium arrays (ela_set_medium) or rho, lambda, mu (ela_set_elastic)")
    ap.add_argument("--target", default="forward", choices=("forward", "gradient"),
                    help="time the forward step, or the three kernel pairs of a gradient step")
    a = ap.parse_args(argv)
    nz, nx, ny = (int(v) for v in a.grid.split(","))
    prop = Propagator(nz, nx, ny, 10, 8, 12.5, 12.5, 12.5, 0.004, 30, 0.99, a.precision)
    dt = torch.float32 if a.precision == "f32" else torch.float64
    if a.medium == "scaled":
        med = torch.empty((8, ny, nx, nz), dtype=dt, device="cuda")
        med[:3] = 1e-4
        med[3] = 3e6
        med[4:] = 1e6
        prop.set_medium(med)
    else:
        med = torch.empty((3, ny, nx, nz), dtype=dt, device="cuda")
        med[0] = 2000.0
        med[1] = 6e9
        med[2] = 4e9
        prop.set_elastic(med, 0.0005)
    del med
    cells = (nz - 8) * (nx - 8) * (ny - 8)
    # algorithmic bytes per cell-update: 8 medium reads, or 3 when the stepper forms the medium
    bpc = (140 if a.medium == "scaled" else 120) * (1 if a.precision == "f32" else 2)
    shapes = _parse_shapes(a.shapes) if a.shapes else list(default_candidates(a.precision))
    best = None
    for tx, yc, order, xg in shapes:
        prop.set_tiling(tx, yc)
        prop.set_tile_order(order, xg)
        if a.target == "gradient":
            fw, rc, tr = prop.time_gradient_kernels(a.reps)
            print(f"[autotune] shape {tx}x{yc}/{order}/{xg}: forward {fw:.4f} ms recompute {rc:.4f} ms "
                  f"transposed {tr:.4f} ms step {fw + rc + tr:.4f} ms", flush=True)
            if best is None or fw + rc + tr < best[0]:
                best = (fw + rc + tr, tx, yc, order, xg)
            continue
        kv, ks = prop.time_step_kernels(a.reps)
        gbs = bpc * cells / ((kv + ks) * 1e-3) / 1e9
        print(f"[autotune] shape {tx}x{yc}/{order}/{xg}: velocity {kv:.4f} ms stress {ks:.4f} ms "
              f"step {kv + ks:.4f} ms ({gbs:.0f} GB/s algorithmic)", flush=True)
        if best is None or kv + ks < best[0]:
            best = (kv + ks, tx, yc, order, xg)
    print(f"[autotune] fastest {best[1]}x{best[2]}/{best[3]}/{best[4]}: {best[0]:.4f} ms per step", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
