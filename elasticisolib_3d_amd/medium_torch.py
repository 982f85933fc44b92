"""Medium staggering, scaling and their transposes on torch tensors (any
device), for the FWI paths that keep the model on the GPU.

The operations are those of the library's host routines, element by element
and in the same order (so the FP64 results are the same numbers):
half-cell averages of elastic_iso_float/src/stagger_3D.cpp:5-208,
the staggering and scaling of fdParamElastic_3D.cpp:130-183
(ela_scale_medium) and its transpose (ela_medium_gradient). Arrays are
[..., ny, nx, nz] (z fastest); axis names: "z" the last dimension, "x" the
one before, "y" the one before that.
"""
from __future__ import annotations

_DIM = {"z": -1, "x": -2, "y": -3}


def half_cell_adjoint(a, axis: str):
    """out[0] = in[0] / 2, out[i] = (in[i] + in[i-1]) / 2 for 0 < i < n-1,
    out[n-1] = in[n-1] + in[n-2] / 2 (the averaging that staggers the
    medium)."""
    d = _DIM[axis]
    n = a.shape[d]
    out = a.clone()
    out.narrow(d, 0, 1).copy_(0.5 * a.narrow(d, 0, 1))
    if n > 2:
        out.narrow(d, 1, n - 2).copy_(0.5 * (a.narrow(d, 1, n - 2) + a.narrow(d, 0, n - 2)))
    out.narrow(d, n - 1, 1).copy_(a.narrow(d, n - 1, 1) + 0.5 * a.narrow(d, n - 2, 1))
    return out


def half_cell_forward(a, axis: str):
    """Transpose of half_cell_adjoint: out[i] = (in[i] + in[i+1]) / 2 for
    i < n-1, out[n-1] = in[n-1]."""
    d = _DIM[axis]
    n = a.shape[d]
    out = a.clone()
    out.narrow(d, 0, n - 1).copy_(0.5 * (a.narrow(d, 0, n - 1) + a.narrow(d, 1, n - 1)))
    return out


def scale_medium(rho, lam, mu, dtw: float):
    """(rho, lambda, mu) -> the 8 staggered, scaled arrays rhoxDtw rhoyDtw
    rhozDtw lamb2MuDtw lambDtw muxzDtw muxyDtw muyzDtw, stacked on a new
    leading dimension (ela_scale_medium)."""
    import torch
    tdt = 2.0 * dtw
    t = half_cell_adjoint(mu, "x")
    s = half_cell_adjoint(mu, "y")
    # a tensor numerator: torch turns `scalar / tensor` into a reciprocal times the scalar (one more rounding)
    num = torch.tensor(tdt, dtype=rho.dtype, device=rho.device)
    return torch.stack([
        num / half_cell_adjoint(rho, "x"),
        num / half_cell_adjoint(rho, "y"),
        num / half_cell_adjoint(rho, "z"),
        tdt * (lam + 2.0 * mu),
        tdt * lam,
        half_cell_adjoint(t, "z") * tdt,
        half_cell_adjoint(t, "y") * tdt,
        half_cell_adjoint(s, "z") * tdt,
    ])


def medium_gradient(g8, med8, dtw: float):
    """Chain rule through scale_medium (ela_medium_gradient): g8 = dJ/d(the
    8 scaled arrays) at the medium med8 -> dJ/d(rho, lambda, mu) stacked on a
    new leading dimension."""
    import torch
    tdt = 2.0 * dtw
    grho = torch.zeros_like(g8[0])
    for k, ax in enumerate(("x", "y", "z")):
        b = med8[k]
        grho = grho + half_cell_forward(-b * b / tdt * g8[k], ax)
    glam = tdt * (g8[3] + g8[4])
    gmu = 2.0 * tdt * g8[3]
    for k, (a0, a1) in enumerate((("x", "z"), ("x", "y"), ("y", "z"))):
        gmu = gmu + half_cell_forward(half_cell_forward(tdt * g8[5 + k], a1), a0)
    return torch.stack([grho, glam, gmu])
