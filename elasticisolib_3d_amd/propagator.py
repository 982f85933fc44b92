"""Regular-grid propagator: a Python handle over the C ABI.

This is the counterpart of the reference's GPU layer
(elastic_iso_float/src/nonlinearPropElasticGpuFunctions_3D.cu,
elastic_iso_float/src/BornElasticGpuFunctions_3D.cu): it propagates traces
that already live on the regular FD grids. Arrays may be numpy arrays (host
memory, synchronous) or torch tensors on the GPU (device memory, queued on
torch's current stream).

Layouts (shared with the reference):
  - medium:    [8, ny, nx, nz]  rhoxDtw rhoyDtw rhozDtw lamb2MuDtw lambDtw muxzDtw muxyDtw muyzDtw
  - positions: 7 int64 arrays   grids centre X Y Z XZ XY YZ (flat [ny][nx][nz] indices)
  - traces:    9 arrays [nReg(grid of component), nts], components vx vy vz sxx syy szz sxz sxy syz
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import ElaConfig, ElaPoints, ElaTraces, check

COMP_GRID = (1, 2, 3, 0, 0, 0, 4, 5, 6)
COMPONENTS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxz", "sxy", "syz")
GRIDS = ("centre", "x", "y", "z", "xz", "xy", "yz")


def _is_torch(x) -> bool:
    return hasattr(x, "data_ptr") and hasattr(x, "is_cuda")


CHECKPOINT_FORMATS = {"native": 0, "bf16": 1, "s16": 2}   # ELA_CK_NATIVE, ELA_CK_BF16, ELA_CK_S16
CHECKPOINT_NAMES = {v: k for k, v in CHECKPOINT_FORMATS.items()}


def checkpoint_format(checkpoint) -> int:
    """ELA_CK_* code of a checkpoint format name (True / False: "bf16" / "native")."""
    if isinstance(checkpoint, (bool, np.bool_)):
        return CHECKPOINT_FORMATS["bf16" if checkpoint else "native"]
    if checkpoint in CHECKPOINT_FORMATS:
        return CHECKPOINT_FORMATS[checkpoint]
    raise ValueError(f"checkpoint format must be one of {sorted(CHECKPOINT_FORMATS)}, got {checkpoint!r}")


class Propagator:
    """One propagator handle (one grid, one time axis, one device)."""

    def __init__(self, nz: int, nx: int, ny: int, nts: int, sub: int, dz: float, dx: float, dy: float,
                 dts: float, min_pad: int, alpha_cos: float = 0.99, precision: str = "f32", device: int = 0):
        if precision not in ("f32", "f64"):
            raise ValueError("precision must be 'f32' or 'f64'")
        self.lib = _lib.load()
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.precision = precision
        self.shape = (ny, nx, nz)
        self.nts = int(nts)
        self.sub = int(sub)
        self.cfg = ElaConfig(nz, nx, ny, nts, sub, dz, dx, dy, dts, min_pad, alpha_cos,
                             _lib.ELA_FP32 if precision == "f32" else _lib.ELA_FP64, device)
        h = ctypes.c_void_p()
        check(self.lib.ela_create(ctypes.byref(self.cfg), ctypes.byref(h)), "ela_create")
        self.h = h
        self._keep = []
        self.medium_kind = None   # "scaled" (set_medium) or "elastic" (set_elastic)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ela_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    def _mem_of(self, *xs) -> int:
        kinds = {_is_torch(x) and x.is_cuda for x in xs if x is not None}
        if len(kinds) > 1:
            raise ValueError("mix of host and device arrays in one call")
        dev = kinds.pop() if kinds else False
        if dev:
            import torch
            self.lib.ela_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
            return _lib.ELA_MEM_DEVICE
        self.lib.ela_set_stream(self.h, None)
        return _lib.ELA_MEM_HOST

    def _ptr(self, x, dtype=None):
        if _is_torch(x):
            if not x.is_contiguous():
                raise ValueError("device arrays must be contiguous")
            if dtype is not None:
                want = {np.float32: "torch.float32", np.float64: "torch.float64", np.int64: "torch.int64"}[dtype]
                if str(x.dtype) != want:
                    raise ValueError(f"expected {want}, got {x.dtype}")
            return ctypes.c_void_p(x.data_ptr())
        if not (x.flags["C_CONTIGUOUS"] and (dtype is None or x.dtype == dtype)):
            raise ValueError("host arrays must be C-contiguous with the handle dtype")
        return ctypes.c_void_p(x.ctypes.data)

    def _points(self, pos: Sequence, mem: int) -> ElaPoints:
        """7 position lists in the memory kind of the call: host arrays are
        moved to the GPU for a device-memory call and device tensors to the
        host for a host-memory call (the library reads device-call positions
        on the GPU)."""
        if len(pos) != 7:
            raise ValueError("need 7 position lists (centre X Y Z XZ XY YZ)")
        p = ElaPoints()
        for g in range(7):
            a = pos[g]
            if mem == _lib.ELA_MEM_DEVICE:
                if not (_is_torch(a) and a.is_cuda):
                    import torch
                    a = torch.as_tensor(np.ascontiguousarray(a, dtype=np.int64), device="cuda")
                elif not a.is_contiguous():
                    a = a.contiguous()
            else:
                if _is_torch(a):
                    a = a.detach().cpu().numpy()
                a = np.ascontiguousarray(a, dtype=np.int64)
            self._keep.append(a)
            n = a.shape[0] if len(a.shape) else 0
            p.n[g] = int(n)
            p.pos[g] = self._ptr(a, np.int64).value if n else None
        return p

    def _traces(self, tr: Sequence, counts) -> ElaTraces:
        if len(tr) != 9:
            raise ValueError("need 9 trace arrays")
        t = ElaTraces()
        for c in range(9):
            a = tr[c]
            n = counts[COMP_GRID[c]]
            if tuple(a.shape) != (n, self.nts):
                raise ValueError(f"component {COMPONENTS[c]}: expected shape {(n, self.nts)}, got {tuple(a.shape)}")
            t.tr[c] = self._ptr(a, self.dtype).value if n else None
        return t

    def _alloc_traces(self, counts, like_device: bool):
        if like_device:
            import torch
            td = torch.float32 if self.dtype == np.float32 else torch.float64
            return [torch.zeros((counts[COMP_GRID[c]], self.nts), dtype=td, device="cuda") for c in range(9)]
        return [np.zeros((counts[COMP_GRID[c]], self.nts), dtype=self.dtype) for c in range(9)]

    @staticmethod
    def _counts(pos):
        return [int(p.shape[0]) for p in pos]

    # ------------------------------------------------------------------
    def set_medium(self, med8) -> None:
        if tuple(med8.shape) != (8,) + self.shape:
            raise ValueError(f"medium must have shape {(8,) + self.shape}")
        mem = self._mem_of(med8)
        if mem == _lib.ELA_MEM_HOST:
            med8 = np.ascontiguousarray(med8, dtype=self.dtype)
        check(self.lib.ela_set_medium(self.h, self._ptr(med8, self.dtype), mem), "ela_set_medium")
        self.medium_kind = "scaled"

    def set_elastic(self, rlm3, dtw: float) -> None:
        """Elastic parameters (rho, lambda, mu) [3, ny, nx, nz] instead of the
        8 scaled arrays (ela_set_elastic): the 8 arrays are derived on the GPU
        with the staggering and scaling of scale_medium(rho, lambda, mu, dtw),
        and the forward / nonlinear-adjoint steps form them on the fly."""
        if tuple(rlm3.shape) != (3,) + self.shape:
            raise ValueError(f"elastic parameters must have shape {(3,) + self.shape}")
        mem = self._mem_of(rlm3)
        if mem == _lib.ELA_MEM_HOST:
            rlm3 = np.ascontiguousarray(rlm3, dtype=self.dtype)
        check(self.lib.ela_set_elastic(self.h, self._ptr(rlm3, self.dtype), float(dtw), mem), "ela_set_elastic")
        self.medium_kind = "elastic"

    def get_medium(self) -> np.ndarray:
        """The handle's 8 scaled medium arrays [8, ny, nx, nz] (those set, or
        those set_elastic derived on the GPU)."""
        out = np.empty((8,) + self.shape, dtype=self.dtype)
        check(self.lib.ela_get_medium(self.h, out.ctypes.data, _lib.ELA_MEM_HOST), "ela_get_medium")
        return out

    def nl_forward(self, src_pos, src_traces, rec_pos, rec_out=None):
        mem = self._mem_of(*src_traces, *(rec_out or []))
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        rc = self._counts(rec_pos)
        if mem == _lib.ELA_MEM_HOST:
            src_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in src_traces]
        rec = rec_out if rec_out is not None else self._alloc_traces(rc, mem == _lib.ELA_MEM_DEVICE)
        st, rt = self._traces(src_traces, self._counts(src_pos)), self._traces(rec, rc)
        check(self.lib.ela_nl_forward(self.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(rp), ctypes.byref(rt), mem),
              "ela_nl_forward")
        self._keep.clear()
        return rec

    def nl_adjoint(self, src_pos, rec_traces, rec_pos, src_out=None):
        mem = self._mem_of(*rec_traces, *(src_out or []))
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        sc = self._counts(src_pos)
        if mem == _lib.ELA_MEM_HOST:
            rec_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in rec_traces]
        src = src_out if src_out is not None else self._alloc_traces(sc, mem == _lib.ELA_MEM_DEVICE)
        st, rt = self._traces(src, sc), self._traces(rec_traces, self._counts(rec_pos))
        check(self.lib.ela_nl_adjoint(self.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(rp), ctypes.byref(rt), mem),
              "ela_nl_adjoint")
        self._keep.clear()
        return src

    def born_forward(self, src_pos, src_traces, dmod8, rec_pos, rec_out=None):
        mem = self._mem_of(*src_traces, dmod8)
        if mem == _lib.ELA_MEM_HOST:
            src_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in src_traces]
            dmod8 = np.ascontiguousarray(dmod8, dtype=self.dtype)
        if tuple(dmod8.shape) != (8,) + self.shape:
            raise ValueError("perturbation must have shape (8, ny, nx, nz)")
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        rc = self._counts(rec_pos)
        rec = rec_out if rec_out is not None else self._alloc_traces(rc, mem == _lib.ELA_MEM_DEVICE)
        st, rt = self._traces(src_traces, self._counts(src_pos)), self._traces(rec, rc)
        check(self.lib.ela_born_forward(self.h, ctypes.byref(sp), ctypes.byref(st), self._ptr(dmod8, self.dtype),
                                        ctypes.byref(rp), ctypes.byref(rt), mem), "ela_born_forward")
        self._keep.clear()
        return rec

    def born_adjoint(self, src_pos, src_traces, rec_traces, rec_pos, dmod8=None):
        """Returns dmod8 with the image ADDED (allocated zero if None)."""
        mem = self._mem_of(*src_traces, *rec_traces, dmod8)
        if mem == _lib.ELA_MEM_HOST:
            src_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in src_traces]
            rec_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in rec_traces]
            if dmod8 is None:
                dmod8 = np.zeros((8,) + self.shape, self.dtype)
        elif dmod8 is None:
            import torch
            dmod8 = torch.zeros((8,) + self.shape, dtype=src_traces[0].dtype, device="cuda")
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        st, rt = self._traces(src_traces, self._counts(src_pos)), self._traces(rec_traces, self._counts(rec_pos))
        check(self.lib.ela_born_adjoint(self.h, ctypes.byref(sp), ctypes.byref(st), self._ptr(dmod8, self.dtype),
                                        ctypes.byref(rp), ctypes.byref(rt), mem), "ela_born_adjoint")
        self._keep.clear()
        return dmod8

    def gradient_forward(self, src_pos, src_traces, rec_pos, rec_out=None):
        """Nonlinear forward that keeps checkpoints for gradient_backward."""
        mem = self._mem_of(*src_traces, *(rec_out or []))
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        rc = self._counts(rec_pos)
        if mem == _lib.ELA_MEM_HOST:
            src_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in src_traces]
        rec = rec_out if rec_out is not None else self._alloc_traces(rc, mem == _lib.ELA_MEM_DEVICE)
        st, rt = self._traces(src_traces, self._counts(src_pos)), self._traces(rec, rc)
        check(self.lib.ela_gradient_forward(self.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(rp),
                                            ctypes.byref(rt), mem), "ela_gradient_forward")
        self._keep.clear()
        return rec

    def gradient_backward(self, src_pos, src_traces, rec_pos, adj_traces, grad8=None, src_grad_out=None):
        """Adjoint propagation of dJ/d(receiver traces). Returns (grad8, src_grad):
        grad8 (FP64 [8, ny, nx, nz], accumulated; allocated zero if None) and
        dJ/d(scaled source traces)."""
        mem = self._mem_of(*src_traces, *adj_traces, grad8)
        if mem == _lib.ELA_MEM_HOST:
            src_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in src_traces]
            adj_traces = [np.ascontiguousarray(s, dtype=self.dtype) for s in adj_traces]
            if grad8 is None:
                grad8 = np.zeros((8,) + self.shape, np.float64)
            if not (grad8.dtype == np.float64 and grad8.flags["C_CONTIGUOUS"]):
                raise ValueError("grad8 must be a C-contiguous float64 array")
        elif grad8 is None:
            import torch
            grad8 = torch.zeros((8,) + self.shape, dtype=torch.float64, device="cuda")
        if tuple(grad8.shape) != (8,) + self.shape:
            raise ValueError("grad8 must have shape (8, ny, nx, nz)")
        sp, rp = self._points(src_pos, mem), self._points(rec_pos, mem)
        sc = self._counts(src_pos)
        gs = src_grad_out if src_grad_out is not None else self._alloc_traces(sc, mem == _lib.ELA_MEM_DEVICE)
        st, at, gt = self._traces(src_traces, sc), self._traces(adj_traces, self._counts(rec_pos)), self._traces(gs, sc)
        gptr = ctypes.c_void_p(grad8.data_ptr()) if _is_torch(grad8) else ctypes.c_void_p(grad8.ctypes.data)
        check(self.lib.ela_gradient_backward(self.h, ctypes.byref(sp), ctypes.byref(st), ctypes.byref(rp),
                                             ctypes.byref(at), gptr, ctypes.byref(gt), mem), "ela_gradient_backward")
        self._keep.clear()
        return grad8, gs

    def set_gradient_checkpoints(self, interval: int = 20, outer_interval: int = 0, checkpoint="s16",
                                 budget_bytes: int = 0) -> None:
        """`checkpoint`: storage of the outer checkpoints, "s16" (int16, one
        scale per 64 values; the default, the format that keeps the FP32
        gradient inside 1e-3 of FP64), "bf16" (bfloat16, one power-of-two
        scale per array and snapshot) or "native" (exact); True / False mean
        "bf16" / "native"."""
        check(self.lib.ela_set_gradient_checkpoints(self.h, int(interval), int(outer_interval),
                                                    checkpoint_format(checkpoint), int(budget_bytes)),
              "ela_set_gradient_checkpoints")

    def set_gradient_derivatives(self, fmt="s16") -> None:
        """Storage of the derivative combinations the gradient's backward
        recomputes and images with: "native" (exact, the default) or "s16"
        (FP32 only: int16 with one scale per 32 values, half the memory of
        the derivative buffers, a slightly slower backward). Takes effect
        from the next gradient forward."""
        check(self.lib.ela_set_gradient_derivatives(self.h, checkpoint_format(fmt)), "ela_set_gradient_derivatives")

    def gradient_derivatives_used(self) -> str:
        """"s16" or "native": the derivative storage of the next gradient."""
        f = ctypes.c_int()
        check(self.lib.ela_gradient_derivatives_used(self.h, ctypes.byref(f)), "ela_gradient_derivatives_used")
        return CHECKPOINT_NAMES[f.value]

    def gradient_plan_used(self):
        """(interval, outer_interval, checkpoint format name, checkpoint bytes) of the last gradient forward."""
        k, k1, b, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int64()
        check(self.lib.ela_gradient_plan_used(self.h, ctypes.byref(k), ctypes.byref(k1), ctypes.byref(b),
                                              ctypes.byref(nb)), "ela_gradient_plan_used")
        return k.value, k1.value, CHECKPOINT_NAMES[b.value], nb.value

    def set_windows(self, y0, ny, nx) -> None:
        """Stack windows along y (see ela_set_windows); empty lists remove them."""
        y0 = np.ascontiguousarray(y0, np.int64)
        ny = np.ascontiguousarray(ny, np.int64)
        nx = np.ascontiguousarray(nx, np.int64)
        if not (len(y0) == len(ny) == len(nx)):
            raise ValueError("y0, ny, nx must have the same length")
        check(self.lib.ela_set_windows(self.h, len(y0), y0.ctypes.data, ny.ctypes.data, nx.ctypes.data),
              "ela_set_windows")

    def set_graph(self, on: bool = True) -> None:
        """Replay the time loops as HIP graphs: the coarse step of nl_forward /
        nl_adjoint, the gradient's pairs of forward steps and its checkpoint
        segments. Each graph is captured once and replayed while the launches
        it holds (kernels, shapes, arguments) stay the same."""
        check(self.lib.ela_set_graph(self.h, int(bool(on))), "ela_set_graph")

    def set_tiling(self, tile_x: int, y_chunk: int) -> None:
        check(self.lib.ela_set_tiling(self.h, int(tile_x), int(y_chunk)), "ela_set_tiling")

    def set_tile_order(self, order: int, x_group: int = 1) -> None:
        """0 hardware order, 1 contiguous per XCD, 2 y-major per XCD in groups
        of x_group columns, 3 hardware order with groups of x_group whole z
        columns per XCD (speed only)."""
        check(self.lib.ela_set_tile_order(self.h, int(order), int(x_group)), "ela_set_tile_order")

    def set_born_storage(self, segment: int = 0, budget_bytes: int = 0) -> None:
        """Coarse-wavefield storage of the Born adjoint: 0 automatic (within
        budget_bytes, 0 = free device memory), -1 every coarse sample kept,
        K > 0 segments of K coarse samples counted back from the end of the
        record, exact checkpoints at the start of all but the last, whose
        samples the forward sweep keeps; the others are recomputed."""
        check(self.lib.ela_set_born_storage(self.h, int(segment), int(budget_bytes)), "ela_set_born_storage")

    def set_born_mode(self, physics: bool) -> None:
        """Yz shear term of the Born operators: False (default) keeps the
        reference's xz derivative pattern (kernelsElaGpu_3D.cu:1419-1431,
        :1622-1634); True uses the yz pattern, the true linearisation in mu."""
        check(self.lib.ela_set_born_mode(self.h, 1 if physics else 0), "ela_set_born_mode")

    def born_storage_used(self) -> int:
        """Plan of the last Born adjoint: -1 all samples kept, else the segment length."""
        k = ctypes.c_int64()
        check(self.lib.ela_born_storage_used(self.h, ctypes.byref(k)), "ela_born_storage_used")
        return k.value

    def get_tiling(self):
        tx, yc = ctypes.c_int(), ctypes.c_int()
        check(self.lib.ela_get_tiling(self.h, ctypes.byref(tx), ctypes.byref(yc)), "ela_get_tiling")
        return tx.value, yc.value

    def set_tiling_transposed(self, tile_x: int = 0) -> None:
        """x tile of the premultiplied transposed steps (Born adjoint,
        gradient backward); 0 uses the forward tile (speed only)."""
        check(self.lib.ela_set_tiling_transposed(self.h, int(tile_x)), "ela_set_tiling_transposed")

    def get_tiling_transposed(self) -> int:
        tx = ctypes.c_int()
        check(self.lib.ela_get_tiling_transposed(self.h, ctypes.byref(tx)), "ela_get_tiling_transposed")
        return tx.value

    def time_step_kernels(self, reps: int = 50):
        """Mean launch duration [ms] of the velocity and stress stencil kernels
        (HIP events on the handle's stream; overwrites the wavefields)."""
        mv, ms = ctypes.c_double(), ctypes.c_double()
        self._mem_of()
        check(self.lib.ela_time_step_kernels(self.h, int(reps), ctypes.byref(mv), ctypes.byref(ms)),
              "ela_time_step_kernels")
        return mv.value, ms.value

    def time_gradient_kernels(self, reps: int = 10):
        """Mean launch durations [ms] of the forward step, the recomputing
        step with derivative store and the transposed step with imaging of
        the adjoint-state gradient (ela_time_gradient_kernels; overwrites the
        wavefields, keeps stored checkpoints)."""
        ms = (ctypes.c_double * 3)()
        self._mem_of()
        check(self.lib.ela_time_gradient_kernels(self.h, int(reps), ms), "ela_time_gradient_kernels")
        return ms[0], ms[1], ms[2]

    def stream_probe(self, reads: int, writes: int, n: int, reps: int = 20) -> float:
        """GB/s of a plain streaming kernel with the given read / in-place
        write mix over n FP32 values (ela_stream_probe)."""
        g = ctypes.c_double()
        check(self.lib.ela_stream_probe(self.h, int(reads), int(writes), int(n), int(reps), ctypes.byref(g)),
              "ela_stream_probe")
        return g.value

    def synchronize(self):
        check(self.lib.ela_synchronize(self.h), "ela_synchronize")

    def last_timing(self):
        ms = ctypes.c_double()
        steps = ctypes.c_int64()
        cells = ctypes.c_int64()
        check(self.lib.ela_last_timing(self.h, ctypes.byref(ms), ctypes.byref(steps), ctypes.byref(cells)),
              "ela_last_timing")
        return ms.value, steps.value, cells.value


def scale_medium(rho, lame, mu, dtw):
    """Stagger and scale (rho, lambda, mu) [ny, nx, nz] into the 8 propagation
    arrays, through the library's host routine (fdParamElastic_3D.cpp:130-183)."""
    lib = _lib.load()
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    lame = np.ascontiguousarray(lame, dtype=np.float64)
    mu = np.ascontiguousarray(mu, dtype=np.float64)
    ny, nx, nz = rho.shape
    out = np.empty((8, ny, nx, nz), np.float64)
    check(lib.ela_scale_medium(nz, nx, ny, float(dtw), rho.ctypes.data, lame.ctypes.data, mu.ctypes.data,
                               out.ctypes.data), "ela_scale_medium")
    return out


def _host3(name, dtw, med8, a, n_out):
    lib = _lib.load()
    med8 = np.ascontiguousarray(med8, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    ny, nx, nz = med8.shape[1:]
    out = np.empty((n_out, ny, nx, nz), np.float64)
    check(getattr(lib, name)(nz, nx, ny, float(dtw), med8.ctypes.data, a.ctypes.data, out.ctypes.data), name)
    return out


def born_stagger(dm3, med8, dtw):
    """(drho, dlambda, dmu) [3, ny, nx, nz] -> the 8 staggered, scaled
    perturbations of born_forward (BornElasticShotsGpu_3D.cpp:167-235)."""
    return _host3("ela_born_stagger", dtw, med8, dm3, 8)


def born_unstagger(img8, med8, dtw):
    """Adjoint of born_stagger (BornElasticShotsGpu_3D.cpp:541-607)."""
    return _host3("ela_born_unstagger", dtw, med8, img8, 3)


def medium_gradient(g8, med8, dtw):
    """dJ/d(8 scaled arrays) -> dJ/d(rho, lambda, mu) through scale_medium."""
    return _host3("ela_medium_gradient", dtw, med8, g8, 3)
