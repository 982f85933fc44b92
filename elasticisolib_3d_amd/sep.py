"""Parameter files, SEP-style headers and numpy-backed vectors.

The reference's Python layer uses SEP's genericIO / SepVector / Hypercube
(not part of the reference repository) for three things, restated here:

* parameters: `genericIO.io(params=sys.argv)` -- `key=value` words from the
  command line and from `par=file` parameter files, read with getInt,
  getFloat, getString, getInts (python_float/Elastic_iso_float_3D.py:24-33,
  400-430);
* vectors on a hypercube: axis 1 is the fastest, so a vector with axes
  (n1, n2, ..., nN) is a numpy array of shape (nN, ..., n2, n1); `getNdArray`,
  `zero`, `clone`, `cloneSpace`, `scale`, `dot`, `rand`, `getHyper`;
* files: a text header with `n1= o1= d1= ... in=<binary file>` and raw
  samples (`esize=4` or 8; `data_format="native_float"` little endian or
  `"xdr_float"` big endian), as genericIO's default SEP IO writes and reads.

Only the behaviour the elastic operators rely on is implemented. genericIO's
exact reader/writer corner cases (SEP history files, `-` continuation lines,
stdin data) are not reproduced.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np


# ---------------------------------------------------------------------------
# Parameters
# ---------------------------------------------------------------------------

_KV = re.compile(r"""([A-Za-z_][A-Za-z0-9_.]*)=("[^"]*"|'[^']*'|\S*)""")


def _parse_kv_text(text: str) -> Dict[str, str]:
    out: Dict[str, str] = {}
    for line in text.splitlines():
        line = line.split("#", 1)[0]
        for k, v in _KV.findall(line):
            if len(v) >= 2 and v[0] == v[-1] and v[0] in "\"'":
                v = v[1:-1]
            out[k] = v
    return out


class _Missing:
    pass


_MISSING = _Missing()


def _to_int(s: str) -> int:
    try:
        return int(s)
    except ValueError:
        f = float(s)
        if f != int(f):
            raise
        return int(f)


class ParObject:
    """Parameters from `key=value` words and `par=<file>` files; later words
    override earlier ones (genericIO.io(params=sys.argv)). A leading word
    without '=' (the program name) is skipped."""

    def __init__(self, params: Optional[Sequence[str]] = None, **kw):
        self._p: Dict[str, str] = {}
        words = [str(w) for w in (params or [])]
        if words and "=" not in words[0]:
            words = words[1:]
        for w in words:
            if w.startswith("par="):
                with open(w[4:].strip("\"'")) as f:
                    self._p.update(_parse_kv_text(f.read()))
            elif "=" in w:
                self._p.update(_parse_kv_text(w))
        for k, v in kw.items():
            self._p[k] = str(v)
        self.param = self

    def set(self, key: str, value) -> None:
        self._p[key] = str(value)

    def has(self, key: str) -> bool:
        return key in self._p

    def _get(self, key, default, conv, kind):
        if key in self._p:
            try:
                return conv(self._p[key])
            except ValueError:
                raise ValueError(f"parameter {key}={self._p[key]!r} is not a valid {kind}") from None
        if default is _MISSING:
            raise KeyError(f"required parameter {key!r} ({kind}) not given")
        return default

    def getInt(self, key: str, default=_MISSING):
        return self._get(key, default, _to_int, "int")

    def getFloat(self, key: str, default=_MISSING):
        return self._get(key, default, float, "float")

    def getString(self, key: str, default=_MISSING):
        return self._get(key, default, str, "string")

    def getInts(self, key: str, default=_MISSING):
        return self._get(key, default, lambda s: [_to_int(x) for x in s.split(",") if x != ""], "int list")

    def getFloats(self, key: str, default=_MISSING):
        return self._get(key, default, lambda s: [float(x) for x in s.split(",") if x != ""], "float list")


# ---------------------------------------------------------------------------
# Axes, hypercubes, vectors
# ---------------------------------------------------------------------------

@dataclass
class Axis:
    n: int = 1
    o: float = 0.0
    d: float = 1.0
    label: str = ""

    def same(self, other: "Axis", tol: float = 1e-6) -> bool:
        return self.n == other.n and abs(self.o - other.o) <= tol and abs(self.d - other.d) <= tol


class Hypercube:
    """Axes fastest first (axis 1 = index 0)."""

    def __init__(self, axes: Optional[Iterable[Axis]] = None, ns: Optional[Sequence[int]] = None):
        if axes is None:
            axes = [Axis(n=int(n)) for n in (ns or [])]
        self.axes: List[Axis] = [Axis(a.n, a.o, a.d, a.label) for a in axes]

    def getAxis(self, i: int) -> Axis:
        """1-based, as SEP's hypercube."""
        return self.axes[i - 1]

    def getNdim(self) -> int:
        return len(self.axes)

    def getNs(self) -> List[int]:
        return [a.n for a in self.axes]

    @property
    def shape(self):
        return tuple(a.n for a in reversed(self.axes))

    def checkSame(self, other: "Hypercube") -> bool:
        return len(self.axes) == len(other.axes) and all(a.same(b) for a, b in zip(self.axes, other.axes))


class Vector:
    """A numpy array on a hypercube (SepVector's float/double vectors)."""

    def __init__(self, hyper: Hypercube, arr: Optional[np.ndarray] = None, dtype=np.float32):
        self.hyper = hyper
        if arr is None:
            arr = np.zeros(hyper.shape, dtype=dtype)
        arr = np.asarray(arr)
        if tuple(arr.shape) != hyper.shape:
            arr = arr.reshape(hyper.shape)
        self.arr = arr

    # SepVector-style accessors
    def getNdArray(self) -> np.ndarray:
        return self.arr

    def getHyper(self) -> Hypercube:
        return self.hyper

    def getCpp(self):
        return self

    @property
    def dtype(self):
        return self.arr.dtype

    def zero(self):
        self.arr[...] = 0
        return self

    def set(self, v):
        self.arr[...] = v
        return self

    def clone(self) -> "Vector":
        return Vector(self.hyper, self.arr.copy())

    def cloneSpace(self) -> "Vector":
        return Vector(self.hyper, np.zeros_like(self.arr))

    def checkSame(self, other: "Vector") -> bool:
        return self.hyper.checkSame(other.hyper)

    def scale(self, s):
        self.arr *= self.arr.dtype.type(s)
        return self

    def scaleAdd(self, other: "Vector", sc1=1.0, sc2=1.0):
        self.arr[...] = sc1 * self.arr + sc2 * other.arr
        return self

    def copy(self, other: "Vector"):
        self.arr[...] = other.arr
        return self

    def dot(self, other: "Vector") -> float:
        """Inner product accumulated in double precision."""
        return float(np.dot(self.arr.ravel().astype(np.float64), other.arr.ravel().astype(np.float64)))

    def norm(self, N: int = 2) -> float:
        return float(np.linalg.norm(self.arr.ravel().astype(np.float64), ord=N))

    def rand(self, seed: Optional[int] = None):
        """Uniform [-1, 1) samples (SEP's random(); its generator is not
        reproduced)."""
        rng = np.random.default_rng(seed)
        self.arr[...] = rng.uniform(-1.0, 1.0, self.arr.shape).astype(self.arr.dtype)
        return self

    random = rand


def getSepVector(hyper: Optional[Hypercube] = None, ns: Optional[Sequence[int]] = None,
                 storage: str = "dataFloat") -> Vector:
    if hyper is None:
        hyper = Hypercube(ns=ns)
    dt = {"dataFloat": np.float32, "dataDouble": np.float64}.get(storage)
    if dt is None:
        raise ValueError(f"unsupported storage {storage!r}")
    return Vector(hyper, dtype=dt)


# ---------------------------------------------------------------------------
# Files
# ---------------------------------------------------------------------------

def read_header(path: str) -> Dict[str, str]:
    with open(path) as f:
        return _parse_kv_text(f.read())


def read_vector(path: str, ndims: Optional[int] = None, dtype=None) -> Vector:
    """Read a SEP header + binary file. `ndims` pads the hypercube with unit
    axes (genericIO getVector(..., ndims=N))."""
    h = read_header(path)
    if "in" not in h:
        raise IOError(f"{path}: no in= entry")
    data_path = h["in"]
    if not os.path.isabs(data_path) and not os.path.exists(data_path):
        data_path = os.path.join(os.path.dirname(os.path.abspath(path)), data_path)
    axes = []
    i = 1
    while f"n{i}" in h:
        axes.append(Axis(int(h[f"n{i}"]), float(h.get(f"o{i}", 0.0)), float(h.get(f"d{i}", 1.0)),
                         h.get(f"label{i}", "")))
        i += 1
    if not axes:
        raise IOError(f"{path}: no n1= entry")
    if ndims is not None:
        if len(axes) > ndims:
            if any(a.n != 1 for a in axes[ndims:]):
                raise IOError(f"{path}: {len(axes)} axes do not fit in {ndims}")
            axes = axes[:ndims]
        while len(axes) < ndims:
            axes.append(Axis(1))
    esize = int(h.get("esize", 4))
    fmt = h.get("data_format", "native_float")
    if esize == 4:
        base = np.float32
    elif esize == 8:
        base = np.float64
    else:
        raise IOError(f"{path}: esize={esize} not supported")
    endian = ">" if fmt.startswith("xdr") else "<"
    hyper = Hypercube(axes)
    n = int(np.prod(hyper.shape))
    raw = np.fromfile(data_path, dtype=np.dtype(base).newbyteorder(endian), count=n)
    if raw.size != n:
        raise IOError(f"{data_path}: expected {n} samples, found {raw.size}")
    arr = raw.astype(dtype or base).reshape(hyper.shape)
    return Vector(hyper, arr)


def write_vector(path: str, vec: Vector, data_path: Optional[str] = None) -> None:
    """Write `vec` as a header at `path` and little-endian samples at
    `data_path` (default: path + "@")."""
    data_path = data_path or path + "@"
    arr = np.ascontiguousarray(vec.arr)
    esize = arr.dtype.itemsize
    arr.astype(arr.dtype.newbyteorder("<")).tofile(data_path)
    lines = []
    for i, a in enumerate(vec.hyper.axes, 1):
        lab = f' label{i}="{a.label}"' if a.label else ""
        lines.append(f"n{i}={a.n} o{i}={a.o!r} d{i}={a.d!r}{lab}")
    lines.append(f"esize={esize} data_format=\"native_{'float' if esize == 4 else 'double'}\"")
    lines.append(f'in="{os.path.abspath(data_path)}"')
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
