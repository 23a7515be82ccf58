"""Binsparse files in the NumPy-directory form (``*.bspnpy``): the on-disk and
interchange format of Finch's ``bspread`` / ``bspwrite``
(src/interface/fileio/binsparse.jl; directory layout ext/NPZExt.jl).

A ``.bspnpy`` directory holds ``binsparse.json`` (``{"binsparse": desc}``)
and one ``.npy`` array per key.  Indices and pointers are 0-based; that is
the backend's ``index_base=0`` layout, so ``read_compressed`` hands the
arrays to the kernels unchanged (the PlusOneVector view of
src/util/vectors.jl:3-37).

Formats (binsparse.jl:93-176, as Finch writes them):

* ``CSC``: a Finch ``Tensor(Dense(SparseList(Element)))``: fibres are
  columns; ``pointers_to_1`` / ``indices_1`` hold ptr and row indices.
* ``CSR``: ``swizzle(Tensor(Dense(SparseList(Element))), 2, 1)``: the same
  buffers read row-wise.
* ``DVEC``: a dense vector ``Tensor(Dense(Element))``.
* ``COO`` (read only): ``indices_0`` = rows, ``indices_1`` = columns; the
  entries are assembled into compressed fibres on the device
  (fgx_coo_to_compressed).

Every file also carries ``values`` and ``fill_value``, the format name,
``shape``, ``number_of_stored_values``, ``version`` "0.1" and ``data_types``.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _lib as L
from . import ops
from .tensors import GPU, Dense, Element, SparseList, SwizzleArray, Tensor, transfer

VERSION = "0.1"
_TYPES = {"uint8": np.uint8, "uint16": np.uint16, "uint32": np.uint32, "uint64": np.uint64,
          "int8": np.int8, "int16": np.int16, "int32": np.int32, "int64": np.int64,
          "float32": np.float32, "float64": np.float64, "bint8": np.bool_}
_NAMES = {np.dtype(v): k for k, v in _TYPES.items()}

_LEVELS = {
    "CSR": {"level": {"level_desc": "dense", "rank": 1,
                      "level": {"level_desc": "sparse", "rank": 1,
                                "level": {"level_desc": "element"}}}},
    "CSC": {"transpose": [1, 0],
            "level": {"level_desc": "dense", "rank": 1,
                      "level": {"level_desc": "sparse", "rank": 1,
                                "level": {"level_desc": "element"}}}},
    "DVEC": {"level": {"level_desc": "dense", "rank": 1, "level": {"level_desc": "element"}}},
}


def _host(a) -> np.ndarray:
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _type_name(a: np.ndarray) -> str:
    try:
        return _NAMES[a.dtype]
    except KeyError as e:
        raise L.UnsupportedError(L.EUNSUP, f"binsparse cannot hold {a.dtype}") from e


def _check_path(path: str):
    if not path.endswith(".bspnpy"):
        raise L.UnsupportedError(L.EUNSUP, f"{path}: only the .bspnpy directory form is supported")


def bspwrite(path: str, t, attrs: dict | None = None) -> str:
    """Writes a CSC tensor, a swizzled CSR view or a dense vector (host or device)."""
    _check_path(path)
    if isinstance(t, SwizzleArray):
        if t.dims != (2, 1) or t.body.format != "csc":
            raise L.UnsupportedError(L.EUNSUP, "only swizzle(A, 2, 1) of a CSC tensor is CSR")
        fmt, body, shape = "CSR", t.body, list(t.shape)
    elif isinstance(t, Tensor) and t.format == "csc":
        fmt, body, shape = "CSC", t, list(t.shape)
    elif isinstance(t, Tensor) and t.format == "dense":
        fmt, body, shape = "DVEC", t, list(t.shape)
    else:
        raise L.UnsupportedError(L.EUNSUP, f"binsparse writing of {t!r} is not supported")
    arrays = {}
    el = body.element
    arrays["values"] = _host(el.val)
    arrays["fill_value"] = np.array([el.fill], dtype=arrays["values"].dtype)
    if fmt != "DVEC":
        sl = body.lvl.lvl
        arrays["pointers_to_1"] = _host(sl.ptr).astype(np.int64) - 1
        arrays["indices_1"] = _host(sl.idx).astype(np.int64) - 1
    desc = {"format": fmt, "fill": True, "shape": [int(x) for x in shape],
            "data_types": {k: _type_name(v) for k, v in arrays.items()},
            "version": VERSION, "number_of_stored_values": int(arrays["values"].size),
            "attrs": attrs or {}}
    os.makedirs(path, exist_ok=True)
    for k, v in arrays.items():
        np.save(os.path.join(path, f"{k}.npy"), v, allow_pickle=False)
    with open(os.path.join(path, "binsparse.json"), "w") as f:
        json.dump({"binsparse": desc}, f, indent=4)
    return path


def _read(path: str):
    _check_path(path)
    with open(os.path.join(path, "binsparse.json")) as f:
        desc = json.load(f)["binsparse"]
    if desc.get("version") != VERSION:
        raise L.UnsupportedError(L.EUNSUP, f"binsparse version {desc.get('version')!r}")
    fmt = desc.get("format")
    if fmt is None:
        custom = desc.get("custom")
        fmt = next((k for k, v in _LEVELS.items() if v == custom), None)
    if fmt not in ("CSR", "CSC", "DVEC", "COO", "COOR"):
        raise L.UnsupportedError(L.EUNSUP, f"binsparse format {fmt!r} is not supported")

    def arr(key):
        a = np.load(os.path.join(path, f"{key}.npy"), allow_pickle=False)
        want = _TYPES[desc["data_types"][key]]
        return a.astype(want, copy=False)

    values = arr("values")
    fill = arr("fill_value")[0] if os.path.exists(os.path.join(path, "fill_value.npy")) \
        else values.dtype.type(0)
    return fmt, desc, arr, values, fill


def _index_array(a: np.ndarray) -> np.ndarray:
    """Binsparse index arrays as the backend's signed Int32 / Int64."""
    if a.dtype in (np.dtype(np.int32), np.dtype(np.int64)):
        return a
    return a.astype(np.int64 if a.dtype.itemsize > 4 or a.dtype == np.uint32 else np.int32)


def read_compressed(path: str, device="cuda"):
    """(kind, ops.Compressed) with the file's 0-based arrays on `device`
    unchanged (index_base 0).  kind is "csr" (fibres are rows) or "csc"."""
    fmt, desc, arr, values, fill = _read(path)
    m, n = (int(x) for x in desc["shape"])
    up = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(device)
    if fmt in ("CSR", "CSC"):
        ptr, idx = _index_array(arr("pointers_to_1")), _index_array(arr("indices_1"))
        n_outer, n_inner = (m, n) if fmt == "CSR" else (n, m)
        if idx.dtype != ptr.dtype and ptr.dtype == np.dtype(np.int32):
            ptr = ptr.astype(np.int64)
        return fmt.lower(), ops.Compressed(up(ptr), up(idx), up(values), n_outer, n_inner, 0)
    if fmt in ("COO", "COOR"):
        rows, cols = _index_array(arr("indices_0")), _index_array(arr("indices_1"))
        t = np.promote_types(rows.dtype, cols.dtype)
        C = ops.coo_to_compressed(up(rows.astype(t)), up(cols.astype(t)), up(values), m, n, 0)
        return "csr", C
    raise L.UnsupportedError(L.EUNSUP, f"{fmt} is not a compressed matrix")


def bspread(path: str, device: GPU | None = None):
    """The file as Finch's mirror types: a CSC Tensor, a swizzled CSR view or a
    dense vector; on `device` if given (COO files need a device: they are
    assembled there)."""
    fmt, desc, arr, values, fill = _read(path)
    if fmt == "DVEC":
        t = Tensor(Dense(Element(fill), int(desc["shape"][0])))
        t.element.val = values
        return transfer(device, t) if device is not None else t
    if fmt in ("COO", "COOR"):
        if device is None:
            raise L.UnsupportedError(L.EUNSUP, "COO files are assembled on a device: pass device=")
        _, C = read_compressed(path, device.torch_device)
        t = Tensor.from_buffers(C.ptr, C.idx, C.val, C.n_inner, C.n_outer, fill, base=0)
        return SwizzleArray(t, (2, 1))
    m, n = (int(x) for x in desc["shape"])
    ptr, idx = arr("pointers_to_1"), arr("indices_1")
    if fmt == "CSC":
        t = Tensor.from_buffers(ptr.astype(np.int64), idx.astype(np.int64), values, m, n, fill, base=0)
        return transfer(device, t) if device is not None else t
    t = Tensor.from_buffers(ptr.astype(np.int64), idx.astype(np.int64), values, n, m, fill, base=0)
    out = SwizzleArray(t, (2, 1))
    return transfer(device, out) if device is not None else out
