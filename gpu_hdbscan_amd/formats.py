"""Reference on-disk formats (tools/filetool.hpp:95-242).

Every row is an int32 count followed by that many payload items:
  .fvecs   d x float32 (points)
  .ivecs   d x int32 (ids, labels)
  .kgraph  K x NNDElement {float32 distance, int32 label} (kNN lists)
Readers check that all rows have the same count.
"""
from __future__ import annotations

import numpy as np

from ._lib import NND_ELEMENT


def _write(path, arr: np.ndarray) -> None:
    arr = np.ascontiguousarray(arr)
    n, d = arr.shape
    rows = np.empty((n,), dtype=[("d", "<i4"), ("v", arr.dtype, (d,))])
    rows["d"] = d
    rows["v"] = arr
    rows.tofile(path)


def _read(path, item: np.dtype) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size == 0:
        return np.empty((0, 0), item)
    d = int(raw[:4].view("<i4")[0])
    if d <= 0:
        raise ValueError(f"{path}: bad row width {d}")
    row = 4 + d * item.itemsize
    if raw.size % row:
        raise ValueError(f"{path}: size {raw.size} is not a multiple of the row size {row}")
    rows = raw.view(np.dtype([("d", "<i4"), ("v", item, (d,))]))
    if np.any(rows["d"] != d):
        raise ValueError(f"{path}: ragged rows")
    return np.ascontiguousarray(rows["v"])


def write_fvecs(path, X) -> None:
    _write(path, np.asarray(X, np.float32))


def read_fvecs(path) -> np.ndarray:
    return _read(path, np.dtype("<f4"))


def write_ivecs(path, A) -> None:
    _write(path, np.asarray(A, np.int32))


def read_ivecs(path) -> np.ndarray:
    return _read(path, np.dtype("<i4"))


def write_kgraph(path, lists) -> None:
    _write(path, np.asarray(lists, NND_ELEMENT))


def read_kgraph(path) -> np.ndarray:
    return _read(path, NND_ELEMENT)
