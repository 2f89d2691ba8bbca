"""Round trips of the reference's row formats (tools/filetool.hpp:95-242)."""
import numpy as np
import pytest

from gpu_hdbscan_amd import formats
from gpu_hdbscan_amd._lib import NND_ELEMENT


def test_fvecs_round_trip(tmp_path):
    X = np.random.default_rng(0).standard_normal((37, 12)).astype(np.float32)
    p = tmp_path / "x.fvecs"
    formats.write_fvecs(p, X)
    raw = np.fromfile(p, np.int32)
    assert raw[0] == 12 and raw.size == 37 * 13  # int32 dim + 12 floats per row
    np.testing.assert_array_equal(formats.read_fvecs(p), X)


def test_kgraph_and_ivecs_round_trip(tmp_path):
    L = np.empty((5, 32), NND_ELEMENT)
    L["distance"] = np.arange(160, dtype=np.float32).reshape(5, 32)
    L["label"] = np.arange(160).reshape(5, 32) % 7
    formats.write_kgraph(tmp_path / "g.kgraph", L)
    np.testing.assert_array_equal(formats.read_kgraph(tmp_path / "g.kgraph"), L)
    lab = np.array([[3], [-1], [0]], np.int32)
    formats.write_ivecs(tmp_path / "l.ivecs", lab)
    np.testing.assert_array_equal(formats.read_ivecs(tmp_path / "l.ivecs"), lab)


def test_ragged_rejected(tmp_path):
    p = tmp_path / "bad.fvecs"
    np.array([2, 0, 0, 3, 0, 0, 0], np.int32).tofile(p)
    with pytest.raises(ValueError):
        formats.read_fvecs(p)
