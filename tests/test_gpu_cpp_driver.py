"""The C++ driver (gpu_hdbscan_amd/cpp/hdbscan_pipeline.cpp over
include/gpu_hdbscan.hpp, the reference-shaped C++ interface) produces the
same labels as the Python pipeline and the labels oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpu_hdbscan_amd import api, datasets, formats
from oracle import tree_oracle

pytestmark = pytest.mark.gpu
DRIVER = Path(__file__).resolve().parents[1] / "gpu_hdbscan_amd" / "hdbscan_pipeline"


def test_cpp_driver_labels(ctx, tmp_path):
    print("case: C++ driver 8000x12", flush=True)
    X = datasets.make("mix10_100k_d12", n=8000, seed=21)
    formats.write_fvecs(tmp_path / "x.fvecs", X)
    out = tmp_path / "labels.ivecs"
    r = subprocess.run([str(DRIVER), str(tmp_path / "x.fvecs"), "32", "1", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    lab_cpp = formats.read_ivecs(out)[:, 0]
    res = api.pipeline(ctx, X)
    lab_py, nc = api.get_clusters(res.edges, len(X), 32)
    np.testing.assert_array_equal(lab_cpp, lab_py)
    ref, rnc = tree_oracle.labels([tuple(e) for e in res.edges], len(X), 32)
    np.testing.assert_array_equal(lab_py, np.array(ref, np.int32))
