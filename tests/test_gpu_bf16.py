"""BF16 MFMA join (distance_mode 1) on the anisotropic d=128 mixture:
stored distances are exact FP32 values of the oracle formula (the BF16 Gram
matrix only shortlists partners), recall@32 >= 0.95 against FP64 brute force,
core distances within 1e-6 relative of FP64."""
import numpy as np
import pytest

from gpu_hdbscan_amd import api, datasets
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_bf16_join_exact_distances_and_recall(ctx):
    print("case: bf16 join aniso50 30000x128", flush=True)
    X = datasets.make("aniso50_2m_d128", n=30000, seed=2)
    p = api.default_params(distance_mode=1)
    L = api.NNDescent(ctx, X, iteration=8, params=p)
    ids = L["label"]
    assert not np.any(ids == np.arange(len(X))[:, None])
    for i in range(0, len(X), 997):
        np.testing.assert_array_equal(L["distance"][i], np.array([O.d2(X[i], X[j]) for j in ids[i]], np.float32))
        assert len(set(ids[i].tolist())) == 32
    q = np.arange(0, len(X), 30, dtype=np.int32)
    gt, _ = O.bruteforce_fp64(X, q, 32)
    r_bf16 = O.recall(ids[q], gt, 32)
    L32 = api.NNDescent(ctx, X, iteration=8)
    r_fp32 = O.recall(L32["label"][q], gt, 32)
    print("recall bf16", r_bf16, "fp32", r_fp32, flush=True)
    assert r_bf16 >= 0.95
    core = api.calculateCoreDistance(ctx, X, L, 32)
    j = ids[:, 31]
    f64 = np.sqrt(((X.astype(np.float64) - X[j].astype(np.float64)) ** 2).sum(1))
    assert np.max(np.abs(core - f64) / f64) < 1e-6
