"""GPU parity of the HDBSCAN graph stage and MST against the CPU oracle:
hub counts, anti-hubs, CSR and MST edges bit-exact; core distances
bit-exact (both sides accumulate in FP64 in dimension order and round once;
the contract tolerance is 1e-6 relative); sorted edges and labels
bit-exact."""
import numpy as np
import pytest

from gpu_hdbscan_amd import api, datasets
from oracle import oracle as O
from oracle import tree_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graph(ctx):
    print("case: graph fixture mix10 30000x12", flush=True)
    X = datasets.make("mix10_100k_d12", n=30000, seed=3)
    L = api.NNDescent(ctx, X, iteration=6)
    return X, L


def test_hub_counts(ctx, graph):
    X, L = graph
    np.testing.assert_array_equal(api.calculate_degrees(ctx, L), O.hub_counts(L))


def test_antihubs(ctx, graph):
    X, L = graph
    c = O.hub_counts(L)
    np.testing.assert_array_equal(api.find_antihubs(ctx, L, c), O.antihubs(L, c))


@pytest.mark.parametrize("n", [3000, 20000])
def test_antihubs_many_ties(ctx, n):
    """Every list holds ids 0..32 (itself skipped), so all other vertices tie
    at in-degree 0 with equal scores: the anti-hubs are the lowest ids among
    them.  n = 3000 sorts the 2,967 tied keys in one block, n = 20000 the
    19,967 tied keys with the radix sort."""
    print(f"case: antihubs many ties n={n}", flush=True)
    L = np.empty((n, 32), api.NND_ELEMENT)
    for i in range(n):
        ids = [j for j in range(33) if j != i][:32]
        L[i]["label"] = ids
        L[i]["distance"] = np.arange(32, dtype=np.float32)
    c = O.hub_counts(L)
    np.testing.assert_array_equal(api.find_antihubs(ctx, L, c), O.antihubs(L, c))


def test_core_distances(ctx, graph):
    X, L = graph
    got = api.calculateCoreDistance(ctx, X, L, 32)
    ref = O.core(X, L, 32)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    # contract: within 1e-6 relative of FP64 recomputed on the returned neighbour
    j = L["label"][:, 31]
    f64 = np.sqrt(((X.astype(np.float64) - X[j].astype(np.float64)) ** 2).sum(1))
    assert np.max(np.abs(got - f64) / f64) < 1e-6


def test_mrd_graph_full_csr(ctx, graph):
    X, L = graph
    g = api.buildEnhancedKNNG(ctx, X, L, 32)
    c = O.core(X, L)
    ah = O.antihubs(L, O.hub_counts(L))
    ni, nl, w = O.mrd_csr(X, L, c, ah)
    np.testing.assert_array_equal(g.nindex, ni)
    np.testing.assert_array_equal(g.nlist, nl)
    np.testing.assert_array_equal(g.eweight.view(np.uint32), w.view(np.uint32))


def test_mst_matches_kruskal(ctx, graph):
    X, L = graph
    c = O.core(X, L)
    ah = O.antihubs(L, O.hub_counts(L))
    ni, nl, w = O.mrd_csr(X, L, c, ah)
    g = api.ECLgraph(len(X), len(nl), ni, nl, w)
    got, comps = api.gpuMST(ctx, g)
    ref, rcomps = O.kruskal(ni, nl, w)
    assert comps == rcomps == 1
    np.testing.assert_array_equal(got, ref)


def test_mst_tie_heavy_and_disconnected(ctx):
    print("case: mst ties + 2 components", flush=True)
    rng = np.random.default_rng(5)
    n = 3000
    rows = [set() for _ in range(n)]
    for i in range(n):
        half = 0 if i < n // 2 else n // 2
        for j in rng.choice(n // 2, 6, replace=False) + half:
            if j != i:
                rows[i].add(int(j))
                rows[j].add(i)
    ni = np.zeros(n + 1, np.int64)
    nl = []
    for i in range(n):
        nl.extend(sorted(rows[i]))
        ni[i + 1] = len(nl)
    nl = np.array(nl, np.int32)
    # symmetric weights from a few levels -> heavy ties
    src = np.repeat(np.arange(n), np.diff(ni))
    w = ((np.minimum(src, nl) * 7 + np.maximum(src, nl) * 13) % 5).astype(np.float32)
    g = api.ECLgraph(n, len(nl), ni, nl, w)
    got, comps = api.gpuMST(ctx, g)
    ref, rcomps = O.kruskal(ni, nl, w)
    assert comps == rcomps == 2
    np.testing.assert_array_equal(got, ref)


def test_mst_long_rows_and_ties(ctx):
    """Rows of a few to two thousand entries (up to 32 steps of 64 entries
    per row) with weights from eleven levels, so ties are decided by the
    vertex pair."""
    print("case: mst long rows 2000 vertices, 40 hubs", flush=True)
    rng = np.random.default_rng(11)
    n, hubs = 2000, 40
    rows = [set() for _ in range(n)]
    for i in range(n):
        for j in rng.choice(n, 5, replace=False):
            if j != i:
                rows[i].add(int(j))
                rows[j].add(i)
    for hb in range(hubs):
        for j in range(n):
            if j != hb:
                rows[hb].add(j)
                rows[j].add(hb)
    ni = np.zeros(n + 1, np.int64)
    nl = []
    for i in range(n):
        nl.extend(sorted(rows[i]))
        ni[i + 1] = len(nl)
    nl = np.array(nl, np.int32)
    src = np.repeat(np.arange(n), np.diff(ni))
    w = ((np.minimum(src, nl) * 7 + np.maximum(src, nl) * 13) % 11).astype(np.float32) * np.float32(0.25) + 1
    g = api.ECLgraph(n, len(nl), ni, nl, w)
    got, comps = api.gpuMST(ctx, g)
    ref, rcomps = O.kruskal(ni, nl, w)
    assert comps == rcomps == 1
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("m", [1, 4095, 4097, 1 << 20, (1 << 20) + 4096])
def test_sort_edges_sizes(ctx, m):
    """Radix tiles of 4096 keys: one partial tile, one full tile plus one
    key, and 256 and 257 tiles (histograms of 65,536 and 65,792 entries)."""
    print(f"case: sort {m} edges", flush=True)
    rng = np.random.default_rng(m)
    e = np.empty(m, api.MST_EDGE)
    e["from_node"] = rng.integers(0, 1 << 24, m)
    e["to_node"] = rng.integers(0, 1 << 24, m)
    e["weight"] = rng.integers(0, 1 << 16, m).astype(np.float32) * np.float32(0.01)
    np.testing.assert_array_equal(api.sort_edges(ctx, e), O.sort_edges(e))


def test_sort_edges(ctx):
    print("case: sort 200000 edges", flush=True)
    rng = np.random.default_rng(7)
    e = np.empty(200000, api.MST_EDGE)
    e["from_node"] = rng.integers(0, 1 << 20, 200000)
    e["to_node"] = rng.integers(0, 1 << 20, 200000)
    e["weight"] = rng.integers(0, 3000, 200000).astype(np.float32) * np.float32(0.37)
    np.testing.assert_array_equal(api.sort_edges(ctx, e), O.sort_edges(e))


def test_pipeline_end_to_end(ctx):
    print("case: pipeline 20000x12", flush=True)
    X = datasets.make("mix10_100k_d12", n=20000, seed=9)
    res = api.pipeline(ctx, X)
    L = res.lists
    np.testing.assert_array_equal(res.counts, O.hub_counts(L))
    c = O.core(X, L)
    np.testing.assert_array_equal(res.core.view(np.uint32), c.view(np.uint32))
    ah = O.antihubs(L, res.counts)
    ni, nl, w = O.mrd_csr(X, L, c, ah)
    ref, rcomps = O.kruskal(ni, nl, w)
    assert res.components == rcomps
    assert res.csr_edges == len(nl)
    np.testing.assert_array_equal(res.edges, O.sort_edges(ref))
    lab, nc = api.get_clusters(res.edges, len(X), 32)
    rlab, rnc = tree_oracle.labels([tuple(x) for x in res.edges], len(X), 32)
    assert nc == rnc
    np.testing.assert_array_equal(lab, np.array(rlab, np.int32))
    print("clusters", nc, "noise", int((lab == -1).sum()), flush=True)


def test_pipeline_blocked(ctx):
    print("case: pipeline blocked 3 blocks 15000x12", flush=True)
    X = datasets.make("mix10_100k_d12", n=15000, seed=10)
    res = api.pipeline(ctx, X, knng_mode=1, blocks=3)
    L = res.lists
    c = O.core(X, L)
    ah = O.antihubs(L, O.hub_counts(L))
    ni, nl, w = O.mrd_csr(X, L, c, ah)
    ref, _ = O.kruskal(ni, nl, w)
    np.testing.assert_array_equal(res.edges, O.sort_edges(ref))
