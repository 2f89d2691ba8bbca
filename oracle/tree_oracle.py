"""Labels oracle (TEST INFRASTRUCTURE, never the product).

A literal pure-Python restatement of HDBSCAN/trees/tree.cu (reference
gorlando04/GPU_HDBSCAN) for small inputs: build_Linkage_tree (:171-202),
BFS_from_hierarchy (:206-253), build_Condensed_tree (:255-442),
compute_stability (:444-528), BFS_from_cluster_tree (:532-567),
do_labelling (:569-625) and get_clusters (:627-746), keeping the reference's
float32 arithmetic (numpy float32 scalars) and its quirks:
lambda = float(INT_MAX) for zero weights (:308-313) and the strict
point-lambda > cluster-lambda test (:609-618).

Only tests/ and the smoke check may import this module.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
INT_MAX = 2147483647


def build_linkage_tree(edges, n):
    """edges: sequence of (from, to, weight) sorted by weight; returns list of
    (left, right, weight(float32), size)."""
    parent = [-1] * (2 * n - 1)
    size = [1] * n + [0] * (n - 1)
    next_label = n

    def fast_find(x):
        # the reference's FastFind returns the root; its partial path
        # compression never changes a root
        while parent[x] != -1:
            x = parent[x]
        return x

    out = []
    for i in range(n - 1):
        a, b, w = int(edges[i][0]), int(edges[i][1]), F32(edges[i][2])
        aa, bb = fast_find(a), fast_find(b)
        out.append((aa, bb, w, size[aa] + size[bb]))
        size[next_label] = size[aa] + size[bb]
        parent[aa] = next_label
        parent[bb] = next_label
        next_label += 1
    return out


def bfs_from_hierarchy(h, bfs_root, num_nodes):
    dim = num_nodes
    max_node = 2 * num_nodes
    num_points = max_node - dim + 1
    to_process = [bfs_root]
    result = []
    while to_process:
        result.extend(to_process)
        aux = [x - num_points for x in to_process if x >= num_points]
        nxt = []
        for a in aux:
            nxt.append(h[a][0])
            nxt.append(h[a][1])
        to_process = nxt
    return result


def build_condensed_tree(h, num_nodes, mpts):
    root = 2 * num_nodes
    num_points = root // 2 + 1
    next_label = num_points + 1
    node_list = bfs_from_hierarchy(h, root, num_nodes)
    relabel = [0] * (root + 1)
    relabel[root] = num_points
    ignore = [0] * len(node_list)
    out = []
    for node in node_list:
        if ignore[node] or node < num_points:
            continue
        left, right, w, _ = h[node - num_points]
        lam = F32(INT_MAX)
        if float(w) > 0.0:
            lam = F32(1.0 / float(w))
        lc = h[left - num_points][3] if left >= num_points else 1
        rc = h[right - num_points][3] if right >= num_points else 1
        if lc >= mpts and rc >= mpts:
            relabel[left] = next_label
            next_label += 1
            out.append((relabel[node], relabel[left], lam, lc))
            relabel[right] = next_label
            next_label += 1
            out.append((relabel[node], relabel[right], lam, rc))
        elif lc < mpts and rc < mpts:
            for sub in bfs_from_hierarchy(h, left, num_nodes):
                if sub < num_points:
                    out.append((relabel[node], sub, lam, 1))
                ignore[sub] = 1
            for sub in bfs_from_hierarchy(h, right, num_nodes):
                if sub < num_points:
                    out.append((relabel[node], sub, lam, 1))
                ignore[sub] = 1
        elif lc < mpts:
            relabel[right] = relabel[node]
            for sub in bfs_from_hierarchy(h, left, num_nodes):
                if sub < num_points:
                    out.append((relabel[node], sub, lam, 1))
                ignore[sub] = 1
        else:
            relabel[left] = relabel[node]
            for sub in bfs_from_hierarchy(h, right, num_nodes):
                if sub < num_points:
                    out.append((relabel[node], sub, lam, 1))
                ignore[sub] = 1
    return out


def compute_stability(ct):
    max_child = max(c[1] for c in ct)
    max_parent = max(c[0] for c in ct)
    min_parent = min(c[0] for c in ct)
    largest_child = max(max_child, min_parent)
    srt = sorted((c[1], float(c[2])) for c in ct)
    birth = [F32(0.0)] * (largest_child + 1)
    current, min_lambda = -1, F32(0.0)
    for child, lam in srt:
        lam = F32(lam)
        if child == current:
            min_lambda = min(min_lambda, lam)
        elif current != -1:
            birth[current] = min_lambda
            current, min_lambda = child, lam
        else:
            current, min_lambda = child, lam
    if current != -1:
        birth[current] = min_lambda
    birth[min_parent] = F32(0.0)
    res = [F32(0.0)] * (max_parent - min_parent + 1)
    for parent, child, lam, csize in ct:
        res[parent - min_parent] = F32(res[parent - min_parent] + F32(F32(lam - birth[parent]) * F32(csize)))
    return [(min_parent + i, res[i]) for i in range(max_parent + 1 - min_parent)]


def bfs_from_cluster_tree(ct, bfs_root):
    min_parent = min(c[0] for c in ct)
    to_process = [bfs_root]
    result = []
    while to_process:
        to_process = sorted(to_process)
        result.extend(to_process)
        tp = set(to_process)
        to_process = [c[1] for c in ct if c[0] in tp and c[1] >= min_parent]
    return sorted(result)


def do_labelling(ct, clusters, cluster_map):
    root_cluster = min(c[0] for c in ct)
    max_parent = max(c[0] for c in ct)
    lam = {}
    for parent, child, l, _ in ct:
        lam[child] = l
        if parent == root_cluster:
            lam[root_cluster] = l
    data0 = list(range(max_parent + 1))
    data1 = [0] * (max_parent + 1)

    def find(x):
        while data0[x] != x:
            data0[x] = data0[data0[x]]
            x = data0[x]
        return x

    def union(x, y):
        xr, yr = find(x), find(y)
        if data1[xr] < data1[yr]:
            data0[xr] = yr
        elif data1[xr] > data1[yr]:
            data0[yr] = xr
        else:
            data0[yr] = xr
            data1[xr] += 1

    cset = set(clusters)
    for parent, child, _, _ in ct:
        if child not in cset:
            union(parent, child)
    labels = []
    for p in range(root_cluster):
        c = find(p)
        if c <= root_cluster:
            labels.append(-1)
        elif lam[p] > lam[c]:
            labels.append(cluster_map[c])
        else:
            labels.append(-1)
    return labels


def get_clusters(ct, stab, n):
    stab = [list(s) for s in stab]
    by_id = {s[0]: s for s in stab}
    node_list = [stab[i][0] for i in range(len(stab) - 1, 0, -1)]
    cluster_tree = [c for c in ct if c[3] > 1]
    max_parent = max(c[0] for c in ct)
    min_parent = min(c[0] for c in ct)
    is_cluster = [i > min_parent for i in range(max_parent + 1)]
    for node in node_list:
        sub = F32(0.0)
        for c in cluster_tree:
            if c[0] == node:
                sub = F32(sub + by_id[c[1]][1])
        if sub > by_id[node][1]:
            is_cluster[node] = False
            by_id[node][1] = sub
        else:
            for s in bfs_from_cluster_tree(ct, node):
                if s != node:
                    is_cluster[s] = False
    clusters = sorted(i for i, v in enumerate(is_cluster) if v)
    cmap = {c: i for i, c in enumerate(clusters)}
    return do_labelling(ct, clusters, cmap), len(clusters)


def labels(edges, n, mpts):
    """Labels for n points from n-1 MST edges sorted by weight."""
    if n == 1:
        return [-1], 0
    h = build_linkage_tree(edges, n)
    ct = build_condensed_tree(h, n - 1, mpts)
    if not ct:
        return [-1] * n, 0
    stab = compute_stability(ct)
    return get_clusters(ct, stab, n)
