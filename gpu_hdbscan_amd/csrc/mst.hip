// Boruvka minimum spanning forest, MST-edge extraction and the stable
// MST-edge sort for gfx950.
//
// Edges are undirected, given once as (src < dst, w) in a canonical order e.
// The strict total order is the 64-bit key (float bits of w) << 32 | e: w >= 0
// so the bit pattern orders like the value, and e breaks ties the way the
// reference's Kruskal sorts (w, CSR position) (HDBSCAN/mst/mst.cu:42-54).
// Weights are never quantised (the reference's GPU path scales to int,
// HDBSCAN/core-sg/core-sg.cu:13-35).
#include <algorithm>

#include "hdb_internal.hpp"

namespace hdb {

__device__ __forceinline__ int find_compress(int v, int* __restrict__ parent) {
  while (true) {
    const int p = parent[v];
    if (p == v) return v;
    const int gp = parent[p];
    if (gp == p) return p;
    parent[v] = gp;  // path halving; values only move to ancestors
    v = gp;
  }
}

__device__ __forceinline__ int find_volatile(int v, volatile int* parent) {
  int p = parent[v];
  while (p != v) {
    v = p;
    p = parent[v];
  }
  return v;
}

// Hook the larger root under the smaller (join, mst.cu:93-100).
__device__ __forceinline__ void link_roots(int a, int b, int* parent) {
  volatile int* vp = parent;
  while (true) {
    a = find_volatile(a, vp);
    b = find_volatile(b, vp);
    if (a == b) return;
    if (a < b) {
      const int t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(&parent[a], a, b) == a) return;
  }
}

__device__ __forceinline__ uint64_t edge_key(uint32_t wbits, uint32_t e) {
  return (static_cast<uint64_t>(wbits) << 32) | e;
}

// Round 0: every vertex is its own component.
__global__ void k_mst_min0(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                           const float* __restrict__ w, int64_t m, unsigned long long* __restrict__ minkey) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const uint64_t key = edge_key(__float_as_uint(w[e]), static_cast<uint32_t>(e));
  atomicMin(&minkey[src[e]], static_cast<unsigned long long>(key));
  atomicMin(&minkey[dst[e]], static_cast<unsigned long long>(key));
}

__global__ void k_mst_hook0(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                            const float* __restrict__ w, int64_t m, const unsigned long long* __restrict__ minkey,
                            int* __restrict__ parent, uint8_t* __restrict__ flags) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const int a = src[e], b = dst[e];
  const uint64_t key = edge_key(__float_as_uint(w[e]), static_cast<uint32_t>(e));
  if (key == minkey[a] || key == minkey[b]) {
    flags[e] = 1;
    link_roots(a, b, parent);
  }
}

// Wave-aggregated append to a worklist.
__device__ __forceinline__ int64_t wave_append(bool want, unsigned long long* counter) {
  const uint64_t bal = __ballot(want);
  if (!bal) return -1;
  const int leader = __builtin_ctzll(bal);
  unsigned long long base = 0;
  if (static_cast<int>(__lane_id()) == leader) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(bal)));
  base = __shfl(base, leader);
  return want ? static_cast<int64_t>(base) + __popcll(bal & lanemask_lt()) : -1;
}

// Round 1 from the edge arrays: drop intra-component edges, compact the
// rest into the worklist with their current roots, offer keys (kernel1,
// mst.cu:155-171).
__global__ void k_mst_min_arr(const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
                              const float* __restrict__ w, int64_t m, int* __restrict__ parent,
                              unsigned long long* __restrict__ minkey, uint4* __restrict__ wl,
                              unsigned long long* __restrict__ wl_cnt) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool keep = false;
  int ra = 0, rb = 0;
  uint32_t wb = 0;
  if (e < m) {
    ra = find_compress(src[e], parent);
    rb = find_compress(dst[e], parent);
    keep = ra != rb;
    wb = __float_as_uint(w[e]);
  }
  const int64_t slot = wave_append(keep, wl_cnt);
  if (keep) {
    wl[slot] = make_uint4(static_cast<uint32_t>(ra), static_cast<uint32_t>(rb), wb, static_cast<uint32_t>(e));
    const uint64_t key = edge_key(wb, static_cast<uint32_t>(e));
    atomicMin(&minkey[ra], static_cast<unsigned long long>(key));
    atomicMin(&minkey[rb], static_cast<unsigned long long>(key));
  }
}

__global__ void k_mst_min_wl(const uint4* __restrict__ wl_in, int64_t cnt, int* __restrict__ parent,
                             unsigned long long* __restrict__ minkey, uint4* __restrict__ wl_out,
                             unsigned long long* __restrict__ wl_cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool keep = false;
  uint4 el = make_uint4(0, 0, 0, 0);
  if (i < cnt) {
    el = wl_in[i];
    el.x = static_cast<uint32_t>(find_compress(static_cast<int>(el.x), parent));
    el.y = static_cast<uint32_t>(find_compress(static_cast<int>(el.y), parent));
    keep = el.x != el.y;
  }
  const int64_t slot = wave_append(keep, wl_cnt);
  if (keep) {
    wl_out[slot] = el;
    const uint64_t key = edge_key(el.z, el.w);
    atomicMin(&minkey[el.x], static_cast<unsigned long long>(key));
    atomicMin(&minkey[el.y], static_cast<unsigned long long>(key));
  }
}

// kernel2 (mst.cu:174-185): an edge that is the minimum of either endpoint
// component joins the forest.
__global__ void k_mst_hook_wl(const uint4* __restrict__ wl, int64_t cnt, const unsigned long long* __restrict__ minkey,
                              int* __restrict__ parent, uint8_t* __restrict__ flags) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const uint4 el = wl[i];
  const uint64_t key = edge_key(el.z, el.w);
  if (key == minkey[el.x] || key == minkey[el.y]) {
    flags[el.w] = 1;
    link_roots(static_cast<int>(el.x), static_cast<int>(el.y), parent);
  }
}

__global__ void k_iota(int* __restrict__ p, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = static_cast<int>(i);
}

__global__ void k_count_roots(const int* __restrict__ parent, int64_t n, unsigned long long* __restrict__ cnt) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool root = v < n && parent[v] == static_cast<int>(v);
  const uint64_t bal = __ballot(root);
  if (__lane_id() == 0 && bal) atomicAdd(cnt, static_cast<unsigned long long>(__popcll(bal)));
}

int mst_edges_boruvka(hdb_ctx* ctx, const int32_t* src, const int32_t* dst, const float* w, int64_t n,
                      int64_t m, uint8_t* flags, int64_t* components_out) {
  if (n < 1 || m < 0) return ctx->fail(HDB_EINVAL, "bad MST sizes");
  if (m >= (1ll << 32)) return ctx->fail(HDB_EUNSUPPORTED, "more than 2^32 undirected edges");
  Stage st(ctx, "mst_boruvka");
  int* parent;
  unsigned long long* minkey;
  unsigned long long* cnt;
  HDB_CALL(scratch_t(ctx, kSlotMstParent, n, &parent));
  HDB_CALL(scratch_t(ctx, kSlotMstMin, n, &minkey));
  HDB_CALL(scratch_t(ctx, kSlotMstCnt, 4, &cnt));
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, parent, n);
  HDB_LAUNCHED();
  HDB_CALL(fill_u64(ctx, reinterpret_cast<uint64_t*>(minkey), n, kKeyMax));
  HDB_CALL(fill_u8(ctx, flags, m, 0));
  if (m > 0) {
    const unsigned g = grid_for(m, 256);
    hipLaunchKernelGGL(k_mst_min0, dim3(g), dim3(256), 0, ctx->stream, src, dst, w, m, minkey);
    HDB_LAUNCHED();
    hipLaunchKernelGGL(k_mst_hook0, dim3(g), dim3(256), 0, ctx->stream, src, dst, w, m, minkey, parent, flags);
    HDB_LAUNCHED();
    uint4 *wl_a, *wl_b;
    HDB_CALL(scratch_t(ctx, kSlotMstWl1, m, &wl_a));
    HDB_CALL(scratch_t(ctx, kSlotMstWl2, m, &wl_b));
    int64_t count = -1;
    for (int round = 1;; ++round) {
      HDB_CALL(fill_u64(ctx, reinterpret_cast<uint64_t*>(minkey), n, kKeyMax));
      HDB_CALL(fill_u64(ctx, reinterpret_cast<uint64_t*>(cnt), 1, 0));
      if (round == 1) {
        hipLaunchKernelGGL(k_mst_min_arr, dim3(g), dim3(256), 0, ctx->stream, src, dst, w, m, parent, minkey, wl_a,
                           cnt);
      } else {
        hipLaunchKernelGGL(k_mst_min_wl, dim3(grid_for(count, 256)), dim3(256), 0, ctx->stream, wl_b, count, parent,
                           minkey, wl_a, cnt);
      }
      HDB_LAUNCHED();
      HDB_CALL(read_i64(ctx, reinterpret_cast<const int64_t*>(cnt), &count));
      if (count == 0) break;
      hipLaunchKernelGGL(k_mst_hook_wl, dim3(grid_for(count, 256)), dim3(256), 0, ctx->stream, wl_a, count, minkey,
                         parent, flags);
      HDB_LAUNCHED();
      std::swap(wl_a, wl_b);
      if (round > 64) return ctx->fail(HDB_EDEVICE, "Boruvka did not converge");
    }
  }
  HDB_CALL(fill_u64(ctx, reinterpret_cast<uint64_t*>(cnt), 1, 0));
  hipLaunchKernelGGL(k_count_roots, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, parent, n, cnt);
  HDB_LAUNCHED();
  return read_i64(ctx, reinterpret_cast<const int64_t*>(cnt), components_out);
}

// ------------------------------------------------------- edge extraction --
// Stable compaction of flagged edges (buildMST, mst.cu:291-334): tiles of
// 4096 edges, 4 waves x 16 chunks of 64.
constexpr int kCmpTile = 4096;

__global__ __launch_bounds__(256) void k_flag_count(const uint8_t* __restrict__ flags, int64_t m,
                                                    int32_t* __restrict__ tilecnt) {
  __shared__ int32_t ws[4];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kCmpTile + (threadIdx.x >> 6) * 1024;
  int c = 0;
  for (int ch = 0; ch < 16; ++ch) {
    const int64_t i = base + ch * 64 + __lane_id();
    c += __popcll(__ballot(i < m && flags[i]));
  }
  if (__lane_id() == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(256) void k_flag_write(const uint8_t* __restrict__ flags, int64_t m,
                                                    const int64_t* __restrict__ tileoff, const int32_t* __restrict__ src,
                                                    const int32_t* __restrict__ dst, const float* __restrict__ w,
                                                    hdb_mst_edge* __restrict__ out, int64_t cap) {
  __shared__ int32_t ws[4];
  const unsigned wv = threadIdx.x >> 6;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kCmpTile + wv * 1024;
  int c = 0;
  for (int ch = 0; ch < 16; ++ch) {
    const int64_t i = base + ch * 64 + __lane_id();
    c += __popcll(__ballot(i < m && flags[i]));
  }
  if (__lane_id() == 0) ws[wv] = c;
  __syncthreads();
  int64_t off = tileoff[blockIdx.x];
  for (unsigned q = 0; q < wv; ++q) off += ws[q];
  for (int ch = 0; ch < 16; ++ch) {
    const int64_t i = base + ch * 64 + __lane_id();
    const bool f = i < m && flags[i];
    const uint64_t bal = __ballot(f);
    if (f) {
      hdb_mst_edge ed;
      ed.from_node = src[i];
      ed.to_node = dst[i];
      ed.weight = w[i];
      const int64_t pos = off + __popcll(bal & lanemask_lt());
      if (pos < cap) out[pos] = ed;
    }
    off += __popcll(bal);
  }
}

int mst_extract(hdb_ctx* ctx, const int32_t* src, const int32_t* dst, const float* w, const uint8_t* flags,
                int64_t m, hdb_mst_edge* out, int64_t cap, int64_t* count_out) {
  Stage st(ctx, "mst_extract");
  const int64_t tiles = std::max<int64_t>((m + kCmpTile - 1) / kCmpTile, 1);
  int32_t* tilecnt;
  int64_t* tileoff;
  HDB_CALL(scratch_t(ctx, kSlotMisc0, tiles, &tilecnt));
  HDB_CALL(scratch_t(ctx, kSlotMisc1, tiles + 1, &tileoff));
  hipLaunchKernelGGL(k_flag_count, dim3(tiles), dim3(256), 0, ctx->stream, flags, m, tilecnt);
  HDB_LAUNCHED();
  HDB_CALL(scan_i32_to_i64(ctx, tilecnt, tiles, tileoff));
  hipLaunchKernelGGL(k_flag_write, dim3(tiles), dim3(256), 0, ctx->stream, flags, m, tileoff, src, dst, w, out, cap);
  HDB_LAUNCHED();
  HDB_CALL(read_i64(ctx, tileoff + tiles, count_out));
  if (*count_out > cap) return ctx->fail(HDB_EDEVICE, "MST edge count exceeds n-1");
  return HDB_OK;
}

// ---------------------------------------------------------- CSR -> upper --
// Entries j of row i with nlist[j] > i, in CSR order (the "n > i" direction
// of cpuMST, mst.cu:42-50): one wave per row.
__global__ __launch_bounds__(256) void k_upper_count(const int64_t* __restrict__ nindex, const int32_t* __restrict__ nlist,
                                                     int64_t n, int32_t* __restrict__ cnt) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  int c = 0;
  for (int64_t j = nindex[i] + __lane_id(); j - __lane_id() < nindex[i + 1]; j += 64) {
    const bool up = j < nindex[i + 1] && nlist[j] > i && nlist[j] < n;
    c += __popcll(__ballot(up));
  }
  if (__lane_id() == 0) cnt[i] = c;
}

__global__ __launch_bounds__(256) void k_upper_fill(const int64_t* __restrict__ nindex, const int32_t* __restrict__ nlist,
                                                    const float* __restrict__ w, int64_t n,
                                                    const int64_t* __restrict__ off, int32_t* __restrict__ usrc,
                                                    int32_t* __restrict__ udst, float* __restrict__ uw) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  int64_t o = off[i];
  for (int64_t j = nindex[i] + __lane_id(); j - __lane_id() < nindex[i + 1]; j += 64) {
    const bool up = j < nindex[i + 1] && nlist[j] > i && nlist[j] < n;
    const uint64_t bal = __ballot(up);
    if (up) {
      const int64_t pos = o + __popcll(bal & lanemask_lt());
      usrc[pos] = static_cast<int32_t>(i);
      udst[pos] = nlist[j];
      uw[pos] = w[j];
    }
    o += __popcll(bal);
  }
}

int csr_upper(hdb_ctx* ctx, const int64_t* nindex, const int32_t* nlist, const float* w, int64_t n, int64_t E,
              int32_t* usrc, int32_t* udst, float* uw, int64_t* m_out) {
  (void)E;
  Stage st(ctx, "csr_upper");
  int32_t* cnt;
  int64_t* off;
  HDB_CALL(scratch_t(ctx, kSlotMisc0, n, &cnt));
  HDB_CALL(scratch_t(ctx, kSlotMisc1, n + 1, &off));
  hipLaunchKernelGGL(k_upper_count, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, nindex, nlist, n, cnt);
  HDB_LAUNCHED();
  HDB_CALL(scan_i32_to_i64(ctx, cnt, n, off));
  if (usrc) {
    hipLaunchKernelGGL(k_upper_fill, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, nindex, nlist, w, n, off, usrc,
                       udst, uw);
    HDB_LAUNCHED();
  }
  return read_i64(ctx, off + n, m_out);
}

// ------------------------------------------------------------ edge sort --
// Stable ascending sort by weight (sort_edges / mergeSort, merge_sort.cu:
// 27-116): LSD radix sort of the weight bits with (from, to) as payload.
__global__ void k_edges_split(const hdb_mst_edge* __restrict__ e, int64_t m, uint32_t* __restrict__ keys,
                              uint64_t* __restrict__ vals) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const hdb_mst_edge x = e[i];
  uint32_t b = __float_as_uint(x.weight);
  // order-preserving map for any sign (weights are >= 0 on this path)
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  keys[i] = b;
  vals[i] = (static_cast<uint64_t>(static_cast<uint32_t>(x.from_node)) << 32) | static_cast<uint32_t>(x.to_node);
}

__global__ void k_edges_join(const uint32_t* __restrict__ keys, const uint64_t* __restrict__ vals, int64_t m,
                             hdb_mst_edge* __restrict__ e) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  uint32_t b = keys[i];
  b = (b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b;
  hdb_mst_edge x;
  x.weight = __uint_as_float(b);
  x.from_node = static_cast<int32_t>(vals[i] >> 32);
  x.to_node = static_cast<int32_t>(static_cast<uint32_t>(vals[i]));
  e[i] = x;
}

int sort_edges(hdb_ctx* ctx, hdb_mst_edge* edges, int64_t m) {
  if (m <= 1) return HDB_OK;
  Stage st(ctx, "sort_edges");
  uint32_t* keys;
  uint64_t* vals;
  HDB_CALL(scratch_t(ctx, kSlotMstSrc, m, &keys));
  HDB_CALL(scratch_t(ctx, kSlotMstPos, m, &vals));
  hipLaunchKernelGGL(k_edges_split, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, edges, m, keys, vals);
  HDB_LAUNCHED();
  HDB_CALL(radix_sort_u32_u64(ctx, keys, vals, m, 0, 32));
  hipLaunchKernelGGL(k_edges_join, dim3(grid_for(m, 256)), dim3(256), 0, ctx->stream, keys, vals, m, edges);
  HDB_LAUNCHED();
  return HDB_OK;
}

}  // namespace hdb
