// HDBSCAN graph stage for gfx950: hub counts, anti-hub selection, core
// distances and the symmetrised mutual-reachability CSR with the anti-hub
// clique.  Integer work is exact; distances that enter weights are FP64
// accumulations rounded once to FP32.
#include <algorithm>
#include <cmath>

#include "hdb_internal.hpp"

namespace hdb {

// ------------------------------------------------------------ hub counts --
// countValues (HDBSCAN/counts/count.cu:10-19) over all n*k labels at once.
__global__ void k_hub_counts(const hdb_nnd_element* __restrict__ lists, int64_t total, int64_t n,
                             int32_t* __restrict__ counts) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < total; i += stride) {
    const uint32_t id = label_id(lists[i].label);
    if (id != kInvalidId && static_cast<int64_t>(id) < n) atomicAdd(&counts[id], 1);
  }
}

int hub_counts(hdb_ctx* ctx, const hdb_nnd_element* lists, int64_t n, int k, int32_t* counts) {
  if (!lists || !counts || n < 1 || k < 1) return ctx->fail(HDB_EINVAL, "bad hub_counts arguments");
  Stage st(ctx, "hub_counts");
  HDB_CALL(fill_i32(ctx, counts, n, 0));
  unsigned grid = grid_for(n * k, 256);
  if (grid > 1u << 20) grid = 1u << 20;
  hipLaunchKernelGGL(k_hub_counts, dim3(grid), dim3(256), 0, ctx->stream, lists, n * k, n, counts);
  HDB_LAUNCHED();
  return HDB_OK;
}

// -------------------------------------------------------------- anti-hubs --
constexpr int kDegBins = 4096;

__global__ void k_degree_hist(const int32_t* __restrict__ counts, int64_t n, int32_t* __restrict__ hist) {
  __shared__ int32_t h[kDegBins];
  for (int i = threadIdx.x; i < kDegBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) atomicAdd(&h[min(counts[i], kDegBins - 1)], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < kDegBins; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// flag[v] = 1 for degree < thr; tied vertices (degree == thr) append the key
// (score << 32 | v) where score = sum of in-degrees over v's list
// (calculateScore, HDBSCAN/calculates/calculates.cu:21-36).
__global__ void k_antihub_mark(const hdb_nnd_element* __restrict__ lists, const int32_t* __restrict__ counts,
                               int64_t n, int k, int32_t thr, int32_t* __restrict__ flag,
                               uint64_t* __restrict__ tied, int32_t* __restrict__ ntied) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int32_t dg = counts[v];
  flag[v] = dg < thr ? 1 : 0;
  if (dg == thr) {
    uint64_t score = 0;
    for (int j = 0; j < k; ++j) {
      const uint32_t id = label_id(lists[v * k + j].label);
      if (id != kInvalidId && static_cast<int64_t>(id) < n) score += static_cast<uint32_t>(counts[id]);
    }
    if (score > 0xFFFFFFFFull) score = 0xFFFFFFFFull;
    const int slot = atomicAdd(ntied, 1);
    tied[slot] = (score << 32) | static_cast<uint64_t>(v);
  }
}

__global__ void k_antihub_take(const uint64_t* __restrict__ sorted, int64_t need, int32_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < need) flag[static_cast<uint32_t>(sorted[i])] = 1;
}

__global__ void k_compact_flags(const int32_t* __restrict__ flag, const int64_t* __restrict__ off, int64_t n,
                                int32_t* __restrict__ out) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n && flag[v]) out[off[v]] = static_cast<int32_t>(v);
}

static int64_t isqrt_floor(int64_t n) {
  int64_t r = static_cast<int64_t>(std::sqrt(static_cast<double>(n)));
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

int antihubs(hdb_ctx* ctx, const hdb_nnd_element* lists, const int32_t* counts, int64_t n, int k, int32_t* ids,
             int64_t* h_out) {
  if (!lists || !counts || !ids || !h_out || n < 1) return ctx->fail(HDB_EINVAL, "bad antihubs arguments");
  Stage st(ctx, "antihubs");
  // get_NumThreshold (HDBSCAN/getters/getters.cu:6-9)
  const int64_t h = isqrt_floor(n);
  *h_out = h;
  if (h == 0) return HDB_OK;
  int32_t *hist, *flag, *aux;
  HDB_CALL(scratch_t(ctx, kSlotAhAux, kDegBins + 8, &aux));
  hist = aux;
  int32_t* ntied = aux + kDegBins;
  HDB_CALL(scratch_t(ctx, kSlotAhFlag, n, &flag));
  HDB_CALL(fill_i32(ctx, aux, kDegBins + 8, 0));
  hipLaunchKernelGGL(k_degree_hist, dim3(std::min<int64_t>(grid_for(n, 256), 1024)), dim3(256), 0, ctx->stream,
                     counts, n, hist);
  HDB_LAUNCHED();
  std::vector<int32_t> hh(kDegBins);
  HDB_HIP(hipMemcpyAsync(hh.data(), hist, kDegBins * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HDB_HIP(hipStreamSynchronize(ctx->stream));
  // threshold degree: the h-th smallest (partial_sort at graph.cu:123-124)
  int64_t cum = 0;
  int thr = -1;
  for (int b = 0; b < kDegBins; ++b) {
    if (cum + hh[b] >= h) {
      thr = b;
      break;
    }
    cum += hh[b];
  }
  if (thr < 0 || thr == kDegBins - 1)
    return ctx->fail(HDB_EUNSUPPORTED, "anti-hub threshold degree beyond the histogram range");
  const int64_t need = h - cum;  // tied vertices to take, 1..hh[thr]
  const int64_t ntied_h = hh[thr];
  uint64_t* tied;
  HDB_CALL(scratch_t(ctx, kSlotAhKeys, std::max<int64_t>(ntied_h, 1), &tied));
  hipLaunchKernelGGL(k_antihub_mark, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, lists, counts, n, k, thr, flag,
                     tied, ntied);
  HDB_LAUNCHED();
  // lowest (score, id) among the tied (calculate_finalAntihubs, calculates.cu:491-523)
  HDB_CALL(radix_sort_u64(ctx, tied, ntied_h, 0, 64));
  hipLaunchKernelGGL(k_antihub_take, dim3(grid_for(need, 256)), dim3(256), 0, ctx->stream, tied, need, flag);
  HDB_LAUNCHED();
  int64_t* off;
  HDB_CALL(scratch_t(ctx, kSlotMisc0, n + 1, &off));
  HDB_CALL(scan_i32_to_i64(ctx, flag, n, off));
  hipLaunchKernelGGL(k_compact_flags, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, flag, off, n, ids);
  HDB_LAUNCHED();
  return HDB_OK;
}

// --------------------------------------------------------- core distance --
// calculateCoreDistance (calculates.cu:185-197) at list index mpts-1
// (graph.cu:77), as a Euclidean distance recomputed in FP64.
__global__ void k_core(const float* __restrict__ X, const hdb_nnd_element* __restrict__ lists, int64_t n, int d,
                       int k, int idx, float* __restrict__ core) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t j = label_id(lists[i * k + idx].label);
  core[i] = (j == kInvalidId || static_cast<int64_t>(j) >= n)
                ? __uint_as_float(0x7F800000u)
                : dist_fp64(X + i * d, X + static_cast<int64_t>(j) * d, d);
}

int core_distances(hdb_ctx* ctx, const float* X, const hdb_nnd_element* lists, int64_t n, int d, int k, int mpts,
                   float* core) {
  if (!X || !lists || !core || n < 1 || d < 1) return ctx->fail(HDB_EINVAL, "bad core_distances arguments");
  if (mpts < 1 || mpts > k) return ctx->fail(HDB_EINVAL, "mpts must be in [1, k]");
  Stage st(ctx, "core_distances");
  hipLaunchKernelGGL(k_core, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, X, lists, n, d, k, mpts - 1, core);
  HDB_LAUNCHED();
  return HDB_OK;
}

// ------------------------------------------------------ MRD graph (CSR) --
// Every kNN entry (i -> v) contributes v to row i and i to row v (full) or
// only max(i,v) to row min(i,v) (upper).  Rows are then sorted and made
// unique, which yields N_i u R_i of calculate_nindex / calculate_nlist
// (calculates.cu:286-409) in neighbour-id order.
template <bool UPPER>
__global__ void k_edge_count(const hdb_nnd_element* __restrict__ lists, int64_t n, int k,
                             int32_t* __restrict__ rowcnt) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  const int64_t i = t / k;
  const uint32_t v = label_id(lists[t].label);
  if (v == kInvalidId || static_cast<int64_t>(v) >= n || static_cast<int64_t>(v) == i) return;
  if (UPPER) {
    atomicAdd(&rowcnt[min(static_cast<int64_t>(v), i)], 1);
  } else {
    atomicAdd(&rowcnt[i], 1);
    atomicAdd(&rowcnt[v], 1);
  }
}

template <bool UPPER>
__global__ void k_edge_scatter(const hdb_nnd_element* __restrict__ lists, int64_t n, int k,
                               const int64_t* __restrict__ rowoff, int32_t* __restrict__ cursor,
                               int32_t* __restrict__ tmp) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * k) return;
  const int64_t i = t / k;
  const uint32_t v = label_id(lists[t].label);
  if (v == kInvalidId || static_cast<int64_t>(v) >= n || static_cast<int64_t>(v) == i) return;
  if (UPPER) {
    const int64_t a = min(static_cast<int64_t>(v), i), bb = max(static_cast<int64_t>(v), i);
    tmp[rowoff[a] + atomicAdd(&cursor[a], 1)] = static_cast<int32_t>(bb);
  } else {
    tmp[rowoff[i] + atomicAdd(&cursor[i], 1)] = static_cast<int32_t>(v);
    tmp[rowoff[v] + atomicAdd(&cursor[v], 1)] = static_cast<int32_t>(i);
  }
}

// Rows of <= 128 entries: one wave per row, 2 keys per lane.  Longer rows
// are appended to `big` for k_row_sort_big.  Writes the sorted unique ids to
// the start of the row segment, uniq[v] and total[v] = uniq + clique share.
constexpr int kSmallRow = 128;
constexpr int kBigRowMax = 16384;  // 64 KiB of LDS

__device__ __forceinline__ int64_t clique_share(int64_t v, const int32_t* ahrank, int64_t h, bool upper) {
  const int32_t r = ahrank[v];
  if (r < 0) return 0;
  return upper ? (h - 1 - r) : (h - 1);
}

template <bool UPPER>
__global__ __launch_bounds__(256) void k_row_sort_small(const int64_t* __restrict__ rowoff, int64_t n,
                                                        int32_t* __restrict__ tmp, int32_t* __restrict__ uniq,
                                                        int32_t* __restrict__ total, const int32_t* __restrict__ ahrank,
                                                        int64_t h, int32_t* __restrict__ big,
                                                        int32_t* __restrict__ nbig, int32_t* __restrict__ maxlen) {
  const unsigned lane = __lane_id();
  const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (v >= n) return;
  const int64_t beg = rowoff[v];
  const int64_t len = rowoff[v + 1] - beg;
  if (len > kSmallRow) {
    if (lane == 0) {
      big[atomicAdd(nbig, 1)] = static_cast<int32_t>(v);
      atomicMax(maxlen, static_cast<int32_t>(min(len, static_cast<int64_t>(0x7FFFFFFF))));
    }
    return;
  }
  uint64_t r0 = lane < len ? static_cast<uint32_t>(tmp[beg + lane]) : kKeyMax;
  uint64_t r1 = 64 + lane < len ? static_cast<uint32_t>(tmp[beg + 64 + lane]) : kKeyMax;
  bitonic128(r0, r1);
  const uint64_t p0 = shfl_u64(r0, (lane + 63) & 63);
  const uint64_t q1 = shfl_u64(r1, (lane + 63) & 63);
  const uint64_t last0 = shfl_u64(r0, 63);
  const uint64_t p1 = lane == 0 ? last0 : q1;
  const bool keep0 = r0 != kKeyMax && (lane == 0 || p0 != r0);
  const bool keep1 = r1 != kKeyMax && p1 != r1;
  const uint64_t b0 = __ballot(keep0), b1 = __ballot(keep1);
  const uint64_t lt = lanemask_lt();
  if (keep0) tmp[beg + __popcll(b0 & lt)] = static_cast<int32_t>(r0);
  if (keep1) tmp[beg + __popcll(b0) + __popcll(b1 & lt)] = static_cast<int32_t>(r1);
  if (lane == 0) {
    const int32_t u = __popcll(b0) + __popcll(b1);
    uniq[v] = u;
    total[v] = static_cast<int32_t>(u + clique_share(v, ahrank, h, UPPER));
  }
}

// One 256-thread block per long row: LDS bitonic sort over the next power
// of two, then unique.
template <bool UPPER>
__global__ __launch_bounds__(256) void k_row_sort_big(const int64_t* __restrict__ rowoff, const int32_t* __restrict__ big,
                                                      int32_t* __restrict__ tmp, int32_t* __restrict__ uniq,
                                                      int32_t* __restrict__ total, const int32_t* __restrict__ ahrank,
                                                      int64_t h) {
  extern __shared__ uint32_t srt[];
  __shared__ int32_t wsum[4];
  const int64_t v = big[blockIdx.x];
  const int64_t beg = rowoff[v];
  const int len = static_cast<int>(rowoff[v + 1] - beg);
  int P = 1;
  while (P < len) P <<= 1;
  for (int i = threadIdx.x; i < P; i += 256) srt[i] = i < len ? static_cast<uint32_t>(tmp[beg + i]) : 0xFFFFFFFFu;
  __syncthreads();
  for (int kk = 2; kk <= P; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += 256) {
        const int l = i ^ j;
        if (l > i) {
          const uint32_t a = srt[i], c = srt[l];
          const bool up = (i & kk) == 0;
          if ((a > c) == up) {
            srt[i] = c;
            srt[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // unique in order: chunked block scan of keep flags
  int base = 0;
  for (int start = 0; start < len; start += 256) {
    const int i = start + threadIdx.x;
    const bool keep = i < len && (i == 0 || srt[i] != srt[i - 1]);
    const uint64_t bal = __ballot(keep);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int off = 0, tot = 0;
    for (int q = 0; q < 4; ++q) {
      if (q < w) off += wsum[q];
      tot += wsum[q];
    }
    if (keep) tmp[beg + base + off + __popcll(bal & lanemask_lt())] = static_cast<int32_t>(srt[i]);
    base += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uniq[v] = base;
    total[v] = static_cast<int32_t>(base + clique_share(v, ahrank, h, UPPER));
  }
}

// Fill row v: merge the unique kNN ids with the clique ids (kNN entry first
// on equal ids) and compute the weights (calculateMRD calculates.cu:61-82;
// clique entries raised to the Euclidean distance as in
// calculate_coreDistance_antihubs :412-446).
template <bool UPPER>
__global__ __launch_bounds__(256) void k_row_fill(const float* __restrict__ X, int d, const int64_t* __restrict__ rowoff,
                                                  const int32_t* __restrict__ tmp, const int32_t* __restrict__ uniq,
                                                  const int64_t* __restrict__ nindex, int64_t n,
                                                  const float* __restrict__ core, const int32_t* __restrict__ ahrank,
                                                  const int32_t* __restrict__ ah, int64_t h,
                                                  int32_t* __restrict__ nlist, float* __restrict__ w,
                                                  int32_t* __restrict__ src) {
  const unsigned lane = __lane_id();
  const int64_t v = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (v >= n) return;
  const int64_t beg = rowoff[v];
  const int u = uniq[v];
  const int64_t out0 = nindex[v];
  const float cv = core[v];
  const int32_t r = ahrank[v];
  // clique slice: ah[c0, c1) without v itself (full mode skips index r)
  int64_t c0 = 0, c1 = 0;
  if (r >= 0) {
    c0 = UPPER ? r + 1 : 0;
    c1 = h;
  }
  const int64_t ncl = (r >= 0) ? (UPPER ? h - 1 - r : h - 1) : 0;
  // kNN part
  for (int i = lane; i < u; i += 64) {
    const int32_t x = tmp[beg + i];
    int64_t before = 0;
    if (ncl > 0) {
      // clique ids < x (lower bound over the slice, minus self if below x)
      int64_t lo = c0, hi = c1;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (ah[mid] < x) lo = mid + 1; else hi = mid;
      }
      before = lo - c0;
      if (!UPPER && r < lo) before -= 1;
    }
    const int64_t pos = out0 + i + before;
    nlist[pos] = x;
    w[pos] = fmaxf(cv, core[x]);
    if (UPPER) src[pos] = static_cast<int32_t>(v);
  }
  // clique part
  for (int64_t q = c0 + lane; q < c1; q += 64) {
    if (!UPPER && q == r) continue;
    const int32_t c = ah[q];
    // kNN ids <= c (upper bound)
    int lo = 0, hi = u;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tmp[beg + mid] <= c) lo = mid + 1; else hi = mid;
    }
    int64_t qi = q - c0;
    if (!UPPER && q > r) qi -= 1;
    const int64_t pos = out0 + qi + lo;
    nlist[pos] = c;
    const float dab = dist_fp64(X + v * d, X + static_cast<int64_t>(c) * d, d);
    w[pos] = fmaxf(fmaxf(cv, core[c]), dab);
    if (UPPER) src[pos] = static_cast<int32_t>(v);
  }
}

__global__ void k_ah_rank(const int32_t* __restrict__ ah, int64_t h, int32_t* __restrict__ rank) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < h) rank[ah[i]] = static_cast<int32_t>(i);
}

int mrd_graph(hdb_ctx* ctx, const float* X, const hdb_nnd_element* lists, const float* core, const int32_t* ah,
              int64_t h, int64_t n, int d, int k, bool upper, int64_t* nindex, int32_t* nlist, float* w, int32_t* src,
              int64_t* edges_out) {
  if (!X || !lists || !core || !nindex || !edges_out || n < 1 || d < 1 || k < 1 || h < 0 || (h > 0 && !ah))
    return ctx->fail(HDB_EINVAL, "bad mrd_graph arguments");
  Stage st(ctx, upper ? "mrd_graph_upper" : "mrd_graph_full");
  const int64_t cap = (upper ? 1 : 2) * n * k;
  int32_t *rowcnt, *cursor, *tmp, *uniq, *ahrank, *misc;
  int64_t* rowoff;
  HDB_CALL(scratch_t(ctx, kSlotRowCnt, n + 1, &rowcnt));
  HDB_CALL(scratch_t(ctx, kSlotRowCursor, n + 2, &cursor));
  HDB_CALL(scratch_t(ctx, kSlotRowTmp, std::max<int64_t>(cap, 1), &tmp));
  HDB_CALL(scratch_t(ctx, kSlotRowUniq, 2 * n + 2, &uniq));
  HDB_CALL(scratch_t(ctx, kSlotMisc1, n + 1, &rowoff));
  HDB_CALL(scratch_t(ctx, kSlotAhFlag, n + 8, &ahrank));
  HDB_CALL(scratch_t(ctx, kSlotMisc2, n + 8, &misc));
  int32_t* total = uniq + n + 1;
  int32_t* big = misc;     // row ids of long rows (n)
  int32_t* nbig = cursor + n;  // two spare cursor slots: long-row count, longest row
  int32_t* maxlen = cursor + n + 1;
  HDB_CALL(fill_i32(ctx, rowcnt, n + 1, 0));
  HDB_CALL(fill_i32(ctx, cursor, n + 1, 0));
  HDB_CALL(fill_i32(ctx, ahrank, n, -1));
  if (h > 0) {
    hipLaunchKernelGGL(k_ah_rank, dim3(grid_for(h, 256)), dim3(256), 0, ctx->stream, ah, h, ahrank);
    HDB_LAUNCHED();
  }
  const unsigned g = grid_for(n * k, 256);
  if (upper) hipLaunchKernelGGL(k_edge_count<true>, dim3(g), dim3(256), 0, ctx->stream, lists, n, k, rowcnt);
  else hipLaunchKernelGGL(k_edge_count<false>, dim3(g), dim3(256), 0, ctx->stream, lists, n, k, rowcnt);
  HDB_LAUNCHED();
  HDB_CALL(scan_i32_to_i64(ctx, rowcnt, n, rowoff));
  if (upper)
    hipLaunchKernelGGL(k_edge_scatter<true>, dim3(g), dim3(256), 0, ctx->stream, lists, n, k, rowoff, cursor, tmp);
  else
    hipLaunchKernelGGL(k_edge_scatter<false>, dim3(g), dim3(256), 0, ctx->stream, lists, n, k, rowoff, cursor, tmp);
  HDB_LAUNCHED();
  HDB_CALL(fill_i32(ctx, nbig, 2, 0));
  if (upper)
    hipLaunchKernelGGL(k_row_sort_small<true>, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, rowoff, n, tmp, uniq,
                       total, ahrank, h, big, nbig, maxlen);
  else
    hipLaunchKernelGGL(k_row_sort_small<false>, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, rowoff, n, tmp, uniq,
                       total, ahrank, h, big, nbig, maxlen);
  HDB_LAUNCHED();
  int32_t nbl[2] = {0, 0};
  HDB_HIP(hipMemcpyAsync(nbl, nbig, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HDB_HIP(hipStreamSynchronize(ctx->stream));
  const int32_t nb = nbl[0];
  if (nb > 0) {
    // the longest row decides the LDS size
    const int64_t longest = nbl[1];
    if (longest > kBigRowMax) return ctx->fail(HDB_EUNSUPPORTED, "vertex degree above 16384 in the symmetrised graph");
    int64_t P = 1;
    while (P < longest) P <<= 1;
    if (upper)
      hipLaunchKernelGGL(k_row_sort_big<true>, dim3(nb), dim3(256), P * sizeof(uint32_t), ctx->stream, rowoff, big,
                         tmp, uniq, total, ahrank, h);
    else
      hipLaunchKernelGGL(k_row_sort_big<false>, dim3(nb), dim3(256), P * sizeof(uint32_t), ctx->stream, rowoff, big,
                         tmp, uniq, total, ahrank, h);
    HDB_LAUNCHED();
  }
  HDB_CALL(scan_i32_to_i64(ctx, total, n, nindex));
  HDB_CALL(read_i64(ctx, nindex + n, edges_out));
  if (!nlist) return HDB_OK;  // size query
  if (!w || (upper && !src)) return ctx->fail(HDB_EINVAL, "null output array");
  if (upper)
    hipLaunchKernelGGL(k_row_fill<true>, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, X, d, rowoff, tmp, uniq,
                       nindex, n, core, ahrank, ah, h, nlist, w, src);
  else
    hipLaunchKernelGGL(k_row_fill<false>, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, X, d, rowoff, tmp, uniq,
                       nindex, n, core, ahrank, ah, h, nlist, w, src);
  HDB_LAUNCHED();
  return HDB_OK;
}

}  // namespace hdb
