// Learned initialiser for NN-Descent on gfx950 (no reference counterpart: it
// replaces the random index init of InitKNNGraphIndexKernel,
// gpuknn/nndescent.cu:1673-1692, with bucket co-members).
//
// Network: d_in -> 128 -> 128 -> 12, GELU after the two hidden layers.  Every
// product runs on v_mfma_f32_16x16x32_bf16: operands (inputs, activations,
// gradients and the FP32 master weights) are rounded to BF16 when the
// fragments are formed, accumulation is FP32.  Lane layout of the
// instruction: A/B lane l holds row (l % 16), k = 8*(l/16) .. +7; the result
// lane l holds rows 4*(l/16) + q (q = 0..3), column l % 16.
//
// Training pairs: anchor a, positive = a kNN neighbour of a (from a plain
// NN-Descent graph of a subsample), negative = a random point.  With
// t = tanh(z):  L = mean_batch[ |t_a - t_p|^2 / 12 + (t_a . t_n / 12)^2 ]
//                   + sum_bits mean_rows(t)^2
// (agreeing signs for neighbours, uncorrelated codes for random pairs,
// balanced bits).  The 12 sign bits of z are the bucket code.
#include <algorithm>
#include <cmath>

#include "hdb_internal.hpp"

namespace hdb {

constexpr int kHid = 128;
constexpr int kCode = 12;
constexpr int kOutP = 16;  // padded output width
constexpr int kMlpWaves = 2;  // waves per forward block (LDS: 16 rows x (d_in + 2 x 128) floats each)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MlpLayout {
  int kin, kinp;  // input width and its multiple of 32
  int64_t w1, b1, w2, b2, w3, b3, total;
};

static MlpLayout mlp_layout(int kin) {
  MlpLayout L;
  L.kin = kin;
  L.kinp = (kin + 31) & ~31;
  L.w1 = 0;
  L.b1 = L.w1 + static_cast<int64_t>(L.kinp) * kHid;
  L.w2 = L.b1 + kHid;
  L.b2 = L.w2 + kHid * kHid;
  L.w3 = L.b2 + kHid;
  L.b3 = L.w3 + kHid * kOutP;
  L.total = L.b3 + kOutP;
  return L;
}

// GELU x * Phi(x) with the normal tail Q(x) = 0.5 erfc(|x| / sqrt 2) from
// Abramowitz & Stegun 7.1.26 (|erfc error| <= 1.5e-7): erfc(z) = t (a1 + t
// (a2 + t (a3 + t (a4 + t a5)))) exp(-z^2), t = 1 / (1 + p z).  Branch-free;
// exp(-x^2 / 2) is shared with the density used by the derivative.
struct NormTail {
  float q;    // Q(|x|)
  float e;    // exp(-x^2 / 2)
};
__device__ __forceinline__ NormTail norm_tail(float x) {
  // z = |x| / sqrt 2 is folded into p, and the factor 0.5 into a1..a5.
  const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.2316418882663604f, __builtin_fabsf(x), 1.0f));
  float poly = __builtin_fmaf(t, 0.5307027145f, -0.7265760135f);
  poly = __builtin_fmaf(t, poly, 0.7107068705f);
  poly = __builtin_fmaf(t, poly, -0.142248368f);
  poly = __builtin_fmaf(t, poly, 0.127414796f);
  NormTail r;
  r.e = __builtin_amdgcn_exp2f(x * x * -0.7213475204444817f);
  r.q = t * poly * r.e;
  return r;
}
__device__ __forceinline__ float gelu(float x) {
  const float q = norm_tail(x).q;
  return x >= 0.0f ? x - x * q : x * q;
}
__device__ __forceinline__ float gelu_grad(float x) {
  const NormTail n = norm_tail(x);
  const float cdf = x >= 0.0f ? 1.0f - n.q : n.q;
  return cdf + x * 0.39894228040143268f * n.e;
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// One wave computes out[16 x N] = act(in[16 x K] . W[K x N] + bias) with
// `in` a row-major LDS tile (stride ldi) and W row-major in global memory
// (N columns, stride N).  Results go to LDS (stride ldo) and optionally to
// global pre-activation / activation rows.
template <bool GELU>
__device__ __forceinline__ void wave_layer(const float* in, int ldi, int K, const float* __restrict__ W,
                                           const float* __restrict__ bias, int N, float* out, int ldo,
                                           float* gpre, float* gact, int64_t grow0, int64_t rows_valid) {
  const unsigned l = __lane_id();
  const int r16 = l & 15, kq = l >> 4;
  for (int n0 = 0; n0 < N; n0 += 16) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32) {
      bf16x8 a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + 8 * kq + e;
        a[e] = static_cast<__bf16>(in[r16 * ldi + k]);
        b[e] = static_cast<__bf16>(W[static_cast<int64_t>(k) * N + n0 + r16]);
      }
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = 4 * kq + q, col = n0 + r16;
      const float pre = acc[q] + bias[col];
      const float v = GELU ? gelu(pre) : pre;
      out[row * ldo + col] = v;
      if (row < rows_valid) {
        if (gpre) gpre[(grow0 + row) * N + col] = pre;
        if (gact) gact[(grow0 + row) * N + col] = v;
      }
    }
  }
}

// Forward pass over R rows (row r = point rows[r], or r when rows is null).
// TRAIN stores P1, H1, P2, H2 (R x 128) for the backward pass.  Z: R x 16.
template <bool TRAIN>
__global__ __launch_bounds__(kMlpWaves * 64) void k_mlp_fwd(const float* __restrict__ X, int d, const int32_t* __restrict__ rows,
                                                 int64_t R, const float* __restrict__ prm, MlpLayout L,
                                                 float* __restrict__ P1, float* __restrict__ H1,
                                                 float* __restrict__ P2, float* __restrict__ H2,
                                                 float* __restrict__ Z) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  const unsigned l = __lane_id();
  const unsigned w = threadIdx.x >> 6;
  const int ldi = L.kinp + 4, lda = kHid + 4;
  float* in = lds + w * (16 * ldi + 2 * 16 * lda);
  float* a1 = in + 16 * ldi;
  float* a2 = a1 + 16 * lda;
  float* zo = in;  // the input tile is dead once layer 1 is done
  const int64_t r0 = (static_cast<int64_t>(blockIdx.x) * kMlpWaves + w) * 16;
  if (r0 >= R) return;  // no block barriers in this kernel
  const int64_t valid = R - r0 < 16 ? R - r0 : 16;
  for (int i = 0; i < 16; ++i) {
    const bool ok = i < valid;
    const int64_t src = ok ? (rows ? rows[r0 + i] : r0 + i) : 0;
    for (int k = l; k < L.kinp; k += 64) in[i * ldi + k] = (ok && k < d) ? X[src * d + k] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  wave_layer<true>(in, ldi, L.kinp, prm + L.w1, prm + L.b1, kHid, a1, lda, TRAIN ? P1 : nullptr,
                   TRAIN ? H1 : nullptr, r0, valid);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  wave_layer<true>(a1, lda, kHid, prm + L.w2, prm + L.b2, kHid, a2, lda, TRAIN ? P2 : nullptr,
                   TRAIN ? H2 : nullptr, r0, valid);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  wave_layer<false>(a2, lda, kHid, prm + L.w3, prm + L.b3, kOutP, zo, kOutP + 4, nullptr, Z, r0, valid);
}

// Backward over rows: dH2 = dZ W3^T, dP2 = dH2 * gelu'(P2), dH1 = dP2 W2^T,
// dP1 = dH1 * gelu'(P1).  Per wave 16 rows.
__global__ __launch_bounds__(256) void k_mlp_bwd_rows(int64_t R, const float* __restrict__ prm, MlpLayout L,
                                                      const float* __restrict__ P1, const float* __restrict__ P2,
                                                      const float* __restrict__ dZ, float* __restrict__ dP2,
                                                      float* __restrict__ dP1) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  const unsigned l = __lane_id();
  const unsigned w = threadIdx.x >> 6;
  const int lda = kHid + 4;
  float* gz = lds + w * (16 * 36 + 16 * lda);  // dZ tile padded to K = 32
  float* g2 = gz + 16 * 36;
  const int64_t r0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * 16;
  if (r0 >= R) return;
  const int64_t valid = R - r0 < 16 ? R - r0 : 16;
  const int r16 = l & 15, kq = l >> 4;
  for (int i = l; i < 16 * 32; i += 64) {
    const int row = i >> 5, c = i & 31;
    gz[row * 36 + c] = (row < valid && c < kOutP) ? dZ[(r0 + row) * kOutP + c] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  const float* W3 = prm + L.w3;  // [128][16]
  const float* W2 = prm + L.w2;  // [128][128]
  for (int n0 = 0; n0 < kHid; n0 += 16) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * kq + e;  // output unit of layer 3
      a[e] = static_cast<__bf16>(gz[r16 * 36 + k]);
      b[e] = static_cast<__bf16>(k < kOutP ? W3[(n0 + r16) * kOutP + k] : 0.0f);
    }
    acc = mfma(a, b, acc);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = 4 * kq + q, col = n0 + r16;
      const float pre = row < valid ? P2[(r0 + row) * kHid + col] : 0.0f;
      const float g = acc[q] * gelu_grad(pre);
      g2[row * lda + col] = row < valid ? g : 0.0f;
      if (row < valid) dP2[(r0 + row) * kHid + col] = g;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int n0 = 0; n0 < kHid; n0 += 16) {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < kHid; k0 += 32) {
      bf16x8 a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = k0 + 8 * kq + e;  // unit of hidden layer 2
        a[e] = static_cast<__bf16>(g2[r16 * lda + k]);
        b[e] = static_cast<__bf16>(W2[(n0 + r16) * kHid + k]);
      }
      acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = 4 * kq + q, col = n0 + r16;
      if (row < valid) dP1[(r0 + row) * kHid + col] = acc[q] * gelu_grad(P1[(r0 + row) * kHid + col]);
    }
  }
}

// Weight gradients G[m][n] = sum_r A[r][m] * B[r][n] over R rows, split in
// chunks of `chunk` rows (gridDim.z chunks); partials[z][M][N].  A rows come
// from X (gathered, width d, zero past d) when rows != null, else from a
// dense R x lda matrix.
__global__ __launch_bounds__(64) void k_mlp_wgrad(const float* __restrict__ A, int lda, const float* __restrict__ X,
                                                  int d, const int32_t* __restrict__ rows, const float* __restrict__ B,
                                                  int ldb, int64_t R, int M, int N, int64_t chunk,
                                                  float* __restrict__ part) {
  const unsigned l = __lane_id();
  const int r16 = l & 15, kq = l >> 4;
  const int n0 = blockIdx.x * 16, m0 = blockIdx.y * 16;
  const int64_t rbeg = blockIdx.z * chunk;
  const int64_t rend = rbeg + chunk < R ? rbeg + chunk : R;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t k0 = rbeg; k0 < rend; k0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int64_t r = k0 + 8 * kq + e;
      float av = 0.0f, bv = 0.0f;
      if (r < rend) {
        const int m = m0 + r16;
        if (rows) av = m < d ? X[static_cast<int64_t>(rows[r]) * d + m] : 0.0f;
        else av = A[r * lda + m];
        bv = B[r * ldb + n0 + r16];
      }
      a[e] = static_cast<__bf16>(av);
      b[e] = static_cast<__bf16>(bv);
    }
    acc = mfma(a, b, acc);
  }
  float* out = part + static_cast<int64_t>(blockIdx.z) * M * N;
#pragma unroll
  for (int q = 0; q < 4; ++q) out[(m0 + 4 * kq + q) * N + n0 + r16] = acc[q];
}

// grad[i] = sum_z part[z][i] (fixed order), i < M*N.
__global__ void k_mlp_reduce(const float* __restrict__ part, int64_t splits, int64_t mn, float* __restrict__ grad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float s = 0.0f;
  for (int64_t z = 0; z < splits; ++z) s += part[z * mn + i];
  grad[i] = s;
}

// Bias gradients, first pass: column sums of rows [z*256, (z+1)*256) of a
// dense R x n matrix into part[z][c] (block z; n divides 256, so 256 / n
// threads share a column, each with four running sums over its rows, then
// combined in a fixed order).
__global__ __launch_bounds__(256) void k_mlp_colsum(const float* __restrict__ G, int64_t R, int n,
                                                    float* __restrict__ part) {
  __shared__ float acc[256];
  const int t = static_cast<int>(threadIdx.x);
  const int c = t % n, h = t / n, hp = 256 / n;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 256;
  float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t r = r0 + h; r < r0 + 256 && r < R; r += 4 * hp) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t ru = r + u * hp;
      if (ru < r0 + 256 && ru < R) s[u] += G[ru * n + c];
    }
  }
  acc[t] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  if (h == 0) {
    float tot = 0.0f;
    for (int q = 0; q < hp; ++q) tot += acc[q * n + c];
    part[static_cast<int64_t>(blockIdx.x) * n + c] = tot;
  }
}

// Mean of tanh(z) per code bit over the R rows: block z sums rows
// [z*256, (z+1)*256) into part[z][b]; k_mlp_tmean_final adds the parts in
// order and divides by R.
__global__ __launch_bounds__(256) void k_mlp_tmean(const float* __restrict__ Z, int64_t R, float* __restrict__ part) {
  __shared__ float acc[kCode][256];
  const int64_t r = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  for (int b = 0; b < kCode; ++b) acc[b][threadIdx.x] = r < R ? tanhf(Z[r * kOutP + b]) : 0.0f;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (static_cast<int>(threadIdx.x) < st)
      for (int b = 0; b < kCode; ++b) acc[b][threadIdx.x] += acc[b][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < kCode) part[static_cast<int64_t>(blockIdx.x) * kCode + threadIdx.x] = acc[threadIdx.x][0];
}

__global__ void k_mlp_tmean_final(const float* __restrict__ part, int64_t blocks, int64_t R, float* __restrict__ mean) {
  const int b = static_cast<int>(threadIdx.x);
  if (b >= kCode) return;
  float s = 0.0f;
  for (int64_t z = 0; z < blocks; ++z) s += part[z * kCode + b];
  mean[b] = s / static_cast<float>(R);
}

// Loss and dZ for B (anchor, positive, negative) triples stored as rows
// [0,B), [B,2B), [2B,3B).  loss_part[b] = per-triple loss (balance added
// once by the host from `mean`).
__global__ void k_mlp_loss(const float* __restrict__ Z, int64_t B, const float* __restrict__ mean,
                           float* __restrict__ dZ, float* __restrict__ loss_part) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float* za = Z + b * kOutP;
  const float* zp = Z + (B + b) * kOutP;
  const float* zn = Z + (2 * B + b) * kOutP;
  float ta[kCode], tp[kCode], tn[kCode];
  float lpos = 0.0f, s = 0.0f;
  for (int i = 0; i < kCode; ++i) {
    ta[i] = tanhf(za[i]);
    tp[i] = tanhf(zp[i]);
    tn[i] = tanhf(zn[i]);
    const float df = ta[i] - tp[i];
    lpos += df * df;
    s += ta[i] * tn[i];
  }
  lpos /= kCode;
  s /= kCode;
  loss_part[b] = lpos + s * s;
  const float inv_b = 1.0f / static_cast<float>(B);
  const float bal = 2.0f / static_cast<float>(3 * B);
  float* ga = dZ + b * kOutP;
  float* gp = dZ + (B + b) * kOutP;
  float* gn = dZ + (2 * B + b) * kOutP;
  for (int i = 0; i < kCode; ++i) {
    const float dta = (2.0f / kCode) * (ta[i] - tp[i]) * inv_b + (2.0f * s / kCode) * tn[i] * inv_b + bal * mean[i];
    const float dtp = -(2.0f / kCode) * (ta[i] - tp[i]) * inv_b + bal * mean[i];
    const float dtn = (2.0f * s / kCode) * ta[i] * inv_b + bal * mean[i];
    ga[i] = dta * (1.0f - ta[i] * ta[i]);
    gp[i] = dtp * (1.0f - tp[i] * tp[i]);
    gn[i] = dtn * (1.0f - tn[i] * tn[i]);
  }
  for (int i = kCode; i < kOutP; ++i) ga[i] = gp[i] = gn[i] = 0.0f;
}

// Loss of the partition-code target (rows [0, 3B), target s_b = +1 when bit
// b of codes[rows[r]] is set, else -1) plus the kNN pair term over rows
// (b, B + b):  L = sum_r sum_b (t_b - s_b)^2 / (12 * 3B)
//                + sum_{b<B} |t_b - t_{B+b}|^2 / (12 * B);  t = tanh(z).
// loss_part[r] is row r's share of L.
__global__ void k_mlp_loss_bits(const float* __restrict__ Z, int64_t B, const int32_t* __restrict__ rows,
                                const uint32_t* __restrict__ codes, float* __restrict__ dZ,
                                float* __restrict__ loss_part) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= 3 * B) return;
  const uint32_t c = codes[rows[r]];
  const bool paired = r < 2 * B;
  const int64_t other = r < B ? r + B : r - B;
  const float inv_r = 1.0f / static_cast<float>(3 * B);
  const float inv_b = 1.0f / static_cast<float>(B);
  float lt = 0.0f, lp = 0.0f;
  float* g = dZ + r * kOutP;
  for (int i = 0; i < kCode; ++i) {
    const float t = tanhf(Z[r * kOutP + i]);
    const float e = t - (((c >> i) & 1u) ? 1.0f : -1.0f);
    lt += e * e;
    float dt = (2.0f / kCode) * e * inv_r;
    if (paired) {
      const float df = t - tanhf(Z[other * kOutP + i]);
      lp += df * df;
      dt += (2.0f / kCode) * df * inv_b;
    }
    g[i] = dt * (1.0f - t * t);
  }
  for (int i = kCode; i < kOutP; ++i) g[i] = 0.0f;
  loss_part[r] = lt / kCode * inv_r + (r < B ? lp / kCode * inv_b : 0.0f);
}

// Balanced random-projection tree codes: at level l every cell (the l code
// bits so far) splits at the median of x . r(l, cell); r is a Gaussian
// direction drawn by Box-Muller from xorshift64*(seed, l, cell).  Keys are
// (cell << 52) | (orderable projection bits << 20) | point id (n <= 2^20).
__global__ void k_rp_keys(const float* __restrict__ X, int64_t n, int d, int level, uint64_t seed,
                          const uint32_t* __restrict__ code, uint64_t* __restrict__ keys) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = code[i];
  uint64_t h = xorshift64star(seed ^ (static_cast<uint64_t>(level + 1) << 44) ^
                              (static_cast<uint64_t>(c) * 0x9E3779B97F4A7C15ull));
  float proj = 0.0f;
  for (int k = 0; k < d; k += 2) {
    h = xorshift64star(h);
    const float u1 = (static_cast<float>(h >> 40) + 1.0f) * (1.0f / 16777217.0f);
    h = xorshift64star(h);
    const float u2 = static_cast<float>(h >> 40) * (1.0f / 16777216.0f);
    const float rad = sqrtf(-2.0f * logf(u1));
    proj = __builtin_fmaf(X[i * d + k], rad * cosf(6.2831853f * u2), proj);
    if (k + 1 < d) proj = __builtin_fmaf(X[i * d + k + 1], rad * sinf(6.2831853f * u2), proj);
  }
  uint32_t b = __float_as_uint(proj);
  b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  keys[i] = (static_cast<uint64_t>(c) << 52) | (static_cast<uint64_t>(b) << 20) | static_cast<uint64_t>(i);
}

__global__ void k_rp_cells(const uint64_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ cell) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < n) cell[s] = static_cast<uint32_t>(keys[s] >> 52);
}

// The lower half (floor) of each cell's sorted run gets bit 0, the rest 1.
__global__ void k_rp_split(const uint64_t* __restrict__ keys, int64_t n, const int32_t* __restrict__ cs,
                           const int32_t* __restrict__ ce, uint32_t* __restrict__ code) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t k = keys[s];
  const uint32_t c = static_cast<uint32_t>(k >> 52);
  const int64_t id = static_cast<int64_t>(k & 0xFFFFFull);
  const int32_t half = (ce[c] - cs[c]) / 2;
  code[id] = (c << 1) | ((s - cs[c]) >= half ? 1u : 0u);
}

// Triples for step `step`: anchor uniform, positive = a uniform entry of the
// anchor's kNN list; the negative is, three times in four, the end of a
// kNegWalk-step random walk over the kNN lists (a point of the anchor's own
// region beyond its neighbours, so decorrelating it splits large clusters
// into several buckets) and otherwise uniform (xorshift64* streams).
constexpr int kNegWalk = 3;

__global__ void k_mlp_batch(const hdb_nnd_element* __restrict__ lists, int64_t n, int64_t B, uint64_t seed,
                            int32_t* __restrict__ rows) {
  const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  uint64_t h = xorshift64star(seed ^ (static_cast<uint64_t>(b) * 0x9E3779B97F4A7C15ull));
  const int64_t a = static_cast<int64_t>(h % static_cast<uint64_t>(n));
  h = xorshift64star(h);
  uint32_t pid = label_id(lists[a * kK + static_cast<int>(h % kK)].label);
  if (pid == kInvalidId || static_cast<int64_t>(pid) >= n) pid = static_cast<uint32_t>(a);
  h = xorshift64star(h);
  int64_t neg;
  if ((h & 3u) != 0) {
    neg = a;
    for (int s = 0; s < kNegWalk; ++s) {
      h = xorshift64star(h);
      const uint32_t nx = label_id(lists[neg * kK + static_cast<int>(h % kK)].label);
      if (nx != kInvalidId && static_cast<int64_t>(nx) < n) neg = nx;
    }
  } else {
    h = xorshift64star(h);
    neg = static_cast<int64_t>(h % static_cast<uint64_t>(n));
  }
  rows[b] = static_cast<int32_t>(a);
  rows[B + b] = static_cast<int32_t>(pid);
  rows[2 * B + b] = static_cast<int32_t>(neg);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t np, float lr, float c1, float c2, const uint8_t* __restrict__ mask) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= np || !mask[i]) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// ------------------------------------------------------- inference path --
// Weights packed into MFMA fragment order (BF16): fragment f holds, for lane
// l, the 8 weights W[k0 + 8*(l/16) + e][n0 + l%16] (e = 0..7) of column tile
// n0 and k step k0.  Layer 1 has (kinp/32) x 8 fragments (column tile major),
// layer 2 4 x 8, layer 3 4 (N = 16).  These are exactly the BF16 operands
// k_mlp_fwd forms, so both kernels compute the same values.
__host__ __device__ inline int infer_frags(int kinp) { return (kinp / 32) * 8 + 32 + 4; }

__global__ void k_mlp_pack(const float* __restrict__ prm, MlpLayout L, __bf16* __restrict__ packed) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int ks1 = L.kinp / 32;
  const int nf1 = ks1 * 8, nf2 = 32;
  if (t >= static_cast<int64_t>(infer_frags(L.kinp)) * 64) return;
  int f = static_cast<int>(t >> 6);
  const int lane = static_cast<int>(t & 63);
  const float* W;
  int N, nt, ks;
  if (f < nf1) {
    W = prm + L.w1, N = kHid, nt = f / ks1, ks = f % ks1;
  } else if (f < nf1 + nf2) {
    f -= nf1;
    W = prm + L.w2, N = kHid, nt = f / 4, ks = f % 4;
  } else {
    f -= nf1 + nf2;
    W = prm + L.w3, N = kOutP, nt = 0, ks = f;
  }
  const int r16 = lane & 15, kq = lane >> 4;
  for (int e = 0; e < 8; ++e) {
    const int k = ks * 32 + 8 * kq + e;
    packed[t * 8 + e] = static_cast<__bf16>(W[static_cast<int64_t>(k) * N + nt * 16 + r16]);
  }
}

// Bucket keys written by the inference kernel in place of Z: key =
// (sign code of outputs 0..11) << qbits | the projection x.u quantised to
// qbits bits over [-lim, lim].
struct InferKeys {
  const float* u;
  float lim;
  int qbits;
  uint32_t* keys;
  uint32_t* ids;
};

__device__ __forceinline__ uint32_t proj_q(const float* __restrict__ x, const float* __restrict__ u, int d,
                                           float lim, int qbits) {
  float s = 0.0f;
  for (int k = 0; k < d; ++k) s = __builtin_fmaf(x[k], u[k], s);
  float t = (s / lim) * 0.5f + 0.5f;
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  return static_cast<uint32_t>(t * static_cast<float>((1u << qbits) - 1u));
}

// Projection bits of the bucket keys: enough levels that a level holds about
// 2^8 points or fewer at n points, 12 at least and 20 at most, so a key of
// 12 + qbits bits sorts in three passes up to about a million points and in
// four beyond.  The order inside a projection level is arbitrary, so coarse
// levels at large n fill a window with unrelated points.
inline int proj_bits(int64_t n) {
  int lg = 0;
  while ((int64_t{1} << lg) < n) ++lg;
  const int q = lg - 8;
  return q < 12 ? 12 : (q > 20 ? 20 : q);
}

// Z = forward(X) for R rows: the block keeps the packed weights and biases in
// LDS and each wave strides over 16-row tiles, with activations stored as
// BF16 (the value they take as the next layer's MFMA operand).  KEYS: the
// rows' bucket keys instead of Z (each row's 12 code bits are one ballot of
// the lanes holding its outputs).
template <bool KEYS>
__global__ __launch_bounds__(1024) void k_mlp_infer(const float* __restrict__ X, int d, int64_t R,
                                                   const __bf16* __restrict__ packed, const float* __restrict__ prm,
                                                   MlpLayout L, float* __restrict__ Z, InferKeys ko) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_b[];
  const int ks1 = L.kinp / 32;
  const int nfrag = infer_frags(L.kinp);
  const int waves = static_cast<int>(blockDim.x >> 6);
  __bf16* sw = reinterpret_cast<__bf16*>(lds_b);
  float* sb = reinterpret_cast<float*>(sw + nfrag * 512);  // b1 | b2 | b3: 128 + 128 + 16
  __bf16* act = reinterpret_cast<__bf16*>(sb + 272);
  for (int i = threadIdx.x; i < nfrag * 64; i += blockDim.x)
    reinterpret_cast<uint4*>(sw)[i] = reinterpret_cast<const uint4*>(packed)[i];
  for (int i = threadIdx.x; i < 272; i += blockDim.x)
    sb[i] = i < kHid ? prm[L.b1 + i] : (i < 2 * kHid ? prm[L.b2 + i - kHid] : prm[L.b3 + i - 2 * kHid]);
  __syncthreads();
  const unsigned l = __lane_id();
  const int w = static_cast<int>(threadIdx.x >> 6);
  const int ldi = L.kinp + 8, lda = kHid + 8;
  // one 16-row activation tile per wave: layer 2 reads all of a1 into its
  // accumulators before writing its output back over it
  __bf16* in = act + static_cast<int64_t>(w) * (16 * ldi + 16 * lda);
  __bf16* a1 = in + 16 * ldi;
  __bf16* a2 = a1;
  const int r16 = static_cast<int>(l & 15), kq = static_cast<int>(l >> 4);
  const int64_t ntiles = (R + 15) / 16;
  for (int64_t tile = static_cast<int64_t>(blockIdx.x) * waves + w; tile < ntiles;
       tile += static_cast<int64_t>(gridDim.x) * waves) {
    const int64_t r0 = tile * 16;
    for (int idx = static_cast<int>(l); idx < 16 * L.kinp; idx += 64) {
      const int i = idx / L.kinp, k = idx - i * L.kinp;
      const int64_t r = r0 + i;
      in[i * ldi + k] = static_cast<__bf16>((r < R && k < d) ? X[r * d + k] : 0.0f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {  // layer 1: the 8 column tiles interleaved so no MFMA waits on the last
      f32x4 acc[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      for (int ks = 0; ks < ks1; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(in + r16 * ldi + ks * 32 + 8 * kq);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt)
          acc[nt] = mfma(a, *reinterpret_cast<const bf16x8*>(sw + ((nt * ks1 + ks) * 64 + l) * 8), acc[nt]);
      }
#pragma unroll
      for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 4 * kq + q, col = nt * 16 + r16;
          a1[row * lda + col] = static_cast<__bf16>(gelu(acc[nt][q] + sb[col]));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {  // layer 2
      f32x4 acc[8];
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) acc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(a1 + r16 * lda + ks * 32 + 8 * kq);
#pragma unroll
        for (int nt = 0; nt < 8; ++nt)
          acc[nt] = mfma(a, *reinterpret_cast<const bf16x8*>(sw + ((ks1 * 8 + nt * 4 + ks) * 64 + l) * 8), acc[nt]);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 4 * kq + q, col = nt * 16 + r16;
          a2[row * lda + col] = static_cast<__bf16>(gelu(acc[nt][q] + sb[kHid + col]));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {  // layer 3
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(a2 + r16 * lda + ks * 32 + 8 * kq);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(sw + ((ks1 * 8 + 32 + ks) * 64 + l) * 8);
        acc = mfma(a, b, acc);
      }
      if (KEYS) {
        // lane 16 kq + c holds output c of rows 4 kq + q
        uint64_t m[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) m[q] = __ballot(r16 < kCode && acc[q] + sb[2 * kHid + r16] > 0.0f);
        const int64_t row = r0 + 4 * kq + r16;
        if (r16 < 4 && row < R) {
          uint64_t mm = m[0];
#pragma unroll
          for (int q = 1; q < 4; ++q)
            if (r16 == q) mm = m[q];
          const uint32_t code = static_cast<uint32_t>(mm >> (16 * kq)) & ((1u << kCode) - 1u);
          ko.keys[row] = (code << ko.qbits) | proj_q(X + row * d, ko.u, d, ko.lim, ko.qbits);
          ko.ids[row] = static_cast<uint32_t>(row);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t row = r0 + 4 * kq + q;
          if (row < R) Z[row * kOutP + r16] = acc[q] + sb[2 * kHid + r16];
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Points permuted into the bucket order: row r of Xs is point order[r].
__global__ void k_mlp_permute(const float* __restrict__ X, const uint32_t* __restrict__ order, int64_t n, int d,
                              float* __restrict__ Xs) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int64_t r = t / d;
  const int k = static_cast<int>(t - r * d);
  Xs[t] = X[static_cast<int64_t>(order[r]) * d + k];
}

constexpr int kSeedTables = 8;  // bucket tables at most
constexpr int kSeedSlots = 64;  // candidate slots per point (one per lane in the merge)
constexpr int kSeedPts = 256;   // points per block, one per lane

template <int W>
__device__ __forceinline__ int64_t seed_lo(int64_t r, int64_t n) {
  int64_t lo = r - W / 2;
  if (lo > n - W - 1) lo = n - W - 1;
  return lo < 0 ? 0 : lo;
}

// Lane r (a position in the table's bucket order) keeps the K nearest of
// the W positions around it (itself skipped), ranked by 32-bit keys: the
// distance bits with the low log2(W) replaced by the window slot, kept in
// registers by a median-of-three insertion.  The winners' exact keys
// are recomputed.  Rows come from LDS when STAGED (stride d + 1, an odd
// stride across lanes); D > 0 fixes d and keeps the point's own row in
// registers.
template <int D, int W, int K, bool STAGED>
__global__ __launch_bounds__(kSeedPts) void k_mlp_seed(const float* __restrict__ Xs,
                                                        const uint32_t* __restrict__ order, int64_t n, int d,
                                                        int slot0, uint64_t* __restrict__ cand) {
  extern __shared__ __attribute__((aligned(16))) float lds_rows[];
  static_assert((W & (W - 1)) == 0 && W <= 512, "window is a power of two");
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kSeedPts;
  const int64_t sb = seed_lo<W>(r0, n);
  const float* rows = Xs + sb * d;
  int dp = d;
  if (STAGED) {
    const int64_t last = r0 + kSeedPts - 1 < n ? r0 + kSeedPts - 1 : n - 1;
    int64_t se = seed_lo<W>(last, n) + W + 1;
    if (se > n) se = n;
    dp = d + 1;
    for (int64_t t = threadIdx.x; t < (se - sb) * d; t += kSeedPts) {
      const int64_t rr = t / d;
      lds_rows[rr * dp + (t - rr * d)] = Xs[sb * d + t];
    }
    __syncthreads();
    rows = lds_rows;
  }
  const int64_t r = r0 + threadIdx.x;
  if (r >= n) return;  // after the block's only barrier
  const int64_t lo = seed_lo<W>(r, n);
  const float* self_row = rows + (r - sb) * dp;
  float own[D > 0 ? D : 1];
  if (D > 0) {
#pragma unroll
    for (int k = 0; k < (D > 0 ? D : 1); ++k) own[k] = self_row[k];
  }
  uint32_t best[K];
#pragma unroll
  for (int j = 0; j < K; ++j) best[j] = 0xFFFFFFFFu;
  for (int slot = 0; slot < W; ++slot) {
    int64_t p = lo + slot;
    if (p >= r) p += 1;
    if (p >= n) break;  // only when n <= W: the remaining slots are past the end too
    const float* pr = rows + (p - sb) * dp;
    float acc = 0.0f;
    if (D > 0) {
#pragma unroll
      for (int k = 0; k < (D > 0 ? D : 1); ++k) {
        const float diff = pr[k] - own[k];
        acc = __builtin_fmaf(diff, diff, acc);
      }
    } else {
      acc = d2_any(pr, self_row, d);
    }
    uint32_t key = (__float_as_uint(acc) & ~static_cast<uint32_t>(W - 1)) | static_cast<uint32_t>(slot);
    top_insert_med3<K>(best, key);
  }
  uint64_t* out = cand + static_cast<int64_t>(order[r]) * kSeedSlots + slot0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    uint64_t key = kKeyMax;
    if (best[j] != 0xFFFFFFFFu) {
      int64_t p = lo + static_cast<int64_t>(best[j] & static_cast<uint32_t>(W - 1));
      if (p >= r) p += 1;
      key = make_key(d2_any(rows + (p - sb) * dp, self_row, d), order[p], 1u);
    }
    out[j] = key;
  }
}

// k_mlp_seed for d = D in {12, 32}: rows staged at a stride of D floats when
// D / 4 is odd, else D + 4 (16-byte rows whose starts fall in distinct
// 4-bank groups for 16 consecutive lanes), read as float4, and ranked by
// |x|^2 + |y|^2 - 2 x.y (clamped at 0) with the rows' norms staged beside
// them; the winners' exact keys are recomputed as in k_mlp_seed.
template <int D, int W, int K>
__global__ __launch_bounds__(kSeedPts) void k_mlp_seed4(const float* __restrict__ Xs,
                                                         const uint32_t* __restrict__ order, int64_t n,
                                                         int slot0, uint64_t* __restrict__ cand) {
  static_assert(D % 4 == 0 && D <= 32, "float4 rows");
  static_assert((W & (W - 1)) == 0 && W <= 512, "window is a power of two");
  constexpr int S = ((D / 4) & 1) ? D : D + 4;
  extern __shared__ __attribute__((aligned(16))) float lds_rows[];
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kSeedPts;
  const int64_t sb = seed_lo<W>(r0, n);
  const int64_t last = r0 + kSeedPts - 1 < n ? r0 + kSeedPts - 1 : n - 1;
  int64_t se = seed_lo<W>(last, n) + W + 1;
  if (se > n) se = n;
  const int nrows = static_cast<int>(se - sb);
  float* s_norm = lds_rows + (kSeedPts + W + 1) * S;
  for (int rr = threadIdx.x; rr < nrows; rr += kSeedPts) {
    const float4* src = reinterpret_cast<const float4*>(Xs + (sb + rr) * D);
    float4* dst = reinterpret_cast<float4*>(lds_rows + rr * S);
    float nrm = 0.0f;
#pragma unroll
    for (int k = 0; k < D / 4; ++k) {
      const float4 x = src[k];
      dst[k] = x;
      nrm = __builtin_fmaf(x.x, x.x, nrm);
      nrm = __builtin_fmaf(x.y, x.y, nrm);
      nrm = __builtin_fmaf(x.z, x.z, nrm);
      nrm = __builtin_fmaf(x.w, x.w, nrm);
    }
    s_norm[rr] = nrm;
  }
  __syncthreads();
  const int64_t r = r0 + threadIdx.x;
  if (r >= n) return;  // after the block's only barrier
  const int64_t lo = seed_lo<W>(r, n);
  const float* self_row = lds_rows + (r - sb) * S;
  float4 own[D / 4];
#pragma unroll
  for (int k = 0; k < D / 4; ++k) own[k] = reinterpret_cast<const float4*>(self_row)[k];
  const float own_norm = s_norm[r - sb];
  uint32_t best[K];
#pragma unroll
  for (int j = 0; j < K; ++j) best[j] = 0xFFFFFFFFu;
  for (int slot = 0; slot < W; ++slot) {
    int64_t p = lo + slot;
    if (p >= r) p += 1;
    if (p >= n) break;  // only when n <= W: the remaining slots are past the end too
    const int pl = static_cast<int>(p - sb);
    const float4* pr = reinterpret_cast<const float4*>(lds_rows + pl * S);
    float dot = 0.0f;
#pragma unroll
    for (int k = 0; k < D / 4; ++k) {
      const float4 y = pr[k];
      dot = __builtin_fmaf(own[k].x, y.x, dot);
      dot = __builtin_fmaf(own[k].y, y.y, dot);
      dot = __builtin_fmaf(own[k].z, y.z, dot);
      dot = __builtin_fmaf(own[k].w, y.w, dot);
    }
    const float acc = fmaxf(__builtin_fmaf(-2.0f, dot, own_norm + s_norm[pl]), 0.0f);
    uint32_t key = (__float_as_uint(acc) & ~static_cast<uint32_t>(W - 1)) | static_cast<uint32_t>(slot);
    top_insert_med3<K>(best, key);
  }
  uint64_t* out = cand + static_cast<int64_t>(order[r]) * kSeedSlots + slot0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    uint64_t key = kKeyMax;
    if (best[j] != 0xFFFFFFFFu) {
      int64_t p = lo + static_cast<int64_t>(best[j] & static_cast<uint32_t>(W - 1));
      if (p >= r) p += 1;
      key = make_key(d2_any(lds_rows + (p - sb) * S, self_row, D), order[p], 1u);
    }
    out[j] = key;
  }
}

// The window search for d = 12 on the matrix cores.  A wave takes 16
// consecutive positions and the span of W + 32 rows that holds all their
// windows (u0 = min(seed_lo(first), n - span)); tile t of the span is 16 rows.
// Rows are staged with two extra features, so that one 16x16x4 FP32 MFMA
// chain over k = 16 gives every ranking distance of the tile directly:
//   A row w = (w_0..w_11, |w|^2, 1, 0, 0), B column p = (-2 p_0..-2 p_11, 1, |p|^2, 0, 0),
//   (A B)[w][p] = |w|^2 + |p|^2 - 2 w.p.
// Feature 4s + g of a row is stored at float 4g + s (row stride 20 floats),
// so lane group g (lanes 16g..16g+15) reads its four k-slices as one
// float4; 20 floats is an odd number of float4s, so 8 consecutive rows fall
// in distinct 4-bank groups.  Lane (g, c) receives rows 4g..4g+3 of each
// tile for position c and keeps the K/2 nearest of them in registers
// (median-of-three insertion); the four lanes of a position then merge
// their lists (xor 16, xor 32) into the K + 1 nearest of the union.  A
// lane's quarter holds more than K/2 of the span's K + 1 nearest rarely
// (then a slightly farther row takes the place): these are initial
// candidates, and every key written is exact.  The position itself
// (distance 0) is dropped when the exact keys are written; each of the four
// lanes writes every fourth winner.  The span rows beyond a window are
// further candidates.
__device__ __forceinline__ float f4_at(const float4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

constexpr int kSmWaves = 16;            // waves per block of k_mlp_seed_mfma
constexpr int kSmPts = 16 * kSmWaves;   // positions per block
constexpr int kSmStride = 20;           // staged row stride, floats

constexpr int seed_mfma_span(int W) { return 16 * (W / 16 + 2); }
constexpr size_t seed_mfma_lds(int W) {
  return static_cast<size_t>(kSmPts - 16 + seed_mfma_span(W)) * kSmStride * sizeof(float);
}

template <int W, int K>
__global__ __launch_bounds__(64 * kSmWaves) void k_mlp_seed_mfma(const float* __restrict__ X,
                                                                  const uint32_t* __restrict__ order, int64_t n,
                                                                  int slot0, uint64_t* __restrict__ cand) {
  constexpr int D = 12;
  constexpr int kTiles = W / 16 + 2;
  constexpr int kSpan = 16 * kTiles;
  constexpr int kRow4 = kSmStride / 4;  // float4s per staged row
  constexpr uint32_t kMask = 2u * W - 1u;
  constexpr int kLane = K / 2;  // nearest kept per lane (a quarter of the span)
  static_assert((W & (W - 1)) == 0 && W >= 64 && W <= 512, "window is a power of two");
  static_assert(kSpan <= 2 * W, "span slots fit the key's low bits");
  static_assert(4 * kLane >= K + 1 && 2 * kLane <= K + 1, "lane lists cover the K + 1 nearest");
  extern __shared__ __attribute__((aligned(16))) float4 lds_span4[];
  auto span_lo = [n](int64_t r) {
    const int64_t lo = seed_lo<W>(r, n);
    return lo < n - kSpan ? lo : n - kSpan;
  };
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kSmPts;
  const int64_t sb = span_lo(r0);
  const int64_t last_wave = r0 + ((n - r0 < kSmPts ? n - r0 : kSmPts) - 1) / 16 * 16;
  const int nrows = static_cast<int>(span_lo(last_wave) + kSpan - sb);
  for (int j = threadIdx.x; j < nrows; j += 64 * kSmWaves) {
    // row j of the span is point order[sb + j], gathered from X directly
    const float4* src = reinterpret_cast<const float4*>(X + static_cast<int64_t>(order[sb + j]) * D);
    const float4 x0 = src[0], x1 = src[1], x2 = src[2];
    float nrm = 0.0f;
    nrm = __builtin_fmaf(x0.x, x0.x, nrm);
    nrm = __builtin_fmaf(x0.y, x0.y, nrm);
    nrm = __builtin_fmaf(x0.z, x0.z, nrm);
    nrm = __builtin_fmaf(x0.w, x0.w, nrm);
    nrm = __builtin_fmaf(x1.x, x1.x, nrm);
    nrm = __builtin_fmaf(x1.y, x1.y, nrm);
    nrm = __builtin_fmaf(x1.z, x1.z, nrm);
    nrm = __builtin_fmaf(x1.w, x1.w, nrm);
    nrm = __builtin_fmaf(x2.x, x2.x, nrm);
    nrm = __builtin_fmaf(x2.y, x2.y, nrm);
    nrm = __builtin_fmaf(x2.z, x2.z, nrm);
    nrm = __builtin_fmaf(x2.w, x2.w, nrm);
    float4* dst = lds_span4 + j * kRow4;
    dst[0] = make_float4(x0.x, x1.x, x2.x, nrm);   // features 0, 4, 8, 12
    dst[1] = make_float4(x0.y, x1.y, x2.y, 1.0f);  // features 1, 5, 9, 13
    dst[2] = make_float4(x0.z, x1.z, x2.z, 0.0f);  // features 2, 6, 10, 14
    dst[3] = make_float4(x0.w, x1.w, x2.w, 0.0f);  // features 3, 7, 11, 15
  }
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int c = lane & 15;
  const int64_t rw = r0 + 16 * wv;
  if (rw >= n) return;  // wave-uniform, after the block's only barrier
  const int j0 = static_cast<int>(span_lo(rw) - sb);
  const int64_t r = rw + c;
  const bool valid = r < n;
  const int jr = valid ? static_cast<int>(r - sb) : j0;
  // B column: this lane's position, k-slices 4s + g scaled by -2, and
  // features 12..15 = (1, |p|^2, 0, 0)
  float4 bq = lds_span4[jr * kRow4 + g];
  const float nr = lds_span4[jr * kRow4].w;
  bq.x *= -2.0f;
  bq.y *= -2.0f;
  bq.z *= -2.0f;
  bq.w = g == 0 ? 1.0f : (g == 1 ? nr : 0.0f);
  if (!valid) bq = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t part[kLane];
#pragma unroll
  for (int j = 0; j < kLane; ++j) part[j] = 0xFFFFFFFFu;
  const float4* arow = lds_span4 + (j0 + c) * kRow4 + g;
#pragma unroll 2
  for (int t = 0; t < kTiles; ++t) {
    const float4 av = arow[t * 16 * kRow4];
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.x, bq.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.y, bq.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.z, bq.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av.w, bq.w, acc, 0, 0, 0);
    const uint32_t s0 = static_cast<uint32_t>(16 * t + 4 * g);
#pragma unroll
    for (int q = 0; q < 4; ++q) top_insert_med3<kLane>(part, (nonneg_bits(acc[q]) & ~kMask) | (s0 + q));
  }
  // lanes g and g ^ 1: the nearest 2 kLane of half the span, sorted
  uint32_t top[K + 1];
#pragma unroll
  for (int j = 0; j <= K; ++j) top[j] = j < kLane ? part[j] : 0xFFFFFFFFu;
#pragma unroll
  for (int j = 0; j < kLane; ++j)
    top_insert_med3<K + 1>(top, static_cast<uint32_t>(__shfl_xor(static_cast<int>(part[j]), 16)));
  // with lanes g ^ 2: over ascending lists A (g < 2) and B (g >= 2),
  // min(A[j], B[K - j]) is the K + 1 smallest of both (in no particular
  // order); all four lanes form the same array, so the every-fourth split
  // below writes each winner once
  const bool hi = (g & 2) != 0;
  uint32_t mrg[K + 1];
#pragma unroll
  for (int j = 0; j <= K; ++j) {
    const uint32_t mine = hi ? top[K - j] : top[j];
    const uint32_t o = static_cast<uint32_t>(__shfl_xor(static_cast<int>(mine), 32));
    mrg[j] = mine < o ? mine : o;
  }
#pragma unroll
  for (int j = 0; j <= K; ++j) top[j] = mrg[j];
  if (!valid) return;
  int selfpos = K + 1;
#pragma unroll
  for (int j = 0; j <= K; ++j)
    if (top[j] != 0xFFFFFFFFu && j0 + static_cast<int>(top[j] & kMask) == jr) selfpos = j;
  uint64_t* out = cand + static_cast<int64_t>(order[r]) * kSeedSlots + slot0;
  // feature k of a staged row is float k / 4 of its float4 k % 4
  float4 self4[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) self4[e] = lds_span4[jr * kRow4 + e];
#pragma unroll
  for (int jj = 0; jj < (K + 4) / 4; ++jj) {
    const int j = 4 * jj + g;
    if (j > K) break;
    uint32_t key = top[4 * jj];
#pragma unroll
    for (int e = 1; e < 4; ++e)
      if (4 * jj + e <= K && g == e) key = top[4 * jj + e];
    const int o = j - (selfpos < j ? 1 : 0);
    if (j == selfpos || o >= K) continue;
    uint64_t v = kKeyMax;
    if (key != 0xFFFFFFFFu) {
      const int js = j0 + static_cast<int>(key & kMask);
      float4 row4[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) row4[e] = lds_span4[js * kRow4 + e];
      // the oracle's operation order: diff = row - self, fmaf over k = 0..11
      float acc = 0.0f;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const float diff = f4_at(row4[k % 4], k / 4) - f4_at(self4[k % 4], k / 4);
        acc = __builtin_fmaf(diff, diff, acc);
      }
      v = make_key(acc, order[sb + js], 1u);
    }
    out[o] = v;
  }
}

// Per point: the 32 nearest distinct ids among its table candidates, sorted
// with exact keys.  Slots no candidate fills take xorshift ids, stepped +1
// past the point itself and past ids already in the list (n > 33 ids leave
// room), with their exact distances.
__global__ __launch_bounds__(256) void k_mlp_merge(const uint64_t* __restrict__ cand, const float* __restrict__ X,
                                                   int64_t n, int d, uint64_t seed,
                                                   hdb_nnd_element* __restrict__ lists) {
  __shared__ uint64_t s_key[4][64];
  const unsigned w = threadIdx.x >> 6;
  const int64_t p = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (p >= n) return;  // uniform across the wave; no block barriers below
  const unsigned lane = __lane_id();
  uint64_t key = cand[p * kSeedSlots + lane];
  key = bitonic_group<64>(key);
  const uint64_t prev = shfl_u64(key, (lane + 63) & 63);
  const bool keep = key != kKeyMax && (lane == 0 || (prev >> 1) != (key >> 1));
  const uint64_t bal = __ballot(keep);
  const int pos = __popcll(bal & lanemask_lt());
  const int total = __popcll(bal);
  s_key[w][lane] = kKeyMax;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (keep && pos < kK) s_key[w][pos] = key;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  key = s_key[w][lane];
  if (total < kK) {  // uniform
    const float* xp = X + p * d;
    for (int j = total; j < kK; ++j) {
      uint32_t id = static_cast<uint32_t>(xorshift64star(seed + static_cast<uint64_t>(p) * kK + j + 1) %
                                          static_cast<uint64_t>(n));
      for (;;) {  // at most 33 ids are excluded, so this ends
        const bool hit = static_cast<int>(lane) < j && key_id(key) == id;
        if (id != static_cast<uint32_t>(p) && __ballot(hit) == 0) break;
        id = static_cast<uint32_t>((id + 1) % static_cast<uint64_t>(n));
      }
      if (static_cast<int>(lane) == j) key = make_key(d2_any(xp, X + static_cast<int64_t>(id) * d, d), id, 1u);
    }
    key = bitonic_group<32>(key);
  }
  if (lane < static_cast<unsigned>(kK))
    lists[p * kK + lane] = hdb_nnd_element{key_d2(key), static_cast<int32_t>(key_id(key))};
}

}  // namespace hdb

// ------------------------------------------------------------- C ABI --
using namespace hdb;

struct hdb_mlp {
  MlpLayout L;
  float* prm = nullptr;   // device: parameters (FP32 master)
  float* grad = nullptr;  // device: gradient
  float* m = nullptr;
  float* v = nullptr;
  uint8_t* mask = nullptr;  // trainable entries (zero padding excluded)
  int64_t step = 0;
};

namespace {

// Work on the context's second stream inside one call: begin() makes
// stream2 wait for everything queued on `stream`; join() (also run at scope
// exit, so an early error return leaves no work outstanding on stream2)
// makes `stream` wait for stream2.
struct StreamFork {
  hdb_ctx* ctx;
  bool forked = false;
  explicit StreamFork(hdb_ctx* c) : ctx(c) {}
  int begin() {
    HDB_HIP(hipEventRecord(ctx->fork_ev, ctx->stream));
    HDB_HIP(hipStreamWaitEvent(ctx->stream2, ctx->fork_ev, 0));
    forked = true;
    return HDB_OK;
  }
  int join() {
    if (!forked) return HDB_OK;
    forked = false;
    HDB_HIP(hipEventRecord(ctx->join_ev, ctx->stream2));
    HDB_HIP(hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
    return HDB_OK;
  }
  ~StreamFork() { (void)join(); }
};

// While alive (and enabled): launches go to stream2 and the per-table slots
// (keys, order, sort and scan scratch, packed weights, permuted points) are
// the second set, so they never alias the buffers of work on `stream`.
struct AltSlots {
  hdb_ctx* ctx;
  hipStream_t saved;
  bool on;
  static void swap_all(hdb_ctx* c) {
    static const Slot a[] = {kSlotMisc1,    kSlotMisc2,   kSlotSortKeys, kSlotSortVals, kSlotSortHist,
                             kSlotSortAux, kSlotScanTmp, kSlotMlpPack,  kSlotSeedX};
    static const Slot b[] = {kSlotAlt0, kSlotAlt1, kSlotAlt2, kSlotAlt3, kSlotAlt4,
                             kSlotAlt5, kSlotAlt6, kSlotAlt7, kSlotAlt8};
    static_assert(sizeof(a) == sizeof(b), "slot pairs");
    for (size_t i = 0; i < sizeof(a) / sizeof(a[0]); ++i) scratch_swap(c, a[i], b[i]);
  }
  AltSlots(hdb_ctx* c, bool enable) : ctx(c), saved(c->stream), on(enable) {
    if (on) {
      ctx->stream = ctx->stream2;
      swap_all(ctx);
    }
  }
  ~AltSlots() {
    if (on) {
      swap_all(ctx);
      ctx->stream = saved;
    }
  }
};

int mlp_alloc(hdb_ctx* ctx, void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return ctx->fail(HDB_ENOMEM, "MLP allocation failed");
  }
  return HDB_OK;
}

// bytes of LDS for the forward kernel
size_t fwd_lds(const MlpLayout& L) {
  return static_cast<size_t>(kMlpWaves) * (16 * (L.kinp + 4) + 2 * 16 * (kHid + 4)) * sizeof(float);
}

int mlp_forward(hdb_ctx* ctx, hdb_mlp* net, const float* X, int d, const int32_t* rows, int64_t R, float* P1,
                float* H1, float* P2, float* H2, float* Z) {
  const unsigned grid = grid_for(R, 16 * kMlpWaves);
  if (P1)
    hipLaunchKernelGGL(k_mlp_fwd<true>, dim3(grid), dim3(kMlpWaves * 64), fwd_lds(net->L), ctx->stream, X, d, rows, R, net->prm,
                       net->L, P1, H1, P2, H2, Z);
  else
    hipLaunchKernelGGL(k_mlp_fwd<false>, dim3(grid), dim3(kMlpWaves * 64), fwd_lds(net->L), ctx->stream, X, d, rows, R, net->prm,
                       net->L, P1, H1, P2, H2, Z);
  HDB_LAUNCHED();
  return HDB_OK;
}

// Inference forward Z = f(X) over R rows (packed-weight kernel).
int mlp_infer(hdb_ctx* ctx, hdb_mlp* net, const float* X, int d, int64_t R, float* Z, const InferKeys* ko) {
  const MlpLayout& L = net->L;
  const int nfrag = infer_frags(L.kinp);
  __bf16* packed;
  HDB_CALL(scratch_t(ctx, kSlotMlpPack, static_cast<int64_t>(nfrag) * 512, &packed));
  hipLaunchKernelGGL(k_mlp_pack, dim3(grid_for(static_cast<int64_t>(nfrag) * 64, 256)), dim3(256), 0, ctx->stream,
                     net->prm, L, packed);
  HDB_LAUNCHED();
  const int waves = L.kinp <= 64 ? 16 : 8;
  const size_t lds = static_cast<size_t>(nfrag) * 1024 + 272 * sizeof(float) +
                     static_cast<size_t>(waves) * (16 * (L.kinp + 8) + 16 * (kHid + 8)) * 2;
  const int64_t tiles = (R + 15) / 16;
  const unsigned grid = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((tiles + waves - 1) / waves, 512)));
  if (ko)
    hipLaunchKernelGGL(k_mlp_infer<true>, dim3(grid), dim3(waves * 64), lds, ctx->stream, X, d, R, packed, net->prm, L,
                       Z, *ko);
  else
    hipLaunchKernelGGL(k_mlp_infer<false>, dim3(grid), dim3(waves * 64), lds, ctx->stream, X, d, R, packed, net->prm,
                       L, Z, InferKeys{nullptr, 1.0f, 12, nullptr, nullptr});
  HDB_LAUNCHED();
  return HDB_OK;
}

// Weight-gradient GEMM M x N over R rows into grad (M*N floats).
int wgrad(hdb_ctx* ctx, const float* A, int lda, const float* X, int d, const int32_t* rows, const float* B, int ldb,
          int64_t R, int M, int N, float* part, float* grad) {
  const int64_t chunk = 256;
  const int64_t splits = (R + chunk - 1) / chunk;
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(N / 16, M / 16, static_cast<unsigned>(splits)), dim3(64), 0, ctx->stream, A,
                     lda, X, d, rows, B, ldb, R, M, N, chunk, part);
  HDB_LAUNCHED();
  hipLaunchKernelGGL(k_mlp_reduce, dim3(grid_for(static_cast<int64_t>(M) * N, 256)), dim3(256), 0, ctx->stream, part,
                     splits, static_cast<int64_t>(M) * N, grad);
  HDB_LAUNCHED();
  return HDB_OK;
}

// Column sums of an R x n matrix (n <= 128) into out, in a fixed order:
// 256-row partial sums, then their sum.  part holds ceil(R/256) * n floats.
int colsum(hdb_ctx* ctx, const float* G, int64_t R, int n, float* part, float* out) {
  if (n < 1 || 256 % n != 0) return ctx->fail(HDB_EINVAL, "column count must divide 256");
  const int64_t splits = (R + 255) / 256;
  hipLaunchKernelGGL(k_mlp_colsum, dim3(static_cast<unsigned>(splits)), dim3(256), 0, ctx->stream, G, R, n, part);
  HDB_LAUNCHED();
  hipLaunchKernelGGL(k_mlp_reduce, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, part, splits,
                     static_cast<int64_t>(n), out);
  HDB_LAUNCHED();
  return HDB_OK;
}

// One loss + gradient evaluation for the triples in `rows` (3B rows).
int mlp_loss_grad(hdb_ctx* ctx, hdb_mlp* net, const float* X, int d, const int32_t* rows, int64_t B,
                  const uint32_t* codes, double* loss) {
  const MlpLayout& L = net->L;
  const int64_t R = 3 * B;
  float *P1, *H1, *P2, *H2, *Z, *dZ, *dP2, *dP1, *part, *mean, *lossp;
  HDB_CALL(scratch_t(ctx, kSlotMisc0, 4 * R * kHid, &P1));
  H1 = P1 + R * kHid;
  P2 = H1 + R * kHid;
  H2 = P2 + R * kHid;
  HDB_CALL(scratch_t(ctx, kSlotMisc1, 2 * R * kHid, &dP2));
  dP1 = dP2 + R * kHid;
  HDB_CALL(scratch_t(ctx, kSlotMisc2, 2 * R * kOutP + R + 64, &Z));
  dZ = Z + R * kOutP;
  lossp = dZ + R * kOutP;
  mean = lossp + R;
  const int64_t splits = (R + 255) / 256;
  HDB_CALL(scratch_t(ctx, kSlotBfTmp, splits * std::max(L.kinp, kHid) * kHid, &part));
  HDB_CALL(mlp_forward(ctx, net, X, d, rows, R, P1, H1, P2, H2, Z));
  if (codes) {
    hipLaunchKernelGGL(k_mlp_loss_bits, dim3(grid_for(R, 256)), dim3(256), 0, ctx->stream, Z, B, rows, codes, dZ,
                       lossp);
  } else {
    hipLaunchKernelGGL(k_mlp_tmean, dim3(static_cast<unsigned>(splits)), dim3(256), 0, ctx->stream, Z, R, part);
    HDB_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_tmean_final, dim3(1), dim3(64), 0, ctx->stream, part, splits, R, mean);
    HDB_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_loss, dim3(grid_for(B, 256)), dim3(256), 0, ctx->stream, Z, B, mean, dZ, lossp);
  }
  HDB_LAUNCHED();
  const size_t blds = static_cast<size_t>(4) * (16 * 36 + 16 * (kHid + 4)) * sizeof(float);
  hipLaunchKernelGGL(k_mlp_bwd_rows, dim3(grid_for(R, 64)), dim3(256), blds, ctx->stream, R, net->prm, L, P1, P2, dZ,
                     dP2, dP1);
  HDB_LAUNCHED();
  float* g = net->grad;
  HDB_CALL(wgrad(ctx, H2, kHid, nullptr, 0, nullptr, dZ, kOutP, R, kHid, kOutP, part, g + L.w3));
  HDB_CALL(wgrad(ctx, H1, kHid, nullptr, 0, nullptr, dP2, kHid, R, kHid, kHid, part, g + L.w2));
  HDB_CALL(wgrad(ctx, nullptr, 0, X, d, rows, dP1, kHid, R, L.kinp, kHid, part, g + L.w1));
  HDB_CALL(colsum(ctx, dZ, R, kOutP, part, g + L.b3));
  HDB_CALL(colsum(ctx, dP2, R, kHid, part, g + L.b2));
  HDB_CALL(colsum(ctx, dP1, R, kHid, part, g + L.b1));
  if (loss && codes) {
    std::vector<float> lp(R);
    HDB_HIP(hipMemcpyAsync(lp.data(), lossp, R * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HDB_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float x : lp) s += x;
    *loss = s;
  } else if (loss) {
    std::vector<float> lp(B);
    float mh[kCode];
    HDB_HIP(hipMemcpyAsync(lp.data(), lossp, B * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HDB_HIP(hipMemcpyAsync(mh, mean, sizeof(mh), hipMemcpyDeviceToHost, ctx->stream));
    HDB_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float x : lp) s += x;
    s /= static_cast<double>(B);
    for (float x : mh) s += static_cast<double>(x) * x;
    *loss = s;
  }
  return HDB_OK;
}

}  // namespace

extern "C" {

int hdb_mlp_create(hdb_ctx* ctx, int32_t d_in, uint64_t seed, hdb_mlp** out) {
  if (!ctx || !out || d_in < 1 || d_in > 128) return ctx ? ctx->fail(HDB_EINVAL, "MLP input width must be 1..128") : HDB_EINVAL;
  HDB_HIP(hipSetDevice(ctx->device));
  hdb_mlp* net = new hdb_mlp();
  net->L = mlp_layout(d_in);
  const MlpLayout& L = net->L;
  std::vector<float> h(L.total, 0.0f);
  std::vector<uint8_t> mask(L.total, 0);
  // He-style uniform init from xorshift64*, zero biases; padding stays zero
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&s]() {
    s = xorshift64star(s);
    return static_cast<float>((s >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
  };
  const float a1 = std::sqrt(6.0f / d_in), a2 = std::sqrt(6.0f / kHid);
  for (int k = 0; k < L.kinp; ++k)
    for (int j = 0; j < kHid; ++j)
      if (k < d_in) {
        h[L.w1 + k * kHid + j] = a1 * uni();
        mask[L.w1 + k * kHid + j] = 1;
      }
  for (int j = 0; j < kHid; ++j) mask[L.b1 + j] = mask[L.b2 + j] = 1;
  for (int k = 0; k < kHid * kHid; ++k) {
    h[L.w2 + k] = a2 * uni();
    mask[L.w2 + k] = 1;
  }
  for (int k = 0; k < kHid; ++k)
    for (int j = 0; j < kCode; ++j) {
      h[L.w3 + k * kOutP + j] = a2 * uni();
      mask[L.w3 + k * kOutP + j] = 1;
    }
  for (int j = 0; j < kCode; ++j) mask[L.b3 + j] = 1;
  int st = HDB_OK;
  if ((st = mlp_alloc(ctx, reinterpret_cast<void**>(&net->prm), L.total * 4)) != HDB_OK ||
      (st = mlp_alloc(ctx, reinterpret_cast<void**>(&net->grad), L.total * 4)) != HDB_OK ||
      (st = mlp_alloc(ctx, reinterpret_cast<void**>(&net->m), L.total * 4)) != HDB_OK ||
      (st = mlp_alloc(ctx, reinterpret_cast<void**>(&net->v), L.total * 4)) != HDB_OK ||
      (st = mlp_alloc(ctx, reinterpret_cast<void**>(&net->mask), L.total)) != HDB_OK) {
    hdb_mlp_destroy(net);
    return st;
  }
  HDB_HIP(hipMemcpy(net->prm, h.data(), L.total * 4, hipMemcpyHostToDevice));
  HDB_HIP(hipMemcpy(net->mask, mask.data(), L.total, hipMemcpyHostToDevice));
  HDB_HIP(hipMemset(net->grad, 0, L.total * 4));
  HDB_HIP(hipMemset(net->m, 0, L.total * 4));
  HDB_HIP(hipMemset(net->v, 0, L.total * 4));
  *out = net;
  return HDB_OK;
}

int hdb_mlp_destroy(hdb_mlp* net) {
  if (!net) return HDB_EINVAL;
  (void)hipFree(net->prm);
  (void)hipFree(net->grad);
  (void)hipFree(net->m);
  (void)hipFree(net->v);
  (void)hipFree(net->mask);
  delete net;
  return HDB_OK;
}

int64_t hdb_mlp_param_count(const hdb_mlp* net) { return net ? net->L.total : -1; }

int hdb_mlp_params(hdb_ctx* ctx, hdb_mlp* net, float* host_in, float* host_out) {
  if (!ctx || !net) return HDB_EINVAL;
  HDB_HIP(hipStreamSynchronize(ctx->stream));
  if (host_in) HDB_HIP(hipMemcpy(net->prm, host_in, net->L.total * 4, hipMemcpyHostToDevice));
  if (host_out) HDB_HIP(hipMemcpy(host_out, net->prm, net->L.total * 4, hipMemcpyDeviceToHost));
  return HDB_OK;
}

int hdb_mlp_forward(hdb_ctx* ctx, hdb_mlp* net, const float* X, int64_t m, int32_t d, float* Z) {
  if (!ctx || !net || !X || !Z || m < 0 || d != net->L.kin) return ctx ? ctx->fail(HDB_EINVAL, "bad forward args") : 1;
  HDB_HIP(hipSetDevice(ctx->device));
  if (m == 0) return HDB_OK;
  return mlp_infer(ctx, net, X, d, m, Z, nullptr);
}

int hdb_mlp_loss_grad(hdb_ctx* ctx, hdb_mlp* net, const float* X, int64_t n, int32_t d, const int32_t* rows,
                      const uint32_t* codes, int64_t batch, double* loss, float* grad_host) {
  if (!ctx || !net || !X || !rows || batch < 1 || d != net->L.kin || n < 1)
    return ctx ? ctx->fail(HDB_EINVAL, "bad loss_grad args") : 1;
  HDB_HIP(hipSetDevice(ctx->device));
  HDB_CALL(mlp_loss_grad(ctx, net, X, d, rows, batch, codes, loss));
  if (grad_host) {
    HDB_HIP(hipMemcpyAsync(grad_host, net->grad, net->L.total * 4, hipMemcpyDeviceToHost, ctx->stream));
    HDB_HIP(hipStreamSynchronize(ctx->stream));
  }
  return HDB_OK;
}

int hdb_mlp_train(hdb_ctx* ctx, hdb_mlp* net, const float* X, int64_t n, int32_t d, const hdb_nnd_element* lists,
                  const uint32_t* codes, int32_t steps, int32_t batch, float lr, uint64_t seed, double* final_loss) {
  if (!ctx || !net || !X || !lists || n < 2 || steps < 0 || batch < 1 || d != net->L.kin)
    return ctx ? ctx->fail(HDB_EINVAL, "bad train args") : 1;
  HDB_HIP(hipSetDevice(ctx->device));
  Stage st(ctx, "mlp_train");
  int32_t* rows;
  HDB_CALL(scratch_t(ctx, kSlotAhKeys2, 3 * static_cast<int64_t>(batch), &rows));
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_mlp_batch, dim3(grid_for(batch, 256)), dim3(256), 0, ctx->stream, lists, n,
                       static_cast<int64_t>(batch), seed + 0x100000001ull * static_cast<uint64_t>(net->step), rows);
    HDB_LAUNCHED();
    HDB_CALL(mlp_loss_grad(ctx, net, X, d, rows, batch, codes, (s == steps - 1) ? final_loss : nullptr));
    net->step += 1;
    const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(net->step));
    const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(net->step));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(net->L.total, 256)), dim3(256), 0, ctx->stream, net->prm, net->grad,
                       net->m, net->v, net->L.total, lr, c1, c2, net->mask);
    HDB_LAUNCHED();
  }
  return HDB_OK;
}

int hdb_rp_codes(hdb_ctx* ctx, const float* X, int64_t n, int32_t d, uint64_t seed, uint32_t* codes) {
  if (!ctx || !X || !codes || n < 2 || d < 1) return ctx ? ctx->fail(HDB_EINVAL, "bad rp_codes args") : 1;
  if (n > (1ll << 20)) return ctx->fail(HDB_EUNSUPPORTED, "rp_codes takes at most 2^20 points");
  HDB_HIP(hipSetDevice(ctx->device));
  Stage st(ctx, "rp_codes");
  uint64_t* keys;
  uint32_t* cell;
  int32_t* bounds;
  HDB_CALL(scratch_t(ctx, kSlotSeedCand, n, &keys));
  HDB_CALL(scratch_t(ctx, kSlotMisc1, n, &cell));
  HDB_CALL(scratch_t(ctx, kSlotMisc2, 2 << kCode, &bounds));
  HDB_CALL(fill_i32(ctx, reinterpret_cast<int32_t*>(codes), n, 0));
  const unsigned g = grid_for(n, 256);
  for (int l = 0; l < kCode; ++l) {
    hipLaunchKernelGGL(k_rp_keys, dim3(g), dim3(256), 0, ctx->stream, X, n, d, l, seed, codes, keys);
    HDB_LAUNCHED();
    HDB_CALL(radix_sort_u64(ctx, keys, n, 20, 52 + l + 1));
    hipLaunchKernelGGL(k_rp_cells, dim3(g), dim3(256), 0, ctx->stream, keys, n, cell);
    HDB_LAUNCHED();
    const int64_t ncell = 1ll << l;
    HDB_CALL(segment_bounds(ctx, cell, n, ncell, bounds, bounds + ncell));
    hipLaunchKernelGGL(k_rp_split, dim3(g), dim3(256), 0, ctx->stream, keys, n, bounds, bounds + ncell, codes);
    HDB_LAUNCHED();
  }
  return HDB_OK;
}

int hdb_learned_init(hdb_ctx* ctx, hdb_mlp* const* nets, int32_t tables, const float* X, int64_t n, int32_t d,
                     int32_t window, uint64_t seed, hdb_nnd_element* lists) {
  if (!ctx || !nets || tables < 1 || tables > kSeedTables || !X || !lists || n < 34)
    return ctx ? ctx->fail(HDB_EINVAL, "bad learned_init args (1..8 networks, n must exceed 33)") : 1;
  if (window != 128 && window != 256 && window != 512) return ctx->fail(HDB_EINVAL, "window must be 128, 256 or 512");
  for (int t = 0; t < tables; ++t)
    if (!nets[t] || nets[t]->L.kin != d) return ctx->fail(HDB_EINVAL, "network input width differs from d");
  if (n >= (1ll << 31)) return ctx->fail(HDB_EUNSUPPORTED, "more than 2^31 points");
  HDB_HIP(hipSetDevice(ctx->device));
  Stage st(ctx, "learned_init");
  const int keep = tables <= 4 ? 16 : 8;  // nearest window entries kept per table
  const int qbits = proj_bits(n);
  float* dirs;
  uint64_t* cand;
  HDB_CALL(scratch_t(ctx, kSlotBfTmp, static_cast<int64_t>(tables) * d, &dirs));
  HDB_CALL(scratch_t(ctx, kSlotSeedCand, n * kSeedSlots, &cand));
  // per table: a unit direction for the order inside a bucket, Box-Muller
  // from xorshift64*
  std::vector<float> hd(static_cast<size_t>(tables) * d);
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 7;
  for (int t = 0; t < tables; ++t) {
    double nrm = 0.0;
    for (int k = 0; k < d; ++k) {
      s = xorshift64star(s);
      const double u1 = ((s >> 11) + 1.0) / 9007199254740993.0;
      s = xorshift64star(s);
      const double u2 = (s >> 11) / 9007199254740992.0;
      hd[t * d + k] = static_cast<float>(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
      nrm += static_cast<double>(hd[t * d + k]) * hd[t * d + k];
    }
    for (int k = 0; k < d; ++k) hd[t * d + k] = static_cast<float>(hd[t * d + k] / std::sqrt(nrm));
  }
  HDB_HIP(hipMemcpyAsync(dirs, hd.data(), hd.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  if (tables * keep < kSeedSlots) HDB_CALL(fill_u64(ctx, cand, n * kSeedSlots, kKeyMax));
  const size_t lds = static_cast<size_t>(kSeedPts + window + 1) * (d + 1) * sizeof(float);
  const bool staged = lds <= 80 * 1024;
  const dim3 grid(static_cast<unsigned>(grid_for(n, kSeedPts)));
  const dim3 gridm(static_cast<unsigned>(grid_for(n, kSmPts)));
  const bool mfma_search = d == 12 && n >= window + 32 && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
  // Tables alternate between the context's two streams, each with its own
  // keys, sort scratch and packed weights, so one table's sort (a few
  // hundred blocks) overlaps the other's network and window search; the
  // window searches write disjoint candidate slots.
  StreamFork fork(ctx);
  HDB_CALL(fork.begin());
  for (int t = 0; t < tables; ++t) {
    AltSlots alt(ctx, (t & 1) != 0);
    uint32_t *keys, *order;
    float* Xs = nullptr;
    HDB_CALL(scratch_t(ctx, kSlotMisc1, n, &keys));
    HDB_CALL(scratch_t(ctx, kSlotMisc2, n, &order));
    // x.u range: 64 covers the spread of a unit direction over typical data
    const InferKeys ko{dirs + t * d, 64.0f, qbits, keys, order};
    HDB_CALL(mlp_infer(ctx, nets[t], X, d, n, nullptr, &ko));
    // the sorted order stays wherever the last pass left it (no copy back)
    uint32_t *skeys, *ids;
    HDB_CALL(radix_sort_u32_u32_out(ctx, keys, order, n, 0, kCode + qbits, &skeys, &ids));
    if (!mfma_search) {  // the other searches read the points in bucket order
      HDB_CALL(scratch_t(ctx, kSlotSeedX, n * d, &Xs));
      hipLaunchKernelGGL(k_mlp_permute, dim3(grid_for(n * d, 256)), dim3(256), 0, ctx->stream, X, ids, n, d, Xs);
      HDB_LAUNCHED();
    }
    const int slot0 = t * keep;
#define HDB_SEED(D_, W_, K_, ST_)                                                                                   \
  hipLaunchKernelGGL((k_mlp_seed<D_, W_, K_, ST_>), grid, dim3(kSeedPts), ST_ ? lds : 0, ctx->stream, Xs, ids, n, d, \
                     slot0, cand)
#define HDB_SEED_K(D_, W_, ST_)                 \
  do {                                          \
    if (keep == 16) HDB_SEED(D_, W_, 16, ST_);  \
    else HDB_SEED(D_, W_, 8, ST_);              \
  } while (0)
#define HDB_SEED_W(D_, ST_)                                \
  do {                                                     \
    if (window == 128) HDB_SEED_K(D_, 128, ST_);           \
    else if (window == 256) HDB_SEED_K(D_, 256, ST_);      \
    else HDB_SEED_K(D_, 512, ST_);                         \
  } while (0)
#define HDB_SEED4(D_, W_)                                                                                         \
  do {                                                                                                            \
    const size_t l4 = static_cast<size_t>(kSeedPts + W_ + 1) * ((((D_) / 4) & 1) ? (D_) : (D_) + 4) * 4 +       \
                      static_cast<size_t>(kSeedPts + W_ + 1) * 4;                                                 \
    if (keep == 16)                                                                                               \
      hipLaunchKernelGGL((k_mlp_seed4<D_, W_, 16>), grid, dim3(kSeedPts), l4, ctx->stream, Xs, ids, n, slot0, cand); \
    else                                                                                                          \
      hipLaunchKernelGGL((k_mlp_seed4<D_, W_, 8>), grid, dim3(kSeedPts), l4, ctx->stream, Xs, ids, n, slot0, cand);  \
  } while (0)
#define HDB_SEED4_W(D_)                 \
  do {                                  \
    if (window == 128) HDB_SEED4(D_, 128); \
    else if (window == 256) HDB_SEED4(D_, 256); \
    else HDB_SEED4(D_, 512);            \
  } while (0)
#define HDB_SEEDM(W_)                                                                                         \
  do {                                                                                                        \
    if (keep == 16)                                                                                           \
      hipLaunchKernelGGL((k_mlp_seed_mfma<W_, 16>), gridm, dim3(64 * kSmWaves), seed_mfma_lds(W_), ctx->stream, \
                         X, ids, n, slot0, cand);                                                             \
    else                                                                                                      \
      hipLaunchKernelGGL((k_mlp_seed_mfma<W_, 8>), gridm, dim3(64 * kSmWaves), seed_mfma_lds(W_), ctx->stream,  \
                         X, ids, n, slot0, cand);                                                             \
  } while (0)
    if (mfma_search) {
      if (window == 128) HDB_SEEDM(128);
      else if (window == 256) HDB_SEEDM(256);
      else HDB_SEEDM(512);
    } else if (d == 12) HDB_SEED4_W(12);
    else if (d == 32) HDB_SEED4_W(32);
    else if (staged) HDB_SEED_W(0, true);
    else HDB_SEED_W(0, false);
#undef HDB_SEEDM
#undef HDB_SEED4_W
#undef HDB_SEED4
#undef HDB_SEED_W
#undef HDB_SEED_K
#undef HDB_SEED
    HDB_LAUNCHED();
  }
  HDB_CALL(fork.join());
  hipLaunchKernelGGL(k_mlp_merge, dim3(grid_for(n, 4)), dim3(256), 0, ctx->stream, cand, X, n, d, seed, lists);
  HDB_LAUNCHED();
  return HDB_OK;
}

}  // extern "C"
