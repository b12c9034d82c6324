// Learned rank index over the globally sorted distinct k-mers (u64 keys,
// k <= 31): a two-level piecewise-linear FP32 model, trained on the GPU by
// gradient descent with the forward / backward kernels below, predicts a
// key's rank within a recorded per-leaf error bound, and compact lines of the
// sorted keys turn the prediction into the exact rank (a monotone minimal
// perfect map). No reference counterpart: it sits beside the BBHash
// (mphf.rs:268-276 is the lookup it can replace).
//
//  sort     : LSD radix sort of the keys, 8-bit digits (k_sort_hist / scan /
//             k_sort_scatter), stable tile ranking in registers.
//  warp     : u = an order-0 arithmetic-coding CDF of the key's first three
//             bytes (symbol counts of the keys themselves, histogrammed on
//             the GPU), non-decreasing in the key and close to uniform, so
//             the linear models below fit the keys' random spacing instead
//             of the genome's base composition (lidx.hpp gives the exact
//             fixed-point recurrence).
//  root     : R = 2^rb pieces by the top rb bits of u. Piece r (keys
//             [p_r, p_{r+1})) owns leaves [leaf0_r, leaf0_{r+1}), leaf0_r =
//             floor(p_r L / D) + r, and maps x = u - (r << s) to
//             leaf0_r + clamp(floor(fma(A_r, x, B_r)), 0, n_r - 1), trained
//             towards y = rank * L / D - floor(p_r L / D).
//  leaves   : L + R linear models of the rank over the keys the root routes
//             to them, trained in leaf coordinates (x = u - first u,
//             y = rank - first rank) and stored in the piece's:
//             pred = p_r + floor(fma(S, (float)x, I) + 0.5), 8 bytes a leaf,
//             so that the root and leaf tables (~2 MB at 2^18 leaves) stay in
//             L2.
//  training : per segment (piece or leaf), inputs standardised with the
//             segment's mean / deviation: z = (x - mu) * isd, f = w z + c,
//             MSE loss; forward writes residuals f - y, backward reduces the
//             gradient (2/n) sum (f - y) (z, 1), update steps w, c. The
//             starting point (and the untrained method) is endpoint
//             interpolation through the segment's first and last key.
//  bounds   : elo_j = max(pred - rank), ehi_j = max(rank - pred) over the
//             leaf's keys, evaluated with the exact lookup arithmetic; a leaf
//             whose window elo + ehi + 1 exceeds the threshold resolves
//             through the BBHash value mapped to its rank.
//  lines    : C consecutive keys by rank (and the next line's first key) in
//             one 64- or 128-byte line: the first key, a cell shift t, the
//             keys' cells (key - k0) >> t in unary, and for keys sharing a
//             cell with their predecessor the highest bit where the two
//             differ (3-bit code below t). ~3.2 bits a key.
//  lookup   : warp (tables in LDS) -> root (L2) -> leaf (L2) -> predicted
//             rank -> line pred / C (one
//             far request) -> the key's cell in the unary, its run, and a
//             blind descent over the run's branch bits; the walk steps to a
//             neighbour line when the key lies before the line's first key
//             or after its last cell, and searches the sorted keys for the
//             rare runs the codes cannot resolve and for lines that overflow.
#include "block.hpp"
#include "classify.hpp"
#include "device.hpp"
#include "launch.hpp"
#include "lidx.hpp"

namespace cf1g {

// ------------------------------------------------------------------ radix sort
constexpr int SRT_WG = 256;
constexpr int SRT_KPT = 16;
constexpr int SRT_TILE = SRT_WG * SRT_KPT;

__global__ void __launch_bounds__(SRT_WG) k_sort_hist(const uint64_t* __restrict__ src, uint64_t n, int sh, uint64_t n_tiles,
                                                      uint64_t* __restrict__ mat) {
  __shared__ uint32_t h[256];
  const uint64_t tile = blockIdx.x;
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t e0 = tile * SRT_TILE;
#pragma unroll
  for (int q = 0; q < SRT_KPT; q++) {
    const uint64_t e = e0 + (uint64_t)q * SRT_WG + threadIdx.x;
    if (e < n) atomicAdd(&h[(src[e] >> sh) & 255], 1u);
  }
  __syncthreads();
  mat[(uint64_t)threadIdx.x * n_tiles + tile] = h[threadIdx.x];
}

// Stable scatter of one tile by the digit at `sh`: element e of the tile
// (wave-blocked: wave*64*KPT + slot*64 + lane, i.e. memory order) goes to
// off[digit * n_tiles + tile] + its rank among the tile's keys of that digit.
__global__ void __launch_bounds__(SRT_WG) k_sort_scatter(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, uint64_t n,
                                                         int sh, uint64_t n_tiles, const uint64_t* __restrict__ off) {
  __shared__ uint32_t cnt[4 * 256];
  __shared__ uint64_t sh4[4];
  const uint64_t tile = blockIdx.x;
  const uint32_t wv = threadIdx.x >> 6, lane = __lane_id();
  uint64_t key[SRT_KPT];
  uint32_t rank[SRT_KPT];
  const uint64_t e0 = tile * SRT_TILE + (uint64_t)wv * 64 * SRT_KPT + lane;
#pragma unroll
  for (int s = 0; s < SRT_KPT; s++) {
    const uint64_t e = e0 + (uint64_t)s * 64;
    key[s] = e < n ? src[e] : ~0ULL;  // pads rank last within digit 255 and are not written
  }
  for (int i = threadIdx.x; i < 4 * 256; i += SRT_WG) cnt[i] = 0;
  __syncthreads();
  rank_tile<uint64_t, SRT_KPT>(key, SRT_KPT, sh, rank, cnt);
  __syncthreads();
  rank_offsets(cnt, sh4, nullptr);  // cnt[w*256+d] = tile-local start of (w, d); cnt[d] = start of d
  __syncthreads();
#pragma unroll
  for (int s = 0; s < SRT_KPT; s++) {
    const uint64_t e = e0 + (uint64_t)s * 64;
    if (e < n) {
      const uint32_t d = (uint32_t)(key[s] >> sh) & 255u;
      dst[off[(uint64_t)d * n_tiles + tile] + (cnt[wv * 256 + d] - cnt[d]) + rank[s]] = key[s];
    }
  }
}

// ------------------------------------------------------------------ warp
// floor(w c / 2^32) for c < 2^32.
__device__ __forceinline__ uint64_t warp_f(uint64_t w, uint32_t c) {
  return (w >> 32) * (uint64_t)c + (uint64_t)__umulhi((uint32_t)w, c);
}
__device__ __forceinline__ uint64_t warp_key(uint64_t key, const uint32_t* __restrict__ cum, int K2, int G) {
  uint64_t lo = 0, w = ~0ULL;
#pragma unroll
  for (int g = 0; g < WARP_MAX_GROUPS; g++) {
    if (g < G) {  // uniform
      const uint32_t sym = (uint32_t)(key >> (K2 - 8 * (g + 1))) & 255u;
      const uint32_t c0 = cum[g * 256 + sym], c1 = cum[g * 256 + (sym < 255u ? sym + 1 : 255u)];
      const uint64_t a = warp_f(w, c0);
      const uint64_t b = sym == 255u ? w : warp_f(w, c1);
      lo += a;
      w = b - a;
    }
  }
  const int rb = K2 - 8 * G;
  if (rb > 0) lo += __umul64hi(w, (key & ((1ULL << rb) - 1)) << (64 - rb));
  return lo;
}

// Byte counts per warp group over the sorted keys: cnt[g * 256 + s].
__global__ void __launch_bounds__(256) k_lidx_warp_hist(const uint64_t* __restrict__ keys, uint64_t D, int K2, int G,
                                                        uint32_t* __restrict__ cnt) {
  __shared__ uint32_t h[WARP_MAX_GROUPS * 256];
  for (int i = threadIdx.x; i < WARP_MAX_GROUPS * 256; i += 256) h[i] = 0;
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < D; i += stride) {
    const uint64_t k = keys[i];
    for (int g = 0; g < G; g++) atomicAdd(&h[g * 256 + ((uint32_t)(k >> (K2 - 8 * (g + 1))) & 255u)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < G * 256; i += 256)
    if (h[i]) atomicAdd(&cnt[i], h[i]);
}
__global__ void __launch_bounds__(256) k_lidx_warp_keys(const uint64_t* __restrict__ keys, uint64_t D, const uint32_t* __restrict__ cum,
                                                        int K2, int G, uint64_t* __restrict__ u) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < D) u[i] = warp_key(keys[i], cum, K2, G);
}

// Leaf of a key in piece r (P = root[r], leaf0n = root[r + 1].leaf0).
__device__ __forceinline__ uint32_t root_leaf(const RootPiece& P, uint32_t leaf0n, float x) {
  const float f = floorf(__fmaf_rn(P.A, x, P.B));
  const float span = (float)(leaf0n - P.leaf0 - 1);
  const float c = f < 0.f ? 0.f : (f > span ? span : f);
  return P.leaf0 + (uint32_t)c;
}
// x = u - (r << s) in FP32 from its top 32 bits (32 < s): (float)(u32)(u >>
// (s - 32)) * 2^(s - 32). The bits dropped lie far below the keys' spacing.
__device__ __forceinline__ float piece_x(uint64_t u, int s) {
  return ldexpf((float)(uint32_t)(u >> (s - 32)), s - 32);
}
__device__ __forceinline__ bool leaf_fallback(const LeafRec& Lf) { return (__float_as_uint(Lf.S) >> 31) != 0; }
// The leaf's rank relative to its piece: floor(fma(|S|, x, I) + 0.5), clamped
// to +-2^30.
__device__ __forceinline__ int32_t leaf_local(const LeafRec& Lf, float x) {
  const float S = __uint_as_float(__float_as_uint(Lf.S) & 0x7FFFFFFFu);
  float p = floorf(__fmaf_rn(S, x, Lf.I) + 0.5f);
  p = p < -1073741824.0f ? -1073741824.0f : (p > 1073741824.0f ? 1073741824.0f : p);
  return (int32_t)p;
}
__device__ __forceinline__ int64_t leaf_pred(const RootPiece& P, const LeafRec& Lf, float x) {
  return (int64_t)P.p + (int64_t)leaf_local(Lf, x);
}

// ------------------------------------------------------------------ training
template <int LEVEL>
__device__ __forceinline__ float seg_x(const SegCtx& c, uint64_t sg, uint64_t i) {
  if (LEVEL == 0) return (float)(c.keys[i] - ((uint64_t)sg << c.s));
  return (float)(c.keys[i] - c.keys[c.seg[sg]]);
}
template <int LEVEL>
__device__ __forceinline__ float seg_y(const SegCtx& c, uint64_t sg, uint64_t i) {
  if (LEVEL == 0) return (float)((double)i * (double)c.L / (double)c.D - (double)c.lo[sg]);
  return (float)(i - c.seg[sg]);
}

// Team of TW threads per segment: TW = 64 (one wave) or 256 (a work-group).
template <int TW>
__device__ __forceinline__ double team_sum(double v, double* sh) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (TW == 64) return v;
  const uint32_t wv = threadIdx.x >> 6;
  __syncthreads();
  if (__lane_id() == 0) sh[wv] = v;
  __syncthreads();
  double t = 0;
  for (int q = 0; q < TW / 64; q++) t += sh[q];
  return t;
}

// Statistics: mu = mean x, isd = 1 / deviation (0 when the deviation is 0),
// in double, then the FP32 values the model uses.
template <int LEVEL, int TW>
__global__ void __launch_bounds__(256) k_lidx_stats(SegCtx c, uint64_t n_segs, float* __restrict__ mu, float* __restrict__ isd) {
  __shared__ double sh[4];
  const uint64_t teams = blockDim.x / TW;
  const uint64_t sg = (uint64_t)blockIdx.x * teams + threadIdx.x / TW;
  const uint32_t t = threadIdx.x % TW;
  if (TW == 64 && sg >= n_segs) return;
  double s1 = 0, s2 = 0;
  uint64_t a = 0, b = 0;
  if (sg < n_segs) {
    a = c.seg[sg];
    b = c.seg[sg + 1];
    for (uint64_t i = a + t; i < b; i += TW) {
      const double x = (double)seg_x<LEVEL>(c, sg, i);
      s1 += x;
      s2 += x * x;
    }
  }
  s1 = team_sum<TW>(s1, sh);
  s2 = team_sum<TW>(s2, sh);
  if (sg < n_segs && t == 0) {
    const double n = (double)(b - a);
    const double m = n > 0 ? s1 / n : 0.0;
    const double v = n > 0 ? s2 / n - m * m : 0.0;
    mu[sg] = (float)m;
    isd[sg] = v > 0 ? (float)(1.0 / sqrt(v)) : 0.f;
  }
}

__device__ __forceinline__ float z_of(float x, float mu, float isd) { return (x - mu) * isd; }

// Endpoint interpolation through the segment's first and last key, in the
// standardised coordinates: w = A0 / isd, c = A0 mu + B0 (A0, B0 the line in
// x). Single-key segments: w = 0, c = y.
template <int LEVEL>
__global__ void __launch_bounds__(256) k_lidx_init(SegCtx c, uint64_t n_segs, const float* __restrict__ mu,
                                                   const float* __restrict__ isd, float* __restrict__ w, float* __restrict__ cc) {
  const uint64_t sg = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= n_segs) return;
  const uint64_t a = c.seg[sg], b = c.seg[sg + 1];
  float W = 0.f, C = 0.f;
  if (b > a) {
    const double x0 = seg_x<LEVEL>(c, sg, a), y0 = seg_y<LEVEL>(c, sg, a);
    const double x1 = seg_x<LEVEL>(c, sg, b - 1), y1 = seg_y<LEVEL>(c, sg, b - 1);
    double A0 = 0, B0 = y0;
    if (b - a >= 2 && x1 > x0) {
      A0 = (y1 - y0) / (x1 - x0);
      B0 = y0 - A0 * x0;
    }
    if (isd[sg] > 0.f) {
      W = (float)(A0 / (double)isd[sg]);
      C = (float)(A0 * (double)mu[sg] + B0);
    } else {
      W = 0.f;
      C = (float)(A0 * (double)mu[sg] + B0);
    }
  }
  w[sg] = W;
  cc[sg] = C;
}

// Forward: residual r_i = f(z_i) - y_i with f = fma(w, z, c), FP32.
template <int LEVEL>
__global__ void __launch_bounds__(256) k_lidx_forward(SegCtx c, uint64_t n_segs, const float* __restrict__ mu,
                                                      const float* __restrict__ isd, const float* __restrict__ w,
                                                      const float* __restrict__ cc, const uint32_t* __restrict__ seg_of_tile,
                                                      float* __restrict__ resid) {
  // tiles of 256 keys; seg_of_tile[t] = segment holding key 256*t
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= c.D) return;
  uint64_t sg = seg_of_tile[blockIdx.x];
  while (sg + 1 < n_segs && c.seg[sg + 1] <= i) sg++;
  const float z = z_of(seg_x<LEVEL>(c, sg, i), mu[sg], isd[sg]);
  resid[i] = __fmaf_rn(w[sg], z, cc[sg]) - seg_y<LEVEL>(c, sg, i);
}

// Backward: per segment g_w = (2/n) sum r z, g_c = (2/n) sum r, loss = (1/n)
// sum r^2 (accumulated in double from the FP32 residuals).
template <int LEVEL, int TW>
__global__ void __launch_bounds__(256) k_lidx_backward(SegCtx c, uint64_t n_segs, const float* __restrict__ mu,
                                                       const float* __restrict__ isd, const float* __restrict__ resid,
                                                       double* __restrict__ gw, double* __restrict__ gc, double* __restrict__ loss) {
  __shared__ double sh[4];
  const uint64_t teams = blockDim.x / TW;
  const uint64_t sg = (uint64_t)blockIdx.x * teams + threadIdx.x / TW;
  const uint32_t t = threadIdx.x % TW;
  if (TW == 64 && sg >= n_segs) return;
  double a1 = 0, a2 = 0, a3 = 0;
  uint64_t a = 0, b = 0;
  if (sg < n_segs) {
    a = c.seg[sg];
    b = c.seg[sg + 1];
    const float m = mu[sg], is = isd[sg];
    for (uint64_t i = a + t; i < b; i += TW) {
      const float r = resid[i];
      const float z = z_of(seg_x<LEVEL>(c, sg, i), m, is);
      a1 += (double)r * (double)z;
      a2 += (double)r;
      a3 += (double)r * (double)r;
    }
  }
  a1 = team_sum<TW>(a1, sh);
  a2 = team_sum<TW>(a2, sh);
  a3 = team_sum<TW>(a3, sh);
  if (sg < n_segs && t == 0) {
    const double n = (double)(b - a);
    gw[sg] = n > 0 ? 2.0 * a1 / n : 0.0;
    gc[sg] = n > 0 ? 2.0 * a2 / n : 0.0;
    loss[sg] = a3;  // sum of squares; the mean is taken by the caller
  }
}

__global__ void __launch_bounds__(256) k_lidx_update(uint64_t n_segs, float lr, const double* __restrict__ gw,
                                                     const double* __restrict__ gc, float* __restrict__ w, float* __restrict__ cc) {
  const uint64_t sg = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= n_segs) return;
  w[sg] = (float)((double)w[sg] - (double)lr * gw[sg]);
  cc[sg] = (float)((double)cc[sg] - (double)lr * gc[sg]);
}

// seg_of_tile[t] = the segment holding key 256*t (largest sg with seg[sg] <= 256 t).
__global__ void __launch_bounds__(256) k_lidx_tile_seg(const uint64_t* __restrict__ seg, uint64_t n_segs, uint64_t D,
                                                       uint32_t* __restrict__ seg_of_tile) {
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t n_tiles = (D + 255) / 256;
  if (t >= n_tiles) return;
  const uint64_t i = t * 256;
  uint64_t lo = 0, hi = n_segs;  // seg[lo] <= i < seg[hi]
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (seg[mid] <= i) lo = mid; else hi = mid;
  }
  seg_of_tile[t] = (uint32_t)lo;
}

// p[r] = first rank with key >= r << s, r in [0, R] (p[R] = D).
__global__ void __launch_bounds__(256) k_lidx_pieces(const uint64_t* __restrict__ keys, uint64_t D, int s, uint32_t R,
                                                     uint64_t* __restrict__ p) {
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r > R) return;
  if (r == R) {
    p[r] = D;
    return;
  }
  const uint64_t target = r << s;
  uint64_t lo = 0, hi = D;  // first index with keys[i] >= target
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1; else hi = mid;
  }
  p[r] = lo;
}
// lo[r] = floor(p_r L / D), r in [0, R] (the root's training offset).
__global__ void __launch_bounds__(256) k_lidx_root_ranges(const uint64_t* __restrict__ p, uint32_t R, uint64_t D, uint32_t L,
                                                          uint32_t* __restrict__ lo) {
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r > R) return;
  lo[r] = (uint32_t)(D ? (p[r] * (uint64_t)L / D) : 0);
}

// Root records in lookup form: A = w isd, B = c - w mu isd (A >= 0), the
// piece's first leaf lo_r + r and first rank; record R closes the table.
__global__ void __launch_bounds__(256) k_lidx_root_final(uint32_t R, uint64_t D, const uint64_t* __restrict__ p,
                                                         const uint32_t* __restrict__ lo, const float* __restrict__ mu,
                                                         const float* __restrict__ isd, const float* __restrict__ w,
                                                         const float* __restrict__ cc, RootPiece* __restrict__ out) {
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r > R) return;
  RootPiece P;
  P.leaf0 = lo[r] + (uint32_t)r;
  P.p = (uint32_t)p[r];
  if (r == R) {
    P.A = P.B = 0.f;
  } else {
    double A = (double)w[r] * (double)isd[r];
    double B = (double)cc[r] - A * (double)mu[r];
    if (!(A >= 0)) A = 0;  // monotone routing (also catches NaN)
    if (!(B == B)) B = 0;
    P.A = (float)A;
    P.B = (float)B;
  }
  out[r] = P;
}

// Leaf boundaries from the root routing of every key (non-decreasing in the
// rank): q[j] = first rank routed to a leaf >= j, q[LT] = D.
__global__ void __launch_bounds__(256) k_lidx_leaf_bounds(const uint64_t* __restrict__ keys, uint64_t D, int s, uint32_t LT,
                                                          const RootPiece* __restrict__ root, uint64_t* __restrict__ q) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i > D) return;
  int64_t prev = -1, cur = (int64_t)LT;
  if (i > 0) {
    const uint64_t k = keys[i - 1];
    const uint32_t r = (uint32_t)(k >> s);
    prev = root_leaf(root[r], root[r + 1].leaf0, piece_x(k, s));
  }
  if (i < D) {
    const uint64_t k = keys[i];
    const uint32_t r = (uint32_t)(k >> s);
    cur = root_leaf(root[r], root[r + 1].leaf0, piece_x(k, s));
  }
  for (int64_t j = prev + 1; j <= cur; j++) q[j] = i;
}

// Leaf records in piece coordinates: S = w isd and, with I = c - S mu the
// intercept in leaf coordinates, I' = I - S (first key - (r << s)) + (q_j -
// p_r); error bounds over the leaf's keys with the lookup arithmetic (one
// wave per leaf); the sign bit of S marks a window above max_window.
__global__ void __launch_bounds__(256) k_lidx_leaf_final(const uint64_t* __restrict__ keys, uint64_t D, uint32_t LT, int s,
                                                         const uint64_t* __restrict__ q, const RootPiece* __restrict__ root,
                                                         const float* __restrict__ mu, const float* __restrict__ isd,
                                                         const float* __restrict__ w, const float* __restrict__ cc,
                                                         uint32_t max_window, LeafRec* __restrict__ out,
                                                         uint32_t* __restrict__ err, uint64_t* __restrict__ sabs) {
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = __lane_id();
  if (j >= LT) return;
  const uint64_t a = q[j], b = q[j + 1];
  LeafRec Lf{0.f, 0.f};
  int64_t elo = 0, ehi = 0;
  uint64_t sa = 0;
  if (b > a) {
    const uint64_t k0 = keys[a];
    const uint32_t r = (uint32_t)(k0 >> s);
    const RootPiece P = root[r];
    double S = (double)w[j] * (double)isd[j];
    double I = (double)cc[j] - S * (double)mu[j];
    I = I - S * (double)(k0 - ((uint64_t)r << s)) + (double)((int64_t)a - (int64_t)P.p);
    if (!(S == S) || !(I == I) || S < 0) S = I = 0;
    Lf.S = (float)S;
    Lf.I = (float)I;
    for (uint64_t i = a + lane; i < b; i += 64) {
      const int64_t e = (int64_t)i - leaf_pred(P, Lf, piece_x(keys[i], s));
      elo = -e > elo ? -e : elo;
      ehi = e > ehi ? e : ehi;
      sa += (uint64_t)(e < 0 ? -e : e);
    }
    for (int d = 32; d >= 1; d >>= 1) {
      const int64_t o1 = __shfl_xor(elo, d, 64), o2 = __shfl_xor(ehi, d, 64);
      elo = o1 > elo ? o1 : elo;
      ehi = o2 > ehi ? o2 : ehi;
      sa += __shfl_xor(sa, d, 64);
    }
    if ((uint64_t)(elo + ehi + 1) > max_window) Lf.S = __uint_as_float(__float_as_uint(Lf.S) | 0x80000000u);
  }
  if (lane == 0) {
    out[j] = Lf;
    err[2 * j] = (uint32_t)(elo > 0xFFFFFFFFLL ? 0xFFFFFFFFLL : elo);
    err[2 * j + 1] = (uint32_t)(ehi > 0xFFFFFFFFLL ? 0xFFFFFFFFLL : ehi);
    sabs[j] = sa;
  }
}

// Fallback leaves: bbh2rank[BBHash value of key i] = i, one wave per leaf.
__global__ void __launch_bounds__(256) k_lidx_bbh2rank(const uint64_t* __restrict__ keys, uint32_t LT, const uint64_t* __restrict__ q,
                                                       const LeafRec* __restrict__ leaf, const MphfView mv,
                                                       uint32_t* __restrict__ bbh2rank) {
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= LT || !leaf_fallback(leaf[j])) return;
  const uint64_t D = q[LT];
  for (uint64_t i = q[j] + __lane_id(); i < q[j + 1]; i += 64) {
    const uint64_t v = mphf_lookup<uint64_t>(mv, keys[i]);
    if (v < D) bbh2rank[v] = (uint32_t)i;
  }
}

// ------------------------------------------------------------------ lines
__device__ __forceinline__ int bitlen64(uint64_t x) { return x ? 64 - __clzll(x) : 0; }

// One thread per line, built in its own LDS slot (padded against bank
// conflicts) and stored as line_words u64 words. stats[2i] = bits used
// (0xFFFFFFFF for an overflow line), stats[2i + 1] = keys of this line's runs
// that resolve by search (a code of 7 or more than 21 codes).
// Top LINE_SP_BITS bits of rel's offset in its cell of 2^t (t <= 61), read as
// a t-bit number (exact, shifted up, when t < LINE_SP_BITS): of two values in
// one cell, the one with the larger prefix is the larger.
__device__ __forceinline__ uint32_t sentinel_prefix(uint64_t rel, uint32_t t) {
  const uint64_t off = rel & ((1ULL << t) - 1);
  return t >= (uint32_t)LINE_SP_BITS ? (uint32_t)(off >> (t - LINE_SP_BITS)) : (uint32_t)(off << (LINE_SP_BITS - t));
}

template <int NW>
__global__ void __launch_bounds__(256) k_lidx_lines(const uint64_t* __restrict__ keys, uint64_t D, uint32_t C, uint64_t n_lines,
                                                    uint64_t* __restrict__ lines, uint32_t* __restrict__ stats) {
  constexpr int NU = 2 * NW, STRIDE = NU + 1;
  __shared__ uint32_t sb[256 * STRIDE];
  uint32_t* b = sb + threadIdx.x * STRIDE;
#pragma unroll
  for (int u = 0; u < NU; u++) b[u] = 0;
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_lines) return;
  auto put = [&](uint32_t pos, uint32_t val, int width) {  // width <= 32
    const uint32_t w = pos >> 5, o = pos & 31;
    b[w] |= val << o;
    if (o + width > 32) b[w + 1] |= val >> (32 - o);
  };
  const uint64_t a = i * C;
  const uint32_t n = (uint32_t)((D - a) < C ? (D - a) : C);
  const uint32_t m = n + ((a + n < D) ? 1u : 0u);
  const uint64_t k0 = keys[a];
  const uint64_t span = keys[a + m - 1] - k0;
  const int tb = max(0, bitlen64(span) - bitlen64(m));
  const int t_lo = max(0, tb - 2), t_hi = min(61, tb + 3);
  uint32_t cnt[6] = {0, 0, 0, 0, 0, 0};
  int min_bl = 64;  // the closest pair: its code is the line's largest
  uint64_t prev = 0;
  for (uint32_t j = 1; j < m; j++) {
    const uint64_t rel = keys[a + j] - k0;
    const int bl = bitlen64(rel ^ prev);
#pragma unroll
    for (int u = 0; u < 6; u++) cnt[u] += (bl <= t_lo + u) ? 1u : 0u;
    min_bl = min(min_bl, bl);
    prev = rel;
  }
  // the first minimum of the bit count over the candidate shifts, for
  // 3-bit and for 4-bit codes
  int t3 = t_lo, t4 = t_lo;
  uint64_t best3 = ~0ULL, best4 = ~0ULL;
#pragma unroll
  for (int u = 0; u < 6; u++) {
    if (t_lo + u > t_hi) break;
    const uint64_t cells = (uint64_t)m + (span >> (t_lo + u)) + 1;
    if (cells + 3ULL * cnt[u] < best3) {
      best3 = cells + 3ULL * cnt[u];
      t3 = t_lo + u;
    }
    if (cells + 4ULL * cnt[u] < best4) {
      best4 = cells + 4ULL * cnt[u];
      t4 = t_lo + u;
    }
  }
  // 4-bit codes where they fit and 3-bit codes would leave a code to search
  // (or would not fit)
  const uint64_t payload = (uint64_t)(NW * 64 - LINE_HDR_BITS);
  const bool wide = best4 <= payload && (best3 > payload || (min_bl <= t3 && t3 - min_bl >= 7));
  const int cw = wide ? 4 : 3, t = wide ? t4 : t3;
  const uint64_t best = wide ? best4 : best3;
  const int csearch = (1 << cw) - 1;
  const uint32_t max_run_codes = 64u / (uint32_t)cw;
  b[0] = (uint32_t)k0;
  b[1] = (uint32_t)(k0 >> 32) & 0x3FFFFFFFu;
  uint32_t n_search = 0;
  if (best > payload) {
    b[1] |= 1u << 30;  // bit 62: overflow
    stats[2 * i] = 0xFFFFFFFFu;
    n_search = n;
  } else {
    if (wide) b[1] |= 1u << 31;  // bit 63: 4-bit codes
    const uint32_t cells = (uint32_t)(span >> t) + 1;
    const uint32_t ulen = m + cells;
    b[2] |= (uint32_t)t | (ulen << 6);
    if (m > n) b[2] |= sentinel_prefix(span, t) << 16;  // bits 80..87
    uint32_t pos = LINE_HDR_BITS, cpos = LINE_HDR_BITS + ulen, run = 1;
    uint64_t prel = 0, pcell = 0;
    bool bad = false;
    for (uint32_t j = 0; j < m; j++) {
      const uint64_t rel = keys[a + j] - k0;
      const uint64_t c = rel >> t;
      pos += (uint32_t)(c - pcell);
      b[pos >> 5] |= 1u << (pos & 31);
      pos++;
      if (j > 0 && c == pcell) {
        const int code = min(t - bitlen64(rel ^ prel), csearch);
        put(cpos, (uint32_t)code, cw);
        cpos += (uint32_t)cw;
        run++;
        bad |= code == csearch;
      } else if (j > 0) {
        if (run > 1 && (bad || run - 1 > max_run_codes)) n_search += run;
        run = 1;
        bad = false;
      }
      pcell = c;
      prel = rel;
    }
    if (run > 1 && (bad || run - 1 > max_run_codes)) n_search += run;
    stats[2 * i] = (uint32_t)best;
  }
  stats[2 * i + 1] = n_search;
  uint64_t* dst = lines + i * NW;
#pragma unroll
  for (int u = 0; u < NW; u++) dst[u] = (uint64_t)b[2 * u] | ((uint64_t)b[2 * u + 1] << 32);
}

// ------------------------------------------------------------------ lookup
// Position of the r-th (0-based) set bit of x (x has more than r).
__device__ __forceinline__ uint32_t select64(uint64_t x, uint32_t r) {
  uint32_t pos = 0, c;
  c = __popc((uint32_t)x);
  if (r >= c) { r -= c; x >>= 32; pos += 32; }
  c = __popc((uint32_t)x & 0xFFFFu);
  if (r >= c) { r -= c; x >>= 16; pos += 16; }
  c = __popc((uint32_t)x & 0xFFu);
  if (r >= c) { r -= c; x >>= 8; pos += 8; }
  c = __popc((uint32_t)x & 0xFu);
  if (r >= c) { r -= c; x >>= 4; pos += 4; }
  c = __popc((uint32_t)x & 0x3u);
  if (r >= c) { r -= c; x >>= 2; pos += 2; }
  c = (uint32_t)x & 1u;
  if (r >= c) pos += 1;
  return pos;
}
enum LineStep : uint32_t { LS_FOUND, LS_ABSENT, LS_SEARCH, LS_RIGHT };
struct LineResult {
  uint32_t kind;
  uint32_t a, b;   // FOUND: index a; SEARCH: indices [a, b] of the line (b = ~0: to the last key)
  bool right_ok;   // SEARCH: a key above them all lies further right
};                 // RIGHT: the key lies above the sentinel (the next line)

// The run of key q in the line whose header admitted it (k0 <= q, q's cell
// below the line's cell count). Register chains pick words by index: a lane
// holds its own line, so no word index is uniform across the wave.
// fetch(w) gives word w of the line for a word index that differs from lane
// to lane (0 past the line): from registers by a select chain, or from the
// lane's slot in LDS by one read. The prefix counts over all words come from
// the register copy W.
template <int NW, typename Fetch>
__device__ __forceinline__ LineResult line_decode_t(const uint64_t (&W)[NW], uint64_t q, bool has_sent, uint32_t C, Fetch fetch) {
  LineResult R{LS_ABSENT, 0, 0, false};
  const uint64_t k0 = W[0] & ((1ULL << 62) - 1);
  if ((W[0] >> 62) & 1) {
    R.kind = LS_SEARCH;
    R.a = 0;
    R.b = has_sent ? C : 0xFFFFFFFFu;
    R.right_ok = has_sent;
    return R;
  }
  const uint32_t t = (uint32_t)W[1] & 63u, ulen = (uint32_t)(W[1] >> 6) & 1023u;
  const uint32_t cw = 3u + (uint32_t)(W[0] >> 63), cmask = (1u << cw) - 1u;
  const uint64_t rel = q - k0;
  const uint32_t cq = (uint32_t)(rel >> t);
  // Zero #(cq - 1) closes the cell before q's: the run starts right after it
  // (or at bit 88 for cq = 0). Zeros are counted from bit 88 on (the header
  // bits of word 1 read as ones); words past the unary hold codes, but the
  // zero sought lies inside the unary, so their counts never decide. A 1
  // right after a 1 is a key in its predecessor's cell (one code each):
  // their counts before each word give the codes before the run.
  uint32_t start = LINE_HDR_BITS;
  uint32_t cz = 0, zbase = 0, wa = 0, ccum = 0;
  uint32_t ccs[NW];
  const uint32_t k = cq - 1;
#pragma unroll
  for (int w = 1; w < NW; w++) {
    const uint64_t Z = ~(w == 1 ? (W[1] | 0xFFFFFFULL) : W[w]);
    const bool in = cq && cz <= k;  // zero #k lies in this word or a later one
    zbase = in ? cz : zbase;
    wa += in ? 1u : 0u;
    cz += __popcll(Z);
    uint64_t Cw = W[w] & ((W[w] << 1) | (W[w - 1] >> 63));
    if (w == 1) Cw &= ~0x1FFFFFFULL;  // bits 64..88: header, and the first key
    ccs[w] = ccum;
    ccum += __popcll(Cw);
  }
  if (cq) {
    const uint64_t zsel = ~(wa == 1 ? (fetch(1) | 0xFFFFFFULL) : fetch(wa));
    start = 64u * wa + select64(zsel, k - zbase) + 1u;
  }
  // the run of ones from `start`
  const uint32_t ws = start >> 6, os = start & 63u;
  const uint64_t w0 = fetch(ws), w1 = fetch(ws + 1), wp = fetch(ws - 1);
  const uint64_t win = os ? ((w0 >> os) | (w1 << (64 - os))) : w0;
  const uint32_t s = (uint32_t)__ffsll((long long)~win) - 1u;  // ones from `start` (64: a longer run)
  const uint32_t j0 = start - LINE_HDR_BITS - cq;
  if (~win == 0ULL) {
    // a run longer than the window: search its keys to the line's end
    R.kind = LS_SEARCH;
    R.a = j0;
    R.b = has_sent ? C : 0xFFFFFFFFu;
    R.right_ok = has_sent;
    return R;
  }
  if (s == 0) return R;  // q's cell is empty: absent
  uint32_t n = s;        // keys of the run below the sentinel
  if (has_sent && j0 + s - 1 >= C) {
    // the run holds the sentinel (the next line's first key): the prefixes
    // of their offsets in the cell decide, a tie is searched
    const uint32_t qp = sentinel_prefix(rel, t), sp = (uint32_t)(W[1] >> 16) & 255u;
    if (qp > sp) {
      R.kind = LS_RIGHT;
      return R;
    }
    if (qp == sp) {
      R.kind = LS_SEARCH;
      R.a = j0;
      R.b = j0 + s - 1;
      R.right_ok = true;
      return R;
    }
    n = s - 1;
    if (n == 0) return R;  // below the sentinel, alone in its cell: absent
  }
  const uint32_t last = j0 + n - 1;
  if (n - 1 > 64u / cw) {
    // more codes than one window holds: search the run's keys
    R.kind = LS_SEARCH;
    R.a = j0;
    R.b = last;
    return R;
  }
  if (n == 1) {
    R.kind = LS_FOUND;
    R.a = j0;
    return R;
  }
  uint64_t csel = w0 & ((w0 << 1) | (wp >> 63));
  if (ws == 1) csel &= ~0x1FFFFFFULL;
  uint32_t cbase = 0;
#pragma unroll
  for (int w = 1; w < NW; w++) cbase = ((uint32_t)w == ws) ? ccs[w] : cbase;
  const uint32_t m0 = cbase + __popcll(csel & ((1ULL << os) - 1));
  const uint32_t cpos = LINE_HDR_BITS + ulen + cw * m0;
  const uint32_t cwd = cpos >> 6, co = cpos & 63u;
  const uint64_t c0 = fetch(cwd), c1 = fetch(cwd + 1);
  const uint64_t codes = co ? ((c0 >> co) | (c1 << (64 - co))) : c0;
  // a search code (all ones) anywhere in the run: search it
  uint64_t sev = cw == 3u ? (codes & (codes >> 1) & (codes >> 2) & 0x1249249249249249ULL)
                          : (codes & (codes >> 1) & (codes >> 2) & (codes >> 3) & 0x1111111111111111ULL);
  const uint32_t used = cw * (n - 1);
  sev &= used >= 64u ? ~0ULL : ((1ULL << used) - 1);
  if (sev) {
    R.kind = LS_SEARCH;
    R.a = j0;
    R.b = last;
    return R;
  }
  // blind descent: the boundary of least code (highest differing bit) splits
  uint32_t lo = 0, hb = n - 1;
  while (lo < hb) {
    uint32_t bc = 16, bp = lo + 1;
    for (uint32_t x = lo + 1; x <= hb; x++) {
      const uint32_t c = (uint32_t)(codes >> (cw * (x - 1))) & cmask;
      if (c < bc) {
        bc = c;
        bp = x;
      }
    }
    if ((rel >> (t - 1 - bc)) & 1) lo = bp; else hb = bp - 1;
  }
  R.kind = LS_FOUND;
  R.a = j0 + lo;
  return R;
}
template <int NW>
__device__ __forceinline__ LineResult line_decode(const uint64_t (&W)[NW], uint64_t q, bool has_sent, uint32_t C) {
  return line_decode_t<NW>(W, q, has_sent, C, [&](uint32_t w) {
    uint64_t v = 0;
#pragma unroll
    for (int x = 1; x < NW; x++) v = ((uint32_t)x == w) ? W[x] : v;
    return v;
  });
}
// The same from the lane's LDS slot (the line's words written there first).
template <int NW>
__device__ __forceinline__ LineResult line_decode_lds(const uint64_t (&W)[NW], uint64_t* slot, uint64_t q, bool has_sent, uint32_t C) {
#pragma unroll
  for (int w = 0; w < NW; w++) slot[w] = W[w];
  return line_decode_t<NW>(W, q, has_sent, C, [&](uint32_t w) { return w < (uint32_t)NW ? slot[w] : 0ULL; });
}

template <int NW>
__device__ __forceinline__ void load_line(const uint64_t* __restrict__ lines, uint32_t i, uint64_t (&W)[NW]) {
  const uint4* lp = (const uint4*)(lines + (uint64_t)i * NW);
#pragma unroll
  for (int u = 0; u < NW / 2; u++) {
    const uint4 v = lp[u];
    W[2 * u] = (uint64_t)v.x | ((uint64_t)v.y << 32);
    W[2 * u + 1] = (uint64_t)v.z | ((uint64_t)v.w << 32);
  }
}

// The line a key's predicted rank falls in (leaves without a BBHash mark;
// no branch, so that two keys' loads can be in flight together). ok = false
// for a key wider than 2k bits or an empty index.
__device__ __forceinline__ uint32_t lidx_pred_line(const LidxView& V, const uint32_t* __restrict__ warp,
                                                   const RootPiece* __restrict__ root, uint64_t key, bool& ok, LeafRec& Lf) {
  ok = !((V.key_bits < 64 && (key >> V.key_bits) != 0) || V.n_lines == 0);
  const uint64_t u = warp_key(key, warp, V.key_bits, V.warp_groups);
  const uint32_t r = (uint32_t)(u >> V.s);
  const RootPiece P = root[r];
  const uint32_t leaf0n = root[r + 1].leaf0;
  const float x = piece_x(u, V.s);
  Lf = V.leaf[root_leaf(P, leaf0n, x)];
  int64_t pred = leaf_pred(P, Lf, x);
  pred = pred < 0 ? 0 : (pred >= (int64_t)V.D ? (int64_t)V.D - 1 : pred);
  const uint32_t pu = (uint32_t)pred, th = __umulhi(pu, V.magic);
  const uint32_t i = (th + ((pu - th) >> 1)) >> V.mshift;
  return i < V.n_lines ? i : (V.n_lines ? V.n_lines - 1 : 0);
}

// lidx_pred_line for NK keys in stages (warp + root records, then leaves,
// then lines), each stage's loads issued for every key before any is used.
template <int NK>
__device__ __forceinline__ void lidx_pred_lines(const LidxView& V, const uint32_t* __restrict__ warp, const RootPiece* __restrict__ root,
                                                const uint64_t (&key)[NK], bool (&ok)[NK], uint32_t (&line)[NK]) {
  uint32_t r[NK], leaf0[NK], leaf0n[NK], p[NK];
  float A[NK], B[NK], x[NK];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    ok[k] = !((V.key_bits < 64 && (key[k] >> V.key_bits) != 0) || V.n_lines == 0);
    const uint64_t u = warp_key(key[k], warp, V.key_bits, V.warp_groups);
    r[k] = (uint32_t)(u >> V.s);
    x[k] = piece_x(u, V.s);
  }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const RootPiece P = root[r[k]];
    A[k] = P.A;
    B[k] = P.B;
    leaf0[k] = P.leaf0;
    p[k] = P.p;
    leaf0n[k] = root[r[k] + 1].leaf0;
  }
  LeafRec Lf[NK];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    RootPiece P;
    P.A = A[k];
    P.B = B[k];
    P.leaf0 = leaf0[k];
    P.p = p[k];
    Lf[k] = V.leaf[root_leaf(P, leaf0n[k], x[k])];
  }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    RootPiece P;
    P.p = p[k];
    int64_t pred = leaf_pred(P, Lf[k], x[k]);
    pred = pred < 0 ? 0 : (pred >= (int64_t)V.D ? (int64_t)V.D - 1 : pred);
    const uint32_t pu = (uint32_t)pred, th = __umulhi(pu, V.magic);
    const uint32_t i = (th + ((pu - th) >> 1)) >> V.mshift;
    line[k] = i < V.n_lines ? i : (V.n_lines ? V.n_lines - 1 : 0);
  }
}

// Waits for a line's words where it stands, so that the compiler does not
// move the loads of the words the walk does not read past the walk (which
// would cost a second round trip to L2 per lookup).
template <int NW>
__device__ __forceinline__ void hold_line(const uint64_t (&W)[NW]) {
#pragma unroll
  for (int w = 0; w < NW; w++) asm volatile("" ::"v"(W[w]));
}

// From line i (words W already loaded) to the key's rank, or D when absent.
// The lines' headers alone decide where the key lies (before the first key:
// left; at or past the cell count: right), so the walk touches no code and
// the run is decoded once; a search that finds the key above every key it
// covers (a run holding the next line's first key, or an overflow line) sends
// the walk on to the next line. Every step visits one line, and the walk from
// the predicted line to the key's own takes at most max_steps of them.
template <int NW>
__device__ __forceinline__ uint64_t lidx_finish(const LidxView& V, uint64_t key, uint32_t i, uint64_t (&W)[NW]) {
  uint32_t steps = 1;
  while (true) {
    bool here = false;
    while (true) {
      const uint64_t k0 = W[0] & ((1ULL << 62) - 1);
      bool move = false;
      if (key < k0) {
        if (i == 0) return V.D;
        i--;
        move = true;
      } else if (!((W[0] >> 62) & 1)) {
        const uint32_t t = (uint32_t)W[1] & 63u, ulen = (uint32_t)(W[1] >> 6) & 1023u;
        const uint64_t base = (uint64_t)i * V.C;
        const uint32_t m = i + 1 < V.n_lines ? V.C + 1 : (uint32_t)(V.D - base);
        if (((key - k0) >> t) >= (uint64_t)(ulen - m)) {  // past the last cell
          if (i + 1 >= V.n_lines) return V.D;
          i++;
          move = true;
        }
      }
      if (!move) {
        here = true;
        break;
      }
      if (steps >= V.max_steps) break;
      load_line<NW>(V.lines, i, W);
      hold_line<NW>(W);
      steps++;
    }
    if (!here) return V.D;
    const bool has_sent = i + 1 < V.n_lines;
    const LineResult res = line_decode<NW>(W, key, has_sent, V.C);
    const uint64_t base = (uint64_t)i * V.C;
    if (res.kind == LS_FOUND) return base + res.a;
    if (res.kind == LS_ABSENT) return V.D;
    if (res.kind == LS_RIGHT) {  // above the sentinel: on to the next line
      if (steps >= V.max_steps || i + 1 >= V.n_lines) return V.D;
      i++;
      load_line<NW>(V.lines, i, W);
      hold_line<NW>(W);
      steps++;
      continue;
    }
    // search ranks [base + a, base + b] of the sorted keys: up to 4 keys
    // in one round of loads, a binary search beyond
    const uint64_t a = base + res.a;
    const uint64_t b = res.b == 0xFFFFFFFFu ? V.D - 1 : base + res.b;
    uint64_t lo = a, hi = b + 1;
    if (b - a < 4) {
      uint64_t kk[4];
#pragma unroll
      for (int x = 0; x < 4; x++) kk[x] = a + x <= b ? V.keys[a + x] : ~0ULL;
      uint32_t below = 0;
      uint64_t hit = ~0ULL;
#pragma unroll
      for (int x = 0; x < 4; x++) {
        below += (a + x <= b && kk[x] < key) ? 1u : 0u;
        hit = (a + x <= b && kk[x] == key) ? a + x : hit;
      }
      lo = a + below;
      if (hit != ~0ULL) return hit;
    } else {
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (V.keys[mid] < key) lo = mid + 1; else hi = mid;
      }
      if (lo <= b && V.keys[lo] == key) return lo;
    }
    if (!(res.right_ok && lo == b + 1) || steps >= V.max_steps) return V.D;
    i++;
    load_line<NW>(V.lines, i, W);
    hold_line<NW>(W);
    steps++;
  }
}

// Run code of a deferred search: a (bits 0-12) and b (bits 13-25, LQ_TO_END:
// to the last key) as indices of the line (C <= 4096), bit 26: a key above
// every key searched goes on to the next line.
constexpr uint32_t LQ_TO_END = 8191;

// Search of ranks [a, b] of the sorted keys for key (up to 4 keys in one
// round of loads, a binary search beyond); when the key lies above them all
// and cont is set, the walk goes on from line i + 1 (a = b + 1: no search, the
// key lies above line i's sentinel). Rank, or D when absent.
template <int NW>
__device__ __forceinline__ uint64_t lidx_search(const LidxView& V, uint64_t key, uint32_t i, uint64_t a, uint64_t b, bool cont) {
  uint64_t lo = a;
  if (a > b) {
    // nothing to search
  } else if (b - a < 4) {
    uint64_t kk[4];
#pragma unroll
    for (int x = 0; x < 4; x++) kk[x] = a + x <= b ? V.keys[a + x] : ~0ULL;
    uint32_t below = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
      if (a + x <= b && kk[x] == key) return a + x;
      below += (a + x <= b && kk[x] < key) ? 1u : 0u;
    }
    lo = a + below;
  } else {
    uint64_t hi = b + 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (V.keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo <= b && V.keys[lo] == key) return lo;
  }
  if (!(cont && lo == b + 1 && i + 1 < V.n_lines)) return V.D;
  uint64_t Wn[NW];
  load_line<NW>(V.lines, i + 1, Wn);
  hold_line<NW>(Wn);
  return lidx_finish<NW>(V, key, i + 1, Wn);
}

// lidx_finish for NK keys in lockstep: the header walk of both, each
// round's neighbour loads issued together; both runs decoded. A key whose
// run must be searched goes to the wave's queue segment (qn entries so far,
// the same count in every lane; deferred[k] set) while it has room: the
// wave then waits for no search loads. The others (a full segment, or no
// queue) are searched in place, one key at a time.
template <int NW, int NK>
__device__ __forceinline__ void lidx_finish_n(const LidxView& V, const uint64_t (&key)[NK], const bool (&ok)[NK], uint32_t (&i)[NK],
                                              uint64_t (&W)[NK][NW], uint64_t* slot, uint64_t (&res)[NK], const LidxQueue& Q,
                                              uint64_t* qseg, uint32_t& qn, uint64_t q0, uint64_t qstride, bool (&deferred)[NK]) {
  bool open[NK];  // still walking
#pragma unroll
  for (int k = 0; k < NK; k++) {
    res[k] = V.D;
    open[k] = ok[k];
  }
  uint32_t steps = 1;
  while (true) {
    bool moved[NK];
    bool any = false;
#pragma unroll
    for (int k = 0; k < NK; k++) moved[k] = false;
#pragma unroll
    for (int k = 0; k < NK; k++) {
      if (!open[k]) continue;
      const uint64_t k0 = W[k][0] & ((1ULL << 62) - 1);
      if (key[k] < k0) {
        if (i[k] == 0) open[k] = false; else { i[k]--; moved[k] = true; }
      } else if (!((W[k][0] >> 62) & 1)) {
        const uint32_t t = (uint32_t)W[k][1] & 63u, ulen = (uint32_t)(W[k][1] >> 6) & 1023u;
        const uint64_t base = (uint64_t)i[k] * V.C;
        const uint32_t m = i[k] + 1 < V.n_lines ? V.C + 1 : (uint32_t)(V.D - base);
        if (((key[k] - k0) >> t) >= (uint64_t)(ulen - m)) {
          if (i[k] + 1 >= V.n_lines) open[k] = false; else { i[k]++; moved[k] = true; }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NK; k++) any |= moved[k];
    if (!any) break;
    if (steps >= V.max_steps) {
#pragma unroll
      for (int k = 0; k < NK; k++) open[k] = open[k] && !moved[k];
      break;
    }
#pragma unroll
    for (int k = 0; k < NK; k++)
      if (moved[k]) load_line<NW>(V.lines, i[k], W[k]);
#pragma unroll
    for (int k = 0; k < NK; k++) hold_line<NW>(W[k]);
    steps++;
  }
  LineResult r[NK];
  uint64_t a[NK], b[NK];
  bool srch[NK];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    srch[k] = false;
    deferred[k] = false;
    a[k] = b[k] = 0;
    if (!open[k]) continue;
    r[k] = line_decode_lds<NW>(W[k], slot, key[k], i[k] + 1 < V.n_lines, V.C);
    const uint64_t base = (uint64_t)i[k] * V.C;
    if (r[k].kind == LS_FOUND) res[k] = base + r[k].a;
    if (r[k].kind == LS_RIGHT) {  // above the sentinel: an empty search that goes on to the next line
      r[k].kind = LS_SEARCH;
      r[k].a = V.C + 1;
      r[k].b = V.C;
      r[k].right_ok = true;
    }
    if (r[k].kind == LS_SEARCH) {
      srch[k] = true;
      a[k] = base + r[k].a;
      b[k] = r[k].b == 0xFFFFFFFFu ? V.D - 1 : base + r[k].b;
    }
  }
  if (Q.cap) {
    const uint64_t below = (1ULL << __lane_id()) - 1;
#pragma unroll
    for (int k = 0; k < NK; k++) {
      const uint64_t m = __ballot(srch[k]);
      if (!m) continue;
      const uint32_t pos = qn + (uint32_t)__popcll(m & below);
      qn = __builtin_amdgcn_readfirstlane(qn + (uint32_t)__popcll(m));
      if (srch[k] && pos < Q.cap) {
        const uint32_t bc = r[k].b == 0xFFFFFFFFu ? LQ_TO_END : r[k].b;
        const uint32_t code = r[k].a | (bc << 13) | ((r[k].right_ok && steps < V.max_steps) ? 1u << 26 : 0u);
        uint64_t* e = qseg + 3 * (uint64_t)pos;
        e[0] = key[k];
        e[1] = q0 + k * qstride;
        e[2] = (uint64_t)i[k] | ((uint64_t)code << 32);
        srch[k] = false;
        deferred[k] = true;
      }
    }
  }
  // searches that found no room in the queue, one key at a time
#pragma unroll
  for (int k = 0; k < NK; k++)
    if (srch[k]) res[k] = lidx_search<NW>(V, key[k], i[k], a[k], b[k], r[k].right_ok && steps < V.max_steps);
}

// Rank of key q, or D when absent, one key at a time (FB: leaves may resolve
// through the BBHash).
template <int NW, bool FB>
__device__ __forceinline__ uint64_t lidx_rank_t(const LidxView& V, const MphfView& mv, const uint32_t* __restrict__ warp,
                                                const RootPiece* __restrict__ root, uint64_t key) {
  bool ok;
  LeafRec Lf;
  const uint32_t i = lidx_pred_line(V, warp, root, key, ok, Lf);
  if (!ok) return V.D;
  if (FB && leaf_fallback(Lf)) {
    const uint64_t v = mphf_lookup<uint64_t>(mv, key);
    const uint64_t rk = v < V.D ? V.bbh2rank[v] : V.D;
    return (rk < V.D && V.keys[rk] == key) ? rk : V.D;
  }
  uint64_t W[NW];
  load_line<NW>(V.lines, i, W);
  hold_line<NW>(W);
  return lidx_finish<NW>(V, key, i, W);
}
__device__ __forceinline__ uint64_t lidx_rank(const LidxView& V, const MphfView& mv, uint64_t key) {
  return V.line_words == 8 ? lidx_rank_t<8, true>(V, mv, V.warp, V.root, key) : lidx_rank_t<16, true>(V, mv, V.warp, V.root, key);
}

// The warp tables (and, with ROOT_LDS, the root records) go to LDS once per
// work-group. Without BBHash leaves a lane carries two keys (i and i +
// stride) through the front of the lookup together: both predictions, then
// both first lines in flight at once; the next two keys are loaded while the
// current ones are resolved.
// Compiled for 6 waves per SIMD (80 VGPRs without spills for 64-byte lines):
// the kernel waits on dependent far loads, so waves in flight set its rate.
constexpr uint32_t LIDX_LDS_ROOT_MAX = 256;  // root pieces held in LDS (4 KiB)
constexpr int LIDX_NK = 2;                    // keys a lane carries through one round
template <int NW, bool FB, bool ROOT_LDS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8))) k_lidx_lookup(const uint64_t* __restrict__ keys, uint64_t n, const LidxView V, const MphfView mv,
                                                     uint64_t* __restrict__ out, const LidxQueue Q) {
  __shared__ uint32_t wt[WARP_MAX_GROUPS * 256];
  __shared__ RootPiece rt[ROOT_LDS ? LIDX_LDS_ROOT_MAX + 1 : 1];
  __shared__ uint64_t lslot[FB ? 1 : 256 * (NW + 1)];
  for (int j = threadIdx.x; j < V.warp_groups * 256; j += 256) wt[j] = V.warp[j];
  if (ROOT_LDS)
    for (uint32_t j = threadIdx.x; j <= V.R; j += 256) rt[j] = V.root[j];
  __syncthreads();
  const RootPiece* root = ROOT_LDS ? rt : V.root;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (FB) {
    for (; i < n; i += stride) {
      const uint64_t r = lidx_rank_t<NW, true>(V, mv, wt, root, keys[i]);
      out[i] = r < V.D ? r : 0;
    }
    return;
  }
  // the wave's queue segment; lane 0 holds the smallest index of its wave,
  // so it runs every round that any lane of the wave runs
  const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint64_t* qseg = Q.ent + wave * Q.cap * 3;
  uint32_t qn = 0;
  // queries and results stream through once: non-temporal loads and stores
  uint64_t next[LIDX_NK];
#pragma unroll
  for (int k = 0; k < LIDX_NK; k++) next[k] = i < n ? __builtin_nontemporal_load(&keys[i + k * stride < n ? i + k * stride : i]) : 0;
  for (; i < n; i += LIDX_NK * stride) {
    uint64_t key[LIDX_NK];
#pragma unroll
    for (int k = 0; k < LIDX_NK; k++) key[k] = next[k];
    const uint64_t i2 = i + LIDX_NK * stride;
    if (i2 < n) {
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) next[k] = __builtin_nontemporal_load(&keys[i2 + k * stride < n ? i2 + k * stride : i2]);
    }
    bool ok[LIDX_NK];
    uint32_t line[LIDX_NK];
    lidx_pred_lines<LIDX_NK>(V, wt, root, key, ok, line);
#pragma unroll
    for (int k = 0; k < LIDX_NK; k++) ok[k] = ok[k] && i + k * stride < n;  // a slot past the end resolves nothing
    uint64_t W[LIDX_NK][NW];
#pragma unroll
    for (int k = 0; k < LIDX_NK; k++) load_line<NW>(V.lines, line[k], W[k]);
#pragma unroll
    for (int k = 0; k < LIDX_NK; k++) hold_line<NW>(W[k]);
    uint64_t res[LIDX_NK];
    bool deferred[LIDX_NK];
    lidx_finish_n<NW, LIDX_NK>(V, key, ok, line, W, lslot + threadIdx.x * (NW + 1), res, Q, qseg, qn, i, stride, deferred);
#pragma unroll
    for (int k = 0; k < LIDX_NK; k++)
      if (i + k * stride < n && !deferred[k]) __builtin_nontemporal_store(res[k] < V.D ? res[k] : 0, &out[i + k * stride]);
  }
  if (Q.cap && __lane_id() == 0) Q.cnt[wave] = qn < Q.cap ? qn : Q.cap;
}

// The searches the lookup deferred: one wave per wave of the lookup's grid,
// over its queue segment.
template <int NW>
__global__ void __launch_bounds__(64) k_lidx_search(const LidxView V, const LidxQueue Q, uint64_t* __restrict__ out) {
  const uint64_t wave = blockIdx.x;
  const uint32_t cnt = Q.cnt[wave];
  const uint64_t* seg = Q.ent + wave * Q.cap * 3;
  for (uint32_t e = threadIdx.x; e < cnt; e += 64) {
    const uint64_t key = seg[3 * e], idx = seg[3 * e + 1], lc = seg[3 * e + 2];
    const uint32_t line = (uint32_t)lc, code = (uint32_t)(lc >> 32);
    const uint64_t base = (uint64_t)line * V.C;
    const uint32_t ar = code & 8191u, bc = (code >> 13) & 8191u;
    const uint64_t r = lidx_search<NW>(V, key, line, base + ar, bc == LQ_TO_END ? V.D - 1 : base + bc, (code >> 26) & 1u);
    out[idx] = r < V.D ? r : 0;
  }
}

// Count of out[i] != expect[i].
__global__ void __launch_bounds__(256) k_count_mismatch(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint64_t n,
                                                        unsigned long long* __restrict__ bad) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) c += a[i] != b[i];
  for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
  if (__lane_id() == 0 && c) atomicAdd(bad, (unsigned long long)c);
}

// ------------------------------------------------------------------ classification through an index
// P4 in the reference's form (classify.rs:524-601) through the learned
// index: one lane per k-mer occurrence of the flat stream, in sequence order,
// looks the canonical k-mer up (its rank) and joins the occurrence's
// observation into the state at that rank by CAS; the persistent grid takes
// chunks of stream positions from a counter, as k_p4_cas does for the
// BBHash. counters: [0] occurrences, [1] CAS issued, [2] CAS lost, [3]
// lookups that found no rank, [4] the chunk ticket.
template <int NW, bool FB>
__global__ void __launch_bounds__(256) k_p4_cas_lidx(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                     uint64_t n_pos, int K, const LidxView V, const MphfView mv,
                                                     unsigned long long* __restrict__ states, unsigned long long* __restrict__ counters,
                                                     uint32_t chunk, uint64_t n_chunks, int force_retry) {
  __shared__ uint32_t wt[WARP_MAX_GROUPS * 256];
  __shared__ RootPiece rt[LIDX_LDS_ROOT_MAX + 1];
  __shared__ uint64_t lslot[FB ? 1 : 256 * (NW + 1)];
  __shared__ uint64_t s_c;
  __shared__ unsigned long long s_sum[4];
  for (int j = threadIdx.x; j < V.warp_groups * 256; j += 256) wt[j] = V.warp[j];
  const bool rl = V.R <= LIDX_LDS_ROOT_MAX;
  if (rl)
    for (uint32_t j = threadIdx.x; j <= V.R; j += 256) rt[j] = V.root[j];
  if (threadIdx.x < 4) s_sum[threadIdx.x] = 0;
  const RootPiece* root = rl ? rt : V.root;
  uint32_t n_occ = 0, n_cas = 0, n_fail = 0, n_miss = 0;
  const LidxQueue noq{nullptr, nullptr, 0};  // searches in place
  uint32_t qn = 0;
  for (;;) {
    if (threadIdx.x == 0) s_c = atomicAdd(&counters[4], 1ULL);
    __syncthreads();
    const uint64_t c = s_c;
    __syncthreads();
    if (c >= n_chunks) break;
    const uint64_t g0 = 1 + c * chunk;
    const uint64_t g1 = g0 + chunk < n_pos ? g0 + chunk : n_pos;
    if (FB) {
      for (uint64_t g = g0 + threadIdx.x; g < g1; g += 256) {
        uint64_t canon;
        const uint32_t m = observe<uint64_t>(seq, nmask, g, K, canon);
        if (!m) continue;
        n_occ++;
        const uint64_t idx = lidx_rank_t<NW, FB>(V, mv, wt, root, canon);
        if (idx >= V.D) {
          n_miss++;
          continue;
        }
        cas_join(states, idx, m, force_retry != 0, n_cas, n_fail);
      }
      continue;
    }
    // LIDX_NK occurrences per lane (positions g + 256 k) through the
    // lookup's lockstep front and walk, then their CAS joins
    for (uint64_t g = g0 + threadIdx.x; g < g1; g += 256 * LIDX_NK) {
      uint64_t key[LIDX_NK];
      uint32_t m[LIDX_NK];
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) {
        key[k] = 0;
        m[k] = g + 256 * k < g1 ? observe<uint64_t>(seq, nmask, g + 256 * k, K, key[k]) : 0u;
      }
      bool ok[LIDX_NK];
      uint32_t line[LIDX_NK];
      lidx_pred_lines<LIDX_NK>(V, wt, root, key, ok, line);
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) ok[k] = ok[k] && m[k] != 0;
      uint64_t W[LIDX_NK][NW];
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) load_line<NW>(V.lines, line[k], W[k]);
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) hold_line<NW>(W[k]);
      uint64_t res[LIDX_NK];
      bool deferred[LIDX_NK];
      lidx_finish_n<NW, LIDX_NK>(V, key, ok, line, W, lslot + threadIdx.x * (NW + 1), res, noq, nullptr, qn, 0, 0, deferred);
#pragma unroll
      for (int k = 0; k < LIDX_NK; k++) {
        if (!m[k]) continue;
        n_occ++;
        if (res[k] >= V.D) {
          n_miss++;
          continue;
        }
        cas_join(states, res[k], m[k], force_retry != 0, n_cas, n_fail);
      }
    }
  }
  atomicAdd(&s_sum[0], (unsigned long long)n_occ);
  atomicAdd(&s_sum[1], (unsigned long long)n_cas);
  atomicAdd(&s_sum[2], (unsigned long long)n_fail);
  atomicAdd(&s_sum[3], (unsigned long long)n_miss);
  __syncthreads();
  if (threadIdx.x < 4) atomicAdd(&counters[threadIdx.x], s_sum[threadIdx.x]);
}

// Per stream position through the learned index: bit0 valid k-mer, bit1 FWD,
// bits2-3 state class (the learned-index counterpart of k_annotate).
__global__ void __launch_bounds__(256) k_annotate_lidx(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                       uint64_t n_pos, int K, const LidxView V, const MphfView mv,
                                                       const uint64_t* __restrict__ states, uint8_t* __restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pos) return;
  uint8_t a = 0;
  if (g >= 1 && window_clear(nmask, g, K)) {
    const uint64_t x = kmer64_at(seq, g, K);
    const uint64_t r = rc64(x, K);
    const uint64_t c = x < r ? x : r;
    const uint64_t idx = lidx_rank(V, mv, c);
    const uint32_t s = idx < V.D ? (uint32_t)(states[idx / 12] >> ((idx % 12) * 5)) & 31u : 0u;
    a = (uint8_t)(1u | (x == c ? 2u : 0u) | (st_class(s) << 2));
  }
  out[g] = a;
}

// ------------------------------------------------------------------ launchers
static unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

uint64_t sort_tiles(uint64_t n) { return (n + SRT_TILE - 1) / SRT_TILE; }
void launch_sort_hist(hipStream_t st, const uint64_t* src, uint64_t n, int sh, uint64_t* mat) {
  const uint64_t nt = sort_tiles(n);
  if (!nt) return;
  CF1G_LAUNCH("lidx_sort_hist", k_sort_hist, dim3((unsigned)nt), dim3(SRT_WG), 0, st, src, n, sh, nt, mat);
}
void launch_sort_scatter(hipStream_t st, const uint64_t* src, uint64_t* dst, uint64_t n, int sh, const uint64_t* off) {
  const uint64_t nt = sort_tiles(n);
  if (!nt) return;
  CF1G_LAUNCH("lidx_sort_scatter", k_sort_scatter, dim3((unsigned)nt), dim3(SRT_WG), 0, st, src, dst, n, sh, nt, off);
}

// One level's training: stats, init, epochs of forward / backward / update.
// Returns through loss_out the mean loss before the first and after the last
// update (sum over segments / D, computed on the host from the loss array).
void launch_lidx_stats_init(hipStream_t st, int level, const SegCtx& c, uint64_t n_segs, const LidxTrainBufs& b) {
  if (level == 0) {
    CF1G_LAUNCH("lidx_stats", (k_lidx_stats<0, 256>), dim3((unsigned)n_segs), dim3(256), 0, st, c, n_segs, b.mu, b.isd);
    CF1G_LAUNCH("lidx_init", k_lidx_init<0>, dim3(blocks_for(n_segs, 256)), dim3(256), 0, st, c, n_segs, b.mu, b.isd, b.w, b.c);
  } else {
    CF1G_LAUNCH("lidx_stats", (k_lidx_stats<1, 64>), dim3(blocks_for(n_segs, 4)), dim3(256), 0, st, c, n_segs, b.mu, b.isd);
    CF1G_LAUNCH("lidx_init", k_lidx_init<1>, dim3(blocks_for(n_segs, 256)), dim3(256), 0, st, c, n_segs, b.mu, b.isd, b.w, b.c);
  }
  CF1G_LAUNCH("lidx_tile_seg", k_lidx_tile_seg, dim3(blocks_for((c.D + 255) / 256, 256)), dim3(256), 0, st, c.seg, n_segs, c.D,
              b.seg_of_tile);
}
void launch_lidx_fwd_bwd(hipStream_t st, int level, const SegCtx& c, uint64_t n_segs, const LidxTrainBufs& b) {
  if (c.D == 0) return;
  if (level == 0) {
    CF1G_LAUNCH("lidx_forward", k_lidx_forward<0>, dim3(blocks_for(c.D, 256)), dim3(256), 0, st, c, n_segs, b.mu, b.isd, b.w, b.c,
                b.seg_of_tile, b.resid);
    CF1G_LAUNCH("lidx_backward", (k_lidx_backward<0, 256>), dim3((unsigned)n_segs), dim3(256), 0, st, c, n_segs, b.mu, b.isd, b.resid,
                b.gw, b.gc, b.loss);
  } else {
    CF1G_LAUNCH("lidx_forward", k_lidx_forward<1>, dim3(blocks_for(c.D, 256)), dim3(256), 0, st, c, n_segs, b.mu, b.isd, b.w, b.c,
                b.seg_of_tile, b.resid);
    CF1G_LAUNCH("lidx_backward", (k_lidx_backward<1, 64>), dim3(blocks_for(n_segs, 4)), dim3(256), 0, st, c, n_segs, b.mu, b.isd,
                b.resid, b.gw, b.gc, b.loss);
  }
}
void launch_lidx_update(hipStream_t st, uint64_t n_segs, float lr, const LidxTrainBufs& b) {
  CF1G_LAUNCH("lidx_update", k_lidx_update, dim3(blocks_for(n_segs, 256)), dim3(256), 0, st, n_segs, lr, b.gw, b.gc, b.w, b.c);
}
void launch_lidx_pieces(hipStream_t st, const uint64_t* keys, uint64_t D, int s, uint32_t R, uint32_t L, uint64_t* p, uint32_t* lo) {
  CF1G_LAUNCH("lidx_pieces", k_lidx_pieces, dim3(blocks_for((uint64_t)R + 1, 256)), dim3(256), 0, st, keys, D, s, R, p);
  CF1G_LAUNCH("lidx_root_ranges", k_lidx_root_ranges, dim3(blocks_for((uint64_t)R + 1, 256)), dim3(256), 0, st, p, R, D, L, lo);
}
void launch_lidx_root_final(hipStream_t st, uint32_t R, uint64_t D, const uint64_t* p, const uint32_t* lo, const LidxTrainBufs& b,
                            RootPiece* out) {
  CF1G_LAUNCH("lidx_root_final", k_lidx_root_final, dim3(blocks_for((uint64_t)R + 1, 256)), dim3(256), 0, st, R, D, p, lo, b.mu, b.isd,
              b.w, b.c, out);
}
void launch_lidx_leaf_bounds(hipStream_t st, const uint64_t* keys, uint64_t D, int s, uint32_t LT, const RootPiece* root, uint64_t* q) {
  CF1G_LAUNCH("lidx_leaf_bounds", k_lidx_leaf_bounds, dim3(blocks_for(D + 1, 256)), dim3(256), 0, st, keys, D, s, LT, root, q);
}
void launch_lidx_leaf_final(hipStream_t st, const uint64_t* keys, uint64_t D, uint32_t LT, int s, const uint64_t* q,
                            const RootPiece* root, const LidxTrainBufs& b, uint32_t max_window, LeafRec* out, uint32_t* err,
                            uint64_t* sabs) {
  CF1G_LAUNCH("lidx_leaf_final", k_lidx_leaf_final, dim3(blocks_for(LT, 4)), dim3(256), 0, st, keys, D, LT, s, q, root, b.mu, b.isd, b.w,
              b.c, max_window, out, err, sabs);
}
void launch_lidx_lines(hipStream_t st, int line_words, const uint64_t* keys, uint64_t D, uint32_t C, uint64_t n_lines, uint64_t* lines,
                       uint32_t* stats) {
  if (!n_lines) return;
  if (line_words == 8)
    CF1G_LAUNCH("lidx_lines", k_lidx_lines<8>, dim3(blocks_for(n_lines, 256)), dim3(256), 0, st, keys, D, C, n_lines, lines, stats);
  else
    CF1G_LAUNCH("lidx_lines", k_lidx_lines<16>, dim3(blocks_for(n_lines, 256)), dim3(256), 0, st, keys, D, C, n_lines, lines, stats);
}
void launch_lidx_warp_hist(hipStream_t st, const uint64_t* keys, uint64_t D, int K2, int G, uint32_t* cnt) {
  if (!D || !G) return;
  uint64_t g = (D + 255) / 256;
  if (g > 4096) g = 4096;
  CF1G_LAUNCH("lidx_warp_hist", k_lidx_warp_hist, dim3((unsigned)g), dim3(256), 0, st, keys, D, K2, G, cnt);
}
void launch_lidx_warp_keys(hipStream_t st, const uint64_t* keys, uint64_t D, const uint32_t* cum, int K2, int G, uint64_t* u) {
  if (!D) return;
  CF1G_LAUNCH("lidx_warp_keys", k_lidx_warp_keys, dim3(blocks_for(D, 256)), dim3(256), 0, st, keys, D, cum, K2, G, u);
}
void launch_lidx_bbh2rank(hipStream_t st, const uint64_t* keys, uint32_t LT, const uint64_t* q, const LeafRec* leaf, const MphfView& mv,
                          uint32_t* bbh2rank) {
  CF1G_LAUNCH("lidx_bbh2rank", k_lidx_bbh2rank, dim3(blocks_for(LT, 4)), dim3(256), 0, st, keys, LT, q, leaf, mv, bbh2rank);
}
static uint64_t lidx_lookup_grid(uint64_t n) {
  uint64_t g = (n + 256 * LIDX_NK - 1) / (256 * LIDX_NK);  // LIDX_NK keys per lane a round
  return g > 8192 ? 8192 : g;  // each work-group fills its LDS tables once
}
// Deferred-search queue of a lookup of n keys: a segment per wave of the
// lookup's grid with room for 1/32 of the wave's keys (a multiple of 64, at
// least 64; searched keys are ~1.1% on the 300 Mbp genome), and at most
// 2^24 entries in all (402 MB) for very large calls. Segments that fill up
// search in place.
void lidx_queue_shape(uint64_t n, uint32_t& waves, uint32_t& cap) {
  waves = (uint32_t)(lidx_lookup_grid(n) * 4);
  if (!waves) {
    cap = 0;
    return;
  }
  const uint64_t per = (n + waves - 1) / waves;
  uint64_t c = (per / 32 + 63) / 64 * 64;
  const uint64_t most = ((1ULL << 24) / waves) / 64 * 64;
  if (c > most) c = most;
  cap = (uint32_t)(c < 64 ? 64 : c);
}
// Q: from lidx_queue_shape for this n (ent: waves x cap x 3 words, cnt:
// waves words), or cap = 0 to search in place; unused for BBHash leaves.
void launch_lidx_lookup(hipStream_t st, const uint64_t* keys, uint64_t n, const LidxView& V, const MphfView& mv, uint64_t* out,
                        LidxQueue Q) {
  if (!n) return;
  const uint64_t g = lidx_lookup_grid(n);
  const bool rl = V.R <= LIDX_LDS_ROOT_MAX;
  if (V.fallback) Q.cap = 0;
#define CF1G_LIDX_LOOKUP(NW_, FB_, RL_) \
  CF1G_LAUNCH("lidx_lookup", (k_lidx_lookup<NW_, FB_, RL_>), dim3((unsigned)g), dim3(256), 0, st, keys, n, V, mv, out, Q)
  if (V.line_words == 8) {
    if (V.fallback) {
      if (rl) CF1G_LIDX_LOOKUP(8, true, true); else CF1G_LIDX_LOOKUP(8, true, false);
    } else {
      if (rl) CF1G_LIDX_LOOKUP(8, false, true); else CF1G_LIDX_LOOKUP(8, false, false);
    }
  } else {
    if (V.fallback) {
      if (rl) CF1G_LIDX_LOOKUP(16, true, true); else CF1G_LIDX_LOOKUP(16, true, false);
    } else {
      if (rl) CF1G_LIDX_LOOKUP(16, false, true); else CF1G_LIDX_LOOKUP(16, false, false);
    }
  }
#undef CF1G_LIDX_LOOKUP
  if (Q.cap) {
    const dim3 sg((unsigned)(g * 4));
    if (V.line_words == 8) CF1G_LAUNCH("lidx_search", k_lidx_search<8>, sg, dim3(64), 0, st, V, Q, out);
    else CF1G_LAUNCH("lidx_search", k_lidx_search<16>, sg, dim3(64), 0, st, V, Q, out);
  }
}
void launch_count_mismatch(hipStream_t st, const uint64_t* a, const uint64_t* b, uint64_t n, unsigned long long* bad) {
  if (!n) return;
  uint64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  CF1G_LAUNCH("count_mismatch", k_count_mismatch, dim3((unsigned)g), dim3(256), 0, st, a, b, n, bad);
}
void launch_p4_cas_lidx(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, const LidxView& V,
                        const MphfView& mv, uint64_t* states, uint64_t* counters, uint32_t chunk, uint32_t grid, int force_retry) {
  if (n_pos <= 1 || chunk == 0 || grid == 0) return;
  const uint64_t n_chunks = (n_pos - 1 + chunk - 1) / chunk;
  auto* sv = (unsigned long long*)states;
  auto* cn = (unsigned long long*)counters;
#define CF1G_P4_LIDX(NW_, FB_)                                                                                                 \
  CF1G_LAUNCH("p4_cas_lidx", (k_p4_cas_lidx<NW_, FB_>), dim3(grid), dim3(256), 0, st, seq, nmask, n_pos, K, V, mv, sv, cn, chunk, \
              n_chunks, force_retry)
  if (V.line_words == 8) {
    if (V.fallback) CF1G_P4_LIDX(8, true); else CF1G_P4_LIDX(8, false);
  } else {
    if (V.fallback) CF1G_P4_LIDX(16, true); else CF1G_P4_LIDX(16, false);
  }
#undef CF1G_P4_LIDX
}
void launch_annotate_lidx(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, const LidxView& V,
                          const MphfView& mv, const uint64_t* states, uint8_t* out) {
  if (!n_pos) return;
  CF1G_LAUNCH("p5_annotate_lidx", k_annotate_lidx, dim3(blocks_for(n_pos, 256)), dim3(256), 0, st, seq, nmask, n_pos, K, V, mv, states,
              out);
}

}  // namespace cf1g
