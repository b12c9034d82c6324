// Learned rank index over the globally sorted distinct k-mers (u64 keys,
// k <= 31): a two-level piecewise-linear FP32 model, trained on the GPU by
// gradient descent with the forward / backward kernels below, that predicts
// a key's rank in the sorted array within a recorded per-leaf error bound; a
// bounded search turns the prediction into the exact rank (a monotone minimal
// perfect map). No reference counterpart: it sits beside the BBHash
// (mphf.rs:268-276 is the lookup it can replace).
//
//  sort     : LSD radix sort of the keys, 8-bit digits (k_sort_hist / scan /
//             k_sort_scatter), stable tile ranking in registers.
//  root     : R = 2^rb pieces by the top rb key bits. Piece r (keys
//             [p_r, p_{r+1})) maps x = key - (r << s) to a leaf:
//             leaf = lo_r + clamp(floor(fma(A_r, x, B_r)), 0, hi_r - lo_r),
//             trained towards y = rank * L / D - lo_r, with
//             lo_r = floor(p_r L / D), hi_r = max(lo_r, ceil(p_{r+1} L / D) - 1).
//  leaves   : L = 2^lb linear models over the keys the root routes to them
//             (contiguous ranks [q_j, q_{j+1}) because routing is monotone):
//             pred = q_j + floor(fma(S_j, (float)(key - base_j), I_j) + 0.5),
//             trained towards y = rank - q_j.
//  training : per segment (piece or leaf), inputs standardised with the
//             segment's mean / deviation: z = (x - mu) * isd, f = w z + c,
//             MSE loss; forward writes residuals f - y, backward reduces the
//             gradient (2/n) sum (f - y) (z, 1), update steps w, c. The
//             starting point (and the untrained method) is endpoint
//             interpolation through the segment's first and last key.
//  bounds   : elo_j = max(pred - rank), ehi_j = max(rank - pred) over the
//             leaf's keys, evaluated with the exact lookup arithmetic, so a
//             present key always lies in [pred - elo, pred + ehi].
//  search   : over a 2-byte array aligned with the sorted keys: the 16-bit
//             fingerprint (key - base_j) >> shift_j, non-decreasing inside a
//             leaf. The walk starts at the 16-byte chunk holding the
//             prediction and moves left / right by comparing fingerprints;
//             a fingerprint shared with a neighbour resolves with the full
//             keys, and leaves
//             whose window exceeds the threshold resolve through the BBHash
//             value mapped to its rank.
#include "block.hpp"
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

__device__ __forceinline__ uint32_t root_leaf(const RootPiece& P, uint64_t key, int s, uint32_t r) {
  const float x = (float)(key - ((uint64_t)r << s));
  const float f = floorf(__fmaf_rn(P.A, x, P.B));
  const float span = (float)(P.hi - P.lo);
  const float c = f < 0.f ? 0.f : (f > span ? span : f);
  return P.lo + (uint32_t)c;
}
__device__ __forceinline__ int64_t leaf_pred(const LeafRec& R, uint64_t key) {
  const float x = (float)(int64_t)(key - R.base);
  float p = floorf(__fmaf_rn(R.S, x, R.I) + 0.5f);
  p = p < -1.0e12f ? -1.0e12f : (p > 1.0e12f ? 1.0e12f : p);
  return (int64_t)R.q + (int64_t)p;
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

// p[r] = first rank with key >= r << s, r in [0, R]; lo / hi leaf ranges.
__global__ void __launch_bounds__(256) k_lidx_pieces(const uint64_t* __restrict__ keys, uint64_t D, int s, uint32_t R, uint32_t L,
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
__global__ void __launch_bounds__(256) k_lidx_root_ranges(const uint64_t* __restrict__ p, uint32_t R, uint64_t D, uint32_t L,
                                                          uint32_t* __restrict__ lo) {
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  lo[r] = (uint32_t)(D ? (p[r] * (uint64_t)L / D) : 0);
  if (lo[r] >= L) lo[r] = L - 1;
}

// Root records in lookup form: A = w isd, B = c - w mu isd (A >= 0), the leaf
// range [lo_r, hi_r].
__global__ void __launch_bounds__(256) k_lidx_root_final(uint32_t R, uint64_t D, uint32_t L, const uint64_t* __restrict__ p,
                                                         const uint32_t* __restrict__ lo, const float* __restrict__ mu,
                                                         const float* __restrict__ isd, const float* __restrict__ w,
                                                         const float* __restrict__ cc, RootPiece* __restrict__ out) {
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  double A = (double)w[r] * (double)isd[r];
  double B = (double)cc[r] - A * (double)mu[r];
  if (!(A >= 0)) A = 0;  // monotone routing (also catches NaN)
  if (!(B == B)) B = 0;
  RootPiece P;
  P.A = (float)A;
  P.B = (float)B;
  P.lo = lo[r];
  uint64_t h = D ? (p[r + 1] * (uint64_t)L + D - 1) / D : 0;  // ceil(p_{r+1} L / D)
  uint32_t hi = h > 0 ? (uint32_t)(h - 1) : 0;
  if (hi < P.lo) hi = P.lo;
  if (hi > L - 1) hi = L - 1;
  P.hi = hi;
  out[r] = P;
}

// Leaf boundaries from the root routing of every key (non-decreasing in the
// rank): q[j] = first rank routed to a leaf >= j, q[L] = D.
__global__ void __launch_bounds__(256) k_lidx_leaf_bounds(const uint64_t* __restrict__ keys, uint64_t D, int s, uint32_t L,
                                                          const RootPiece* __restrict__ root, uint64_t* __restrict__ q) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i > D) return;
  int64_t prev = -1, cur = (int64_t)L;
  if (i > 0) {
    const uint64_t k = keys[i - 1];
    const uint32_t r = (uint32_t)(k >> s);
    prev = root_leaf(root[r], k, s, r);
  }
  if (i < D) {
    const uint64_t k = keys[i];
    const uint32_t r = (uint32_t)(k >> s);
    cur = root_leaf(root[r], k, s, r);
  }
  for (int64_t j = prev + 1; j <= cur; j++) q[j] = i;
}

// Leaf records: S = w isd, I = c - w mu isd; fingerprint shift from the span;
// error bounds over the leaf's keys (one wave per leaf); flags.
__global__ void __launch_bounds__(256) k_lidx_leaf_final(const uint64_t* __restrict__ keys, uint64_t D, uint32_t L,
                                                         const uint64_t* __restrict__ q, const float* __restrict__ mu,
                                                         const float* __restrict__ isd, const float* __restrict__ w,
                                                         const float* __restrict__ cc, uint32_t max_window,
                                                         LeafRec* __restrict__ out, uint64_t* __restrict__ sabs) {
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = __lane_id();
  if (j >= L) return;
  const uint64_t a = q[j], b = q[j + 1];
  LeafRec R;
  R.base = a < D ? keys[a] : (D ? keys[D - 1] : 0);
  R.q = (uint32_t)a;
  R.n = (uint32_t)(b - a);
  double S = (double)w[j] * (double)isd[j];
  double I = (double)cc[j] - S * (double)mu[j];
  if (!(S == S) || !(I == I)) S = I = 0;
  R.S = (float)S;
  R.I = (float)I;
  int sh = 0;
  if (b > a + 1) {
    const uint64_t span = keys[b - 1] - keys[a];
    const int bits = 64 - __clzll(span);
    sh = bits > 16 ? bits - 16 : 0;
  }
  R.shift = (uint8_t)sh;
  R.pad = 0;
  int64_t elo = 0, ehi = 0;
  uint64_t sa = 0;
  for (uint64_t i = a + lane; i < b; i += 64) {
    const int64_t e = (int64_t)i - leaf_pred(R, keys[i]);
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
  const bool wide = elo > 65535 || ehi > 65535 || (uint64_t)(elo + ehi + 1) > max_window;
  R.elo = (uint16_t)(elo > 65535 ? 65535 : elo);
  R.ehi = (uint16_t)(ehi > 65535 ? 65535 : ehi);
  R.flags = (b > a && wide) ? 1 : 0;
  if (lane == 0) {
    out[j] = R;
    sabs[j] = sa;
  }
}

// Fingerprints (key - base) >> shift, one wave per leaf; ties[j] = keys of
// leaf j whose fingerprint equals a neighbour's.
__device__ __forceinline__ uint32_t fp_of(const LeafRec& R, uint64_t key) { return (uint32_t)((key - R.base) >> R.shift) & 0xFFFFu; }
__global__ void __launch_bounds__(256) k_lidx_fingerprints(const uint64_t* __restrict__ keys, uint32_t L,
                                                           const uint64_t* __restrict__ q, const LeafRec* __restrict__ leaf,
                                                           uint16_t* __restrict__ fp, uint32_t* __restrict__ ties) {
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= L) return;
  const LeafRec R = leaf[j];
  const uint64_t a = q[j], b = q[j + 1];
  uint32_t t = 0;
  for (uint64_t i = a + __lane_id(); i < b; i += 64) {
    const uint32_t f = fp_of(R, keys[i]);
    bool tie = false;
    if (i > a) tie |= fp_of(R, keys[i - 1]) == f;
    if (i + 1 < b) tie |= fp_of(R, keys[i + 1]) == f;
    fp[i] = (uint16_t)f;
    t += tie ? 1u : 0u;
  }
  for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
  if (__lane_id() == 0) ties[j] = t;
}

// Fallback leaves: bbh2rank[BBHash value of key i] = i, one wave per leaf.
__global__ void __launch_bounds__(256) k_lidx_bbh2rank(const uint64_t* __restrict__ keys, uint32_t L, const uint64_t* __restrict__ q,
                                                       const LeafRec* __restrict__ leaf, const MphfView mv,
                                                       uint32_t* __restrict__ bbh2rank) {
  const uint64_t j = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= L || !(leaf[j].flags & 1)) return;
  const uint64_t D = q[L];
  for (uint64_t i = q[j] + __lane_id(); i < q[j + 1]; i += 64) {
    const uint64_t v = mphf_lookup<uint64_t>(mv, keys[i]);
    if (v < D) bbh2rank[v] = (uint32_t)i;
  }
}

// ------------------------------------------------------------------ lookup
// Exact rank of a present key (absent keys give D, "not found"; callers map
// it to 0 like mphf.rs:269). One lane per key: the walk reads the 16-byte
// chunk of 8 fingerprints holding the prediction and steps one chunk left
// or right while every fingerprint in it is larger / smaller than the key's
// (adjacent chunks usually share the cache line).
__device__ __forceinline__ uint64_t lidx_rank(const LidxView& V, const MphfView& mv, uint64_t key) {
  if ((key >> V.s) >= V.R) return V.D;  // wider than a 2k-bit key
  const uint32_t r = (uint32_t)(key >> V.s);
  const RootPiece P = V.root[r];
  const uint32_t j = root_leaf(P, key, V.s, r);
  const LeafRec R = V.leaf[j];
  if (R.n == 0) return V.D;
  if (R.flags & 1) {
    const uint64_t v = mphf_lookup<uint64_t>(mv, key);
    const uint64_t rk = v < V.D ? V.bbh2rank[v] : V.D;
    return (rk < V.D && V.keys[rk] == key) ? rk : V.D;
  }
  const int64_t qa = (int64_t)R.q, qb = qa + (int64_t)R.n;
  const int64_t pr = leaf_pred(R, key);
  int64_t lo = pr - (int64_t)R.elo, hi = pr + (int64_t)R.ehi;
  if (lo < qa) lo = qa;
  if (hi > qb - 1) hi = qb - 1;
  if (lo > hi) return V.D;
  const uint32_t f = fp_of(R, key);
  int64_t c = (pr < lo ? lo : (pr > hi ? hi : pr)) & ~7LL;
  bool resolve = false;
  int64_t found = -1;
  for (int it = 0; it < 16384; it++) {
    const uint4 v = *(const uint4*)(V.fp + c);
    const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
    uint32_t n_eq = 0, lt = 0, gt = 0;
    int64_t first_eq = -1;
#pragma unroll
    for (int u = 7; u >= 0; u--) {
      const int64_t pos = c + u;
      const uint32_t e = (wd[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
      const bool in = pos >= lo && pos <= hi;
      const bool eq = in && e == f;
      n_eq += eq ? 1u : 0u;
      first_eq = eq ? pos : first_eq;
      lt |= (in && e < f) ? 1u : 0u;
      gt |= (in && e > f) ? 1u : 0u;
    }
    if (n_eq == 1) {
      // a neighbour just outside the chunk could share the fingerprint
      bool tie = false;
      if (first_eq == c && c > lo) tie = V.fp[c - 1] == f;
      if (first_eq == c + 7 && c + 7 < hi) tie |= V.fp[c + 8] == f;
      if (tie) resolve = true;
      else found = first_eq;
      break;
    }
    if (n_eq > 1) {
      resolve = true;
      break;
    }
    if (!gt && c + 8 <= hi) {  // every fingerprint here is smaller: go right
      c += 8;
      continue;
    }
    if (!lt && c > lo) {  // every fingerprint here is larger: go left
      c -= 8;
      continue;
    }
    break;  // between two fingerprints: absent
  }
  if (resolve) {  // exact search over the window's full keys
    int64_t a = lo, b = hi + 1;
    while (a < b) {
      const int64_t mid = (a + b) >> 1;
      if (V.keys[mid] < key) a = mid + 1; else b = mid;
    }
    return (a <= hi && V.keys[a] == key) ? (uint64_t)a : V.D;
  }
  return found >= 0 ? (uint64_t)found : V.D;
}

__global__ void __launch_bounds__(256) k_lidx_lookup(const uint64_t* __restrict__ keys, uint64_t n, const LidxView V, const MphfView mv,
                                                     uint64_t* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t r = lidx_rank(V, mv, keys[i]);
    out[i] = r < V.D ? r : 0;
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
// Every distinct k-mer (P3 keys with their joined masks, both in MPHF value
// order) is looked up through the index and its 5-bit state stored as a byte
// at that slot; k_states_pack packs the bytes 12 per word.
template <bool LEARNED>
__global__ void __launch_bounds__(256) k_classify_via(const uint64_t* __restrict__ keys, const uint16_t* __restrict__ vmask,
                                                      const LidxView V, const MphfView mv, uint8_t* __restrict__ st8,
                                                      unsigned int* __restrict__ err) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V.D; v += stride) {
    const uint64_t k = keys[v];
    uint64_t idx;
    if (LEARNED) idx = lidx_rank(V, mv, k);
    else idx = mphf_lookup<uint64_t>(mv, k);
    if (idx >= V.D) {
      atomicOr(err, 1u);
      continue;
    }
    st8[idx] = (uint8_t)mask_state(vmask[v]);
  }
}
__global__ void __launch_bounds__(256) k_states_pack(const uint8_t* __restrict__ st8, uint64_t D, uint64_t n_words,
                                                     uint64_t* __restrict__ states) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  uint64_t out = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const uint64_t x = 12 * w + (uint64_t)i;
    if (x < D) out |= (uint64_t)(st8[x] & 31u) << (5 * i);
  }
  states[w] = out;
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
  CF1G_LAUNCH("lidx_pieces", k_lidx_pieces, dim3(blocks_for((uint64_t)R + 1, 256)), dim3(256), 0, st, keys, D, s, R, L, p);
  CF1G_LAUNCH("lidx_root_ranges", k_lidx_root_ranges, dim3(blocks_for(R, 256)), dim3(256), 0, st, p, R, D, L, lo);
}
void launch_lidx_root_final(hipStream_t st, uint32_t R, uint64_t D, uint32_t L, const uint64_t* p, const uint32_t* lo,
                            const LidxTrainBufs& b, RootPiece* out) {
  CF1G_LAUNCH("lidx_root_final", k_lidx_root_final, dim3(blocks_for(R, 256)), dim3(256), 0, st, R, D, L, p, lo, b.mu, b.isd, b.w, b.c,
              out);
}
void launch_lidx_leaf_bounds(hipStream_t st, const uint64_t* keys, uint64_t D, int s, uint32_t L, const RootPiece* root, uint64_t* q) {
  CF1G_LAUNCH("lidx_leaf_bounds", k_lidx_leaf_bounds, dim3(blocks_for(D + 1, 256)), dim3(256), 0, st, keys, D, s, L, root, q);
}
void launch_lidx_leaf_final(hipStream_t st, const uint64_t* keys, uint64_t D, uint32_t L, const uint64_t* q, const LidxTrainBufs& b,
                            uint32_t max_window, LeafRec* out, uint64_t* sabs) {
  CF1G_LAUNCH("lidx_leaf_final", k_lidx_leaf_final, dim3(blocks_for(L, 4)), dim3(256), 0, st, keys, D, L, q, b.mu, b.isd, b.w, b.c,
              max_window, out, sabs);
}
void launch_lidx_fingerprints(hipStream_t st, const uint64_t* keys, uint32_t L, const uint64_t* q, const LeafRec* leaf, uint16_t* fp,
                              uint32_t* ties) {
  CF1G_LAUNCH("lidx_fingerprints", k_lidx_fingerprints, dim3(blocks_for(L, 4)), dim3(256), 0, st, keys, L, q, leaf, fp, ties);
}
void launch_lidx_bbh2rank(hipStream_t st, const uint64_t* keys, uint32_t L, const uint64_t* q, const LeafRec* leaf, const MphfView& mv,
                          uint32_t* bbh2rank) {
  CF1G_LAUNCH("lidx_bbh2rank", k_lidx_bbh2rank, dim3(blocks_for(L, 4)), dim3(256), 0, st, keys, L, q, leaf, mv, bbh2rank);
}
void launch_lidx_lookup(hipStream_t st, const uint64_t* keys, uint64_t n, const LidxView& V, const MphfView& mv, uint64_t* out) {
  if (!n) return;
  uint64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  CF1G_LAUNCH("lidx_lookup", k_lidx_lookup, dim3((unsigned)g), dim3(256), 0, st, keys, n, V, mv, out);
}
void launch_count_mismatch(hipStream_t st, const uint64_t* a, const uint64_t* b, uint64_t n, unsigned long long* bad) {
  if (!n) return;
  uint64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  CF1G_LAUNCH("count_mismatch", k_count_mismatch, dim3((unsigned)g), dim3(256), 0, st, a, b, n, bad);
}
void launch_classify_via(hipStream_t st, bool learned, const uint64_t* keys, const uint16_t* vmask, const LidxView& V,
                         const MphfView& mv, uint8_t* st8, unsigned int* err) {
  if (!V.D) return;
  uint64_t g = (V.D + 255) / 256;
  if (g > 65536) g = 65536;
  if (learned)
    CF1G_LAUNCH("p4_classify_lidx", k_classify_via<true>, dim3((unsigned)g), dim3(256), 0, st, keys, vmask, V, mv, st8, err);
  else
    CF1G_LAUNCH("p4_classify_bbh", k_classify_via<false>, dim3((unsigned)g), dim3(256), 0, st, keys, vmask, V, mv, st8, err);
}
void launch_states_pack(hipStream_t st, const uint8_t* st8, uint64_t D, uint64_t* states) {
  const uint64_t nw = (D + 11) / 12;
  if (!nw) return;
  CF1G_LAUNCH("p4_states_pack", k_states_pack, dim3(blocks_for(nw, 256)), dim3(256), 0, st, st8, D, nw, states);
}
void launch_annotate_lidx(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, const LidxView& V,
                          const MphfView& mv, const uint64_t* states, uint8_t* out) {
  if (!n_pos) return;
  CF1G_LAUNCH("p5_annotate_lidx", k_annotate_lidx, dim3(blocks_for(n_pos, 256)), dim3(256), 0, st, seq, nmask, n_pos, K, V, mv, states,
              out);
}

}  // namespace cf1g
