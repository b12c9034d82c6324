// P3 kernels: per-bin canonical k-mer extraction, sort and dedup
// (mphf.rs:323-363 dedup_bin) and the BBHash build (the project's
// replacement for ph::fmph, mphf.rs:144-190).
//
// P3a layout: the k-mers of bin b are split by the top `sb` bits of the
// canonical key into 2^sb segments (b, s). Segments are counted (k_p3_count),
// scattered into one occurrence array (k_p3_scatter), then each segment is
// sorted and deduplicated by one work-group (LDS radix sort, or an LDS
// bitmap when only <= 18 key bits vary); segments too large for LDS go to a
// global-memory radix sort (k_p3_bigsort). Uniques are compacted in
// (bin, segment) order, so every bin's keys come out sorted, bins
// concatenated: the order of the reference's cf1_dedup_all.bin.
#include "device.hpp"
#include "launch.hpp"

namespace cf1g {

// ------------------------------------------------------------------ scan
constexpr int SC_WG = 256;
constexpr int SC_IPT = 16;
constexpr int SC_TILE = SC_WG * SC_IPT;

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
  uint32_t lane = __lane_id();
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if ((int)lane >= d) x += y;
  }
  return x;
}
// Exclusive scan across a 256-thread block; *total gets the block sum.
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* sh4, uint64_t* total) {
  uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  uint64_t x = wave_incl_scan_u64(v);
  if (lane == 63) sh4[wv] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int q = 0; q < 4; q++) {
    uint64_t t = sh4[q];
    if (q < (int)wv) off += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return off + x - v;
}

__global__ void __launch_bounds__(SC_WG) k_scan_tile(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n,
                                                     uint64_t* __restrict__ block_sums) {
  __shared__ uint64_t sh4[4];
  uint64_t base = (uint64_t)blockIdx.x * SC_TILE + (uint64_t)threadIdx.x * SC_IPT;
  uint64_t v[SC_IPT];
  uint64_t s = 0;
#pragma unroll
  for (int q = 0; q < SC_IPT; q++) {
    v[q] = (base + q < n) ? in[base + q] : 0;
    s += v[q];
  }
  uint64_t tot;
  uint64_t off = block_excl_scan_u64(s, sh4, &tot);
#pragma unroll
  for (int q = 0; q < SC_IPT; q++) {
    if (base + q < n) out[base + q] = off;
    off += v[q];
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(SC_WG) k_scan_add(uint64_t* __restrict__ out, uint64_t n, const uint64_t* __restrict__ offs,
                                                    uint64_t* __restrict__ total_out, const uint64_t* __restrict__ total_src) {
  uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
  uint64_t add = offs[blockIdx.x];
  for (int q = threadIdx.x; q < SC_TILE; q += SC_WG)
    if (base + q < n) out[base + q] += add;
  if (total_out && blockIdx.x == 0 && threadIdx.x == 0) *total_out = *total_src;
}

uint64_t scan_temp_words(uint64_t n) {
  uint64_t t = 0;
  while (n > 1) {
    n = (n + SC_TILE - 1) / SC_TILE;
    t += n + 1;
  }
  return t + 2;
}
// Exclusive scan of n u64 values (in may equal out). If total != nullptr the
// grand total is written to *total (device). tmp: scan_temp_words(n) words.
void scan_excl_u64(hipStream_t st, const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total, uint64_t* tmp) {
  if (n == 0) {
    if (total) hip_check(hipMemsetAsync(total, 0, 8, st), "memset");
    return;
  }
  uint64_t nb = (n + SC_TILE - 1) / SC_TILE;
  uint64_t* sums = tmp;
  CF1G_LAUNCH("scan_tile", k_scan_tile, dim3((unsigned)nb), dim3(SC_WG), 0, st, in, out, n, sums);
  if (nb == 1) {
    if (total) hip_check(hipMemcpyAsync(total, sums, 8, hipMemcpyDeviceToDevice, st), "copy");
    return;
  }
  uint64_t* sub_total = sums + nb;  // one word
  scan_excl_u64(st, sums, sums, nb, sub_total, sub_total + 1);
  CF1G_LAUNCH("scan_add", k_scan_add, dim3((unsigned)nb), dim3(SC_WG), 0, st, out, n, sums, total, sub_total);
}

// ------------------------------------------------------------------ record k-mers
// Records hold up to 60 (RW=2) or 124 (RW=4) bases MSB-first; base count in
// the low byte of the last word. k-mers are rolled forward with their
// reverse complements (kmer.rs:240-262).
template <int RW>
__device__ __forceinline__ uint32_t rec_base(const uint64_t (&w)[RW], int t) {
  return (uint32_t)((w[t >> 5] >> (62 - 2 * (t & 31))) & 3);
}

template <typename KeyT>
struct Roll;
template <>
struct Roll<uint64_t> {
  uint64_t x, r, mask;
  int K;
  __device__ void init(int K_) {
    K = K_;
    mask = (K == 32) ? ~0ULL : ((1ULL << (2 * K)) - 1);
    x = 0;
    r = 0;
  }
  __device__ void push(uint32_t b) {
    x = ((x << 2) | b) & mask;
    r = (r >> 2) | ((uint64_t)(3 - b) << (2 * K - 2));
  }
  __device__ uint64_t canon() const { return x < r ? x : r; }
};
template <>
struct Roll<Key128> {
  Key128 x, r, mask;
  int K;
  __device__ void init(int K_) {
    K = K_;
    mask.lo = ~0ULL;
    mask.hi = (1ULL << (2 * K - 64)) - 1;
    x = {0, 0};
    r = {0, 0};
  }
  __device__ void push(uint32_t b) {
    x.hi = ((x.hi << 2) | (x.lo >> 62)) & mask.hi;
    x.lo = (x.lo << 2) | b;
    // r >>= 2, then complement of b at bit 2K-2 (in the high word)
    r.lo = (r.lo >> 2) | (r.hi << 62);
    r.hi = (r.hi >> 2) | ((uint64_t)(3 - b) << (2 * K - 66));
  }
  __device__ Key128 canon() const { return key_lt(r, x) ? r : x; }
};

// Iterate the canonical k-mers of one record.
template <typename KeyT, int RW, typename F>
__device__ __forceinline__ void for_each_rec_kmer(const uint64_t* __restrict__ rec, uint64_t idx, int K, F f) {
  uint64_t w[RW];
#pragma unroll
  for (int t = 0; t < RW; t++) w[t] = rec[idx * RW + t];
  int L = (int)(w[RW - 1] & 255);
  if (L > 32 * RW - 4) L = 32 * RW - 4;  // record capacity (bases above the count byte)
  Roll<KeyT> ro;
  ro.init(K);
  for (int t = 0; t < K - 1; t++) ro.push(rec_base<RW>(w, t));
  for (int t = K - 1; t < L; t++) {
    ro.push(rec_base<RW>(w, t));
    f(ro.canon());
  }
}

constexpr int P3_WG = 256;
constexpr int P3_MAX_SB = 12;

// Work-group -> (bin, record range) for the per-bin passes.
__device__ __forceinline__ bool p3_chunk(const uint64_t* __restrict__ chunk_off, const uint64_t* __restrict__ bin_start,
                                         int num_bins, uint32_t chunk_records, int& bin, uint64_t& r0, uint64_t& r1) {
  uint64_t wg = blockIdx.x;
  if (wg >= chunk_off[num_bins]) return false;
  int lo = 0, hi = num_bins;  // largest b with chunk_off[b] <= wg
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (chunk_off[mid] <= wg) lo = mid; else hi = mid;
  }
  // skip empty bins sharing the same chunk offset
  while (lo + 1 < num_bins && chunk_off[lo + 1] <= wg) lo++;
  bin = lo;
  uint64_t c = wg - chunk_off[lo];
  r0 = bin_start[lo] + c * chunk_records;
  r1 = r0 + chunk_records;
  if (r1 > bin_start[lo + 1]) r1 = bin_start[lo + 1];
  return true;
}

template <typename KeyT, int RW>
__global__ void __launch_bounds__(P3_WG) k_p3_count(const uint64_t* __restrict__ rec, const uint64_t* __restrict__ bin_start,
                                                    const uint64_t* __restrict__ chunk_off, int num_bins, uint32_t chunk_records,
                                                    int K, int sb, uint64_t* __restrict__ counts) {
  __shared__ uint32_t hist[1 << P3_MAX_SB];
  int bin;
  uint64_t r0, r1;
  if (!p3_chunk(chunk_off, bin_start, num_bins, chunk_records, bin, r0, r1)) return;
  const int ns = 1 << sb;
  for (int s = threadIdx.x; s < ns; s += P3_WG) hist[s] = 0;
  __syncthreads();
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += P3_WG)
    for_each_rec_kmer<KeyT, RW>(rec, r, K, [&](const KeyT& c) { atomicAdd(&hist[key_top(c, 2 * K, sb)], 1u); });
  __syncthreads();
  uint64_t* dst = counts + (uint64_t)bin * ns;
  for (int s = threadIdx.x; s < ns; s += P3_WG)
    if (hist[s]) atomicAdd((unsigned long long*)&dst[s], (unsigned long long)hist[s]);
}

template <typename KeyT, int RW>
__global__ void __launch_bounds__(P3_WG) k_p3_scatter(const uint64_t* __restrict__ rec, const uint64_t* __restrict__ bin_start,
                                                      const uint64_t* __restrict__ chunk_off, int num_bins, uint32_t chunk_records,
                                                      int K, int sb, uint64_t* __restrict__ cursor, KeyT* __restrict__ occ) {
  __shared__ uint32_t hist[1 << P3_MAX_SB];
  __shared__ unsigned long long gb[1 << P3_MAX_SB];
  int bin;
  uint64_t r0, r1;
  if (!p3_chunk(chunk_off, bin_start, num_bins, chunk_records, bin, r0, r1)) return;
  const int ns = 1 << sb;
  for (int s = threadIdx.x; s < ns; s += P3_WG) hist[s] = 0;
  __syncthreads();
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += P3_WG)
    for_each_rec_kmer<KeyT, RW>(rec, r, K, [&](const KeyT& c) { atomicAdd(&hist[key_top(c, 2 * K, sb)], 1u); });
  __syncthreads();
  uint64_t* cur = cursor + (uint64_t)bin * ns;
  for (int s = threadIdx.x; s < ns; s += P3_WG)
    if (hist[s]) gb[s] = atomicAdd((unsigned long long*)&cur[s], (unsigned long long)hist[s]);
  __syncthreads();
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += P3_WG)
    for_each_rec_kmer<KeyT, RW>(rec, r, K, [&](const KeyT& c) {
      uint32_t s = key_top(c, 2 * K, sb);
      uint32_t slot = atomicSub(&hist[s], 1u) - 1u;
      occ[gb[s] + slot] = c;
    });
}

// ------------------------------------------------------------------ tile ranking
// A tile of up to 256*KPT keys lives in registers, wave-blocked: element
// e = wave*(64*kpt) + slot*64 + lane. rank_tile computes, for one 8-bit
// digit, each element's rank among equal digits of its wave (stable, in
// element order) and the per-wave digit counts in cnt[wave*256 + d].
constexpr int RK_WG = 256;

template <typename KeyT, int KPT>
__device__ __forceinline__ void rank_tile(const KeyT (&key)[KPT], int kpt, int sh, uint32_t (&rank)[KPT], uint32_t* cnt) {
  const uint32_t wv = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  uint32_t* wc = cnt + wv * 256;
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    if (s < kpt) {
      uint32_t d = key_digit(key[s], sh);
      uint64_t peers = ~0ULL;
#pragma unroll
      for (int b = 0; b < 8; b++) {
        uint64_t bal = __ballot((d >> b) & 1);
        peers &= ((d >> b) & 1) ? bal : ~bal;
      }
      uint32_t below = (uint32_t)__popcll(peers & lt);
      uint32_t base = wc[d];
      rank[s] = base + below;
      __builtin_amdgcn_wave_barrier();
      if (below == 0) wc[d] = base + (uint32_t)__popcll(peers);
      __builtin_amdgcn_wave_barrier();
    }
  }
}
// Turn per-wave counts into scatter bases: cnt[w*256+d] = start of (w, d).
// Optional: tile totals per digit into tot[d] (when tot != nullptr).
__device__ __forceinline__ void rank_offsets(uint32_t* cnt, uint64_t* sh4, uint32_t* tot) {
  uint32_t d = threadIdx.x;  // 256 threads == 256 digits
  uint32_t c0 = cnt[d], c1 = cnt[256 + d], c2 = cnt[512 + d], c3 = cnt[768 + d];
  uint32_t t = c0 + c1 + c2 + c3;
  uint64_t ex = block_excl_scan_u64(t, sh4, nullptr);
  cnt[d] = (uint32_t)ex;
  cnt[256 + d] = (uint32_t)ex + c0;
  cnt[512 + d] = (uint32_t)ex + c0 + c1;
  cnt[768 + d] = (uint32_t)ex + c0 + c1 + c2;
  if (tot) tot[d] = t;
}

template <typename KeyT>
__device__ __forceinline__ KeyT key_pad();
template <>
__device__ __forceinline__ uint64_t key_pad<uint64_t>() { return ~0ULL; }
template <>
__device__ __forceinline__ Key128 key_pad<Key128>() { return {~0ULL, ~0ULL}; }

__device__ __forceinline__ void key_xor_or(uint64_t& acc_lo, uint64_t& acc_hi, uint64_t k, uint64_t k0) { acc_lo |= k ^ k0; }
__device__ __forceinline__ void key_xor_or(uint64_t& acc_lo, uint64_t& acc_hi, const Key128& k, const Key128& k0) {
  acc_lo |= k.lo ^ k0.lo;
  acc_hi |= k.hi ^ k0.hi;
}
__device__ __forceinline__ int varying_bits(uint64_t lo, uint64_t hi) {
  if (hi) return 128 - __clzll(hi);
  if (lo) return 64 - __clzll(lo);
  return 0;
}
// Block OR-reduce of two words.
__device__ __forceinline__ void block_or2(uint64_t& a, uint64_t& b, uint64_t* sh8) {
  for (int d = 32; d >= 1; d >>= 1) {
    a |= __shfl_xor(a, d, 64);
    b |= __shfl_xor(b, d, 64);
  }
  uint32_t wv = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    sh8[wv] = a;
    sh8[4 + wv] = b;
  }
  __syncthreads();
  a = sh8[0] | sh8[1] | sh8[2] | sh8[3];
  b = sh8[4] | sh8[5] | sh8[6] | sh8[7];
  __syncthreads();
}

// ------------------------------------------------------------------ segment sort
template <typename KeyT>
struct SegCfg;
template <>
struct SegCfg<uint64_t> {
  static constexpr int KPT = 32;
};
template <>
struct SegCfg<Key128> {
  static constexpr int KPT = 16;
};
constexpr int BM_MAX_BITS = 18;  // LDS bitmap dedup when <= 18 low key bits vary

template <typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_p3_segsort(KeyT* __restrict__ occ, const uint64_t* __restrict__ seg_off, uint64_t n_segs,
                                                      int K2, int sb, uint64_t* __restrict__ ucount,
                                                      uint32_t* __restrict__ over_list, unsigned int* __restrict__ n_over) {
  constexpr int KPT = SegCfg<KeyT>::KPT;
  constexpr int CAP = RK_WG * KPT;
  __shared__ KeyT buf[CAP];
  __shared__ uint32_t cnt[4 * 256];
  __shared__ uint64_t sh8[8];
  const uint64_t seg = blockIdx.x;
  if (seg >= n_segs) return;
  const uint64_t start = seg_off[seg];
  const uint64_t n = seg_off[seg + 1] - start;
  const int rbits = K2 - sb;  // low bits that may vary in this segment
  if (n == 0) {
    if (threadIdx.x == 0) ucount[seg] = 0;
    return;
  }
  if (sizeof(KeyT) == 8 && rbits <= BM_MAX_BITS) {
    // ---- bitmap dedup: any number of duplicates, sorted output.
    uint64_t* bm = (uint64_t*)buf;  // 2^18 bits = 32 KB <= sizeof(buf)
    const uint32_t nwords = rbits <= 6 ? 1u : (1u << (rbits - 6));
    for (uint32_t i = threadIdx.x; i < nwords; i += RK_WG) bm[i] = 0;
    __syncthreads();
    const uint64_t lowmask = (rbits >= 64) ? ~0ULL : ((1ULL << rbits) - 1);
    for (uint64_t i = threadIdx.x; i < n; i += RK_WG) {
      uint64_t low = key_lo(occ[start + i]) & lowmask;
      atomicOr((unsigned long long*)&bm[low >> 6], 1ULL << (low & 63));
    }
    __syncthreads();
    const uint32_t wpt = (nwords + RK_WG - 1) / RK_WG;
    const uint32_t w0 = threadIdx.x * wpt;
    uint64_t c = 0;
    for (uint32_t q = 0; q < wpt; q++)
      if (w0 + q < nwords) c += (uint64_t)__popcll(bm[w0 + q]);
    uint64_t tot;
    uint64_t o = block_excl_scan_u64(c, sh8, &tot);
    const uint64_t prefix = (uint64_t)(seg & ((1ULL << sb) - 1)) << rbits;
    for (uint32_t q = 0; q < wpt; q++) {
      if (w0 + q >= nwords) break;
      uint64_t x = bm[w0 + q];
      while (x) {
        int b = __ffsll((long long)x) - 1;
        x &= x - 1;
        uint64_t kv = prefix | ((uint64_t)(w0 + q) * 64 + (uint64_t)b);
        KeyT kk;
        if constexpr (sizeof(KeyT) == 8) kk = (KeyT)kv;
        occ[start + o++] = kk;
      }
    }
    if (threadIdx.x == 0) ucount[seg] = tot;
    return;
  }
  if (n > (uint64_t)CAP) {
    if (threadIdx.x == 0) {
      unsigned int idx = atomicAdd(n_over, 1u);
      over_list[idx] = (uint32_t)seg;
      ucount[seg] = 0;
    }
    return;
  }
  // ---- LDS LSD radix sort of <= CAP keys.
  const int kpt = (int)((n + RK_WG - 1) / RK_WG);
  const uint32_t wv = threadIdx.x >> 6, lane = __lane_id();
  KeyT key[KPT];
  uint32_t rank[KPT];
  const KeyT k0 = occ[start];
  uint64_t xl = 0, xh = 0;
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    if (s < kpt) {
      uint64_t e = (uint64_t)wv * 64 * kpt + (uint64_t)s * 64 + lane;
      if (e < n) {
        key[s] = occ[start + e];
        key_xor_or(xl, xh, key[s], k0);
      } else {
        key[s] = key_pad<KeyT>();
      }
    }
  }
  block_or2(xl, xh, sh8);
  const int vb = varying_bits(xl, xh);
  if (vb == 0) {
    if (threadIdx.x == 0) {
      occ[start] = k0;
      ucount[seg] = 1;
    }
    return;
  }
  const int passes = (vb + 7) >> 3;
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
    for (int i = threadIdx.x; i < 4 * 256; i += RK_WG) cnt[i] = 0;
    __syncthreads();
    rank_tile<KeyT, KPT>(key, kpt, sh, rank, cnt);
    __syncthreads();
    rank_offsets(cnt, sh8, nullptr);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < KPT; s++)
      if (s < kpt) buf[cnt[wv * 256 + key_digit(key[s], sh)] + rank[s]] = key[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < KPT; s++)
      if (s < kpt) key[s] = buf[(uint64_t)wv * 64 * kpt + (uint64_t)s * 64 + lane];
    __syncthreads();
  }
  // buf[0..n) is sorted; dedup in thread-contiguous ranges.
  const uint64_t e0 = (uint64_t)threadIdx.x * kpt;
  uint64_t c = 0;
  for (int q = 0; q < kpt; q++) {
    uint64_t e = e0 + q;
    if (e < n && (e == 0 || !key_eq(buf[e], buf[e - 1]))) c++;
  }
  uint64_t tot;
  uint64_t o = block_excl_scan_u64(c, sh8, &tot);
  for (int q = 0; q < kpt; q++) {
    uint64_t e = e0 + q;
    if (e < n && (e == 0 || !key_eq(buf[e], buf[e - 1]))) occ[start + o++] = buf[e];
  }
  if (threadIdx.x == 0) ucount[seg] = tot;
}

// Oversized segments: one work-group per segment runs a stable LSD radix
// sort through global memory (ping-pong with `aux`), tile by tile, then
// deduplicates into occ[start...].
template <typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_p3_bigsort(KeyT* __restrict__ occ, KeyT* __restrict__ aux,
                                                      const uint64_t* __restrict__ seg_off, const uint32_t* __restrict__ over_list,
                                                      uint64_t* __restrict__ ucount) {
  constexpr int KPT = SegCfg<KeyT>::KPT;
  constexpr int CAP = RK_WG * KPT;
  __shared__ uint32_t cnt[4 * 256];
  __shared__ unsigned long long dbase[256];
  __shared__ uint32_t ttot[256];
  __shared__ uint64_t sh8[8];
  __shared__ KeyT carry;
  const uint32_t seg = over_list[blockIdx.x];
  const uint64_t start = seg_off[seg];
  const uint64_t n = seg_off[seg + 1] - start;
  const uint32_t wv = threadIdx.x >> 6, lane = __lane_id();
  KeyT* src = occ + start;
  KeyT* dst = aux + start;
  const KeyT k0 = src[0];
  uint64_t xl = 0, xh = 0;
  for (uint64_t i = threadIdx.x; i < n; i += RK_WG) key_xor_or(xl, xh, src[i], k0);
  block_or2(xl, xh, sh8);
  const int vb = varying_bits(xl, xh);
  const int passes = (vb + 7) >> 3;
  KeyT key[KPT];
  uint32_t rank[KPT];
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
    // digit histogram of the whole segment
    for (int i = threadIdx.x; i < 256; i += RK_WG) ttot[i] = 0;
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < n; i += RK_WG) atomicAdd(&ttot[key_digit(src[i], sh)], 1u);
    __syncthreads();
    {
      uint64_t ex = block_excl_scan_u64(ttot[threadIdx.x], sh8, nullptr);
      dbase[threadIdx.x] = ex;
    }
    __syncthreads();
    for (uint64_t t0 = 0; t0 < n; t0 += CAP) {
      uint64_t tn = n - t0 < (uint64_t)CAP ? n - t0 : (uint64_t)CAP;
      int kpt = (int)((tn + RK_WG - 1) / RK_WG);
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s < kpt) {
          uint64_t e = (uint64_t)wv * 64 * kpt + (uint64_t)s * 64 + lane;
          key[s] = e < tn ? src[t0 + e] : key_pad<KeyT>();
        }
      }
      for (int i = threadIdx.x; i < 4 * 256; i += RK_WG) cnt[i] = 0;
      __syncthreads();
      rank_tile<KeyT, KPT>(key, kpt, sh, rank, cnt);
      __syncthreads();
      rank_offsets(cnt, sh8, ttot);
      __syncthreads();
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s < kpt) {
          uint64_t e = (uint64_t)wv * 64 * kpt + (uint64_t)s * 64 + lane;
          if (e < tn) {
            uint32_t d = key_digit(key[s], sh);
            uint32_t local = cnt[wv * 256 + d] + rank[s];
            // position inside the tile's digit-d run = local - (start of digit d in tile)
            uint32_t dstart = cnt[d];  // wave 0 base of digit d == tile start of digit d
            dst[dbase[d] + (local - dstart)] = key[s];
          }
        }
      }
      __syncthreads();
      dbase[threadIdx.x] += ttot[threadIdx.x];
      __syncthreads();
    }
    KeyT* t = src;
    src = dst;
    dst = t;
    __syncthreads();
  }
  // dedup sorted src into occ[start...] (src may be occ itself: writes trail reads)
  KeyT* out = occ + start;
  uint64_t u = 0;
  for (uint64_t t0 = 0; t0 < n; t0 += CAP) {
    uint64_t tn = n - t0 < (uint64_t)CAP ? n - t0 : (uint64_t)CAP;
    int kpt = (int)((tn + RK_WG - 1) / RK_WG);
    // thread-contiguous: thread t owns [t*kpt, (t+1)*kpt)
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      uint64_t e = (uint64_t)threadIdx.x * kpt + s;
      if (s < kpt && e < tn) key[s] = src[t0 + e];
    }
    KeyT prev_carry = carry;
    __syncthreads();
    uint64_t c = 0;
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      uint64_t e = (uint64_t)threadIdx.x * kpt + s;
      if (s < kpt && e < tn) {
        bool first;
        if (s > 0) first = !key_eq(key[s], key[s - 1]);
        else if (e == 0) first = (t0 == 0) || !key_eq(key[s], prev_carry);
        else first = !key_eq(key[s], src[t0 + e - 1]);
        rank[s] = first ? 1u : 0u;
        c += first;
      }
    }
    uint64_t tot;
    uint64_t o = block_excl_scan_u64(c, sh8, &tot);
    // all reads of this tile happened before the scan's barriers
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      uint64_t e = (uint64_t)threadIdx.x * kpt + s;
      if (s < kpt && e < tn && rank[s]) out[u + o++] = key[s];
      if (s < kpt && e + 1 == tn) carry = key[s];
    }
    u += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) ucount[seg] = u;
}

template <typename KeyT>
__global__ void __launch_bounds__(256) k_p3_compact(const KeyT* __restrict__ occ, const uint64_t* __restrict__ seg_off,
                                                    const uint64_t* __restrict__ ucount, const uint64_t* __restrict__ out_off,
                                                    uint64_t n_segs, KeyT* __restrict__ keys) {
  uint64_t seg = blockIdx.x;
  if (seg >= n_segs) return;
  uint64_t u = ucount[seg], s = seg_off[seg], o = out_off[seg];
  for (uint64_t i = threadIdx.x; i < u; i += 256) keys[o + i] = occ[s + i];
}

// ------------------------------------------------------------------ BBHash build
// One work-group per P3a segment builds that segment's BBHash (see
// device.hpp) from its sorted distinct keys: keys live in registers, the
// level's seen/collision bit arrays in LDS. MODE 0 only measures the
// segment's block count; after a scan of those counts MODE 1 rebuilds the
// same levels and writes the blocks with their segment-local rank words.
// Segments with more keys than fit in registers (> CAP) are handed to
// k_bbh_big through big_list.
template <int MODE, typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_bbh_seg(const KeyT* __restrict__ keys, const uint64_t* __restrict__ key_off,
                                                   uint64_t n_segs, uint64_t* __restrict__ seg_nblk,
                                                   const uint64_t* __restrict__ blk_off, uint64_t* __restrict__ blocks,
                                                   uint32_t* __restrict__ big_list, unsigned int* __restrict__ n_big,
                                                   unsigned int* __restrict__ err) {
  constexpr int KPT = SegCfg<KeyT>::KPT;
  constexpr int CAP = RK_WG * KPT;
  constexpr int SEEN_W = (2 * CAP) / 64 + 1;
  constexpr int REG_W = (4 * CAP + 8192) / 64;
  __shared__ uint64_t seen[SEEN_W];
  __shared__ uint64_t col[SEEN_W];
  __shared__ uint64_t region[MODE ? REG_W : 1];
  __shared__ uint64_t sh8[8];
  const uint64_t seg = blockIdx.x;
  if (seg >= n_segs) return;
  const uint64_t ko = key_off[seg];
  const uint64_t D = key_off[seg + 1] - ko;
  if (D == 0) {
    if (MODE == 0 && threadIdx.x == 0) seg_nblk[seg] = 0;
    return;
  }
  if (D > (uint64_t)CAP) {
    if (MODE == 0 && threadIdx.x == 0) {
      big_list[atomicAdd(n_big, 1u)] = (uint32_t)seg;
      seg_nblk[seg] = 0;
    }
    return;
  }
  KeyT key[KPT];
  uint32_t alive = 0;
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    uint64_t e = (uint64_t)s * RK_WG + threadIdx.x;
    if (e < D) {
      key[s] = keys[ko + e];
      alive |= 1u << s;
    }
  }
  if (MODE)
    for (int i = threadIdx.x; i < REG_W; i += RK_WG) region[i] = 0;
  uint64_t off = 0, t = D, r = D;
  int l = 0;
  while (r > 0) {
    const uint64_t n = bbh_level_n(t);
    if (l >= BBH_MAX_LEVELS || off + n > (uint64_t)REG_W * 64) {
      if (threadIdx.x == 0) atomicOr(err, 1u);
      break;
    }
    const uint32_t nw = (uint32_t)(n >> 6);
    for (uint32_t i = threadIdx.x; i < nw; i += RK_WG) {
      seen[i] = 0;
      col[i] = 0;
    }
    __syncthreads();
    const uint64_t seed = bbh_seed(l);
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if ((alive >> s) & 1) {
        uint64_t pos = __umul64hi(bbh_hash(key[s], seed), n);
        unsigned long long bit = 1ULL << (pos & 63);
        unsigned long long old = atomicOr((unsigned long long*)&seen[pos >> 6], bit);
        if (old & bit) atomicOr((unsigned long long*)&col[pos >> 6], bit);
      }
    }
    __syncthreads();
    if (MODE)
      for (uint32_t i = threadIdx.x; i < nw; i += RK_WG) region[(off >> 6) + i] = seen[i] & ~col[i];
    uint64_t cnt = 0;
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if ((alive >> s) & 1) {
        uint64_t pos = __umul64hi(bbh_hash(key[s], seed), n);
        if (((seen[pos >> 6] & ~col[pos >> 6]) >> (pos & 63)) & 1) alive &= ~(1u << s);
        else cnt++;
      }
    }
    block_excl_scan_u64(cnt, sh8, &r);  // r = keys still unplaced (barriers inside)
    off += n;
    t = bbh_next_t(t);
    l++;
  }
  const uint64_t nblk = (off + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS;
  if (MODE == 0) {
    if (threadIdx.x == 0) seg_nblk[seg] = nblk;
    return;
  }
  __syncthreads();
  // nblk <= REG_W*64/448 + 1 < 256: one block per thread
  uint64_t w[7];
  uint64_t pc = 0;
  const uint64_t j = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 7; u++) {
    uint64_t q = j * 7 + u;
    w[u] = (j < nblk && q < (uint64_t)REG_W) ? region[q] : 0;
    pc += (uint64_t)__popcll(w[u]);
  }
  uint64_t tot;
  uint64_t rank = block_excl_scan_u64(j < nblk ? pc : 0, sh8, &tot);
  if (j < nblk) {
    uint64_t* B = blocks + (blk_off[seg] + j) * 8;
    B[0] = rank;
#pragma unroll
    for (int u = 0; u < 7; u++) B[1 + u] = w[u];
  }
}

// Segments above the register capacity: the same levels, with seen/collision
// bits in global scratch (device-scope atomics) and per-key alive bits in a
// global bitmap. One work-group per big segment.
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int MODE, typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_bbh_big(const KeyT* __restrict__ keys, const uint64_t* __restrict__ key_off,
                                                   const uint32_t* __restrict__ big_list, uint64_t* __restrict__ seg_nblk,
                                                   const uint64_t* __restrict__ blk_off, uint64_t* __restrict__ blocks,
                                                   uint64_t* __restrict__ scratch, uint64_t* __restrict__ alive_g,
                                                   unsigned int* __restrict__ err) {
  __shared__ uint64_t sh8[8];
  __shared__ unsigned long long carry;
  const uint32_t seg = big_list[blockIdx.x];
  const uint64_t ko = key_off[seg];
  const uint64_t D = key_off[seg + 1] - ko;
  uint64_t* seen = scratch + 4 * (ko / 64) + 8 * (uint64_t)seg;
  uint64_t* col = seen + (D / 32 + 2);
  uint64_t* region = blocks + (MODE ? blk_off[seg] * 8 : 0);
  if (MODE) {
    uint64_t nwords = (blk_off[seg + 1] - blk_off[seg]) * 8;
    for (uint64_t i = threadIdx.x; i < nwords; i += RK_WG) region[i] = 0;
  }
  uint64_t off = 0, t = D, r = D;
  int l = 0;
  while (r > 0) {
    const uint64_t n = bbh_level_n(t);
    if (l >= BBH_MAX_LEVELS) {
      if (threadIdx.x == 0) atomicOr(err, 2u);
      break;
    }
    const uint64_t nw = n >> 6;
    for (uint64_t i = threadIdx.x; i < nw; i += RK_WG) {
      seen[i] = 0;
      col[i] = 0;
    }
    __threadfence();
    __syncthreads();
    const uint64_t seed = bbh_seed(l);
    for (uint64_t e = threadIdx.x; e < D; e += RK_WG) {
      uint64_t gi = ko + e;
      if (!((ld_agent(&alive_g[gi >> 6]) >> (gi & 63)) & 1)) continue;
      uint64_t pos = __umul64hi(bbh_hash(keys[gi], seed), n);
      unsigned long long bit = 1ULL << (pos & 63);
      unsigned long long old = atomicOr((unsigned long long*)&seen[pos >> 6], bit);
      if (old & bit) atomicOr((unsigned long long*)&col[pos >> 6], bit);
    }
    __threadfence();
    __syncthreads();
    if (MODE) {
      for (uint64_t i = threadIdx.x; i < nw; i += RK_WG) {
        uint64_t q = (off >> 6) + i;
        region[(q / 7) * 8 + 1 + (q % 7)] = ld_agent(&seen[i]) & ~ld_agent(&col[i]);
      }
    }
    uint64_t cnt = 0;
    for (uint64_t e = threadIdx.x; e < D; e += RK_WG) {
      uint64_t gi = ko + e;
      if (!((ld_agent(&alive_g[gi >> 6]) >> (gi & 63)) & 1)) continue;
      uint64_t pos = __umul64hi(bbh_hash(keys[gi], seed), n);
      if (((ld_agent(&seen[pos >> 6]) & ~ld_agent(&col[pos >> 6])) >> (pos & 63)) & 1)
        atomicAnd((unsigned long long*)&alive_g[gi >> 6], ~(1ULL << (gi & 63)));
      else
        cnt++;
    }
    __threadfence();
    block_excl_scan_u64(cnt, sh8, &r);
    off += n;
    t = bbh_next_t(t);
    l++;
  }
  const uint64_t nblk = (off + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS;
  if (MODE == 0) {
    if (threadIdx.x == 0) seg_nblk[seg] = nblk;
    return;
  }
  __threadfence();
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint64_t j0 = 0; j0 < nblk; j0 += RK_WG) {
    uint64_t j = j0 + threadIdx.x;
    uint64_t pc = 0;
    if (j < nblk)
      for (int u = 0; u < 7; u++) pc += (uint64_t)__popcll(ld_agent(&region[j * 8 + 1 + u]));
    uint64_t tot;
    uint64_t rk = block_excl_scan_u64(pc, sh8, &tot);
    if (j < nblk) region[j * 8] = carry + rk;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launchers
static inline unsigned grid_for(uint64_t n, unsigned wg, unsigned cap) {
  uint64_t g = (n + wg - 1) / wg;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

void launch_p3_count(hipStream_t st, bool wide, const uint64_t* rec, const uint64_t* bin_start, const uint64_t* chunk_off,
                     int num_bins, uint32_t chunk_records, uint64_t n_chunks, int K, int sb, uint64_t* counts) {
  if (n_chunks == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_count", (k_p3_count<uint64_t, 2>), dim3((unsigned)n_chunks), dim3(P3_WG), 0, st, rec, bin_start, chunk_off,
                num_bins, chunk_records, K, sb, counts);
  else
    CF1G_LAUNCH("p3_count", (k_p3_count<Key128, 4>), dim3((unsigned)n_chunks), dim3(P3_WG), 0, st, rec, bin_start, chunk_off,
                num_bins, chunk_records, K, sb, counts);
}
void launch_p3_scatter(hipStream_t st, bool wide, const uint64_t* rec, const uint64_t* bin_start, const uint64_t* chunk_off,
                       int num_bins, uint32_t chunk_records, uint64_t n_chunks, int K, int sb, uint64_t* cursor, void* occ) {
  if (n_chunks == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_scatter", (k_p3_scatter<uint64_t, 2>), dim3((unsigned)n_chunks), dim3(P3_WG), 0, st, rec, bin_start,
                chunk_off, num_bins, chunk_records, K, sb, cursor, (uint64_t*)occ);
  else
    CF1G_LAUNCH("p3_scatter", (k_p3_scatter<Key128, 4>), dim3((unsigned)n_chunks), dim3(P3_WG), 0, st, rec, bin_start,
                chunk_off, num_bins, chunk_records, K, sb, cursor, (Key128*)occ);
}
void launch_p3_segsort(hipStream_t st, bool wide, void* occ, const uint64_t* seg_off, uint64_t n_segs, int K, int sb,
                       uint64_t* ucount, uint32_t* over_list, unsigned int* n_over) {
  if (n_segs == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_segsort", k_p3_segsort<uint64_t>, dim3((unsigned)n_segs), dim3(RK_WG), 0, st, (uint64_t*)occ, seg_off,
                n_segs, 2 * K, sb, ucount, over_list, n_over);
  else
    CF1G_LAUNCH("p3_segsort", k_p3_segsort<Key128>, dim3((unsigned)n_segs), dim3(RK_WG), 0, st, (Key128*)occ, seg_off, n_segs,
                2 * K, sb, ucount, over_list, n_over);
}
void launch_p3_bigsort(hipStream_t st, bool wide, void* occ, void* aux, const uint64_t* seg_off, const uint32_t* over_list,
                       uint32_t n_over, uint64_t* ucount) {
  if (n_over == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_bigsort", k_p3_bigsort<uint64_t>, dim3(n_over), dim3(RK_WG), 0, st, (uint64_t*)occ, (uint64_t*)aux,
                seg_off, over_list, ucount);
  else
    CF1G_LAUNCH("p3_bigsort", k_p3_bigsort<Key128>, dim3(n_over), dim3(RK_WG), 0, st, (Key128*)occ, (Key128*)aux, seg_off,
                over_list, ucount);
}
void launch_p3_compact(hipStream_t st, bool wide, const void* occ, const uint64_t* seg_off, const uint64_t* ucount,
                       const uint64_t* out_off, uint64_t n_segs, void* keys) {
  if (n_segs == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_compact", k_p3_compact<uint64_t>, dim3((unsigned)n_segs), dim3(256), 0, st, (const uint64_t*)occ, seg_off,
                ucount, out_off, n_segs, (uint64_t*)keys);
  else
    CF1G_LAUNCH("p3_compact", k_p3_compact<Key128>, dim3((unsigned)n_segs), dim3(256), 0, st, (const Key128*)occ, seg_off,
                ucount, out_off, n_segs, (Key128*)keys);
}
void launch_bbh_seg(hipStream_t st, bool wide, int mode, const void* keys, const uint64_t* key_off, uint64_t n_segs,
                    uint64_t* seg_nblk, const uint64_t* blk_off, uint64_t* blocks, uint32_t* big_list, unsigned int* n_big,
                    unsigned int* err) {
  if (n_segs == 0) return;
  dim3 g((unsigned)n_segs), b(RK_WG);
  if (!wide) {
    if (mode == 0)
      CF1G_LAUNCH("p3b_bbh_size", (k_bbh_seg<0, uint64_t>), g, b, 0, st, (const uint64_t*)keys, key_off, n_segs, seg_nblk, blk_off,
                  blocks, big_list, n_big, err);
    else
      CF1G_LAUNCH("p3b_bbh_build", (k_bbh_seg<1, uint64_t>), g, b, 0, st, (const uint64_t*)keys, key_off, n_segs, seg_nblk, blk_off,
                  blocks, big_list, n_big, err);
  } else {
    if (mode == 0)
      CF1G_LAUNCH("p3b_bbh_size", (k_bbh_seg<0, Key128>), g, b, 0, st, (const Key128*)keys, key_off, n_segs, seg_nblk, blk_off,
                  blocks, big_list, n_big, err);
    else
      CF1G_LAUNCH("p3b_bbh_build", (k_bbh_seg<1, Key128>), g, b, 0, st, (const Key128*)keys, key_off, n_segs, seg_nblk, blk_off,
                  blocks, big_list, n_big, err);
  }
}
void launch_bbh_big(hipStream_t st, bool wide, int mode, const void* keys, const uint64_t* key_off, const uint32_t* big_list,
                    uint32_t n_big, uint64_t* seg_nblk, const uint64_t* blk_off, uint64_t* blocks, uint64_t* scratch,
                    uint64_t* alive_g, unsigned int* err) {
  if (n_big == 0) return;
  dim3 g(n_big), b(RK_WG);
  if (!wide) {
    if (mode == 0)
      CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<0, uint64_t>), g, b, 0, st, (const uint64_t*)keys, key_off, big_list, seg_nblk, blk_off,
                  blocks, scratch, alive_g, err);
    else
      CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<1, uint64_t>), g, b, 0, st, (const uint64_t*)keys, key_off, big_list, seg_nblk, blk_off,
                  blocks, scratch, alive_g, err);
  } else {
    if (mode == 0)
      CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<0, Key128>), g, b, 0, st, (const Key128*)keys, key_off, big_list, seg_nblk, blk_off,
                  blocks, scratch, alive_g, err);
    else
      CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<1, Key128>), g, b, 0, st, (const Key128*)keys, key_off, big_list, seg_nblk, blk_off,
                  blocks, scratch, alive_g, err);
  }
}

}  // namespace cf1g
