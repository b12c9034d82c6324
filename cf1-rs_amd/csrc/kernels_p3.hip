// P3 kernels: per-bin canonical k-mer extraction, dedup (mphf.rs:323-363
// dedup_bin) and the BBHash build (the project's replacement for ph::fmph,
// mphf.rs:144-190).
//
// P3 layout: bin b's records (super k-mers) are split into 2^sb segments
// (b, s) by sb bits of their minimizer's hash, so that a segment is a
// function of the super k-mer and its k-mers never span two segments.
// k_p3_rcount counts records and k-mers per segment, k_p3_gate checks the
// counted k-mers against the P1 total, and k_p3_rscatter sorts the records
// into their segments' runs. k_p3_dedup deduplicates one segment per
// work-group, decoding its records itself (LDS hash table, or a
// direct-address table for keys of <= 12 bits), OR-joins the observation
// masks of equal keys and builds the segment's BBHash into a padded block
// slot, writing the masks in MPHF value order; k_p3_dedup_list does the
// same at 512 threads for the segments above its capacity, and larger ones
// take a global-memory radix sort (k_p3_expand, k_p3_bigsort) and
// k_bbh_big. k_p3_compact then packs the blocks densely; P4 turns the masks,
// in value order within each segment's run, into states.
#include <algorithm>

#include "block.hpp"
#include "device.hpp"
#include "launch.hpp"

namespace cf1g {

// ------------------------------------------------------------------ scan
constexpr int SC_WG = 256;
constexpr int SC_IPT = 16;
constexpr int SC_TILE = SC_WG * SC_IPT;

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

// Zero `bytes` bytes at p (p 8-byte aligned) / copy one word: kernels, so
// that a captured launch sequence holds no memset or copy nodes.
__global__ void __launch_bounds__(256) k_zero(uint64_t* __restrict__ p, uint64_t words, uint8_t* __restrict__ tail, uint32_t tail_bytes) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < tail_bytes) tail[threadIdx.x] = 0;
}
__global__ void k_copy_word(uint64_t* __restrict__ dst, const uint64_t* __restrict__ src) { *dst = *src; }
void launch_zero(hipStream_t st, void* p, uint64_t bytes) {
  if (bytes == 0) return;
  const uint64_t words = bytes / 8;
  const uint32_t tail = (uint32_t)(bytes % 8);
  const uint64_t g = std::min<uint64_t>(std::max<uint64_t>((words + 255) / 256, 1), 4096);
  CF1G_LAUNCH("zero", k_zero, dim3((unsigned)g), dim3(256), 0, st, (uint64_t*)p, words, (uint8_t*)p + 8 * words, tail);
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
    if (total) launch_zero(st, total, 8);
    return;
  }
  uint64_t nb = (n + SC_TILE - 1) / SC_TILE;
  uint64_t* sums = tmp;
  CF1G_LAUNCH("scan_tile", k_scan_tile, dim3((unsigned)nb), dim3(SC_WG), 0, st, in, out, n, sums);
  if (nb == 1) {
    if (total) CF1G_LAUNCH("scan_total", k_copy_word, dim3(1), dim3(1), 0, st, total, (const uint64_t*)sums);
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

// Iterate the k-mer occurrences of one record: f(canonical, observation mask).
// Record positions: 0 = left flank, 1..L = super k-mer, L+1 = right flank.
// The observation of k-mer j follows classify.rs:65-177: neighbours are the
// bases at j and j+K+1 (valid inside the record, or by the flank flags),
// isolated -> MIMO, no right -> rightmost, self loop (canonical of the next
// k-mer equal) -> MIMO or rightmost, else leftmost / internal.
template <typename KeyT, int RW, bool OBS, typename F>
__device__ __forceinline__ void for_each_rec_occ(const uint64_t* __restrict__ rec, uint64_t idx, int K, F f) {
  uint64_t w[RW];
#pragma unroll
  for (int t = 0; t < RW; t++) w[t] = rec[idx * RW + t];
  const uint64_t tail = w[RW - 1];
  int L = (int)(tail & 255);
  if (L > rec_cap(RW)) L = rec_cap(RW);
  const bool lf = (tail >> 8) & 1, rf = (tail >> 9) & 1;
  const int n = L - K + 1;
  if (n <= 0) return;
  Roll<KeyT> ro;
  ro.init(K);
  for (int t = 1; t <= K; t++) ro.push(rec_base<RW>(w, t));
  for (int j = 0; j < n; j++) {
    const KeyT c = ro.canon();
    if (OBS) {
      const bool pv = j > 0 || lf, nv = j < n - 1 || rf;
      const uint32_t pb = rec_base<RW>(w, j), nb = rec_base<RW>(w, j + K + 1);
      const bool fwd = key_eq(ro.x, c);
      Roll<KeyT> nx = ro;
      nx.push(nb);
      uint32_t type;
      if (!pv && !nv) type = OBS_MIMO;
      else if (!nv) type = OBS_RIGHT;
      else if (key_eq(nx.canon(), c)) type = (!pv || key_eq(ro.x, nx.x)) ? OBS_MIMO : OBS_RIGHT;
      else type = pv ? OBS_INTERNAL : OBS_LEFT;
      f(c, obs_mask(type, fwd, pb, nb));
      ro = nx;
    } else {
      f(c, 0u);
      if (j + 1 < n) ro.push(rec_base<RW>(w, j + K + 1));
    }
  }
}

constexpr int P3_WG = 256;
constexpr int P3_MAX_SB = 12;

// ---- one lane per k-mer for u64-key records (RW = 2, K <= 31). The record
// is the 128-bit value W = w0:w1, position t at bits [126-2t, 127-2t]; k-mer
// j spans positions j+1..j+K.
__device__ __forceinline__ uint64_t rec2_bits(uint64_t w0, uint64_t w1, int sh) {  // low 64 bits of W >> sh, 0 < sh < 128
  // both halves of the funnel, then a select (lanes of a wave take either
  // side; branches would run both with exec masks)
  const uint64_t lo = (w1 >> (sh & 63)) | (w0 << ((64 - sh) & 63));  // 0 < sh < 64
  const uint64_t hi = w0 >> (sh & 63);                               // 64 <= sh < 128
  return sh >= 64 ? hi : lo;
}
__device__ __forceinline__ int rec2_kmers(uint64_t w1, int K) {
  int L = (int)(w1 & 255);
  if (L > rec_cap(2)) L = rec_cap(2);
  return L >= K ? L - K + 1 : 0;
}
// Canonical k-mer j of a record of n k-mers and (OBS) its observation mask,
// the same as the j-th call of for_each_rec_occ.
template <bool OBS>
__device__ __forceinline__ uint64_t rec2_occ(uint64_t w0, uint64_t w1, int j, int n, int K, uint64_t mask, uint32_t& m) {
  const int sh = 126 - 2 * (j + K);
  const uint64_t x = rec2_bits(w0, w1, sh) & mask;
  const uint64_t r = rc64(x, K);
  const uint64_t c = x < r ? x : r;
  if (OBS) {
    const bool lf = (w1 >> 8) & 1, rf = (w1 >> 9) & 1;
    const bool pv = j > 0 || lf, nv = j < n - 1 || rf;
    const uint32_t pb = (uint32_t)rec2_bits(w0, w1, 126 - 2 * j) & 3u;
    const uint32_t nb = (uint32_t)rec2_bits(w0, w1, sh - 2) & 3u;
    const uint64_t nx = ((x << 2) | nb) & mask;
    const uint64_t nr = (r >> 2) | ((uint64_t)(3 - nb) << (2 * K - 2));
    const uint64_t nc = nx < nr ? nx : nr;
    uint32_t type;
    if (!pv && !nv) type = OBS_MIMO;
    else if (!nv) type = OBS_RIGHT;
    else if (nc == c) type = (!pv || x == nx) ? OBS_MIMO : OBS_RIGHT;
    else type = pv ? OBS_INTERNAL : OBS_LEFT;
    m = obs_mask(type, x == c, pb, nb);
  }
  return c;
}
// ---- one lane per k-mer for wide records (RW = 4, 32 < K <= 63). The
// record is the 256-bit value W = w0:w1:w2:w3, position t at bits
// [254-2t, 255-2t]; k-mer j spans positions j+1..j+K.
struct Rec4 {
  uint64_t w0, w1, w2, w3;
};
// 64 bits of W from bit s up (bit 0 = least significant of w3), 0 <= s < 256.
__device__ __forceinline__ uint64_t rec4_bits(const Rec4& R, int s) {
  const int q = s >> 6, r = s & 63;
  const uint64_t a = q == 0 ? R.w3 : q == 1 ? R.w2 : q == 2 ? R.w1 : R.w0;
  const uint64_t b = q == 0 ? R.w2 : q == 1 ? R.w1 : q == 2 ? R.w0 : 0;
  return r ? (a >> r) | (b << (64 - r)) : a;
}
__device__ __forceinline__ int rec4_kmers(uint64_t w3, int K) {
  int L = (int)(w3 & 255);
  if (L > rec_cap(4)) L = rec_cap(4);
  return L >= K ? L - K + 1 : 0;
}
// Canonical k-mer j of a wide record of n k-mers and (OBS) its observation
// mask, the same as the j-th call of for_each_rec_occ.
template <bool OBS>
__device__ __forceinline__ Key128 rec4_occ(const Rec4& R, int j, int n, int K, uint32_t& m) {
  const uint64_t mh = (1ULL << (2 * K - 64)) - 1;
  const int sh = 254 - 2 * (j + K);  // bit of k-mer j's last base
  Key128 x;
  x.lo = rec4_bits(R, sh);
  x.hi = rec4_bits(R, sh + 64) & mh;
  const Key128 r = rc128(x, K);
  const Key128 c = kmer_min(x, r);
  if (OBS) {
    const bool lf = (R.w3 >> 8) & 1, rf = (R.w3 >> 9) & 1;
    const bool pv = j > 0 || lf, nv = j < n - 1 || rf;
    const uint32_t pb = (uint32_t)rec4_bits(R, 254 - 2 * j) & 3u;
    const uint32_t nb = (uint32_t)rec4_bits(R, sh - 2) & 3u;
    Key128 nx, nr;
    nx.hi = ((x.hi << 2) | (x.lo >> 62)) & mh;
    nx.lo = (x.lo << 2) | nb;
    nr.lo = (r.lo >> 2) | (r.hi << 62);
    nr.hi = (r.hi >> 2) | ((uint64_t)(3 - nb) << (2 * K - 66));
    const Key128 nc = kmer_min(nx, nr);
    uint32_t type;
    if (!pv && !nv) type = OBS_MIMO;
    else if (!nv) type = OBS_RIGHT;
    else if (key_eq(nc, c)) type = (!pv || key_eq(x, nx)) ? OBS_MIMO : OBS_RIGHT;
    else type = pv ? OBS_INTERNAL : OBS_LEFT;
    m = obs_mask(type, key_eq(x, c), pb, nb);
  }
  return c;
}

// k-mers j0 .. j0+cnt-1 (cnt <= PK) of a record of n k-mers with their
// observation masks, the same as rec2_occ / rec4_occ for each j: the first
// k-mer and its reverse complement are cut out of the record, the rest are
// rolled one base at a time (kmer.rs:240-262), and the base before k-mer j
// (j > j0) is the top base of k-mer j-1.
template <int PK, typename FK>
__device__ __forceinline__ void rec2_piece(uint64_t w0, uint64_t w1, int j0, int cnt, int n, int K, uint64_t mask, FK out) {
  static_assert(PK <= 4, "the next bases of a piece come from one byte of the record");
  uint64_t x = rec2_bits(w0, w1, 126 - 2 * (j0 + K)) & mask;
  uint64_t r = rc64(x, K);
  const bool lf = (w1 >> 8) & 1, rf = (w1 >> 9) & 1;
  uint32_t pb = (uint32_t)rec2_bits(w0, w1, 126 - 2 * j0) & 3u;
  // positions j0 + K + 1 .. j0 + K + 4 (the base after each k-mer of the
  // piece) as one byte, the first in its top bits; a record holds at most
  // rec_cap(2) + 2 <= 64 positions, so the shift is >= 0 for j0 + K + 4 <= 63
  const uint32_t nbw = (uint32_t)rec2_bits(w0, w1, 126 - 2 * (j0 + K + 4)) & 0xFFu;
#pragma unroll
  for (int i = 0; i < PK; i++) {
    if (i < cnt) {
      const int j = j0 + i;
      const uint64_t c = x < r ? x : r;
      const bool pv = i > 0 || j > 0 || lf, nv = j < n - 1 || rf;
      const uint32_t nb = (nbw >> (6 - 2 * i)) & 3u;
      const uint64_t nx = ((x << 2) | nb) & mask;
      const uint64_t nr = (r >> 2) | ((uint64_t)(3 - nb) << (2 * K - 2));
      const uint64_t nc = nx < nr ? nx : nr;
      // the type of rec2_occ as (has a previous, has a next) neighbour, no
      // branches: a self loop (next canonical equal) drops the next one, and
      // the previous one too when the next k-mer is this one itself
      const bool self = nv && nc == c;
      const bool hn = nv && !self, hp = pv && !(self && x == nx);
      out(i, c, obs_mask((hp ? OBS_RIGHT : 0u) | (hn ? OBS_LEFT : 0u), x == c, pb, nb));
      pb = (uint32_t)(x >> (2 * K - 2)) & 3u;
      x = nx;
      r = nr;
    }
  }
}
template <int PK, typename FK>
__device__ __forceinline__ void rec4_piece(const Rec4& R, int j0, int cnt, int n, int K, FK out) {
  const uint64_t mh = (1ULL << (2 * K - 64)) - 1;
  const int sh = 254 - 2 * (j0 + K);
  Key128 x;
  x.lo = rec4_bits(R, sh);
  x.hi = rec4_bits(R, sh + 64) & mh;
  Key128 r = rc128(x, K);
  const bool lf = (R.w3 >> 8) & 1, rf = (R.w3 >> 9) & 1;
  uint32_t pb = (uint32_t)rec4_bits(R, 254 - 2 * j0) & 3u;
#pragma unroll
  for (int i = 0; i < PK; i++) {
    if (i < cnt) {
      const int j = j0 + i;
      const Key128 c = kmer_min(x, r);
      const bool pv = j > 0 || lf, nv = j < n - 1 || rf;
      const uint32_t nb = (uint32_t)rec4_bits(R, sh - 2 - 2 * i) & 3u;
      Key128 nx, nr;
      nx.hi = ((x.hi << 2) | (x.lo >> 62)) & mh;
      nx.lo = (x.lo << 2) | nb;
      nr.lo = (r.lo >> 2) | (r.hi << 62);
      nr.hi = (r.hi >> 2) | ((uint64_t)(3 - nb) << (2 * K - 66));
      const Key128 nc = kmer_min(nx, nr);
      uint32_t type;
      if (!pv && !nv) type = OBS_MIMO;
      else if (!nv) type = OBS_RIGHT;
      else if (key_eq(nc, c)) type = (!pv || key_eq(x, nx)) ? OBS_MIMO : OBS_RIGHT;
      else type = pv ? OBS_INTERNAL : OBS_LEFT;
      out(i, c, obs_mask(type, key_eq(x, c), pb, nb));
      pb = (uint32_t)(x.hi >> (2 * K - 66)) & 3u;
      x = nx;
      r = nr;
    }
  }
}

// Round-local k-mer -> record map for P3_WG records of at most
// OwnKmers<RW>::N k-mers each (P2 emits super k-mers of at most w k-mers:
// w <= 17 for K <= 31 and w <= 49 for K <= 63 with the default minimizer
// length; a round with longer records takes the per-record loop).
constexpr int OWN_KMERS = 17;
template <int RW>
struct OwnKmers {
  static constexpr int N = RW == 2 ? OWN_KMERS : 49;
};
// k-mer j of record r (n k-mers) for the one-lane-per-k-mer paths.
template <typename KeyT, int RW, bool OBS>
__device__ __forceinline__ KeyT rec_occ_at(const uint64_t* __restrict__ rec, uint64_t r, int j, int K, uint64_t kmask, uint32_t& m) {
  if constexpr (RW == 2) {
    const uint64_t x0 = rec[r * 2], x1 = rec[r * 2 + 1];
    return (KeyT)rec2_occ<OBS>(x0, x1, j, rec2_kmers(x1, K), K, kmask, m);
  } else {
    const Rec4 R{rec[r * 4], rec[r * 4 + 1], rec[r * 4 + 2], rec[r * 4 + 3]};
    return rec4_occ<OBS>(R, j, rec4_kmers(R.w3, K), K, m);
  }
}
template <int RW>
__device__ __forceinline__ int rec_kmers_at(const uint64_t* __restrict__ rec, uint64_t r, int K) {
  if constexpr (RW == 2) return rec2_kmers(rec[r * 2 + 1], K);
  else return rec4_kmers(rec[r * 4 + 3], K);
}


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

// ---- records grouped by MPHF segment. A record's segment inside its bin is
// fixed by its minimizer: the sb bits just below the bucket in the
// minimizer's hash (rec_sub holds the 16 hash bits below the bucket), so
// every k-mer of a super k-mer lands in the same segment and P3 moves
// records (16 or 32 bytes per super k-mer) instead of occurrences.

// Records and k-mers per (bin, segment), counted per chunk of a bin's records
// in LDS (2^sb counters each, dynamic) and added to rcnt / kcnt.
template <int RW>
__global__ void __launch_bounds__(P3_WG) k_p3_rcount(const uint64_t* __restrict__ rec, const uint16_t* __restrict__ rec_sub,
                                                     const uint64_t* __restrict__ bin_start, const uint64_t* __restrict__ chunk_off,
                                                     int num_bins, uint32_t chunk_records, int K, int sb,
                                                     unsigned long long* __restrict__ rcnt, unsigned long long* __restrict__ kcnt) {
  extern __shared__ __align__(16) uint32_t rk_lds[];  // records, then k-mers, 2^sb each
  int bin;
  uint64_t r0, r1;
  if (!p3_chunk(chunk_off, bin_start, num_bins, chunk_records, bin, r0, r1)) return;
  const int ns = 1 << sb;
  uint32_t* rc = rk_lds;
  uint32_t* kc = rk_lds + ns;
  for (int i = threadIdx.x; i < 2 * ns; i += P3_WG) rk_lds[i] = 0;
  __syncthreads();
  for (uint64_t r = r0 + threadIdx.x; r < r1; r += P3_WG) {
    const uint32_t sg = rec_seg(rec_sub[r], sb);
    const int n = rec_kmers_at<RW>(rec, r, K);
    atomicAdd(&rc[sg], 1u);
    atomicAdd(&kc[sg], (uint32_t)n);
  }
  __syncthreads();
  const uint64_t s0 = (uint64_t)bin << sb;
  for (int sg = threadIdx.x; sg < ns; sg += P3_WG) {
    if (rc[sg]) {
      atomicAdd(&rcnt[s0 + sg], (unsigned long long)rc[sg]);
      atomicAdd(&kcnt[s0 + sg], (unsigned long long)kc[sg]);
    }
  }
}

// Records of each chunk into their segments' runs: a tile of RS_TILE records
// is ranked per segment in LDS, each (tile, segment) run is reserved from
// the segment's cursor, and the tile is staged in segment order so that
// consecutive lanes store consecutive records of one run. The order of
// records inside a segment depends on the schedule; nothing downstream
// depends on it (the dedup joins and ranks by key).
// Records per staged tile: 2,048 two-word records when the tile and the
// bin's segment counters fit 64 KiB of LDS (sb <= 10; on the 300 Mbp genome
// at 256 bins this pass takes ~0.44 against 0.57 ms with 1,024, each
// (tile, segment) run holding twice the records), else 1,024; 512 four-word
// records.
template <int RW, int TILE>
size_t rscatter_lds(int sb) {
  return (size_t)TILE * RW * 8 + (size_t)TILE * 2 + ((size_t)12 << sb) + 64;
}
constexpr size_t RS_LDS_MAX = 65536;
template <int RW, int TILE>
__global__ void __launch_bounds__(P3_WG) k_p3_rscatter(const uint64_t* __restrict__ rec, const uint16_t* __restrict__ rec_sub,
                                                       const uint64_t* __restrict__ bin_start, const uint64_t* __restrict__ chunk_off,
                                                       int num_bins, uint32_t chunk_records, int sb, const uint64_t* __restrict__ roff,
                                                       unsigned int* __restrict__ rcur, uint64_t* __restrict__ srec,
                                                       uint32_t* __restrict__ rec_dest, const unsigned int* __restrict__ stop) {
  constexpr int PT = TILE / P3_WG;
  extern __shared__ __align__(16) uint64_t rs_lds[];
  __shared__ uint64_t sh4[4];
  int bin;
  uint64_t r0, r1;
  if (*stop || !p3_chunk(chunk_off, bin_start, num_bins, chunk_records, bin, r0, r1)) return;
  const int ns = 1 << sb;
  uint64_t* tr = rs_lds;                          // staged records, TILE * RW words
  uint16_t* ts = (uint16_t*)(tr + TILE * RW);     // their segments
  uint32_t* cnt = (uint32_t*)(ts + TILE);         // per segment: records in the tile
  uint32_t* beg = cnt + ns;                       //   their first staged index
  uint32_t* gb = beg + ns;                        //   their run's offset in the segment
  const uint64_t s0 = (uint64_t)bin << sb;
  const int per = ns > P3_WG ? ns / P3_WG : 1;    // counters per thread in the scan
  for (uint64_t t0 = r0; t0 < r1; t0 += TILE) {
    const uint32_t tn = (uint32_t)(r1 - t0 < (uint64_t)TILE ? r1 - t0 : (uint64_t)TILE);
    for (int i = threadIdx.x; i < ns; i += P3_WG) cnt[i] = 0;
    __syncthreads();
    uint64_t w[PT][RW];
    uint32_t sg[PT], rk[PT];
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const uint32_t i = (uint32_t)q * P3_WG + threadIdx.x;
      sg[q] = 0;
      rk[q] = 0;
      if (i < tn) {
#pragma unroll
        for (int t = 0; t < RW; t++) w[q][t] = rec[(t0 + i) * RW + t];
        sg[q] = rec_seg(rec_sub[t0 + i], sb);
        rk[q] = atomicAdd(&cnt[sg[q]], 1u);
      }
    }
    __syncthreads();
    // exclusive scan of the counts (per consecutive counters per thread) and
    // one reservation per non-empty (tile, segment) run
    uint32_t loc = 0;
    const int a = (int)threadIdx.x * per;
    if (a < ns)
      for (int q = 0; q < per; q++) loc += cnt[a + q];
    uint32_t run = (uint32_t)block_excl_scan_u64(loc, sh4, nullptr);
    if (a < ns)
      for (int q = 0; q < per; q++) {
        const uint32_t c = cnt[a + q];
        beg[a + q] = run;
        gb[a + q] = c ? atomicAdd(&rcur[s0 + a + q], c) : 0u;
        run += c;
      }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const uint32_t i = (uint32_t)q * P3_WG + threadIdx.x;
      if (i < tn) {
        const uint32_t j = beg[sg[q]] + rk[q];
#pragma unroll
        for (int t = 0; t < RW; t++) tr[j * RW + t] = w[q][t];
        ts[j] = (uint16_t)sg[q];
        rec_dest[t0 + i] = (uint32_t)(roff[s0 + sg[q]] + gb[sg[q]] + rk[q]);  // bin order -> segment order (for P5)
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tn; i += P3_WG) {
      const uint32_t g = ts[i];
      const uint64_t out = roff[s0 + g] + gb[g] + (i - beg[g]);
#pragma unroll
      for (int t = 0; t < RW; t++) srec[out * RW + t] = tr[i * RW + t];
    }
    __syncthreads();
  }
}

// Occurrences (canonical key, observation mask) of the listed segments'
// records, in record order, at occ / obs[seg_off[seg] ...] (all segments
// when list is null): the input of the global-memory sort and of the
// direct-address dedup. One work-group per segment, 256 records at a time.
template <typename KeyT, int RW>
__global__ void __launch_bounds__(P3_WG) k_p3_expand(const uint64_t* __restrict__ srec, const uint64_t* __restrict__ roff,
                                                     const uint64_t* __restrict__ seg_off, const uint32_t* __restrict__ list,
                                                     int K, KeyT* __restrict__ occ, uint16_t* __restrict__ obs,
                                                     const unsigned int* __restrict__ stop) {
  __shared__ uint64_t sh4[4];
  if (*stop) return;
  const uint64_t seg = list ? (uint64_t)list[blockIdx.x] : (uint64_t)blockIdx.x;
  const uint64_t ra = roff[seg], rb = roff[seg + 1];
  uint64_t at = seg_off[seg];
  for (uint64_t rr = ra; rr < rb; rr += P3_WG) {
    const uint64_t r = rr + threadIdx.x;
    const int n = r < rb ? rec_kmers_at<RW>(srec, r, K) : 0;
    uint64_t T;
    uint64_t o = at + block_excl_scan_u64((uint64_t)n, sh4, &T);
    if (n)
      for_each_rec_occ<KeyT, RW, true>(srec, r, K, [&](const KeyT& c, uint32_t m) {
        occ[o] = c;
        obs[o] = (uint16_t)m;
        o++;
      });
    at += T;
  }
}

// ------------------------------------------------------------------ P5 annotation
// What AnnotatedKmer::new computes per k-mer for P5 (src/directed_kmer.rs:81-96)
// — valid bit, forward bit, class of the k-mer's state — without a lookup per
// stream position: one work-group per MPHF segment decodes the k-mers of the
// segment's records (P3 left them grouped by segment, each with its P1 entry:
// stream position and length), finds each key's value through that
// segment's BBHash blocks staged in LDS, and writes the annotation at the
// k-mer's stream position.
// The P1 entry of each record in segment order: the route kept where each
// entry's record went (ent_dest, in entry order) and the record sort where
// each record went (rec_dest), so list[rec_dest[ent_dest[slot]]] = entry.
__global__ void __launch_bounds__(256) k_p5_list(const uint64_t* __restrict__ entries, const uint32_t* __restrict__ ent_dest,
                                                 const uint32_t* __restrict__ tile_cnt, uint64_t n_tiles, uint32_t tile_slots,
                                                 const uint32_t* __restrict__ rec_dest, uint64_t S, uint64_t* __restrict__ list,
                                                 unsigned int* __restrict__ err) {
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint32_t c = tile_cnt[t];
    for (uint32_t i = threadIdx.x; i < c; i += 256) {
      const uint64_t slot = t * tile_slots + i;
      const uint32_t rb = ent_dest[slot];
      const uint32_t rs = rb < S ? rec_dest[rb] : ~0u;
      if (rs < S) list[rs] = entries[slot];
      else atomicOr(err, 1u);
    }
  }
}
template <int RW>
struct P5Own {
  static constexpr int N = RW == 2 ? 17 : 49;  // k-mers per record in the owner map (w with the default m)
};
constexpr int P5_LB = 64;  // BBHash blocks staged in LDS (a segment of up to ~3,000 keys)
template <typename KeyT, int RW>
__global__ void __launch_bounds__(P3_WG) k_p5_annot(const uint64_t* __restrict__ srec, const uint64_t* __restrict__ srec_ent,
                                                    const uint64_t* __restrict__ roff, const uint64_t* __restrict__ key_off,
                                                    const uint64_t* __restrict__ blk_off, const uint64_t* __restrict__ blocks,
                                                    const uint64_t* __restrict__ states, int K, uint64_t n_segs,
                                                    uint8_t* __restrict__ annot, unsigned int* __restrict__ err) {
  constexpr int OWN = P5Own<RW>::N;
  constexpr int LW = P5_LB * 7;
  __shared__ uint64_t lw[LW];  // the segment's level bits (7 words per block), if it has at most P5_LB blocks
  __shared__ uint32_t lr[LW];  // rank before each word
  __shared__ uint64_t srl[P3_WG * RW];
  __shared__ uint64_t sent[P3_WG];
  __shared__ uint16_t oo[P3_WG];
  __shared__ uint8_t own[P3_WG * OWN];
  __shared__ uint64_t sh4[4];
  const uint64_t seg = blockIdx.x;
  if (seg >= n_segs) return;
  const uint64_t ko = key_off[seg], u = key_off[seg + 1] - ko;
  const uint64_t b0 = blk_off[seg], nb = blk_off[seg + 1] - b0;
  const uint64_t* base = blocks + b0 * 8;
  const bool staged = nb <= (uint64_t)P5_LB;
  if (staged) {
    // block j word q (1..7) is level word 7j + q - 1 (448 = 7 * 64 bits per
    // block); its rank = the block's rank word + popcounts of the words before
    for (uint64_t j = threadIdx.x; j < nb; j += P3_WG) {
      uint32_t rk = (uint32_t)base[j * 8];
#pragma unroll
      for (int q = 0; q < 7; q++) {
        const uint64_t x = base[j * 8 + 1 + q];
        lw[j * 7 + q] = x;
        lr[j * 7 + q] = rk;
        rk += (uint32_t)__popcll(x);
      }
    }
  }
  __syncthreads();
  const uint64_t lim = nb * BBH_BLOCK_BITS;
  // local rank of a present key through the staged level words (the levels
  // of mphf_local_rank, read from LDS)
  auto lds_rank = [&](const KeyT& c) -> uint64_t {
    uint64_t t = u, off = 0, h = bbh_h1(c);
    const uint64_t d = bbh_d(h);
    for (int l = 0; l < BBH_MAX_LEVELS && t > 0; l++, h += d) {
      const uint64_t n = bbh_level_n(t);
      if (off + n > lim) break;
      const uint32_t p = (uint32_t)(off + bbh_pos(h, (uint32_t)n));
      const uint64_t x = lw[p >> 6];
      if ((x >> (p & 63)) & 1) return (uint64_t)lr[p >> 6] + (uint64_t)__popcll(x & ((1ULL << (p & 63)) - 1));
      off += n;
      t = bbh_next_t(t);
    }
    return ~0ULL;
  };
  const uint64_t kmask = K < 32 ? (1ULL << (2 * K)) - 1 : ~0ULL;
  // k-mer j of the record at w (n k-mers), at stream position p
  auto one = [&](const uint64_t* w, int j, int n, uint64_t p) {
    KeyT x, c;
    if constexpr (RW == 2) {
      x = rec2_bits(w[0], w[1], 126 - 2 * (j + K)) & kmask;
      c = kmer_min(x, rc64(x, K));
    } else {
      const Rec4 R{w[0], w[1], w[2], w[3]};
      const int sh = 254 - 2 * (j + K);
      x.lo = rec4_bits(R, sh);
      x.hi = rec4_bits(R, sh + 64) & ((1ULL << (2 * K - 64)) - 1);
      c = kmer_min(x, rc128(x, K));
    }
    (void)n;
    const uint64_t rk = staged ? lds_rank(c) : mphf_local_rank<KeyT>(base, u, c);
    uint8_t a = 0;
    if (rk < u) {
      const uint64_t v = ko + rk;
      const uint32_t s = (uint32_t)(states[v / 12] >> ((v % 12) * 5)) & 31u;
      a = (uint8_t)(1u | (key_eq(x, c) ? 2u : 0u) | (st_class(s) << 2));
    } else {
      atomicOr(err, 2u);
    }
    annot[p] = a;
  };
  const uint64_t ra = roff[seg], rb = roff[seg + 1];
  for (uint64_t rr = ra; rr < rb; rr += P3_WG) {
    const uint64_t r = rr + threadIdx.x;
    int n = 0;
    if (r < rb) {
#pragma unroll
      for (int t = 0; t < RW; t++) srl[threadIdx.x * RW + t] = srec[r * RW + t];
      sent[threadIdx.x] = srec_ent[r];
      n = rec_kmers_at<RW>(srec, r, K);
    }
    uint64_t T;
    const uint32_t o = (uint32_t)block_excl_scan_u64((uint64_t)n | ((uint64_t)(n > OWN) << 32), sh4, &T);
    if (T >> 32) {  // a record longer than the owner map: each lane walks its own
      for (int j = 0; j < n; j++) one(srl + threadIdx.x * RW, j, n, (sent[threadIdx.x] >> 24) + (uint64_t)j);
      __syncthreads();
      continue;
    }
    if (n) {
      oo[threadIdx.x] = (uint16_t)o;
      for (int q = 0; q < n; q++) own[o + q] = (uint8_t)threadIdx.x;
    }
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < (uint32_t)T; g += P3_WG) {
      const uint32_t rl = own[g];
      const int j = (int)(g - oo[rl]);
      one(srl + rl * RW, j, 0, (sent[rl] >> 24) + (uint64_t)j);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ segment dedup
// Per-segment capacity of the register/LDS paths: keys are held KPT per
// thread. Larger segments take the global-memory paths (k_p3_bigsort,
// k_bbh_big).
template <typename KeyT>
struct SegCfg;
template <>
struct SegCfg<uint64_t> {
  static constexpr int KPT = 8;
};
template <>
struct SegCfg<Key128> {
  static constexpr int KPT = 8;
};
constexpr int BM_MAX_BITS = 12;   // direct-address mask table for keys of <= 12 bits (k <= 6)
// k_p3_dedup: DD_WG threads, DD_KPT occurrences each, at most 2,048 per
// segment; k_p3_dedup_list takes the segments of up to 4,096 occurrences
// that exceed it, DD2_WG threads each (the 300 Mbp genome's largest segment
// holds ~2,600)
constexpr int DD_WG = 256;
constexpr int DD_KPT = 8;
constexpr int DD2_WG = 512;
constexpr int DEC_PK = 4;  // k-mers per lane in the dedup's record decode

// A segment's MPHF is built by k_bbh_big when its distinct keys exceed the
// register capacity, or when it was sorted by k_p3_bigsort (more
// occurrences than capacity, keys too wide for the direct-address table).
template <typename KeyT, int CAP>
__device__ __forceinline__ bool seg_is_big(uint64_t n, uint64_t u, int K2) {
  const bool table = sizeof(KeyT) == 8 && K2 <= BM_MAX_BITS;
  return u > (uint64_t)CAP || (n > (uint64_t)CAP && !table);
}

// Each segment builds its BBHash blocks into a padded slot of
// floor(n/64) + 16 blocks (n = the segment's occurrences) at block offset
// floor(seg_off/64) + 16*seg; k_p3_compact then packs them densely. The
// slot bounds the region: sum_l roundup64(2 t_l) over t_{l+1} = ceil(9 t_l/20)
// is < 3.7u + 67L bits, i.e. < u/121 + 10 blocks for L <= 64 levels.
constexpr uint64_t PAD_BLOCKS = 16;
__device__ __forceinline__ uint64_t pad_slot_off(uint64_t seg_start, uint64_t seg) { return seg_start / 64 + PAD_BLOCKS * seg; }
__device__ __forceinline__ uint64_t pad_slot_blocks(uint64_t n_occ) { return n_occ / 64 + PAD_BLOCKS; }

// ---- BBHash levels over keys held in registers (see device.hpp), key slot s
// of thread t being element s*WG + t. Writes the level words into `region`
// and each element's bit position into ppos[element] (LDS); returns the
// region size in bits. While more than TAIL_KEYS keys remain a level is
// block-wide: the seen / collision bitmaps are double-buffered (A for even
// levels, B for odd ones), so it takes two barriers, after marking and after
// placing (which also clears the next level's buffers and counts the keys
// left). The last TAIL_KEYS or fewer keys go to a dense list in LDS (level
// state, step, element) that wave 0 walks 64 keys at a time, with no block
// barriers (the LDS operations of one wave execute in order); each level
// compacts the unplaced keys to the front of the list. seenA / colA must be
// clear for level 0. A block-wide level costs every slot of every wave however
// few keys are left, the list one pass per 64 keys but on one wave: on the
// 300 Mbp genome (~1,100 keys per segment) 256 measured best (dedup 3.21 ms,
// against 3.33 ms with 128 and 3.39 ms with 512).
constexpr int TAIL_KEYS = 256;
struct BbhList {
  uint64_t* h;  // level state
  uint64_t* d;  // step
  uint16_t* e;  // element
};
template <int WG>
struct BbhShared {
  uint32_t left[2][WG / 64];
  uint64_t off;
};
// Sum over the wave of a per-lane count in [0, 15].
__device__ __forceinline__ uint32_t wave_sum4(uint32_t c) {
  uint32_t s = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) s += (uint32_t)__popcll(__ballot((c >> b) & 1)) << b;
  return s;
}
// Exclusive prefix over the lanes below this one of a per-lane count in [0, 15].
__device__ __forceinline__ uint32_t wave_prefix4(uint32_t c) {
  const uint64_t below = (1ULL << __lane_id()) - 1;
  uint32_t s = 0;
#pragma unroll
  for (int b = 0; b < 4; b++) s += (uint32_t)__popcll(__ballot((c >> b) & 1) & below) << b;
  return s;
}

template <int KPT, int WG>
__device__ uint64_t bbh_levels(const uint64_t (&h1)[KPT], uint32_t alive, int ns, uint64_t D, uint64_t* seenA, uint64_t* colA,
                               uint64_t* seenB, uint64_t* colB, uint64_t* region, uint64_t max_bits, uint16_t* ppos, BbhList L,
                               BbhShared<WG>* sh, unsigned int* err) {
  // level l state of a slot: h1 + l * bbh_d(h1) (device.hpp), recomputed per
  // level so that only h1 stays in registers
  const uint32_t wv = threadIdx.x >> 6;
  uint64_t off = 0, t = D;
  int l = 0;
  uint64_t* seen = seenA;
  uint64_t* col = colA;
  uint64_t* nseen = seenB;
  uint64_t* ncol = colB;
  uint32_t total = (uint32_t)D;
  while (total > TAIL_KEYS) {
    const uint64_t n = bbh_level_n(t);
    if (l >= BBH_MAX_LEVELS || off + n > max_bits) {
      if (threadIdx.x == 0) atomicOr(err, 1u);
      return off;
    }
    uint32_t pp[KPT];
    // mark: as a lambda, the register allocator keeps the kernel free of
    // spills (the same loop written inline spills 32 bytes per lane at
    // 64 VGPRs, and dedup takes 3.35 instead of 3.27 ms)
    auto mark = [&]() {
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        if ((alive >> s) & 1) {
          const uint32_t pos = bbh_pos(h1[s] + (uint64_t)l * bbh_d(h1[s]), (uint32_t)n);
          pp[s] = pos;
          // 32-bit views of the bitmaps: bit pos is bit pos & 31 of word pos >> 5
          const uint32_t bit = 1u << (pos & 31);
          const uint32_t old = atomicOr((uint32_t*)seen + (pos >> 5), bit);
          if (old & bit) atomicOr((uint32_t*)col + (pos >> 5), bit);
        }
      }
    };
    mark();
    __syncthreads();
    const uint32_t nw = (uint32_t)(n >> 6);
    for (uint32_t i = threadIdx.x; i < nw; i += WG) region[(off >> 6) + i] = seen[i] & ~col[i];
    const uint32_t nw2 = (uint32_t)(bbh_level_n(bbh_next_t(t)) >> 6);
    for (uint32_t i = threadIdx.x; i < nw2; i += WG) {
      nseen[i] = 0;
      ncol[i] = 0;
    }
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if (s >= ns) break;
      if ((alive >> s) & 1) {
        const uint32_t pos = pp[s];
        if (((((const uint32_t*)seen)[pos >> 5] & ~((const uint32_t*)col)[pos >> 5]) >> (pos & 31)) & 1) {
          alive &= ~(1u << s);
          ppos[s * WG + threadIdx.x] = (uint16_t)(off + pos);
        }
      }
    }
    const uint32_t wl = wave_sum4((uint32_t)__popc(alive));
    if (__lane_id() == 0) sh->left[l & 1][wv] = wl;
    __syncthreads();
    total = 0;
#pragma unroll
    for (int w = 0; w < WG / 64; w++) total += sh->left[l & 1][w];
    uint64_t* ts = seen;
    seen = nseen;
    nseen = ts;
    ts = col;
    col = ncol;
    ncol = ts;
    off += n;
    t = bbh_next_t(t);
    l++;
  }
  if (total == 0) return off;
  // ---- tail: list the remaining keys (wave counts from the last level, or
  // a fresh count when the segment starts small), then wave 0 alone.
  if (l == 0) {
    const uint32_t wl = wave_sum4((uint32_t)__popc(alive));
    if (__lane_id() == 0) sh->left[0][wv] = wl;
    __syncthreads();
  }
  const int lb = l == 0 ? 0 : ((l - 1) & 1);
  uint32_t base = 0;
  for (uint32_t w = 0; w < wv; w++) base += sh->left[lb][w];
  base += wave_prefix4((uint32_t)__popc(alive));
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    if (s >= ns) break;
    if ((alive >> s) & 1) {
      const uint64_t d = bbh_d(h1[s]);
      L.h[base] = h1[s] + (uint64_t)l * d;
      L.d[base] = d;
      L.e[base] = (uint16_t)(s * WG + threadIdx.x);
      base++;
    }
  }
  __syncthreads();
  if (wv == 0) {
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ULL << lane) - 1;
    uint32_t cnt = total;
    for (;;) {  // seen / col are clear for this level's n
      const uint64_t n = bbh_level_n(t);
      if (l >= BBH_MAX_LEVELS || off + n > max_bits) {
        if (lane == 0) atomicOr(err, 1u);
        break;
      }
      for (uint32_t i = lane; i < cnt; i += 64) {
        const uint32_t pos = bbh_pos(L.h[i], (uint32_t)n);
        const uint32_t bit = 1u << (pos & 31);
        const uint32_t old = atomicOr((uint32_t*)seen + (pos >> 5), bit);
        if (old & bit) atomicOr((uint32_t*)col + (pos >> 5), bit);
      }
      wave_lds_order();
      const uint32_t nw = (uint32_t)(n >> 6);
      for (uint32_t i = lane; i < nw; i += 64) region[(off >> 6) + i] = seen[i] & ~col[i];
      // place; unplaced keys move to the front with the next level's state
      // (a key's new index never exceeds its old one, and each 64-key chunk
      // is read before it is written)
      uint32_t kept = 0;
      for (uint32_t c0 = 0; c0 < cnt; c0 += 64) {
        const uint32_t i = c0 + lane;
        const bool act = i < cnt;
        uint64_t hh = 0, dd = 0;
        uint32_t ee = 0, pos = 0;
        bool placed = false;
        if (act) {
          hh = L.h[i];
          dd = L.d[i];
          ee = L.e[i];
          pos = bbh_pos(hh, (uint32_t)n);
          placed = ((((const uint32_t*)seen)[pos >> 5] & ~((const uint32_t*)col)[pos >> 5]) >> (pos & 31)) & 1;
        }
        if (placed) ppos[ee] = (uint16_t)(off + pos);
        const bool keep = act && !placed;
        const uint64_t km = __ballot(keep);
        wave_lds_order();
        if (keep) {
          const uint32_t j = kept + (uint32_t)__popcll(km & below);
          L.h[j] = hh + dd;
          L.d[j] = dd;
          L.e[j] = (uint16_t)ee;
        }
        kept += (uint32_t)__popcll(km);
      }
      wave_lds_order();
      const uint32_t nw2 = (uint32_t)(bbh_level_n(bbh_next_t(t)) >> 6);
      for (uint32_t i = lane; i < nw2; i += 64) {
        seen[i] = 0;
        col[i] = 0;
      }
      wave_lds_order();
      off += n;
      t = bbh_next_t(t);
      l++;
      cnt = kept;
      if (cnt == 0) break;
    }
    if (lane == 0) sh->off = off;
  }
  __syncthreads();
  return sh->off;
}

// The dedup holds each u64 key as bbh_h1(key), a bijection (an odd
// multiplier after an xor): the occurrence filter, the hash table and the
// BBHash levels then share one hash per key, and two held values are equal
// exactly when their keys are. 16-byte keys are held as they are (their h1
// is not one to one) and hashed by seg_hash where the filter and the table
// need bits.
__device__ __forceinline__ uint64_t dd_hold(uint64_t k) { return bbh_h1(k); }
__device__ __forceinline__ Key128 dd_hold(const Key128& k) { return k; }
__device__ __forceinline__ uint64_t dd_bits(uint64_t h, int) { return h; }
__device__ __forceinline__ uint64_t dd_bits(const Key128& k, int sb) { return seg_hash(k) << sb; }
__device__ __forceinline__ uint64_t dd_h1(uint64_t h) { return h; }
__device__ __forceinline__ uint64_t dd_h1(const Key128& k) { return bbh_h1(k); }
// held value against the key of an occurrence in occ
__device__ __forceinline__ bool dd_same(uint64_t h, uint64_t k) { return bbh_h1(k) == h; }
__device__ __forceinline__ bool dd_same(const Key128& h, const Key128& k) { return key_eq(h, k); }

// Position of a held key in the dedup's occurrence filter of FILT_BITS bits
// (the top 16 of its bits), its slot in the segment's hash table of 2^tb
// entries (the next tb bits) and its FPB-bit fingerprint (the FPB after).
constexpr int FILT_BITS = 1 << 16;
template <typename HeldT>
__device__ __forceinline__ uint32_t filt_pos(const HeldT& k, int sb) {
  return (uint32_t)(dd_bits(k, sb) >> 48);
}
template <int FPB, typename HeldT>
__device__ __forceinline__ void tab_slot(const HeldT& k, int sb, int tb, uint32_t& slot, uint32_t& fp) {
  const uint64_t h = dd_bits(k, sb) << 16;
  slot = (uint32_t)(h >> (64 - tb));
  fp = (uint32_t)((h << tb) >> (64 - FPB));
}
constexpr int ceil_log2(int x) { return x <= 1 ? 0 : 1 + ceil_log2((x + 1) / 2); }

// Work-groups per CU the dedup kernel is compiled for: 8 with u64 keys (64
// VGPRs and 8 x 20 KiB of LDS fill a CU; 3.99 -> 3.60 ms on the 300 Mbp
// genome against 7 per CU at 70 VGPRs); 6 with 16-byte keys (77 VGPRs;
// 5.22 -> 4.74 ms at k = 63 against the compiler's 5 per CU, and 7 per CU
// spills). The u64 build spills 8 bytes per lane since the level bitmaps are
// addressed as 32-bit words (which saved more than the spill costs); a
// spill of 80 bytes per lane cost 15% of the kernel's time.
template <typename KeyT>
struct DdOccupancy {
  static constexpr int WGS_PER_CU = sizeof(KeyT) == 8 ? 8 : 6;
};

// Dedup of one segment and its BBHash, WG threads holding KPT occurrences
// each (element e = slot * WG + thread), at most CAP = WG * KPT per segment.
// Each occurrence carries its 10-bit observation mask; the masks of equal
// keys are OR-joined (the DFA join of classify.rs:147-518). An occurrence
// filter (2^16 positions, seen / shared bits in LDS) settles every
// occurrence alone at its position: its key occurs once. The rest (repeated
// keys and ~n / 2^16 false sharers) are deduplicated with an LDS hash table
// of 2 CAP entries (linear probing; entries = representative occurrence + 1
// in EB = log2(2 CAP) bits | joined mask << EB | (22 - EB)-bit key
// fingerprint << (EB + 10); keys are compared, from occ, only when the
// fingerprints match); which occurrence represents a key depends on timing,
// but the set of keys, their joined masks and the BBHash built over them do
// not. Output: occ[start + e] = the key of occurrence e (equal keys land on
// one MPHF value, so the dense key array is built from these on request,
// k_p3_keys); the joined masks in MPHF value order at vpad[start + local
// rank], their count in ucount, the BBHash blocks in the segment's padded
// slot of blocks_pad and their count in nblk. Segments with more
// occurrences than CAP, or a record longer than the decode's owner map, go
// to over_list (count flags[0], flags[4] with LIST); segments whose distinct keys exceed CAP go
// to big_list (BBHash built by k_bbh_big). Keys of <= 12 bits (k <= 6) use
// a direct-address mask table instead, which takes any number of
// occurrences (not with LIST: that variant sees only the hash path's
// oversized segments).
template <typename KeyT, int WG, int KPT, bool LIST>
__device__ __forceinline__ void dd_segment(const uint64_t seg, KeyT* __restrict__ occ, uint16_t* __restrict__ obs,
                                           const uint64_t* __restrict__ srec, const uint64_t* __restrict__ roff,
                                           const uint64_t* __restrict__ seg_off, int K, int sb, uint64_t* __restrict__ ucount,
                                           uint64_t* __restrict__ nblk, uint32_t* __restrict__ over_list, unsigned int* __restrict__ flags,
                                           uint32_t* __restrict__ big_list,
                                           uint64_t* __restrict__ blocks_pad, uint16_t* __restrict__ vpad) {
  constexpr int CAP = WG * KPT;
  constexpr int TAB_BITS = ceil_log2(2 * CAP);
  constexpr int EB = TAB_BITS;      // entry: element + 1
  constexpr int FPB = 22 - EB;      // entry: fingerprint bits above the 10-bit mask
  constexpr uint32_t EMASK = (1u << EB) - 1;
  constexpr int SEEN_W = (2 * CAP) / 64;  // level 0 has roundup64(2u) <= 2 CAP bits
  constexpr int REG_W = (4 * CAP + 8192) / 64;
  constexpr int TAB_BYTES = 4 << TAB_BITS;
  // tab, byte offsets: hash table (or direct-address mask table) over all of
  // it; then ppos (u16) [0, 2 CAP), word ranks, seen / collision bitmaps
  // and the tail list
  constexpr int T_WRANK = 2 * CAP, T_SEEN = T_WRANK + 4 * REG_W, T_LIST = T_SEEN + 4 * 8 * SEEN_W;
  constexpr int T_END = T_LIST + TAIL_KEYS * 18;
  // the joined masks by element, from the join until they are staged in rank
  // order (the level structures end below them)
  constexpr int T_MJ = TAB_BYTES - 2 * CAP;
  // aux: region (level words), the block-scan and level-count words behind
  // it; then the staged masks (all of it)
  constexpr int A_SH = REG_W * 8, A_BSH = A_SH + 8 * (WG / 64);
  constexpr int A_END = A_BSH + (int)sizeof(BbhShared<WG>);
  constexpr int AUX_BYTES = A_END > CAP * 2 ? A_END : CAP * 2;
  static_assert(2 * CAP == (1 << TAB_BITS) && CAP < (1 << EB) && FPB >= 8, "hash table entries hold element + 1, mask and fingerprint");
  static_assert(2 * (FILT_BITS / 8) <= TAB_BYTES, "the occurrence filter fits the table space");
  static_assert(T_END <= T_MJ && (1 << BM_MAX_BITS) * 4 <= TAB_BYTES, "LDS reuse fits");
  static_assert(REG_W / 7 + 2 <= SEEN_W && CAP <= 65536 && REG_W * 64 <= 65536 && T_SEEN % 8 == 0 && T_LIST % 8 == 0,
                "LDS reuse fits");
  __shared__ __align__(16) uint8_t tab_b[TAB_BYTES];
  __shared__ __align__(16) uint8_t aux_b[AUX_BYTES];
  uint64_t* seenA = (uint64_t*)(tab_b + T_SEEN);
  uint64_t* colA = seenA + SEEN_W;
  uint64_t* seenB = colA + SEEN_W;
  uint64_t* colB = seenB + SEEN_W;
  uint64_t* sh8 = (uint64_t*)(aux_b + A_SH);
  BbhShared<WG>& bsh = *(BbhShared<WG>*)(aux_b + A_BSH);
  uint32_t* tab = (uint32_t*)tab_b;
  uint16_t* mj = (uint16_t*)(tab_b + T_MJ);
  constexpr int RW = sizeof(KeyT) == 8 ? 2 : 4;
  const int K2 = 2 * K;
  unsigned int* n_over = flags + (LIST ? 4 : 0);
  unsigned int* n_big = flags + 1;
  unsigned int* err = flags + 2;
  const uint64_t start = seg_off[seg];
  const uint64_t n = seg_off[seg + 1] - start;
  if (n == 0) {
    if (threadIdx.x == 0) {
      ucount[seg] = 0;
      nblk[seg] = 0;
    }
    return;
  }
  KeyT key[KPT];  // held values (dd_hold)
  uint32_t msk[KPT];
  uint32_t alive = 0;
  uint64_t u = 0;  // distinct keys
  // slots that hold elements (uniform): the slot loops stop there, so a
  // segment of n elements costs ceil(n / WG) slots of work, not KPT
  int ns = KPT;
  if (!LIST && sizeof(KeyT) == 8 && K2 <= BM_MAX_BITS) {
    // ---- direct-address table of masks over all 2^K2 keys: any number of
    // occurrences; distinct keys come out sorted at occ / obs[start...].
    const uint32_t nt = 1u << K2;
    for (uint32_t i = threadIdx.x; i < nt; i += WG) tab[i] = 0;
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < n; i += WG) atomicOr(&tab[key_lo(occ[start + i])], (uint32_t)obs[start + i]);
    __syncthreads();
    const uint32_t per = (nt + WG - 1) / WG;
    const uint32_t t0 = threadIdx.x * per;
    uint64_t c = 0;
    for (uint32_t q = 0; q < per; q++)
      if (t0 + q < nt && tab[t0 + q]) c++;
    uint64_t tot;
    uint64_t o = block_excl_scan<WG>(c, sh8, &tot);
    for (uint32_t q = 0; q < per; q++) {
      if (t0 + q >= nt) break;
      uint32_t m = tab[t0 + q];
      if (!m) continue;
      if constexpr (sizeof(KeyT) == 8) occ[start + o] = (KeyT)(t0 + q);
      obs[start + o] = (uint16_t)m;
      o++;
    }
    u = tot;
    if (threadIdx.x == 0) ucount[seg] = u;
    if (seg_is_big<KeyT, CAP>(n, u, K2)) {
      if (threadIdx.x == 0) {
        big_list[atomicAdd(n_big, 1u)] = (uint32_t)seg;
        nblk[seg] = 0;
      }
      return;
    }
    __threadfence_block();
    __syncthreads();
    ns = (int)((u + WG - 1) / WG);
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if (s >= ns) break;
      const uint64_t e = (uint64_t)s * WG + threadIdx.x;
      if (e < u) {
        key[s] = dd_hold(occ[start + e]);
        mj[e] = obs[start + e];
        alive |= 1u << s;
      }
    }
    for (uint64_t e = u + threadIdx.x; e < n; e += WG) occ[start + e] = key_pad<KeyT>();
  } else if (n > (uint64_t)CAP) {
    if (threadIdx.x == 0) {
      over_list[atomicAdd(n_over, 1u)] = (uint32_t)seg;
      ucount[seg] = 0;
      nblk[seg] = 0;
    }
    return;
  } else {
    // ---- hash-table dedup: table of 2^tb >= 2n entries (load <= 1/2).
    int tb = 6;
    while ((1ull << tb) < 2 * n) tb++;
    const uint32_t tmask = (1u << tb) - 1;
    uint32_t pend = 0;
    ns = (int)((n + WG - 1) / WG);
    // the segment's k-mers and observation masks straight from its records:
    // rounds of RR records staged in LDS (in the table's space, before the
    // table is cleared), each record cut into pieces of up to DEC_PK k-mers,
    // one lane per piece through an LDS owner map: the piece's first k-mer is
    // cut out of the record, the others rolled. Element e is k-mer e of the
    // segment in record order. Every key goes to occ (the table reads a
    // representative's key there when fingerprints match, and the dense key
    // array is built from it on request), every mask to mj; the slots then
    // load their elements back (a decode per slot, unrolled, would hold every
    // slot's record words at once: 86 VGPRs instead of 64).
    {
      constexpr int RR = RW == 2 ? 256 : 128;
      constexpr int OWN = OwnKmers<RW>::N;
      constexpr int PK = DEC_PK;
      constexpr int OWNP = (OWN + PK - 1) / PK;  // pieces per record at most
      static_assert(RR * RW * 8 + RR * 4 + RR * OWNP <= T_MJ, "decode scratch fits below the joined masks");
      uint64_t* srl = (uint64_t*)tab_b;
      uint16_t* kofs = (uint16_t*)(tab_b + RR * RW * 8);  // per staged record: first k-mer in the round
      uint16_t* pofs = kofs + RR;                          //   first piece in the round
      uint8_t* ownp = (uint8_t*)(pofs + RR);               // per piece: its record
      const uint64_t kmask = K < 32 ? (1ULL << (2 * K)) - 1 : ~0ULL;
      const uint64_t ra = roff[seg], rb = roff[seg + 1];
      uint64_t done = 0;
      for (uint64_t rr = ra; rr < rb; rr += RR) {
        const uint64_t r = rr + threadIdx.x;
        uint32_t nk = 0;
        if ((int)threadIdx.x < RR && r < rb) {
#pragma unroll
          for (int t = 0; t < RW; t++) srl[threadIdx.x * RW + t] = srec[r * RW + t];
          nk = (uint32_t)rec_kmers_at<RW>(srec, r, K);
        }
        const uint32_t np = (nk + PK - 1) / PK;
        // one scan for the k-mer offsets (bits 0-15), the piece offsets (bits
        // 16-31) and the records longer than the owner map (counted in the
        // high half; such a segment takes the global-memory path)
        uint64_t T;
        const uint64_t o = block_excl_scan<WG>((uint64_t)nk | ((uint64_t)np << 16) | ((uint64_t)(nk > (uint32_t)OWN) << 32), sh8, &T);
        if (T >> 32) {
          if (threadIdx.x == 0) {
            over_list[atomicAdd(n_over, 1u)] = (uint32_t)seg;
            ucount[seg] = 0;
            nblk[seg] = 0;
          }
          return;
        }
        if (nk) {
          kofs[threadIdx.x] = (uint16_t)o;
          pofs[threadIdx.x] = (uint16_t)(o >> 16);
          for (uint32_t q = 0; q < np; q++) ownp[(uint32_t)(o >> 16) + q] = (uint8_t)threadIdx.x;
        }
        __syncthreads();
        const uint32_t TP = (uint32_t)(T >> 16) & 0xFFFFu;
        for (uint32_t p = threadIdx.x; p < TP; p += WG) {
          const uint32_t rl = ownp[p];
          const int j0 = (int)(p - pofs[rl]) * PK;
          const uint64_t* w = srl + rl * RW;
          const uint64_t e0 = done + kofs[rl] + (uint32_t)j0;
          KeyT* ko = occ + start + e0;
          uint16_t* mo = mj + e0;
          auto out = [&](int i, const KeyT& c, uint32_t m) {
            ko[i] = c;
            mo[i] = (uint16_t)m;
          };
          if constexpr (RW == 2) {
            const int nr = rec2_kmers(w[1], K);
            rec2_piece<PK>(w[0], w[1], j0, nr - j0 < PK ? nr - j0 : PK, nr, K, kmask, out);
          } else {
            const Rec4 R{w[0], w[1], w[2], w[3]};
            const int nr = rec4_kmers(R.w3, K);
            rec4_piece<PK>(R, j0, nr - j0 < PK ? nr - j0 : PK, nr, K, out);
          }
        }
        __syncthreads();
        done += T & 0xFFFFu;
      }
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        const uint64_t e = (uint64_t)s * WG + threadIdx.x;
        const bool v = e < n;
        key[s] = v ? dd_hold(occ[start + e]) : KeyT{};
        msk[s] = v ? (uint32_t)mj[e] : 0u;
        pend |= (uint32_t)v << s;
      }
      __syncthreads();  // every element is loaded before the table (over mj) is cleared
    }
    // filter: an occurrence whose filter position no other occurrence of the
    // segment shares has a key no other occurrence has (equal keys share
    // their position), so it represents its key with its own mask; only the
    // others (~n / 2^16 of them, and every repeated key) go through the table
    {
      uint32_t* seenF = (uint32_t*)tab_b;
      uint32_t* colF = seenF + (FILT_BITS / 32);
      for (uint32_t i = threadIdx.x; i < FILT_BITS / 64; i += WG) ((uint4*)tab_b)[i] = make_uint4(0, 0, 0, 0);
      __syncthreads();
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        if ((pend >> s) & 1) {
          const uint32_t f = filt_pos(key[s], sb);
          const uint32_t bit = 1u << (f & 31);
          if (atomicOr(&seenF[f >> 5], bit) & bit) atomicOr(&colF[f >> 5], bit);
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        if ((pend >> s) & 1) {
          const uint32_t f = filt_pos(key[s], sb);
          if (!((colF[f >> 5] >> (f & 31)) & 1)) {
            alive |= 1u << s;
            pend &= ~(1u << s);
          }
        }
      }
      __syncthreads();  // the filter is read before the table is cleared over it
    }
    const uint32_t cand = pend;  // slots that go through the table
    for (uint32_t i = threadIdx.x; i <= tmask / 4; i += WG) ((uint4*)tab)[i] = make_uint4(0, 0, 0, 0);  // 2^tb >= 64 entries
    __syncthreads();
    uint32_t hs[KPT], ent[KPT];  // probe slot; this occurrence's entry
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if (s >= ns) break;
      uint32_t slot = 0, fp = 0;
      const bool v = (pend >> s) & 1;
      if (v) tab_slot<FPB>(key[s], sb, tb, slot, fp);
      hs[s] = v ? slot : 0u;
      ent[s] = v ? ((uint32_t)(s * WG + threadIdx.x + 1) | (msk[s] << EB) | (fp << (EB + 10))) : 0u;
    }
    while (pend) {
      uint32_t old[KPT];
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        if ((pend >> s) & 1) old[s] = atomicCAS(&tab[hs[s]], 0u, ent[s]);
      }
#pragma unroll
      for (int s = 0; s < KPT; s++) {
        if (s >= ns) break;
        if ((pend >> s) & 1) {
          if (old[s] == 0) {
            alive |= 1u << s;
            pend &= ~(1u << s);
          } else if ((old[s] >> (EB + 10)) == (ent[s] >> (EB + 10)) && dd_same(key[s], occ[start + (old[s] & EMASK) - 1])) {
            atomicOr(&tab[hs[s]], msk[s] << EB);
            pend &= ~(1u << s);
          } else {
            hs[s] = (hs[s] + 1) & tmask;
          }
        }
      }
    }
    __syncthreads();  // every mask is joined
    // each representative's joined mask: read from the table, then (the
    // table being free after the scan's barriers) kept by element in mj
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if (s >= ns) break;
      if ((alive & cand) >> s & 1) msk[s] = (tab[hs[s]] >> EB) & 0x3FFu;
    }
    uint64_t tot;
    block_excl_scan<WG>((uint64_t)__popc(alive), sh8, &tot);  // barriers: tab is free after it
    u = tot;
    if (threadIdx.x == 0) ucount[seg] = u;
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      if (s >= ns) break;
      if ((alive >> s) & 1) mj[s * WG + threadIdx.x] = (uint16_t)msk[s];
    }
  }
  // ---- BBHash of the segment, into its padded block slot.
  uint64_t* region = (uint64_t*)aux_b;
  uint16_t* ppos = (uint16_t*)tab_b;  // bit positions < REG_W * 64 <= 2^16
  uint32_t* wrank = (uint32_t*)(tab_b + T_WRANK);
  for (int i = threadIdx.x; i < REG_W; i += WG) region[i] = 0;
  for (int i = threadIdx.x; i < SEEN_W; i += WG) {
    seenA[i] = 0;
    colA[i] = 0;
  }
  __syncthreads();
  const uint64_t slot = pad_slot_blocks(n);
  const uint64_t max_bits = slot * BBH_BLOCK_BITS < (uint64_t)REG_W * 64 ? slot * BBH_BLOCK_BITS : (uint64_t)REG_W * 64;
  BbhList tl;
  tl.h = (uint64_t*)(tab_b + T_LIST);
  tl.d = tl.h + TAIL_KEYS;
  tl.e = (uint16_t*)(tl.d + TAIL_KEYS);
  uint64_t h1[KPT];
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    if (s >= ns) break;
    h1[s] = ((alive >> s) & 1) ? dd_h1(key[s]) : 0;
  }
  const uint64_t bits = bbh_levels<KPT, WG>(h1, alive, ns, u, seenA, colA, seenB, colB, region, max_bits, ppos, tl, &bsh, err);
  const uint64_t nb = (bits + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS;  // <= ceil(REG_W*64/448) < WG
  __syncthreads();
  // blocks {rank before, 7 words} and the word ranks for the local ranks
  // below, by wave 0 (nb <= ceil(REG_W * 64 / 448) <= 64 blocks)
  static_assert((REG_W * 64 + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS <= 64, "a segment's blocks fit one wave");
  if (threadIdx.x < 64) {
    const uint64_t j = threadIdx.x;
    uint64_t w[7];
    uint64_t pc = 0;
#pragma unroll
    for (int q = 0; q < 7; q++) {
      const uint64_t x = j * 7 + q;
      w[q] = (j < nb && x < (uint64_t)REG_W) ? region[x] : 0;
      pc += (uint64_t)__popcll(w[q]);
    }
    const uint64_t rk = wave_incl_scan_u64(pc) - pc;
    if (j < nb) {
      uint64_t* B = blocks_pad + (pad_slot_off(start, seg) + j) * 8;
      B[0] = rk;
      uint64_t run = rk;
#pragma unroll
      for (int q = 0; q < 7; q++) {
        B[1 + q] = w[q];
        if (j * 7 + q < (uint64_t)REG_W) wrank[j * 7 + q] = (uint32_t)run;
        run += (uint64_t)__popcll(w[q]);
      }
    }
  }
  if (threadIdx.x == 0) nblk[seg] = nb;
  __syncthreads();
  // local rank of each distinct key, then keys and masks staged in rank order
  uint32_t rr[KPT];
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    rr[s] = ~0u;
    if (s >= ns) continue;
    if ((alive >> s) & 1) {
      const uint32_t p = ppos[s * WG + threadIdx.x];
      const uint32_t wi = p >> 6;
      const uint32_t r = p < bits ? wrank[wi] + (uint32_t)__popcll(region[wi] & ((1ULL << (p & 63)) - 1)) : ~0u;
      if (r < u) rr[s] = r;
      else atomicOr(err, 8u);
    }
  }
  __syncthreads();
  uint16_t* stage_m = (uint16_t*)aux_b;
#pragma unroll
  for (int s = 0; s < KPT; s++) {
    if (s >= ns) break;
    if (rr[s] != ~0u) stage_m[rr[s]] = mj[s * WG + threadIdx.x];
  }
  __syncthreads();
  for (uint64_t i = threadIdx.x; i < u; i += WG) vpad[start + i] = stage_m[i];
}

template <typename KeyT>
__global__ void __launch_bounds__(DD_WG, DdOccupancy<KeyT>::WGS_PER_CU) k_p3_dedup(KeyT* __restrict__ occ, uint16_t* __restrict__ obs,
                                                    const uint64_t* __restrict__ srec, const uint64_t* __restrict__ roff,
                                                    const uint64_t* __restrict__ seg_off, uint64_t n_segs, int K, int sb,
                                                    uint64_t* __restrict__ ucount, uint64_t* __restrict__ nblk,
                                                    uint32_t* __restrict__ over_list, unsigned int* __restrict__ flags,
                                                    uint32_t* __restrict__ big_list, uint64_t* __restrict__ blocks_pad,
                                                    uint16_t* __restrict__ vpad) {
  const uint64_t seg = blockIdx.x;
  if (seg >= n_segs || flags[3]) return;  // flags[3]: the occurrence counts disagreed with W (k_p3_gate)
  dd_segment<KeyT, DD_WG, DD_KPT, false>(seg, occ, obs, srec, roff, seg_off, K, sb, ucount, nblk, over_list, flags, big_list,
                                         blocks_pad, vpad);
}

// The segments k_p3_dedup put in over_list (flags[0] of them), DD2_WG
// threads each, one work-group per entry over a grid of at most DD2_GRID
// (the list's length is known on the device only; a grid-stride loop around
// the segment's dedup would keep its loop-invariant values live through it
// and spill at any occupancy). Entries beyond the grid, segments above
// DD2_WG * DD_KPT occurrences and those with a record longer than the decode's owner
// map go on to over2 (count flags[4]) for k_p3_bigsort.
constexpr unsigned DD2_GRID = 4096;
template <typename KeyT>
__global__ void __launch_bounds__(DD2_WG, 4) k_p3_dedup_list(KeyT* __restrict__ occ, uint16_t* __restrict__ obs,
                                                             const uint64_t* __restrict__ srec, const uint64_t* __restrict__ roff,
                                                             const uint64_t* __restrict__ seg_off, int K, int sb,
                                                             uint64_t* __restrict__ ucount, uint64_t* __restrict__ nblk,
                                                             const uint32_t* __restrict__ over_list, uint32_t* __restrict__ over2,
                                                             unsigned int* __restrict__ flags, uint64_t* __restrict__ blocks_pad,
                                                             uint16_t* __restrict__ vpad) {
  if (flags[3]) return;
  const uint32_t n_list = flags[0];
  if (threadIdx.x == 0)
    for (uint32_t i = blockIdx.x + gridDim.x; i < n_list; i += gridDim.x) over2[atomicAdd(flags + 4, 1u)] = over_list[i];
  if (blockIdx.x >= n_list) return;
  dd_segment<KeyT, DD2_WG, DD_KPT, true>(over_list[blockIdx.x], occ, obs, srec, roff, seg_off, K, sb, ucount, nblk, over2, flags,
                                         nullptr, blocks_pad, vpad);
}

// Oversized segments: one work-group per segment runs a stable LSD radix
// sort of (key, mask) through global memory (ping-pong with aux), tile by
// tile, then deduplicates (OR-joining masks) into occ/obs[start...]. Their
// MPHF is built by k_bbh_big (big_list).
template <typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_p3_bigsort(KeyT* __restrict__ occ, uint16_t* __restrict__ obs, KeyT* __restrict__ aux,
                                                      uint16_t* __restrict__ aux_obs, const uint64_t* __restrict__ seg_off,
                                                      const uint32_t* __restrict__ over_list, uint64_t* __restrict__ ucount,
                                                      unsigned int* __restrict__ flags, uint32_t* __restrict__ big_list) {
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
  uint16_t* msrc = obs + start;
  uint16_t* mdst = aux_obs + start;
  const KeyT k0 = src[0];
  uint64_t xl = 0, xh = 0;
  for (uint64_t i = threadIdx.x; i < n; i += RK_WG) key_xor_or(xl, xh, src[i], k0);
  block_or2(xl, xh, sh8);
  const int vb = varying_bits(xl, xh);
  const int passes = (vb + 7) >> 3;
  KeyT key[KPT];
  uint32_t msk[KPT];
  uint32_t rank[KPT];
  for (int p = 0; p < passes; p++) {
    const int sh = 8 * p;
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
          msk[s] = e < tn ? msrc[t0 + e] : 0u;
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
            uint64_t at = dbase[d] + (local - cnt[d]);  // cnt[d]: tile start of digit d
            dst[at] = key[s];
            mdst[at] = (uint16_t)msk[s];
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
    uint16_t* mt = msrc;
    msrc = mdst;
    mdst = mt;
    __threadfence_block();
    __syncthreads();
  }
  // Dedup sorted src into occ/obs[start...] (src may be occ itself: writes
  // trail reads). Thread t owns elements [t*kpt, (t+1)*kpt) of each tile; a
  // run head reads ahead through its whole run to join the masks. Positions
  // before the current tile may already hold uniques, so the key preceding a
  // tile is carried over from the previous tile.
  KeyT* out = occ + start;
  uint16_t* mout = obs + start;
  uint64_t u = 0;
  for (uint64_t t0 = 0; t0 < n; t0 += CAP) {
    uint64_t tn = n - t0 < (uint64_t)CAP ? n - t0 : (uint64_t)CAP;
    int kpt = (int)((tn + RK_WG - 1) / RK_WG);
    const KeyT prevc = carry;
    uint64_t c = 0;
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      uint64_t e = (uint64_t)threadIdx.x * kpt + s;
      rank[s] = 0;
      if (s < kpt && e < tn) {
        key[s] = src[t0 + e];
        bool head = (t0 + e == 0) || !key_eq(key[s], e == 0 ? prevc : src[t0 + e - 1]);
        if (head) {
          uint32_t m = msrc[t0 + e];
          for (uint64_t f = t0 + e + 1; f < n && key_eq(src[f], key[s]); f++) m |= msrc[f];
          msk[s] = m;
          rank[s] = 1;
          c++;
        }
      }
    }
    uint64_t tot;
    uint64_t o = block_excl_scan_u64(c, sh8, &tot);  // all reads of this tile precede its barriers
#pragma unroll
    for (int s = 0; s < KPT; s++) {
      uint64_t e = (uint64_t)threadIdx.x * kpt + s;
      if (rank[s]) {
        out[u + o] = key[s];
        mout[u + o] = (uint16_t)msk[s];
        o++;
      }
      if (s < kpt && e + 1 == tn) carry = key[s];
    }
    u += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ucount[seg] = u;
    big_list[atomicAdd(flags + 1, 1u)] = seg;
  }
}

// ------------------------------------------------------------------ compaction
// A segment's BBHash blocks go from its padded slot to the dense block order
// (blk_off), one wave per segment (CP_WG / 64 segments per work-group),
// copied 8 loads deep per lane, loads before stores, so that a wave keeps
// several requests in flight (with one work-group per segment the copies
// were bound by work-group turnover).
constexpr int CP_WG = 256;
constexpr int CP_DEPTH = 8;
template <typename T>
__device__ __forceinline__ void wave_copy(const T* __restrict__ src, T* __restrict__ dst, uint64_t n) {
  uint64_t i = __lane_id();
  for (; i + (CP_DEPTH - 1) * 64 < n; i += CP_DEPTH * 64) {
    T v[CP_DEPTH];
#pragma unroll
    for (int q = 0; q < CP_DEPTH; q++) v[q] = src[i + q * 64];
#pragma unroll
    for (int q = 0; q < CP_DEPTH; q++) dst[i + q * 64] = v[q];
  }
  for (; i < n; i += 64) dst[i] = src[i];
}
template <typename KeyT>
__global__ void __launch_bounds__(CP_WG) k_p3_compact(const uint64_t* __restrict__ seg_off, const uint64_t* __restrict__ blk_off,
                                                      const uint64_t* __restrict__ blocks_pad, uint64_t n_segs,
                                                      uint64_t* __restrict__ blocks) {
  const uint64_t seg = (uint64_t)blockIdx.x * (CP_WG / 64) + (threadIdx.x >> 6);
  if (seg >= n_segs) return;
  // the segment's words loaded together: no test between them (an empty
  // segment copies nothing)
  const uint64_t s0 = seg_off[seg], b0 = blk_off[seg], b1 = blk_off[seg + 1];
  wave_copy(blocks_pad + pad_slot_off(s0, seg) * 8, blocks + b0 * 8, (b1 - b0) * 8);
}

// The dense key array on request: every key P3 left in its segment's run of
// kout (pads skipped) goes to keys[its MPHF value], the local rank coming
// from the segment's (dense) BBHash blocks.
template <typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_p3_keys(const KeyT* __restrict__ kout, const uint64_t* __restrict__ seg_off,
                                                   const uint64_t* __restrict__ ucount, const uint64_t* __restrict__ key_off,
                                                   const uint64_t* __restrict__ blk_off, const uint64_t* __restrict__ blocks,
                                                   uint64_t n_segs, KeyT* __restrict__ keys, unsigned int* __restrict__ err) {
  const uint64_t seg = blockIdx.x;
  if (seg >= n_segs) return;
  const uint64_t s0 = seg_off[seg], n = seg_off[seg + 1] - s0, u = ucount[seg], ko = key_off[seg];
  const uint64_t* base = blocks + blk_off[seg] * 8;
  for (uint64_t e = threadIdx.x; e < n; e += RK_WG) {
    const KeyT k = kout[s0 + e];
    if (key_eq(k, key_pad<KeyT>())) continue;
    const uint64_t r = mphf_local_rank<KeyT>(base, u, k);
    if (r < u) keys[ko + r] = k;
    else atomicOr(err, 1u);
  }
}

// ------------------------------------------------------------------ BBHash of big segments
// Segments above the register capacity, one work-group each: the same levels
// as bbh_levels with keys read from occ (sorted, distinct, in place). LDSB:
// the seen / collision / alive bitmaps sit in LDS (segments of <= MED_CAP
// distinct keys); otherwise in global scratch (seen / collision words of the
// segment at 4*floor(seg_off/64) + 8*seg, alive bits per occurrence position)
// with device-scope atomics. Blocks go to the padded slot; each key's mask
// is written to vpad[seg_off + local rank], the keys (then pads up to the
// segment's occurrence count) to kout[seg_off ...].
constexpr uint64_t MED_CAP = 1ULL << 17;
constexpr int MED_W = (int)(MED_CAP / 64);
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
  return __hip_atomic_load((uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool LDSB, typename KeyT>
__global__ void __launch_bounds__(RK_WG) k_bbh_big(const KeyT* __restrict__ occ, const uint16_t* __restrict__ obs,
                                                   uint16_t* __restrict__ vpad, const uint64_t* __restrict__ seg_off,
                                                   const uint64_t* __restrict__ ucount, const uint32_t* __restrict__ big_list,
                                                   uint64_t* __restrict__ nblk, uint64_t* __restrict__ blocks_pad,
                                                   uint64_t* __restrict__ scratch, uint64_t* __restrict__ alive_g,
                                                   KeyT* __restrict__ kout, unsigned int* __restrict__ err) {
  __shared__ uint64_t sh8[8];
  __shared__ unsigned long long carry;
  const uint32_t seg = big_list[blockIdx.x];
  const uint64_t so = seg_off[seg];
  const uint64_t n_occ = seg_off[seg + 1] - so;
  const uint64_t D = ucount[seg];
  if (LDSB != (D <= MED_CAP)) return;  // the other variant builds it
  const KeyT* keys = occ + so;
  uint64_t* seen;
  uint64_t* col;
  uint64_t* alive;
  uint64_t abase;  // bit index of key 0 in alive
  if constexpr (LDSB) {
    __shared__ uint64_t l_alive[MED_W];
    __shared__ uint64_t l_seen[2 * MED_W + 1];
    __shared__ uint64_t l_col[2 * MED_W + 1];
    seen = l_seen;
    col = l_col;
    alive = l_alive;
    abase = 0;
    for (uint64_t i = threadIdx.x; i < (D + 63) / 64; i += RK_WG) {
      uint64_t rem = D - 64 * i;
      alive[i] = rem >= 64 ? ~0ULL : ((1ULL << rem) - 1);
    }
  } else {
    seen = scratch + 4 * (so / 64) + 8 * (uint64_t)seg;
    col = seen + (D / 32 + 2);
    alive = alive_g;
    abase = so;
  }
  auto ld = [&](const uint64_t* p) -> uint64_t {
    if constexpr (LDSB) return *p;
    else return ld_agent(p);
  };
  uint64_t* region = blocks_pad + pad_slot_off(so, seg) * 8;
  const uint64_t slot = pad_slot_blocks(n_occ);
  uint64_t off = 0, t = D, r = D;
  int l = 0;
  while (r > 0) {
    const uint64_t n = bbh_level_n(t);
    if (l >= BBH_MAX_LEVELS || (n >> 32) != 0 || (off + n + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS > slot) {
      if (threadIdx.x == 0) atomicOr(err, 2u);
      break;
    }
    const uint64_t nw = n >> 6;
    for (uint64_t i = threadIdx.x; i < nw; i += RK_WG) {
      seen[i] = 0;
      col[i] = 0;
    }
    if constexpr (!LDSB) __threadfence();
    __syncthreads();
    for (uint64_t e = threadIdx.x; e < D; e += RK_WG) {
      const uint64_t gi = abase + e;
      if (!((ld(&alive[gi >> 6]) >> (gi & 63)) & 1)) continue;
      const uint64_t h1 = bbh_h1(keys[e]);
      const uint64_t pos = bbh_pos(h1 + (uint64_t)l * bbh_d(h1), (uint32_t)n);
      unsigned long long bit = 1ULL << (pos & 63);
      unsigned long long old = atomicOr((unsigned long long*)&seen[pos >> 6], bit);
      if (old & bit) atomicOr((unsigned long long*)&col[pos >> 6], bit);
    }
    if constexpr (!LDSB) __threadfence();
    __syncthreads();
    for (uint64_t i = threadIdx.x; i < nw; i += RK_WG) {
      uint64_t q = (off >> 6) + i;
      region[(q / 7) * 8 + 1 + (q % 7)] = ld(&seen[i]) & ~ld(&col[i]);
    }
    uint64_t cnt = 0;
    for (uint64_t e = threadIdx.x; e < D; e += RK_WG) {
      const uint64_t gi = abase + e;
      if (!((ld(&alive[gi >> 6]) >> (gi & 63)) & 1)) continue;
      const uint64_t h1 = bbh_h1(keys[e]);
      const uint64_t pos = bbh_pos(h1 + (uint64_t)l * bbh_d(h1), (uint32_t)n);
      if (((ld(&seen[pos >> 6]) & ~ld(&col[pos >> 6])) >> (pos & 63)) & 1)
        atomicAnd((unsigned long long*)&alive[gi >> 6], ~(1ULL << (gi & 63)));
      else
        cnt++;
    }
    if constexpr (!LDSB) __threadfence();
    block_excl_scan_u64(cnt, sh8, &r);  // barriers inside
    off += n;
    t = bbh_next_t(t);
    l++;
  }
  const uint64_t nb = (off + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS;
  // zero the tail of the last block
  for (uint64_t q = (off >> 6) + threadIdx.x; q < nb * 7; q += RK_WG) region[(q / 7) * 8 + 1 + (q % 7)] = 0;
  __threadfence();
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint64_t j0 = 0; j0 < nb; j0 += RK_WG) {
    uint64_t j = j0 + threadIdx.x;
    uint64_t pc = 0;
    if (j < nb)
      for (int q = 0; q < 7; q++) pc += (uint64_t)__popcll(ld_agent(&region[j * 8 + 1 + q]));
    uint64_t tot;
    uint64_t rk = block_excl_scan_u64(pc, sh8, &tot);
    if (j < nb) region[j * 8] = carry + rk;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) nblk[seg] = nb;
  __threadfence();
  __syncthreads();
  for (uint64_t e = threadIdx.x; e < D; e += RK_WG) {
    const KeyT kv = keys[e];
    uint64_t rk = mphf_local_rank<KeyT>(region, D, kv);
    kout[so + e] = kv;
    if (rk >= D) atomicOr(err, 4u);
    else vpad[so + rk] = obs[so + e];
  }
  for (uint64_t e = D + threadIdx.x; e < n_occ; e += RK_WG) kout[so + e] = key_pad<KeyT>();
}

// ------------------------------------------------------------------ launchers
static inline unsigned grid_for(uint64_t n, unsigned wg, unsigned cap) {
  uint64_t g = (n + wg - 1) / wg;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

void launch_p3_rcount(hipStream_t st, bool wide, const uint64_t* rec, const uint16_t* rec_sub, const uint64_t* bin_start,
                      const uint64_t* chunk_off, int num_bins, uint32_t chunk_records, uint64_t n_chunks, int K, int sb,
                      unsigned long long* rcnt, unsigned long long* kcnt) {
  if (n_chunks == 0) return;
  const size_t lds = (size_t)8 << sb;
  if (!wide)
    CF1G_LAUNCH("p3_rcount", k_p3_rcount<2>, dim3((unsigned)n_chunks), dim3(P3_WG), lds, st, rec, rec_sub, bin_start, chunk_off,
                num_bins, chunk_records, K, sb, rcnt, kcnt);
  else
    CF1G_LAUNCH("p3_rcount", k_p3_rcount<4>, dim3((unsigned)n_chunks), dim3(P3_WG), lds, st, rec, rec_sub, bin_start, chunk_off,
                num_bins, chunk_records, K, sb, rcnt, kcnt);
}
void launch_p3_rscatter(hipStream_t st, bool wide, const uint64_t* rec, const uint16_t* rec_sub, const uint64_t* bin_start,
                        const uint64_t* chunk_off, int num_bins, uint32_t chunk_records, uint64_t n_chunks, int sb, const uint64_t* roff,
                        unsigned int* rcur, uint64_t* srec, uint32_t* rec_dest, const unsigned int* stop) {
  if (n_chunks == 0) return;
  const size_t l2048 = rscatter_lds<2, 2048>(sb), l1024 = rscatter_lds<2, 1024>(sb), l512 = rscatter_lds<4, 512>(sb);
  if (!wide && l2048 <= RS_LDS_MAX)
    CF1G_LAUNCH("p3_rscatter", (k_p3_rscatter<2, 2048>), dim3((unsigned)n_chunks), dim3(P3_WG), l2048, st, rec, rec_sub, bin_start,
                chunk_off, num_bins, chunk_records, sb, roff, rcur, srec, rec_dest, stop);
  else if (!wide)
    CF1G_LAUNCH("p3_rscatter", (k_p3_rscatter<2, 1024>), dim3((unsigned)n_chunks), dim3(P3_WG), l1024, st, rec, rec_sub, bin_start,
                chunk_off, num_bins, chunk_records, sb, roff, rcur, srec, rec_dest, stop);
  else
    CF1G_LAUNCH("p3_rscatter", (k_p3_rscatter<4, 512>), dim3((unsigned)n_chunks), dim3(P3_WG), l512, st, rec, rec_sub, bin_start,
                chunk_off, num_bins, chunk_records, sb, roff, rcur, srec, rec_dest, stop);
}
void launch_p3_expand(hipStream_t st, bool wide, const uint64_t* srec, const uint64_t* roff, const uint64_t* seg_off,
                      const uint32_t* list, uint64_t n, int K, void* occ, uint16_t* obs, const unsigned int* stop) {
  if (n == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_expand", (k_p3_expand<uint64_t, 2>), dim3((unsigned)n), dim3(P3_WG), 0, st, srec, roff, seg_off, list, K,
                (uint64_t*)occ, obs, stop);
  else
    CF1G_LAUNCH("p3_expand", (k_p3_expand<Key128, 4>), dim3((unsigned)n), dim3(P3_WG), 0, st, srec, roff, seg_off, list, K,
                (Key128*)occ, obs, stop);
}
void launch_p5_list(hipStream_t st, const uint64_t* entries, const uint32_t* ent_dest, const uint32_t* tile_cnt, uint64_t n_tiles,
                    uint32_t tile_slots, const uint32_t* rec_dest, uint64_t S, uint64_t* list, unsigned int* err) {
  if (n_tiles == 0) return;
  const uint64_t g = n_tiles < 8192 ? n_tiles : 8192;
  CF1G_LAUNCH("p5_list", k_p5_list, dim3((unsigned)g), dim3(256), 0, st, entries, ent_dest, tile_cnt, n_tiles, tile_slots, rec_dest, S,
              list, err);
}
void launch_p5_annot(hipStream_t st, bool wide, const uint64_t* srec, const uint64_t* srec_ent, const uint64_t* roff,
                     const uint64_t* key_off, const uint64_t* blk_off, const uint64_t* blocks, const uint64_t* states, int K,
                     uint64_t n_segs, uint8_t* annot, unsigned int* err) {
  if (n_segs == 0) return;
  if (!wide)
    CF1G_LAUNCH("p5_annot", (k_p5_annot<uint64_t, 2>), dim3((unsigned)n_segs), dim3(P3_WG), 0, st, srec, srec_ent, roff, key_off,
                blk_off, blocks, states, K, n_segs, annot, err);
  else
    CF1G_LAUNCH("p5_annot", (k_p5_annot<Key128, 4>), dim3((unsigned)n_segs), dim3(P3_WG), 0, st, srec, srec_ent, roff, key_off,
                blk_off, blocks, states, K, n_segs, annot, err);
}
// Compares the scanned occurrence total with W (the valid k-mers P1 counted)
// before the occurrence pass: on a mismatch flags[3] is set, and scatter,
// split and dedup return at once instead of writing past buffers sized by W.
__global__ void k_p3_gate(const uint64_t* __restrict__ total, uint64_t expect, unsigned int* __restrict__ flags) {
  flags[3] = *total == expect ? 0u : 1u;
}
void launch_p3_gate(hipStream_t st, const uint64_t* total, uint64_t expect, unsigned int* flags) {
  CF1G_LAUNCH("p3_gate", k_p3_gate, dim3(1), dim3(1), 0, st, total, expect, flags);
}
void launch_p3_dedup(hipStream_t st, bool wide, void* occ, uint16_t* obs, const uint64_t* srec, const uint64_t* roff,
                     const uint64_t* seg_off, uint64_t n_segs, int K, int sb, uint64_t* ucount, uint64_t* nblk, uint32_t* over_list,
                     unsigned int* flags, uint32_t* big_list, uint64_t* blocks_pad, uint16_t* vpad) {
  if (n_segs == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_dedup", k_p3_dedup<uint64_t>, dim3((unsigned)n_segs), dim3(DD_WG), 0, st, (uint64_t*)occ, obs, srec, roff, seg_off,
                n_segs, K, sb, ucount, nblk, over_list, flags, big_list, blocks_pad, vpad);
  else
    CF1G_LAUNCH("p3_dedup", k_p3_dedup<Key128>, dim3((unsigned)n_segs), dim3(DD_WG), 0, st, (Key128*)occ, obs, srec, roff, seg_off,
                n_segs, K, sb, ucount, nblk, over_list, flags, big_list, blocks_pad, vpad);
}
void launch_p3_dedup_list(hipStream_t st, bool wide, void* occ, uint16_t* obs, const uint64_t* srec, const uint64_t* roff,
                          const uint64_t* seg_off, uint64_t n_segs, int K, int sb, uint64_t* ucount, uint64_t* nblk,
                          const uint32_t* over_list, uint32_t* over2, unsigned int* flags, uint64_t* blocks_pad, uint16_t* vpad) {
  if (n_segs == 0 || (!wide && 2 * K <= BM_MAX_BITS)) return;  // the direct-address table leaves no list
  const dim3 g(n_segs < DD2_GRID ? (unsigned)n_segs : DD2_GRID);
  if (!wide)
    CF1G_LAUNCH("p3_dedup_list", k_p3_dedup_list<uint64_t>, g, dim3(DD2_WG), 0, st, (uint64_t*)occ, obs, srec, roff, seg_off, K, sb,
                ucount, nblk, over_list, over2, flags, blocks_pad, vpad);
  else
    CF1G_LAUNCH("p3_dedup_list", k_p3_dedup_list<Key128>, g, dim3(DD2_WG), 0, st, (Key128*)occ, obs, srec, roff, seg_off, K, sb,
                ucount, nblk, over_list, over2, flags, blocks_pad, vpad);
}
void launch_p3_bigsort(hipStream_t st, bool wide, void* occ, uint16_t* obs, void* aux, uint16_t* aux_obs, const uint64_t* seg_off,
                       const uint32_t* over_list, uint32_t n_over, uint64_t* ucount, unsigned int* flags, uint32_t* big_list) {
  if (n_over == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_bigsort", k_p3_bigsort<uint64_t>, dim3(n_over), dim3(RK_WG), 0, st, (uint64_t*)occ, obs, (uint64_t*)aux, aux_obs,
                seg_off, over_list, ucount, flags, big_list);
  else
    CF1G_LAUNCH("p3_bigsort", k_p3_bigsort<Key128>, dim3(n_over), dim3(RK_WG), 0, st, (Key128*)occ, obs, (Key128*)aux, aux_obs,
                seg_off, over_list, ucount, flags, big_list);
}
// Both variants are launched over the whole big list; each builds the
// segments of its size class and returns at once for the others.
void launch_bbh_big(hipStream_t st, bool wide, const void* occ, const uint16_t* obs, uint16_t* vpad, const uint64_t* seg_off,
                    const uint64_t* ucount, const uint32_t* big_list, uint32_t n_big, uint64_t* nblk, uint64_t* blocks_pad,
                    uint64_t* scratch, uint64_t* alive_g, void* kout, unsigned int* err) {
  if (n_big == 0) return;
  dim3 g(n_big), b(RK_WG);
  if (!wide) {
    CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<true, uint64_t>), g, b, 0, st, (const uint64_t*)occ, obs, vpad, seg_off, ucount, big_list,
                nblk, blocks_pad, scratch, alive_g, (uint64_t*)kout, err);
    if (alive_g)
      CF1G_LAUNCH("p3b_bbh_huge", (k_bbh_big<false, uint64_t>), g, b, 0, st, (const uint64_t*)occ, obs, vpad, seg_off, ucount,
                  big_list, nblk, blocks_pad, scratch, alive_g, (uint64_t*)kout, err);
  } else {
    CF1G_LAUNCH("p3b_bbh_big", (k_bbh_big<true, Key128>), g, b, 0, st, (const Key128*)occ, obs, vpad, seg_off, ucount, big_list,
                nblk, blocks_pad, scratch, alive_g, (Key128*)kout, err);
    if (alive_g)
      CF1G_LAUNCH("p3b_bbh_huge", (k_bbh_big<false, Key128>), g, b, 0, st, (const Key128*)occ, obs, vpad, seg_off, ucount, big_list,
                  nblk, blocks_pad, scratch, alive_g, (Key128*)kout, err);
  }
}
void launch_p3_compact(hipStream_t st, bool wide, const uint64_t* seg_off, const uint64_t* blk_off, const uint64_t* blocks_pad,
                       uint64_t n_segs, uint64_t* blocks) {
  if (n_segs == 0) return;
  const dim3 grid((unsigned)((n_segs + CP_WG / 64 - 1) / (CP_WG / 64)));
  if (!wide)
    CF1G_LAUNCH("p3_compact", k_p3_compact<uint64_t>, grid, dim3(CP_WG), 0, st, seg_off, blk_off, blocks_pad, n_segs, blocks);
  else
    CF1G_LAUNCH("p3_compact", k_p3_compact<Key128>, grid, dim3(CP_WG), 0, st, seg_off, blk_off, blocks_pad, n_segs, blocks);
}
void launch_p3_keys(hipStream_t st, bool wide, const void* kout, const uint64_t* seg_off, const uint64_t* ucount,
                    const uint64_t* key_off, const uint64_t* blk_off, const uint64_t* blocks, uint64_t n_segs, void* keys,
                    unsigned int* err) {
  if (n_segs == 0) return;
  if (!wide)
    CF1G_LAUNCH("p3_keys", k_p3_keys<uint64_t>, dim3((unsigned)n_segs), dim3(RK_WG), 0, st, (const uint64_t*)kout, seg_off, ucount,
                key_off, blk_off, blocks, n_segs, (uint64_t*)keys, err);
  else
    CF1G_LAUNCH("p3_keys", k_p3_keys<Key128>, dim3((unsigned)n_segs), dim3(RK_WG), 0, st, (const Key128*)kout, seg_off, ucount,
                key_off, blk_off, blocks, n_segs, (Key128*)keys, err);
}

}  // namespace cf1g
