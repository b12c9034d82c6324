// P1 / P2 kernels: 2-bit packing of the flat stream, canonical-minimizer
// histogram (minimizer.rs:27-110), bucket->bin partition
// (minimizer.rs:116-184) and super k-mer routing into per-bin HBM segments
// (superkmer.rs:21-210).
#include "device.hpp"
#include "launch.hpp"

namespace cf1g {

// ------------------------------------------------------------------ pack
// One thread per 64 stream positions: one N-mask word and two sequence words.
// gstart[r] = stream position of record r's first base; gstart[n_rec] = the
// position just past the last separator. Non-ACGT bytes and separators are
// placeholders (dna.rs:64-92); their 2-bit code is 0.
__global__ void __launch_bounds__(256) k_pack(const uint8_t* __restrict__ ascii, const uint64_t* __restrict__ rec_off,
                                              const uint64_t* __restrict__ gstart, uint64_t n_rec, uint64_t n_words64,
                                              uint64_t* __restrict__ seq, uint64_t* __restrict__ nmask) {
  uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_words64) return;
  uint64_t g0 = t * 64;
  // largest r with gstart[r] <= g0 (r = n_rec when past the last record)
  uint64_t lo = 0, hi = n_rec + 1;
  while (hi - lo > 1) {
    uint64_t mid = (lo + hi) >> 1;
    if (gstart[mid] <= g0) lo = mid; else hi = mid;
  }
  uint64_t r = lo;
  uint64_t s0 = 0, s1 = 0, nm = 0;
  for (int q = 0; q < 64; q++) {
    uint64_t g = g0 + (uint64_t)q;
    while (r < n_rec && g >= gstart[r + 1]) r++;
    uint32_t code = 4;
    if (r < n_rec && g >= gstart[r] && g + 1 < gstart[r + 1]) {
      uint8_t c = ascii[rec_off[r] + (g - gstart[r])];
      uint8_t u = c & 0xDF;  // upper-case letters
      code = (u == 'A') ? 0u : (u == 'C') ? 1u : (u == 'G') ? 2u : (u == 'T') ? 3u : 4u;
      if (c < 'A') code = 4;
    }
    uint64_t isn = code >> 2;
    uint64_t b = code & 3;
    nm |= isn << (63 - q);
    if (q < 32) s0 |= b << (62 - 2 * q);
    else s1 |= b << (62 - 2 * (q - 32));
  }
  seq[2 * t] = s0;
  seq[2 * t + 1] = s1;
  nmask[t] = nm;
}

// ------------------------------------------------------------------ minimizer tiles
// Project minimizer scheme (same on the CPU oracle): order of an m-mer =
// fmix32(canonical m-mer + c); the minimizer of the k-mer at i is the leftmost
// least-order m-mer in [i, i+w); a super k-mer is a maximal run of valid
// k-mers sharing the minimizer position; it routes by bucket_hash(value).
constexpr int MZ_WG = 256;
constexpr int MZ_T = 1024;         // k-mer positions per tile
constexpr int MZ_MAXW = 64;        // w <= 49 for k <= 63
constexpr int MZ_NJ = MZ_T + 2 * MZ_MAXW;
constexpr int MZ_PPT = MZ_T / MZ_WG;
constexpr int MAX_BINS = 4096;

struct MzTile {
  uint32_t ord[MZ_NJ];
  uint32_t can[MZ_NJ];
  uint16_t pidx[MZ_T + MZ_MAXW + 1];
  uint8_t valid[MZ_T + MZ_MAXW + 1];
};

// Fills the tile for k-mer positions [t0, t0+MZ_T): orders of m-mers at
// [t0-1, t0-1+NJ), minimizer positions (relative to t0-1) and validity of
// k-mers at [t0-1, t0+T+w).
__device__ __forceinline__ void mz_fill(MzTile& S, const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                        uint64_t t0, int K, int m, int w) {
  const uint64_t base = t0 - 1;
  const int nj = MZ_T + 2 * w;
  for (int j = threadIdx.x; j < nj; j += MZ_WG) {
    uint64_t g = base + (uint64_t)j;
    uint32_t f = (uint32_t)(stream_bits64(seq, 2 * g) >> (64 - 2 * m));
    uint32_t r = (uint32_t)rc64(f, m);
    uint32_t c = f < r ? f : r;
    S.can[j] = c;
    S.ord[j] = mmer_order(c);
  }
  __syncthreads();
  const int np = MZ_T + w + 1;
  for (int j = threadIdx.x; j < np; j += MZ_WG) {
    uint64_t g = base + (uint64_t)j;
    uint32_t best = S.ord[j];
    int bi = j;
    for (int q = 1; q < w; q++) {
      uint32_t o = S.ord[j + q];
      if (o < best) { best = o; bi = j + q; }
    }
    S.pidx[j] = (uint16_t)bi;
    S.valid[j] = window_clear(nmask, g, K) ? 1 : 0;
  }
  __syncthreads();
}

__device__ __forceinline__ bool mz_is_start(const MzTile& S, int j) {  // j relative to t0-1, j >= 1
  return S.valid[j] && (!S.valid[j - 1] || S.pidx[j] != S.pidx[j - 1]);
}

// P1: histogram of super k-mers per bucket; also counts valid k-mers (W) and
// super k-mers (S).
__global__ void __launch_bounds__(MZ_WG) k_p1_hist(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                   uint64_t n_pos, int K, int m, unsigned long long* __restrict__ hist,
                                                   unsigned long long* __restrict__ totals) {
  __shared__ MzTile S;
  __shared__ uint32_t red[2];
  const int w = K - m + 1;
  uint64_t t0 = 1 + (uint64_t)blockIdx.x * MZ_T;
  if (threadIdx.x < 2) red[threadIdx.x] = 0;
  mz_fill(S, seq, nmask, t0, K, m, w);
  uint32_t nw = 0, ns = 0;
  for (int q = 0; q < MZ_PPT; q++) {
    int li = threadIdx.x + q * MZ_WG;
    uint64_t i = t0 + (uint64_t)li;
    if (i >= n_pos) break;
    int j = li + 1;
    if (!S.valid[j]) continue;
    nw++;
    if (mz_is_start(S, j)) {
      ns++;
      atomicAdd(&hist[bucket_hash((uint64_t)S.can[S.pidx[j]])], 1ULL);
    }
  }
  atomicAdd(&red[0], nw);
  atomicAdd(&red[1], ns);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&totals[0], (unsigned long long)red[0]);
    atomicAdd(&totals[1], (unsigned long long)red[1]);
  }
}

// P2 route: each super k-mer [i, i+n) becomes one fixed-size record in its
// bin (the bin record of superkmer.rs:181-202 in fixed width): its L = n-1+K
// bases plus one flanking base on each side, MSB-first in RW words; the last
// word's low byte holds L and bits 8/9 say whether the left/right flank is a
// real base (not a placeholder or record end). The flanks carry the
// neighbours P4's observations need (classify.rs:76-77).
template <int RW>
__global__ void __launch_bounds__(MZ_WG) k_p2_route(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                    uint64_t n_pos, int K, int m, const uint16_t* __restrict__ b2b,
                                                    int num_bins, const uint64_t* __restrict__ bin_start,
                                                    unsigned long long* __restrict__ bin_cursor, uint64_t* __restrict__ rec,
                                                    unsigned int* __restrict__ overflow) {
  __shared__ MzTile S;
  __shared__ uint32_t cnt[MAX_BINS];
  __shared__ unsigned long long gbase[MAX_BINS];
  const int w = K - m + 1;
  uint64_t t0 = 1 + (uint64_t)blockIdx.x * MZ_T;
  for (int b = threadIdx.x; b < num_bins; b += MZ_WG) cnt[b] = 0;
  mz_fill(S, seq, nmask, t0, K, m, w);
  int my_bin[MZ_PPT];
  uint32_t my_rank[MZ_PPT];
  for (int q = 0; q < MZ_PPT; q++) {
    my_bin[q] = -1;
    int li = threadIdx.x + q * MZ_WG;
    uint64_t i = t0 + (uint64_t)li;
    if (i >= n_pos) continue;
    int j = li + 1;
    if (!mz_is_start(S, j)) continue;
    int bin = b2b[bucket_hash((uint64_t)S.can[S.pidx[j]])];
    my_bin[q] = bin;
    my_rank[q] = atomicAdd(&cnt[bin], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < num_bins; b += MZ_WG)
    if (cnt[b]) gbase[b] = atomicAdd(&bin_cursor[b], (unsigned long long)cnt[b]);
  __syncthreads();
  for (int q = 0; q < MZ_PPT; q++) {
    if (my_bin[q] < 0) continue;
    int li = threadIdx.x + q * MZ_WG;
    int j = li + 1;
    uint32_t p = S.pidx[j];
    int n = 1;
    while (n < w && S.valid[j + n] && S.pidx[j + n] == p) n++;
    uint32_t L = (uint32_t)(n + K - 1);
    uint64_t i = t0 + (uint64_t)li;
    uint64_t out = bin_start[my_bin[q]] + gbase[my_bin[q]] + my_rank[q];
    if (out >= bin_start[my_bin[q] + 1]) {  // never write past the bin
      atomicOr(overflow, 1u);
      continue;
    }
    // bases [i-1, i+L]: the super k-mer with one flanking base each side
    const uint32_t LS = L + 2;
    uint64_t wv[RW];
#pragma unroll
    for (int t = 0; t < RW; t++) {
      int lo_bit = 64 * t;  // record bit offset of this word
      uint64_t x = stream_bits64(seq, 2 * (i - 1) + (uint64_t)lo_bit);
      int keep = (int)(2 * LS) - lo_bit;  // record bits living in this word
      if (keep <= 0) x = 0;
      else if (keep < 64) x &= ~0ULL << (64 - keep);
      wv[t] = x;
    }
    const uint64_t lf = is_n(nmask, i - 1) ? 0 : 1, rf = is_n(nmask, i + L) ? 0 : 1;
    wv[RW - 1] |= (uint64_t)L | (lf << 8) | (rf << 9);
    uint64_t* dst = rec + out * RW;
#pragma unroll
    for (int t = 0; t < RW; t++) dst[t] = wv[t];
  }
}

// ------------------------------------------------------------------ partition
// partition_minimizers (minimizer.rs:116-144) in closed form. With
// t = total/num_bins, f(j) = floor(P(j)/t) (P = inclusive prefix sum) and
// cap = num_bins-1, the number of boundaries placed up to bucket b is
//   nb(b) = min(cap, b + min(1, min_{j<=b} (f(j) - j)))     (t > 0)
//   nb(b) = min(cap, b + 1)                                  (t == 0)
// because the greedy loop places at most one boundary per bucket and one
// whenever f(b) >= nb(b-1)+1. bucket_to_bin[b] = nb(b-1) (from_boundaries,
// minimizer.rs:156-172). One work-group of 1024 threads, 64 buckets each.
constexpr int PT_WG = 1024;
constexpr int PT_PER = 65536 / PT_WG;

__device__ __forceinline__ uint64_t blk_scan_excl_u64(uint64_t v, uint64_t* sh, uint64_t* total) {
  // sh: PT_WG/64 slots. Inclusive wave scan then block offsets.
  uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  uint64_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if ((int)lane >= d) x += y;
  }
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t run = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      uint64_t t = sh[q];
      sh[q] = run;
      run += t;
    }
    sh[blockDim.x >> 6] = run;
  }
  __syncthreads();
  uint64_t r = sh[wv] + x - v;
  if (total) *total = sh[blockDim.x >> 6];
  __syncthreads();
  return r;
}
__device__ __forceinline__ int64_t blk_scan_excl_min(int64_t v, int64_t* sh, int64_t ident) {
  uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  int64_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    int64_t y = __shfl_up(x, d, 64);
    if ((int)lane >= d) x = y < x ? y : x;
  }
  int64_t ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = ident;
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = ident;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
      int64_t t = sh[q];
      sh[q] = run;
      run = t < run ? t : run;
    }
  }
  __syncthreads();
  int64_t pre = sh[wv];
  int64_t r = ex < pre ? ex : pre;
  __syncthreads();
  return r;
}

// Boundaries come from `hist` (the caller's histogram); record counts per
// bin come from `hist_true` (the super k-mers P1 actually counted), so the
// routed segments are sized by what P2 will write.
// Outputs: b2b[65536]; bin_records[B] (super k-mers per bin), bin_start[B+1]
// (exclusive scan), bin_cursor[B] = 0, chunk_off[B+1] (exclusive scan of
// ceil(records/chunk)), info[0..3] = {num_bins, S, chunks, unused}.
__global__ void __launch_bounds__(PT_WG) k_partition(const unsigned long long* __restrict__ hist,
                                                     const unsigned long long* __restrict__ hist_true, uint32_t num_bins_req,
                                                     uint32_t chunk_records, uint16_t* __restrict__ b2b,
                                                     uint64_t* __restrict__ bin_records, uint64_t* __restrict__ bin_start,
                                                     unsigned long long* __restrict__ bin_cursor, uint64_t* __restrict__ chunk_off,
                                                     uint64_t* __restrict__ info) {
  __shared__ uint64_t shu[PT_WG / 64 + 1];
  __shared__ int64_t shm[PT_WG / 64 + 1];
  __shared__ unsigned long long binrec[MAX_BINS];
  __shared__ uint32_t s_nbins;
  const int tid = threadIdx.x;
  const int b0 = tid * PT_PER;
  uint64_t loc = 0;
  for (int q = 0; q < PT_PER; q++) loc += hist[b0 + q];
  uint64_t total;
  uint64_t pre = blk_scan_excl_u64(loc, shu, &total);
  const uint64_t t = total / (uint64_t)num_bins_req;
  const int64_t cap = (int64_t)num_bins_req - 1;
  // running min of g(j) = f(j) - j over this thread's buckets (inclusive, local)
  int64_t gmin_loc = INT64_MAX;
  if (t > 0) {
    uint64_t P = pre;
    for (int q = 0; q < PT_PER; q++) {
      P += hist[b0 + q];
      int64_t g = (int64_t)(P / t) - (int64_t)(b0 + q);
      gmin_loc = g < gmin_loc ? g : gmin_loc;
    }
  }
  int64_t gmin_pre = blk_scan_excl_min(gmin_loc, shm, INT64_MAX);  // min over j < b0
  for (int b = tid; b < (int)num_bins_req && b < MAX_BINS; b += PT_WG) binrec[b] = 0;
  __syncthreads();
  {
    uint64_t P = pre;
    int64_t run = gmin_pre;
    // nb(b0 - 1)
    int64_t prev_nb;
    if (b0 == 0) {
      prev_nb = 0;
    } else if (t == 0) {
      prev_nb = (int64_t)b0 < cap ? (int64_t)b0 : cap;  // min(cap, (b0-1)+1)
    } else {
      int64_t mm = run < 1 ? run : 1;
      int64_t v = (int64_t)(b0 - 1) + mm;
      prev_nb = v < cap ? v : cap;
    }
    for (int q = 0; q < PT_PER; q++) {
      int b = b0 + q;
      P += hist[b];
      int64_t nb;
      if (t == 0) {
        nb = (int64_t)b + 1 < cap ? (int64_t)b + 1 : cap;
      } else {
        int64_t g = (int64_t)(P / t) - (int64_t)b;
        run = g < run ? g : run;
        int64_t mm = run < 1 ? run : 1;
        int64_t v = (int64_t)b + mm;
        nb = v < cap ? v : cap;
      }
      uint16_t bin = (uint16_t)prev_nb;
      b2b[b] = bin;
      uint64_t ht = hist_true[b];
      if (ht) atomicAdd(&binrec[bin], (unsigned long long)ht);
      prev_nb = nb;
    }
    if (b0 + PT_PER == 65536) s_nbins = (uint32_t)(prev_nb + 1);
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t nb = s_nbins;
    info[0] = nb;
    uint64_t run = 0, crun = 0;
    for (uint32_t b = 0; b < nb; b++) {
      bin_start[b] = run;
      chunk_off[b] = crun;
      uint64_t r = binrec[b];
      bin_records[b] = r;
      bin_cursor[b] = 0;
      run += r;
      crun += (r + chunk_records - 1) / chunk_records;
    }
    bin_start[nb] = run;
    chunk_off[nb] = crun;
    info[1] = run;
    info[2] = crun;
  }
}

// ------------------------------------------------------------------ launchers
void launch_pack(hipStream_t st, const uint8_t* ascii, const uint64_t* rec_off, const uint64_t* gstart, uint64_t n_rec,
                 uint64_t n_words64, uint64_t* seq, uint64_t* nmask) {
  if (n_words64 == 0) return;
  dim3 grid((unsigned)((n_words64 + 255) / 256));
  CF1G_LAUNCH("pack", k_pack, grid, dim3(256), 0, st, ascii, rec_off, gstart, n_rec, n_words64, seq, nmask);
}
void launch_p1_hist(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, int m,
                    unsigned long long* hist, unsigned long long* totals) {
  if (n_pos <= 1) return;
  dim3 grid((unsigned)((n_pos - 1 + MZ_T - 1) / MZ_T));
  CF1G_LAUNCH("p1_minimizer_hist", k_p1_hist, grid, dim3(MZ_WG), 0, st, seq, nmask, n_pos, K, m, hist, totals);
}
void launch_partition(hipStream_t st, const unsigned long long* hist, const unsigned long long* hist_true, uint32_t num_bins,
                      uint32_t chunk_records, uint16_t* b2b, uint64_t* bin_records, uint64_t* bin_start,
                      unsigned long long* bin_cursor, uint64_t* chunk_off, uint64_t* info) {
  CF1G_LAUNCH("p2_partition", k_partition, dim3(1), dim3(PT_WG), 0, st, hist, hist_true, num_bins, chunk_records, b2b,
              bin_records, bin_start, bin_cursor, chunk_off, info);
}
void launch_p2_route(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, int m,
                     const uint16_t* b2b, int num_bins, const uint64_t* bin_start, unsigned long long* bin_cursor,
                     uint64_t* rec, int rw, unsigned int* overflow) {
  if (n_pos <= 1) return;
  dim3 grid((unsigned)((n_pos - 1 + MZ_T - 1) / MZ_T));
  if (rw == 2)
    CF1G_LAUNCH("p2_route", k_p2_route<2>, grid, dim3(MZ_WG), 0, st, seq, nmask, n_pos, K, m, b2b, num_bins, bin_start,
                bin_cursor, rec, overflow);
  else
    CF1G_LAUNCH("p2_route", k_p2_route<4>, grid, dim3(MZ_WG), 0, st, seq, nmask, n_pos, K, m, b2b, num_bins, bin_start,
                bin_cursor, rec, overflow);
}

}  // namespace cf1g
