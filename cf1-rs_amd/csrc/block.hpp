// Block-level building blocks shared by the kernels: wave / block scans,
// stable per-digit ranking of register tiles, key padding and OR-reductions.
#pragma once
#include "device.hpp"

namespace cf1g {

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t x) {
  uint32_t lane = __lane_id();
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if ((int)lane >= d) x += y;
  }
  return x;
}
// Exclusive scan across a block of WG threads (sh: WG/64 words); *total
// gets the block sum.
template <int WG>
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* sh, uint64_t* total) {
  uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  uint64_t x = wave_incl_scan_u64(v);
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int q = 0; q < WG / 64; q++) {
    uint64_t t = sh[q];
    if (q < (int)wv) off += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return off + x - v;
}
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t* sh4, uint64_t* total) {
  return block_excl_scan<256>(v, sh4, total);
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

}  // namespace cf1g
