// Per-occurrence classification shared by the CAS kernels of both indexes
// (kernels_p4.hip: BBHash, kernels_lidx.hip: learned index): k-mer rolling,
// the observation of one occurrence (classify.rs:65-177) and its join into
// the packed 5-bit state by a 64-bit CAS loop (state_vector.rs:47-70).
#pragma once
#include "device.hpp"

namespace cf1g {

template <typename KeyT>
struct KOps;
template <>
struct KOps<uint64_t> {
  static __device__ __forceinline__ uint64_t at(const uint64_t* seq, uint64_t g, int K) { return kmer64_at(seq, g, K); }
  static __device__ __forceinline__ uint64_t rc(uint64_t x, int K) { return rc64(x, K); }
  static __device__ __forceinline__ void roll(uint64_t& x, uint64_t& r, uint32_t b, int K) {
    uint64_t mask = (K == 32) ? ~0ULL : ((1ULL << (2 * K)) - 1);
    x = ((x << 2) | b) & mask;
    r = (r >> 2) | ((uint64_t)(3 - b) << (2 * K - 2));
  }
};
template <>
struct KOps<Key128> {
  static __device__ __forceinline__ Key128 at(const uint64_t* seq, uint64_t g, int K) { return kmer128_at(seq, g, K); }
  static __device__ __forceinline__ Key128 rc(const Key128& x, int K) { return rc128(x, K); }
  static __device__ __forceinline__ void roll(Key128& x, Key128& r, uint32_t b, int K) {
    uint64_t mh = (1ULL << (2 * K - 64)) - 1;
    x.hi = ((x.hi << 2) | (x.lo >> 62)) & mh;
    x.lo = (x.lo << 2) | b;
    r.lo = (r.lo >> 2) | (r.hi << 62);
    r.hi = (r.hi >> 2) | ((uint64_t)(3 - b) << (2 * K - 66));
  }
};

// Observation mask of the k-mer occurrence at stream position g >= 1, or 0
// when no valid k-mer starts there. Neighbours are the bases at g-1 and g+K
// (a placeholder or record end is "none"); the occurrence type follows
// classify.rs:65-177: isolated -> MIMO, no right neighbour -> rightmost,
// self loop (the next k-mer has the same canonical form, :143-145) -> MIMO
// when leftmost or a palindromic step, else rightmost (:160-177); otherwise
// leftmost / internal.
template <typename KeyT>
__device__ __forceinline__ uint32_t observe(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask, uint64_t g, int K,
                                            KeyT& canon) {
  if (!window_clear(nmask, g, K)) return 0;
  const bool pv = !is_n(nmask, g - 1), nv = !is_n(nmask, g + K);
  const KeyT x = KOps<KeyT>::at(seq, g, K);
  const KeyT r = KOps<KeyT>::rc(x, K);
  canon = kmer_min(x, r);
  const uint32_t pb = base_at(seq, g - 1), nb = base_at(seq, g + K);
  uint32_t type;
  if (!pv && !nv) {
    type = OBS_MIMO;
  } else if (!nv) {
    type = OBS_RIGHT;
  } else {
    KeyT x2 = x, r2 = r;
    KOps<KeyT>::roll(x2, r2, nb, K);
    if (key_eq(kmer_min(x2, r2), canon)) type = (!pv || key_eq(x, x2)) ? OBS_MIMO : OBS_RIGHT;
    else type = pv ? OBS_INTERNAL : OBS_LEFT;
  }
  return obs_mask(type, key_eq(x, canon), pb, nb);
}

// Joins observation m into the state at value idx of the packed vector:
// reread-and-retry until the CAS holds the word it read, or the state
// already holds the observation. inject makes the first attempt compare
// against a word with bit 63 set, which no state word has (bits 60-63 are
// zero), so that the retry path runs.
__device__ __forceinline__ void cas_join(unsigned long long* __restrict__ states, uint64_t idx, uint32_t m, bool inject,
                                         uint32_t& n_cas, uint32_t& n_fail) {
  unsigned long long* wp = states + idx / 12;
  const uint32_t sh = (uint32_t)(idx % 12) * 5;
  unsigned long long cur = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    const uint32_t s = (uint32_t)(cur >> sh) & 31u;
    const uint32_t ns = mask_state(state_mask(s) | m);
    if (ns == s) break;
    const unsigned long long want = (cur & ~(31ULL << sh)) | ((unsigned long long)ns << sh);
    const unsigned long long expect = inject ? (cur | (1ULL << 63)) : cur;
    inject = false;
    n_cas++;
    const unsigned long long seen = atomicCAS(wp, expect, want);
    if (seen == expect) break;
    n_fail++;
    cur = seen;
  }
}

}  // namespace cf1g
