// P4: DFA vertex classification (classify.rs:25-601) into the packed state
// vector (state_vector.rs:7-71), and the per-position annotation consumed by
// the host P5 (directed_kmer.rs:70-109).
//
// Every occurrence's observation (isolated / leftmost / rightmost / internal
// with the self-loop cases, classify.rs:65-177) was computed in P3a from the
// routed super k-mers and OR-joined per distinct k-mer while P3a sorted and
// deduplicated the segment; P3b placed each joined mask at the k-mer's MPHF
// value. The reference reaches the same state by one CAS per occurrence,
// because its threads meet the occurrences in sequence order; the join is
// order-free, so P4 here is a streaming pass that encodes each joined mask
// as its 5-bit state (State::from_vertex, state.rs:92-113) and packs 12 per
// word.
//
// k_p4_cas is the reference's own form of P4, for callers that hold an MPHF
// and a zeroed state vector and classify the records against them: one lane
// per k-mer occurrence of the flat stream derives the observation from the
// two neighbouring bases, looks the canonical k-mer up in the BBHash, and
// joins the observation into the 5-bit state by a 64-bit CAS loop on the
// packed word (state_vector.rs:47-70). A persistent grid of work-groups
// takes fixed-size chunks of stream positions from an atomic counter (work
// stealing over records of any length; a chunk reads up to K bases past its
// end, the k-1 overlap). Its states equal k_p4_pack's word for word.
#include "device.hpp"
#include "launch.hpp"

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

// One thread per state word: the 12 joined masks of values [12w, 12w+12)
// (three 8-byte loads; the mask array is value-ordered by P3) become 5-bit
// states, packed 12 per word with bits 60-63 zero (state_vector.rs:36-71).
__global__ void __launch_bounds__(256) k_p4_pack(const uint64_t* __restrict__ vmask, uint64_t D, uint64_t n_words,
                                                 uint64_t* __restrict__ states) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  uint64_t m[3] = {vmask[3 * w], vmask[3 * w + 1], vmask[3 * w + 2]};
  const uint64_t valid = D - 12 * w;  // >= 1
  uint64_t out = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const uint32_t mk = (uint32_t)(m[i >> 2] >> (16 * (i & 3))) & 0xFFFFu;
    if ((uint64_t)i < valid) out |= (uint64_t)mask_state(mk) << (5 * i);
  }
  states[w] = out;
}

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

// Persistent work-stealing classification by CAS (see the file comment).
// counters: [0] occurrences, [1] CAS issued, [2] CAS lost to a concurrent
// update, [3] lookups outside [0, D) (a key missing from the MPHF: never
// written), [4] the chunk ticket. force_retry makes every lane's first CAS
// compare against a word with bit 63 set, which no state word has (bits
// 60-63 are zero), so each occurrence's first attempt fails once and the
// retry path runs.
constexpr int P4_WG = 256;
template <typename KeyT>
__global__ void __launch_bounds__(P4_WG) k_p4_cas(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask, uint64_t n_pos,
                                                  int K, const MphfView mv, uint64_t D, unsigned long long* __restrict__ states,
                                                  unsigned long long* __restrict__ counters, uint32_t chunk, uint64_t n_chunks,
                                                  int force_retry) {
  __shared__ uint64_t s_c;
  __shared__ unsigned long long s_sum[4];
  if (threadIdx.x < 4) s_sum[threadIdx.x] = 0;
  uint32_t n_occ = 0, n_cas = 0, n_fail = 0, n_miss = 0;
  for (;;) {
    if (threadIdx.x == 0) s_c = atomicAdd(&counters[4], 1ULL);
    __syncthreads();
    const uint64_t c = s_c;
    __syncthreads();
    if (c >= n_chunks) break;
    const uint64_t g0 = 1 + c * chunk;
    const uint64_t g1 = g0 + chunk < n_pos ? g0 + chunk : n_pos;
    for (uint64_t g = g0 + threadIdx.x; g < g1; g += P4_WG) {
      KeyT canon;
      const uint32_t m = observe<KeyT>(seq, nmask, g, K, canon);
      if (!m) continue;
      n_occ++;
      const uint64_t idx = mphf_lookup<KeyT>(mv, canon);
      if (idx >= D) {
        n_miss++;
        continue;
      }
      unsigned long long* wp = states + idx / 12;
      const uint32_t sh = (uint32_t)(idx % 12) * 5;
      unsigned long long cur = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bool inject = force_retry != 0;
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
  }
  atomicAdd(&s_sum[0], (unsigned long long)n_occ);
  atomicAdd(&s_sum[1], (unsigned long long)n_cas);
  atomicAdd(&s_sum[2], (unsigned long long)n_fail);
  atomicAdd(&s_sum[3], (unsigned long long)n_miss);
  __syncthreads();
  if (threadIdx.x < 4) atomicAdd(&counters[threadIdx.x], s_sum[threadIdx.x]);
}

// Per stream position: bit0 valid k-mer, bit1 FWD, bits2-3 state class.
template <typename KeyT>
__global__ void __launch_bounds__(256) k_annotate(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask, uint64_t n_pos,
                                                  int K, const MphfView mv, const uint64_t* __restrict__ states,
                                                  uint8_t* __restrict__ out) {
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pos) return;
  uint8_t a = 0;
  if (g >= 1 && window_clear(nmask, g, K)) {
    KeyT x = KOps<KeyT>::at(seq, g, K);
    KeyT r = KOps<KeyT>::rc(x, K);
    KeyT c = kmer_min(x, r);
    uint64_t idx = mphf_lookup<KeyT>(mv, c);
    uint32_t s = (uint32_t)(states[idx / 12] >> ((idx % 12) * 5)) & 31u;
    a = (uint8_t)(1u | (key_eq(x, c) ? 2u : 0u) | (st_class(s) << 2));
  }
  out[g] = a;
}

template <typename KeyT>
__global__ void __launch_bounds__(256) k_lookup(const KeyT* __restrict__ keys, uint64_t n, const MphfView mv,
                                                uint64_t* __restrict__ out) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = mphf_lookup<KeyT>(mv, keys[i]);
}

void launch_p4_pack(hipStream_t st, const uint16_t* vmask, uint64_t D, uint64_t* states) {
  const uint64_t nw = (D + 11) / 12;
  if (nw == 0) return;
  CF1G_LAUNCH("p4_classify", k_p4_pack, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, (const uint64_t*)vmask, D, nw, states);
}
// counters: 5 zeroed words (see k_p4_cas). Chunks cover stream positions
// [1, n_pos); grid = persistent work-groups.
void launch_p4_cas(hipStream_t st, bool wide, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, const MphfView& mv,
                   uint64_t D, uint64_t* states, uint64_t* counters, uint32_t chunk, uint32_t grid, int force_retry) {
  if (n_pos <= 1 || chunk == 0 || grid == 0) return;
  const uint64_t n_chunks = (n_pos - 1 + chunk - 1) / chunk;
  auto* sv = (unsigned long long*)states;
  auto* cn = (unsigned long long*)counters;
  if (!wide)
    CF1G_LAUNCH("p4_cas", k_p4_cas<uint64_t>, dim3(grid), dim3(P4_WG), 0, st, seq, nmask, n_pos, K, mv, D, sv, cn, chunk, n_chunks,
                force_retry);
  else
    CF1G_LAUNCH("p4_cas", k_p4_cas<Key128>, dim3(grid), dim3(P4_WG), 0, st, seq, nmask, n_pos, K, mv, D, sv, cn, chunk, n_chunks,
                force_retry);
}
void launch_annotate(hipStream_t st, bool wide, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K,
                     const MphfView& mv, const uint64_t* states, uint8_t* out) {
  if (n_pos == 0) return;
  dim3 grid((unsigned)((n_pos + 255) / 256));
  if (!wide)
    CF1G_LAUNCH("p5_annotate", k_annotate<uint64_t>, grid, dim3(256), 0, st, seq, nmask, n_pos, K, mv, states, out);
  else
    CF1G_LAUNCH("p5_annotate", k_annotate<Key128>, grid, dim3(256), 0, st, seq, nmask, n_pos, K, mv, states, out);
}
void launch_lookup(hipStream_t st, bool wide, const void* keys, uint64_t n, const MphfView& mv, uint64_t* out) {
  if (n == 0) return;
  uint64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  if (!wide)
    CF1G_LAUNCH("bbh_lookup", k_lookup<uint64_t>, dim3((unsigned)g), dim3(256), 0, st, (const uint64_t*)keys, n, mv, out);
  else
    CF1G_LAUNCH("bbh_lookup", k_lookup<Key128>, dim3((unsigned)g), dim3(256), 0, st, (const Key128*)keys, n, mv, out);
}

}  // namespace cf1g
