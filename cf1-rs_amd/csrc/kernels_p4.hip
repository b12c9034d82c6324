// P4: lock-free DFA vertex classification (classify.rs:25-601) and the
// per-position annotation consumed by the host P5 (directed_kmer.rs:70-109).
//
// One lane per k-mer position of the flat stream. Each valid occurrence
// derives its observation from its two neighbours (classify.rs:65-141 reduces
// to: isolated / leftmost / rightmost / internal, with the self-loop cases of
// :143-177), looks its canonical k-mer up in the BBHash and joins the
// observation into the 5-bit state with a 64-bit CAS loop on the packed word
// (state_vector.rs:47-70). The join is commutative and idempotent, so the
// final states do not depend on the interleaving.
//
// Scheduling: a persistent grid; work-groups take fixed chunks of stream
// positions from an atomic counter (work stealing across records of any
// length; a chunk reads up to K bases past its end, the k-1 overlap).
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

// Observation of the valid k-mer at g; returns false when g holds no valid k-mer.
template <typename KeyT>
__device__ __forceinline__ bool observe(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask, uint64_t g, int K,
                                       KeyT& canon, bool& fwd, uint32_t& type, uint32_t& pb, uint32_t& nb) {
  if (!window_clear(nmask, g, K)) return false;
  const bool pv = !is_n(nmask, g - 1);
  const bool nv = !is_n(nmask, g + K);
  KeyT x = KOps<KeyT>::at(seq, g, K);
  KeyT r = KOps<KeyT>::rc(x, K);
  canon = kmer_min(x, r);
  fwd = key_eq(x, canon);
  pb = base_at(seq, g - 1);
  nb = base_at(seq, g + K);
  if (!pv && !nv) {
    type = OBS_MIMO;  // isolated (classify.rs:147-158)
  } else if (!nv) {
    type = OBS_RIGHT;  // classify.rs:83-86, :124-125
  } else {
    KeyT x2 = x, r2 = r;
    KOps<KeyT>::roll(x2, r2, nb, K);
    KeyT c2 = kmer_min(x2, r2);
    if (key_eq(c2, canon)) {  // self loop (classify.rs:143-145, 160-177)
      type = (!pv || key_eq(x, x2)) ? OBS_MIMO : OBS_RIGHT;
    } else {
      type = pv ? OBS_INTERNAL : OBS_LEFT;
    }
  }
  return true;
}

__device__ __forceinline__ void state_join(unsigned long long* __restrict__ states, uint64_t idx, uint32_t type, bool fwd,
                                           uint32_t pb, uint32_t nb) {
  unsigned long long* wp = states + idx / 12;
  const uint32_t sh = (uint32_t)(idx % 12) * 5;
  unsigned long long old = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    uint32_t s = (uint32_t)(old >> sh) & 31u;
    uint32_t ns = dfa_next(s, type, fwd, pb, nb);
    if (ns == s) return;
    unsigned long long nw = (old & ~(31ULL << sh)) | ((unsigned long long)ns << sh);
    unsigned long long prev = atomicCAS(wp, old, nw);
    if (prev == old) return;
    old = prev;
  }
}

template <typename KeyT>
__global__ void __launch_bounds__(256) k_p4_classify(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                     uint64_t n_pos, int K, const MphfView mv, unsigned long long* __restrict__ states,
                                                     unsigned int* __restrict__ next_chunk, uint32_t chunk, uint32_t n_chunks) {
  __shared__ uint32_t s_c;
  for (;;) {
    if (threadIdx.x == 0) s_c = atomicAdd(next_chunk, 1u);
    __syncthreads();
    const uint32_t c = s_c;
    __syncthreads();
    if (c >= n_chunks) break;
    uint64_t g0 = 1 + (uint64_t)c * chunk;
    uint64_t g1 = g0 + chunk;
    if (g1 > n_pos) g1 = n_pos;
    for (uint64_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
      KeyT canon;
      bool fwd;
      uint32_t type, pb, nb;
      if (!observe<KeyT>(seq, nmask, g, K, canon, fwd, type, pb, nb)) continue;
      uint64_t idx = mphf_lookup<KeyT>(mv, canon);
      state_join(states, idx, type, fwd, pb, nb);
    }
  }
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

void launch_p4_classify(hipStream_t st, bool wide, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K,
                        const MphfView& mv, uint64_t* states, unsigned int* next_chunk, uint32_t chunk, uint32_t n_chunks,
                        uint32_t grid) {
  if (n_chunks == 0) return;
  if (!wide)
    CF1G_LAUNCH("p4_classify", k_p4_classify<uint64_t>, dim3(grid), dim3(256), 0, st, seq, nmask, n_pos, K, mv,
                (unsigned long long*)states, next_chunk, chunk, n_chunks);
  else
    CF1G_LAUNCH("p4_classify", k_p4_classify<Key128>, dim3(grid), dim3(256), 0, st, seq, nmask, n_pos, K, mv,
                (unsigned long long*)states, next_chunk, chunk, n_chunks);
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
