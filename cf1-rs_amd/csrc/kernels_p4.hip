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
#include "classify.hpp"
#include "device.hpp"
#include "launch.hpp"

namespace cf1g {

// One wave per P3 segment: its joined masks (vpad[seg_off ...], value order
// within the segment) become 5-bit states, one byte each, at vstate[key_off
// ...]: the dense state of every value, copied eight loads deep. The states
// of all 2^10 masks sit in an LDS table (one entry per thread).
constexpr int PS_WG = 1024;
__global__ void __launch_bounds__(PS_WG) k_p4_states(const uint16_t* __restrict__ vpad, const uint64_t* __restrict__ seg_off,
                                                     const uint64_t* __restrict__ ucount, const uint64_t* __restrict__ key_off,
                                                     uint64_t n_segs, uint8_t* __restrict__ vstate) {
  __shared__ uint8_t st_of[1024];
  st_of[threadIdx.x] = (uint8_t)mask_state(threadIdx.x);
  __syncthreads();
  const uint64_t seg = (uint64_t)blockIdx.x * (PS_WG / 64) + (threadIdx.x >> 6);
  if (seg >= n_segs) return;
  // the segment's words loaded together: no test between them
  const uint64_t u = ucount[seg], s0 = seg_off[seg], ko = key_off[seg];
  const uint16_t* src = vpad + s0;
  uint8_t* dst = vstate + ko;
  uint64_t i = __lane_id();
  for (; i + 7 * 64 < u; i += 8 * 64) {
    uint32_t v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = src[i + q * 64];
#pragma unroll
    for (int q = 0; q < 8; q++) dst[i + q * 64] = st_of[v[q] & 1023u];
  }
  for (; i < u; i += 64) dst[i] = st_of[src[i] & 1023u];
}

// One thread per state word: the 12 states of values [12w, 12w+12) (three
// 4-byte loads) packed 12 per word with bits 60-63 zero
// (state_vector.rs:36-71).
__global__ void __launch_bounds__(256) k_p4_pack(const uint32_t* __restrict__ vstate, uint64_t D, uint64_t n_words,
                                                 uint64_t* __restrict__ states) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_words) return;
  const uint32_t s[3] = {vstate[3 * w], vstate[3 * w + 1], vstate[3 * w + 2]};
  const uint64_t valid = D - 12 * w;  // >= 1
  uint64_t out = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const uint64_t st = (s[i >> 2] >> (8 * (i & 3))) & 31u;
    out |= (uint64_t)i < valid ? st << (5 * i) : 0;
  }
  states[w] = out;
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
      const uint64_t idx = mphf_find<KeyT>(mv, canon);
      if (idx >= D) {
        n_miss++;
        continue;
      }
      cas_join(states, idx, m, force_retry != 0, n_cas, n_fail);
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
// err |= 1 when a k-mer misses the MPHF (the position is left invalid).
template <typename KeyT>
__global__ void __launch_bounds__(256) k_annotate(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask, uint64_t n_pos,
                                                  int K, const MphfView mv, const uint64_t* __restrict__ states, uint64_t D,
                                                  uint8_t* __restrict__ out, unsigned int* __restrict__ err) {
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_pos) return;
  uint8_t a = 0;
  if (g >= 1 && window_clear(nmask, g, K)) {
    KeyT x = KOps<KeyT>::at(seq, g, K);
    KeyT r = KOps<KeyT>::rc(x, K);
    KeyT c = kmer_min(x, r);
    const uint64_t idx = mphf_find<KeyT>(mv, c);
    if (idx < D) {
      const uint32_t s = (uint32_t)(states[idx / 12] >> ((idx % 12) * 5)) & 31u;
      a = (uint8_t)(1u | (key_eq(x, c) ? 2u : 0u) | (st_class(s) << 2));
    } else {
      atomicOr(err, 1u);
    }
  }
  out[g] = a;
}

// ---------------------------------------------------------------- unitig spans (P5)
// A unitig of output_unitigs_contiguous (src/output.rs:506-709) is a maximal
// run of valid k-mers cut wherever is_unipath_start(curr, prev) holds
// (src/output.rs:79-110); is_unipath_end(prev, curr) (:114-145) is the same
// predicate with the roles swapped, so the unitigs partition every run of
// valid k-mers at exactly those cuts. From the per-position annotation
// (bit0 valid, bit1 forward, bits 2-3 class): a unitig starts at a valid g
// when g-1 holds no valid k-mer or the cut holds between g-1 and g, and ends
// at g when g+1 holds no valid k-mer or the cut holds between g and g+1.
// Work-groups take tiles of SP_T positions; starts and ends are compacted
// separately in stream order, so the k-th start and the k-th end bound the
// k-th unitig.
constexpr int SP_WG = 256, SP_T = 4096;
__device__ __forceinline__ bool unipath_cut(uint32_t c, uint32_t p) {  // is_unipath_start(c, p)
  const uint32_t cc = (c >> 2) & 3u, pc = (p >> 2) & 3u;
  const bool cf = (c >> 1) & 1u, pf = (p >> 1) & 1u;
  return cc == CLS_MIMO || cc == (cf ? CLS_MISO : CLS_SIMO) || pc == CLS_MIMO || pc == (pf ? CLS_SIMO : CLS_MISO);
}
__device__ __forceinline__ void span_flags(const uint8_t* __restrict__ ann, uint64_t g, uint64_t n_pos, bool& st, bool& en) {
  st = en = false;
  if (g == 0 || g >= n_pos) return;
  const uint32_t a = ann[g];
  if (!(a & 1u)) return;
  // a valid k-mer ends before the closing separator, so g + 1 < n_pos
  const uint32_t p = ann[g - 1], n = ann[g + 1];
  st = !(p & 1u) || unipath_cut(a, p);
  en = !(n & 1u) || unipath_cut(n, a);
}
__global__ void __launch_bounds__(SP_WG) k_span_count(const uint8_t* __restrict__ ann, uint64_t n_pos, uint64_t* __restrict__ cnt_s,
                                                      uint64_t* __restrict__ cnt_e) {
  __shared__ uint32_t sh[2][SP_WG / 64];
  const uint64_t base = (uint64_t)blockIdx.x * SP_T;
  uint32_t ns = 0, ne = 0;
  for (int q = 0; q < SP_T / SP_WG; q++) {
    bool st, en;
    span_flags(ann, base + (uint64_t)q * SP_WG + threadIdx.x, n_pos, st, en);
    ns += st;
    ne += en;
  }
  for (int d = 32; d > 0; d >>= 1) {
    ns += __shfl_down(ns, d, 64);
    ne += __shfl_down(ne, d, 64);
  }
  if (__lane_id() == 0) {
    sh[0][threadIdx.x >> 6] = ns;
    sh[1][threadIdx.x >> 6] = ne;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t a = 0, b = 0;
    for (int w = 0; w < SP_WG / 64; w++) {
      a += sh[0][w];
      b += sh[1][w];
    }
    cnt_s[blockIdx.x] = a;
    cnt_e[blockIdx.x] = b;
  }
}
__global__ void __launch_bounds__(SP_WG) k_span_emit(const uint8_t* __restrict__ ann, uint64_t n_pos, const uint64_t* __restrict__ off_s,
                                                     const uint64_t* __restrict__ off_e, uint64_t* __restrict__ S,
                                                     uint64_t* __restrict__ E) {
  __shared__ uint32_t sh[2][SP_WG / 64];
  const uint64_t base = (uint64_t)blockIdx.x * SP_T;
  const uint32_t wv = threadIdx.x >> 6;
  uint64_t os = off_s[blockIdx.x], oe = off_e[blockIdx.x];
  for (int q = 0; q < SP_T / SP_WG; q++) {
    const uint64_t g = base + (uint64_t)q * SP_WG + threadIdx.x;
    bool st, en;
    span_flags(ann, g, n_pos, st, en);
    const uint64_t bs = __ballot(st), be = __ballot(en);
    if (__lane_id() == 0) {
      sh[0][wv] = (uint32_t)__popcll(bs);
      sh[1][wv] = (uint32_t)__popcll(be);
    }
    __syncthreads();
    uint32_t ws = 0, we = 0, ts = 0, te = 0;
    for (uint32_t w = 0; w < SP_WG / 64; w++) {
      if (w < wv) {
        ws += sh[0][w];
        we += sh[1][w];
      }
      ts += sh[0][w];
      te += sh[1][w];
    }
    const uint64_t lt = lanemask_lt();
    if (st) S[os + ws + (uint32_t)__popcll(bs & lt)] = g;
    if (en) E[oe + we + (uint32_t)__popcll(be & lt)] = g;
    os += ts;
    oe += te;
    __syncthreads();
  }
}
// Per unitig: the key its id is looked up by, min(canonical(first),
// canonical(last)), and its orientation, forward iff first < rc(last)
// (output_gfa_unitig, src/output.rs:728-736).
template <typename KeyT>
__global__ void __launch_bounds__(256) k_span_keys(const uint64_t* __restrict__ seq, int K, const uint64_t* __restrict__ S,
                                                   const uint64_t* __restrict__ E, uint64_t n, KeyT* __restrict__ keys,
                                                   uint8_t* __restrict__ fl) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const KeyT xs = KOps<KeyT>::at(seq, S[u], K), xe = KOps<KeyT>::at(seq, E[u], K);
  const KeyT rs = KOps<KeyT>::rc(xs, K), re = KOps<KeyT>::rc(xe, K);
  const KeyT cs = kmer_min(xs, rs), ce = kmer_min(xe, re);
  keys[u] = kmer_min(cs, ce);
  fl[u] = key_lt(xs, re) ? 1u : 0u;
}
// The outputted claim of output_gfa_unitig (src/output.rs:747-751): a unitig
// whose id is not yet in an outputted state moves it there by CAS
// (OUTPUTTED_TABLE, src/state.rs:272-298); the one that succeeds writes the
// unitig (flag bit 1). states is a copy, so the P4 state vector stays as it
// was. err: ids outside [0, D).
__device__ __forceinline__ uint32_t outputted_state(uint32_t s) {
  if (s == 3 || s == 15) return 15;
  if ((s >= 4 && s <= 7) || s == 13) return 13;
  if ((s >= 8 && s <= 11) || s == 14) return 14;
  if (s == 12 || s >= 16) return 12;
  return 0;
}
__global__ void __launch_bounds__(256) k_span_claim(const uint64_t* __restrict__ ids, uint64_t n, uint64_t D,
                                                    unsigned long long* __restrict__ states, uint8_t* __restrict__ fl,
                                                    unsigned int* __restrict__ err) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const uint64_t id = ids[u];
  if (id >= D) {
    atomicOr(err, 1u);
    return;
  }
  unsigned long long* wp = states + id / 12;
  const uint32_t sh = (uint32_t)(id % 12) * 5;
  unsigned long long cur = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  bool won = false;
  for (;;) {
    const uint32_t s = (uint32_t)(cur >> sh) & 31u;
    if (s >= 12 && s <= 15) break;
    const unsigned long long want = (cur & ~(31ULL << sh)) | ((unsigned long long)outputted_state(s) << sh);
    const unsigned long long seen = atomicCAS(wp, cur, want);
    if (seen == cur) {
      won = true;
      break;
    }
    cur = seen;
  }
  if (won) fl[u] |= 2u;
}

// Batch Mphf::hash; FIND: a key that misses every level gets ~0 instead of
// the reference's 0 (for callers that check the ids). Queries and results
// stream through once: non-temporal loads (u64 keys) and stores.
template <typename KeyT, bool FIND>
__global__ void __launch_bounds__(256) k_lookup(const KeyT* __restrict__ keys, uint64_t n, const MphfView mv,
                                                uint64_t* __restrict__ out) {
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    KeyT q;
    if constexpr (sizeof(KeyT) == 8) q = (KeyT)__builtin_nontemporal_load((const uint64_t*)&keys[i]);
    else q = keys[i];
    __builtin_nontemporal_store(FIND ? mphf_find<KeyT>(mv, q) : mphf_lookup<KeyT>(mv, q), &out[i]);
  }
}

// vstate: 12 * ceil(D/12) bytes, 4-byte aligned.
void launch_p4_pack(hipStream_t st, const uint16_t* vpad, const uint64_t* seg_off, const uint64_t* ucount, const uint64_t* key_off,
                    uint64_t n_segs, uint64_t D, uint8_t* vstate, uint64_t* states) {
  const uint64_t nw = (D + 11) / 12;
  if (nw == 0) return;
  CF1G_LAUNCH("p4_states", k_p4_states, dim3((unsigned)((n_segs + PS_WG / 64 - 1) / (PS_WG / 64))), dim3(PS_WG), 0, st, vpad, seg_off,
              ucount, key_off, n_segs, vstate);
  CF1G_LAUNCH("p4_classify", k_p4_pack, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, (const uint32_t*)vstate, D, nw, states);
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
uint64_t span_tiles(uint64_t n_pos) { return (n_pos + SP_T - 1) / SP_T; }
void launch_span_count(hipStream_t st, const uint8_t* ann, uint64_t n_pos, uint64_t* cnt_s, uint64_t* cnt_e) {
  const uint64_t nt = span_tiles(n_pos);
  if (!nt) return;
  CF1G_LAUNCH("p5_span_count", k_span_count, dim3((unsigned)nt), dim3(SP_WG), 0, st, ann, n_pos, cnt_s, cnt_e);
}
void launch_span_emit(hipStream_t st, const uint8_t* ann, uint64_t n_pos, const uint64_t* off_s, const uint64_t* off_e, uint64_t* S,
                      uint64_t* E) {
  const uint64_t nt = span_tiles(n_pos);
  if (!nt) return;
  CF1G_LAUNCH("p5_span_emit", k_span_emit, dim3((unsigned)nt), dim3(SP_WG), 0, st, ann, n_pos, off_s, off_e, S, E);
}
void launch_span_keys(hipStream_t st, bool wide, const uint64_t* seq, int K, const uint64_t* S, const uint64_t* E, uint64_t n, void* keys,
                      uint8_t* fl) {
  if (!n) return;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (!wide)
    CF1G_LAUNCH("p5_span_keys", k_span_keys<uint64_t>, grid, dim3(256), 0, st, seq, K, S, E, n, (uint64_t*)keys, fl);
  else
    CF1G_LAUNCH("p5_span_keys", k_span_keys<Key128>, grid, dim3(256), 0, st, seq, K, S, E, n, (Key128*)keys, fl);
}
void launch_span_claim(hipStream_t st, const uint64_t* ids, uint64_t n, uint64_t D, uint64_t* states, uint8_t* fl, unsigned int* err) {
  if (!n) return;
  CF1G_LAUNCH("p5_span_claim", k_span_claim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ids, n, D,
              (unsigned long long*)states, fl, err);
}
void launch_annotate(hipStream_t st, bool wide, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K,
                     const MphfView& mv, const uint64_t* states, uint64_t D, uint8_t* out, unsigned int* err) {
  if (n_pos == 0) return;
  dim3 grid((unsigned)((n_pos + 255) / 256));
  if (!wide)
    CF1G_LAUNCH("p5_annotate", k_annotate<uint64_t>, grid, dim3(256), 0, st, seq, nmask, n_pos, K, mv, states, D, out, err);
  else
    CF1G_LAUNCH("p5_annotate", k_annotate<Key128>, grid, dim3(256), 0, st, seq, nmask, n_pos, K, mv, states, D, out, err);
}
void launch_lookup(hipStream_t st, bool wide, const void* keys, uint64_t n, const MphfView& mv, uint64_t* out, bool find) {
  if (n == 0) return;
  uint64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  const dim3 gr((unsigned)g);
  if (!wide) {
    if (find) CF1G_LAUNCH("bbh_lookup", (k_lookup<uint64_t, true>), gr, dim3(256), 0, st, (const uint64_t*)keys, n, mv, out);
    else CF1G_LAUNCH("bbh_lookup", (k_lookup<uint64_t, false>), gr, dim3(256), 0, st, (const uint64_t*)keys, n, mv, out);
  } else {
    if (find) CF1G_LAUNCH("bbh_lookup", (k_lookup<Key128, true>), gr, dim3(256), 0, st, (const Key128*)keys, n, mv, out);
    else CF1G_LAUNCH("bbh_lookup", (k_lookup<Key128, false>), gr, dim3(256), 0, st, (const Key128*)keys, n, mv, out);
  }
}

}  // namespace cf1g
