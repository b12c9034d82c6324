// Device-side primitives shared by the cf1g kernels (gfx950 / CDNA4).
//
// Sequence layout in HBM ("flat stream" G): one separator, then every record
// followed by one separator, padded with separators:
//     G = N r0 N r1 N ... r_{n-1} N N...
// Bases are 2-bit (A0 C1 G2 T3), 32 per u64 word, MSB-first (base g sits at
// bits 63-2*(g%32)..62-2*(g%32) of word g/32). A parallel bit array marks
// placeholders (non-ACGT input and separators), 64 per word, MSB-first.
// A k-mer at g is valid iff its K mask bits are clear; record ends behave
// exactly like an N (classify.rs:76-77, output.rs:517-518).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf1g {

struct Key128 {
  uint64_t lo, hi;
};

__host__ __device__ __forceinline__ bool key_eq(uint64_t a, uint64_t b) { return a == b; }
__host__ __device__ __forceinline__ bool key_eq(const Key128& a, const Key128& b) { return a.lo == b.lo && a.hi == b.hi; }
__host__ __device__ __forceinline__ bool key_lt(uint64_t a, uint64_t b) { return a < b; }
__host__ __device__ __forceinline__ bool key_lt(const Key128& a, const Key128& b) {
  return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo);
}
__host__ __device__ __forceinline__ uint64_t key_lo(uint64_t a) { return a; }
__host__ __device__ __forceinline__ uint64_t key_hi(uint64_t) { return 0; }
__host__ __device__ __forceinline__ uint64_t key_lo(const Key128& a) { return a.lo; }
__host__ __device__ __forceinline__ uint64_t key_hi(const Key128& a) { return a.hi; }

// ------------------------------------------------------------ stream access
// 64 bits of the MSB-first bit stream starting at bit position `bit`.
__device__ __forceinline__ uint64_t stream_bits64(const uint64_t* __restrict__ w, uint64_t bit) {
  uint64_t wi = bit >> 6;
  uint32_t sh = (uint32_t)(bit & 63);
  uint64_t a = w[wi];
  uint64_t b = w[wi + 1];
  return (a << sh) | ((b >> (63 - sh)) >> 1);
}
// Forward k-mer (K <= 32) whose first base is at g: first base in the high bits.
__device__ __forceinline__ uint64_t kmer64_at(const uint64_t* __restrict__ seq, uint64_t g, int K) {
  return stream_bits64(seq, 2 * g) >> (64 - 2 * K);
}
// Forward k-mer for 32 < K <= 63 as (hi, lo) of the 2K-bit value.
__device__ __forceinline__ Key128 kmer128_at(const uint64_t* __restrict__ seq, uint64_t g, int K) {
  Key128 r;
  r.lo = stream_bits64(seq, 2 * g + 2 * (uint64_t)K - 64);
  r.hi = stream_bits64(seq, 2 * g) >> (128 - 2 * K);
  return r;
}
__device__ __forceinline__ uint32_t base_at(const uint64_t* __restrict__ seq, uint64_t g) {
  return (uint32_t)((seq[g >> 5] >> (62 - 2 * (g & 31))) & 3);
}
__device__ __forceinline__ bool is_n(const uint64_t* __restrict__ nmask, uint64_t g) {
  return (nmask[g >> 6] >> (63 - (g & 63))) & 1;
}
// True iff bases [g, g+K) hold no placeholder (1 <= K <= 63).
__device__ __forceinline__ bool window_clear(const uint64_t* __restrict__ nmask, uint64_t g, int K) {
  return (stream_bits64(nmask, g) >> (64 - K)) == 0;
}

// ------------------------------------------------------------ reverse complement
__device__ __forceinline__ uint64_t swap_pairs(uint64_t x) {
  return ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
}
// Reverse complement of a 2K-bit k-mer (K <= 32): complement = XOR 3 per
// base; reversing bit order then swapping each bit pair reverses the bases.
__device__ __forceinline__ uint64_t rc64(uint64_t x, int K) {
  return swap_pairs(__brevll(~x)) >> (64 - 2 * K);
}
__device__ __forceinline__ Key128 rc128(const Key128& x, int K) {
  uint64_t nh = swap_pairs(__brevll(~x.lo));  // reversed value, high word
  uint64_t nl = swap_pairs(__brevll(~x.hi));
  uint32_t s = 128 - 2 * K;  // 2..62 for K in [33,63]
  Key128 r;
  r.lo = (nl >> s) | (nh << (64 - s));
  r.hi = nh >> s;
  return r;
}
__device__ __forceinline__ uint64_t kmer_min(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __forceinline__ Key128 kmer_min(const Key128& a, const Key128& b) { return key_lt(b, a) ? b : a; }

// 8-bit radix digit at bit `sh` (sh multiple of 8).
__device__ __forceinline__ uint32_t key_digit(uint64_t k, int sh) { return (uint32_t)(k >> sh) & 255u; }
__device__ __forceinline__ uint32_t key_digit(const Key128& k, int sh) {
  return (uint32_t)(sh >= 64 ? (k.hi >> (sh - 64)) : (k.lo >> sh)) & 255u;
}

// ------------------------------------------------------------ hashes
// Minimizer order: fmix32 (a bijection on u32) of the canonical m-mer.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ uint32_t mmer_order(uint32_t canon) { return fmix32(canon + 0x9E3779B9u); }

// xxh3_64 (seed 0) of the 8 little-endian bytes of v: the XXH3 4..8-byte
// path (bitflip of default-secret bytes 8..23, rrmxmx finalizer).
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
__device__ __forceinline__ uint64_t xxh3_u64le(uint64_t v) {
  uint64_t in64 = (v >> 32) + (v << 32);
  uint64_t h = in64 ^ (0x1cad21f72c81017cULL ^ 0xdb979083e96dd4deULL);
  h ^= rotl64(h, 49) ^ rotl64(h, 24);
  h *= 0x9FB21C651E98DF25ULL;
  h ^= (h >> 35) + 8;
  h *= 0x9FB21C651E98DF25ULL;
  return h ^ (h >> 28);
}
__device__ __forceinline__ uint32_t bucket_hash(uint64_t v) { return (uint32_t)(xxh3_u64le(v) >> 48); }

// splitmix64 finalizer; seeds from a Weyl sequence.
__host__ __device__ constexpr uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__host__ __device__ constexpr uint64_t bbh_seed(int i) { return mix64(0x243F6A8885A308D3ULL + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL); }
constexpr uint64_t BBH_S1 = bbh_seed(0);
constexpr uint64_t FIB_MUL = 0x9E3779B97F4A7C15ULL, WIDE_MUL = 0xC2B2AE3D27D4EB4FULL;

// Segment of a key inside its bin: the top sb bits of a multiplicative
// (Fibonacci) hash, which spreads each bin evenly over its 2^sb segments.
__device__ __forceinline__ uint64_t seg_hash(uint64_t k) { return k * FIB_MUL; }
__device__ __forceinline__ uint64_t seg_hash(const Key128& k) { return (k.lo ^ (k.hi * WIDE_MUL)) * FIB_MUL; }
template <typename KeyT>
__device__ __forceinline__ uint32_t seg_sub(const KeyT& k, int sb) {
  return sb ? (uint32_t)(seg_hash(k) >> (64 - sb)) : 0u;
}
// BBHash level hashes: one 64-bit state per key and level, g_l = h1 + l * d
// (mod 2^64), h1 = (key ^ S1) * M1 (wide keys: (lo ^ mix64(hi ^ S1)) *
// M1), d = rotl(h1, 21) | 1. The position among n < 2^32 bits is
// floor(x * n / 2^32) with x = fmix32(hi32(g_l) ^ lo32(g_l)); the fmix32 step
// keeps the levels independent (top bits of g_l alone would keep two keys
// with close h1 and close d on one bit at every level).
constexpr uint64_t BBH_M1 = 0xD6E8FEB86659FD93ULL;
__device__ __forceinline__ uint64_t bbh_h1(uint64_t k) { return (k ^ BBH_S1) * BBH_M1; }
// For wide keys the high word goes through a full mixer before the fold: a
// linear fold (lo ^ hi * C) gives one h1 to pairs of 63-mers that differ only
// near the top of both words (bases 0 and 32, say: hi * C then changes in its
// top bits alone, which a change of lo can cancel), and two keys with one h1
// share every level position, so no level could separate them.
__device__ __forceinline__ uint64_t bbh_h1(const Key128& k) { return (k.lo ^ mix64(k.hi ^ BBH_S1)) * BBH_M1; }
__device__ __forceinline__ uint64_t bbh_d(uint64_t h1) { return ((h1 << 21) | (h1 >> 43)) | 1ULL; }
__device__ __forceinline__ uint32_t bbh_pos(uint64_t g, uint32_t n) {
  const uint32_t x = fmix32((uint32_t)(g >> 32) ^ (uint32_t)g);
  return __umulhi(x, n);
}

constexpr int BBH_MAX_LEVELS = 64;
constexpr uint64_t BBH_BLOCK_BITS = 448;  // 7 words of bits + 1 rank word = one 64-byte line

// ------------------------------------------------------------ MPHF (minimizer-partitioned BBHash)
// The distinct keys are partitioned exactly as P3a left them: with h =
// xxh3_64(minimizer value), segment (bin, s), bin = bucket_to_bin[h >> 48]
// (its bucket) and s = the next sb bits of h, so every k-mer of a super
// k-mer lies in one segment.
// Each segment holds its own BBHash: level l has n_l = roundup64(max(2 t_l, 1))
// bits with t_0 = D_s (the segment's key count) and t_{l+1} = ceil(9 t_l / 20);
// a key lands at pos = bbh_pos(g_l(key), n_l) and is placed at the first
// level where no other remaining key lands on the same bit. Levels are
// concatenated bit-contiguously in 64-byte blocks {u64 rank_before, u64 bits[7]}
// (rank local to the segment). value = key_off[segment] + local rank.
__host__ __device__ __forceinline__ uint64_t bbh_level_n(uint64_t t) {
  uint64_t n = 2 * t;
  if (n == 0) n = 1;
  return (n + 63) & ~63ULL;
}
__host__ __device__ __forceinline__ uint64_t bbh_next_t(uint64_t t) { return (9 * t + 19) / 20; }

struct MphfView {
  const uint16_t* b2b;      // 65536 bucket -> bin
  const uint64_t* key_off;  // n_segs + 1
  const uint64_t* blk_off;  // n_segs + 1 (in blocks)
  const uint64_t* blocks;   // 8 words per block
  int K, m, sb;
};

// 2K-bit value shifted right by s, low 32 bits (s < 128).
__device__ __forceinline__ uint32_t key_shr32(uint64_t k, int s) { return (uint32_t)(k >> s); }
__device__ __forceinline__ uint32_t key_shr32(const Key128& k, int s) {
  if (s >= 64) return (uint32_t)(k.hi >> (s - 64));
  if (s == 0) return (uint32_t)k.lo;
  return (uint32_t)((k.lo >> s) | (k.hi << (64 - s)));
}
__device__ __forceinline__ uint64_t rc_any(uint64_t x, int K) { return rc64(x, K); }
__device__ __forceinline__ Key128 rc_any(const Key128& x, int K) { return rc128(x, K); }

// Minimizer value of a k-mer given in canonical form (the value is the same
// for both orientations): the canonical m-mer of least order among its w m-mers.
template <typename KeyT>
__device__ __forceinline__ uint32_t minimizer_value(const KeyT& c, int K, int m) {
  const KeyT r = rc_any(c, K);
  const uint32_t M = (uint32_t)((1ULL << (2 * m)) - 1);
  const int w = K - m + 1;
  uint32_t best = 0xFFFFFFFFu, bestv = 0;
  for (int j = 0; j < w; j++) {
    uint32_t f = key_shr32(c, 2 * (K - m - j)) & M;
    uint32_t b = key_shr32(r, 2 * j) & M;
    uint32_t cv = f < b ? f : b;
    uint32_t o = mmer_order(cv);
    if (o < best) {
      best = o;
      bestv = cv;
    }
  }
  return bestv;
}
// A super k-mer's segment inside its bin from the 16 hash bits below its
// bucket (bits 47..32 of xxh3_64(minimizer)): their top sb bits.
__device__ __forceinline__ uint32_t rec_seg(uint32_t sub16, int sb) { return sb ? sub16 >> (16 - sb) : 0u; }
// Segment bits of a minimizer value: the sb bits of xxh3_64(v) just below
// its 16 bucket bits.
__device__ __forceinline__ uint32_t seg_sub_min(uint64_t h, int sb) { return sb ? (uint32_t)(h >> (48 - sb)) & ((1u << sb) - 1) : 0u; }
template <typename KeyT>
__device__ __forceinline__ uint64_t mphf_segment(const MphfView& v, const KeyT& c) {
  const uint64_t h = xxh3_u64le((uint64_t)minimizer_value(c, v.K, v.m));
  const uint32_t bin = v.b2b[h >> 48];
  return ((uint64_t)bin << v.sb) | seg_sub_min(h, v.sb);
}
// Rank of a present key inside its segment (segment already known).
template <typename KeyT>
__device__ __forceinline__ uint64_t mphf_local_rank(const uint64_t* __restrict__ base, uint64_t t, const KeyT& c) {
  uint64_t off = 0;
  uint64_t h = bbh_h1(c);
  const uint64_t d = bbh_d(h);
  for (int l = 0; l < BBH_MAX_LEVELS && t > 0; l++, h += d) {
    const uint64_t n = bbh_level_n(t);
    const uint64_t pos = off + bbh_pos(h, (uint32_t)n);
    const uint64_t q = pos / BBH_BLOCK_BITS;
    const uint32_t b = (uint32_t)(pos - q * BBH_BLOCK_BITS);
    const uint64_t* B = base + q * 8;
    const uint32_t wi = b >> 6;
    const uint64_t wd = B[1 + wi];
    if ((wd >> (b & 63)) & 1) {
      uint64_t r = B[0] + (uint64_t)__popcll(wd & ((1ULL << (b & 63)) - 1));
      for (uint32_t u = 0; u < wi; u++) r += (uint64_t)__popcll(B[1 + u]);
      return r;
    }
    off += n;
    t = bbh_next_t(t);
  }
  return ~0ULL;
}
// Exact value of a present key; for an absent key, some value, or ~0 if it
// misses every level (callers that must notice a missing key test >= D).
// Levels are read only while they lie inside the segment's blocks: a segment
// stores the levels its build needed (until every key was placed), and the
// level sizes never reach 0, so an absent key would otherwise walk on into
// the next segment's blocks or past the array.
template <typename KeyT>
__device__ __forceinline__ uint64_t mphf_find(const MphfView& v, const KeyT& c) {
  const uint64_t seg = mphf_segment(v, c);
  const uint64_t ko = v.key_off[seg];
  uint64_t t = v.key_off[seg + 1] - ko;
  const uint64_t b0 = v.blk_off[seg];
  const uint64_t lim = (v.blk_off[seg + 1] - b0) * BBH_BLOCK_BITS;
  const uint64_t* base = v.blocks + b0 * 8;
  uint64_t off = 0;
  uint64_t h = bbh_h1(c);
  const uint64_t d = bbh_d(h);
  for (int l = 0; l < BBH_MAX_LEVELS && t > 0; l++, h += d) {
    const uint64_t n = bbh_level_n(t);
    if (off + n > lim) break;
    const uint64_t pos = off + bbh_pos(h, (uint32_t)n);
    const uint64_t q = pos / BBH_BLOCK_BITS;
    const uint32_t b = (uint32_t)(pos - q * BBH_BLOCK_BITS);
    const uint64_t* B = base + q * 8;
    const uint32_t wi = b >> 6;
    const uint64_t wd = B[1 + wi];
    if ((wd >> (b & 63)) & 1) {
      uint64_t r = ko + B[0] + (uint64_t)__popcll(wd & ((1ULL << (b & 63)) - 1));
      for (uint32_t u = 0; u < wi; u++) r += (uint64_t)__popcll(B[1 + u]);
      return r;
    }
    off += n;
    t = bbh_next_t(t);
  }
  return ~0ULL;
}
// Mphf::hash (mphf.rs:268-270): get(&key).unwrap_or(0) — a key that misses
// every level maps to 0.
template <typename KeyT>
__device__ __forceinline__ uint64_t mphf_lookup(const MphfView& v, const KeyT& c) {
  const uint64_t r = mphf_find(v, c);
  return r == ~0ULL ? 0 : r;
}

// ------------------------------------------------------------ DFA (state.rs / classify.rs)
// 5-bit codes (state.rs:92-113): 0 unvisited, 3 MIMO, 4+b MISO, 8+f SIMO,
// 12..15 outputted, 16+4f+b SISO.
enum : uint32_t { CLS_SISO = 0, CLS_MISO = 1, CLS_SIMO = 2, CLS_MIMO = 3 };
__host__ __device__ __forceinline__ uint32_t st_class(uint32_t s) {  // STATE_CLASS_TABLE state.rs:249-270
  if (s == 3 || s == 15) return CLS_MIMO;
  if ((s >= 4 && s <= 7) || s == 13) return CLS_MISO;
  if ((s >= 8 && s <= 11) || s == 14) return CLS_SIMO;
  return CLS_SISO;
}
// Occurrence types of classify.rs:65-177: isolated or MIMO self loop,
// leftmost, rightmost (also the inverted-repeat self loop), internal.
enum : uint32_t { OBS_MIMO = 0, OBS_LEFT = 1, OBS_RIGHT = 2, OBS_INTERNAL = 3 };
// ------------------------------------------------------------ observation masks
// The DFA of classify.rs:147-518 is a join: a state is a function of the set
// of in-labels and the set of out-labels observed for the k-mer (canonical
// orientation; "none" marks a missing neighbour or a forced MIMO), and every
// transition adds one (in, out) pair. An observation is therefore a 10-bit
// mask (bits 0-4 in-label A,C,G,T,none; bits 5-9 out-label) and the joined
// state of a k-mer is the OR of its occurrences' masks.
constexpr uint32_t LBL_NONE = 4;
// The type's bits say which neighbours count: bit 1 (RIGHT, INTERNAL) the
// previous base, bit 0 (LEFT, INTERNAL) the next one; in reverse orientation
// they swap sides and complement. Selects only, no branches.
__host__ __device__ __forceinline__ uint32_t obs_mask(uint32_t type, bool fwd, uint32_t pb, uint32_t nb) {
  const bool hp = (type >> 1) & 1u, hn = type & 1u;
  const uint32_t a = fwd ? (hp ? pb : LBL_NONE) : (hn ? 3 - nb : LBL_NONE);
  const uint32_t b = fwd ? (hn ? nb : LBL_NONE) : (hp ? 3 - pb : LBL_NONE);
  return (1u << a) | (1u << (5 + b));
}
// Orders one wave's LDS accesses before and after it (a wave needs no block
// barrier to see its own LDS writes).
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
// State::from_vertex (state.rs:92-113) of the joined label sets.
// Arithmetic only (lanes of a wave hold different masks, and the compiler
// turns chained conditions into branches that run every path): a side is
// single when its label set is one of bits 0-3 (bits 1, 2, 4, 8 of 0x116);
// the base of the state comes from (single in, single out) as a byte of
// 0x10040803, and the labels are added on the single sides.
__host__ __device__ __forceinline__ uint32_t mask_state(uint32_t m) {
  const uint32_t I = m & 31u, O = (m >> 5) & 31u;
  const uint32_t si = (0x116u >> I) & 1u, so = (0x116u >> O) & 1u;
  const uint32_t f = (uint32_t)__builtin_ctz(I | 32u) & 3u, b = (uint32_t)__builtin_ctz(O | 32u) & 3u;
  const uint32_t s = ((0x10040803u >> (8 * (si | (so << 1)))) & 0xFFu) | ((f * si) << (2 * so)) | (b * so);
  return (uint32_t)(m != 0) * s;
}
// The label sets a 5-bit state stands for, as a mask that joins like the
// DFA: a single side keeps its label bit, a multi side is the "none" bit
// (absorbing under OR). mask_state(state_mask(s) | m) is the state after
// observation m (classify.rs:202-511 branch by branch); unvisited is 0.
__host__ __device__ __forceinline__ uint32_t state_mask(uint32_t s) {
  if (s == 0) return 0;
  uint32_t I = 1u << LBL_NONE, O = 1u << LBL_NONE;  // MIMO (3); outputted states never reach P4
  if (s >= 16) {
    I = 1u << ((s >> 2) & 3u);
    O = 1u << (s & 3u);
  } else if (s >= 4 && s <= 7) {
    O = 1u << (s & 3u);  // MISO: in multi, out single
  } else if (s >= 8 && s <= 11) {
    I = 1u << (s & 3u);  // SIMO: in single, out multi
  }
  return I | (O << 5);
}

// ------------------------------------------------------------ P2 records
// A routed super k-mer of L bases is a record of RW words holding positions
// 0..L+1 (one flanking base each side) MSB-first, 32 per word, with L in bits
// 0-7 of the last word and the flank flags in bits 8-9. Position
// 32 * RW - 6 is the last one clear of those bits, so L <= 32 * RW - 7. A
// super k-mer has at most w = K - m + 1 k-mers, so L <= 2K - m (the context
// refuses a minimizer length that would exceed the record).
__host__ __device__ constexpr int rec_cap(int rw) { return 32 * rw - 7; }

// ------------------------------------------------------------ wave helpers (wave64)
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  uint32_t l = __lane_id();
  return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

}  // namespace cf1g
