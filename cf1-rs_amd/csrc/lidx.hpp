// Learned rank index: model records and views shared by the kernels and the
// C ABI (plain structs; cf1g_lidx_export hands the records out as they are).
#pragma once
#include <stdint.h>

namespace cf1g {

// ------------------------------------------------------------------ key warp
// Before the linear models a key is mapped to u in [0, 2^64) by a learned
// order-0 model of its bytes (four bases each): interval (lo, w) starts at
// (0, 2^64 - 1); group g (bits [K2 - 8(g+1), K2 - 8g), up to 3 groups) with
// symbol s moves lo by f(w, cum_g[s]) and sets w = f(w, cum_g[s+1]) -
// f(w, cum_g[s]), f(w, c) = floor(w c / 2^32), cum_g[256] = 2^32; the bits
// below the groups add mulhi(w, rest << (64 - rest_bits)). The intervals
// nest, so u is non-decreasing in the key, and with cum_g from the keys' own
// byte counts (+1 each) u is close to uniform: the leaves then see the
// keys' random spacing rather than the genome's base composition.
constexpr int WARP_MAX_GROUPS = 3;

// ------------------------------------------------------------------ model records
// Root piece r (warped keys with u >> s == r): leaf = leaf0 + clamp(floor(fma(A,
// x, B)), 0, n_r - 1), x = key - (r << s), n_r = leaf0 of piece r + 1 minus
// leaf0 (at least 1; no leaf holds keys of two pieces); p = rank of the
// piece's first key. R + 1 records, the last one {0, 0, L + R, D}.
struct RootPiece {
  float A, B;
  uint32_t leaf0;
  uint32_t p;
};
// Leaf: predicted rank = p_r + floor(fma(|S|, x, I) + 0.5) in the piece's
// coordinates (x = u - (r << s) in FP32 from its top 32 bits, the local rank
// clamped to +-2^30). The sign bit of S marks a leaf whose recorded error window
// exceeds the threshold: its keys resolve through the BBHash.
struct LeafRec {
  float S, I;
};
static_assert(sizeof(LeafRec) == 8, "leaf record is 8 bytes");

// Lines: line i holds ranks [iC, iC + n_i) and the sentinel (rank iC + C)
// when there is one, as a little-endian bit stream of u64 words:
//   bits 0..61 k0 (key of rank iC), bit 62 overflow (keys searched instead),
//   bit 63 code width w (0: 3 bits, 1: 4 bits)
//   bits 64..69 t (cell shift), bits 70..79 ulen
//   bits 80..87 the sentinel's prefix: the top 8 bits of its offset in its
//   cell ((k_sent - k0) mod 2^t, shifted up when t < 8; 0 without sentinel)
//   unary [88, 88 + ulen): per cell c = 0 .. (k_last - k0) >> t, a 1 per key
//   with (key - k0) >> t == c, then a 0
//   w-bit codes after the unary, one per key that shares the cell of the key
//   before it: min(t - bitlen(rel_j ^ rel_{j-1}), 2^w - 1); the highest bit
//   where the two differ is t - 1 - code, and 2^w - 1 means "search the
//   run's keys". A line takes 4-bit codes when they fit and 3-bit codes would
//   leave a code to search; a run of more than 64 / w codes is searched.
// In the run that holds the sentinel the prefixes decide: a query with a
// larger one lies above the sentinel (the next line), a smaller one below it
// (the run without the sentinel), an equal one searches the run.
constexpr int LINE_HDR_BITS = 88;
constexpr int LINE_SP_BITS = 8;

struct LidxView {
  const RootPiece* root;     // R + 1
  const LeafRec* leaf;       // LT = L + R
  const uint64_t* keys;      // sorted (D)
  const uint64_t* lines;     // n_lines x line_words
  const uint32_t* bbh2rank;  // rank of the key whose BBHash value is v (fallback leaves only)
  const uint32_t* warp;      // warp_groups x 256 cumulative byte frequencies (x 2^-32)
  uint64_t D;
  uint32_t R, LT;
  uint32_t C, n_lines;
  uint32_t magic, mshift;  // line = (t + ((pred - t) >> 1)) >> mshift, t = umulhi(pred, magic)
  uint32_t max_steps;      // lines one lookup may visit
  int s;                   // root shift: piece = u >> s
  int line_words;          // 8 (64-byte lines) or 16 (128-byte lines)
  int key_bits;            // 2k: wider keys are absent
  int warp_groups;
  bool fallback;           // some leaf resolves through the BBHash
};

// Searches deferred by the lookup kernel: each wave of its grid owns a
// segment of cap entries (3 words: key, query index, line | run code << 32)
// and writes its count to cnt[wave]; k_lidx_search resolves them afterwards.
struct LidxQueue {
  uint64_t* ent;  // waves x cap x 3 words
  uint32_t* cnt;  // entries per wave
  uint32_t cap;   // 0: every search in place
};

// Segment views over the warped keys. LEVEL 0 = root pieces (x = u - (r << s),
// y = rank*L/D - lo_r); LEVEL 1 = leaves (x = u - first u of the leaf, y =
// rank - q_j).
struct SegCtx {
  const uint64_t* keys;  // warped, non-decreasing
  const uint64_t* seg;  // segment starts (n_segs + 1)
  const uint32_t* lo;   // root: lo_r = floor(p_r L / D)
  uint64_t D;
  uint32_t L;
  int s;
};
struct LidxTrainBufs {
  float *mu, *isd, *w, *c, *resid;
  double *gw, *gc, *loss;
  uint32_t* seg_of_tile;
};
}  // namespace cf1g
