// Learned rank index: model records and views shared by the kernels and the
// C ABI (plain structs; cf1g_lidx_export hands the records out as they are).
#pragma once
#include <stdint.h>

namespace cf1g {

// ------------------------------------------------------------------ model records
struct RootPiece {
  float A, B;
  uint32_t lo, hi;
};
struct LeafRec {
  uint64_t base;  // first key of the leaf
  uint32_t q;     // rank of the first key
  uint32_t n;     // keys in the leaf (ranks [q, q + n))
  float S, I;     // pred = q + floor(fma(S, (float)(key - base), I) + 0.5)
  uint16_t elo, ehi;
  uint8_t shift;  // fingerprint = (key - base) >> shift (< 2^16 inside the leaf)
  uint8_t flags;  // bit 0: window above threshold, resolve through the BBHash
  uint16_t pad;
};
static_assert(sizeof(LeafRec) == 32, "leaf record is 32 bytes");

struct LidxView {
  const RootPiece* root;
  const LeafRec* leaf;
  const uint64_t* keys;      // sorted
  const uint16_t* fp;        // fingerprint of each sorted key (monotone inside a leaf)
  const uint32_t* bbh2rank;  // rank of the key whose BBHash value is v (fallback leaves only)
  uint64_t D;
  uint32_t L, R;
  int s;  // root shift: piece = key >> s (< R for keys of 2k bits)
};

// Segment views. LEVEL 0 = root pieces (x = key - (r << s), y = rank*L/D -
// lo_r); LEVEL 1 = leaves (x = key - base_j, y = rank - q_j).
struct SegCtx {
  const uint64_t* keys;
  const uint64_t* seg;  // segment starts (n_segs + 1)
  const uint32_t* lo;   // root: lo_r
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
