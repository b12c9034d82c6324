// P1 / P2 kernels: 2-bit packing of the flat stream, canonical-minimizer
// histogram (minimizer.rs:27-110), bucket->bin partition
// (minimizer.rs:116-184) and super k-mer routing into per-bin HBM segments
// (superkmer.rs:21-210).
#include "device.hpp"
#include "launch.hpp"

namespace cf1g {

// ------------------------------------------------------------------ pack
// Per 64 stream positions: one N-mask word and two sequence words.
// gstart[r] = stream position of record r's first base; gstart[n_rec] = the
// position just past the last separator. Non-ACGT bytes and separators are
// placeholders (dna.rs:64-92); their 2-bit code is 0.

// 2-bit code of an ACGT byte (either case: A 0, C 1, G 2, T 3), 4 for any
// other byte.
__device__ __forceinline__ uint32_t base_code(uint32_t c) {
  // upper-cased letter minus 'A' in {0 (A), 2 (C), 6 (G), 19 (T)}
  const uint32_t d = (c & 0xDFu) - 0x41u;
  const uint32_t ok = (uint32_t)(d < 32u) & (0x80045u >> (d & 31u));
  return (ok & 1u) ? ((c >> 1) ^ (c >> 2)) & 3u : 4u;
}
// Four bytes at once: the 2-bit codes (MSB-first, base 0 in bits 7..6) and
// the placeholder bits (base 0 in bit 3). ((x >> 1) ^ (x >> 2)) & 3 is the
// code of an ACGT byte of either case; a byte is ACGT iff its upper-cased
// value equals the letter of that code (v_perm of the codes into "ACGT").
// The multiplies gather one bit (two for codes) from each byte lane into the
// top byte; the partial products below it never overlap.
__device__ __forceinline__ void pack4(uint32_t x, uint32_t& cb, uint32_t& n4) {
  const uint32_t u = x & 0xDFDFDFDFu;
  uint32_t c = ((x >> 1) ^ (x >> 2)) & 0x03030303u;
  const uint32_t bad = u ^ __builtin_amdgcn_perm(0u, 0x54474341u, c);  // bytes 'A' 'C' 'G' 'T' by code
  const uint32_t nz = ((((bad & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | bad) & 0x80808080u) >> 7;  // 1 per non-ACGT byte
  n4 = (nz * 0x08040201u) >> 24;
  c &= ~(nz * 3u);
  cb = (c * 0x40100401u) >> 24;
}
// The 64 bytes at ascii[a ..] as one packed word: 17 aligned dword loads
// (bytes up to a + 67 are read), 4 bytes at a time.
__device__ __forceinline__ void pack64(const uint8_t* __restrict__ ascii, uint64_t a, uint64_t& hw, uint64_t& lw, uint64_t& nm) {
  const uint32_t* d = (const uint32_t*)(ascii + (a & ~3ULL));
  const uint32_t sh = (uint32_t)(a & 3);
  uint32_t v[17];
#pragma unroll
  for (int i = 0; i < 17; i++) v[i] = d[i];
  uint32_t cw[4] = {0, 0, 0, 0}, mw[2] = {0, 0};
#pragma unroll
  for (int i = 0; i < 16; i++) {
    uint32_t cb, n4;
    pack4(__builtin_amdgcn_alignbyte(v[i + 1], v[i], sh), cb, n4);  // bytes a + 4i ..
    cw[i >> 2] |= cb << (24 - 8 * (i & 3));
    mw[i >> 3] |= n4 << (28 - 4 * (i & 7));
  }
  hw = ((uint64_t)cw[0] << 32) | cw[1];
  lw = ((uint64_t)cw[2] << 32) | cw[3];
  nm = ((uint64_t)mw[0] << 32) | mw[1];
}
// Mask of the code bits of positions s .. 63 of a packed word (position i in
// bits 127-2i .. 126-2i of hw:lw), as (high word, low word).
__device__ __forceinline__ void tail_mask(int s, uint64_t& mh, uint64_t& ml) {
  const int cb = 128 - 2 * s;
  if (cb >= 64) {
    ml = ~0ULL;
    mh = cb == 128 ? ~0ULL : ((1ULL << (cb - 64)) - 1);
  } else {
    ml = (1ULL << cb) - 1;
    mh = 0;
  }
}
// A separator (placeholder) inserted at position s: positions s .. 62 move
// up by one, the last one drops out.
__device__ __forceinline__ void insert_sep(int s, uint64_t& hw, uint64_t& lw, uint64_t& nm) {
  uint64_t mh, ml;
  tail_mask(s, mh, ml);
  const uint64_t ah = hw & mh, al = lw & ml;
  hw = (hw & ~mh) | (ah >> 2);
  lw = (lw & ~ml) | (al >> 2) | (ah << 62);
  const uint64_t na = s == 0 ? ~0ULL : ((1ULL << (64 - s)) - 1);
  nm = (nm & ~na) | ((nm & na) >> 1) | (1ULL << (63 - s));
}
// One lane per 64-position word. Records lie back to back in `ascii`, so
// position p holds byte p - (separators before p): a word whose positions
// are all bases of one record reads its 64 bytes and converts them 4 at a
// time; a word that holds up to 4 separators converts the bytes from its
// first position's byte on and inserts a placeholder at each separator, the
// positions past the stream end becoming placeholders too. The first word,
// words whose bytes would run past the buffer's last 64 bytes and words with
// more separators go position by position.
__global__ void __launch_bounds__(256) k_pack(const uint8_t* __restrict__ ascii, const uint64_t* __restrict__ rec_off,
                                              const uint64_t* __restrict__ gstart, uint64_t n_rec, uint64_t n_words64,
                                              uint64_t* __restrict__ seq, uint64_t* __restrict__ nmask) {
  const uint64_t wv0 = (((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64;  // first word of the wave
  if (wv0 >= n_words64) return;
  // largest r with gstart[r] <= the wave's first position (r = n_rec past the
  // last record), then each lane moves on to its own word's record
  const uint64_t gw = wv0 * 64;
  uint64_t lo = 0, hi = n_rec + 1;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (gstart[mid] <= gw) lo = mid; else hi = mid;
  }
  const uint64_t w = wv0 + __lane_id();
  if (w >= n_words64) return;
  const uint64_t g0 = w * 64;
  uint64_t r = lo;
  while (r < n_rec && g0 >= gstart[r + 1]) r++;
  uint64_t hw = 0, lw = 0, nm = 0;
  bool done = false;
  if (r < n_rec && g0 >= gstart[r] && g0 + 64 < gstart[r + 1]) {
    pack64(ascii, rec_off[r] + (g0 - gstart[r]), hw, lw, nm);
    done = true;
  } else if (r < n_rec && g0 >= gstart[r]) {
    const uint64_t a = rec_off[r] + (g0 - gstart[r]);  // byte of position g0 (or of the next base, at a separator)
    int sep[4], k = 0;
    bool few = true;
    for (uint64_t q = r + 1; q <= n_rec && gstart[q] - 1 < g0 + 64; q++) {
      if (k == 4) {
        few = false;
        break;
      }
      sep[k++] = (int)(gstart[q] - 1 - g0);
    }
    if (few && a + 64 <= rec_off[n_rec]) {
      pack64(ascii, a, hw, lw, nm);
      for (int t = 0; t < k; t++) insert_sep(sep[t], hw, lw, nm);
      const uint64_t e = gstart[n_rec];  // positions from here on lie past the stream
      if (g0 + 64 > e) {
        const int er = (int)(e - g0);
        uint64_t mh, ml;
        tail_mask(er, mh, ml);
        hw &= ~mh;
        lw &= ~ml;
        nm |= er == 0 ? ~0ULL : ((1ULL << (64 - er)) - 1);
      }
      done = true;
    }
  }
  if (!done) {
    uint64_t g = g0;
    for (int i = 0; i < 64; i++, g++) {
      while (r < n_rec && g >= gstart[r + 1]) r++;
      uint32_t code = 4;
      if (r < n_rec && g >= gstart[r] && g + 1 < gstart[r + 1]) code = base_code(ascii[rec_off[r] + (g - gstart[r])]);
      const uint64_t b = code & 3u;
      if (i < 32) hw |= b << (62 - 2 * i);
      else lw |= b << (62 - 2 * (i - 32));
      nm |= (uint64_t)(code >> 2) << (63 - i);
    }
  }
  seq[2 * w] = hw;
  seq[2 * w + 1] = lw;
  nmask[w] = nm;
}

// ------------------------------------------------------------------ minimizer scan (P1)
// Project minimizer scheme (same on the CPU oracle): order of an m-mer =
// fmix32(canonical m-mer + c); the minimizer of the k-mer at i is the leftmost
// least-order m-mer in [i, i+w); a super k-mer is a maximal run of valid
// k-mers sharing the minimizer position, at most w long; it routes by
// bucket_hash(value).
//
// k_p1_scan walks tiles of SK_TP k-mer positions (work-groups loop over
// tiles). Per tile it computes the m-mer orders of positions [t0-1, t0+T+2w)
// into LDS (8 consecutive m-mers per thread from one 64-bit window and its
// reverse complement), each position's minimizer index (sliding window) and
// validity (no placeholder in its K bases), and emits one entry per super
// k-mer starting in the tile into the tile's slot of the entry array:
//   entry = start position << 24 | length (k-mers) << 16 | bucket.
// tile_cnt[tile] = entries of the tile, tile_w[tile] = its valid k-mers.
//
// Chunk edges: the reference scans a record of CHUNK_BASES (4,000,000) bases
// or more in chunks that start every CHUNK_BASES bases (minimizer.rs:57-86,
// superkmer.rs:57-88), each as a sequence of its own. P2's chunks overlap by
// K-1 bases, so a super k-mer always starts at a chunk's first k-mer: the
// scan breaks runs at each edge (record start + j * CHUNK_BASES, j >= 1).
// P1's chunks do not overlap, so the k-mers across an edge are not counted:
// a super k-mer that starts less than K positions before an edge (all its
// k-mers span it) is routed but left out of the histogram; its bucket goes
// to the `excl` list, which k_p1_excl takes off the counts. `edges` is
// sorted and ends with a UINT64_MAX sentinel.
constexpr int SK_WG = 256;
constexpr int SK_T = 2048;  // slots per tile: entries and buckets arrays, LDS
constexpr int SK_Q = 8;     // consecutive m-mers / positions per thread step
constexpr int SK_MAXW = 64;  // w <= 49 for k <= 63
// k-mer positions per tile: the tile's m-mer orders cover T + 2w slots,
// which at T = SK_T - 2 SK_MAXW fit one step of 8 slots per thread (with
// SK_T positions the first wave made a second pass over the last 2w slots
// while the other three waited at the barrier)
constexpr int SK_TP = SK_T - 2 * SK_MAXW;
constexpr int SK_NJ = SK_T + 2 * SK_MAXW;
constexpr int MAX_BINS = 4096;

// LDS index of tile slot j: a gap after every 8 slots, so that lanes working
// on 8 consecutive slots each (stride 9 words apart) hit distinct banks.
__device__ __forceinline__ int PJ(int j) { return j + (j >> 3); }
constexpr int SK_NJP = SK_NJ + SK_NJ / 8 + 8;
constexpr int SK_NPP = (SK_T + SK_MAXW + 8) * 9 / 8 + 8;
constexpr int SK_BRK = (SK_T + 2 * SK_MAXW) / 8 + 16;  // bytes of break flags (bit j%8 of byte j/8)
// pidx: the minimizer slot of the k-mer at each slot, bit 15 set when the
// k-mer is not valid (a placeholder among its bases). The canonical m-mer
// values are not kept: a super k-mer start reads its minimizer's m-mer from
// the stream again, which leaves the tile at ~15 KiB. The kernel is compiled
// for 8 work-groups per CU (64 VGPRs, a 12-byte spill): 1.10 ms on the
// 300 Mbp genome, against 1.16 ms at 7 per CU and 1.48 ms at 5 with a
// 27 KiB tile.
constexpr uint16_t SK_INVALID = 0x8000;
struct SkTile {
  uint32_t ord[SK_NJP];
  uint16_t pidx[SK_NPP];
  uint64_t brk[SK_BRK / 8];
  uint64_t sh[4];
  uint64_t edge;  // first chunk edge at or after the tile's first k-mer position
};
static_assert(2 * SK_T <= 4 * SK_NJP, "the start list of a tile (u16 slots) fits in SkTile::ord");

__device__ __forceinline__ uint64_t sk_entry(uint64_t pos, uint32_t n, uint32_t bucket) {
  return (pos << 24) | ((uint64_t)n << 16) | (uint64_t)bucket;
}
__device__ __forceinline__ uint64_t sk_pos(uint64_t e) { return e >> 24; }
__device__ __forceinline__ uint32_t sk_len(uint64_t e) { return (uint32_t)(e >> 16) & 255u; }
__device__ __forceinline__ uint32_t sk_bucket(uint64_t e) { return (uint32_t)e & 0xFFFFu; }

__device__ __forceinline__ uint64_t sk_scan_excl(uint64_t v, uint64_t* sh4, uint64_t* total) {
  uint32_t lane = __lane_id(), wv = threadIdx.x >> 6;
  uint64_t x = v;
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up(x, d, 64);
    if ((int)lane >= d) x += y;
  }
  if (lane == 63) sh4[wv] = x;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int q = 0; q < SK_WG / 64; q++) {
    uint64_t t = sh4[q];
    if (q < (int)wv) off += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return off + x - v;
}
__device__ __forceinline__ uint64_t min64(uint64_t a, uint64_t b) { return a < b ? a : b; }

__global__ void __launch_bounds__(SK_WG, 8) k_p1_scan(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                   uint64_t n_pos, uint64_t tile_begin, uint64_t n_tiles, int K, int m,
                                                   uint64_t* __restrict__ entries, uint32_t* __restrict__ buckets,
                                                   uint32_t* __restrict__ tile_cnt, uint32_t* __restrict__ tile_w,
                                                   const uint64_t* __restrict__ edges, uint32_t n_edges,
                                                   uint32_t* __restrict__ excl_cnt, uint16_t* __restrict__ excl, uint32_t excl_cap) {
  __shared__ SkTile S;
  const int w = K - m + 1;
  const uint32_t M = (uint32_t)((1ULL << (2 * m)) - 1);
  uint8_t* brk8 = (uint8_t*)S.brk;
  for (uint64_t tile = tile_begin + blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const uint64_t t0 = 1 + tile * SK_TP;  // first k-mer position of the tile
    const uint64_t B = t0 - 1;            // tile slot 0 <-> stream position B
    if (threadIdx.x == 0) {  // first edge >= t0 (edges are >= 4M apart: at most one in the tile's reach)
      uint32_t lo = 0, hi = n_edges;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (edges[mid] < t0) lo = mid + 1; else hi = mid;
      }
      S.edge = edges[lo];
    }
    // (1) m-mer orders of slots [0, T + 2w)
    const int nj = SK_TP + 2 * w;
    for (int j0 = threadIdx.x * SK_Q; j0 < nj; j0 += SK_WG * SK_Q) {
      const uint64_t x = stream_bits64(seq, 2 * (B + (uint64_t)j0));
      const uint64_t rx = rc64(x, 32);
#pragma unroll
      for (int q = 0; q < SK_Q; q++) {
        const uint32_t f = (uint32_t)(x >> (64 - 2 * m - 2 * q)) & M;
        const uint32_t r = (uint32_t)(rx >> (2 * q)) & M;
        const uint32_t c = f < r ? f : r;
        if (j0 + q < nj) S.ord[PJ(j0 + q)] = mmer_order(c);
      }
    }
    __syncthreads();
    // (2) minimizer slot and validity of the k-mers at slots [0, T + w + 1).
    // Window minima of (order, slot), leftmost on equal orders: for 8
    // consecutive windows [q, q+w-1] (w >= 8), the least of the suffix
    // [q, 7), the common core [7, w) and the prefix [w, q+w). Orders and
    // slots are tracked as two 32-bit values: scans towards higher slots keep
    // the incumbent on ties, scans towards lower slots take the newcomer, and
    // a left part beats a right part on ties.
    const int np = SK_TP + w + 1;
    const int span = K + SK_Q - 1;  // bases under the 8 windows of a step
    for (int j0 = threadIdx.x * SK_Q; j0 < np; j0 += SK_WG * SK_Q) {
      const uint64_t g = B + (uint64_t)j0;
      const uint64_t n0 = stream_bits64(nmask, g), n1 = stream_bits64(nmask, g + 64);
      // validity of the 8 k-mers: bases [g+q, g+q+K) of the N mask are bits
      // q .. q+K-1 of n0:n1; one test when no N lies under any of them
      const bool clean = span <= 64 ? (n0 >> (64 - span)) == 0 : (n0 == 0 && (n1 >> (128 - span)) == 0);
      uint32_t vbits = 0xFFu;
      if (!clean) {
        vbits = 0;
#pragma unroll
        for (int q = 0; q < SK_Q; q++) {
          const uint64_t hi = q ? ((n0 << q) | (n1 >> (64 - q))) : n0;
          vbits |= (uint32_t)((hi >> (64 - K)) == 0) << q;
        }
      }
      auto emit = [&](int q, uint32_t slot) {
        const int j = j0 + q;
        if (j < np) S.pidx[PJ(j)] = (uint16_t)(slot | ((vbits >> q) & 1u ? 0u : SK_INVALID));
      };
      if (w >= SK_Q) {
        // slot j0 + u is ord[PJ(j0) + u + (u >> 3)] (j0 is a multiple of 8)
        const uint32_t* op = &S.ord[PJ(j0)];
        uint32_t cb = op[SK_Q - 1], ci = SK_Q - 1;
        // (a rolled loop: fully unrolled with a guard per slot it ran
        // slower, 1.72 against 1.48 ms)
        for (int u = SK_Q; u < w; u++) {
          const uint32_t o = op[u + (u >> 3)];
          if (o < cb) {
            cb = o;
            ci = (uint32_t)u;
          }
        }
        uint32_t Lb[SK_Q], Li[SK_Q];  // least of [q, w)
        Lb[SK_Q - 1] = cb;
        Li[SK_Q - 1] = ci;
        uint32_t sb = 0, si = 0;
#pragma unroll
        for (int q = SK_Q - 2; q >= 0; q--) {
          const uint32_t o = op[q];
          if (q == SK_Q - 2 || o <= sb) {
            sb = o;
            si = (uint32_t)q;
          }
          const bool core = cb < sb;
          Lb[q] = core ? cb : sb;
          Li[q] = core ? ci : si;
        }
        // slot j0 + w + v is ord[PJ(j0 + w) + v + (((w & 7) + v) >> 3)]
        const int w7 = w & 7;
        const uint32_t* opw = &S.ord[PJ(j0 + w)];
        uint32_t pb = 0, pi = 0;
#pragma unroll
        for (int q = 0; q < SK_Q; q++) {
          const uint32_t bi = (q > 0 && pb < Lb[q]) ? pi : Li[q];
          emit(q, (uint32_t)j0 + bi);
          if (q < SK_Q - 1) {
            const uint32_t o = opw[q + ((w7 + q) >> 3)];
            if (q == 0 || o < pb) {
              pb = o;
              pi = (uint32_t)(w + q);
            }
          }
        }
      } else {
        auto key = [&](int u) -> uint64_t { return ((uint64_t)S.ord[PJ(j0 + u)] << 16) | (uint64_t)(j0 + u); };
#pragma unroll
        for (int q = 0; q < SK_Q; q++) {
          uint64_t b = ~0ULL;
          for (int u = 0; u < w; u++) b = min64(b, key(q + u));
          emit(q, (uint32_t)(b & 0xFFFF));
        }
      }
    }
    __syncthreads();
    // (3) break flags of slots [1, T + w]: slot j breaks a run unless it and
    // slot j-1 are valid with the same minimizer slot, or a chunk edge lies
    // at slot j (slot jE). Byte j/8, bit j%8.
    const uint64_t jE = S.edge - B;
    for (int j0 = threadIdx.x * SK_Q; j0 < np; j0 += SK_WG * SK_Q) {
      uint32_t bits = 0;
#pragma unroll
      for (int q = 0; q < SK_Q; q++) {
        const int j = j0 + q;
        bool br = true;
        if (j >= 1 && j < np) {
          const uint32_t a = S.pidx[PJ(j)], b = S.pidx[PJ(j - 1)];
          br = (a & SK_INVALID) || a != b || (uint64_t)j == jE;  // unequal when either is invalid and the other not
        }
        bits |= (uint32_t)br << q;
      }
      brk8[j0 >> 3] = (uint8_t)bits;
    }
    __syncthreads();
    // (4) super k-mer starts at positions t0 + [0, T) (slot j = 1 + li): a
    // valid slot with its break flag set. Its length runs to the next break.
    // The start slots are listed in scan order (in S.ord, free after step
    // 2) and emitted from the list by every lane: starts are about one slot
    // in w/2, so emitting them slot by slot would leave most lanes idle.
    uint32_t starts = 0, valid_cnt = 0, smask = 0;
    const int li0 = threadIdx.x * SK_Q;
#pragma unroll
    for (int q = 0; q < SK_Q; q++) {
      const int j = li0 + q + 1;
      if (li0 + q >= SK_TP || t0 + (uint64_t)(li0 + q) >= n_pos || (S.pidx[PJ(j)] & SK_INVALID)) continue;
      valid_cnt++;
      if ((brk8[j >> 3] >> (j & 7)) & 1) {
        starts++;
        smask |= 1u << q;
      }
    }
    // one scan of both counts: valid k-mers in the high half, starts low
    uint64_t both;
    uint32_t o = (uint32_t)sk_scan_excl(((uint64_t)valid_cnt << 32) | starts, S.sh, &both);
    const uint64_t wt = both >> 32, tot = both & 0xFFFFFFFFu;
    uint16_t* list = (uint16_t*)S.ord;  // <= SK_T entries
    while (smask) {
      const int q = __builtin_ctz(smask);
      smask &= smask - 1;
      list[o++] = (uint16_t)(li0 + q + 1);
    }
    __syncthreads();
    uint64_t* dst = entries + tile * SK_T;
    for (uint32_t i = threadIdx.x; i < (uint32_t)tot; i += SK_WG) {
      const int j = list[i];
      // next break after j: bits j+1 .. of the break flags (within w <= 49)
      const int wi = (j + 1) >> 6, sh = (j + 1) & 63;
      const uint64_t lo = S.brk[wi], hi = S.brk[wi + 1];
      const uint64_t after = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
      const uint32_t n = (uint32_t)__builtin_ctzll(after) + 1;
      // the canonical m-mer at the minimizer slot, from the stream (as in step 1)
      const uint64_t x = stream_bits64(seq, 2 * (B + (uint64_t)S.pidx[PJ(j)]));
      const uint32_t f = (uint32_t)(x >> (64 - 2 * m)) & M, r = (uint32_t)rc64(x, 32) & M;
      const uint64_t h = xxh3_u64le((uint64_t)(f < r ? f : r));
      const uint32_t bk = (uint32_t)(h >> 48);  // bucket_hash
      dst[i] = sk_entry(t0 + (uint64_t)(j - 1), n, bk);
      // the top 32 hash bits: the bucket (for the histogram and the route)
      // and the 16 bits below it, which pick the MPHF segment in P3
      buckets[tile * SK_T + i] = (uint32_t)(h >> 32);
      if ((uint64_t)j < jE && jE - (uint64_t)j < (uint64_t)K) {  // every k-mer spans the edge: not in P1's counts
        const uint32_t x = atomicAdd(excl_cnt, 1u);
        if (x < excl_cap) excl[x] = (uint16_t)bk;
      }
    }
    if (threadIdx.x == 0) {
      tile_cnt[tile] = (uint32_t)tot;
      tile_w[tile] = (uint32_t)wt;
    }
    __syncthreads();  // LDS reuse by the next tile
  }
}

// P1 histogram from the entries' buckets (the top 16 of the 32 hash bits the
// scan writes beside each entry), without device atomics: work-group (slice,
// range r) counts the buckets in [r*16384, (r+1)*16384) of a slice of tiles
// in LDS (each wave takes two tiles at a time) and writes its partial counts
// to part[(slice*4 + r)*16384 ...]; the r = 0 work-groups write the slice's
// W and S to sums[2*slice ...]. k_p1_reduce adds the partials up.
constexpr int HB_RANGE = 16384;
constexpr int HB_SLICES = 512;
constexpr int HB_WG = 1024;  // 16 waves share one 64 KiB LDS range
__global__ void __launch_bounds__(HB_WG) k_p1_hist(const uint32_t* __restrict__ buckets, const uint32_t* __restrict__ tile_cnt,
                                                   const uint32_t* __restrict__ tile_w, uint64_t n_tiles, uint64_t tiles_per_slice,
                                                   uint32_t* __restrict__ part, uint64_t* __restrict__ sums) {
  constexpr int NW = HB_WG / 64;
  __shared__ uint32_t h[HB_RANGE];
  __shared__ uint64_t red[2 * NW];
  const uint32_t r = blockIdx.y;
  for (int i = threadIdx.x; i < HB_RANGE; i += HB_WG) h[i] = 0;
  __syncthreads();
  const uint64_t ta = (uint64_t)blockIdx.x * tiles_per_slice;
  uint64_t tb = ta + tiles_per_slice;
  if (tb > n_tiles) tb = n_tiles;
  const uint32_t wv = threadIdx.x >> 6, lane = __lane_id();
  uint64_t s = 0, wsum = 0;
  // two tiles per step and four entries of each per lane, loaded before the
  // LDS updates, so that a lane has eight loads in flight
  for (uint64_t t = ta + wv; t < tb; t += 2 * NW) {
    const uint64_t t2 = t + NW;
    const uint32_t c1 = tile_cnt[t], c2 = t2 < tb ? tile_cnt[t2] : 0u;
    const uint32_t cm = c1 > c2 ? c1 : c2;
    for (uint32_t i0 = 0; i0 < cm; i0 += 4 * 64) {
      uint32_t bk[8];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t i = i0 + (uint32_t)q * 64 + lane;
        bk[q] = i < c1 ? buckets[t * SK_T + i] >> 16 : ~0u;
        bk[4 + q] = i < c2 ? buckets[t2 * SK_T + i] >> 16 : ~0u;
      }
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (bk[q] != ~0u && bk[q] / HB_RANGE == r) atomicAdd(&h[bk[q] % HB_RANGE], 1u);
    }
    s += c1 + c2;
    wsum += tile_w[t] + (t2 < tb ? tile_w[t2] : 0u);
  }
  if (lane == 0) {
    red[wv] = s;
    red[NW + wv] = wsum;
  }
  __syncthreads();
  uint32_t* dst = part + ((uint64_t)blockIdx.x * (65536 / HB_RANGE) + r) * HB_RANGE;
  for (int i = threadIdx.x; i < HB_RANGE; i += HB_WG) dst[i] = h[i];
  if (r == 0 && threadIdx.x == 0) {
    uint64_t ws = 0, ss = 0;
    for (int q = 0; q < NW; q++) {
      ss += red[q];
      ws += red[NW + q];
    }
    sums[2 * blockIdx.x] = ws;
    sums[2 * blockIdx.x + 1] = ss;
  }
}
// hist[b] = sum over slices of the partial counts; totals = {W, S}. Four
// threads per bucket take every fourth slice; LDS adds their sums.
__global__ void __launch_bounds__(256) k_p1_reduce(const uint32_t* __restrict__ part, const uint64_t* __restrict__ sums,
                                                   uint32_t n_slices, unsigned long long* __restrict__ hist,
                                                   unsigned long long* __restrict__ hist_all,
                                                   unsigned long long* __restrict__ totals) {
  __shared__ uint64_t acc4[4][64];
  const uint32_t g = threadIdx.x >> 6, bl = __lane_id();
  const uint32_t b = blockIdx.x * 64 + bl;  // grid covers 65536 buckets, 64 per work-group
  uint64_t acc = 0;
  uint32_t sl = g;
  for (; sl + 60 < n_slices; sl += 64) {  // 16 loads in flight per thread
    uint32_t v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = part[(uint64_t)(sl + 4 * q) * 65536 + b];
#pragma unroll
    for (int q = 0; q < 16; q++) acc += v[q];
  }
  for (; sl < n_slices; sl += 4) acc += part[(uint64_t)sl * 65536 + b];
  acc4[g][bl] = acc;
  __syncthreads();
  if (g == 0) hist[b] = hist_all[b] = acc4[0][bl] + acc4[1][bl] + acc4[2][bl] + acc4[3][bl];
  if (blockIdx.x == 0) {  // totals {W, S}: every thread of the block adds a stripe of the slices
    __syncthreads();
    uint64_t t = 0;
    for (uint32_t q = threadIdx.x >> 1; q < n_slices; q += 128) t += sums[2 * q + (threadIdx.x & 1)];
    uint64_t* red = &acc4[0][0];  // 256 words
    red[threadIdx.x] = t;
    __syncthreads();
    for (int d = 128; d >= 2; d >>= 1) {
      if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
      __syncthreads();
    }
    if (threadIdx.x < 2) totals[threadIdx.x] = red[threadIdx.x];
  }
}

// P1's counts without the super k-mers that only span a chunk edge (the
// histogram the partition uses; hist_all keeps every routed super k-mer and
// sizes the bins). flag |= 1 when the list overflowed (cannot happen: at
// most K-1 starts per edge).
__global__ void __launch_bounds__(256) k_p1_excl(const uint16_t* __restrict__ excl, const uint32_t* __restrict__ excl_cnt,
                                                 uint32_t excl_cap, unsigned long long* __restrict__ hist,
                                                 unsigned int* __restrict__ flag) {
  const uint32_t n = *excl_cnt;
  if (n > excl_cap) {
    if (threadIdx.x == 0) atomicOr(flag, 1u);
    return;
  }
  for (uint32_t i = threadIdx.x; i < n; i += 256) atomicAdd(&hist[excl[i]], ~0ULL);
}

// 64-bit digest of the packed stream (a save file records the records it
// belongs to): the sum mod 2^64 of mix64(word ^ index * golden) over the
// bases words, then the placeholder words (indices continuing).
__global__ void __launch_bounds__(256) k_stream_digest(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                       uint64_t n_words64, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4];
  const uint64_t n = 3 * n_words64;
  unsigned long long acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const uint64_t w = i < 2 * n_words64 ? seq[i] : nmask[i - 2 * n_words64];
    acc += mix64(w ^ (i * 0x9E3779B97F4A7C15ULL));
  }
  for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, 64);
  if (__lane_id() == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}
void launch_stream_digest(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_words64, unsigned long long* out) {
  const uint64_t n = 3 * n_words64;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g == 0) return;
  CF1G_LAUNCH("stream_digest", k_stream_digest, dim3((unsigned)g), dim3(256), 0, st, seq, nmask, n_words64, out);
}

// ------------------------------------------------------------------ route (P2)
// One work-group per chunk of `chunk_tiles` tiles. It counts its records per
// bin in LDS, reserves one run per bin from the bin's cursor (bins are sized
// exactly by the partition, from the true histogram), then writes the
// records. Both passes take RT_U tiles per step, one entry of each per
// thread, and issue each dependent load (entry, bucket -> bin, stream words)
// for all of them before the next, so that a thread has RT_U chains in
// flight. With 2-word records RT_U = 16 (157 VGPRs, 3 waves per SIMD): on
// the 300 Mbp genome at 256 bins the route takes 1.15 ms, against 1.28 ms
// with 8 (87 VGPRs, 5 waves) and 1.52 ms with 4 (55 VGPRs, 8 waves): the
// chains in flight per thread matter more than waves per SIMD. 4-word records
// keep 8 (16 would take 223 VGPRs). The order of records inside
// a bin follows the schedule. Each super k-mer [i, i+n) becomes one fixed-size record in its bin
// (the bin record of superkmer.rs:181-202 in fixed width): its L = n-1+K
// bases plus one flanking base on each side, MSB-first in RW words; the last
// word's low byte holds L and bits 8/9 say whether the left/right flank is a
// real base (not a placeholder or record end). The flanks carry the
// neighbours P4's observations need (classify.rs:76-77).
template <int RW>
struct RouteTiles {
  static constexpr int N = RW == 2 ? 16 : 8;
};
template <int RW>
__global__ void __launch_bounds__(256) k_p2_route(const uint64_t* __restrict__ seq, const uint64_t* __restrict__ nmask,
                                                  const uint64_t* __restrict__ entries, const uint32_t* __restrict__ buckets,
                                                  const uint32_t* __restrict__ tile_cnt, uint64_t n_tiles, uint64_t chunk_tiles, int K,
                                                  const uint16_t* __restrict__ b2b,
                                                  int num_bins, const uint64_t* __restrict__ bin_start,
                                                  unsigned long long* __restrict__ bin_cursor, uint64_t* __restrict__ rec,
                                                  uint16_t* __restrict__ rec_sub, uint32_t* __restrict__ ent_dest,
                                                  unsigned int* __restrict__ overflow,
                                                  unsigned long long* __restrict__ routed) {
  constexpr int RT_U = RouteTiles<RW>::N;
  // per-bin bases and cursors of this chunk, num_bins of each (dynamic LDS)
  extern __shared__ __align__(8) uint8_t route_lds[];
  uint64_t* base = (uint64_t*)route_lds;
  uint32_t* cnt = (uint32_t*)(base + num_bins);
  const uint64_t chunk = blockIdx.x;
  for (int b = threadIdx.x; b < num_bins; b += 256) cnt[b] = 0;
  __syncthreads();
  const uint64_t ta = chunk * chunk_tiles;
  uint64_t tb = ta + chunk_tiles;
  if (tb > n_tiles) tb = n_tiles;
  // the chunk's records per bin, then one range per bin reserved from its
  // cursor (the bin sizes are exact from the histogram)
  uint64_t mine = 0;
  for (uint64_t t0 = ta; t0 < tb; t0 += RT_U) {
    uint32_t c[RT_U], cmax = 0;
#pragma unroll
    for (int u = 0; u < RT_U; u++) {
      c[u] = t0 + u < tb ? tile_cnt[t0 + u] : 0u;
      cmax = c[u] > cmax ? c[u] : cmax;
      mine += threadIdx.x == 0 ? c[u] : 0u;
    }
    for (uint32_t i = threadIdx.x; i < cmax; i += 256) {
      uint32_t bk[RT_U];
#pragma unroll
      for (int u = 0; u < RT_U; u++) bk[u] = i < c[u] ? b2b[buckets[(t0 + u) * SK_T + i] >> 16] : ~0u;
#pragma unroll
      for (int u = 0; u < RT_U; u++)
        if (bk[u] != ~0u) atomicAdd(&cnt[bk[u]], 1u);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < num_bins; b += 256) {
    const uint32_t n = cnt[b];
    if (n) {
      base[b] = bin_start[b] + atomicAdd(&bin_cursor[b], (unsigned long long)n);
      if (base[b] + n > bin_start[b + 1]) atomicOr(overflow, 1u);
    }
    cnt[b] = 0;
  }
  if (threadIdx.x == 0 && mine) atomicAdd(routed, (unsigned long long)mine);
  __syncthreads();
  constexpr uint64_t NONE = ~0ULL;  // no entry (an entry's position is < 2^40)
  for (uint64_t t0 = ta; t0 < tb; t0 += RT_U) {
    uint32_t c[RT_U], cmax = 0;
#pragma unroll
    for (int u = 0; u < RT_U; u++) {
      c[u] = t0 + u < tb ? tile_cnt[t0 + u] : 0u;
      cmax = c[u] > cmax ? c[u] : cmax;
    }
    for (uint32_t i = threadIdx.x; i < cmax; i += 256) {
      uint64_t en[RT_U];
#pragma unroll
      for (int u = 0; u < RT_U; u++) en[u] = i < c[u] ? entries[(t0 + u) * SK_T + i] : NONE;
      uint32_t bin[RT_U];
#pragma unroll
      for (int u = 0; u < RT_U; u++) bin[u] = en[u] != NONE ? b2b[sk_bucket(en[u])] : 0u;
      uint64_t out[RT_U];
#pragma unroll
      for (int u = 0; u < RT_U; u++) {
        if (en[u] == NONE) continue;
        out[u] = base[bin[u]] + atomicAdd(&cnt[bin[u]], 1u);
        if (out[u] >= bin_start[bin[u] + 1] || out[u] < bin_start[bin[u]]) {  // never write outside the bin
          atomicOr(overflow, 1u);
          en[u] = NONE;
        }
      }
      uint64_t wv[RT_U][RW];
#pragma unroll
      for (int u = 0; u < RT_U; u++) {
        if (en[u] == NONE) continue;
        const uint64_t pos = sk_pos(en[u]);
        const uint32_t L = sk_len(en[u]) + (uint32_t)K - 1;
        // bases [pos-1, pos+L]: the super k-mer with one flanking base each side
        const uint32_t LS = L + 2;
#pragma unroll
        for (int q = 0; q < RW; q++) {
          const int lo_bit = 64 * q;  // record bit offset of this word
          uint64_t x = stream_bits64(seq, 2 * (pos - 1) + (uint64_t)lo_bit);
          const int keep = (int)(2 * LS) - lo_bit;  // record bits living in this word
          if (keep <= 0) x = 0;
          else if (keep < 64) x &= ~0ULL << (64 - keep);
          wv[u][q] = x;
        }
        const uint64_t lf = is_n(nmask, pos - 1) ? 0 : 1, rf = is_n(nmask, pos + L) ? 0 : 1;
        wv[u][RW - 1] |= (uint64_t)L | (lf << 8) | (rf << 9);
      }
      uint32_t hb[RT_U];
#pragma unroll
      for (int u = 0; u < RT_U; u++) hb[u] = en[u] != NONE ? buckets[(t0 + u) * SK_T + i] : 0u;
#pragma unroll
      for (int u = 0; u < RT_U; u++) {
        if (en[u] == NONE) continue;
        uint64_t* dst = rec + out[u] * RW;
#pragma unroll
        for (int q = 0; q < RW; q++) dst[q] = wv[u][q];
        rec_sub[out[u]] = (uint16_t)hb[u];  // the segment bits below the bucket
        ent_dest[(t0 + u) * SK_T + i] = (uint32_t)out[u];  // where the entry's record went (P5 finds it again)
      }
    }
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
// minimizer.rs:156-172).
constexpr int PT_WG = 1024;
static_assert(MAX_BINS % PT_WG == 0, "bins split evenly over the threads");
// floor(p / d) for d > 0 from a double estimate p * (1/d), corrected exactly
// (the estimate is within one of the quotient for p, d < 2^52).
__device__ __forceinline__ uint64_t div_floor(uint64_t p, uint64_t d, double inv_d) {
  uint64_t q = (uint64_t)((double)p * inv_d);
  while (q > 0 && q * d > p) q--;
  while ((q + 1) * d <= p) q++;
  return q;
}

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
// Four launches over 64 blocks of 1024 buckets (one per thread, coalesced):
// block sums of hist; g(b) = f(b) - b per bucket with each block's minimum;
// bucket_to_bin[b] = nb(b-1) from the running minimum (earlier blocks' minima
// and the block's own exclusive prefix minimum), with each bucket's true
// count added to its bin's record count; then one work-group scans the
// bins into record starts and chunk offsets.
constexpr int PT_NBLK = 65536 / PT_WG;
// scratch words: block sums, block minima, g per bucket, the bin count
constexpr uint64_t PT_SCRATCH_WORDS = 2 * PT_NBLK + 65536 + 1;
__global__ void __launch_bounds__(PT_WG) k_part_sums(const unsigned long long* __restrict__ hist, uint64_t* __restrict__ scratch,
                                                     uint64_t* __restrict__ bin_records) {
  __shared__ uint64_t shu[PT_WG / 64 + 1];
  uint64_t tot;
  blk_scan_excl_u64(hist[blockIdx.x * PT_WG + threadIdx.x], shu, &tot);
  if (threadIdx.x == 0) scratch[blockIdx.x] = tot;
  if (blockIdx.x == 0)
    for (int b = threadIdx.x; b < MAX_BINS; b += PT_WG) bin_records[b] = 0;
}
// total, this block's offset and t = total / num_bins from the block sums
__device__ __forceinline__ void part_totals(const uint64_t* __restrict__ scratch, uint32_t num_bins_req, uint64_t& pre, uint64_t& t) {
  uint64_t total = 0;
  pre = 0;
  for (int q = 0; q < PT_NBLK; q++) {
    const uint64_t v = scratch[q];
    total += v;
    if (q < (int)blockIdx.x) pre += v;
  }
  t = total / (uint64_t)num_bins_req;
}
__global__ void __launch_bounds__(PT_WG) k_part_g(const unsigned long long* __restrict__ hist, uint32_t num_bins_req,
                                                  uint64_t* __restrict__ scratch) {
  __shared__ uint64_t shu[PT_WG / 64 + 1];
  __shared__ int64_t shm[PT_WG / 64 + 1];
  uint64_t pre, t;
  part_totals(scratch, num_bins_req, pre, t);
  const int b = blockIdx.x * PT_WG + threadIdx.x;
  const uint64_t h = hist[b];
  const uint64_t P = pre + blk_scan_excl_u64(h, shu, nullptr) + h;  // inclusive prefix
  const int64_t g = t > 0 ? (int64_t)div_floor(P, t, 1.0 / (double)t) - (int64_t)b : INT64_MAX;
  int64_t* gs = (int64_t*)(scratch + 2 * PT_NBLK);
  gs[b] = g;
  const int64_t ex = blk_scan_excl_min(g, shm, INT64_MAX);
  if (threadIdx.x == PT_WG - 1) ((int64_t*)scratch)[PT_NBLK + blockIdx.x] = g < ex ? g : ex;
}
__global__ void __launch_bounds__(PT_WG) k_part_assign(const unsigned long long* __restrict__ hist_true, uint32_t num_bins_req,
                                                       uint64_t* __restrict__ scratch, uint16_t* __restrict__ b2b,
                                                       uint64_t* __restrict__ bin_records) {
  __shared__ int64_t shm[PT_WG / 64 + 1];
  uint64_t pre, t;
  part_totals(scratch, num_bins_req, pre, t);
  const int64_t cap = (int64_t)num_bins_req - 1;
  int64_t carry = INT64_MAX;  // min of g over the earlier blocks
  for (int q = 0; q < (int)blockIdx.x; q++) {
    const int64_t m = ((const int64_t*)scratch)[PT_NBLK + q];
    carry = m < carry ? m : carry;
  }
  const int b = blockIdx.x * PT_WG + threadIdx.x;
  const int64_t g = ((const int64_t*)(scratch + 2 * PT_NBLK))[b];
  const int64_t ex = blk_scan_excl_min(g, shm, INT64_MAX);
  const int64_t run = ex < carry ? ex : carry;  // min of g over j < b
  // nb(b - 1): boundaries placed up to the previous bucket
  int64_t prev_nb;
  if (b == 0) {
    prev_nb = 0;
  } else if (t == 0) {
    prev_nb = (int64_t)b < cap ? (int64_t)b : cap;
  } else {
    const int64_t v = (int64_t)(b - 1) + (run < 1 ? run : 1);
    prev_nb = v < cap ? v : cap;
  }
  b2b[b] = (uint16_t)prev_nb;
  const uint64_t ht = hist_true[b];
  if (ht) atomicAdd((unsigned long long*)&bin_records[prev_nb], (unsigned long long)ht);
  if (b == 65535) {  // nb(65535) + 1 bins
    int64_t nb;
    if (t == 0) {
      nb = 65536 < cap ? 65536 : cap;
    } else {
      const int64_t r2 = g < run ? g : run;
      const int64_t v = 65535 + (r2 < 1 ? r2 : 1);
      nb = v < cap ? v : cap;
    }
    scratch[2 * PT_NBLK + 65536] = (uint64_t)(nb + 1);
  }
}
__global__ void __launch_bounds__(PT_WG) k_part_bins(const uint64_t* __restrict__ scratch, uint32_t chunk_records,
                                                     uint64_t* __restrict__ bin_records, uint64_t* __restrict__ bin_start,
                                                     unsigned long long* __restrict__ bin_cursor, uint64_t* __restrict__ chunk_off,
                                                     uint64_t* __restrict__ info) {
  __shared__ uint64_t shu[PT_WG / 64 + 1];
  const int tid = threadIdx.x;
  // bin starts and chunk offsets: block scans over MAX_BINS / PT_WG bins per thread
  constexpr int BPT = MAX_BINS / PT_WG;
  const uint32_t nb = (uint32_t)scratch[2 * PT_NBLK + 65536];
  const double inv_c = 1.0 / (double)chunk_records;
  uint64_t rl = 0, cl = 0;
  uint64_t rv[BPT];
  for (int q = 0; q < BPT; q++) {
    const uint32_t b = (uint32_t)tid * BPT + q;
    rv[q] = b < nb ? bin_records[b] : 0;
    rl += rv[q];
    if (b < nb) cl += div_floor(rv[q] + chunk_records - 1, chunk_records, inv_c);
  }
  uint64_t rtot, ctot;
  uint64_t run = blk_scan_excl_u64(rl, shu, &rtot);
  uint64_t crun = blk_scan_excl_u64(cl, shu, &ctot);
  for (int q = 0; q < BPT; q++) {
    const uint32_t b = (uint32_t)tid * BPT + q;
    if (b < nb) {
      const uint64_t r = rv[q];
      bin_start[b] = run;
      chunk_off[b] = crun;
      bin_cursor[b] = 0;
      run += r;
      crun += div_floor(r + chunk_records - 1, chunk_records, inv_c);
    }
  }
  if (tid == 0) {
    bin_start[nb] = rtot;
    chunk_off[nb] = ctot;
    info[0] = nb;
    info[1] = rtot;
    info[2] = ctot;
  }
}

// ------------------------------------------------------------------ launchers
void launch_pack(hipStream_t st, const uint8_t* ascii, const uint64_t* rec_off, const uint64_t* gstart, uint64_t n_rec,
                 uint64_t n_words64, uint64_t* seq, uint64_t* nmask) {
  if (n_words64 == 0) return;
  const uint64_t waves = (n_words64 + 63) / 64;  // 64 words per wave, 4 waves per work-group
  dim3 grid((unsigned)((waves + 3) / 4));
  CF1G_LAUNCH("pack", k_pack, grid, dim3(256), 0, st, ascii, rec_off, gstart, n_rec, n_words64, seq, nmask);
}
uint64_t p1_tiles(uint64_t n_pos) { return n_pos > 1 ? (n_pos - 1 + SK_TP - 1) / SK_TP : 0; }
// Slots per tile in the entry, bucket and destination arrays (a tile emits at
// most one entry per position).
uint64_t p1_tile_slots() { return SK_T; }
uint64_t p1_slices(uint64_t n_pos) {
  const uint64_t nt = p1_tiles(n_pos);
  return nt < (uint64_t)HB_SLICES ? nt : (uint64_t)HB_SLICES;
}
// P1 scan of tiles [t_begin, t_end) (a tile reads the stream up to
// p1_tile_reach(t_end - 1) positions; the streaming load scans each range as
// its words arrive).
void launch_p1_scan(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, uint64_t t_begin, uint64_t t_end, int K,
                    int m, uint64_t* entries, uint32_t* buckets, uint32_t* tile_cnt, uint32_t* tile_w, const uint64_t* edges,
                    uint32_t n_edges, uint32_t* excl_cnt, uint16_t* excl, uint32_t excl_cap) {
  if (t_end <= t_begin) return;
  const uint64_t nt = t_end - t_begin;
  const uint64_t grid = nt < 4096 ? nt : 4096;
  CF1G_LAUNCH("p1_scan", k_p1_scan, dim3((unsigned)grid), dim3(SK_WG), 0, st, seq, nmask, n_pos, t_begin, t_end, K, m, entries, buckets,
              tile_cnt, tile_w, edges, n_edges, excl_cnt, excl, excl_cap);
}
// First stream position past what the scan of tile t reads (orders over
// [t0-1, t0+T+2w) from 64-position windows, validity from two placeholder
// words per step, each start's minimizer m-mer).
uint64_t p1_tile_reach(uint64_t t) { return t * SK_TP + SK_TP + 2 * SK_MAXW + 512; }
// The histogram from every scanned tile: partial counts per slice, their sum,
// then the chunk-edge exclusions.
void launch_p1_finish(hipStream_t st, uint64_t n_pos, uint32_t* buckets, uint32_t* tile_cnt, uint32_t* tile_w, uint32_t* part,
                      uint64_t* sums, unsigned long long* hist, unsigned long long* hist_all, unsigned long long* totals,
                      uint32_t n_edges, uint32_t* excl_cnt, uint16_t* excl, uint32_t excl_cap, unsigned int* flag) {
  const uint64_t nt = p1_tiles(n_pos);
  if (nt == 0) return;
  const uint64_t slices = p1_slices(n_pos);
  const uint64_t per = (nt + slices - 1) / slices;
  const uint64_t used = (nt + per - 1) / per;
  CF1G_LAUNCH("p1_hist", k_p1_hist, dim3((unsigned)used, 65536 / HB_RANGE), dim3(HB_WG), 0, st, buckets, tile_cnt, tile_w, nt, per,
              part, sums);
  CF1G_LAUNCH("p1_reduce", k_p1_reduce, dim3(65536 / 64), dim3(256), 0, st, part, sums, (uint32_t)used, hist, hist_all, totals);
  if (n_edges) CF1G_LAUNCH("p1_excl", k_p1_excl, dim3(1), dim3(256), 0, st, excl, excl_cnt, excl_cap, hist, flag);
}
void launch_p1(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, uint64_t n_pos, int K, int m, uint64_t* entries,
               uint32_t* buckets, uint32_t* tile_cnt, uint32_t* tile_w, uint32_t* part, uint64_t* sums, unsigned long long* hist,
               unsigned long long* hist_all, unsigned long long* totals, const uint64_t* edges, uint32_t n_edges,
               uint32_t* excl_cnt, uint16_t* excl, uint32_t excl_cap, unsigned int* flag) {
  const uint64_t nt = p1_tiles(n_pos);
  if (nt == 0) return;
  launch_p1_scan(st, seq, nmask, n_pos, 0, nt, K, m, entries, buckets, tile_cnt, tile_w, edges, n_edges, excl_cnt, excl, excl_cap);
  launch_p1_finish(st, n_pos, buckets, tile_cnt, tile_w, part, sums, hist, hist_all, totals, n_edges, excl_cnt, excl, excl_cap, flag);
}
// Route chunking: chunk_tiles tiles per chunk, one route work-group each.
void p2_chunking(uint64_t n_pos, uint32_t num_bins, uint64_t* chunk_tiles, uint64_t* n_chunks) {
  (void)num_bins;
  const uint64_t nt = p1_tiles(n_pos);
  const uint64_t ct = 16;  // route on the 300 Mbp genome: 0.90 ms, 0.96 with 32 tiles, 1.00 with 64
  *chunk_tiles = ct;
  *n_chunks = (nt + ct - 1) / ct;
}
void launch_p2_route(hipStream_t st, const uint64_t* seq, const uint64_t* nmask, const uint64_t* entries, const uint32_t* buckets,
                     const uint32_t* tile_cnt, uint64_t n_pos, uint64_t chunk_tiles, uint64_t n_chunks, int K, const uint16_t* b2b,
                     int num_bins, const uint64_t* bin_start, unsigned long long* bin_cursor, uint64_t* rec, uint16_t* rec_sub,
                     uint32_t* ent_dest, int rw,
                     unsigned int* overflow, unsigned long long* routed) {
  if (n_chunks == 0) return;
  const uint64_t nt = p1_tiles(n_pos);
  const size_t lds = (size_t)num_bins * 12;
  if (rw == 2)
    CF1G_LAUNCH("p2_route", k_p2_route<2>, dim3((unsigned)n_chunks), dim3(256), lds, st, seq, nmask, entries, buckets, tile_cnt, nt,
                chunk_tiles, K, b2b, num_bins, bin_start, bin_cursor, rec, rec_sub, ent_dest, overflow, routed);
  else
    CF1G_LAUNCH("p2_route", k_p2_route<4>, dim3((unsigned)n_chunks), dim3(256), lds, st, seq, nmask, entries, buckets, tile_cnt, nt,
                chunk_tiles, K, b2b, num_bins, bin_start, bin_cursor, rec, rec_sub, ent_dest, overflow, routed);
}
uint64_t partition_scratch_words() { return PT_SCRATCH_WORDS; }
void launch_partition(hipStream_t st, const unsigned long long* hist, const unsigned long long* hist_true, uint32_t num_bins,
                      uint32_t chunk_records, uint16_t* b2b, uint64_t* bin_records, uint64_t* bin_start,
                      unsigned long long* bin_cursor, uint64_t* chunk_off, uint64_t* info, uint64_t* scratch) {
  CF1G_LAUNCH("p2_partition", k_part_sums, dim3(PT_NBLK), dim3(PT_WG), 0, st, hist, scratch, bin_records);
  CF1G_LAUNCH("p2_partition", k_part_g, dim3(PT_NBLK), dim3(PT_WG), 0, st, hist, num_bins, scratch);
  CF1G_LAUNCH("p2_partition", k_part_assign, dim3(PT_NBLK), dim3(PT_WG), 0, st, hist_true, num_bins, scratch, b2b, bin_records);
  CF1G_LAUNCH("p2_partition", k_part_bins, dim3(1), dim3(PT_WG), 0, st, scratch, chunk_records, bin_records, bin_start, bin_cursor,
              chunk_off, info);
}
}  // namespace cf1g
