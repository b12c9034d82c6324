// cf1ref — CPU oracle for the cf1-rs phase-1..5 path.
//
// TEST INFRASTRUCTURE ONLY. This restatement is the checker for the HIP
// product path and the timed CPU baseline. Only tests/, the smoke check and
// bench.py's cpu_baseline leg may load it; the product never links it.
//
// Every function names the reference file:line it restates (paths relative to
// the cf1-rs v0.3.1 source tree). Where the reference delegates to a
// third-party crate that is not present (simd-minimizers, ph), the oracle
// implements the project's own scheme (the same one the GPU implements) and
// says so; graph output is invariant to both choices.
//
// Pinning (what checks each part):
//   dna tables, Kmer ops     : the reference's own unit tests (known answers,
//                              tests/golden/reference_kat.json) + string-level
//                              brute force over random k-mers for every odd K
//   xxh3_64 (8-byte input)   : python xxhash 3.8.1 golden vectors
//   partition_minimizers     : closed-form properties (monotone map, bin count
//                              <= num_bins, target==0 edge) + hand cases
//   minimizer scheme         : brute force definition on small inputs
//   dedup                    : sorted/unique closed form vs python set
//   BBHash                   : bijection onto [0,D) + structural invariants
//   DFA / classify           : brute-force observation-set oracle (python)
//   P5 output                : brute-force unitig compaction + tiling spelling
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace cf1ref {

typedef unsigned __int128 u128;

// ---------------------------------------------------------------- dna.rs
// src/dna.rs:64-105 — ASCII -> 2-bit (A0 C1 G2 T3, lowercase accepted),
// placeholder test, complement char.
struct DnaTables {
  uint8_t mapped_base[256];
  bool is_placeholder[256];
  uint8_t complemented_char[256];
  DnaTables() {
    for (int i = 0; i < 256; i++) {
      mapped_base[i] = 4;
      is_placeholder[i] = true;
      complemented_char[i] = 'N';
    }
    const char* up = "ACGT";
    const char* lo = "acgt";
    const char* cm = "TGCA";
    for (int b = 0; b < 4; b++) {
      mapped_base[(uint8_t)up[b]] = (uint8_t)b;
      mapped_base[(uint8_t)lo[b]] = (uint8_t)b;
      is_placeholder[(uint8_t)up[b]] = false;
      is_placeholder[(uint8_t)lo[b]] = false;
      complemented_char[(uint8_t)up[b]] = (uint8_t)cm[b];
      complemented_char[(uint8_t)lo[b]] = (uint8_t)cm[b];
    }
  }
};
extern const DnaTables kDna;

inline uint8_t map_base(uint8_t c) { return kDna.mapped_base[c]; }            // dna.rs:30
inline bool is_placeholder(uint8_t c) { return kDna.is_placeholder[c]; }      // dna.rs:42
inline uint8_t complement_base(uint8_t b) { return (uint8_t)(3 - b); }        // dna.rs:13 (T G C A)
inline uint8_t complement_char(uint8_t c) { return kDna.complemented_char[c]; } // dna.rs:58
inline uint8_t to_upper(uint8_t c) { return c <= 'T' ? c : (uint8_t)(c - ('a' - 'A')); } // dna.rs:48-54
inline char base_char(uint8_t b) { return "ACGT"[b & 3]; }

// ---------------------------------------------------------------- kmer.rs
// REVERSE_COMPLEMENT_BYTE (kmer.rs:7-21): a byte holds 4 bases (2 bits each,
// first base in the high bits); the table entry reverses the 4 bases and
// complements each. Built here from that definition rather than as a literal.
struct RcByteTable {
  uint8_t t[256];
  RcByteTable() {
    for (int v = 0; v < 256; v++) {
      int out = 0;
      for (int j = 0; j < 4; j++) {
        int b = (v >> (2 * j)) & 3;        // base at bit pair j
        out |= (3 - b) << (2 * (3 - j));   // goes to mirrored pair, complemented
      }
      t[v] = (uint8_t)out;
    }
  }
};
extern const RcByteTable kRcByte;

// Kmer<K> (kmer.rs:33-388): u64 storage for K<=32, u128 for K<=63. Word 0 is
// the low 64 bits; the first base sits in the most significant bits. Ordering
// compares the high word first (kmer.rs:73-89) == numeric order of the u128.
template <int K>
struct Kmer {
  static_assert(K >= 1 && K <= 63, "K in [1,63]");
  typedef typename std::conditional<(K <= 32), uint64_t, u128>::type Storage;
  static constexpr int NUM_WORDS = (K + 31) / 32;
  static constexpr int MSN_SHIFT = 2 * ((K - 1) % 32);
  Storage bits = 0;

  uint64_t word(int idx) const {
    if (NUM_WORDS == 1) return (uint64_t)bits;
    return idx == 0 ? (uint64_t)bits : (uint64_t)((u128)bits >> 64);
  }
  void set_word(int idx, uint64_t v) {
    if (NUM_WORDS == 1) {
      bits = (Storage)v;
    } else {
      u128 b = (u128)bits;
      if (idx == 0)
        b = (b & ((u128)~0ULL << 64)) | (u128)v;
      else
        b = (b & (u128)~0ULL) | ((u128)v << 64);
      bits = (Storage)b;
    }
  }
  bool operator==(const Kmer& o) const { return bits == o.bits; }
  bool operator!=(const Kmer& o) const { return bits != o.bits; }
  // kmer.rs:73-89 — high word first.
  bool operator<(const Kmer& o) const {
    for (int idx = NUM_WORDS - 1; idx >= 0; idx--) {
      uint64_t a = word(idx), b = o.word(idx);
      if (a != b) return a < b;
    }
    return false;
  }
  bool operator<=(const Kmer& o) const { return !(o < *this); }

  // kmer.rs:110-131 — encode K ASCII bases; full 32-base words from the tail,
  // the partial (highest) word from the head.
  static Kmer from_ascii(const uint8_t* seq, size_t offset) {
    const uint8_t* label = seq + offset;
    Kmer km;
    int packed_word_count = K / 32;
    for (int d = 0; d < packed_word_count; d++) {
      int start = K - (d << 5) - 32;
      uint64_t w = 0;
      for (int t = 0; t < 32; t++) w = (w << 2) | map_base(label[start + t]);
      km.set_word(d, w);
    }
    int rem = K & 31;
    if (rem > 0) {
      uint64_t w = 0;
      for (int t = 0; t < rem; t++) w = (w << 2) | map_base(label[t]);
      km.set_word(NUM_WORDS - 1, w);
    }
    return km;
  }

  // kmer.rs:136-158 — decode the k-mer at base offset i of an MSB-first
  // 2-bit buffer. Restated by its documented meaning (base j at bits
  // 6-2*(j%4) of byte j/4). The reference's u128 accumulator underflows at
  // K=63 when (i%4) >= 2 (kmer.rs:148); the oracle is the semantic decode.
  static Kmer from_packed_2bit(const uint8_t* packed, size_t base_offset) {
    u128 v = 0;
    for (int t = 0; t < K; t++) {
      size_t j = base_offset + (size_t)t;
      v = (v << 2) | (u128)((packed[j / 4] >> (6 - 2 * (j % 4))) & 3);
    }
    Kmer km;
    km.bits = (Storage)v;
    return km;
  }

  // kmer.rs:161-200 — byte-level reverse complement over the little-endian
  // storage bytes, then a partial-byte fix-up for K % 4 != 0.
  Kmer reverse_complement() const {
    uint8_t data[16];
    uint8_t rc[16];
    std::memset(data, 0, sizeof data);
    std::memset(rc, 0, sizeof rc);
    std::memcpy(data, &bits, sizeof(Storage));
    int packed_byte_count = K / 4;
    for (int bi = 0; bi < packed_byte_count; bi++) rc[packed_byte_count - 1 - bi] = kRcByte.t[data[bi]];
    Kmer out;
    std::memcpy(&out.bits, rc, sizeof(Storage));
    int rem = K % 4;
    if (rem == 0) return out;
    // rc byte [packed_byte_count] = 0 (already), then left shift by rem bases.
    out.left_shift_by(rem);
    std::memcpy(rc, &out.bits, sizeof(Storage));
    uint8_t partial = data[packed_byte_count];
    for (int i = 0; i < rem; i++) {
      uint8_t b = (partial >> (2 * i)) & 3;
      rc[0] |= (uint8_t)(complement_base(b) << (2 * (rem - 1 - i)));
    }
    std::memcpy(&out.bits, rc, sizeof(Storage));
    return out;
  }
  // kmer.rs:203-219
  Kmer canonical() const {
    Kmer rc = reverse_complement();
    return (*this <= rc) ? *this : rc;
  }
  Kmer canonical_with_rc(const Kmer& rc) const { return (*this <= rc) ? *this : rc; }
  bool in_forward(const Kmer& hat) const { return *this == hat; }   // kmer.rs:222-224
  uint8_t front() const { return (uint8_t)((word(NUM_WORDS - 1) >> MSN_SHIFT) & 3); }  // kmer.rs:227-230
  uint8_t back() const { return (uint8_t)(word(0) & 3); }           // kmer.rs:233-236

  // kmer.rs:240-262 — chop the front base, append at the back; keep rc in step.
  void roll_to_next_kmer(uint8_t base, Kmer& rev_compl) {
    uint64_t clear_msn = ~(3ULL << MSN_SHIFT);
    set_word(NUM_WORDS - 1, word(NUM_WORDS - 1) & clear_msn);
    left_shift();
    set_word(0, word(0) | base);
    rev_compl.right_shift();
    rev_compl.set_word(NUM_WORDS - 1,
                       rev_compl.word(NUM_WORDS - 1) | ((uint64_t)complement_base(base) << MSN_SHIFT));
  }

  std::string string_label() const {  // kmer.rs:277-289
    std::string s(K, 'A');
    Kmer c = *this;
    for (int t = K - 1; t >= 0; t--) {
      s[t] = base_char((uint8_t)(c.word(0) & 3));
      c.right_shift();
    }
    return s;
  }

  void left_shift() {  // kmer.rs:292-306
    for (int idx = NUM_WORDS - 1; idx >= 1; idx--)
      set_word(idx, (word(idx) << 2) | (word(idx - 1) >> 62));
    set_word(0, word(0) << 2);
  }
  void left_shift_by(int b) {  // kmer.rs:309-327
    if (b == 0) return;
    int nb = 2 * b;
    uint64_t mask_msns = ((1ULL << nb) - 1) << (64 - nb);
    for (int idx = NUM_WORDS - 1; idx >= 1; idx--)
      set_word(idx, (word(idx) << nb) | ((word(idx - 1) & mask_msns) >> (64 - nb)));
    set_word(0, word(0) << nb);
  }
  void right_shift() {  // kmer.rs:330-343
    for (int idx = 0; idx < NUM_WORDS - 1; idx++)
      set_word(idx, (word(idx) >> 2) | ((word(idx + 1) & 3) << 62));
    set_word(NUM_WORDS - 1, word(NUM_WORDS - 1) >> 2);
  }
};

// ---------------------------------------------------------------- directed_kmer.rs
// directed_kmer.rs:14-64 — (kmer, rc, canonical, dir); FWD iff kmer == canonical.
template <int K>
struct DirectedKmer {
  Kmer<K> kmer, rev_compl, canonical;
  bool dir;  // true = FWD
  DirectedKmer() : dir(true) {}
  explicit DirectedKmer(const Kmer<K>& k) : kmer(k) {  // :26-36
    rev_compl = kmer.reverse_complement();
    canonical = kmer.canonical_with_rc(rev_compl);
    dir = kmer.in_forward(canonical);
  }
  void roll_to_next_kmer(uint8_t next_char) {  // :38-43
    kmer.roll_to_next_kmer(map_base(next_char), rev_compl);
    canonical = kmer.canonical_with_rc(rev_compl);
    dir = kmer.in_forward(canonical);
  }
};

// ---------------------------------------------------------------- state.rs
// state.rs:6-298 — 5-bit DFA state codes and decode tables.
enum StateClass : uint8_t { SISO = 0, MISO = 1, SIMO = 2, MIMO = 3 };
struct Vertex {
  StateClass state_class;
  uint8_t front, back;
  bool visited, outputted;
};
inline Vertex v_unvisited() { return {SISO, 0, 0, false, false}; }           // :27-35
inline Vertex v_siso(uint8_t f, uint8_t b) { return {SISO, f, b, true, false}; }  // :38-46
inline Vertex v_miso(uint8_t b) { return {MISO, 0, b, true, false}; }        // :49-57
inline Vertex v_simo(uint8_t f) { return {SIMO, f, 0, true, false}; }        // :60-68
inline Vertex v_mimo() { return {MIMO, 0, 0, true, false}; }                 // :71-77

// State::from_vertex — state.rs:92-113
inline uint8_t state_from_vertex(const Vertex& v) {
  if (!v.visited) return 0;
  if (!v.outputted) {
    switch (v.state_class) {
      case SISO: return (uint8_t)(0x10 | (v.front << 2) | v.back);
      case MISO: return (uint8_t)(0x04 | v.back);
      case SIMO: return (uint8_t)(0x08 | v.front);
      default: return 0x03;
    }
  }
  switch (v.state_class) {
    case SISO: return 0x0C;
    case MISO: return 0x0D;
    case SIMO: return 0x0E;
    default: return 0x0F;
  }
}

struct StateTables {
  Vertex decode[32];
  StateClass cls[32];
  uint8_t outputted[32];
  StateTables() {
    // DECODE_TABLE state.rs:136-247
    for (int i = 0; i < 32; i++) decode[i] = {SISO, 0, 0, false, false};
    decode[3] = {MIMO, 0, 0, true, false};
    for (int i = 0; i < 4; i++) decode[4 + i] = {MISO, 0, (uint8_t)i, true, false};
    for (int i = 0; i < 4; i++) decode[8 + i] = {SIMO, (uint8_t)i, 0, true, false};
    decode[12] = {SISO, 0, 0, true, true};
    decode[13] = {MISO, 0, 0, true, true};
    decode[14] = {SIMO, 0, 0, true, true};
    decode[15] = {MIMO, 0, 0, true, true};
    for (int c = 16; c < 32; c++) decode[c] = {SISO, (uint8_t)((c >> 2) & 3), (uint8_t)(c & 3), true, false};
    // STATE_CLASS_TABLE state.rs:249-270
    for (int i = 0; i < 32; i++) cls[i] = SISO;
    cls[3] = MIMO;
    for (int i = 4; i < 8; i++) cls[i] = MISO;
    for (int i = 8; i < 12; i++) cls[i] = SIMO;
    cls[12] = SISO; cls[13] = MISO; cls[14] = SIMO; cls[15] = MIMO;
    // OUTPUTTED_TABLE state.rs:272-298
    for (int i = 0; i < 32; i++) outputted[i] = 0;
    outputted[3] = 0x0F;
    for (int i = 4; i < 8; i++) outputted[i] = 0x0D;
    for (int i = 8; i < 12; i++) outputted[i] = 0x0E;
    outputted[12] = 0x0C; outputted[13] = 0x0D; outputted[14] = 0x0E; outputted[15] = 0x0F;
    for (int i = 16; i < 32; i++) outputted[i] = 0x0C;
  }
};
extern const StateTables kState;

inline bool st_is_visited(uint8_t s) { return s != 0; }                                        // :122-124
inline StateClass st_class(uint8_t s) { return kState.cls[s & 31]; }                            // :133-135
inline bool st_is_dead_end(uint8_t s) { return st_is_visited(s) && st_class(s) == MIMO; }       // :128-130
inline bool st_is_outputted(uint8_t s) { return s >= 12 && s <= 15; }                           // :138-140
inline uint8_t st_outputted(uint8_t s) { return kState.outputted[s & 31]; }                     // :143-145
inline Vertex st_decode(uint8_t s) { return kState.decode[s & 31]; }                            // :117-119

// ---------------------------------------------------------------- state_vector.rs
// state_vector.rs:7-71 — 12 x 5-bit states per u64 word, no straddling.
struct AtomicStateVector {
  static constexpr size_t STATES_PER_WORD = 12;
  static constexpr size_t BITS_PER_STATE = 5;
  static constexpr uint64_t STATE_MASK = 0x1F;
  std::vector<std::atomic<uint64_t>> data;
  size_t len = 0;
  // Test hook: number of leading CAS attempts to fail on purpose (each forced
  // failure behaves like a lost race; callers must retry and converge).
  std::atomic<int64_t> forced_cas_failures{0};

  explicit AtomicStateVector(size_t n = 0) { reset(n); }
  void reset(size_t n) {  // :17-24
    len = n;
    std::vector<std::atomic<uint64_t>> d((n + STATES_PER_WORD - 1) / STATES_PER_WORD);
    for (auto& w : d) w.store(0, std::memory_order_relaxed);
    data.swap(d);
  }
  size_t num_words() const { return data.size(); }
  uint8_t get(size_t i) const {  // :36-42
    size_t wi = i / STATES_PER_WORD;
    size_t sh = (i % STATES_PER_WORD) * BITS_PER_STATE;
    return (uint8_t)((data[wi].load(std::memory_order_acquire) >> sh) & STATE_MASK);
  }
  bool compare_and_swap(size_t i, uint8_t expected, uint8_t nv) {  // :47-70
    size_t wi = i / STATES_PER_WORD;
    size_t sh = (i % STATES_PER_WORD) * BITS_PER_STATE;
    uint64_t mask = STATE_MASK << sh;
    for (;;) {
      uint64_t cur = data[wi].load(std::memory_order_acquire);
      uint8_t cs = (uint8_t)((cur >> sh) & STATE_MASK);
      if (cs != expected) return false;
      if (forced_cas_failures.load(std::memory_order_relaxed) > 0) {
        if (forced_cas_failures.fetch_sub(1) > 0) return false;
      }
      uint64_t nw = (cur & ~mask) | ((uint64_t)nv << sh);
      if (data[wi].compare_exchange_weak(cur, nw, std::memory_order_acq_rel, std::memory_order_relaxed))
        return true;
    }
  }
};

// ---------------------------------------------------------------- params.rs
// params.rs:221-228 — m = min(k/2, 15), made odd by subtracting one (min 1).
inline int minimizer_len(int k) {
  int m = std::min(k / 2, 15);
  if (m % 2 == 0) m = std::max(m - 1, 1);
  return m;
}

// ---------------------------------------------------------------- xxh3
// xxhash-rust 0.8.15 xxh3_64 (seed 0) for an 8-byte input, restated from the
// published XXH3 specification (len 4..8 path: bitflip from the default
// secret bytes 8..23, then the rrmxmx finalizer). Pinned by python xxhash.
inline uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t xxh3_64_u64le(uint64_t v) {
  uint32_t in1 = (uint32_t)v, in2 = (uint32_t)(v >> 32);
  uint64_t bitflip = 0x1cad21f72c81017cULL ^ 0xdb979083e96dd4deULL;
  uint64_t in64 = (uint64_t)in2 + ((uint64_t)in1 << 32);
  uint64_t h = in64 ^ bitflip;
  h ^= rotl64(h, 49) ^ rotl64(h, 24);
  h *= 0x9FB21C651E98DF25ULL;
  h ^= (h >> 35) + 8;
  h *= 0x9FB21C651E98DF25ULL;
  h ^= h >> 28;
  return h;
}
// minimizer.rs:18-21 — top 16 bits of xxh3 of the value's LE bytes.
constexpr size_t NUM_HISTOGRAM_BUCKETS = 65536;  // minimizer.rs:8
inline uint32_t bucket_hash(uint64_t val) { return (uint32_t)(xxh3_64_u64le(val) >> 48); }

// ---------------------------------------------------------------- minimizer scheme
// The reference delegates canonical minimizers to simd-minimizers 2.3.1
// (minimizer.rs:103-104, superkmer.rs:118-121), which is absent here. The
// project's scheme (identical on CPU and GPU):
//   order of an m-mer = fmix32(canonical value + 0x9E3779B9), a bijection on
//   u32, so equal order <=> equal canonical m-mer (m <= 15 => value < 2^30);
//   minimizer of the k-mer at i = leftmost m-mer position in [i, i+w) with the
//   least order; its value v = that canonical m-mer;
//   a super k-mer is a maximal run of consecutive valid k-mers that share the
//   minimizer position; it routes to bin bucket_to_bin[bucket_hash(v)].
// v depends on the k-mer's bases only and is the same for its reverse
// complement, so per-bin dedup is exact (mphf.rs:27-28).
inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
inline uint32_t mmer_order(uint32_t canon) { return fmix32(canon + 0x9E3779B9u); }

// ---------------------------------------------------------------- MPHF
// The reference builds ph::fmph (ph 0.11.0, gamma = relative_level_size 200%,
// mphf.rs:155-160). ph is absent; the project's MPHF (the same structure on
// CPU and GPU; exact values parity-unpinned against ph; graph output is
// invariant to the MPHF once unitig ids are canonicalised) is a
// minimizer-partitioned BBHash:
//   segment of key c = (bin, s): with h = xxh3_64(v(c)) (v(c) the minimizer
//   value of c, the hash whose top 16 bits are its bucket), bin =
//   bucket_to_bin[h >> 48] and s = the sb bits of h just below the bucket,
//   (h >> (48 - sb)) & (2^sb - 1): every k-mer of a super k-mer shares its
//   segment, and a bin's minimizers spread evenly over its 2^sb segments.
//   Segments are
//   numbered (bin << sb) | s and hold consecutive value ranges key_off[seg]..;
//   the distinct keys are stored in value order (keys[v] has value v).
//   sb = fewest bits (<= 12) with W <= bins * 2^sb * target
//   (target 1536; W = valid k-mer occurrences).
//   per segment with D_s keys, level l has n_l = roundup64(max(2 t_l, 1)) bits,
//   t_0 = D_s, t_{l+1} = ceil(9 t_l / 20); a key lands on off_l + bbh_pos(key,
//   l, n_l) (below) and is placed at the first level where no other remaining
//   key of the segment lands on that bit;
//   levels are bit-contiguous in 64-byte blocks {u64 rank_before, u64 bits[7]}
//   (ranks local to the segment); value = key_off[segment] + local rank.
constexpr int BBH_MAX_LEVELS = 64;
constexpr uint64_t BBH_BLOCK_BITS = 448;
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline uint64_t bbh_seed(int i) { return mix64(0x243F6A8885A308D3ULL + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL); }
// Level hashes of a key come from one 64-bit state per key and level,
// g_l = h1 + l * d (mod 2^64), with h1 = (lo ^ S) * M1 for keys of <= 64
// bits, (lo ^ mix64(hi ^ S)) * M1 for wider ones (S =
// bbh_seed(0), M1 = 0xD6E8FEB86659FD93), and d = rotl(h1, 21) | 1. The
// level position among n bits (n < 2^32) is floor(x * n / 2^32) with x =
// fmix32(hi32(g_l) ^ lo32(g_l)). The fmix32 step is what keeps levels
// independent: reading the position straight from the top bits of g_l would
// keep two keys whose h1 and d are both close together on one bit at every
// level. (A rotation by 32 would make the fold symmetric in the two halves
// and is avoided for the same reason. A linear fold of the wide key,
// lo ^ hi * C, is avoided too: it gives one h1 to pairs of 63-mers that
// differ only near the top of both words, and such a pair can never be
// separated.)
constexpr uint64_t BBH_M1 = 0xD6E8FEB86659FD93ULL;
inline uint64_t bbh_h1(uint64_t lo, uint64_t hi, bool wide) {
  uint64_t s = bbh_seed(0);
  if (!wide) return (lo ^ s) * BBH_M1;
  return (lo ^ mix64(hi ^ s)) * BBH_M1;
}
inline uint64_t bbh_d(uint64_t h1) { return ((h1 << 21) | (h1 >> 43)) | 1ULL; }
inline uint64_t bbh_pos_g(uint64_t g, uint64_t n) {
  uint32_t x = fmix32((uint32_t)(g >> 32) ^ (uint32_t)g);
  return ((uint64_t)x * n) >> 32;
}
inline uint64_t bbh_pos(uint64_t lo, uint64_t hi, bool wide, int level, uint64_t n) {
  uint64_t h1 = bbh_h1(lo, hi, wide);
  return bbh_pos_g(h1 + (uint64_t)level * bbh_d(h1), n);
}
// Segment hash: multiplicative (Fibonacci) hash of the key.
inline uint64_t seg_hash(uint64_t lo, uint64_t hi, bool wide) {
  return (lo ^ (wide ? hi * 0xC2B2AE3D27D4EB4FULL : 0)) * 0x9E3779B97F4A7C15ULL;
}
inline uint64_t bbh_level_bits(uint64_t t) { return (std::max<uint64_t>(2 * t, 1) + 63) / 64 * 64; }
inline uint64_t bbh_next_t(uint64_t t) { return (9 * t + 19) / 20; }

// Reverse complement of a 2K-bit value, base by base.
inline u128 rc_value(u128 v, int K) {
  u128 r = 0;
  for (int t = 0; t < K; t++) {
    r = (r << 2) | (u128)(3 - (uint64_t)(v & 3));
    v >>= 2;
  }
  return r;
}
// Minimizer value of a k-mer from its canonical form: the canonical m-mer of
// least order among its w = K-m+1 m-mers.
inline uint32_t minimizer_value(u128 c, int K, int m) {
  u128 r = rc_value(c, K);
  uint32_t M = (uint32_t)((1ULL << (2 * m)) - 1);
  uint32_t best = 0xFFFFFFFFu, bv = 0;
  for (int j = 0; j + m <= K; j++) {
    uint32_t f = (uint32_t)(c >> (2 * (K - m - j))) & M;
    uint32_t b = (uint32_t)(r >> (2 * j)) & M;
    uint32_t cv = std::min(f, b);
    uint32_t o = mmer_order(cv);
    if (o < best) {
      best = o;
      bv = cv;
    }
  }
  return bv;
}
inline int segment_bits(uint64_t W, uint64_t bins, uint64_t target) {
  int sb = 0;
  while (sb < 12 && W > bins * (1ULL << sb) * target) sb++;
  return sb;
}
// The segment bits of a minimizer value (below its bucket bits).
inline uint64_t seg_sub_min(uint32_t v, int sb) { return sb ? (xxh3_64_u64le(v) >> (48 - sb)) & ((1ULL << sb) - 1) : 0; }

struct Mphf {
  int K = 0, m = 0, sb = 0;
  bool wide = false;
  uint64_t n_keys = 0;
  std::vector<uint16_t> b2b;
  std::vector<uint64_t> key_off, blk_off, blocks;
  uint64_t segment(u128 c) const {
    uint32_t v = minimizer_value(c, K, m);
    uint64_t bin = b2b[bucket_hash(v)];
    return (bin << sb) | seg_sub_min(v, sb);
  }
  uint64_t lookup(uint64_t lo, uint64_t hi) const {
    u128 c = ((u128)hi << 64) | lo;
    uint64_t seg = segment(c);
    uint64_t ko = key_off[seg], t = key_off[seg + 1] - ko;
    const uint64_t* base = blocks.data() + blk_off[seg] * 8;
    // only the levels inside the segment's blocks (the build stores levels
    // until every key is placed; the level sizes never reach 0)
    const uint64_t lim = (blk_off[seg + 1] - blk_off[seg]) * BBH_BLOCK_BITS;
    uint64_t off = 0;
    for (int l = 0; l < BBH_MAX_LEVELS && t > 0; l++) {
      uint64_t n = bbh_level_bits(t);
      if (off + n > lim) break;
      uint64_t pos = off + bbh_pos(lo, hi, wide, l, n);
      const uint64_t* B = base + (pos / BBH_BLOCK_BITS) * 8;
      uint64_t b = pos % BBH_BLOCK_BITS;
      uint64_t wd = B[1 + b / 64];
      if ((wd >> (b % 64)) & 1) {
        uint64_t r = ko + B[0];
        for (uint64_t u = 0; u < b / 64; u++) r += (uint64_t)__builtin_popcountll(B[1 + u]);
        return r + (uint64_t)__builtin_popcountll(wd & ((1ULL << (b % 64)) - 1));
      }
      off += n;
      t = bbh_next_t(t);
    }
    return 0;  // absent key: the reference maps absent keys to 0 (mphf.rs:269)
  }
};

// Build over the consolidated keys (bins in order, each bin sorted) given the
// per-bin distinct counts; lo/hi are reordered in place into MPHF value
// order (lo[v], hi[v] = the key whose value is v). Threads split the segments.
void mphf_build(Mphf& out, std::vector<uint64_t>& lo, std::vector<uint64_t>& hi, const std::vector<uint64_t>& per_bin,
                const std::vector<uint16_t>& b2b, int K, int m, uint64_t W, uint64_t target, int threads);

// ---------------------------------------------------------------- threading
void parallel_for(size_t n_tasks, int threads, const std::function<void(size_t)>& fn);

}  // namespace cf1ref
