// cf1ref — CPU oracle pipeline (P1..P5) and its C entry points for ctypes.
// TEST INFRASTRUCTURE ONLY (see cf1ref.hpp): loaded by tests/, the smoke check
// and bench.py's cpu_baseline leg; never by the product.
#include <atomic>
#include "cf1ref.hpp"

#include <chrono>
#include <cstdio>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>

namespace cf1ref {

const DnaTables kDna;
const RcByteTable kRcByte;
const StateTables kState;

void parallel_for(size_t n_tasks, int threads, const std::function<void(size_t)>& fn) {
  if (threads <= 1 || n_tasks <= 1) {
    for (size_t i = 0; i < n_tasks; i++) fn(i);
    return;
  }
  std::atomic<size_t> next{0};
  int nt = (int)std::min<size_t>((size_t)threads, n_tasks);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&]() {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= n_tasks) break;
        fn(i);
      }
    });
  for (auto& th : pool) th.join();
}

// ------------------------------------------------------------------ MPHF build
static void build_segment(const uint64_t* lo, const uint64_t* hi, uint64_t D, bool wide, std::vector<uint64_t>& blocks_out) {
  std::vector<char> alive(D, 1);
  std::vector<uint64_t> region;  // level words, bit-contiguous
  uint64_t off = 0, t = D, r = D;
  int l = 0;
  while (r > 0) {
    if (l >= BBH_MAX_LEVELS) throw std::runtime_error("MPHF: too many levels");
    if (t >= (1ULL << 31)) throw std::runtime_error("MPHF: segment too large");
    uint64_t n = bbh_level_bits(t), nw = n / 64;
    std::vector<uint64_t> seen(nw, 0), col(nw, 0);
    for (uint64_t i = 0; i < D; i++) {
      if (!alive[i]) continue;
      uint64_t pos = bbh_pos(lo[i], wide ? hi[i] : 0, wide, l, n);
      uint64_t bit = 1ULL << (pos % 64);
      if (seen[pos / 64] & bit) col[pos / 64] |= bit;
      seen[pos / 64] |= bit;
    }
    for (uint64_t i = 0; i < nw; i++) region.push_back(seen[i] & ~col[i]);
    uint64_t left = 0;
    for (uint64_t i = 0; i < D; i++) {
      if (!alive[i]) continue;
      uint64_t pos = bbh_pos(lo[i], wide ? hi[i] : 0, wide, l, n);
      if (((seen[pos / 64] & ~col[pos / 64]) >> (pos % 64)) & 1) alive[i] = 0;
      else left++;
    }
    r = left;
    off += n;
    t = bbh_next_t(t);
    l++;
  }
  uint64_t nblk = (off + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS;
  blocks_out.assign(nblk * 8, 0);
  uint64_t rank = 0;
  for (uint64_t j = 0; j < nblk; j++) {
    blocks_out[j * 8] = rank;
    for (int u = 0; u < 7; u++) {
      uint64_t q = j * 7 + (uint64_t)u;
      uint64_t w = q < region.size() ? region[q] : 0;
      blocks_out[j * 8 + 1 + (uint64_t)u] = w;
      rank += (uint64_t)__builtin_popcountll(w);
    }
  }
  if (rank != D) throw std::runtime_error("MPHF: rank total mismatch");
}

void mphf_build(Mphf& out, std::vector<uint64_t>& lo, std::vector<uint64_t>& hi, const std::vector<uint64_t>& per_bin,
                const std::vector<uint16_t>& b2b, int K, int m, uint64_t W, uint64_t target, int threads) {
  out = Mphf();
  out.K = K;
  out.m = m;
  out.wide = K > 32;
  out.n_keys = lo.size();
  out.b2b = b2b;
  const uint64_t bins = per_bin.size();
  out.sb = segment_bits(W, bins, target);
  const int sb = out.sb;
  const uint64_t ns = bins << sb;
  out.key_off.assign(ns + 1, 0);
  // group each sorted bin by segment (stable counting sort)
  std::vector<uint64_t> bin_start(bins + 1, 0);
  for (uint64_t b = 0; b < bins; b++) bin_start[b + 1] = bin_start[b] + per_bin[b];
  if (bin_start[bins] != lo.size()) throw std::runtime_error("MPHF: per-bin counts do not cover the keys");
  std::vector<uint64_t> nlo(lo.size()), nhi(hi.size());
  std::vector<uint64_t> cnt(ns + 1, 0);
  std::vector<uint16_t> sub(lo.size());  // segment bits of each key's minimizer
  parallel_for(bins, threads, [&](size_t b) {
    const uint64_t a = bin_start[b], e = bin_start[b + 1];
    uint64_t* c = cnt.data() + (b << sb);
    for (uint64_t i = a; i < e; i++) {
      if (i > a && !(((u128)hi[i - 1] << 64 | lo[i - 1]) < ((u128)hi[i] << 64 | lo[i])))
        throw std::runtime_error("MPHF: keys of a bin are not sorted and distinct");
      sub[i] = (uint16_t)(sb ? seg_sub_min(minimizer_value(((u128)hi[i] << 64) | lo[i], K, m), sb) : 0);
      c[sub[i]]++;
    }
    std::vector<uint64_t> at(1ULL << sb);
    uint64_t run = a;
    for (uint64_t s = 0; s < (1ULL << sb); s++) {
      at[s] = run;
      run += c[s];
    }
    for (uint64_t i = a; i < e; i++) {
      uint64_t d = at[sub[i]]++;
      nlo[d] = lo[i];
      nhi[d] = hi[i];
    }
  });
  uint64_t pos = 0;
  for (uint64_t s = 0; s < ns; s++) {
    out.key_off[s] = pos;
    pos += cnt[s];
  }
  out.key_off[ns] = pos;
  lo.swap(nlo);
  hi.swap(nhi);
  std::vector<std::vector<uint64_t>> blk(ns);
  parallel_for(ns, threads, [&](size_t s) {
    uint64_t a = out.key_off[s], d = out.key_off[s + 1] - a;
    if (d) build_segment(lo.data() + a, hi.data() + a, d, out.wide, blk[s]);
  });
  out.blk_off.assign(ns + 1, 0);
  uint64_t tot = 0;
  for (uint64_t s = 0; s < ns; s++) {
    out.blk_off[s] = tot;
    tot += blk[s].size() / 8;
  }
  out.blk_off[ns] = tot;
  out.blocks.assign(tot * 8, 0);
  for (uint64_t s = 0; s < ns; s++) std::copy(blk[s].begin(), blk[s].end(), out.blocks.begin() + (ptrdiff_t)(out.blk_off[s] * 8));
  // Value order: key i moves to its MPHF value; the values of each segment
  // must be a permutation of the segment's key range.
  std::vector<uint64_t> vlo(lo.size()), vhi(hi.size());
  std::vector<uint8_t> hit(lo.size(), 0);
  std::atomic<bool> bad{false};
  parallel_for(ns, threads, [&](size_t s) {
    for (uint64_t i = out.key_off[s]; i < out.key_off[s + 1]; i++) {
      const uint64_t v = out.lookup(lo[i], hi[i]);
      if (v < out.key_off[s] || v >= out.key_off[s + 1] || hit[v]) {
        bad = true;
        return;
      }
      hit[v] = 1;
      vlo[v] = lo[i];
      vhi[v] = hi[i];
    }
  });
  if (bad) throw std::runtime_error("MPHF: values are not a permutation of the keys");
  lo.swap(vlo);
  hi.swap(vhi);
}

// ------------------------------------------------------------------ minimizer scan
// Emits every super k-mer whose first k-mer position lies in [lo, hi) of one
// N-free run [a, b) of a record (b - a >= K). cb(start, n_kmers, value).
template <typename CB>
static void superkmers_in_run(const uint8_t* seq, size_t a, size_t b, int K, int m, size_t lo, size_t hi, CB cb) {
  const int w = K - m + 1;
  const size_t last_kmer = b - (size_t)K;  // inclusive
  if (lo < a) lo = a;
  if (hi > last_kmer + 1) hi = last_kmer + 1;
  if (lo >= hi) return;
  // m-mer orders are computed lazily over [first_m, ...).
  size_t first_i = (lo > a) ? lo - 1 : lo;  // need p(lo-1) to decide if lo starts
  uint64_t mask = (m == 32) ? ~0ULL : ((1ULL << (2 * m)) - 1);
  std::deque<std::pair<size_t, uint32_t>> dq;  // (position, order) increasing order
  std::vector<uint32_t> ord;
  std::vector<uint32_t> canon;
  size_t mbase = first_i;
  uint64_t fw = 0, rc = 0;
  // prime rolling m-mer over [first_i, first_i + m - 1)
  for (int t = 0; t < m - 1; t++) {
    uint8_t c = map_base(seq[first_i + (size_t)t]);
    fw = ((fw << 2) | c) & mask;
    rc = (rc >> 2) | ((uint64_t)(3 - c) << (2 * (m - 1)));
  }
  size_t next_m = first_i;  // next m-mer position to push
  auto push_mmer = [&]() {
    uint8_t c = map_base(seq[next_m + (size_t)m - 1]);
    fw = ((fw << 2) | c) & mask;
    rc = (rc >> 2) | ((uint64_t)(3 - c) << (2 * (m - 1)));
    uint32_t cv = (uint32_t)std::min(fw, rc);
    uint32_t o = mmer_order(cv);
    while (!dq.empty() && dq.back().second > o) dq.pop_back();
    dq.emplace_back(next_m, o);
    ord.push_back(o);
    canon.push_back(cv);
    next_m++;
  };
  auto p_of = [&](size_t i) -> size_t {  // minimizer position of k-mer i
    while (next_m < i + (size_t)w) push_mmer();
    while (dq.front().first < i) dq.pop_front();
    return dq.front().first;
  };
  size_t prev_p = (size_t)-1;
  if (first_i < lo) prev_p = p_of(first_i);
  size_t cur_start = (size_t)-1, cur_p = 0;
  for (size_t i = lo;; i++) {
    bool in_range = i <= last_kmer;
    size_t p = in_range ? p_of(i) : (size_t)-1;
    bool is_start = in_range && (i == a || p != prev_p);
    if (is_start || !in_range) {
      if (cur_start != (size_t)-1) cb(cur_start, i - cur_start, (uint64_t)canon[cur_p - mbase]);
      cur_start = (size_t)-1;
      if (!in_range || i >= hi) break;
      cur_start = i;
      cur_p = p;
    }
    prev_p = p;
  }
}

// Invoke fn(run_a, run_b) for each maximal N-free stretch of length >= K.
template <typename FN>
static void for_each_run(const uint8_t* seq, size_t len, int K, FN fn) {
  size_t i = 0;
  while (i < len) {
    while (i < len && is_placeholder(seq[i])) i++;
    size_t a = i;
    while (i < len && !is_placeholder(seq[i])) i++;
    if (i - a >= (size_t)K) fn(a, i);
  }
}

// ------------------------------------------------------------------ pipeline
struct Records {
  std::vector<uint8_t> ascii;
  std::vector<uint64_t> off;  // n+1
  std::vector<std::string> names;
  std::vector<uint32_t> file;  // input file index per record (output.rs:176 file_idx)
  size_t n() const { return off.empty() ? 0 : off.size() - 1; }
  const uint8_t* seq(size_t r) const { return ascii.data() + off[r]; }
  size_t len(size_t r) const { return (size_t)(off[r + 1] - off[r]); }
};

struct P5Meta {  // output.rs:34-75 UnipathsMeta
  uint64_t unipath_count = 0, kmer_count = 0;
  uint64_t max_len = 0, min_len = UINT64_MAX, sum_len = 0;
  void add_unitig(uint64_t vertex_count, int k) {
    unipath_count++;
    kmer_count += vertex_count;
    uint64_t l = vertex_count + (uint64_t)k - 1;
    max_len = std::max(max_len, l);
    min_len = std::min(min_len, l);
    sum_len += l;
  }
  void aggregate(const P5Meta& o) {
    unipath_count += o.unipath_count;
    kmer_count += o.kmer_count;
    max_len = std::max(max_len, o.max_len);
    min_len = std::min(min_len, o.min_len);
    sum_len += o.sum_len;
  }
};

struct PhaseTimes {
  double p1 = 0, p2 = 0, p3a = 0, p3b = 0, p4 = 0, p5 = 0;
};

class PipelineBase {
 public:
  virtual ~PipelineBase() {}
  Records rec;
  int k = 31, m = 15, threads = 1, num_bins_req = 128;
  std::vector<uint64_t> hist;
  std::vector<uint16_t> bucket_to_bin;
  int num_bins = 0;
  std::vector<std::vector<uint8_t>> bins;  // reference bin record format
  std::vector<uint64_t> keys_lo, keys_hi;  // consolidated dedup keys
  std::vector<uint64_t> per_bin_distinct;
  Mphf mphf;
  uint64_t target_seg = 0;  // 0 = default (1536)
  uint64_t n_kmers = 0;     // W, valid k-mer occurrences decoded in P3a
  AtomicStateVector states;
  PhaseTimes times;
  P5Meta meta;
  std::string out_seg, out_seq;
  std::vector<std::pair<std::string, uint64_t>> short_seqs;

  virtual void p1() = 0;
  void p2_partition();
  virtual void p2_route() = 0;
  virtual void p3a() = 0;
  void p3b() {
    uint64_t tgt = target_seg ? target_seg : 1536;
    mphf_build(mphf, keys_lo, keys_hi, per_bin_distinct, bucket_to_bin, k, m, n_kmers, tgt, threads);
  }
  virtual void p4() = 0;
  virtual void p5(bool poly_n_stretch) = 0;
  virtual uint64_t kmer_from_ascii(const char* s, uint64_t* hi) = 0;
};

// minimizer.rs:116-144 (partition_minimizers) and :156-172 (from_boundaries).
static void partition_minimizers(const uint64_t* histogram, size_t n, int num_bins, std::vector<uint16_t>& b2b, int& out_bins) {
  uint64_t total = 0;
  for (size_t i = 0; i < n; i++) total += histogram[i];
  uint64_t target = total / (uint64_t)num_bins;
  std::vector<uint16_t> boundaries;
  uint64_t current = 0;
  for (size_t bucket = 0; bucket < n; bucket++) {
    current += histogram[bucket];
    if (current >= target && boundaries.size() < (size_t)(num_bins - 1)) {
      boundaries.push_back((uint16_t)bucket);
      current -= target;
    }
  }
  out_bins = (int)boundaries.size() + 1;
  b2b.assign(NUM_HISTOGRAM_BUCKETS, 0);
  uint16_t bin = 0;
  size_t bi = 0;
  for (size_t bucket = 0; bucket < NUM_HISTOGRAM_BUCKETS; bucket++) {
    if (bi < boundaries.size() && (uint16_t)bucket > boundaries[bi]) {
      bin++;
      bi++;
    }
    b2b[bucket] = bin;
  }
}

void PipelineBase::p2_partition() {
  partition_minimizers(hist.data(), hist.size(), num_bins_req, bucket_to_bin, num_bins);
}

constexpr size_t CHUNK_BASES = 4000000;   // minimizer.rs:13, superkmer.rs:16
constexpr size_t CHUNK_KMERS = 4000000;   // classify.rs:13

template <int K>
class Pipeline : public PipelineBase {
 public:
  typedef Kmer<K> KM;

  // Tasks over (record, k-mer-position chunk) for record-scale work.
  struct Task {
    size_t r, lo, hi;
  };
  std::vector<Task> make_tasks(size_t chunk) const {
    std::vector<Task> t;
    for (size_t r = 0; r < rec.n(); r++) {
      size_t len = rec.len(r);
      if (len < (size_t)K) continue;
      size_t last = len - K;
      for (size_t lo = 0; lo <= last; lo += chunk) t.push_back({r, lo, std::min(last + 1, lo + chunk)});
    }
    return t;
  }

  // The sequences P1 and P2 scan (minimizer.rs:57-86, superkmer.rs:57-88):
  // a record shorter than K is skipped, one shorter than CHUNK_BASES is
  // scanned whole, and a longer one in chunks starting every CHUNK_BASES
  // bases. P1's chunks are CHUNK_BASES long (no overlap: k-mers across a
  // chunk edge are not counted); P2's extend `overlap` = K-1 bases into the
  // next chunk, so every k-mer is routed once and a super k-mer always
  // starts at a chunk's first k-mer. Each piece is scanned as a sequence of
  // its own.
  struct Piece {
    size_t r, a, b;  // bases [a, b) of record r
  };
  std::vector<Piece> scan_pieces(size_t overlap) const {
    std::vector<Piece> out;
    for (size_t r = 0; r < rec.n(); r++) {
      const size_t len = rec.len(r);
      if (len < (size_t)K) continue;
      if (len < CHUNK_BASES) {
        out.push_back({r, 0, len});
        continue;
      }
      for (size_t a = 0; a < len; a += CHUNK_BASES) out.push_back({r, a, std::min(a + CHUNK_BASES + overlap, len)});
    }
    return out;
  }
  // cb(record-relative start, n_kmers, minimizer value) for every super
  // k-mer of a piece.
  template <typename CB>
  void superkmers_of_piece(const Piece& p, CB cb) const {
    const uint8_t* s = rec.seq(p.r) + p.a;
    for_each_run(s, p.b - p.a, K, [&](size_t a, size_t b) {
      superkmers_in_run(s, a, b, K, m, a, b, [&](size_t st, size_t n, uint64_t v) { cb(p.a + st, n, v); });
    });
  }

  // P1 — minimizer.rs:27-98: one count per super k-mer at bucket_hash(v).
  void p1() override {
    auto pieces = scan_pieces(0);
    std::vector<std::atomic<uint64_t>> h(NUM_HISTOGRAM_BUCKETS);
    for (auto& x : h) x.store(0);
    parallel_for(pieces.size(), threads, [&](size_t pi) {
      superkmers_of_piece(pieces[pi], [&](size_t, size_t, uint64_t v) { h[bucket_hash(v)].fetch_add(1, std::memory_order_relaxed); });
    });
    hist.resize(NUM_HISTOGRAM_BUCKETS);
    for (size_t i = 0; i < NUM_HISTOGRAM_BUCKETS; i++) hist[i] = h[i].load();
  }

  // superkmer.rs:181-210 pack_and_buffer: u16 LE base count, then bases
  // packed 4 per byte, base j at bits 6-2*(j%4) of byte j/4.
  static void pack_and_buffer(const uint8_t* ascii, size_t len, std::vector<uint8_t>& buf) {
    uint16_t bl = (uint16_t)len;
    buf.push_back((uint8_t)(bl & 0xFF));
    buf.push_back((uint8_t)(bl >> 8));
    size_t nb = (len + 3) / 4;
    size_t o = buf.size();
    buf.resize(o + nb, 0);
    for (size_t i = 0; i < len; i++) buf[o + i / 4] |= (uint8_t)(map_base(ascii[i]) << (6 - 2 * (i % 4)));
  }
  // superkmer.rs:156-178 write_packed_segments: split at N, keep runs >= K.
  static void write_packed_segments(const uint8_t* sub, size_t len, std::vector<uint8_t>& buf) {
    size_t seg_start = 0;
    for (size_t i = 0; i < len; i++) {
      if (is_placeholder(sub[i])) {
        if (i - seg_start >= (size_t)K) pack_and_buffer(sub + seg_start, i - seg_start, buf);
        seg_start = i + 1;
      }
    }
    if (len - seg_start >= (size_t)K) pack_and_buffer(sub + seg_start, len - seg_start, buf);
  }

  // P2 — superkmer.rs:21-153: each super k-mer [s, s+n-1+K) of a piece to its bin.
  void p2_route() override {
    auto pieces = scan_pieces((size_t)K - 1);
    bins.assign((size_t)num_bins, {});
    std::vector<std::mutex> mu((size_t)num_bins);
    parallel_for(pieces.size(), threads, [&](size_t pi) {
      std::vector<std::vector<uint8_t>> local((size_t)num_bins);
      const uint8_t* s = rec.seq(pieces[pi].r);
      superkmers_of_piece(pieces[pi], [&](size_t start, size_t n, uint64_t v) {
        int bin = bucket_to_bin[bucket_hash(v)];
        write_packed_segments(s + start, n - 1 + K, local[(size_t)bin]);
      });
      for (int b = 0; b < num_bins; b++) {
        if (local[(size_t)b].empty()) continue;
        std::lock_guard<std::mutex> g(mu[(size_t)b]);
        bins[(size_t)b].insert(bins[(size_t)b].end(), local[(size_t)b].begin(), local[(size_t)b].end());
      }
    });
  }

  // P3a — mphf.rs:323-363 dedup_bin: decode every k-mer of every record,
  // canonicalise, sort, dedup; bins concatenated in order (mphf.rs:120-132).
  void p3a() override {
    std::vector<std::vector<typename KM::Storage>> per((size_t)num_bins);
    std::vector<uint64_t> occ((size_t)num_bins, 0);
    parallel_for((size_t)num_bins, threads, [&](size_t b) {
      const std::vector<uint8_t>& d = bins[b];
      auto& out = per[b];
      size_t p = 0;
      while (p + 2 <= d.size()) {  // BinReader::read_packed_superkmer superkmer.rs:276-288
        size_t bc = (size_t)d[p] | ((size_t)d[p + 1] << 8);
        p += 2;
        const uint8_t* packed = d.data() + p;
        if (bc >= (size_t)K)
          for (size_t i = 0; i + K <= bc; i++) out.push_back(KM::from_packed_2bit(packed, i).canonical().bits);
        p += (bc + 3) / 4;
      }
      occ[b] = out.size();
      std::sort(out.begin(), out.end());
      out.erase(std::unique(out.begin(), out.end()), out.end());
    });
    n_kmers = 0;
    for (auto x : occ) n_kmers += x;
    keys_lo.clear();
    keys_hi.clear();
    per_bin_distinct.assign((size_t)num_bins, 0);
    for (size_t b = 0; b < (size_t)num_bins; b++) {
      per_bin_distinct[b] = per[b].size();
      for (auto v : per[b]) {
        keys_lo.push_back((uint64_t)v);
        keys_hi.push_back((uint64_t)((u128)v >> 64));
      }
    }
  }

  uint64_t hash(const KM& canon) const {  // mphf.rs:268-270
    u128 v = (u128)canon.bits;
    return mphf.lookup((uint64_t)v, (uint64_t)(v >> 64));
  }

  // ---------------------------------------------------------------- P4
  // classify.rs:25-519 Classifier, restated with its control flow and CAS loops.
  void process_substring(const uint8_t* seq, size_t seq_len, size_t left_end, size_t right_end) {  // :34-43
    size_t kmer_idx = left_end;
    while (kmer_idx <= right_end) {
      kmer_idx = search_valid_kmer(seq, kmer_idx, right_end);
      if (kmer_idx > right_end) break;
      kmer_idx = process_contiguous_subseq(seq, seq_len, right_end, kmer_idx);
    }
  }
  size_t search_valid_kmer(const uint8_t* seq, size_t start, size_t right_end) {  // :46-61
    while (start <= right_end) {
      bool valid = true;
      for (size_t j = 0; j < (size_t)K; j++) {
        if (is_placeholder(seq[start + j])) {
          start = start + j + 1;
          valid = false;
          break;
        }
      }
      if (valid) return start;
    }
    return right_end + 1;
  }
  size_t process_contiguous_subseq(const uint8_t* seq, size_t seq_len, size_t right_end, size_t start_idx) {  // :65-141
    size_t kmer_idx = start_idx;
    DirectedKmer<K> curr(KM::from_ascii(seq, kmer_idx));
    bool no_left = kmer_idx == 0 || is_placeholder(seq[kmer_idx - 1]);
    bool no_right = kmer_idx + K == seq_len || is_placeholder(seq[kmer_idx + K]);
    if (no_left && no_right) {
      while (!process_isolated_kmer(curr)) {
      }
    } else {
      if (no_right) {
        while (!process_rightmost_kmer(curr, seq[kmer_idx - 1])) {
        }
        return kmer_idx + K;
      }
      DirectedKmer<K> next = curr;
      next.roll_to_next_kmer(seq[kmer_idx + K]);
      if (no_left) {
        while (!process_leftmost_kmer(curr, next, seq[kmer_idx + K])) {
        }
      } else {
        while (!process_internal_kmer(curr, next, seq[kmer_idx - 1], seq[kmer_idx + K])) {
        }
      }
      kmer_idx += 1;
      while (kmer_idx < right_end && !is_placeholder(seq[kmer_idx + K])) {
        curr = next;
        next.roll_to_next_kmer(seq[kmer_idx + K]);
        while (!process_internal_kmer(curr, next, seq[kmer_idx - 1], seq[kmer_idx + K])) {
        }
        kmer_idx += 1;
      }
      if (kmer_idx <= right_end) {
        curr = next;
        if (kmer_idx + K == seq_len || is_placeholder(seq[kmer_idx + K])) {
          while (!process_rightmost_kmer(curr, seq[kmer_idx - 1])) {
          }
        } else {
          next.roll_to_next_kmer(seq[kmer_idx + K]);
          while (!process_internal_kmer(curr, next, seq[kmer_idx - 1], seq[kmer_idx + K])) {
          }
        }
      } else {
        kmer_idx -= 1;
      }
    }
    return kmer_idx + K;
  }
  bool process_isolated_kmer(const DirectedKmer<K>& km) {  // :147-158
    size_t idx = (size_t)hash(km.canonical);
    uint8_t s = states.get(idx);
    if (st_is_dead_end(s)) return true;
    return states.compare_and_swap(idx, s, state_from_vertex(v_mimo()));
  }
  bool process_loop(const DirectedKmer<K>& km, const DirectedKmer<K>& next, uint8_t prev_char) {  // :160-177
    if (prev_char == 0 || km.kmer == next.kmer) {
      size_t idx = (size_t)hash(km.canonical);
      uint8_t s = states.get(idx);
      return states.compare_and_swap(idx, s, state_from_vertex(v_mimo()));
    }
    return process_rightmost_kmer(km, prev_char);
  }
  static uint8_t with_class(const Vertex& v, StateClass c) {
    Vertex x = v;
    x.state_class = c;
    return state_from_vertex(x);
  }
  bool process_leftmost_kmer(const DirectedKmer<K>& km, const DirectedKmer<K>& next, uint8_t next_char) {  // :179-285
    size_t idx = (size_t)hash(km.canonical);
    uint8_t s = states.get(idx);
    if (st_is_dead_end(s)) return true;
    if (km.canonical == next.canonical) return process_loop(km, next, 0);
    uint8_t old = s, ns;
    uint8_t nb = map_base(next_char);
    if (km.dir) {
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_miso(nb));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO)
          ns = (v.back == nb) ? with_class(v, MISO) : with_class(v, MIMO);
        else if (v.state_class == MISO)
          ns = (v.back != nb) ? with_class(v, MIMO) : old;
        else
          ns = with_class(v, MIMO);
      }
    } else {
      uint8_t cn = complement_base(nb);
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_simo(cn));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO)
          ns = (v.front == cn) ? with_class(v, SIMO) : with_class(v, MIMO);
        else if (v.state_class == MISO)
          ns = with_class(v, MIMO);
        else
          ns = (v.front != cn) ? with_class(v, MIMO) : old;
      }
    }
    if (ns == old) return true;
    return states.compare_and_swap(idx, old, ns);
  }
  bool process_rightmost_kmer(const DirectedKmer<K>& km, uint8_t prev_char) {  // :287-383
    size_t idx = (size_t)hash(km.canonical);
    uint8_t s = states.get(idx);
    if (st_is_dead_end(s)) return true;
    uint8_t old = s, ns;
    uint8_t pb = map_base(prev_char);
    if (km.dir) {
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_simo(pb));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO)
          ns = (v.front == pb) ? with_class(v, SIMO) : with_class(v, MIMO);
        else if (v.state_class == MISO)
          ns = with_class(v, MIMO);
        else
          ns = (v.front != pb) ? with_class(v, MIMO) : old;
      }
    } else {
      uint8_t cp = complement_base(pb);
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_miso(cp));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO)
          ns = (v.back == cp) ? with_class(v, MISO) : with_class(v, MIMO);
        else if (v.state_class == MISO)
          ns = (v.back != cp) ? with_class(v, MIMO) : old;
        else
          ns = with_class(v, MIMO);
      }
    }
    if (ns == old) return true;
    return states.compare_and_swap(idx, old, ns);
  }
  bool process_internal_kmer(const DirectedKmer<K>& km, const DirectedKmer<K>& next, uint8_t prev_char, uint8_t next_char) {  // :385-518
    size_t idx = (size_t)hash(km.canonical);
    uint8_t s = states.get(idx);
    if (st_is_dead_end(s)) return true;
    if (km.canonical == next.canonical) return process_loop(km, next, prev_char);
    uint8_t old = s, ns;
    uint8_t pb = map_base(prev_char), nb = map_base(next_char);
    if (km.dir) {
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_siso(pb, nb));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO) {
          if (v.front == pb && v.back == nb) return true;
          if (v.front != pb && v.back != nb)
            ns = with_class(v, MIMO);
          else if (v.front != pb)
            ns = with_class(v, MISO);
          else
            ns = with_class(v, SIMO);
        } else if (v.state_class == MISO) {
          ns = (v.back != nb) ? with_class(v, MIMO) : old;
        } else {
          ns = (v.front != pb) ? with_class(v, MIMO) : old;
        }
      }
    } else {
      uint8_t cn = complement_base(nb), cp = complement_base(pb);
      if (!st_is_visited(s)) {
        ns = state_from_vertex(v_siso(cn, cp));
      } else {
        Vertex v = st_decode(s);
        if (v.state_class == SISO) {
          if (v.front == cn && v.back == cp) return true;
          if (v.front != cn && v.back != cp)
            ns = with_class(v, MIMO);
          else if (v.front != cn)
            ns = with_class(v, MISO);
          else
            ns = with_class(v, SIMO);
        } else if (v.state_class == MISO) {
          ns = (v.back != cp) ? with_class(v, MIMO) : old;
        } else {
          ns = (v.front != cn) ? with_class(v, MIMO) : old;
        }
      }
    }
    if (ns == old) return true;
    return states.compare_and_swap(idx, old, ns);
  }

  // classify.rs:524-601 — records < 4M k-mers as one task, larger ones in
  // CHUNK_KMERS ranges reading neighbours from the full record.
  void p4() override {
    states.reset(keys_lo.size());
    auto tasks = make_tasks(CHUNK_KMERS);
    parallel_for(tasks.size(), threads, [&](size_t ti) {
      const Task& t = tasks[ti];
      process_substring(rec.seq(t.r), rec.len(t.r), t.lo, t.hi - 1);
    });
  }

  // ---------------------------------------------------------------- P5
  // output.rs:14-782, restated.
  struct Annot {  // directed_kmer.rs:70-109 AnnotatedKmer
    DirectedKmer<K> d;
    size_t idx = 0;
    StateClass cls = SISO;
  };
  Annot annotate(const KM& km, size_t idx) const {
    Annot a;
    a.d = DirectedKmer<K>(km);
    a.idx = idx;
    a.cls = st_class(states.get((size_t)hash(a.d.canonical)));
    return a;
  }
  void roll(Annot& a, uint8_t next_char) const {
    a.d.roll_to_next_kmer(next_char);
    a.idx += 1;
    a.cls = st_class(states.get((size_t)hash(a.d.canonical)));
  }
  static bool is_unipath_start(StateClass c, bool dir, StateClass pc, bool pdir) {  // output.rs:79-110
    if (c == MIMO) return true;
    if (dir) {
      if (c == MISO) return true;
    } else if (c == SIMO) {
      return true;
    }
    if (pc == MIMO) return true;
    if (pdir) {
      if (pc == SIMO) return true;
    } else if (pc == MISO) {
      return true;
    }
    return false;
  }
  static bool is_unipath_end(StateClass c, bool dir, StateClass nc, bool ndir) {  // output.rs:114-145
    if (c == MIMO) return true;
    if (dir) {
      if (c == SIMO) return true;
    } else if (c == MISO) {
      return true;
    }
    if (nc == MIMO) return true;
    if (ndir) {
      if (nc == MISO) return true;
    } else if (nc == SIMO) {
      return true;
    }
    return false;
  }
  struct OU {
    uint64_t id;
    bool dir;
    size_t s, e;
  };
  void output_gfa_unitig(const uint8_t* seq, const Annot& st, const Annot& en, std::string& seg, std::vector<OU>& u, P5Meta& mt) {  // output.rs:714-782
    KM mfl = std::min(st.d.canonical, en.d.canonical);
    uint64_t id = hash(mfl);
    bool dir = st.d.kmer < en.d.rev_compl;
    u.push_back({id, dir, st.idx, en.idx});
    uint8_t s = states.get((size_t)id);
    if (!st_is_outputted(s)) {
      uint8_t o = st_outputted(s);
      if (states.compare_and_swap((size_t)id, s, o)) {
        size_t L = en.idx - st.idx + K;
        seg += std::to_string(id);
        seg.push_back('\t');
        if (dir)
          for (size_t t = 0; t < L; t++) seg.push_back((char)to_upper(seq[st.idx + t]));
        else
          for (size_t t = 0; t < L; t++) seg.push_back((char)complement_char(seq[en.idx + K - 1 - t]));
        seg.push_back('\n');
        mt.add_unitig(en.idx - st.idx + 1, K);
      }
    }
  }
  size_t output_unitigs_contiguous(const uint8_t* seq, size_t seq_len, size_t right_end, size_t start_idx, std::string& seg, std::vector<OU>& u, P5Meta& mt) {  // output.rs:506-709
    size_t kmer_idx = start_idx;
    Annot curr = annotate(KM::from_ascii(seq, kmer_idx), kmer_idx);
    bool no_left = kmer_idx == 0 || is_placeholder(seq[kmer_idx - 1]);
    bool no_right = kmer_idx + K == seq_len || is_placeholder(seq[kmer_idx + K]);
    if (no_left && no_right) {
      output_gfa_unitig(seq, curr, curr, seg, u, mt);
    } else {
      if (no_right) {
        if (no_left) {
          output_gfa_unitig(seq, curr, curr, seg, u, mt);
        } else {
          Annot prev = annotate(KM::from_ascii(seq, kmer_idx - 1), kmer_idx);
          if (is_unipath_start(curr.cls, curr.d.dir, prev.cls, prev.d.dir)) output_gfa_unitig(seq, curr, curr, seg, u, mt);
        }
        return kmer_idx + K;
      }
      Annot next = curr;
      roll(next, seq[kmer_idx + K]);
      bool on = false;
      Annot ustart = curr;
      Annot prev = curr;
      if (no_left) {
        on = true;
        ustart = curr;
      } else {
        prev = annotate(KM::from_ascii(seq, kmer_idx - 1), kmer_idx);
        if (is_unipath_start(curr.cls, curr.d.dir, prev.cls, prev.d.dir)) {
          on = true;
          ustart = curr;
        }
      }
      if (on && is_unipath_end(curr.cls, curr.d.dir, next.cls, next.d.dir)) {
        output_gfa_unitig(seq, ustart, curr, seg, u, mt);
        on = false;
      }
      kmer_idx += 1;
      for (;;) {
        if (!on && kmer_idx > right_end) break;
        prev = curr;
        curr = next;
        if (is_unipath_start(curr.cls, curr.d.dir, prev.cls, prev.d.dir)) {
          on = true;
          ustart = curr;
        }
        if (kmer_idx + K == seq_len || is_placeholder(seq[kmer_idx + K])) {
          if (on) output_gfa_unitig(seq, ustart, curr, seg, u, mt);
          return kmer_idx + K;
        }
        roll(next, seq[kmer_idx + K]);
        if (on && is_unipath_end(curr.cls, curr.d.dir, next.cls, next.d.dir)) {
          output_gfa_unitig(seq, ustart, curr, seg, u, mt);
          on = false;
        }
        kmer_idx += 1;
      }
    }
    return kmer_idx + K;
  }
  size_t search_valid_kmer_p5(const uint8_t* seq, size_t start, size_t right_end) { return search_valid_kmer(seq, start, right_end); }

  // output.rs:416-482 extract_unitigs_from_chunk (extract_unitigs_from_seq
  // is the single chunk [0, right_end]).
  void extract_chunk(const uint8_t* seq, size_t seq_len, size_t cs, size_t cre, std::string& seg, std::vector<OU>& u, P5Meta& mt) {
    size_t abs_right_end = seq_len - K;
    size_t kmer_idx = cs;
    while (kmer_idx <= abs_right_end) {
      kmer_idx = search_valid_kmer(seq, kmer_idx, cre);
      if (kmer_idx > cre) break;
      kmer_idx = output_unitigs_contiguous(seq, seq_len, cre, kmer_idx, seg, u, mt);
      if (kmer_idx > cre) break;
    }
  }

  void p5(bool poly_n_stretch) override {  // output.rs:152-354
    constexpr size_t LARGE = 4000000;
    size_t nrec = rec.n();
    std::vector<std::string> tiling(nrec), segs(nrec);
    std::vector<P5Meta> metas(nrec);
    std::vector<char> present(nrec, 0);
    auto tiling_line = [&](size_t r, const std::vector<OU>& us) {  // output.rs:208-259
      std::string t = "Reference:" + std::to_string((uint64_t)(r < rec.file.size() ? rec.file[r] : 0) + 1) + "_Sequence:";  // output.rs:176-177
      t += rec.names.size() > r ? rec.names[r] : std::string();
      t.push_back('\t');
      if (!us.empty()) {
        const OU& f = us[0];
        if (poly_n_stretch && f.s > 0) t += "N" + std::to_string(f.s) + " ";
        t += std::to_string(f.id);
        t.push_back(f.dir ? '+' : '-');
        for (size_t i = 1; i < us.size(); i++) {
          const OU& l = us[i - 1];
          const OU& rr = us[i];
          if (poly_n_stretch && (l.e + 1) != rr.s) {
            int64_t gap = (int64_t)rr.s - (int64_t)l.e;
            if (gap >= (int64_t)K + 1) t += " N" + std::to_string((uint64_t)gap - K);
          }
          t.push_back(' ');
          t += std::to_string(rr.id);
          t.push_back(rr.dir ? '+' : '-');
        }
      }
      t.push_back('\n');
      return t;
    };
    // small records in parallel, large ones chunked per thread
    std::vector<size_t> small, large;
    for (size_t r = 0; r < nrec; r++) {
      if (rec.len(r) < (size_t)K) continue;
      present[r] = 1;
      (rec.len(r) < LARGE ? small : large).push_back(r);
    }
    parallel_for(small.size(), threads, [&](size_t i) {
      size_t r = small[i];
      std::vector<OU> us;
      extract_chunk(rec.seq(r), rec.len(r), 0, rec.len(r) - K, segs[r], us, metas[r]);
      tiling[r] = tiling_line(r, us);
    });
    for (size_t r : large) {
      size_t len = rec.len(r);
      size_t total = len - K + 1;  // output.rs:406-423 compute_chunk_boundaries
      size_t nch = (size_t)std::max(threads, 1);
      std::vector<std::pair<size_t, size_t>> ch;
      for (size_t i = 0; i < nch; i++) {
        size_t s = i * total / nch, e = (i + 1) * total / nch;
        if (s < e) ch.push_back({s, e - 1});
      }
      std::vector<std::vector<OU>> cu(ch.size());
      std::vector<std::string> cs(ch.size());
      std::vector<P5Meta> cm(ch.size());
      parallel_for(ch.size(), threads, [&](size_t c) { extract_chunk(rec.seq(r), len, ch[c].first, ch[c].second, cs[c], cu[c], cm[c]); });
      std::vector<OU> all;
      for (size_t c = 0; c < ch.size(); c++) {
        all.insert(all.end(), cu[c].begin(), cu[c].end());
        segs[r] += cs[c];
        metas[r].aggregate(cm[c]);
      }
      tiling[r] = tiling_line(r, all);
    }
    out_seg.clear();
    out_seq.clear();
    meta = P5Meta();
    for (size_t r = 0; r < nrec; r++) {
      if (!present[r]) continue;
      out_seg += segs[r];
      out_seq += tiling[r];
      meta.aggregate(metas[r]);
    }
  }

  uint64_t kmer_from_ascii(const char* s, uint64_t* hi) override {
    KM km = KM::from_ascii((const uint8_t*)s, 0);
    u128 v = (u128)km.bits;
    if (hi) *hi = (uint64_t)(v >> 64);
    return (uint64_t)v;
  }
};

static PipelineBase* make_pipeline(int k) {
  switch (k) {
#define CF1REF_K(KK) \
  case KK:           \
    return new Pipeline<KK>();
    CF1REF_K(1) CF1REF_K(3) CF1REF_K(5) CF1REF_K(7) CF1REF_K(9) CF1REF_K(11) CF1REF_K(13) CF1REF_K(15)
    CF1REF_K(17) CF1REF_K(19) CF1REF_K(21) CF1REF_K(23) CF1REF_K(25) CF1REF_K(27) CF1REF_K(29) CF1REF_K(31)
    CF1REF_K(33) CF1REF_K(35) CF1REF_K(37) CF1REF_K(39) CF1REF_K(41) CF1REF_K(43) CF1REF_K(45) CF1REF_K(47)
    CF1REF_K(49) CF1REF_K(51) CF1REF_K(53) CF1REF_K(55) CF1REF_K(57) CF1REF_K(59) CF1REF_K(61) CF1REF_K(63)
#undef CF1REF_K
    default:
      return nullptr;
  }
}

// write_json — output.rs:785-834 (serde_json pretty, keys in BTreeMap order).
static std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') {
      o.push_back('\\');
      o.push_back(c);
    } else if ((unsigned char)c < 0x20) {
      char b[8];
      std::snprintf(b, sizeof b, "\\u%04x", (unsigned)(unsigned char)c);
      o += b;
    } else {
      o.push_back(c);
    }
  }
  return o;
}
static std::string make_json(const std::string& input, int k, const std::string& prefix, const P5Meta& m,
                             const std::vector<std::pair<std::string, uint64_t>>& shorts) {
  std::string j = "{\n";
  j += "  \"basic info\": {\n    \"vertex count\": " + std::to_string(m.kmer_count) + "\n  },\n";
  j += "  \"contigs info\": {\n";
  j += "    \"_comment\": \"lengths are in bases\",\n";
  j += "    \"avg. maximal unitig length\": " + std::to_string(m.unipath_count ? m.sum_len / m.unipath_count : 0) + ",\n";
  j += "    \"longest maximal unitig length\": " + std::to_string(m.max_len) + ",\n";
  j += "    \"maximal unitig count\": " + std::to_string(m.unipath_count) + ",\n";
  j += "    \"shortest maximal unitig length\": " + std::to_string(m.min_len == UINT64_MAX ? 0 : m.min_len) + ",\n";
  j += "    \"sum maximal unitig length\": " + std::to_string(m.sum_len) + ",\n";
  j += "    \"vertex count in the maximal unitigs\": " + std::to_string(m.kmer_count) + "\n  },\n";
  j += "  \"parameters info\": {\n";
  j += "    \"input\": \"" + json_escape(input) + "\",\n";
  j += "    \"k\": " + std::to_string(k) + ",\n";
  j += "    \"output prefix\": \"" + json_escape(prefix) + "\"\n  }";
  if (!shorts.empty()) {
    j += ",\n  \"short seqs\": [\n";
    for (size_t i = 0; i < shorts.size(); i++) {
      j += "    [\n      \"" + json_escape(shorts[i].first) + "\",\n      " + std::to_string(shorts[i].second) + "\n    ]";
      j += (i + 1 < shorts.size()) ? ",\n" : "\n";
    }
    j += "  ]";
  }
  j += "\n}";
  return j;
}

}  // namespace cf1ref

// ====================================================================== C entry points
using namespace cf1ref;

namespace {
struct Handle {
  std::unique_ptr<PipelineBase> p;
  std::string json;
};
double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }
}  // namespace

extern "C" {

void* ref_create(int k, int num_bins, int threads) {
  if (k < 1 || k > 63 || k % 2 == 0 || num_bins < 1) return nullptr;
  Handle* h = new Handle();
  h->p.reset(make_pipeline(k));
  h->p->k = k;
  h->p->m = minimizer_len(k);
  h->p->threads = threads < 1 ? 1 : threads;
  h->p->num_bins_req = num_bins;
  return h;
}
void ref_destroy(void* hp) { delete (Handle*)hp; }

int ref_load(void* hp, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec) {
  Handle* h = (Handle*)hp;
  Records& r = h->p->rec;
  r.off.assign(rec_off, rec_off + n_rec + 1);
  uint64_t base = r.off[0];
  for (auto& o : r.off) o -= base;
  r.ascii.assign(ascii + base, ascii + base + r.off.back());
  r.names.clear();
  r.file.assign(n_rec, 0);
  return 0;
}
int ref_set_files(void* hp, const uint32_t* file_idx, uint64_t n) {
  Handle* h = (Handle*)hp;
  if (n != h->p->rec.n()) return -1;
  h->p->rec.file.assign(file_idx, file_idx + n);
  return 0;
}
int ref_set_names(void* hp, const char* blob, const uint64_t* name_off, uint64_t n) {
  Handle* h = (Handle*)hp;
  h->p->rec.names.clear();
  for (uint64_t i = 0; i < n; i++) h->p->rec.names.emplace_back(blob + name_off[i], blob + name_off[i + 1]);
  return 0;
}

int ref_p1(void* hp, uint64_t* hist_out) {
  Handle* h = (Handle*)hp;
  double t = now_s();
  h->p->p1();
  h->p->times.p1 = now_s() - t;
  if (hist_out) std::memcpy(hist_out, h->p->hist.data(), NUM_HISTOGRAM_BUCKETS * 8);
  return 0;
}
int ref_p2(void* hp, uint16_t* b2b_out, uint32_t* nbins_out) {
  Handle* h = (Handle*)hp;
  double t = now_s();
  h->p->p2_partition();
  h->p->p2_route();
  h->p->times.p2 = now_s() - t;
  if (b2b_out) std::memcpy(b2b_out, h->p->bucket_to_bin.data(), NUM_HISTOGRAM_BUCKETS * 2);
  if (nbins_out) *nbins_out = (uint32_t)h->p->num_bins;
  return 0;
}
// Size of bin b in the reference bin-file format; copies it when out != NULL.
uint64_t ref_bin_bytes(void* hp, uint32_t b, uint8_t* out) {
  Handle* h = (Handle*)hp;
  if (b >= h->p->bins.size()) return 0;
  const auto& d = h->p->bins[b];
  if (out) std::memcpy(out, d.data(), d.size());
  return d.size();
}
// All routed records of every bin in a fixed-width form for bulk
// comparison (test infrastructure): per record its bin, its base count L and
// `rw` u64 words holding the bases MSB-first from bit 63 of word 0 (base t at
// bits 63-2(t%32) .. 62-2(t%32) of word t/32), zero past base L. Parses the
// reference byte records (BinReader, superkmer.rs:276-288). Returns the
// record count; outputs may be NULL.
uint64_t ref_bin_records(void* hp, uint32_t rw, uint64_t* out_bin, uint64_t* out_len, uint64_t* out_words) {
  Handle* h = (Handle*)hp;
  uint64_t n = 0;
  for (size_t b = 0; b < h->p->bins.size(); b++) {
    const std::vector<uint8_t>& d = h->p->bins[b];
    size_t p = 0;
    while (p + 2 <= d.size()) {
      const size_t bc = (size_t)d[p] | ((size_t)d[p + 1] << 8);
      p += 2;
      if (out_bin) {
        out_bin[n] = b;
        out_len[n] = bc;
        uint64_t* w = out_words + n * rw;
        for (uint32_t q = 0; q < rw; q++) w[q] = 0;
        for (size_t t = 0; t < bc && t < 32 * (size_t)rw; t++) {
          const uint64_t code = (d[p + t / 4] >> (6 - 2 * (t % 4))) & 3u;
          w[t / 32] |= code << (62 - 2 * (t % 32));
        }
      }
      p += (bc + 3) / 4;
      n++;
    }
  }
  return n;
}
int ref_p3(void* hp, uint64_t* d_out) {
  Handle* h = (Handle*)hp;
  double t = now_s();
  h->p->p3a();
  double t2 = now_s();
  h->p->p3b();
  h->p->times.p3a = t2 - t;
  h->p->times.p3b = now_s() - t2;
  if (d_out) *d_out = h->p->keys_lo.size();
  return 0;
}
// Consolidated keys: lo words (and hi words for K > 32).
uint64_t ref_keys(void* hp, uint64_t* lo, uint64_t* hi) {
  Handle* h = (Handle*)hp;
  size_t n = h->p->keys_lo.size();
  if (lo) std::memcpy(lo, h->p->keys_lo.data(), n * 8);
  if (hi) std::memcpy(hi, h->p->keys_hi.data(), n * 8);
  return n;
}
uint64_t ref_per_bin_distinct(void* hp, uint64_t* out) {
  Handle* h = (Handle*)hp;
  size_t n = h->p->per_bin_distinct.size();
  if (out) std::memcpy(out, h->p->per_bin_distinct.data(), n * 8);
  return n;
}
// MPHF export (same layout as the device export): 16 header words, the
// bucket->bin table as 65536 u16, key offsets, block offsets, blocks.
static uint64_t export_mphf(const Mphf& b, uint64_t D, uint64_t bins, uint64_t* out) {
  uint64_t ns = b.key_off.empty() ? 0 : b.key_off.size() - 1;
  uint64_t total = 16 + 65536 / 4 + 2 * (ns + 1) + b.blocks.size();
  if (out) {
    std::memset(out, 0, 16 * 8);
    out[0] = 0x6366316D70686631ULL;
    out[1] = (uint64_t)b.K;
    out[2] = (uint64_t)b.m;
    out[3] = (uint64_t)b.sb;
    out[4] = bins;
    out[5] = ns;
    out[6] = D;
    out[7] = b.blocks.size() / 8;
    out[8] = b.wide ? 1 : 0;
    std::memcpy(out + 16, b.b2b.data(), 65536 * 2);
    std::memcpy(out + 16 + 16384, b.key_off.data(), (ns + 1) * 8);
    std::memcpy(out + 16 + 16384 + ns + 1, b.blk_off.data(), (ns + 1) * 8);
    std::memcpy(out + 16 + 16384 + 2 * (ns + 1), b.blocks.data(), b.blocks.size() * 8);
  }
  return total;
}
uint64_t ref_mphf_export(void* hp, uint64_t* out) {
  Handle* h = (Handle*)hp;
  return export_mphf(h->p->mphf, h->p->keys_lo.size(), (uint64_t)h->p->num_bins, out);
}
int ref_set_target_seg(void* hp, uint64_t t) {
  ((Handle*)hp)->p->target_seg = t;
  return 0;
}
int ref_mphf_lookup(void* hp, const uint64_t* lo, const uint64_t* hi, uint64_t n, uint64_t* out) {
  Handle* h = (Handle*)hp;
  for (uint64_t i = 0; i < n; i++) out[i] = h->p->mphf.lookup(lo[i], hi ? hi[i] : 0);
  return 0;
}
int ref_p4(void* hp) {
  Handle* h = (Handle*)hp;
  double t = now_s();
  h->p->p4();
  h->p->times.p4 = now_s() - t;
  return 0;
}
uint64_t ref_states(void* hp, uint64_t* out) {
  Handle* h = (Handle*)hp;
  auto& sv = h->p->states;
  if (out)
    for (size_t i = 0; i < sv.num_words(); i++) out[i] = sv.data[i].load();
  return sv.num_words();
}
int ref_force_cas_failures(void* hp, int64_t n) {
  ((Handle*)hp)->p->states.forced_cas_failures.store(n);
  return 0;
}
int64_t ref_forced_cas_failures_left(void* hp) { return ((Handle*)hp)->p->states.forced_cas_failures.load(); }
// P5: writes prefix.cf_seg / .cf_seq / .json. Short records are reported in
// JSON when track_short (name, length) — classify.rs:550-561.
int ref_p5(void* hp, const char* prefix, const char* input_desc, int poly_n_stretch, int track_short) {
  Handle* h = (Handle*)hp;
  double t = now_s();
  h->p->p5(poly_n_stretch != 0);
  h->p->times.p5 = now_s() - t;
  h->p->short_seqs.clear();
  if (track_short)
    for (size_t r = 0; r < h->p->rec.n(); r++)
      if (h->p->rec.len(r) < (size_t)h->p->k)
        h->p->short_seqs.push_back({r < h->p->rec.names.size() ? h->p->rec.names[r] : std::string(), h->p->rec.len(r)});
  std::string p(prefix);
  FILE* f = std::fopen((p + ".cf_seg").c_str(), "wb");
  if (!f) return -1;
  std::fwrite(h->p->out_seg.data(), 1, h->p->out_seg.size(), f);
  std::fclose(f);
  f = std::fopen((p + ".cf_seq").c_str(), "wb");
  if (!f) return -1;
  std::fwrite(h->p->out_seq.data(), 1, h->p->out_seq.size(), f);
  std::fclose(f);
  h->json = make_json(input_desc ? input_desc : "", h->p->k, p, h->p->meta, h->p->short_seqs);
  f = std::fopen((p + ".json").c_str(), "wb");
  if (!f) return -1;
  std::fwrite(h->json.data(), 1, h->json.size(), f);
  std::fclose(f);
  return 0;
}
int ref_meta(void* hp, uint64_t* out5) {
  Handle* h = (Handle*)hp;
  const P5Meta& m = h->p->meta;
  out5[0] = m.unipath_count;
  out5[1] = m.kmer_count;
  out5[2] = m.max_len;
  out5[3] = m.min_len == UINT64_MAX ? 0 : m.min_len;
  out5[4] = m.sum_len;
  return 0;
}
int ref_times(void* hp, double* out6) {
  Handle* h = (Handle*)hp;
  const PhaseTimes& t = h->p->times;
  out6[0] = t.p1;
  out6[1] = t.p2;
  out6[2] = t.p3a;
  out6[3] = t.p3b;
  out6[4] = t.p4;
  out6[5] = t.p5;
  return 0;
}

// ---- primitives for known-answer and property tests
uint64_t ref_xxh3_u64(uint64_t v) { return xxh3_64_u64le(v); }
uint32_t ref_bucket_hash(uint64_t v) { return bucket_hash(v); }
uint32_t ref_mmer_order(uint32_t c) { return mmer_order(c); }
int ref_minimizer_len(int k) { return minimizer_len(k); }
int ref_partition(const uint64_t* hist, uint32_t num_bins, uint16_t* b2b, uint32_t* nbins) {
  std::vector<uint16_t> v;
  int nb = 0;
  partition_minimizers(hist, NUM_HISTOGRAM_BUCKETS, (int)num_bins, v, nb);
  std::memcpy(b2b, v.data(), NUM_HISTOGRAM_BUCKETS * 2);
  *nbins = (uint32_t)nb;
  return 0;
}
}  // extern "C"

// Kmer ops on an ASCII label of length k (k odd or even, 1..63).
template <int K>
static void kmer_ops_t(const char* s, int op, uint64_t* out, char* lab) {
  Kmer<K> km = Kmer<K>::from_ascii((const uint8_t*)s, 0);
  Kmer<K> r;
  switch (op) {
    case 0: r = km; break;
    case 1: r = km.reverse_complement(); break;
    case 2: r = km.canonical(); break;
    case 3: {  // roll by s[K]
      Kmer<K> rc = km.reverse_complement();
      km.roll_to_next_kmer(map_base((uint8_t)s[K]), rc);
      r = km;
      out[2] = (uint64_t)(u128)rc.bits;
      out[3] = (uint64_t)((u128)rc.bits >> 64);
      if (lab) {
        std::string l = rc.string_label();
        std::memcpy(lab + K + 1, l.data(), (size_t)K);
      }
      break;
    }
    case 4: out[2] = km.front(); out[3] = km.back(); r = km; break;
    default: r = km;
  }
  out[0] = (uint64_t)(u128)r.bits;
  out[1] = (uint64_t)((u128)r.bits >> 64);
  if (lab) {
    std::string l = r.string_label();
    std::memcpy(lab, l.data(), (size_t)K);
  }
}
extern "C" {
int ref_kmer_ops(int k, const char* s, int op, uint64_t* out4, char* label_out) {
  switch (k) {
#define CF1REF_KO(KK) \
  case KK:            \
    kmer_ops_t<KK>(s, op, out4, label_out); \
    return 0;
    CF1REF_KO(1) CF1REF_KO(3) CF1REF_KO(4) CF1REF_KO(5) CF1REF_KO(7) CF1REF_KO(9) CF1REF_KO(11) CF1REF_KO(13)
    CF1REF_KO(15) CF1REF_KO(17) CF1REF_KO(19) CF1REF_KO(21) CF1REF_KO(23) CF1REF_KO(25) CF1REF_KO(27)
    CF1REF_KO(29) CF1REF_KO(31) CF1REF_KO(33) CF1REF_KO(35) CF1REF_KO(37) CF1REF_KO(39) CF1REF_KO(41)
    CF1REF_KO(43) CF1REF_KO(45) CF1REF_KO(47) CF1REF_KO(49) CF1REF_KO(51) CF1REF_KO(53) CF1REF_KO(55)
    CF1REF_KO(57) CF1REF_KO(59) CF1REF_KO(61) CF1REF_KO(63)
#undef CF1REF_KO
    default:
      return -1;
  }
}
// from_packed_2bit at base offset i (semantic decode).
int ref_kmer_from_packed(int k, const uint8_t* packed, uint64_t off, uint64_t* out2) {
  u128 v = 0;
  for (int t = 0; t < k; t++) {
    uint64_t j = off + (uint64_t)t;
    v = (v << 2) | (u128)((packed[j / 4] >> (6 - 2 * (j % 4))) & 3);
  }
  out2[0] = (uint64_t)v;
  out2[1] = (uint64_t)(v >> 64);
  return 0;
}
// state.rs: from_vertex(class, front, back, visited, outputted); tables.
int ref_state_from_vertex(int cls, int front, int back, int visited, int outputted) {
  Vertex v{(StateClass)cls, (uint8_t)front, (uint8_t)back, visited != 0, outputted != 0};
  return state_from_vertex(v);
}
int ref_state_tables(uint8_t* cls32, uint8_t* out32, uint8_t* dead32, uint8_t* front32, uint8_t* back32) {
  for (int s = 0; s < 32; s++) {
    cls32[s] = (uint8_t)st_class((uint8_t)s);
    out32[s] = st_outputted((uint8_t)s);
    dead32[s] = st_is_dead_end((uint8_t)s) ? 1 : 0;
    front32[s] = st_decode((uint8_t)s).front;
    back32[s] = st_decode((uint8_t)s).back;
  }
  return 0;
}
// One DFA step as the reference's process_* functions compute it for a
// state `s` (classify.rs:147-518): type 0 isolated, 1 leftmost, 2 rightmost,
// 3 internal, 4 self loop with the leftmost or direct-repeat rule (MIMO),
// 5 self loop of an inverted repeat (handled as rightmost with prev). Bases
// are 2-bit codes; fwd = direction of the occurrence. Returns the new state.
int ref_dfa_step(int s, int type, int fwd, int pb, int nb) {
  AtomicStateVector sv(1);
  if (s) sv.compare_and_swap(0, 0, (uint8_t)s);
  // a one-key pipeline context is not needed: restate the branch bodies on sv
  struct Local {
    static uint8_t with_class(const Vertex& v, StateClass c) {
      Vertex x = v;
      x.state_class = c;
      return state_from_vertex(x);
    }
  };
  uint8_t st = sv.get(0);
  if (st_is_dead_end(st)) return st;
  if (type == 0 || type == 4) return state_from_vertex(v_mimo());
  if (type == 5) type = 2;
  uint8_t P = (uint8_t)pb, N = (uint8_t)nb, ns = st;
  if (type == 1) {  // process_leftmost_kmer :199-278
    if (fwd) {
      if (!st_is_visited(st)) ns = state_from_vertex(v_miso(N));
      else {
        Vertex v = st_decode(st);
        if (v.state_class == SISO) ns = v.back == N ? Local::with_class(v, MISO) : Local::with_class(v, MIMO);
        else if (v.state_class == MISO) ns = v.back != N ? Local::with_class(v, MIMO) : st;
        else ns = Local::with_class(v, MIMO);
      }
    } else {
      uint8_t cn = complement_base(N);
      if (!st_is_visited(st)) ns = state_from_vertex(v_simo(cn));
      else {
        Vertex v = st_decode(st);
        if (v.state_class == SISO) ns = v.front == cn ? Local::with_class(v, SIMO) : Local::with_class(v, MIMO);
        else if (v.state_class == MISO) ns = Local::with_class(v, MIMO);
        else ns = v.front != cn ? Local::with_class(v, MIMO) : st;
      }
    }
  } else if (type == 2) {  // process_rightmost_kmer :297-376
    if (fwd) {
      if (!st_is_visited(st)) ns = state_from_vertex(v_simo(P));
      else {
        Vertex v = st_decode(st);
        if (v.state_class == SISO) ns = v.front == P ? Local::with_class(v, SIMO) : Local::with_class(v, MIMO);
        else if (v.state_class == MISO) ns = Local::with_class(v, MIMO);
        else ns = v.front != P ? Local::with_class(v, MIMO) : st;
      }
    } else {
      uint8_t cp = complement_base(P);
      if (!st_is_visited(st)) ns = state_from_vertex(v_miso(cp));
      else {
        Vertex v = st_decode(st);
        if (v.state_class == SISO) ns = v.back == cp ? Local::with_class(v, MISO) : Local::with_class(v, MIMO);
        else if (v.state_class == MISO) ns = v.back != cp ? Local::with_class(v, MIMO) : st;
        else ns = Local::with_class(v, MIMO);
      }
    }
  } else {  // process_internal_kmer :406-511
    uint8_t a = fwd ? P : complement_base(N), c = fwd ? N : complement_base(P);
    if (!st_is_visited(st)) ns = state_from_vertex(v_siso(a, c));
    else {
      Vertex v = st_decode(st);
      if (v.state_class == SISO) {
        if (v.front == a && v.back == c) ns = st;
        else if (v.front != a && v.back != c) ns = Local::with_class(v, MIMO);
        else if (v.front != a) ns = Local::with_class(v, MISO);
        else ns = Local::with_class(v, SIMO);
      } else if (v.state_class == MISO) ns = v.back != c ? Local::with_class(v, MIMO) : st;
      else ns = v.front != a ? Local::with_class(v, MIMO) : st;
    }
  }
  return ns;
}

// Stand-alone AtomicStateVector for the state_vector.rs unit tests.
void* ref_sv_new(uint64_t n) { return new AtomicStateVector(n); }
void ref_sv_free(void* p) { delete (AtomicStateVector*)p; }
int ref_sv_get(void* p, uint64_t i) { return ((AtomicStateVector*)p)->get(i); }
int ref_sv_cas(void* p, uint64_t i, int e, int n) { return ((AtomicStateVector*)p)->compare_and_swap(i, (uint8_t)e, (uint8_t)n) ? 1 : 0; }
uint64_t ref_sv_word(void* p, uint64_t w) { return ((AtomicStateVector*)p)->data[w].load(); }

}  // extern "C"
