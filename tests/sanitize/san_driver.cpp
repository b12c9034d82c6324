// Host code under AddressSanitizer / ThreadSanitizer (built by the Makefile
// beside this file, run by tests/test_sanitizers.py). No GPU: the product's
// host-side C++ (FASTA ingest and host packing, the host thread pool, the
// bin / dedup file writers, the P5 line formatter) and the CPU oracle are
// linked in directly and driven on synthetic records, each with one thread
// and with several, and the results of the two runs are compared. P5's
// formatter reads its unitig spans through four C-ABI calls; this driver
// answers them itself (synthetic spans over the loaded records), so the
// formatter runs without a device context.
//
// Usage: san_driver <scratch dir> <threads> <oracle bases>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <zlib.h>

#include <atomic>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../cf1-rs_amd/csrc/formats.hpp"
#include "../../cf1-rs_amd/csrc/ingest.hpp"
#include "../../include/cf1g.h"

extern "C" {
void* ref_create(int k, int num_bins, int threads);
void ref_destroy(void* hp);
int ref_load(void* hp, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec);
int ref_p1(void* hp, uint64_t* hist_out);
int ref_p2(void* hp, uint16_t* b2b_out, uint32_t* nbins_out);
int ref_p3(void* hp, uint64_t* d_out);
int ref_p4(void* hp);
uint64_t ref_states(void* hp, uint64_t* out);
int ref_p5(void* hp, const char* prefix, const char* input_desc, int poly_n_stretch, int track_short);
int ref_meta(void* hp, uint64_t* out5);
}

namespace {

int g_fail = 0;
#define CHECK(c, ...)                                   \
  do {                                                  \
    if (!(c)) {                                         \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);                     \
      fputc('\n', stderr);                              \
      g_fail = 1;                                       \
    }                                                   \
  } while (0)

std::string slurp(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  std::stringstream s;
  s << f.rdbuf();
  return s.str();
}

// Records of random bases with lowercase stretches, N runs and other
// placeholders, short records between them.
struct Recs {
  std::vector<std::string> seq, name;
};
Recs make_records(uint64_t total, uint64_t seed) {
  std::mt19937_64 rng(seed);
  Recs R;
  uint64_t got = 0;
  int i = 0;
  while (got < total) {
    uint64_t len = 1 + rng() % (total / 4 + 64);
    if (i % 5 == 3) len = 1 + rng() % 40;  // around and below k
    if (i == 0 && total > 4100000) len = 4005000;  // across the 4 Mbp chunk edge of P1 / P2
    if (got + len > total) len = total - got + 1;
    std::string s(len, 'A');
    for (auto& c : s) c = "ACGT"[rng() & 3];
    if (len > 500) {
      const uint64_t a = rng() % (len - 200);
      for (uint64_t t = a; t < a + 1 + rng() % 150; t++) s[t] = 'N';
      const uint64_t b = rng() % (len - 200);
      for (uint64_t t = b; t < b + 100; t++) s[t] = (char)(s[t] | 0x20);
      s[rng() % len] = 'X';
    }
    R.seq.push_back(s);
    R.name.push_back("rec" + std::to_string(i) + " desc");
    got += len;
    i++;
  }
  return R;
}

void flat(const Recs& R, std::vector<uint8_t>& ascii, std::vector<uint64_t>& off) {
  ascii.clear();
  off.assign(1, 0);
  for (auto& s : R.seq) {
    ascii.insert(ascii.end(), s.begin(), s.end());
    off.push_back(ascii.size());
  }
}

void write_fasta(const Recs& R, size_t a, size_t b, const std::string& path, bool gz) {
  std::string text;
  for (size_t r = a; r < b; r++) {
    text += ">" + R.name[r] + "\n";
    for (size_t p = 0; p < R.seq[r].size(); p += 61) text += R.seq[r].substr(p, 61) + (p % 3 ? "\r\n" : "\n");
    if (r % 4 == 1) text += "\n";
  }
  if (gz) {
    gzFile f = gzopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("gzopen " + path);
    gzwrite(f, text.data(), (unsigned)text.size());
    gzclose(f);
  } else {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("fopen " + path);
    fwrite(text.data(), 1, text.size(), f);
    fclose(f);
  }
}

// ---- FASTA ingest, host packing and the host pool
void check_ingest(const std::string& dir, int threads) {
  const Recs R = make_records(1500000, 7);
  const size_t h = R.seq.size() / 2;
  const std::vector<std::string> paths = {dir + "/a.fa", dir + "/b.fa.gz"};
  write_fasta(R, 0, h, paths[0], false);
  write_fasta(R, h, R.seq.size(), paths[1], true);
  cf1g::HostRecords one, many;
  cf1g::parse_fasta(paths, 1, one);
  cf1g::parse_fasta(paths, threads, many);
  std::vector<uint8_t> ascii;
  std::vector<uint64_t> off;
  flat(R, ascii, off);
  CHECK(one.off == off, "parse_fasta: record offsets");
  CHECK(one.ascii.size() == ascii.size() && memcmp(one.ascii.data(), ascii.data(), ascii.size()) == 0, "parse_fasta: bases");
  CHECK(many.off == one.off && many.names == one.names && many.name_off == one.name_off && many.file == one.file,
        "parse_fasta: %d threads differ from one", threads);
  CHECK(many.ascii.size() == one.ascii.size() && memcmp(many.ascii.data(), one.ascii.data(), one.ascii.size()) == 0,
        "parse_fasta: bases with %d threads", threads);
  CHECK(one.file.size() == R.seq.size() && one.file[0] == 0 && one.file.back() == 1, "parse_fasta: file indices");

  // stream: separator at 0 and after each record
  const uint64_t n_rec = off.size() - 1;
  std::vector<uint64_t> gstart(n_rec + 1);
  uint64_t pos = 1;
  for (uint64_t r = 0; r <= n_rec; r++) {
    gstart[r] = pos;
    if (r < n_rec) pos += off[r + 1] - off[r] + 1;
  }
  const uint64_t n_words64 = (pos + 63) / 64 + 1;
  std::vector<uint64_t> seq1(2 * n_words64), nm1(n_words64), seqN(2 * n_words64), nmN(n_words64);
  cf1g::pack_stream(ascii.data(), off.data(), n_rec, n_words64, seq1.data(), nm1.data(), 1);
  cf1g::pack_stream(ascii.data(), off.data(), n_rec, n_words64, seqN.data(), nmN.data(), threads);
  CHECK(seq1 == seqN && nm1 == nmN, "pack_stream: %d threads differ from one", threads);
  // the same words in pieces through the pool, two pieces in flight
  std::vector<uint64_t> seqP(2 * n_words64), nmP(n_words64);
  const uint64_t piece = 4096;
  const uint64_t n_pieces = (n_words64 + piece - 1) / piece;
  cf1g::HostPool pool(threads);
  std::atomic<uint64_t> calls{0};
  pool.run(n_pieces, [&](uint64_t i) {
    const uint64_t w0 = i * piece, w1 = std::min(n_words64, w0 + piece);
    cf1g::pack_words(ascii.data(), off.data(), gstart.data(), n_rec, w0, w1, seqP.data() + 2 * w0, nmP.data() + w0);
    calls++;
  });
  CHECK(calls.load() == n_pieces, "HostPool::run: %llu of %llu calls", (unsigned long long)calls.load(), (unsigned long long)n_pieces);
  CHECK(seqP == seq1 && nmP == nm1, "pack_words pieces differ from pack_stream");
  for (int rep = 0; rep < 20; rep++) {
    std::atomic<uint64_t> sum{0};
    auto job = [&](uint64_t i) { sum += i + 1; };
    pool.start(1000 + rep, job);
    pool.wait();
    const uint64_t n = 1000 + rep;
    CHECK(sum.load() == n * (n + 1) / 2, "HostPool::start/wait: sum");
  }
}

// ---- bin and dedup file writers
void check_formats(const std::string& dir, int threads) {
  std::mt19937_64 rng(11);
  const uint32_t num_bins = 37;
  for (int rw : {2, 4}) {
    std::vector<uint64_t> bin_start(num_bins + 1, 0);
    for (uint32_t b = 0; b < num_bins; b++) bin_start[b + 1] = bin_start[b] + (b % 7 == 3 ? 0 : rng() % 900);
    std::vector<uint64_t> recs(bin_start[num_bins] * rw);
    const int cap = rw == 2 ? 60 : 124;
    for (uint64_t r = 0; r < bin_start[num_bins]; r++) {
      for (int t = 0; t < rw; t++) recs[r * rw + t] = rng();
      recs[r * rw + rw - 1] = (recs[r * rw + rw - 1] & ~0x3FFULL) | (uint64_t)(1 + rng() % cap) | ((rng() & 3) << 8);
    }
    const std::string d1 = dir + "/bins1_" + std::to_string(rw), dN = dir + "/binsN_" + std::to_string(rw);
    mkdir(d1.c_str(), 0755);
    mkdir(dN.c_str(), 0755);
    const uint64_t b1 = cf1g::write_bin_files(d1, num_bins, rw, bin_start.data(), recs.data(), 1);
    const uint64_t bN = cf1g::write_bin_files(dN, num_bins, rw, bin_start.data(), recs.data(), threads);
    CHECK(b1 == bN && b1 > 0, "write_bin_files: byte counts");
    for (uint32_t b = 0; b < num_bins; b++)
      CHECK(slurp(cf1g::bin_file_path(d1, b)) == slurp(cf1g::bin_file_path(dN, b)), "write_bin_files: bin %u", b);
  }
  for (bool wide : {false, true}) {
    const uint64_t D = 50000;
    std::vector<uint64_t> bko = {0, 17, 17, 20000, 31000, D};
    std::vector<uint64_t> k1(D * (wide ? 2 : 1));
    for (auto& x : k1) x = rng();
    std::vector<uint64_t> kN = k1;
    cf1g::write_dedup_file(dir + "/dd1.bin", k1.data(), D, wide, bko, 1);
    cf1g::write_dedup_file(dir + "/ddN.bin", kN.data(), D, wide, bko, threads);
    CHECK(k1 == kN && slurp(dir + "/dd1.bin") == slurp(dir + "/ddN.bin"), "write_dedup_file: %s keys", wide ? "wide" : "u64");
  }
}

// ---- the P5 formatter over synthetic spans (its four C-ABI reads are
// answered below from g_spans)
struct Spans {
  uint64_t k = 31, n_rec = 0;
  std::vector<uint64_t> S, E, ID, gstart;
  std::vector<uint8_t> FL;
  std::vector<uint32_t> file;
} g_spans;

void make_spans(const std::vector<uint64_t>& off, uint64_t k, uint64_t seed) {
  std::mt19937_64 rng(seed);
  Spans& P = g_spans;
  P = Spans();
  P.k = k;
  P.n_rec = off.size() - 1;
  uint64_t pos = 1, id = 0;
  for (uint64_t r = 0; r < P.n_rec; r++) {
    P.gstart.push_back(pos);
    P.file.push_back((uint32_t)(r * 3 / P.n_rec));
    const uint64_t len = off[r + 1] - off[r];
    if (len >= k) {
      uint64_t i = rng() % 3 == 0 ? rng() % 40 : 0;
      while (i + k <= len) {
        const uint64_t run = 1 + rng() % 200;
        const uint64_t last = std::min(len - k, i + run - 1);
        P.S.push_back(pos + i);
        P.E.push_back(pos + last);
        P.ID.push_back(rng() % 4 == 0 && id ? rng() % id : id++);
        P.FL.push_back((uint8_t)((rng() & 1) | (P.ID.back() + 1 == id ? 2 : 0)));
        i = last + 1 + (rng() % 5 == 0 ? k + rng() % 60 : 0);
      }
    }
    pos += len + 1;
  }
  P.gstart.push_back(pos);
}

void check_p5(const std::string& dir, int threads) {
  const Recs R = make_records(800000, 5);
  std::vector<uint8_t> ascii;
  std::vector<uint64_t> off;
  flat(R, ascii, off);
  std::string blob;
  std::vector<uint64_t> noff = {0};
  for (auto& n : R.name) {
    blob += n;
    noff.push_back(blob.size());
  }
  make_spans(off, 31, 3);
  const uint64_t n_rec = off.size() - 1;
  uint64_t m1[5] = {0}, mN[5] = {0};
  cf1g_ctx* ctx = (cf1g_ctx*)&g_spans;
  int rc1 = cf1g_p5_write(ctx, ascii.data(), off.data(), n_rec, blob.data(), noff.data(), (dir + "/p5one").c_str(), "in.fa", 1, 1, 1, m1);
  int rcN = cf1g_p5_write(ctx, ascii.data(), off.data(), n_rec, blob.data(), noff.data(), (dir + "/p5many").c_str(), "in.fa", 1, 1,
                          threads, mN);
  CHECK(rc1 == 0 && rcN == 0, "cf1g_p5_write: %d %d", rc1, rcN);
  CHECK(memcmp(m1, mN, sizeof m1) == 0 && m1[0] > 0, "cf1g_p5_write: meta");
  for (const char* ext : {".cf_seg", ".cf_seq"})
    CHECK(slurp(dir + "/p5one" + ext) == slurp(dir + "/p5many" + ext), "cf1g_p5_write: %s with %d threads", ext, threads);
}

// ---- the CPU oracle, P1..P5
struct OracleRun {
  uint64_t D = 0;
  std::vector<uint64_t> hist, states;
  uint64_t meta[5] = {0};
  std::string seg, seq;
};
OracleRun run_oracle(const std::vector<uint8_t>& ascii, const std::vector<uint64_t>& off, int k, int threads, const std::string& pre) {
  OracleRun o;
  void* h = ref_create(k, 16, threads);
  if (!h) throw std::runtime_error("ref_create");
  CHECK(ref_load(h, ascii.data(), off.data(), off.size() - 1) == 0, "ref_load");
  o.hist.assign((size_t)1 << 16, 0);
  CHECK(ref_p1(h, o.hist.data()) == 0, "ref_p1");
  std::vector<uint16_t> b2b((size_t)1 << 16);
  uint32_t nb = 0;
  CHECK(ref_p2(h, b2b.data(), &nb) == 0, "ref_p2");
  CHECK(ref_p3(h, &o.D) == 0, "ref_p3");
  CHECK(ref_p4(h) == 0, "ref_p4");
  o.states.assign(ref_states(h, nullptr), 0);
  ref_states(h, o.states.data());
  CHECK(ref_p5(h, pre.c_str(), "in.fa", 1, 1) == 0, "ref_p5");
  ref_meta(h, o.meta);
  o.seg = slurp(pre + ".cf_seg");
  o.seq = slurp(pre + ".cf_seq");
  ref_destroy(h);
  return o;
}

void check_oracle(const std::string& dir, int threads, uint64_t bases) {
  // one record past the 4 Mbp chunk length when asked for that many bases
  Recs R = make_records(bases, 9);
  std::vector<uint8_t> ascii;
  std::vector<uint64_t> off;
  flat(R, ascii, off);
  for (int k : {31, 63}) {
    const OracleRun a = run_oracle(ascii, off, k, 1, dir + "/or1");
    const OracleRun b = run_oracle(ascii, off, k, threads, dir + "/orN");
    CHECK(a.D > 0 && a.D == b.D, "oracle k=%d: D %llu vs %llu", k, (unsigned long long)a.D, (unsigned long long)b.D);
    CHECK(a.hist == b.hist, "oracle k=%d: P1 histogram", k);
    CHECK(a.states == b.states, "oracle k=%d: state words", k);
    CHECK(memcmp(a.meta, b.meta, sizeof a.meta) == 0, "oracle k=%d: P5 meta", k);
    CHECK(a.seg.size() == b.seg.size() && a.seq == b.seq, "oracle k=%d: P5 files", k);
  }
}

}  // namespace

// The C-ABI reads of the P5 formatter, answered from g_spans.
extern "C" int cf1g_get_info(cf1g_ctx* ctx, cf1g_info* out) {
  const Spans& P = *(const Spans*)ctx;
  memset(out, 0, sizeof *out);
  out->n_records = P.n_rec;
  out->k = (uint32_t)P.k;
  return 0;
}
extern "C" int cf1g_unitig_spans(cf1g_ctx* ctx, uint64_t* n_spans, uint64_t* start, uint64_t* end, uint64_t* id, uint8_t* flags,
                                 uint64_t cap) {
  const Spans& P = *(const Spans*)ctx;
  const uint64_t U = P.S.size();
  *n_spans = U;
  if (!start && !end && !id && !flags) return 0;
  if (cap < U) return CF1G_E_INVALID;
  for (uint64_t u = 0; u < U; u++) {
    start[u] = P.S[u];
    end[u] = P.E[u];
    id[u] = P.ID[u];
    flags[u] = P.FL[u];
  }
  return 0;
}
extern "C" int cf1g_record_starts(cf1g_ctx* ctx, uint64_t* out, uint64_t n) {
  const Spans& P = *(const Spans*)ctx;
  if (n < P.gstart.size()) return CF1G_E_INVALID;
  memcpy(out, P.gstart.data(), P.gstart.size() * 8);
  return 0;
}
extern "C" int cf1g_record_files(cf1g_ctx* ctx, uint32_t* out, uint64_t n) {
  const Spans& P = *(const Spans*)ctx;
  if (n < P.n_rec) return CF1G_E_INVALID;
  memcpy(out, P.file.data(), P.n_rec * 4);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s <scratch dir> <threads> <oracle bases>\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const int threads = atoi(argv[2]);
  const uint64_t bases = strtoull(argv[3], nullptr, 10);
  try {
    check_ingest(dir, threads);
    printf("ingest ok\n");
    check_formats(dir, threads);
    printf("formats ok\n");
    check_p5(dir, threads);
    printf("p5 ok\n");
    check_oracle(dir, threads, bases);
    printf("oracle ok\n");
  } catch (const std::exception& e) {
    fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  fflush(stdout);
  return g_fail;
}
