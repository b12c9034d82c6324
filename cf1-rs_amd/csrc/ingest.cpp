// Host ingest: FASTA files parsed once, in parallel, and the flat 2-bit
// stream packed on the host (the layout of device.hpp), so that only the
// packed bases and the placeholder bits cross PCIe.
//
// The reference parses every input with needletail four times, once per
// phase (src/minimizer.rs:44-89, src/superkmer.rs:44-91,
// src/classify.rs:540-597, src/output.rs:166-206). A record is a '>' header
// line (its id is the text up to the first whitespace,
// src/classify.rs:552-557) followed by sequence lines; line breaks (LF or
// CRLF) are dropped and every other byte is kept, so that non-ACGT bytes stay
// placeholders (src/dna.rs:64-92). Files ending in .gz are inflated with zlib.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <immintrin.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "ingest.hpp"

namespace cf1g {
namespace {

void run_threads(int threads, const std::function<void(int)>& f) {
  std::vector<std::thread> th;
  for (int t = 1; t < threads; t++) th.emplace_back(f, t);
  f(0);
  for (auto& x : th) x.join();
}

// The bytes of one input file: mapped, or inflated for .gz.
struct FileBytes {
  const uint8_t* p = nullptr;
  size_t n = 0;
  void* map = nullptr;
  std::vector<uint8_t> buf;
  explicit FileBytes(const std::string& path) {
    const bool gz = path.size() > 3 && path.compare(path.size() - 3, 3, ".gz") == 0;
    if (gz) {
      gzFile f = gzopen(path.c_str(), "rb");
      if (!f) throw std::runtime_error("failed to open FASTA file " + path);
      gzbuffer(f, 1 << 20);
      size_t used = 0;
      buf.resize(1 << 24);
      for (;;) {
        if (buf.size() - used < (1u << 22)) buf.resize(buf.size() * 2);
        const int got = gzread(f, buf.data() + used, (unsigned)std::min<size_t>(buf.size() - used, 1u << 30));
        if (got < 0) {
          int e = 0;
          std::string msg = gzerror(f, &e);
          gzclose(f);
          throw std::runtime_error("failed to inflate " + path + ": " + msg);
        }
        if (got == 0) break;
        used += (size_t)got;
      }
      gzclose(f);
      buf.resize(used);
      p = buf.data();
      n = used;
      return;
    }
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("failed to open FASTA file " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) {
      close(fd);
      throw std::runtime_error("failed to stat " + path);
    }
    n = (size_t)st.st_size;
    if (n) {
      map = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
      if (map == MAP_FAILED) {
        close(fd);
        throw std::runtime_error("failed to map " + path);
      }
      p = (const uint8_t*)map;
    }
    close(fd);
  }
  ~FileBytes() {
    if (map) munmap(map, n);
  }
};

struct Rec {
  size_t name_b, name_e;  // id bytes in the file
  size_t seq_b, seq_e;    // sequence lines (line breaks included)
};

// Header positions ('>' at a line start) of the file, found by `threads`
// scanners over byte ranges.
std::vector<size_t> headers(const FileBytes& F, int threads) {
  const size_t n = F.n;
  std::vector<std::vector<size_t>> part(threads);
  run_threads(threads, [&](int t) {
    const size_t a = n * (size_t)t / threads, b = n * (size_t)(t + 1) / threads;
    for (size_t i = a; i < b;) {
      const void* q = memchr(F.p + i, '>', b - i);
      if (!q) break;
      const size_t j = (const uint8_t*)q - F.p;
      if (j == 0 || F.p[j - 1] == '\n') part[t].push_back(j);
      i = j + 1;
    }
  });
  std::vector<size_t> out;
  for (auto& v : part) out.insert(out.end(), v.begin(), v.end());
  return out;
}

// Sequence bytes of [b, e) without line breaks.
size_t strip_count(const uint8_t* p, size_t b, size_t e) {
  size_t c = 0;
  for (size_t i = b; i < e; i++) c += (p[i] != '\n' && p[i] != '\r');
  return c;
}
// Copies exactly the bytes strip_count counts: runs between LF / CR bytes.
size_t strip_copy(const uint8_t* p, size_t b, size_t e, uint8_t* out) {
  size_t c = 0;
  for (size_t i = b; i < e;) {
    const void* q = memchr(p + i, '\n', e - i);
    const size_t j = q ? (size_t)((const uint8_t*)q - p) : e;
    for (size_t s = i; s < j;) {  // [s, j) may hold CRs (normally one, last)
      const void* cr = memchr(p + s, '\r', j - s);
      const size_t t = cr ? (size_t)((const uint8_t*)cr - p) : j;
      std::memcpy(out + c, p + s, t - s);
      c += t - s;
      s = t + 1;
    }
    i = j + 1;
  }
  return c;
}

}  // namespace

void parse_fasta(const std::vector<std::string>& paths, int threads, HostRecords& R) {
  threads = std::max(1, threads);
  R.ascii.clear();
  R.off.assign(1, 0);
  R.names.clear();
  R.name_off.assign(1, 0);
  R.file.clear();
  for (size_t file_idx = 0; file_idx < paths.size(); file_idx++) {
    const std::string& path = paths[file_idx];
    FileBytes F(path);
    const uint8_t* p = F.p;
    std::vector<size_t> hs = headers(F, threads);
    std::vector<Rec> recs(hs.size());
    for (size_t r = 0; r < hs.size(); r++) {
      const size_t h = hs[r];
      const void* nl = memchr(p + h, '\n', F.n - h);
      const size_t le = nl ? (size_t)((const uint8_t*)nl - p) : F.n;
      size_t ne = h + 1;
      while (ne < le && p[ne] != ' ' && p[ne] != '\t' && p[ne] != '\r') ne++;
      recs[r] = Rec{h + 1, ne, std::min(le + 1, F.n), r + 1 < hs.size() ? hs[r + 1] : F.n};
    }
    // pieces of ~16 MB of input, each starting at a line start and inside
    // one record, so that one chromosome spreads over every thread: bytes
    // counted per piece in parallel, offsets by a prefix sum, then copied
    struct Piece {
      size_t r, b, e, n, out;
    };
    std::vector<Piece> pieces;
    for (size_t r = 0; r < recs.size(); r++) {
      size_t b = recs[r].seq_b;
      const size_t e = recs[r].seq_e;
      while (b < e) {
        size_t m = std::min(e, b + (16u << 20));
        if (m < e) {
          const void* q = memchr(p + m, '\n', e - m);
          m = q ? (size_t)((const uint8_t*)q - p) + 1 : e;
        }
        pieces.push_back({r, b, m, 0, 0});
        b = m;
      }
    }
    run_threads(threads, [&](int t) {
      for (size_t i = t; i < pieces.size(); i += threads) pieces[i].n = strip_count(p, pieces[i].b, pieces[i].e);
    });
    const size_t base = R.ascii.size();
    std::vector<size_t> at(recs.size() + 1, base);
    {
      size_t out = base, pi = 0;
      for (size_t r = 0; r < recs.size(); r++) {
        at[r] = out;
        for (; pi < pieces.size() && pieces[pi].r == r; pi++) {
          pieces[pi].out = out;
          out += pieces[pi].n;
        }
      }
      at[recs.size()] = out;
    }
    R.ascii.resize(at[recs.size()]);
    run_threads(threads, [&](int t) {
      for (size_t i = t; i < pieces.size(); i += threads) strip_copy(p, pieces[i].b, pieces[i].e, R.ascii.data() + pieces[i].out);
    });
    for (size_t r = 0; r < recs.size(); r++) {
      R.off.push_back(at[r + 1]);
      R.names.append((const char*)p + recs[r].name_b, recs[r].name_e - recs[r].name_b);
      R.name_off.push_back(R.names.size());
      R.file.push_back((uint32_t)file_idx);
    }
  }
}

// 2-bit code of a byte: A/a 0, C/c 1, G/g 2, T/t 3, anything else 4.
static uint8_t g_code[256];
static bool g_code_init = [] {
  for (int i = 0; i < 256; i++) g_code[i] = 4;
  g_code['A'] = g_code['a'] = 0;
  g_code['C'] = g_code['c'] = 1;
  g_code['G'] = g_code['g'] = 2;
  g_code['T'] = g_code['t'] = 3;
  return true;
}();

namespace {
// 64 bases of one record (a[0..63]) -> the two sequence words (32 bases
// each, base 0 in bits 63..62) and the placeholder word (base 0 in bit 63).
void pack64_table(const uint8_t* a, uint64_t& hw, uint64_t& lw, uint64_t& nm) {
  hw = lw = nm = 0;
  for (int i = 0; i < 32; i++) {
    const uint64_t c0 = g_code[a[i]], c1 = g_code[a[i + 32]];
    hw |= (c0 & 3u) << (62 - 2 * i);
    lw |= (c1 & 3u) << (62 - 2 * i);
    nm |= ((c0 >> 2) << (63 - i)) | ((c1 >> 2) << (31 - i));
  }
}
// The same 32 bases at a time with AVX2. ((x >> 1) ^ (x >> 2)) & 3 is the
// code of an ACGT byte of either case; the byte is ACGT iff its upper-cased
// value equals the letter of that code (a byte shuffle of "ACGT"). Codes of
// other bytes are cleared; pairs of multiply-adds with weights 64, 16, 4, 1
// put four codes in a byte, MSB-first, and one byte shuffle gathers them.
__attribute__((target("avx2"))) void pack32_avx2(const uint8_t* a, uint64_t& w, uint32_t& nbits) {
  const __m256i v = _mm256_loadu_si256((const __m256i*)a);
  const __m256i m3 = _mm256_set1_epi8(3);
  const __m256i c0 = _mm256_and_si256(_mm256_xor_si256(_mm256_srli_epi16(v, 1), _mm256_srli_epi16(v, 2)), m3);
  const __m256i lut = _mm256_setr_epi8('A', 'C', 'G', 'T', 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 'A', 'C', 'G', 'T', 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                       0, 0, 0);
  const __m256i ok = _mm256_cmpeq_epi8(_mm256_and_si256(v, _mm256_set1_epi8((char)0xDF)), _mm256_shuffle_epi8(lut, c0));
  const __m256i c = _mm256_and_si256(c0, ok);
  const __m256i p16 = _mm256_maddubs_epi16(c, _mm256_set1_epi32(0x01041040));  // bytes 64, 16, 4, 1
  const __m256i p32 = _mm256_madd_epi16(p16, _mm256_set1_epi16(1));
  const __m256i g = _mm256_shuffle_epi8(p32, _mm256_setr_epi8(0, 4, 8, 12, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 4, 8, 12,
                                                              -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1));
  const uint64_t bytes = (uint64_t)(uint32_t)_mm256_extract_epi32(g, 0) | ((uint64_t)(uint32_t)_mm256_extract_epi32(g, 4) << 32);
  w = __builtin_bswap64(bytes);
  // placeholder bits MSB-first: the flags in reverse byte order, then their signs
  const __m256i rev = _mm256_setr_epi8(15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3,
                                       2, 1, 0);
  nbits = ~(uint32_t)_mm256_movemask_epi8(_mm256_permute4x64_epi64(_mm256_shuffle_epi8(ok, rev), 0x4E));
}
__attribute__((target("avx2"))) void pack64_avx2(const uint8_t* a, uint64_t& hw, uint64_t& lw, uint64_t& nm) {
  uint32_t n0, n1;
  pack32_avx2(a, hw, n0);
  pack32_avx2(a + 32, lw, n1);
  nm = ((uint64_t)n0 << 32) | n1;
}
const bool g_avx2 = __builtin_cpu_supports("avx2");
}  // namespace

void pack_words(const uint8_t* ascii, const uint64_t* rec_off, const uint64_t* gstart, uint64_t n_rec, uint64_t w0, uint64_t w1,
                uint64_t* seq, uint64_t* nmask) {
  (void)g_code_init;
  if (w0 >= w1) return;
  auto pack64 = g_avx2 ? pack64_avx2 : pack64_table;
  // the record whose range [gstart[r], gstart[r+1]) holds the position, or none
  uint64_t r = (uint64_t)(std::upper_bound(gstart, gstart + n_rec + 1, w0 * 64) - gstart);
  r = r ? r - 1 : n_rec;
  if (w0 * 64 < gstart[0]) r = 0;
  for (uint64_t w = w0; w < w1; w++) {
    uint64_t hi_w = 0, lo_w = 0, nm = 0;
    const uint64_t g0 = w * 64;
    while (r < n_rec && g0 >= gstart[r + 1]) r++;
    uint64_t* sq = seq + 2 * (w - w0);
    if (r < n_rec && g0 >= gstart[r] && g0 + 64 < gstart[r + 1]) {
      // all 64 positions are bases of record r: straight from the bytes
      pack64(ascii + (rec_off[r] - rec_off[0]) + (g0 - gstart[r]), hi_w, lo_w, nm);
    } else {
      for (int i = 0; i < 64; i++) {
        const uint64_t g = g0 + (uint64_t)i;
        while (r < n_rec && g >= gstart[r + 1]) r++;
        uint32_t code = 4;
        if (r < n_rec && g >= gstart[r] && g + 1 < gstart[r + 1]) code = g_code[ascii[rec_off[r] - rec_off[0] + (g - gstart[r])]];
        const uint64_t b = code & 3u;
        if (i < 32) hi_w |= b << (62 - 2 * i);
        else lo_w |= b << (62 - 2 * (i - 32));
        nm |= (uint64_t)(code >> 2) << (63 - i);
      }
    }
    // the packed words are written once and read by the copy engine (or the
    // next copy): non-temporal stores skip each line's read for ownership,
    // a third of the packing's host memory traffic
    _mm_stream_si64((long long*)&sq[0], (long long)hi_w);
    _mm_stream_si64((long long*)&sq[1], (long long)lo_w);
    _mm_stream_si64((long long*)&nmask[w - w0], (long long)nm);
  }
  _mm_sfence();  // the stores are complete before the caller hands the words on
}

void pack_stream(const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint64_t n_words64, uint64_t* seq, uint64_t* nmask,
                 int threads) {
  threads = std::max(1, threads);
  // gstart[r] = off[r] - off[0] + r + 1: stream position of record r's first
  // base; position gstart[r+1] - 1 is the separator after record r
  std::vector<uint64_t> gstart(n_rec + 1);
  for (uint64_t r = 0; r <= n_rec; r++) gstart[r] = rec_off[r] - rec_off[0] + r + 1;
  run_threads(threads, [&](int t) {
    const uint64_t w0 = n_words64 * (uint64_t)t / threads, w1 = n_words64 * (uint64_t)(t + 1) / threads;
    pack_words(ascii, rec_off, gstart.data(), n_rec, w0, w1, seq + 2 * w0, nmask + w0);
  });
}

// ---------------------------------------------------------------- host pool
// Each run() is a job of its own (shared with the workers that take part):
// a worker still leaving an earlier job only sees that job's exhausted
// counter.
struct HostPool::Job {
  std::function<void(uint64_t)> fn;  // a copy: the caller's may be a temporary
  uint64_t n;
  std::atomic<uint64_t> next{0}, done{0};
};
HostPool::HostPool(int threads) {
  for (int t = 1; t < threads; t++) th_.emplace_back([this] { worker(); });
}
HostPool::~HostPool() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& x : th_) x.join();
}
void HostPool::drain(Job& j) {
  for (;;) {
    const uint64_t i = j.next.fetch_add(1);
    if (i >= j.n) break;
    j.fn(i);
    if (j.done.fetch_add(1) + 1 == j.n) {
      std::lock_guard<std::mutex> g(mu_);
      cv_done_.notify_all();
    }
  }
}
// Workers spin briefly on the job counter before sleeping: a pipeline that
// hands out a job every ~0.1 ms (the streaming load's chunks) would
// otherwise pay a condition-variable wake-up per job and thread.
void HostPool::worker() {
  uint64_t seen = 0;
  for (;;) {
    for (int spin = 0; spin < 200000 && gen_.load(std::memory_order_acquire) == seen; spin++) _mm_pause();
    std::shared_ptr<Job> j;
    {
      std::unique_lock<std::mutex> g(mu_);
      cv_.wait(g, [&] { return stop_ || gen_.load() != seen; });
      if (stop_) return;
      seen = gen_.load();
      j = job_;
    }
    if (j) drain(*j);
  }
}
void HostPool::start(uint64_t n, const std::function<void(uint64_t)>& fn) {
  auto j = std::make_shared<Job>();
  j->fn = fn;
  j->n = n;
  {
    std::lock_guard<std::mutex> g(mu_);
    job_ = j;
    gen_++;
  }
  cv_.notify_all();
}
void HostPool::wait() {
  std::unique_lock<std::mutex> g(mu_);
  if (!job_) return;
  std::shared_ptr<Job> j = job_;
  cv_done_.wait(g, [&] { return j->done.load() == j->n; });
  job_.reset();
}
void HostPool::run(uint64_t n, const std::function<void(uint64_t)>& fn) {
  if (n == 0) return;
  start(n, fn);
  std::shared_ptr<Job> j;
  {
    std::lock_guard<std::mutex> g(mu_);
    j = job_;
  }
  drain(*j);
  wait();
}

}  // namespace cf1g
