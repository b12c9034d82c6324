// On-disk formats of the reference's intermediate files, written from the
// products of the GPU phases:
//   cf1_bin_%05d.bin   routed super k-mers of one bin (src/superkmer.rs:181-226):
//                      records of a u16 LE base count and ceil(n/4) bytes,
//                      base j at bits 6-2(j%4) of byte j/4 (A0 C1 G2 T3);
//   cf1_dedup_all.bin  the distinct canonical keys, each bin sorted, bins in
//                      order, as raw native-endian k-mer storage: u64, or
//                      u128 (low word first) for k > 32 (src/mphf.rs:120-132,
//                      :340-360).
#include <algorithm>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "formats.hpp"

namespace cf1g {
namespace {

void par(int threads, size_t n, const std::function<void(size_t)>& f) {
  threads = std::max(1, std::min<int>(threads, (int)std::max<size_t>(n, 1)));
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++)
    th.emplace_back([&, t] {
      for (size_t i = (size_t)t; i < n; i += (size_t)threads) f(i);
    });
  for (auto& x : th) x.join();
}

struct File {
  FILE* f;
  explicit File(const std::string& path) : f(std::fopen(path.c_str(), "wb")) {
    if (!f) throw std::runtime_error("cannot create " + path);
  }
  void write(const void* p, size_t n, const std::string& path) {
    if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("write failed: " + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
};

}  // namespace

std::string bin_file_path(const std::string& dir, uint32_t bin) {
  char name[32];
  std::snprintf(name, sizeof name, "cf1_bin_%05u.bin", bin);
  return dir + "/" + name;
}

uint64_t write_bin_files(const std::string& dir, uint32_t num_bins, int rw, const uint64_t* bin_start, const uint64_t* recs, int threads) {
  std::vector<uint64_t> bytes(num_bins, 0);
  std::vector<std::string> errs(num_bins);
  par(threads, num_bins, [&](size_t b) {
    try {
      const std::string path = bin_file_path(dir, (uint32_t)b);
      File out(path);
      std::vector<uint8_t> buf;
      buf.reserve(1 << 20);
      for (uint64_t r = bin_start[b]; r < bin_start[b + 1]; r++) {
        const uint64_t* w = recs + r * (uint64_t)rw;
        const uint32_t L = (uint32_t)(w[rw - 1] & 255);
        buf.push_back((uint8_t)(L & 255));
        buf.push_back((uint8_t)(L >> 8));
        const size_t at = buf.size();
        buf.resize(at + (L + 3) / 4, 0);
        // device record position t (1..L are the super k-mer's bases) at
        // bits 62-2(t%32) of word t/32
        for (uint32_t j = 0; j < L; j++) {
          const uint32_t t = j + 1;
          const uint32_t base = (uint32_t)(w[t >> 5] >> (62 - 2 * (t & 31))) & 3u;
          buf[at + j / 4] |= (uint8_t)(base << (6 - 2 * (j % 4)));
        }
        if (buf.size() >= (1u << 20)) {
          out.write(buf.data(), buf.size(), path);
          bytes[b] += buf.size();
          buf.clear();
        }
      }
      out.write(buf.data(), buf.size(), path);
      bytes[b] += buf.size();
    } catch (const std::exception& e) {
      errs[b] = e.what();
    }
  });
  for (auto& e : errs)
    if (!e.empty()) throw std::runtime_error(e);
  uint64_t tot = 0;
  for (uint64_t x : bytes) tot += x;
  return tot;
}

void write_dedup_file(const std::string& path, uint64_t* keys, uint64_t D, bool wide, const std::vector<uint64_t>& bin_key_off,
                      int threads) {
  const size_t nb = bin_key_off.size() - 1;
  if (!wide) {
    par(threads, nb, [&](size_t b) { std::sort(keys + bin_key_off[b], keys + bin_key_off[b + 1]); });
  } else {
    struct K128 {
      uint64_t lo, hi;
      bool operator<(const K128& o) const { return hi < o.hi || (hi == o.hi && lo < o.lo); }
    };
    K128* k = (K128*)keys;
    par(threads, nb, [&](size_t b) { std::sort(k + bin_key_off[b], k + bin_key_off[b + 1]); });
  }
  File out(path);
  out.write(keys, D * (wide ? 16 : 8), path);
}

}  // namespace cf1g
