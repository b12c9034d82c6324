// Writers of the reference's intermediate file formats (formats.cpp).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace cf1g {

std::string bin_file_path(const std::string& dir, uint32_t bin);
// cf1_bin_%05d.bin for bins [0, num_bins) from fixed-width device records
// (rw words each, bins at bin_start); returns the bytes written. Throws
// std::runtime_error.
uint64_t write_bin_files(const std::string& dir, uint32_t num_bins, int rw, const uint64_t* bin_start, const uint64_t* recs, int threads);
// cf1_dedup_all.bin: keys (D u64, or D (lo, hi) pairs when wide), sorted in
// place within each bin [bin_key_off[b], bin_key_off[b+1]). Throws.
void write_dedup_file(const std::string& path, uint64_t* keys, uint64_t D, bool wide, const std::vector<uint64_t>& bin_key_off,
                      int threads);

}  // namespace cf1g
