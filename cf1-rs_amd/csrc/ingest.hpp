// Host ingest (ingest.cpp): FASTA parsing and host packing of the flat stream.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace cf1g {

// Allocator that leaves new elements uninitialised (the parser writes every
// byte), so that sizing the record buffer does not clear it first.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) {}
  template <typename U>
  void construct(U* p) {
    ::new ((void*)p) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};

// Records as the C ABI takes them: ascii[off[r] .. off[r+1]) is record r,
// names[name_off[r] .. name_off[r+1]) its id.
struct HostRecords {
  std::vector<uint8_t, NoInitAlloc<uint8_t>> ascii;
  std::vector<uint64_t> off{0};
  std::string names;
  std::vector<uint64_t> name_off{0};
  std::vector<uint32_t> file;  // file[r] = index of the input file record r came from
};

// Parse FASTA files (plain or .gz) in order into R; throws std::runtime_error.
void parse_fasta(const std::vector<std::string>& paths, int threads, HostRecords& R);

// The flat stream of device.hpp for records ascii/rec_off: n_words64 words of
// placeholder bits and 2 * n_words64 words of bases, position 0 and each
// record's end a separator, positions past the last record padding.
void pack_stream(const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint64_t n_words64, uint64_t* seq, uint64_t* nmask,
                 int threads);

}  // namespace cf1g
