// Host ingest (ingest.cpp): FASTA parsing and host packing of the flat stream.
#pragma once
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
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
// Words [w0, w1) of that stream into seq[0 .. 2(w1-w0)) and nmask[0 .. w1-w0);
// gstart[r] = stream position of record r's first base (n_rec + 1 entries).
void pack_words(const uint8_t* ascii, const uint64_t* rec_off, const uint64_t* gstart, uint64_t n_rec, uint64_t w0, uint64_t w1,
                uint64_t* seq, uint64_t* nmask);

// A fixed set of host threads: run(n, fn) calls fn(0 .. n-1) on them and on
// the calling thread, and returns when every call has returned; start(n, fn)
// hands the calls to the pool's threads alone (on a copy of fn) and returns
// at once; wait() returns when they have all returned. Calls are taken in
// index order.
class HostPool {
 public:
  explicit HostPool(int threads);
  ~HostPool();
  void run(uint64_t n, const std::function<void(uint64_t)>& fn);
  void start(uint64_t n, const std::function<void(uint64_t)>& fn);
  void wait();
  int threads() const { return (int)th_.size() + 1; }

 private:
  struct Job;
  void worker();
  void drain(Job& j);
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::shared_ptr<Job> job_;
  std::atomic<uint64_t> gen_{0};
  bool stop_ = false;
};

}  // namespace cf1g
