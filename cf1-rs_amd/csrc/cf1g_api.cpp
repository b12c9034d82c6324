// cf1g context, phase orchestration and the extern "C" boundary.
// Mirrors the phase sequence of run_pipeline (src/pipeline.rs:66-118); all
// intermediates (packed stream, bins, keys, MPHF, states) stay in HBM.
#include <hip/hip_runtime.h>
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cf1g.h"
#include "device.hpp"
#include "formats.hpp"
#include "ingest.hpp"
#include "launch.hpp"
#include "lidx.hpp"

namespace cf1g {
// kernels_p1p2.hip
void launch_pack(hipStream_t, const uint8_t*, const uint64_t*, const uint64_t*, uint64_t, uint64_t, uint64_t*, uint64_t*);
uint64_t p1_tiles(uint64_t);
uint64_t p1_tile_slots();
uint64_t p1_slices(uint64_t);
void launch_p1_scan(hipStream_t, const uint64_t*, const uint64_t*, uint64_t, uint64_t, uint64_t, int, int, uint64_t*, uint32_t*,
                    uint32_t*, uint32_t*, const uint64_t*, uint32_t, uint32_t*, uint16_t*, uint32_t);
uint64_t p1_tile_reach(uint64_t);
void launch_p1_finish(hipStream_t, uint64_t, uint32_t*, uint32_t*, uint32_t*, uint32_t*, uint64_t*, unsigned long long*,
                      unsigned long long*, unsigned long long*, uint32_t, uint32_t*, uint16_t*, uint32_t, unsigned int*);
void launch_p1(hipStream_t, const uint64_t*, const uint64_t*, uint64_t, int, int, uint64_t*, uint32_t*, uint32_t*, uint32_t*, uint32_t*,
               uint64_t*, unsigned long long*, unsigned long long*, unsigned long long*, const uint64_t*, uint32_t, uint32_t*,
               uint16_t*, uint32_t, unsigned int*);
void p2_chunking(uint64_t, uint32_t, uint64_t*, uint64_t*);
uint64_t partition_scratch_words();
void launch_partition(hipStream_t, const unsigned long long*, const unsigned long long*, uint32_t, uint32_t, uint16_t*, uint64_t*,
                      uint64_t*, unsigned long long*, uint64_t*, uint64_t*, uint64_t*);
void launch_p2_route(hipStream_t, const uint64_t*, const uint64_t*, const uint64_t*, const uint32_t*, const uint32_t*, uint64_t, uint64_t,
                     uint64_t, int, const uint16_t*, int, const uint64_t*, unsigned long long*, uint64_t*, uint16_t*, uint32_t*, int,
                     unsigned int*, unsigned long long*);
// kernels_p3.hip
uint64_t scan_temp_words(uint64_t n);
void launch_zero(hipStream_t, void*, uint64_t);
void scan_excl_u64(hipStream_t, const uint64_t*, uint64_t*, uint64_t, uint64_t*, uint64_t*);
void launch_p3_rcount(hipStream_t, bool, const uint64_t*, const uint16_t*, const uint64_t*, const uint64_t*, int, uint32_t, uint64_t, int,
                      int, unsigned long long*, unsigned long long*);
void launch_p3_rscatter(hipStream_t, bool, const uint64_t*, const uint16_t*, const uint64_t*, const uint64_t*, int, uint32_t, uint64_t,
                        int, const uint64_t*, unsigned int*, uint64_t*, uint32_t*, const unsigned int*);
void launch_p5_annot(hipStream_t, bool, const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*,
                     const uint64_t*, const uint64_t*, int, uint64_t, uint8_t*, unsigned int*);
void launch_p5_list(hipStream_t, const uint64_t*, const uint32_t*, const uint32_t*, uint64_t, uint32_t, const uint32_t*, uint64_t,
                    uint64_t*, unsigned int*);
void launch_p3_expand(hipStream_t, bool, const uint64_t*, const uint64_t*, const uint64_t*, const uint32_t*, uint64_t, int, void*,
                      uint16_t*, const unsigned int*);
void launch_p3_gate(hipStream_t, const uint64_t*, uint64_t, unsigned int*);
void launch_p3_dedup(hipStream_t, bool, void*, uint16_t*, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t, int, int, uint64_t*,
                     uint64_t*, uint32_t*, unsigned int*, uint32_t*, uint64_t*, uint16_t*);
void launch_p3_dedup_list(hipStream_t, bool, void*, uint16_t*, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t, int, int,
                          uint64_t*, uint64_t*, const uint32_t*, uint32_t*, unsigned int*, uint64_t*, uint16_t*);
void launch_p3_bigsort(hipStream_t, bool, void*, uint16_t*, void*, uint16_t*, const uint64_t*, const uint32_t*, uint32_t, uint64_t*,
                       unsigned int*, uint32_t*);
void launch_bbh_big(hipStream_t, bool, const void*, const uint16_t*, uint16_t*, const uint64_t*, const uint64_t*, const uint32_t*,
                    uint32_t, uint64_t*, uint64_t*, uint64_t*, uint64_t*, void*, unsigned int*);
void launch_p3_compact(hipStream_t, bool, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t, uint64_t*);
void launch_p3_keys(hipStream_t, bool, const void*, const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*, const uint64_t*,
                    uint64_t, void*, unsigned int*);
void launch_p4_pack(hipStream_t, const uint16_t*, const uint64_t*, const uint64_t*, const uint64_t*, uint64_t, uint64_t, uint8_t*,
                    uint64_t*);
void launch_stream_digest(hipStream_t, const uint64_t*, const uint64_t*, uint64_t, unsigned long long*);
void noalloc_setup(int device);
void* noalloc_scratch(int device, size_t* bytes);
void launch_p4_cas(hipStream_t, bool, const uint64_t*, const uint64_t*, uint64_t, int, const MphfView&, uint64_t, uint64_t*, uint64_t*,
                   uint32_t, uint32_t, int);
uint64_t sort_tiles(uint64_t);
void launch_sort_hist(hipStream_t, const uint64_t*, uint64_t, int, uint64_t*);
void launch_sort_scatter(hipStream_t, const uint64_t*, uint64_t*, uint64_t, int, const uint64_t*);
void launch_lidx_stats_init(hipStream_t, int, const SegCtx&, uint64_t, const LidxTrainBufs&);
void launch_lidx_fwd_bwd(hipStream_t, int, const SegCtx&, uint64_t, const LidxTrainBufs&);
void launch_lidx_update(hipStream_t, uint64_t, float, const LidxTrainBufs&);
void launch_lidx_pieces(hipStream_t, const uint64_t*, uint64_t, int, uint32_t, uint32_t, uint64_t*, uint32_t*);
void launch_lidx_root_final(hipStream_t, uint32_t, uint64_t, const uint64_t*, const uint32_t*, const LidxTrainBufs&, RootPiece*);
void launch_lidx_leaf_bounds(hipStream_t, const uint64_t*, uint64_t, int, uint32_t, const RootPiece*, uint64_t*);
void launch_lidx_leaf_final(hipStream_t, const uint64_t*, uint64_t, uint32_t, int, const uint64_t*, const RootPiece*,
                            const LidxTrainBufs&, uint32_t, LeafRec*, uint32_t*, uint64_t*);
void launch_lidx_lines(hipStream_t, int, const uint64_t*, uint64_t, uint32_t, uint64_t, uint64_t*, uint32_t*);
void launch_lidx_warp_hist(hipStream_t, const uint64_t*, uint64_t, int, int, uint32_t*);
void launch_lidx_warp_keys(hipStream_t, const uint64_t*, uint64_t, const uint32_t*, int, int, uint64_t*);
void launch_lidx_bbh2rank(hipStream_t, const uint64_t*, uint32_t, const uint64_t*, const LeafRec*, const MphfView&, uint32_t*);
void launch_lidx_lookup(hipStream_t, const uint64_t*, uint64_t, const LidxView&, const MphfView&, uint64_t*, LidxQueue);
void lidx_queue_shape(uint64_t, uint32_t&, uint32_t&);
void launch_p4_cas_lidx(hipStream_t, const uint64_t*, const uint64_t*, uint64_t, int, const LidxView&, const MphfView&, uint64_t*,
                        uint64_t*, uint32_t, uint32_t, int);
void launch_annotate_lidx(hipStream_t, const uint64_t*, const uint64_t*, uint64_t, int, const LidxView&, const MphfView&,
                          const uint64_t*, uint8_t*);
void launch_annotate(hipStream_t, bool, const uint64_t*, const uint64_t*, uint64_t, int, const MphfView&, const uint64_t*, uint64_t,
                     uint8_t*, unsigned int*);
void launch_lookup(hipStream_t, bool, const void*, uint64_t, const MphfView&, uint64_t*, bool);

uint64_t span_tiles(uint64_t);
void launch_span_count(hipStream_t, const uint8_t*, uint64_t, uint64_t*, uint64_t*);
void launch_span_emit(hipStream_t, const uint8_t*, uint64_t, const uint64_t*, const uint64_t*, uint64_t*, uint64_t*);
void launch_span_keys(hipStream_t, bool, const uint64_t*, int, const uint64_t*, const uint64_t*, uint64_t, void*, uint8_t*);
void launch_span_claim(hipStream_t, const uint64_t*, uint64_t, uint64_t, uint64_t*, uint8_t*, unsigned int*);

// ------------------------------------------------------------------ profiling
namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_ev_pool;
hipEvent_t ev_get_locked() {
  if (!g_ev_pool.empty()) {
    hipEvent_t e = g_ev_pool.back();
    g_ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  return e;
}
thread_local hipEvent_t t_pending = nullptr;
}  // namespace

void ev_put(hipEvent_t e) {
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_ev_pool.push_back(e);
}

hipEvent_t ev_get() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  return ev_get_locked();
}

void prof_begin(const char* name, hipStream_t st) {
  if (!g_prof_on) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  t_pending = ev_get_locked();
  hip_check(hipEventRecord(t_pending, st), "hipEventRecord");
  (void)name;
}
void prof_end(const char* name, hipStream_t st) {
  if (!g_prof_on || !t_pending) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  hipEvent_t e = ev_get_locked();
  hip_check(hipEventRecord(e, st), "hipEventRecord");
  g_prof.push_back({name, t_pending, e});
  t_pending = nullptr;
}

}  // namespace cf1g

using namespace cf1g;

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// A captured launch sequence and the sizes / allocation generation it was
// captured with.
struct GraphSeg {
  std::vector<uint64_t> key;
  hipGraphExec_t exec = nullptr;
  size_t nodes = 0;
  bool valid = false;
};

struct cf1g_ctx {
  cf1g_params prm{};
  int k = 31, m = 15, w = 17;
  bool wide = false;
  int rw = 2;
  size_t key_bytes = 8;
  hipStream_t st = nullptr;
  int device = 0;
  int cu_count = 256;
  std::string err;
  int stage = 0;  // 1 loaded, 2 P1, 3 P2, 4 P3, 5 P4
  // state_gen changes whenever records, phase products, states or the
  // selected index change; the unitig spans are kept until it does (a count
  // query and the fill that follows it mark the spans once)
  uint64_t state_gen = 1, spans_gen = 0, spans_U = 0;
  void set_stage(int s) {
    stage = s;
    state_gen++;
  }
  std::map<std::string, DevBuf> bufs;

  // sizes
  uint64_t n_rec = 0, n_ascii = 0, n_pos = 0, n_words64 = 0;
  std::vector<uint64_t> gstart_h;
  // chunk edges of records of CHUNK_BASES or more (P1 / P2 scan them in
  // chunks, src/minimizer.rs:57-86, src/superkmer.rs:57-88)
  std::vector<uint64_t> edges_h;
  // input file of each record (cf_seq's Reference:<file+1>, src/output.rs:176-177)
  std::vector<uint32_t> rec_file;
  uint64_t W = 0, S = 0, D = 0;
  uint32_t num_bins = 0;
  int sb = 0;
  uint64_t n_segs = 0, n_chunks = 0;
  uint32_t n_over = 0;
  uint64_t n_blocks = 0;
  uint32_t n_big = 0;
  // learned index
  bool lidx_ready = false, lidx_fallback = false, lidx_states = false;
  bool keys_ready = false;  // dense keys compacted since the last P3
  uint32_t lidx_R = 0, lidx_L = 0, lidx_Lb = 0;  // root pieces; leaves in all (L + R); L
  int lidx_s = 0;
  uint32_t lidx_C = 0, lidx_lines = 0, lidx_lw = 16, lidx_magic = 1, lidx_mshift = 0, lidx_steps = 1;
  int lidx_groups = 0;
  int use_index = 0;
  // launch shapes the autotuner may change (cf1g_autotune); the params'
  // non-zero fields take precedence
  uint32_t tuned_p4_chunk = 8192, tuned_p4_wgs = 8;
  // hipGraph segments (cf1g_set_graphs); alloc_gen changes whenever a
  // device buffer is (re)allocated, which invalidates every captured graph
  bool graphs_on = true;
  uint64_t alloc_gen = 0, graph_captures = 0, graph_replays = 0;
  std::map<std::string, GraphSeg> graphs;
  // records from cf1g_load_fasta, kept on the host for P5; the stream was
  // packed on the host (P1 then starts from the packed stream)
  HostRecords host;
  bool host_records = false, host_packed = false;
  // MPHF and states restored from a cf1g_save file: no keys or joined masks
  bool restored = false;
  void* pinned = nullptr;
  size_t pinned_cap = 0;
  // streaming P1 (cf1g_p1_stream): host packing threads, a copy stream, and
  // pinned staging slots with the events that free them
  std::unique_ptr<HostPool> pool;
  hipStream_t st_copy = nullptr;
  static constexpr int N_SLOTS = 3;
  void* slot_mem = nullptr;
  size_t slot_words = 0;  // packed words per slot
  hipEvent_t slot_ev[N_SLOTS] = {};

  void* get(const std::string& name, size_t bytes) {
    DevBuf& b = bufs[name];
    if (bytes == 0) bytes = 8;
    if (b.cap < bytes) {
      if (b.p) hip_check(hipFree(b.p), "hipFree");
      b.p = nullptr;
      b.cap = 0;
      size_t want = bytes + bytes / 8;  // headroom for repeated runs
      hip_check(hipMalloc(&b.p, want), ("hipMalloc " + name).c_str());
      b.cap = want;
      alloc_gen++;
    }
    return b.p;
  }
  template <typename T>
  T* getT(const std::string& name, size_t count) {
    return (T*)get(name, count * sizeof(T));
  }
  MphfView view() {
    MphfView v;
    v.b2b = getT<uint16_t>("b2b", CF1G_NUM_BUCKETS);
    v.key_off = getT<uint64_t>("seg_out_off", n_segs + 1);
    v.blk_off = getT<uint64_t>("bbh_blk_off", n_segs + 1);
    v.blocks = getT<uint64_t>("blocks", n_blocks * 8 + 8);
    v.K = k;
    v.m = m;
    v.sb = sb;
    return v;
  }
  LidxView lview() {
    LidxView v;
    v.root = getT<RootPiece>("lidx_root", lidx_R + 1);
    v.leaf = getT<LeafRec>("lidx_leaf", lidx_L + 1);
    v.keys = getT<uint64_t>("lidx_keys", D + 1);
    v.lines = getT<uint64_t>("lidx_lines", (uint64_t)lidx_lines * lidx_lw + 2);
    v.bbh2rank = lidx_fallback ? getT<uint32_t>("lidx_bbh2rank", D + 1) : nullptr;
    v.warp = getT<uint32_t>("lidx_warp", WARP_MAX_GROUPS * 256);
    v.D = D;
    v.R = lidx_R;
    v.LT = lidx_L;
    v.C = lidx_C;
    v.n_lines = lidx_lines;
    v.magic = lidx_magic;
    v.mshift = lidx_mshift;
    v.max_steps = lidx_steps;
    v.s = lidx_s;
    v.line_words = (int)lidx_lw;
    v.key_bits = 2 * k;
    v.warp_groups = lidx_groups;
    v.fallback = lidx_fallback;
    return v;
  }
  // Deferred-search queue of a lookup of n keys through the learned index
  // (none when leaves resolve through the BBHash).
  LidxQueue lqueue(uint64_t n) {
    LidxQueue q{nullptr, nullptr, 0};
    if (lidx_fallback || n == 0) return q;
    uint32_t waves = 0, cap = 0;
    lidx_queue_shape(n, waves, cap);
    if (!cap) return q;
    q.ent = getT<uint64_t>("lidx_queue", (uint64_t)waves * cap * 3);
    q.cnt = getT<uint32_t>("lidx_queue_cnt", waves);
    q.cap = cap;
    return q;
  }
  void sync() { hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"); }
  void zero(void* p, size_t bytes) { launch_zero(st, p, bytes); }
  // Run `launches` (kernel launches on st only: no allocation, copy, memset
  // or synchronisation) through a hipGraph captured under `name`, replayed
  // while `key` and the allocation generation are unchanged. Eager when
  // graphs are off or per-kernel timing is on.
  template <typename F>
  void seg(const char* name, std::vector<uint64_t> key, F&& launches) {
    if (!graphs_on || g_prof_on) {
      launches();
      return;
    }
    key.push_back(alloc_gen);
    key.push_back((uint64_t)wide);
    GraphSeg& g = graphs[name];
    if (!g.valid || g.key != key) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
      g.valid = false;
      hipGraph_t graph = nullptr;
      hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
      try {
        launches();
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(st, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      hip_check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
      size_t n = 0;
      hipError_t e = hipGraphGetNodes(graph, nullptr, &n);
      if (e == hipSuccess && n) e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(e, "hipGraphInstantiate");
      g.nodes = n;
      g.key = key;
      g.valid = true;
      graph_captures++;
    } else {
      graph_replays++;
    }
    if (g.nodes) hip_check(hipGraphLaunch(g.exec, st), "hipGraphLaunch");
  }
  void d2h(void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st), "hipMemcpy D2H");
    sync();
  }
  void h2d(void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "hipMemcpy H2D");
    sync();
  }
};

namespace {
// Records per work-group in the P3 count / scatter passes (~37K k-mers, so
// each (chunk, part) run of the scatter is ~hundreds of keys).
constexpr uint32_t CHUNK_RECORDS = 4096;
constexpr uint64_t STREAM_PAD = 4096;

struct Fail {
  int code;
  std::string msg;
};
[[noreturn]] void fail(int code, const std::string& m) { throw Fail{code, m}; }

template <typename F>
int guarded(cf1g_ctx* c, F f) {
  if (!c) return CF1G_E_INVALID;
  try {
    f();
    return CF1G_OK;
  } catch (const Fail& e) {
    c->err = e.msg;
    return e.code;
  } catch (const HipError& e) {
    c->err = e.what();
    return CF1G_E_HIP;
  } catch (const std::exception& e) {
    c->err = e.what();
    return CF1G_E_INTERNAL;
  }
}
void need(cf1g_ctx* c, int stage, const char* what) {
  if (c->stage < stage) fail(CF1G_E_ORDER, std::string(what) + " called before its inputs exist");
}
int minimizer_len(int k) {  // src/params.rs:221-228
  int m = std::min(k / 2, 15);
  if (m % 2 == 0) m = std::max(m - 1, 1);
  return m;
}
uint64_t level_bits(uint64_t remaining) {
  uint64_t n = std::max<uint64_t>(2 * remaining, 1);
  return (n + BBH_BLOCK_BITS - 1) / BBH_BLOCK_BITS * BBH_BLOCK_BITS;
}

// ---------------------------------------------------------------- phases
// Each phase's launches between host synchronisations form one segment
// (cf1g_ctx::seg): captured into a hipGraph the first time and replayed
// while the sizes and buffers it was captured with are unchanged.
void do_p1(cf1g_ctx* c) {
  const uint64_t nt = p1_tiles(c->n_pos), nsl = p1_slices(c->n_pos);
  uint8_t* ascii = c->host_packed ? nullptr : c->getT<uint8_t>("ascii", c->n_ascii + 8);
  uint64_t* roff = c->host_packed ? nullptr : c->getT<uint64_t>("rec_off", c->n_rec + 1);
  uint64_t* gst = c->host_packed ? nullptr : c->getT<uint64_t>("gstart", c->n_rec + 1);
  uint64_t* seq = c->getT<uint64_t>("seq", 2 * c->n_words64 + 8);
  uint64_t* nm = c->getT<uint64_t>("nmask", c->n_words64 + 8);
  auto* hist = c->getT<unsigned long long>("hist", CF1G_NUM_BUCKETS);
  auto* tot = c->getT<unsigned long long>("totals", 4);
  // super k-mer entries: one slot of tile-size per scan tile; partial
  // histograms per slice of tiles
  uint64_t* ent = c->getT<uint64_t>("sk_entries", nt * p1_tile_slots() + 1);
  uint32_t* tcnt = c->getT<uint32_t>("sk_tile_cnt", nt + 1);
  uint32_t* tw = c->getT<uint32_t>("sk_tile_w", nt + 1);
  uint32_t* bkt = c->getT<uint32_t>("sk_buckets", nt * p1_tile_slots() + 1);
  uint32_t* part = c->getT<uint32_t>("p1_part", nsl * CF1G_NUM_BUCKETS + 1);
  uint64_t* sums = c->getT<uint64_t>("p1_sums", 2 * nsl + 2);
  auto* hist_all = c->getT<unsigned long long>("hist_all", CF1G_NUM_BUCKETS);
  const uint32_t n_edges = (uint32_t)(c->edges_h.size() - 1);
  const uint64_t* edges = c->getT<uint64_t>("edges", c->edges_h.size());
  const uint32_t excl_cap = n_edges * 64 + 1;
  uint32_t* excl_cnt = c->getT<uint32_t>("p1_excl_cnt", 2);  // [0] count, [1] overflow flag
  uint16_t* excl = c->getT<uint16_t>("p1_excl", excl_cap);
  c->seg("p1", {c->n_rec, c->n_pos, c->n_words64, (uint64_t)c->k, (uint64_t)c->m, (uint64_t)c->host_packed, n_edges}, [&] {
    hipStream_t st = c->st;
    if (!c->host_packed) launch_pack(st, ascii, roff, gst, c->n_rec, c->n_words64, seq, nm);
    c->zero(excl_cnt, 8);
    if (nt == 0) {  // otherwise k_p1_reduce writes every bucket and both totals
      c->zero(hist, CF1G_NUM_BUCKETS * 8);
      c->zero(hist_all, CF1G_NUM_BUCKETS * 8);
      c->zero(tot, 32);
    }
    launch_p1(st, seq, nm, c->n_pos, c->k, c->m, ent, bkt, tcnt, tw, part, sums, hist, hist_all, tot, edges, n_edges, excl_cnt, excl,
              excl_cap, excl_cnt + 1);
  });
  c->set_stage(2);
}

// Chunk edges (stream positions) of the loaded records, plus a UINT64_MAX
// sentinel, into edges_h and the device buffer "edges".
void set_chunk_edges(cf1g_ctx* c) {
  constexpr uint64_t CHUNK_BASES = 4000000;  // src/minimizer.rs:13, src/superkmer.rs:16
  c->edges_h.clear();
  for (uint64_t r = 0; r < c->n_rec; r++) {
    const uint64_t len = c->gstart_h[r + 1] - 1 - c->gstart_h[r];
    for (uint64_t a = CHUNK_BASES; a < len; a += CHUNK_BASES) c->edges_h.push_back(c->gstart_h[r] + a);
  }
  c->edges_h.push_back(UINT64_MAX);
  hip_check(hipMemcpyAsync(c->getT<uint64_t>("edges", c->edges_h.size()), c->edges_h.data(), c->edges_h.size() * 8,
                           hipMemcpyHostToDevice, c->st),
            "H2D chunk edges");
}

// Record sizes, record starts and chunk edges of host records (what
// cf1g_load_records and cf1g_p1_stream share).
void set_records(cf1g_ctx* c, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec) {
  if (!rec_off) fail(CF1G_E_INVALID, "rec_off is NULL");
  for (uint64_t r = 0; r < n_rec; r++)
    if (rec_off[r + 1] < rec_off[r]) fail(CF1G_E_INVALID, "rec_off must be non-decreasing");
  const uint64_t base = rec_off[0];
  if (rec_off[n_rec] - base > 0 && !ascii) fail(CF1G_E_INVALID, "ascii is NULL");
  c->n_rec = n_rec;
  c->n_ascii = rec_off[n_rec] - base;
  c->gstart_h.assign(n_rec + 1, 0);
  for (uint64_t r = 0; r <= n_rec; r++) c->gstart_h[r] = rec_off[r] - base + r + 1;
  c->n_pos = c->n_ascii + n_rec + 1;
  c->n_words64 = (c->n_pos + STREAM_PAD + 63) / 64;
  c->host_records = c->host_packed = c->restored = false;
  c->rec_file.assign(n_rec, 0);
  c->W = c->S = c->D = 0;
}

// P1 over host records while they stream to the device: host threads pack
// chunks of the 2-bit stream into pinned slots, each chunk's copy (on the
// copy stream) overlaps the packing of the next, and the P1 scan runs over
// the tiles whose stream words have landed; the histogram follows the last
// chunk. The products equal cf1g_load_records + cf1g_p1_histogram's.
void do_p1_stream(cf1g_ctx* c, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint32_t threads) {
  set_records(c, ascii, rec_off, n_rec);
  c->set_stage(0);
  set_chunk_edges(c);
  // the calling thread issues the copies and launches; th - 1 pool threads pack
  const int th = std::max(2, threads ? (int)threads : (int)std::thread::hardware_concurrency());
  if (!c->pool || c->pool->threads() != th) c->pool.reset(new HostPool(th));
  if (!c->st_copy) hip_check(hipStreamCreateWithFlags(&c->st_copy, hipStreamNonBlocking), "hipStreamCreate (copy)");
  // stream words per chunk (default 2^18: 16.8 M positions, 6 MiB packed)
  const uint32_t cl = c->prm.stream_chunk_log2 ? c->prm.stream_chunk_log2 : 18;
  if (cl < 10 || cl > 24) fail(CF1G_E_INVALID, "stream_chunk_log2 must be in [10, 24]");
  const size_t CW = size_t(1) << cl;
  if (c->slot_words != CW) {
    if (c->slot_mem) hip_check(hipHostFree(c->slot_mem), "hipHostFree (staging)");
    c->slot_mem = nullptr;
    c->slot_words = 0;
    hip_check(hipHostMalloc(&c->slot_mem, cf1g_ctx::N_SLOTS * 3 * CW * 8, hipHostMallocDefault), "hipHostMalloc (staging)");
    c->slot_words = CW;
  }
  if (!c->slot_ev[0])
    for (auto& e : c->slot_ev) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  const uint64_t nw = c->n_words64, nt = p1_tiles(c->n_pos), nsl = p1_slices(c->n_pos);
  uint64_t* seq = c->getT<uint64_t>("seq", 2 * nw + 8);
  uint64_t* nm = c->getT<uint64_t>("nmask", nw + 8);
  auto* hist = c->getT<unsigned long long>("hist", CF1G_NUM_BUCKETS);
  auto* hist_all = c->getT<unsigned long long>("hist_all", CF1G_NUM_BUCKETS);
  auto* tot = c->getT<unsigned long long>("totals", 4);
  uint64_t* ent = c->getT<uint64_t>("sk_entries", nt * p1_tile_slots() + 1);
  uint32_t* tcnt = c->getT<uint32_t>("sk_tile_cnt", nt + 1);
  uint32_t* tw = c->getT<uint32_t>("sk_tile_w", nt + 1);
  uint32_t* bkt = c->getT<uint32_t>("sk_buckets", nt * p1_tile_slots() + 1);
  uint32_t* part = c->getT<uint32_t>("p1_part", nsl * CF1G_NUM_BUCKETS + 1);
  uint64_t* sums = c->getT<uint64_t>("p1_sums", 2 * nsl + 2);
  const uint32_t n_edges = (uint32_t)(c->edges_h.size() - 1);
  const uint64_t* edges = c->getT<uint64_t>("edges", c->edges_h.size());
  const uint32_t excl_cap = n_edges * 64 + 1;
  uint32_t* excl_cnt = c->getT<uint32_t>("p1_excl_cnt", 2);
  uint16_t* excl = c->getT<uint16_t>("p1_excl", excl_cap);
  hipStream_t st = c->st;
  c->zero(excl_cnt, 8);
  if (nt == 0) {
    c->zero(hist, CF1G_NUM_BUCKETS * 8);
    c->zero(hist_all, CF1G_NUM_BUCKETS * 8);
    c->zero(tot, 32);
  }
  // the first records' offset is subtracted by pack_words (rec_off[0] is the base)
  const uint8_t* a = ascii ? ascii + rec_off[0] : (const uint8_t*)"";
  uint64_t t_done = 0;
  // The pool packs every chunk's parts in order, each chunk into staging slot
  // chunk % N_SLOTS once that slot's previous copy has completed
  // (slot_free[s] = the first chunk that may use slot s); the calling thread
  // waits for each chunk's parts, issues its copies and the scan of its
  // tiles, and frees the slot when the copy is done.
  constexpr int NS = cf1g_ctx::N_SLOTS;
  const uint64_t n_chunks = (nw + CW - 1) / CW;
  const uint64_t parts = (uint64_t)(th - 1) * 2;
  std::unique_ptr<std::atomic<uint64_t>[]> left(new std::atomic<uint64_t>[n_chunks + 1]);
  for (uint64_t i = 0; i < n_chunks; i++) left[i].store(parts);
  std::atomic<uint64_t> slot_free[NS];
  for (int s = 0; s < NS; s++) slot_free[s].store((uint64_t)s);
  std::atomic<bool> abandon{false};
  const std::function<void(uint64_t)> pack_part = [&](uint64_t task) {
    const uint64_t ci = task / parts, p = task % parts;
    const int s = (int)(ci % NS);
    while (slot_free[s].load(std::memory_order_acquire) < ci)
      if (abandon.load(std::memory_order_relaxed)) return;
    const uint64_t w0 = ci * CW, w1 = std::min<uint64_t>(nw, w0 + CW);
    uint64_t* sseq = (uint64_t*)c->slot_mem + (size_t)s * 3 * CW;
    uint64_t* snm = sseq + 2 * CW;
    const uint64_t a0 = w0 + (w1 - w0) * p / parts, a1 = w0 + (w1 - w0) * (p + 1) / parts;
    pack_words(a, rec_off, c->gstart_h.data(), n_rec, a0, a1, sseq + 2 * (a0 - w0), snm + (a0 - w0));
    left[ci].fetch_sub(1, std::memory_order_release);
  };
  c->pool->start(n_chunks * parts, pack_part);
  // a HIP error below must not leave the pool waiting on a slot forever
  struct Join {
    HostPool* pool;
    std::atomic<bool>* abandon;
    ~Join() {
      abandon->store(true);
      pool->wait();
    }
  } join{c->pool.get(), &abandon};
  for (uint64_t i = 0; i < n_chunks; i++) {
    const uint64_t w0 = i * CW, w1 = std::min<uint64_t>(nw, w0 + CW);
    const int s = (int)(i % NS);
    uint64_t* sseq = (uint64_t*)c->slot_mem + (size_t)s * 3 * CW;
    uint64_t* snm = sseq + 2 * CW;
    while (left[i].load(std::memory_order_acquire) != 0) _mm_pause();
    hip_check(hipMemcpyAsync(seq + 2 * w0, sseq, 2 * (w1 - w0) * 8, hipMemcpyHostToDevice, c->st_copy), "H2D stream chunk");
    hip_check(hipMemcpyAsync(nm + w0, snm, (w1 - w0) * 8, hipMemcpyHostToDevice, c->st_copy), "H2D placeholder chunk");
    hip_check(hipEventRecord(c->slot_ev[s], c->st_copy), "hipEventRecord");
    hip_check(hipStreamWaitEvent(st, c->slot_ev[s], 0), "hipStreamWaitEvent");
    // tiles whose reads lie inside the words that have landed
    uint64_t t_ready = nt;
    if (w1 < nw) {
      const uint64_t have = w1 * 64;
      t_ready = t_done;
      while (t_ready < nt && p1_tile_reach(t_ready) <= have) t_ready++;
    }
    launch_p1_scan(st, seq, nm, c->n_pos, t_done, t_ready, c->k, c->m, ent, bkt, tcnt, tw, edges, n_edges, excl_cnt, excl, excl_cap);
    t_done = t_ready;
    hip_check(hipEventSynchronize(c->slot_ev[s]), "hipEventSynchronize (staging slot)");
    slot_free[s].store(i + NS, std::memory_order_release);
  }
  launch_p1_finish(st, c->n_pos, bkt, tcnt, tw, part, sums, hist, hist_all, tot, n_edges, excl_cnt, excl, excl_cap, excl_cnt + 1);
  c->sync();
  c->host_packed = true;
  c->set_stage(2);
}

void do_p2(cf1g_ctx* c, const uint64_t* hist_in, uint32_t num_bins) {
  hipStream_t st = c->st;
  if (num_bins < 1 || num_bins > CF1G_MAX_BINS) fail(CF1G_E_INVALID, "num_bins must be in [1, 4096]");
  auto* hpart = c->getT<unsigned long long>("hist", CF1G_NUM_BUCKETS);     // P1's histogram: what the partition splits
  auto* hist = c->getT<unsigned long long>("hist_all", CF1G_NUM_BUCKETS);  // every routed super k-mer: the bin sizes
  if (hist_in) {
    hpart = c->getT<unsigned long long>("hist_in", CF1G_NUM_BUCKETS);
    hip_check(hipMemcpyAsync(hpart, hist_in, CF1G_NUM_BUCKETS * 8, hipMemcpyHostToDevice, st), "H2D hist");
  }
  uint16_t* b2b = c->getT<uint16_t>("b2b", CF1G_NUM_BUCKETS);
  uint64_t* brec = c->getT<uint64_t>("bin_records", CF1G_MAX_BINS + 1);
  uint64_t* bstart = c->getT<uint64_t>("bin_start", CF1G_MAX_BINS + 1);
  auto* bcur = c->getT<unsigned long long>("bin_cursor", CF1G_MAX_BINS + 1);
  uint64_t* coff = c->getT<uint64_t>("chunk_off", CF1G_MAX_BINS + 1);
  uint64_t* info = c->getT<uint64_t>("info", 4);
  uint64_t* totals = c->getT<uint64_t>("totals", 4);
  uint64_t* pscr = c->getT<uint64_t>("partition_scratch", partition_scratch_words());
  c->seg("p2a", {num_bins, hist_in ? 1u : 0u}, [&] {
    launch_partition(c->st, hpart, hist, num_bins, CHUNK_RECORDS, b2b, brec, bstart, bcur, coff, info, pscr);
  });
  uint64_t hinfo[4], htot[4];
  uint32_t hexcl[2];
  hip_check(hipMemcpyAsync(hinfo, info, 32, hipMemcpyDeviceToHost, st), "D2H info");
  hip_check(hipMemcpyAsync(htot, totals, 32, hipMemcpyDeviceToHost, st), "D2H totals");
  hip_check(hipMemcpyAsync(hexcl, c->getT<uint32_t>("p1_excl_cnt", 2), 8, hipMemcpyDeviceToHost, st), "D2H P1 edge list");
  c->sync();
  if (hexcl[1]) fail(CF1G_E_INTERNAL, "P1 chunk-edge list overflowed");
  c->num_bins = (uint32_t)hinfo[0];
  c->S = hinfo[1];
  c->n_chunks = hinfo[2];
  c->W = htot[0];
  if (c->S != htot[1]) fail(CF1G_E_INTERNAL, "bin record total differs from the P1 super k-mer count");
  if (c->S >= (1ULL << 32)) fail(CF1G_E_INVALID, "more than 2^32 super k-mers (record indices are 32-bit)");
  // route: each chunk counts its records per bin, reserves a run from each
  // bin's cursor and writes them
  const uint64_t nt = p1_tiles(c->n_pos);
  uint64_t chunk_tiles = 0, n_chunks = 0;
  p2_chunking(c->n_pos, c->num_bins, &chunk_tiles, &n_chunks);
  uint64_t* rec = c->getT<uint64_t>("rec", c->S * (uint64_t)c->rw + 1);
  uint16_t* rec_sub = c->getT<uint16_t>("rec_sub", c->S + 8);
  const uint64_t nslots = nt * p1_tile_slots();
  uint32_t* ent_dest = c->getT<uint32_t>("ent_dest", nslots + 1);
  // one read after the route: the routed-record total, then the overflow flag
  auto* route_chk = c->getT<unsigned long long>("route_check", 2);
  auto* routed_d = route_chk;
  unsigned int* ovf = (unsigned int*)(route_chk + 1);
  const uint64_t* ent = c->getT<uint64_t>("sk_entries", nt * p1_tile_slots() + 1);
  const uint32_t* bkt = c->getT<uint32_t>("sk_buckets", nt * p1_tile_slots() + 1);
  const uint32_t* tcnt = c->getT<uint32_t>("sk_tile_cnt", nt + 1);
  uint64_t* seq = c->getT<uint64_t>("seq", 2 * c->n_words64 + 8);
  uint64_t* nmask = c->getT<uint64_t>("nmask", c->n_words64 + 8);
  c->seg("p2b", {c->n_pos, c->num_bins, c->S, chunk_tiles, n_chunks, (uint64_t)c->rw, (uint64_t)c->k}, [&] {
    hipStream_t s2 = c->st;
    c->zero(route_chk, 16);
    c->zero(bcur, (uint64_t)c->num_bins * 8);
    launch_p2_route(s2, seq, nmask, ent, bkt, tcnt, c->n_pos, chunk_tiles, n_chunks, c->k, b2b, (int)c->num_bins, bstart, bcur, rec,
                    rec_sub, ent_dest, c->rw, ovf, routed_d);
  });
  unsigned long long chk[2] = {0, 0};
  c->d2h(chk, route_chk, 16);
  if (chk[1] || chk[0] != c->S) fail(CF1G_E_INTERNAL, "P2 routing disagrees with the bin sizes from the histogram");
  c->set_stage(3);
}

// P3: records and k-mers per MPHF segment (a segment is fixed by a super
// k-mer's minimizer, so P2's records are grouped, not their k-mers), scans,
// the records into their segments' runs, then per segment: its k-mers
// decoded from the records, dedup / mask join / BBHash into padded block
// slots (masks in local-rank order), scans of the distinct counts (MPHF key
// offsets) and block counts — one segment up to here — then, for the rare
// segments above LDS capacity, their occurrences expanded into global
// memory, the global-memory sort / BBHash and the scans again; then one
// compaction of the masks (MPHF-value order) and blocks. The dense key array
// is built on first use (dense_keys) from the keys the dedup left per
// occurrence.
void do_p3(cf1g_ctx* c) {
  hipStream_t st = c->st;
  const size_t kb = c->key_bytes;
  const int K = c->k;
  // Segment bits: the fewest (<= 12) that bring the mean segment to <= target
  // k-mer occurrences. The MPHF partition depends on it (it is exported).
  uint64_t target = c->prm.p3_target_seg ? c->prm.p3_target_seg : 1536;
  int sb = 0;
  while (sb < 12 && c->W > (uint64_t)c->num_bins * (1ULL << sb) * target) sb++;
  c->sb = sb;
  c->n_segs = (uint64_t)c->num_bins << sb;
  const uint64_t ns = c->n_segs;
  auto* kcnt = c->getT<unsigned long long>("seg_counts", ns + 1);
  auto* rcnt = c->getT<unsigned long long>("seg_rcounts", ns + 1);
  uint64_t* seg_off = c->getT<uint64_t>("seg_off", ns + 1);
  uint64_t* roff = c->getT<uint64_t>("seg_roff", ns + 1);
  unsigned int* rcur = c->getT<unsigned int>("seg_rcur", ns + 1);
  uint64_t* ucount = c->getT<uint64_t>("seg_ucount", ns + 1);
  uint64_t* key_off = c->getT<uint64_t>("seg_out_off", ns + 1);
  uint64_t* nblk = c->getT<uint64_t>("bbh_seg_nblk", ns + 1);
  uint64_t* blk_off = c->getT<uint64_t>("bbh_blk_off", ns + 1);
  uint32_t* over = c->getT<uint32_t>("seg_over", ns + 1);
  uint32_t* over2 = c->getT<uint32_t>("seg_over2", ns + 1);
  uint32_t* big = c->getT<uint32_t>("bbh_big", ns + 1);
  // [0] n_over, [1] n_big, [2] error bits, [3] count gate, [4] n_over2
  unsigned int* flags = c->getT<unsigned int>("p3_flags", 8);
  uint64_t* tmp = c->getT<uint64_t>("scan_tmp", scan_temp_words(ns + 1) + 16);
  const uint64_t* rec = c->getT<uint64_t>("rec", c->S * (uint64_t)c->rw + 1);
  const uint16_t* rec_sub = c->getT<uint16_t>("rec_sub", c->S + 8);
  uint64_t* srec = c->getT<uint64_t>("rec_seg", c->S * (uint64_t)c->rw + 1);
  uint32_t* rec_dest = c->getT<uint32_t>("rec_dest", c->S + 1);
  const uint64_t* bstart = c->getT<uint64_t>("bin_start", CF1G_MAX_BINS + 1);
  const uint64_t* coff = c->getT<uint64_t>("chunk_off", CF1G_MAX_BINS + 1);
  // occ: every occurrence's key (the dedup's output; the expanded
  // occurrences of oversized segments); obs: masks of expanded occurrences;
  // aux: the global-memory sort's other buffer
  void* occ = c->get("occ", c->W * kb + 16);
  uint16_t* occ_obs = c->getT<uint16_t>("occ_obs", c->W + 8);
  // padded BBHash block slots: floor(seg_off/64) + 16*seg, floor(n/64) + 16 blocks each
  uint64_t* blocks_pad = c->getT<uint64_t>("blocks_pad", 8 * (c->W / 64 + 16 * (ns + 1)) + 8);
  uint16_t* vpad = c->getT<uint16_t>("vpad", c->W + 8);
  auto scans = [&] {
    scan_excl_u64(st, ucount, key_off, ns, key_off + ns, tmp);
    scan_excl_u64(st, nblk, blk_off, ns, blk_off + ns, tmp);
  };
  const bool direct = !c->wide && 2 * K <= 12;  // k <= 6: the dedup's direct-address table reads expanded occurrences
  const std::vector<uint64_t> shape = {c->S, c->W, c->n_chunks, c->num_bins, (uint64_t)sb, (uint64_t)K, (uint64_t)c->rw};
  // counts and scans, then (unless the gate finds the counted occurrences
  // differ from W) the records into segments and the dedup: one launch
  // sequence, checked afterwards with the flags
  c->seg("p3ab", shape, [&] {
    c->zero(kcnt, (ns + 1) * 8);
    c->zero(rcnt, (ns + 1) * 8);
    c->zero(rcur, (ns + 1) * 4);
    c->zero(flags, 32);
    launch_p3_rcount(st, c->wide, rec, rec_sub, bstart, coff, (int)c->num_bins, CHUNK_RECORDS, c->n_chunks, K, sb, rcnt, kcnt);
    scan_excl_u64(st, (const uint64_t*)kcnt, seg_off, ns, seg_off + ns, tmp);
    scan_excl_u64(st, (const uint64_t*)rcnt, roff, ns, roff + ns, tmp);
    launch_p3_gate(st, seg_off + ns, c->W, flags);
    launch_p3_rscatter(st, c->wide, rec, rec_sub, bstart, coff, (int)c->num_bins, CHUNK_RECORDS, c->n_chunks, sb, roff, rcur, srec,
                       rec_dest, flags + 3);
    if (direct) launch_p3_expand(st, c->wide, srec, roff, seg_off, nullptr, ns, K, occ, occ_obs, flags + 3);
    launch_p3_dedup(st, c->wide, occ, occ_obs, srec, roff, seg_off, ns, K, sb, ucount, nblk, over, flags, big, blocks_pad, vpad);
    launch_p3_dedup_list(st, c->wide, occ, occ_obs, srec, roff, seg_off, ns, K, sb, ucount, nblk, over, over2, flags, blocks_pad, vpad);
    scans();
  });
  unsigned int hf[8];
  uint64_t tot[3] = {0, 0, 0};
  hip_check(hipMemcpyAsync(hf, flags, 32, hipMemcpyDeviceToHost, st), "D2H flags");
  hip_check(hipMemcpyAsync(&tot[1], key_off + ns, 8, hipMemcpyDeviceToHost, st), "D2H D");
  c->d2h(&tot[2], blk_off + ns, 8);
  if (hf[3]) fail(CF1G_E_INTERNAL, "k-mers in bins differ from the valid k-mer count");
  // segments the work-group dedups could not hold: sorted through global memory
  c->n_over = hf[4];
  if (c->n_over > 0) {
    void* aux = c->get("occ_aux", c->W * kb + 16);
    uint16_t* aux_obs = c->getT<uint16_t>("occ_aux_obs", c->W + 8);
    launch_p3_expand(st, c->wide, srec, roff, seg_off, over2, c->n_over, K, occ, occ_obs, flags + 3);
    launch_p3_bigsort(st, c->wide, occ, occ_obs, aux, aux_obs, seg_off, over2, c->n_over, ucount, flags, big);
    c->d2h(hf, flags, 12);
  }
  c->n_big = hf[1];
  if (c->n_big) {
    // segments above LDS capacity: seen / collision words in scratch (at
    // 4*floor(seg_off/64) + 8*seg, 2*(D_s/32 + 2) words), alive bits per
    // occurrence position
    uint64_t* scratch = c->getT<uint64_t>("bbh_scratch", 4 * (c->W / 64 + 2) + 8 * (ns + 1) + 1024);
    const uint64_t alive_words = c->W / 64 + 2;
    uint64_t* alive = c->getT<uint64_t>("bbh_alive", alive_words);
    hip_check(hipMemsetAsync(alive, 0xFF, alive_words * 8, st), "memset");
    launch_bbh_big(st, c->wide, occ, occ_obs, vpad, seg_off, ucount, big, c->n_big, nblk, blocks_pad, scratch, alive, occ, flags + 2);
  }
  if (c->n_over || c->n_big) {
    scans();
    hip_check(hipMemcpyAsync(hf, flags, 12, hipMemcpyDeviceToHost, st), "D2H flags");
    hip_check(hipMemcpyAsync(&tot[1], key_off + ns, 8, hipMemcpyDeviceToHost, st), "D2H D");
    c->d2h(&tot[2], blk_off + ns, 8);
  }
  if (hf[2]) fail(CF1G_E_INTERNAL, "BBHash build exceeded its level or region limit (flags " + std::to_string(hf[2]) + ")");
  const uint64_t D = tot[1];
  c->D = D;
  c->n_blocks = tot[2];
  const uint64_t nsw = (D + 11) / 12;
  uint64_t* blocks = c->getT<uint64_t>("blocks", c->n_blocks * 8 + 8);
  c->getT<uint64_t>("states", nsw + 1);  // allocated with the MPHF; P4 writes every word
  c->getT<uint8_t>("vstate", 12 * nsw + 16);
  // dense blocks now (the masks stay in the segments' runs of vpad for P4);
  // the dense key array (which the reference's MPHF does not keep) is built
  // from occ on first use, dense_keys()
  c->seg("p3c", {ns, D, c->n_blocks}, [&] { launch_p3_compact(st, c->wide, seg_off, blk_off, blocks_pad, ns, blocks); });
  c->keys_ready = false;
  c->lidx_ready = c->lidx_states = false;
  c->use_index = 0;
  c->state_gen++;
  c->restored = false;
  c->set_stage(4);
}

// The distinct keys in MPHF value order (keys[v] has value v), built on
// first use from the per-segment key runs P3 left in occ (each key placed by
// its BBHash value).
void* dense_keys(cf1g_ctx* c) {
  if (c->restored) fail(CF1G_E_ORDER, "a restored MPHF keeps no keys (run P1..P3 for them)");
  const uint64_t kb = c->key_bytes;
  void* keys = c->get("keys", c->D * kb + 16);
  if (!c->keys_ready) {
    const uint64_t ns = c->n_segs;
    unsigned int* err = c->getT<unsigned int>("keys_err", 2);
    c->zero(err, 4);
    launch_p3_keys(c->st, c->wide, c->get("occ", c->W * kb + 16), c->getT<uint64_t>("seg_off", ns + 1),
                   c->getT<uint64_t>("seg_ucount", ns + 1), c->getT<uint64_t>("seg_out_off", ns + 1),
                   c->getT<uint64_t>("bbh_blk_off", ns + 1), c->getT<uint64_t>("blocks", c->n_blocks * 8 + 8), ns, keys, err);
    unsigned int he = 0;
    c->d2h(&he, err, 4);
    if (he) fail(CF1G_E_INTERNAL, "a P3 key is missing from its MPHF segment");
    c->keys_ready = true;
  }
  return keys;
}

// P4: the joined masks, in MPHF-value order within each segment's run of
// vpad, become one state byte per value (dense), then packed states.
void do_p4(cf1g_ctx* c) {
  if (c->restored) fail(CF1G_E_ORDER, "a restored MPHF keeps no joined observations: classify with cf1g_p4_classify_cas");
  const uint64_t nsw = (c->D + 11) / 12, ns = c->n_segs;
  uint8_t* vstate = c->getT<uint8_t>("vstate", 12 * nsw + 16);
  uint64_t* states = c->getT<uint64_t>("states", nsw + 1);
  const uint16_t* vpad = c->getT<uint16_t>("vpad", c->W + 8);
  const uint64_t* seg_off = c->getT<uint64_t>("seg_off", ns + 1);
  const uint64_t* ucount = c->getT<uint64_t>("seg_ucount", ns + 1);
  const uint64_t* key_off = c->getT<uint64_t>("seg_out_off", ns + 1);
  c->seg("p4", {ns, c->D}, [&] { launch_p4_pack(c->st, vpad, seg_off, ucount, key_off, ns, c->D, vstate, states); });
  c->sync();
  c->set_stage(5);
}

// P4 in the reference's form: zeroed states, then one MPHF lookup and CAS
// join per k-mer occurrence of the loaded records (k_p4_cas).
// P4 in the reference's form: one lookup per k-mer occurrence and a CAS join,
// through the BBHash (index 0, into "states") or the learned index (1, into
// "states_lidx", in rank order).
void do_p4_cas_via(cf1g_ctx* c, uint32_t flags, cf1g_cas_stats* out, int index) {
  hipStream_t st = c->st;
  const uint64_t nsw = (c->D + 11) / 12;
  uint64_t* states = c->getT<uint64_t>(index == 1 ? "states_lidx" : "states", nsw + 1);
  uint64_t* cnt = c->getT<uint64_t>("p4_counters", 8);
  c->zero(states, (nsw + 1) * 8);
  c->zero(cnt, 64);
  const uint32_t chunk = c->prm.p4_chunk ? c->prm.p4_chunk : c->tuned_p4_chunk;
  const uint32_t wpc = c->prm.p4_wgs_per_cu ? c->prm.p4_wgs_per_cu : c->tuned_p4_wgs;
  if (chunk < 256 || chunk > (1u << 24)) fail(CF1G_E_INVALID, "p4_chunk must be in [256, 2^24]");
  if (wpc < 1 || wpc > 32) fail(CF1G_E_INVALID, "p4_wgs_per_cu must be in [1, 32]");
  const uint64_t n_chunks = c->n_pos > 1 ? (c->n_pos - 1 + chunk - 1) / chunk : 0;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)c->cu_count * wpc, std::max<uint64_t>(n_chunks, 1));
  hipEvent_t e0 = ev_get(), e1 = ev_get();
  hip_check(hipEventRecord(e0, st), "hipEventRecord");
  const uint64_t* seq = c->getT<uint64_t>("seq", 2 * c->n_words64 + 8);
  const uint64_t* nm = c->getT<uint64_t>("nmask", c->n_words64 + 8);
  const int fr = (flags & CF1G_CAS_FORCE_RETRY) ? 1 : 0;
  if (index == 1) launch_p4_cas_lidx(st, seq, nm, c->n_pos, c->k, c->lview(), c->view(), states, cnt, chunk, grid, fr);
  else launch_p4_cas(st, c->wide, seq, nm, c->n_pos, c->k, c->view(), c->D, states, cnt, chunk, grid, fr);
  hip_check(hipEventRecord(e1, st), "hipEventRecord");
  uint64_t h[5] = {0, 0, 0, 0, 0};
  c->d2h(h, cnt, 40);
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  ev_put(e0);
  ev_put(e1);
  if (h[3]) fail(CF1G_E_INTERNAL, std::to_string(h[3]) + " k-mer occurrences are missing from the index");
  if (h[0] != c->W) fail(CF1G_E_INTERNAL, "P4 classified a different number of occurrences than P1 counted");
  if (out) {
    cf1g_cas_stats o{};
    o.occurrences = h[0];
    o.cas_issued = h[1];
    o.cas_failed = h[2];
    o.chunk = chunk;
    o.wgs_per_cu = wpc;
    o.grid = grid;
    o.n_chunks = n_chunks;
    o.kernel_ms = ms;
    *out = o;
  }
  c->state_gen++;
}
void do_p4_cas(cf1g_ctx* c, uint32_t flags, cf1g_cas_stats* out) {
  do_p4_cas_via(c, flags, out, 0);
  c->use_index = 0;
  c->set_stage(5);
}

// ---------------------------------------------------------------- autotuner
// Times candidate launch configurations on the loaded records and keeps the
// fastest: bin count x P3 segment target over whole P1..P4 steps (host wall
// clock around the synchronised phases, the minimum of `reps` runs after
// one warm-up), then P4 chunk x work-groups per CU over the CAS
// classification kernel (HIP events). The context keeps the chosen values
// and ends with the products of the chosen configuration (stage 5).
void do_autotune_run(cf1g_ctx* c, uint32_t reps, cf1g_tune_entry* ent, uint32_t cap, cf1g_tune_result* out);
void do_autotune(cf1g_ctx* c, uint32_t reps, cf1g_tune_entry* ent, uint32_t cap, cf1g_tune_result* out) {
  const cf1g_params saved = c->prm;
  try {
    do_autotune_run(c, reps, ent, cap, out);
  } catch (...) {
    c->prm = saved;  // a failed measurement leaves the caller's settings
    throw;
  }
}
void do_autotune_run(cf1g_ctx* c, uint32_t reps, cf1g_tune_entry* ent, uint32_t cap, cf1g_tune_result* out) {
  static const uint32_t kBins[] = {128, 256, 512};
  static const uint32_t kTarget[] = {1024, 1536, 2048};
  static const uint32_t kChunk[] = {2048, 8192, 32768};
  static const uint32_t kWgs[] = {4, 8, 16};
  if (reps < 1 || reps > 16) fail(CF1G_E_INVALID, "reps must be in [1, 16]");
  std::vector<cf1g_tune_entry> all;
  const cf1g_params saved = c->prm;
  double best_step = 1e300, best_cas = 1e300;
  uint32_t bb = saved.num_bins, bt = saved.p3_target_seg ? saved.p3_target_seg : 1536;
  for (uint32_t nb : kBins) {
    for (uint32_t tg : kTarget) {
      c->prm.num_bins = nb;
      c->prm.p3_target_seg = tg;
      double best = 1e300;
      for (uint32_t r = 0; r <= reps; r++) {
        c->sync();
        auto t0 = std::chrono::steady_clock::now();
        do_p1(c);
        do_p2(c, nullptr, nb);
        do_p3(c);
        do_p4(c);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (r > 0) best = std::min(best, ms);
      }
      all.push_back(cf1g_tune_entry{CF1G_TUNE_STEP, nb, tg, c->num_bins, best});
      if (best < best_step) {
        best_step = best;
        bb = nb;
        bt = tg;
      }
    }
  }
  c->prm.num_bins = bb;
  c->prm.p3_target_seg = bt;
  do_p1(c);
  do_p2(c, nullptr, bb);
  do_p3(c);
  do_p4(c);
  uint32_t bc = c->tuned_p4_chunk, bw = c->tuned_p4_wgs;
  for (uint32_t ch : kChunk) {
    for (uint32_t wg : kWgs) {
      c->prm.p4_chunk = ch;
      c->prm.p4_wgs_per_cu = wg;
      double best = 1e300;
      for (uint32_t r = 0; r <= reps; r++) {
        cf1g_cas_stats cs{};
        do_p4_cas(c, 0, &cs);
        if (r > 0) best = std::min(best, cs.kernel_ms);
      }
      all.push_back(cf1g_tune_entry{CF1G_TUNE_P4_CAS, ch, wg, 0, best});
      if (best < best_cas) {
        best_cas = best;
        bc = ch;
        bw = wg;
      }
    }
  }
  c->prm.p4_chunk = saved.p4_chunk;
  c->prm.p4_wgs_per_cu = saved.p4_wgs_per_cu;
  c->tuned_p4_chunk = bc;
  c->tuned_p4_wgs = bw;
  do_p4(c);  // the state vector of the chosen step configuration
  if (ent)
    for (uint32_t i = 0; i < cap && i < all.size(); i++) ent[i] = all[i];
  if (out) {
    cf1g_tune_result o{};
    o.num_bins = bb;
    o.p3_target_seg = bt;
    o.p4_chunk = bc;
    o.p4_wgs_per_cu = bw;
    o.step_ms = best_step;
    o.p4_cas_ms = best_cas;
    o.n_entries = (uint32_t)all.size();
    *out = o;
  }
}

// ---------------------------------------------------------------- learned index
LidxTrainBufs lidx_bufs(cf1g_ctx* c, uint64_t n_segs) {
  LidxTrainBufs b;
  b.mu = c->getT<float>("lidx_mu", n_segs + 1);
  b.isd = c->getT<float>("lidx_isd", n_segs + 1);
  b.w = c->getT<float>("lidx_w", n_segs + 1);
  b.c = c->getT<float>("lidx_c", n_segs + 1);
  b.gw = c->getT<double>("lidx_gw", n_segs + 1);
  b.gc = c->getT<double>("lidx_gc", n_segs + 1);
  b.loss = c->getT<double>("lidx_loss", n_segs + 1);
  b.resid = c->getT<float>("lidx_resid", c->D + 1);
  b.seg_of_tile = c->getT<uint32_t>("lidx_tile_seg", (c->D + 255) / 256 + 1);
  return b;
}
double lidx_mse(cf1g_ctx* c, const LidxTrainBufs& b, uint64_t n_segs) {
  std::vector<double> l(n_segs);
  c->d2h(l.data(), b.loss, n_segs * 8);
  double t = 0;
  for (double v : l) t += v;
  return c->D ? t / (double)c->D : 0.0;
}
// Stats, endpoint-interpolation start, then epochs of forward / backward /
// update; loss[0] = MSE at the start, loss[1] = after the last update.
void lidx_train_level(cf1g_ctx* c, int level, const SegCtx& sc, uint64_t n_segs, uint32_t epochs, float lr, bool train,
                      double loss[2]) {
  const LidxTrainBufs b = lidx_bufs(c, n_segs);
  launch_lidx_stats_init(c->st, level, sc, n_segs, b);
  launch_lidx_fwd_bwd(c->st, level, sc, n_segs, b);
  loss[0] = loss[1] = lidx_mse(c, b, n_segs);
  if (!train) return;
  for (uint32_t e = 0; e < epochs; e++) {
    launch_lidx_update(c->st, n_segs, lr, b);
    launch_lidx_fwd_bwd(c->st, level, sc, n_segs, b);
  }
  loss[1] = lidx_mse(c, b, n_segs);
}

void do_lidx_train(cf1g_ctx* c, const cf1g_lidx_params* prm, cf1g_lidx_info* info) {
  if (c->wide) fail(CF1G_E_INVALID, "the learned index takes u64 keys (k <= 31)");
  cf1g_lidx_params p{};
  if (prm) p = *prm;
  const uint32_t leaf_bits = p.leaf_bits ? p.leaf_bits : 17, root_bits = p.root_bits ? p.root_bits : 8;
  const uint32_t epochs = p.epochs ? p.epochs : 4, max_window = p.max_window ? p.max_window : 4096;
  const float lr = p.lr > 0 ? p.lr : 0.5f;
  const uint32_t line_bytes = p.line_bytes ? p.line_bytes : 64;
  if (leaf_bits > 26 || root_bits > 20) fail(CF1G_E_INVALID, "leaf_bits <= 26 and root_bits <= 20");
  if (line_bytes != 64 && line_bytes != 128) fail(CF1G_E_INVALID, "line_bytes is 64 or 128");
  const uint32_t C = p.line_keys ? p.line_keys : (line_bytes == 128 ? 256u : 108u);
  if (C < 2 || C > 4096) fail(CF1G_E_INVALID, "line_keys is in [2, 4096]");
  if (c->D >= (1ULL << 32)) fail(CF1G_E_INVALID, "the learned index holds fewer than 2^32 keys");
  hipStream_t st = c->st;
  const uint64_t D = c->D;
  const int K2 = 2 * c->k;
  auto t0 = std::chrono::steady_clock::now();
  // ---- global sort of the distinct keys (LSD, 8-bit digits)
  uint64_t* ka = c->getT<uint64_t>("lidx_keys", D + 1);
  uint64_t* kt = c->getT<uint64_t>("lidx_tmp", D + 1);
  if (D) hip_check(hipMemcpyAsync(ka, dense_keys(c), D * 8, hipMemcpyDeviceToDevice, st), "D2D keys");
  const uint64_t nt = sort_tiles(D);
  uint64_t* mat = c->getT<uint64_t>("lidx_mat", 256 * nt + 1);
  uint64_t* off = c->getT<uint64_t>("lidx_off", 256 * nt + 1);
  uint64_t* tmp = c->getT<uint64_t>("scan_tmp", scan_temp_words(256 * nt + 1) + 16);
  uint64_t* src = ka;
  uint64_t* dst = kt;
  for (int sh = 0; sh < K2 && D; sh += 8) {
    launch_sort_hist(st, src, D, sh, mat);
    scan_excl_u64(st, mat, off, 256 * nt, off + 256 * nt, tmp);
    launch_sort_scatter(st, src, dst, D, sh, off);
    std::swap(src, dst);
  }
  if (src != ka) hip_check(hipMemcpyAsync(ka, src, D * 8, hipMemcpyDeviceToDevice, st), "D2D sorted");
  c->sync();
  auto t1 = std::chrono::steady_clock::now();
  // ---- key warp: byte counts per group -> cumulative tables -> warped keys
  const int G = std::min(WARP_MAX_GROUPS, K2 / 8);
  uint32_t* wcnt = c->getT<uint32_t>("lidx_warp_cnt", WARP_MAX_GROUPS * 256);
  uint32_t* wtab = c->getT<uint32_t>("lidx_warp", WARP_MAX_GROUPS * 256);
  c->zero(wcnt, WARP_MAX_GROUPS * 256 * 4);
  launch_lidx_warp_hist(st, ka, D, K2, G, wcnt);
  std::vector<uint32_t> hcnt(WARP_MAX_GROUPS * 256, 0), htab(WARP_MAX_GROUPS * 256, 0);
  c->d2h(hcnt.data(), wcnt, hcnt.size() * 4);
  for (int g = 0; g < G; g++) {
    // cum_g[s] = floor(2^32 (sum_{t < s} (cnt_t + 1)) / (D + 256)); cum_g[256] = 2^32 is implied
    const unsigned __int128 tot = (unsigned __int128)D + 256;
    unsigned __int128 run = 0;
    for (int x = 0; x < 256; x++) {
      htab[g * 256 + x] = (uint32_t)((run << 32) / tot);
      run += (unsigned __int128)hcnt[g * 256 + x] + 1;
    }
  }
  c->h2d(wtab, htab.data(), htab.size() * 4);
  uint64_t* uk = c->getT<uint64_t>("lidx_u", D + 1);
  launch_lidx_warp_keys(st, ka, D, wtab, K2, G, uk);
  // ---- root pieces by the top bits of the warped keys
  const uint32_t rb = std::max<uint32_t>(1, std::min<uint32_t>(root_bits, 20));
  const uint32_t R = 1u << rb;
  const int s = 64 - (int)rb;
  uint32_t lb = 0;
  while (lb < leaf_bits && (2ULL << lb) <= std::max<uint64_t>(D, 1)) lb++;
  const uint32_t L = 1u << lb, LT = L + R;
  uint64_t* pcs = c->getT<uint64_t>("lidx_p", R + 1);
  uint32_t* lo = c->getT<uint32_t>("lidx_lo", R + 1);
  launch_lidx_pieces(st, uk, D, s, R, L, pcs, lo);
  SegCtx sc0{uk, pcs, lo, D, L, s};
  double rl[2] = {0, 0}, ll[2] = {0, 0};
  lidx_train_level(c, 0, sc0, R, epochs, lr, !p.untrained, rl);
  RootPiece* root = c->getT<RootPiece>("lidx_root", R + 1);
  launch_lidx_root_final(st, R, D, pcs, lo, lidx_bufs(c, R), root);
  // ---- leaves
  uint64_t* q = c->getT<uint64_t>("lidx_q", LT + 1);
  launch_lidx_leaf_bounds(st, uk, D, s, LT, root, q);
  SegCtx sc1{uk, q, nullptr, D, L, s};
  lidx_train_level(c, 1, sc1, LT, epochs, lr, !p.untrained, ll);
  LeafRec* leaf = c->getT<LeafRec>("lidx_leaf", LT + 1);
  uint32_t* err = c->getT<uint32_t>("lidx_err", 2 * (uint64_t)LT + 2);
  uint64_t* sabs = c->getT<uint64_t>("lidx_sabs", LT + 1);
  launch_lidx_leaf_final(st, uk, D, LT, s, q, root, lidx_bufs(c, LT), max_window, leaf, err, sabs);
  // ---- lines
  const uint32_t lw = line_bytes / 8;
  const uint64_t nl = (D + C - 1) / C;
  uint64_t* lines = c->getT<uint64_t>("lidx_lines", nl * lw + 2);
  uint32_t* lst = c->getT<uint32_t>("lidx_line_stats", 2 * nl + 2);
  launch_lidx_lines(st, (int)lw, ka, D, C, nl, lines, lst);
  // ---- host-side summary; fallback table when needed
  std::vector<LeafRec> hl(LT);
  std::vector<uint32_t> he(2 * (size_t)LT);
  std::vector<uint64_t> hs(LT), hq(LT + 1);
  std::vector<uint32_t> hst(2 * nl);
  c->d2h(hl.data(), leaf, (size_t)LT * sizeof(LeafRec));
  c->d2h(he.data(), err, (size_t)LT * 8);
  c->d2h(hs.data(), sabs, (size_t)LT * 8);
  c->d2h(hq.data(), q, ((size_t)LT + 1) * 8);
  if (nl) c->d2h(hst.data(), lst, nl * 8);
  uint64_t fb_leaves = 0, fb_keys = 0, abs_sum = 0, n_over = 0, n_search = 0, bits_sum = 0;
  uint32_t mlo = 0, mhi = 0;
  for (uint32_t j = 0; j < LT; j++) {
    const uint64_t n = hq[j + 1] - hq[j];
    abs_sum += hs[j];
    if (!n) continue;
    uint32_t sbits;
    memcpy(&sbits, &hl[j].S, 4);
    if (sbits >> 31) {
      fb_leaves++;
      fb_keys += n;
    } else {
      mlo = std::max(mlo, he[2 * j]);
      mhi = std::max(mhi, he[2 * j + 1]);
    }
  }
  for (uint64_t i = 0; i < nl; i++) {
    if (hst[2 * i] == 0xFFFFFFFFu) n_over++;
    else bits_sum += hst[2 * i];
    n_search += hst[2 * i + 1];
  }
  // line = pred / C by multiply-high: l = ceil(log2 C), magic = floor(2^32 (2^l - C) / C) + 1
  uint32_t l = 0;
  while ((1ULL << l) < C) l++;
  c->lidx_magic = (uint32_t)((((1ULL << l) - C) << 32) / C + 1);
  c->lidx_mshift = l - 1;
  c->lidx_R = R;
  c->lidx_L = LT;
  c->lidx_Lb = L;
  c->lidx_groups = G;
  c->lidx_s = s;
  c->lidx_C = C;
  c->lidx_lines = (uint32_t)nl;
  c->lidx_lw = lw;
  c->lidx_steps = 3 + (std::max(mlo, mhi) + C - 1) / C;
  c->lidx_fallback = fb_leaves > 0;
  if (c->lidx_fallback) launch_lidx_bbh2rank(st, ka, LT, q, leaf, c->view(), c->getT<uint32_t>("lidx_bbh2rank", D + 1));
  c->sync();
  auto t2 = std::chrono::steady_clock::now();
  c->lidx_ready = true;
  c->lidx_states = false;
  if (info) {
    cf1g_lidx_info o{};
    o.n_keys = D;
    o.n_leaves = LT;
    o.n_pieces = R;
    o.fallback_leaves = fb_leaves;
    o.fallback_keys = fb_keys;
    o.n_lines = nl;
    o.overflow_lines = n_over;
    o.search_keys = n_search;
    o.line_bytes = line_bytes;
    o.line_keys = C;
    o.max_err_lo = mlo;
    o.max_err_hi = mhi;
    o.mean_abs_err = D ? (double)abs_sum / (double)D : 0.0;
    o.line_bits_per_key = D ? (double)bits_sum / (double)D : 0.0;
    o.bytes_total = (uint64_t)(R + 1) * sizeof(RootPiece) + (uint64_t)LT * sizeof(LeafRec) + nl * line_bytes;
    o.root_loss[0] = rl[0];
    o.root_loss[1] = rl[1];
    o.leaf_loss[0] = ll[0];
    o.leaf_loss[1] = ll[1];
    o.sort_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    o.train_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    *info = o;
  }
}

// Per stream position (cf1g_annotate's layout) through the selected index,
// into the device buffer "annot".
uint8_t* annotate_device(cf1g_ctx* c) {
  uint8_t* d = c->getT<uint8_t>("annot", c->n_pos + 8);
  const uint64_t nsw = (c->D + 11) / 12;
  uint64_t* seq = c->getT<uint64_t>("seq", 2 * c->n_words64 + 8);
  uint64_t* nm = c->getT<uint64_t>("nmask", c->n_words64 + 8);
  if (c->use_index == 1) {
    launch_annotate_lidx(c->st, seq, nm, c->n_pos, c->k, c->lview(), c->view(), c->getT<uint64_t>("states_lidx", nsw + 1), d);
  } else if (!c->restored) {
    // by segment: each segment's records (P3 left them grouped, with their
    // stream positions) against the segment's own blocks (no lookup per
    // position); positions without a valid k-mer stay 0
    const uint64_t ns = c->n_segs, nt = p1_tiles(c->n_pos);
    unsigned int* err = c->getT<unsigned int>("annot_err", 2);
    uint64_t* list = c->getT<uint64_t>("p5_list", c->S + 1);
    c->zero(err, 8);
    c->zero(d, c->n_pos);
    launch_p5_list(c->st, c->getT<uint64_t>("sk_entries", nt * p1_tile_slots() + 1),
                   c->getT<uint32_t>("ent_dest", nt * p1_tile_slots() + 1), c->getT<uint32_t>("sk_tile_cnt", nt + 1), nt,
                   (uint32_t)p1_tile_slots(), c->getT<uint32_t>("rec_dest", c->S + 1), c->S, list, err);
    launch_p5_annot(c->st, c->wide, c->getT<uint64_t>("rec_seg", c->S * (uint64_t)c->rw + 1), list,
                    c->getT<uint64_t>("seg_roff", ns + 1), c->getT<uint64_t>("seg_out_off", ns + 1),
                    c->getT<uint64_t>("bbh_blk_off", ns + 1), c->getT<uint64_t>("blocks", c->n_blocks * 8 + 8),
                    c->getT<uint64_t>("states", nsw + 1), c->k, ns, d, err);
    unsigned int he = 0;
    c->d2h(&he, err, 4);
    if (he) fail(CF1G_E_INTERNAL, "P5 annotation: a super k-mer or k-mer missing from its MPHF segment (flags " + std::to_string(he) + ")");
  } else {
    unsigned int* err = c->getT<unsigned int>("annot_err", 2);
    c->zero(err, 8);
    launch_annotate(c->st, c->wide, seq, nm, c->n_pos, c->k, c->view(), c->getT<uint64_t>("states", nsw + 1), c->D, d, err);
    unsigned int he = 0;
    c->d2h(&he, err, 4);
    if (he) fail(CF1G_E_INTERNAL, "a k-mer of the loaded records is missing from the MPHF (a restored file of other records?)");
  }
  return d;
}

// Unitig spans (k_span_* in kernels_p4.hip) into "span_s" / "span_e" (stream
// positions of the first and last k-mer), "span_id" (MPHF value, or rank
// through the learned index when selected) and "span_fl" (bit0 forward,
// bit1 this unitig claims its id), in stream order. Returns the count.
uint64_t build_spans(cf1g_ctx* c) {
  if (c->spans_gen == c->state_gen) return c->spans_U;
  hipStream_t st = c->st;
  uint8_t* ann = annotate_device(c);
  const uint64_t nt = span_tiles(c->n_pos);
  uint64_t* cs = c->getT<uint64_t>("span_cnt_s", nt + 1);
  uint64_t* ce = c->getT<uint64_t>("span_cnt_e", nt + 1);
  uint64_t* os = c->getT<uint64_t>("span_off_s", nt + 1);
  uint64_t* oe = c->getT<uint64_t>("span_off_e", nt + 1);
  uint64_t* tmp = c->getT<uint64_t>("scan_tmp", scan_temp_words(nt + 1) + 16);
  launch_span_count(st, ann, c->n_pos, cs, ce);
  scan_excl_u64(st, cs, os, nt, os + nt, tmp);
  scan_excl_u64(st, ce, oe, nt, oe + nt, tmp);
  uint64_t tot[2] = {0, 0};
  if (nt) {
    hip_check(hipMemcpyAsync(&tot[0], os + nt, 8, hipMemcpyDeviceToHost, st), "D2H spans");
    c->d2h(&tot[1], oe + nt, 8);
  }
  if (tot[0] != tot[1]) fail(CF1G_E_INTERNAL, "unitig starts and ends do not pair up");
  const uint64_t U = tot[0];
  uint64_t* S = c->getT<uint64_t>("span_s", U + 1);
  uint64_t* E = c->getT<uint64_t>("span_e", U + 1);
  uint64_t* ids = c->getT<uint64_t>("span_id", U + 1);
  uint8_t* fl = c->getT<uint8_t>("span_fl", U + 8);
  void* keys = c->get("span_keys", U * c->key_bytes + 16);
  const uint64_t nsw = (c->D + 11) / 12;
  uint64_t* claim = c->getT<uint64_t>("p5_states", nsw + 1);
  unsigned int* err = c->getT<unsigned int>("span_err", 2);
  launch_span_emit(st, ann, c->n_pos, os, oe, S, E);
  launch_span_keys(st, c->wide, c->getT<uint64_t>("seq", 2 * c->n_words64 + 8), c->k, S, E, U, keys, fl);
  if (c->use_index == 1) launch_lidx_lookup(st, (const uint64_t*)keys, U, c->lview(), c->view(), ids, c->lqueue(U));
  else launch_lookup(st, c->wide, keys, U, c->view(), ids, true);
  if (nsw)
    hip_check(hipMemcpyAsync(claim, c->get(c->use_index == 1 ? "states_lidx" : "states", (nsw + 1) * 8), nsw * 8,
                             hipMemcpyDeviceToDevice, st),
              "D2D states");
  c->zero(err, 8);
  launch_span_claim(st, ids, U, c->D, claim, fl, err);
  unsigned int he = 0;
  c->d2h(&he, err, 4);
  if (he) fail(CF1G_E_INTERNAL, "a unitig's k-mer is missing from the index");
  c->spans_U = U;
  c->spans_gen = c->state_gen;
  return U;
}

void check_k(const cf1g_params* p) {
  if (!p) fail(CF1G_E_INVALID, "params is NULL");
  if (p->k < 1 || p->k > 63 || p->k % 2 == 0) fail(CF1G_E_INVALID, "k must be odd and in [1, 63], got " + std::to_string(p->k));
  if (p->num_bins < 1 || p->num_bins > CF1G_MAX_BINS) fail(CF1G_E_INVALID, "num_bins must be in [1, 4096]");
}
}  // namespace

// ====================================================================== C ABI
extern "C" {

int cf1g_create(const cf1g_params* params, cf1g_ctx** out) {
  if (!out) return CF1G_E_INVALID;
  *out = nullptr;
  cf1g_ctx* c = new cf1g_ctx();
  int rc = guarded(c, [&]() {
    check_k(params);
    c->prm = *params;
    c->k = (int)params->k;
    c->m = params->m ? (int)params->m : minimizer_len(c->k);
    if (c->m < 1 || c->m > 15 || c->m > c->k || c->m % 2 == 0) fail(CF1G_E_INVALID, "minimizer length must be odd, <= 15 and <= k");
    if (2 * c->k - c->m > rec_cap(c->k > 32 ? 4 : 2))
      fail(CF1G_E_INVALID, "minimizer length " + std::to_string(c->m) + " is too short for k = " + std::to_string(c->k) +
                               ": super k-mers of up to 2k - m bases would not fit a routed record");
    c->w = c->k - c->m + 1;
    c->wide = c->k > 32;
    c->rw = c->wide ? 4 : 2;
    c->key_bytes = c->wide ? 16 : 8;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) fail(CF1G_E_NODEVICE, "no HIP device visible");
    hip_check(hipSetDevice(0), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(CF1G_E_NODEVICE, std::string("cf1g is built for gfx950, device is ") + prop.gcnArchName);
    c->cu_count = prop.multiProcessorCount;
    hip_check(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking), "hipStreamCreate");
    c->device = 0;
    noalloc_setup(c->device);  // the fixed scratch of cf1g_mphf_lookup_noalloc
  });
  if (rc != CF1G_OK) {
    // keep the context so the caller can read the error
    *out = c;
    return rc;
  }
  *out = c;
  return CF1G_OK;
}

void cf1g_destroy(cf1g_ctx* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto& kv : c->graphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto& kv : c->bufs)
    if (kv.second.p) (void)hipFree(kv.second.p);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->st_copy) (void)hipStreamSynchronize(c->st_copy);
  for (auto& e : c->slot_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->slot_mem) (void)hipHostFree(c->slot_mem);
  if (c->st_copy) (void)hipStreamDestroy(c->st_copy);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

const char* cf1g_last_error(const cf1g_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cf1g_get_info(cf1g_ctx* c, cf1g_info* o) {
  return guarded(c, [&]() {
    if (!o) fail(CF1G_E_INVALID, "out is NULL");
    std::memset(o, 0, sizeof *o);
    o->n_records = c->n_rec;
    o->stream_len = c->n_pos;
    o->n_kmers = c->W;
    o->n_superkmers = c->S;
    o->n_distinct = c->D;
    o->n_state_words = (c->D + 11) / 12;
    o->mphf_words = c->stage >= 4 ? 16 + CF1G_NUM_BUCKETS / 4 + 2 * (c->n_segs + 1) + c->n_blocks * 8 : 0;
    o->num_bins = c->num_bins;
    o->seg_bits = (uint32_t)c->sb;
    o->mphf_big_segments = c->n_big;
    o->n_oversized = c->n_over;
    o->k = (uint32_t)c->k;
    o->m = (uint32_t)c->m;
    o->graph_captures = c->graph_captures;
    o->graph_replays = c->graph_replays;
  });
}

int cf1g_load_records(cf1g_ctx* c, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec) {
  return guarded(c, [&]() {
    if (!c->st) fail(CF1G_E_NODEVICE, "context has no device");
    set_records(c, ascii, rec_off, n_rec);
    c->set_stage(0);
    const uint64_t base = rec_off[0];
    std::vector<uint64_t> off(n_rec + 1);
    for (uint64_t r = 0; r <= n_rec; r++) off[r] = rec_off[r] - base;
    uint8_t* da = c->getT<uint8_t>("ascii", c->n_ascii + 8);
    if (c->n_ascii) hip_check(hipMemcpyAsync(da, ascii + base, c->n_ascii, hipMemcpyHostToDevice, c->st), "H2D ascii");
    hip_check(hipMemcpyAsync(c->getT<uint64_t>("rec_off", n_rec + 1), off.data(), (n_rec + 1) * 8, hipMemcpyHostToDevice, c->st),
              "H2D rec_off");
    hip_check(hipMemcpyAsync(c->getT<uint64_t>("gstart", n_rec + 1), c->gstart_h.data(), (n_rec + 1) * 8, hipMemcpyHostToDevice,
                             c->st),
              "H2D gstart");
    set_chunk_edges(c);
    c->sync();
    c->W = c->S = c->D = 0;
    c->set_stage(1);
  });
}

int cf1g_load_fasta(cf1g_ctx* c, const char* const* paths, uint32_t n_paths, uint32_t threads, cf1g_ingest_stats* out) {
  return guarded(c, [&]() {
    if (!c->st) fail(CF1G_E_NODEVICE, "context has no device");
    if (!paths && n_paths) fail(CF1G_E_INVALID, "paths is NULL");
    std::vector<std::string> ps;
    for (uint32_t i = 0; i < n_paths; i++) {
      if (!paths[i]) fail(CF1G_E_INVALID, "a path is NULL");
      ps.emplace_back(paths[i]);
    }
    const int th = threads ? (int)threads : 1;
    auto t0 = std::chrono::steady_clock::now();
    c->host_records = c->host_packed = c->restored = false;
    try {
      parse_fasta(ps, th, c->host);
    } catch (const std::runtime_error& e) {
      fail(CF1G_E_INVALID, e.what());
    }
    auto t1 = std::chrono::steady_clock::now();
    const HostRecords& R = c->host;
    const uint64_t n_rec = R.off.size() - 1;
    c->n_rec = n_rec;
    c->rec_file = R.file;
    c->n_ascii = R.off[n_rec];
    c->gstart_h.assign(n_rec + 1, 0);
    for (uint64_t r = 0; r <= n_rec; r++) c->gstart_h[r] = R.off[r] + r + 1;
    c->n_pos = c->n_ascii + n_rec + 1;
    c->n_words64 = (c->n_pos + STREAM_PAD + 63) / 64;
    // pack into pinned memory, then one copy of the packed stream
    const size_t bytes = 3 * c->n_words64 * 8;
    if (c->pinned_cap < bytes) {
      if (c->pinned) hip_check(hipHostFree(c->pinned), "hipHostFree");
      c->pinned = nullptr;
      c->pinned_cap = 0;
      hip_check(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault), "hipHostMalloc");
      c->pinned_cap = bytes;
    }
    uint64_t* hseq = (uint64_t*)c->pinned;
    uint64_t* hnm = hseq + 2 * c->n_words64;
    pack_stream(R.ascii.data(), R.off.data(), n_rec, c->n_words64, hseq, hnm, th);
    auto t2 = std::chrono::steady_clock::now();
    uint64_t* seq = c->getT<uint64_t>("seq", 2 * c->n_words64 + 8);
    uint64_t* nm = c->getT<uint64_t>("nmask", c->n_words64 + 8);
    hip_check(hipMemcpyAsync(seq, hseq, 2 * c->n_words64 * 8, hipMemcpyHostToDevice, c->st), "H2D stream");
    hip_check(hipMemcpyAsync(nm, hnm, c->n_words64 * 8, hipMemcpyHostToDevice, c->st), "H2D placeholder bits");
    set_chunk_edges(c);
    c->sync();
    auto t3 = std::chrono::steady_clock::now();
    c->host_records = c->host_packed = true;
    c->W = c->S = c->D = 0;
    c->set_stage(1);
    if (out) {
      auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
      cf1g_ingest_stats o{};
      o.n_records = n_rec;
      o.bases = c->n_ascii;
      o.stream_bytes = bytes;
      o.parse_ms = ms(t0, t1);
      o.pack_ms = ms(t1, t2);
      o.h2d_ms = ms(t2, t3);
      *out = o;
    }
  });
}

int cf1g_set_record_files(cf1g_ctx* c, const uint32_t* file_idx, uint64_t n_rec) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_set_record_files");
    if (n_rec != c->n_rec || (n_rec && !file_idx)) fail(CF1G_E_INVALID, "one file index per loaded record");
    c->rec_file.assign(file_idx, file_idx + n_rec);
  });
}

int cf1g_record_files(cf1g_ctx* c, uint32_t* out, uint64_t n) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_record_files");
    if (n < c->n_rec || (c->n_rec && !out)) fail(CF1G_E_INVALID, "output too small");
    for (uint64_t r = 0; r < c->n_rec; r++) out[r] = r < c->rec_file.size() ? c->rec_file[r] : 0;
  });
}

int cf1g_loaded_records(cf1g_ctx* c, const uint8_t** ascii, const uint64_t** rec_off, uint64_t* n_rec, const char** names,
                        const uint64_t** name_off) {
  return guarded(c, [&]() {
    if (!c->host_records) fail(CF1G_E_ORDER, "cf1g_loaded_records needs records from cf1g_load_fasta");
    if (ascii) *ascii = c->host.ascii.data();
    if (rec_off) *rec_off = c->host.off.data();
    if (n_rec) *n_rec = c->host.off.size() - 1;
    if (names) *names = c->host.names.data();
    if (name_off) *name_off = c->host.name_off.data();
  });
}

int cf1g_stream_export(cf1g_ctx* c, uint64_t* seq, uint64_t* nmask, uint64_t n_words) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_stream_export");
    const uint64_t nw = (c->n_pos + 63) / 64;
    if (n_words < nw) fail(CF1G_E_INVALID, "stream buffers too small");
    if (!c->host_packed) need(c, 2, "cf1g_stream_export of records packed on the device (after P1)");
    if (seq) c->d2h(seq, c->get("seq", (2 * c->n_words64 + 8) * 8), 2 * nw * 8);
    if (nmask) c->d2h(nmask, c->get("nmask", (c->n_words64 + 8) * 8), nw * 8);
  });
}

int cf1g_p1_stream(cf1g_ctx* c, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint32_t threads, uint64_t* out_hist) {
  return guarded(c, [&]() {
    if (!c->st) fail(CF1G_E_NODEVICE, "context has no device");
    do_p1_stream(c, ascii, rec_off, n_rec, threads);
    if (out_hist) c->d2h(out_hist, c->get("hist", CF1G_NUM_BUCKETS * 8), CF1G_NUM_BUCKETS * 8);
  });
}

int cf1g_p1_histogram(cf1g_ctx* c, uint64_t* out_hist) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_p1_histogram");
    do_p1(c);
    if (out_hist) c->d2h(out_hist, c->get("hist", CF1G_NUM_BUCKETS * 8), CF1G_NUM_BUCKETS * 8);
    else c->sync();
  });
}

int cf1g_p2_partition_route(cf1g_ctx* c, const uint64_t* hist, uint32_t num_bins, uint16_t* out_b2b, uint32_t* out_nb) {
  return guarded(c, [&]() {
    need(c, 2, "cf1g_p2_partition_route");
    do_p2(c, hist, num_bins);
    if (out_b2b) c->d2h(out_b2b, c->get("b2b", CF1G_NUM_BUCKETS * 2), CF1G_NUM_BUCKETS * 2);
    else c->sync();
    if (out_nb) *out_nb = c->num_bins;
  });
}

int cf1g_p3_build(cf1g_ctx* c, uint64_t* out_D) {
  return guarded(c, [&]() {
    need(c, 3, "cf1g_p3_build");
    do_p3(c);
    if (out_D) *out_D = c->D;
  });
}

int cf1g_p4_classify(cf1g_ctx* c) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_p4_classify");
    do_p4(c);
  });
}

int cf1g_p4_classify_cas(cf1g_ctx* c, uint32_t flags, cf1g_cas_stats* out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_p4_classify_cas");
    do_p4_cas(c, flags, out);
    c->state_gen++;
  });
}

int cf1g_autotune(cf1g_ctx* c, uint32_t reps, cf1g_tune_entry* entries, uint32_t cap, cf1g_tune_result* out) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_autotune");
    do_autotune(c, reps, entries, cap, out);
  });
}

int cf1g_run_p1_p4(cf1g_ctx* c, uint64_t* out_D) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_run_p1_p4");
    do_p1(c);
    do_p2(c, nullptr, c->prm.num_bins);
    do_p3(c);
    do_p4(c);
    if (out_D) *out_D = c->D;
  });
}

int cf1g_run_p1_p4_stream(cf1g_ctx* c, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint32_t threads,
                          uint64_t* out_D) {
  return guarded(c, [&]() {
    if (!c->st) fail(CF1G_E_NODEVICE, "context has no device");
    do_p1_stream(c, ascii, rec_off, n_rec, threads);
    do_p2(c, nullptr, c->prm.num_bins);
    do_p3(c);
    do_p4(c);
    if (out_D) *out_D = c->D;
  });
}

int cf1g_states_export(cf1g_ctx* c, uint64_t* words, uint64_t n_words) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_states_export");
    uint64_t nsw = (c->D + 11) / 12;
    if (!words || n_words < nsw) fail(CF1G_E_INVALID, "state buffer too small");
    c->d2h(words, c->get(c->use_index == 1 ? "states_lidx" : "states", (nsw + 1) * 8), nsw * 8);
  });
}

int cf1g_keys_export(cf1g_ctx* c, uint64_t* out, uint64_t n_words) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_keys_export");
    uint64_t need_w = c->D * (c->key_bytes / 8);
    if (!out || n_words < need_w) fail(CF1G_E_INVALID, "key buffer too small");
    c->d2h(out, dense_keys(c), need_w * 8);
  });
}

int cf1g_mphf_export(cf1g_ctx* c, uint64_t* buf, uint64_t* n_words) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_mphf_export");
    if (!n_words) fail(CF1G_E_INVALID, "n_words is NULL");
    const uint64_t ns = c->n_segs;
    const uint64_t hdr = 16, tab = CF1G_NUM_BUCKETS / 4;
    const uint64_t total = hdr + tab + 2 * (ns + 1) + c->n_blocks * 8;
    if (!buf) {
      *n_words = total;
      return;
    }
    if (*n_words < total) fail(CF1G_E_INVALID, "MPHF buffer too small");
    std::memset(buf, 0, hdr * 8);
    buf[0] = 0x6366316D70686631ULL;  // "cf1mphf1"
    buf[1] = (uint64_t)c->k;
    buf[2] = (uint64_t)c->m;
    buf[3] = (uint64_t)c->sb;
    buf[4] = c->num_bins;
    buf[5] = ns;
    buf[6] = c->D;
    buf[7] = c->n_blocks;
    buf[8] = c->wide ? 1 : 0;
    MphfView v = c->view();
    hip_check(hipMemcpyAsync(buf + hdr, v.b2b, CF1G_NUM_BUCKETS * 2, hipMemcpyDeviceToHost, c->st), "D2H");
    hip_check(hipMemcpyAsync(buf + hdr + tab, v.key_off, (ns + 1) * 8, hipMemcpyDeviceToHost, c->st), "D2H");
    hip_check(hipMemcpyAsync(buf + hdr + tab + ns + 1, v.blk_off, (ns + 1) * 8, hipMemcpyDeviceToHost, c->st), "D2H");
    if (c->n_blocks)
      hip_check(hipMemcpyAsync(buf + hdr + tab + 2 * (ns + 1), v.blocks, c->n_blocks * 64, hipMemcpyDeviceToHost, c->st), "D2H");
    c->sync();
    *n_words = total;
  });
}

int cf1g_bins_write(cf1g_ctx* c, const char* dir, uint32_t threads, uint64_t* out_bytes) {
  return guarded(c, [&]() {
    need(c, 3, "cf1g_bins_write");
    if (!dir) fail(CF1G_E_INVALID, "dir is NULL");
    std::vector<uint64_t> bstart(c->num_bins + 1);
    c->d2h(bstart.data(), c->get("bin_start", (CF1G_MAX_BINS + 1) * 8), (c->num_bins + 1) * 8);
    std::vector<uint64_t> recs(c->S * (uint64_t)c->rw + 1);
    if (c->S) c->d2h(recs.data(), c->get("rec", (c->S * (uint64_t)c->rw + 1) * 8), c->S * (uint64_t)c->rw * 8);
    uint64_t bytes = 0;
    try {
      bytes = write_bin_files(dir, c->num_bins, c->rw, bstart.data(), recs.data(), threads ? (int)threads : 1);
    } catch (const std::runtime_error& e) {
      fail(CF1G_E_INVALID, e.what());
    }
    if (out_bytes) *out_bytes = bytes;
  });
}

int cf1g_dedup_write(cf1g_ctx* c, const char* path, uint32_t threads) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_dedup_write");
    if (!path) fail(CF1G_E_INVALID, "path is NULL");
    const uint64_t kw = c->key_bytes / 8;
    std::vector<uint64_t> keys(c->D * kw + 2);
    if (c->D) c->d2h(keys.data(), dense_keys(c), c->D * c->key_bytes);
    // bin b's keys are those of its segments [b << sb, (b + 1) << sb)
    std::vector<uint64_t> key_off(c->n_segs + 1);
    c->d2h(key_off.data(), c->get("seg_out_off", (c->n_segs + 1) * 8), (c->n_segs + 1) * 8);
    std::vector<uint64_t> bko(c->num_bins + 1);
    for (uint32_t b = 0; b <= c->num_bins; b++) bko[b] = key_off[(uint64_t)b << c->sb];
    try {
      write_dedup_file(path, keys.data(), c->D, c->wide, bko, threads ? (int)threads : 1);
    } catch (const std::runtime_error& e) {
      fail(CF1G_E_INVALID, e.what());
    }
  });
}

namespace {
constexpr uint64_t SAVE_MAGIC = 0x3176617367316663ULL;  // "cf1gsav1"
constexpr uint64_t SAVE_WORDS = 16;                      // header words
constexpr uint64_t SAVE_VERSION = 2;
// Digest of a run of words (the save body): sum mod 2^64 of
// mix64(word ^ index * golden), as k_stream_digest computes it.
uint64_t words_digest(const uint64_t* w, uint64_t n) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; i++) acc += mix64(w[i] ^ (i * 0x9E3779B97F4A7C15ULL));
  return acc;
}
// Digest of the context's packed stream (packed by P1 for ASCII records).
uint64_t stream_digest(cf1g_ctx* c) {
  auto* d = c->getT<unsigned long long>("stream_digest", 1);
  c->zero(d, 8);
  launch_stream_digest(c->st, c->getT<uint64_t>("seq", 2 * c->n_words64 + 8), c->getT<uint64_t>("nmask", c->n_words64 + 8),
                       c->n_words64, d);
  uint64_t h = 0;
  c->d2h(&h, d, 8);
  return h;
}
// The MPHF words of a save file, checked before anything reaches the device:
// the layout sizes (bounded before any arithmetic), bins of the
// bucket->bin table, and monotone key and block offsets ending at D and the
// block count (a lookup reads only its segment's blocks).
void check_mphf_words(const uint64_t* m, uint64_t mw, int k, int mm, bool wide) {
  auto bad = [](const char* why) { fail(CF1G_E_INVALID, std::string("the save file's MPHF is not valid: ") + why); };
  const uint64_t tab = CF1G_NUM_BUCKETS / 4;
  if (mw < 16 + tab) bad("too short");
  if (m[0] != 0x6366316D70686631ULL || m[1] != (uint64_t)k || m[2] != (uint64_t)mm || m[8] != (wide ? 1u : 0u))
    fail(CF1G_E_INVALID, "the save file's MPHF does not match this context (k, m or layout)");
  const uint64_t sb = m[3], bins = m[4], ns = m[5], D = m[6], nblk = m[7];
  if (sb > 12 || bins < 1 || bins > CF1G_MAX_BINS || ns != (bins << sb) || D >= (1ULL << 40) || nblk >= (1ULL << 40)) bad("header");
  if (mw != 16 + tab + 2 * (ns + 1) + nblk * 8) bad("size");
  const uint16_t* b2b = (const uint16_t*)(m + 16);
  for (uint64_t b = 0; b < CF1G_NUM_BUCKETS; b++)
    if (b2b[b] >= bins) bad("bucket->bin table");
  const uint64_t* ko = m + 16 + tab;
  const uint64_t* bo = ko + ns + 1;
  if (ko[0] != 0 || bo[0] != 0 || ko[ns] != D || bo[ns] != nblk) bad("offsets");
  for (uint64_t s = 0; s < ns; s++)
    if (ko[s + 1] < ko[s] || bo[s + 1] < bo[s]) bad("offsets");
}
}  // namespace

int cf1g_save(cf1g_ctx* c, const char* path) {
  return guarded(c, [&]() {
    need(c, 5, "cf1g_save");
    if (!path) fail(CF1G_E_INVALID, "path is NULL");
    uint64_t mw = 0;
    if (cf1g_mphf_export(c, nullptr, &mw) != CF1G_OK) fail(CF1G_E_INTERNAL, c->err);
    const uint64_t nsw = (c->D + 11) / 12;
    std::vector<uint64_t> buf(SAVE_WORDS + mw + nsw, 0);
    buf[0] = SAVE_MAGIC;
    buf[1] = SAVE_VERSION;
    buf[2] = c->n_rec;
    buf[3] = c->n_pos;
    buf[4] = c->W;
    buf[5] = c->S;
    buf[6] = mw;
    buf[7] = nsw;
    buf[8] = stream_digest(c);
    if (cf1g_mphf_export(c, buf.data() + SAVE_WORDS, &mw) != CF1G_OK) fail(CF1G_E_INTERNAL, c->err);
    if (nsw) c->d2h(buf.data() + SAVE_WORDS + mw, c->get("states", (nsw + 1) * 8), nsw * 8);
    buf[9] = words_digest(buf.data() + SAVE_WORDS, mw + nsw);
    FILE* f = std::fopen(path, "wb");
    if (!f) fail(CF1G_E_INVALID, std::string("cannot create ") + path);
    const size_t n = buf.size() * 8;
    const bool ok = std::fwrite(buf.data(), 1, n, f) == n;
    if (std::fclose(f) != 0 || !ok) fail(CF1G_E_INVALID, std::string("write failed: ") + path);
  });
}

int cf1g_restore(cf1g_ctx* c, const char* path) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_restore");
    if (!path) fail(CF1G_E_INVALID, "path is NULL");
    FILE* f = std::fopen(path, "rb");
    if (!f) fail(CF1G_E_INVALID, std::string("cannot open ") + path);
    std::vector<uint64_t> buf;
    uint64_t hdr[SAVE_WORDS];
    bool ok = std::fread(hdr, 8, SAVE_WORDS, f) == SAVE_WORDS && hdr[0] == SAVE_MAGIC && hdr[1] == SAVE_VERSION;
    if (ok) {  // the header's sizes must match the file before anything is allocated
      const long here = std::ftell(f);
      ok = std::fseek(f, 0, SEEK_END) == 0;
      const long size = ok ? std::ftell(f) : -1;
      ok = ok && here >= 0 && size >= here && std::fseek(f, here, SEEK_SET) == 0 && hdr[6] < (1ULL << 40) && hdr[7] < (1ULL << 40) &&
           (uint64_t)(size - here) == 8 * (hdr[6] + hdr[7]);
    }
    if (ok) {
      buf.resize(hdr[6] + hdr[7]);
      ok = std::fread(buf.data(), 8, buf.size(), f) == buf.size();
    }
    std::fclose(f);
    if (!ok) fail(CF1G_E_INVALID, std::string("not a complete cf1g save file: ") + path);
    if (words_digest(buf.data(), buf.size()) != hdr[9]) fail(CF1G_E_INVALID, "the save file is damaged (body checksum)");
    if (hdr[2] != c->n_rec || hdr[3] != c->n_pos) fail(CF1G_E_INVALID, "the save file belongs to other records (count or length differ)");
    const uint64_t* m = buf.data();
    const uint64_t mw = hdr[6], nsw = hdr[7];
    check_mphf_words(m, mw, c->k, c->m, c->wide);
    const uint64_t ns = m[5], D = m[6], nblk = m[7], tab = CF1G_NUM_BUCKETS / 4;
    if (nsw != (D + 11) / 12) fail(CF1G_E_INVALID, "the save file's state vector does not match its MPHF");
    if (!c->host_packed) {  // ASCII records: the stream annotation and P5 read is packed in P1
      launch_pack(c->st, c->getT<uint8_t>("ascii", c->n_ascii + 8), c->getT<uint64_t>("rec_off", c->n_rec + 1),
                  c->getT<uint64_t>("gstart", c->n_rec + 1), c->n_rec, c->n_words64, c->getT<uint64_t>("seq", 2 * c->n_words64 + 8),
                  c->getT<uint64_t>("nmask", c->n_words64 + 8));
    }
    if (stream_digest(c) != hdr[8]) fail(CF1G_E_INVALID, "the save file belongs to other records (stream digest differs)");
    c->sb = (int)m[3];
    c->num_bins = (uint32_t)m[4];
    c->n_segs = ns;
    c->D = D;
    c->n_blocks = nblk;
    c->W = hdr[4];
    c->S = hdr[5];
    MphfView v = c->view();
    c->h2d((void*)v.b2b, m + 16, CF1G_NUM_BUCKETS * 2);
    c->h2d((void*)v.key_off, m + 16 + tab, (ns + 1) * 8);
    c->h2d((void*)v.blk_off, m + 16 + tab + ns + 1, (ns + 1) * 8);
    if (nblk) c->h2d((void*)v.blocks, m + 16 + tab + 2 * (ns + 1), nblk * 64);
    uint64_t* states = c->getT<uint64_t>("states", nsw + 1);
    if (nsw) c->h2d(states, m + mw, nsw * 8);
    c->keys_ready = c->lidx_ready = c->lidx_states = false;
    c->use_index = 0;
    c->state_gen++;
    c->restored = true;
    c->set_stage(5);
  });
}

int cf1g_mphf_lookup(cf1g_ctx* c, const uint64_t* keys, uint64_t n, uint64_t* out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_mphf_lookup");
    if (n == 0) return;
    if (!keys || !out) fail(CF1G_E_INVALID, "NULL buffer");
    size_t kb = c->key_bytes;
    void* dk = c->get("lookup_keys", n * kb);
    uint64_t* dout = c->getT<uint64_t>("lookup_out", n);
    hip_check(hipMemcpyAsync(dk, keys, n * kb, hipMemcpyHostToDevice, c->st), "H2D keys");
    launch_lookup(c->st, c->wide, dk, n, c->view(), dout, false);
    c->d2h(out, dout, n * 8);
  });
}

int cf1g_mphf_lookup_noalloc(cf1g_ctx* c, const uint64_t* keys, uint64_t n, uint64_t* out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_mphf_lookup_noalloc");
    if (n == 0) return;
    if (!keys || !out) fail(CF1G_E_INVALID, "NULL buffer");
    size_t cap = 0;
    uint8_t* s = (uint8_t*)noalloc_scratch(c->device, &cap);
    if (!s) fail(CF1G_E_INTERNAL, "the fixed scratch of this device was not set up");
    const uint64_t kb = c->key_bytes;
    const uint64_t max_n = cap / (kb + 8) - 64;
    if (n > max_n)
      fail(CF1G_E_INVALID, "cf1g_mphf_lookup_noalloc takes at most " + std::to_string(max_n) + " keys per call, got " + std::to_string(n));
    // the MPHF's buffers exist since P3 (or a restore); nothing is (re)allocated
    auto have = [&](const char* name, uint64_t bytes) -> void* {
      auto it = c->bufs.find(name);
      if (it == c->bufs.end() || !it->second.p || it->second.cap < bytes) fail(CF1G_E_ORDER, std::string("no MPHF buffer ") + name);
      return it->second.p;
    };
    MphfView v;
    v.b2b = (const uint16_t*)have("b2b", CF1G_NUM_BUCKETS * 2);
    v.key_off = (const uint64_t*)have("seg_out_off", (c->n_segs + 1) * 8);
    v.blk_off = (const uint64_t*)have("bbh_blk_off", (c->n_segs + 1) * 8);
    v.blocks = (const uint64_t*)have("blocks", (c->n_blocks * 8 + 8) * 8);
    v.K = c->k;
    v.m = c->m;
    v.sb = c->sb;
    uint8_t* dk = s;                                               // keys
    uint64_t* dout = (uint64_t*)(s + ((n * kb + 255) & ~uint64_t(255)));  // values
    hip_check(hipMemcpyAsync(dk, keys, n * kb, hipMemcpyHostToDevice, c->st), "H2D keys");
    launch_lookup(c->st, c->wide, dk, n, v, dout, false);
    hip_check(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, c->st), "D2H values");
    c->sync();
  });
}

int cf1g_mphf_lookup_device(cf1g_ctx* c, const void* d_keys, uint64_t n, uint64_t* d_out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_mphf_lookup_device");
    launch_lookup(c->st, c->wide, d_keys, n, c->view(), d_out, false);
    c->sync();
  });
}

int cf1g_annotate(cf1g_ctx* c, uint8_t* out, uint64_t n) {
  return guarded(c, [&]() {
    need(c, 5, "cf1g_annotate");
    if (!out || n < c->n_pos) fail(CF1G_E_INVALID, "annotation buffer too small");
    c->d2h(out, annotate_device(c), c->n_pos);
  });
}

int cf1g_unitig_spans(cf1g_ctx* c, uint64_t* n_spans, uint64_t* start, uint64_t* end, uint64_t* id, uint8_t* flags, uint64_t cap) {
  return guarded(c, [&]() {
    need(c, 5, "cf1g_unitig_spans");
    if (!n_spans) fail(CF1G_E_INVALID, "n_spans is NULL");
    const uint64_t U = build_spans(c);
    *n_spans = U;
    if (!start && !end && !id && !flags) return;
    if (cap < U) fail(CF1G_E_INVALID, "span buffers too small");
    if (!U) return;
    if (start) c->d2h(start, c->get("span_s", (U + 1) * 8), U * 8);
    if (end) c->d2h(end, c->get("span_e", (U + 1) * 8), U * 8);
    if (id) c->d2h(id, c->get("span_id", (U + 1) * 8), U * 8);
    if (flags) c->d2h(flags, c->get("span_fl", U + 8), U);
  });
}

int cf1g_record_starts(cf1g_ctx* c, uint64_t* out, uint64_t n) {
  return guarded(c, [&]() {
    need(c, 1, "cf1g_record_starts");
    if (!out || n < c->n_rec + 1) fail(CF1G_E_INVALID, "buffer too small");
    std::memcpy(out, c->gstart_h.data(), (c->n_rec + 1) * 8);
  });
}

int cf1g_bins_export(cf1g_ctx* c, uint32_t* rec_words, uint64_t* bin_start, uint64_t* records, uint64_t n_words) {
  return guarded(c, [&]() {
    need(c, 3, "cf1g_bins_export");
    if (rec_words) *rec_words = (uint32_t)c->rw;
    if (bin_start) c->d2h(bin_start, c->get("bin_start", (CF1G_MAX_BINS + 1) * 8), (c->num_bins + 1) * 8);
    if (records) {
      uint64_t nw = c->S * (uint64_t)c->rw;
      if (n_words < nw) fail(CF1G_E_INVALID, "record buffer too small");
      c->d2h(records, c->get("rec", nw * 8 + 8), nw * 8);
    }
  });
}

int cf1g_set_graphs(cf1g_ctx* c, int on) {
  return guarded(c, [&]() { c->graphs_on = on != 0; });
}

int cf1g_sync(cf1g_ctx* c) {
  return guarded(c, [&]() { c->sync(); });
}

int cf1g_set_profiling(cf1g_ctx* c, int on) {
  return guarded(c, [&]() {
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_on = on != 0;
  });
}

size_t cf1g_kernel_times(cf1g_ctx* c, char* buf, size_t cap, int reset) {
  if (!c) return 0;
  std::string txt;
  try {
    c->sync();
    std::lock_guard<std::mutex> g(g_prof_mu);
    std::map<std::string, std::pair<uint64_t, double>> agg;
    std::vector<std::string> order;
    for (auto& r : g_prof) {
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
      if (!agg.count(r.name)) order.push_back(r.name);
      agg[r.name].first += 1;
      agg[r.name].second += ms;
    }
    for (auto& nme : order) {
      char line[256];
      std::snprintf(line, sizeof line, "%s %llu %.6f\n", nme.c_str(), (unsigned long long)agg[nme].first, agg[nme].second);
      txt += line;
    }
    if (reset) {
      for (auto& r : g_prof) {
        g_ev_pool.push_back(r.a);
        g_ev_pool.push_back(r.b);
      }
      g_prof.clear();
    }
  } catch (const std::exception& e) {
    c->err = e.what();
    return 0;
  }
  if (buf && cap > 0) {
    size_t n = std::min(cap - 1, txt.size());
    std::memcpy(buf, txt.data(), n);
    buf[n] = 0;
  }
  return txt.size() + 1;
}

const void* cf1g_device_keys(cf1g_ctx* c) {
  if (!c || c->stage < 4) return nullptr;
  const void* p = nullptr;
  if (guarded(c, [&]() { p = dense_keys(c); }) != CF1G_OK) return nullptr;
  return p;
}

// ---------------------------------------------------------------- learned index
int cf1g_lidx_train(cf1g_ctx* c, const cf1g_lidx_params* params, cf1g_lidx_info* out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_lidx_train");
    do_lidx_train(c, params, out);
  });
}

int cf1g_lidx_lookup(cf1g_ctx* c, const uint64_t* keys, uint64_t n, uint64_t* out) {
  return guarded(c, [&]() {
    if (!c->lidx_ready) fail(CF1G_E_ORDER, "cf1g_lidx_lookup called before cf1g_lidx_train");
    if (n == 0) return;
    if (!keys || !out) fail(CF1G_E_INVALID, "NULL buffer");
    uint64_t* dk = c->getT<uint64_t>("lookup_in", n + 1);
    uint64_t* dout = c->getT<uint64_t>("lookup_out", n + 1);
    c->h2d(dk, keys, n * 8);
    launch_lidx_lookup(c->st, dk, n, c->lview(), c->view(), dout, c->lqueue(n));
    c->d2h(out, dout, n * 8);
  });
}

int cf1g_lidx_lookup_device(cf1g_ctx* c, const void* d_keys, uint64_t n, uint64_t* d_out) {
  return guarded(c, [&]() {
    if (!c->lidx_ready) fail(CF1G_E_ORDER, "cf1g_lidx_lookup_device called before cf1g_lidx_train");
    if (n == 0) return;
    if (!d_keys || !d_out) fail(CF1G_E_INVALID, "NULL buffer");
    launch_lidx_lookup(c->st, (const uint64_t*)d_keys, n, c->lview(), c->view(), d_out, c->lqueue(n));
    c->sync();
  });
}

int cf1g_lidx_export(cf1g_ctx* c, cf1g_lidx_shape* shape, uint64_t* sorted_keys, uint32_t* warp, void* root, void* leaves,
                     uint32_t* leaf_err, void* lines) {
  return guarded(c, [&]() {
    if (!c->lidx_ready) fail(CF1G_E_ORDER, "cf1g_lidx_export called before cf1g_lidx_train");
    if (shape) {
      *shape = cf1g_lidx_shape{};
      shape->n_keys = c->D;
      shape->n_pieces = c->lidx_R;
      shape->n_leaves = c->lidx_L;
      shape->root_shift = c->lidx_s;
      shape->line_bytes = c->lidx_lw * 8;
      shape->line_keys = c->lidx_C;
      shape->n_lines = c->lidx_lines;
      shape->key_bits = 2 * c->k;
      shape->warp_groups = (uint32_t)c->lidx_groups;
    }
    const LidxView v = c->lview();
    if (sorted_keys && c->D) c->d2h(sorted_keys, v.keys, c->D * 8);
    if (warp && c->lidx_groups) c->d2h(warp, v.warp, (size_t)c->lidx_groups * 256 * 4);
    if (root) c->d2h(root, v.root, ((size_t)c->lidx_R + 1) * sizeof(RootPiece));
    if (leaves) c->d2h(leaves, v.leaf, (size_t)c->lidx_L * sizeof(LeafRec));
    if (leaf_err) c->d2h(leaf_err, c->getT<uint32_t>("lidx_err", 2 * (uint64_t)c->lidx_L + 2), (size_t)c->lidx_L * 8);
    if (lines && c->lidx_lines) c->d2h(lines, v.lines, (size_t)c->lidx_lines * c->lidx_lw * 8);
  });
}

int cf1g_lidx_gradients(cf1g_ctx* c, int level, const float* w, const float* cc, uint64_t n_segs, double* out_gw, double* out_gc,
                        double* out_sq, uint64_t* out_seg, float* out_mu, float* out_isd) {
  return guarded(c, [&]() {
    if (!c->lidx_ready) fail(CF1G_E_ORDER, "cf1g_lidx_gradients called before cf1g_lidx_train");
    if (level != 0 && level != 1) fail(CF1G_E_INVALID, "level is 0 (root) or 1 (leaves)");
    const uint64_t ns = level == 0 ? c->lidx_R : c->lidx_L;
    if (n_segs != ns || !w || !cc) fail(CF1G_E_INVALID, "one (w, c) pair per segment is required");
    uint64_t* ka = c->getT<uint64_t>("lidx_u", c->D + 1);
    SegCtx sc{ka, level == 0 ? c->getT<uint64_t>("lidx_p", ns + 1) : c->getT<uint64_t>("lidx_q", ns + 1),
              level == 0 ? c->getT<uint32_t>("lidx_lo", ns + 1) : nullptr, c->D, c->lidx_Lb, c->lidx_s};
    const LidxTrainBufs b = lidx_bufs(c, ns);
    launch_lidx_stats_init(c->st, level, sc, ns, b);
    c->h2d(b.w, w, ns * 4);
    c->h2d(b.c, cc, ns * 4);
    launch_lidx_fwd_bwd(c->st, level, sc, ns, b);
    if (out_gw) c->d2h(out_gw, b.gw, ns * 8);
    if (out_gc) c->d2h(out_gc, b.gc, ns * 8);
    if (out_sq) c->d2h(out_sq, b.loss, ns * 8);
    if (out_seg) c->d2h(out_seg, sc.seg, (ns + 1) * 8);
    if (out_mu) c->d2h(out_mu, b.mu, ns * 4);
    if (out_isd) c->d2h(out_isd, b.isd, ns * 4);
  });
}

int cf1g_classify_via(cf1g_ctx* c, int index, cf1g_cas_stats* out) {
  return guarded(c, [&]() {
    need(c, 4, "cf1g_classify_via");
    if (c->wide) fail(CF1G_E_INVALID, "classification through an index takes u64 keys (k <= 31)");
    if (index == 1 && !c->lidx_ready) fail(CF1G_E_ORDER, "cf1g_classify_via(1) called before cf1g_lidx_train");
    if (index != 0 && index != 1) fail(CF1G_E_INVALID, "index is 0 (BBHash) or 1 (learned)");
    do_p4_cas_via(c, 0, out, index);
    if (index == 1) c->lidx_states = true;
  });
}

int cf1g_select_index(cf1g_ctx* c, int index) {
  return guarded(c, [&]() {
    if (index == 0) {
      c->use_index = 0;
      c->state_gen++;
      return;
    }
    if (index != 1) fail(CF1G_E_INVALID, "index is 0 (BBHash) or 1 (learned)");
    if (!c->lidx_states) fail(CF1G_E_ORDER, "cf1g_select_index(1) needs cf1g_classify_via(ctx, 1)");
    c->use_index = 1;
    c->state_gen++;
  });
}
int cf1g_selected_index(const cf1g_ctx* c) { return c ? c->use_index : 0; }

}  // extern "C"
