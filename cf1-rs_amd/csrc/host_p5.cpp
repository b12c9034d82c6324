// Host P5 of the cf1g product: unitig extraction and GFA-reduced output
// (src/output.rs:152-834) over the products of the GPU phases.
//
// Where the reference re-looks-up every k-mer's MPHF slot and state on the
// host (AnnotatedKmer, src/directed_kmer.rs:81-109), this P5 reads one byte
// per stream position computed on the GPU (valid, direction, state class)
// and only looks up the two flanking k-mers of each unitig in the exported
// BBHash (src/output.rs:728-731). Unitig claiming uses the exported state
// words with the same outputted-state CAS (src/output.rs:747-751).
#include <atomic>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cf1g.h"
#include "host.hpp"

using namespace cf1g;

namespace {

inline uint8_t outputted_state(uint8_t s) {  // OUTPUTTED_TABLE, src/state.rs:272-298
  if (s == 3 || s == 15) return 15;
  if ((s >= 4 && s <= 7) || s == 13) return 13;
  if ((s >= 8 && s <= 11) || s == 14) return 14;
  if (s == 12 || s >= 16) return 12;
  return 0;
}

struct Meta {  // UnipathsMeta, src/output.rs:34-75
  uint64_t count = 0, kmers = 0, maxl = 0, minl = UINT64_MAX, suml = 0;
  void add(uint64_t vc, int k) {
    count++;
    kmers += vc;
    uint64_t l = vc + (uint64_t)k - 1;
    if (l > maxl) maxl = l;
    if (l < minl) minl = l;
    suml += l;
  }
  void agg(const Meta& o) {
    count += o.count;
    kmers += o.kmers;
    if (o.maxl > maxl) maxl = o.maxl;
    if (o.minl < minl) minl = o.minl;
    suml += o.suml;
  }
};

struct OU {
  uint64_t id;
  bool fwd;
  uint64_t s, e;
};

enum : uint8_t { C_SISO = 0, C_MISO = 1, C_SIMO = 2, C_MIMO = 3 };

struct AK {  // an annotated k-mer position
  uint64_t idx;
  uint8_t cls;
  bool fwd;
};

bool unipath_start(const AK& c, const AK& p) {  // src/output.rs:79-110
  if (c.cls == C_MIMO) return true;
  if (c.fwd ? c.cls == C_MISO : c.cls == C_SIMO) return true;
  if (p.cls == C_MIMO) return true;
  if (p.fwd ? p.cls == C_SIMO : p.cls == C_MISO) return true;
  return false;
}
bool unipath_end(const AK& c, const AK& n) {  // src/output.rs:114-145
  if (c.cls == C_MIMO) return true;
  if (c.fwd ? c.cls == C_SIMO : c.cls == C_MISO) return true;
  if (n.cls == C_MIMO) return true;
  if (n.fwd ? n.cls == C_MISO : n.cls == C_SIMO) return true;
  return false;
}
inline bool placeholder(uint8_t c) { return h_code(c) > 3; }

struct P5 {
  int K;
  const uint8_t* annot;  // per stream position
  HostMphf mphf;
  const HostLidx* lidx = nullptr;  // lookups through the learned index when set
  std::atomic<uint64_t>* states;

  AK at(uint64_t g0, uint64_t idx) const {
    uint8_t a = annot[g0 + idx];
    return AK{idx, (uint8_t)((a >> 2) & 3), (a & 2) != 0};
  }
  uint8_t get_state(uint64_t i) const { return (uint8_t)((states[i / 12].load(std::memory_order_acquire) >> ((i % 12) * 5)) & 31); }
  bool cas_state(uint64_t i, uint8_t expected, uint8_t nv) {  // src/state_vector.rs:47-70
    uint64_t sh = (i % 12) * 5;
    std::atomic<uint64_t>& w = states[i / 12];
    for (;;) {
      uint64_t cur = w.load(std::memory_order_acquire);
      if (((cur >> sh) & 31) != expected) return false;
      uint64_t nw = (cur & ~(31ULL << sh)) | ((uint64_t)nv << sh);
      if (w.compare_exchange_weak(cur, nw, std::memory_order_acq_rel, std::memory_order_relaxed)) return true;
    }
  }

  // output_gfa_unitig, src/output.rs:714-782
  void unitig(const uint8_t* seq, const AK& st, const AK& en, std::string& seg, std::vector<OU>& us, Meta& mt) {
    u128 xs = h_encode(seq + st.idx, K), xe = h_encode(seq + en.idx, K);
    u128 rs = h_rc(xs, K), re = h_rc(xe, K);
    u128 cs = xs < rs ? xs : rs, ce = xe < re ? xe : re;
    u128 mf = cs < ce ? cs : ce;
    uint64_t id = lidx ? lidx->lookup((uint64_t)mf) : mphf.lookup(mf);
    bool fwd = xs < re;
    us.push_back({id, fwd, st.idx, en.idx});
    uint8_t s = get_state(id);
    if (!(s >= 12 && s <= 15)) {
      if (cas_state(id, s, outputted_state(s))) {
        uint64_t L = en.idx - st.idx + (uint64_t)K;
        char num[32];
        int nn = std::snprintf(num, sizeof num, "%llu\t", (unsigned long long)id);
        seg.append(num, (size_t)nn);
        if (fwd) {
          for (uint64_t t = 0; t < L; t++) {
            uint8_t c = seq[st.idx + t];
            seg.push_back((char)(c <= 'T' ? c : c - ('a' - 'A')));
          }
        } else {
          for (uint64_t t = 0; t < L; t++) {
            uint8_t c = seq[en.idx + (uint64_t)K - 1 - t];
            seg.push_back("TGCA"[h_code(c) & 3]);
          }
        }
        seg.push_back('\n');
        mt.add(en.idx - st.idx + 1, K);
      }
    }
  }

  uint64_t search_valid(const uint8_t* seq, uint64_t start, uint64_t right_end) const {  // src/output.rs:484-501
    while (start <= right_end) {
      bool ok = true;
      for (int j = 0; j < K; j++)
        if (placeholder(seq[start + j])) {
          start = start + j + 1;
          ok = false;
          break;
        }
      if (ok) return start;
    }
    return right_end + 1;
  }

  // output_unitigs_contiguous, src/output.rs:506-709
  uint64_t contiguous(const uint8_t* seq, uint64_t g0, uint64_t len, uint64_t right_end, uint64_t start_idx, std::string& seg,
                      std::vector<OU>& us, Meta& mt) {
    uint64_t i = start_idx;
    AK curr = at(g0, i);
    bool no_left = i == 0 || placeholder(seq[i - 1]);
    bool no_right = i + K == len || placeholder(seq[i + K]);
    if (no_left && no_right) {
      unitig(seq, curr, curr, seg, us, mt);
      return i + K;
    }
    if (no_right) {
      AK prev = at(g0, i - 1);
      if (unipath_start(curr, prev)) unitig(seq, curr, curr, seg, us, mt);
      return i + K;
    }
    AK next = at(g0, i + 1);
    bool on = false;
    AK ustart = curr;
    if (no_left) {
      on = true;
    } else {
      AK prev = at(g0, i - 1);
      if (unipath_start(curr, prev)) on = true;
    }
    if (on && unipath_end(curr, next)) {
      unitig(seq, ustart, curr, seg, us, mt);
      on = false;
    }
    i += 1;
    for (;;) {
      if (!on && i > right_end) break;
      AK prev = curr;
      curr = next;
      if (unipath_start(curr, prev)) {
        on = true;
        ustart = curr;
      }
      if (i + K == len || placeholder(seq[i + K])) {
        if (on) unitig(seq, ustart, curr, seg, us, mt);
        return i + K;
      }
      next = at(g0, i + 1);
      if (on && unipath_end(curr, next)) {
        unitig(seq, ustart, curr, seg, us, mt);
        on = false;
      }
      i += 1;
    }
    return i + K;
  }

  // extract_unitigs_from_chunk, src/output.rs:439-482
  void chunk(const uint8_t* seq, uint64_t g0, uint64_t len, uint64_t cs, uint64_t cre, std::string& seg, std::vector<OU>& us,
             Meta& mt) {
    uint64_t abs_re = len - K;
    uint64_t i = cs;
    while (i <= abs_re) {
      i = search_valid(seq, i, cre);
      if (i > cre) break;
      i = contiguous(seq, g0, len, cre, i, seg, us, mt);
      if (i > cre) break;
    }
  }
};

void par_for(size_t n, int threads, const std::function<void(size_t)>& f) {
  if (threads <= 1 || n <= 1) {
    for (size_t i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<size_t> nx{0};
  std::vector<std::thread> th;
  int nt = (int)std::min<size_t>((size_t)threads, n);
  for (int t = 0; t < nt; t++)
    th.emplace_back([&]() {
      for (;;) {
        size_t i = nx.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
    });
  for (auto& x : th) x.join();
}

std::string json_escape(const std::string& s) {
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

}  // namespace

extern "C" int cf1g_p5_write(cf1g_ctx* ctx, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, const char* names_blob,
                             const uint64_t* name_off, const char* out_prefix, const char* input_desc, int poly_n_stretch,
                             int track_short, int threads, uint64_t* out_meta5) {
  if (!ctx || !rec_off || !out_prefix) return CF1G_E_INVALID;
  cf1g_info info;
  int rc = cf1g_get_info(ctx, &info);
  if (rc) return rc;
  if (info.n_records != n_rec) return CF1G_E_INVALID;
  const int K = (int)info.k;
  std::vector<uint8_t> annot(info.stream_len);
  if ((rc = cf1g_annotate(ctx, annot.data(), annot.size()))) return rc;
  std::vector<uint64_t> words(info.n_state_words + 1);
  if ((rc = cf1g_states_export(ctx, words.data(), words.size()))) return rc;
  uint64_t mw = 0;
  if ((rc = cf1g_mphf_export(ctx, nullptr, &mw))) return rc;
  std::vector<uint64_t> mbuf(mw);
  if ((rc = cf1g_mphf_export(ctx, mbuf.data(), &mw))) return rc;
  std::vector<uint64_t> gstart(n_rec + 1);
  if ((rc = cf1g_record_starts(ctx, gstart.data(), gstart.size()))) return rc;
  // learned index selected: unitig ids are ranks in the sorted key order
  HostLidx hl;
  std::vector<uint64_t> lkeys;
  std::vector<RootPiece> lroot;
  std::vector<LeafRec> lleaf;
  const bool learned = cf1g_selected_index(ctx) == 1;
  if (learned) {
    cf1g_lidx_shape sh;
    if ((rc = cf1g_lidx_export(ctx, &sh, nullptr, nullptr, nullptr))) return rc;
    lkeys.resize(sh.n_keys + 1);
    lroot.resize(sh.n_pieces);
    lleaf.resize(sh.n_leaves);
    if ((rc = cf1g_lidx_export(ctx, nullptr, lkeys.data(), lroot.data(), lleaf.data()))) return rc;
    hl.keys = lkeys.data();
    hl.D = sh.n_keys;
    hl.root = lroot.data();
    hl.leaf = lleaf.data();
    hl.L = sh.n_leaves;
    hl.R = sh.n_pieces;
    hl.s = sh.root_shift;
  }

  std::vector<std::atomic<uint64_t>> st(words.size());
  for (size_t i = 0; i < words.size(); i++) st[i].store(words[i], std::memory_order_relaxed);
  P5 p5;
  p5.K = K;
  p5.annot = annot.data();
  p5.mphf.attach(mbuf.data());
  if (learned) p5.lidx = &hl;
  p5.states = st.data();

  const uint64_t base = rec_off[0];
  auto name_of = [&](uint64_t r) -> std::string {
    if (!names_blob || !name_off) return std::string();
    return std::string(names_blob + name_off[r], names_blob + name_off[r + 1]);
  };
  std::vector<std::string> segs(n_rec), tiles(n_rec);
  std::vector<Meta> metas(n_rec);
  std::vector<char> present(n_rec, 0);
  const int nth = threads < 1 ? 1 : threads;
  auto tiling = [&](uint64_t r, const std::vector<OU>& us) {  // write_tiling, src/output.rs:208-259
    std::string t = "Reference:1_Sequence:" + name_of(r) + "\t";
    if (!us.empty()) {
      if (poly_n_stretch && us[0].s > 0) t += "N" + std::to_string(us[0].s) + " ";
      t += std::to_string(us[0].id);
      t.push_back(us[0].fwd ? '+' : '-');
      for (size_t i = 1; i < us.size(); i++) {
        const OU& l = us[i - 1];
        const OU& q = us[i];
        if (poly_n_stretch && l.e + 1 != q.s) {
          int64_t gap = (int64_t)q.s - (int64_t)l.e;
          if (gap >= (int64_t)K + 1) t += " N" + std::to_string((uint64_t)gap - (uint64_t)K);
        }
        t.push_back(' ');
        t += std::to_string(q.id);
        t.push_back(q.fwd ? '+' : '-');
      }
    }
    t.push_back('\n');
    return t;
  };
  constexpr uint64_t LARGE = 4000000;  // LARGE_SEQ_THRESHOLD, src/output.rs:205
  std::vector<uint64_t> small, large;
  for (uint64_t r = 0; r < n_rec; r++) {
    uint64_t len = rec_off[r + 1] - rec_off[r];
    if (len < (uint64_t)K) continue;
    present[r] = 1;
    (len < LARGE ? small : large).push_back(r);
  }
  try {
    par_for(small.size(), nth, [&](size_t i) {
      uint64_t r = small[i];
      const uint8_t* seq = ascii + (rec_off[r] - base);
      uint64_t len = rec_off[r + 1] - rec_off[r];
      std::vector<OU> us;
      p5.chunk(seq, gstart[r], len, 0, len - K, segs[r], us, metas[r]);
      tiles[r] = tiling(r, us);
    });
    for (uint64_t r : large) {
      const uint8_t* seq = ascii + (rec_off[r] - base);
      uint64_t len = rec_off[r + 1] - rec_off[r];
      uint64_t total = len - K + 1;  // compute_chunk_boundaries, src/output.rs:406-423
      std::vector<std::pair<uint64_t, uint64_t>> ch;
      for (uint64_t i = 0; i < (uint64_t)nth; i++) {
        uint64_t s = i * total / nth, e = (i + 1) * total / nth;
        if (s < e) ch.push_back({s, e - 1});
      }
      std::vector<std::vector<OU>> cu(ch.size());
      std::vector<std::string> cs(ch.size());
      std::vector<Meta> cm(ch.size());
      par_for(ch.size(), nth, [&](size_t c) { p5.chunk(seq, gstart[r], len, ch[c].first, ch[c].second, cs[c], cu[c], cm[c]); });
      std::vector<OU> all;
      for (size_t c = 0; c < ch.size(); c++) {
        all.insert(all.end(), cu[c].begin(), cu[c].end());
        segs[r] += cs[c];
        metas[r].agg(cm[c]);
      }
      tiles[r] = tiling(r, all);
    }
  } catch (const std::exception&) {
    return CF1G_E_INTERNAL;
  }
  Meta meta;
  std::string pre(out_prefix);
  FILE* fs = std::fopen((pre + ".cf_seg").c_str(), "wb");
  FILE* fq = std::fopen((pre + ".cf_seq").c_str(), "wb");
  if (!fs || !fq) {
    if (fs) std::fclose(fs);
    if (fq) std::fclose(fq);
    return CF1G_E_INVALID;
  }
  for (uint64_t r = 0; r < n_rec; r++) {
    if (!present[r]) continue;
    std::fwrite(segs[r].data(), 1, segs[r].size(), fs);
    std::fwrite(tiles[r].data(), 1, tiles[r].size(), fq);
    meta.agg(metas[r]);
  }
  std::fclose(fs);
  std::fclose(fq);
  // write_json, src/output.rs:785-834 (serde_json pretty output, sorted keys)
  std::string j = "{\n";
  j += "  \"basic info\": {\n    \"vertex count\": " + std::to_string(meta.kmers) + "\n  },\n";
  j += "  \"contigs info\": {\n";
  j += "    \"_comment\": \"lengths are in bases\",\n";
  j += "    \"avg. maximal unitig length\": " + std::to_string(meta.count ? meta.suml / meta.count : 0) + ",\n";
  j += "    \"longest maximal unitig length\": " + std::to_string(meta.maxl) + ",\n";
  j += "    \"maximal unitig count\": " + std::to_string(meta.count) + ",\n";
  j += "    \"shortest maximal unitig length\": " + std::to_string(meta.minl == UINT64_MAX ? 0 : meta.minl) + ",\n";
  j += "    \"sum maximal unitig length\": " + std::to_string(meta.suml) + ",\n";
  j += "    \"vertex count in the maximal unitigs\": " + std::to_string(meta.kmers) + "\n  },\n";
  j += "  \"parameters info\": {\n";
  j += "    \"input\": \"" + json_escape(input_desc ? input_desc : "") + "\",\n";
  j += "    \"k\": " + std::to_string(K) + ",\n";
  j += "    \"output prefix\": \"" + json_escape(pre) + "\"\n  }";
  std::vector<uint64_t> shorts;
  if (track_short)
    for (uint64_t r = 0; r < n_rec; r++)
      if (rec_off[r + 1] - rec_off[r] < (uint64_t)K) shorts.push_back(r);
  if (!shorts.empty()) {
    j += ",\n  \"short seqs\": [\n";
    for (size_t i = 0; i < shorts.size(); i++) {
      uint64_t r = shorts[i];
      j += "    [\n      \"" + json_escape(name_of(r)) + "\",\n      " + std::to_string(rec_off[r + 1] - rec_off[r]) + "\n    ]";
      j += (i + 1 < shorts.size()) ? ",\n" : "\n";
    }
    j += "  ]";
  }
  j += "\n}";
  FILE* fj = std::fopen((pre + ".json").c_str(), "wb");
  if (!fj) return CF1G_E_INVALID;
  std::fwrite(j.data(), 1, j.size(), fj);
  std::fclose(fj);
  if (out_meta5) {
    out_meta5[0] = meta.count;
    out_meta5[1] = meta.kmers;
    out_meta5[2] = meta.maxl;
    out_meta5[3] = meta.minl == UINT64_MAX ? 0 : meta.minl;
    out_meta5[4] = meta.suml;
  }
  return CF1G_OK;
}
