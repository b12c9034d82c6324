// Host P5 of the cf1g product: GFA-reduced output (src/output.rs:152-834)
// from the unitigs the GPU marked.
//
// The reference walks every record on the host, re-looking-up each k-mer's
// MPHF slot and state (AnnotatedKmer, src/directed_kmer.rs:81-109) to find
// unitig boundaries (output_unitigs_contiguous, src/output.rs:506-709), looks
// up each unitig's id and claims it with a CAS on the state vector
// (output_gfa_unitig, :714-782). Here all of that runs on the GPU
// (cf1g_unitig_spans: per-unitig first / last k-mer position, id,
// orientation and claim), and the host only formats the lines: cf_seg for
// the unitigs that claimed their id, cf_seq tilings per record
// (write_tiling, :208-259) and the JSON summary (write_json, :785-834).
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
  const uint64_t K = info.k;
  uint64_t U = 0;
  if ((rc = cf1g_unitig_spans(ctx, &U, nullptr, nullptr, nullptr, nullptr, 0))) return rc;
  std::vector<uint64_t> S(U + 1), E(U + 1), ID(U + 1);
  std::vector<uint8_t> FL(U + 1);
  if (U && (rc = cf1g_unitig_spans(ctx, &U, S.data(), E.data(), ID.data(), FL.data(), U + 1))) return rc;
  std::vector<uint64_t> gstart(n_rec + 1);
  if ((rc = cf1g_record_starts(ctx, gstart.data(), gstart.size()))) return rc;
  std::vector<uint32_t> file(n_rec + 1, 0);
  if ((rc = cf1g_record_files(ctx, file.data(), file.size()))) return rc;
  const uint64_t base = rec_off[0];
  auto name_of = [&](uint64_t r) -> std::string {
    if (!names_blob || !name_off) return std::string();
    return std::string(names_blob + name_off[r], names_blob + name_off[r + 1]);
  };
  // the spans of record r are [first[r], first[r + 1]) (stream order)
  std::vector<uint64_t> first(n_rec + 1, U);
  {
    uint64_t u = 0;
    for (uint64_t r = 0; r < n_rec; r++) {
      while (u < U && S[u] < gstart[r]) u++;
      first[r] = u;
    }
  }
  // pieces of at most 65536 spans inside one record, formatted in parallel
  struct Piece {
    uint64_t r, u0, u1;
    std::string seg, tile;
    Meta mt;
  };
  std::vector<Piece> pieces;
  for (uint64_t r = 0; r < n_rec; r++)
    for (uint64_t u = first[r]; u < first[r + 1]; u += 65536) pieces.push_back({r, u, std::min(first[r + 1], u + 65536), {}, {}, {}});
  const int nth = threads < 1 ? 1 : threads;
  par_for(pieces.size(), nth, [&](size_t i) {
    Piece& P = pieces[i];
    const uint64_t r = P.r, g0 = gstart[r];
    const uint8_t* seq = ascii + (rec_off[r] - base);
    char num[48];
    for (uint64_t u = P.u0; u < P.u1; u++) {
      const uint64_t s = S[u] - g0, e = E[u] - g0;
      const bool fwd = FL[u] & 1;
      // tiling token (write_tiling): N-stretch before the first unitig or
      // between unitigs, then id and orientation
      if (u == first[r]) {
        if (poly_n_stretch && s > 0) P.tile += "N" + std::to_string(s) + " ";
      } else {
        const uint64_t pe = E[u - 1] - g0;
        if (poly_n_stretch && pe + 1 != s) {
          const int64_t gap = (int64_t)s - (int64_t)pe;
          if (gap >= (int64_t)K + 1) P.tile += " N" + std::to_string((uint64_t)gap - K);
        }
        P.tile.push_back(' ');
      }
      int nn = std::snprintf(num, sizeof num, "%llu%c", (unsigned long long)ID[u], fwd ? '+' : '-');
      P.tile.append(num, (size_t)nn);
      if (!(FL[u] & 2)) continue;  // another occurrence of this unitig writes its line
      // cf_seg line (output_gfa_unitig): id, then the unitig's bases in its
      // orientation, upper-case
      const uint64_t L = e - s + K;
      nn = std::snprintf(num, sizeof num, "%llu\t", (unsigned long long)ID[u]);
      P.seg.append(num, (size_t)nn);
      if (fwd) {
        for (uint64_t t = 0; t < L; t++) {
          const uint8_t c = seq[s + t];
          P.seg.push_back((char)(c <= 'T' ? c : c - ('a' - 'A')));
        }
      } else {
        for (uint64_t t = 0; t < L; t++) P.seg.push_back("TGCA"[h_code(seq[e + K - 1 - t]) & 3]);
      }
      P.seg.push_back('\n');
      P.mt.add(e - s + 1, (int)K);
    }
  });
  Meta meta;
  std::string pre(out_prefix);
  FILE* fs = std::fopen((pre + ".cf_seg").c_str(), "wb");
  FILE* fq = std::fopen((pre + ".cf_seq").c_str(), "wb");
  if (!fs || !fq) {
    if (fs) std::fclose(fs);
    if (fq) std::fclose(fq);
    return CF1G_E_INVALID;
  }
  size_t pi = 0;
  for (uint64_t r = 0; r < n_rec; r++) {
    // records shorter than k have no tiling line (src/output.rs:166-206)
    const bool present = rec_off[r + 1] - rec_off[r] >= K;
    // ref_id = file_idx + 1 (src/output.rs:176-177, 214-217)
    std::string head = "Reference:" + std::to_string((uint64_t)file[r] + 1) + "_Sequence:" + name_of(r) + "\t";
    if (present) std::fwrite(head.data(), 1, head.size(), fq);
    for (; pi < pieces.size() && pieces[pi].r == r; pi++) {
      Piece& P = pieces[pi];
      std::fwrite(P.seg.data(), 1, P.seg.size(), fs);
      std::fwrite(P.tile.data(), 1, P.tile.size(), fq);
      meta.agg(P.mt);
    }
    if (present) std::fputc('\n', fq);
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
      if (rec_off[r + 1] - rec_off[r] < K) shorts.push_back(r);
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
