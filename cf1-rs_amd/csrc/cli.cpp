// cf1g — command-line driver with the `cf1-rs build` flag surface
// (src/params.rs:11-68, src/main.rs:6-77). Runs P1..P4 on the GPU through the
// C ABI and P5 on the host, like run_pipeline (src/pipeline.rs:66-118).
#include <dirent.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/cf1g.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: cf1g build (-s FILE | -l LIST | -d DIR) -o PREFIX [-k 31] [-t 1] [-f 3] [-w DIR]\n"
               "                  [--track-short-seqs] [--poly-N-stretch] [--collate-output-in-mem]\n"
               "                  [--num-bins 128] [--memory-budget 4.0]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2 || std::strcmp(argv[1], "build") != 0) return usage();
  std::vector<std::string> inputs;
  std::string out;
  int k = 31, threads = 1, num_bins = 128;
  bool track_short = false, poly_n = false;
  for (int i = 2; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "-s") inputs.push_back(val());
    else if (a == "-l") {
      std::ifstream lf(val());
      std::string ln;
      while (std::getline(lf, ln)) {
        size_t b = ln.find_first_not_of(" \t\r"), e = ln.find_last_not_of(" \t\r");
        if (b != std::string::npos) inputs.push_back(ln.substr(b, e - b + 1));
      }
    } else if (a == "-d") {
      std::string d = val();
      if (DIR* dp = opendir(d.c_str())) {
        while (dirent* de = readdir(dp)) {
          std::string n = de->d_name;
          size_t dot = n.rfind('.');
          std::string ext = dot == std::string::npos ? "" : n.substr(dot + 1);
          if (ext == "fa" || ext == "fasta" || ext == "fna") inputs.push_back(d + "/" + n);
        }
        closedir(dp);
      }
    } else if (a == "-k") k = std::atoi(val().c_str());
    else if (a == "-t") threads = std::atoi(val().c_str());
    else if (a == "-o") out = val();
    else if (a == "-f" || a == "-w" || a == "--work-dir" || a == "--memory-budget") val();  // accepted, unused (as in the reference)
    else if (a == "--track-short-seqs") track_short = true;
    else if (a == "--poly-N-stretch") poly_n = true;
    else if (a == "--collate-output-in-mem") {
    } else if (a == "--num-bins") num_bins = std::atoi(val().c_str());
    else return usage();
  }
  if (inputs.empty()) {
    std::fprintf(stderr, "No input files specified\n");
    return 2;
  }
  if (out.empty()) return usage();
  std::string desc;
  std::vector<const char*> paths;
  for (size_t i = 0; i < inputs.size(); i++) {
    paths.push_back(inputs[i].c_str());
    desc += (i ? ", " : "") + inputs[i];
  }
  cf1g_params p{};
  p.k = (uint32_t)k;
  p.num_bins = (uint32_t)num_bins;
  cf1g_ctx* ctx = nullptr;
  int rc = cf1g_create(&p, &ctx);
  auto die = [&](const char* what) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, cf1g_last_error(ctx));
    cf1g_destroy(ctx);
    return 1;
  };
  if (rc) return die("cf1g_create");
  auto t0 = std::chrono::steady_clock::now();
  // FASTA parsed once on the host (at least 8 threads for the parse and the
  // packing), the packed stream copied to HBM
  cf1g_ingest_stats ist{};
  if ((rc = cf1g_load_fasta(ctx, paths.data(), (uint32_t)paths.size(), (uint32_t)std::max(threads, 8), &ist)))
    return die("load");
  auto t1 = std::chrono::steady_clock::now();
  uint64_t D = 0;
  if ((rc = cf1g_run_p1_p4(ctx, &D))) return die("P1..P4");
  auto t2 = std::chrono::steady_clock::now();
  const uint8_t* ascii = nullptr;
  const uint64_t *off = nullptr, *name_off = nullptr;
  const char* names = nullptr;
  uint64_t n_rec = 0;
  if ((rc = cf1g_loaded_records(ctx, &ascii, &off, &n_rec, &names, &name_off))) return die("records");
  uint64_t meta[5];
  if ((rc = cf1g_p5_write(ctx, ascii, off, n_rec, names, name_off, out.c_str(), desc.c_str(), poly_n, track_short, threads, meta)))
    return die("P5");
  auto t3 = std::chrono::steady_clock::now();
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  std::fprintf(stderr,
               "cf1g: %llu records, %llu bases, D=%llu unitigs=%llu  load %.1f ms (parse %.1f, pack %.1f, copy %.1f), "
               "P1-P4 %.1f ms, P5 %.1f ms\n",
               (unsigned long long)ist.n_records, (unsigned long long)ist.bases, (unsigned long long)D, (unsigned long long)meta[0],
               ms(t0, t1), ist.parse_ms, ist.pack_ms, ist.h2d_ms, ms(t1, t2), ms(t2, t3));
  cf1g_destroy(ctx);
  return 0;
}
