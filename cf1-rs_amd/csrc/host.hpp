// Host-side helpers of the cf1g product: BBHash lookup over the exported
// structure (identical to the device lookup) and k-mer encoding for P5.
#pragma once
#include <stdint.h>

#include <vector>

namespace cf1g {

typedef unsigned __int128 u128;

inline uint64_t h_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline uint64_t h_seed(int level) { return h_mix64(0x243F6A8885A308D3ULL + (uint64_t)(level + 1) * 0x9E3779B97F4A7C15ULL); }

// View over a cf1g_mphf_export buffer.
struct HostMphf {
  const uint64_t* buf = nullptr;
  int levels = 0;
  bool wide = false;
  const uint64_t* level_bits = nullptr;
  const uint64_t* level_off = nullptr;
  const uint64_t* blocks = nullptr;
  void attach(const uint64_t* b) {
    buf = b;
    levels = (int)b[0];
    wide = b[2] != 0;
    level_bits = b + 4;
    level_off = b + 4 + 64;
    blocks = b + 4 + 128;
  }
  uint64_t lookup(uint64_t lo, uint64_t hi) const {
    for (int l = 0; l < levels; l++) {
      uint64_t s = h_seed(l);
      uint64_t h = wide ? h_mix64(lo ^ h_mix64(hi ^ s)) : h_mix64(lo ^ s);
      uint64_t pos = (uint64_t)(((u128)h * level_bits[l]) >> 64);
      uint64_t q = pos / 448, b = pos % 448;
      const uint64_t* B = blocks + (level_off[l] + q) * 8;
      uint64_t wd = B[1 + b / 64];
      if ((wd >> (b % 64)) & 1) {
        uint64_t r = B[0] + (uint64_t)__builtin_popcountll(wd & ((1ULL << (b % 64)) - 1));
        for (uint64_t t = 0; t < b / 64; t++) r += (uint64_t)__builtin_popcountll(B[1 + t]);
        return r;
      }
    }
    return 0;
  }
};

// 2-bit encode of K ASCII bases (first base most significant) and its
// reverse complement, as a 2K-bit value.
inline uint8_t h_code(uint8_t c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}
inline u128 h_encode(const uint8_t* s, int K) {
  u128 v = 0;
  for (int t = 0; t < K; t++) v = (v << 2) | (u128)(h_code(s[t]) & 3);
  return v;
}
inline u128 h_rc(u128 v, int K) {
  u128 r = 0;
  for (int t = 0; t < K; t++) {
    r = (r << 2) | (u128)(3 - (uint64_t)(v & 3));
    v >>= 2;
  }
  return r;
}

}  // namespace cf1g
