// Host-side helpers of the cf1g product: BBHash lookup over the exported
// structure (identical to the device lookup) and k-mer encoding for P5.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lidx.hpp"

namespace cf1g {

typedef unsigned __int128 u128;

inline uint64_t h_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline uint64_t h_seed(int i) { return h_mix64(0x243F6A8885A308D3ULL + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL); }
// Segment and level hashes of the device (device.hpp): Fibonacci hash for
// the segment; state h1 + l * d folded through fmix32 for BBHash level l.
inline uint64_t h_seg_hash(uint64_t lo, uint64_t hi, bool wide) {
  return (lo ^ (wide ? hi * 0xC2B2AE3D27D4EB4FULL : 0)) * 0x9E3779B97F4A7C15ULL;
}
inline uint64_t h_bbh_h1(uint64_t lo, uint64_t hi, bool wide) {
  const uint64_t s = h_seed(0), m1 = 0xD6E8FEB86659FD93ULL;
  return wide ? (lo ^ h_mix64(hi ^ s)) * m1 : (lo ^ s) * m1;
}
inline uint64_t h_bbh_d(uint64_t h1) { return ((h1 << 21) | (h1 >> 43)) | 1ULL; }

inline uint32_t h_fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
inline uint64_t h_xxh3_u64le(uint64_t v) {
  uint64_t in64 = (v >> 32) + (v << 32);
  uint64_t h = in64 ^ (0x1cad21f72c81017cULL ^ 0xdb979083e96dd4deULL);
  h ^= ((h << 49) | (h >> 15)) ^ ((h << 24) | (h >> 40));
  h *= 0x9FB21C651E98DF25ULL;
  h ^= (h >> 35) + 8;
  h *= 0x9FB21C651E98DF25ULL;
  return h ^ (h >> 28);
}
inline u128 h_rc(u128 v, int K);

// View over a cf1g_mphf_export buffer (the minimizer-partitioned BBHash of
// the device, same lookup).
struct HostMphf {
  int K = 0, m = 0, sb = 0;
  bool wide = false;
  uint64_t n_segs = 0;
  const uint16_t* b2b = nullptr;
  const uint64_t* key_off = nullptr;
  const uint64_t* blk_off = nullptr;
  const uint64_t* blocks = nullptr;
  void attach(const uint64_t* b) {
    K = (int)b[1];
    m = (int)b[2];
    sb = (int)b[3];
    n_segs = b[5];
    wide = b[8] != 0;
    b2b = (const uint16_t*)(b + 16);
    key_off = b + 16 + 65536 / 4;
    blk_off = key_off + n_segs + 1;
    blocks = blk_off + n_segs + 1;
  }
  uint64_t segment(u128 c) const {
    u128 r = h_rc(c, K);
    uint32_t M = (uint32_t)((1ULL << (2 * m)) - 1);
    uint32_t best = 0xFFFFFFFFu, bv = 0;
    for (int j = 0; j < K - m + 1; j++) {
      uint32_t f = (uint32_t)(c >> (2 * (K - m - j))) & M;
      uint32_t b = (uint32_t)(r >> (2 * j)) & M;
      uint32_t cv = f < b ? f : b;
      uint32_t o = h_fmix32(cv + 0x9E3779B9u);
      if (o < best) {
        best = o;
        bv = cv;
      }
    }
    uint32_t bin = b2b[h_xxh3_u64le((uint64_t)bv) >> 48];
    const uint64_t h = h_seg_hash((uint64_t)c, (uint64_t)(c >> 64), wide);
    uint32_t sub = sb ? (uint32_t)(h >> (64 - sb)) : 0;
    return ((uint64_t)bin << sb) | sub;
  }
  uint64_t lookup(u128 c) const {
    uint64_t seg = segment(c);
    uint64_t ko = key_off[seg], t = key_off[seg + 1] - ko;
    const uint64_t* base = blocks + blk_off[seg] * 8;
    uint64_t off = 0;
    uint64_t lo = (uint64_t)c, hi = (uint64_t)(c >> 64);
    uint64_t h = h_bbh_h1(lo, hi, wide);
    const uint64_t d = h_bbh_d(h);
    for (int l = 0; l < 64 && t > 0; l++, h += d) {
      uint64_t n = (2 * t + 63) & ~63ULL;
      uint64_t pos = off + (((uint64_t)h_fmix32((uint32_t)(h >> 32) ^ (uint32_t)h) * n) >> 32);
      uint64_t q = pos / 448, b = pos % 448;
      const uint64_t* B = base + q * 8;
      uint64_t wd = B[1 + b / 64];
      if ((wd >> (b % 64)) & 1) {
        uint64_t rk = ko + B[0] + (uint64_t)__builtin_popcountll(wd & ((1ULL << (b % 64)) - 1));
        for (uint64_t u = 0; u < b / 64; u++) rk += (uint64_t)__builtin_popcountll(B[1 + u]);
        return rk;
      }
      off += n;
      t = (9 * t + 19) / 20;
    }
    return 0;
  }
};

// Host lookup through the learned index (u64 keys): the root routes the key
// to its leaf with the device's FP32 arithmetic, then an exact search over
// the leaf's keys gives the rank (the same map as the device lookup).
struct HostLidx {
  const uint64_t* keys = nullptr;
  uint64_t D = 0;
  const RootPiece* root = nullptr;
  const LeafRec* leaf = nullptr;
  uint32_t L = 0, R = 0;
  int s = 0;
  uint64_t lookup(uint64_t key) const {
    if (!D || (key >> s) >= R) return 0;
    const uint32_t r = (uint32_t)(key >> s);
    const RootPiece& P = root[r];
    const float x = (float)(key - ((uint64_t)r << s));
    const float f = std::floor(std::fmaf(P.A, x, P.B));
    const float span = (float)(P.hi - P.lo);
    const float cl = f < 0.f ? 0.f : (f > span ? span : f);
    const uint32_t j = P.lo + (uint32_t)cl;
    const uint64_t qa = leaf[j].q, qb = qa + leaf[j].n;
    const uint64_t* it = std::lower_bound(keys + qa, keys + qb, key);
    return (it != keys + qb && *it == key) ? (uint64_t)(it - keys) : 0;
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
