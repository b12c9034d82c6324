// Host-side helpers of the cf1g product (host P5).
#pragma once
#include <stdint.h>

namespace cf1g {

// 2-bit code of an ASCII base (src/dna.rs:64-92): A/a 0, C/c 1, G/g 2,
// T/t 3, anything else 4 (a placeholder).
inline uint8_t h_code(uint8_t c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

}  // namespace cf1g
