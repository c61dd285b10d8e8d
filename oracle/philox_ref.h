/*
 * ORACLE — test infrastructure only.
 *
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11), written out directly from the paper's
 * definition with 64-bit products.  Pinned by the paper's / Random123's
 * published known-answer vectors (tests/test_philox.py).
 */
#ifndef MCREF_PHILOX_REF_H
#define MCREF_PHILOX_REF_H
#include <stdint.h>

typedef struct {
  uint32_t v[4];
} ref_u4;

static inline ref_u4 ref_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  ref_u4 x = {{c0, c1, c2, c3}};
  for (int r = 0; r < 10; r++) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * x.v[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * x.v[2];
    ref_u4 y;
    y.v[0] = (uint32_t)(p1 >> 32) ^ x.v[1] ^ k0;
    y.v[1] = (uint32_t)p1;
    y.v[2] = (uint32_t)(p0 >> 32) ^ x.v[3] ^ k1;
    y.v[3] = (uint32_t)p0;
    x = y;
  }
  return x;
}
#endif
