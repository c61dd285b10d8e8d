// Deterministic FP32 primitives shared by every MI355X kernel of the engine.
//
// Every function here is a fixed sequence of IEEE-754 single-precision
// operations (explicit fmaf, correctly rounded sqrt/div, no contraction: the
// library is compiled with -ffp-contract=off), so its result does not depend
// on how the compiler schedules it.  That is what lets a CPU checker that
// performs the same sequence reproduce the GPU sweep bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcg {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11): counter-based, 4x32-bit output per
// (counter, key).  Multipliers and Weyl increments are the paper's.
// ---------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = u32x4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Philox counter tags: which random decision a draw belongs to.
enum : uint32_t {
  TAG_METRO = 0,   // single-site Metropolis update (site, sweep)
  TAG_MAJORITY = 1,// block-spin majority tie break (site, frame)
  TAG_WOLFF = 2,   // Wolff seed + direction + cluster accept (cluster)
  TAG_BOND = 3,    // Wolff bond activation (link, cluster)
  TAG_PT = 4,      // parallel-tempering pair (pair, sweep)
};

// uniform in [0,1) on the 2^-24 grid, exact in float
__host__ __device__ __forceinline__ float u24(uint32_t v) { return (float)(v >> 8) * 0x1p-24f; }

// exp(x) for x <= 0.  Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, then
// the degree-7 Taylor polynomial of e^r (truncation < 1e-8 relative) and an
// exact scaling by 2^k.  Returns 0 below -80 (e^-80 < 2^-115).
__host__ __device__ __forceinline__ float exp_neg(float x) {
  if (x < -80.0f) return 0.0f;
  const float k = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
  float r = fmaf(k, -0.693145751953125f, x);
  r = fmaf(k, -1.42860682030941723e-6f, r);
  float p = 1.98412698412698413e-4f;
  p = fmaf(p, r, 1.38888888888888889e-3f);
  p = fmaf(p, r, 8.33333333333333333e-3f);
  p = fmaf(p, r, 4.16666666666666667e-2f);
  p = fmaf(p, r, 1.66666666666666667e-1f);
  p = fmaf(p, r, 0.5f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  const int ki = (int)k;
  const float scale = __builtin_bit_cast(float, (uint32_t)(ki + 127) << 23);
  return p * scale;
}

// (cos phi, sin phi) for phi = 2*pi*(v / 2^32): octant from the top three
// bits, a 24-bit fraction inside it, Taylor polynomials of sin and cos on
// [0, pi/4] and exact octant symmetries.
__host__ __device__ __forceinline__ void sincos_turn(uint32_t v, float* c, float* s) {
  const uint32_t o = v >> 29;
  float f = (float)((v >> 5) & 0xFFFFFFu) * 0x1p-24f;
  if (o & 1u) f = 1.0f - f;
  const float a = f * 0.785398163397448310f;
  const float a2 = a * a;
  float ps = 2.75573192239858907e-6f;
  ps = fmaf(ps, a2, -1.98412698412698413e-4f);
  ps = fmaf(ps, a2, 8.33333333333333333e-3f);
  ps = fmaf(ps, a2, -1.66666666666666667e-1f);
  const float sn = fmaf(a, a2 * ps, a);
  float pc = -2.75573192239858907e-7f;
  pc = fmaf(pc, a2, 2.48015873015873016e-5f);
  pc = fmaf(pc, a2, -1.38888888888888889e-3f);
  pc = fmaf(pc, a2, 4.16666666666666667e-2f);
  pc = fmaf(pc, a2, -0.5f);
  const float cs = fmaf(pc, a2, 1.0f);
  // even octant 2m: phi = m*pi/2 + a ; odd octant: phi = ((o+1)/2)*pi/2 - a.
  // Octants 1,2,5,6 swap (cos, sin); cos < 0 in 2..5, sin < 0 in 4..7.
  const bool swp = ((o + 1u) & 2u) != 0u;
  const float x = swp ? sn : cs;
  const float y = swp ? cs : sn;
  const uint32_t cneg = ((o + 2u) & 4u) << 29;
  const uint32_t sneg = (o & 4u) << 29;
  *c = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) ^ cneg);
  *s = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, y) ^ sneg);
}

// Renormalisation factor to bring a vector of squared length nn back to
// length |S|, given inv_s2 = 1/S^2 (FP32, correctly rounded): one Newton
// step for 1/sqrt(nn/S^2) from 1, f = 1.5 - 0.5 * nn/S^2.  The reflection
// move keeps nn within ~1e-7 (FP32) or ~1e-3 (FP16 storage) of S^2, so the
// renormalised length is exact to O(delta^2).
__host__ __device__ __forceinline__ float renorm_factor(float nn, float inv_s2) {
  return fmaf(-0.5f, nn * inv_s2, 1.5f);
}

}  // namespace mcg
