// libmcgpu — MI355X (gfx950) Monte Carlo sweep engine for Ising / XY /
// Heisenberg spins on arbitrary bond graphs.  See include/mcgpu.h for the ABI.
//
// Kernels
//   k_lds<NC,ST,MODE,FULL,PERSIST>
//                     LDS sweep: each workgroup owns R replicas of one
//                     lattice (or one part of a replica split over 2-4
//                     workgroups), stages their spins in LDS once, runs many
//                     graph-coloured Metropolis sweeps with in-kernel fused
//                     measurement, writes spins back once.  PERSIST: every
//                     workgroup resident, a grid barrier per sweep, tempering
//                     and part exchange inside the kernel.  HBM traffic per
//                     sweep ~1 MB of bookkeeping; bound by VALU issue and the
//                     per-sweep synchronisation.
//   k_pass<NC,ST>     streaming colour pass over all replicas (spins in HBM /
//                     L2), for lattices too large for LDS; also the "plain"
//                     one-launch-per-replica-per-colour baseline.
//   k_measure<NC,ST>  one workgroup per replica: FP64 full-Hamiltonian
//                     recomputation and/or one measured frame.
//   k_wolff<NC,ST>    one workgroup per replica: level-synchronous frontier
//                     BFS Wolff clusters (embedded-spin for O(n)).
//   k_pt              parallel-tempering swaps of temperature rows.
//
// Every single-site update is the same deterministic function of
// (spins, site, sweep index, replica key), with a counter-based Philox
// stream, so all kernels produce the same trajectory bit for bit.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <type_traits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <queue>
#include <tuple>
#include <vector>

#include "../../include/mcgpu.h"
#include "mc_math.h"

using namespace mcg;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail("%s failed: %s", #x, hipGetErrorString(e_));       \
  } while (0)

extern "C" const char* mcg_last_error(void) { return g_err.c_str(); }
extern "C" int mcg_version(void) { return 1; }

// ---------------------------------------------------------------------------
// spin storage
// ---------------------------------------------------------------------------
// pointers into HBM with their address space spelled out, so that the
// compiler emits global (not flat) memory instructions
#define GAS __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ GAS T* gas(T* p) {
  return (GAS T*)p;
}
template <typename T>
__device__ __forceinline__ const GAS T* gas(const T* p) {
  return (const GAS T*)p;
}

template <typename ST>
struct SIO;
template <>
struct SIO<float> {
  template <class P>
  static __device__ __forceinline__ float ld(P p) { return *p; }
  template <class P>
  static __device__ __forceinline__ void st(P p, float v) { *p = v; }
};
template <>
struct SIO<__half> {
  static __device__ __forceinline__ float ld(const __half* p) { return __half2float(*p); }
  static __device__ __forceinline__ void st(__half* p, float v) { *p = __float2half_rn(v); }
  // HBM: move the 16-bit pattern as a scalar
  static __device__ __forceinline__ float ld(const GAS __half* p) {
    return __half2float(__ushort_as_half(*(const GAS unsigned short*)p));
  }
  static __device__ __forceinline__ void st(GAS __half* p, float v) {
    *(GAS unsigned short*)p = __half_as_ushort(__float2half_rn(v));
  }
};
template <>
struct SIO<int8_t> {
  template <class P>
  static __device__ __forceinline__ float ld(P p) { return (float)*p; }
  template <class P>
  static __device__ __forceinline__ void st(P p, float v) { *p = (int8_t)v; }
};

// one replica's spins; PT is ST* (LDS, inferred) or GAS ST* (HBM)
template <typename ST, typename PT = ST*>
struct Spins {
  typedef ST value_type;
  PT p;
  int n;
  __device__ __forceinline__ float get(int i, int c) const { return SIO<ST>::ld(p + (size_t)c * n + i); }
  // the spin `off` bytes away from site i's (same component)
  __device__ __forceinline__ float get_off(int i, int c, int off) const {
    return SIO<ST>::ld((PT)((char*)(p + (size_t)c * n + i) + off));
  }
  __device__ __forceinline__ void set(int i, int c, float v) const { SIO<ST>::st(p + (size_t)c * n + i, v); }
};
template <typename ST>
using GSpins = Spins<ST, GAS ST*>;

// neighbour accessors: the global link table (site order) or a pattern row
// (colour-permuted positions, LDS)
struct GlobalLinks {
  static constexpr bool BYTE_OFFSETS = false;
  const GAS uint32_t* lk;
  const GAS int32_t* lc;  // per-link coupling indices (wide graphs) or nullptr
  __device__ __forceinline__ int nbr(int l) const { return (int)(lk[l] & 0xFFFFFFu); }
  __device__ __forceinline__ uint32_t cid(int l) const { return lc ? (uint32_t)lc[l] : lk[l] >> 24; }
  // neighbours of links l .. l+3
  __device__ __forceinline__ void nbr4(int l, int* j) const {
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = nbr(l + u);
  }
};
// LDS pattern rows are padded to a multiple of four words (16 bytes), so
// that four links come in one 128-bit read
__host__ __device__ constexpr int pat_stride(int zmax) { return (zmax + 4) & ~3; }
struct PatLinks {
  static constexpr bool BYTE_OFFSETS = false;
  const int32_t* row;  // 16-byte aligned
  int p;
  __device__ __forceinline__ int nbr(int l) const { return p + ((int)((uint32_t)row[l] << 8) >> 8); }
  __device__ __forceinline__ uint32_t cid(int l) const { return (uint32_t)row[l] >> 24; }
  // neighbours of links l .. l+3 (l a multiple of 4)
  __device__ __forceinline__ void nbr4(int l, int* j) const {
    const int4 w = *(const int4*)(row + l);
    j[0] = p + ((int)((uint32_t)w.x << 8) >> 8);
    j[1] = p + ((int)((uint32_t)w.y << 8) >> 8);
    j[2] = p + ((int)((uint32_t)w.z << 8) >> 8);
    j[3] = p + ((int)((uint32_t)w.w << 8) >> 8);
  }
};
// pattern rows of a uniform-coupling graph in LDS: each link is the byte
// offset from the site's spin to the neighbour's (same component), so that a
// neighbour's address is one add
struct PatOffLinks {
  static constexpr bool BYTE_OFFSETS = true;
  const int32_t* row;  // 16-byte aligned
  int p;
  __device__ __forceinline__ int off(int l) const { return row[l]; }
  __device__ __forceinline__ void off4(int l, int* o) const {
    const int4 w = *(const int4*)(row + l);
    o[0] = w.x;
    o[1] = w.y;
    o[2] = w.z;
    o[3] = w.w;
  }
};

static constexpr double REF_PI = 3.1415926535;  // heisenbergLib.c:6

// The acceptance of a row: beta, and for exact Ising the row's threshold
// table thr[k] = floor(2^32 exp(-beta k u)) for dE = k u (u = 1 / inv_unit, a
// power of two; thr0 / nthr: the whole table, for replicas of other rows).
// A float converts to an Acc without a table (the FP32 exp path).
struct Acc {
  float beta;
  const uint32_t* thr;
  float inv_unit;
  const uint32_t* thr0;
  int nthr;
  __device__ __forceinline__ Acc(float b) : beta(b), thr(nullptr), inv_unit(0.0f), thr0(nullptr), nthr(0) {}
  __device__ __forceinline__ Acc(float b, const uint32_t* t, float iu, const uint32_t* t0, int n)
      : beta(b), thr(t), inv_unit(iu), thr0(t0), nthr(n) {}
  // the same table at another row (beta b)
  __device__ __forceinline__ Acc at(float b, int row) const {
    return Acc(b, thr0 ? thr0 + (size_t)row * nthr : nullptr, inv_unit, thr0, nthr);
  }
};
// the acceptance of row `row` of the batch
__device__ __forceinline__ Acc acc_of(const mcg_batch& B, float beta, int row) {
  return B.ithr ? Acc(beta, B.ithr + (size_t)row * B.nthr, B.e_unit_inv, B.ithr, B.nthr) : Acc(beta);
}

// ---------------------------------------------------------------------------
// the single-site Metropolis update
// ---------------------------------------------------------------------------
// Energy: E = sum_i [ 1/2 sum_l s_i^T J_il s_j(i,l) + sum_a D_a s_ia^2 - h s_i,axis ]
// (axis z for Heisenberg, x for XY; Ising s = sigma*S).  Proposal: reflection
// s' = s - 2 (s.r) r with r uniform on the sphere / circle
// (heisenbergLib.c:441-473, xyLib.c:382-409); Ising: flip (isingLib.c:238-254).
// Accept if beta*dE <= 0 or u < exp(-beta*dE).  Accepted O(n) spins are
// renormalised to |S|.
// Random words: w_a = r_z (Heisenberg), w_b = phi (XY, Heisenberg),
// w_c = acceptance.
// UNI: every link of the graph carries the tensor at coup (a uniform
// graph); the field is that tensor times the sum of the neighbour spins.
// aniso = false: no site class has an on-site term (D = 0), so the D terms,
// each an exact +0 to dE, are skipped.
template <int NC, typename SV, typename LK, typename CP, bool UNI = false>
__device__ __forceinline__ void metro_core(const SV& sp, int i, const LK& lk, int zmax, CP coup, int ci, CP clsdat,
                                           Acc ac, float h, uint32_t w_a, uint32_t w_b, uint32_t w_c,
                                           float& dEacc, float* dMacc, bool aniso = true) {
  const float beta = ac.beta;
  const float S = clsdat[ci * 8 + 0];
  if constexpr (NC == 1) {
    float f = 0.0f;
    if constexpr (UNI) {
      // links four at a time: the neighbour indices are read together, then
      // the spins, so that each site waits for two LDS round trips per four
      // links rather than two per link (the sum keeps link order)
      float m = 0.0f;
      int l = 0;
      for (; l + 4 <= zmax; l += 4) {
        int j[4];
        float a[4];
        if constexpr (LK::BYTE_OFFSETS) {
          lk.off4(l, j);
#pragma unroll
          for (int u = 0; u < 4; ++u) a[u] = sp.get_off(i, 0, j[u]);
        } else {
          lk.nbr4(l, j);
#pragma unroll
          for (int u = 0; u < 4; ++u) a[u] = sp.get(j[u], 0);
        }
        m += a[0]; m += a[1]; m += a[2]; m += a[3];
      }
      for (; l < zmax; ++l) {
        if constexpr (LK::BYTE_OFFSETS)
          m += sp.get_off(i, 0, lk.off(l));
        else
          m += sp.get(lk.nbr(l), 0);
      }
      f = coup[0] * m;
    } else {
      for (int l = 0; l < zmax; ++l) f = fmaf(coup[lk.cid(l) * 9], sp.get(lk.nbr(l), 0), f);
    }
    const float sg = sp.get(i, 0);
    const float s = sg * S;
    const float t = f - h;
    const float dE = (-2.0f * s) * t;
    bool acc;
    if (ac.thr) {
      // exact: couplings, fields and spin lengths are multiples of powers of
      // two small enough that every operation above is exact, dE = k u with
      // 0 < k < nthr, and thr[k] = floor(2^32 exp(-beta k u))
      acc = dE <= 0.0f || w_c < ac.thr[(int)(dE * ac.inv_unit)];
    } else {
      const float db = dE * beta;
      acc = db <= 0.0f || u24(w_c) < exp_neg(-db);
    }
    if (acc) {
      sp.set(i, 0, -sg);
      dEacc += dE;
      dMacc[0] += -2.0f * s;
    }
  } else {
    const float s0 = sp.get(i, 0), s1 = sp.get(i, 1);
    const float s2 = NC == 3 ? sp.get(i, 2) : 0.0f;
    float H0 = 0.0f, H1 = 0.0f, H2 = 0.0f;
    if constexpr (UNI) {
      // links four at a time, as for Ising
      float S0 = 0.0f, S1 = 0.0f, S2 = 0.0f;
      int l = 0;
      for (; l + 4 <= zmax; l += 4) {
        int j[4];
        float t0[4], t1[4], t2[4];
        if constexpr (LK::BYTE_OFFSETS) {
          lk.off4(l, j);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            t0[u] = sp.get_off(i, 0, j[u]);
            t1[u] = sp.get_off(i, 1, j[u]);
            t2[u] = NC == 3 ? sp.get_off(i, 2, j[u]) : 0.0f;
          }
        } else {
          lk.nbr4(l, j);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            t0[u] = sp.get(j[u], 0);
            t1[u] = sp.get(j[u], 1);
            t2[u] = NC == 3 ? sp.get(j[u], 2) : 0.0f;
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          S0 += t0[u];
          S1 += t1[u];
          if constexpr (NC == 3) S2 += t2[u];
        }
      }
      for (; l < zmax; ++l) {
        if constexpr (LK::BYTE_OFFSETS) {
          const int o = lk.off(l);
          S0 += sp.get_off(i, 0, o);
          S1 += sp.get_off(i, 1, o);
          if constexpr (NC == 3) S2 += sp.get_off(i, 2, o);
        } else {
          const int j = lk.nbr(l);
          S0 += sp.get(j, 0);
          S1 += sp.get(j, 1);
          if constexpr (NC == 3) S2 += sp.get(j, 2);
        }
      }
      H0 = coup[0] * S0; H0 = fmaf(coup[1], S1, H0);
      H1 = coup[3] * S0; H1 = fmaf(coup[4], S1, H1);
      if constexpr (NC == 3) {
        H0 = fmaf(coup[2], S2, H0);
        H1 = fmaf(coup[5], S2, H1);
        H2 = coup[6] * S0; H2 = fmaf(coup[7], S1, H2); H2 = fmaf(coup[8], S2, H2);
      }
    } else {
      for (int l = 0; l < zmax; ++l) {
        const int j = lk.nbr(l);
        const CP J = coup + lk.cid(l) * 9;
        const float t0 = sp.get(j, 0), t1 = sp.get(j, 1);
        if constexpr (NC == 3) {
          const float t2 = sp.get(j, 2);
          H0 = fmaf(J[0], t0, H0); H0 = fmaf(J[1], t1, H0); H0 = fmaf(J[2], t2, H0);
          H1 = fmaf(J[3], t0, H1); H1 = fmaf(J[4], t1, H1); H1 = fmaf(J[5], t2, H1);
          H2 = fmaf(J[6], t0, H2); H2 = fmaf(J[7], t1, H2); H2 = fmaf(J[8], t2, H2);
        } else {
          H0 = fmaf(J[0], t0, H0); H0 = fmaf(J[1], t1, H0);
          H1 = fmaf(J[3], t0, H1); H1 = fmaf(J[4], t1, H1);
        }
      }
    }
    float rx, ry, rz = 0.0f;
    if constexpr (NC == 3) {
      rz = fmaf(u24(w_a), 2.0f, -1.0f);
      const float rho = sqrtf(fmaxf(fmaf(-rz, rz, 1.0f), 0.0f));
      float c, sn;
      sincos_turn(w_b, &c, &sn);
      rx = rho * c;
      ry = rho * sn;
    } else {
      sincos_turn(w_b, &rx, &ry);
    }
    float p = s0 * rx;
    p = fmaf(s1, ry, p);
    if constexpr (NC == 3) p = fmaf(s2, rz, p);
    const float a = -2.0f * p;
    const float d0 = a * rx, d1 = a * ry, d2 = a * rz;
    float dE = d0 * H0;
    dE = fmaf(d1, H1, dE);
    if constexpr (NC == 3) dE = fmaf(d2, H2, dE);
    float n0 = s0 + d0, n1 = s1 + d1, n2 = s2 + d2;
    if (aniso) {
      const float Dx = clsdat[ci * 8 + 1], Dy = clsdat[ci * 8 + 2], Dz = clsdat[ci * 8 + 3];
      const float a0 = n0 * n0, b0 = s0 * s0, q0 = a0 - b0;
      const float a1 = n1 * n1, b1 = s1 * s1, q1 = a1 - b1;
      dE = fmaf(Dx, q0, dE);
      dE = fmaf(Dy, q1, dE);
      if constexpr (NC == 3) {
        const float a2 = n2 * n2, b2 = s2 * s2, q2 = a2 - b2;
        dE = fmaf(Dz, q2, dE);
      }
    }
    if (h != 0.0f) dE = fmaf(-h, NC == 3 ? d2 : d0, dE);  // h = 0 would add an exact 0
    const float db = dE * beta;
    if (db <= 0.0f || u24(w_c) < exp_neg(-db)) {
      float nn = n0 * n0;
      nn = fmaf(n1, n1, nn);
      if constexpr (NC == 3) nn = fmaf(n2, n2, nn);
      const float f = renorm_factor(nn, clsdat[ci * 8 + 4]);
      n0 = n0 * f;
      n1 = n1 * f;
      n2 = n2 * f;
      sp.set(i, 0, n0);
      sp.set(i, 1, n1);
      if constexpr (NC == 3) sp.set(i, 2, n2);
      dEacc += dE;
      if constexpr (std::is_same<typename SV::value_type, float>::value) {
        // FP32 storage keeps the value as computed
        dMacc[0] += n0 - s0;
        dMacc[1] += n1 - s1;
        if constexpr (NC == 3) dMacc[2] += n2 - s2;
      } else {
        dMacc[0] += sp.get(i, 0) - s0;
        dMacc[1] += sp.get(i, 1) - s1;
        if constexpr (NC == 3) dMacc[2] += sp.get(i, 2) - s2;
      }
    }
  }
}

// One Philox draw (128 bits) serves G sites of a colour class: 4 Ising
// updates (one acceptance word each), 2 XY updates (angle + acceptance), or
// 1 Heisenberg update (r_z, angle, acceptance).  Group q of colour c covers
// colour-list positions q*G .. q*G+G-1; counter (q, sweep, c<<8 | TAG).
template <int NC>
struct SitesPerDraw {
  static constexpr int G = NC == 1 ? 4 : NC == 2 ? 2 : 1;
};

template <int NC, bool UNI = false, typename SV, typename CP>
__device__ __forceinline__ void metro_group(const SV& sp, int b, int cnt, int q, int colour,
                                            const GAS uint32_t* __restrict__ link,
                                            const GAS int32_t* __restrict__ lcid, int zmax, CP coup,
                                            const GAS int32_t* __restrict__ cls, CP clsdat, Acc beta, float h,
                                            uint64_t sweep, uint32_t key0, uint32_t key1, float& dEacc,
                                            float* dMacc) {
  constexpr int G = SitesPerDraw<NC>::G;
  const u32x4 x = philox4x32_10(
      u32x4{(uint32_t)q, (uint32_t)sweep, (uint32_t)(sweep >> 32), TAG_METRO | ((uint32_t)colour << 8)}, key0, key1);
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int p = q * G + g;
    if (p < cnt) {
      const int i = b + p;
      const GlobalLinks lk{link + (size_t)i * zmax, lcid ? lcid + (size_t)i * zmax : nullptr};
      const uint32_t wa = NC == 3 ? w[0] : 0u;
      const uint32_t wb = NC == 3 ? w[1] : NC == 2 ? w[2 * g] : 0u;
      const uint32_t wc = NC == 3 ? w[2] : NC == 2 ? w[2 * g + 1] : w[g];
      metro_core<NC, SV, GlobalLinks, CP, UNI>(sp, i, lk, zmax, coup, (int)cls[i], clsdat, beta, h, wa, wb, wc, dEacc,
                                               dMacc);
    }
  }
}

// same, on the colour-permuted LDS replica: colour c occupies positions
// b .. b+cnt-1, the neighbours come from the pattern row of each position
template <int NC, typename ST, bool UNI = false>
__device__ __forceinline__ void metro_group_pat(const Spins<ST>& sp, int b, int cnt, int q, int colour,
                                                const uint8_t* pat, const int32_t* pattab, int zmax,
                                                const float* coup, const float* clsdat, Acc beta, float h,
                                                uint64_t sweep, uint32_t key0, uint32_t key1, float& dEacc,
                                                float* dMacc, bool aniso = true) {
  constexpr int G = SitesPerDraw<NC>::G;
  const u32x4 x = philox4x32_10(
      u32x4{(uint32_t)q, (uint32_t)sweep, (uint32_t)(sweep >> 32), TAG_METRO | ((uint32_t)colour << 8)}, key0, key1);
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int lp = q * G + g;
    if (lp < cnt) {
      const int p = b + lp;
      // 24-bit multiply: full rate, and pattern ids and strides are small
      const int32_t* row = pattab + __umul24((uint32_t)pat[p], (uint32_t)pat_stride(zmax));
      typedef typename std::conditional<UNI, PatOffLinks, PatLinks>::type LK;
      const LK lk{row, p};
      const uint32_t wa = NC == 3 ? w[0] : 0u;
      const uint32_t wb = NC == 3 ? w[1] : NC == 2 ? w[2 * g] : 0u;
      const uint32_t wc = NC == 3 ? w[2] : NC == 2 ? w[2 * g + 1] : w[g];
      metro_core<NC, Spins<ST>, LK, const float*, UNI>(sp, p, lk, zmax, coup, row[zmax], clsdat, beta, h, wa, wb, wc,
                                                       dEacc, dMacc, aniso);
    }
  }
}

// same, with the spins in HBM (streaming pass): pattern ids read from HBM,
// pattern rows (raw position deltas | coupling << 24, then the class) in LDS
template <int NC, typename ST, bool UNI>
__device__ __forceinline__ void metro_group_gpat(const GSpins<ST>& sp, int b, int cnt, int q, int colour,
                                                 const GAS uint8_t* __restrict__ pat, const int32_t* pattab, int zmax,
                                                 const GAS float* coup, const GAS float* clsdat, Acc beta, float h,
                                                 uint64_t sweep, uint32_t key0, uint32_t key1, float& dEacc,
                                                 float* dMacc, bool aniso) {
  constexpr int G = SitesPerDraw<NC>::G;
  const u32x4 x = philox4x32_10(
      u32x4{(uint32_t)q, (uint32_t)sweep, (uint32_t)(sweep >> 32), TAG_METRO | ((uint32_t)colour << 8)}, key0, key1);
  const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int lp = q * G + g;
    if (lp < cnt) {
      const int p = b + lp;
      const int32_t* row = pattab + __umul24((uint32_t)pat[p], (uint32_t)pat_stride(zmax));
      const PatLinks lk{row, p};
      const uint32_t wa = NC == 3 ? w[0] : 0u;
      const uint32_t wb = NC == 3 ? w[1] : NC == 2 ? w[2 * g] : 0u;
      const uint32_t wc = NC == 3 ? w[2] : NC == 2 ? w[2 * g + 1] : w[g];
      metro_core<NC, GSpins<ST>, PatLinks, const GAS float*, UNI>(sp, p, lk, zmax, coup, row[zmax], clsdat, beta, h,
                                                                  wa, wb, wc, dEacc, dMacc, aniso);
    }
  }
}

// spin vector of site i (Ising: sigma*S)
template <int NC, typename SV>
__device__ __forceinline__ void spin_vec(const SV& sp, int i, const int32_t* __restrict__ cls, const float* clsdat,
                                         float* v) {
  if constexpr (NC == 1) {
    v[0] = sp.get(i, 0) * clsdat[cls[i] * 8];
    v[1] = v[2] = 0.0f;
  } else {
    v[0] = sp.get(i, 0);
    v[1] = sp.get(i, 1);
    v[2] = NC == 3 ? sp.get(i, 2) : 0.0f;
  }
}

// per-site energy e_i (FP32): 1/2 s_i.H_i + onsite - h s_axis
template <int NC, typename SV>
__device__ __forceinline__ float site_energy(const SV& sp, int i, const uint32_t* __restrict__ link,
                                             const int32_t* __restrict__ lcid, int zmax, const float* coup,
                                             const int32_t* __restrict__ cls, const float* clsdat, float h) {
  const int ci = cls[i];
  const GlobalLinks lk{gas(link) + (size_t)i * zmax, lcid ? gas(lcid) + (size_t)i * zmax : nullptr};
  if constexpr (NC == 1) {
    float f = 0.0f;
    for (int l = 0; l < zmax; ++l) f = fmaf(coup[lk.cid(l) * 9], sp.get(lk.nbr(l), 0), f);
    const float s = sp.get(i, 0) * clsdat[ci * 8];
    return fmaf(0.5f * s, f, -h * s);
  } else {
    const float s0 = sp.get(i, 0), s1 = sp.get(i, 1), s2 = NC == 3 ? sp.get(i, 2) : 0.0f;
    float H0 = 0.0f, H1 = 0.0f, H2 = 0.0f;
    for (int l = 0; l < zmax; ++l) {
      const int j = lk.nbr(l);
      const float* J = coup + lk.cid(l) * 9;
      const float t0 = sp.get(j, 0), t1 = sp.get(j, 1), t2 = NC == 3 ? sp.get(j, 2) : 0.0f;
      H0 = fmaf(J[0], t0, H0); H0 = fmaf(J[1], t1, H0); H0 = fmaf(J[2], t2, H0);
      H1 = fmaf(J[3], t0, H1); H1 = fmaf(J[4], t1, H1); H1 = fmaf(J[5], t2, H1);
      H2 = fmaf(J[6], t0, H2); H2 = fmaf(J[7], t1, H2); H2 = fmaf(J[8], t2, H2);
    }
    float e = s0 * H0;
    e = fmaf(s1, H1, e);
    e = fmaf(s2, H2, e);
    e = 0.5f * e;
    e = fmaf(clsdat[ci * 8 + 1] * s0, s0, e);
    e = fmaf(clsdat[ci * 8 + 2] * s1, s1, e);
    if constexpr (NC == 3) e = fmaf(clsdat[ci * 8 + 3] * s2, s2, e);
    e = fmaf(-h, NC == 3 ? s2 : s0, e);
    return e;
  }
}

// ---------------------------------------------------------------------------
// group reductions (a group = TPR consecutive threads, TPR a multiple of 64)
// ---------------------------------------------------------------------------
// accesses at the device coherence point (device-scope atomics): visible to
// every workgroup of the launch whatever XCD it runs on
template <typename T>
__device__ __forceinline__ void dev_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T dev_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void dev_add(double* p, double v) { unsafeAtomicAdd(p, v); }

static constexpr int RK = 8;  // values per reduction call
static constexpr int FRM_ACC = 96;  // frame scratch: accumulator staging starts here

// FP64 value moved across lanes by DPP (two 32-bit halves); lanes of rows
// outside ROWMASK get 0.  With every row enabled (the quad permutes and row
// mirrors used here, where each lane reads a lane of its own row) there is
// no old value to keep, so the destination is not zeroed first: those two
// moves per step were a third of the instructions of a wave's FP64 sum,
// which all waves of a workgroup run at once at the end of every sweep.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __double2loint(x), hi = __double2hiint(x);
  int l2, h2;
  if constexpr (ROWMASK == 0xF) {
    l2 = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    h2 = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  } else {
    l2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWMASK, 0xF, false);
    h2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWMASK, 0xF, false);
  }
  return __hiloint2double(h2, l2);
}

// sum over the 64 lanes of a wave (every lane active); the total lands in
// lane 63.  Quad swaps, half-row and row mirrors give every lane its 16-lane
// row sum; row broadcasts 15 and 31 then add the rows up.  DPP moves are
// VALU operations: no LDS traffic, unlike a shuffle butterfly.
__device__ __forceinline__ double wave_sum_f64(double x) {
  x += dpp_f64<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141, 0xF>(x);  // row_half_mirror
  x += dpp_f64<0x140, 0xF>(x);  // row_mirror
  x += dpp_f64<0x142, 0xA>(x);  // row_bcast:15 into rows 1, 3
  x += dpp_f64<0x143, 0xC>(x);  // row_bcast:31 into rows 2, 3
  return x;
}

// lane src's value (src the same in every lane), in every lane of the wave
__device__ __forceinline__ double lane_f64(double x, int src) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63(double x) { return lane_f64(x, 63); }

// sum of N (<= RK) values over the TPR threads of group g (TPR a multiple of
// 64); the result is valid in every lane of the group's first wave, so in
// the group's first thread.  All threads call it; it ends with a workgroup
// barrier.  Two levels of DPP wave sums: the waves' totals go through LDS
// and the group's first wave adds them up (at most 16 waves).
template <int N>
__device__ __forceinline__ void group_sum(double* v, double* s_red, int g, int TPR) {
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = wave_sum_f64(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = TPR >> 6;
  if (nw == 1) {  // the same for every group of the workgroup
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = lane63(v[k]);
    __syncthreads();
    return;
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < N; ++k) s_red[wave * RK + k] = v[k];
  __syncthreads();
  const int w0 = g * nw;
  if (wave == w0) {  // the whole wave, so that the DPP sums see every lane
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double x = lane < nw ? s_red[(w0 + lane) * RK + k] : 0.0;
      v[k] = lane63(wave_sum_f64(x));
    }
  }
  __syncthreads();
}
__device__ __forceinline__ void group_sum8(double* v, double* s_red, int g, int TPR) { group_sum<RK>(v, s_red, g, TPR); }

// ---------------------------------------------------------------------------
// measurement of one frame (all flagged observables) for one replica
// ---------------------------------------------------------------------------
struct FrameCtx {
  const mcg_graph* G;
  const float* coup;
  const float* clsdat;
  double* s_red;   // reduction scratch (nwaves*RK)
  double* s_frm;   // per-group frame values (>= 64 doubles per group)
  int g, lt, TPR;
};

// majority spin of block-spin site (heisenbergLib.c:255-267, isingLib.c:133-150)
template <int NC, typename SV>
__device__ __forceinline__ void majority(const SV& sp, const mcg_graph& G, const float* clsdat, int site,
                                         uint64_t frame, uint32_t key0, uint32_t key1, float* v) {
  const int ri = G.rrank[site];
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
  for (int k = 0; k < G.nclu; ++k) {
    float t[3];
    spin_vec<NC>(sp, G.rclu[ri * G.nclu + k], G.cls, clsdat, t);
    a0 += t[0];
    a1 += t[1];
    a2 += t[2];
  }
  const float S = clsdat[G.cls[site] * 8];
  if constexpr (NC == 1) {
    float sg;
    if (a0 > 0.0f)
      sg = 1.0f;
    else if (a0 < 0.0f)
      sg = -1.0f;
    else {
      const u32x4 x = philox4x32_10(u32x4{(uint32_t)site, (uint32_t)frame, (uint32_t)(frame >> 32), TAG_MAJORITY},
                                    key0, key1);
      sg = u24(x.x) > 0.5f ? 1.0f : -1.0f;
    }
    v[0] = sg * fabsf(S);
    v[1] = v[2] = 0.0f;
  } else {
    const float l = sqrtf(a0 * a0 + a1 * a1 + a2 * a2);
    if (l >= 1e-5f) {
      a0 /= l;
      a1 /= l;
      a2 /= l;
    }
    v[0] = a0 * S;
    v[1] = a1 * S;
    v[2] = a2 * S;
  }
}

// solid angle of a spin triangle, heisenbergLib.c:114-127 (atan, +-pi branch)
// (ia, ib, ic: 1/|S| of the three sites).  An observable, compared within
// statistics: products with reciprocal lengths and one fast division stand
// in for the reference's seven divisions.
__device__ __forceinline__ float signed_area(const float* a, const float* b, const float* c, float ia, float ib,
                                             float ic) {
  const float ab = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * (ia * ib);
  const float bc = (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * (ib * ic);
  const float ca = (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]) * (ic * ia);
  const float x0 = b[1] * c[2] - b[2] * c[1], x1 = b[2] * c[0] - b[0] * c[2], x2 = b[0] * c[1] - b[1] * c[0];
  const float re = 1.0f + ab + bc + ca;
  const float im = (a[0] * x0 + a[1] * x1 + a[2] * x2) * (ia * ib * ic);
  if (fabsf(re) < 1e-6f) return im > 0.0f ? (float)REF_PI : -(float)REF_PI;
  return 2.0f * atanf(__fdividef(im, re));
}

// 1/|S| of a site class: sqrt of the stored 1/S^2
__device__ __forceinline__ float inv_len(const float* clsdat, const int32_t* cls, int i) {
  return sqrtf(clsdat[cls[i] * 8 + 4]);
}

// Measures one frame.  Tracked E and M (FP64) of the replica are passed in
// st[]; everything else is computed from the spins.  Only the group leader
// (lt == 0) writes the accumulators.
// Observables taken from the spins themselves (pairs, Q, block spins,
// groups); a kernel compiled without them (FULL = false) is much smaller.
static constexpr int MEAS_SPIN_OBS = MCG_MEAS_PAIRS | MCG_MEAS_Q | MCG_MEAS_RNORM | MCG_MEAS_GROUPS;

template <int NC>
__device__ __forceinline__ int frame_groups(const mcg_graph& G, int flags) {
  return (NC > 1 && (flags & MCG_MEAS_GROUPS)) ? G.ngroup : 0;
}

// apply a frame staged by measure_frame to the accumulators of its row: one
// device-scope FP64 atomic per non-zero term, spread over the group's
// threads.  Fire-and-forget: nothing here waits for the atomics.
template <int NC>
__device__ __forceinline__ void flush_frame(const FrameCtx& C, const mcg_batch& B, int row, int flags) {
  const mcg_graph& G = *C.G;
  const int lt = C.lt, TPR = C.TPR;
  const double* F = C.s_frm;
  const double* Cn = F + FRM_ACC;
  const int ngq = frame_groups<NC>(G, flags);
  double* A = B.acc + (size_t)row * MCG_NACC;
  const int nge = ngq > 0 ? (ngq + 1) * (ngq + 1) : 0;
  for (int q = lt; q < MCG_NACC + MCG_NBINQ + nge; q += TPR) {
    if (q < MCG_NACC) {
      if (q != MCG_A_MPREV && Cn[q] != 0.0) dev_add(A + q, Cn[q]);  // adding 0 changes nothing
    } else if (q < MCG_NACC + MCG_NBINQ) {
      const int64_t bin = (int64_t)Cn[MCG_NACC + MCG_NBINQ];
      if (Cn[q] != 0.0) dev_add(B.bins + ((size_t)row * B.nbins + bin) * MCG_NBINQ + (q - MCG_NACC), Cn[q]);
    } else {
      // orbital groups (heisenbergLib.c:805-830): <g_a . g_b> and <(g_a . g_a)^2>
      const int j = q - MCG_NACC - MCG_NBINQ, ga = j / (ngq + 1), gb = j % (ngq + 1);
      double* GA = B.gacc + (size_t)row * ((ngq + 2) * (ngq + 1));
      const double* va = F + 32 + 3 * ga;
      const double* vb = F + 32 + 3 * gb;
      const double d = va[0] * vb[0] + va[1] * vb[1] + va[2] * vb[2];
      dev_add(GA + j, d);
      if (ga == gb) dev_add(GA + (ngq + 1) * (ngq + 1) + ga, d * d);
    }
  }
}

// the previous frame's M of a row (for <M M'>), read ahead of the frame by
// the thread that applies it
__device__ __forceinline__ double frame_mprev(const mcg_batch& B, int row) {
  return dev_load(B.acc + (size_t)row * MCG_NACC + MCG_A_MPREV);
}

// One frame's contribution to every FP64 accumulator of a row
// (isingLib.c:388-431, heisenbergLib.c:677-830 / xyLib.c:605-757), staged in
// Cn for flush_frame: [0, NACC) accumulators ([MM] holds M itself until the
// frame's M' product replaces it), [NACC, +NBINQ) the bin terms, then the bin
// index.  st: the replica's state; F: the frame's spin-derived values.  One
// thread per replica; every term is written (zeros included).
template <int NC>
__device__ __forceinline__ void frame_terms(const mcg_graph& G, const mcg_batch& B, const double* st, double* F,
                                            int flags, int64_t frame, double* Cn) {
  const int ngq = frame_groups<NC>(G, flags);
  for (int q = 0; q < MCG_NACC + MCG_NBINQ + 1; ++q) Cn[q] = 0.0;
  const double n = G.n;
  const double nLat = G.nlat > 0 ? G.nlat : G.n;
  const double inv_nLat = 1.0 / nLat;
  const bool pairs = (flags & MCG_MEAS_PAIRS) && G.nlat > 0;
  double E = st[MCG_S_E];
  if (flags & MCG_ISING_RELE) E -= st[MCG_S_EREF];
  const double e = E * (1.0 / n);
  const double M0 = st[MCG_S_M], M1 = st[MCG_S_M + 1], M2 = st[MCG_S_M + 2];
  if (pairs && G.self_pairs) {
    // pairs (i, i) over every site: sum s_i is the tracked magnetisation,
    // sum s_i.s_i the sum of the squared spin lengths
    F[0] = F[3] = M0;
    F[1] = F[4] = M1;
    F[2] = F[5] = M2;
    F[6] = G.pair_s2;
  }
  double Mf;
  const int ax = NC == 3 ? 2 : 0;
  if constexpr (NC == 1) {
    Mf = (pairs ? F[0] : M0) * inv_nLat;
    if (pairs) {
      Cn[MCG_A_SI] = fabs(F[0]) * inv_nLat;
      Cn[MCG_A_SJ] = fabs(F[3]) * inv_nLat;
      Cn[MCG_A_SIJ] = F[6] * inv_nLat;
      Cn[MCG_A_LAST] = F[7];
    }
    Cn[MCG_A_TOT] = M0;
  } else {
    const double Mv[3] = {M0, M1, NC == 3 ? M2 : 0.0};
    double dir[3] = {Mv[0], Mv[1], Mv[2]};
    const double ml = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (ml >= 1e-5) {
      dir[0] /= ml; dir[1] /= ml;
      if constexpr (NC == 3) dir[2] /= ml;  // XY: 0 / ml = 0
    }
    if (NC == 3 || !pairs)
      Mf = ml * inv_nLat;
    else
      Mf = sqrt(F[0] * F[0] + F[1] * F[1]) * inv_nLat;
    if (pairs) {
      for (int k = 0; k < NC; ++k) {
        Cn[MCG_A_SI + k] = fabs(F[k] * inv_nLat);
        Cn[MCG_A_SJ + k] = fabs(F[3 + k] * inv_nLat);
      }
      Cn[MCG_A_SIJ] = F[6] * inv_nLat;
      Cn[MCG_A_ZI] = (dir[0] * F[0] + dir[1] * F[1] + dir[2] * F[2]) * inv_nLat;
      Cn[MCG_A_ZJ] = (dir[0] * F[3] + dir[1] * F[4] + dir[2] * F[5]) * inv_nLat;
      Cn[MCG_A_HI] = F[ax] * inv_nLat;
      Cn[MCG_A_HJ] = F[3 + ax] * inv_nLat;
    }
    Cn[MCG_A_ZT] = (dir[0] * Mv[0] + dir[1] * Mv[1] + dir[2] * Mv[2]) * inv_nLat;
    Cn[MCG_A_HT] = Mv[ax] * inv_nLat;
    if ((flags & MCG_MEAS_Q) && NC == 3 && G.ncirc > 0) Cn[MCG_A_Q] = F[8] / REF_PI / 4;
    if ((flags & MCG_MEAS_RNORM) && G.nr > 0) {
      for (int k = 0; k < NC; ++k) {
        Cn[MCG_A_SIR + k] = fabs(F[16 + k] / F[24]);
        Cn[MCG_A_SJR + k] = fabs(F[19 + k] / F[25]);
      }
      Cn[MCG_A_SIJR] = F[22] / F[26];
    }
    // groups: the total (all sites / nLat) is group ng
    if (ngq > 0)
      for (int k = 0; k < 3; ++k) F[32 + 3 * ngq + k] = Mv[k] * inv_nLat;
  }
  if ((flags & MCG_MEAS_RNORM) && G.nr > 0) {
    const double er = (F[27] / 2 + F[28]) / G.nr;
    Cn[MCG_A_ER] = er;
    Cn[MCG_A_ER2] = er * er;
  }
  Cn[MCG_A_E] = e;
  Cn[MCG_A_E2] = e * e;
  Cn[MCG_A_M] = Mf;
  Cn[MCG_A_M2] = Mf * Mf;
  Cn[MCG_A_M4] = Mf * Mf * Mf * Mf;
  Cn[MCG_A_MM] = Mf;  // applied as A[MM] += A[MPREV] * Mf; A[MPREV] = Mf
  Cn[MCG_A_COUNT] = 1.0;
  Cn[MCG_A_ABSM] = fabs(Mf);
  // bins for error bars: bin = floor(frame * nbins / nframes), from an FP64
  // quotient corrected to the exact integer floor
  const int64_t nf = B.nframes > 0 ? B.nframes : 1;
  const int64_t num = frame * (int64_t)B.nbins;
  int64_t bin = (int64_t)floor((double)num * (1.0 / (double)nf));  // corrected to the exact floor below
  if (bin * nf > num) --bin;
  if ((bin + 1) * nf <= num) ++bin;
  if (bin >= B.nbins) bin = B.nbins - 1;
  if (bin < 0) bin = 0;
  Cn[MCG_NACC + 0] = e;
  Cn[MCG_NACC + 1] = e * e;
  Cn[MCG_NACC + 2] = Mf;
  Cn[MCG_NACC + 3] = Mf * Mf;
  Cn[MCG_NACC + 4] = Mf * Mf * Mf * Mf;
  Cn[MCG_NACC + MCG_NBINQ] = (double)bin;
}

// One measured frame of the group's replica at `row`: the observables are
// computed and staged in the group's frame area, <M M'> and M' are brought
// up to date at once (thread MCG_A_MM; mprev, if have_prev, is M' read ahead
// by that thread), and the rest is applied to the accumulators now or, with
// defer, left staged for the caller's flush_frame.
// [lo, hi): share `part` of `np` of a loop over cnt items
__device__ __forceinline__ void slice(int cnt, int part, int np, int& lo, int& hi) {
  lo = (int)((int64_t)cnt * part / np);
  hi = (int)((int64_t)cnt * (part + 1) / np);
}

// The frame's share of one part of a split replica: the correlation pairs
// and circuits whose first site the part owns (indices into G.pairs and
// G.circ, from the split tables), and whether it owns the last site.
struct PartFrame {
  const int32_t* pairs;
  const int32_t* circs;
  int npairs, ncircs;
  bool last;
};

// The spin-derived sums of one frame over share `part` of `np` of every
// loop (pairs, circuits, block spins, groups), left in F (valid in the
// group's first thread): F[0..8) pairs, F[8] solid angles, F[16..29) block
// spins, F[32..) groups.  np = 1: the whole frame.  With `own`, the share
// is instead own's pairs and circuits (their other sites being current in
// the part's copy), and block spins and groups are not measured.
static constexpr int FSUMS = 88;     // frame sums per replica (F[0, 88))
static constexpr int FSUMS_PQ = 9;   // of which pairs and circuits (F[0, 9))
template <int NC, typename SV, bool FULL = true>
__device__ void frame_sums(const SV& sp, const FrameCtx& C, const mcg_batch& B, int64_t frame, uint32_t key0,
                           uint32_t key1, float h, int part, int np, const PartFrame* own = nullptr) {
  const mcg_graph& G = *C.G;
  const int lt = C.lt, TPR = C.TPR;
  double* F = C.s_frm;  // leader-only use
  double v[RK];
  const int flags = FULL ? B.meas : (B.meas & ~MEAS_SPIN_OBS);
  int lo, hi;
  // --- correlation pairs: S_i, S_j, sum s_i.s_j, last-site spin
  if ((flags & MCG_MEAS_PAIRS) && G.nlat > 0) {
    for (int k = 0; k < RK; ++k) v[k] = 0.0;
    slice(G.nlat, part, np, lo, hi);
    if (own) lo = 0, hi = own->npairs;
    if (G.self_pairs) hi = lo;  // the sums follow from the state (frame_terms)
    for (int q = lo + lt; q < hi; q += TPR) {
      const int p = own ? own->pairs[q] : q;
      float a[3], b[3];
      spin_vec<NC>(sp, G.pairs[2 * p], G.cls, C.clsdat, a);
      spin_vec<NC>(sp, G.pairs[2 * p + 1], G.cls, C.clsdat, b);
      v[0] += a[0]; v[1] += a[1]; v[2] += a[2];
      v[3] += b[0]; v[4] += b[1]; v[5] += b[2];
      v[6] += (double)(a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
    }
    if (lt == 0 && (own ? own->last : part == 0)) {
      float t[3];
      spin_vec<NC>(sp, G.pos[G.n - 1], G.cls, C.clsdat, t);
      v[7] = t[0];
    }
    // self pairs: every other lane holds zeros, so the group reduction (a
    // microsecond of FP64 DPP steps and barriers per frame) is not needed
    if (!G.self_pairs) group_sum8(v, C.s_red, C.g, TPR);
    if (lt == 0)
      for (int k = 0; k < 8; ++k) F[k] = v[k];
  }
  // --- topological charge
  if ((flags & MCG_MEAS_Q) && NC == 3 && G.ncirc > 0) {
    for (int k = 0; k < RK; ++k) v[k] = 0.0;
    slice(G.ncirc, part, np, lo, hi);
    if (own) lo = 0, hi = own->ncircs;
    // one spin length for every class: no per-site class lookups
    float il = sqrtf(C.clsdat[4]);
    bool one_len = G.ncls <= 64;
    for (int q = 1; q < G.ncls && one_len; ++q) one_len = C.clsdat[q * 8 + 4] == C.clsdat[4];
    for (int q = lo + lt; q < hi; q += TPR) {
      const int p = own ? own->circs[q] : q;
      const int ia = G.circ[3 * p], ib = G.circ[3 * p + 1], ic = G.circ[3 * p + 2];
      float a[3], b[3], c[3];
      spin_vec<NC>(sp, ia, G.cls, C.clsdat, a);
      spin_vec<NC>(sp, ib, G.cls, C.clsdat, b);
      spin_vec<NC>(sp, ic, G.cls, C.clsdat, c);
      if (one_len)
        v[0] += (double)signed_area(a, b, c, il, il, il);
      else
        v[0] += (double)signed_area(a, b, c, inv_len(C.clsdat, G.cls, ia), inv_len(C.clsdat, G.cls, ib),
                                    inv_len(C.clsdat, G.cls, ic));
    }
    group_sum<1>(v, C.s_red, C.g, TPR);
    if (lt == 0) F[8] = v[0];
  }
  // --- block-spin observables
  if ((flags & MCG_MEAS_RNORM) && G.nr > 0 && !own) {
    for (int k = 0; k < RK; ++k) v[k] = 0.0;
    double w[RK];
    for (int k = 0; k < RK; ++k) w[k] = 0.0;
    slice(G.nlat, part, np, lo, hi);
    for (int p = lo + lt; p < hi; p += TPR) {
      const int ia = G.pairs[2 * p], ib = G.pairs[2 * p + 1];
      const bool ca = G.rrank[ia] >= 0, cb = G.rrank[ib] >= 0;
      float ma[3] = {0, 0, 0}, mb[3] = {0, 0, 0};
      if (ca) {
        majority<NC>(sp, G, C.clsdat, ia, frame, key0, key1, ma);
        v[0] += ma[0]; v[1] += ma[1]; v[2] += ma[2];
        w[0] += 1.0;
      }
      if (cb) {
        majority<NC>(sp, G, C.clsdat, ib, frame, key0, key1, mb);
        v[3] += mb[0]; v[4] += mb[1]; v[5] += mb[2];
        w[1] += 1.0;
      }
      if (ca && cb) {
        v[6] += (double)(ma[0] * mb[0] + ma[1] * mb[1] + ma[2] * mb[2]);
        w[2] += 1.0;
      }
    }
    slice(G.nr, part, np, lo, hi);
    for (int ri = lo + lt; ri < hi; ri += TPR) {
      const int j = G.rorb[ri];
      float m[3];
      majority<NC>(sp, G, C.clsdat, j, frame, key0, key1, m);
      float cc = 0.0f;
      for (int l = 0; l < G.zmax; ++l) {
        const uint32_t wd = G.rlink[(size_t)ri * G.zmax + l];
        const int t = (int)(wd & 0xFFFFFFu);
        // wide graphs: the block-spin link takes the coupling of the site's
        // link l; a set high byte marks an absent link
        if (G.lcid && (wd >> 24)) continue;
        const float* J = C.coup + (G.lcid ? G.lcid[(size_t)j * G.zmax + l] : (int)(wd >> 24)) * 9;
        float mt[3];
        majority<NC>(sp, G, C.clsdat, t, frame, key0, key1, mt);
        if constexpr (NC == 1) {
          cc += J[1] * m[0] * mt[0];
        } else {
          float H0 = J[0] * mt[0] + J[1] * mt[1] + J[2] * mt[2];
          float H1 = J[3] * mt[0] + J[4] * mt[1] + J[5] * mt[2];
          float H2 = J[6] * mt[0] + J[7] * mt[1] + J[8] * mt[2];
          cc += m[0] * H0 + m[1] * H1 + m[2] * H2;
        }
      }
      w[3] += (double)cc;
      // on-site part with the site's actual spin (heisenbergLib.c:798-800,
      // isingLib.c:390)
      float s[3];
      spin_vec<NC>(sp, j, G.cls, C.clsdat, s);
      const int ci = G.cls[j];
      float eo;
      if constexpr (NC == 1)
        eo = -h * s[0];
      else
        eo = C.clsdat[ci * 8 + 1] * s[0] * s[0] + C.clsdat[ci * 8 + 2] * s[1] * s[1] +
             C.clsdat[ci * 8 + 3] * s[2] * s[2] - h * (NC == 3 ? s[2] : s[0]);
      w[4] += (double)eo;
    }
    group_sum<7>(v, C.s_red, C.g, TPR);
    if (lt == 0)
      for (int k = 0; k < 7; ++k) F[16 + k] = v[k];
    group_sum<5>(w, C.s_red, C.g, TPR);
    if (lt == 0)
      for (int k = 0; k < 5; ++k) F[24 + k] = w[k];
  }
  // --- orbital groups (heisenbergLib.c:805-830): 2 groups per reduction
  const int ng = (flags & MCG_MEAS_GROUPS) && !own ? G.ngroup : 0;
  for (int g0 = 0; g0 < ng; g0 += 2) {
    for (int k = 0; k < RK; ++k) v[k] = 0.0;
    slice(G.gmax, part, np, lo, hi);
    for (int q = 0; q < 2 && g0 + q < ng; ++q) {
      for (int k = lo + lt; k < hi; k += TPR) {
        const int o = G.groups[(g0 + q) * G.gmax + k];
        if (o < 0) continue;
        float s[3];
        spin_vec<NC>(sp, o, G.cls, C.clsdat, s);
        v[3 * q] += s[0]; v[3 * q + 1] += s[1]; v[3 * q + 2] += s[2];
      }
    }
    group_sum<6>(v, C.s_red, C.g, TPR);
    if (lt == 0)
      for (int k = 0; k < 6; ++k) F[32 + 3 * g0 + k] = v[k];
  }
}

// One measured frame of the group's replica at `row`: the spin-derived sums
// (frame_sums), then frame_apply.
template <int NC>
__device__ void frame_apply(const FrameCtx& C, const mcg_batch& B, const double* st, int row, int64_t frame,
                            int flags, double mprev, bool have_prev, bool defer);
template <int NC, typename SV, bool FULL = true>
__device__ void measure_frame(const SV& sp, const FrameCtx& C, const mcg_batch& B, const double* st, int row,
                              int64_t frame, uint32_t key0, uint32_t key1, float h, double mprev = 0.0,
                              bool have_prev = false, bool defer = false) {
  frame_sums<NC, SV, FULL>(sp, C, B, frame, key0, key1, h, 0, 1);
  frame_apply<NC>(C, B, st, row, frame, FULL ? B.meas : (B.meas & ~MEAS_SPIN_OBS), mprev, have_prev, defer);
}

// This frame's contribution to every FP64 accumulator of the row, from the
// tracked state st and the spin-derived sums in C.s_frm (valid in the
// group's first thread); called by every thread of the group.
template <int NC>
__device__ void frame_apply(const FrameCtx& C, const mcg_batch& B, const double* st, int row, int64_t frame,
                            int flags, double mprev, bool have_prev, bool defer) {
  const mcg_graph& G = *C.G;
  const int lt = C.lt;
  double* F = C.s_frm;
  // --- leader: this frame's contribution to every FP64 accumulator of the
  //     row (isingLib.c:388-431, heisenbergLib.c:677-830 / xyLib.c:605-757),
  //     staged in LDS; the group's threads then apply them to global memory
  //     in parallel (one read-modify-write each instead of a serial chain)
  double* Cn = F + FRM_ACC;  // [0, NACC) accumulators, [NACC, +NBINQ) bin, then bin index
  const int ngq = frame_groups<NC>(G, flags);
  if (lt == 0) frame_terms<NC>(G, B, st, F, flags, frame, Cn);
  __syncthreads();
  // A row's accumulators are written by whichever replica holds the row, so
  // within one persistent launch by workgroups on different XCDs: the
  // updates go to the device coherence point (device-scope atomics).  M' is
  // the one value read back: its store must be complete before the row's
  // next frame (a workgroup barrier or the launch boundary orders it).
  if (lt == MCG_A_MM) {
    double* A = B.acc + (size_t)row * MCG_NACC;
    const double prev = have_prev ? mprev : dev_load(A + MCG_A_MPREV);
    const double Mf = Cn[MCG_A_MM];
    Cn[MCG_A_MM] = prev * Mf;
    dev_store(A + MCG_A_MPREV, Mf);
  }
  if (!defer) flush_frame<NC>(C, B, row, flags);
}

// full FP64 recomputation of one replica (group-parallel): out[0] E,
// out[1..3] M, out[4] the summed relative spin-length deviation
// sum_i | |s_i|^2 / S_i^2 - 1 | (0 for Ising)
static constexpr int NREC = 5;
template <int NC, typename SV>
__device__ void recompute(const SV& sp, const mcg_graph& G, const float* coup, const float* clsdat, float h,
                          double* s_red, int g, int lt, int TPR, double* out, int part = 0, int np = 1) {
  double v[RK];
  for (int k = 0; k < RK; ++k) v[k] = 0.0;
  int lo, hi;
  slice(G.n, part, np, lo, hi);
  for (int i = lo + lt; i < hi; i += TPR) {
    v[0] += (double)site_energy<NC>(sp, i, G.link, G.lcid, G.zmax, coup, G.cls, clsdat, h);
    float s[3];
    spin_vec<NC>(sp, i, G.cls, clsdat, s);
    v[1] += s[0];
    v[2] += s[1];
    v[3] += s[2];
    if constexpr (NC > 1) {
      const double l2 = (double)s[0] * s[0] + (double)s[1] * s[1] + (double)s[2] * s[2];
      v[4] += fabs(fma(l2, (double)clsdat[G.cls[i] * 8 + 4], -1.0));
    }
  }
  group_sum<NREC>(v, s_red, g, TPR);
  if (lt == 0)
    for (int k = 0; k < NREC; ++k) out[k] = v[k];
}

// Health checks on a recompute (called by the thread holding its result;
// drift = |E_tracked - E_full|): NaN / Inf in the spins makes E or M
// non-finite; spin lengths whose mean relative deviation exceeds the
// storage's tolerance (FP32 keeps |s|^2 to ~1e-7 by the renormalisation,
// FP16 rounds each component to 11 bits) mean the renormalisation no longer
// holds; a drift above the storage's tolerance per site (FP32 tracking
// wanders by ~1e-9 per site per sweep, FP16 by the untracked rounding of
// stored spins) means the spins changed behind the sweep's back.  Each
// raises a bit of the device error flag.
template <typename ST>
__device__ __forceinline__ void recompute_checks(const mcg_batch& B, const mcg_graph& G, const double* r,
                                                 double drift) {
  const bool f16 = sizeof(ST) == 2;
  if (!isfinite(r[0] + r[1] + r[2] + r[3] + r[4])) {
    atomicOr(B.err, MCG_ERR_NONFINITE);
    return;
  }
  if (r[4] > (f16 ? 1e-2 : 1e-4) * G.n) atomicOr(B.err, MCG_ERR_LENGTH);
  if (drift > (f16 ? 1e-2 : 1e-4) * G.n) atomicOr(B.err, MCG_ERR_DRIFT);
}

// ---------------------------------------------------------------------------
// k_lds: persistent LDS-resident sweeps
// ---------------------------------------------------------------------------
// One workgroup of the LDS kernel: nrep replicas of one graph, or one part
// of a replica split over nparts workgroups (part, nparts, slot of its
// counter and partial sums, offset of its extra-group table).
struct WorkItem {
  int32_t graph, nrep, off;
  uint8_t part, nparts;
  uint16_t slot;
  int32_t hoff;
};

// what a workgroup needs of each of its replicas, resolved on the host so
// that staging starts one dependent load after the work item
struct ItemRep {
  int64_t spin_off;
  int32_t rep, serial, row0, pad;
};

static constexpr int MAX_PARTS = 4;
static constexpr int SYNC_STRIDE = 16;  // uint32 per replica counter (one 64-byte line)

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

struct SweepArgs {
  int64_t k_begin;      // run-relative index of the first sweep of this launch
  int64_t philox_base;  // added to k for the Philox counter
  int32_t nsweeps;      // 1 whenever the launch has split replicas
  int32_t nthermal;     // run-relative: sweeps < nthermal are thermal
  int32_t spf;          // sweeps per frame
  int32_t frame0;       // index of the first frame of the run
  uint32_t* sync;       // split replicas: arrival counters, zero at run start
  double* xpart;        // split replicas: per-part (dE, dM) partial sums
  const int32_t* halo;  // split replicas: extra-group tables
  // persistent launches (many sweeps, tempering inside the kernel)
  uint32_t* bar;             // grid barrier counter, zero at run start
  int32_t bar_base;          // barriers passed before this launch
  int32_t pt_every;          // tempering period in sweeps (0: none)
  int32_t has_split;         // the launch holds split replicas
  int32_t nladder;
  const int32_t* lad_row0;
  const int32_t* lad_n;
  const int32_t* blk_lad;    // per workgroup: [nblocks+1] offsets, then ladder
                             // indices (bit 31: the workgroup owns the ladder)
  const uint8_t* rep_np;     // parts of each replica (1: whole)
  double* eslot;             // [2][nrep][MAX_PARTS] published energies
  double* mslot;             // [2][nrep][MAX_PARTS][3] published magnetisation changes
  void* xbuf;                // [2][nslots][MAX_PARTS][max_export][NC] exported sites
  double* mprev3;            // [3][nrow] M' of each row's last frames (frame f in slot f % 3)
  double* fpart;             // [nslots][MAX_PARTS][FSUMS] split replicas' partial frame sums
                             // (spin-derived observables); their arrival count is sync[slot]
  void* fspin;               // [2][nslots][NC][fspin_n] split replicas' sites as published for a
                             // frame, by sweep parity: the next sweep's publication must not
                             // overwrite sites a slower part still reads
  int32_t fspin_n;
  int32_t max_export;
  int32_t nslots;
  int32_t rec_host;          // the full recompute is the host's (a launch ends at each
                             // recompute sweep, then k_measure recomputes every replica)
};

// The schedules of the full recompute and of tempering follow the global
// sweep index (Philox counter), not the index within a call, so that a run
// split into calls (frames, checkpoints) does the same thing at the same
// sweeps as one call.
__host__ __device__ inline bool recompute_due(const mcg_batch& B, int64_t global_sweep) {
  return B.recompute_every > 0 && ((global_sweep + 1) % B.recompute_every) == 0;
}

// LDS of the tempering state a persistent launch keeps in every workgroup:
// who and accepted-swap counts per row, rung per replica, beta and h per row
// (and the workgroup's ladder list: at most one ladder per replica, so at
// most MAX_LADW entries of (ladder | owner bit, first row, rungs))
static constexpr int MAX_LADW = 16;
__host__ __device__ inline size_t persist_lds_bytes(int nrow, int nrep) {
  return align16((size_t)(2 * nrow + nrep) * 4) + align16((size_t)2 * nrow * 4) + align16((size_t)MAX_LADW * 3 * 4) +
         align16((size_t)nrep);
}

static constexpr int FRM = 96 + 48;  // frame scratch doubles per group: scalars, then accumulator staging


// copy one replica between HBM (16-byte aligned) and LDS in 16-byte vectors,
// four in flight per thread
typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));

template <typename ST, bool IN>
__device__ __forceinline__ void stage(ST* lds, ST* g, int count, int lt, int TPR) {
  const int nbytes = count * (int)sizeof(ST);
  const int nvec = nbytes >> 4;
  v4u32* lv = (v4u32*)lds;
  GAS v4u32* gv = (GAS v4u32*)gas(g);
  int t = lt;
  for (; t + 3 * TPR < nvec; t += 4 * TPR) {
    if constexpr (IN) {
      const v4u32 a0 = gv[t], a1 = gv[t + TPR], a2 = gv[t + 2 * TPR], a3 = gv[t + 3 * TPR];
      lv[t] = a0; lv[t + TPR] = a1; lv[t + 2 * TPR] = a2; lv[t + 3 * TPR] = a3;
    } else {
      const v4u32 a0 = lv[t], a1 = lv[t + TPR], a2 = lv[t + 2 * TPR], a3 = lv[t + 3 * TPR];
      gv[t] = a0; gv[t + TPR] = a1; gv[t + 2 * TPR] = a2; gv[t + 3 * TPR] = a3;
    }
  }
  for (; t < nvec; t += TPR) {
    if constexpr (IN) lv[t] = gv[t];
    else gv[t] = lv[t];
  }
  unsigned char* lb = (unsigned char*)lds;
  GAS unsigned char* gb = (GAS unsigned char*)gas(g);
  for (int e = (nvec << 4) + lt; e < nbytes; e += TPR) {
    if constexpr (IN) lb[e] = gb[e];
    else gb[e] = lb[e];
  }
}

__host__ __device__ inline size_t lds_fixed_bytes(int ncoup, int ncls, int ncol, int npat, int zmax, int n,
                                                  int nwaves, int R) {
  size_t b = align16((size_t)(ncoup * 9 + ncls * 8) * 4) + align16((size_t)(ncol + 1) * 4);
  if (npat > 0) b += align16((size_t)npat * pat_stride(zmax) * 4) + align16((size_t)n);
  return b + align16((size_t)nwaves * RK * 8) + align16((size_t)R * (MCG_NSTATE + FRM) * 8) + align16((size_t)R * 16);
}

__host__ __device__ inline size_t lds_replica_bytes(int nc, int n, int ssize) {
  return align16((size_t)nc * n * ssize);
}

// Split replicas.  A replica too large for one CU's share of a sweep runs as
// P workgroups (parts), each holding the whole replica in LDS.  Part p owns
// the Philox groups [ngrp*p/P, ngrp*(p+1)/P) of every colour class.  Its own
// sites need, when they are updated, the current values of their neighbours;
// those of lower colour that another part owns are recomputed here too: the
// host lists, per part and colour, the extra groups that close the own
// sites under "neighbour of lower colour".  Every part starts the sweep from
// the same spins and draws the same Philox words, so the recomputed sites
// come out bit for bit as their owner computes them, and no data moves
// between parts during a sweep.  A launch with split replicas does one
// sweep.  At its end each part publishes its own sites and its partial
// sums (dE, dM) at the device coherence point and counts itself on the
// replica's counter; the part that arrives last reads the other parts' sites
// back, combines the sums and does the measurement.  Nothing waits.
//
// Device coherence: the parts may sit on different XCDs, whose L2 caches are
// not coherent with each other for ordinary accesses.  Published data goes
// through device-scope (agent) atomic stores and loads, which the compiler
// emits at the device coherence point (sc1), and each wave waits for its
// stores to be acknowledged (s_waitcnt 0) before the workgroup barrier that
// precedes the counter increment.
__device__ __forceinline__ void part_range(int b, int cnt, int ngrp, int gd, int p, int np, int& e0, int& e1) {
  const int qa = (int)((int64_t)ngrp * p / np), qb = (int)((int64_t)ngrp * (p + 1) / np);
  e0 = b + qa * gd;
  e1 = b + min(qb * gd, cnt);
}

// every store this wave issued has been acknowledged
__device__ __forceinline__ void stores_done() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Grid-wide barrier of a persistent launch (every workgroup resident), in
// two levels, because arrivals on one counter serialise at its memory
// channel (256 workgroups on one counter: ~3.9 us per barrier on MI355X;
// 16 counters of 16 plus a top counter: ~1.6 us).  Workgroup b counts in
// on group counter b % BAR_GROUPS; the group's last arrival counts in on the
// top counter; the last of those writes the barrier's epoch (its index,
// from 1) to every group's release word, which the group's workgroups poll.
// All words sit in their own 128-byte lines.  Bounded: on time-out bit 16
// of the error flag is raised and false returned, and the caller leaves the
// kernel.
static constexpr int BAR_GROUPS = 16;
static constexpr int BAR_LINE = 32;  // uint32 per line
static constexpr int BAR_WORDS = (1 + 2 * BAR_GROUPS) * BAR_LINE;
// (grid_arrive then grid_wait: work placed between them overlaps the wait)
__device__ __forceinline__ void grid_arrive(uint32_t* bar, uint32_t epoch) {
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nb = (int)gridDim.x;
    const int grp = (int)blockIdx.x % BAR_GROUPS;
    const int ngrp = nb < BAR_GROUPS ? nb : BAR_GROUPS;
    const uint32_t gsz = (uint32_t)((nb - grp + BAR_GROUPS - 1) / BAR_GROUPS);
    uint32_t* gc = bar + (1 + grp) * BAR_LINE;
    const uint32_t old = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1 == epoch * gsz) {
      const uint32_t o2 = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (o2 + 1 == epoch * (uint32_t)ngrp)
        for (int q = 0; q < ngrp; ++q) dev_store(bar + (1 + BAR_GROUPS + q) * BAR_LINE, epoch);
    }
  }
}
__device__ __forceinline__ bool grid_wait(uint32_t* bar, uint32_t epoch, int32_t* err, int* s_flag) {
  if (threadIdx.x == 0) {
    uint32_t* rel = bar + (1 + BAR_GROUPS + (int)blockIdx.x % BAR_GROUPS) * BAR_LINE;
    int ok = 1;
    uint32_t it = 0;
    while ((int32_t)(dev_load(rel) - epoch) < 0) {
      if (++it > (1u << 22)) {
        atomicOr(err, MCG_ERR_BARRIER);
        ok = 0;
        break;
      }
    }
    *s_flag = ok;
  }
  __syncthreads();
  return *s_flag != 0;
}

template <int S>
struct RawBits;
template <>
struct RawBits<1> {
  typedef uint8_t T;
};
template <>
struct RawBits<2> {
  typedef uint16_t T;
};
template <>
struct RawBits<4> {
  typedef uint32_t T;
};

// copy the sites of part p (every colour) between LDS and global memory, as
// raw bit patterns of the storage type (ordinary accesses; the caller
// orders them with fences where another workgroup reads them)
template <typename ST, bool OUT>
__device__ void part_copy(ST* lds_s, ST* gsp, int n, int nc, const int32_t* colptr, int ncol, int gd, int p, int np,
                          int lt, int TPR) {
  typedef typename RawBits<sizeof(ST)>::T U;
  U* lds = (U*)lds_s;
  GAS U* g = (GAS U*)gas((U*)gsp);
  for (int c = 0; c < ncol; ++c) {
    const int b = colptr[c], cnt = colptr[c + 1] - b;
    const int ngrp = (cnt + gd - 1) / gd;
    int e0, e1;
    part_range(b, cnt, ngrp, gd, p, np, e0, e1);
    for (int k = 0; k < nc; ++k) {
      U* l = lds + (size_t)k * n;
      GAS U* q = g + (size_t)k * n;
      for (int e = e0 + lt; e < e1; e += TPR) {
        if constexpr (OUT) q[e] = l[e];
        else l[e] = q[e];
      }
    }
  }
}

// the sites of part p (every colour) from LDS (component stride n) to a
// copy of the replica in global memory (component stride dn) with
// device-scope stores, so that any workgroup of the launch can read them
// (DevSpins) after the next grid barrier.  The copy is not the replica's
// spin array: other parts still stage from that at the launch start.
template <typename ST>
__device__ void part_publish(const ST* lds_s, typename RawBits<sizeof(ST)>::T* dst, int n, int dn, int nc,
                             const int32_t* colptr, int ncol, int gd, int p, int np, int lt, int TPR) {
  typedef typename RawBits<sizeof(ST)>::T U;
  const U* lds = (const U*)lds_s;
  for (int c = 0; c < ncol; ++c) {
    const int b = colptr[c], cnt = colptr[c + 1] - b;
    const int ngrp = (cnt + gd - 1) / gd;
    int e0, e1;
    part_range(b, cnt, ngrp, gd, p, np, e0, e1);
    for (int k = 0; k < nc; ++k)
      for (int e = e0 + lt; e < e1; e += TPR) dev_store(dst + (size_t)k * dn + e, lds[(size_t)k * n + e]);
  }
}

// one replica's spins in global memory read with device-scope loads: the
// sites other workgroups of the same launch published with part_publish
template <typename ST>
__device__ __forceinline__ float raw_spin(typename RawBits<sizeof(ST)>::T u);
template <>
__device__ __forceinline__ float raw_spin<float>(uint32_t u) { return __uint_as_float(u); }
template <>
__device__ __forceinline__ float raw_spin<__half>(uint16_t u) { return __half2float(__ushort_as_half(u)); }
template <>
__device__ __forceinline__ float raw_spin<int8_t>(uint8_t u) { return (float)(int8_t)u; }
template <typename ST>
struct DevSpins {
  typedef ST value_type;
  typedef typename RawBits<sizeof(ST)>::T U;
  const U* p;
  int n;
  __device__ __forceinline__ float get(int i, int c) const { return raw_spin<ST>(dev_load(p + (size_t)c * n + i)); }
};

// ints per part in the split tables: (start, count) of the extra groups per
// colour, (start, count) of the export list, of the correlation pairs and
// of the circuits whose first site the part owns, 1 if the part owns the
// last site, then 1 if no pair or circuit of the replica has sites of two
// parts (the frame then needs only each part's own sites)
__host__ __device__ inline int halo_desc_ints(int ncol) { return 2 * ncol + 8; }

// The leftover groups of a workgroup whose replicas each have one wave
// (64 threads per replica, R replicas, every colour with the same number of
// Philox groups): each replica's last, partial round of rem groups goes to
// a pool of R*rem groups dealt over the workgroup's first threads, so that
// it costs ceil(R*rem/64) wave-rounds instead of R.  Thread t takes group
// ngrp-rem + t%rem of replica t/rem (pr; -1: no pool group).  A group's
// result does not depend on the thread that takes it; the ΔE, ΔM of the
// pool groups are summed per replica in a fixed order (pool_sums).
template <typename ST>
struct PoolInfo {
  int rem = 0;             // 0: no pool
  int pr = -1, pq = 0;     // this thread's pool group: replica pr (group index), group ngrp-rem+pq
  const float4* par = nullptr;  // per group: beta, h, key1 bits
  ST* spin0 = nullptr;     // the workgroup's replicas in LDS, stride sstride
  int sstride = 0;
  double* pp = nullptr;    // pool partials [NC+1][nr], nr = R*rem
  int nr = 0;
};

struct PartInfo {
  int part = 0, np = 1;            // this workgroup's part of the replica
  const int32_t* hx = nullptr;     // extra groups: [2c] start, [2c+1] count per colour, into halo
  const int32_t* halo = nullptr;
};

// end-of-sweep work after the state is current: full recompute (drift
// check), energy reference at the end of thermalisation, measured frame
template <int NC, typename ST, bool FULL, typename SV>
__device__ __forceinline__ void tail_rest(const SV& sv, const mcg_graph& G, const mcg_batch& B, const float* s_coup,
                                          const float* s_cls, double* s_red, double* st, FrameCtx& C, int g, int lt,
                                          int TPR, int row, float h, int64_t k, const SweepArgs& a, bool meas,
                                          bool therm_end, bool rec, uint32_t key0, uint32_t key1,
                                          double mprev = 0.0, bool have_prev = false, bool defer = false) {
  if (rec) {
    double r4[NREC];
    recompute<NC>(sv, G, s_coup, s_cls, h, s_red, g, lt, TPR, r4);
    if (lt == 0) {
      const double d = fabs(st[MCG_S_E] - r4[0]);
      if (d > st[MCG_S_DRIFT]) st[MCG_S_DRIFT] = d;
      st[MCG_S_E] = r4[0];
      st[MCG_S_M] = r4[1];
      st[MCG_S_M + 1] = r4[2];
      st[MCG_S_M + 2] = r4[3];
      recompute_checks<ST>(B, G, r4, d);
    }
  }
  if (therm_end && lt == 0) st[MCG_S_EREF] = st[MCG_S_E];
  __syncthreads();
  if (meas)
    measure_frame<NC, SV, FULL>(sv, C, B, st, row, (k - a.nthermal + 1) / a.spf - 1 + a.frame0, key0, key1, h, mprev,
                                have_prev, defer);
  __syncthreads();
}

// FP64 value of lane src (any lane) in every lane
__device__ __forceinline__ double bperm_f64(double x, int src) {
  const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(x));
  const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(x));
  return __hiloint2double(hi, lo);
}

// One exported site of a split replica: its NC components (raw bits, SoA in
// LDS) as one device-scope store when they fill 4 or 8 bytes, else one per
// component.
template <int NC, typename U>
__device__ __forceinline__ void export_site(U* dst, const U* lraw, int n, int site) {
  constexpr int W = NC * (int)sizeof(U);
  if constexpr (W == 8 || W == 4) {
    typedef typename std::conditional<W == 8, uint64_t, uint32_t>::type P;
    P w = 0;
#pragma unroll
    for (int q = 0; q < NC; ++q) w |= (P)lraw[(size_t)q * n + site] << (q * 8 * sizeof(U));
    dev_store((P*)dst, w);
  } else {
#pragma unroll
    for (int q = 0; q < NC; ++q) dev_store(dst + q, lraw[(size_t)q * n + site]);
  }
}
// The other parts' exported sites, into this part's LDS copy (replicas
// split into three or more parts).  The parts' lists form one round-robin:
// thread lt takes flat indices lt, lt + TPR, ... over the lists of parts
// 0..np-1 (its own list has no entries here), so that it imports about
// (sites of all lists) / TPR sites in all instead of working through every
// part's list in turn, one device-scope round trip per list.  K sites are
// loaded before any is stored (K = 1 measured faster than 4 on the
// honeycomb set).  xs: part 0's export area of this replica and sweep
// parity; parts are max_export * NC words apart.  dsc: this part's
// descriptor; part p's is (p - part) * ds ints away.
template <int NC, typename U>
__device__ __forceinline__ void import_parts(const U* xs, int max_export, const int32_t* halo, const int32_t* dsc,
                                             int ds, int ncol, int part, int np, U* lraw, int n, int lt, int TPR) {
  constexpr int W = NC * (int)sizeof(U);
  constexpr bool PACKED = W == 8 || W == 4;
  typedef typename std::conditional<W == 8, uint64_t, uint32_t>::type P;
  constexpr int K = 1;
  int ofs[MAX_PARTS], lst[MAX_PARTS], tot = 0;
#pragma unroll
  for (int p = 0; p < MAX_PARTS; ++p) {
    const int32_t* d = dsc + (p - part) * ds;
    const bool other = p < np && p != part;
    ofs[p] = tot;
    lst[p] = other ? d[2 * ncol] : 0;
    tot += other ? d[2 * ncol + 1] : 0;
  }
  for (int j0 = lt; j0 < tot; j0 += K * TPR) {
    P w[K];
    U v[K][NC];
    int site[K];
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int j = j0 + u * TPR;
      site[u] = -1;
      if (j < tot) {
        // the part of flat index j: the last one whose list starts at or
        // before j (a part with an empty list starts where the next one does)
        int p = 0;
#pragma unroll
        for (int q = 1; q < MAX_PARTS; ++q)
          if (ofs[q] <= j) p = q;
        const int i = j - ofs[p];
        const U* src = xs + ((size_t)p * max_export + i) * NC;
        site[u] = halo[lst[p] + i];
        if constexpr (PACKED) {
          w[u] = dev_load((const P*)src);
        } else {
#pragma unroll
          for (int q = 0; q < NC; ++q) v[u][q] = dev_load(src + q);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < K; ++u) {
      if (site[u] < 0) continue;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        if constexpr (PACKED)
          lraw[(size_t)q * n + site[u]] = (U)(w[u] >> (q * 8 * sizeof(U)));
        else
          lraw[(size_t)q * n + site[u]] = v[u][q];
      }
    }
  }
}

// Frame f of the replica of lane g (persistent launches): its terms are
// staged for flush_frame; mp is the row's previous M', slot (f-1) % 3 of
// a.mprev3 (frame_mprev3), and the frame's own M goes to slot f % 3.
template <int NC>
__device__ __forceinline__ double frame_mprev3(const SweepArgs& a, int64_t f, int row, int nrow) {
  return dev_load(a.mprev3 + (size_t)((f + 2) % 3) * nrow + row);
}
template <int NC>
__device__ __forceinline__ void take_frame(const mcg_graph& G, const mcg_batch& B, const SweepArgs& a, double* s_st,
                                           double* s_frm, int lane, int64_t f, int row, int nrow, double mp,
                                           int flags) {
  double* F = s_frm + lane * FRM;
  double* Cn = F + FRM_ACC;
  frame_terms<NC>(G, B, s_st + lane * MCG_NSTATE, F, flags, f, Cn);
  const double Mf = Cn[MCG_A_MM];
  Cn[MCG_A_MM] = mp * Mf;
  dev_store(a.mprev3 + (size_t)(f % 3) * nrow + row, Mf);
}

// Split replica, measured frame with spin-derived observables: each part's
// share F[0, FSUMS) of the sums (frame_sums) goes to a.fpart and counts in
// on the replica's counter a.sync[slot]; part 0 waits for the other parts of
// this frame (the counter counts from the run's first frame, a.frame0), then
// adds their sums to its own in part order, so the result does not depend on
// arrival order.  Bounded like the grid barrier.  Called by every thread of
// the part's workgroup; false: time-out (the caller leaves the kernel).
// A part overwrites its slot only after the next grid barrier, which part 0
// passes after reading it.
__device__ bool gather_frame_sums(const SweepArgs& a, const mcg_batch& B, double* F, int slot, int part, int np,
                                  int64_t frame, int* s_flag) {
  double* fp = a.fpart + (size_t)slot * MAX_PARTS * FSUMS;
  uint32_t* cnt = a.sync + (size_t)slot * SYNC_STRIDE;
  __syncthreads();  // F is written by the first thread
  if (part != 0) {
    if (threadIdx.x < 64)
      for (int q = threadIdx.x; q < FSUMS; q += 64) dev_store(fp + (size_t)part * FSUMS + q, F[q]);
    stores_done();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  if (threadIdx.x == 0) {
    const uint32_t want = (uint32_t)((frame - a.frame0 + 1) * (int64_t)(np - 1));
    int ok = 1;
    uint32_t it = 0;
    while ((int32_t)(dev_load(cnt) - want) < 0) {
      if (++it > (1u << 22)) {
        atomicOr(B.err, MCG_ERR_BARRIER);
        ok = 0;
        break;
      }
    }
    *s_flag = ok;
  }
  __syncthreads();
  if (!*s_flag) return false;
  if (threadIdx.x < 64)
    for (int q = threadIdx.x; q < FSUMS; q += 64) {
      // the parts' values loaded together, then added in part order
      double y[MAX_PARTS];
#pragma unroll
      for (int p = 1; p < MAX_PARTS; ++p) y[p] = p < np ? dev_load(fp + (size_t)p * FSUMS + q) : 0.0;
      double x = F[q];
#pragma unroll
      for (int p = 1; p < MAX_PARTS; ++p)
        if (p < np) x += y[p];
      F[q] = x;
    }
  __syncthreads();
  return true;
}

// End of a persistent sweep for a workgroup of unsplit replicas: the state
// update (and published energy) of every replica is done by the
// workgroup's first wave, lane g for group g, so that this serial
// per-replica work runs once per workgroup instead of once in every group.
// The sums are group_sum's (each wave's DPP total, then the same DPP steps
// over a group's waves).  rep_l: lane g's replica; eslot_par: the energy
// slots of this sweep's parity, or nullptr to publish nothing.  Called by
// every thread; ends with a workgroup barrier.
// Pool partials (PoolInfo): thread t's ΔE, ΔM of its pool groups are kept
// in pp[k*nr + t] (nr = R*rem, k < NC+1; pp is the frame area, free between
// a sweep's frames and the next tail); pool_sums gives replica g the sum of
// its rem entries in thread order.
template <int NC>
__device__ __forceinline__ void pool_sums(const double* pp, int nr, int rem, int g, double* out) {
  for (int k = 0; k <= NC; ++k) {
    double x = 0.0;
    for (int i = 0; i < rem; ++i) x += pp[k * nr + g * rem + i];
    out[k] = x;
  }
}

// The sweep's ΔE, ΔM of every group, in the first wave: lane g gets group
// g's totals in t (the other waves get zeros).  The sums are group_sum's
// (each wave's DPP total, then the same DPP steps over a group's waves).
// Called by every thread; s_red is free again only after the caller's next
// workgroup barrier.
template <int NC>
__device__ __forceinline__ void wave0_sums(double* s_red, int TPR, float fE, const float* fM, double* t) {
  constexpr int NV = NC + 1;  // E and the spin components (the rest stay 0)
  double v[4] = {(double)fE, (double)fM[0], (double)fM[1], (double)fM[2]};
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum_f64(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < NV; ++k) s_red[wave * RK + k] = v[k];
  __syncthreads();
  t[0] = t[1] = t[2] = t[3] = 0.0;
  if (wave == 0) {
    const int nw = TPR >> 6, nwaves = (int)blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double x = lane < nwaves ? s_red[lane * RK + k] : 0.0;
      if (nw >= 2) x += dpp_f64<0xB1, 0xF>(x);   // quad_perm [1,0,3,2]
      if (nw >= 4) x += dpp_f64<0x4E, 0xF>(x);   // quad_perm [2,3,0,1]
      if (nw >= 8) x += dpp_f64<0x141, 0xF>(x);  // row_half_mirror
      if (nw >= 16) x += dpp_f64<0x140, 0xF>(x); // row_mirror
      t[k] = bperm_f64(x, lane * nw);
    }
  }
}

template <int NC>
__device__ __forceinline__ void tail_wave0(const mcg_batch& B, double* s_red, double* s_st, int R, int TPR, float fE,
                                           const float* fM, bool therm_end, int rep_l, double* eslot_par,
                                           const double* pp, int prem) {
  constexpr int NV = NC + 1;
  double t[4];
  wave0_sums<NC>(s_red, TPR, fE, fM, t);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave == 0) {
    if (prem > 0) {
      // a replica's sweep change: its groups' sum plus its pool groups' sum
      // (pool_sums' order); lane 16k+g adds term k of replica g, R <= 16
      const int gk = lane & 15, kk = lane >> 4;
      double x = 0.0;
      if (gk < R && kk < NV)
        for (int i = 0; i < prem; ++i) x += pp[kk * R * prem + gk * prem + i];
#pragma unroll
      for (int k = 0; k < NV; ++k) t[k] += bperm_f64(x, 16 * k + (lane & 15));
    }
    if (lane < R) {
      double* st = s_st + lane * MCG_NSTATE;
#pragma unroll
      for (int k = 0; k < NV; ++k) st[MCG_S_E + k] += t[k];
      if (therm_end) st[MCG_S_EREF] = st[MCG_S_E];
      if (eslot_par) dev_store(eslot_par + (size_t)rep_l * MAX_PARTS, st[MCG_S_E]);
    }
  }
  __syncthreads();
}

template <int NC, typename ST, bool FULL, typename SV>
__device__ __forceinline__ void lds_sweep_tail(const SV& sv, const mcg_graph& G, const mcg_batch& B, const float* s_coup,
                                               const float* s_cls, double* s_red, double* st, FrameCtx& C, int g,
                                               int lt, int TPR, int row, float h, int64_t k, const SweepArgs& a,
                                               bool meas, bool therm_end, bool rec, double& dE, double* dM,
                                               uint32_t key0, uint32_t key1, double mprev, bool have_prev,
                                               const PoolInfo<ST>& pl) {
  double v[RK] = {dE, dM[0], dM[1], dM[2], 0, 0, 0, 0};
  group_sum<4>(v, s_red, g, TPR);
  if (lt == 0) {
    if (pl.rem > 0) {
      double ps[4];
      pool_sums<NC>(pl.pp, pl.nr, pl.rem, g, ps);
      for (int k = 0; k <= NC; ++k) v[k] += ps[k];
    }
    st[MCG_S_E] += v[0];
    st[MCG_S_M] += v[1];
    st[MCG_S_M + 1] += v[2];
    st[MCG_S_M + 2] += v[3];
  }

  dE = 0.0;
  dM[0] = dM[1] = dM[2] = 0.0;
  tail_rest<NC, ST, FULL>(sv, G, B, s_coup, s_cls, s_red, st, C, g, lt, TPR, row, h, k, a, meas, therm_end, rec, key0,
                          key1, mprev, have_prev);
}

// one sweep of a workgroup's replica group: every colour class in turn,
// own Philox groups (changes counted in fE, fM) and, for a split replica,
// the other parts' groups this part depends on (changes not counted)
template <int NC, typename ST, int MODE>
__device__ __forceinline__ void sweep_colours(const mcg_graph& G, const Spins<ST>& sp, const float* coup,
                                              const float* clsdat, const int32_t* s_colptr, const uint8_t* s_pid,
                                              const int32_t* s_pat, int lt, int TPR, Acc beta, float h, uint64_t kk,
                                              uint32_t key0, uint32_t key1, const PartInfo& pi, float& fE, float* fM,
                                              bool aniso, const PoolInfo<ST>& pl, bool pfresh) {
  constexpr int GD = SitesPerDraw<NC>::G;
  const int zmax = G.zmax;
  const bool split = pi.np > 1;
  // A colour's groups are dealt out in rounds of TPR; the last, partial
  // round falls to the first lanes of the deal.  The deal starts at wave r
  // of the group, chosen so that the groups' partial rounds land on
  // different SIMDs (wave w of a workgroup runs on SIMD w % 4).  Which
  // thread takes a Philox group does not change its result.
  {
    const int nw = TPR >> 6, g = (int)threadIdx.x / TPR;
    const int r = nw >= 4 ? g & 3 : nw == 2 ? (g >> 1) & 1 : 0;
    lt = lt - 64 * r;
    if (lt < 0) lt += TPR;
  }
  for (int c = 0; c < G.ncol; ++c) {
    const int b = s_colptr[c], cnt = s_colptr[c + 1] - b;
    const int ngrp = (cnt + GD - 1) / GD;
    const int q0 = split ? (int)((int64_t)ngrp * pi.part / pi.np) : 0;
    const int q1 = split ? (int)((int64_t)ngrp * (pi.part + 1) / pi.np) : ngrp;
    const int nx = split ? pi.hx[2 * c + 1] : 0;
    const int32_t* xl = split ? pi.halo + pi.hx[2 * c] : nullptr;
    // A split part's recomputed groups continue the round-robin of its own
    // groups: a thread whose own loop ends at q (the first q >= q1 of its
    // deal) takes xl[q - q1], xl[q - q1 + TPR], ...  They fill the part's
    // last own round: a round of them alone would cost about as much as a
    // full one (one wave's dependent chain).
    float xE = 0.0f, xM[3] = {0.0f, 0.0f, 0.0f};
    if constexpr (MODE == 0 || MODE == 3) {
      constexpr bool UNI = MODE == 3;
      const float* cp = UNI ? coup + G.ucid * 9 : coup;
      const int qm = q1 - pl.rem;  // the pool takes the last rem groups (unsplit replicas only)
      int q = q0 + lt;
      for (; q < qm; q += TPR)
        metro_group_pat<NC, ST, UNI>(sp, b, cnt, q, c, s_pid, s_pat, zmax, cp, clsdat, beta, h, kk, key0, key1, fE,
                                     fM, aniso);
      if (pl.pr >= 0) {
        const float4 pv = pl.par[pl.pr];
        float pe = 0.0f, pm[3] = {0.0f, 0.0f, 0.0f};
        const Spins<ST> psp{pl.spin0 + (size_t)pl.sstride * pl.pr, G.n};
        metro_group_pat<NC, ST, UNI>(psp, b, cnt, qm + pl.pq, c, s_pid, s_pat, zmax, cp, clsdat,
                                     beta.at(pv.x, __float_as_int(pv.w)), pv.y, kk, key0,
                                     __float_as_uint(pv.z), pe, pm, aniso);
        // the first pool group after a tail stores, later ones add
        double* pp = pl.pp + threadIdx.x;
        const bool first = pfresh && c == 0;
        pp[0] = first ? (double)pe : pp[0] + (double)pe;
#pragma unroll
        for (int k = 1; k <= NC; ++k) pp[k * pl.nr] = first ? (double)pm[k - 1] : pp[k * pl.nr] + (double)pm[k - 1];
      }
      for (int j = q - qm; j < nx; j += TPR)
        metro_group_pat<NC, ST, UNI>(sp, b, cnt, xl[j], c, s_pid, s_pat, zmax, cp, clsdat, beta, h, kk, key0, key1,
                                     xE, xM, aniso);
    } else if constexpr (MODE == 1) {
      const GAS int32_t* lcid = G.lcid ? gas(G.lcid) : nullptr;
      if (G.ucid >= 0) {
        const float* cp = coup + G.ucid * 9;
        int q = q0 + lt;
        for (; q < q1; q += TPR)
          metro_group<NC, true>(sp, b, cnt, q, c, gas(G.link), lcid, zmax, cp, gas(G.cls), clsdat, beta, h, kk, key0,
                                key1, fE, fM);
        for (int j = q - q1; j < nx; j += TPR)
          metro_group<NC, true>(sp, b, cnt, xl[j], c, gas(G.link), lcid, zmax, cp, gas(G.cls), clsdat, beta, h, kk,
                                key0, key1, xE, xM);
      } else {
        int q = q0 + lt;
        for (; q < q1; q += TPR)
          metro_group<NC>(sp, b, cnt, q, c, gas(G.link), lcid, zmax, coup, gas(G.cls), clsdat, beta, h, kk, key0, key1,
                          fE, fM);
        for (int j = q - q1; j < nx; j += TPR)
          metro_group<NC>(sp, b, cnt, xl[j], c, gas(G.link), lcid, zmax, coup, gas(G.cls), clsdat, beta, h, kk, key0,
                          key1, xE, xM);
      }
    } else {
      const GAS int32_t* lcid = G.lcid ? gas(G.lcid) : nullptr;
      const bool uni = G.ucid >= 0;
      const GAS float* cp = gas(coup) + (uni ? G.ucid * 9 : 0);
      if (uni) {
        int q = q0 + lt;
        for (; q < q1; q += TPR)
          metro_group<NC, true>(sp, b, cnt, q, c, gas(G.link), lcid, zmax, cp, gas(G.cls), gas(clsdat), beta, h, kk,
                                key0, key1, fE, fM);
        for (int j = q - q1; j < nx; j += TPR)
          metro_group<NC, true>(sp, b, cnt, xl[j], c, gas(G.link), lcid, zmax, cp, gas(G.cls), gas(clsdat), beta, h,
                                kk, key0, key1, xE, xM);
      } else {
        int q = q0 + lt;
        for (; q < q1; q += TPR)
          metro_group<NC>(sp, b, cnt, q, c, gas(G.link), lcid, zmax, cp, gas(G.cls), gas(clsdat), beta, h, kk, key0,
                          key1, fE, fM);
        for (int j = q - q1; j < nx; j += TPR)
          metro_group<NC>(sp, b, cnt, xl[j], c, gas(G.link), lcid, zmax, cp, gas(G.cls), gas(clsdat), beta, h, kk,
                          key0, key1, xE, xM);
      }
    }
    __syncthreads();
  }
}

// sweep loop of one LDS block; MODE 0: pattern links (colour-permuted
// replica), 1: global link table with LDS-staged couplings/classes, 2:
// global link table with couplings/classes read from HBM (> 256 tensors or
// classes).  Each mode is its own instantiation so that every table pointer
// has one address space.
template <int NC, typename ST, int MODE, bool FULL>
__device__ __forceinline__ void lds_sweeps(const mcg_graph& G, const mcg_batch& B, const SweepArgs& a,
                                           const Spins<ST>& sp, const float* coup,
                                           const float* clsdat, const int32_t* s_colptr, const uint8_t* s_pid,
                                           const int32_t* s_pat, double* s_red, double* st, FrameCtx& C, int g,
                                           int lt, int TPR, int row, float beta, float h, uint32_t key0,
                                           uint32_t key1, const PartInfo& pi, double& dE, double* dM, bool aniso,
                                           const PoolInfo<ST>& pl) {
  const bool split = pi.np > 1;
  dE = 0.0;
  dM[0] = dM[1] = dM[2] = 0.0;
  bool pfresh = true;  // the next pool group starts the pool partials
  for (int s = 0; s < a.nsweeps; ++s) {
    const int64_t k = a.k_begin + s;
    const uint64_t kk = (uint64_t)(a.philox_base + k);
    const bool meas = k >= a.nthermal && ((k - a.nthermal + 1) % a.spf) == 0;
    double mprev = 0.0;  // read ahead, so that the frame does not wait for it
    if (meas && !split && lt == MCG_A_MM) mprev = frame_mprev(B, row);
    float fE = 0.0f, fM[3] = {0.0f, 0.0f, 0.0f};  // this sweep's accepted changes
    sweep_colours<NC, ST, MODE>(G, sp, coup, clsdat, s_colptr, s_pid, s_pat, lt, TPR, acc_of(B, beta, row), h, kk, key0,
                                key1, pi, fE,
                                fM, aniso, pl, pfresh);
    pfresh = false;
    dE += (double)fE;
    dM[0] += (double)fM[0];
    dM[1] += (double)fM[1];
    dM[2] += (double)fM[2];
    if (split) continue;  // one sweep per launch; the last part to finish runs the tail
    const bool therm_end = (k + 1 == a.nthermal);
    const bool rec = recompute_due(B, a.philox_base + k);
    if (meas || therm_end || rec || s == a.nsweeps - 1) {
      lds_sweep_tail<NC, ST, FULL>(sp, G, B, coup, clsdat, s_red, st, C, g, lt, TPR, row, h, k, a, meas, therm_end,
                                   rec, dE, dM, key0, key1, mprev, true, pl);
      pfresh = true;
    }
  }
}

// MODE: link mode of every graph of the launch (0 pattern, 1 global links
// with LDS tables, 2 wide); FULL: spin-derived observables are measured.
// Both are compile-time so that each instantiation carries only the code it
// runs (instruction-cache misses dominate the per-launch cost otherwise).
// PERSIST: a persistent launch (every workgroup resident, see below).
template <int NC, typename ST, int MODE, bool FULL, bool PERSIST>
__global__ __launch_bounds__(1024) void k_lds(mcg_batch B, const WorkItem* __restrict__ items,
                                              const ItemRep* __restrict__ recs, SweepArgs a) {
  const WorkItem w = items[blockIdx.x];
  const mcg_graph G = B.graphs[w.graph];
  const int R = w.nrep;
  const int TPR = blockDim.x / R;
  const int g = threadIdx.x / TPR;
  const int lt = threadIdx.x - g * TPR;
  const ItemRep ir = recs[w.off + g];
  const int rep = ir.rep;
  const int row = ir.row0 + B.rung[rep];
  const float beta = B.row_beta[row], h = B.row_h[row];
  const int nwaves = blockDim.x / 64;
  const bool perm = MODE == 0 || MODE == 3;
  const int n = G.n, zmax = G.zmax;
  const bool wide = MODE == 2;
  const int ncl = wide ? 0 : G.ncoup;  // couplings staged in LDS
  const int ncs = wide ? 0 : G.ncls;   // site classes staged in LDS

  extern __shared__ __align__(16) unsigned char smem[];
  size_t off = 0;
  float* s_coup = (float*)smem;
  float* s_clsl = s_coup + ncl * 9;
  off += align16((size_t)(ncl * 9 + ncs * 8) * 4);
  int32_t* s_colptr = (int32_t*)(smem + off);
  off += align16((size_t)(G.ncol + 1) * 4);
  int32_t* s_pat = (int32_t*)(smem + off);
  if (perm) off += align16((size_t)G.npat * pat_stride(zmax) * 4);
  uint8_t* s_pid = (uint8_t*)(smem + off);
  if (perm) off += align16((size_t)n);
  double* s_red = (double*)(smem + off);
  off += align16((size_t)nwaves * RK * 8);
  double* s_st = (double*)(smem + off);  // R * NSTATE
  double* s_frm = s_st + R * MCG_NSTATE;  // R * FRM
  off += align16((size_t)R * (MCG_NSTATE + FRM) * 8);
  float4* s_par = (float4*)(smem + off);  // per group: beta, h, key1 bits (the pool's parameters)
  off += align16((size_t)R * 16);
  unsigned char* s_pers = smem + off;  // tempering state of a persistent launch
  if (PERSIST) off += persist_lds_bytes(B.nrow, B.nrep);
  ST* s_spin = (ST*)(smem + off);

  for (int t = threadIdx.x; t < ncl * 9; t += blockDim.x) s_coup[t] = G.coup[t];
  for (int t = threadIdx.x; t < ncs * 8; t += blockDim.x) s_clsl[t] = G.clsdat[t];
  // does any site class carry an on-site term? (metro_core skips them if not)
  int anz = 0;
  for (int t = threadIdx.x; t < G.ncls; t += blockDim.x)
    anz |= (G.clsdat[t * 8 + 1] != 0.0f) | (G.clsdat[t * 8 + 2] != 0.0f) | (G.clsdat[t * 8 + 3] != 0.0f);
  const bool aniso = __syncthreads_or(anz) != 0;
  for (int t = threadIdx.x; t <= G.ncol; t += blockDim.x) s_colptr[t] = G.col_ptr[t];
  if (perm) {
    for (int t = threadIdx.x; t < G.npat * (zmax + 1); t += blockDim.x) {
      const int l = t % (zmax + 1);
      int32_t v = G.pattab[t];
      // uniform graphs: links as byte offsets between spins (PatOffLinks)
      if (MODE == 3 && l < zmax) v = ((int32_t)((uint32_t)v << 8) >> 8) * (int32_t)sizeof(ST);
      s_pat[(t / (zmax + 1)) * pat_stride(zmax) + l] = v;
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) s_pid[t] = G.pat[t];
  }

  ST* gsp = (ST*)B.spins + ir.spin_off;
  ST* my = s_spin + lds_replica_bytes(NC, n, sizeof(ST)) / sizeof(ST) * g;
  stage<ST, true>(my, gsp, NC * n, lt, TPR);
  double* st = s_st + g * MCG_NSTATE;
  if (lt < MCG_NSTATE) st[lt] = B.state[(size_t)rep * MCG_NSTATE + lt];
  if (lt == 0) s_par[g] = make_float4(beta, h, __uint_as_float(B.seed_hi ^ (uint32_t)ir.serial), __int_as_float(row));
  __syncthreads();

  const uint32_t key0 = B.seed_lo, key1 = B.seed_hi ^ (uint32_t)ir.serial;
  const Spins<ST> sp{my, n};
  // pool of the replicas' partial last rounds (PoolInfo)
  PoolInfo<ST> pl;
  if (perm && TPR == 64 && R > 1 && w.nparts <= 1) {
    constexpr int GDp = SitesPerDraw<NC>::G;
    int rem = -1;
    for (int c = 0; c < G.ncol; ++c) {
      const int rc = ((s_colptr[c + 1] - s_colptr[c] + GDp - 1) / GDp) % 64;
      rem = rem < 0 ? rc : rem == rc ? rem : 0;
    }
    if (rem > 0 && rem * 4 <= FRM) pl.rem = rem;  // partials fit the frame area
  }
  if (pl.rem > 0) {
    pl.par = s_par;
    pl.spin0 = s_spin;
    pl.sstride = (int)(lds_replica_bytes(NC, n, sizeof(ST)) / sizeof(ST));
    pl.pp = s_frm;
    pl.nr = R * pl.rem;
    if ((int)threadIdx.x < R * pl.rem) {
      pl.pr = (int)threadIdx.x / pl.rem;
      pl.pq = (int)threadIdx.x - pl.pr * pl.rem;
    }
  }
  PartInfo pi;
  if (w.nparts > 1) {
    pi.part = w.part;
    pi.np = w.nparts;
    pi.hx = a.halo + w.hoff;
    pi.halo = a.halo;
  }
  FrameCtx C{&G, wide ? G.coup : s_coup, wide ? G.clsdat : s_clsl, s_red, s_frm + g * FRM, g, lt, TPR};
  if constexpr (PERSIST) {
    // Persistent launch: many sweeps with the replicas in LDS.  After each
    // sweep every replica's energy is published at the device coherence
    // point (double-buffered by sweep parity); split parts publish their
    // partial sums and the sites another part reads (export lists).  A grid
    // barrier follows whenever tempering is due or parts must exchange.
    // Then every workgroup computes the swap decisions of the ladders its
    // replicas belong to itself, on its own copy of who/rung (identical in
    // all workgroups that share a ladder: same energies, same Philox words),
    // so no second barrier is needed: a workgroup can get at most one sweep
    // ahead, and the parity buffers keep what the slower ones still read.
    // Each ladder's owner writes its tables back at the end.
    typedef typename RawBits<sizeof(ST)>::T U;
    constexpr int GD = SitesPerDraw<NC>::G;
    const int nrow = B.nrow, nrepb = B.nrep;
    int32_t* s_who = (int32_t*)s_pers;
    int32_t* s_ptacc = s_who + nrow;
    int32_t* s_rung = s_ptacc + nrow;
    float* s_beta = (float*)(s_pers + align16((size_t)(2 * nrow + nrepb) * 4));
    float* s_h = s_beta + nrow;
    int32_t* s_lad = (int32_t*)(s_pers + align16((size_t)(2 * nrow + nrepb) * 4) + align16((size_t)2 * nrow * 4));
    uint8_t* s_np = (uint8_t*)s_lad + align16((size_t)MAX_LADW * 3 * 4);  // parts of each replica
    for (int t = threadIdx.x; t < nrow; t += blockDim.x) {
      s_who[t] = B.who[t];
      s_ptacc[t] = 0;
      s_beta[t] = B.row_beta[t];
      s_h[t] = B.row_h[t];
    }
    for (int t = threadIdx.x; t < nrepb; t += blockDim.x) {
      s_rung[t] = B.rung[t];
      s_np[t] = a.rep_np[t];
    }
    // this workgroup's ladders, kept in LDS (read every tempering pass)
    const int nlad = min(a.blk_lad[blockIdx.x + 1] - a.blk_lad[blockIdx.x], MAX_LADW);
    if ((int)threadIdx.x < nlad) {
      const int e = a.blk_lad[a.blk_lad[blockIdx.x] + threadIdx.x];
      const int L = e & 0x7FFFFFFF;
      s_lad[3 * threadIdx.x] = e;
      s_lad[3 * threadIdx.x + 1] = a.lad_row0[L];
      s_lad[3 * threadIdx.x + 2] = a.lad_n[L];
    }
    __syncthreads();
    U* lraw = (U*)my;
    U* xb = (U*)a.xbuf;
    const int DS = halo_desc_ints(G.ncol);
    int* s_flag = (int*)s_red;
    int nbar = 0;
    // Lane g of the first wave stands for group g (unsplit replicas) or for
    // the replica (part 0 of a split one) in the end-of-sweep work.  M' of a
    // row (for <M M'>) passes from frame to frame through slot f % 3 of
    // a.mprev3: frame f of a row is taken after the barrier of its sweep (or
    // before it, see frames_after), reads slot (f-1) % 3 and writes slot
    // f % 3; the writer of frame f-1 stored before it arrived at a barrier
    // the reader has passed, and a workgroup is never more than one sweep
    // ahead, so three slots never collide.  A[MPREV] seeds the slots of the
    // rows held at the start and takes the last frame's M'.
    const int lane = threadIdx.x & 63;
    const int nfr = pi.np == 1 ? R : pi.part == 0 ? 1 : 0;  // groups whose frames this workgroup takes
    const bool l_grp = threadIdx.x < 64 && lane < nfr;
    const int pt0 = blockDim.x >= 128 ? 64 : 0;  // tempering on the waves after the first
    // Frames are taken after the barrier when every sweep has one (split
    // replicas, tempering every sweep); otherwise before it, so that a
    // barrier always separates two frames of a row taken by different
    // workgroups.
    const bool frames_after = a.has_split || a.pt_every == 1;
    const int fflags = FULL ? B.meas : (B.meas & ~MEAS_SPIN_OBS);  // what a frame measures
    // A split part measures the pairs and circuits whose first site it owns
    // from its own copy before the grid barrier, when none of them has
    // sites of two parts and no block spins or orbital groups are measured;
    // its sums reach part 0 through the barrier (a.fpart, by sweep parity).
    // Otherwise every part publishes its sites and, after the barrier, sums
    // its share over the published copy; part 0 waits for the others' sums.
    const bool own_frame = FULL && pi.np > 1 && !(B.meas & (MCG_MEAS_RNORM | MCG_MEAS_GROUPS)) &&
                           pi.hx[2 * G.ncol + 7] != 0;
    int rep_l = 0, row0_l = 0, last_row = -1;
    if (l_grp) {
      const ItemRep irl = recs[w.off + lane];
      rep_l = irl.rep;
      row0_l = irl.row0;
      const int r = row0_l + s_rung[rep_l];
      const double m0 = B.acc[(size_t)r * MCG_NACC + MCG_A_MPREV];
      for (int q = 0; q < 3; ++q) dev_store(a.mprev3 + (size_t)q * nrow + r, m0);
    }
    for (int s = 0; s < a.nsweeps; ++s) {
      const int64_t k = a.k_begin + s;
      const uint64_t kk = (uint64_t)(a.philox_base + k);
      const int par = (int)(k & 1);
      const int rowg = ir.row0 + s_rung[rep];
      const float bg = s_beta[rowg], hg = s_h[rowg];
      const bool meas = k >= a.nthermal && ((k - a.nthermal + 1) % a.spf) == 0;
      const bool therm_end = (k + 1 == a.nthermal);
      const bool rec = !a.rec_host && recompute_due(B, a.philox_base + k);
      const bool pt_now = a.pt_every > 0 && ((a.philox_base + k + 1) % a.pt_every) == 0;
      const int rowg_l = row0_l + s_rung[rep_l];
      const int64_t fcur = (k - a.nthermal + 1) / a.spf - 1 + a.frame0;
      if (pl.rem > 0) {  // the pool's parameters of this sweep (rows move with tempering)
        if (lt == 0) s_par[g] = make_float4(bg, hg, __uint_as_float(key1), __int_as_float(rowg));
        __syncthreads();
      }
      float fE = 0.0f, fM[3] = {0.0f, 0.0f, 0.0f};
      sweep_colours<NC, ST, MODE>(G, sp, C.coup, C.clsdat, s_colptr, s_pid, s_pat, lt, TPR, acc_of(B, bg, rowg), hg, kk,
                                  key0, key1, pi,
                                  fE, fM, aniso, pl, true);
      double v[RK] = {(double)fE, (double)fM[0], (double)fM[1], (double)fM[2], 0.0, 0.0, 0.0, 0.0};
      double* es = a.eslot + ((size_t)par * nrepb + rep) * MAX_PARTS;
      double* ms = a.mslot + ((size_t)par * nrepb + rep) * MAX_PARTS * 3;
      if (pi.np == 1) {
        // state update and published energy before the barrier (with a full
        // recompute due, the recomputed state is published instead)
        tail_wave0<NC>(B, s_red, s_st, R, TPR, fE, fM, therm_end && !rec, rep_l,
                       rec ? nullptr : a.eslot + (size_t)par * nrepb * MAX_PARTS, s_frm, pl.rem);
        if (rec) {
          double r4[NREC];
          recompute<NC>(sp, G, C.coup, C.clsdat, hg, s_red, g, lt, TPR, r4);
          if (lt == 0) {
            const double d = fabs(st[MCG_S_E] - r4[0]);
            if (d > st[MCG_S_DRIFT]) st[MCG_S_DRIFT] = d;
            recompute_checks<ST>(B, G, r4, d);
            for (int q = 0; q < 4; ++q) st[MCG_S_E + q] = r4[q];
            if (therm_end) st[MCG_S_EREF] = st[MCG_S_E];
            dev_store(es, st[MCG_S_E]);
          }
          __syncthreads();
        }
        if constexpr (FULL) {
          // the spin-derived sums of every group's frame, from its LDS
          // replica (the frame area is free after the tail)
          if (meas) frame_sums<NC, Spins<ST>, true>(sp, C, B, fcur, key0, key1, hg, 0, 1);
        }
      } else {
        // one replica: lane 0 of the first wave gets the part's sums
        wave0_sums<NC>(s_red, TPR, fE, fM, v);
        if (threadIdx.x == 0) {
          dev_store(es + pi.part, (pi.part == 0 ? st[MCG_S_E] : 0.0) + v[0]);
          for (int q = 0; q < 3; ++q) dev_store(ms + pi.part * 3 + q, v[1 + q]);
        }
        const int32_t* ex = a.halo + pi.hx[2 * G.ncol];
        const int nex = pi.hx[2 * G.ncol + 1];
        U* xo = xb + (((size_t)par * a.nslots + w.slot) * MAX_PARTS + pi.part) * (size_t)a.max_export * NC;
        for (int i = lt; i < nex; i += TPR)
          export_site<NC>(xo + (size_t)i * NC, lraw, n, ex[i]);
        if constexpr (FULL) {
          if (meas && own_frame) {
            const int32_t* d = pi.hx + 2 * G.ncol;
            const PartFrame pf{a.halo + d[2], a.halo + d[4], d[3], d[5], d[6] != 0};
            frame_sums<NC, Spins<ST>, true>(sp, C, B, fcur, key0, key1, hg, pi.part, pi.np, &pf);
            if (pi.part != 0) {
              __syncthreads();  // F is written by the first thread
              double* fp = a.fpart + (((size_t)par * a.nslots + w.slot) * MAX_PARTS + pi.part) * FSUMS;
              if ((int)threadIdx.x < FSUMS_PQ) dev_store(fp + threadIdx.x, s_frm[threadIdx.x]);
            }
          }
          // the frame reads every part's sites: publish this part's own
          if (meas && !own_frame)
            part_publish<ST>(my, (U*)a.fspin + ((size_t)par * a.nslots + w.slot) * NC * a.fspin_n, n, a.fspin_n,
                             NC, s_colptr, G.ncol, GD, pi.part, pi.np, lt, TPR);
        }
      }
      if (meas && l_grp && !frames_after) {
        take_frame<NC>(G, B, a, s_st, s_frm, lane, fcur, rowg_l, nrow, frame_mprev3<NC>(a, fcur, rowg_l, nrow),
                       fflags);
        last_row = rowg_l;
      }
      // the swap words of the first ladder are drawn while the barrier
      // completes (they do not depend on the energies)
      float pt_u0 = 0.0f;
      if (a.has_split || pt_now) {
        const uint32_t epoch = (uint32_t)(a.bar_base + nbar + 1);
        grid_arrive(a.bar, epoch);
        if (pt_now && nlad > 0) {
          const int64_t swap = a.philox_base + k;
          const int i = (int)threadIdx.x - pt0;
          const int t = (int)(swap & 1) + 2 * i;
          if (i >= 0 && t + 1 < s_lad[2]) {
            const u32x4 x = philox4x32_10(u32x4{(uint32_t)t, (uint32_t)swap, (uint32_t)((uint64_t)swap >> 32), TAG_PT},
                                          B.seed_lo, B.seed_hi ^ (0x80000000u | (uint32_t)(s_lad[0] & 0x7FFFFFFF)));
            pt_u0 = u24(x.x);
          }
        }
        if (!grid_wait(a.bar, epoch, B.err, s_flag)) return;
        ++nbar;
      }
      // M' of the frame's row, read first so that it overlaps the import
      double mp = 0.0;
      if (meas && l_grp && frames_after) mp = frame_mprev3<NC>(a, fcur, rowg_l, nrow);
      if (pi.np > 1) {
        // part 0 reads part p's sums in lane p, overlapping the import
        const bool w0 = pi.part == 0 && threadIdx.x < 64;
        double pe = 0.0, pm[3] = {0.0, 0.0, 0.0};
        if (w0 && lane >= 1 && lane < pi.np) {
          pe = dev_load(es + lane);
          for (int q = 0; q < 3; ++q) pm[q] = dev_load(ms + lane * 3 + q);
        }
        if (pi.np == 2) {
          // the other part's list directly (measured faster than the flat
          // round-robin on the XY scan's two-part replicas)
          for (int p = 0; p < pi.np; ++p) {
            if (p == pi.part) continue;
            const int32_t* hq = pi.hx + (p - pi.part) * DS;
            const int32_t* ex = a.halo + hq[2 * G.ncol];
            const int nex = hq[2 * G.ncol + 1];
            const U* xi = xb + (((size_t)par * a.nslots + w.slot) * MAX_PARTS + p) * (size_t)a.max_export * NC;
            for (int i = lt; i < nex; i += TPR) {
              const U* src = xi + (size_t)i * NC;
              const int site = ex[i];
              constexpr int W = NC * (int)sizeof(U);
              if constexpr (W == 8 || W == 4) {
                typedef typename std::conditional<W == 8, uint64_t, uint32_t>::type P;
                const P w = dev_load((const P*)src);
#pragma unroll
                for (int q = 0; q < NC; ++q) lraw[(size_t)q * n + site] = (U)(w >> (q * 8 * sizeof(U)));
              } else {
#pragma unroll
                for (int q = 0; q < NC; ++q) lraw[(size_t)q * n + site] = dev_load(src + q);
              }
            }
          }
        } else {
          import_parts<NC>(xb + ((size_t)par * a.nslots + w.slot) * MAX_PARTS * (size_t)a.max_export * NC,
                           a.max_export, a.halo, pi.hx, DS, G.ncol, pi.part, pi.np, lraw, n, lt, TPR);
        }
        if (w0) {
          double E = st[MCG_S_E] + v[0], dm[3] = {v[1], v[2], v[3]};
          for (int p = 1; p < pi.np; ++p) {
            E += lane_f64(pe, p);
            for (int q = 0; q < 3; ++q) dm[q] += lane_f64(pm[q], p);
          }
          if (lane == 0) {
            st[MCG_S_E] = E;
            for (int q = 0; q < 3; ++q) st[MCG_S_M + q] += dm[q];
            if (therm_end) st[MCG_S_EREF] = st[MCG_S_E];
          }
        }
        // the imported sites are read by other threads: by the frame, and by
        // the next sweep's first colour when no frame or tempering pass
        // (which end with a workgroup barrier) comes between
        __syncthreads();
      }
      if constexpr (FULL) {
        if (meas && pi.np > 1) {
          // this part's share of the spin-derived sums: from its own copy, or
          // over every part's published sites; part 0 gathers the others'
          if (own_frame) {
            // the other parts' sums, stored before the barrier, added in
            // part order
            if (pi.part == 0 && (int)threadIdx.x < FSUMS_PQ) {
              const double* fp = a.fpart + ((size_t)par * a.nslots + w.slot) * MAX_PARTS * FSUMS + threadIdx.x;
              double y[MAX_PARTS];
#pragma unroll
              for (int p = 1; p < MAX_PARTS; ++p) y[p] = p < pi.np ? dev_load(fp + (size_t)p * FSUMS) : 0.0;
              double x = s_frm[threadIdx.x];
#pragma unroll
              for (int p = 1; p < MAX_PARTS; ++p)
                if (p < pi.np) x += y[p];
              s_frm[threadIdx.x] = x;
            }
            __syncthreads();
          } else {
            const DevSpins<ST> dsp{(const U*)a.fspin + ((size_t)par * a.nslots + w.slot) * NC * a.fspin_n,
                                   a.fspin_n};
            frame_sums<NC, DevSpins<ST>, true>(dsp, C, B, fcur, key0, key1, hg, pi.part, pi.np);
            if (!gather_frame_sums(a, B, s_frm, w.slot, pi.part, pi.np, fcur, s_flag)) return;
          }
        }
      }
      // the frame of every replica (first wave, lane g), while the other
      // waves run the tempering pass
      if (meas && l_grp && frames_after) {
        take_frame<NC>(G, B, a, s_st, s_frm, lane, fcur, rowg_l, nrow, mp, fflags);
        last_row = rowg_l;
      }
      if (pt_now) {
        // the swaps of the ladders this workgroup's replicas belong to (the
        // other ladders' rows are never read here)
        const int64_t swap = a.philox_base + k;
        const int pp = (int)(swap & 1);
        for (int li = 0; li < nlad; ++li) {
        const int L = s_lad[3 * li] & 0x7FFFFFFF;
        const int r0 = s_lad[3 * li + 1];
        const int npairs = max(0, (s_lad[3 * li + 2] - pp) / 2);
        for (int i = (int)threadIdx.x - pt0; i < npairs; i += (int)blockDim.x - pt0) {
          if (i < 0) break;
          const int t = pp + 2 * i;
          const int ra = s_who[r0 + t], rb = s_who[r0 + t + 1];
          const double* ea = a.eslot + ((size_t)par * nrepb + ra) * MAX_PARTS;
          const double* eb = a.eslot + ((size_t)par * nrepb + rb) * MAX_PARTS;
          // the parts' slots at once (a replica's energy is the sum of its
          // parts' slots, in part order)
          const int npa = s_np[ra], npb = s_np[rb];
          double xa[MAX_PARTS], xb[MAX_PARTS];
#pragma unroll
          for (int q = 0; q < MAX_PARTS; ++q) {
            xa[q] = q < npa ? dev_load(ea + q) : 0.0;
            xb[q] = q < npb ? dev_load(eb + q) : 0.0;
          }
          double Ea = xa[0], Eb = xb[0];
#pragma unroll
          for (int q = 1; q < MAX_PARTS; ++q) {
            Ea += xa[q];
            Eb += xb[q];
          }
          const double ba = s_beta[r0 + t], bb = s_beta[r0 + t + 1];
          const double d = (ba - bb) * (Ea - Eb);
          bool acc = d >= 0.0;
          if (!acc) {
            float u = pt_u0;  // drawn during the barrier for the first ladder
            if (li > 0 || i >= (int)blockDim.x - pt0) {
              const u32x4 x =
                  philox4x32_10(u32x4{(uint32_t)t, (uint32_t)swap, (uint32_t)((uint64_t)swap >> 32), TAG_PT},
                                B.seed_lo, B.seed_hi ^ (0x80000000u | (uint32_t)L));
              u = u24(x.x);
            }
            acc = u < exp_neg((float)d);
          }
          if (acc) {
            s_who[r0 + t] = rb;
            s_who[r0 + t + 1] = ra;
            s_rung[ra] = t + 1;
            s_rung[rb] = t;
            s_ptacc[r0 + t] += 1;
          }
        }
        }
        __syncthreads();
      }
      if (meas && nfr > 0) {
        __syncthreads();
        if (g < nfr) flush_frame<NC>(C, B, rowg, fflags);
      }
    }
    __syncthreads();
    if (last_row >= 0) {
      const int64_t fl = (a.k_begin + a.nsweeps - a.nthermal) / a.spf - 1 + a.frame0;  // last frame
      B.acc[(size_t)last_row * MCG_NACC + MCG_A_MPREV] = dev_load(a.mprev3 + (size_t)(fl % 3) * nrow + last_row);
    }
    if (pi.np == 1)
      stage<ST, false>(my, gsp, NC * n, lt, TPR);
    else
      part_copy<ST, true>(my, gsp, n, NC, s_colptr, G.ncol, GD, pi.part, pi.np, lt, TPR);
    if (pi.part == 0 && lt < MCG_NSTATE) B.state[(size_t)rep * MCG_NSTATE + lt] = st[lt];
    // each ladder's state goes back from its owner (flag bit 31 of the
    // workgroup's ladder entry): the first workgroup holding one of its
    // replicas
    for (int li = 0; li < nlad; ++li) {
      if (!(s_lad[3 * li] & 0x80000000)) continue;
      const int r0 = s_lad[3 * li + 1];
      for (int t = threadIdx.x; t < s_lad[3 * li + 2]; t += blockDim.x) {
        const int r = s_who[r0 + t];
        B.who[r0 + t] = r;
        B.pt_acc[r0 + t] += s_ptacc[r0 + t];
        B.rung[r] = s_rung[r];
      }
    }
    return;
  } else {
  double dE, dM[3];
  if constexpr (MODE == 0 || MODE == 3)
    lds_sweeps<NC, ST, MODE, FULL>(G, B, a, sp, s_coup, s_clsl, s_colptr, s_pid, s_pat, s_red, st, C, g, lt, TPR, row,
                                   beta, h, key0, key1, pi, dE, dM, aniso, pl);
  else if constexpr (MODE == 1)
    lds_sweeps<NC, ST, 1, FULL>(G, B, a, sp, s_coup, s_clsl, s_colptr, s_pid, s_pat, s_red, st, C, g, lt, TPR, row,
                                beta, h, key0, key1, pi, dE, dM, aniso, pl);
  else
    lds_sweeps<NC, ST, 2, FULL>(G, B, a, sp, G.coup, G.clsdat, s_colptr, s_pid, s_pat, s_red, st, C, g, lt, TPR,
                                row, beta, h, key0, key1, pi, dE, dM, aniso, pl);
  __syncthreads();
  if (pi.np == 1) {
    stage<ST, false>(my, gsp, NC * n, lt, TPR);
    if (lt < MCG_NSTATE) B.state[(size_t)rep * MCG_NSTATE + lt] = st[lt];
    return;
  }
  // split replica (one workgroup per part, R = 1): write back own sites,
  // publish the partial sums, count in; the last part to arrive finishes
  // the sweep.  It needs the other parts' sites only when this sweep is
  // measured with observables taken from the spins, or recomputed; then the
  // parts make their sites visible device-wide (one L2 write-back per
  // workgroup) and the last part reads them after an acquire.
  constexpr int GD = SitesPerDraw<NC>::G;
  const int64_t k = a.k_begin;
  const bool meas = k >= a.nthermal && ((k - a.nthermal + 1) % a.spf) == 0;
  const bool therm_end = (k + 1 == a.nthermal);
  const bool rec = recompute_due(B, a.philox_base + k);
  const bool need_all = rec || (FULL && meas && (B.meas & MEAS_SPIN_OBS));
  uint32_t* ctr = a.sync + (size_t)w.slot * SYNC_STRIDE;
  double* xp = a.xpart + (size_t)w.slot * MAX_PARTS * 4;
  part_copy<ST, true>(my, gsp, n, NC, s_colptr, G.ncol, GD, pi.part, pi.np, lt, TPR);
  double v[RK] = {dE, dM[0], dM[1], dM[2], 0.0, 0.0, 0.0, 0.0};
  group_sum<4>(v, s_red, g, TPR);
  if (lt == 0)
    for (int q = 0; q < 4; ++q) dev_store(xp + pi.part * 4 + q, v[q]);
  stores_done();
  __syncthreads();
  int* s_last = (int*)s_red;
  if (lt == 0) {
    if (need_all) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    *s_last = (__hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) % pi.np) ==
              (uint32_t)(pi.np - 1);
  }
  __syncthreads();
  if (!*s_last) return;
  if (need_all) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    for (int p = 0; p < pi.np; ++p)
      if (p != pi.part) part_copy<ST, false>(my, gsp, n, NC, s_colptr, G.ncol, GD, p, pi.np, lt, TPR);
  }
  double tE = 0.0, tM[3] = {0.0, 0.0, 0.0};
  if (lt == 0)
    for (int p = 0; p < pi.np; ++p) {
      tE += p == pi.part ? v[0] : dev_load(xp + p * 4);
      for (int q = 0; q < 3; ++q) tM[q] += p == pi.part ? v[1 + q] : dev_load(xp + p * 4 + 1 + q);
    }
  __syncthreads();
  lds_sweep_tail<NC, ST, FULL>(sp, G, B, C.coup, C.clsdat, s_red, st, C, g, lt, TPR, row, h, k, a, meas, therm_end,
                               rec, tE, tM, key0, key1, 0.0, false, PoolInfo<ST>{});
  if (lt < MCG_NSTATE) B.state[(size_t)rep * MCG_NSTATE + lt] = st[lt];
  }
}

// ---------------------------------------------------------------------------
// k_pass: one colour class of every listed replica, spins in HBM / L2.
// grid = (chunks, nrep) of 256 threads; thread t of chunk x takes the Philox
// groups (x * PASS_GPT + u) * 256 + t of the colour, u < PASS_GPT, so that
// a wave touches consecutive positions and each thread's reduction is shared
// by PASS_GPT groups.  Graphs with link patterns read their neighbours as
// position deltas from pattern rows staged in LDS (no link-table traffic and
// no load that waits on another); the others read the global link table.
// Changes of E and M: FP32 per thread, FP64 DPP wave sums, one device
// atomic per value per block.
// ---------------------------------------------------------------------------
// Philox groups per thread and the unrolling of that loop (measured on the
// cubic 64^3 and triangular 128^2 streaming sets: 8 groups, not unrolled,
// is 7 % faster than 4 unrolled and 22 % faster than 2)
#ifndef MCG_PASS_GPT
#define MCG_PASS_GPT 8
#endif
#ifndef MCG_PASS_UNROLL
#define MCG_PASS_UNROLL 1
#endif
static constexpr int PASS_GPT = MCG_PASS_GPT;
static constexpr int PASS_PAT_INTS = 4096;  // LDS for pattern rows (16 KB)
// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b % 8),
// and each XCD has its own L2.  xcd_item maps workgroup b to a work item so
// that every XCD takes one contiguous run of the `total` items: the items of
// a run are neighbouring chunks of a replica, whose neighbour sites then
// come from that XCD's L2 instead of being fetched again by another XCD.
// The grid has 8 * ceil(total / 8) workgroups; -1: no item.
static constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_item(int b, int total) {
  const int per = (total + NXCD - 1) / NXCD;
  const int w = (b % NXCD) * per + b / NXCD;
  return w < total ? w : -1;
}
__host__ __device__ inline int xcd_grid(int total) { return (total + NXCD - 1) / NXCD * NXCD; }

// at least 8 waves per SIMD: the pass is latency-bound, and a 64-VGPR cap
// hides more of it (measured per sweep: cubic 64^3 0.702 -> 0.648 ms,
// triangular 128^2 0.470 -> 0.436 ms; 6 waves: 0.628 and 0.497 ms)
#ifndef MCG_PASS_MIN_WAVES
#define MCG_PASS_MIN_WAVES 8
#endif
template <int NC, typename ST>
__global__ __launch_bounds__(256, MCG_PASS_MIN_WAVES) void k_pass(mcg_batch B, const int32_t* __restrict__ rep_list, int colour,
                                              int64_t sweep, int nchunk, int nitems) {
  __shared__ int32_t s_pat[PASS_PAT_INTS];
  __shared__ double s_red[4 * 4];
  const int item = xcd_item(blockIdx.x, nitems);
  if (item < 0) return;  // the whole workgroup
  const int chunk = item % nchunk;
  const int rep = rep_list[item / nchunk];
  const mcg_replica rp = B.reps[rep];
  const mcg_graph G = B.graphs[rp.graph];
  const int zmax = G.zmax;
  const bool pats = G.npat > 0 && G.lcid == nullptr && G.npat * pat_stride(zmax) <= PASS_PAT_INTS;
  int anz = 0;
  if (pats) {
    for (int t = threadIdx.x; t < G.npat * (zmax + 1); t += blockDim.x)
      s_pat[(t / (zmax + 1)) * pat_stride(zmax) + t % (zmax + 1)] = G.pattab[t];
    // does any site class carry an on-site term? (metro_core skips them if not)
    for (int t = threadIdx.x; t < G.ncls; t += blockDim.x)
      anz |= (G.clsdat[t * 8 + 1] != 0.0f) | (G.clsdat[t * 8 + 2] != 0.0f) | (G.clsdat[t * 8 + 3] != 0.0f);
  }
  const bool aniso = __syncthreads_or(anz) != 0 || !pats;
  float dE = 0.0f, dM[3] = {0.0f, 0.0f, 0.0f};
  if (colour < G.ncol) {
    const int b = G.col_ptr[colour], cnt = G.col_ptr[colour + 1] - b;
    const int ngrp = (cnt + SitesPerDraw<NC>::G - 1) / SitesPerDraw<NC>::G;
    const int row = rp.row0 + B.rung[rep];
    const float h = B.row_h[row];
    const Acc beta = acc_of(B, B.row_beta[row], row);
    const uint32_t key1 = B.seed_hi ^ rp.serial;
    const GSpins<ST> sp{gas((ST*)B.spins + rp.spin_off), G.n};
    const GAS int32_t* lcid = G.lcid ? gas(G.lcid) : nullptr;
#pragma unroll MCG_PASS_UNROLL
    for (int u = 0; u < PASS_GPT; ++u) {
      const int q = (chunk * PASS_GPT + u) * (int)blockDim.x + (int)threadIdx.x;
      if (q >= ngrp) break;
      if (pats) {
        if (G.ucid >= 0)
          metro_group_gpat<NC, ST, true>(sp, b, cnt, q, colour, gas(G.pat), s_pat, zmax, gas(G.coup) + G.ucid * 9,
                                         gas(G.clsdat), beta, h, (uint64_t)sweep, B.seed_lo, key1, dE, dM, aniso);
        else
          metro_group_gpat<NC, ST, false>(sp, b, cnt, q, colour, gas(G.pat), s_pat, zmax, gas(G.coup),
                                          gas(G.clsdat), beta, h, (uint64_t)sweep, B.seed_lo, key1, dE, dM, aniso);
      } else if (G.ucid >= 0) {
        metro_group<NC, true>(sp, b, cnt, q, colour, gas(G.link), lcid, zmax, gas(G.coup) + G.ucid * 9, gas(G.cls),
                              gas(G.clsdat), beta, h, (uint64_t)sweep, B.seed_lo, key1, dE, dM);
      } else {
        metro_group<NC>(sp, b, cnt, q, colour, gas(G.link), lcid, zmax, gas(G.coup), gas(G.cls), gas(G.clsdat),
                        beta, h, (uint64_t)sweep, B.seed_lo, key1, dE, dM);
      }
    }
  }
  // every lane reaches here: DPP sums per wave, the waves' totals through LDS
  double v[4] = {(double)dE, (double)dM[0], (double)dM[1], (double)dM[2]};
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = wave_sum_f64(v[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 4; ++k) s_red[wave * 4 + k] = v[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_red[w * 4 + threadIdx.x];
    if (t != 0.0) atomicAdd(B.state + (size_t)rep * MCG_NSTATE + MCG_S_E + threadIdx.x, t);
  }
}

// ---------------------------------------------------------------------------
// k_measure: one block per replica.  mode bits: 1 recompute, 2 set E_ref,
// 4 measure frame, 8 reset drift
// ---------------------------------------------------------------------------
template <int NC, typename ST>
__global__ __launch_bounds__(1024) void k_measure(mcg_batch B, const int32_t* __restrict__ rep_list, int mode,
                                                  int64_t frame) {
  const int rep = rep_list[blockIdx.x];
  const mcg_replica rp = B.reps[rep];
  const mcg_graph G = B.graphs[rp.graph];
  __shared__ double s_red[16 * RK];
  __shared__ double s_frm[FRM];
  __shared__ double st[MCG_NSTATE];
  const Spins<ST> sp{(ST*)B.spins + rp.spin_off, G.n};
  const int row = rp.row0 + B.rung[rep];
  const float h = B.row_h[row];
  if (threadIdx.x < MCG_NSTATE) st[threadIdx.x] = B.state[(size_t)rep * MCG_NSTATE + threadIdx.x];
  __syncthreads();
  if (mode & 1) {
    double r4[NREC];
    recompute<NC>(sp, G, G.coup, G.clsdat, h, s_red, 0, threadIdx.x, blockDim.x, r4);
    if (threadIdx.x == 0) {
      double d = 0.0;
      if (!(mode & 8)) {
        d = fabs(st[MCG_S_E] - r4[0]);
        if (d > st[MCG_S_DRIFT]) st[MCG_S_DRIFT] = d;
      }
      recompute_checks<ST>(B, G, r4, d);
      st[MCG_S_E] = r4[0];
      st[MCG_S_M] = r4[1];
      st[MCG_S_M + 1] = r4[2];
      st[MCG_S_M + 2] = r4[3];
    }
  }
  if ((mode & 2) && threadIdx.x == 0) st[MCG_S_EREF] = st[MCG_S_E];
  __syncthreads();
  if (mode & 4) {
    FrameCtx C{&G, G.coup, G.clsdat, s_red, s_frm, 0, (int)threadIdx.x, (int)blockDim.x};
    measure_frame<NC>(sp, C, B, st, row, frame, B.seed_lo, B.seed_hi ^ rp.serial, h);
  }
  __syncthreads();
  if (threadIdx.x < MCG_NSTATE) B.state[(size_t)rep * MCG_NSTATE + threadIdx.x] = st[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Split measurement (streaming and Wolff paths): k_frame_part gives np
// workgroups to each replica, each reducing its share of the frame sums and
// of the full recompute and adding them into the replica's row of fsum;
// k_frame_end (one wave per replica) turns a row into the recompute
// bookkeeping and the frame's accumulator terms, and clears it.  mode bits
// as k_measure: 1 recompute, 2 set E_ref, 4 measure frame, 8 reset drift.
// ---------------------------------------------------------------------------
static constexpr int FSLOTS = FSUMS + 8;  // per replica: frame sums, then NREC recompute sums
template <int NC, typename ST>
__global__ __launch_bounds__(256) void k_frame_part(mcg_batch B, const int32_t* __restrict__ rep_list, int mode,
                                                    int64_t frame, double* __restrict__ fsum, int np, int nitems) {
  const int item = xcd_item(blockIdx.x, nitems);
  if (item < 0) return;  // the whole workgroup
  const int rep = rep_list[item / np];
  const mcg_replica rp = B.reps[rep];
  const mcg_graph G = B.graphs[rp.graph];
  __shared__ double s_red[4 * RK];
  __shared__ double s_frm[FSUMS];
  const Spins<ST> sp{(ST*)B.spins + rp.spin_off, G.n};
  const int row = rp.row0 + B.rung[rep];
  const float h = B.row_h[row];
  const int part = item % np;
  double* out = fsum + (size_t)rep * FSLOTS;
  if (mode & 1) {
    double r[NREC];
    recompute<NC>(sp, G, G.coup, G.clsdat, h, s_red, 0, threadIdx.x, blockDim.x, r, part, np);
    if (threadIdx.x == 0)
      for (int k = 0; k < NREC; ++k)
        if (r[k] != 0.0) atomicAdd(out + FSUMS + k, r[k]);
  }
  if (mode & 4) {
    for (int t = threadIdx.x; t < FSUMS; t += blockDim.x) s_frm[t] = 0.0;
    __syncthreads();
    FrameCtx C{&G, G.coup, G.clsdat, s_red, s_frm, 0, (int)threadIdx.x, (int)blockDim.x};
    frame_sums<NC>(sp, C, B, frame, B.seed_lo, B.seed_hi ^ rp.serial, h, part, np);
    if (threadIdx.x == 0)
      for (int k = 0; k < FSUMS; ++k)
        if (s_frm[k] != 0.0) atomicAdd(out + k, s_frm[k]);
  }
}

template <int NC, typename ST>
__global__ __launch_bounds__(64) void k_frame_end(mcg_batch B, const int32_t* __restrict__ rep_list, int mode,
                                                  int64_t frame, double* __restrict__ fsum) {
  const int rep = rep_list[blockIdx.x];
  const mcg_replica rp = B.reps[rep];
  const mcg_graph G = B.graphs[rp.graph];
  __shared__ double s_red[RK];
  __shared__ double s_frm[FRM];
  __shared__ double st[MCG_NSTATE];
  const int lt = threadIdx.x;
  const int row = rp.row0 + B.rung[rep];
  double* in = fsum + (size_t)rep * FSLOTS;
  if (lt < MCG_NSTATE) st[lt] = B.state[(size_t)rep * MCG_NSTATE + lt];
  for (int t = lt; t < FSUMS; t += blockDim.x) s_frm[t] = in[t];
  __syncthreads();
  if ((mode & 1) && lt == 0) {
    double r[NREC];
    for (int k = 0; k < NREC; ++k) r[k] = in[FSUMS + k];
    double d = 0.0;
    if (!(mode & 8)) {
      d = fabs(st[MCG_S_E] - r[0]);
      if (d > st[MCG_S_DRIFT]) st[MCG_S_DRIFT] = d;
    }
    recompute_checks<ST>(B, G, r, d);
    for (int k = 0; k < 4; ++k) st[MCG_S_E + k] = r[k];
  }
  if ((mode & 2) && lt == 0) st[MCG_S_EREF] = st[MCG_S_E];
  __syncthreads();
  if (mode & 4) {
    FrameCtx C{&G, G.coup, G.clsdat, s_red, s_frm, 0, lt, (int)blockDim.x};
    frame_apply<NC>(C, B, st, row, frame, B.meas, 0.0, false, false);
  }
  __syncthreads();
  for (int t = lt; t < FSLOTS; t += blockDim.x) in[t] = 0.0;
  if (lt < MCG_NSTATE) B.state[(size_t)rep * MCG_NSTATE + lt] = st[lt];
}

// ---------------------------------------------------------------------------
// k_pt: tempering swaps, one thread per ladder.  Pairs (t, t+1) with
// t = parity, parity+2, ...; parity alternates with the swap count.
// ---------------------------------------------------------------------------
// Replica-exchange swaps of one ladder (rows r0 .. r0+nr-1): pairs of one
// parity, alternating with the swap index; threads tid, tid+nthreads, ...
// take disjoint pairs.  Accept with min(1, exp((b_a - b_b)(E_a - E_b))).
__device__ void pt_ladder(const mcg_batch& B, int L, int r0, int nr, int64_t swap_index, int tid, int nthreads) {
  const int parity = (int)(swap_index & 1);
  for (int t = parity + 2 * tid; t + 1 < nr; t += 2 * nthreads) {
    const int ra = B.who[r0 + t], rb = B.who[r0 + t + 1];
    const double Ea = B.state[(size_t)ra * MCG_NSTATE + MCG_S_E], Eb = B.state[(size_t)rb * MCG_NSTATE + MCG_S_E];
    const double ba = B.row_beta[r0 + t], bb = B.row_beta[r0 + t + 1];
    const double d = (ba - bb) * (Ea - Eb);
    bool acc = d >= 0.0;
    if (!acc) {
      const u32x4 x = philox4x32_10(
          u32x4{(uint32_t)t, (uint32_t)swap_index, (uint32_t)((uint64_t)swap_index >> 32), TAG_PT}, B.seed_lo,
          B.seed_hi ^ (0x80000000u | (uint32_t)L));
      acc = u24(x.x) < exp_neg((float)d);
    }
    if (acc) {
      B.who[r0 + t] = rb;
      B.who[r0 + t + 1] = ra;
      B.rung[ra] = t + 1;
      B.rung[rb] = t;
      B.pt_acc[r0 + t] += 1;
    }
  }
}

__global__ void k_pt(mcg_batch B, const int32_t* __restrict__ ladder_row0, const int32_t* __restrict__ ladder_n,
                     int nladder, int64_t swap_index) {
  const int L = blockIdx.x;
  if (L >= nladder) return;
  pt_ladder(B, L, ladder_row0[L], ladder_n[L], swap_index, (int)threadIdx.x, (int)blockDim.x);
}

// ---------------------------------------------------------------------------
// k_wolff: one 1024-thread block per replica, ncl clusters, frontier BFS in
// global memory with one work item per (frontier site, link)
// (heisenbergLib.c:310-439, xyLib.c:256-380, isingLib.c:165-236)
// ---------------------------------------------------------------------------
// accesses to a workgroup's private scratch in global memory (Wolff stamps
// and cluster lists): ordered among the workgroup's threads only
template <typename T>
__device__ __forceinline__ T wg_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void wg_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <int NC, typename ST>
__global__ __launch_bounds__(1024) void k_wolff(mcg_batch B, const int32_t* __restrict__ rep_list, int64_t cl0,
                                                int ncl) {
  const int rep = rep_list[blockIdx.x];
  const mcg_replica rp = B.reps[rep];
  const mcg_graph G = B.graphs[rp.graph];
  const Spins<ST> sp{(ST*)B.spins + rp.spin_off, G.n};
  uint32_t* stamp = B.wstamp + rp.site_off;
  int32_t* mem = B.wlist + rp.site_off;
  const int row = rp.row0 + B.rung[rep];
  const float beta = B.row_beta[row], h = B.row_h[row];
  const uint32_t key0 = B.seed_lo, key1 = B.seed_hi ^ rp.serial;
  __shared__ int s_nc[3], s_seed;
  __shared__ float s_r[3];
  __shared__ double s_red[16];
  __shared__ int s_acc;
  const int tid = threadIdx.x;
  // isotropic exchange (every tensor c*I on the spin components) and no
  // single-ion term: with h = 0 the cluster's acceptance energy
  // (heisenbergLib.c:403-423) is zero in exact arithmetic, so the cluster is
  // accepted without evaluating it (Ising: zero whenever h = 0)
  bool iso_ok = true;
  if constexpr (NC > 1) {
    for (int c = tid; c < G.ncoup; c += blockDim.x) {
      const float* J = G.coup + c * 9;
      iso_ok = iso_ok && J[1] == 0.0f && J[3] == 0.0f && J[0] == J[4];
      if (NC == 3) iso_ok = iso_ok && J[2] == 0.0f && J[5] == 0.0f && J[6] == 0.0f && J[7] == 0.0f && J[8] == J[0];
    }
    for (int c = tid; c < G.ncls; c += blockDim.x)
      for (int k = 0; k < NC; ++k) iso_ok = iso_ok && G.clsdat[c * 8 + 1 + k] == 0.0f;
  }
  const bool skip_acc = __syncthreads_and(iso_ok ? 1 : 0) != 0 && h == 0.0f;

  for (int c = 0; c < ncl; ++c) {
    const uint64_t cl = (uint64_t)(cl0 + c);
    const uint32_t cur = (uint32_t)(cl & 0x7FFFFFFFu) * 2u + 1u;  // odd: never 0
    if (tid == 0) {
      const u32x4 x = philox4x32_10(u32x4{0u, (uint32_t)cl, (uint32_t)(cl >> 32), TAG_WOLFF}, key0, key1);
      const int seed = (int)(((uint64_t)x.x * (uint64_t)G.n) >> 32);
      float rx = 0, ry = 0, rz = 0;
      if constexpr (NC == 3) {
        rz = fmaf(u24(x.y), 2.0f, -1.0f);
        const float rho = sqrtf(fmaxf(fmaf(-rz, rz, 1.0f), 0.0f));
        float cc, sn;
        sincos_turn(x.z, &cc, &sn);
        rx = rho * cc;
        ry = rho * sn;
      } else if constexpr (NC == 2) {
        sincos_turn(x.y, &rx, &ry);
      }
      s_r[0] = rx; s_r[1] = ry; s_r[2] = rz;
      s_seed = seed;
      wg_store(stamp + seed, cur);
      wg_store(mem, seed);
      s_nc[0] = s_nc[1] = s_nc[2] = 0;
      // cluster accept draw: x.w
      s_acc = (int)(x.w >> 8);
    }
    __syncthreads();
    const float r0 = s_r[0], r1 = s_r[1], r2 = s_r[2];
    // Level-synchronous growth, one workgroup barrier per level: the sites
    // level k adds go to mem[hi + s_nc[k % 3]++], so the frontier of level
    // k+1 is [hi, hi + s_nc[k % 3]).  During level k the first thread clears
    // the counter of level k+1, last read (level k-2's) before the barrier
    // that ended level k-1.  One workgroup owns the cluster, so the stamps
    // and the list need workgroup scope only: their accesses stay in the
    // CU's caches.
    int lo = 0, hi = 1, cnt = 1;
    for (int k = 0;; ++k) {
      // one work item per (frontier site, link): every item is one short
      // dependent chain (stamp test, bond draw, stamp exchange)
      const uint32_t zl = (uint32_t)G.zmax;
      const uint32_t nitem = (uint32_t)(hi - lo) * zl;
      for (uint32_t it = tid; it < nitem; it += blockDim.x) {
        const uint32_t fq = it / zl;
        const int l = (int)(it - fq * zl);
        const int o = wg_load(mem + lo + (int)fq);
        float so[3];
        spin_vec<NC>(sp, o, G.cls, G.clsdat, so);
        const float sdo = -(so[0] * r0 + so[1] * r1 + so[2] * r2);
        {
          const uint32_t wd = G.link[(size_t)o * G.zmax + l];
          const int t = (int)(wd & 0xFFFFFFu);
          if (t == o) continue;
          if (wg_load(stamp + t) == cur) continue;
          const uint32_t cw = G.lcid ? (uint32_t)G.lcid[(size_t)o * G.zmax + l] : (wd >> 24);
          const float* J = G.coup + cw * 9;
          float corr;
          if constexpr (NC == 1) {
            corr = (J[0] * sp.get(t, 0)) * so[0] * beta;  // J*S_t*sigma_t * s_o
            corr = 2.0f * corr;                            // add if J s s < 0, p = 1-exp(2 corr)
          } else {
            float st3[3];
            spin_vec<NC>(sp, t, G.cls, G.clsdat, st3);
            const float sdt = -(st3[0] * r0 + st3[1] * r1 + st3[2] * r2);
            const float rJr = r0 * (J[0] * r0 + J[1] * r1 + J[2] * r2) + r1 * (J[3] * r0 + J[4] * r1 + J[5] * r2) +
                              r2 * (J[6] * r0 + J[7] * r1 + J[8] * r2);
            corr = 2.0f * sdo * sdt * rJr * beta;
          }
          if (corr < 0.0f) {
            const u32x4 x = philox4x32_10(
                u32x4{(uint32_t)(o * G.zmax + l), (uint32_t)cl, (uint32_t)(cl >> 32), TAG_BOND}, key0, key1);
            if (1.0f - exp_neg(corr) > u24(x.x)) {
              const uint32_t old = __hip_atomic_exchange(stamp + t, cur, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_WORKGROUP);
              if (old != cur) {
                const int pos = hi + atomicAdd(&s_nc[k % 3], 1);
                wg_store(mem + pos, t);
              }
            }
          }
        }
      }
      if (tid == 0) s_nc[(k + 1) % 3] = 0;
      __syncthreads();
      const int add = s_nc[k % 3];
      if (add == 0) {
        cnt = hi;
        break;
      }
      lo = hi;
      hi += add;
    }
    // cluster acceptance energy (field / anisotropy correction)
    // O(n): one work item per (cluster site, link) plus one per site for the
    // single-ion and field terms
    double d = 0.0;
    const uint32_t zi = NC == 1 ? 1u : (uint32_t)G.zmax + 1u;
    const uint32_t nacc = skip_acc ? 0u : (uint32_t)cnt * zi;
    for (uint32_t it = tid; it < nacc; it += blockDim.x) {
      const uint32_t fq = it / zi;
      const int l = (int)(it - fq * zi);
      const int b = wg_load(mem + (int)fq);
      float sb[3];
      spin_vec<NC>(sp, b, G.cls, G.clsdat, sb);
      if constexpr (NC == 1) {
        d += (double)(2.0f * h * sb[0]);
      } else {
        const float sdb = -(sb[0] * r0 + sb[1] * r1 + sb[2] * r2);
        const float pb[3] = {sb[0] + sdb * r0, sb[1] + sdb * r1, sb[2] + sdb * r2};
        if (l < G.zmax) {
          const uint32_t wd = G.link[(size_t)b * G.zmax + l];
          const int t = (int)(wd & 0xFFFFFFu);
          const uint32_t cw = G.lcid ? (uint32_t)G.lcid[(size_t)b * G.zmax + l] : (wd >> 24);
          const float* J = G.coup + cw * 9;
          if (cw == 0u) continue;
          float stt[3];
          spin_vec<NC>(sp, t, G.cls, G.clsdat, stt);
          const float sdt = -(stt[0] * r0 + stt[1] * r1 + stt[2] * r2);
          const float pt[3] = {stt[0] + sdt * r0, stt[1] + sdt * r1, stt[2] + sdt * r2};
          // r^T J p_t
          const float src = sdb * (r0 * (J[0] * pt[0] + J[1] * pt[1] + J[2] * pt[2]) +
                                   r1 * (J[3] * pt[0] + J[4] * pt[1] + J[5] * pt[2]) +
                                   r2 * (J[6] * pt[0] + J[7] * pt[1] + J[8] * pt[2]));
          d += (double)src;
          if (wg_load(stamp + t) == cur)
            d += (double)(sdt * (pb[0] * (J[0] * r0 + J[1] * r1 + J[2] * r2) +
                                 pb[1] * (J[3] * r0 + J[4] * r1 + J[5] * r2) +
                                 pb[2] * (J[6] * r0 + J[7] * r1 + J[8] * r2)));
          else
            d += (double)src;
        } else {
          // single-ion term with the half step s + sdn*r (heisenbergLib.c:418)
          const int ci = G.cls[b];
          const float* Dd = G.clsdat + ci * 8 + 1;
          const float tr[3] = {sdb * r0, sdb * r1, sdb * r2};
          float o = 0.0f;
          for (int k = 0; k < NC; ++k) {
            const float q = sb[k] + tr[k];
            o += Dd[k] * (q * q - sb[k] * sb[k]);
          }
          o -= h * tr[NC == 3 ? 2 : 0];
          d += (double)o;
        }
      }
    }
    for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = d;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += s_red[q];
      const float db = (float)(t * (double)beta);
      const float u = (float)s_acc * 0x1p-24f;
      s_acc = (db <= 0.0f || u < exp_neg(-db)) ? 1 : 0;
      // cluster statistics: sites flipped and sites grown
      double* st = B.state + (size_t)rep * MCG_NSTATE;
      if (s_acc) st[MCG_S_ACC] += (double)cnt;
      st[MCG_S_TRY] += (double)cnt;
    }
    __syncthreads();
    if (s_acc) {
      for (int f = tid; f < cnt; f += blockDim.x) {
        const int b = wg_load(mem + f);
        if constexpr (NC == 1) {
          sp.set(b, 0, -sp.get(b, 0));
        } else {
          float sb[3];
          spin_vec<NC>(sp, b, G.cls, G.clsdat, sb);
          const float a2 = -2.0f * (sb[0] * r0 + sb[1] * r1 + sb[2] * r2);
          float n0 = sb[0] + a2 * r0, n1 = sb[1] + a2 * r1, n2 = sb[2] + a2 * r2;
          float nn = n0 * n0 + n1 * n1 + n2 * n2;
          const float fct = fabsf(G.clsdat[G.cls[b] * 8]) / sqrtf(nn);
          sp.set(b, 0, n0 * fct);
          sp.set(b, 1, n1 * fct);
          if constexpr (NC == 3) sp.set(b, 2, n2 * fct);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// small utility kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_f64(double* p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, size_t n, int32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int upload(const void* src, size_t n, hipStream_t s) {
    if (n == 0) return 0;
    if (!p || bytes < n) {
      if (p) (void)hipFree(p);
      p = nullptr;
      if (hipMalloc(&p, n) != hipSuccess) return fail("hipMalloc(%zu) failed", n);
      bytes = n;
    }
    if (hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s) != hipSuccess) return fail("hipMemcpyAsync failed");
    return 0;
  }
};

int store_size(int store) { return store == MCG_STORE_F32 ? 4 : store == MCG_STORE_F16 ? 2 : 1; }

int check_batch(const mcg_batch* B) {
  if (!B) return fail("null batch");
  if (B->nc < 1 || B->nc > 3) return fail("nc must be 1, 2 or 3 (got %d)", B->nc);
  if (B->nc == 1 && B->store != MCG_STORE_I8) return fail("Ising batches use int8 storage");
  if (B->nc > 1 && B->store == MCG_STORE_I8) return fail("int8 storage is for Ising only");
  if (!B->graphs_host || !B->reps_host) return fail("batch lacks host mirrors of graphs/reps");
  if (B->nbins < 1) return fail("nbins must be >= 1");
  for (int gi = 0; gi < B->ngraph; ++gi) {
    const mcg_graph& G = B->graphs_host[gi];
    if (G.n <= 0 || G.n > (1 << 24)) return fail("graph %d: n=%d outside 1..2^24", gi, G.n);
    if (G.ncoup < 1 || (G.ncoup > 256 && !G.lcid))
      return fail("graph %d: ncoup=%d needs per-link coupling indices (lcid) above 256", gi, G.ncoup);
    if (G.ncls < 1 || G.ncls > G.n) return fail("graph %d: ncls=%d outside 1..n", gi, G.ncls);
    if (G.ncol < 1) return fail("graph %d: no colours", gi);
    if ((B->meas & MCG_MEAS_GROUPS) && G.ngroup > 16) return fail("graph %d: more than 16 orbital groups", gi);
  }
  for (int r = 0; r < B->nrep; ++r) {
    const mcg_replica& R = B->reps_host[r];
    if (R.graph < 0 || R.graph >= B->ngraph) return fail("replica %d: bad graph index", r);
    if (R.row0 < 0 || R.row0 + R.nrung > B->nrow) return fail("replica %d: rows outside the batch", r);
  }
  return 0;
}

// A prepared run: the launch plan (host), and the device scratch the
// launches use beyond the batch's own arrays, carved from one region in
// 256-byte aligned pieces:
//   tables  work items, their replicas, replica lists, ladders, split and
//           workgroup-ladder tables; written once when the run is prepared
//   state   arrival counters, barrier words, published energies and M',
//           exported sites; every run starts them from zero (one fill kernel)
// Preparing plans and uploads; running only launches kernels, so repeated
// runs of one batch pay no planning, allocation or host copy.
struct RunCtx {
  // plan (host)
  std::vector<WorkItem> hitems;
  std::vector<int32_t> hitem_reps, hstream, hall, hhalo;
  std::vector<int32_t> hl_row0, hl_n;
  int block = 0;
  int ncu = 0;
  int nslots = 0;  // split replicas (arrival counters)
  int max_export = 0;  // longest export list of a part
  bool persist = false;  // persistent launches (barrier + tempering inside)
  int mode_begin[5] = {0, 0, 0, 0, 0};  // work items of link mode m: [mode_begin[m], mode_begin[m+1])
  size_t lds_max = 0;
  double lds_updates_per_sweep = 0;
  int device = 0;
  int stream_ncol = 0, stream_maxcnt = 0;  // colours and largest colour class of the streamed graphs
  int frame_np_stream = 1, frame_np_all = 1;  // workgroups per replica of the split measurement
  bool frame_spins = true;  // a measured frame reads spins (launch_frame)
  // what the plan was made for: a run whose plan or batch differs in any of
  // these is refused
  int32_t k_kernel = 0, k_block = 0, k_rpb = 0, k_parts = 0, k_nopersist = 0, k_pt = 0;
  int32_t k_nc = 0, k_store = 0, k_ngraph = 0, k_nrep = 0, k_nrow = 0, k_meas = 0, k_rec = 0;
  const void *k_graphs = nullptr, *k_reps = nullptr, *k_graphs_host = nullptr, *k_reps_host = nullptr;
  // device region: [0, tab_bytes) tables, [reset_off, reset_off + reset_bytes) state
  unsigned char* dev = nullptr;
  size_t dev_bytes = 0, tab_bytes = 0, reset_off = 0, reset_bytes = 0;
  std::vector<unsigned char> hstage;  // the tables as uploaded
  // offsets of the pieces, then their device addresses
  size_t o_items = 0, o_item_reps = 0, o_stream = 0, o_all = 0, o_lrow0 = 0, o_ln = 0, o_halo = 0, o_repnp = 0,
         o_blklad = 0, o_sync = 0, o_xpart = 0, o_bar = 0, o_eslot = 0, o_mslot = 0, o_xbuf = 0, o_mprev3 = 0,
         o_fsum = 0, o_fpart = 0, o_fspin = 0;
  int fspin_n = 0;  // component stride of the published-site copies
  const WorkItem* items = nullptr;
  const ItemRep* item_reps = nullptr;
  const int32_t *stream_reps = nullptr, *all_reps = nullptr, *lad_row0 = nullptr, *lad_n = nullptr,
                *halo = nullptr, *blk_lad = nullptr;
  const uint8_t* rep_np = nullptr;
  uint32_t *sync = nullptr, *bar = nullptr;
  double *xpart = nullptr, *eslot = nullptr, *mslot = nullptr, *mprev3 = nullptr;
  double* fsum = nullptr;  // split measurement: FSLOTS per replica, zero between frames
  double* fpart = nullptr;  // persistent split replicas: partial frame sums of each part
  void* fspin = nullptr;    // persistent split replicas: sites published for a frame
  void* xbuf = nullptr;
};

// HIP-event bracketing of the sweep-kernel launches (plan->timing)
// Per-launch kernel timing: a start/stop event pair per timed launch,
// stamped by the kernel's own dispatch (hipExtLaunchKernelGGL), so the
// measured span is the kernel's execution and not the queue around it.
struct LaunchTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;
  double launches = 0, sweeps = 0, updates = 0;
  ~LaunchTimer() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  // two fresh events when timing is on, null otherwise
  int pair(hipEvent_t* a, hipEvent_t* b) {
    *a = *b = nullptr;
    if (!on) return 0;
    HIPCHK(hipEventCreate(a));
    ev.push_back(*a);
    HIPCHK(hipEventCreate(b));
    ev.push_back(*b);
    return 0;
  }
  int finish(double* out) {
    if (!on || !out) return 0;
    double ms = 0;
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      HIPCHK(hipEventSynchronize(ev[i + 1]));
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      ms += t;
    }
    out[0] = ms;
    out[1] = launches;
    out[2] = sweeps;
    out[3] = updates;
    return 0;
  }
};

int device_lds_max(size_t* out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  int v = 0;
  HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  *out = (size_t)v;
  return 0;
}

// link mode of a graph on the LDS path (the MODE of k_lds)
int lds_mode(const mcg_graph& G) {
  const bool wide = G.lcid != nullptr || G.ncls > 256;
  return wide ? 2 : G.npat > 0 ? (G.ucid >= 0 ? 3 : 0) : 1;
}

template <int NC, typename ST, int MODE, bool FULL, bool PERSIST>
int set_lds_attr(size_t bytes) {
  hipError_t e = hipFuncSetAttribute((const void*)k_lds<NC, ST, MODE, FULL, PERSIST>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  (void)e;
  (void)hipGetLastError();
  return 0;
}

// Tables of a replica split into np parts (see the comment on split
// replicas at the LDS kernel).  For each part p:
//   * extra groups: the other parts' groups of each colour that p's own
//     sites depend on, closed under "neighbour of lower colour";
//   * export list: p's own sites that another part reads during a sweep
//     (a recomputed site or a neighbour of one) or in a frame (the second
//     site of a correlation pair, a vertex of a circuit, whose first site
//     that part owns), which the persistent kernel hands over between
//     sweeps.
// The np descriptors are contiguous, DESC(ncol) ints each: per colour
// (start, count) of the extra groups, then (start, count) of the export
// list, all indices into `halo`.  Returns the offset of the first
// descriptor, or -1 when `limit` is set and a part would recompute more
// than a quarter of its own groups (then the replica is not worth
// splitting), or -2 when the link table cannot be read.  *max_export gets
// the longest export list.
int halo_tables(const mcg_graph& G, int gd, int np, bool limit, std::vector<int32_t>& halo, int* max_export) {
  const int n = G.n, zmax = G.zmax, ncol = G.ncol;
  std::vector<uint32_t> link((size_t)n * zmax);
  std::vector<int32_t> colptr(ncol + 1);
  if (zmax > 0 && hipMemcpy(link.data(), G.link, link.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  if (hipMemcpy(colptr.data(), G.col_ptr, colptr.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  std::vector<int32_t> col(n), grp(n), own(n);
  for (int c = 0; c < ncol; ++c) {
    const int b = colptr[c], cnt = colptr[c + 1] - b, ngrp = (cnt + gd - 1) / gd;
    for (int q = 0, part = 0; q < ngrp; ++q) {
      while (part + 1 < np && q >= (int)((int64_t)ngrp * (part + 1) / np)) ++part;
      for (int e = b + q * gd; e < std::min(b + (q + 1) * gd, b + cnt); ++e) {
        col[e] = c;
        grp[e] = q;
        own[e] = part;
      }
    }
  }
  auto nbr = [&](int e, int l) { return (int)(link[(size_t)e * zmax + l] & 0xFFFFFFu); };
  const int DS = halo_desc_ints(ncol);
  const size_t base = halo.size();
  halo.resize(base + (size_t)np * DS);
  std::vector<char> wanted(n, 0);  // read by a part that does not own it
  for (int part = 0; part < np; ++part) {
    std::vector<char> need(n);
    for (int e = 0; e < n; ++e) need[e] = own[e] == part;
    for (int c = ncol - 1; c >= 0; --c)
      for (int e = colptr[c]; e < colptr[c + 1]; ++e) {
        if (!need[e]) continue;
        for (int l = 0; l < zmax; ++l) {
          const int y = nbr(e, l);
          if (y != e && col[y] < c) need[y] = 1;
        }
      }
    for (int e = 0; e < n; ++e) {
      if (!need[e]) continue;
      if (own[e] != part) wanted[e] = 1;
      for (int l = 0; l < zmax; ++l) {
        const int y = nbr(e, l);
        if (own[y] != part) wanted[y] = 1;
      }
    }
    const size_t d = base + (size_t)part * DS;
    int64_t own_groups = 0, extra = 0;
    for (int c = 0; c < ncol; ++c) {
      const int b = colptr[c], cnt = colptr[c + 1] - b, ngrp = (cnt + gd - 1) / gd;
      own_groups += (int64_t)(ngrp * (part + 1) / np) - (int64_t)ngrp * part / np;
      const size_t start = halo.size();
      int last = -1;
      for (int e = b; e < b + cnt; ++e)
        if (need[e] && own[e] != part && grp[e] != last) {
          last = grp[e];
          halo.push_back(last);
        }
      halo[d + 2 * c] = (int32_t)start;
      halo[d + 2 * c + 1] = (int32_t)(halo.size() - start);
      extra += (int64_t)(halo.size() - start);
    }
    if (limit && extra * 4 > own_groups) return -1;
  }
  // the frame's pairs and circuits: local when none has sites of two parts
  std::vector<int32_t> pr((size_t)2 * G.nlat), cr((size_t)3 * G.ncirc);
  int32_t last = 0;  // position of the last site
  if (G.nlat > 0 && hipMemcpy(pr.data(), G.pairs, pr.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  if (G.ncirc > 0 && hipMemcpy(cr.data(), G.circ, cr.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  if (n > 0 && hipMemcpy(&last, G.pos + (n - 1), 4, hipMemcpyDeviceToHost) != hipSuccess) return -2;
  bool local = true;
  for (int k = 0; k < G.nlat; ++k) local = local && own[pr[2 * k + 1]] == own[pr[2 * k]];
  for (int k = 0; k < G.ncirc; ++k)
    for (int v = 1; v < 3; ++v) local = local && own[cr[3 * k + v]] == own[cr[3 * k]];
  int mx = 0;
  for (int part = 0; part < np; ++part) {
    const size_t d = base + (size_t)part * DS;
    const size_t start = halo.size();
    for (int e = 0; e < n; ++e)
      if (wanted[e] && own[e] == part) halo.push_back(e);
    halo[d + 2 * ncol] = (int32_t)start;
    halo[d + 2 * ncol + 1] = (int32_t)(halo.size() - start);
    mx = std::max(mx, (int)(halo.size() - start));
    // the part's share of the frame
    const size_t ps = halo.size();
    for (int k = 0; k < G.nlat; ++k)
      if (own[pr[2 * k]] == part) halo.push_back(k);
    halo[d + 2 * ncol + 2] = (int32_t)ps;
    halo[d + 2 * ncol + 3] = (int32_t)(halo.size() - ps);
    const size_t cs = halo.size();
    for (int k = 0; k < G.ncirc; ++k)
      if (own[cr[3 * k]] == part) halo.push_back(k);
    halo[d + 2 * ncol + 4] = (int32_t)cs;
    halo[d + 2 * ncol + 5] = (int32_t)(halo.size() - cs);
    halo[d + 2 * ncol + 6] = n > 0 && own[last] == part;
    halo[d + 2 * ncol + 7] = local;
  }
  *max_export = std::max(*max_export, mx);
  return (int)base;
}

// Plan the LDS work items.
//
// Replicas of one graph are packed R to a workgroup (R divides the wave
// count, so each replica gets whole waves).  R is bounded by the LDS budget
// and by a work target t (site updates per workgroup per sweep): t starts at
// the batch's updates per sweep over the CU count, so that the slowest
// workgroup carries about one CU's share.  A replica of n >= 8192 sites
// larger than t is split over P = min(4, ceil(n/t), n/2048) workgroups
// (parts); a launch with parts does one sweep.  The plan keeps at most one
// workgroup per CU (t grows by 5 % until it does), because the LDS the
// largest item asks for is given to every workgroup.  P->parts = 1 disables
// splitting, 2..4 splits every LDS replica into that many parts.  Parts of a
// replica are placed eight workgroup indices apart: workgroups are dealt to
// the eight XCDs in turn, so the parts share an XCD and its L2.
int plan_lds(const mcg_batch* B, const mcg_plan* P, RunCtx& X, bool force_stream) {
  const int ss = store_size(B->store);
  const int gd = B->nc == 1 ? 4 : B->nc == 2 ? 2 : 1;  // sites per Philox draw
  X.hitems.clear();
  X.hitem_reps.clear();
  X.hstream.clear();
  X.hall.clear();
  X.hhalo.clear();
  X.nslots = 0;
  for (int r = 0; r < B->nrep; ++r) X.hall.push_back(r);
  if (P->parts < 0 || P->parts > MAX_PARTS) return fail("parts must be in 0..%d", MAX_PARTS);
  int block = P->block;
  if (block <= 0) {
    int maxcol = 0;
    for (int gi = 0; gi < B->ngraph; ++gi) {
      const mcg_graph& G = B->graphs_host[gi];
      maxcol = std::max(maxcol, (G.n + G.ncol - 1) / std::max(G.ncol, 1));
    }
    block = maxcol >= 1024 ? 1024 : 256;
  }
  if (block % 64 || block < 64 || block > 1024) return fail("block must be a multiple of 64 in [64,1024]");
  X.block = block;
  const int nwaves = block / 64;
  // group replicas by graph, preserving order; Rmax from the LDS budget
  std::vector<std::vector<int>> by(B->ngraph);
  for (int r = 0; r < B->nrep; ++r) by[B->reps_host[r].graph].push_back(r);
  std::vector<int> rmax(B->ngraph, 0);
  double work = 0;
  for (int gi = 0; gi < B->ngraph; ++gi) {
    const mcg_graph& G = B->graphs_host[gi];
    if (by[gi].empty()) continue;
    const size_t per = lds_replica_bytes(B->nc, G.n, ss);
    if (!force_stream) {
      for (int R = nwaves; R >= 1; --R) {
        if (nwaves % R) continue;
        if (P->reps_per_block > 0 && R > P->reps_per_block) continue;
        const bool wide = G.lcid != nullptr || G.ncls > 256;
        if (lds_fixed_bytes(wide ? 0 : G.ncoup, wide ? 0 : G.ncls, G.ncol, G.npat, G.zmax, G.n, nwaves, R) + R * per <=
            X.lds_max) {
          rmax[gi] = R;
          break;
        }
      }
    }
    if (rmax[gi] == 0)
      for (int r : by[gi]) X.hstream.push_back(r);
    else
      work += (double)G.n * by[gi].size();
  }
  if (work == 0) return 0;
  const int ncu = std::max(1, X.ncu);
  std::vector<char> nosplit(B->ngraph, 0);
  // automatic splits only where the launch can be persistent: otherwise a
  // launch with split replicas runs one sweep and restages them every sweep,
  // which costs more than the imbalance it removes.  With spin-derived
  // observables a persistent launch needs pattern links (modes 0 and 3).
  // A part recomputes every group its own sites depend on through lower
  // colours, a cone that widens by a link per colour; halo_tables refuses
  // the split when that is more than a quarter of the part's own groups.
  // Parts keep at least 2048 sites.
  const bool auto_split = !P->no_persist;
  const bool spin_obs = (B->meas & MEAS_SPIN_OBS) != 0;
  struct Split {
    int gi, rep, np;
  };
  for (int attempt = 0; attempt <= B->ngraph; ++attempt) {
    double t = std::max(1.0, std::ceil(work / ncu));
    std::vector<WorkItem> items;
    std::vector<int32_t> ireps;
    std::vector<Split> splits;
    std::vector<int> gnp(B->ngraph, 1);
    for (int iter = 0; iter < 400; ++iter) {
      items.clear();
      ireps.clear();
      splits.clear();
      size_t nsplit_blocks = 0;
      for (int gi = 0; gi < B->ngraph; ++gi) {
        gnp[gi] = 1;
        if (rmax[gi] == 0) continue;
        const mcg_graph& G = B->graphs_host[gi];
        const std::vector<int>& rs = by[gi];
        int np = 1;
        if (P->parts >= 2)
          np = P->parts;
        else if (P->parts == 0 && auto_split && !nosplit[gi] && G.n >= 8192 && G.n > t &&
                 (!spin_obs || lds_mode(G) == 0 || lds_mode(G) == 3))
          np = std::min({MAX_PARTS, (int)std::ceil(G.n / t), G.n / 2048});
        if (np >= 2) {
          gnp[gi] = np;
          for (int r : rs) splits.push_back(Split{gi, r, np});
          nsplit_blocks += (size_t)np * rs.size();
          continue;
        }
        int R = rmax[gi];
        if (P->reps_per_block <= 0) R = std::min(R, std::max(1, (int)(t / G.n)));
        while (nwaves % R) --R;
        for (size_t i = 0; i < rs.size(); i += R) {
          const int cnt = (int)std::min<size_t>(R, rs.size() - i);
          // the tail item takes the largest divisor of nwaves that fits, repeatedly
          for (int q = 0; q < cnt;) {
            int Rq = cnt - q;
            while (nwaves % Rq) --Rq;
            items.push_back(WorkItem{gi, Rq, 0, 0, 1, 0, 0});
            for (int k = 0; k < Rq; ++k) ireps.push_back(rs[i + q + k]);
            q += Rq;
          }
        }
      }
      if (P->parts == 0 && !splits.empty() && items.size() + nsplit_blocks > (size_t)ncu) {
        t = t * 1.05 + 1.0;
        continue;
      }
      break;
    }
    if (splits.size() > 65535) return fail("too many split replicas");
    // split tables, one set per split graph
    std::vector<int> hbase(B->ngraph, -1);
    bool redo = false;
    X.hhalo.clear();
    X.max_export = 0;
    for (int gi = 0; gi < B->ngraph; ++gi) {
      if (gnp[gi] < 2) continue;
      hbase[gi] = halo_tables(B->graphs_host[gi], gd, gnp[gi], P->parts == 0, X.hhalo, &X.max_export);
      if (hbase[gi] == -2) return fail("graph %d: reading the link table for the split failed", gi);
      if (hbase[gi] < 0) {
        nosplit[gi] = 1;
        redo = true;
      }
    }
    if (redo) continue;
    // place parts: in each run of up to 8 split replicas with np parts,
    // part p of the run's replica j goes to index base + j + 8 p
    std::vector<WorkItem> out;
    std::vector<int32_t> oreps;
    X.nslots = 0;
    for (size_t j = 0; j < splits.size();) {
      const int np = splits[j].np;
      size_t e = j;
      while (e < splits.size() && e - j < 8 && splits[e].np == np) ++e;
      std::vector<WorkItem> run((size_t)8 * np, WorkItem{-1, 0, 0, 0, 0, 0, 0});
      std::vector<int32_t> run_rep((size_t)8 * np, -1);
      for (size_t q = j; q < e; ++q) {
        const uint16_t slot = (uint16_t)X.nslots++;
        const int gi = splits[q].gi;
        for (int p = 0; p < np; ++p) {
          run[(q - j) + (size_t)8 * p] =
              WorkItem{gi, 1, 0, (uint8_t)p, (uint8_t)np, slot,
                       hbase[gi] + p * halo_desc_ints(B->graphs_host[gi].ncol)};
          run_rep[(q - j) + (size_t)8 * p] = splits[q].rep;
        }
      }
      for (size_t q = 0; q < run.size(); ++q) {
        if (run[q].graph < 0) continue;  // holes of a short run
        run[q].off = (int)oreps.size();
        oreps.push_back(run_rep[q]);
        out.push_back(run[q]);
      }
      j = e;
    }
    size_t ir = 0;
    for (WorkItem w : items) {
      w.off = (int)oreps.size();
      for (int k = 0; k < w.nrep; ++k) oreps.push_back(ireps[ir + k]);
      ir += w.nrep;
      out.push_back(w);
    }
    // one kernel per link mode: sort the items by mode (stable, so the
    // parts of a split replica stay eight indices apart)
    std::stable_sort(out.begin(), out.end(), [&](const WorkItem& x, const WorkItem& y) {
      return lds_mode(B->graphs_host[x.graph]) < lds_mode(B->graphs_host[y.graph]);
    });
    for (int m = 0; m <= 4; ++m) X.mode_begin[m] = 0;
    for (const WorkItem& w : out) X.mode_begin[lds_mode(B->graphs_host[w.graph]) + 1]++;
    for (int m = 1; m <= 4; ++m) X.mode_begin[m] += X.mode_begin[m - 1];
    X.hitems = out;
    X.hitem_reps = oreps;
    return 0;
  }
  return fail("internal: LDS plan did not converge");
}

template <int NC, typename ST>
size_t lds_bytes_for(const mcg_batch* B, const RunCtx& X, int i0, int i1) {
  size_t mx = 0;
  const int nwaves = X.block / 64;
  for (int i = i0; i < i1; ++i) {
    const WorkItem& w = X.hitems[i];
    const mcg_graph& G = B->graphs_host[w.graph];
    const bool wide = G.lcid != nullptr || G.ncls > 256;
    const size_t b = lds_fixed_bytes(wide ? 0 : G.ncoup, wide ? 0 : G.ncls, G.ncol, G.npat, G.zmax, G.n, nwaves,
                                     w.nrep) +
                     w.nrep * lds_replica_bytes(NC, G.n, sizeof(ST));
    mx = std::max(mx, b);
  }
  return mx;
}

template <int NC, typename ST, int MODE, bool FULL, bool PERSIST>
int launch_lds_range(const mcg_batch* B, RunCtx& X, const SweepArgs& a, hipStream_t s, int i0, int i1, hipEvent_t e0,
                     hipEvent_t e1) {
  size_t bytes = lds_bytes_for<NC, ST>(B, X, i0, i1);
  if (PERSIST) bytes += persist_lds_bytes(B->nrow, B->nrep);
  set_lds_attr<NC, ST, MODE, FULL, PERSIST>(bytes);
  hipExtLaunchKernelGGL((k_lds<NC, ST, MODE, FULL, PERSIST>), dim3((unsigned)(i1 - i0)), dim3(X.block),
                        (uint32_t)bytes, s, e0, e1, 0, *B, X.items + i0, X.item_reps, a);
  HIPCHK(hipGetLastError());
  return 0;
}

// one launch per link mode present (work items are sorted by mode)
template <int NC, typename ST>
int launch_lds(const mcg_batch* B, RunCtx& X, const SweepArgs& a, hipStream_t s, LaunchTimer& T) {
  const bool full = (B->meas & MEAS_SPIN_OBS) != 0;
  for (int m = 0; m < 4; ++m) {
    const int i0 = X.mode_begin[m], i1 = X.mode_begin[m + 1];
    if (i1 <= i0) continue;
    hipEvent_t e0, e1;
    if (T.pair(&e0, &e1)) return -1;
    int r;
    if (X.persist && full)  // single pattern mode (checked when planned)
      r = m == 0 ? launch_lds_range<NC, ST, 0, true, true>(B, X, a, s, i0, i1, e0, e1)
                 : launch_lds_range<NC, ST, 3, true, true>(B, X, a, s, i0, i1, e0, e1);
    else if (X.persist)  // single mode (checked when planned)
      r = m == 0   ? launch_lds_range<NC, ST, 0, false, true>(B, X, a, s, i0, i1, e0, e1)
          : m == 1 ? launch_lds_range<NC, ST, 1, false, true>(B, X, a, s, i0, i1, e0, e1)
          : m == 2 ? launch_lds_range<NC, ST, 2, false, true>(B, X, a, s, i0, i1, e0, e1)
                   : launch_lds_range<NC, ST, 3, false, true>(B, X, a, s, i0, i1, e0, e1);
    else if (m == 0)
      r = full ? launch_lds_range<NC, ST, 0, true, false>(B, X, a, s, i0, i1, e0, e1)
               : launch_lds_range<NC, ST, 0, false, false>(B, X, a, s, i0, i1, e0, e1);
    else if (m == 1)
      r = full ? launch_lds_range<NC, ST, 1, true, false>(B, X, a, s, i0, i1, e0, e1)
               : launch_lds_range<NC, ST, 1, false, false>(B, X, a, s, i0, i1, e0, e1);
    else if (m == 2)
      r = full ? launch_lds_range<NC, ST, 2, true, false>(B, X, a, s, i0, i1, e0, e1)
               : launch_lds_range<NC, ST, 2, false, false>(B, X, a, s, i0, i1, e0, e1);
    else
      r = full ? launch_lds_range<NC, ST, 3, true, false>(B, X, a, s, i0, i1, e0, e1)
               : launch_lds_range<NC, ST, 3, false, false>(B, X, a, s, i0, i1, e0, e1);
    if (r) return -1;
    T.launches += 1;
  }
  return 0;
}

// Can this run use a persistent launch?  All replicas on the LDS path, one
// link mode (a pattern mode when spin-derived observables are measured),
// tempering or split replicas to synchronise (otherwise a plain multi-sweep
// launch already is persistent), the tempering state fits in LDS, and every
// workgroup resident at once.
template <int NC, typename ST>
int check_persist(const mcg_batch* B, const mcg_plan* P, RunCtx& X, bool* ok) {
  *ok = false;
  if (X.hitems.empty() || !X.hstream.empty()) return 0;
  int modes = 0, m = -1;
  for (int q = 0; q < 4; ++q)
    if (X.mode_begin[q + 1] > X.mode_begin[q]) ++modes, m = q;
  if (modes != 1) return 0;
  const bool full = (B->meas & MEAS_SPIN_OBS) != 0;
  if (full && m != 0 && m != 3) return 0;
  const bool pt = P->pt_every > 0 && !X.hl_row0.empty();
  if (!pt && X.nslots == 0) return 0;
  const size_t bytes = lds_bytes_for<NC, ST>(B, X, 0, (int)X.hitems.size()) + persist_lds_bytes(B->nrow, B->nrep);
  if (bytes > X.lds_max) return 0;
  const void* f = full ? (m == 0 ? (const void*)k_lds<NC, ST, 0, true, true> : (const void*)k_lds<NC, ST, 3, true, true>)
                  : m == 0 ? (const void*)k_lds<NC, ST, 0, false, true>
                  : m == 1 ? (const void*)k_lds<NC, ST, 1, false, true>
                  : m == 2 ? (const void*)k_lds<NC, ST, 2, false, true>
                           : (const void*)k_lds<NC, ST, 3, false, true>;
  if (full && m == 0) set_lds_attr<NC, ST, 0, true, true>(bytes);
  else if (full) set_lds_attr<NC, ST, 3, true, true>(bytes);
  else if (m == 0) set_lds_attr<NC, ST, 0, false, true>(bytes);
  else if (m == 1) set_lds_attr<NC, ST, 1, false, true>(bytes);
  else if (m == 2) set_lds_attr<NC, ST, 2, false, true>(bytes);
  else set_lds_attr<NC, ST, 3, false, true>(bytes);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, X.block, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  *ok = (int64_t)per_cu * X.ncu >= (int64_t)X.hitems.size();
  return 0;
}


template <int NC, typename ST>
int launch_pass(const mcg_batch* B, const int32_t* dreps, int nreps, int maxcnt, int colour, int64_t sweep,
                hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (nreps == 0) return 0;
  const int ngrp = (maxcnt + SitesPerDraw<NC>::G - 1) / SitesPerDraw<NC>::G;
  const int chunks = std::max(1, (ngrp + 256 * PASS_GPT - 1) / (256 * PASS_GPT));
  const int64_t items = (int64_t)chunks * nreps;
  if (items > (1 << 30)) return fail("streaming pass: %lld work items", (long long)items);
  hipExtLaunchKernelGGL((k_pass<NC, ST>), dim3((unsigned)xcd_grid((int)items)), dim3(256), 0, s, e0, e1, 0, *B, dreps,
                        colour, sweep, chunks, (int)items);
  HIPCHK(hipGetLastError());
  return 0;
}

template <int NC, typename ST>
int launch_measure(const mcg_batch* B, const int32_t* dreps, int nreps, int mode, int64_t frame, hipStream_t s) {
  if (nreps == 0) return 0;
  hipLaunchKernelGGL((k_measure<NC, ST>), dim3((unsigned)nreps), dim3(1024), 0, s, *B, dreps, mode, frame);
  HIPCHK(hipGetLastError());
  return 0;
}

// workgroups per replica of the split measurement: about MCG_FRAME_WG_PER_CU
// per CU over the batch, at least 1024 sites each
#ifndef MCG_FRAME_WG_PER_CU
#define MCG_FRAME_WG_PER_CU 16
#endif
int frame_parts(int nreps, int nmax, int ncu) {
  if (nreps <= 0) return 1;
  const int want = (MCG_FRAME_WG_PER_CU * std::max(ncu, 1) + nreps - 1) / nreps;
  return std::max(1, std::min({want, std::max(1, nmax / 1024), 65535}));
}

template <int NC, typename ST>
// spins: a measured frame reads spins (a spin-derived observable other than
// self pairs of an O(n) model); the recompute (mode bit 1) always does
int launch_frame(const mcg_batch* B, const int32_t* dreps, int nreps, int mode, int64_t frame, double* fsum, int np,
                 bool spins, hipStream_t s) {
  if (nreps == 0) return 0;
  if ((mode & 1) || ((mode & 4) && spins)) {
    const int64_t items = (int64_t)np * nreps;
    if (items > (1 << 30)) return fail("split measurement: %lld work items", (long long)items);
    hipLaunchKernelGGL((k_frame_part<NC, ST>), dim3((unsigned)xcd_grid((int)items)), dim3(256), 0, s, *B, dreps,
                       mode, frame, fsum, np, (int)items);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL((k_frame_end<NC, ST>), dim3((unsigned)nreps), dim3(64), 0, s, *B, dreps, mode, frame, fsum);
  HIPCHK(hipGetLastError());
  return 0;
}

int launch_fill_f64(double* p, size_t n, double v, hipStream_t s) {
  if (!p || n == 0) return 0;
  const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f64, dim3(g), dim3(256), 0, s, p, n, v);
  HIPCHK(hipGetLastError());
  return 0;
}
int launch_fill_i32(int32_t* p, size_t n, int32_t v, hipStream_t s) {
  if (!p || n == 0) return 0;
  const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, p, n, v);
  HIPCHK(hipGetLastError());
  return 0;
}

// planned sweep schedule for the streaming / plain paths
template <int NC, typename ST>
int stream_sweeps(const mcg_batch* B, const mcg_plan* P, RunCtx& X, int64_t k0, int64_t nsw, int64_t nthermal,
                  int spf, bool plain, hipStream_t s, LaunchTimer& T) {
  if (X.hstream.empty()) return 0;
  const int ncolmax = X.stream_ncol;
  const int maxn = X.stream_maxcnt;  // largest colour class of the streamed graphs
  const int32_t* dreps = X.stream_reps;
  const int nrs = (int)X.hstream.size();
  for (int64_t q = 0; q < nsw; ++q) {
    const int64_t k = k0 + q;
    const int64_t kk = P->sweep0 + k;
    for (int c = 0; c < ncolmax; ++c) {
      if (plain) {
        for (int i = 0; i < nrs; ++i) {
          hipEvent_t e0, e1;
          if (T.pair(&e0, &e1) || launch_pass<NC, ST>(B, dreps + i, 1, maxn, c, kk, s, e0, e1)) return -1;
          T.launches += 1;
        }
      } else {
        hipEvent_t e0, e1;
        if (T.pair(&e0, &e1) || launch_pass<NC, ST>(B, dreps, nrs, maxn, c, kk, s, e0, e1)) return -1;
        T.launches += 1;
      }
    }
    if (T.on) {
      T.sweeps += 1;
      for (int r : X.hstream) T.updates += B->graphs_host[B->reps_host[r].graph].n;
    }
    const bool meas = k >= nthermal && ((k - nthermal + 1) % spf) == 0;
    const bool therm_end = (k + 1 == nthermal);
    const bool rec = recompute_due(*B, P->sweep0 + k);
    int mode = (rec ? 1 : 0) | (therm_end ? 2 : 0) | (meas ? 4 : 0);
    if (mode) {
      const int64_t frame = meas ? (k - nthermal + 1) / spf - 1 + P->frame0 : -1;
      if (plain) {
        for (int i = 0; i < nrs; ++i)
          if (launch_measure<NC, ST>(B, dreps + i, 1, mode, frame, s)) return -1;
      } else if (launch_frame<NC, ST>(B, dreps, nrs, mode, frame, X.fsum, X.frame_np_stream, X.frame_spins, s)) {
        return -1;
      }
    }
  }
  return 0;
}

// piece of the device region, 256-byte aligned
size_t take(size_t* off, size_t bytes) {
  const size_t o = *off;
  *off += (bytes + 255) & ~(size_t)255;
  return o;
}

// Plan the run of B under P and lay out its device region: fills everything
// in X but the device addresses (bind_region).  X.dev_bytes is the size of
// the region the run needs.
template <int NC, typename ST>
int prepare_typed(const mcg_batch* B, const mcg_plan* P, RunCtx& X) {
  HIPCHK(hipGetDevice(&X.device));
  if (device_lds_max(&X.lds_max)) return -1;
  HIPCHK(hipDeviceGetAttribute(&X.ncu, hipDeviceAttributeMultiprocessorCount, X.device));
  const bool force_stream = P->kernel == 2 || P->kernel == 3;
  if (plan_lds(B, P, X, force_stream)) return -1;
  if (P->kernel == 1 && !X.hstream.empty())
    return fail("LDS-resident kernel requested but %zu replica(s) do not fit in LDS", X.hstream.size());
  // tempering ladders (row0, rungs), each of one lattice
  X.hl_row0.clear();
  X.hl_n.clear();
  {
    std::vector<int32_t> seen(B->nrow, 0), lgraph(B->nrow, -1);
    for (int r = 0; r < B->nrep; ++r) {
      const mcg_replica& R = B->reps_host[r];
      if (R.ladder < 0 || R.nrung < 2) continue;
      if (lgraph[R.row0] >= 0 && lgraph[R.row0] != R.graph)
        return fail("replica %d: a tempering ladder holds replicas of one lattice", r);
      lgraph[R.row0] = R.graph;
      if (seen[R.row0]) continue;
      seen[R.row0] = 1;
      X.hl_row0.push_back(R.row0);
      X.hl_n.push_back(R.nrung);
    }
  }
  // colour classes of the streamed graphs (the streaming grid covers the largest)
  X.stream_ncol = X.stream_maxcnt = 0;
  {
    std::vector<char> seen(B->ngraph, 0);
    for (int r : X.hstream) {
      const int gi = B->reps_host[r].graph;
      if (seen[gi]) continue;
      seen[gi] = 1;
      const mcg_graph& G = B->graphs_host[gi];
      std::vector<int32_t> cp(G.ncol + 1);
      HIPCHK(hipMemcpy(cp.data(), G.col_ptr, cp.size() * 4, hipMemcpyDeviceToHost));
      X.stream_ncol = std::max(X.stream_ncol, G.ncol);
      for (int c = 0; c < G.ncol; ++c) X.stream_maxcnt = std::max(X.stream_maxcnt, cp[c + 1] - cp[c]);
    }
  }
  {
    int nmax = 1;
    for (int gi = 0; gi < B->ngraph; ++gi) nmax = std::max(nmax, B->graphs_host[gi].n);
    X.frame_np_stream = frame_parts((int)X.hstream.size(), nmax, X.ncu);
    // the frame reads spins unless its only spin-derived observable is the
    // self pairs of an O(n) model
    X.frame_spins = (B->meas & (MCG_MEAS_Q | MCG_MEAS_RNORM | MCG_MEAS_GROUPS)) != 0;
    if (B->meas & MCG_MEAS_PAIRS)
      for (int gi = 0; gi < B->ngraph; ++gi)
        X.frame_spins = X.frame_spins || B->nc == 1 || !B->graphs_host[gi].self_pairs;
    X.frame_np_all = frame_parts(B->nrep, nmax, X.ncu);
  }
  X.lds_updates_per_sweep = 0;
  for (const WorkItem& w : X.hitems)
    if (w.part == 0) X.lds_updates_per_sweep += (double)w.nrep * B->graphs_host[w.graph].n;
  // persistent launches when possible (P->no_persist turns them off)
  X.persist = false;
  if (P->algorithm == MCG_METROPOLIS && !P->no_persist && !force_stream && check_persist<NC, ST>(B, P, X, &X.persist))
    return -1;
  // workgroup ladder lists of a persistent launch: the ladders of each
  // workgroup's replicas; the first workgroup of a ladder owns it
  std::vector<int32_t> bl;
  std::vector<uint8_t> np;
  if (X.persist) {
    np.assign(B->nrep, 1);
    for (const WorkItem& w : X.hitems)
      if (w.nparts > 1) np[X.hitem_reps[w.off]] = w.nparts;
    std::vector<int32_t> lad_of_row(B->nrow, -1);
    for (size_t L = 0; L < X.hl_row0.size(); ++L) lad_of_row[X.hl_row0[L]] = (int32_t)L;
    const size_t nb = X.hitems.size();
    bl.assign(nb + 1, 0);
    std::vector<char> owned(X.hl_row0.size(), 0);
    for (size_t i = 0; i < nb; ++i) {
      const WorkItem& w = X.hitems[i];
      std::vector<int32_t> ls;
      for (int q = 0; q < w.nrep; ++q) {
        const mcg_replica& R = B->reps_host[X.hitem_reps[w.off + q]];
        const int L = R.ladder >= 0 && R.nrung >= 2 ? lad_of_row[R.row0] : -1;
        if (L >= 0 && std::find(ls.begin(), ls.end(), L) == ls.end()) ls.push_back(L);
      }
      std::sort(ls.begin(), ls.end());
      if (ls.size() > (size_t)MAX_LADW) return fail("internal: more than %d ladders in one workgroup", MAX_LADW);
      for (int32_t L : ls) {
        bl.push_back(owned[L] ? L : (int32_t)(L | 0x80000000u));
        owned[L] = 1;
      }
      bl[i + 1] = (int32_t)(bl.size() - (nb + 1));
    }
    for (size_t i = 0; i <= nb; ++i) bl[i] += (int32_t)(nb + 1);
  }
  std::vector<ItemRep> recs(X.hitem_reps.size());
  for (size_t i = 0; i < recs.size(); ++i) {
    const mcg_replica& R = B->reps_host[X.hitem_reps[i]];
    recs[i] = ItemRep{R.spin_off, X.hitem_reps[i], (int32_t)R.serial, R.row0, 0};
  }
  // layout: tables first, then the state every run clears
  size_t off = 0;
  X.o_items = take(&off, X.hitems.size() * sizeof(WorkItem));
  X.o_item_reps = take(&off, recs.size() * sizeof(ItemRep));
  X.o_stream = take(&off, X.hstream.size() * 4);
  X.o_all = take(&off, X.hall.size() * 4);
  X.o_lrow0 = take(&off, X.hl_row0.size() * 4);
  X.o_ln = take(&off, X.hl_n.size() * 4);
  X.o_halo = take(&off, X.hhalo.size() * 4);
  X.o_repnp = take(&off, np.size());
  X.o_blklad = take(&off, bl.size() * 4);
  X.tab_bytes = off;
  X.reset_off = off;
  X.o_fsum = take(&off, (size_t)B->nrep * FSLOTS * 8);
  X.o_sync = take(&off, (size_t)X.nslots * SYNC_STRIDE * 4);
  X.o_xpart = take(&off, (size_t)X.nslots * MAX_PARTS * 4 * 8);
  if (X.persist) {
    X.o_bar = take(&off, (size_t)BAR_WORDS * 4);
    X.o_eslot = take(&off, (size_t)2 * B->nrep * MAX_PARTS * 8);
    X.o_mslot = take(&off, (size_t)2 * B->nrep * MAX_PARTS * 3 * 8);
    X.o_xbuf = take(&off, (size_t)2 * std::max(X.nslots, 1) * MAX_PARTS * std::max(X.max_export, 1) * NC * sizeof(ST));
    X.o_mprev3 = take(&off, (size_t)3 * B->nrow * 8);
    X.o_fpart = take(&off, (size_t)2 * std::max(X.nslots, 1) * MAX_PARTS * FSUMS * 8);
  }
  X.reset_bytes = off - X.reset_off;
  if (X.persist && X.nslots > 0 && (B->meas & MEAS_SPIN_OBS)) {
    // written in full before every read: outside the reset region
    int nmax = 0;
    for (int gi = 0; gi < B->ngraph; ++gi) nmax = std::max(nmax, B->graphs_host[gi].n);
    X.fspin_n = (int)((nmax + 63) / 64 * 64);
    X.o_fspin = take(&off, (size_t)2 * X.nslots * NC * X.fspin_n * sizeof(ST));
  }
  X.dev_bytes = off;
  X.hstage.assign(X.tab_bytes, 0);
  auto put = [&](size_t o, const void* src, size_t n) {
    if (n) std::memcpy(X.hstage.data() + o, src, n);
  };
  put(X.o_items, X.hitems.data(), X.hitems.size() * sizeof(WorkItem));
  put(X.o_item_reps, recs.data(), recs.size() * sizeof(ItemRep));
  put(X.o_stream, X.hstream.data(), X.hstream.size() * 4);
  put(X.o_all, X.hall.data(), X.hall.size() * 4);
  put(X.o_lrow0, X.hl_row0.data(), X.hl_row0.size() * 4);
  put(X.o_ln, X.hl_n.data(), X.hl_n.size() * 4);
  put(X.o_halo, X.hhalo.data(), X.hhalo.size() * 4);
  put(X.o_repnp, np.data(), np.size());
  put(X.o_blklad, bl.data(), bl.size() * 4);
  // the key
  X.k_kernel = P->kernel;
  X.k_block = P->block;
  X.k_rpb = P->reps_per_block;
  X.k_parts = P->parts;
  X.k_nopersist = P->no_persist;
  X.k_pt = P->pt_every > 0;
  X.k_nc = B->nc;
  X.k_store = B->store;
  X.k_ngraph = B->ngraph;
  X.k_nrep = B->nrep;
  X.k_nrow = B->nrow;
  X.k_meas = B->meas;
  X.k_rec = B->recompute_every;
  X.k_graphs = B->graphs;
  X.k_reps = B->reps;
  X.k_graphs_host = B->graphs_host;
  X.k_reps_host = B->reps_host;
  return 0;
}

// Give X its device region (X.dev_bytes at dev) and copy the tables there,
// on stream s.  The caller keeps X.hstage alive until the copy is done.
int bind_region(RunCtx& X, unsigned char* dev, hipStream_t s) {
  X.dev = dev;
  auto at = [&](size_t o) { return (void*)(dev + o); };
  X.items = (const WorkItem*)at(X.o_items);
  X.item_reps = (const ItemRep*)at(X.o_item_reps);
  X.stream_reps = (const int32_t*)at(X.o_stream);
  X.all_reps = (const int32_t*)at(X.o_all);
  X.lad_row0 = (const int32_t*)at(X.o_lrow0);
  X.lad_n = (const int32_t*)at(X.o_ln);
  X.halo = (const int32_t*)at(X.o_halo);
  X.rep_np = (const uint8_t*)at(X.o_repnp);
  X.blk_lad = (const int32_t*)at(X.o_blklad);
  X.sync = (uint32_t*)at(X.o_sync);
  X.xpart = (double*)at(X.o_xpart);
  X.bar = (uint32_t*)at(X.o_bar);
  X.eslot = (double*)at(X.o_eslot);
  X.mslot = (double*)at(X.o_mslot);
  X.xbuf = at(X.o_xbuf);
  X.mprev3 = (double*)at(X.o_mprev3);
  X.fsum = (double*)at(X.o_fsum);
  X.fpart = (double*)at(X.o_fpart);
  X.fspin = at(X.o_fspin);
  if (X.tab_bytes && hipMemcpyAsync(dev, X.hstage.data(), X.tab_bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail("hipMemcpyAsync of the run tables failed");
  return 0;
}

// Is X the plan of this batch and plan?
int check_key(const mcg_batch* B, const mcg_plan* P, const RunCtx& X) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev != X.device) return fail("run prepared on device %d, called on device %d", X.device, dev);
  if (P->kernel != X.k_kernel || P->block != X.k_block || P->reps_per_block != X.k_rpb || P->parts != X.k_parts ||
      P->no_persist != X.k_nopersist || (P->pt_every > 0) != (X.k_pt != 0))
    return fail("plan differs from the prepared one in kernel, block, reps_per_block, parts, no_persist or "
                "tempering on/off; prepare a new run for it");
  if (B->nc != X.k_nc || B->store != X.k_store || B->ngraph != X.k_ngraph || B->nrep != X.k_nrep ||
      B->nrow != X.k_nrow || B->meas != X.k_meas || B->recompute_every != X.k_rec || B->graphs != X.k_graphs ||
      B->reps != X.k_reps || B->graphs_host != X.k_graphs_host || B->reps_host != X.k_reps_host)
    return fail("batch differs from the one the run was prepared for");
  return 0;
}

// The launches of one run of a prepared context (no allocation, no host
// copy): clear the run state, then the sweep / cluster / tempering launches.
template <int NC, typename ST>
int exec_typed(const mcg_batch* B, const mcg_plan* P, hipStream_t s, RunCtx& X) {
  const bool plain = P->kernel == 3;
  const int spf = P->sweeps_per_frame > 0 ? P->sweeps_per_frame : 1;
  const int64_t nth = P->nthermal;
  const int64_t total = nth + (int64_t)P->nsweep * spf;
  const int nall = B->nrep;
  if (X.nslots > 0 && total * MAX_PARTS >= ((int64_t)1 << 32))
    return fail("too many sweeps in one call for split replicas (%lld); split the call", (long long)total);
  if (X.reset_bytes) {
    const size_t nw = X.reset_bytes / 4;
    const unsigned g = (unsigned)std::min<size_t>((nw + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, (uint32_t*)(X.dev + X.reset_off), nw, 0u);
    HIPCHK(hipGetLastError());
  }
  const bool pt_on = P->pt_every > 0 && !X.hl_row0.empty();

  if (P->algorithm == MCG_WOLFF) {
    if (!B->wstamp || !B->wlist) return fail("Wolff needs wstamp/wlist scratch");
    int64_t nmax = 0, wmax = 0;
    for (int gi = 0; gi < B->ngraph; ++gi) {
      nmax = std::max<int64_t>(nmax, B->graphs_host[gi].n);
      wmax = std::max<int64_t>(wmax, (int64_t)B->graphs_host[gi].n * (B->graphs_host[gi].zmax + 1));
    }
    if (wmax >= ((int64_t)1 << 32)) return fail("Wolff: sites x (links + 1) must stay below 2^32");
    for (int64_t k = 0; k < total; ++k) {
      const int64_t kk = P->sweep0 + k;
      // clusters per sweep: plan value or n of the replica's graph (use max n)
      const int ncl = P->clusters_per_sweep > 0 ? P->clusters_per_sweep : (int)nmax;
      hipLaunchKernelGGL((k_wolff<NC, ST>), dim3((unsigned)nall), dim3(1024), 0, s, *B, X.all_reps,
                         kk * (int64_t)ncl, ncl);
      HIPCHK(hipGetLastError());
      const bool meas = k >= nth && ((k - nth + 1) % spf) == 0;
      const bool therm_end = (k + 1 == nth);
      const int mode = 1 | 8 | (therm_end ? 2 : 0) | (meas ? 4 : 0);
      if (meas || therm_end || k + 1 == total) {
        const int64_t frame = meas ? (k - nth + 1) / spf - 1 + P->frame0 : -1;
        if (launch_frame<NC, ST>(B, X.all_reps, nall, mode, frame, X.fsum, X.frame_np_all, X.frame_spins, s)) return -1;
      }
      if (pt_on && ((P->sweep0 + k + 1) % P->pt_every) == 0) {
        if (launch_frame<NC, ST>(B, X.all_reps, nall, 1 | 8, -1, X.fsum, X.frame_np_all, X.frame_spins, s)) return -1;
        hipLaunchKernelGGL(k_pt, dim3((unsigned)X.hl_row0.size()), dim3(64), 0, s, *B, X.lad_row0, X.lad_n,
                           (int)X.hl_row0.size(), P->sweep0 + k);
        HIPCHK(hipGetLastError());
      }
    }
    return 0;
  }

  LaunchTimer T;
  T.on = P->timing != nullptr;
  // a persistent launch runs up to 4096 sweeps; otherwise split replicas
  // finish every launch (one sweep per launch) and tempering needs a launch
  // boundary every pt_every sweeps
  const int64_t chunk = X.persist ? 4096 : X.nslots > 0 ? 1 : P->pt_every > 0 ? P->pt_every : total;
  int64_t bar_base = 0;
  for (int64_t k = 0; k < total;) {
    int64_t ns = std::min<int64_t>(chunk, total - k);
    if (!X.persist && X.nslots == 0 && P->pt_every > 0)  // launches end where swaps are due
      ns = std::min<int64_t>(ns, P->pt_every - (P->sweep0 + k) % P->pt_every);
    // split replicas in a persistent launch: launches end where a full
    // recompute is due, and the host recomputes every replica after them
    const bool rec_host = X.persist && X.nslots > 0 && B->recompute_every > 0;
    if (rec_host) ns = std::min<int64_t>(ns, B->recompute_every - (P->sweep0 + k) % B->recompute_every);
    SweepArgs a{k, P->sweep0, (int32_t)ns, (int32_t)nth, spf, P->frame0, X.sync, X.xpart, X.halo};
    if (X.persist) {
      a.bar = X.bar;
      a.bar_base = (int32_t)bar_base;
      a.pt_every = pt_on ? P->pt_every : 0;
      a.has_split = X.nslots > 0;
      a.nladder = (int32_t)X.hl_row0.size();
      a.lad_row0 = X.lad_row0;
      a.lad_n = X.lad_n;
      a.blk_lad = X.blk_lad;
      a.rep_np = X.rep_np;
      a.eslot = X.eslot;
      a.mslot = X.mslot;
      a.xbuf = X.xbuf;
      a.mprev3 = X.mprev3;
      a.fpart = X.fpart;
      a.fspin = X.fspin;
      a.fspin_n = X.fspin_n;
      a.max_export = std::max(X.max_export, 1);
      a.nslots = std::max(X.nslots, 1);
      a.rec_host = rec_host ? 1 : 0;
      for (int64_t q = k; q < k + ns; ++q)
        if (X.nslots > 0 || (pt_on && ((P->sweep0 + q + 1) % P->pt_every) == 0)) ++bar_base;
      if (bar_base * (int64_t)std::max<size_t>(X.hitems.size(), BAR_GROUPS) >= ((int64_t)1 << 31))
        return fail("too many barriers in one call (%lld); split the call", (long long)bar_base);
    }
    if (!X.hitems.empty()) {
      if (launch_lds<NC, ST>(B, X, a, s, T)) return -1;
      T.sweeps += (double)ns;
      T.updates += X.lds_updates_per_sweep * (double)ns;
    }
    if (stream_sweeps<NC, ST>(B, P, X, k, ns, nth, spf, plain, s, T)) return -1;
    if (rec_host && recompute_due(*B, P->sweep0 + k + ns - 1)) {
      const int mode = 1 | (k + ns == nth ? 2 : 0);  // recompute (+ energy reference at thermalisation end)
      if (launch_frame<NC, ST>(B, X.all_reps, nall, mode, -1, X.fsum, X.frame_np_all, X.frame_spins, s)) return -1;
    }
    k += ns;
    if (!X.persist && pt_on && ((P->sweep0 + k) % P->pt_every) == 0) {
      hipLaunchKernelGGL(k_pt, dim3((unsigned)X.hl_row0.size()), dim3(64), 0, s, *B, X.lad_row0, X.lad_n,
                         (int)X.hl_row0.size(), P->sweep0 + k - 1);
      HIPCHK(hipGetLastError());
    }
  }
  return T.finish(P->timing);
}

// mcg_run: prepare into a region of its own, run, wait, release
template <int NC, typename ST>
int run_typed(const mcg_batch* B, const mcg_plan* P, hipStream_t s) {
  RunCtx X;
  if (prepare_typed<NC, ST>(B, P, X)) return -1;
  void* dev = nullptr;
  if (X.dev_bytes && hipMalloc(&dev, X.dev_bytes) != hipSuccess) return fail("hipMalloc(%zu) failed", X.dev_bytes);
  int rc = bind_region(X, (unsigned char*)dev, s);
  if (!rc) rc = exec_typed<NC, ST>(B, P, s, X);
  // the region and the staged tables must outlive the launches
  const hipError_t e = hipStreamSynchronize(s);
  if (dev) (void)hipFree(dev);
  if (!rc && e != hipSuccess) rc = fail("hipStreamSynchronize: %s", hipGetErrorString(e));
  return rc;
}

// ---------------------------------------------------------------------------
// mcg_run_noalloc: mcg_run's launches with their device scratch in a
// file-scope array (allocated with the code object on every device that
// loads the library, held until the process ends), for callers that cannot
// allocate.  Host code never passes `scratch` itself to a kernel (in host
// code the name stands for a host-side shadow): each device's address is
// looked up once with hipGetSymbolAddress and kept in g_scratch_dev.
// ---------------------------------------------------------------------------
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)4 << 20;
static_assert(NOALLOC_SCRATCH_BYTES == (size_t)MCG_NOALLOC_SCRATCH_BYTES, "header and library disagree");
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[MAX_DEVICES];

// the current device's copy of `scratch` (looked up on its first use)
int scratch_region(unsigned char** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail("device %d outside 0..%d", dev, MAX_DEVICES - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_dev[dev] = (unsigned char*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

template <int NC, typename ST>
int run_noalloc_typed(const mcg_batch* B, const mcg_plan* P, hipStream_t s) {
  if (P->timing) return fail("mcg_run_noalloc does not time launches: plan->timing must be NULL");
  RunCtx X;
  if (prepare_typed<NC, ST>(B, P, X)) return -1;
  if (X.dev_bytes > NOALLOC_SCRATCH_BYTES)
    return fail("this run needs %zu bytes of device scratch; mcg_run_noalloc has %zu (use mcg_run or a runner)",
                X.dev_bytes, NOALLOC_SCRATCH_BYTES);
  unsigned char* dev = nullptr;
  if (scratch_region(&dev)) return -1;
  if (bind_region(X, dev, s)) return -1;
  // the tables are staged in X, which ends with this call: wait for their copy
  HIPCHK(hipStreamSynchronize(s));
  return exec_typed<NC, ST>(B, P, s, X);
}

template <int NC, typename ST>
int prepare_runner(const mcg_batch* B, const mcg_plan* P, hipStream_t s, RunCtx& X) {
  if (prepare_typed<NC, ST>(B, P, X)) return -1;
  void* dev = nullptr;
  if (X.dev_bytes && hipMalloc(&dev, X.dev_bytes) != hipSuccess) return fail("hipMalloc(%zu) failed", X.dev_bytes);
  X.dev = (unsigned char*)dev;
  if (bind_region(X, (unsigned char*)dev, s)) return -1;
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

template <int NC, typename ST>
int init_typed(const mcg_batch* B, hipStream_t s) {
  std::vector<int32_t> all(B->nrep);
  for (int r = 0; r < B->nrep; ++r) all[r] = r;
  DevBuf d;
  if (d.upload(all.data(), all.size() * 4, s)) return -1;
  if (launch_fill_f64(B->state, (size_t)B->nrep * MCG_NSTATE, 0.0, s)) return -1;
  if (launch_fill_f64(B->acc, (size_t)B->nrow * MCG_NACC, 0.0, s)) return -1;
  if (launch_fill_f64(B->bins, (size_t)B->nrow * B->nbins * MCG_NBINQ, 0.0, s)) return -1;
  if (B->gacc) {
    int ng = 0;
    for (int gi = 0; gi < B->ngraph; ++gi) ng = std::max(ng, B->graphs_host[gi].ngroup);
    if (launch_fill_f64(B->gacc, (size_t)B->nrow * (ng + 2) * (ng + 1), 0.0, s)) return -1;
  }
  if (launch_fill_i32(B->pt_acc, (size_t)B->nrow, 0, s)) return -1;
  if (launch_fill_i32(B->err, 1, 0, s)) return -1;
  if (B->wstamp) {
    size_t tot = 0;
    for (int r = 0; r < B->nrep; ++r)
      tot = std::max<size_t>(tot, (size_t)B->reps_host[r].site_off + B->graphs_host[B->reps_host[r].graph].n);
    const unsigned g = (unsigned)std::min<size_t>((tot + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, B->wstamp, tot, 0u);
    HIPCHK(hipGetLastError());
  }
  if (launch_measure<NC, ST>(B, (const int32_t*)d.p, B->nrep, 1 | 2 | 8, -1, s)) return -1;
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

template <int NC, typename ST>
int measure_typed(const mcg_batch* B, int64_t frame, hipStream_t s) {
  std::vector<int32_t> all(B->nrep);
  for (int r = 0; r < B->nrep; ++r) all[r] = r;
  DevBuf d;
  if (d.upload(all.data(), all.size() * 4, s)) return -1;
  if (launch_measure<NC, ST>(B, (const int32_t*)d.p, B->nrep, 1 | 4, frame, s)) return -1;
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

}  // namespace

#define DISPATCH(fn, ...)                                                                    \
  do {                                                                                       \
    if (B->nc == 1) return fn<1, int8_t>(__VA_ARGS__);                                       \
    if (B->nc == 2 && B->store == MCG_STORE_F32) return fn<2, float>(__VA_ARGS__);           \
    if (B->nc == 2 && B->store == MCG_STORE_F16) return fn<2, __half>(__VA_ARGS__);          \
    if (B->nc == 3 && B->store == MCG_STORE_F32) return fn<3, float>(__VA_ARGS__);           \
    if (B->nc == 3 && B->store == MCG_STORE_F16) return fn<3, __half>(__VA_ARGS__);          \
    return fail("unsupported nc/store combination");                                         \
  } while (0)

extern "C" int mcg_init(const mcg_batch* B, void* stream) {
  if (check_batch(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(init_typed, B, s);
}

extern "C" int mcg_run(const mcg_batch* B, const mcg_plan* P, void* stream) {
  if (check_batch(B)) return -1;
  if (!P) return fail("null plan");
  if (P->nthermal < 0 || P->nsweep < 0) return fail("negative sweep counts");
  if (P->algorithm != MCG_METROPOLIS && P->algorithm != MCG_WOLFF) return fail("unknown algorithm %d", P->algorithm);
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(run_typed, B, P, s);
}

// A prepared run (mcg_runner_*): the RunCtx and the region it owns.
struct mcg_runner {
  RunCtx X;
};

namespace {
template <int NC, typename ST>
int runner_create_typed(const mcg_batch* B, const mcg_plan* P, hipStream_t s, mcg_runner* R) {
  return prepare_runner<NC, ST>(B, P, s, R->X);
}
template <int NC, typename ST>
int runner_run_typed(const mcg_batch* B, const mcg_plan* P, hipStream_t s, mcg_runner* R) {
  if (check_key(B, P, R->X)) return -1;
  return exec_typed<NC, ST>(B, P, s, R->X);
}
void runner_free(mcg_runner* R) {
  if (!R) return;
  if (R->X.dev) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(R->X.device);
    (void)hipDeviceSynchronize();  // no launch of this run may still read the region
    (void)hipFree(R->X.dev);
    (void)hipSetDevice(cur);
  }
  delete R;
}
}  // namespace

extern "C" int mcg_runner_create(const mcg_batch* B, const mcg_plan* P, void* stream, mcg_runner** out) {
  if (!out) return fail("null runner out-pointer");
  *out = nullptr;
  if (check_batch(B)) return -1;
  if (!P) return fail("null plan");
  if (P->algorithm != MCG_METROPOLIS && P->algorithm != MCG_WOLFF) return fail("unknown algorithm %d", P->algorithm);
  mcg_runner* R = new (std::nothrow) mcg_runner();
  if (!R) return fail("out of host memory");
  hipStream_t s = (hipStream_t)stream;
  const int rc = [&]() -> int { DISPATCH(runner_create_typed, B, P, s, R); }();
  if (rc) {
    runner_free(R);
    return -1;
  }
  *out = R;
  return 0;
}

extern "C" int mcg_runner_run(mcg_runner* R, const mcg_batch* B, const mcg_plan* P, void* stream) {
  if (!R) return fail("null runner");
  if (check_batch(B)) return -1;
  if (!P) return fail("null plan");
  if (P->nthermal < 0 || P->nsweep < 0) return fail("negative sweep counts");
  if (P->algorithm != MCG_METROPOLIS && P->algorithm != MCG_WOLFF) return fail("unknown algorithm %d", P->algorithm);
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(runner_run_typed, B, P, s, R);
}

extern "C" int mcg_runner_destroy(mcg_runner* R) {
  runner_free(R);
  return 0;
}

extern "C" int mcg_run_noalloc(const mcg_batch* B, const mcg_plan* P, void* stream) {
  if (check_batch(B)) return -1;
  if (!P) return fail("null plan");
  if (P->nthermal < 0 || P->nsweep < 0) return fail("negative sweep counts");
  if (P->algorithm != MCG_METROPOLIS && P->algorithm != MCG_WOLFF) return fail("unknown algorithm %d", P->algorithm);
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(run_noalloc_typed, B, P, s);
}

extern "C" int mcg_sweeps(const mcg_batch* B, const mcg_plan* P, int64_t sweep0, int32_t nsweeps, void* stream) {
  if (check_batch(B)) return -1;
  if (!P) return fail("null plan");
  mcg_plan q = *P;
  q.nthermal = nsweeps;  // all sweeps unmeasured
  q.nsweep = 0;
  q.sweep0 = sweep0;
  q.pt_every = 0;
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(run_typed, B, &q, s);
}

extern "C" int mcg_measure(const mcg_batch* B, int64_t frame, void* stream) {
  if (check_batch(B)) return -1;
  hipStream_t s = (hipStream_t)stream;
  DISPATCH(measure_typed, B, frame, s);
}

// greedy colouring in site order: each site takes the smallest colour not
// used by an already-coloured neighbour (in- and out-links both count)
// DSatur (Brelaz 1979): colour next the uncoloured vertex with the most
// distinct colours among its neighbours (ties: most neighbours, then lowest
// index) with the lowest colour none of them has.  Returns the colour count.
static int dsatur_colour(int32_t n, const std::vector<std::vector<int32_t>>& adj, std::vector<int32_t>& col) {
  col.assign(n, -1);
  std::vector<std::vector<char>> seen(n);
  std::vector<int32_t> sat(n, 0);
  // max-heap on (saturation, degree, -index)
  typedef std::tuple<int32_t, int32_t, int32_t> Key;
  std::priority_queue<Key> q;
  for (int32_t i = 0; i < n; ++i) q.push(Key(0, (int32_t)adj[i].size(), -i));
  int ncol = 0;
  while (!q.empty()) {
    const Key k = q.top();
    q.pop();
    const int32_t i = -std::get<2>(k);
    if (col[i] >= 0 || std::get<0>(k) != sat[i]) continue;  // stale entry
    int c = 0;
    while (c < (int)seen[i].size() && seen[i][c]) ++c;
    col[i] = c;
    ncol = std::max(ncol, c + 1);
    for (int32_t j : adj[i]) {
      if (col[j] >= 0) continue;
      if ((int)seen[j].size() <= c) seen[j].resize(c + 1, 0);
      if (!seen[j][c]) {
        seen[j][c] = 1;
        ++sat[j];
        q.push(Key(sat[j], (int32_t)adj[j].size(), -j));
      }
    }
  }
  return ncol;
}

// Proper colouring of the site graph: greedy in site order (the checkerboard
// of square and cubic lattices), or DSatur where that needs fewer colours
// (the honeycomb with second neighbours: 4 colours of 2048 sites against
// greedy's 6 on 64x64 cells).  Fewer colours mean fewer synchronised rounds
// per sweep and, for the streaming pass, fewer launches and neighbour reads.
extern "C" int mcg_colour_graph(int32_t n, int32_t zmax, const int32_t* nbr, int32_t* colour) {
  if (n < 0 || zmax < 0 || (!nbr && n * zmax) || !colour) return fail("bad arguments to mcg_colour_graph");
  std::vector<std::vector<int32_t>> adj(n);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t l = 0; l < zmax; ++l) {
      const int32_t j = nbr[(size_t)i * zmax + l];
      if (j < 0 || j == i) continue;
      if (j >= n) return fail("neighbour index %d out of range", j);
      adj[i].push_back(j);
      adj[j].push_back(i);
    }
  std::vector<int32_t> mark;
  int ncol = 0;
  for (int32_t i = 0; i < n; ++i) colour[i] = -1;
  for (int32_t i = 0; i < n; ++i) {
    mark.assign(ncol + 1, 0);
    for (int32_t j : adj[i])
      if (colour[j] >= 0 && colour[j] <= ncol) mark[colour[j]] = 1;
    int c = 0;
    while (c <= ncol && mark[c]) ++c;
    colour[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  if (ncol > 2) {
    std::vector<int32_t> ds;
    const int nd = dsatur_colour(n, adj, ds);
    if (nd < ncol) {
      for (int32_t i = 0; i < n; ++i) colour[i] = ds[i];
      ncol = nd;
    }
  }
  return ncol;
}
