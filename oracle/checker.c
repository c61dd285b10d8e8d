/*
 * ORACLE — test infrastructure only (tests/, smoke(), bench cpu_baseline).
 *
 * CPU checker of the GPU engine: graph-coloured (checkerboard) Metropolis and
 * frontier Wolff with the engine's Philox4x32-10 counter scheme, in FP32 with
 * the same operation sequence, so that the GPU trajectory can be compared bit
 * for bit.  It reads the raw per-site tables (neighbours, 9-component
 * couplings in the reference order xx,yy,zz,xy,xz,yz,yx,zx,zy, spin lengths,
 * anisotropy) and a colour per site, and builds its own coupling view; it
 * shares no code with the product.
 *
 * Engine conventions restated here:
 *   key = (seed_lo, seed_hi ^ serial)
 *   Metropolis: colour c, positions q*G..q*G+G-1 of its site list (sites in
 *     site order) share the draw ctr = (q, k_lo, k_hi, c << 8), G = 1
 *     (Heisenberg: r_z = 2 u(x0) - 1, phi from x1, accept u(x2)),
 *     G = 2 (XY: site g takes phi from x_{2g}, accept u(x_{2g+1})),
 *     G = 4 (Ising: site g accepts with u(x_g))
 *   u(x) = (x >> 8) * 2^-24
 *   Wolff cluster c: ctr (0, c_lo, c_hi, 2): seed position = x0*n >> 32,
 *     direction from x1 (XY phi; Heisenberg r_z) and x2 (Heisenberg phi),
 *     accept x3>>8.  Positions are indices in colour order.
 *   Wolff bond (o,l): ctr (pos(o)*zmax + l, c_lo, c_hi, 3)
 *   tempering pair t of ladder L at swap s: ctr (t, s_lo, s_hi, 4),
 *     key (seed_lo, seed_hi ^ (0x80000000 | L))
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "philox_ref.h"

static float uu(uint32_t v) { return (float)(v >> 8) * 0x1p-24f; }

/* exp(x), x <= 0: Cody-Waite reduction, degree-7 Taylor, exact 2^k scaling */
static float cexp_neg(float x) {
  if (x < -80.0f) return 0.0f;
  float k = floorf(fmaf(x, 1.44269504088896341f, 0.5f));
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
  int ki = (int)k;
  uint32_t bits = (uint32_t)(ki + 127) << 23;
  float sc;
  memcpy(&sc, &bits, 4);
  return p * sc;
}

/* (cos, sin) of 2*pi*v/2^32 via octant symmetry and Taylor polynomials */
static void csincos(uint32_t v, float *c, float *s) {
  uint32_t o = v >> 29;
  float f = (float)((v >> 5) & 0xFFFFFFu) * 0x1p-24f;
  if (o & 1u) f = 1.0f - f;
  float a = f * 0.785398163397448310f;
  float a2 = a * a;
  float ps = 2.75573192239858907e-6f;
  ps = fmaf(ps, a2, -1.98412698412698413e-4f);
  ps = fmaf(ps, a2, 8.33333333333333333e-3f);
  ps = fmaf(ps, a2, -1.66666666666666667e-1f);
  float sn = fmaf(a, a2 * ps, a);
  float pc = -2.75573192239858907e-7f;
  pc = fmaf(pc, a2, 2.48015873015873016e-5f);
  pc = fmaf(pc, a2, -1.38888888888888889e-3f);
  pc = fmaf(pc, a2, 4.16666666666666667e-2f);
  pc = fmaf(pc, a2, -0.5f);
  float cs = fmaf(pc, a2, 1.0f);
  float cc[8] = {cs, sn, -sn, -cs, -cs, -sn, sn, cs};
  float ss[8] = {sn, cs, cs, sn, -sn, -cs, -cs, -sn};
  *c = cc[o];
  *s = ss[o];
}

/* IEEE binary16 round-to-nearest-even of a float, returned as float */
static float f16r(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  uint32_t sign = u & 0x80000000u;
  uint32_t ax = u & 0x7FFFFFFFu;
  float a;
  memcpy(&a, &ax, 4);
  if (a >= 65520.0f) { /* overflow (or inf/nan) */
    float r = isnan(x) ? x : (sign ? -INFINITY : INFINITY);
    return r;
  }
  if (a < 6.103515625e-05f) { /* subnormal half: quantum 2^-24 */
    float q = nearbyintf(a * 16777216.0f) / 16777216.0f;
    return sign ? -q : q;
  }
  /* normal: keep 10 mantissa bits, round to nearest even */
  uint32_t lsb = (ax >> 13) & 1u;
  uint32_t r = ax + 0x0FFFu + lsb;
  r &= ~0x1FFFu;
  float y;
  memcpy(&y, &r, 4);
  return sign ? -y : y;
}

typedef struct {
  int nc, store; /* store: 0 f32, 1 f16, 2 int8 */
  long n;
  int zmax;
  const int *nbr;
  float *cj;    /* [n*zmax*9] row-major coupling per link (Ising: [0] folded, [1] raw) */
  float *S, *D; /* [n], [n*3] */
  int uni;      /* every site has zmax links, all with the tensor cj[0..8] */
} cb_graph;

static void cb_build(cb_graph *g, int nc, int store, long n, int zmax, const int *nbr, const double *J,
                     const double *S, const double *D, int diag_only) {
  static const int rm[9] = {0, 3, 4, 6, 1, 5, 7, 8, 2}; /* rowmajor[k] = ref index */
  g->nc = nc;
  g->store = store;
  g->n = n;
  g->zmax = zmax;
  g->nbr = nbr;
  g->cj = calloc((size_t)n * zmax * 9, sizeof(float));
  g->S = malloc(sizeof(float) * n);
  g->D = malloc(sizeof(float) * n * 3);
  for (long i = 0; i < n; i++) {
    g->S[i] = (float)S[i];
    for (int k = 0; k < 3; k++) g->D[i * 3 + k] = (float)D[i * 3 + k];
    for (int l = 0; l < zmax; l++) {
      long t = nbr[i * zmax + l];
      float *c = g->cj + (i * zmax + l) * 9;
      if (t < 0) continue;
      const double *j = J + (i * zmax + l) * 9;
      if (nc == 1) {
        c[0] = (float)(j[0] * S[t]);
        c[1] = (float)j[0];
      } else {
        for (int k = 0; k < 9; k++) {
          int ri = rm[k];
          double v = j[ri];
          if (diag_only && ri >= 3) v = 0;
          if (nc == 2 && (ri == 2 || ri == 4 || ri == 5 || ri == 7 || ri == 8)) v = 0;
          c[k] = (float)v;
        }
      }
    }
  }
  /* uniform graph: the Metropolis field is J . (sum of neighbour spins) */
  g->uni = zmax > 0;
  for (long i = 0; i < n && g->uni; i++)
    for (int l = 0; l < zmax && g->uni; l++) {
      if (nbr[i * zmax + l] < 0) g->uni = 0;
      for (int k = 0; k < 9; k++)
        if (g->cj[(i * zmax + l) * 9 + k] != g->cj[k]) g->uni = 0;
    }
}

static void cb_free(cb_graph *g) {
  free(g->cj);
  free(g->S);
  free(g->D);
}

static float stor(const cb_graph *g, float v) { return g->store == 1 ? f16r(v) : v; }

/* The engine's exact Ising acceptance (mcg_batch.ithr): thresholds
 * [nrep][nthr] for the replicas of the following mcref_cb_sweeps calls, in
 * their order; a flip with dE = k / inv_unit > 0 is accepted when the uint32
 * acceptance word is below thr[k].  NULL: the FP32 exp(-beta dE) rule. */
static const uint32_t *cb_ithr;
static int cb_nthr;
static float cb_inv_unit;
static const uint32_t *cb_thr_row;
void mcref_cb_ising_table(const uint32_t *thr, int nthr, float inv_unit) {
  cb_ithr = thr;
  cb_nthr = nthr;
  cb_inv_unit = inv_unit;
}

/* spins: [nc][n] floats for one replica (Ising: sigma = +-1).
 * wa: r_z word (Heisenberg), wb: angle word, wc: acceptance word */
static void cb_site(const cb_graph *g, float *sp, long i, float beta, float h, uint32_t wa, uint32_t wb, uint32_t wc,
                    double *dE, double *dM) {
  long n = g->n;
  int nc = g->nc;
  const float S = g->S[i];
  if (nc == 1) {
    float f = 0.0f;
    if (g->uni) {
      float m = 0.0f;
      for (int l = 0; l < g->zmax; l++) m += sp[g->nbr[i * g->zmax + l]];
      f = g->cj[0] * m;
    } else {
      for (int l = 0; l < g->zmax; l++) {
        long t = g->nbr[i * g->zmax + l];
        if (t < 0) t = i;
        f = fmaf(g->cj[(i * g->zmax + l) * 9], sp[t], f);
      }
    }
    float sg = sp[i];
    float s = sg * S;
    float t = f - h;
    float de = (-2.0f * s) * t;
    int acc;
    if (cb_thr_row) {
      acc = de <= 0.0f || wc < cb_thr_row[(int)(de * cb_inv_unit)];
    } else {
      float db = de * beta;
      acc = db <= 0.0f || uu(wc) < cexp_neg(-db);
    }
    if (acc) {
      sp[i] = -sg;
      *dE += de;
      dM[0] += -2.0f * s;
    }
    return;
  }
  float s0 = sp[i], s1 = sp[n + i], s2 = nc == 3 ? sp[2 * n + i] : 0.0f;
  float H0 = 0, H1 = 0, H2 = 0;
  if (g->uni) {
    const float *J = g->cj;
    float S0 = 0, S1 = 0, S2 = 0;
    for (int l = 0; l < g->zmax; l++) {
      long t = g->nbr[i * g->zmax + l];
      S0 += sp[t];
      S1 += sp[n + t];
      if (nc == 3) S2 += sp[2 * n + t];
    }
    H0 = J[0] * S0; H0 = fmaf(J[1], S1, H0);
    H1 = J[3] * S0; H1 = fmaf(J[4], S1, H1);
    if (nc == 3) {
      H0 = fmaf(J[2], S2, H0);
      H1 = fmaf(J[5], S2, H1);
      H2 = J[6] * S0; H2 = fmaf(J[7], S1, H2); H2 = fmaf(J[8], S2, H2);
    }
  }
  for (int l = 0; l < g->zmax && !g->uni; l++) {
    long t = g->nbr[i * g->zmax + l];
    const float *J = g->cj + (i * g->zmax + l) * 9;
    if (t < 0) t = i;
    float t0 = sp[t], t1 = sp[n + t];
    if (nc == 3) {
      float t2 = sp[2 * n + t];
      H0 = fmaf(J[0], t0, H0); H0 = fmaf(J[1], t1, H0); H0 = fmaf(J[2], t2, H0);
      H1 = fmaf(J[3], t0, H1); H1 = fmaf(J[4], t1, H1); H1 = fmaf(J[5], t2, H1);
      H2 = fmaf(J[6], t0, H2); H2 = fmaf(J[7], t1, H2); H2 = fmaf(J[8], t2, H2);
    } else {
      H0 = fmaf(J[0], t0, H0); H0 = fmaf(J[1], t1, H0);
      H1 = fmaf(J[3], t0, H1); H1 = fmaf(J[4], t1, H1);
    }
  }
  float rx, ry, rz = 0.0f;
  if (nc == 3) {
    rz = fmaf(uu(wa), 2.0f, -1.0f);
    float rho = sqrtf(fmaxf(fmaf(-rz, rz, 1.0f), 0.0f));
    float c, sn;
    csincos(wb, &c, &sn);
    rx = rho * c;
    ry = rho * sn;
  } else {
    csincos(wb, &rx, &ry);
  }
  float p = s0 * rx;
  p = fmaf(s1, ry, p);
  if (nc == 3) p = fmaf(s2, rz, p);
  float a = -2.0f * p;
  float d0 = a * rx, d1 = a * ry, d2 = a * rz;
  float de = d0 * H0;
  de = fmaf(d1, H1, de);
  if (nc == 3) de = fmaf(d2, H2, de);
  float n0 = s0 + d0, n1 = s1 + d1, n2 = s2 + d2;
  const float *Dv = g->D + i * 3;
  {
    float a0 = n0 * n0, b0 = s0 * s0, q0 = a0 - b0;
    float a1 = n1 * n1, b1 = s1 * s1, q1 = a1 - b1;
    de = fmaf(Dv[0], q0, de);
    de = fmaf(Dv[1], q1, de);
    if (nc == 3) {
      float a2 = n2 * n2, b2 = s2 * s2, q2 = a2 - b2;
      de = fmaf(Dv[2], q2, de);
    }
  }
  de = fmaf(-h, nc == 3 ? d2 : d0, de);
  float db = de * beta;
  if (db <= 0.0f || uu(wc) < cexp_neg(-db)) {
    float nn = n0 * n0;
    nn = fmaf(n1, n1, nn);
    if (nc == 3) nn = fmaf(n2, n2, nn);
    /* one Newton step for 1/sqrt(nn/S^2) starting from 1 */
    float inv_s2 = 1.0f / (S * S);
    float f = fmaf(-0.5f, nn * inv_s2, 1.5f);
    n0 = n0 * f;
    n1 = n1 * f;
    n2 = n2 * f;
    sp[i] = stor(g, n0);
    sp[n + i] = stor(g, n1);
    if (nc == 3) sp[2 * n + i] = stor(g, n2);
    *dE += de;
    dM[0] += sp[i] - s0;
    dM[1] += sp[n + i] - s1;
    if (nc == 3) dM[2] += sp[2 * n + i] - s2;
  }
}

/*
 * Run nsweeps colour-ordered Metropolis sweeps (global sweep indices
 * sweep0 .. sweep0+nsweeps-1) on nrep replicas of one graph.
 * spins: [nrep][nc][n] in/out (f16 storage: values already representable).
 * dE/dM: per replica accumulated energy change (unscaled) and [3] moment change.
 */
int mcref_cb_sweeps(int nc, int store, long n, int zmax, const int *nbr, const double *J, const double *S,
                    const double *D, int diag_only, const int *colour, int ncol, int nrep, const float *beta,
                    const float *h, const uint32_t *serial, uint32_t seed_lo, uint32_t seed_hi, long sweep0,
                    long nsweeps, float *spins, double *dE, double *dM) {
  cb_graph g;
  cb_build(&g, nc, store, n, zmax, nbr, J, S, D, diag_only);
  /* sites by colour, in site order */
  long *order = malloc(sizeof(long) * n);
  long *ptr = calloc((size_t)ncol + 1, sizeof(long));
  for (long i = 0; i < n; i++) ptr[colour[i] + 1]++;
  for (int c = 0; c < ncol; c++) ptr[c + 1] += ptr[c];
  long *fill = malloc(sizeof(long) * (ncol + 1));
  memcpy(fill, ptr, sizeof(long) * (ncol + 1));
  for (long i = 0; i < n; i++) order[fill[colour[i]]++] = i;
  for (int r = 0; r < nrep; r++) {
    float *sp = spins + (size_t)r * nc * n;
    uint32_t k0 = seed_lo, k1 = seed_hi ^ serial[r];
    cb_thr_row = (nc == 1 && cb_ithr) ? cb_ithr + (size_t)r * cb_nthr : NULL;
    for (long s = 0; s < nsweeps; s++) {
      uint64_t k = (uint64_t)(sweep0 + s);
      for (int c = 0; c < ncol; c++) {
        /* one Philox draw per group of G consecutive colour-list positions:
           G = 4 (Ising), 2 (XY), 1 (Heisenberg) */
        long cnt = ptr[c + 1] - ptr[c];
        int G = nc == 1 ? 4 : nc == 2 ? 2 : 1;
        for (long q = 0; q * G < cnt; q++) {
          ref_u4 x = ref_philox((uint32_t)q, (uint32_t)k, (uint32_t)(k >> 32), (uint32_t)c << 8, k0, k1);
          for (int gg = 0; gg < G && q * G + gg < cnt; gg++) {
            long i = order[ptr[c] + q * G + gg];
            if (nc == 3)
              cb_site(&g, sp, i, beta[r], h[r], x.v[0], x.v[1], x.v[2], dE + r, dM + 3 * r);
            else if (nc == 2)
              cb_site(&g, sp, i, beta[r], h[r], 0u, x.v[2 * gg], x.v[2 * gg + 1], dE + r, dM + 3 * r);
            else
              cb_site(&g, sp, i, beta[r], h[r], 0u, 0u, x.v[gg], dE + r, dM + 3 * r);
          }
        }
      }
    }
  }
  cb_thr_row = NULL;
  free(order);
  free(ptr);
  free(fill);
  cb_free(&g);
  return 0;
}

/* FP64 full energy and moment of one replica from FP32 site energies */
int mcref_cb_energy(int nc, long n, int zmax, const int *nbr, const double *J, const double *S, const double *D,
                    int diag_only, float h, const float *sp, double *out4) {
  cb_graph g;
  cb_build(&g, nc, 0, n, zmax, nbr, J, S, D, diag_only);
  double E = 0, M[3] = {0, 0, 0};
  for (long i = 0; i < n; i++) {
    float s[3] = {0, 0, 0}, H[3] = {0, 0, 0};
    if (nc == 1) {
      float f = 0.0f;
      for (int l = 0; l < zmax; l++) {
        long t = nbr[i * zmax + l];
        if (t < 0) t = i;
        f = fmaf(g.cj[(i * zmax + l) * 9], sp[t], f);
      }
      float si = sp[i] * g.S[i];
      E += (double)fmaf(0.5f * si, f, -h * si);
      M[0] += si;
      continue;
    }
    for (int c = 0; c < nc; c++) s[c] = sp[c * n + i];
    for (int l = 0; l < zmax; l++) {
      long t = nbr[i * zmax + l];
      const float *Jm = g.cj + (i * zmax + l) * 9;
      if (t < 0) t = i;
      float tv[3] = {sp[t], sp[n + t], nc == 3 ? sp[2 * n + t] : 0.0f};
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) H[a] = fmaf(Jm[a * 3 + b], tv[b], H[a]);
    }
    float e = s[0] * H[0];
    e = fmaf(s[1], H[1], e);
    e = fmaf(s[2], H[2], e);
    e = 0.5f * e;
    e = fmaf(g.D[i * 3 + 0] * s[0], s[0], e);
    e = fmaf(g.D[i * 3 + 1] * s[1], s[1], e);
    if (nc == 3) e = fmaf(g.D[i * 3 + 2] * s[2], s[2], e);
    e = fmaf(-h, nc == 3 ? s[2] : s[0], e);
    E += e;
    for (int c = 0; c < 3; c++) M[c] += s[c];
  }
  out4[0] = E;
  out4[1] = M[0];
  out4[2] = M[1];
  out4[3] = M[2];
  cb_free(&g);
  return 0;
}

/* one tempering swap pass of one ladder (rows r0..r0+nr-1), in place */
int mcref_cb_pt(int nr, const float *beta, const double *E_of_rep, int *who, int *rung, uint32_t seed_lo,
                uint32_t seed_hi, int ladder, long swap_index, int *acc) {
  int parity = (int)(swap_index & 1);
  for (int t = parity; t + 1 < nr; t += 2) {
    int ra = who[t], rb = who[t + 1];
    double d = ((double)beta[t] - (double)beta[t + 1]) * (E_of_rep[ra] - E_of_rep[rb]);
    int a = d >= 0.0;
    if (!a) {
      ref_u4 x = ref_philox((uint32_t)t, (uint32_t)swap_index, (uint32_t)((uint64_t)swap_index >> 32), 4u, seed_lo,
                            seed_hi ^ (0x80000000u | (uint32_t)ladder));
      a = uu(x.v[0]) < cexp_neg((float)d);
    }
    if (a) {
      who[t] = rb;
      who[t + 1] = ra;
      rung[ra] = t + 1;
      rung[rb] = t;
      acc[t]++;
    }
  }
  return 0;
}

/*
 * Wolff clusters c0 .. c0+ncl-1 on one replica (engine convention).  The
 * cluster is the set reachable from the seed through activated links, which
 * does not depend on the order the frontier is processed in.
 */
int mcref_cb_wolff(int nc, int store, long n, int zmax, const int *nbr, const double *J, const double *S,
                   const double *D, int diag_only, const int *colour, int ncol, float beta, float h,
                   uint32_t serial, uint32_t seed_lo, uint32_t seed_hi, long c0, long ncl, float *sp, long *sizes) {
  cb_graph g;
  cb_build(&g, nc, store, n, zmax, nbr, J, S, D, diag_only);
  /* isotropic exchange and no single-ion term: with h = 0 the acceptance
     energy is zero in exact arithmetic and is not evaluated (as the engine) */
  int iso = 1;
  for (long b = 0; b < n && nc > 1; b++) {
    for (int l = 0; l < zmax; l++) {
      if (nbr[b * zmax + l] < 0) continue;
      const float *Jm = g.cj + (b * zmax + l) * 9;
      if (Jm[1] != 0.0f || Jm[3] != 0.0f || Jm[0] != Jm[4]) iso = 0;
      if (nc == 3 && (Jm[2] != 0.0f || Jm[5] != 0.0f || Jm[6] != 0.0f || Jm[7] != 0.0f || Jm[8] != Jm[0])) iso = 0;
    }
    for (int k = 0; k < nc; k++)
      if (g.D[b * 3 + k] != 0.0f) iso = 0;
  }
  char *in = malloc((size_t)n);
  long *q = malloc(sizeof(long) * n);
  /* the engine indexes sites by their position in colour order: the seed is
     drawn as a position and bond counters use the position of the site */
  long *order = malloc(sizeof(long) * n), *pos = malloc(sizeof(long) * n);
  {
    long *cnt = calloc((size_t)ncol + 1, sizeof(long));
    for (long i = 0; i < n; i++) cnt[colour[i] + 1]++;
    for (int c = 0; c < ncol; c++) cnt[c + 1] += cnt[c];
    for (long i = 0; i < n; i++) order[cnt[colour[i]]++] = i;
    for (long p = 0; p < n; p++) pos[order[p]] = p;
    free(cnt);
  }
  uint32_t k0 = seed_lo, k1 = seed_hi ^ serial;
  for (long c = 0; c < ncl; c++) {
    uint64_t cl = (uint64_t)(c0 + c);
    memset(in, 0, (size_t)n);
    ref_u4 x = ref_philox(0u, (uint32_t)cl, (uint32_t)(cl >> 32), 2u, k0, k1);
    long seed = order[(long)(((uint64_t)x.v[0] * (uint64_t)n) >> 32)];
    float r[3] = {0, 0, 0};
    if (nc == 3) {
      r[2] = fmaf(uu(x.v[1]), 2.0f, -1.0f);
      float rho = sqrtf(fmaxf(fmaf(-r[2], r[2], 1.0f), 0.0f));
      float cc, sn;
      csincos(x.v[2], &cc, &sn);
      r[0] = rho * cc;
      r[1] = rho * sn;
    } else if (nc == 2) {
      csincos(x.v[1], &r[0], &r[1]);
    }
    long head = 0, tail = 0;
    q[tail++] = seed;
    in[seed] = 1;
    while (head < tail) {
      long o = q[head++];
      float so[3] = {0, 0, 0};
      if (nc == 1)
        so[0] = sp[o] * g.S[o];
      else
        for (int k = 0; k < nc; k++) so[k] = sp[k * n + o];
      float sdo = -(so[0] * r[0] + so[1] * r[1] + so[2] * r[2]);
      for (int l = 0; l < zmax; l++) {
        long t = nbr[o * zmax + l];
        if (t < 0 || t == o || in[t]) continue;
        const float *Jm = g.cj + (o * zmax + l) * 9;
        float corr;
        if (nc == 1) {
          corr = (Jm[0] * sp[t]) * so[0] * beta;
          corr = 2.0f * corr;
        } else {
          float st[3] = {sp[t], sp[n + t], nc == 3 ? sp[2 * n + t] : 0.0f};
          float sdt = -(st[0] * r[0] + st[1] * r[1] + st[2] * r[2]);
          float rJr = r[0] * (Jm[0] * r[0] + Jm[1] * r[1] + Jm[2] * r[2]) +
                      r[1] * (Jm[3] * r[0] + Jm[4] * r[1] + Jm[5] * r[2]) +
                      r[2] * (Jm[6] * r[0] + Jm[7] * r[1] + Jm[8] * r[2]);
          corr = 2.0f * sdo * sdt * rJr * beta;
        }
        if (corr < 0.0f) {
          ref_u4 y = ref_philox((uint32_t)(pos[o] * zmax + l), (uint32_t)cl, (uint32_t)(cl >> 32), 3u, k0, k1);
          if (1.0f - cexp_neg(corr) > uu(y.v[0])) {
            in[t] = 1;
            q[tail++] = t;
          }
        }
      }
    }
    if (sizes) sizes[c] = tail;
    double d = 0.0;
    for (long m = 0; m < (iso && h == 0.0f ? 0 : tail); m++) {
      long b = q[m];
      if (nc == 1) {
        d += (double)(2.0f * h * (sp[b] * g.S[b]));
        continue;
      }
      float sb[3] = {sp[b], sp[n + b], nc == 3 ? sp[2 * n + b] : 0.0f};
      float sdb = -(sb[0] * r[0] + sb[1] * r[1] + sb[2] * r[2]);
      float pb[3] = {sb[0] + sdb * r[0], sb[1] + sdb * r[1], sb[2] + sdb * r[2]};
      for (int l = 0; l < zmax; l++) {
        long t = nbr[b * zmax + l];
        if (t < 0) continue;
        const float *Jm = g.cj + (b * zmax + l) * 9;
        float st[3] = {sp[t], sp[n + t], nc == 3 ? sp[2 * n + t] : 0.0f};
        float sdt = -(st[0] * r[0] + st[1] * r[1] + st[2] * r[2]);
        float pt[3] = {st[0] + sdt * r[0], st[1] + sdt * r[1], st[2] + sdt * r[2]};
        float src = sdb * (r[0] * (Jm[0] * pt[0] + Jm[1] * pt[1] + Jm[2] * pt[2]) +
                           r[1] * (Jm[3] * pt[0] + Jm[4] * pt[1] + Jm[5] * pt[2]) +
                           r[2] * (Jm[6] * pt[0] + Jm[7] * pt[1] + Jm[8] * pt[2]));
        d += (double)src;
        if (in[t])
          d += (double)(sdt * (pb[0] * (Jm[0] * r[0] + Jm[1] * r[1] + Jm[2] * r[2]) +
                               pb[1] * (Jm[3] * r[0] + Jm[4] * r[1] + Jm[5] * r[2]) +
                               pb[2] * (Jm[6] * r[0] + Jm[7] * r[1] + Jm[8] * r[2])));
        else
          d += (double)src;
      }
      const float *Dd = g.D + b * 3;
      float tr[3] = {sdb * r[0], sdb * r[1], sdb * r[2]};
      float o = 0.0f;
      for (int k = 0; k < nc; k++) {
        float qq = sb[k] + tr[k];
        o += Dd[k] * (qq * qq - sb[k] * sb[k]);
      }
      o -= h * tr[nc == 3 ? 2 : 0];
      d += (double)o;
    }
    float db = (float)(d * (double)beta);
    float u = (float)(x.v[3] >> 8) * 0x1p-24f;
    if (db <= 0.0f || u < cexp_neg(-db)) {
      for (long m = 0; m < tail; m++) {
        long b = q[m];
        if (nc == 1) {
          sp[b] = -sp[b];
          continue;
        }
        float sb[3] = {sp[b], sp[n + b], nc == 3 ? sp[2 * n + b] : 0.0f};
        float a2 = -2.0f * (sb[0] * r[0] + sb[1] * r[1] + sb[2] * r[2]);
        float n0 = sb[0] + a2 * r[0], n1 = sb[1] + a2 * r[1], n2 = sb[2] + a2 * r[2];
        float nn = n0 * n0 + n1 * n1 + n2 * n2;
        float fct = fabsf(g.S[b]) / sqrtf(nn);
        sp[b] = stor(&g, n0 * fct);
        sp[n + b] = stor(&g, n1 * fct);
        if (nc == 3) sp[2 * n + b] = stor(&g, n2 * fct);
      }
    }
  }
  free(in);
  free(q);
  free(order);
  free(pos);
  cb_free(&g);
  return 0;
}

/* test hooks for the primitives */
void mcref_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t *out) {
  ref_u4 x = ref_philox(c0, c1, c2, c3, k0, k1);
  memcpy(out, x.v, 16);
}
float mcref_exp_neg(float x) { return cexp_neg(x); }
void mcref_sincos(uint32_t v, float *cs) { csincos(v, cs, cs + 1); }
float mcref_f16_round(float x) { return f16r(x); }
