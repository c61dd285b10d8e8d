/*
 * ORACLE — test infrastructure only.  Nothing in the product links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only to verify results or to time the CPU baseline.
 *
 * FP64 CPU restatement of the reference Monte Carlo engines
 *   mcsolver/isingLib.c       (Ising,      MCMainFunction  isingLib.c:259-451)
 *   mcsolver/heisenbergLib.c  (Heisenberg, MCMainFunction  heisenbergLib.c:478-885)
 *   mcsolver/xyLib.c          (XY,         MCMainFunction  xyLib.c:413-809)
 * driven by glibc rand() in the same consumption order as the reference, so
 * that with the same srand() seed it reproduces the reference's outputs.
 *
 * Written from the reference's behaviour, not its text: per-site records are
 * flat arrays, the BFS is an index queue, vectors are double[3].  The FP64
 * expression order of every energy / observable term follows the reference
 * exactly (that is what makes replay bit-exact); the order of rand() calls
 * (site pick r1,r2; proposal rejection sampler; accept draw only when the
 * move is uphill; Wolff bond draw only for frustrated-by-sign bonds; cluster
 * acceptance draw only when uphill; majority tie draw) is the reference's.
 *
 * Deliberate departures, each documented where it happens:
 *   - XY site pick uses the unsigned two-draw form of isingLib.c:239-241;
 *     the reference's xyLib.c:324,384 overflows `int` (negative site index).
 *   - nthermal*ninterval uses 64-bit arithmetic (isingLib.c:352 uses int).
 *   - ninterval > n in the Ising totSpin initialisation is clamped to n
 *     (isingLib.c:350 reads past initSpin).
 *   - spinFrame > nsweep is rejected (reference divides by zero).
 * Reproduced quirks (they define the reference's numbers):
 *   - Ising per-site coupling pointer is linkStrength + i, not
 *     linkStrength + i*maxNLinking (isingLib.c:30): link l of site i uses
 *     flat entry i+l.
 *   - Ising Metropolis energy is reset to 0 after thermalisation
 *     (isingLib.c:359); Ising totSpin starts as the sum of the first
 *     ninterval spins (isingLib.c:350); Wolff skips the last site's flip in
 *     totSpin (isingLib.c:228).
 *   - O(n) Wolff single-ion correction uses the half-step s+trans, i.e. the
 *     projection, not the reflected spin (heisenbergLib.c:418 + 299-308).
 *   - PI = 3.1415926535 and atan (not atan2) in the solid angle
 *     (heisenbergLib.c:6, 114-127).
 *   - U4 = <M^2>^2/<M^4> (isingLib.c:433).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REF_PI 3.1415926535

/* The integer source of every draw.  Mode 0 (default): glibc rand(), the
 * reference's own generator (replay).  Mode 1: xoshiro256** (Blackman and
 * Vigna, "Scrambled linear pseudorandom number generators", 2018; state
 * seeded by SplitMix64 from the srand seed), its top 31 bits, i.e. values
 * on rand()'s range [0, RAND_MAX = 2^31 - 1], consumed in the same order.
 * glibc's rand() is an additive lagged-Fibonacci generator with lags 3 and
 * 31 (r_i = r_{i-3} + r_{i-31}, up to a carry), and one update consumes 3 to
 * 6 consecutive values; mode 1 removes that generator from the algorithm to
 * tell its effect on the sampled averages apart from the algorithm's. */
static int g_rng_mode = 0;
static uint64_t g_xs[4];

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

static uint64_t xoshiro256ss(void) {
  const uint64_t out = rotl64(g_xs[1] * 5, 7) * 9;
  const uint64_t t = g_xs[1] << 17;
  g_xs[2] ^= g_xs[0];
  g_xs[3] ^= g_xs[1];
  g_xs[1] ^= g_xs[2];
  g_xs[0] ^= g_xs[3];
  g_xs[2] ^= t;
  g_xs[3] = rotl64(g_xs[3], 45);
  return out;
}

static int ref_rand(void) { return g_rng_mode ? (int)(xoshiro256ss() >> 33) : rand(); }

static double urand(void) { return ref_rand() / (double)RAND_MAX; }

/* isingLib.c:239-241 / heisenbergLib.c:443-445: two draws, r1 first. */
static long pick_site(long n) {
  unsigned long long r1 = (unsigned long long)ref_rand();
  unsigned long long r2 = (unsigned long long)ref_rand();
  return (long)((r1 * RAND_MAX + r2) % (unsigned long long)n);
}

void mcref_srand(unsigned int seed) {
  srand(seed);
  uint64_t x = 0x5EED000000000000ull ^ seed;
  for (int k = 0; k < 4; ++k) g_xs[k] = splitmix64(&x);
}
void mcref_rng_mode(int mode) { g_rng_mode = mode ? 1 : 0; }
/* known-answer access to the xoshiro256** state (tests) */
void mcref_xoshiro_set(uint64_t s0, uint64_t s1, uint64_t s2, uint64_t s3) {
  g_xs[0] = s0; g_xs[1] = s1; g_xs[2] = s2; g_xs[3] = s3;
}
uint64_t mcref_xoshiro_next(void) { return xoshiro256ss(); }
int mcref_rand(void) { return ref_rand(); }

/* ========================================================================== */
/* Ising                                                                       */
/* ========================================================================== */

typedef struct {
  long n, maxl;
  double *s;           /* spins (+-|S|) */
  double h;            /* beta*h */
  const int *nlink;
  const double *J;     /* flat linkStrength; link l of site i = J[i + l] */
  const int *nbr;      /* flat linkedOrb [n*maxl] */
  /* renormalised (block-spin) lattice */
  int *chosen;         /* per site: index into rOrb list + 1, or 0 */
  long nclu;
  const int *rclu;     /* [nr*nclu] */
  const int *rlink;    /* [nr*maxl] */
  /* Wolff scratch */
  char *inblock;
  long *queue;
} ising_t;

/* isingLib.c:121-127 */
static double ising_corr(const ising_t *m, long i) {
  double c = 0;
  for (int l = 0; l < m->nlink[i]; l++)
    c += m->J[i + l] * m->s[i] * m->s[m->nbr[i * m->maxl + l]];
  return c;
}

/* isingLib.c:238-254 */
static void ising_local(ising_t *m, double *E, double *tot) {
  long i = pick_site(m->n);
  double corr = 2 * (ising_corr(m, i) - m->h * m->s[i]);
  if (corr >= 0 || exp(corr) > urand()) {
    m->s[i] *= -1;
    *tot += m->s[i] * 2;
    *E -= corr;
  }
}

/* isingLib.c:165-236 */
static void ising_wolff(ising_t *m, double *E, double *tot) {
  long n = m->n;
  memset(m->inblock, 0, (size_t)n);
  long seed = pick_site(n);
  long head = 0, tail = 0;
  m->queue[tail++] = seed;
  m->inblock[seed] = 1;
  while (head < tail) {
    long o = m->queue[head++];
    for (int l = 0; l < m->nlink[o]; l++) {
      long t = m->nbr[o * m->maxl + l];
      if (m->inblock[t]) continue;
      double corr = m->J[o + l] * m->s[o] * m->s[t];
      if (corr < 0 && (1 - exp(2 * corr)) > urand()) {
        m->inblock[t] = 1;
        m->queue[tail++] = t;
      }
    }
  }
  double dh = 0;
  for (long k = 0; k < tail; k++) dh += 2 * m->h * m->s[m->queue[k]];
  if (dh <= 0 || exp(-dh) > urand()) {
    for (long k = 0; k < tail; k++) {
      long i = m->queue[k];
      m->s[i] *= -1;
      if (i < n - 1) *tot += m->s[i] * 2;
    }
    *E = 0.;
    for (long i = 0; i < n; i++) *E += ising_corr(m, i) / 2 - m->h * m->s[i];
  }
}

/* isingLib.c:133-150 */
static double ising_majority(const ising_t *m, long site) {
  long ri = m->chosen[site] - 1;
  double avg = 0.0;
  for (long k = 0; k < m->nclu; k++) avg += m->s[m->rclu[ri * m->nclu + k]];
  if (avg > 0) return fabs(m->s[site]);
  if (avg < 0) return -fabs(m->s[site]);
  return urand() > 0.5 ? fabs(m->s[site]) : -fabs(m->s[site]);
}

/* isingLib.c:152-163 */
static double ising_corr_rnorm(const ising_t *m, long site) {
  long ri = m->chosen[site] - 1;
  double c = 0;
  double ms = ising_majority(m, site);
  for (int l = 0; l < m->nlink[site]; l++) {
    /* a -1 padded block-spin link (small or odd L) is undefined behaviour in
       the reference; here, as in the engine, an absent link adds nothing */
    const int t = m->rlink[ri * m->maxl + l];
    if (t < 0) continue;
    double mt = ising_majority(m, t);
    c += m->J[site + l] * ms * mt;
  }
  return c;
}

/*
 * Restates isinglib.MCMainFunction (isingLib.c:259-451).
 * out[10] = spin_i, spin_j, spin_ij, autoCorr, E, E2, E_rnorm, E2_rnorm, U4,
 *           spin_tot  (energies in the caller's units, i.e. beta-scaled).
 * frames: NULL or [spinFrame * n].  Returns 0, or <0 on invalid input.
 */
int mcref_ising_run(int algorithm, long n, const double *initSpin, long nthermal,
                    long nsweep, long ninterval, int maxNLinking, const int *nlink,
                    const double *linkStrength, const int *linkedOrb, long nLat,
                    const int *corrOrbPair, double h, long nr, long nOrbInCluster,
                    const int *rOrb, const int *rOrbCluster, const int *linkedOrb_rnorm,
                    int spinFrame, double *out, double *frames) {
  if (n <= 0 || nsweep <= 0 || ninterval <= 0 || nLat <= 0) return -1;
  if (spinFrame > nsweep) return -2;
  ising_t m;
  m.n = n;
  m.maxl = maxNLinking;
  m.s = (double *)malloc(sizeof(double) * n);
  m.h = h;
  m.nlink = nlink;
  m.J = linkStrength;
  m.nbr = linkedOrb;
  m.chosen = (int *)calloc((size_t)n, sizeof(int));
  m.nclu = nOrbInCluster;
  m.rclu = rOrbCluster;
  m.rlink = linkedOrb_rnorm;
  m.inblock = (char *)malloc((size_t)n);
  m.queue = (long *)malloc(sizeof(long) * n);
  for (long i = 0; i < n; i++) m.s[i] = initSpin[i];
  for (long k = 0; k < nr; k++) m.chosen[rOrb[k]] = (int)(k + 1);

  double E = 0, tot = 0, Er = 0;
  long ninit = ninterval < n ? ninterval : n;
  for (long i = 0; i < ninit; i++) tot += initSpin[i];
  void (*upd)(ising_t *, double *, double *) = algorithm == 1 ? ising_wolff : ising_local;
  long long nth = (long long)nthermal * (long long)ninterval;
  for (long long i = 0; i < nth; i++) upd(&m, &E, &tot);

  double si = 0, sj = 0, sij = 0, totE = 0, totEr = 0, E2 = 0, E2r = 0;
  double M = 0, M2 = 0, M4 = 0, Mprev = 0, MM = 0, Mtot = 0, stot = 0;
  E = 0;
  long ops = spinFrame > 0 ? nsweep / spinFrame : nsweep;
  int iframe = 0;
  for (long t = 0; t < nsweep; t++) {
    for (long k = 0; k < ninterval; k++) upd(&m, &E, &tot);
    if (spinFrame > 0 && t % ops == 0) {
      if (frames && iframe < spinFrame) memcpy(frames + (size_t)iframe * n, m.s, sizeof(double) * n);
      iframe++;
    }
    Er = 0;
    for (long j = 0; j < n; j++)
      if (m.chosen[j]) Er += ising_corr_rnorm(&m, j) / 2 - m.h * m.s[j];
    double ai = 0, aj = 0, c = 0;
    for (long p = 0; p < nLat; p++) {
      double a = m.s[corrOrbPair[2 * p]], b = m.s[corrOrbPair[2 * p + 1]];
      ai += a;
      aj += b;
      c += a * b;
    }
    stot += tot;
    M = ai / nLat;
    M2 += M * M;
    M4 += M * M * M * M;
    Mtot += M;
    MM += Mprev * M;
    Mprev = M;
    si += fabs(ai) / nLat;
    sj += fabs(aj) / nLat;
    sij += c / nLat;
    double e = E / n, er = Er / nr;
    totE += e;
    totEr += er;
    E2 += e * e;
    E2r += er * er;
  }
  out[0] = si / nsweep;
  out[1] = sj / nsweep;
  out[2] = sij / nsweep;
  out[3] = MM / nsweep - (Mtot / nsweep) * (Mtot / nsweep);
  out[4] = totE / nsweep;
  out[5] = E2 / nsweep;
  out[6] = totEr / nsweep;
  out[7] = E2r / nsweep;
  out[8] = (M2 / nsweep) * (M2 / nsweep) / (M4 / nsweep);
  out[9] = stot / nsweep / nLat;
  free(m.s);
  free(m.chosen);
  free(m.inblock);
  free(m.queue);
  return 0;
}

/* ========================================================================== */
/* O(n): XY (nc = 2) and Heisenberg (nc = 3)                                   */
/* ========================================================================== */

typedef struct {
  int nc;              /* 2 or 3 */
  int diag_only;       /* ignoreNonDiagonalJ */
  long n, maxl;
  double (*s)[3];      /* spin */
  double (*tr)[3];     /* transSpin */
  double (*pp)[3];     /* perpenSpin */
  double *sdn;         /* sDotN */
  double *S;           /* initSpin (signed length) */
  double *len;         /* |initSpin| */
  const double *D;     /* [n*3] */
  double h;
  const int *nlink;
  const double *J;     /* [n*maxl*9] order xx,yy,zz,xy,xz,yz,yx,zx,zy */
  const int *nbr;
  int *chosen;
  long nclu;
  const int *rclu, *rlink;
  char *inblock, *proj;
  long *queue;
} on_t;

/* heisenbergLib.c:77-93 (nc=3), xyLib.c:59-68 (nc=2: slots xx,yy,xy,yx) */
static double dot9(const on_t *m, const double *a, const double *b, const double *J) {
  if (m->nc == 3) {
    if (m->diag_only) return a[0] * b[0] * J[0] + a[1] * b[1] * J[1] + a[2] * b[2] * J[2];
    return a[0] * b[0] * J[0] + a[1] * b[1] * J[1] + a[2] * b[2] * J[2] + a[0] * b[1] * J[3] +
           a[0] * b[2] * J[4] + a[1] * b[2] * J[5] + a[1] * b[0] * J[6] + a[2] * b[0] * J[7] +
           a[2] * b[1] * J[8];
  }
  if (m->diag_only) return a[0] * b[0] * J[0] + a[1] * b[1] * J[1];
  return a[0] * b[0] * J[0] + a[1] * b[1] * J[1] + a[0] * b[1] * J[3] + a[1] * b[0] * J[6];
}

static double vdot(int nc, const double *a, const double *b) {
  return nc == 3 ? a[0] * b[0] + a[1] * b[1] + a[2] * b[2] : a[0] * b[0] + a[1] * b[1];
}

/* heisenbergLib.c:19-25 */
static void vnormalize(int nc, double *v) {
  double l = nc == 3 ? sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) : sqrt(v[0] * v[0] + v[1] * v[1]);
  if (l < 1e-5) return;
  for (int k = 0; k < nc; k++) v[k] /= l;
}

/* heisenbergLib.c:95-112, xyLib.c:73-88: rejection in the ball / disc */
static void random_dir(int nc, double *r) {
  for (;;) {
    double x = urand() - 0.5;
    double y = urand() - 0.5;
    double z = 0, l2;
    if (nc == 3) {
      z = urand() - 0.5;
      l2 = x * x + y * y + z * z;
    } else {
      l2 = x * x + y * y;
    }
    if (l2 > 0.25) continue;
    double l = sqrt(l2);
    r[0] = x / l;
    r[1] = y / l;
    r[2] = nc == 3 ? z / l : 0;
    return;
  }
}

/* heisenbergLib.c:238-247 */
static double on_corr(const on_t *m, long i) {
  double c = 0;
  for (int l = 0; l < m->nlink[i]; l++)
    c += dot9(m, m->s[i], m->s[m->nbr[i * m->maxl + l]], m->J + (i * m->maxl + l) * 9);
  return c;
}

/* heisenbergLib.c:249-253, xyLib.c:200-204 (field on z / on x) */
static double on_onsite(const on_t *m, long i) {
  const double *s = m->s[i], *D = m->D + 3 * i;
  if (m->nc == 3) return D[0] * s[0] * s[0] + D[1] * s[1] * s[1] + D[2] * s[2] * s[2] - m->h * s[2];
  return D[0] * s[0] * s[0] + D[1] * s[1] * s[1] - m->h * s[0];
}

/* heisenbergLib.c:299-308, xyLib.c:248-254 */
static double on_delta_onsite(const on_t *m, long i) {
  const double *s = m->s[i], *t = m->tr[i], *D = m->D + 3 * i;
  double a = s[0] + t[0], b = s[1] + t[1];
  if (m->nc == 3) {
    double c = s[2] + t[2];
    return D[0] * (a * a - s[0] * s[0]) + D[1] * (b * b - s[1] * s[1]) + D[2] * (c * c - s[2] * s[2]) -
           m->h * t[2];
  }
  return D[0] * (a * a - s[0] * s[0]) + D[1] * (b * b - s[1] * s[1]) - m->h * t[0];
}

static double on_energy(const on_t *m) {
  double E = 0;
  for (long i = 0; i < m->n; i++) E += on_corr(m, i);
  E /= 2;
  for (long i = 0; i < m->n; i++) E += on_onsite(m, i);
  return E;
}

/* heisenbergLib.c:441-473, xyLib.c:382-409 */
static void on_local(on_t *m, double *E, double *tot) {
  int nc = m->nc;
  long i = pick_site(m->n);
  double r[3];
  random_dir(nc, r);
  double s1n = -2 * vdot(nc, m->s[i], r);
  for (int k = 0; k < nc; k++) m->tr[i][k] = r[k] * s1n;
  double corr = 0;
  for (int l = 0; l < m->nlink[i]; l++)
    corr += dot9(m, m->tr[i], m->s[m->nbr[i * m->maxl + l]], m->J + (i * m->maxl + l) * 9);
  corr += on_delta_onsite(m, i);
  if (corr <= 0 || exp(-corr) > urand()) {
    for (int k = 0; k < nc; k++) tot[k] += m->tr[i][k];
    for (int k = 0; k < nc; k++) m->s[i][k] += m->tr[i][k];
    *E += corr;
  }
}

static void on_project(on_t *m, long i, const double *r) {
  m->sdn[i] = -vdot(m->nc, m->s[i], r);
  for (int k = 0; k < m->nc; k++) m->tr[i][k] = r[k] * m->sdn[i];
  for (int k = 0; k < m->nc; k++) m->pp[i][k] = m->s[i][k] + m->tr[i][k];
  m->proj[i] = 1;
}

/* heisenbergLib.c:310-439, xyLib.c:256-380 */
static void on_wolff(on_t *m, double *E, double *tot) {
  int nc = m->nc;
  long n = m->n;
  memset(m->inblock, 0, (size_t)n);
  memset(m->proj, 0, (size_t)n);
  long seed = pick_site(n);
  long head = 0, tail = 0;
  m->queue[tail++] = seed;
  m->inblock[seed] = 1;
  double r[3];
  random_dir(nc, r);
  while (head < tail) {
    long o = m->queue[head++];
    if (!m->proj[o]) on_project(m, o, r);
    for (int l = 0; l < m->nlink[o]; l++) {
      long t = m->nbr[o * m->maxl + l];
      if (m->inblock[t]) continue;
      if (!m->proj[t]) on_project(m, t, r);
      double corr = 2 * m->sdn[o] * m->sdn[t] * dot9(m, r, r, m->J + (o * m->maxl + l) * 9);
      if (corr < 0 && (1 - exp(corr)) > urand()) {
        m->inblock[t] = 1;
        m->queue[tail++] = t;
      }
    }
  }
  double d = 0;
  for (long k = 0; k < tail; k++) {
    long b = m->queue[k];
    for (int l = 0; l < m->nlink[b]; l++) {
      long t = m->nbr[b * m->maxl + l];
      const double *J = m->J + (b * m->maxl + l) * 9;
      double src = m->sdn[b] * dot9(m, r, m->pp[t], J);
      d += src;
      if (m->inblock[t])
        d += m->sdn[t] * dot9(m, m->pp[b], r, J);
      else
        d += src;
    }
  }
  for (long k = 0; k < tail; k++) d += on_delta_onsite(m, m->queue[k]);
  if (d <= 0 || exp(-d) > urand()) {
    for (long k = 0; k < tail; k++) {
      long b = m->queue[k];
      for (int c = 0; c < nc; c++) m->tr[b][c] *= 2;
      for (int c = 0; c < nc; c++) m->s[b][c] += m->tr[b][c];
      for (int c = 0; c < nc; c++) tot[c] += m->tr[b][c];
    }
    *E = on_energy(m);
  }
}

/* heisenbergLib.c:255-267 */
static void on_majority(const on_t *m, long site, double *v) {
  long ri = m->chosen[site] - 1;
  v[0] = v[1] = v[2] = 0;
  for (long k = 0; k < m->nclu; k++) {
    const double *s = m->s[m->rclu[ri * m->nclu + k]];
    for (int c = 0; c < m->nc; c++) v[c] += s[c];
  }
  vnormalize(m->nc, v);
  for (int c = 0; c < m->nc; c++) v[c] *= m->S[site];
}

/* heisenbergLib.c:114-127 */
static double signed_area(const on_t *m, long a, long b, long c) {
  const double *s1 = m->s[a], *s2 = m->s[b], *s3 = m->s[c];
  double l1 = m->len[a], l2 = m->len[b], l3 = m->len[c];
  double s1s2 = vdot(3, s1, s2) / l1 / l2;
  double s2s3 = vdot(3, s2, s3) / l2 / l3;
  double s3s1 = vdot(3, s3, s1) / l3 / l1;
  double x[3] = {s2[1] * s3[2] - s2[2] * s3[1], s2[2] * s3[0] - s2[0] * s3[2], s2[0] * s3[1] - s2[1] * s3[0]};
  double re = 1 + s1s2 + s2s3 + s3s1;
  double im = vdot(3, s1, x) / l1 / l2 / l3;
  if (fabs(re) < 1e-6) return im > 0 ? REF_PI : -REF_PI;
  return 2 * atan(im / re);
}

/*
 * Restates {xylib,heisenberglib}.MCMainFunction (xyLib.c:413-809,
 * heisenbergLib.c:478-885).
 * out[27] = tuple slots 0..26 of the reference's 29-tuple (spin_i xyz,
 *   spin_j xyz, spin_ij, autoCorr, E, E2, U4, spin_i_r xyz, spin_j_r xyz,
 *   spin_ij_r, E_r, E2_r, spin_i_z, spin_j_z, spin_tot_z, spin_i_h, spin_j_h,
 *   spin_tot_h, Q).
 * frames: NULL or [spinFrame*n*3]; groups: NULL or [(nG+2)*(nG+1)].
 */
int mcref_on_run(int nc, int algorithm, long n, const double *initSpin, const double *initD,
                 long nthermal, long nsweep, long ninterval, int maxNLinking, const int *nlink,
                 const double *linkStrength, const int *linkedOrb, long nCirc, const int *circuits,
                 long nLat, const int *corrOrbPair, long nOrbGroup, long maxOrbGroupSize,
                 const int *orbGroupList, double flunc, double h, long nr, long nOrbInCluster,
                 const int *rOrb, const int *rOrbCluster, const int *linkedOrb_rnorm, int spinFrame,
                 int ignoreNonDiagonalJ, double *out, double *frames, double *groups) {
  if (nc != 2 && nc != 3) return -1;
  if (n <= 0 || nsweep <= 0 || ninterval <= 0 || nLat <= 0) return -1;
  if (spinFrame > nsweep) return -2;
  on_t m;
  memset(&m, 0, sizeof m);
  m.nc = nc;
  m.diag_only = ignoreNonDiagonalJ > 0;
  m.n = n;
  m.maxl = maxNLinking;
  m.s = calloc((size_t)n, sizeof(double[3]));
  m.tr = calloc((size_t)n, sizeof(double[3]));
  m.pp = calloc((size_t)n, sizeof(double[3]));
  m.sdn = calloc((size_t)n, sizeof(double));
  m.S = malloc(sizeof(double) * n);
  m.len = malloc(sizeof(double) * n);
  m.D = initD;
  m.h = h;
  m.nlink = nlink;
  m.J = linkStrength;
  m.nbr = linkedOrb;
  m.chosen = calloc((size_t)n, sizeof(int));
  m.nclu = nOrbInCluster;
  m.rclu = rOrbCluster;
  m.rlink = linkedOrb_rnorm;
  m.inblock = malloc((size_t)n);
  m.proj = malloc((size_t)n);
  m.queue = malloc(sizeof(long) * n);

  /* heisenbergLib.c:157-172: start along x with |S|, flunc tilt (consumes rand) */
  for (long i = 0; i < n; i++) {
    m.S[i] = initSpin[i];
    m.len[i] = fabs(initSpin[i]);
    m.s[i][0] = initSpin[i];
    m.s[i][1] = 0;
    m.s[i][2] = 0;
    double r[3];
    random_dir(nc, r);
    for (int c = 0; c < nc; c++) r[c] *= flunc;
    for (int c = 0; c < nc; c++) m.s[i][c] += r[c];
    vnormalize(nc, m.s[i]);
    for (int c = 0; c < nc; c++) m.s[i][c] *= fabs(initSpin[i]);
  }
  for (long k = 0; k < nr; k++) m.chosen[rOrb[k]] = (int)(k + 1);

  double E = on_energy(&m);
  double tot[3] = {0, 0, 0};
  for (long i = 0; i < n; i++)
    for (int c = 0; c < nc; c++) tot[c] += m.s[i][c];
  void (*upd)(on_t *, double *, double *) = algorithm == 1 ? on_wolff : on_local;
  long long nth = (long long)nthermal * (long long)ninterval;
  for (long long i = 0; i < nth; i++) upd(&m, &E, tot);

  double si[3] = {0, 0, 0}, sj[3] = {0, 0, 0}, sir[3] = {0, 0, 0}, sjr[3] = {0, 0, 0};
  double sij = 0, sijr = 0, totE = 0, totEr = 0, E2 = 0, E2r = 0;
  double M, M2 = 0, M4 = 0, Mprev = 0, MM = 0, Mtot = 0, Q = 0;
  double siz = 0, sjz = 0, stz = 0, sih = 0, sjh = 0, sth = 0;
  long ng = nOrbGroup;
  double *SS = calloc((size_t)((ng + 1) * (ng + 1)), sizeof(double));
  double *S4 = calloc((size_t)(ng + 1), sizeof(double));
  double (*G)[3] = calloc((size_t)(ng + 1), sizeof(double[3]));
  int hax = nc == 3 ? 2 : 0; /* field axis: z (Heisenberg) / x (XY) */
  long ops = spinFrame > 0 ? nsweep / spinFrame : nsweep;
  int iframe = 0;

  for (long t = 0; t < nsweep; t++) {
    for (long k = 0; k < ninterval; k++) upd(&m, &E, tot);
    if (spinFrame > 0 && t % ops == 0) {
      if (frames && iframe < spinFrame)
        for (long i = 0; i < n; i++)
          for (int c = 0; c < 3; c++) frames[((size_t)iframe * n + i) * 3 + c] = c < nc ? m.s[i][c] : 0.0;
      iframe++;
    }
    double dir[3] = {tot[0], tot[1], nc == 3 ? tot[2] : 0};
    vnormalize(nc, dir);
    double ai[3] = {0, 0, 0}, aj[3] = {0, 0, 0}, c = 0, zi = 0, zj = 0, hi = 0, hj = 0;
    for (long p = 0; p < nLat; p++) {
      const double *a = m.s[corrOrbPair[2 * p]], *b = m.s[corrOrbPair[2 * p + 1]];
      for (int k = 0; k < nc; k++) ai[k] += a[k];
      for (int k = 0; k < nc; k++) aj[k] += b[k];
      c += vdot(nc, a, b);
      zi += vdot(nc, dir, a);
      zj += vdot(nc, dir, b);
      hi += a[hax];
      hj += b[hax];
    }
    if (nc == 3) {
      double q = 0;
      for (long k = 0; k < nCirc; k++) q += signed_area(&m, circuits[3 * k], circuits[3 * k + 1], circuits[3 * k + 2]);
      Q += q / REF_PI / 4;
    }
    siz += zi / nLat;
    sjz += zj / nLat;
    stz += vdot(nc, dir, tot) / nLat;
    sih += hi / nLat;
    sjh += hj / nLat;
    sth += tot[hax] / nLat;
    M = nc == 3 ? sqrt(vdot(3, tot, tot)) / nLat : sqrt(vdot(2, ai, ai)) / nLat;
    M2 += M * M;
    M4 += M * M * M * M;
    Mtot += M;
    MM += Mprev * M;
    Mprev = M;
    for (int k = 0; k < nc; k++) {
      ai[k] /= nLat;
      aj[k] /= nLat;
      si[k] += fabs(ai[k]);
      sj[k] += fabs(aj[k]);
    }
    sij += c / nLat;
    double e = E / n;
    totE += e;
    E2 += e * e;

    /* block-spin observables, heisenbergLib.c:746-803 */
    double ari[3] = {0, 0, 0}, arj[3] = {0, 0, 0}, cr = 0;
    int ci = 0, cj = 0, cij = 0;
    for (long p = 0; p < nLat; p++) {
      long a = corrOrbPair[2 * p], b = corrOrbPair[2 * p + 1];
      double v[3], w[3];
      if (m.chosen[a]) {
        ci++;
        on_majority(&m, a, v);
        for (int k = 0; k < nc; k++) ari[k] += v[k];
      }
      if (m.chosen[b]) {
        cj++;
        on_majority(&m, b, v);
        for (int k = 0; k < nc; k++) arj[k] += v[k];
      }
      if (m.chosen[a] && m.chosen[b]) {
        cij++;
        on_majority(&m, a, v);
        on_majority(&m, b, w);
        cr += vdot(nc, v, w);
      }
    }
    for (int k = 0; k < nc; k++) {
      ari[k] /= ci;
      arj[k] /= cj;
      sir[k] += fabs(ari[k]);
      sjr[k] += fabs(arj[k]);
    }
    sijr += cr / cij;
    double er = 0;
    for (long j = 0; j < n; j++) {
      if (!m.chosen[j]) continue;
      long ri = m.chosen[j] - 1;
      double v[3], w[3], cc = 0;
      on_majority(&m, j, v);
      for (int l = 0; l < m.nlink[j]; l++) {
        const int t = m.rlink[ri * m.maxl + l];
        if (t < 0) continue;  /* absent block-spin link: adds nothing (UB in the reference) */
        on_majority(&m, t, w);
        cc += dot9(&m, v, w, m.J + (j * m.maxl + l) * 9);
      }
      er += cc;
    }
    er /= 2;
    for (long j = 0; j < n; j++)
      if (m.chosen[j]) er += on_onsite(&m, j);
    er /= nr;
    totEr += er;
    E2r += er * er;

    /* orbital groups, heisenbergLib.c:805-830 */
    long g;
    for (g = 0; g < ng; g++) {
      G[g][0] = G[g][1] = G[g][2] = 0;
      for (long k = 0; k < maxOrbGroupSize; k++) {
        int o = orbGroupList[g * maxOrbGroupSize + k];
        if (o < 0) break;
        for (int cc = 0; cc < nc; cc++) G[g][cc] += m.s[o][cc];
      }
    }
    for (int cc = 0; cc < nc; cc++) G[g][cc] = tot[cc] / nLat;
    for (long a = 0; a < ng + 1; a++)
      for (long b = 0; b < ng + 1; b++) {
        double d = vdot(nc, G[a], G[b]);
        SS[a * (ng + 1) + b] += d;
        if (a == b) S4[a] += d * d;
      }
  }
  double *o = out;
  o[0] = si[0] / nsweep; o[1] = si[1] / nsweep; o[2] = nc == 3 ? si[2] / nsweep : 0;
  o[3] = sj[0] / nsweep; o[4] = sj[1] / nsweep; o[5] = nc == 3 ? sj[2] / nsweep : 0;
  o[6] = sij / nsweep;
  /* heisenbergLib.c:834 vs xyLib.c:761 group the product differently */
  o[7] = nc == 3 ? MM / nsweep - Mtot / nsweep * Mtot / nsweep
                 : MM / nsweep - (Mtot / nsweep) * (Mtot / nsweep);
  o[8] = totE / nsweep;
  o[9] = E2 / nsweep;
  o[10] = (M2 / nsweep) * (M2 / nsweep) / (M4 / nsweep);
  o[11] = sir[0] / nsweep; o[12] = sir[1] / nsweep; o[13] = nc == 3 ? sir[2] / nsweep : 0;
  o[14] = sjr[0] / nsweep; o[15] = sjr[1] / nsweep; o[16] = nc == 3 ? sjr[2] / nsweep : 0;
  o[17] = sijr / nsweep;
  o[18] = totEr / nsweep;
  o[19] = E2r / nsweep;
  o[20] = siz / nsweep; o[21] = sjz / nsweep; o[22] = stz / nsweep;
  o[23] = sih / nsweep; o[24] = sjh / nsweep; o[25] = sth / nsweep;
  o[26] = Q / nsweep;
  if (groups) {
    for (long a = 0; a < (ng + 1) * (ng + 1); a++) groups[a] = SS[a] / nsweep;
    for (long a = 0; a < ng + 1; a++) groups[(ng + 1) * (ng + 1) + a] = S4[a] / nsweep;
  }
  free(SS); free(S4); free(G);
  free(m.s); free(m.tr); free(m.pp); free(m.sdn); free(m.S); free(m.len);
  free(m.chosen); free(m.inblock); free(m.proj); free(m.queue);
  return 0;
}
