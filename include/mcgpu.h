/*
 * libmcgpu — MI355X-native Monte Carlo sweep engine for classical spin models
 * (Ising, XY, Heisenberg).  Plain C ABI: POD structs, device pointers owned by
 * the caller, an explicit HIP stream, an int status (0 = ok), no exceptions
 * and no process-global mutable state across calls (except the per-device
 * scratch array of mcg_run_noalloc).
 *
 * Replaces the single entry point of the reference's three CPython
 * extensions,
 *   isinglib.MCMainFunction      mcsolver/isingLib.c:259-451      (module 453-471)
 *   xylib.MCMainFunction         mcsolver/xyLib.c:413-809         (module 811-829)
 *   heisenberglib.MCMainFunction mcsolver/heisenbergLib.c:478-885 (module 887-905)
 * which mcsolver/mcMain.py:119-124 and :239-248 call once per (T, h) point.
 * The Python modules mcsolver_amd/{isinglib,xylib,heisenberglib}.py keep that
 * exact positional signature and return tuple; they compile the per-site
 * tuples into the tables below and call mcg_run().  mcg_run() also runs many
 * (lattice, T, h, replica) instances in one launch, which replaces the
 * one-process-per-point multiprocessing.Pool of mcsolver/win.py:81-103.
 */
#ifndef MCGPU_H
#define MCGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* spin storage of the spin pool */
enum { MCG_STORE_F32 = 0, MCG_STORE_F16 = 1, MCG_STORE_I8 = 2 };

/* update algorithm (isingLib.c:339-340: 0 = Metropolis, 1 = Wolff) */
enum { MCG_METROPOLIS = 0, MCG_WOLFF = 1 };

/* measurement flags */
enum {
  MCG_MEAS_PAIRS = 1,   /* spin_i/spin_j/spin_ij over correlation pairs      */
  MCG_MEAS_Q = 2,       /* topological charge over triangle circuits          */
  MCG_MEAS_RNORM = 4,   /* block-spin (majority) observables                  */
  MCG_MEAS_GROUPS = 8,  /* orbital-group spin.spin statistics                 */
  MCG_ISING_RELE = 16,  /* Ising: energy relative to end of thermalisation    */
};

/* accumulator row layout (FP64), one row per measured temperature slot */
enum {
  MCG_A_E = 0, MCG_A_E2, MCG_A_M, MCG_A_M2, MCG_A_M4, MCG_A_MM, MCG_A_MPREV,
  MCG_A_SI = 7,   /* 3: sum |S_i/nLat| per component */
  MCG_A_SJ = 10,  /* 3 */
  MCG_A_SIJ = 13, MCG_A_ZI, MCG_A_ZJ, MCG_A_ZT, MCG_A_HI, MCG_A_HJ, MCG_A_HT, MCG_A_Q,
  MCG_A_SIR = 21, /* 3 */
  MCG_A_SJR = 24, /* 3 */
  MCG_A_SIJR = 27, MCG_A_ER, MCG_A_ER2, MCG_A_COUNT, MCG_A_LAST, MCG_A_TOT, MCG_A_ABSM,
  MCG_NACC = 34
};
/* tracked per-replica state (FP64); Wolff adds the sizes of accepted
 * clusters (spins flipped) to MCG_S_ACC and of all grown clusters to MCG_S_TRY */
enum { MCG_S_E = 0, MCG_S_M = 1 /* 3 */, MCG_S_EREF = 4, MCG_S_DRIFT = 5, MCG_S_ACC = 6, MCG_S_TRY = 7,
       MCG_NSTATE = 8 };
/* per-bin accumulators: e, e^2, M, M^2, M^4 */
enum { MCG_NBINQ = 5 };

/* bits of the device error flag (batch->err), raised by the kernels:
 * BARRIER    a grid barrier of a persistent launch timed out (the launch left)
 * NONFINITE  a full FP64 recompute found a non-finite energy or magnetisation
 *            (NaN / Inf in the spins)
 * LENGTH     a full FP64 recompute found the mean relative spin-length
 *            deviation | |s|^2 / S^2 - 1 | above 1e-4 (FP32 storage) or 1e-2
 *            (FP16 storage)
 * DRIFT      a full FP64 recompute found |E_tracked - E_full| above 1e-4 n
 *            (FP32 storage, Ising) or 1e-2 n (FP16 storage), n sites: the
 *            spins changed outside the sweeps
 * The full recompute runs every batch->recompute_every sweeps; it also
 * records the largest |E_tracked - E_full| in MCG_S_DRIFT. */
enum { MCG_ERR_BARRIER = 16, MCG_ERR_NONFINITE = 32, MCG_ERR_LENGTH = 64, MCG_ERR_DRIFT = 128 };

/*
 * One lattice geometry ("graph"): the per-site link tables of
 * mcMain.py:60-78 / 162-174 compiled for the device.  All pointers are device
 * pointers.  link words pack the neighbour site (low 24 bits) and a coupling
 * index (high 8 bits); padding links point at the site itself with coupling
 * 0, which is the zero tensor.
 */
typedef struct mcg_graph {
  int32_t n, zmax, ncol, ncoup, ncls;
  int32_t nlat, ncirc, nr, nclu, ngroup, gmax;
  int32_t npat;            /* link patterns (0 = none: LDS kernel reads link[]) */
  int32_t ucid;            /* >= 0: every site has zmax links, all with coupling ucid; the
                              Metropolis field is then J . (sum of neighbour spins) */
  int32_t pad_;
  const uint32_t *link;    /* [n*zmax] */
  const float *coup;       /* [ncoup*9] row-major 3x3 (E = s_i^T J s_j), unscaled */
  const int32_t *cls;      /* [n] site class */
  const float *clsdat;     /* [ncls*8] : S (signed spin length), Dx, Dy, Dz, 1/S^2
                              (FP32, correctly rounded), 0, 0, 0 */
  const int32_t *col_ptr;  /* [ncol+1] */
  const int32_t *col_site; /* [n] sites grouped by colour */
  const int32_t *pairs;    /* [nlat*2] correlation pairs */
  const int32_t *circ;     /* [ncirc*3] triangle circuits */
  const int32_t *rorb;     /* [nr] block-spin sites */
  const int32_t *rclu;     /* [nr*nclu] their clusters */
  const uint32_t *rlink;   /* [nr*zmax] block-spin links, packed like link */
  const int32_t *groups;   /* [ngroup*gmax], -1 padded */
  const int32_t *rrank;    /* [n] index into rorb, or -1 */
  /* colour-permuted view used by the LDS-resident kernel: position p holds
     site col_site[p]; pos[site] = p.  pat[p] selects a row of pattab:
     zmax words (signed position delta in the low 24 bits | coupling << 24)
     followed by the site class.  Periodic lattices have few patterns. */
  const int32_t *pos;      /* [n] */
  const uint8_t *pat;      /* [n] */
  const int32_t *pattab;   /* [npat*(zmax+1)] */
  const int32_t *lcid;     /* [n*zmax] coupling index per link when there are more
                              than 256 distinct tensors (link words then carry
                              coupling 0); NULL otherwise */
  /* self_pairs != 0: the correlation pairs are (i, i) once for every site (one
     orbital per cell, mcMain.py's default ki_s = ki_t): the pair sums of a
     frame are then sum_i s_i = the tracked magnetisation and
     sum_i s_i.s_i = pair_s2 = sum_i S_i^2, and no spin is read for them */
  double pair_s2;
  int32_t self_pairs, pad2_;
} mcg_graph;
/* Ising couplings: coup[9c+0] = J*S_target (folded, used by the sweep and the
 * energy), coup[9c+1] = J (raw, used by the block-spin energy). */

/* one replica (spin configuration) */
typedef struct mcg_replica {
  int32_t graph;           /* index into the batch's graph array */
  int32_t ladder;          /* parallel-tempering ladder, -1 = none */
  int32_t row0;            /* first accumulator/temperature row of its ladder (or its own row) */
  int32_t nrung;           /* rungs in its ladder (1 without tempering) */
  uint32_t serial;         /* replica identity in the Philox key */
  int32_t pad_;
  int64_t spin_off;        /* element offset of component 0 in the spin pool;
                              component c lives at spin_off + c*n */
  int64_t site_off;        /* offset of this replica's per-site Wolff scratch */
} mcg_replica;

/* a batch: many replicas, possibly of different graphs, in one launch */
typedef struct mcg_batch {
  int32_t nc;              /* spin components: 1 Ising, 2 XY, 3 Heisenberg */
  int32_t store;           /* MCG_STORE_* */
  int32_t ngraph, nrep, nrow;
  int32_t meas;            /* MCG_MEAS_* flags */
  int32_t nbins;           /* bins per row for error bars (>= 1) */
  int32_t nframes;         /* measured frames planned per row (for binning) */
  uint32_t seed_lo, seed_hi;
  int32_t recompute_every; /* full FP64 energy recomputation period (sweeps) */
  int32_t pad_;
  const mcg_graph *graphs; /* device array [ngraph] */
  const mcg_replica *reps; /* device array [nrep] */
  void *spins;             /* device spin pool */
  const float *row_beta;   /* [nrow] inverse temperature per row */
  const float *row_h;      /* [nrow] field per row */
  int32_t *rung;           /* [nrep] current rung (row = row0 + rung) */
  int32_t *who;            /* [nrow] replica at each row (tempering) */
  double *state;           /* [nrep*MCG_NSTATE] */
  double *acc;             /* [nrow*MCG_NACC] */
  double *bins;            /* [nrow*nbins*MCG_NBINQ] */
  double *gacc;            /* [nrow*((ngroup+2)*(ngroup+1))] or NULL */
  int32_t *pt_acc;         /* [nrow] accepted swaps of the pair (row, row+1) */
  int32_t *err;            /* [1] device error flag */
  uint32_t *wstamp;        /* [sum n] Wolff cluster stamps (or NULL without Wolff) */
  int32_t *wlist;          /* [sum n] Wolff cluster members (or NULL) */
  const mcg_graph *graphs_host;  /* host mirror of graphs (launch planning) */
  const mcg_replica *reps_host;  /* host mirror of reps */
  /* Exact Ising acceptance (NULL: FP32 exp(-beta dE) against a 24-bit
   * uniform).  When every folded coupling J*S_target, field and spin length
   * is a multiple of a power of two small enough that the local field and dE
   * are exact in FP32, dE = k e_unit with 0 < k < nthr for every uphill flip,
   * and the flip is accepted when the uint32 draw is below
   * ithr[row*nthr + k] = floor(2^32 exp(-row_beta[row] k e_unit))
   * (isingLib.c:238-254 accepts when exp(corr) > rand()/RAND_MAX). */
  const uint32_t *ithr;    /* [nrow*nthr] device */
  int32_t nthr;
  float e_unit_inv;        /* 1 / e_unit (a power of two) */
} mcg_batch;

/* run plan for mcg_run */
typedef struct mcg_plan {
  int32_t algorithm;       /* MCG_METROPOLIS / MCG_WOLFF */
  int32_t nthermal;        /* thermalisation sweeps (Wolff: cluster rounds) */
  int32_t nsweep;          /* measured frames */
  int32_t sweeps_per_frame;/* sweeps between frames (>= 1) */
  int32_t pt_every;        /* tempering swap period in sweeps, 0 = off */
  int32_t kernel;          /* 0 = auto, 1 = LDS-resident persistent, 2 = streaming
                              (one launch per colour over all replicas), 3 = plain
                              (one launch per replica per colour) */
  int32_t block;           /* threads per block for the LDS kernel (0 = auto) */
  int32_t reps_per_block;  /* replicas packed per LDS block (0 = auto) */
  int32_t clusters_per_sweep; /* Wolff: clusters per replica per sweep (0 = n) */
  int32_t frame0;          /* index of the first measured frame of this call */
  int32_t parts;           /* LDS kernel: workgroups per replica (0 = auto: split replicas of
                              8192+ sites that exceed one CU's share of the batch; 1 = never
                              split; 2..4 = split every LDS replica).  Without a persistent
                              launch, a launch with split replicas runs one sweep. */
  int32_t no_persist;      /* nonzero: never use persistent launches of the LDS kernel (see
                              mcg_run); 0: use them when every workgroup can be resident */
  int64_t sweep0;          /* global sweep index of the first sweep (Philox counter) */
  double *timing;          /* host double[4] or NULL: when set, mcg_run brackets every
                              launch of the sweep kernel with HIP events on the stream
                              and returns [0] total ms, [1] launches, [2] sweeps,
                              [3] site updates of those launches */
} mcg_plan;

/*
 * Initialise the tracked state (energy, magnetisation) of every replica from
 * the current spins: FP64 full-Hamiltonian evaluation.  Clears accumulators,
 * bins, group accumulators, swap counters and the error flag.
 */
int mcg_init(const mcg_batch *host_batch, void *stream);

/*
 * Run plan->nthermal thermalisation sweeps then plan->nsweep measured frames
 * (each plan->sweeps_per_frame sweeps) on every replica of the batch.
 * Stands in for the thermalisation + measurement loops of MCMainFunction
 * (isingLib.c:352-432, xyLib.c:549-758, heisenbergLib.c:614-831).
 * Plans the launches, allocates their scratch, runs, and returns when the
 * work on `stream` is done (it then releases the scratch): the same as
 * mcg_runner_create + mcg_runner_run + mcg_runner_destroy.
 */
int mcg_run(const mcg_batch *host_batch, const mcg_plan *plan, void *stream);

/*
 * A prepared run, for callers that run one batch many times (frame
 * snapshots, chunked production runs, timing loops).  The same loops as
 * mcg_run (isingLib.c:352-432, xyLib.c:549-758, heisenbergLib.c:614-831);
 * the reference has no counterpart because it plans nothing: one call runs
 * one (T, h) point.
 *
 * mcg_runner_create plans the launches of `plan` on `batch` (work items,
 * tempering ladders, split-replica tables), allocates one device region for
 * them, uploads the tables and waits for the upload.  mcg_runner_run then
 * only queues kernel launches on `stream` (no allocation, no host copy, no
 * host wait) and returns.  nthermal, nsweep, sweeps_per_frame, the pt_every
 * period, clusters_per_sweep, sweep0, frame0, algorithm and timing may
 * differ from run to run (a runner prepared for Wolff runs Metropolis
 * without persistent launches).  A run whose plan differs in kernel, block,
 * reps_per_block, parts, no_persist or tempering on/off, or whose batch
 * differs in shape, device arrays, meas or recompute_every, is refused with
 * an error.  Runs of one runner must not overlap in time (they share its
 * region): queue them on one stream.  mcg_runner_destroy waits for the
 * device, then releases the region.
 */
typedef struct mcg_runner mcg_runner;
int mcg_runner_create(const mcg_batch *host_batch, const mcg_plan *plan, void *stream, mcg_runner **out);
int mcg_runner_run(mcg_runner *runner, const mcg_batch *host_batch, const mcg_plan *plan, void *stream);
int mcg_runner_destroy(mcg_runner *runner);

/*
 * mcg_run without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same plan, launches and
 * results as mcg_run; its device scratch (launch tables, arrival counters,
 * barrier words, published energies) lives in a file-scope array of
 * MCG_NOALLOC_SCRATCH_BYTES per device, which every device that loads the
 * library holds until the process ends.  A run that needs more is refused
 * with an error before anything is launched (configs of a few thousand
 * replicas need well under 1 MiB).  Every call on a device shares that
 * device's array and overwrites it: calls that overlap in time on one device
 * (two streams, two threads) race; queue them on one stream.  It uploads
 * the launch tables and waits for that copy (not for the launches), then
 * returns when the launches are queued.  plan->timing must be NULL.
 */
#define MCG_NOALLOC_SCRATCH_BYTES (4u << 20)
int mcg_run_noalloc(const mcg_batch *host_batch, const mcg_plan *plan, void *stream);

/* Advance plan->nsweep sweeps without measuring (used between spin frames). */
int mcg_sweeps(const mcg_batch *host_batch, const mcg_plan *plan, int64_t sweep0, int32_t nsweeps,
               void *stream);

/* Measure one frame (all batch->meas observables) into the accumulators. */
int mcg_measure(const mcg_batch *host_batch, int64_t frame, void *stream);

/* Greedy proper colouring of a site graph given as padded neighbour lists
 * (host memory; -1 = no link).  Writes colour[n]; returns the number of
 * colours, or <0 on error.  Host-side helper for the table compiler. */
int mcg_colour_graph(int32_t n, int32_t zmax, const int32_t *nbr, int32_t *colour);

/* Text of the last error on this thread (never NULL). */
const char *mcg_last_error(void);

/* ABI version of this header. */
int mcg_version(void);

#ifdef __cplusplus
}
#endif
#endif
