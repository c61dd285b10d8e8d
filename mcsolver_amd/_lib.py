"""ctypes binding of libmcgpu.so (the C ABI declared in include/mcgpu.h).

The library is the only compute path: there is no CPU fallback.  Loading
fails loudly when the .so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcgpu.so")

# --- enums (include/mcgpu.h) -------------------------------------------------
STORE_F32, STORE_F16, STORE_I8 = 0, 1, 2
METROPOLIS, WOLFF = 0, 1
MEAS_PAIRS, MEAS_Q, MEAS_RNORM, MEAS_GROUPS, ISING_RELE = 1, 2, 4, 8, 16
(A_E, A_E2, A_M, A_M2, A_M4, A_MM, A_MPREV) = range(7)
A_SI, A_SJ, A_SIJ, A_ZI, A_ZJ, A_ZT, A_HI, A_HJ, A_HT, A_Q = 7, 10, 13, 14, 15, 16, 17, 18, 19, 20
A_SIR, A_SJR, A_SIJR, A_ER, A_ER2, A_COUNT, A_LAST, A_TOT, A_ABSM = 21, 24, 27, 28, 29, 30, 31, 32, 33
NACC = 34
S_E, S_M, S_EREF, S_DRIFT, S_ACC, S_TRY = 0, 1, 4, 5, 6, 7
NSTATE = 8
NBINQ = 5
KERNEL_AUTO, KERNEL_LDS, KERNEL_STREAM, KERNEL_PLAIN = 0, 1, 2, 3
ERR_BARRIER, ERR_NONFINITE, ERR_LENGTH, ERR_DRIFT = 16, 32, 64, 128
# production period of the full FP64 recompute (tracked-energy drift, NaN /
# Inf and spin-length checks): a recompute costs about one sweep
RECOMPUTE_EVERY = 1000

P = C.c_void_p


class Graph(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("n", "zmax", "ncol", "ncoup", "ncls", "nlat", "ncirc", "nr", "nclu",
                                         "ngroup", "gmax", "npat", "ucid", "pad_")] + \
               [(k, P) for k in ("link", "coup", "cls", "clsdat", "col_ptr", "col_site", "pairs", "circ", "rorb",
                                 "rclu", "rlink", "groups", "rrank", "pos", "pat", "pattab", "lcid")] + \
               [("pair_s2", C.c_double), ("self_pairs", C.c_int32), ("pad2_", C.c_int32)]


class Replica(C.Structure):
    _fields_ = [("graph", C.c_int32), ("ladder", C.c_int32), ("row0", C.c_int32), ("nrung", C.c_int32),
                ("serial", C.c_uint32), ("pad_", C.c_int32), ("spin_off", C.c_int64), ("site_off", C.c_int64)]


class Batch(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nc", "store", "ngraph", "nrep", "nrow", "meas", "nbins", "nframes")] + \
               [("seed_lo", C.c_uint32), ("seed_hi", C.c_uint32), ("recompute_every", C.c_int32),
                ("pad_", C.c_int32)] + \
               [(k, P) for k in ("graphs", "reps", "spins", "row_beta", "row_h", "rung", "who", "state", "acc",
                                 "bins", "gacc", "pt_acc", "err", "wstamp", "wlist", "graphs_host", "reps_host")] + \
               [("ithr", P), ("nthr", C.c_int32), ("e_unit_inv", C.c_float)]


class Plan(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("algorithm", "nthermal", "nsweep", "sweeps_per_frame", "pt_every", "kernel",
                                         "block", "reps_per_block", "clusters_per_sweep", "frame0", "parts",
                                         "no_persist")] + \
               [("sweep0", C.c_int64), ("timing", C.POINTER(C.c_double))]


assert C.sizeof(Graph) == 56 + 17 * 8 + 16
assert C.sizeof(Replica) == 40
assert C.sizeof(Plan) == 64

_lib = None


def lib() -> C.CDLL:
    """Load libmcgpu.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmcgpu.so not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python -m mcsolver_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.mcg_init.argtypes = [C.POINTER(Batch), P]
    L.mcg_run.argtypes = [C.POINTER(Batch), C.POINTER(Plan), P]
    L.mcg_sweeps.argtypes = [C.POINTER(Batch), C.POINTER(Plan), C.c_int64, C.c_int32, P]
    L.mcg_measure.argtypes = [C.POINTER(Batch), C.c_int64, P]
    L.mcg_colour_graph.argtypes = [C.c_int32, C.c_int32, P, P]
    L.mcg_runner_create.argtypes = [C.POINTER(Batch), C.POINTER(Plan), P, C.POINTER(C.c_void_p)]
    L.mcg_runner_run.argtypes = [C.c_void_p, C.POINTER(Batch), C.POINTER(Plan), P]
    L.mcg_runner_destroy.argtypes = [C.c_void_p]
    L.mcg_run_noalloc.argtypes = [C.POINTER(Batch), C.POINTER(Plan), P]
    L.mcg_last_error.restype = C.c_char_p
    L.mcg_version.restype = C.c_int
    for f in (L.mcg_init, L.mcg_run, L.mcg_sweeps, L.mcg_measure, L.mcg_colour_graph, L.mcg_runner_create,
              L.mcg_runner_run, L.mcg_runner_destroy, L.mcg_run_noalloc):
        f.restype = C.c_int
    _lib = L
    return L


EXPORTS = ("mcg_init", "mcg_run", "mcg_runner_create", "mcg_runner_run", "mcg_runner_destroy", "mcg_run_noalloc",
           "mcg_sweeps", "mcg_measure", "mcg_colour_graph", "mcg_last_error", "mcg_version")
NOALLOC_SCRATCH_BYTES = 4 << 20  # MCG_NOALLOC_SCRATCH_BYTES


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError("%s failed: %s" % (what, lib().mcg_last_error().decode()))
    return rc
