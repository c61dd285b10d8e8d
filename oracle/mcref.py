"""ORACLE — test infrastructure only.  ctypes wrapper of oracle/libmcref.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to verify results or to time the CPU baseline.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmcref.so")
_l = None

P = C.c_void_p
L_ = C.c_long
I = C.c_int
D = C.c_double
F = C.c_float
U = C.c_uint32


def lib() -> C.CDLL:
    global _l
    if _l is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _l = C.CDLL(LIB)
        _l.mcref_srand.argtypes = [C.c_uint]
        _l.mcref_rng_mode.argtypes = [I]
        _l.mcref_xoshiro_set.argtypes = [C.c_uint64] * 4
        _l.mcref_xoshiro_next.restype = C.c_uint64
        _l.mcref_ising_run.argtypes = [I, L_, P, L_, L_, L_, I, P, P, P, L_, P, D, L_, L_, P, P, P, I, P, P]
        _l.mcref_on_run.argtypes = [I, I, L_, P, P, L_, L_, L_, I, P, P, P, L_, P, L_, P, L_, L_, P, D, D, L_, L_, P,
                                    P, P, I, I, P, P, P]
        _l.mcref_cb_sweeps.argtypes = [I, I, L_, I, P, P, P, P, I, P, I, I, P, P, P, U, U, L_, L_, P, P, P]
        _l.mcref_cb_ising_table.argtypes = [P, I, F]
        _l.mcref_cb_energy.argtypes = [I, L_, I, P, P, P, P, I, F, P, P]
        _l.mcref_cb_pt.argtypes = [I, P, P, P, P, U, U, I, L_, P]
        _l.mcref_cb_wolff.argtypes = [I, I, L_, I, P, P, P, P, I, P, I, F, F, U, U, U, L_, L_, P, P]
        _l.mcref_philox.argtypes = [U, U, U, U, U, U, P]
        _l.mcref_exp_neg.argtypes = [F]
        _l.mcref_exp_neg.restype = F
        _l.mcref_sincos.argtypes = [U, P]
        _l.mcref_f16_round.argtypes = [F]
        _l.mcref_f16_round.restype = F
    return _l


def _a(x, dt):
    return np.ascontiguousarray(np.asarray(x, dtype=dt))


def ptr(a: np.ndarray):
    return a.ctypes.data_as(P)


@contextlib.contextmanager
def rng(mode: str):
    """Integer draws from `mode` inside the block, glibc rand() after it."""
    rng_mode(mode)
    try:
        yield
    finally:
        rng_mode("glibc")


def rng_mode(mode: str) -> None:
    """Integer source of the restatement's draws: "glibc" (rand(), the
    reference's generator, the default) or "xoshiro" (xoshiro256**, top 31
    bits on rand()'s range, same consumption order)."""
    lib().mcref_rng_mode({"glibc": 0, "xoshiro": 1}[mode])


# ---------------------------------------------------------------- glibc replay
def ising_run(args: tuple, seed: int | None = None):
    """Run the Ising restatement on MCMainFunction-style positional args.
    Returns (10 floats, frames or None)."""
    (alg, initSpin, nth, nsw, nint, maxl, nlink, J, nbr, pairs, h, rOrb, rClu, rLink, spinFrame, _cb) = args
    n = len(initSpin)
    s = _a(initSpin, np.float64); nl = _a(nlink, np.int32); Jv = _a(J, np.float64); nb = _a(nbr, np.int32)
    pr = _a(pairs, np.int32); ro = _a(rOrb, np.int32); rc = _a(rClu, np.int32); rl = _a(rLink, np.int32)
    nr = len(rOrb)
    nclu = len(rClu) // max(nr, 1)
    out = np.zeros(10)
    frames = np.zeros((max(spinFrame, 1), n)) if spinFrame > 0 else None
    if seed is not None:
        lib().mcref_srand(seed)
    rc_ = lib().mcref_ising_run(alg, n, ptr(s), nth, nsw, nint, maxl, ptr(nl), ptr(Jv), ptr(nb), len(pairs) // 2,
                                ptr(pr), h, nr, nclu, ptr(ro), ptr(rc), ptr(rl), spinFrame, ptr(out),
                                ptr(frames) if frames is not None else None)
    if rc_ != 0:
        raise ValueError("mcref_ising_run rejected the input (%d)" % rc_)
    return out, frames


def on_run(nc: int, args: tuple, seed: int | None = None):
    """O(n) restatement on {xylib,heisenberglib}.MCMainFunction args.
    Returns (27 floats, frames or None, groups or None)."""
    (alg, initSpin, initD, nth, nsw, nint, maxl, nlink, J, nbr, circ, pairs, ng, maxg, glist, flunc, h, rOrb, rClu,
     rLink, spinFrame, diag, _cb) = args
    n = len(initSpin)
    s = _a(initSpin, np.float64); Dv = _a(initD, np.float64); nl = _a(nlink, np.int32); Jv = _a(J, np.float64)
    nb = _a(nbr, np.int32); ci = _a(circ if len(circ) else [0, 0, 0], np.int32); pr = _a(pairs, np.int32)
    gl = _a(glist if len(glist) else [-1], np.int32)
    ro = _a(rOrb, np.int32); rc = _a(rClu, np.int32); rl = _a(rLink, np.int32)
    nr = len(rOrb)
    nclu = len(rClu) // max(nr, 1)
    out = np.zeros(27)
    frames = np.zeros((max(spinFrame, 1), n, 3)) if spinFrame > 0 else None
    groups = np.zeros((ng + 2) * (ng + 1))
    if seed is not None:
        lib().mcref_srand(seed)
    rc_ = lib().mcref_on_run(nc, alg, n, ptr(s), ptr(Dv), nth, nsw, nint, maxl, ptr(nl), ptr(Jv), ptr(nb),
                             len(circ) // 3, ptr(ci), len(pairs) // 2, ptr(pr), ng, maxg, ptr(gl), flunc, h, nr, nclu,
                             ptr(ro), ptr(rc), ptr(rl), spinFrame, diag, ptr(out),
                             ptr(frames) if frames is not None else None, ptr(groups))
    if rc_ != 0:
        raise ValueError("mcref_on_run rejected the input (%d)" % rc_)
    return out, frames, (groups if ng > 0 else None)


def ising_tuple(args, seed=None):
    """The reference's 11-tuple (isingLib.c:436-447) from the restatement."""
    out, frames = ising_run(args, seed)
    sf = args[14]
    fr = tuple(tuple(float(v) for v in f) for f in frames) if sf > 0 else 0.0
    return tuple(float(v) for v in out) + (fr,)


def on_tuple(nc, args, seed=None):
    """The reference's 29-tuple (heisenbergLib.c:854-883)."""
    out, frames, groups = on_run(nc, args, seed)
    sf = args[20]
    fr = tuple(tuple(tuple(float(c) for c in v) for v in f) for f in frames) if sf > 0 else 0.0
    gr = tuple(float(v) for v in groups) if groups is not None else 0.0
    return tuple(float(v) for v in out) + (fr, gr)


# ---------------------------------------------------------------- checker
def cb_sweeps(nc, store, nbr, J, S, D, colour, beta, h, serial, seed, sweep0, nsweeps, spins, diag_only=0,
              ising_table=None):
    """Engine-convention coloured Metropolis on replicas of one graph (in place).
    spins: float32 [nrep, nc, n].  ising_table: (thr [nrep, nthr] uint32,
    1/unit), the engine's exact Ising acceptance for these replicas (None:
    the FP32 exp rule).  Returns (dE[nrep], dM[nrep,3])."""
    nbr = _a(nbr, np.int32); n, z = nbr.shape
    J = _a(J, np.float64).reshape(n, z, -1)
    if J.shape[2] < 9:
        J = np.concatenate([J, np.zeros((n, z, 9 - J.shape[2]))], 2)
    J = _a(J, np.float64); S = _a(S, np.float64); D = _a(D, np.float64)
    colour = _a(colour, np.int32)
    nrep = spins.shape[0]
    beta = _a(beta, np.float32); h = _a(h, np.float32); serial = _a(serial, np.uint32)
    assert spins.dtype == np.float32 and spins.flags.c_contiguous
    dE = np.zeros(nrep); dM = np.zeros((nrep, 3))
    thr = None
    if ising_table is not None:
        thr = np.ascontiguousarray(ising_table[0], np.uint32)
        assert thr.shape[0] == nrep
        lib().mcref_cb_ising_table(ptr(thr), thr.shape[1], float(ising_table[1]))
    try:
        lib().mcref_cb_sweeps(nc, store, n, z, ptr(nbr), ptr(J), ptr(S), ptr(D), diag_only, ptr(colour),
                              int(colour.max()) + 1, nrep, ptr(beta), ptr(h), ptr(serial), seed & 0xFFFFFFFF,
                              (seed >> 32) & 0xFFFFFFFF, sweep0, nsweeps, ptr(spins), ptr(dE), ptr(dM))
    finally:
        if thr is not None:
            lib().mcref_cb_ising_table(None, 0, 0.0)
    return dE, dM


def cb_energy(nc, nbr, J, S, D, h, spins1, diag_only=0):
    nbr = _a(nbr, np.int32); n, z = nbr.shape
    J = _a(J, np.float64).reshape(n, z, -1)
    if J.shape[2] < 9:
        J = np.concatenate([J, np.zeros((n, z, 9 - J.shape[2]))], 2)
    J = _a(J, np.float64)
    out = np.zeros(4)
    sp = _a(spins1, np.float32)
    lib().mcref_cb_energy(nc, n, z, ptr(nbr), ptr(J), ptr(_a(S, np.float64)), ptr(_a(D, np.float64)), diag_only,
                          float(h), ptr(sp), ptr(out))
    return out


def cb_pt(beta, E_of_rep, who, rung, seed, ladder, swap_index, acc):
    beta = _a(beta, np.float32); E = _a(E_of_rep, np.float64)
    assert who.dtype == np.int32 and rung.dtype == np.int32 and acc.dtype == np.int32
    lib().mcref_cb_pt(len(beta), ptr(beta), ptr(E), ptr(who), ptr(rung), seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF,
                      ladder, swap_index, ptr(acc))


def cb_wolff(nc, store, nbr, J, S, D, colour, beta, h, serial, seed, c0, ncl, spins1, diag_only=0):
    nbr = _a(nbr, np.int32); n, z = nbr.shape
    J = _a(J, np.float64).reshape(n, z, -1)
    if J.shape[2] < 9:
        J = np.concatenate([J, np.zeros((n, z, 9 - J.shape[2]))], 2)
    J = _a(J, np.float64)
    assert spins1.dtype == np.float32 and spins1.flags.c_contiguous
    sizes = np.zeros(ncl, np.int64)
    colour = _a(colour, np.int32)
    lib().mcref_cb_wolff(nc, store, n, z, ptr(nbr), ptr(J), ptr(_a(S, np.float64)), ptr(_a(D, np.float64)), diag_only,
                         ptr(colour), int(colour.max()) + 1, float(beta), float(h), serial, seed & 0xFFFFFFFF,
                         (seed >> 32) & 0xFFFFFFFF, c0, ncl, ptr(spins1), ptr(sizes))
    return sizes


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().mcref_philox(*[int(c) for c in ctr], *[int(k) for k in key], ptr(out))
    return out


def exp_neg(x: float) -> float:
    return lib().mcref_exp_neg(x)


def sincos(v: int):
    out = np.zeros(2, np.float32)
    lib().mcref_sincos(v, ptr(out))
    return out


def f16_round(x: float) -> float:
    return lib().mcref_f16_round(x)
