"""Host side of the batched engine: compiles per-site link tables into the
device layout of include/mcgpu.h, owns the device buffers (torch tensors on
the ROCm device), and drives libmcgpu through the C ABI.

Device layout (HBM), per lattice ("graph"), shared by all its replicas:
  link[n*zmax]   uint32  neighbour | coupling-id << 24   (4 B / link)
  coup[ncoup*9]  fp32    distinct 3x3 tensors, row 0 = zero (<= 256)
  cls[n] int32 + clsdat[ncls*8] fp32: (S, Dx, Dy, Dz, 1/S^2, 0, 0, 0) per site class
  col_ptr/col_site: sites grouped by colour of a proper graph colouring
Per replica: spins as structure-of-arrays, component c of replica r at
spin_off + c*n (fp32, fp16 or int8 for Ising).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L

# reference tensor order [xx,yy,zz,xy,xz,yz,yx,zx,zy] -> row-major 3x3
_REF_TO_ROWMAJOR = np.array([0, 3, 4, 6, 1, 5, 7, 8, 2])


def colour_graph(nbr: np.ndarray) -> np.ndarray:
    """Proper colouring of the site graph (libmcgpu: greedy in site order, or
    DSatur where that needs fewer colours)."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    n = nbr.shape[0]
    z = nbr.shape[1] if nbr.ndim == 2 else 0
    col = np.empty(n, np.int32)
    L.check(L.lib().mcg_colour_graph(n, z, nbr.ctypes.data, col.ctypes.data), "mcg_colour_graph")
    return col


@dataclass
class GraphTables:
    """Host copy of one compiled graph (numpy), plus its device tensors."""
    n: int
    zmax: int
    link: np.ndarray
    coup: np.ndarray
    cls: np.ndarray
    clsdat: np.ndarray
    col_ptr: np.ndarray
    col_site: np.ndarray
    pairs: np.ndarray
    circ: np.ndarray
    rorb: np.ndarray
    rclu: np.ndarray
    rlink: np.ndarray
    groups: np.ndarray
    gmax: int
    rrank: np.ndarray
    colour: np.ndarray
    pos: Optional[np.ndarray] = None
    pat: Optional[np.ndarray] = None
    pattab: Optional[np.ndarray] = None
    npat: int = 0
    lcid: Optional[np.ndarray] = None
    ucid: int = -1
    dev: Optional[dict] = None
    struct: Optional[L.Graph] = None

    @property
    def ncol(self) -> int:
        return len(self.col_ptr) - 1


# transpose in the reference tensor order [xx,yy,zz,xy,xz,yz,yx,zx,zy]
_T9 = np.array([0, 1, 2, 6, 7, 8, 3, 4, 5])


def reciprocal_couplings(nbr: np.ndarray, J: np.ndarray, ising: bool, diag_only: bool = False,
                         rtol: float = 1e-9) -> bool:
    """True when every coupling has its reciprocal: the summed tensor of the
    links i -> j equals the transpose of the summed tensor of the links
    j -> i (Ising: the same number).  Only then do the single-site updates
    change one energy function, which the full FP64 recompute can check the
    tracked energy against.  Lattices built from a bond list are reciprocal
    (Lattice.py:256-262: the target gets J^T); tuples read with the Ising
    linkStrength + i convention (isingLib.c:30) and unequal couplings, or
    in-place merged tensors of tiny lattices, are not."""
    nbr = np.asarray(nbr, np.int64)
    n, z = nbr.shape
    if z == 0 or n == 0:
        return True
    Jv = np.asarray(J, np.float64).reshape(n, z, -1)
    Jv = Jv[..., :1] if ising else Jv[..., :9].copy()
    if diag_only and not ising:
        Jv[..., 3:] = 0.0
    i = np.repeat(np.arange(n), z)
    j = nbr.reshape(-1)
    m = j >= 0
    i, j, v = i[m], j[m], Jv.reshape(n * z, -1)[m]
    key = i * n + j
    uk, inv = np.unique(key, return_inverse=True)
    sums = np.zeros((len(uk), v.shape[1]))
    np.add.at(sums, inv, v)
    rk = (uk % n) * n + uk // n
    pos = np.minimum(np.searchsorted(uk, rk), len(uk) - 1)
    if not np.all(uk[pos] == rk):
        return False
    rev = sums[pos] if ising else sums[pos][:, _T9]
    scale = max(float(np.abs(sums).max()), 1e-300)
    return bool(np.all(np.abs(sums - rev) <= rtol * scale))


def compile_graph(nc: int, nbr: np.ndarray, J: np.ndarray, S: np.ndarray, D: np.ndarray,
                  pairs: Optional[np.ndarray] = None, circuits: Optional[np.ndarray] = None,
                  rOrb: Optional[np.ndarray] = None, rOrbCluster: Optional[np.ndarray] = None,
                  rlink: Optional[np.ndarray] = None, groups: Optional[Sequence[Sequence[int]]] = None,
                  diag_only: bool = False, colour: Optional[np.ndarray] = None) -> GraphTables:
    """Per-site tables -> device tables.

    nbr [n,zmax] (-1 padded), J [n,zmax,9] in reference order (Ising: J[...,0]),
    S [n] signed spin length, D [n,3].  Couplings are unscaled (the engine
    multiplies dE by beta) or pre-scaled with beta = 1, as the caller wishes.
    """
    nbr = np.asarray(nbr, np.int64)
    n, zmax = nbr.shape
    if n > (1 << 24):
        raise ValueError("more than 2^24 sites per replica")
    J = np.asarray(J, np.float64).reshape(n, zmax, -1) if zmax else np.zeros((n, 0, 9))
    S = np.asarray(S, np.float64).reshape(n)
    D = np.asarray(D, np.float64).reshape(n, 3)
    valid = nbr >= 0
    if nbr.size and (nbr.max() >= n or nbr.min() < -1):
        raise ValueError("neighbour index outside [-1, n)")
    tgt = np.where(valid, nbr, np.arange(n)[:, None])
    if nc == 1:
        jr = np.where(valid, J[:, :, 0], 0.0)
        folded = jr * S[tgt]
        rows = np.zeros((n, zmax, 9), np.float32)
        rows[:, :, 0] = folded
        rows[:, :, 1] = jr
    else:
        jj = np.where(valid[:, :, None], J[:, :, :9], 0.0)
        if diag_only:
            jj = jj.copy()
            jj[:, :, 3:] = 0.0
        if nc == 2:  # XY reads xx, yy, xy, yx only (xyLib.c:150-153)
            jj = jj.copy()
            jj[:, :, [2, 4, 5, 7, 8]] = 0.0
        rows = jj[:, :, _REF_TO_ROWMAJOR].astype(np.float32)
    flat = rows.reshape(-1, 9)
    allrows = np.concatenate([np.zeros((1, 9), np.float32), flat])
    uniq, inv = np.unique(allrows, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    zero_id = inv[0]
    # put the zero tensor at index 0
    perm = np.arange(len(uniq))
    perm[[0, zero_id]] = perm[[zero_id, 0]]
    uniq = uniq[perm]
    remap = np.empty_like(perm)
    remap[perm] = np.arange(len(perm))
    cid = remap[inv[1:]].reshape(n, zmax)
    wide = len(uniq) > 256
    # uniform: every site has zmax links and all carry one coupling tensor
    # (equal FP32 rows); the kernels then sum the neighbour spins first
    ucid = int(cid[0, 0]) if (zmax > 0 and valid.all() and (cid == cid[0, 0]).all()) else -1
    if wide:
        # per-link coupling indices; link words carry the neighbour only
        lcid = cid.astype(np.int32).reshape(-1)
        link = tgt.astype(np.uint32)
    else:
        lcid = None
        link = (tgt.astype(np.uint32) | (cid.astype(np.uint32) << 24)).astype(np.uint32)

    cd = np.concatenate([S[:, None], D], 1).astype(np.float32)
    cu, cinv = np.unique(cd, axis=0, return_inverse=True)
    cls = cinv.reshape(-1).astype(np.int32)
    # per class: S, Dx, Dy, Dz, 1/S^2 in FP32 (correctly rounded), padding
    s32 = cu[:, 0].astype(np.float32)
    with np.errstate(divide="ignore"):
        inv_s2 = np.where(s32 != 0, np.float32(1.0) / (s32 * s32), np.float32(0.0)).astype(np.float32)
    cu = np.concatenate([cu, inv_s2[:, None], np.zeros((len(cu), 3), np.float32)], 1).astype(np.float32)

    if colour is None:
        colour = colour_graph(np.where(valid, nbr, -1).astype(np.int32))
    colour = np.asarray(colour, np.int32)
    order = np.argsort(colour, kind="stable").astype(np.int32)
    ncol = int(colour.max()) + 1 if n else 0
    col_ptr = np.zeros(ncol + 1, np.int32)
    col_ptr[1:] = np.cumsum(np.bincount(colour, minlength=ncol))

    pairs = np.zeros((0, 2), np.int32) if pairs is None else np.asarray(pairs, np.int32).reshape(-1, 2)
    circ = np.zeros((0, 3), np.int32) if circuits is None else np.asarray(circuits, np.int32).reshape(-1, 3)
    rorb = np.zeros(0, np.int32) if rOrb is None else np.asarray(rOrb, np.int32).reshape(-1)
    nr = len(rorb)
    rclu = np.zeros((nr, 1), np.int32) if rOrbCluster is None or nr == 0 else \
        np.asarray(rOrbCluster, np.int32).reshape(nr, -1)
    rrank = -np.ones(n, np.int32)
    rrank[rorb] = np.arange(nr, dtype=np.int32)
    if nr and rlink is not None:
        rl = np.asarray(rlink, np.int64).reshape(nr, -1)[:, :zmax]
        if rl.shape[1] < zmax:
            rl = np.concatenate([rl, -np.ones((nr, zmax - rl.shape[1]), np.int64)], 1)
        # a block-spin link uses the coupling of the site's normal link l
        # (heisenbergLib.c:269-279); the link must exist and point at a
        # block-spin site, otherwise it contributes nothing
        rc = cid[rorb]
        ok = (rl >= 0) & (np.arange(zmax)[None, :] < (nbr[rorb] >= 0).sum(1)[:, None])
        ok &= np.where(rl >= 0, rrank[np.clip(rl, 0, n - 1)] >= 0, False)
        rt = np.where(ok, rl, rorb[:, None])
        rc = np.where(ok, rc, 0)
        if wide:  # the kernel reads the coupling of link l from lcid
            rt = np.where(ok, rt, -1)
            rlk = np.where(rt >= 0, rt, 0).astype(np.uint32) | np.where(ok, 0, 0xFF000000).astype(np.uint32)
        else:
            rlk = (rt.astype(np.uint32) | (rc.astype(np.uint32) << 24)).astype(np.uint32)
    else:
        rlk = np.zeros((nr, zmax), np.uint32)
    for name, arr in (("corrOrbPair", pairs), ("localCircuits", circ), ("rOrb", rorb), ("rOrbCluster", rclu)):
        if arr.size and (arr.min() < 0 or arr.max() >= n):
            raise ValueError("%s index outside [0, n)" % name)
    groups = list(groups or [])
    for gg in groups:
        if len(gg) and (min(gg) < 0 or max(gg) >= n):
            raise ValueError("orbital group index outside [0, n)")
    if len(groups) > 16:
        raise ValueError("at most 16 orbital groups are supported")
    gmax = max([len(g) for g in groups] + [1])
    garr = -np.ones((max(len(groups), 1), gmax), np.int32)
    for k, g in enumerate(groups):
        garr[k, :len(g)] = g
    # ---- position space.  Spins live in HBM and LDS in colour order:
    # position p holds site order[p]; colour c is the contiguous range
    # col_ptr[c] .. col_ptr[c+1].  Every table the kernels read is
    # translated to positions here, so no kernel indexes by site.
    pos = np.empty(n, np.int32)
    pos[order] = np.arange(n, dtype=np.int32)
    site_of = order.astype(np.int64)
    link = link.reshape(n, zmax)[site_of]
    link = ((pos[(link & 0xFFFFFF).astype(np.int64)].astype(np.uint32)) | (link & np.uint32(0xFF000000))).astype(
        np.uint32)
    if lcid is not None:
        lcid = lcid.reshape(n, zmax)[site_of].reshape(-1)
    cls_pos = cls[site_of]
    pairs = pos[pairs] if len(pairs) else pairs
    circ = pos[circ] if len(circ) else circ
    rrank_pos = -np.ones(n, np.int32)
    if nr:
        rrank_pos[pos[rorb]] = np.arange(nr, dtype=np.int32)
        rclu = pos[rclu]
        rt = (rlk & 0xFFFFFF).astype(np.int64)
        rlk = (pos[rt].astype(np.uint32) | (rlk & np.uint32(0xFF000000))).astype(np.uint32)
        rorb = pos[rorb]
    garr = np.where(garr >= 0, pos[np.clip(garr, 0, n - 1)], -1).astype(np.int32)
    delta = pos[tgt[site_of]].astype(np.int64) - np.arange(n)[:, None]
    keyrows = np.concatenate([delta, cid[site_of].astype(np.int64), cls_pos.astype(np.int64)[:, None]], 1)
    prow, pinv = np.unique(keyrows, axis=0, return_inverse=True)
    if len(prow) <= 256 and n < (1 << 23) and not wide:
        npat = len(prow)
        words = (prow[:, :zmax] & 0xFFFFFF) | (prow[:, zmax:2 * zmax] << 24)
        pattab = np.concatenate([words, prow[:, 2 * zmax:]], 1).astype(np.uint32).view(np.int32).reshape(-1)
        pat = pinv.reshape(-1).astype(np.uint8)
    else:
        npat, pattab, pat = 0, np.zeros(1, np.int32), np.zeros(n, np.uint8)
    return GraphTables(n, zmax, link.reshape(-1), uniq.astype(np.float32).reshape(-1), cls_pos, cu.reshape(-1),
                       col_ptr, order, pairs, circ, rorb, rclu, rlk.reshape(-1), garr, gmax, rrank_pos, colour,
                       pos=pos, pat=pat, pattab=pattab, npat=npat, lcid=lcid, ucid=ucid)


def _dev(a: np.ndarray, device) -> torch.Tensor:
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    if a.size == 0:
        a = np.zeros(1, a.dtype)
    return torch.from_numpy(a).to(device)


def upload_graph(g: GraphTables, device) -> None:
    if g.dev is not None:
        return
    d = {k: _dev(getattr(g, k), device) for k in ("link", "coup", "cls", "clsdat", "col_ptr", "col_site", "pairs",
                                                  "circ", "rorb", "rclu", "rlink", "groups", "rrank", "pos", "pat",
                                                  "pattab")}
    if g.lcid is not None:
        d["lcid"] = _dev(g.lcid, device)
    g.dev = d
    s = L.Graph()
    s.n, s.zmax, s.ncol = g.n, g.zmax, g.ncol
    s.ncoup, s.ncls = len(g.coup) // 9, len(g.clsdat) // 8
    s.nlat, s.ncirc, s.nr = len(g.pairs), len(g.circ), len(g.rorb)
    s.nclu = g.rclu.shape[1] if len(g.rorb) else 0
    s.ngroup = 0 if (g.groups < 0).all() else g.groups.shape[0]
    s.gmax = g.gmax
    s.npat = g.npat
    s.ucid = g.ucid
    # every site paired with itself once: the frame's pair sums follow from
    # the tracked magnetisation and the spin lengths
    pr = np.asarray(g.pairs).reshape(-1, 2)
    if len(pr) == g.n and g.n > 0 and np.array_equal(pr[:, 0], pr[:, 1]) and \
            np.array_equal(np.sort(pr[:, 0]), np.arange(g.n)):
        sl = np.asarray(g.clsdat, np.float32).reshape(-1, 8)[:, 0].astype(np.float64)[np.asarray(g.cls)]
        s.self_pairs, s.pair_s2 = 1, float(np.sum(sl * sl))
    for k, t in d.items():
        setattr(s, k, t.data_ptr())
    g.struct = s


_STORE = {"f32": (L.STORE_F32, torch.float32), "f16": (L.STORE_F16, torch.float16), "i8": (L.STORE_I8, torch.int8)}


def _pow2_exponent(x: np.ndarray, kmax: int) -> Optional[int]:
    """Smallest k <= kmax with every x * 2^k an integer, or None."""
    x = np.asarray(x, np.float64).ravel()
    for k in range(kmax + 1):
        y = x * 2.0 ** k
        if np.all(y == np.round(y)):
            return k
    return None


def exact_ising_table(graphs: List["GraphTables"], beta: Sequence[float], h: Sequence[float],
                      max_entries: int = 1 << 16):
    """The exact Ising acceptance of a batch (include/mcgpu.h, mcg_batch.ithr),
    or None.  When every folded coupling J*S_target and every field is a
    multiple of 2^-kf (kf <= 8) and every spin length a multiple of 2^-ks
    (ks <= 4), the sweep's FP32 local field f = sum J S_t sigma_j and
    dE = -2 sigma_i S_i (f - h) are exact multiples of u = 2^(1-kf-ks), and a
    flip with dE = k u > 0 is accepted when the uint32 draw is below
    thr[row, k] = floor(2^32 exp(-beta_row k u)), beta_row the FP32 inverse
    temperature the kernel holds.  Returns (thr [nrow, nthr] uint32, 1/u);
    None when the inputs are not such multiples or the largest |dE|/u needs
    more than max_entries entries per row (the sweep then keeps the FP32
    exp(-beta dE) acceptance).  isingLib.c:238-254 accepts when
    exp(corr) > rand()/RAND_MAX."""
    J = np.concatenate([g.coup.reshape(-1, 9)[:, 0] for g in graphs]).astype(np.float32).astype(np.float64)
    S = np.concatenate([g.clsdat.reshape(-1, 8)[:, 0] for g in graphs]).astype(np.float32).astype(np.float64)
    hh = np.asarray(h, np.float32).astype(np.float64)
    kf = _pow2_exponent(np.concatenate([J, hh]), 8)
    ks = _pow2_exponent(S, 4)
    if kf is None or ks is None:
        return None
    unit = 2.0 ** (1 - kf - ks)
    hmax = float(np.abs(hh).max()) if hh.size else 0.0
    kmax = 0
    for g in graphs:
        cid = g.lcid.reshape(g.n, -1) if g.lcid is not None else (g.link.reshape(g.n, -1) >> 24)
        cj = g.coup.reshape(-1, 9)[:, 0].astype(np.float64)
        fsum = np.abs(cj[cid.astype(np.int64)]).sum(1) if g.zmax else np.zeros(g.n)
        sa = np.abs(g.clsdat.reshape(-1, 8)[:, 0].astype(np.float64)[g.cls])
        kmax = max(kmax, int(np.ceil(np.max(2.0 * sa * (fsum + hmax)) / unit)))
    nthr = kmax + 1
    if nthr > max_entries:
        return None
    b32 = np.asarray(beta, np.float32).astype(np.float64)
    k = np.arange(nthr, dtype=np.float64)
    thr = np.floor(np.exp(-b32[:, None] * k[None, :] * unit) * 2.0 ** 32)
    return np.minimum(thr, 2.0 ** 32 - 1).astype(np.uint32), 1.0 / unit


@dataclass
class ReplicaSpec:
    graph: int
    row0: int
    nrung: int = 1
    ladder: int = -1
    rung: int = 0          # initial rung (row = row0 + rung)


class Batch:
    """Many replicas (of possibly different lattices) run by one engine call."""

    def __init__(self, nc: int, graphs: List[GraphTables], reps: List[ReplicaSpec], beta: Sequence[float],
                 h: Sequence[float], init_spins: List[np.ndarray], store: str = "f32", meas: int = 0,
                 nbins: int = 32, nframes: int = 1, seed: int = 0x5EED2026, wolff: bool = False,
                 recompute_every: int = 0, device="cuda", serial_base: int = 0, exact_ising: bool = True):
        self.nc = nc
        self.device = torch.device(device)
        if nc == 1:
            store = "i8"
        self.store_code, self.dtype = _STORE[store]
        self.graphs = graphs
        self.reps = reps
        for g in graphs:
            upload_graph(g, self.device)
        nrep = len(reps)
        nrow = len(beta)
        self.nrow = nrow
        # spin pool: per replica nc component arrays in colour order
        # (position p = site col_site[p]), each replica 16-byte aligned
        esz = {"f32": 4, "f16": 2, "i8": 1}[store]
        al = 16 // esz
        offs, soffs, tot, stot = [], [], 0, 0
        for r in reps:
            n = graphs[r.graph].n
            offs.append(tot)
            soffs.append(stot)
            tot += (nc * n + al - 1) // al * al
            stot += n
        self.spin_off = offs
        pool = np.zeros(max(tot, 1), np.float32)
        for k, (r, sp) in enumerate(zip(reps, init_spins)):
            g = graphs[r.graph]
            n = g.n
            a = np.asarray(sp, np.float64).reshape(n, -1)[g.col_site]
            if nc == 1:
                pool[offs[k]:offs[k] + n] = np.where(a[:, 0] >= 0, 1.0, -1.0)
            else:
                pool[offs[k]:offs[k] + nc * n] = a[:, :nc].T.reshape(-1)
        self.spins = torch.from_numpy(pool).to(self.device).to(self.dtype)
        f64 = dict(dtype=torch.float64, device=self.device)
        i32 = dict(dtype=torch.int32, device=self.device)
        self.row_beta = torch.tensor(np.asarray(beta, np.float32), device=self.device)
        self.row_h = torch.tensor(np.asarray(h, np.float32), device=self.device)
        rung = np.array([r.rung for r in reps], np.int32)
        who = -np.ones(nrow, np.int32)
        for k, r in enumerate(reps):
            if not (0 <= r.rung < r.nrung) or r.row0 < 0 or r.row0 + r.nrung > nrow:
                raise ValueError("replica %d: rung/rows outside the batch" % k)
            if who[r.row0 + r.rung] >= 0:
                raise ValueError("two replicas on row %d" % (r.row0 + r.rung))
            who[r.row0 + r.rung] = k
        for r in reps:
            if r.ladder >= 0 and (who[r.row0:r.row0 + r.nrung] < 0).any():
                raise ValueError("tempering ladder at row %d is not fully occupied" % r.row0)
        self.rung = torch.from_numpy(rung).to(self.device)
        self.who = torch.from_numpy(who).to(self.device)
        self.state = torch.zeros(nrep * L.NSTATE, **f64)
        self.acc = torch.zeros(nrow * L.NACC, **f64)
        self.nbins = max(1, nbins)
        self.bins = torch.zeros(nrow * self.nbins * L.NBINQ, **f64)
        ng = max([g.groups.shape[0] if not (g.groups < 0).all() else 0 for g in graphs] + [0])
        self.ngroup = ng
        self.gacc = torch.zeros(max(1, nrow * (ng + 2) * (ng + 1)), **f64)
        self.pt_acc = torch.zeros(nrow, **i32)
        self.err = torch.zeros(1, **i32)
        self.wstamp = torch.zeros(max(stot, 1), **i32) if wolff else None
        self.wlist = torch.zeros(max(stot, 1), **i32) if wolff else None
        # structs
        self.h_graphs = (L.Graph * len(graphs))(*[g.struct for g in graphs])
        self.h_reps = (L.Replica * nrep)()
        for k, r in enumerate(reps):
            x = self.h_reps[k]
            x.graph, x.ladder, x.row0, x.nrung = r.graph, r.ladder, r.row0, r.nrung
            x.serial = (serial_base + k) & 0x7FFFFFFF
            x.spin_off, x.site_off = offs[k], soffs[k]
        self.d_graphs = torch.frombuffer(bytearray(bytes(self.h_graphs)), dtype=torch.uint8).to(self.device)
        self.d_reps = torch.frombuffer(bytearray(bytes(self.h_reps)), dtype=torch.uint8).to(self.device)
        b = L.Batch()
        b.nc, b.store, b.ngraph, b.nrep, b.nrow = nc, self.store_code, len(graphs), nrep, nrow
        b.meas, b.nbins, b.nframes = meas, self.nbins, max(1, nframes)
        b.seed_lo, b.seed_hi = seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF
        b.recompute_every = recompute_every
        b.graphs, b.reps, b.spins = self.d_graphs.data_ptr(), self.d_reps.data_ptr(), self.spins.data_ptr()
        b.row_beta, b.row_h = self.row_beta.data_ptr(), self.row_h.data_ptr()
        b.rung, b.who, b.state = self.rung.data_ptr(), self.who.data_ptr(), self.state.data_ptr()
        b.acc, b.bins, b.gacc = self.acc.data_ptr(), self.bins.data_ptr(), self.gacc.data_ptr()
        b.pt_acc, b.err = self.pt_acc.data_ptr(), self.err.data_ptr()
        b.wstamp = self.wstamp.data_ptr() if wolff else None
        b.wlist = self.wlist.data_ptr() if wolff else None
        b.graphs_host = C.addressof(self.h_graphs)
        b.reps_host = C.addressof(self.h_reps)
        # exact Ising acceptance where the inputs allow it (exact_ising_table)
        self.ithr, self.e_unit = None, None
        tab = exact_ising_table(graphs, beta, h) if (nc == 1 and exact_ising) else None
        if tab is not None:
            thr, inv_unit = tab
            self.ithr = torch.from_numpy(thr.view(np.int32).reshape(-1).copy()).to(self.device)
            self.e_unit = 1.0 / inv_unit
            self._ithr_host = (thr, inv_unit)
            b.ithr, b.nthr, b.e_unit_inv = self.ithr.data_ptr(), thr.shape[1], inv_unit
        self.b = b
        self.frames_per_row = max(1, nframes)
        self.wolff = wolff
        self._runners = {}        # plan shape -> prepared run (mcg_runner_*)
        self._next_sweep = 0      # Philox sweep index after the last run
        self._wolff_cl_end = 0    # first Wolff cluster id not yet used

    def ising_table(self, rows=None):
        """The exact Ising acceptance this batch runs with (exact_ising_table):
        (thresholds of `rows` (all rows by default), 1/unit), or None when it
        uses the FP32 exp rule."""
        if self.ithr is None:
            return None
        thr, inv_unit = self._ithr_host
        return (thr if rows is None else thr[np.asarray(rows)]), inv_unit

    def __del__(self):
        self.close()

    def close(self) -> None:
        """Release the prepared runs (their device scratch)."""
        rs, self._runners = getattr(self, "_runners", {}), {}
        for h in rs.values():
            try:
                L.lib().mcg_runner_destroy(h)
            except Exception:
                pass

    # ---------------------------------------------------------------------
    @staticmethod
    def _stream():
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def init(self) -> None:
        """Tracked state from the current spins; clears accumulators, bins,
        swap counts, the error flag and the Wolff stamps."""
        L.check(L.lib().mcg_init(C.byref(self.b), self._stream()), "mcg_init")
        self._wolff_cl_end = 0

    @staticmethod
    def plan(nthermal: int, nsweep: int, sweeps_per_frame: int = 1, algorithm: int = L.METROPOLIS,
             pt_every: int = 0, kernel: int = L.KERNEL_AUTO, block: int = 0, reps_per_block: int = 0,
             clusters_per_sweep: int = 0, sweep0: int = 0, frame0: int = 0, parts: int = 0,
             no_persist: bool = False, timing=None) -> L.Plan:
        p = L.Plan()
        p.algorithm, p.nthermal, p.nsweep, p.sweeps_per_frame = algorithm, nthermal, nsweep, sweeps_per_frame
        p.pt_every, p.kernel, p.block, p.reps_per_block = pt_every, kernel, block, reps_per_block
        p.clusters_per_sweep, p.sweep0, p.frame0, p.parts = clusters_per_sweep, sweep0, frame0, parts
        p.no_persist = int(bool(no_persist))
        if timing is not None:
            p.timing = C.cast(timing, C.POINTER(C.c_double))
        return p

    def runner(self, p: L.Plan):
        """The prepared run (mcg_runner_create) for this plan's shape: made on
        first use, then reused, so that repeated runs only launch kernels."""
        key = (p.kernel, p.block, p.reps_per_block, p.parts, p.no_persist, p.pt_every > 0, p.algorithm,
               self.b.meas, self.b.recompute_every)
        h = self._runners.get(key)
        if h is None:
            out = C.c_void_p()
            L.check(L.lib().mcg_runner_create(C.byref(self.b), C.byref(p), self._stream(), C.byref(out)),
                    "mcg_runner_create")
            h = self._runners[key] = out
        return h

    def run_plan(self, p: L.Plan, prepared: bool = True) -> None:
        """Run plan p: through its prepared run (default), or with mcg_run
        (plans, allocates and waits every call)."""
        ncl = p.clusters_per_sweep if p.clusters_per_sweep > 0 else max(g.n for g in self.graphs)
        total = p.nthermal + p.nsweep * max(1, p.sweeps_per_frame)
        if p.algorithm == L.WOLFF and self.wstamp is not None:
            # cluster ids are (sweep index) x ncl + c; stamps of an earlier run
            # with the same ids would be taken for members of the new cluster
            if p.sweep0 * ncl < self._wolff_cl_end:
                self.wstamp.zero_()
            self._wolff_cl_end = (p.sweep0 + total) * ncl
        if prepared:
            L.check(L.lib().mcg_runner_run(self.runner(p), C.byref(self.b), C.byref(p), self._stream()),
                    "mcg_runner_run")
        else:
            L.check(L.lib().mcg_run(C.byref(self.b), C.byref(p), self._stream()), "mcg_run")
        self._next_sweep = p.sweep0 + total

    def run(self, nthermal: int, nsweep: int, sweeps_per_frame: int = 1, algorithm: int = L.METROPOLIS,
            pt_every: int = 0, kernel: int = L.KERNEL_AUTO, block: int = 0, reps_per_block: int = 0,
            clusters_per_sweep: int = 0, sweep0: Optional[int] = None, frame0: int = 0, parts: int = 0,
            no_persist: bool = False, prepared: bool = True) -> None:
        """parts: workgroups per LDS replica (0 = balance automatically, 1 =
        never split, 2..4 = split every LDS replica).  no_persist: never run
        the LDS kernel as a persistent launch with tempering inside.  sweep0:
        Philox sweep index of the first sweep; None continues after the last
        run of this batch (0 for the first)."""
        if sweep0 is None:
            sweep0 = self._next_sweep
        self.run_plan(self.plan(nthermal, nsweep, sweeps_per_frame, algorithm, pt_every, kernel, block,
                                reps_per_block, clusters_per_sweep, sweep0, frame0, parts, no_persist),
                      prepared=prepared)

    def sweeps(self, sweep0: Optional[int], nsweeps: int, algorithm: int = L.METROPOLIS,
               kernel: int = L.KERNEL_AUTO, clusters_per_sweep: int = 0, block: int = 0, parts: int = 0,
               prepared: bool = True) -> None:
        """nsweeps unmeasured sweeps (the same as mcg_sweeps)."""
        if sweep0 is None:
            sweep0 = self._next_sweep
        self.run_plan(self.plan(nsweeps, 0, 1, algorithm, 0, kernel, block, 0, clusters_per_sweep, sweep0, 0,
                                parts), prepared=prepared)

    def measure(self, frame: int) -> None:
        L.check(L.lib().mcg_measure(C.byref(self.b), frame, self._stream()), "mcg_measure")

    # ---------------------------------------------------------------------
    def set_nframes(self, nframes: int) -> None:
        self.b.nframes = max(1, nframes)

    def accumulators(self) -> np.ndarray:
        torch.cuda.synchronize()
        return self.acc.cpu().numpy().reshape(self.nrow, L.NACC)

    def bin_data(self) -> np.ndarray:
        torch.cuda.synchronize()
        return self.bins.cpu().numpy().reshape(self.nrow, self.nbins, L.NBINQ)

    def group_acc(self) -> np.ndarray:
        torch.cuda.synchronize()
        ng = self.ngroup
        return self.gacc.cpu().numpy()[: self.nrow * (ng + 2) * (ng + 1)].reshape(self.nrow, -1)

    def state_data(self) -> np.ndarray:
        torch.cuda.synchronize()
        return self.state.cpu().numpy().reshape(-1, L.NSTATE)

    def replica_spins(self, k: int) -> np.ndarray:
        """[n, nc] spins of replica k in site order (Ising: sigma, S not applied)."""
        g = self.graphs[self.reps[k].graph]
        n = g.n
        o = self.spin_off[k]
        a = self.spins[o:o + self.nc * n].float().cpu().numpy().reshape(self.nc, n).T
        out = np.empty_like(a)
        out[g.col_site] = a
        return out

    def error_flag(self) -> int:
        torch.cuda.synchronize()
        return int(self.err.item())

    # ---------------------------------------------------------------------
    _CKPT = ("spins", "state", "acc", "bins", "gacc", "pt_acc", "rung", "who", "err")

    def checkpoint(self) -> dict:
        """Everything a run continues from, in host memory: spins, tracked
        state, accumulators, bins, group sums, ladder state (rung, who, swap
        counts), the error flag, the next Philox sweep index and the Wolff
        cluster stamps.  Philox is counter based, so a run restored from it
        (restore, then run with sweep0 = next_sweep) continues the trajectory
        bit for bit, in this process or another."""
        torch.cuda.synchronize()
        ck = {k: getattr(self, k).cpu().numpy().copy() for k in self._CKPT}
        ck["next_sweep"] = np.int64(self._next_sweep)
        ck["wolff_cl_end"] = np.int64(self._wolff_cl_end)
        if self.wstamp is not None:
            ck["wstamp"] = self.wstamp.cpu().numpy().copy()
        return ck

    def restore(self, ck: dict) -> None:
        """Load a checkpoint taken from a batch of the same shape."""
        names = list(self._CKPT) + (["wstamp"] if self.wstamp is not None else [])
        for k in names:
            t = getattr(self, k)
            a = np.asarray(ck[k])
            if a.shape != tuple(t.shape) or a.dtype != t.cpu().numpy().dtype:
                raise ValueError("checkpoint %s: shape %s / %s does not match the batch" % (k, a.shape, a.dtype))
            t.copy_(torch.from_numpy(a).to(t.device))
        self._next_sweep = int(ck["next_sweep"])
        self._wolff_cl_end = int(ck["wolff_cl_end"])
        torch.cuda.synchronize()
