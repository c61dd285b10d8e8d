"""Lattice builder: bond list -> per-site link tables, in the reference's layout.

Restates, vectorised with numpy, what ``mcsolver/Lattice.py`` builds and how
``mcsolver/mcMain.py`` flattens it for ``MCMainFunction``:

* site id = ((x*Ly + y)*Lz + z)*norb + o                (Lattice.py:172-184)
* per cell, per orbital o, per bond with source == o, in bond-list order:
  the source gets (target, J) and, unless it is the same site, the target
  gets (source, J^T)                                    (Lattice.py:247-262)
* a second link to an already-linked target is skipped when its coupling
  differs by < 1e-5 (summed |diff| for tensors) and added onto the existing
  link otherwise                                         (Lattice.py:34-57)
* block-spin ("renormalised") sites are cells with x,y,z all even, each with
  the 2x2(x2) cluster of same-orbital sites and links at twice the bond
  offset, duplicates skipped                             (Lattice.py:185-205,
                                                          60-66, 265-271)
* correlation pairs (ki_s in cell c, ki_t in cell c+ki_overLat) per cell
                                                         (Lattice.py:273)
* topological circuits per cell                          (Lattice.py:223-234)
* orbital groups                                         (Lattice.py:208-211)
* beta scaling: J -> (1/T)*J, D -> D/T, h -> h/T, T clamped to >= 0.1
                                                         (mcMain.py:21-31,
                                                          Lattice.py:141-147,
                                                          mcMain.py:121,244)
* padding: links to maxNLinking with -1 / 0.0, rnorm links with -1,
  groups with -1                                         (mcMain.py:60-98,
                                                          150-221)

Repeated targets (only for L <= 2*max|overLat|) follow the reference's
object semantics exactly (reference_links): a link holds a reference to its
bond's scaled tensor; a merge adds the new tensor into that array *in place*
(Lattice.py:51), so every link sharing the array -- the bond's own strength
or invStrength -- carries the sum, and later skip tests see it; a float
(Ising) merge rebinds only that link.  The scaled values are built in the
reference's order (each bond scaled by 1/T, then added), and the Ising skip
test applies only to Python floats (Lattice.py:45: `type(strength) is
float`), i.e. not when T is a numpy scalar, as in loadMC's np.linspace
temperatures.  Lattices with repeats keep their bonds (link_bonds) so that
ising_args / on_args redo this at each T.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

# tensor component order used everywhere (Lattice.py:136, heisenbergLib.c:193-201)
J_ORDER = ("xx", "yy", "zz", "xy", "xz", "yz", "yx", "zx", "zy")
# index permutation giving the transpose in that order (Lattice.py:137)
J_TRANSPOSE = np.array([0, 1, 2, 6, 7, 8, 3, 4, 5])


def tensor9_to_mat(j9) -> np.ndarray:
    """[xx,yy,zz,xy,xz,yz,yx,zx,zy] -> 3x3 matrix M with E = s_i^T M s_j."""
    j = np.asarray(j9, dtype=np.float64)
    return np.array([[j[0], j[3], j[4]], [j[6], j[1], j[5]], [j[7], j[8], j[2]]])


def mat_to_tensor9(m) -> np.ndarray:
    m = np.asarray(m, dtype=np.float64)
    return np.array([m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2], m[1, 0], m[2, 0], m[2, 1]])


@dataclass
class Bond:
    """One bond of the input list (Lattice.py:123-153): source orbital,
    target orbital, cell offset and the 9-component coupling (energy units,
    not beta-scaled).  For Ising only J[0] (Jxx) is used."""

    source: int
    target: int
    overLat: Tuple[int, int, int]
    J: Sequence[float]

    def j9(self) -> np.ndarray:
        j = np.zeros(9)
        jj = np.atleast_1d(np.asarray(self.J, dtype=np.float64))
        j[: len(jj)] = jj
        return j


@dataclass
class Lattice:
    """Per-site tables of one lattice, unscaled couplings, reference order."""

    Lx: int
    Ly: int
    Lz: int
    norb: int
    n: int
    nlink: np.ndarray            # [n] int32
    nbr: np.ndarray              # [n, maxl] int32, -1 padded
    J: np.ndarray                # [n, maxl, 9] float64, 0 padded (unscaled)
    S: np.ndarray                # [n] spin length per site (signed, as given)
    D: np.ndarray                # [n, 3] single-ion anisotropy (unscaled)
    corr_pairs: np.ndarray       # [nLat, 2] int32
    rOrb: np.ndarray             # [nr] int32
    rOrbCluster: np.ndarray      # [nr, nclu] int32
    rlink: np.ndarray            # [nr, maxl] int32, -1 padded
    circuits: np.ndarray         # [nCirc, 3] int32
    groups: List[List[int]]      # orbital groups (site ids)
    pos: np.ndarray              # [n, 3] cartesian positions
    cell: np.ndarray = field(default=None)  # [n, 3] cell coords
    orb: np.ndarray = field(default=None)   # [n] orbital index
    link_bonds: Optional[list] = field(default=None)  # bonds, kept when some site links one target twice
    ising: bool = False

    @property
    def maxl(self) -> int:
        return int(self.nbr.shape[1])

    @property
    def nLat(self) -> int:
        return int(self.corr_pairs.shape[0])


_CLUSTER_OFFSETS = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1)]


def _group_ordered(site: np.ndarray, other: np.ndarray, seq: np.ndarray, n: int):
    """Stable grouping of insertions by site, ordered by seq within a site."""
    order = np.lexsort((seq, site))
    return order, np.bincount(site, minlength=n)


def reference_links(Lx: int, Ly: int, Lz: int, norb: int, bonds: Sequence[Bond], T, ising: bool):
    """The normal links of Lattice.establishLinking (Lattice.py:237-262) with
    the object semantics of Orbital.addLinking (Lattice.py:34-57), for a
    temperature T (the bonds scaled by (1/T) first, Lattice.py:141-147, with
    T as given: a numpy T makes Ising strengths numpy scalars).

    Returns (nbr [n, maxl] -1 padded, vals [n, maxl] list of float or 9-arrays
    as the flattening reads them after all merges).  Pure Python: used only
    for lattices where a site links one target twice."""
    n = Lx * Ly * Lz * norb
    objs = []   # float / numpy scalar (immutable) or ndarray (mutable, shared)
    s_ref, i_ref = [], []
    for b in bonds:
        j = b.j9()
        if ising:
            jx = float(j[0])
            objs.append((1 / T) * jx)
            s_ref.append(len(objs) - 1)
            objs.append((1 / T) * jx)
            i_ref.append(len(objs) - 1)
        else:
            objs.append((1 / T) * np.array([float(v) for v in j]))
            s_ref.append(len(objs) - 1)
            objs.append((1 / T) * np.array([float(v) for v in j[J_TRANSPOSE]]))
            i_ref.append(len(objs) - 1)
    links = [[] for _ in range(n)]

    def add(site, target, oi):
        new = objs[oi]
        for lk in links[site]:
            if lk[0] != target:
                continue
            old = objs[lk[1]]
            d = abs(old - new)
            if type(new) is float and d < 1e-5:
                return
            if type(new) is np.ndarray and sum(d) < 1e-5:
                return
            if isinstance(old, np.ndarray):
                old += new          # in place: every link holding this array sees the sum
            else:
                objs.append(old + new)
                lk[1] = len(objs) - 1
            return
        links[site].append([target, oi])

    for x in range(Lx):
        for y in range(Ly):
            for z in range(Lz):
                for o in range(norb):
                    src = ((x * Ly + y) * Lz + z) * norb + o
                    for bi, b in enumerate(bonds):
                        if o != b.source:
                            continue
                        dx, dy, dz = (int(v) for v in b.overLat)
                        tgt = ((((x + dx) % Lx) * Ly + (y + dy) % Ly) * Lz + (z + dz) % Lz) * norb + b.target
                        add(src, tgt, s_ref[bi])
                        if src != tgt:
                            add(tgt, src, i_ref[bi])
    maxl = max((len(lk) for lk in links), default=0)
    nbr = -np.ones((n, maxl), np.int32)
    vals = [[None] * maxl for _ in range(n)]
    for i, lk in enumerate(links):
        for l, (t, oi) in enumerate(lk):
            nbr[i, l] = t
            vals[i][l] = objs[oi]
    return nbr, vals


def build_lattice(Lx: int, Ly: int, Lz: int, LMatrix, pos, bonds: Sequence[Bond], S: Sequence[float],
                  D: Optional[Sequence[Sequence[float]]] = None, ki_s: int = 0, ki_t: int = 0,
                  ki_overLat=(0, 0, 0), orbGroupList: Sequence[Sequence[int]] = (), groupInSC: bool = False,
                  localCircuitList: Sequence = (), ising: bool = False, numpy_T: bool = False) -> Lattice:
    """Build the per-site tables the way Lattice.establishLattice +
    establishLinking do (see module docstring for the rules and citations).
    numpy_T: repeated Ising links merge as they do when the reference's T is a
    numpy scalar (loadMC), instead of as for a Python float."""
    LMatrix = np.asarray(LMatrix, dtype=np.float64)
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    norb = len(S)
    if len(pos) < norb:
        raise ValueError("not enough basis positions for the orbitals")
    if ki_s >= norb or ki_t >= norb:
        raise ValueError("index out of range ki_s=%d, ki_t=%d, norb=%d" % (ki_s, ki_t, norb))
    ncell = Lx * Ly * Lz
    n = ncell * norb
    L = np.array([Lx, Ly, Lz])
    c = np.arange(ncell)
    cx, cy, cz = c // (Ly * Lz), (c // Lz) % Ly, c % Lz
    cxyz = np.stack([cx, cy, cz], 1)

    def cell_of(dx, dy, dz):
        return (((cx + dx) % Lx) * Ly + (cy + dy) % Ly) * Lz + (cz + dz) % Lz

    nb = len(bonds)
    # ---- normal links: events ordered by (cell, orbital, bond index)
    src_l, tgt_l, j_l, seq_l = [], [], [], []
    for bi, b in enumerate(bonds):
        if b.source >= norb or b.target >= norb:
            raise ValueError("bond orbital index out of range")
        dx, dy, dz = (int(v) for v in b.overLat)
        src = c * norb + b.source
        tgt = cell_of(dx, dy, dz) * norb + b.target
        base = (c * norb + b.source) * nb + bi
        j9 = b.j9()
        if ising:
            j9 = np.array([j9[0]] + [0.0] * 8)
        jt = j9[J_TRANSPOSE] if not ising else j9
        src_l.append(src); tgt_l.append(tgt); j_l.append(np.broadcast_to(j9, (ncell, 9))); seq_l.append(2 * base)
        keep = src != tgt
        src_l.append(tgt[keep]); tgt_l.append(src[keep])
        j_l.append(np.broadcast_to(jt, (int(keep.sum()), 9))); seq_l.append(2 * base[keep] + 1)
    if nb:
        site = np.concatenate(src_l); other = np.concatenate(tgt_l)
        jj = np.concatenate(j_l).astype(np.float64); seq = np.concatenate(seq_l)
    else:
        site = np.zeros(0, np.int64); other = np.zeros(0, np.int64); jj = np.zeros((0, 9)); seq = np.zeros(0, np.int64)
    order = np.lexsort((seq, site))
    site, other, jj = site[order], other[order], jj[order].copy()
    # duplicates (same site, same target): keep first, skip or merge later ones
    key = site.astype(np.int64) * n + other
    _, first_idx, inv = np.unique(key, return_index=True, return_inverse=True)
    dup = np.ones(len(key), bool)
    dup[first_idx] = False
    keep = ~dup
    repeats = bool(dup.any())
    if repeats:
        # some site links one target twice: the reference's merge / skip with
        # its in-place tensor additions, unscaled (T = 1)
        nbr, vals = reference_links(Lx, Ly, Lz, norb, bonds, np.float64(1.0) if numpy_T else 1.0, ising)
        maxl = nbr.shape[1]
        nlink = (nbr >= 0).sum(1).astype(np.int32)
        J = np.zeros((n, maxl, 9))
        for i in range(n):
            for l in range(int(nlink[i])):
                v = vals[i][l]
                if ising:
                    J[i, l, 0] = float(v)
                else:
                    J[i, l] = v
    else:
        site, other, jj = site[keep], other[keep], jj[keep]
        nlink = np.bincount(site, minlength=n).astype(np.int32)
        maxl = int(nlink.max()) if n else 0
        start = np.concatenate([[0], np.cumsum(nlink)[:-1]])
        slot = np.arange(len(site)) - start[site]
        nbr = -np.ones((n, max(maxl, 1)), np.int32)
        J = np.zeros((n, max(maxl, 1), 9))
        nbr[site, slot] = other
        J[site, slot] = jj
        if maxl == 0:
            nbr = nbr[:, :0]; J = J[:, :0]

    # ---- block-spin sites: chosen cells (all coords even), Lattice.py:185-205
    chosen_cell = (cx % 2 + cy % 2 + cz % 2) == 0
    cc = np.nonzero(chosen_cell)[0]
    rOrb = (cc[:, None] * norb + np.arange(norb)[None, :]).reshape(-1)
    clusters = []
    for o in range(norb):
        rows = []
        for (ox, oy, oz) in _CLUSTER_OFFSETS:
            rows.append(cell_of(ox, oy, oz)[cc] * norb + o)
        clusters.append(np.stack(rows, 1))
    # dedupe preserving order (identical pattern for every cell)
    clu = np.stack(clusters, 1).reshape(len(rOrb), len(_CLUSTER_OFFSETS)) if len(rOrb) else np.zeros((0, 8), int)
    if len(rOrb):
        uniq_cols = []
        probe = clu[0]
        seen = set()
        for k, v in enumerate(probe):
            if v not in seen:
                seen.add(v); uniq_cols.append(k)
        clu = clu[:, uniq_cols]
    # rnorm links: events (chosen cell, orbital, bond), target at 2x offset
    rs, rt, rq = [], [], []
    for bi, b in enumerate(bonds):
        dx, dy, dz = (int(v) for v in b.overLat)
        src = cc * norb + b.source
        tgt = cell_of(2 * dx, 2 * dy, 2 * dz)[cc] * norb + b.target
        base = (cc * norb + b.source) * nb + bi
        rs.append(src); rt.append(tgt); rq.append(2 * base)
        k2 = src != tgt
        rs.append(tgt[k2]); rt.append(src[k2]); rq.append(2 * base[k2] + 1)
    rlink = -np.ones((len(rOrb), max(maxl, 1)), np.int32)
    if nb and len(rOrb):
        rsite = np.concatenate(rs); roth = np.concatenate(rt); rseq = np.concatenate(rq)
        o2 = np.lexsort((rseq, rsite))
        rsite, roth = rsite[o2], roth[o2]
        rkey = rsite.astype(np.int64) * n + roth
        _, rfirst = np.unique(rkey, return_index=True)
        m = np.zeros(len(rkey), bool); m[rfirst] = True
        rsite, roth = rsite[m], roth[m]
        rcount = np.bincount(rsite, minlength=n)
        rstart = np.concatenate([[0], np.cumsum(rcount)[:-1]])
        rslot = np.arange(len(rsite)) - rstart[rsite]
        rank = -np.ones(n, np.int64); rank[rOrb] = np.arange(len(rOrb))
        sel = (rank[rsite] >= 0) & (rslot < rlink.shape[1])
        rlink[rank[rsite[sel]], rslot[sel]] = roth[sel]
    if maxl == 0:
        rlink = rlink[:, :0]

    # ---- correlation pairs, circuits, groups
    kx, ky, kz = (int(v) for v in ki_overLat)
    corr = np.stack([c * norb + ki_s, cell_of(kx, ky, kz) * norb + ki_t], 1).astype(np.int32)
    circ = []
    for circuit in localCircuitList:
        (o1, l1), (o2, l2), (o3, l3) = circuit
        circ.append(np.stack([cell_of(*l1) * norb + o1, cell_of(*l2) * norb + o2, cell_of(*l3) * norb + o3], 1))
    circuits = (np.stack(circ, 1).reshape(-1, 3) if circ else np.zeros((0, 3), int)).astype(np.int32)
    groups: List[List[int]] = []
    for sub in orbGroupList:
        if groupInSC:
            groups.append([int(cell * norb + o) for o in sub for cell in c])
        else:
            groups.append([int(o) for o in sub])

    Sarr = np.repeat(np.asarray(S, np.float64)[None, :], ncell, 0).reshape(-1)
    if D is None:
        D = [[0.0, 0.0, 0.0]] * norb
    Darr = np.repeat(np.asarray(D, np.float64).reshape(norb, 3)[None], ncell, 0).reshape(-1, 3)
    orbv = np.tile(np.arange(norb), ncell)
    cellv = np.repeat(cxyz, norb, 0)
    posv = (cellv + pos[orbv]) @ LMatrix
    return Lattice(Lx, Ly, Lz, norb, n, nlink, nbr, J, Sarr, Darr, corr, rOrb.astype(np.int32),
                   clu.astype(np.int32), rlink, circuits, groups, posv, cellv, orbv,
                   link_bonds=list(bonds) if repeats else None, ising=ising)


def dipolar_bonds(lat: Lattice, alpha: float, cutoff: Optional[float] = None, ising: bool = False,
                  max_links: int = 1 << 24) -> Lattice:
    """Dipolar couplings as extra bond entries, truncated at |r| <= cutoff.

    What Lattice.generateDipoleBondings (Lattice.py:286-308) sets out to add:
    for every ordered pair of distinct sites (i, j), with r = r_i - r_j the
    Cartesian separation and open boundaries (no periodic images), a link
    i -> j of strength alpha/|r|^3 (Ising: only this isotropic part) or of
    tensor alpha/|r|^3 (I - 3 r^ r^T) (O(n)), appended after the site's other
    links even when i already links j.  The reference's call lacks an
    argument and raises TypeError, and its O(n) branch exits, so only the
    intent can be followed.  cutoff=None keeps every pair (N(N-1) links);
    the energy is then the full open-boundary dipole sum
    sum_{i<j} alpha/|r|^3 (s_i.s_j - 3 (s_i.r^)(s_j.r^)).
    """
    n = lat.n
    pos = np.asarray(lat.pos, np.float64)
    rc2 = np.inf if cutoff is None else float(cutoff) ** 2
    src_l, tgt_l, j_l = [], [], []
    total = 0
    chunk = max(1, (1 << 22) // max(n, 1))
    for a in range(0, n, chunk):
        r = pos[a:a + chunk, None, :] - pos[None, :, :]          # [c, n, 3]
        d2 = np.einsum("ijk,ijk->ij", r, r)
        ii = np.arange(a, min(a + chunk, n))
        ok = d2 <= rc2
        ok[np.arange(len(ii)), ii] = False                      # no self link
        si, tj = np.nonzero(ok)
        if si.size == 0:
            continue
        total += si.size
        if total > max_links:
            raise ValueError("dipolar couplings would add %d+ links; give a cutoff" % total)
        rv = r[si, tj]
        dd = np.sqrt(d2[si, tj])
        a3 = alpha / dd ** 3
        j9 = np.zeros((si.size, 9))
        if ising:
            j9[:, 0] = a3
        else:
            u = rv / dd[:, None]
            m = a3[:, None, None] * (np.eye(3)[None] - 3.0 * u[:, :, None] * u[:, None, :])
            j9[:, 0], j9[:, 1], j9[:, 2] = m[:, 0, 0], m[:, 1, 1], m[:, 2, 2]
            j9[:, 3], j9[:, 4], j9[:, 5] = m[:, 0, 1], m[:, 0, 2], m[:, 1, 2]
            j9[:, 6], j9[:, 7], j9[:, 8] = m[:, 1, 0], m[:, 2, 0], m[:, 2, 1]
        src_l.append(ii[si]); tgt_l.append(tj); j_l.append(j9)
    if not src_l:
        return lat
    site = np.concatenate(src_l); other = np.concatenate(tgt_l); jj = np.concatenate(j_l)
    order = np.lexsort((other, site))                            # target order within a site
    site, other, jj = site[order], other[order], jj[order]
    extra = np.bincount(site, minlength=n)
    nlink = lat.nlink.astype(np.int64) + extra
    maxl = int(nlink.max())
    nbr = -np.ones((n, maxl), np.int32)
    J = np.zeros((n, maxl, 9))
    nbr[:, :lat.maxl] = lat.nbr
    J[:, :lat.maxl] = lat.J
    start = np.concatenate([[0], np.cumsum(extra)[:-1]])
    slot = lat.nlink[site] + (np.arange(len(site)) - start[site])
    nbr[site, slot] = other
    J[site, slot] = jj
    rlink = -np.ones((lat.rlink.shape[0], maxl), np.int32)
    rlink[:, :lat.rlink.shape[1]] = lat.rlink
    out = Lattice(**{k: getattr(lat, k) for k in lat.__dataclass_fields__})
    out.link_bonds = None  # the dipolar links are not in the bond list
    out.nlink, out.nbr, out.J, out.rlink = nlink.astype(np.int32), nbr, J, rlink
    return out


def clamp_T(T: float) -> float:
    """mcMain.py:21 — temperatures below 0.1 run at 0.1."""
    return 0.1 if T < 0.1 else T


def ising_args(lat: Lattice, T: float, h: float, nthermal: int, nsweep: int, ninterval: int = -1,
               algorithm: str = "Metropolis", spinFrame: int = 0) -> tuple:
    """Positional arguments of isinglib.MCMainFunction exactly as
    MC.mainLoopViaCLib builds them (mcMain.py:51-124)."""
    T = clamp_T(T)
    n = lat.n
    ninterval = n if ninterval <= 0 else ninterval
    nbr, nlink, flat = _scaled_links(lat, T, True)
    return (1 if algorithm == "Wolff" else 0, tuple(float(v) for v in lat.S), int(nthermal), int(nsweep),
            int(ninterval), nbr.shape[1], tuple(int(v) for v in nlink), flat,
            tuple(int(v) for v in nbr.reshape(-1)), tuple(int(v) for v in lat.corr_pairs.reshape(-1)),
            h / T, tuple(int(v) for v in lat.rOrb), tuple(int(v) for v in lat.rOrbCluster.reshape(-1)),
            tuple(int(v) for v in _pad(lat.rlink, nbr.shape[1]).reshape(-1)), int(spinFrame), None)


def _pad(a: np.ndarray, width: int) -> np.ndarray:
    """-1 padded to `width` columns (maxNLinking; mcMain.py:81-87, 209-215)."""
    out = -np.ones((a.shape[0], width), np.int32)
    w = min(width, a.shape[1])
    out[:, :w] = a[:, :w]
    return out


def _scaled_links(lat: Lattice, T, ising: bool):
    """(nbr, nlink, flattened beta-scaled couplings) as mcMain.py:60-78 /
    162-174 read them at temperature T: (1/T)*J per component, or, for a
    lattice with repeated targets, the reference's merge semantics at T."""
    if lat.link_bonds is not None:
        nbr, vals = reference_links(lat.Lx, lat.Ly, lat.Lz, lat.norb, lat.link_bonds, T, ising)
        nlink = (nbr >= 0).sum(1)
        flat = []
        for i in range(lat.n):
            for l in range(nbr.shape[1]):
                v = vals[i][l]
                if ising:
                    flat.append(float(v) if v is not None else 0.0)
                elif v is None:
                    flat += [0.0] * 9
                else:
                    flat += [float(x) for x in v]
        return nbr, nlink, tuple(flat)
    if ising:
        Js = (1 / T) * lat.J[:, :, 0]
        Js = np.where(lat.nbr >= 0, Js, 0.0)
    else:
        Js = (1 / T) * lat.J
        Js = np.where((lat.nbr >= 0)[:, :, None], Js, 0.0)
    return lat.nbr, lat.nlink, tuple(float(v) for v in Js.reshape(-1))


def on_args(lat: Lattice, T: float, h: float, nthermal: int, nsweep: int, ninterval: int = -1,
            algorithm: str = "Metropolis", spinFrame: int = 0, flunc: float = 0.0) -> tuple:
    """Positional arguments of {xylib,heisenberglib}.MCMainFunction exactly
    as MC.mainLoopViaCLib_On builds them (mcMain.py:137-248)."""
    T = clamp_T(T)
    n = lat.n
    ninterval = n if ninterval <= 0 else ninterval
    nbr, nlink, flat = _scaled_links(lat, T, False)
    maxl = nbr.shape[1]
    Js = np.asarray(flat, np.float64).reshape(n, maxl, 9)
    valid = nbr >= 0
    ignoreNonDiagonalJ = 0 if np.any(np.abs(Js[valid][:, 3:]) > 1e-6) else 1
    DT = lat.D / T
    ng = len(lat.groups)
    maxg = 1
    if ng > 0:
        maxg = len(lat.groups[0])
    if ng > 1:
        maxg = int(np.max([len(g) for g in lat.groups]))
    glist = []
    for g in lat.groups:
        glist += list(g) + [-1] * (maxg - len(g))
    return (1 if algorithm == "Wolff" else 0, tuple(float(v) for v in lat.S), tuple(float(v) for v in DT.reshape(-1)),
            int(nthermal), int(nsweep), int(ninterval), maxl, tuple(int(v) for v in nlink), flat,
            tuple(int(v) for v in nbr.reshape(-1)),
            tuple(int(v) for v in lat.circuits.reshape(-1)), tuple(int(v) for v in lat.corr_pairs.reshape(-1)),
            ng, maxg, tuple(glist), float(flunc), h / T, tuple(int(v) for v in lat.rOrb),
            tuple(int(v) for v in lat.rOrbCluster.reshape(-1)), tuple(int(v) for v in _pad(lat.rlink, maxl).reshape(-1)),
            int(spinFrame), ignoreNonDiagonalJ, None)


# ----------------------------------------------------------------------------
# Standard lattices used by the workloads
# ----------------------------------------------------------------------------

def square(L: int, J=-1.0, S=1.0, D=(0.0, 0.0, 0.0), ising=False, **kw) -> Lattice:
    Jv = [J, J, J] if not ising else [J]
    bonds = [Bond(0, 0, (1, 0, 0), Jv), Bond(0, 0, (0, 1, 0), Jv)]
    return build_lattice(L, L, 1, np.eye(3), [[0, 0, 0]], bonds, [S], [list(D)], ising=ising, **kw)


def cubic(L: int, J=-1.0, S=1.0, D=(0.0, 0.0, 0.0), ising=False, **kw) -> Lattice:
    Jv = [J, J, J] if not ising else [J]
    bonds = [Bond(0, 0, (1, 0, 0), Jv), Bond(0, 0, (0, 1, 0), Jv), Bond(0, 0, (0, 0, 1), Jv)]
    return build_lattice(L, L, L, np.eye(3), [[0, 0, 0]], bonds, [S], [list(D)], ising=ising, **kw)
