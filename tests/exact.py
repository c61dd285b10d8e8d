"""Closed forms and quadratures used to pin the oracle and the engine.

* 2D Ising on the L x L torus: Kaufman (1949) exact partition function in the
  form of Beale, PRL 76, 78 (1996).  Pinned against brute-force enumeration
  of the 4x4 torus (tests/test_exact.py).
* 1D rings, N -> infinity limits: Ising <s s> = tanh K; XY <cos> = I1(K)/I0(K);
  Heisenberg <s.s> = coth K - 1/K  (K = beta J).
* single classical spin with anisotropy and field; two-spin dimer with an
  arbitrary 3x3 coupling: direct numerical quadrature over the sphere(s).
Sign convention of the reference: E = sum_bonds s_i^T J s_j + sum D s^2 - h s,
so a ferromagnet has J < 0.
"""
from __future__ import annotations

import numpy as np
from scipy import special


def ising_logz(L: int, K: float) -> float:
    """ln Z of the L x L periodic Ising ferromagnet, E = -sum s s, K = beta."""
    k = np.arange(2 * L)
    c2 = np.cosh(2 * K) / np.tanh(2 * K) - np.cos(np.pi * k / L)
    gam = np.arccosh(np.maximum(c2, 1.0))
    gam[0] = 2 * K + np.log(np.tanh(K))
    odd = gam[1::2][:L]
    even = gam[0::2][:L]

    def lc(x):  # ln(2 cosh(L x / 2))
        y = L * np.abs(x) / 2
        return np.sum(y + np.log1p(np.exp(-2 * y)))

    def ls(x):  # ln|2 sinh(L x / 2)| and sign
        y = L * x / 2
        sign = np.prod(np.sign(y))
        ay = np.abs(y)
        return np.sum(ay + np.log(-np.expm1(-2 * ay))), sign

    t1 = lc(odd)
    t2, s2 = ls(odd)
    t3 = lc(even)
    t4, s4 = ls(even)
    terms = np.array([t1, t2, t3, t4])
    signs = np.array([1.0, s2, 1.0, s4])
    m = terms.max()
    tot = np.sum(signs * np.exp(terms - m))
    return np.log(0.5) + (L * L / 2) * np.log(2 * np.sinh(2 * K)) + m + np.log(tot)


def ising_e_c(L: int, T: float, h: float = 1e-4):
    """Energy and specific heat per site of the L x L torus at temperature T
    (J = 1): e = -d lnZ/d beta / N, c = beta^2 d^2 lnZ / d beta^2 / N."""
    b = 1.0 / T
    N = L * L
    f = [ising_logz(L, b + d * h) for d in (-2, -1, 0, 1, 2)]
    d1 = (f[0] - 8 * f[1] + 8 * f[3] - f[4]) / (12 * h)
    d2 = (-f[0] + 16 * f[1] - 30 * f[2] + 16 * f[3] - f[4]) / (12 * h * h)
    return -d1 / N, b * b * d2 / N


def ising_brute_logz(L: int, K: float) -> float:
    N = L * L
    st = ((np.arange(2 ** N)[:, None] >> np.arange(N)) & 1) * 2 - 1
    s = st.reshape(-1, L, L)
    e = -(s * np.roll(s, 1, 1)).sum((1, 2)) - (s * np.roll(s, 1, 2)).sum((1, 2))
    return np.log(np.sum(np.exp(-K * e - np.max(-K * e)))) + np.max(-K * e)


def ring_bond_energy(model: str, K: float) -> float:
    """<s_i . s_j> per bond of the infinite ferro chain, K = beta |J|."""
    if model == "ising":
        return np.tanh(K)
    if model == "xy":
        return special.i1(K) / special.i0(K)
    if model == "heis":
        return 1 / np.tanh(K) - 1 / K
    raise ValueError(model)


def single_spin(D, h, beta, nc=3, S=1.0, n=160):
    """<E>, <s_axis> of one spin of length S with E = sum D_a s_a^2 - h s_axis
    (axis z for nc = 3, x for nc = 2), by quadrature (Gauss-Legendre in
    cos theta, uniform in phi)."""
    D = np.asarray(D, float)
    if nc == 3:
        ct, wc = np.polynomial.legendre.leggauss(n)
        ph = np.linspace(0, 2 * np.pi, 2 * n, endpoint=False)
        C, P = np.meshgrid(ct, ph, indexing="ij")
        st = np.sqrt(1 - C * C)
        s = S * np.stack([st * np.cos(P), st * np.sin(P), C])
        ax = 2
        w = np.outer(wc, np.ones_like(ph))
    else:
        ph = np.linspace(0, 2 * np.pi, 8 * n, endpoint=False)
        s = S * np.stack([np.cos(ph), np.sin(ph), 0 * ph])
        ax = 0
        w = np.ones_like(ph)
    E = D[0] * s[0] ** 2 + D[1] * s[1] ** 2 + D[2] * s[2] ** 2 - h * s[ax]
    wt = w * np.exp(-beta * (E - E.min()))
    Z = wt.sum()
    return (wt * E).sum() / Z, (wt * s[ax]).sum() / Z


def dimer(Jm, beta, n=36):
    """<s1^T J s2> and <s1 x s2> of two unit spins coupled by E = s1^T J s2
    (both link directions of the pair, J and J^T, summed once), quadrature
    over both spheres (Gauss-Legendre in cos theta, uniform in phi)."""
    Jm = np.asarray(Jm, float)
    x, wx = np.polynomial.legendre.leggauss(n)
    ph = np.linspace(0, 2 * np.pi, 2 * n, endpoint=False)
    C, P = np.meshgrid(x, ph, indexing="ij")
    W = np.outer(wx, np.ones_like(ph)).reshape(-1)
    st = np.sqrt(1 - C * C)
    s = np.stack([st * np.cos(P), st * np.sin(P), C]).reshape(3, -1)
    E = np.einsum("ai,ab,bj->ij", s, Jm, s)
    wt = np.outer(W, W) * np.exp(-beta * (E - E.min()))
    Z = wt.sum()
    cr = np.cross(s.T[:, None, :], s.T[None, :, :])
    return (wt * E).sum() / Z, np.einsum("ij,ijk->k", wt, cr) / Z
