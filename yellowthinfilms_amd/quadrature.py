"""k-rho quadrature rules for the Sommerfeld integral (host side).

  * gl_rule -- composite Gauss-Legendre over panels, with the u^2
    substitution on a panel that ends on a branch point (mod_KxGrid.jl:361-384);
  * evanescent_umax, evanescent_band_nodes -- the cosh-substituted tail
    kmax -> infinity (mod_KxGrid.jl:126-145);
  * light_lines -- the isotropic/uniaxial in-plane light lines of a stack
    (sqrt(Re eps_t) omega, the breaks of kρ_grid, Gz_ClosedForm_TestHelpers.jl:10-24).

Gauss-Legendre nodes come from numpy (the reference uses FastGaussQuadrature;
both are the textbook rule to rounding).
"""
import math

import numpy as np


def gauss_legendre(n):
    return np.polynomial.legendre.leggauss(int(n))


def gl_rule(panels, order=8, singular=()):
    """Composite GL over `panels` [(a, b), ...]; on a panel with an end in
    `singular`, k = L + sg u^2 (dk = 2u du), u in [0, sqrt(b - a)]."""
    x, w = gauss_legendre(order)
    S = set(float(s) for s in singular)
    nodes, wts = [], []
    for a, b in panels:
        a, b = float(a), float(b)
        if a in S or b in S:
            L, sg = (a, 1.0) if a in S else (b, -1.0)
            U = math.sqrt(b - a)
            u = 0.5 * U * (1 + x)
            nodes.append(L + sg * u ** 2)
            wts.append(0.5 * U * w * 2 * u)
        else:
            h, c = 0.5 * (b - a), 0.5 * (a + b)
            nodes.append(c + h * x)
            wts.append(h * w)
    if not nodes:
        return np.zeros(0), np.zeros(0)
    return np.concatenate(nodes), np.concatenate(wts)


def evanescent_umax(kmax, Leff, target_decay=25.0, kx_ratio_max=80.0):
    """u_max of the tail k = kmax cosh(u): decay exp(-kmax Leff sinh u) reaches
    `target_decay`, capped at cosh(u) = kx_ratio_max."""
    umax_phys = math.asinh(target_decay / (kmax * Leff))
    umax_cap = math.acosh(max(kx_ratio_max, 1.0 + 1e-10))
    return min(umax_phys, umax_cap)


def evanescent_band_nodes(kmax, umax, N, delta=1e-6):
    """N-point GL band on u in [delta, umax], k = kmax cosh u, w = wu kmax sinh u."""
    if umax <= delta:
        return np.zeros(0), np.zeros(0)
    x, w = gauss_legendre(N)
    u = (umax - delta) / 2 * x + (umax + delta) / 2
    wu = w * (umax - delta) / 2
    return kmax * np.cosh(u), wu * (kmax * np.sinh(u))


def light_lines(stack, omega):
    """sqrt(Re eps_t) * omega of every layer (transverse diagonal entries)."""
    out = set()
    for l in stack:
        e = np.asarray(l.eps.eps)
        for i in range(2):
            out.add(float(np.sqrt(complex(e[i, i])).real * omega))
    return sorted(out)


def panel_rule(breaks, n_nodes, order=16, singular=()):
    """About `n_nodes` GL nodes on [breaks[0], breaks[-1]]: each interval
    between breaks gets panels in proportion to its length (at least one),
    `order` nodes per panel, u^2 substitution at the ends in `singular`."""
    br = sorted(set(float(b) for b in breaks))
    total = br[-1] - br[0]
    npan = max(len(br) - 1, int(round(n_nodes / order)))
    counts = [max(1, int(round(npan * (b - a) / total))) for a, b in zip(br[:-1], br[1:])]
    # trim or pad the longest intervals so the panel count is exact
    while sum(counts) > npan:
        i = int(np.argmax(counts))
        counts[i] -= 1
    while sum(counts) < npan:
        i = int(np.argmax([(b - a) / c for (a, b), c in zip(zip(br[:-1], br[1:]), counts)]))
        counts[i] += 1
    pans = []
    for (a, b), c in zip(zip(br[:-1], br[1:]), counts):
        edges = np.linspace(a, b, c + 1)
        edges[0], edges[-1] = a, b
        pans += [(edges[j], edges[j + 1]) for j in range(c)]
    return gl_rule(pans, order=order, singular=singular)
