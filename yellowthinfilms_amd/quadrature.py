"""k-rho quadrature rules for the Sommerfeld integral (host side).

  * gl_rule -- composite Gauss-Legendre over panels, with the u^2
    substitution on a panel that ends on a branch point (mod_KxGrid.jl:361-384);
  * evanescent_umax, evanescent_band_nodes -- the cosh-substituted tail
    kmax -> infinity (mod_KxGrid.jl:126-145);
  * light_lines -- the isotropic/uniaxial in-plane light lines of a stack
    (sqrt(Re eps_t) omega, the breaks of kρ_grid, Gz_ClosedForm_TestHelpers.jl:10-24);
  * layer_light_lines, dedup, decay_cutoff, tail_bound, graded_panels,
    halve_panels, build_kphi_grid -- the pieces of the feature-driven rule
    (mod_KxGrid.jl:41-114, 331-394).

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


# ---------------------------------------------------------------------------
# Feature-driven rule pieces (mod_KxGrid.jl)
# ---------------------------------------------------------------------------
def dedup(ks, rtol=1e-9):
    """_dedup (mod_KxGrid.jl:83-89): values closer than rtol (relative) merged,
    keeping the first of each cluster."""
    out = []
    for k in ks:
        if not out or k - out[-1] > rtol * max(abs(k), 1e-300):
            out.append(float(k))
    return out


def layer_light_lines(layer, omega, phi):
    """layer_light_lines (mod_KxGrid.jl:41-71): the real-axis branch points
    k_rho = Re sqrt(q) of the layer along in-plane direction phi, q the roots of
    det(omega^2 eps + k_rho^2 K mu^-1 K) = 0 (K the cross-product matrix of the
    unit in-plane direction), a quadratic in q read off three evaluations;
    roots with Re q <= 0 have no light line."""
    eps = np.asarray(layer.eps.eps, np.complex128)
    mu = np.asarray(layer.mu.eps, np.complex128)
    c, s = math.cos(phi), math.sin(phi)
    K = np.array([[0, 0, s], [0, 0, -c], [-s, c, 0]], np.complex128)
    M = K @ np.linalg.inv(mu) @ K
    A = float(omega) ** 2 * eps
    sc = float(omega) ** 2 * max(np.max(np.abs(eps)) * np.max(np.abs(mu)), np.finfo(float).eps)
    p0 = np.linalg.det(A)
    p1 = np.linalg.det(A + sc * M)
    p2 = np.linalg.det(A + 2 * sc * M)
    c2 = (p2 - 2 * p1 + p0) / (2 * sc ** 2)
    c1 = (p1 - p0) / sc - c2 * sc
    c0 = p0
    if abs(c2) * sc ** 2 <= 1e-12 * (abs(c1) * sc + abs(c0)):
        roots = [-c0 / c1] if abs(c1) > 0 else []
    else:
        D = c1 ** 2 - 4 * c2 * c0
        Dl = 0.0 if abs(D) <= 1e-10 * abs(c1) ** 2 else np.sqrt(complex(D))
        roots = [(-c1 + Dl) / (2 * c2), (-c1 - Dl) / (2 * c2)]
    return dedup(sorted(np.sqrt(complex(q)).real for q in roots if complex(q).real > 0), rtol=1e-7)


def decay_cutoff(h, target_decay=25.0):
    """decay_cutoff (mod_KxGrid.jl:100): k_rho where e^{-k_rho h} has decayed by target_decay e-folds."""
    return float(target_decay) / float(h)


def tail_bound(K, kmax, krho_cap=math.inf, kx_ratio_min=80.0):
    """tail_bound (mod_KxGrid.jl:113-114): at least kx_ratio_min kmax, at most krho_cap."""
    return min(max(float(K), float(kx_ratio_min) * kmax), float(krho_cap))


def graded_panels(breaks, width, alpha=0.5, max_panels=10 ** 7):
    """graded_panels (mod_KxGrid.jl:331-350): panels ending on every break, each
    no wider than width(k) anywhere inside it (width Lipschitz with constant <=
    alpha: a step width(a) / (1 + alpha) from a satisfies that)."""
    bs = sorted(set(float(b) for b in breaks))
    panels = []
    for a, b in zip(bs[:-1], bs[1:]):
        x = a
        while x < b:
            s = width(x) / (1 + alpha)
            if not s > 0:
                raise ValueError(f"graded_panels: non-positive width {s} at k={x}")
            y = b if x + s >= b - 1e-14 * abs(b) else x + s
            panels.append((x, y))
            if len(panels) > max_panels:
                raise ValueError(f"graded_panels: more than {max_panels} panels; check the width function")
            x = y
    return panels


def halve_panels(panels):
    """halve_panels (mod_KxGrid.jl:387-388): every panel split in two (the self-check's rule)."""
    return [p for (a, b) in panels for p in ((float(a), (a + b) / 2), ((a + b) / 2, float(b)))]


def build_kphi_grid(n_phi):
    """build_kphi_grid (mod_KxGrid.jl:390-394): uniform phi nodes and weights on [0, 2 pi)."""
    return 2 * math.pi * np.arange(n_phi) / n_phi, np.full(n_phi, 2 * math.pi / n_phi)
