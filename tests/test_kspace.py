"""The real-axis pole search (mod_KSpaceFeatures.jl) on the CPU: the oracle's
cavity response and step-by-step real_axis_poles pinned by closed forms and
by the reference's own numbers, and the host logic of the product module
(clusters, material classes, engine choice).

Pins:
  * a metal half-space under vacuum has the surface-plasmon pole
    k = k0 sqrt(eps / (1 + eps)): the scan finds Re k to 2e-4 and a
    half-width within 10% of Im k (the response is a Lorentzian of width
    Im k on a smooth background);
  * the `base` stack of comparisons/gila_airy (README.md:57-71, stack of
    plane_wave_slice.jl:34-52) has its strongest reflection pole at
    1.0624 k0 and a second one at 1.66 k0 (the reference's numbers, to the
    digits it prints)."""
import math

import numpy as np
import pytest

from helpers import po, ostack
from yellowthinfilms_amd.materials import (IsotropicMaterial, UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial,
                                           FlatOpticalLayer)

W0 = 2 * math.pi


def spp_stack(em):
    return [FlatOpticalLayer(IsotropicMaterial(1.0), 1.0), FlatOpticalLayer(IsotropicMaterial(em), 1.0)]


def base_stack():
    films = [(2.25, 1 / 4), (4.0, 3 / 16), (2.0 + 0.1j, 1 / 8)]  # bottom -> top
    return ([FlatOpticalLayer(IsotropicMaterial(1.0), 0.0)] +
            [FlatOpticalLayer(IsotropicMaterial(e), d) for e, d in reversed(films)] +
            [FlatOpticalLayer(IsotropicMaterial(1.0), 4 / 64)])


@pytest.mark.parametrize("em", [-4 + 0.1j, -2.5 + 0.05j, -9 + 0.5j])
def test_oracle_surface_plasmon_pole(em):
    poles = po.real_axis_poles(ostack(spp_stack(em)), W0, 1, k_lo=0.5 * W0, k_hi=6 * W0, exclude=[W0])
    k = W0 * np.sqrt(em / (1 + em))
    assert len(poles) == 1, poles
    q, G = poles[0][0], poles[0][1]
    assert abs(q - k.real) <= 2e-4 * k.real, (q / W0, k / W0)
    assert abs(G - k.imag) <= 0.1 * k.imag, (G / W0, k.imag / W0)


def test_oracle_base_stack_poles_of_the_reference():
    poles = po.real_axis_poles(ostack(base_stack()), W0, 1, k_lo=0.5 * W0, k_hi=3 * W0,
                               exclude=[W0, 1.5 * W0, 2 * W0, math.sqrt(2.0) * W0])
    q = sorted(p[0] / W0 for p in poles)
    strongest = sorted(poles, key=lambda p: -p[3])[:2]
    assert round(strongest[0][0] / W0, 4) == 1.0624
    assert round(strongest[1][0] / W0, 2) == 1.66
    assert all(1.0 < x < 3.0 for x in q)


def test_oracle_response_of_a_homogeneous_stack():
    """No interface reflects: RL = RR = 0 to round-off, so the response is 0
    for the top layer and exactly sigma(inv(I)) = 1 inside the stack."""
    st = [FlatOpticalLayer(IsotropicMaterial(2.0 + 0.01j), 1.0)] * 3
    k = np.linspace(0.1, 5, 50) * W0
    assert np.max(po.cavity_response(ostack(st), W0, 1, k, 0 * k)) < 1e-12
    assert np.max(np.abs(po.cavity_response(ostack(st), W0, 2, k, 0.3 * k) - 1.0)) < 1e-12


def test_clusters_and_material_classes():
    from yellowthinfilms_amd.kspace import clusters, material_class, choose_engine
    poles = [dict(q=1.0, Gamma=0.01, phi=0.0, prom=3.0), dict(q=1.02, Gamma=0.005, phi=0.1, prom=2.0),
             dict(q=2.0, Gamma=0.02, phi=0.0, prom=1.0)]
    cl = clusters(poles, 0.05)
    assert dict(a=1.0, b=1.02, Gamma=0.005) in cl and dict(a=2.0, b=2.0, Gamma=0.02) in cl
    assert not any(c["a"] == 1.02 and c["b"] == 2.0 for c in cl)
    L = lambda m: FlatOpticalLayer(m, 1.0)  # noqa: E731
    assert material_class(L(IsotropicMaterial(2.25))) == "dielectric"
    assert material_class(L(IsotropicMaterial(-4 + 0.1j))) == "metallic"
    assert material_class(L(UniaxialMaterial(4.0, -3.2))) == "hyperbolic"
    assert material_class(L(BiaxialMaterial(2.0, -1.0, 3.0))) == "hyperbolic"
    gy = np.array([[3, 0.3j, 0], [-0.3j, 3, 0], [0, 0, 2.8]])
    assert material_class(L(GyrotropicMaterial(gy))) == "gyrotropic"
    assert choose_engine([L(IsotropicMaterial(1.0)), L(IsotropicMaterial(2.25))]) == "adaptive"
    assert choose_engine([L(IsotropicMaterial(1.0)), L(UniaxialMaterial(4.0, -3.2))]) == "fixed"
    assert choose_engine([L(IsotropicMaterial(1.0))], "fixed") == "fixed"
    with pytest.raises(ValueError):
        choose_engine([L(IsotropicMaterial(1.0))], "bogus")


# ---------------------------------------------------------------------------
# the feature-driven rule (mod_KxGrid.jl, fixed_kρ_rule)
# ---------------------------------------------------------------------------
def test_layer_light_lines_closed_forms():
    """Light lines from the determinant (mod_KxGrid.jl:41-71) against the
    closed forms: isotropic omega Re sqrt(eps); uniaxial TE omega sqrt(eps_t)
    and TM omega sqrt(eps_z); diagonal biaxial along x: omega sqrt(eps_y),
    omega sqrt(eps_z), along y: omega sqrt(eps_x), omega sqrt(eps_z); a metal or
    the negative axis of a hyperbolic layer: none."""
    from yellowthinfilms_amd.quadrature import layer_light_lines
    L = lambda m: FlatOpticalLayer(m, 1.0)  # noqa: E731
    r = lambda v: [x / W0 for x in v]  # noqa: E731
    assert np.allclose(r(layer_light_lines(L(IsotropicMaterial(2.25 + 0.01j)), W0, 0.3)),
                       [np.sqrt(2.25 + 0.01j).real], rtol=1e-9)
    assert np.allclose(r(layer_light_lines(L(UniaxialMaterial(2.0, 3.5)), W0, 0.7)), [math.sqrt(2), math.sqrt(3.5)],
                       rtol=1e-9)
    bi = L(BiaxialMaterial(2.0, 3.0, 4.0))
    assert np.allclose(r(layer_light_lines(bi, W0, 0.0)), [math.sqrt(3), 2.0], rtol=1e-9)
    assert np.allclose(r(layer_light_lines(bi, W0, math.pi / 2)), [math.sqrt(2), 2.0], rtol=1e-9)
    assert layer_light_lines(L(IsotropicMaterial(-4 + 0.1j)), W0, 0.0) == []
    assert np.allclose(r(layer_light_lines(L(UniaxialMaterial(4.0, -3.2)), W0, 0.0)), [2.0], rtol=1e-9)


def _features(stack, h_min=0.05, Kcut=None):
    from yellowthinfilms_amd.kspace import kspace_features
    return kspace_features(stack, W0, 1, h_min=h_min, scan=False, Kcut=Kcut)


def test_graded_panels_respect_the_width():
    from yellowthinfilms_amd.quadrature import graded_panels, halve_panels
    # Lipschitz with constant 0.5 (the condition of the grading)
    width = lambda k: min(max(1e-3, 0.5 * abs(k - 3.0)), max(0.05 * k, 0.2), 2.0)  # noqa: E731
    panels = graded_panels([0.0, 3.0, 5.0, 40.0], width)
    ends = {b for _, b in panels} | {a for a, _ in panels}
    assert {0.0, 3.0, 5.0, 40.0} <= ends
    assert all(b > a for a, b in panels) and panels[0][0] == 0.0 and panels[-1][1] == 40.0
    for a, b in panels:
        assert all(b - a <= width(x) * (1 + 1e-12) for x in np.linspace(a, b, 11))
    assert all(abs(panels[i][1] - panels[i + 1][0]) == 0 for i in range(len(panels) - 1))
    h = halve_panels(panels)
    assert len(h) == 2 * len(panels) and sum(b - a for a, b in h) == pytest.approx(40.0, rel=1e-14)


def test_fixed_rule_integrates_the_features():
    """fixed_kρ_rule on a vacuum / glass / vacuum stack (features without a
    scan, plus a hand-made pole cluster): a smooth decaying integrand, a
    1/sqrt|k - L| branch point at the singular light line L = k0 (u^2
    substitution), and a Lorentzian pole of width Gamma all integrate to the
    closed forms at the rule's rtol; the halved rule agrees."""
    from yellowthinfilms_amd.kspace import fixed_krho_rule
    from yellowthinfilms_amd.quadrature import gl_rule, halve_panels
    st = [FlatOpticalLayer(IsotropicMaterial(1.0), 1.0), FlatOpticalLayer(IsotropicMaterial(2.25), 0.3),
          FlatOpticalLayer(IsotropicMaterial(1.0), 1.0)]
    feat = _features(st, h_min=0.05)
    assert feat["singular_lines"] == pytest.approx([W0]) and feat["light_lines"] == pytest.approx([W0, 1.5 * W0])
    assert feat["Kcut"] == pytest.approx(25.0 / 0.05)
    q, G = 2.3 * W0, 1e-3 * W0
    feat["clusters"] = [dict(a=q, b=q, Gamma=G)]
    nodes, wts, panels = fixed_krho_rule(feat)
    K, h, Lw = feat["Kcut"], 0.05, W0
    smooth = np.sum(wts * nodes * np.exp(-nodes * h))
    assert smooth == pytest.approx((1 - (1 + K * h) * np.exp(-K * h)) / h ** 2, rel=1e-10)
    branch = np.sum(wts / np.sqrt(np.abs(nodes - Lw)))
    assert branch == pytest.approx(2 * math.sqrt(Lw) + 2 * math.sqrt(K - Lw), rel=1e-9)
    pole = np.sum(wts * G / ((nodes - q) ** 2 + G ** 2))
    assert pole == pytest.approx(math.atan((K - q) / G) + math.atan(q / G), rel=1e-6)
    n2, w2 = gl_rule(halve_panels(panels), order=8, singular=feat["singular_lines"])
    assert np.sum(w2 * G / ((n2 - q) ** 2 + G ** 2)) == pytest.approx(pole, rel=1e-6)
    # nodes stay away from the branch point (the general eigen path fails within ~1e-14 of it)
    assert np.min(np.abs(nodes - Lw)) > 1e-12 * Lw


def test_features_cut_off_and_phi_count():
    from yellowthinfilms_amd.kspace import kspace_features, adaptive_seeds
    st = [FlatOpticalLayer(IsotropicMaterial(1.0), 1.0), FlatOpticalLayer(BiaxialMaterial(2.0, 3.0, 4.0), 0.3),
          FlatOpticalLayer(IsotropicMaterial(2.25), 1.0)]
    f = kspace_features(st, W0, 1, h_min=0.01, rho_max=0.5, dims=2, scan=False, krho_cap=1000.0)
    assert f["phi_scan"] == pytest.approx(list(np.linspace(0, math.pi / 2, 9)))   # diagonal: [0, pi/2]
    assert f["Kcut"] == 1000.0 and f["capped"] and f["K_geom"] == pytest.approx(2500.0)
    assert f["N_phi"] % 2 == 0 and f["N_phi"] == 2 * math.ceil(2.5 * 1000.0 * 0.5 / 2)
    assert f["kmax"] == pytest.approx(2.0 * W0)           # the biaxial layer's sqrt(eps_z) line
    assert f["singular_lines"] == pytest.approx([W0, 1.5 * W0])
    s = adaptive_seeds(f)
    assert all(0 < x < f["Kcut"] for x in s) and s == sorted(s)
    g = kspace_features(st[:1] + [FlatOpticalLayer(IsotropicMaterial(2.25), 1.0)], W0, 1, h_min=10.0, scan=False)
    assert g["Kcut"] == pytest.approx(2 * 1.5 * W0)        # at least 2 kmax (K_geom = 2.5 is below it)
