"""CPU tests of the adaptive k_rho refinement (adaptive.py): the G7-K15 rule
against polynomial exactness and QuadGK's published example, the lock-step
refinement against closed-form integrals, stalls and errors, and the box
grid's host helpers (box_tail_cutoff against the Lambert-W closed form,
branch_points, is_lossless)."""
import math

import numpy as np
import pytest
from scipy.special import erf, lambertw

from helpers import lay
from yellowthinfilms_amd.adaptive import (kronrod_reference_rule, adaptive_kx_grid, box_tail_cutoff, branch_points,
                                          is_lossless, _points, _evalrule)
from yellowthinfilms_amd.materials import IsotropicMaterial


def test_kronrod_rule_exactness():
    xi, w = kronrod_reference_rule(7)
    assert xi.size == 15 and w.size == 15
    for deg in range(0, 24):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        err = abs(np.sum(w * xi ** deg) - exact)
        if deg <= 22:   # K15 is exact to degree 3n + 1 = 22
            assert err < 1e-14, (deg, err)
    assert abs(np.sum(w * xi ** 24) - 2.0 / 25) > 1e-9    # and not beyond
    # the embedded 7-point Gauss rule is exact to degree 13: E of a degree-13 polynomial vanishes
    a, b = 0.3, 1.7
    f = lambda k: (k - 0.9) ** 13 + 2 * k ** 4
    I, E = _evalrule(a, b, f(_points(a, b))[:, None])
    exact = ((b - 0.9) ** 14 - (a - 0.9) ** 14) / 14 + 2 * (b ** 5 - a ** 5) / 5
    assert abs(I[0] - exact) < 1e-13 * abs(exact) and E < 1e-13 * abs(exact)
    I, E = _evalrule(a, b, ((_points(a, b) - 0.9) ** 16)[:, None])
    assert E > 1e-10


def test_quadgk_published_example():
    """quadgk(x -> exp(-x^2), 0, 1, rtol=1e-8) = (0.746824132812427, 7.887024366937112e-13)
    in QuadGK's README: one segment, no refinement."""
    F = lambda k: np.exp(-k ** 2)[:, None, None]
    nodes, w, info = adaptive_kx_grid(F, 1, [], 1.0, rtol=1e-8)
    assert info["n_panels"] == 1 and info["n_batches"] == 1
    assert abs(np.sum(w * np.exp(-nodes ** 2)) - math.sqrt(math.pi) / 2 * erf(1.0)) < 1e-15
    assert abs(info["per_probe"][0]["E"] - 7.887024366937112e-13) < 1e-3 * 7.887024366937112e-13


def _lorentz(k, a, g):
    return 1.0 / ((k - a) ** 2 + g ** 2)


def _lorentz_int(K, a, g):
    return (math.atan((K - a) / g) + math.atan(a / g)) / g


@pytest.mark.parametrize("rtol", [1e-4, 1e-7])
def test_lockstep_refinement_closed_forms(rtol):
    """Three probes with different structure (two narrow resonances at
    different places, a smooth decay): every probe converges, the merged rule
    integrates each to its own tolerance, and each probe's panels are a
    subset of the merged partition."""
    K = 40.0
    probes = [(3.0, 1e-3), (17.5, 3e-2)]
    exact = [_lorentz_int(K, a, g) for a, g in probes] + [(1 - math.exp(-2 * K)) / 2]

    def F(k):
        cols = [_lorentz(k, a, g) for a, g in probes] + [np.exp(-2 * k)]
        return np.stack(cols, axis=1)[:, :, None]

    nodes, w, info = adaptive_kx_grid(F, 3, [1.0, 2.0, 30.0, 55.0], K, rtol=rtol)
    assert not info["any_stalled"]
    assert info["n_nodes"] == 15 * info["n_panels"] and nodes.size == info["n_nodes"]
    vals = F(nodes)[:, :, 0]
    for p in range(3):
        got = np.sum(w * vals[:, p])
        assert abs(got - exact[p]) < 10 * rtol * abs(exact[p]), (p, got, exact[p])
    bps = {a for a, _ in info["panels"]} | {info["panels"][-1][1]}
    assert {0.0, 1.0, 2.0, 30.0, K} <= bps and 55.0 not in bps
    # the narrowest panels sit on the narrowest resonance
    a, b = min(info["panels"], key=lambda p: p[1] - p[0])
    assert abs(0.5 * (a + b) - 3.0) < 0.05
    # every probe's own run (one probe at a time) gives panels contained in the merged set
    for p in range(3):
        _, _, own = adaptive_kx_grid(lambda k: F(k)[:, p:p + 1], 1, [1.0, 2.0, 30.0, 55.0], K, rtol=rtol)
        own_bps = {a for a, _ in own["panels"]} | {own["panels"][-1][1]}
        assert own_bps <= bps
        assert own["per_probe"][0]["E"] == info["per_probe"][p]["E"]
    # the rule evaluations read at most 15 points per segment ever evaluated
    assert info["n_krho_used"] <= 15 * (4 + 2 * (info["n_batches"] - 1) * 3)
    # speculation changes only which calls evaluate the points: the same rule, fewer calls
    n2, w2, plain = adaptive_kx_grid(F, 3, [1.0, 2.0, 30.0, 55.0], K, rtol=rtol, speculate=False)
    assert np.array_equal(n2, nodes) and np.array_equal(w2, w)
    assert [q["E"] for q in plain["per_probe"]] == [q["E"] for q in info["per_probe"]]
    assert plain["n_calls"] <= plain["n_batches"]
    assert info["n_calls"] <= (plain["n_calls"] + 1) // 2 + 1 and info["n_krho_used"] == plain["n_krho_used"]


def test_stall_and_errors():
    F = lambda k: np.stack([_lorentz(k, 3.0, 1e-7)], axis=1)[:, :, None]
    _, _, info = adaptive_kx_grid(F, 1, [], 10.0, rtol=1e-10, maxevals=300)
    assert info["any_stalled"] and info["per_probe"][0]["stalled"]
    # at most maxevals plus one split's points
    assert info["n_krho_used"] <= 300 + 30
    with pytest.raises(ValueError):
        adaptive_kx_grid(F, 0, [], 10.0)
    with pytest.raises(ValueError):
        adaptive_kx_grid(F, 1, [], 0.0)
    with pytest.raises(ValueError):
        adaptive_kx_grid(F, 1, [], 10.0, kr_order=5)
    with pytest.raises(FloatingPointError):
        adaptive_kx_grid(lambda k: (1.0 / (k - k))[:, None, None], 1, [], 1.0)


@pytest.mark.parametrize("delta,rtol", [(0.1, 1e-4), (1 / 8, 1e-1), (0.02, 1e-6)])
def test_box_tail_cutoff_lambert(delta, rtol):
    """log(x)/x = rtol at x = K delta, the larger root x = -W_{-1}(-rtol)/rtol."""
    K = box_tail_cutoff(delta, rtol)
    x = -lambertw(-rtol, -1).real / rtol
    assert abs(K * delta - x) < 1e-9 * x
    assert abs(math.log(K * delta) / (K * delta) - rtol) < 1e-9 * rtol


def test_branch_points_and_lossless():
    w = 2 * math.pi
    st = [lay(IsotropicMaterial(1.0 + 0.001j), 1.0), lay(IsotropicMaterial(2.25 + 0.05j), 0.6),
          lay(IsotropicMaterial(1.0 + 0.001j), 1.0)]
    bp = branch_points(st, w)
    want = sorted({round(w * np.sqrt(1.0 + 0.001j).real, 12), round(w * np.sqrt(2.25 + 0.05j).real, 12)})
    assert len(bp) == 2 and np.allclose(bp, want, rtol=1e-10)
    assert not is_lossless(st)
    assert is_lossless([lay(IsotropicMaterial(1.0), 1.0), lay(IsotropicMaterial(4.0), 1.0)])


def test_phi_groups():
    """One group for a symmetric stack; mirror classes (closed under phi -> -phi
    and phi -> pi - phi, in order of smallest phi) for diagonal tensors; one phi
    per group otherwise; classes_per_group merges neighbours; every result
    partitions the grid."""
    from helpers import FOUNDATION_STACKS
    from yellowthinfilms_amd.adaptive import phi_groups
    assert phi_groups(FOUNDATION_STACKS["uniaxial 4-layer"], 16) == [list(range(16))]
    for n in (16, 12, 9):
        g = phi_groups(FOUNDATION_STACKS["biaxial 4-layer"], n)
        assert sorted(i for c in g for i in c) == list(range(n))
        phi = 2 * np.pi * np.arange(n) / n
        for c in g:
            for p in c:
                assert (-p) % n in c
                if n % 2 == 0:
                    assert (n // 2 - p) % n in c
            # a class is one orbit: the same |sin|, |cos| for all members
            assert np.ptp(np.abs(np.cos(phi[c]))) < 1e-12
        assert [min(c) for c in g] == sorted(min(c) for c in g)
    assert phi_groups(FOUNDATION_STACKS["biaxial 4-layer"], 16)[:3] == [[0, 8], [1, 7, 9, 15], [2, 6, 10, 14]]
    assert phi_groups(FOUNDATION_STACKS["gyrotropic 4-layer"], 6) == [[0], [1], [2], [3], [4], [5]]
    assert phi_groups(FOUNDATION_STACKS["gyrotropic 4-layer"], 5, classes_per_group=2) == [[0, 1], [2, 3], [4]]
    g2 = phi_groups(FOUNDATION_STACKS["biaxial 4-layer"], 16, classes_per_group=2)
    assert g2[0] == [0, 1, 7, 8, 9, 15] and sorted(i for c in g2 for i in c) == list(range(16))
