"""The cavity response (ytf_cavity_response) and the batched real-axis pole
search on the GPU against the oracle (mod_KSpaceFeatures.jl:109-127,
151-214).  Tolerances.  The response: 1e-9 relative in norm and 3e-8 at every node.
Both sides use the reference's closed form of the largest singular value,
sqrt((t + sqrt(t^2 - 4 det)) / 2), which loses half the digits where the two
singular values nearly coincide (a response near 1 inside the stack, normal
incidence): FP64 rounding differences of the star products (same
formulation, other order) then show up at 1e-8.  The poles: the same list,
q to 1e-4 of the half-width (a maximum located from values good to d f / f
sits within sqrt(d f / f) widths), the half-width to 1e-3 of itself (it is
bisected between the peak and the valley; in a dense comb of poles the
response is not monotone there and the crossing found can move with the
last digits; the rule only sizes panels by it), the prominence to 1e-7."""
import math

import numpy as np
import pytest

from helpers import po, ostack, FOUNDATION_STACKS, lay
from test_kspace import spp_stack, base_stack
from yellowthinfilms_amd.materials import IsotropicMaterial, UniaxialMaterial, BiaxialMaterial, GyrotropicMaterial

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def tilted_stack():
    """A uniaxial layer with its axis tilted by 30 degrees in the xz plane (xz
    coupling: the general eigen-solve)."""
    c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    eps = R @ np.diag([2.0 + 0.05j, 2.0 + 0.05j, 4.5 + 0.05j]) @ R.T
    return [lay(IsotropicMaterial(1.0 + 0.001j), 1.0), lay(GyrotropicMaterial(eps), 0.4),
            lay(IsotropicMaterial(2.25 + 0.01j), 0.3), lay(IsotropicMaterial(1.0 + 0.001j), 1.0)]


def hyperbolic_stack():
    return [lay(IsotropicMaterial(1.0 + 1e-3j), 0.5), lay(BiaxialMaterial(2.25 + 1e-2j, 4.6 + 1e-2j, -3.2 + 1e-2j), 0.5),
            lay(IsotropicMaterial(1.0 + 1e-3j), 0.5)]


CASES = [("spp", spp_stack(-4 + 0.1j), 1), ("base", base_stack(), 1), ("iso5 s=2", FOUNDATION_STACKS["iso 5-layer lossy"], 2),
         ("iso5 s=3", FOUNDATION_STACKS["iso 5-layer lossy"], 3), ("biaxial s=2", FOUNDATION_STACKS["biaxial 4-layer"], 2),
         ("gyro s=1", FOUNDATION_STACKS["gyrotropic 4-layer"], 1), ("gyro s=3", FOUNDATION_STACKS["gyrotropic 4-layer"], 3),
         ("tilted s=2", tilted_stack(), 2), ("tilted s=3", tilted_stack(), 3), ("hyperbolic s=2", hyperbolic_stack(), 2)]


@pytest.mark.parametrize("label,stack,s", CASES, ids=[c[0] for c in CASES])
def test_cavity_response_against_oracle(label, stack, s):
    from yellowthinfilms_amd.kspace import cavity_response
    print(f"case: cavity response {label}", flush=True)
    rng = np.random.default_rng(len(label))
    kr = np.concatenate([rng.uniform(0, 4, 2000), rng.uniform(4, 40, 500)]) * W0
    ph = rng.uniform(0, 2 * np.pi, kr.size)
    kx, ky = kr * np.cos(ph), kr * np.sin(ph)
    g = cavity_response(stack, W0, s, kx, ky).cpu().numpy()
    o = po.cavity_response(ostack(stack), W0, s, kx, ky)
    ok = o > 0
    assert ok.mean() > 0.99
    rel = np.abs(g[ok] - o[ok]) / np.abs(o[ok])
    err = float(np.linalg.norm(g[ok] - o[ok]) / np.linalg.norm(o[ok]))
    worst = np.argsort(rel)[-3:]
    print(f"  rel err {err:.2e} (norm), max {rel.max():.2e} over {ok.sum()} nodes; response range "
          f"{o.min():.2e}..{o.max():.2e}; worst at k/k0 =", np.round(kr[ok][worst] / W0, 5), "response",
          o[ok][worst], flush=True)
    assert err < 1e-9
    assert rel.max() < 3e-8


@pytest.mark.parametrize("label,stack,s,phis,k_lo,k_hi,exclude", [
    ("spp", spp_stack(-4 + 0.1j), 1, [0.0], 0.5, 6.0, [1.0]),
    ("base", base_stack(), 1, [0.0], 0.5, 3.0, [1.0, 1.5, 2.0, math.sqrt(2.0)]),
    ("hyperbolic, 5 directions", hyperbolic_stack(), 2, list(np.linspace(0, np.pi / 2, 5)), 0.5, 30.0, [1.0, 1.5]),
    ("gyrotropic, 3 directions", FOUNDATION_STACKS["gyrotropic 4-layer"], 2, [0.0, 1.0, 2.5], 0.5, 8.0, []),
], ids=["spp", "base", "hyperbolic", "gyrotropic"])
def test_real_axis_poles_against_oracle(label, stack, s, phis, k_lo, k_hi, exclude):
    from yellowthinfilms_amd.kspace import real_axis_poles
    print(f"case: real_axis_poles {label}", flush=True)
    ex = [x * W0 for x in exclude]
    g = real_axis_poles(stack, W0, s, phis=phis, k_lo=k_lo * W0, k_hi=k_hi * W0, exclude=ex)
    o = po.real_axis_poles(ostack(stack), W0, s, phis=phis, k_lo=k_lo * W0, k_hi=k_hi * W0, exclude=ex)
    print(f"  {len(g)} poles (oracle {len(o)}):", [(round(p['q'] / W0, 5), round(p['Gamma'] / W0, 6)) for p in g],
          flush=True)
    assert len(g) == len(o) and len(g) > 0
    dq = 0.0
    for pg, (q, G, phi, prom) in zip(g, o):
        assert pg["phi"] == phi
        assert abs(pg["q"] - q) <= 1e-4 * G + 1e-12 * q, (pg["q"], q, G)
        assert abs(pg["Gamma"] - G) <= 1e-3 * G, (pg["Gamma"], G)
        assert abs(pg["prom"] - prom) <= 1e-7, (pg["prom"], prom)
        dq = max(dq, abs(pg["q"] - q) / G)
    print(f"  max |dq| / Gamma = {dq:.2e}", flush=True)
    if label == "base":
        strongest = sorted(g, key=lambda p: -p["prom"])[:2]
        assert round(strongest[0]["q"] / W0, 4) == 1.0624 and round(strongest[1]["q"] / W0, 2) == 1.66
