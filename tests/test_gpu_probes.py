"""GPU parity of the batched probe integrands of the adaptive k_rho grids
(probes.py: ytf_green_nodes + ytf_segment_lateral_sum) against the oracle's
per-node spectral dyadic and exact lateral sum, and against this build's own
2D closed_form_G on one-k_rho-node caches (the reference's per-k_rho route,
mod_ClosedFormGrid.jl:204-218).  Tolerance: 1e-10 relative to each probe's
largest entry (the oracle comparisons of the 2D path elsewhere use the same)."""
import math

import numpy as np
import pytest

from helpers import po, ostack, lay

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def glass_stack():
    """The reference's grid validation stack (Gz_ClosedForm_Grid_Validation.jl:25-28)."""
    from yellowthinfilms_amd.materials import IsotropicMaterial
    return [lay(IsotropicMaterial(1.0 + 0.001j), 1.0), lay(IsotropicMaterial(2.25 + 0.05j), 0.6),
            lay(IsotropicMaterial(1.0 + 0.001j), 1.0)]


def gyro_stack():
    from helpers import FOUNDATION_STACKS
    return FOUNDATION_STACKS["gyrotropic 4-layer"]


@pytest.mark.parametrize("label,stack,src,box,zbar,rho_far", [
    ("glass box", glass_stack(), 2, (0.1, 0.1, 0.1), [0.3], math.hypot(0.5, 0.5)),
    ("gyrotropic box, default depths", gyro_stack(), 2, (0.05, 0.08, 0.04), None, 0.3),
    ("source layer 1", glass_stack(), 1, (0.1, 0.1, 0.2), [0.5], 0.0),
])
def test_probe_values_against_oracle_and_per_krho_caches(label, stack, src, box, zbar, rho_far):
    print(f"case: probe values, {label}", flush=True)
    from yellowthinfilms_amd.probes import box_probes, probe_values
    from yellowthinfilms_amd.quadrature import build_kphi_grid
    from yellowthinfilms_amd.green_operator import CavityTransfer2D, closed_form_G_2d, box_lateral_factor
    probes = box_probes(stack, src, box, zbar=zbar, rho_far=rho_far, lateral_dir=0.3)
    phis, wph = build_kphi_grid(24)
    kr = np.array([0.02, 0.7 * W0, 1.49 * W0, 3.3 * W0, 40.0, 700.0])
    V = probe_values(stack, W0, src, kr, phis, wph, probes, box)
    assert V.shape == (len(kr), len(probes), 3, 3)
    ost = ostack(stack)
    for i, k in enumerate(kr):
        kx, ky = k * np.cos(phis), k * np.sin(phis)
        W = k * wph / (4 * math.pi ** 2) * box_lateral_factor(kx, ky, box[0], box[1])
        cache = CavityTransfer2D(stack, W0, src, [k], [1.0], phis, wph)
        for m, (obs, zo, x, y, h, (zb, tag)) in enumerate(probes):
            g = po.ghat_batch(ost, W0, kx, ky, src, obs, [zb], [zo], dz_box=box[2], ez_local=True)[:, 0]
            ref = po.lateral_sum_points(g, W, kx, ky, [(x, y)])[0]                      # (c, q)
            own = closed_form_G_2d(cache, obs, zb, [zo], [x], [y], box_lateral=box[:2], dz_box=box[2])
            own = own[:, :, 0, 0, 0].cpu().numpy().T                                       # (c, q)
            scale = max(np.abs(ref).max(), 1e-300)
            e_ref = np.abs(V[i, m] - ref).max() / scale
            e_own = np.abs(V[i, m] - own).max() / scale
            assert e_ref < 1e-10, (tag, k, e_ref)
            assert e_own < 1e-12, (tag, k, e_own)
    print(f"  {len(probes)} probes x {len(kr)} k_rho within 1e-10 of the oracle", flush=True)


def test_probe_values_empty_and_errors():
    print("case: probe values, empty batch and bad boxes", flush=True)
    from yellowthinfilms_amd.probes import box_probes, probe_values
    st = glass_stack()
    probes = box_probes(st, 2, (0.1, 0.1, 0.1), zbar=[0.3])
    assert probe_values(st, W0, 2, [], [0.0], [2 * math.pi], probes, (0.1, 0.1, 0.1)).shape == (0, len(probes), 3, 3)
    with pytest.raises(ValueError):
        box_probes(st, 2, (0.1, 0.1, 0.7))       # thicker than the layer
    with pytest.raises(ValueError):
        box_probes(st, 3, (0.1, 0.1, 0.1))       # the last layer cannot hold a source
    with pytest.raises(ValueError):
        box_probes(st, 2, (0.1, 0.1, 0.1), zbar=[0.01])  # no depth fits


def test_point_probe_values_against_one_node_caches():
    """Per-node point probe values (ytf_green_nodes + ytf_radial_terms) against
    this build's 1D closed_form_G on one-k_x-node caches (the reference's
    cache_at / closed_form_pixel, mod_ClosedFormGrid.jl:407-411), and their
    weighted sum against closed_form_G on the whole rule (linearity)."""
    print("case: point probe values, uniaxial 4-layer, source layer 2 and 1", flush=True)
    from helpers import FOUNDATION_STACKS
    from yellowthinfilms_amd.probes import point_probes, point_probe_values
    from yellowthinfilms_amd.synthesis import CavityTransfer1D, closed_form_G
    st = FOUNDATION_STACKS["uniaxial 4-layer"]
    kx = np.array([0.05, 0.7 * W0, 1.3 * W0, 5 * W0, 60.0, 400.0])
    wt = np.array([0.3, 0.2, 0.1, 0.7, 1.1, 0.4])
    for src, zf in ((2, [0.3, 0.5]), (1, [0.5])):
        probes = point_probes(st, src, zf, 0.02, 0.4)
        V = point_probe_values(st, W0, src, kx, probes)
        assert V.shape == (kx.size, len(probes), 3)

        def field(cache, p):
            obs, zo, z, rho, mom = p
            G = closed_form_G(cache, obs, z, [zo], [rho], [0.0])[:, :, 0, 0, 0].cpu().numpy()   # [q, c]
            return np.einsum("qc,q->c", G, np.asarray(mom, np.complex128))
        for i, k in enumerate(kx):
            cache = CavityTransfer1D(st, W0, [k], [1.0], src)
            for m, p in enumerate(probes):
                own = field(cache, p)
                scale = max(np.abs(own).max(), 1e-300)
                assert np.abs(V[i, m] - own).max() / scale < 1e-12, (src, k, p)
        S = point_probe_values(st, W0, src, kx, probes, wt=wt).sum(axis=0)
        full = CavityTransfer1D(st, W0, kx, wt, src)
        for m, p in enumerate(probes):
            own = field(full, p)
            assert np.abs(S[m] - own).max() <= 1e-12 * max(np.abs(own).max(), 1e-300), p


def test_point_kx_grid_engines():
    """point_kx_grid: the adaptive engine on a dielectric stack converges every
    probe and its check (every panel halved) moves the probes by < 10 rtol;
    the fixed engine on a gyrotropic stack builds the feature rule and checks it."""
    print("case: point_kx_grid, iso 5-layer (adaptive) and gyrotropic 4-layer (fixed)", flush=True)
    import time
    from helpers import FOUNDATION_STACKS
    from yellowthinfilms_amd.adaptive import point_kx_grid
    st = FOUNDATION_STACKS["iso 5-layer lossy"]
    for rho_max, rtol in ((0.0, 1e-6), (0.1, 1e-4)):
        t0 = time.perf_counter()
        kx, wt, info = point_kx_grid(st, W0, 2, rho_max=rho_max, z_fracs=(0.3, 0.7), rtol=rtol, check=True)
        print(f"  adaptive rho_max {rho_max}: {info['n_panels']} panels, {info['n_kx_solves']} k_x in "
              f"{info['n_batches']} batches, {info['n_probes']} probes, check {info['check']['rel_change']:.2e}, "
              f"{time.perf_counter() - t0:.2f} s", flush=True)
        assert info["engine"] == "adaptive" and not info["any_stalled"]
        assert kx.size == wt.size == 15 * info["n_panels"] and info["tail_bound"] == info["features"]["Kcut"]
        assert info["check"]["rel_change"] < 10 * rtol
    gy = FOUNDATION_STACKS["gyrotropic 4-layer"]
    kx, wt, info = point_kx_grid(gy, W0, 2, z_fracs=(0.5,), check=True)
    print(f"  fixed: {info['n_nodes']} nodes, check {info['check']['rel_change']:.2e}", flush=True)
    assert info["engine"] == "fixed" and kx.size == info["n_nodes"] > 0 and np.isfinite(info["check"]["rel_change"])
