"""GPU parity of the radial Bessel synthesis (ytf_radial_synthesis,
ytf_closed_form_G_1d) against the oracle's closed_form_G 1D restatement and
against the free-space dipole closed form, through the C-ABI.

Tolerances: FP64 throughout, <= 1e-10 relative over the whole output (the
plain-path bar); the free-space closed form at the oracle test's 1e-10."""
import math

import numpy as np
import pytest

from helpers import po, relerr, ostack, lay, IsotropicMaterial, E_free, kr_rule, NM

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def _julia(G):
    """(3, 3, nz, ny, nx) torch -> numpy (nx, ny, nz, 3, 3)."""
    return G.cpu().numpy().transpose(4, 3, 2, 1, 0)


def _oracle_G(stack, kx, wt, src, obs, zs, zo, xs, ys):
    c = po.nodes_batch(ostack(stack), W0, kx, np.zeros_like(kx), src, [obs])
    return po.closed_form_G_1d(c, obs, wt, zs, zo, xs, ys)


def test_free_space_dipole_closed_form():
    print("case: free-space dipole, 1D synthesis", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G
    ef = 1.0 + 1e-4j
    stack = [lay(IsotropicMaterial(ef), 3.0) for _ in range(3)]
    zs, zo = 1.0, 1.7
    k = W0 * np.sqrt(ef)
    kx, wt = kr_rule([W0 * np.sqrt(ef).real], kmax=36 / 0.7 + W0, width=0.5, order=16)
    cache = CavityTransfer1D(stack, W0, kx, wt, 2)
    xs, ys = np.array([0.4, -0.25]), np.array([0.0, 0.3])
    G = _julia(closed_form_G(cache, 2, zs, [zo], xs, ys))
    for i, x in enumerate(xs):
        for j, y in enumerate(ys):
            Gref = np.stack([E_free([x, y, zo - zs], np.eye(3)[q], k, ef) for q in range(3)], axis=1)
            assert relerr(G[i, j, 0], Gref) < 1e-10, (x, y, relerr(G[i, j, 0], Gref))


@pytest.mark.parametrize("obs", [1, 2, 3])
def test_uniaxial_stack_every_obs_layer_against_oracle(obs):
    print(f"case: uniaxial film stack, obs layer {obs}", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, workloads as W
    stack = W.point_uniaxial_stack()
    kx, wt = W.point_uniaxial_rule(dict(W.POINT_UNIAXIAL_1D, n_band=1024, n_tail=128))
    d = stack[obs - 1].d
    zs = 10 * NM
    zo = [1.0 - 20 * NM, 1.0 - 80 * NM, 0.5] if obs == 1 else ([0.3 * d, 0.8 * d] if obs == 2 else [0.0, 0.05])
    xs = np.array([0.0, 0.1, -0.35, 0.8])
    ys = np.array([0.0, 0.2, -0.6])
    cache = CavityTransfer1D(stack, W0, kx, wt, 1)
    G = _julia(closed_form_G(cache, obs, zs, zo, xs, ys))
    Gref = _oracle_G(stack, kx, wt, 1, obs, zs, zo, xs, ys)
    assert relerr(G, Gref) < 1e-10, relerr(G, Gref)


def test_source_inside_the_film():
    print("case: source in the film, obs in film and air", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, workloads as W
    stack = W.point_uniaxial_stack()
    kx, wt = W.point_uniaxial_rule(dict(W.POINT_UNIAXIAL_1D, n_band=512, n_tail=64))
    xs = np.array([0.05, 0.4])
    ys = np.array([0.0, 0.1])
    cache = CavityTransfer1D(stack, W0, kx, wt, 2)
    for obs, zo in [(2, [10 * NM, 40 * NM]), (1, [1.0 - 30 * NM])]:
        G = _julia(closed_form_G(cache, obs, 25 * NM, zo, xs, ys))
        Gref = _oracle_G(stack, kx, wt, 2, obs, 25 * NM, zo, xs, ys)
        assert relerr(G, Gref) < 1e-10, (obs, relerr(G, Gref))


def test_point_uniaxial_workload_full_size():
    """The full 4608-node rule, 256 radii x 8 heights, against the oracle."""
    print("case: point_uniaxial_1d full size", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, radial_synthesis, workloads as W
    wl = W.POINT_UNIAXIAL_1D
    stack = W.point_uniaxial_stack()
    kx, wt = W.point_uniaxial_rule()
    zo, rho = W.point_uniaxial_geometry()
    cache = CavityTransfer1D(stack, W0, kx, wt, wl["source_layer"])
    G = _julia(closed_form_G(cache, wl["obs_layer"], wl["zsrc"], zo, rho, [0.0]))
    Gref = _oracle_G(stack, kx, wt, 1, 1, wl["zsrc"], zo, rho, [0.0])
    assert relerr(G, Gref) < 1e-10, relerr(G, Gref)
    # the radial response on the x axis is G itself there
    g = cache.ghat(wl["obs_layer"], wl["zsrc"], zo)
    R = radial_synthesis(cache, g, rho).cpu().numpy()  # (nz, 9, nr)
    Rj = R.reshape(len(zo), 3, 3, len(rho)).transpose(3, 0, 2, 1)  # (nr, nz, c, q)
    assert relerr(Rj, G[:, 0]) < 1e-13


def test_no_nodes_and_bad_mode():
    print("case: synthesis edge cases", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, radial_synthesis, closed_form_G
    from yellowthinfilms_amd._lib import YtfError
    stack = [lay(IsotropicMaterial(1.0 + 0.01j), 1.0), lay(IsotropicMaterial(2.25), 1.0)]
    cache = CavityTransfer1D(stack, W0, np.zeros(0), np.zeros(0), 1)
    g = torch.zeros((2, 9, 0), dtype=torch.complex128, device="cuda")
    R = radial_synthesis(cache, g, [0.1, 0.2])
    assert R.shape == (2, 9, 2) and float(R.abs().max()) == 0.0
    with pytest.raises(ValueError):
        radial_synthesis(cache, g, [0.1], mode="fp8")
    cache2 = CavityTransfer1D(stack, W0, [1.0, 2.0], [0.5, 0.5], 1)
    with pytest.raises(YtfError):
        closed_form_G(cache2, 1, 2.0, [0.5], [0.1], [0.0])  # zsrc outside layer 1
