"""GPU parity of the Green k-node kernel (ytf_green_nodes) and the spectral
apply (ytf_spectral_apply) against the CPU oracle, called through the C-ABI.

Tolerances: FP64 everywhere; the spectral dyadic g~ and the cache entries must
agree with the oracle to <= 1e-10 relative (the plain-path bar); eigenvector-
dependent entries are compared where both sides normalise the same way."""
import math

import numpy as np
import pytest

from helpers import (po, relerr, ostack, c1_stack, c2_stack, FOUNDATION_STACKS, FOUNDATION_KTS, NM, lay,
                     IsotropicMaterial)

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def _gpu(stack, kx, ky, src, obs, zs, zo):
    from yellowthinfilms_amd import green_nodes
    nc, g = green_nodes(stack, W0, kx, ky, src, obs, zs, zo, cache=True, ghat=True)
    torch.cuda.synchronize()
    return nc.to_numpy(), g.cpu().numpy()


def _compare(stack, kx, ky, src, obs, zs, zo, tol=1e-10):
    kx = np.asarray(kx, float)
    ky = np.asarray(ky, float)
    c, g = _gpu(stack, kx, ky, src, obs, zs, zo)
    oc = po.nodes_batch(ostack(stack), W0, kx, ky, src, [obs])
    og = po.ghat_from_cache(oc, obs, zs, zo)  # (n, npairs, 3, 3)
    # flags
    assert np.array_equal(c["valid"], oc.valid), "valid mask differs"
    assert np.array_equal(c["Tvalid"], oc.Tv[obs] & oc.valid), "Tvalid mask differs"
    v = oc.valid
    # spectral dyadic: (npairs, 9, n) with entry c + 3q
    gg = g.reshape(len(zs), 3, 3, -1).transpose(3, 0, 2, 1)  # (n, npairs, c, q)
    assert relerr(gg, og) < tol, relerr(gg, og)
    for k in np.nonzero(v)[0]:
        scale = max(np.linalg.norm(og[k]), 1e-300)
        assert np.linalg.norm(gg[k] - og[k]) <= 1e3 * tol * scale + 1e-300, (k, kx[k], ky[k])
    # cache entries (same LAPACK-style normalisation on both sides)
    assert relerr(c["Q"][v], np.concatenate([oc.Qf, oc.Qb], axis=1)[v]) < tol
    assert relerr(c["sf"][v], oc.sf[v]) < tol
    assert relerr(c["sb"][v], oc.sb[v]) < tol
    tv = v & oc.Tv[obs]
    assert relerr(c["T"][tv], oc.T[obs][tv]) < tol
    assert relerr(c["basis"][v], oc.basis[obs][v]) < tol
    return c, g


def test_c2_stack_subgrid_same_layer():
    """5-layer gyrotropic stack on a 64x64 subsample of the 512^2 uniform grid."""
    print("case: C2 stack, 64x64 subgrid, src=obs=1", flush=True)
    from yellowthinfilms_amd import uniform_kxky_grid
    grid = uniform_kxky_grid(16.0, 1 / 16)
    kx1 = grid["kx"][::8]
    ky1 = grid["ky"][::8]
    KX, KY = np.meshgrid(kx1, ky1, indexing="ij")
    _compare(c2_stack(), KX.ravel(), KY.ravel(), 1, 1, [20 * NM], [40 * NM])


@pytest.mark.parametrize("label", list(FOUNDATION_STACKS))
def test_foundation_stacks_every_layer_pair(label):
    stack = FOUNDATION_STACKS[label]
    N = len(stack)
    rng = np.random.default_rng(7)
    kts = list(FOUNDATION_KTS) + [tuple(rng.uniform(-12, 12, 2)) for _ in range(12)] + [(0.0, 0.0)]
    kx = np.array([k[0] for k in kts])
    ky = np.array([k[1] for k in kts])
    for src in range(1, N):
        ds = stack[src - 1].d
        for obs in range(1, N + 1):
            do = stack[obs - 1].d
            print(f"case: {label} src={src} obs={obs}", flush=True)
            zs = [0.3 * ds, 0.7 * ds]
            zo = [0.2 * do, 0.9 * do]
            _compare(stack, kx, ky, src, obs, zs, zo)


@pytest.mark.parametrize("src,obs", [(1, 1), (2, 4), (3, 3), (4, 2)])
def test_depth_pairs_split_over_groups(src, obs):
    """9,600 nodes (150 one-wave blocks) and 7 depth pairs: on a 256-CU GPU
    the launch splits the pairs into 4 groups on grid.y (2, 2, 2, 1 pairs),
    each group solving the node again; every pair's g~ and the cache written
    by the first group must match the oracle."""
    print(f"case: depth-pair groups, C2 stack src={src} obs={obs}", flush=True)
    stack = c2_stack()
    rng = np.random.default_rng(19)
    kr = rng.uniform(0.0, 4.0, 9600) * W0
    ph = rng.uniform(0, 2 * np.pi, 9600)
    ds, do = stack[src - 1].d, stack[obs - 1].d
    zs = list(np.linspace(0.1, 0.9, 7) * min(ds, 0.5))
    zo = list(np.linspace(0.85, 0.15, 7) * min(do, 0.5))
    _compare(stack, kr * np.cos(ph), kr * np.sin(ph), src, obs, zs, zo)


def test_c1_stack_radial_nodes_all_obs_layers():
    stack = c1_stack()
    # near (not on) the light lines of air (k0) and glass (1.5 k0): exactly on a
    # lossless light line kz is rounding noise and the reference's own mode
    # sort is rounding-dependent there (its rules keep nodes off, mod_KxGrid.jl:355-359)
    near = [W0 * (1 + s * e) for s in (-1, 1) for e in (1e-3, 1e-6)]
    near += [1.5 * W0 * (1 + s * e) for s in (-1, 1) for e in (1e-3, 1e-6)]
    kr = np.concatenate([np.linspace(0, 6 * W0, 400), near])
    for obs in (1, 2, 3):
        print(f"case: C1 stack obs={obs}", flush=True)
        do = stack[obs - 1].d
        _compare(stack, kr, np.zeros_like(kr), 1, obs, [10 * NM, 10 * NM], [20 * NM, 0.5 * do])


def test_deep_evanescent_nodes_are_masked_like_the_reference():
    # Im Q d > 709 makes T non-finite; the node keeps valid but loses Tvalid
    stack = [lay(IsotropicMaterial(1.0 + 0.01j), 1.0), lay(IsotropicMaterial(2.25 + 0.01j), 2.0),
             lay(IsotropicMaterial(1.0 + 0.01j), 1.0)]
    kr = np.array([50.0, 300.0, 400.0, 800.0, 2000.0])
    for src, obs in [(1, 1), (1, 3), (2, 1), (2, 2)]:
        print(f"case: evanescent src={src} obs={obs}", flush=True)
        c, _ = _compare(stack, kr, 0.3 * kr, src, obs, [0.5], [0.5])
        oc = po.nodes_batch(ostack(stack), W0, kr, 0.3 * kr, src, [obs])
        assert np.array_equal(c["Tvalid"], oc.Tv[obs] & oc.valid)


def test_spectral_apply_small_grid_against_oracle():
    """g~ and the FFT convolution on a 32x64 grid against the oracle route."""
    print("case: spectral apply 32x64", flush=True)
    from yellowthinfilms_amd import uniform_kxky_grid, SpectralGreenOperator
    grid = uniform_kxky_grid(1.0, 1 / 16, Ly=2.0)
    nx, ny = grid["Nx"], grid["Ny"]
    assert (nx, ny) == (32, 64)
    stack = c2_stack()
    op = SpectralGreenOperator(stack, W0, grid, 1, 20 * NM, 1, 40 * NM).build()
    KX, KY = np.meshgrid(grid["kx"], grid["ky"], indexing="ij")
    og = po.ghat_batch(ostack(stack), W0, KX.ravel(), KY.ravel(), 1, 1, [20 * NM], [40 * NM])[:, 0]
    og = og.reshape(nx, ny, 3, 3)
    gpu_g = op.ghat[0].cpu().numpy().reshape(3, 3, nx, ny).transpose(2, 3, 1, 0)  # (x, y, c, q)
    assert relerr(gpu_g, og) < 1e-10
    rng = np.random.default_rng(1)
    J = (rng.normal(size=(nx, ny, 3)) + 1j * rng.normal(size=(nx, ny, 3))) / math.sqrt(2)
    Jt = torch.as_tensor(np.ascontiguousarray(J.transpose(2, 1, 0)), device="cuda")
    E = op.apply(Jt).cpu().numpy().transpose(2, 1, 0)
    Eref = po.spectral_apply(og, J, grid)
    assert relerr(E, Eref) < 1e-10


@pytest.mark.parametrize("L,dx,Ly,dy", [
    (0.5, 0.1, None, None),        # 10 x 10: uniform_kxky_grid(L=0.5, dx=0.1) of the reference's validations
    (1.0, 0.1, None, None),        # 20 x 20
    (0.3, 0.1, None, None),        # 6 x 6
    (0.25, 0.1, 0.35, 0.1),        # 5 x 7 (odd: the wavenumbers start at -ceil(N/2) dk)
    (0.55, 0.1, 0.65, 0.1),        # 11 x 13 (primes above 7)
    (3.0, 1 / 16, 1.875, 1 / 16),  # 96 x 60
    (0.5, 1 / 16, 0.3125, 1 / 16),  # 16 x 10 (radix-8 x pass, ragged row blocks)
    (31.25, 1 / 16, 0.1875, 1 / 16),  # 1000 x 6
    (1 / 32, 1 / 64, 1021 / 128, 1 / 64),  # 4 x 1021 (a prime length)
])
def test_spectral_apply_any_grid_size(L, dx, Ly, dy):
    """g~ and the FFT convolution on grids that are not powers of two, against
    the oracle's g~ and numpy's transform pair (library check of the
    mixed-radix LDS transforms)."""
    from yellowthinfilms_amd import uniform_kxky_grid, SpectralGreenOperator
    grid = uniform_kxky_grid(L, dx, Ly=Ly, dy=dy)
    nx, ny = grid["Nx"], grid["Ny"]
    print(f"case: spectral apply {nx}x{ny}", flush=True)
    stack = c2_stack()
    op = SpectralGreenOperator(stack, W0, grid, 1, 20 * NM, 1, 40 * NM).build()
    KX, KY = np.meshgrid(grid["kx"], grid["ky"], indexing="ij")
    og = po.ghat_batch(ostack(stack), W0, KX.ravel(), KY.ravel(), 1, 1, [20 * NM], [40 * NM])[:, 0]
    og = og.reshape(nx, ny, 3, 3)
    rng = np.random.default_rng(nx * ny)
    J = (rng.normal(size=(nx, ny, 3)) + 1j * rng.normal(size=(nx, ny, 3))) / math.sqrt(2)
    Jt = torch.as_tensor(np.ascontiguousarray(J.transpose(2, 1, 0)), device="cuda")
    E = op.apply(Jt).cpu().numpy().transpose(2, 1, 0)
    Eref = po.spectral_apply(og, J, grid)
    assert relerr(E, Eref) < 1e-10, relerr(E, Eref)


def test_spectral_apply_full_512_grid():
    """Full 512^2 grid: the GPU apply equals numpy's transform pair on the GPU g~
    (size-independent check of the FFT passes), g~ parity on a node sample,
    and linearity of the apply."""
    print("case: spectral apply 512x512", flush=True)
    from yellowthinfilms_amd import uniform_kxky_grid, SpectralGreenOperator
    grid = uniform_kxky_grid(16.0, 1 / 16)
    nx, ny = grid["Nx"], grid["Ny"]
    stack = c2_stack()
    op = SpectralGreenOperator(stack, W0, grid, 1, 20 * NM, 1, 40 * NM).build()
    g = op.ghat[0].cpu().numpy().reshape(3, 3, nx, ny).transpose(2, 3, 1, 0)
    rng = np.random.default_rng(2)
    idx = rng.choice(nx * ny, 1024, replace=False)
    xi, yi = idx // ny, idx % ny
    og = po.ghat_batch(ostack(stack), W0, grid["kx"][xi], grid["ky"][yi], 1, 1, [20 * NM], [40 * NM])[:, 0]
    err = relerr(g[xi, yi], og)
    if err >= 1e-10:
        d = np.array([np.linalg.norm(g[xi[k], yi[k]] - og[k]) / np.linalg.norm(og[k]) for k in range(len(xi))])
        for k in np.argsort(-d)[:5]:
            print("worst node", k, xi[k], yi[k], d[k], np.linalg.norm(og[k]), np.linalg.norm(g[xi[k], yi[k]]),
                  flush=True)
        g2 = op.build().ghat[0].cpu().numpy().reshape(3, 3, nx, ny).transpose(2, 3, 1, 0)
        print("rebuild: rel err", relerr(g2[xi, yi], og), "vs first build", relerr(g2, g), flush=True)
    assert err < 1e-10
    J = (rng.normal(size=(nx, ny, 3)) + 1j * rng.normal(size=(nx, ny, 3))) / math.sqrt(2)
    Jt = torch.as_tensor(np.ascontiguousarray(J.transpose(2, 1, 0)), device="cuda")
    E = op.apply(Jt).cpu().numpy().transpose(2, 1, 0)
    assert relerr(E, po.spectral_apply(g, J, grid)) < 1e-12
    E2 = op.apply(2 * Jt + 1j * Jt).cpu().numpy().transpose(2, 1, 0)
    assert relerr(E2, (2 + 1j) * E) < 1e-13


@pytest.mark.parametrize("label", ["iso 5-layer lossy", "gyrotropic 4-layer"])
def test_term_classes_against_oracle(label):
    """terms = all / direct / scattered / toeplitz / hankel (closed_form_G's
    term classes) in the source layer, and the same-layer requirement."""
    from yellowthinfilms_amd import green_nodes
    from yellowthinfilms_amd._lib import YtfError
    stack = FOUNDATION_STACKS[label]
    rng = np.random.default_rng(21)
    kx, ky = rng.uniform(-9, 9, 48), rng.uniform(-9, 9, 48)
    for src in (1, 2):
        d = stack[src - 1].d
        zs, zo = [0.3 * d, 0.6 * d], [0.7 * d, 0.2 * d]
        for t in ("all", "direct", "scattered", "toeplitz", "hankel"):
            print(f"case: {label} src={src} terms={t}", flush=True)
            _, g = green_nodes(stack, W0, kx, ky, src, src, zs, zo, cache=False, terms=t)
            gg = g.cpu().numpy().reshape(2, 3, 3, -1).transpose(3, 0, 2, 1)
            og = po.ghat_batch(ostack(stack), W0, kx, ky, src, src, zs, zo, terms=t)
            assert relerr(gg, og) < 1e-10, (t, relerr(gg, og))
    with pytest.raises(YtfError):
        green_nodes(stack, W0, kx, ky, 1, 2, [0.1], [0.1], cache=False, terms="direct")
    with pytest.raises(ValueError):
        green_nodes(stack, W0, kx, ky, 1, 1, [0.1], [0.1], cache=False, terms="reflected")


@pytest.mark.parametrize("label,src,obs,zc,dz,zos", [
    ("iso 5-layer lossy", 2, 2, 0.35, 0.2, [0.6, 0.1, 0.3, 0.44, 0.25]),  # exterior and interior planes
    ("iso 5-layer lossy", 1, 1, 0.3, 0.2, [0.2, 0.35, 0.6, 0.05]),         # source above the stack (CASE_TOP)
    ("gyrotropic 4-layer", 2, 2, 0.3, 0.2, [0.5, 0.25, 0.1]),
    ("uniaxial 4-layer", 2, 3, 0.4, 0.3, [0.25]),                           # another layer
    ("biaxial 4-layer", 2, 1, 0.4, 0.3, [0.3]),
])
def test_box_sources_against_oracle(label, src, obs, zc, dz, zos):
    """Box sources (closed_form_G's Δz_box): box seeds, the exterior direct term
    and, inside the slab, the face ramps and the local Ez, for every term class."""
    from yellowthinfilms_amd import green_nodes
    stack = FOUNDATION_STACKS[label]
    rng = np.random.default_rng(4)
    kr = np.concatenate([rng.uniform(0, 3 * W0, 96), rng.uniform(3 * W0, 40 * W0, 32)])
    ph = rng.uniform(0, 2 * np.pi, 128)
    kx, ky = kr * np.cos(ph), kr * np.sin(ph)
    st = ostack(stack)
    classes = ["all", "scattered"] + (["direct", "toeplitz", "hankel"] if obs == src else [])
    for terms in classes:
        for ez in (True, False):
            print(f"case: box {label} src={src} obs={obs} dz={dz} terms={terms} ez_local={ez}", flush=True)
            _, g = green_nodes(stack, W0, kx, ky, src, obs, [zc] * len(zos), zos, cache=False, terms=terms,
                               dz_box=dz, ez_local=ez)
            torch.cuda.synchronize()
            gg = g.cpu().numpy().reshape(len(zos), 3, 3, -1).transpose(3, 0, 2, 1)
            og = po.ghat_batch(st, W0, kx, ky, src, obs, [zc] * len(zos), zos, terms=terms, dz_box=dz, ez_local=ez)
            assert relerr(gg, og) < 1e-10, (terms, ez, relerr(gg, og))


def test_box_must_fit_its_layer_gpu():
    from yellowthinfilms_amd import green_nodes
    from yellowthinfilms_amd._lib import YtfError
    with pytest.raises(YtfError, match="does not fit inside its layer"):
        green_nodes(FOUNDATION_STACKS["iso 5-layer lossy"], W0, [1.0], [0.0], 2, 2, [0.05], [0.3], dz_box=0.2)
