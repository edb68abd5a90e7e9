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


# ---------------------------------------------------------------------------
# Matrix-core path (split fp16 products, FP32 accumulation).  Error bar: the
# rounding bound of _f16x3_error_bound (each product ~2^-22 relative); on sums
# without strong cancellation that is <= 2e-6 relative to the FP64 result.
# ---------------------------------------------------------------------------
F16X3_TOL = 2e-6


@pytest.mark.parametrize("n,nr,npairs", [(1, 1, 1), (63, 5, 1), (130, 129, 3), (1000, 300, 8)])
def test_f16x3_ragged_sizes_against_fp64(n, nr, npairs):
    """Sizes that are not multiples of the 128 x 128 x 64 tiles (padding,
    the MFMA lane layout) against the FP64 contraction of the same g~ (random
    well-separated nodes: little cancellation, so a plain relative bar)."""
    print(f"case: f16x3 ragged n={n} nr={nr} pairs={npairs}", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, radial_synthesis, workloads as W
    stack = W.point_uniaxial_stack()
    rng = np.random.default_rng(n + nr)
    kx = np.sort(rng.uniform(0.01, 40.0, n))
    wt = rng.uniform(0.001, 0.05, n)
    cache = CavityTransfer1D(stack, W0, kx, wt, 1)
    zo = 1.0 - NM * rng.uniform(15, 200, npairs)
    g = cache.ghat(1, 10 * NM, zo)
    rho = np.sort(rng.uniform(0.0, 2.0, nr))
    R64 = radial_synthesis(cache, g, rho, mode="fp64").cpu().numpy()
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    assert relerr(R16, R64) < F16X3_TOL, relerr(R16, R64)


def _f16x3_error_bound(kx, wt, g, rho):
    """|| B ||_2 for the per-entry rounding bound of the split-fp16 products:
    a product of two hi+lo splits carries <= ~2^-22 relative error where both
    factors are normal fp16 numbers; below that the splits keep an absolute
    2^-25 (Bessel values) and 2^-25 * 2^e (coefficient column scaled by 2^-e,
    2^e <= 2 colmax).  Per Bessel term of an entry:
        B = 2^-20 sum_k |J||C| + 2^-23 colmax sum_k |J| + 2^-24 sum_k |C|."""
    from scipy.special import jv
    g = np.asarray(g)  # (np, 9, n)
    w = kx * wt / (2 * np.pi)
    gxx, gyx, gzx, gxy, gyy, gzy, gxz, gyz, gzz = (g[:, i] for i in range(9))
    cols = [[(gxx + gyy) / 2, (gxy - gyx) / 2, gzz], [gxz, gyz, gzx, gzy], [(gxx - gyy) / 2, (gxy + gyx) / 2]]
    X = np.outer(rho, kx)
    B = []
    for o in range(3):
        J = np.abs(jv(o, X))
        out = []
        for c in cols[o]:
            C = np.abs(c * w[None, :])  # (np, n)
            cm = C.max(axis=1)
            out.append(2.0 ** -20 * (J @ C.T) + 2.0 ** -23 * J.sum(axis=1)[:, None] * cm[None, :]
                       + 2.0 ** -24 * C.sum(axis=1)[None, :])
        B.append(out)
    (A, Bc, E), (XZ, YZ, ZX, ZY), (Cc, D) = B
    entries = [A + Cc, D + Bc, ZX, Bc + D, A + Cc, ZY, XZ, YZ, E]
    return np.linalg.norm(np.stack(entries))


def test_f16x3_bessel_arguments_in_every_regime():
    """The split table's J0/J1 (power series below x = 12, Hankel's
    expansion above) over k rho from 0 to ~2e4: within the rounding bound of
    the FP64 contraction."""
    print("case: f16x3 Bessel regimes", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, radial_synthesis, workloads as W
    rng = np.random.default_rng(11)
    kx = np.sort(np.geomspace(1e-3, 60.0, 700) * rng.uniform(0.9, 1.1, 700))
    wt = rng.uniform(0.001, 0.05, 700)
    cache = CavityTransfer1D(W.point_uniaxial_stack(), W0, kx, wt, 1)
    g = cache.ghat(1, 10 * NM, 1.0 - NM * np.array([20.0, 90.0]))
    rho = np.concatenate([[0.0], np.geomspace(1e-3, 300.0, 200)])
    X = np.outer(rho, kx)
    assert (X < 12).sum() > 1000 and ((X >= 12) & (X < 30)).sum() > 1000 and (X >= 30).sum() > 1000
    R64 = radial_synthesis(cache, g, rho, mode="fp64").cpu().numpy()
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    bound = _f16x3_error_bound(kx, wt, g.cpu().numpy(), rho)
    err = np.linalg.norm(R16 - R64)
    print("  rel err", relerr(R16, R64), "abs err", err, "bound", bound, flush=True)
    assert err <= bound


def test_f16x3_point_uniaxial_workload():
    """Split-fp16 error within its rounding bound on the full
    point_uniaxial_1d rule (whose direct term cancels strongly), and
    closed_form_G with the same mode."""
    print("case: f16x3 point_uniaxial_1d", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, radial_synthesis, workloads as W
    wl = W.POINT_UNIAXIAL_1D
    kx, wt = W.point_uniaxial_rule()
    zo, rho = W.point_uniaxial_geometry()
    cache = CavityTransfer1D(W.point_uniaxial_stack(), W0, kx, wt, 1)
    g = cache.ghat(1, wl["zsrc"], zo)
    R64 = radial_synthesis(cache, g, rho, mode="fp64").cpu().numpy()
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    bound = _f16x3_error_bound(kx, wt, g.cpu().numpy(), rho)
    err = np.linalg.norm(R16 - R64)
    print("f16x3 point_uniaxial_1d: rel err", relerr(R16, R64), "abs err", err, "bound", bound, flush=True)
    assert err <= bound
    xs = np.linspace(-1.0, 1.0, 12)
    G64 = _julia(closed_form_G(cache, 1, wl["zsrc"], zo, xs, xs, mode="fp64"))
    G16 = _julia(closed_form_G(cache, 1, wl["zsrc"], zo, xs, xs, mode="f16x3"))
    assert relerr(G16, G64) < 2e-5, relerr(G16, G64)


def test_f16x3_gemm_workload_scaled_down():
    """The mixed-precision GEMM workload's stack, pairs and rule at 1/4 of its
    node count and 1/16 of its radii, against the FP64 path."""
    print("case: f16x3 bessel_gemm scaled", flush=True)
    from yellowthinfilms_amd import workloads as W, green_nodes
    from yellowthinfilms_amd.synthesis import CavityTransfer1D, radial_synthesis
    wl = dict(W.BESSEL_GEMM, n_band=3584, n_tail=512, n_rho=512)
    kx, wt = W.bessel_gemm_rule(wl)
    zs, zo, rho = W.bessel_gemm_geometry(wl)
    cache = CavityTransfer1D(W.point_uniaxial_stack(), W0, kx, wt, 1)
    _, g = green_nodes(cache.stack, W0, cache.kx, cache.ky, 1, 1, zs, zo, cache=False)
    R64 = radial_synthesis(cache, g, rho, mode="fp64").cpu().numpy()
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    bound = _f16x3_error_bound(kx, wt, g.cpu().numpy(), rho)
    err = np.linalg.norm(R16 - R64)
    print("f16x3 bessel_gemm (scaled): rel err", relerr(R16, R64), "bound/|R|", bound / np.linalg.norm(R64), flush=True)
    assert err <= bound


def _oracle_R(stack, kx, wt, zs, zo, rho):
    """The oracle's closed_form_G (mod_ClosedFormField.jl:435-487) at pixels
    (rho_r, 0) for every depth pair, in R's layout (np, 9 = c + 3q, nr)."""
    c = po.nodes_batch(ostack(stack), W0, kx, np.zeros_like(kx), 1, [1])
    out = np.zeros((len(zs), 9, len(rho)), complex)
    for s in np.unique(zs):
        idx = np.nonzero(zs == s)[0]
        G = po.closed_form_G_1d(c, 1, wt, float(s), zo[idx], rho, [0.0])  # (nr, 1, nz, 3, 3)
        out[idx] = G[:, 0].transpose(1, 3, 2, 0).reshape(len(idx), 9, len(rho))
    return out


def test_f16x3_point_uniaxial_workload_against_the_oracle():
    """The matrix-core path on the whole point_uniaxial_1d configuration (4608
    nodes, 256 radii, 8 heights) against the oracle's FP64 closed_form_G: the
    difference stays within the split-fp16 rounding bound plus the FP64 path's
    own 1e-10 bar."""
    print("case: f16x3 point_uniaxial_1d vs oracle", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, radial_synthesis, workloads as W
    wl = W.POINT_UNIAXIAL_1D
    kx, wt = W.point_uniaxial_rule()
    zo, rho = W.point_uniaxial_geometry()
    zs = np.full(len(zo), wl["zsrc"])
    stack = W.point_uniaxial_stack()
    cache = CavityTransfer1D(stack, W0, kx, wt, 1)
    g = cache.ghat(1, wl["zsrc"], zo)
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    Ro = _oracle_R(stack, kx, wt, zs, zo, rho)
    bound = _f16x3_error_bound(kx, wt, g.cpu().numpy(), rho)
    err = np.linalg.norm(R16 - Ro)
    print("  rel err vs oracle", relerr(R16, Ro), "abs err", err, "bound", bound, flush=True)
    assert err <= bound + 1e-10 * np.linalg.norm(Ro)


def test_f16x3_bessel_gemm_full_size_against_the_oracle():
    """The mixed-precision configuration at its full size: 16384 nodes x 8192
    radii x 64 depth pairs through the matrix-core GEMM, every 128th radius (and
    the first and last) against the oracle's FP64 closed_form_G for all 64 pairs,
    within the rounding bound (the fp32 accumulation runs over all 16384 nodes)."""
    print("case: f16x3 bessel_gemm full size vs oracle", flush=True)
    from yellowthinfilms_amd import workloads as W, green_nodes
    from yellowthinfilms_amd.synthesis import CavityTransfer1D, radial_synthesis
    wl = W.BESSEL_GEMM
    kx, wt = W.bessel_gemm_rule(wl)
    zs, zo, rho = W.bessel_gemm_geometry(wl)
    assert len(kx) == 16384 and len(rho) == 8192 and len(zs) == 64
    stack = W.point_uniaxial_stack()
    cache = CavityTransfer1D(stack, W0, kx, wt, 1)
    _, g = green_nodes(cache.stack, W0, cache.kx, cache.ky, 1, 1, zs, zo, cache=False)
    R16 = radial_synthesis(cache, g, rho, mode="f16x3").cpu().numpy()
    pick = np.unique(np.concatenate([np.arange(0, len(rho), 128), [len(rho) - 1]]))
    Ro = _oracle_R(stack, kx, wt, zs, zo, rho[pick])
    bound = _f16x3_error_bound(kx, wt, g.cpu().numpy(), rho[pick])
    err = np.linalg.norm(R16[:, :, pick] - Ro)
    print("  rel err vs oracle", relerr(R16[:, :, pick], Ro), "abs err", err, "bound", bound, flush=True)
    assert np.isfinite(R16).all()
    assert err <= bound + 1e-10 * np.linalg.norm(Ro)


def test_kernel_timing_counts_launches():
    print("case: kernel timing", flush=True)
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, workloads as W
    from yellowthinfilms_amd._lib import Context
    kx, wt = W.point_uniaxial_rule(dict(W.POINT_UNIAXIAL_1D, n_band=256, n_tail=32))
    cache = CavityTransfer1D(W.point_uniaxial_stack(), W0, kx, wt, 1)
    ctx = Context.get(0)
    ctx.set_timing(True)
    try:
        for _ in range(2):
            closed_form_G(cache, 1, 10 * NM, [0.9], [0.1, 0.2], [0.0], mode="f16x3")
        for name in ("node_solve", "synth_coef", "bessel_table", "split_columns", "bessel_gemm", "mfma_combine",
                     "synth_rotate"):
            ms, n = ctx.kernel_time(name)
            assert n == 2 and ms > 0.0, (name, ms, n)
        assert ctx.kernel_time("synth_fp64") == (0.0, 0)
    finally:
        ctx.set_timing(False)
    assert ctx.kernel_time("node_solve") == (0.0, 0)
