"""The tabulated k_rho surrogate (ytf_ktable_ghat) through the C-ABI against
the exact FP64 solve of every node (ytf_green_nodes, scattered terms, source
and observation point in the air above the stack).

Tolerances: the table is built to 1e-10 relative to its largest |t|; the
dyadic on every node is then within 1e-8 of the exact solve, norm-wise over
the node set and node by node (relative to the node set's largest entry);
flagged nodes equal the exact solve bit for bit."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

W0 = 2 * math.pi


def _exact(stack, kx, ky, hs, ho):
    from yellowthinfilms_amd.green import green_nodes
    from yellowthinfilms_amd.ktable import NM
    _, g = green_nodes(stack, W0, torch.as_tensor(kx, device="cuda"), torch.as_tensor(ky, device="cuda"), 1, 1,
                       [hs * NM], [stack[0].d - ho * NM], cache=False, terms="scattered")
    return g[0].T


def _same(a, b):
    """Bit-identical, NaN where the exact solve gives NaN (its own light-line value)."""
    return a.shape == b.shape and bool(torch.all((a == b) | (torch.isnan(a) & torch.isnan(b))))


def _nodes(seed, n, kmax):
    rng = np.random.default_rng(seed)
    kx, ky = rng.uniform(-kmax, kmax, (2, n))
    # the air light line from both sides, the origin, exact axis points
    ph = rng.uniform(0, 2 * math.pi, 64)
    for f in (1 - 1e-3, 1 - 1e-7, 1 + 1e-7, 1 + 1e-3):
        kx = np.concatenate([kx, f * W0 * np.cos(ph)])
        ky = np.concatenate([ky, f * W0 * np.sin(ph)])
    kx = np.concatenate([kx, [0.0, 1e-9, 0.3 * W0, 0.0]])
    ky = np.concatenate([ky, [0.0, 0.0, 0.0, -2.2 * W0]])
    return kx, ky


@pytest.mark.parametrize("label,loss", [("gyrotropic film, loss 0.1", 0.1), ("gyrotropic film, loss 0.01", 0.01),
                                        ("uniaxial film", None)])
def test_ktable_against_the_exact_solve(label, loss):
    from yellowthinfilms_amd.ktable import KrhoTable
    from yellowthinfilms_amd.surrogate import GyroFamily
    from helpers import c1_stack
    print(f"case: k_rho table, {label}", flush=True)
    stack = GyroFamily(seed=1).stack(loss) if loss is not None else c1_stack()
    kmax = 3.0 * W0
    tab = KrhoTable(stack, W0, kmax * math.sqrt(2) * 1.001, tol=1e-10, deg=16)
    print(f"  {tab.n_panels} panels ({tab.n_unconverged} unconverged), {tab.n_exact} exact samples, "
          f"{tab.val.shape[0]} table rows, built in {tab.build_s:.2f} s", flush=True)
    assert tab.n_unconverged == 0
    kx, ky = _nodes(3, 20000, kmax)
    for hs, ho in ((40.0, 50.0), (20.0, 80.0)):
        g, mask = tab.ghat(kx, ky, hs, ho)
        ex = _exact(stack, kx, ky, hs, ho)
        assert _same(g[mask], ex[mask])
        ok = torch.isfinite(ex).all(dim=1)
        d = (g - ex)[ok]
        e_norm = float(torch.linalg.norm(d) / torch.linalg.norm(ex[ok]))
        e_node = float(d.abs().max() / ex[ok].abs().max())
        print(f"  heights {hs}/{ho} nm: rel err {e_norm:.2e} (norm), {e_node:.2e} (largest entry); "
              f"{int(mask.sum())} flagged of {len(kx)}", flush=True)
        assert e_norm < 1e-8 and e_node < 1e-8
        # only the exact light-line nodes (none here) and nothing else are flagged
        assert int(mask.sum()) == 0


def test_ktable_flags_and_fallback():
    from yellowthinfilms_amd.ktable import KrhoTable
    from yellowthinfilms_amd.surrogate import GyroFamily
    print("case: k_rho table flags", flush=True)
    stack = GyroFamily(seed=1).stack(0.1)
    tab = KrhoTable(stack, W0, 2.0 * W0, tol=1e-10, deg=12)
    kx = np.array([W0, 0.0, 2.5 * W0, 1.0 * W0, 1.9 * W0, -3.0 * W0])
    ky = np.array([0.0, W0, 0.0, 0.5 * W0, 0.2 * W0, 0.0])
    g, mask = tab.ghat(kx, ky, 40.0, 50.0)
    assert tab.last_flagged == 4
    assert mask.cpu().tolist() == [True, True, True, False, False, True]  # k = omega twice, beyond k_max twice
    ex = _exact(stack, kx, ky, 40.0, 50.0)
    assert _same(g[mask], ex[mask])
    assert float((g[~mask] - ex[~mask]).abs().max() / ex[~mask].abs().max()) < 1e-8
    # without the fallback the flagged rows are left alone and counted
    g2 = torch.full((9, len(kx)), 7.0 + 0j, dtype=torch.complex128, device="cuda")
    from yellowthinfilms_amd import _lib
    import ctypes as ct
    from yellowthinfilms_amd.green import layer_structs
    flags = torch.empty(len(kx), dtype=torch.uint8, device="cuda")
    kxd, kyd = torch.as_tensor(kx, device="cuda"), torch.as_tensor(ky, device="cuda")
    arrs = [ct.c_void_p(t.data_ptr()) for t in tab._dev]
    nf = ct.c_int64()
    p = lambda t: ct.c_void_p(t.data_ptr())
    ctx = _lib.Context.get(0)
    ctx.bind_stream()
    _lib.check(_lib.lib().ytf_ktable_ghat(ctx.handle, layer_structs(stack), len(stack), W0, tab.n_panels, *arrs,
                                          p(kxd), p(kyd), len(kx), 0.15, 0.0667, 0.9, 0, p(g2), p(flags),
                                          ct.byref(nf)), ctx.handle)
    assert nf.value == 4 and torch.all(g2[:, flags.bool()] == 7.0)
    with pytest.raises(_lib.YtfError, match="bad table"):
        _lib.check(_lib.lib().ytf_ktable_ghat(ctx.handle, layer_structs(stack), len(stack), W0, 0, *arrs, p(kxd),
                                              p(kyd), len(kx), 0.15, 0.0667, 0.9, 0, p(g2), p(flags), None),
                   ctx.handle)


def test_ktable_on_the_surrogate_workload_grid():
    """The 1024 x 1024 grid of the surrogate workload (|kx|, |ky| <= 3 omega,
    gyrotropic family at loss 0.1, heights 40 / 50 nm): within 1e-8 of the
    exact solve of every node and at least 3x its speed (the table's own
    build, ~10^3 exact solves, is not in the timed region, like a network's
    training)."""
    from yellowthinfilms_amd.ktable import KrhoTable
    from yellowthinfilms_amd.surrogate import GyroFamily
    from yellowthinfilms_amd import uniform_kxky_grid, workloads as W
    print("case: k_rho table on the 1024^2 surrogate grid", flush=True)
    wl = W.SURROGATE_1024
    grid = uniform_kxky_grid(wl["L"] / wl["f"], wl["dx"] / wl["f"])
    kx = torch.as_tensor(np.repeat(grid["kx"], grid["Ny"]), device="cuda")
    ky = torch.as_tensor(np.tile(grid["ky"], grid["Nx"]), device="cuda")
    stack = GyroFamily(seed=wl["family_seed"]).stack(wl["loss"])
    kmax = float(torch.sqrt(kx * kx + ky * ky).max()) * 1.001
    tab = KrhoTable(stack, W0, kmax, tol=wl["ktable_tol"], deg=wl["ktable_deg"])
    hs, ho = wl["h_src_nm"], wl["h_obs_nm"]
    ex = _exact(stack, kx, ky, hs, ho)
    out = (torch.empty((9, len(kx)), dtype=torch.complex128, device="cuda"),
           torch.empty(len(kx), dtype=torch.uint8, device="cuda"))

    def ms(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps
    t_tab = ms(lambda: tab.ghat(kx, ky, hs, ho, count=False, out=out))
    t_ex = ms(lambda: _exact(stack, kx, ky, hs, ho))
    g = out[0].T
    err = float(torch.linalg.norm(g - ex) / torch.linalg.norm(ex))
    print(f"  {tab.n_panels} panels, {tab.val.shape[0]} rows; rel err {err:.2e}; {t_tab:.3f} ms against exact "
          f"{t_ex:.3f} ms = {t_ex / t_tab:.1f}x; {int(out[1].sum())} flagged", flush=True)
    assert err < 1e-8
    assert t_ex / t_tab >= 3.0
