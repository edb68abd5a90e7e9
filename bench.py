"""Benchmarks of the GPU Green-operator path on one MI355X.

    python bench.py [--workload W] [--steps K] [--warmup W] [--no-cpu-baseline]

Workloads (yellowthinfilms_amd.workloads):
  batched_gyro_512 (default, the headline line): a 5-layer stack with a 40 nm
    gyrotropic film on the uniform 512 x 512 (kx, ky) grid.  One step = one
    Green k-node per grid node (262,144 node solves -> spectral dyadic g~,
    ytf_green_nodes) + one FFT convolution of a 512 x 512 x 3 complex N(0,1)
    source (ytf_spectral_apply).  value = Green k-nodes/s of the whole step.
  point_uniaxial_1d: air / uniaxial film / glass, 4608 k-rho nodes, the 1D
    closed-form G at 256 radii x 8 heights (node solve + FP64 Bessel
    synthesis).  value = Green k-nodes/s of the whole step.
  bessel_gemm: 16384 k-rho nodes x 8192 radii x 64 depth pairs; node solve
    (FP64) + the Bessel contraction on the matrix cores (split fp16 products).
    value = contraction TFLOP/s (algorithmic) of the whole step.
  uneven_256: 256 stacks of 2..24 hyperbolic layers with 1k..16k nodes each,
    modes + total S-matrix per node in one persistent (graph) launch whose
    shape the autotuner picks first.  value = stack-nodes/s.
  surrogate_1024: the learned surrogate of the scattered spectral dyadic,
    trained first (2^22 solver samples, untimed), then on a 1024 x 1024 grid:
    network + flags + exact FP64 fallback, against the plain GPU solve of the
    same nodes.  value = Green k-nodes/s of the surrogate path.
  nufft_1e6: type-1 NUFFT of 10^6 seeded points onto 1024^2 modes, eps 1e-9.
    value = applies/s.
  operator_512: the padded (linear-convolution) Green operator of the
    batched_gyro_512 stack and grid, one source plane onto one observation
    plane, built once (untimed) from 262,144 polar nodes with the NUFFT
    lateral sum; one step = one apply_operator of a 512 x 512 x 3 complex
    N(0,1) moment sheet.  value = operator applies/s.
  operator_volume_64: the reference's own operator benchmark at size L
    (Bench_GreenOperator.jl, iso vac/glass/vac): 16 box-source planes of
    layer 2 onto the 16 planes of each of layers 1-3 on a 64 x 64 grid
    (padded 128^2), z skeletons on, built on the benchmark's own grid
    (box_krho_grid_grouped, N_phi 256, rtol 1e-3, k_rho cap 3000) with the
    phi-rotation fill; one step = one apply_operator of a full
    64 x 64 x 16 x 3 moment volume.  value = operator applies/s;
    vs_baseline against the reference's 90 ms per apply.
  pole_scan_hyperbolic: the real-axis pole search of a source in the
    hyperbolic biaxial layer of the reference's operator benchmark, 9
    directions, k_rho from 0.5 k0 to 3000 at a relative step of 1e-3, then
    the refinement of every maximum.  value = cavity-response evaluations/s
    of the whole search.
  probe_batch_box: the probe integrands of the adaptive box k_rho grid on
    the reference's grid validation setup (8 box probes, 96 phi nodes) for
    one batch of 4096 k_rho.  value = k_rho probe sets/s (the reference
    builds one one-node cache per visited k_rho, ~1.3 ms each).
  box_krho_grid: the whole adaptive box k_rho grid of a 1/8-wavelength
    voxel in the glass layer of the same stack (22 probes, 64 phi, rtol
    1e-6, k_rho cap 2000): pole scan, seeds and the lock-step refinement.
    value = distinct k_rho solved per second of grid building.

Inputs are resident in HBM before the timed region.  Each line carries a
roofline record of the dominant kernel, timed live with HIP events on the
stream it is launched on (the library's kernel timing), and the CPU baseline
(the FP64 oracle on this host's cores), whose leg also gives the relative
error of the GPU result against the oracle.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64 = 78.6e12    # MI355X FP64 (vector and matrix), FLOP/s
PEAK_FP16 = 2.5e15     # MI355X dense fp16/bf16 matrix, FLOP/s
PEAK_HBM = 8.0e12      # bytes/s
# FP64 LDS atomic updates per second: 7 ds_add_f64 lane-updates per clock per CU
# (an LDS probe, 64 lanes on 64 distinct doubles), 256 CUs, 2.4 GHz
LDS_F64_ATOMIC_PEAK = 7 * 256 * 2.4e9
METRIC = "Green k-nodes/s + operator applies/s, surrogate vs plain GPU, <=1e-6 rel err"


def _cores():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def _timed_steps(a, step, torch):
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _timed_kernels(a, step, torch, ctx, roof, others=(), events_in_wall=True):
    """The K timed steps with HIP events on the launches named in `roof` (a
    name or a tuple of names: the roofline's kernels) and on no other launch,
    since every event pair adds a few microseconds to the wall clock; then,
    if `others` names more kernels, a second pass of K steps outside the wall
    clock with every launch bracketed, for their per-kernel times.  With
    events_in_wall=False the timed steps carry no events at all and the
    `roof` kernels are timed in the second pass too (for a rate that is
    compared with one measured without events).  Returns (wall seconds of the
    timed steps, {name: (total ms, launches)})."""
    roof = (roof,) if isinstance(roof, str) else tuple(roof)
    if not events_in_wall:
        ctx.set_timing(False)
        wall = _timed_steps(dict_ns(a, warmup=0), step, torch)
        ctx.set_timing(True)
        for _ in range(a.steps):
            step()
        t = {k: ctx.kernel_time(k) for k in roof + tuple(others)}
        ctx.set_timing(False)
        return wall, t
    ctx.set_timing_only(",".join(roof))
    wall = _timed_steps(dict_ns(a, warmup=0), step, torch)
    t = {k: ctx.kernel_time(k) for k in roof}
    if others:
        ctx.set_timing(True)
        for _ in range(a.steps):
            step()
        t.update({k: ctx.kernel_time(k) for k in others if k not in roof})
    ctx.set_timing(False)
    return wall, t


def _calib_copy(a, torch, dev):
    # one device-to-device copy of known size: the scale of the FETCH_SIZE /
    # WRITE_SIZE counters in a profiling run
    if a.calib_copy_mb > 0:
        src = torch.empty(a.calib_copy_mb * 2 ** 20 // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst


def _base(a, wl, value, unit, wall, dtype):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": wall * 1e3 / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic (seeded inputs; stack, rule and grid as configured)",
            "config": {"workload": wl["description"], "name": wl["name"]}}


# ---------------------------------------------------------------------------
def bench_gyro(a):
    import torch
    from yellowthinfilms_amd import uniform_kxky_grid, SpectralGreenOperator, workloads as W
    from yellowthinfilms_amd._lib import Context

    wl = W.BATCHED_GYRO_512
    grid = uniform_kxky_grid(wl["L"], wl["dx"])
    nx, ny = grid["Nx"], grid["Ny"]
    nn = nx * ny
    stack = W.gyro_512_stack()
    dev = torch.device("cuda", 0)
    op = SpectralGreenOperator(stack, W.OMEGA, grid, wl["source_layer"], wl["zsrc"], wl["obs_layer"], wl["zobs"])
    J = torch.as_tensor(W.uniform_c128_source(nx, ny, wl["seed"]), device=dev)
    E = torch.empty_like(J)
    ctx = Context.get(0)

    def step():
        op.build()
        op.apply(J, out=E)

    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "node_solve", ("spectral_apply",))
    node_ms, nlaunch = t["node_solve"]
    apply_ms, _ = t["spectral_apply"]
    node_ms /= nlaunch
    apply_ms /= a.steps
    _calib_copy(a, torch, dev)

    flops = W.BATCHED_GYRO_512_FLOPS_PER_NODE * nn
    achieved = flops / (node_ms * 1e-3)
    executed = W.BATCHED_GYRO_512_EXECUTED_FLOPS_PER_NODE * nn / (node_ms * 1e-3)
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        cpu, rel_err = cpu_gyro(stack, grid, wl, a.cpu_repeats, op.ghat[0].cpu().numpy())
    apply_bytes = 16.0 * (3 * nn + 9 * nn + 3 * nn)
    out = _base(a, wl, nn * a.steps / wall, "Green k-nodes/s", wall, "f64")
    out["config"].update({"nodes_per_step": nn, "grid": [nx, ny], "path": "plain GPU (ytf_green_nodes + ytf_spectral_apply)"})
    out.update({
        "node_solve_ms": node_ms, "apply_ms": apply_ms, "applies_per_s": 1e3 / apply_ms,
        "node_solve_knodes_per_s": nn / (node_ms * 1e-3),
        "apply_hbm_gbps_algorithmic": apply_bytes / (apply_ms * 1e-3) / 1e9,
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "node_solve (ytf node_kernel, Green k-node solve, FP64)", "bound": "valu-fp64",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": W.BATCHED_GYRO_512_NODE_HBM_BYTES,
                     "flops_per_node": W.BATCHED_GYRO_512_FLOPS_PER_NODE, "nodes_per_launch": nn,
                     "flops_basis": "reference formulation (oracle op count); the kernel executes fewer",
                     "executed_flops_per_node": W.BATCHED_GYRO_512_EXECUTED_FLOPS_PER_NODE,
                     "achieved_executed": executed / 1e12, "frac_executed": executed / PEAK_FP64},
        "cpu_baseline": cpu,
    })
    if not a.no_surrogate:
        # the metric's other half: the configured surrogate (surrogate_1024) against the plain GPU solve
        sw = dict_ns(a, steps=min(a.steps, 10), warmup=min(a.warmup, 2))
        sr, sf = surrogate_runs(sw, [(W.SURROGATE_1024["width"], W.SURROGATE_1024["depth"])])
        r, kt = sr[0], sf["ktable"]
        # the surrogate that meets the metric's error bar is the tabulated k_rho surrogate
        # (exact solves, no learning); the learned network is reported beside it
        out.update({
            "surrogate_speedup_vs_plain": kt["speedup_vs_plain_gpu"],
            "surrogate_rel_err_vs_fp64": kt["rel_err_ghat_vs_fp64"],
            "surrogate_field_rel_err_vs_fp64": kt["rel_err_field_vs_fp64"],
            "flagged_fraction": kt["flagged_fraction"],
            "surrogate_knodes_per_s": kt["knodes_per_s"],
            "surrogate_kernel_ms": kt["kernel_ms"],
            "surrogate_config": (f"surrogate_1024 grid ({sf['grid'][0]}^2 nodes), tabulated k_rho surrogate: t(k_rho) "
                                 f"from {kt['exact_samples']} exact solves at phi = 0 on {kt['panels']} Chebyshev "
                                 f"panels (tol {kt['tol']:g}), cubic interpolation on {kt['table_rows']} rows, "
                                 "rotation about z, exact fallback on flagged nodes; no learning"),
            "mlp_surrogate": {
                "speedup_vs_plain": r["speedup_vs_plain_gpu"], "rel_err_vs_fp64": r["rel_err_ghat_vs_fp64"],
                "field_rel_err_vs_fp64": r["rel_err_field_vs_fp64"], "flagged_fraction": r["flagged_fraction"],
                "knodes_per_s": r["knodes_per_s"],
                "config": (f"MLP 6->{r['width']}x{r['depth']}->18 BF16, {sf['train_samples']} samples x "
                           f"{sf['epochs']} epochs, exact fallback on light line, origin, cavity-response pole "
                           "windows, large |t|")},
        })
    return out


def cpu_gyro(stack, grid, wl, repeats, gpu_ghat):
    """The FP64 oracle (C++/OpenMP restatement of the reference node loop and
    strength) over the same 262,144 grid nodes on this host's cores: median of
    `repeats` full passes after one warm-up pass on 4096 nodes.  Also returns
    the relative error of the GPU's g~ (all nodes) against the oracle's."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "Green k-nodes/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    from yellowthinfilms_amd.materials import stack_arrays
    from yellowthinfilms_amd import workloads as W
    cores = _cores()
    nx, ny = grid["Nx"], grid["Ny"]
    KX = np.repeat(grid["kx"], ny)
    KY = np.tile(grid["ky"], nx)
    st = po.Stack(*stack_arrays(stack))
    args = (wl["source_layer"], wl["obs_layer"], [wl["zsrc"]], [wl["zobs"]])
    po.ghat_batch(st, W.OMEGA, KX[:4096], KY[:4096], *args, nthreads=cores)
    ts = []
    og = None
    for _ in range(repeats):
        t = time.perf_counter()
        og = po.ghat_batch(st, W.OMEGA, KX, KY, *args, nthreads=cores)
        ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    g = gpu_ghat.reshape(3, 3, -1).transpose(2, 1, 0)  # (node, c, q)
    rel_err = float(np.linalg.norm(g - og[:, 0]) / np.linalg.norm(og[:, 0]))
    return ({"value": nx * ny / dt, "unit": "Green k-nodes/s", "cores": cores, "kind": "port",
             "sample": f"all {nx * ny} nodes of the same grid (node solve + spectral dyadic), median of {repeats}"},
            rel_err)


# ---------------------------------------------------------------------------
def bench_point(a):
    import torch
    from yellowthinfilms_amd import CavityTransfer1D, closed_form_G, workloads as W
    from yellowthinfilms_amd._lib import Context

    wl = W.POINT_UNIAXIAL_1D
    kx, wt = W.point_uniaxial_rule()
    zo, rho = W.point_uniaxial_geometry()
    stack = W.point_uniaxial_stack()
    cache = CavityTransfer1D(stack, W.OMEGA, kx, wt, wl["source_layer"])
    res = {}

    def step():
        res["G"] = closed_form_G(cache, wl["obs_layer"], wl["zsrc"], zo, rho, [0.0], mode="fp64")

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "synth_fp64", ("node_solve", "synth_coef", "synth_reduce",
                                                                  "synth_rotate"))
    n, nr, nz = len(kx), len(rho), len(zo)
    syn_ms = t["synth_fp64"][0] / t["synth_fp64"][1]
    flops = 36.0 * n * nr * nz  # 9 complex coefficients x real J per (node, radius, depth pair)
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        G = res["G"].cpu().numpy().transpose(4, 3, 2, 1, 0)
        cpu, rel_err = cpu_point(stack, kx, wt, wl, zo, rho, G, a.cpu_repeats)
    out = _base(a, wl, n * a.steps / wall, "Green k-nodes/s", wall, "f64")
    out["config"].update({"nodes": n, "radii": nr, "heights": nz, "path": "ytf_green_nodes + ytf_closed_form_G_1d"})
    out.update({
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in t.items()},
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "synth_fp64 (radial Bessel contraction, FP64: Bessel table, node-chunk contraction, chunk sum)", "bound": "valu-fp64",
                     "achieved": flops / (syn_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": flops / (syn_ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "flops_per_launch": flops, "bessel_evals_per_launch": 2.0 * n * nr},
        "cpu_baseline": cpu,
    })
    return out


def cpu_point(stack, kx, wt, wl, zo, rho, G, repeats):
    """The FP64 oracle over the whole workload (node solve + 1D synthesis),
    median of `repeats`; also the relative error of the GPU's G against it."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "Green k-nodes/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    from yellowthinfilms_amd.materials import stack_arrays
    from yellowthinfilms_amd import workloads as W
    cores = _cores()
    st = po.Stack(*stack_arrays(stack))
    ts, Go = [], None
    for _ in range(max(1, repeats)):
        t = time.perf_counter()
        c = po.nodes_batch(st, W.OMEGA, kx, np.zeros_like(kx), wl["source_layer"], [wl["obs_layer"]], nthreads=cores)
        Go = po.closed_form_G_1d(c, wl["obs_layer"], wt, wl["zsrc"], zo, rho, [0.0], nthreads=cores)
        ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    rel = float(np.linalg.norm(G - Go) / np.linalg.norm(Go))
    return ({"value": len(kx) / dt, "unit": "Green k-nodes/s", "cores": cores, "kind": "port",
             "sample": "the whole workload (node solve + 1D synthesis at all radii and heights)"}, rel)


# ---------------------------------------------------------------------------
def bench_gemm(a):
    import torch
    from yellowthinfilms_amd import green_nodes, workloads as W
    from yellowthinfilms_amd.synthesis import CavityTransfer1D, radial_synthesis
    from yellowthinfilms_amd._lib import Context

    wl = W.BESSEL_GEMM
    kx, wt = W.bessel_gemm_rule()
    zs, zo, rho = W.bessel_gemm_geometry()
    stack = W.point_uniaxial_stack()
    cache = CavityTransfer1D(stack, W.OMEGA, kx, wt, wl["source_layer"])
    n, nr, npairs = len(kx), len(rho), len(zs)
    res = {}
    g = torch.empty((npairs, 9, n), dtype=torch.complex128, device="cuda")

    def step():
        green_nodes(stack, W.OMEGA, cache.kx, cache.ky, 1, wl["obs_layer"], zs, zo, cache=False, ghat_out=g)
        res["R"] = radial_synthesis(cache, g, rho, mode="f16x3")

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    names = ("node_solve", "synth_coef", "bessel_table", "split_columns", "bessel_gemm", "mfma_combine")
    wall, t = _timed_kernels(a, step, torch, ctx, "bessel_gemm", names)
    gemm_ms = t["bessel_gemm"][0] / t["bessel_gemm"][1]
    alg = 4.0 * nr * n * 9 * npairs          # real J x complex coefficient, 9 complex columns per pair
    issued = 3 * 2.0 * nr * n * 18 * npairs  # three fp16 products per term on the real GEMM
    # accuracy against the FP64 contraction of the same g~ (GPU path pinned to the oracle)
    R64 = radial_synthesis(cache, g, rho, mode="fp64")
    rel_fp64 = float(torch.linalg.norm(res["R"] - R64) / torch.linalg.norm(R64))
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        cpu, rel_err = cpu_gemm(stack, kx, wt, wl, zs, zo, rho, res["R"].cpu().numpy())
    out = _base(a, wl, alg * a.steps / wall / 1e12, "TFLOP/s (algorithmic contraction, whole step)", wall,
                "f16x3 (split fp16, fp32 accumulate) + f64")
    out["config"].update({"nodes": n, "radii": nr, "depth_pairs": npairs,
                          "path": "ytf_green_nodes (FP64) + ytf_radial_synthesis(YTF_SYNTH_F16X3)"})
    out.update({
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in t.items()},
        "rel_err_vs_gpu_fp64": rel_fp64,
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "bessel_gemm (split-fp16 MFMA GEMM)", "bound": "mfma",
                     "achieved": alg / (gemm_ms * 1e-3) / 1e12, "peak": PEAK_FP16 / 1e12, "unit": "TFLOP/s",
                     "frac": alg / (gemm_ms * 1e-3) / PEAK_FP16, "traffic": None,
                     "flops_per_launch": alg, "mfma_flops_issued_per_launch": issued,
                     "mfma_issued_tflops": issued / (gemm_ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
    })
    return out


def cpu_gemm(stack, kx, wt, wl, zs, zo, rho, R):
    """The oracle's 1D synthesis (FP64, OpenMP over radii) on a bounded sample:
    every 32nd radius and the 8 observation heights of one source height."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "synthesis terms/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    from yellowthinfilms_amd.materials import stack_arrays
    from yellowthinfilms_amd import workloads as W
    cores = _cores()
    st = po.Stack(*stack_arrays(stack))
    nobs = len(wl["obs_heights_nm"])
    sel = np.arange(0, len(rho), 32)
    src_i = 3  # source height index
    pairs = np.arange(src_i * nobs, (src_i + 1) * nobs)
    c = po.nodes_batch(st, W.OMEGA, kx, np.zeros_like(kx), wl["source_layer"], [wl["obs_layer"]], nthreads=cores)
    t = time.perf_counter()
    Go = po.closed_form_G_1d(c, wl["obs_layer"], wt, zs[pairs[0]], zo[pairs], rho[sel], [0.0], nthreads=cores)
    dt = time.perf_counter() - t
    Go = Go[:, 0]  # (nr_s, nz, 3, 3)
    Rg = R[pairs][:, :, sel].reshape(nobs, 3, 3, len(sel)).transpose(3, 0, 2, 1)  # (nr_s, nz, c, q)
    rel = float(np.linalg.norm(Rg - Go) / np.linalg.norm(Go))
    terms = len(sel) * nobs * len(kx)
    return ({"value": terms / dt, "unit": "synthesis terms (radius x node x depth pair)/s", "cores": cores,
             "kind": "port", "sample": f"{len(sel)} radii x {nobs} depth pairs x {len(kx)} nodes (one source height)"},
            rel)


# ---------------------------------------------------------------------------
def bench_uneven(a):
    import torch
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd.batch import StackBatch

    wl = W.UNEVEN_256
    stacks, counts, kx, ky = W.uneven_batch()
    dev = torch.device("cuda", 0)
    kxd = torch.as_tensor(kx, device=dev)
    kyd = torch.as_tensor(ky, device=dev)
    b = StackBatch(stacks, counts, W.OMEGA)
    report = b.tune(kxd, kyd, tune_db=a.tune_db)  # before the timed region
    out = b.outputs()

    def step():
        b.solve(kxd, kyd, out=out)

    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, b.ctx, "batch_solve")
    ms, nl = t["batch_solve"]
    ms /= nl
    n = len(kx)
    flops = W.UNEVEN_256_FLOPS_PER_LAUNCH
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        cpu, rel_err = cpu_uneven(stacks, counts, kx, ky, out)
    valid = float((out["flags"] & 1).sum().item()) / n
    res = _base(a, wl, n * a.steps / wall, "stack-nodes/s (modes + total S-matrix)", wall, "f64")
    res["config"].update({"stacks": len(stacks), "nodes": n, "launch": b.config, "tuner": report[:400]})
    res.update({
        "batch_solve_ms": ms, "valid_fraction": valid,
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "batch_solve (persistent uneven-batch kernel, one graph launch, FP64)",
                     "bound": "valu-fp64", "achieved": flops / (ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": flops / (ms * 1e-3) / PEAK_FP64, "traffic": None,
                     "flops_per_launch": flops, "nodes_per_launch": n},
        "cpu_baseline": cpu,
    })
    return res


def cpu_uneven(stacks, counts, kx, ky, out):
    """The oracle (modes + S_all per node, OpenMP) on every 64th node of every
    stack; the GPU's S on the same nodes is compared with it."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "stack-nodes/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    from yellowthinfilms_amd.materials import stack_arrays
    from yellowthinfilms_amd import workloads as W
    cores = _cores()
    S = out["S"].cpu().numpy()
    off = np.concatenate([[0], np.cumsum(counts)])
    dt, m, num, den = 0.0, 0, 0.0, 0.0
    for s, st in enumerate(stacks):
        idx = np.arange(off[s], off[s + 1], 64)
        o = po.Stack(*stack_arrays(st))
        t = time.perf_counter()
        r = po.stack_S_batch(o, W.OMEGA, kx[idx], ky[idx], nthreads=cores)
        dt += time.perf_counter() - t
        m += len(idx)
        g = S[:, idx].T.reshape(-1, 4, 4).transpose(0, 2, 1)
        v = r["ok"]
        num += float(np.sum(np.abs(g[v] - r["S"][v]) ** 2))
        den += float(np.sum(np.abs(r["S"][v]) ** 2))
    return ({"value": m / dt, "unit": "stack-nodes/s", "cores": cores, "kind": "port",
             "sample": f"every 64th node of every stack ({m} nodes)"}, float(np.sqrt(num / den)))


# ---------------------------------------------------------------------------
def surrogate_runs(a, configs, windows=None):
    """The configured surrogate workload: 2^22 exact-solver samples (generated
    once), then per (width, depth) in `configs` a network trained on them and
    the surrogate path (network + flags + exact fallback) timed against the
    plain GPU solve of the same 1024^2 grid, with g~ and field errors against
    the exact FP64 path.  Returns (per-config results, shared facts)."""
    import torch
    from yellowthinfilms_amd import uniform_kxky_grid, green_nodes, workloads as W
    from yellowthinfilms_amd.surrogate import GyroFamily, SurrogateGreen
    from yellowthinfilms_amd._lib import Context, lib, check

    wl = W.SURROGATE_1024
    fam = GyroFamily(seed=wl["family_seed"])
    gen = SurrogateGreen(fam, width=wl["width"], depth=wl["depth"], seed=wl["seed"])
    t0 = time.perf_counter()
    x, t = gen.generate(wl["n_train"], seed=wl["data_seed"])
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    omega = 2 * math.pi * wl["f"]
    grid = uniform_kxky_grid(wl["L"] / wl["f"], wl["dx"] / wl["f"])
    nx, ny = grid["Nx"], grid["Ny"]
    dev = torch.device("cuda", 0)
    KX = torch.as_tensor(np.repeat(grid["kx"], ny), device=dev)
    KY = torch.as_tensor(np.tile(grid["ky"], nx), device=dev)
    hs, ho, loss = wl["h_src_nm"], wl["h_obs_nm"], wl["loss"]
    st = fam.stack(loss)
    nn = nx * ny
    g_exact = torch.empty((1, 9, nn), dtype=torch.complex128, device=dev)
    ctx = Context.get(0)

    def plain():
        green_nodes(st, omega, KX, KY, 1, 1, [hs * W.NM], [1.0 - ho * W.NM], cache=False, ghat_out=g_exact,
                    terms="scattered")

    for _ in range(a.warmup):
        plain()
    wall_plain = _timed_steps(dict_ns(a, warmup=0), plain, torch)
    plain_rate = nn * a.steps / wall_plain
    ge = g_exact[0].T.contiguous()            # (nn, 9)
    J = torch.as_tensor(W.uniform_c128_source(nx, ny, wl["seed"]), device=dev)
    work = torch.empty((3, nx, ny), dtype=torch.complex128, device=dev)
    scale = grid["dx"] * grid["dy"] * grid["dkx"] * grid["dky"] / (2 * math.pi) ** 2

    def field(g9n):
        E = torch.empty_like(J)
        gg = g9n.T.contiguous()  # [9][nn]
        ctx.bind_stream()
        import ctypes as ct
        check(lib().ytf_spectral_apply(ctx.handle, nx, ny, ct.c_void_p(gg.data_ptr()), ct.c_void_p(J.data_ptr()),
                                       ct.c_void_p(E.data_ptr()), ct.c_void_p(work.data_ptr()), scale), ctx.handle)
        return E

    E_ex = field(ge)
    kr = torch.sqrt(KX * KX + KY * KY)
    # the tabulated k_rho surrogate of the same stack (rotation invariant about z): table built
    # from exact solves at phi = 0, then one evaluation kernel over the grid
    from yellowthinfilms_amd.ktable import KrhoTable
    tab = KrhoTable(st, omega, float(kr.max()) * 1.001, tol=wl["ktable_tol"], deg=wl["ktable_deg"])
    gt_out = (torch.empty((9, nn), dtype=torch.complex128, device=dev), torch.empty(nn, dtype=torch.uint8, device=dev))

    def table():
        tab.ghat(KX, KY, hs, ho, fallback=True, count=False, out=gt_out)

    for _ in range(a.warmup):
        table()
    # the plain solve above is timed without events, and so is this
    wall_tab, tk = _timed_kernels(a, table, torch, ctx, "ktable", events_in_wall=False)
    ktab_ms, _ = tk["ktable"]
    gt, mt = gt_out[0].T, gt_out[1].bool()
    ktable = {
        "knodes_per_s": nn * a.steps / wall_tab, "speedup_vs_plain_gpu": (nn * a.steps / wall_tab) / plain_rate,
        "rel_err_ghat_vs_fp64": float(torch.linalg.norm(gt - ge) / torch.linalg.norm(ge)),
        "rel_err_field_vs_fp64": float(torch.linalg.norm(field(gt) - E_ex) / torch.linalg.norm(E_ex)),
        "flagged_fraction": float(mt.float().mean()), "kernel_ms": ktab_ms / a.steps,
        "write_gbps": 9 * 16 * nn / (ktab_ms / a.steps * 1e-3) / 1e9,
        "build_s": tab.build_s, "panels": tab.n_panels, "unconverged_panels": tab.n_unconverged,
        "table_rows": int(tab.val.shape[0]), "table_bytes": int(tab.val.nbytes),
        "exact_samples": tab.n_exact, "tol": wl["ktable_tol"], "deg": wl["ktable_deg"],
    }
    del gt_out, gt
    windows = [wl["pole_window"]] if windows is None else windows
    poles, win_conf = gen.pole_windows(omega, loss, wl["pole_window"])
    in_conf = torch.zeros(nn, dtype=torch.bool, device=dev)
    for lo, hi in win_conf:
        in_conf |= (kr >= lo) & (kr <= hi)
    results = []
    for width, depth in configs:
        sg = gen if (width, depth) == (wl["width"], wl["depth"]) else SurrogateGreen(fam, width=width, depth=depth,
                                                                                      seed=wl["seed"])
        t0 = time.perf_counter()
        hist = sg.fit(x, t, epochs=wl["epochs"], batch=wl["batch"], lr=wl["lr"])
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        for pw in windows:
            res = {}
            _, win = gen.pole_windows(omega, loss, pw) if pw > 0 else (None, np.zeros((0, 2)))

            def surrogate():
                res["g"], res["mask"] = sg.scattered_ghat(KX, KY, hs, ho, omega, loss, fallback=True, count=False,
                                                          pole_window=pw)

            for _ in range(a.warmup):
                surrogate()
            wall_sur, tk = _timed_kernels(a, surrogate, torch, ctx, "mlp_fused", events_in_wall=False)
            gemm_ms, _ = tk["mlp_fused"]
            gs, mask = res["g"], res["mask"]
            in_win = torch.zeros_like(mask)
            for lo, hi in win:
                in_win |= (kr >= lo) & (kr <= hi)
            resp = mask & ~in_win & ~(((kr - omega).abs() < sg.BRANCH_TOL * omega) | (kr < sg.ORIGIN_TOL * omega))
            results.append({
                "width": width, "depth": depth, "pole_window": pw, "flops_per_sample": sg.mlp.flops_per_sample,
                "knodes_per_s": nn * a.steps / wall_sur,
                "speedup_vs_plain_gpu": (nn * a.steps / wall_sur) / plain_rate,
                "rel_err_ghat_vs_fp64": float(torch.linalg.norm(gs - ge) / torch.linalg.norm(ge)),
                "rel_err_field_vs_fp64": float(torch.linalg.norm(field(gs) - E_ex) / torch.linalg.norm(E_ex)),
                "flagged_fraction": float(mask.float().mean()),
                "flagged_pole_windows": float(in_win.float().mean()),
                "flagged_response": float(resp.float().mean()),
                # of the nodes within the configured pole windows, the share this run flags
                "flagged_of_configured_window_nodes": float((mask & in_conf).float().sum() /
                                                            in_conf.float().sum().clamp_min(1)),
                "mlp_ms_per_forward": gemm_ms / a.steps, "train_s": t_train,
                "train_loss_per_epoch": [round(h, 6) for h in hist],
            })
        del sg
    facts = {"nodes": nn, "grid": [nx, ny], "plain_gpu_knodes_per_s": plain_rate, "data_s": t_data,
             "train_samples": wl["n_train"], "epochs": wl["epochs"], "ktable": ktable,
             "poles": [(round(p["q"] / omega, 6), round(p["Gamma"] / omega, 6)) for p in poles]}
    return results, facts


def bench_surrogate(a):
    from yellowthinfilms_amd import workloads as W
    wl = W.SURROGATE_1024
    configs = [(wl["width"], wl["depth"])]
    for c in (a.sur_sweep.split(",") if a.sur_sweep else []):
        w, d = (int(v) for v in c.lower().split("x"))
        if (w, d) not in configs:
            configs.append((w, d))
    wins = [wl["pole_window"]] + [float(v) for v in a.sur_windows.split(",") if v.strip() and
                                  float(v) != wl["pole_window"]] if a.sur_windows else None
    results, facts = surrogate_runs(a, configs, wins)
    main = results[0]
    wall = a.steps * facts["nodes"] / main["knodes_per_s"]
    out = _base(a, wl, main["knodes_per_s"], "Green k-nodes/s (surrogate + exact fallback)", wall,
                "bf16 (MLP, fp32 accumulate) + f64 (fallback)")
    out["config"].update({"nodes": facts["nodes"], "grid": facts["grid"],
                          "mlp": f"6->{wl['width']}x{wl['depth']}->18 SiLU", "train_samples": wl["n_train"],
                          "epochs": wl["epochs"]})
    out.update({k: v for k, v in main.items() if k not in ("width", "depth")})
    out["config"]["pole_window_halfwidths"] = main["pole_window"]
    out.update({"plain_gpu_knodes_per_s": facts["plain_gpu_knodes_per_s"], "data_s": facts["data_s"],
                "poles": facts["poles"], "tradeoff": results[1:], "ktable": facts["ktable"]})
    out["roofline"] = {"kernel": "mlp_fused (the whole network in one BF16 matrix-core kernel, with the exact "
                                 "factors and flags of the surrogate)",
                       "bound": "mfma", "achieved": main["flops_per_sample"] * facts["nodes"] /
                       (main["mlp_ms_per_forward"] * 1e-3) / 1e12, "peak": PEAK_FP16 / 1e12, "unit": "TFLOP/s",
                       "frac": main["flops_per_sample"] * facts["nodes"] / (main["mlp_ms_per_forward"] * 1e-3) /
                       PEAK_FP16, "traffic": None, "flops_per_sample": main["flops_per_sample"],
                       "samples_per_launch_set": facts["nodes"], "mlp_ms_per_forward": main["mlp_ms_per_forward"]}
    out["cpu_baseline"] = None
    return out


# ---------------------------------------------------------------------------
def bench_nufft(a):
    import torch
    from yellowthinfilms_amd import nufft2d1, workloads as W
    from yellowthinfilms_amd._lib import Context

    wl = W.NUFFT_1E6
    x, y, c = W.nufft_points(wl)
    dev = torch.device("cuda", 0)
    xd, yd, cd_ = (torch.as_tensor(v, device=dev) for v in (x, y, c))
    N = wl["N"]
    f = torch.empty((1, N, N), dtype=torch.complex128, device=dev)

    def step():
        nufft2d1(xd, yd, cd_, N, N, eps=wl["eps"], out=f)

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "nufft_apply", ("nufft_sort", "nufft_spread", "nufft_fft"))
    M = len(x)
    alg_bytes = M * (16 + 16) + N * N * 16  # points + strengths in, modes out
    ms = t.pop("nufft_apply")[0] / a.steps  # the whole transform, one event pair per apply
    w_nu = min(16, max(2, math.ceil(math.log10(1.0 / wl["eps"])) + 1))
    w_nu += w_nu & 1                             # the library's spreading width (even)
    spread_updates = 2.0 * w_nu * w_nu * M       # ds_add_f64 lane-updates (re, im) per transform
    spread_ms = t["nufft_spread"][0] / max(t["nufft_spread"][1], 1)
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        cpu, rel_err = cpu_nufft(x, y, c, f[0].cpu().numpy(), N)
    out = _base(a, wl, a.steps / wall, "NUFFT applies/s (10^6 points -> 1024^2 modes)", wall, "f64")
    out["config"].update({"points": M, "modes": [N, N], "eps": wl["eps"]})
    out.update({
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in t.items()},
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "nufft (sort + LDS-tiled spreading + FFT passes, one apply)", "bound": "hbm",
                     "achieved": alg_bytes / (ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": alg_bytes / (ms * 1e-3) / PEAK_HBM, "traffic": None, "algorithmic_bytes": alg_bytes},
        # what binds the spreading kernel: its FP64 LDS atomics (2 w^2 per point), against the
        # rate an LDS probe reaches (7 ds_add_f64 lane-updates per clock per CU, 256 CUs, 2.4 GHz)
        "spread_lds_atomics": {"updates": spread_updates, "achieved": spread_updates / (spread_ms * 1e-3),
                               "peak": LDS_F64_ATOMIC_PEAK, "unit": "lane-updates/s",
                               "frac": spread_updates / (spread_ms * 1e-3) / LDS_F64_ATOMIC_PEAK,
                               "width": w_nu, "spread_ms": spread_ms},
        "cpu_baseline": cpu,
    })
    return out


def cpu_nufft(x, y, c, f, N):
    """The oracle's direct sum (OpenMP) at 256 seeded modes; the GPU's value
    at those modes is compared with it."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "point-mode terms/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    cores = _cores()
    rng = np.random.default_rng(5)
    modes = np.stack([rng.integers(-N // 2, N // 2, 256), rng.integers(-N // 2, N // 2, 256)], axis=1)
    t = time.perf_counter()
    ref = po.nudft2d1(x, y, c, modes=modes, nthreads=cores)
    dt = time.perf_counter() - t
    got = f[modes[:, 1] + N // 2, modes[:, 0] + N // 2]
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return ({"value": len(x) * len(modes) / dt, "unit": "point-mode terms/s (direct sum)", "cores": cores,
             "kind": "port", "sample": f"{len(modes)} seeded modes x {len(x)} points"}, rel)


def bench_operator(a):
    import torch
    from yellowthinfilms_amd import workloads as W, uniform_kxky_grid
    from yellowthinfilms_amd.green_operator import CavityTransfer2D, build_layer_operator, apply_operator
    from yellowthinfilms_amd._lib import Context

    wl = W.OPERATOR_512
    stack = W.gyro_512_stack()
    grid = uniform_kxky_grid(wl["L"], wl["dx"])
    Nx, Ny = grid["Nx"], grid["Ny"]
    dev = torch.device("cuda", 0)
    cache = CavityTransfer2D(stack, W.OMEGA, wl["source_layer"], *W.operator_nodes(wl))
    t0 = time.perf_counter()
    op = build_layer_operator(cache, grid, [wl["zsrc"]], {wl["obs_layer"]: [wl["zobs"]]}, lateral="nufft",
                              nufft_eps=wl["eps"])
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    P = torch.as_tensor(W.uniform_c128_source(Nx, Ny, wl["seed"])[:, None], device=dev)  # (3, 1, Ny, Nx)

    def step():
        apply_operator(op, P)

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "op_apply", ("op_fft_fwd", "op_mix", "op_fft_inv"))
    Px, Py = 2 * Nx, 2 * Ny
    alg_bytes = 16 * (3 * Nx * Ny + 9 * Px * Py + 3 * Nx * Ny)  # read J and Ĝ, write E
    ms = t.pop("op_apply")[0] / a.steps  # the whole apply, one event pair
    cpu, rel_err, rel_apply = None, None, None
    if not a.no_cpu_baseline:
        cpu, rel_err, rel_apply = cpu_operator(stack, cache, op, P, grid, wl)
    out = _base(a, wl, a.steps / wall, "operator applies/s (512x512x3 sheet, padded 1024^2)", wall, "f64")
    out["config"].update({"grid": [Nx, Ny], "padded": [Px, Py], "nodes": cache.n_nodes,
                          "block": op.blocks[0]["kind"], "lateral": "nufft", "nufft_eps": wl["eps"]})
    out.update({
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in t.items()},
        "build_s": build_s,
        "rel_err_vs_fp64_oracle": rel_err,
        "rel_err_apply_vs_oracle_fft_route": rel_apply,
        "roofline": {"kernel": "operator apply (x FFT + fused y FFT / 3x3 product / inverse y FFT + inverse x FFT)",
                     "bound": "hbm", "achieved": alg_bytes / (ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": alg_bytes / (ms * 1e-3) / PEAK_HBM, "traffic": None,
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
    })
    return out


def cpu_operator(stack, cache, op, P, grid, wl):
    """Two checks and the CPU time of one apply.  (1) Four seeded dipoles:
    the GPU field at 64 seeded pixels against the oracle's exact lateral sum
    of its own spectral dyadics (mod_ClosedFormField.jl:743-753) at those
    offsets.  (2) The oracle's FFT route of the apply (numpy FFT,
    single-threaded) on the GPU-built Ĝ and the benchmark sheet: timed, and
    compared with the GPU apply."""
    import torch
    from yellowthinfilms_amd.green_operator import apply_operator
    try:
        from oracle import pyoracle as po
        from yellowthinfilms_amd.materials import stack_arrays
    except Exception as exc:
        return ({"value": None, "unit": "operator applies/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None, None)
    Nx, Ny = grid["Nx"], grid["Ny"]
    rng = np.random.default_rng(3)
    dip = [(int(rng.integers(Nx)), int(rng.integers(Ny)), rng.normal(size=3) + 1j * rng.normal(size=3))
           for _ in range(4)]
    Pv = np.zeros((3, 1, Ny, Nx), np.complex128)
    for i, j, p in dip:
        Pv[:, 0, j, i] += p
    Ev = apply_operator(op, torch.as_tensor(Pv, device=P.device))[wl["obs_layer"]][:, 0].cpu().numpy()  # (3, Ny, Nx)
    st = po.Stack(*stack_arrays(stack))
    g = po.ghat_batch(st, cache.omega, cache.kx_host, cache.ky_host, wl["source_layer"], wl["obs_layer"],
                      [wl["zsrc"]], [wl["zobs"]], nthreads=_cores())[:, 0]
    pix = [(int(rng.integers(Nx)), int(rng.integers(Ny))) for _ in range(64)]
    pts = np.array([((x - i) * grid["dx"], (y - j) * grid["dy"]) for (x, y) in pix for (i, j, _) in dip])
    Gp = po.lateral_sum_points(g, cache.W_host, cache.kx_host, cache.ky_host, pts).reshape(len(pix), len(dip), 3, 3)
    ref = np.array([sum(Gp[m, d] @ dip[d][2] for d in range(len(dip))) for m in range(len(pix))])
    got = np.array([Ev[:, y, x] for (x, y) in pix])
    rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # the apply on the CPU
    Gh = op.ghat_plane(0, "toeplitz", 0).cpu().numpy().transpose(3, 2, 1, 0)[:, :, None]  # (Px, Py, 1, 3, 3)
    J = P[:, 0].cpu().numpy().transpose(2, 1, 0)  # (Nx, Ny, 3) moments
    Eg = apply_operator(op, P)[wl["obs_layer"]][:, 0].cpu().numpy().transpose(2, 1, 0)
    reps, t = 0, time.perf_counter()
    while True:
        Ec = po.apply_ghat(Gh, J)[:, :, 0]
        reps += 1
        if time.perf_counter() - t > 10.0 or reps >= 5:
            break
    dt = (time.perf_counter() - t) / reps
    rel_apply = float(np.linalg.norm(Eg - Ec) / np.linalg.norm(Ec))
    return ({"value": 1.0 / dt, "unit": "operator applies/s", "cores": 1, "kind": "port",
             "sample": f"{reps} applies of the oracle's FFT route (numpy FFT) on the GPU-built Ĝ"}, rel, rel_apply)


def _group_nodes(groups, box):
    """Host kx, ky and box-weighted node weights of a list of (k_rho, phi)
    groups, in the layout CavityTransfer2D gives each group (k_rho fastest)."""
    from yellowthinfilms_amd.green_operator import box_lateral_factor
    kx, ky, w = [], [], []
    for g in groups:
        K, Ph = np.meshgrid(g["krho"], g["phi"], indexing="ij")
        kx.append((K * np.cos(Ph)).ravel(order="F"))
        ky.append((K * np.sin(Ph)).ravel(order="F"))
        w.append((np.outer(g["krho"] * g["wrho"], g["wphi"]) / (4 * np.pi ** 2)).ravel(order="F"))
    kx, ky, w = (np.concatenate(v) for v in (kx, ky, w))
    return kx, ky, w * box_lateral_factor(kx, ky, *box)


def bench_operator_volume(a):
    import torch
    from yellowthinfilms_amd import workloads as W, uniform_kxky_grid
    from yellowthinfilms_amd.green_operator import build_layer_operator_grouped, apply_operator
    from yellowthinfilms_amd.kspace import _is_azimuthally_symmetric
    from yellowthinfilms_amd._lib import Context

    wl = W.OPERATOR_VOLUME_64
    stack = W.operator_volume_stack(wl)
    grid = uniform_kxky_grid(wl["L"], wl["dx"])
    Nx, Ny = grid["Nx"], grid["Ny"]
    zs, zobs = W.operator_volume_planes(wl)
    Ns = len(zs)
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    groups, ginfo = W.operator_volume_groups(wl)
    t_grid = time.perf_counter() - t0
    kw = dict(lateral="nufft", nufft_eps=wl["eps"], skeleton=True, skeleton_tol=wl["skeleton_tol"],
              box_lateral=(wl["dx"], wl["dx"]), dz_box=wl["dx"])
    ctx = Context.get(0)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    op, t_cache, t_build = build_layer_operator_grouped(stack, W.OMEGA, wl["source_layer"], groups, grid, zs, zobs,
                                                        node_budget=wl["node_budget"], **kw)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    build_kernels = {k: ctx.kernel_time(k) for k in ("node_solve", "rot_fill", "nufft_sort", "nufft_spread",
                                                      "nufft_fft")}
    ctx.set_timing(False)
    rng = np.random.default_rng(wl["seed"])
    P = torch.as_tensor((rng.standard_normal((3, Ns, Ny, Nx)) + 1j * rng.standard_normal((3, Ns, Ny, Nx)))
                        / math.sqrt(2.0), device=dev)  # full volume, column-major (Nx, Ny, Ns, 3)

    def step():
        apply_operator(op, P)

    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "op_apply", ("op_fft_fwd", "op_mix", "op_fft_inv"))
    n_obs = sum(b["n_obs"] for b in op.blocks)
    ghat_bytes = sum(b["bytes"] for b in op.blocks)
    # read J, read every stored plane and skeleton matrix once, write E
    alg_bytes = 16 * 3 * Nx * Ny * (Ns + n_obs) + ghat_bytes
    ms = t.pop("op_apply")[0] / a.steps  # the whole apply, one event pair
    mix_ms = t["op_mix"][0] / a.steps
    cpu, rel_err, rel_apply = None, None, None
    if not a.no_cpu_baseline:
        cpu, rel_err, rel_apply = cpu_operator_volume(stack, groups, op, P, grid, wl, zs, zobs, kw)
    value = a.steps / wall
    out = _base(a, wl, value, "operator applies/s (64x64x16x3 volume onto 3 x 16 planes, padded 128^2)", wall,
                "f64")
    out["vs_baseline"] = value / wl["reference_applies_per_s"]
    out["config"].update({"grid": [Nx, Ny], "padded": [2 * Nx, 2 * Ny], "source_planes": Ns,
                          "obs_planes": {b["obs_layer"]: b["n_obs"] for b in op.blocks},
                          "krho_x_phi": [ginfo["nkrho"], wl["n_phi"]], "nodes": ginfo["n_nodes"],
                          "phi_groups": ginfo["n_groups"], "any_stalled": ginfo["any_stalled"],
                          "rotation_fill": bool(_is_azimuthally_symmetric(stack)),
                          "blocks": {b["obs_layer"]: [b["kind"], b["n_src_skel"], b["n_obs_skel"]]
                                     for b in op.blocks},
                          "lateral": "nufft", "nufft_eps": wl["eps"], "skeleton_tol": wl["skeleton_tol"]})
    out.update({
        "reference_applies_per_s": wl["reference_applies_per_s"],
        "reference_note": "apply_operator at size L, iso vac/glass/vac: 90 ms on an Apple M5 with 4 Julia threads "
                          "(STATUS_closed_form_rewrite.md:75); vs_baseline = value / 11.1",
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in t.items()},
        "grid_s": t_grid,
        "build_s": build_s,
        "build_kernel_ms": {k: v[0] for k, v in build_kernels.items()},
        "ghat_bytes": ghat_bytes,
        "rel_err_vs_fp64_oracle": rel_err,
        "rel_err_apply_vs_oracle_fft_route": rel_apply,
        "roofline": {"kernel": "operator apply (x and y FFTs of 48 source planes, z coupling of 3 blocks, inverse FFTs "
                               "of 144 observation planes)",
                     "bound": "hbm", "achieved": alg_bytes / (ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": alg_bytes / (ms * 1e-3) / PEAK_HBM, "traffic": None,
                     "algorithmic_bytes": alg_bytes, "mix_ms": mix_ms},
        "cpu_baseline": cpu,
    })
    return out


def _export_blocks(op):
    """The stored planes (each (3, 3, Px, Py), [c][q]) and skeleton matrices of every block."""
    out = []
    for bi, b in enumerate(op.blocks):
        def plane(part, slot):
            return op.ghat_plane(bi, part, slot).cpu().numpy().transpose(1, 0, 3, 2)
        e = dict(obs_layer=b["obs_layer"], No=b["n_obs"], skeleton=b["kind"].endswith("skeleton"),
                 T=[plane("toeplitz", k) for k in range(b["n_toeplitz"])],
                 H=[plane("hankel", k) for k in range(b["n_hankel"])],
                 D=[plane("dense", k) for k in range(b["n_dense"])])
        if e["skeleton"]:
            e["Ps"] = op.skeleton(bi, "source")[1]
            e["Po"] = op.skeleton(bi, "observation")[1]
        out.append(e)
    return out


def cpu_operator_volume(stack, groups, op, P, grid, wl, zs, zobs, kw):
    """(1) Two seeded voxel dipoles: the GPU field at 6 seeded (pixel, plane)
    points of every observation layer against the oracle's exact lateral sum
    of its own box spectral dyadics (mod_ClosedFormField.jl:743-753, 815-862)
    over every node of the grid (no rotation: the oracle solves each node).
    (2) The apply on all cores: the GPU-built planes read back, the lateral
    FFTs by scipy.fft on every core and the z coupling by the oracle's OpenMP
    ora_op_zmix; timed on the benchmark volume and compared with the GPU apply."""
    import torch
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd.green_operator import apply_operator
    try:
        import scipy.fft as sfft
        from oracle import pyoracle as po
        from yellowthinfilms_amd.materials import stack_arrays
    except Exception as exc:
        return ({"value": None, "unit": "operator applies/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None, None)
    cores = _cores()
    Nx, Ny = grid["Nx"], grid["Ny"]
    Ns = len(zs)
    kx, ky, Wn = _group_nodes(groups, kw["box_lateral"])
    rng = np.random.default_rng(3)
    dip = [(int(rng.integers(Nx)), int(rng.integers(Ny)), int(rng.integers(Ns)),
            rng.normal(size=3) + 1j * rng.normal(size=3)) for _ in range(2)]
    Pv = np.zeros((3, Ns, Ny, Nx), np.complex128)
    for i, j, k, p in dip:
        Pv[:, k, j, i] += p
    Ev = {o: e.cpu().numpy() for o, e in apply_operator(op, torch.as_tensor(Pv, device=P.device)).items()}
    st = po.Stack(*stack_arrays(stack))
    got, ref = [], []
    for o, zo in zobs.items():
        pts = [(int(rng.integers(Nx)), int(rng.integers(Ny)), int(rng.integers(len(zo)))) for _ in range(6)]
        pairs = sorted({(k, m) for (_, _, k, _) in dip for (_, _, m) in pts})
        g = po.ghat_batch(st, W.OMEGA, kx, ky, wl["source_layer"], o, [zs[k] for k, _ in pairs],
                          [zo[m] for _, m in pairs], nthreads=cores, dz_box=kw["dz_box"])
        for (x, y, m) in pts:
            e = np.zeros(3, np.complex128)
            for (i, j, k, p) in dip:
                G = po.lateral_sum_points(g[:, pairs.index((k, m))], Wn, kx, ky,
                                          [((x - i) * grid["dx"], (y - j) * grid["dy"])])[0]
                e += G @ p
            ref.append(e)
            got.append(Ev[o][:, m, y, x])
    rel = float(np.linalg.norm(np.array(got) - np.array(ref)) / np.linalg.norm(np.array(ref)))
    # the apply on every core, on the benchmark volume
    Px, Py = 2 * Nx, 2 * Ny
    nxy = Px * Py
    Pn = P.cpu().numpy()  # (3, Ns, Ny, Nx)
    Eg = {o: e.cpu().numpy() for o, e in apply_operator(op, P).items()}
    blocks = [{k: (np.ascontiguousarray(np.array(v).reshape(len(v), 3, 3, nxy)) if k in ("T", "H", "D") and v else v)
               for k, v in b.items()} for b in _export_blocks(op)]
    reps, t = 0, time.perf_counter()
    while True:
        Jp = np.zeros((Ns, 3, Px, Py), np.complex128)
        Jp[:, :, :Nx, :Ny] = Pn.transpose(1, 0, 3, 2)
        Jh = sfft.fft2(Jp, axes=(2, 3), workers=cores).reshape(Ns, 3, nxy)
        Ec = {}
        for b in blocks:
            E = po.op_zmix(nxy, Jh, b["No"], T=b["T"] if len(b["T"]) else None, H=b["H"] if len(b["H"]) else None,
                           D=b["D"] if len(b["D"]) else None, Ps=b.get("Ps"), Po=b.get("Po"), nthreads=cores)
            Ec[b["obs_layer"]] = (sfft.ifft2(E.reshape(b["No"], 3, Px, Py), axes=(2, 3), workers=cores)
                                  * nxy)[:, :, :Nx, :Ny]
        reps += 1
        if time.perf_counter() - t > 20.0 or reps >= 5:
            break
    dt = (time.perf_counter() - t) / reps
    num = sum(np.linalg.norm(Eg[o].transpose(1, 0, 3, 2) - Ec[o]) ** 2 for o in Ec)
    den = sum(np.linalg.norm(Ec[o]) ** 2 for o in Ec)
    rel_apply = float(math.sqrt(num / den))
    return ({"value": 1.0 / dt, "unit": "operator applies/s", "cores": cores, "kind": "port",
             "sample": f"{reps} applies on the benchmark volume: the GPU-built planes read back, lateral FFTs by "
                       f"scipy.fft on {cores} workers, z coupling by the oracle's OpenMP ora_op_zmix"},
            rel, rel_apply)


def bench_pole_scan(a):
    import torch
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd import kspace
    from yellowthinfilms_amd.kspace import real_axis_poles
    from yellowthinfilms_amd._lib import Context

    wl = W.POLE_SCAN_HYPERBOLIC
    stack = W.pole_scan_stack(wl)
    phis, k_lo, k_hi, step, exclude = W.pole_scan_args(wl)
    res = {}
    calls = {"n": 0}
    orig = kspace.cavity_response

    def counting(stack_, omega_, s_, kx, ky, device=0):  # nodes evaluated per scan (scan + refinement)
        calls["n"] += int(np.asarray(kx).size)
        return orig(stack_, omega_, s_, kx, ky, device=device)

    def step_fn():
        res["poles"] = real_axis_poles(stack, W.OMEGA, wl["source_layer"], phis=phis, k_lo=k_lo, k_hi=k_hi,
                                       scan_step=step, exclude=exclude)

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step_fn()
    kspace.cavity_response = counting
    try:
        wall, kms = _timed_kernels(a, step_fn, torch, ctx, ("cavity_response", "pole_refine"))
    finally:
        kspace.cavity_response = orig
    ctx.set_timing_only("pole_candidates")  # the candidate kernel, in a pass outside the wall clock
    for _ in range(a.steps):
        step_fn()
    kms["pole_candidates"] = ctx.kernel_time("pole_candidates")
    ctx.set_timing(False)
    kt = kms["cavity_response"][0] + kms["pole_refine"][0]  # every response evaluation
    kn = kms["cavity_response"][1] + kms["pole_refine"][1]
    # scan nodes counted at the launch, refinement evaluations counted in the kernel
    nodes_per_scan = calls["n"] / a.steps + kspace.search_stats["refine_evaluations"]
    n_scan = len(phis) * (int(math.floor(math.log(k_hi / k_lo) / step + 1e-10)) + 1)
    # the scan launch alone (every direction's log grid in one launch), timed on its own
    from yellowthinfilms_amd.kspace import cavity_response
    L = math.log(k_hi / k_lo)
    ks = k_lo * np.exp(np.arange(int(math.floor(L / step + 1e-10)) + 1) * step)
    KK = np.repeat(ks[None, :], len(phis), axis=0).ravel()
    PP = np.repeat(np.asarray(phis)[:, None], ks.size, axis=1).ravel()
    kxd = torch.as_tensor(KK * np.cos(PP), device="cuda")
    kyd = torch.as_tensor(KK * np.sin(PP), device="cuda")
    cavity_response(stack, W.OMEGA, wl["source_layer"], kxd, kyd)
    torch.cuda.synchronize()
    ctx.set_timing(True)
    for _ in range(5):
        cavity_response(stack, W.OMEGA, wl["source_layer"], kxd, kyd)
    st_ms, st_n = ctx.kernel_time("cavity_response")
    ctx.set_timing(False)
    scan_ms = st_ms / st_n
    # the scan launch is the largest cavity_response launch of a step
    flops = _pole_scan_flops(stack, wl)
    cpu, cmp = None, None
    if not a.no_cpu_baseline:
        cpu, cmp = cpu_pole_scan(stack, wl, res["poles"])
    out = _base(a, wl, nodes_per_scan * a.steps / wall, "cavity-response k-nodes/s (whole pole search)", wall, "f64")
    out["config"].update({"directions": len(phis), "scan_nodes": n_scan, "k_lo": k_lo, "k_hi": k_hi,
                          "scan_step": step})
    out.update({
        "scans_per_s": a.steps / wall, "nodes_per_scan": nodes_per_scan, "poles_found": len(res["poles"]),
        "response_launches_per_scan": kn / a.steps, "response_kernel_ms_per_scan": kt / a.steps,
        "kernel_ms_per_scan": {k: v[0] / a.steps for k, v in kms.items()},
        "poles_vs_oracle": cmp,
        "all_launches_tflops": flops * nodes_per_scan / (kt / a.steps * 1e-3) / 1e12,
        "roofline": {"kernel": "cavity_response scan launch (every direction's log grid: Berreman modes + S cascade "
                               "+ cavity per node, FP64)", "bound": "valu-fp64",
                     "achieved": flops * n_scan / (scan_ms * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": flops * n_scan / (scan_ms * 1e-3) / PEAK_FP64,
                     "traffic": None, "flops_per_node": flops, "nodes_per_launch": n_scan, "launch_ms": scan_ms},
        "cpu_baseline": cpu,
    })
    return out


def _pole_scan_flops(stack, wl):
    """Mean op count of one cavity response (the oracle's reference-order
    arithmetic with an op-counting real type) over 16 seeded nodes of the scan range."""
    try:
        from oracle import pyoracle as po
        from yellowthinfilms_amd.materials import stack_arrays
        from yellowthinfilms_amd import workloads as W
    except Exception:
        return float("nan")
    st = po.Stack(*stack_arrays(stack))
    rng = np.random.default_rng(0)
    tot = []
    for k in np.exp(rng.uniform(math.log(wl["k_lo_k0"] * W.OMEGA), math.log(wl["k_hi"]), 16)):
        ph = rng.uniform(0, math.pi / 2)
        c = po.count_response_ops(st, W.OMEGA, wl["source_layer"], k * math.cos(ph), k * math.sin(ph))
        tot.append(c["add"] + c["mul"] + c["div"])
    return float(np.mean(tot))


def cpu_pole_scan(stack, wl, poles):
    """The oracle's step-by-step real_axis_poles (its cavity response on the
    host cores, the reference's sequential refinement): timed once, and its
    poles compared with the GPU's."""
    try:
        from oracle import pyoracle as po
        from yellowthinfilms_amd.materials import stack_arrays
        from yellowthinfilms_amd import workloads as W
    except Exception as exc:
        return ({"value": None, "unit": "cavity-response k-nodes/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    st = po.Stack(*stack_arrays(stack))
    phis, k_lo, k_hi, step, exclude = W.pole_scan_args(wl)
    n_calls = {"n": 0}
    orig = po.cavity_response

    def counting(*args, **kw):
        n_calls["n"] += len(np.atleast_1d(args[3]))
        return orig(*args, **kw)

    po.cavity_response = counting
    try:
        t = time.perf_counter()
        o = po.real_axis_poles(st, W.OMEGA, wl["source_layer"], phis=phis, k_lo=k_lo, k_hi=k_hi, scan_step=step,
                               exclude=exclude)
        dt = time.perf_counter() - t
    finally:
        po.cavity_response = orig
    cmp = None
    if len(o) == len(poles):
        dq = np.array([abs(p["q"] - q) / G for p, (q, G, _, _) in zip(poles, o)])
        dG = np.array([abs(p["Gamma"] - G) / G for p, (q, G, _, _) in zip(poles, o)])
        cmp = {"poles": len(o), "max_dq_over_gamma": float(dq.max()), "median_dgamma_rel": float(np.median(dG)),
               "n_gamma_off_1e-3": int(np.sum(dG > 1e-3)), "n_q_off_1e-4_gamma": int(np.sum(dq > 1e-4))}
    else:
        cmp = {"poles": len(o), "gpu_poles": len(poles)}
    return ({"value": n_calls["n"] / dt, "unit": "cavity-response k-nodes/s", "cores": _cores(), "kind": "port",
             "sample": f"one whole pole search ({n_calls['n']} evaluations, {len(o)} poles; scan on all cores, "
                       "refinement one evaluation at a time as the reference takes it)"}, cmp)


# ---------------------------------------------------------------------------
def dict_ns(a, **kw):
    b = argparse.Namespace(**vars(a))
    for k, v in kw.items():
        setattr(b, k, v)
    return b


def bench_probes(a):
    import torch
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd.probes import probe_values
    from yellowthinfilms_amd._lib import Context

    wl = W.PROBE_BATCH_BOX
    st = W.probe_batch_stack(wl)
    kr, phis, wph, probes = W.probe_batch_args(wl)
    res = {}

    def step():
        res["v"] = probe_values(st, W.OMEGA, wl["source_layer"], kr, phis, wph, probes, wl["box"])

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "node_solve", ("lateral_segments",))
    node_ms, node_n = t["node_solve"]
    lat_ms, _ = t["lateral_segments"]
    nodes = len(kr) * len(phis)
    flops = W.PROBE_BATCH_BOX_FLOPS_PER_NODE * nodes  # the 3 node-solve launches of one step
    node_step = node_ms / a.steps
    cpu = cmp = None
    if not a.no_cpu_baseline:
        cpu, cmp = cpu_probes(st, wl, kr, phis, wph, probes, res["v"])
    out = _base(a, wl, len(kr) * a.steps / wall, "k_rho probe sets/s (8 box probes, 96 phi nodes each)", wall, "f64")
    out["config"].update({"krho": len(kr), "phi": len(phis), "probes": len(probes), "nodes": nodes})
    out.update({
        "kernel_ms": {"node_solve": node_step, "lateral_segments": lat_ms / a.steps},
        "node_launches_per_step": node_n / a.steps,
        "ms_per_krho": wall * 1e3 / (a.steps * len(kr)),
        "rel_err_vs_fp64_oracle": cmp,
        "roofline": {"kernel": "node_solve (one launch per observation layer, box source, FP64)", "bound": "valu-fp64",
                     "achieved": flops / (node_step * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": flops / (node_step * 1e-3) / PEAK_FP64, "traffic": None,
                     "flops_per_node": W.PROBE_BATCH_BOX_FLOPS_PER_NODE, "nodes_per_step": nodes},
        "cpu_baseline": cpu,
    })
    return out


def cpu_probes(st, wl, kr, phis, wph, probes, v_gpu):
    """The oracle's probe values (per-node spectral dyadic with the box
    factors on all host cores, then the exact lateral sum per k_rho) over the
    whole batch, timed; and the GPU's largest error relative to each probe's
    largest entry."""
    try:
        from oracle import pyoracle as po
    except Exception as exc:
        return ({"value": None, "unit": "k_rho probe sets/s", "cores": 0, "kind": "port",
                 "sample": f"oracle unavailable: {exc}"}, None)
    from yellowthinfilms_amd.green_operator import box_lateral_factor
    from yellowthinfilms_amd.materials import stack_arrays
    ost = po.Stack(*stack_arrays(st))
    cores = _cores()
    box = wl["box"]
    kx = (kr[:, None] * np.cos(phis)[None, :]).ravel()
    ky = (kr[:, None] * np.sin(phis)[None, :]).ravel()
    W = (kr[:, None] * wph[None, :]).ravel() / (4 * math.pi ** 2) * box_lateral_factor(kx, ky, box[0], box[1])
    L = len(phis)
    t0 = time.perf_counter()
    ref = np.zeros((len(kr), len(probes), 3, 3), np.complex128)
    for obs in sorted({p[0] for p in probes}):
        idx = [m for m, p in enumerate(probes) if p[0] == obs]
        g = po.ghat_batch(ost, 2 * math.pi, kx, ky, wl["source_layer"], obs, [probes[m][5][0] for m in idx],
                          [probes[m][1] for m in idx], nthreads=cores, dz_box=box[2], ez_local=True)
        S = g * W[:, None, None, None]
        S = np.where(np.isfinite(S), S, 0.0)
        for j, m in enumerate(idx):
            ph = np.exp(1j * (kx * probes[m][2] + ky * probes[m][3]))
            ref[:, m] = (S[:, j] * ph[:, None, None]).reshape(len(kr), L, 3, 3).sum(axis=1)
    dt = time.perf_counter() - t0
    scale = np.maximum(np.abs(ref).reshape(len(kr), len(probes), 9).max(axis=2), 1e-300)
    err = float((np.abs(v_gpu - ref).reshape(len(kr), len(probes), 9).max(axis=2) / scale).max())
    return ({"value": len(kr) / dt, "unit": "k_rho probe sets/s", "cores": cores, "kind": "port",
             "sample": f"the whole batch ({len(kr)} k_rho x {L} phi, {len(probes)} probes)"}, err)


def bench_box_grid(a):
    import torch
    from yellowthinfilms_amd import workloads as W
    from yellowthinfilms_amd.adaptive import box_krho_grid
    from yellowthinfilms_amd.probes import probe_values
    from yellowthinfilms_amd._lib import Context

    wl = W.BOX_KRHO_GRID
    st = W.probe_batch_stack()
    res = {}

    def step():
        res["g"] = box_krho_grid(st, W.OMEGA, wl["source_layer"], wl["box"], n_phi=wl["n_phi"], rtol=wl["rtol"],
                                 krho_cap=wl["krho_cap"], rho_far=wl["rho_far"])

    ctx = Context.get(0)
    for _ in range(a.warmup):
        step()
    wall, t = _timed_kernels(a, step, torch, ctx, "node_solve", ("lateral_segments", "cavity_response"))
    node_ms, node_n = t["node_solve"]
    lat_ms, _ = t["lateral_segments"]
    scan_ms, _ = t["cavity_response"]
    kr, wr, phis, wph, info = res["g"]
    nk = info["n_krho_solves"]
    nodes = info["n_krho_evaluated"] * len(phis)          # every solved node, the speculative ones included
    flops = W.BOX_KRHO_GRID_FLOPS_PER_NODE * nodes
    node_step = node_ms / a.steps
    cpu = cmp = None
    if not a.no_cpu_baseline:
        probes = W.box_grid_probes(wl)
        ks = info["krho_visited"][:: max(1, nk // 192)]
        v = probe_values(st, W.OMEGA, wl["source_layer"], ks, phis, wph, probes, wl["box"])
        cpu, cmp = cpu_probes(st, dict(wl, box=wl["box"]), ks, phis, wph, probes, v)
        cpu["sample"] = f"{len(ks)} of the {nk} k_rho the grid visits (every {max(1, nk // 192)}th), " \
                        f"{len(phis)} phi x {len(probes)} probes each; the reference builds one one-node cache per k_rho"
    out = _base(a, wl, nk * a.steps / wall, "k_rho probe sets/s (whole adaptive grid build, 22 box probes, 64 phi)",
                wall, "f64")
    out["config"].update({"phi": len(phis), "probes": len(info["probes"]), "rtol": wl["rtol"],
                          "krho_cap": wl["krho_cap"]})
    out.update({
        "grid": {"n_panels": info["n_panels"], "n_nodes": info["n_nodes"], "n_krho_solves": nk,
                 "n_krho_evaluated": info["n_krho_evaluated"], "n_batches": info["n_batches"],
                 "n_probe_calls": info["n_calls"], "stalled": info["any_stalled"], "tail_bound": info["tail_bound"]},
        "ms_per_grid": wall * 1e3 / a.steps,
        "kernel_ms": {"node_solve": node_step, "lateral_segments": lat_ms / a.steps,
                      "cavity_response (pole scan)": scan_ms / a.steps},
        "node_launches_per_step": node_n / a.steps,
        "rel_err_vs_fp64_oracle": cmp,
        "roofline": {"kernel": "node_solve (one launch per observation layer and refinement batch, box source, FP64)",
                     "bound": "valu-fp64", "achieved": flops / (node_step * 1e-3) / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": flops / (node_step * 1e-3) / PEAK_FP64, "traffic": None,
                     "flops_per_node": W.BOX_KRHO_GRID_FLOPS_PER_NODE, "nodes_per_step": nodes},
        "cpu_baseline": cpu,
    })
    return out


WORKLOADS = {"batched_gyro_512": bench_gyro, "point_uniaxial_1d": bench_point, "bessel_gemm": bench_gemm,
             "uneven_256": bench_uneven, "surrogate_1024": bench_surrogate, "nufft_1e6": bench_nufft,
             "operator_512": bench_operator, "operator_volume_64": bench_operator_volume,
             "pole_scan_hyperbolic": bench_pole_scan, "probe_batch_box": bench_probes,
             "box_krho_grid": bench_box_grid}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="batched_gyro_512", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--tune-db", default=None, help="autotuner database file (uneven_256)")
    ap.add_argument("--sur-sweep", default="",
                    help="surrogate_1024: further network sizes to train and time on the same data, e.g. 128x2,256x2")
    ap.add_argument("--sur-windows", default="",
                    help="surrogate_1024: further pole-window widths (half-widths) to evaluate, e.g. 0,2,4")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="batched_gyro_512: leave out the surrogate-vs-plain comparison of the metric")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed region, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, to scale the FETCH_SIZE/WRITE_SIZE counters of a profiling run)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding of the node solve "
              "is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    out = WORKLOADS[a.workload](a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
