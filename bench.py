"""Benchmark of the batched plain GPU Green operator on one MI355X.

Workload (yellowthinfilms_amd.workloads.BATCHED_GYRO_512): a 5-layer stack
with a 40 nm gyrotropic film on the uniform 512 x 512 (kx, ky) grid.  One
step = one Green k-node per grid node (262,144 node solves -> spectral
dyadic g~, the ytf_green_nodes kernel) + one FFT convolution of a
512 x 512 x 3 complex N(0,1) source (ytf_spectral_apply).  Inputs are
resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]

Prints one JSON line: value = Green k-nodes/s of the whole step; the apply
rate, a roofline record of the dominant kernel (the node solve, FP64) measured
with HIP events on the stream the kernels run on, and the CPU baseline (the
oracle timed on this host's cores over the same grid, which also gives the
relative error of the GPU's g~ against it on every node).
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

PEAK_FP64 = 78.6e12   # MI355X FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12     # bytes/s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed region, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, to scale the FETCH_SIZE/WRITE_SIZE counters of a profiling run)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU sharding of the node solve "
              "is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from yellowthinfilms_amd import uniform_kxky_grid, SpectralGreenOperator, workloads as W

    wl = W.BATCHED_GYRO_512
    grid = uniform_kxky_grid(wl["L"], wl["dx"])
    nx, ny = grid["Nx"], grid["Ny"]
    nn = nx * ny
    stack = W.gyro_512_stack()
    dev = torch.device("cuda", 0)
    op = SpectralGreenOperator(stack, W.OMEGA, grid, wl["source_layer"], wl["zsrc"], wl["obs_layer"], wl["zobs"])
    J = torch.as_tensor(W.uniform_c128_source(nx, ny, wl["seed"]), device=dev)
    E = torch.empty_like(J)
    stream = torch.cuda.current_stream(dev)  # the ytf context is bound to this stream

    for _ in range(a.warmup):
        op.build()
        op.apply(J, out=E)
    torch.cuda.synchronize()

    ev = lambda: torch.cuda.Event(enable_timing=True)
    t_node, t_apply = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        e0, e1, e2 = ev(), ev(), ev()
        e0.record(stream)
        op.build()
        e1.record(stream)
        op.apply(J, out=E)
        e2.record(stream)
        t_node.append((e0, e1))
        t_apply.append((e1, e2))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    node_ms = float(np.mean([s.elapsed_time(e) for s, e in t_node]))
    apply_ms = float(np.mean([s.elapsed_time(e) for s, e in t_apply]))
    ms_per_step = wall * 1e3 / a.steps

    if a.calib_copy_mb > 0:
        src = torch.empty(a.calib_copy_mb * 2 ** 20 // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    flops = W.BATCHED_GYRO_512_FLOPS_PER_NODE * nn
    achieved = flops / (node_ms * 1e-3)
    cpu, rel_err = None, None
    if not a.no_cpu_baseline:
        g = op.ghat[0].cpu().numpy()  # [9][node], node = y + ny*x
        cpu, rel_err = cpu_baseline(stack, grid, wl, a.cpu_repeats, g)

    apply_bytes = 16.0 * (3 * nn + 9 * nn + 3 * nn)
    out = {
        "metric": "Green k-nodes/s + operator applies/s, surrogate vs plain GPU, <=1e-6 rel err",
        "value": nn * a.steps / wall,
        "unit": "Green k-nodes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) complex source; stack and grid as configured)",
        "config": {"workload": wl["description"], "nodes_per_step": nn, "grid": [nx, ny],
                   "path": "plain GPU (ytf_green_nodes + ytf_spectral_apply)"},
        "node_solve_ms": node_ms,
        "apply_ms": apply_ms,
        "applies_per_s": 1e3 / apply_ms,
        "node_solve_knodes_per_s": nn / (node_ms * 1e-3),
        "apply_hbm_gbps_algorithmic": apply_bytes / (apply_ms * 1e-3) / 1e9,
        "rel_err_vs_fp64_oracle": rel_err,
        "roofline": {"kernel": "ytf node_kernel (Green k-node solve, FP64 vector ALU)", "bound": "mfma",
                     "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": W.BATCHED_GYRO_512_NODE_HBM_BYTES,
                     "flops_per_node": W.BATCHED_GYRO_512_FLOPS_PER_NODE, "nodes_per_launch": nn},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(stack, grid, wl, repeats, gpu_ghat):
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
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
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


if __name__ == "__main__":
    main()
