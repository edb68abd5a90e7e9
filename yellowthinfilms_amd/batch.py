"""Uneven batches of independent stacks on the GPU (host side).

StackBatch mirrors the per-stack, per-node loop of Modes_Stack +
S_all_Matrix (mod_ModeSolver.jl:358-413, mod_ScatteringMatrix.jl:112-154) over
many stacks: every stack has its own layers and its own node list, all solved
in one persistent launch (ytf_batch_solve, replayed as a HIP graph) whose
launch shape an autotuner picks (ytf_batch_tune).
"""
import ctypes as ct

import numpy as np

from . import _lib
from .green import layer_structs


def _torch():
    import torch
    return torch


def _ptr(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else None


class StackBatch:
    def __init__(self, stacks, node_counts, omega, device=0):
        if len(stacks) != len(node_counts) or not stacks:
            raise ValueError("StackBatch: one node count per stack")
        self.device = device
        self.n_stacks = len(stacks)
        self.layer_counts = np.array([len(s) for s in stacks], np.int32)
        self.node_counts = np.array(node_counts, np.int64)
        self.n_nodes = int(self.node_counts.sum())
        self.node_offsets = np.concatenate([[0], np.cumsum(self.node_counts)])
        flat = [l for s in stacks for l in s]
        self._layers = layer_structs(flat)
        self.ctx = _lib.Context.get(device)
        h = ct.c_void_p()
        rc = _lib.lib().ytf_batch_create(self.ctx.handle, self._layers,
                                         self.layer_counts.ctypes.data_as(ct.POINTER(ct.c_int32)),
                                         self.node_counts.ctypes.data_as(ct.POINTER(ct.c_int64)), self.n_stacks,
                                         float(omega), ct.byref(h))
        _lib.check(rc, self.ctx.handle)
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().ytf_batch_destroy(self.ctx.handle, h)
            except Exception:
                pass
            self.handle = None

    @property
    def config(self):
        a, b, c = ct.c_int(), ct.c_int(), ct.c_int()
        _lib.check(_lib.lib().ytf_batch_get_config(self.ctx.handle, self.handle, ct.byref(a), ct.byref(b),
                                                   ct.byref(c)), self.ctx.handle)
        return dict(nodes_per_item=a.value, waves_per_block=b.value, blocks=c.value)

    def set_config(self, nodes_per_item, waves_per_block, blocks):
        _lib.check(_lib.lib().ytf_batch_set_config(self.ctx.handle, self.handle, int(nodes_per_item),
                                                   int(waves_per_block), int(blocks)), self.ctx.handle)

    def _nodes(self, kx, ky):
        torch = _torch()
        dev = torch.device("cuda", self.device)
        kx = torch.as_tensor(kx, dtype=torch.float64, device=dev).contiguous()
        ky = torch.as_tensor(ky, dtype=torch.float64, device=dev).contiguous()
        if kx.shape != (self.n_nodes,) or ky.shape != (self.n_nodes,):
            raise ValueError(f"StackBatch: kx, ky must hold {self.n_nodes} nodes (stack after stack)")
        return kx, ky

    def outputs(self, modes=False):
        torch = _torch()
        dev = torch.device("cuda", self.device)
        n = self.n_nodes
        z = lambda e: torch.empty((e, n), dtype=torch.complex128, device=dev)
        out = dict(S=z(16), flags=torch.empty(n, dtype=torch.uint8, device=dev))
        if modes:
            out.update(V1=z(16), Q1=z(4), VN=z(16), QN=z(4))
        return out

    def solve(self, kx, ky, modes=False, out=None):
        """Total S (16, n) and flags (n,) [and V1, Q1, VN, QN] per node."""
        kx, ky = self._nodes(kx, ky)
        out = out if out is not None else self.outputs(modes)
        self.ctx.bind_stream()
        rc = _lib.lib().ytf_batch_solve(self.ctx.handle, self.handle, _ptr(kx), _ptr(ky), _ptr(out["S"]),
                                        _ptr(out.get("V1")), _ptr(out.get("Q1")), _ptr(out.get("VN")),
                                        _ptr(out.get("QN")), _ptr(out["flags"]))
        _lib.check(rc, self.ctx.handle)
        return out

    def tune(self, kx, ky, tune_db=None):
        """Autotune the launch shape on these inputs; returns the tuner's report."""
        kx, ky = self._nodes(kx, ky)
        out = self.outputs()
        self.ctx.bind_stream()
        buf = ct.create_string_buffer(4096)
        rc = _lib.lib().ytf_batch_tune(self.ctx.handle, self.handle, _ptr(kx), _ptr(ky), _ptr(out["S"]),
                                       _ptr(out["flags"]), tune_db.encode() if tune_db else None, buf, len(buf))
        _lib.check(rc, self.ctx.handle)
        return buf.value.decode()
