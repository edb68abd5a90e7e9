"""Learned surrogate of the spectral dyadic, with the exact solve as fallback
(host side).  No reference counterpart: the reference evaluates every node
with the exact Berreman + S-matrix solve (mod_CavityTransfer.jl,
mod_ClosedFormField.jl), which here generates the training data, checks the
surrogate, and replaces it on flagged nodes.

  MLP             -- the network (ytf_mlp_*): BF16 matrix-core training and
                     inference, FP32 accumulation and master weights.
  GyroFamily      -- a seeded family of 4-layer gyrotropic stacks and its
                     6 inputs (kx, ky, z, z', omega, loss).
  SurrogateGreen  -- training data from the node solver, training, and the
                     scattered spectral dyadic on a node set: network
                     prediction, flags, exact FP64 solve on the flagged nodes.

What the network learns.  With source and observation in the air above the
stack, the scattered dyadic is g~_s = P(k) t(k) omega (1 + k^2/omega^2) / kz0
with the air propagator P = exp(i kz0 (h_s + h_o)), kz0 = sqrt(omega^2 - k^2)
(Im >= 0): the network predicts the reflection-like factor t (18 reals), of
order one over the whole band; P, 1/kz0 and the quasi-static growth
(1 + k^2/omega^2) are applied exactly.  Nodes are flagged for the exact
solve near the air light line k = omega (the branch point, where 1/kz0
diverges), near k = 0 (where the dyadic's angular factors are
discontinuous), within a few half-widths of a pole of the cavity response
(_cavity_response's sigma_max indicator, mod_KSpaceFeatures.jl:109-127: the
poles come from real_axis_poles on the stack, once per (omega, loss)), and
where the predicted |t| is large; training samples exclude the first two
neighbourhoods.
"""
import ctypes as ct
import math

import numpy as np

from . import _lib
from .green import green_nodes
from .materials import IsotropicMaterial, GyrotropicMaterial, FlatOpticalLayer

NM = 1.0 / 600.0


def _torch():
    import torch
    return torch


def _p(t):
    return ct.c_void_p(t.data_ptr())


def _fp(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(ct.POINTER(ct.c_float))


class MLP:
    """n_in -> width x depth (SiLU) -> n_out, BF16 on the matrix cores.
    kz_features: the first layer also sees Re, Im sqrt(1 - x0^2 - x1^2)."""

    def __init__(self, n_in=6, width=256, depth=4, n_out=18, seed=1, device=0, kz_features=False):
        self.device = device
        self.n_in, self.width, self.depth, self.n_out = n_in, width, depth, n_out
        self.kz_features = bool(kz_features)
        self.ctx = _lib.Context.get(device)
        h = ct.c_void_p()
        _lib.check(_lib.lib().ytf_mlp_create(self.ctx.handle, n_in, width, depth, n_out, 1 if kz_features else 0,
                                             seed, ct.byref(h)), self.ctx.handle)
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().ytf_mlp_destroy(self.ctx.handle, h)
            except Exception:
                pass
            self.handle = None

    @property
    def num_params(self):
        return int(_lib.lib().ytf_mlp_num_params(self.handle))

    @property
    def flops_per_sample(self):
        """Forward multiply-adds x 2 of one sample."""
        w, d = self.width, self.depth
        return 2 * ((self.n_in + 2 * self.kz_features) * w + (d - 1) * w * w + w * self.n_out)

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        _lib.check(_lib.lib().ytf_mlp_get_params(self.ctx.handle, self.handle,
                                                 out.ctypes.data_as(ct.POINTER(ct.c_float))), self.ctx.handle)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (self.num_params,):
            raise ValueError(f"MLP.set_params: need {self.num_params} values")
        _lib.check(_lib.lib().ytf_mlp_set_params(self.ctx.handle, self.handle, _fp(p)), self.ctx.handle)

    def set_normalization(self, in_shift, in_scale, out_shift, out_scale):
        _lib.check(_lib.lib().ytf_mlp_set_normalization(self.ctx.handle, self.handle, _fp(in_shift), _fp(in_scale),
                                                        _fp(out_shift), _fp(out_scale)), self.ctx.handle)

    def _rows(self, x, n_cols):
        torch = _torch()
        x = torch.as_tensor(x, dtype=torch.float32, device=torch.device("cuda", self.device)).contiguous()
        if x.dim() != 2 or x.shape[1] != n_cols:
            raise ValueError(f"MLP: expected (n, {n_cols}) rows, got {tuple(x.shape)}")
        return x

    def forward(self, x, out=None):
        torch = _torch()
        x = self._rows(x, self.n_in)
        y = out if out is not None else torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        self.ctx.bind_stream()
        _lib.check(_lib.lib().ytf_mlp_forward(self.ctx.handle, self.handle, x.shape[0], _p(x), _p(y)),
                   self.ctx.handle)
        return y

    def loss_grad(self, x, y):
        x = self._rows(x, self.n_in)
        y = self._rows(y, self.n_out)
        loss = ct.c_float()
        g = np.zeros(self.num_params, np.float32)
        self.ctx.bind_stream()
        _lib.check(_lib.lib().ytf_mlp_loss_grad(self.ctx.handle, self.handle, x.shape[0], _p(x), _p(y),
                                                ct.byref(loss), g.ctypes.data_as(ct.POINTER(ct.c_float))),
                   self.ctx.handle)
        return loss.value, g

    def train_step(self, x, y, lr, loss_out=None):
        self.ctx.bind_stream()
        _lib.check(_lib.lib().ytf_mlp_train_step(self.ctx.handle, self.handle, x.shape[0], _p(x), _p(y), float(lr),
                                                 _p(loss_out) if loss_out is not None else None), self.ctx.handle)


# ---------------------------------------------------------------------------
class GyroFamily:
    """Air / gyrotropic film / spacer / substrate; the gyration g and the
    film, spacer and substrate permittivities are drawn once from the seed
    (g ~ U[0, 0.5]); every sample varies (kx, ky, h_s, h_o, omega, loss), with
    loss the imaginary part added to every layer below the air."""

    K_RANGE = 3.0                 # |kx|, |ky| <= 3 omega
    H_RANGE = (20.0, 80.0)        # source / observation heights above the film, nm
    F_RANGE = (0.9, 1.1)          # omega / 2 pi
    LOSS_RANGE = (0.01, 0.5)

    def __init__(self, seed=1):
        rng = np.random.default_rng(seed)
        self.seed = seed
        self.g = float(rng.uniform(0.0, 0.5))
        self.eps_film = (float(rng.uniform(2.5, 3.5)), float(rng.uniform(2.5, 3.5)))
        self.eps_spacer = float(rng.uniform(1.8, 2.6))
        self.eps_sub = float(rng.uniform(3.0, 5.0))
        self.d_film, self.d_spacer = 40 * NM, 60 * NM

    def stack(self, loss):
        et, ez = self.eps_film
        gy = np.array([[et + 1j * loss, 1j * self.g, 0], [-1j * self.g, et + 1j * loss, 0], [0, 0, ez + 1j * loss]])
        return [FlatOpticalLayer(IsotropicMaterial(1.0), 1.0),
                FlatOpticalLayer(GyrotropicMaterial(gy), self.d_film),
                FlatOpticalLayer(IsotropicMaterial(self.eps_spacer + 1j * loss), self.d_spacer),
                FlatOpticalLayer(IsotropicMaterial(self.eps_sub + 1j * loss), 1.0)]

    def input_normalization(self):
        lo = np.array([-self.K_RANGE, -self.K_RANGE, self.H_RANGE[0], self.H_RANGE[0], self.F_RANGE[0],
                       self.LOSS_RANGE[0]])
        hi = np.array([self.K_RANGE, self.K_RANGE, self.H_RANGE[1], self.H_RANGE[1], self.F_RANGE[1],
                       self.LOSS_RANGE[1]])
        return (lo + hi) / 2, 2.0 / (hi - lo)

    @staticmethod
    def features(kx, ky, hs_nm, ho_nm, omega, loss):
        """Raw network inputs (n, 6): kx/omega, ky/omega, h_s [nm], h_o [nm], omega/2pi, loss."""
        n = np.broadcast(kx, ky, hs_nm, ho_nm, omega, loss).shape
        w = np.broadcast_to(omega, n)
        return np.stack([np.broadcast_to(kx, n) / w, np.broadcast_to(ky, n) / w, np.broadcast_to(hs_nm, n),
                         np.broadcast_to(ho_nm, n), w / (2 * math.pi), np.broadcast_to(loss, n)],
                        axis=-1).reshape(-1, 6).astype(np.float32)


def _kz0(k2, omega):
    kz = np.sqrt((omega ** 2 - k2).astype(np.complex128))
    return np.where(kz.imag < 0, -kz, kz)


def _growth(k2, omega):
    return 1.0 + k2 / (omega * omega)


def target_from_ghat(g, kx, ky, hs, ho, omega):
    """g~_s (n, 9) complex -> t (n, 18) reals: t = g~_s kz0 / (omega P (1 + k^2/omega^2))."""
    k2 = kx ** 2 + ky ** 2
    kz = _kz0(k2, omega)
    P = np.exp(1j * kz * (hs + ho))
    t = g * (kz / (omega * P * _growth(k2, omega)))[:, None]
    return np.concatenate([t.real, t.imag], axis=1)


def ghat_from_target(t, kx, ky, hs, ho, omega, torch_mod=None):
    """Inverse of target_from_ghat (torch tensors on the device)."""
    torch = torch_mod or _torch()
    k2 = kx * kx + ky * ky
    kz = torch.sqrt((omega * omega - k2).to(torch.complex128))
    kz = torch.where(kz.imag < 0, -kz, kz)
    P = torch.exp(1j * kz * (hs + ho))
    tc = torch.complex(t[:, :9].double(), t[:, 9:].double())
    return tc * (omega * P * _growth(k2, omega) / kz)[:, None]


class SurrogateGreen:
    """The surrogate for one GyroFamily: data, training, prediction with fallback."""

    BRANCH_TOL = 0.03   # |k - omega| < BRANCH_TOL omega: the air branch point
    ORIGIN_TOL = 0.02   # k < ORIGIN_TOL omega: discontinuous angular factors

    def __init__(self, family, width=256, depth=4, seed=1, device=0):
        self.family = family
        self.device = device
        self.mlp = MLP(6, width, depth, 18, seed=seed, device=device, kz_features=True)
        self.out_shift = np.zeros(18, np.float32)
        self.out_scale = np.ones(18, np.float32)
        self.t_ref = 1.0  # typical |t|, for the response flag
        self._windows = {}  # (omega, loss, factor) -> (poles, k_rho windows)

    # -- data from the exact solver ------------------------------------------
    def exact_scattered(self, kx, ky, hs_nm, ho_nm, omega, loss):
        """g~_s (n, 9) complex (torch, device) for nodes sharing (h_s, h_o, omega, loss)."""
        st = self.family.stack(loss)
        _, g = green_nodes(st, omega, kx, ky, 1, 1, [hs_nm * NM], [1.0 - ho_nm * NM], cache=False,
                           terms="scattered", device=self.device)
        return g[0].T  # (n, 9), entry c + 3q

    def generate(self, n, seed=2, group=1024, pairs=4):
        """n training samples (x (n, 6), t (n, 18) float32 on the device).
        Samples come in groups sharing (omega, loss); inside a group, nodes
        (kx, ky) ~ U[-3 omega, 3 omega]^2 outside the flagged neighbourhoods of
        the light line and the origin, and `pairs` (h_s, h_o) draws."""
        torch = _torch()
        f = self.family
        rng = np.random.default_rng(seed)
        dev = torch.device("cuda", self.device)
        xs, ts = [], []
        per = group // pairs
        got = 0
        while got < n:
            omega = 2 * math.pi * rng.uniform(*f.F_RANGE)
            loss = rng.uniform(*f.LOSS_RANGE)
            kx = np.empty(0)
            ky = np.empty(0)
            while kx.size < per:
                a = rng.uniform(-f.K_RANGE, f.K_RANGE, 2 * per) * omega
                b = rng.uniform(-f.K_RANGE, f.K_RANGE, 2 * per) * omega
                kr = np.hypot(a, b)
                ok = (np.abs(kr - omega) >= self.BRANCH_TOL * omega) & (kr >= self.ORIGIN_TOL * omega)
                kx = np.concatenate([kx, a[ok]])
                ky = np.concatenate([ky, b[ok]])
            kx, ky = kx[:per], ky[:per]
            hs = rng.uniform(*f.H_RANGE, pairs)
            ho = rng.uniform(*f.H_RANGE, pairs)
            st = f.stack(loss)
            _, g = green_nodes(st, omega, kx, ky, 1, 1, hs * NM, 1.0 - ho * NM, cache=False, terms="scattered",
                               device=self.device)
            g = g.cpu().numpy()  # (pairs, 9, per)
            for p in range(pairs):
                xs.append(f.features(kx, ky, hs[p], ho[p], omega, loss))
                ts.append(target_from_ghat(g[p].T, kx, ky, hs[p] * NM, ho[p] * NM, omega).astype(np.float32))
            got += per * pairs
        x = torch.as_tensor(np.concatenate(xs)[:n], device=dev)
        t = torch.as_tensor(np.concatenate(ts)[:n], device=dev)
        return x, t

    # -- training ---------------------------------------------------------------
    def fit(self, x, t, epochs=8, batch=8192, lr=2e-3, lr_end=5e-5, seed=3, log_every=0):
        """Adam with cosine-decayed learning rate; returns the loss history (per epoch)."""
        torch = _torch()
        sh, sc = self.family.input_normalization()
        tn = t.double()
        self.out_shift = tn.mean(0).float().cpu().numpy()
        self.out_scale = tn.std(0).clamp_min(1e-12).float().cpu().numpy()
        self.mlp.set_normalization(sh, sc, self.out_shift, self.out_scale)
        mag = torch.linalg.norm(t.double(), dim=1)
        self.t_ref = float(torch.quantile(mag[: min(len(mag), 1 << 20)], 0.99))
        n = x.shape[0]
        steps_per = n // batch
        total = epochs * steps_per
        g = torch.Generator(device=x.device)
        g.manual_seed(seed)
        loss_dev = torch.zeros(1, dtype=torch.float32, device=x.device)
        hist = []
        step = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=x.device, generator=g)
            acc = torch.zeros((), dtype=torch.float64, device=x.device)
            for b in range(steps_per):
                idx = perm[b * batch:(b + 1) * batch]
                xb, tb = x[idx].contiguous(), t[idx].contiguous()
                lr_t = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * step / max(total - 1, 1)))
                self.mlp.train_step(xb, tb, lr_t, loss_out=loss_dev)
                acc += loss_dev[0]
                step += 1
            hist.append(float(acc) / steps_per)
            if log_every and (ep + 1) % log_every == 0:
                print(f"epoch {ep + 1}: loss {hist[-1]:.3e}", flush=True)
        return hist

    # -- the pole indicator -------------------------------------------------------
    def pole_windows(self, omega, loss, factor=8.0):
        """k_rho windows [q - factor Gamma, q + factor Gamma] around the poles of
        the cavity response of the family's stack at (omega, loss):
        real_axis_poles (mod_KSpaceFeatures.jl:151-214) on _cavity_response
        (sigma_max of RL, RR and (I - PF RL PB RR)^-1, :109-127) with the source in
        the air, from half the air light line to the evaluation range's corner,
        light lines excluded.  The film is gyrotropic about z, so the stack is
        symmetric about z and the poles are rings: one direction suffices.
        Returns (poles, windows (m, 2))."""
        key = (float(omega), float(loss), float(factor))
        if key not in self._windows:
            from . import quadrature as Q
            from .kspace import real_axis_poles
            st = self.family.stack(loss)
            ll = Q.dedup(sorted(x for lay in st for x in Q.layer_light_lines(lay, omega, 0.0)))
            k_hi = math.sqrt(2.0) * self.family.K_RANGE * omega
            poles = real_axis_poles(st, omega, 1, phis=(0.0,), k_lo=0.5 * omega, k_hi=k_hi, exclude=ll,
                                    device=self.device)
            win = np.array([[p["q"] - factor * p["Gamma"], p["q"] + factor * p["Gamma"]] for p in poles],
                           np.float64).reshape(-1, 2)
            self._windows[key] = (poles, win)
        return self._windows[key]

    def set_pole_windows(self, windows):
        """Load k_rho windows into the fused path's input flags (ytf_mlp_set_pole_windows)."""
        w = np.ascontiguousarray(np.asarray(windows, np.float64).reshape(-1, 2))
        _lib.check(_lib.lib().ytf_mlp_set_pole_windows(self.mlp.ctx.handle, self.mlp.handle, w.shape[0],
                                                       w.ctypes.data_as(ct.POINTER(ct.c_double))),
                   self.mlp.ctx.handle)

    # -- prediction -------------------------------------------------------------
    def flags(self, kx, ky, t_pred, omega, branch_tol=None, origin_tol=None, pole_factor=3.0, windows=()):
        """Nodes to solve exactly: near the air light line, near k = 0, inside a
        pole window (k_rho in [lo, hi]), or with a predicted |t| above
        pole_factor x the 99th percentile of the training targets."""
        torch = _torch()
        branch_tol = self.BRANCH_TOL if branch_tol is None else branch_tol
        origin_tol = self.ORIGIN_TOL if origin_tol is None else origin_tol
        kr = torch.sqrt(kx * kx + ky * ky)
        near_branch = (kr - omega).abs() < branch_tol * omega
        near_origin = kr < origin_tol * omega
        near_pole = torch.zeros_like(near_origin)
        for lo, hi in np.asarray(windows, np.float64).reshape(-1, 2):
            near_pole |= (kr >= lo) & (kr <= hi)
        big = torch.linalg.norm(t_pred.double(), dim=1) > pole_factor * self.t_ref
        return near_branch | near_origin | near_pole | big

    def scattered_ghat(self, kx, ky, hs_nm, ho_nm, omega, loss, fallback=True, branch_tol=None, origin_tol=None,
                       pole_factor=3.0, count=True, pole_window=8.0):
        """g~_s (n, 9) complex on the device; returns (g, flagged mask).
        One call of ytf_mlp_surrogate_ghat: the flags of the nodes near the
        light line or the origin, the network with the exact factors
        (ghat_from_target) and the response flags (`flags`) in one kernel, the
        exact FP64 solve of the input-flagged nodes before it and that of the
        response-flagged nodes after it.  count=True waits for the stream and
        sets `last_flagged`; count=False only queues the work.  pole_window > 0:
        nodes within that many half-widths of a cavity-response pole
        (pole_windows) are flagged from their inputs; 0 leaves the poles to the
        response flag alone."""
        torch = _torch()
        from .green import layer_structs
        wkey = (float(omega), float(loss), float(pole_window)) if pole_window > 0 else None
        if wkey != getattr(self, "_loaded_windows", "none"):
            self.set_pole_windows(self.pole_windows(omega, loss, pole_window)[1] if wkey else np.zeros((0, 2)))
            self._loaded_windows = wkey
        dev = torch.device("cuda", self.device)
        kx = torch.as_tensor(kx, dtype=torch.float64, device=dev).contiguous()
        ky = torch.as_tensor(ky, dtype=torch.float64, device=dev).contiguous()
        n = kx.shape[0]
        branch_tol = self.BRANCH_TOL if branch_tol is None else branch_tol
        origin_tol = self.ORIGIN_TOL if origin_tol is None else origin_tol
        g = torch.empty((9, n), dtype=torch.complex128, device=dev)
        flags = torch.empty(n, dtype=torch.uint8, device=dev)
        st = self.family.stack(loss)
        layers = layer_structs(st)
        xc = (ct.c_float * 4)(float(hs_nm), float(ho_nm), float(omega / (2 * math.pi)), float(loss))
        nflag = ct.c_int64()
        ctx = self.mlp.ctx
        ctx.bind_stream()
        _lib.check(_lib.lib().ytf_mlp_surrogate_ghat(
            ctx.handle, self.mlp.handle, layers, len(st), float(omega), _p(kx), _p(ky), n, xc,
            float((hs_nm + ho_nm) * NM), float(hs_nm * NM), float(st[0].d - ho_nm * NM), float(branch_tol),
            float(origin_tol), float(pole_factor * self.t_ref), 1 if fallback else 0, _p(g), _p(flags),
            ct.byref(nflag) if count else None), ctx.handle)
        if count:
            self.last_flagged = int(nflag.value)
        return g.T, flags.view(torch.bool)
