"""Tabulated k_rho surrogate of the scattered spectral dyadic (host side).

No reference counterpart: the reference solves every (kx, ky) node exactly
(build_cavity_transfer_2d + closed_form_G's terms=:scattered).  For a stack
that is invariant under rotations about z, the dyadic at any node is a
rotation of its value on the kx axis,

    g~_s(kx, ky) = R(phi) g~_s(k_rho, 0) R(phi)^T,

the symmetry behind the reference's 1D k_rho treatment of such stacks
(CavityTransferCache1D, mod_CavityTransfer.jl; _rotate_modes,
mod_ModeSolver.jl:471-482).  With source and observation point in layer 1
above the stack, g~_s(k_rho, 0) = P omega (1 + k_rho^2/omega^2) / kz0 t(k_rho)
with the air propagator P = exp(i kz0 (h_s + h_o)) applied exactly, so the
factor t depends on k_rho alone (for one omega and stack), not on the heights.

KrhoTable samples t with the exact FP64 node solve at phi = 0 and fits it by
Chebyshev series on adaptively bisected k_rho panels.  t has square-root
branch points on the real k_rho axis at the light lines of the lossless
semi-infinite layers (the air above, k_b = omega, and a lossless substrate),
and is analytic in that layer's k_z there, so around each such k_b the panel
variable is that k_z:
    kind 0: s = k_rho (from 0 to the first k_b / sqrt 2);
    kind 1: s = sqrt(k_b^2 - k_rho^2) (below k_b);
    kind 2: s = sqrt(k_rho^2 - k_b^2) (above k_b, to midway to the next one
            or to k_max).
Lossy layers put their branch points off the axis; the bisection resolves
them, and the poles of guided modes, like any other sharp feature.
A panel is accepted when the series matches exact solves at points between
its samples within tol (relative to the largest |t| sampled); a panel that
cannot be split further is marked for the exact solve.  For evaluation
each panel's series is sampled on a uniform grid in s, fine enough that
four-point (cubic) Lagrange interpolation stays within tol of the series;
the kernel (ytf_ktable_ghat) then reads four rows per node from the
entry-major device copy: panel search, interpolation, exact factors,
rotation, and the exact solve of the flagged nodes.
"""
import ctypes as ct
import math
import time

import numpy as np

from . import _lib
from .green import green_nodes, layer_structs

NM = 1.0 / 600.0


def _torch():
    import torch
    return torch


def rotation_invariant(stack):
    """Every layer's eps and mu commute with the rotations about z: the form
    [[a, b, 0], [-b, a, 0], [0, 0, c]] (isotropic, uniaxial about z,
    gyrotropic about z)."""
    for lay in stack:
        for m in (np.asarray(lay.eps.eps), np.asarray(lay.mu.eps)):
            if (m[0, 2] != 0 or m[1, 2] != 0 or m[2, 0] != 0 or m[2, 1] != 0 or m[0, 0] != m[1, 1]
                    or m[0, 1] != -m[1, 0]):
                return False
    return True


def branch_points(stack, omega):
    """The real k_rho branch points of t: the light lines of the lossless
    semi-infinite layers (first and last), sqrt(eps mu) omega for the
    in-plane and the z permittivity."""
    out = set()
    for lay in (stack[0], stack[-1]):
        mu = complex(lay.mu.eps[0, 0])
        for e in (complex(lay.eps.eps[0, 0]), complex(lay.eps.eps[2, 2])):
            v = np.sqrt(e * mu)
            if v.imag == 0.0 and v.real > 0.0:
                out.add(float(v.real) * omega)
    return sorted(out)


def _kz0(k2, omega):
    kz = np.sqrt((omega ** 2 - k2).astype(np.complex128))
    return np.where(kz.imag < 0, -kz, kz)


class KrhoTable:
    """The table of t(k_rho) of one stack at one omega, k_rho in [0, k_max]."""

    def __init__(self, stack, omega, k_max, tol=1e-10, deg=16, h_nm=(40.0, 50.0), device=0, min_width=1e-7,
                 max_panels=2048):
        if not rotation_invariant(stack):
            raise ValueError("KrhoTable: every layer must be invariant under rotations about z "
                             "(eps, mu of the form [[a, b, 0], [-b, a, 0], [0, 0, c]])")
        if not (np.array_equal(stack[0].eps.eps, np.eye(3)) and np.array_equal(stack[0].mu.eps, np.eye(3))):
            raise ValueError("KrhoTable: the exact factors take layer 1 (source and observation) as vacuum")
        if not (k_max > omega > 0):
            raise ValueError("KrhoTable: need 0 < omega < k_max")
        if not (1 <= deg <= 64):
            raise ValueError("KrhoTable: deg must be in 1..64")
        self.stack, self.omega, self.k_max, self.tol, self.deg = stack, float(omega), float(k_max), tol, int(deg)
        self.h_nm = h_nm
        self.device = device
        t0 = time.perf_counter()
        self._build(min_width, max_panels)
        self.build_s = time.perf_counter() - t0

    # -- the panel variable --------------------------------------------------
    @staticmethod
    def _krho(kind, s, kb):
        if kind == 0:
            return s
        if kind == 1:
            return np.sqrt(np.maximum(kb * kb - s * s, 0.0))
        return np.sqrt(kb * kb + s * s)

    @staticmethod
    def _s(kind, kr, kb):
        if kind == 0:
            return kr
        return math.sqrt(max(kb * kb - kr * kr if kind == 1 else kr * kr - kb * kb, 0.0))

    def _t_exact(self, kr):
        """t at (k_rho, 0) from the exact scattered solve (n, 9) complex."""
        torch = _torch()
        hs, ho = self.h_nm
        st = self.stack
        kx = torch.as_tensor(np.asarray(kr, np.float64), device=torch.device("cuda", self.device))
        ky = torch.zeros_like(kx)
        _, g = green_nodes(st, self.omega, kx, ky, 1, 1, [hs * NM], [st[0].d - ho * NM], cache=False,
                           terms="scattered", device=self.device)
        g = g[0].T.cpu().numpy()
        k2 = np.asarray(kr, np.float64) ** 2
        kz = _kz0(k2, self.omega)
        P = np.exp(1j * kz * (hs + ho) * NM)
        return g * (kz / (self.omega * P * (1.0 + k2 / self.omega ** 2)))[:, None]

    def _build(self, min_width, max_panels):
        w, deg = self.omega, self.deg
        m = deg + 1
        xs = np.cos(math.pi * (np.arange(m) + 0.5) / m)       # first-kind Chebyshev points (interior)
        xc = np.cos(math.pi * np.arange(1, m) / m)            # check points between them
        Tk = np.cos(np.outer(np.arange(m), math.pi * (np.arange(m) + 0.5) / m))   # [k][j]
        # segments of k_rho: [0, b1/sqrt 2] kind 0; around every branch point b,
        # kind 1 down to b/sqrt 2 (or midway from the previous one) and kind 2 up
        # to midway to the next one (or k_max)
        bps = [b for b in branch_points(self.stack, w) if b < self.k_max]
        segs = []  # (kind, k_b, s_lo, s_hi)
        lo = bps[0] / math.sqrt(2.0)
        segs.append((0, 0.0, 0.0, lo))
        for q, b in enumerate(bps):
            a = lo
            hi = 0.5 * (b + bps[q + 1]) if q + 1 < len(bps) else self.k_max
            segs.append((1, b, 0.0, math.sqrt(b * b - a * a)))
            segs.append((2, b, 0.0, math.sqrt(hi * hi - b * b)))
            lo = hi
        # initial panels: breaks at the other layers' light lines, at most omega wide in k_rho
        lls = set()
        for lay in self.stack:
            for e in (lay.eps.eps[0, 0], lay.eps.eps[2, 2]):
                lls.add(float(np.sqrt(complex(e) * complex(lay.mu.eps[0, 0])).real * w))
        pending = []
        for kind, b, sa, sb in segs:
            k_a, k_b = sorted((float(self._krho(kind, sa, b)), float(self._krho(kind, sb, b))))
            cuts = sorted({sa, sb} | {self._s(kind, k, b) for k in lls if k_a < k < k_b})
            for a0, b0 in zip(cuts[:-1], cuts[1:]):
                k0, k1 = sorted((float(self._krho(kind, a0, b)), float(self._krho(kind, b0, b))))
                nsub = max(2, int(math.ceil((k1 - k0) / w)))
                for q in range(nsub):
                    pending.append((kind, b, a0 + (b0 - a0) * q / nsub, a0 + (b0 - a0) * (q + 1) / nsub))
        done = []
        scale = 0.0
        self.n_exact = 0
        while pending:
            S = []
            for kind, kb, a, b in pending:
                c, h = 0.5 * (a + b), 0.5 * (b - a)
                S.append(self._krho(kind, np.concatenate([c + h * xs, c + h * xc]), kb))
            t = self._t_exact(np.concatenate(S))
            self.n_exact += t.shape[0]
            if not np.all(np.isfinite(t)):
                raise RuntimeError("KrhoTable: the exact solve returned non-finite values")
            t = t.reshape(len(pending), 2 * m - 1, 9)
            scale = max(scale, float(np.abs(t).max()))
            nxt = []
            for q, (kind, kb, a, b) in enumerate(pending):
                vals = np.concatenate([t[q, :m].real, t[q, :m].imag], axis=1)      # (m, 18)
                coef = (2.0 / m) * (Tk @ vals)
                coef[0] *= 0.5
                chk = np.concatenate([t[q, m:].real, t[q, m:].imag], axis=1)
                fit = np.polynomial.chebyshev.chebval(xc, coef).T              # (m - 1, 18)
                err = float(np.abs(fit - chk).max())
                if err <= self.tol * scale:
                    done.append((kind, kb, a, b, coef, 0))
                elif (b - a) < min_width * w or len(done) + len(pending) + len(nxt) >= max_panels:
                    done.append((kind, kb, a, b, coef, 4))
                else:
                    mid = 0.5 * (a + b)
                    nxt += [(kind, kb, a, mid), (kind, kb, mid, b)]
            pending = nxt
        # panels in increasing k_rho
        rows = []
        for kind, kb, a, b, coef, fl in done:
            k1, k2 = sorted((float(self._krho(kind, a, kb)), float(self._krho(kind, b, kb))))
            rows.append((k1, k2, kind | fl, a, b, kb, coef))
        rows.sort(key=lambda r: r[0])
        for r0, r1 in zip(rows[:-1], rows[1:]):
            if abs(r0[1] - r1[0]) > 1e-12 * w:
                raise RuntimeError("KrhoTable: panels are not contiguous")
        self.n_panels = len(rows)
        self.n_unconverged = sum(1 for r in rows if r[2] & 4)
        self.scale = scale
        self.kb = np.array([r[0] for r in rows] + [rows[-1][1]], np.float64)
        self.kind = np.array([r[2] for r in rows], np.int32)
        self.sab = np.array([[r[3], r[4], r[5]] for r in rows], np.float64).ravel()
        self.coef = np.ascontiguousarray(np.stack([r[6] for r in rows]), np.float64)   # (np, deg + 1, 18)
        self._fine_grids()
        torch = _torch()
        dev = torch.device("cuda", self.device)
        # the device copy of the rows is entry-major: entry e of row r at e * n_rows + r
        self._dev = [torch.as_tensor(a, device=dev).contiguous()
                     for a in (self.kb, self.kind, self.sab, self.row0, self.m, np.ascontiguousarray(self.val.T))]

    @staticmethod
    def _cubic(f, u):
        """The kernel's four-point Lagrange interpolation of rows f (M, 18) at
        grid coordinates u in [0, M - 1]."""
        M = f.shape[0]
        j = np.clip(np.floor(u).astype(np.int64), 1, M - 3)
        r = u - j
        rm, rp, r2 = r - 1.0, r + 1.0, r - 2.0
        w = np.stack([-r * rm * r2 / 6.0, 0.5 * rp * rm * r2, -0.5 * rp * r * r2, rp * r * rm / 6.0])
        return sum(w[q][:, None] * f[j - 1 + q] for q in range(4))

    def _fine_grids(self, m_max=1 << 16):
        """Uniform samples of every panel's series, doubled until cubic
        interpolation matches the series within tol at three points in every
        interval; a panel that needs more than m_max samples is left to the
        exact solve."""
        from numpy.polynomial.chebyshev import chebval
        vals, row0, ms = [], [], []
        r = 0
        for q in range(self.n_panels):
            c = self.coef[q]
            M = 9
            while True:
                x = np.linspace(-1.0, 1.0, M)
                f = chebval(x, c).T
                ut = (np.arange(M - 1)[:, None] + np.array([0.25, 0.5, 0.75])[None, :]).ravel()
                err = float(np.abs(self._cubic(f, ut) - chebval(ut / (M - 1) * 2.0 - 1.0, c).T).max())
                if err <= self.tol * self.scale or M > m_max:
                    break
                M = 2 * M - 1
            if M > m_max:
                self.kind[q] |= 4
            vals.append(f)
            row0.append(r)
            ms.append(M)
            r += M
        self.n_unconverged = int(np.count_nonzero(self.kind & 4))
        self.val = np.ascontiguousarray(np.concatenate(vals), np.float64)
        self.row0 = np.array(row0, np.int64)
        self.m = np.array(ms, np.int32)

    def t(self, kr):
        """t from the table at k_rho values (host numpy, for checks), (n, 9) complex."""
        kr = np.atleast_1d(np.asarray(kr, np.float64))
        out = np.full((kr.size, 9), np.nan + 0j)
        p = np.clip(np.searchsorted(self.kb, kr, side="right") - 1, 0, self.n_panels - 1)
        for i, (k, q) in enumerate(zip(kr, p)):
            a, b, kb = self.sab[3 * q], self.sab[3 * q + 1], self.sab[3 * q + 2]
            s = self._s(int(self.kind[q] & 3), k, kb)
            x = min(1.0, max(-1.0, (2 * s - a - b) / (b - a)))
            v = np.polynomial.chebyshev.chebval(x, self.coef[q])
            out[i] = v[:9] + 1j * v[9:]
        return out

    def ghat(self, kx, ky, hs_nm, ho_nm, fallback=True, count=True, out=None):
        """g~_s (n, 9) complex on the device and the flag mask; one call of
        ytf_ktable_ghat.  count=True waits and sets `last_flagged`; out: a
        ((9, n) complex128, (n,) uint8) pair of device tensors to write into."""
        torch = _torch()
        dev = torch.device("cuda", self.device)
        kx = torch.as_tensor(kx, dtype=torch.float64, device=dev).contiguous()
        ky = torch.as_tensor(ky, dtype=torch.float64, device=dev).contiguous()
        n = kx.shape[0]
        if out is None:
            g = torch.empty((9, n), dtype=torch.complex128, device=dev)
            flags = torch.empty(n, dtype=torch.uint8, device=dev)
        else:
            g, flags = out
            if g.shape != (9, n) or g.dtype != torch.complex128 or flags.shape != (n,) or flags.dtype != torch.uint8 \
                    or not (g.is_contiguous() and flags.is_contiguous()):
                raise ValueError("KrhoTable.ghat: out must be contiguous (9, n) complex128 and (n,) uint8 tensors")
        st = self.stack
        if not hasattr(self, "_layers"):
            self._layers = layer_structs(st)
        nflag = ct.c_int64()
        ctx = _lib.Context.get(self.device)
        ctx.bind_stream()
        p = lambda t: ct.c_void_p(t.data_ptr())
        _lib.check(_lib.lib().ytf_ktable_ghat(
            ctx.handle, self._layers, len(st), self.omega, self.n_panels, *[p(a) for a in self._dev], p(kx), p(ky),
            n, float((hs_nm + ho_nm) * NM), float(hs_nm * NM), float(st[0].d - ho_nm * NM), 1 if fallback else 0,
            p(g), p(flags), ct.byref(nflag) if count else None), ctx.handle)
        if count:
            self.last_flagged = int(nflag.value)
        return g.T, flags.view(torch.bool)
