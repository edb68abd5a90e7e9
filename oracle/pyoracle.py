"""ctypes wrapper of the CPU oracle (ytf_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Stacks are passed as flat arrays: eps, mu (n, 3, 3) complex (converted to the
column-major interleaved layout the C side reads), d (n,), ek (n,) eps kind
(0 isotropic, 1 uniaxial, 2 biaxial, 3 general 3x3), mk (n,) mu kind (0 iso).
Layer indices are 1-based, as in the reference.
"""
import contextlib
import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ct.POINTER(ct.c_double)
_i = ct.POINTER(ct.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        _LIB = ct.CDLL(path)
    return _LIB


@contextlib.contextmanager
def using(path):
    """Run the block against another build of this same source (a variant with
    one planted mistake, oracle/_mutants/); the default build is restored after."""
    global _LIB
    saved = _LIB
    _LIB = ct.CDLL(path)  # RTLD_LOCAL: its symbols do not interpose on the default build's
    try:
        yield
    finally:
        _LIB = saved


def _p(a, typ=_d):
    return a.ctypes.data_as(typ)


def _flat33(m):
    """(n,3,3) complex -> (n*18,) doubles, column-major 3x3, interleaved."""
    m = np.asarray(m, dtype=np.complex128)
    colmaj = np.transpose(m, (0, 2, 1)).reshape(m.shape[0], 9)
    return np.ascontiguousarray(colmaj).view(np.float64).reshape(-1).copy()


class Stack:
    def __init__(self, eps, mu, d, ek, mk):
        self.eps = np.asarray(eps, dtype=np.complex128).reshape(-1, 3, 3)
        self.mu = np.asarray(mu, dtype=np.complex128).reshape(-1, 3, 3)
        self.d = np.ascontiguousarray(d, dtype=np.float64)
        self.ek = np.ascontiguousarray(ek, dtype=np.int32)
        self.mk = np.ascontiguousarray(mk, dtype=np.int32)
        self.n = len(self.d)
        self._e = _flat33(self.eps)
        self._m = _flat33(self.mu)

    def args(self):
        return (ct.c_int(self.n), _p(self._e), _p(self._m), _p(self.d), _p(self.ek, _i), _p(self.mk, _i))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def eigen4(A):
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    buf = A.reshape(-1, order="F").copy()
    Q = np.zeros(4, np.complex128)
    V = np.zeros(16, np.complex128)
    rc = lib().ora_eigen4(_p(buf.view(np.float64)), _p(Q.view(np.float64)), _p(V.view(np.float64)))
    if rc != 0:
        raise RuntimeError("oracle eigen4 did not converge")
    return Q, V.reshape(4, 4, order="F")


def interface_solve(L, Rh):
    """L \\ Rh as _S_Interface_Matrix solves it: LU, and on an exactly zero pivot
    the pivoted-QR minimum-norm solve.  Returns (S, fallback_taken)."""
    a = np.asarray(L, dtype=np.complex128).reshape(-1, order="F").copy()
    b = np.asarray(Rh, dtype=np.complex128).reshape(-1, order="F").copy()
    S = np.zeros(16, np.complex128)
    fb = lib().ora_interface_solve(_p(a.view(np.float64)), _p(b.view(np.float64)), _p(S.view(np.float64)))
    return S.reshape(4, 4, order="F"), bool(fb)


def qr_solve4(L, Rh):
    """The pivoted-QR minimum-norm solve of the fallback, taken unconditionally."""
    a = np.asarray(L, dtype=np.complex128).reshape(-1, order="F").copy()
    b = np.asarray(Rh, dtype=np.complex128).reshape(-1, order="F").copy()
    S = np.zeros(16, np.complex128)
    lib().ora_qr_solve4(_p(a.view(np.float64)), _p(b.view(np.float64)), _p(S.view(np.float64)))
    return S.reshape(4, 4, order="F")


def berreman(eps, mu, omega, kx, ky):
    e = _flat33(np.asarray(eps)[None])
    m = _flat33(np.asarray(mu)[None])
    B = np.zeros(16, np.complex128)
    lib().ora_berreman(_p(e), _p(m), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky), _p(B.view(np.float64)))
    return B.reshape(4, 4, order="F")


def modes_layer(eps, mu, ek, mk, omega, kx, ky):
    e = _flat33(np.asarray(eps)[None])
    m = _flat33(np.asarray(mu)[None])
    Q = np.zeros(4, np.complex128)
    V = np.zeros(16, np.complex128)
    M = np.zeros(36, np.complex128)
    rc = lib().ora_modes_layer(_p(e), _p(m), ct.c_int(ek), ct.c_int(mk), ct.c_double(omega), ct.c_double(kx),
                               ct.c_double(ky), _p(Q.view(np.float64)), _p(V.view(np.float64)), _p(M.view(np.float64)))
    if rc != 0:
        raise RuntimeError("oracle modes_layer failed (%d)" % rc)
    return Q, V.reshape(4, 4, order="F"), _unpack_modes(M)


def _unpack_modes(M):
    return dict(Vf=M[0:8].reshape(4, 2, order="F"), Vb=M[8:16].reshape(4, 2, order="F"),
                Wf=M[16:24].reshape(4, 2, order="F"), Wb=M[24:32].reshape(4, 2, order="F"),
                Qf=M[32:34].copy(), Qb=M[34:36].copy())


def stack_smatrix(st, omega, kx, ky):
    n = st.n
    modes = np.zeros(36 * n, np.complex128)
    Sint = np.zeros(16 * max(n - 1, 1), np.complex128)
    Sprop = np.zeros(16 * n, np.complex128)
    SL = np.zeros(16 * n, np.complex128)
    SR = np.zeros(16 * n, np.complex128)
    rc = lib().ora_stack_smatrix(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky),
                                 _p(modes.view(np.float64)), _p(Sint.view(np.float64)), _p(Sprop.view(np.float64)),
                                 _p(SL.view(np.float64)), _p(SR.view(np.float64)))
    if rc != 0:
        raise RuntimeError("oracle stack_smatrix failed")
    M = [_unpack_modes(modes[36 * l:36 * (l + 1)]) for l in range(n)]
    r = lambda a, l: a[16 * l:16 * (l + 1)].reshape(4, 4, order="F")
    return (M, [r(Sint, l) for l in range(n - 1)], [r(Sprop, l) for l in range(n)], [r(SL, l) for l in range(n)],
            [r(SR, l) for l in range(n)])


class Cache:
    """Oracle node cache: per node Q (4), sf, sb (2x3), and per obs layer T
    (4x4) + basis (16: Exf Eyf Ezf Exb Eyb Ezb KZf KZb, 2 modes each)."""

    def __init__(self, st, omega, src, obs, kx, ky, rec, valid, Tvalid):
        self.st, self.omega, self.src, self.obs = st, omega, src, list(obs)
        self.kx, self.ky = kx, ky
        self.rec, self.valid, self.Tvalid = rec, valid, Tvalid
        nobs = len(self.obs)
        rl = 16 + 32 * nobs
        R = rec.reshape(-1, rl)
        self.Qf = R[:, 0:2]
        self.Qb = R[:, 2:4]
        self.sf = R[:, 4:10].reshape(-1, 3, 2).transpose(0, 2, 1)
        self.sb = R[:, 10:16].reshape(-1, 3, 2).transpose(0, 2, 1)
        self.T = {}
        self.basis = {}
        for i, o in enumerate(self.obs):
            b = 16 + 32 * i
            self.T[o] = R[:, b:b + 16].reshape(-1, 4, 4).transpose(0, 2, 1)
            self.basis[o] = R[:, b + 16:b + 32].reshape(-1, 8, 2)
        self.Tv = {o: Tvalid.reshape(-1, nobs)[:, i].astype(bool) for i, o in enumerate(self.obs)}


def nodes_batch(st, omega, kx, ky, src, obs, nthreads=0):
    kx = np.ascontiguousarray(kx, np.float64)
    ky = np.ascontiguousarray(ky, np.float64)
    nn = len(kx)
    ob = np.ascontiguousarray(obs, np.int32)
    nobs = len(ob)
    rec = np.zeros(nn * (16 + 32 * nobs), np.complex128)
    valid = np.zeros(nn, np.int32)
    Tvalid = np.zeros(nn * nobs, np.int32)
    lib().ora_nodes_batch(*st.args(), ct.c_double(omega), _p(kx), _p(ky), ct.c_int64(nn), ct.c_int(src),
                          _p(ob, _i), ct.c_int(nobs), _p(rec.view(np.float64)), _p(valid, _i), _p(Tvalid, _i),
                          ct.c_int(nthreads))
    return Cache(st, omega, src, obs, kx, ky, rec, valid.astype(bool), Tvalid)


def ghat_from_cache(cache, obs_layer, zsrc, zobs, nthreads=0):
    zsrc = np.ascontiguousarray(np.atleast_1d(zsrc), np.float64)
    zobs = np.ascontiguousarray(np.atleast_1d(zobs), np.float64)
    npairs = len(zsrc)
    nn = len(cache.kx)
    g = np.zeros(nn * npairs * 9, np.complex128)
    ob = np.ascontiguousarray(cache.obs, np.int32)
    valid = np.ascontiguousarray(cache.valid, np.int32)
    rc = lib().ora_ghat_from_cache(ct.c_int(cache.st.n), _p(cache.st.d), ct.c_int(cache.src), _p(ob, _i),
                                   ct.c_int(len(ob)), ct.c_int(obs_layer), _p(cache.rec.view(np.float64)),
                                   _p(valid, _i), _p(cache.Tvalid, _i), ct.c_int64(nn), ct.c_int(npairs), _p(zsrc),
                                   _p(zobs), _p(g.view(np.float64)), ct.c_int(nthreads))
    if rc != 0:
        raise RuntimeError("ghat_from_cache: obs layer not in cache")
    # [node][pair][c + 3q] -> (nn, npairs, 3(c), 3(q))
    return g.reshape(nn, npairs, 3, 3).transpose(0, 1, 3, 2)


TERMS = {"all": 0, "direct": 1, "scattered": 2, "toeplitz": 3, "hankel": 4}


def ghat_batch(st, omega, kx, ky, src, obs_layer, zsrc, zobs, nthreads=0, terms="all", dz_box=0.0, ez_local=True):
    """Per-node spectral dyadic (n, npairs, c, q).  dz_box > 0: a box source of
    that thickness centred at each zsrc (_boxphase, interior_slab_at,
    ez_local_density, mod_ClosedFormField.jl:86-98, 815-862)."""
    kx = np.ascontiguousarray(kx, np.float64)
    ky = np.ascontiguousarray(ky, np.float64)
    zsrc = np.ascontiguousarray(np.atleast_1d(zsrc), np.float64)
    zobs = np.ascontiguousarray(np.atleast_1d(zobs), np.float64)
    nn, npairs = len(kx), len(zsrc)
    g = np.zeros(nn * npairs * 9, np.complex128)
    rc = lib().ora_ghat_batch_box(*st.args(), ct.c_double(omega), _p(kx), _p(ky), ct.c_int64(nn), ct.c_int(src),
                                  ct.c_int(obs_layer), ct.c_int(npairs), _p(zsrc), _p(zobs), _p(g.view(np.float64)),
                                  ct.c_int(nthreads), ct.c_int(TERMS[terms]), ct.c_double(dz_box),
                                  ct.c_int(1 if ez_local else 0))
    if rc != 0:
        raise ValueError("ghat_batch: the box does not fit inside its layer")
    return g.reshape(nn, npairs, 3, 3).transpose(0, 1, 3, 2)


def stack_S_batch(st, omega, kx, ky, nthreads=0):
    """Total S-matrix (SLeft[N]) and first/last layer modes per node:
    dict S (n,4,4), V1, VN (n,4,4) = [Vf Vb], Q1, QN (n,4), ok (n,) bool."""
    kx = np.ascontiguousarray(kx, np.float64)
    ky = np.ascontiguousarray(ky, np.float64)
    nn = len(kx)
    S = np.zeros(nn * 16, np.complex128)
    V1 = np.zeros(nn * 16, np.complex128)
    VN = np.zeros(nn * 16, np.complex128)
    Q1 = np.zeros(nn * 4, np.complex128)
    QN = np.zeros(nn * 4, np.complex128)
    ok = np.zeros(nn, np.int32)
    lib().ora_stack_S_batch(*st.args(), ct.c_double(omega), _p(kx), _p(ky), ct.c_int64(nn), _p(S.view(np.float64)),
                            _p(V1.view(np.float64)), _p(Q1.view(np.float64)), _p(VN.view(np.float64)),
                            _p(QN.view(np.float64)), _p(ok, _i), ct.c_int(nthreads))
    m = lambda a: a.reshape(nn, 4, 4).transpose(0, 2, 1)
    return dict(S=m(S), V1=m(V1), VN=m(VN), Q1=Q1.reshape(nn, 4), QN=QN.reshape(nn, 4), ok=ok.astype(bool))


def nudft2d1(x, y, c, modes=None, N1=None, N2=None, nthreads=0):
    """sum_j c_j exp(+i (k1 x_j + k2 y_j)) at the listed modes ((n, 2) ints),
    or on the whole (N1, N2) mode grid k = -floor(N/2).. (returned (N1, N2))."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    c = np.ascontiguousarray(c, np.complex128)
    grid = modes is None
    if grid:
        K1, K2 = np.meshgrid(np.arange(N1) - N1 // 2, np.arange(N2) - N2 // 2, indexing="ij")
        modes = np.stack([K1.ravel(), K2.ravel()], axis=1)
    modes = np.asarray(modes)
    k1 = np.ascontiguousarray(modes[:, 0], np.int32)
    k2 = np.ascontiguousarray(modes[:, 1], np.int32)
    f = np.zeros(len(k1), np.complex128)
    lib().ora_nudft2d1_modes(ct.c_int64(len(x)), _p(x), _p(y), _p(c.view(np.float64)), ct.c_int64(len(k1)),
                             _p(k1, _i), _p(k2, _i), _p(f.view(np.float64)), ct.c_int(nthreads))
    return f.reshape(N1, N2) if grid else f


def count_node_ops(st, omega, kx, ky, src, obs_layer, zsrc, zobs):
    out = np.zeros(10, np.uint64)
    lib().ora_count_node_ops(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky), ct.c_int(src),
                             ct.c_int(obs_layer), ct.c_double(zsrc), ct.c_double(zobs),
                             out.ctypes.data_as(ct.POINTER(ct.c_uint64)))
    keys = ["add", "mul", "div", "sqrt", "trans"]
    return {"solve": dict(zip(keys, map(int, out[:5]))), "ghat": dict(zip(keys, map(int, out[5:])))}


def count_stack_ops(st, omega, kx, ky):
    """Real add/mul/div, sqrt and transcendental counts of Modes_Stack + the
    total S (SLeft[N]) for one node, in the reference's formulation."""
    out = np.zeros(5, np.uint64)
    lib().ora_count_stack_ops(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky),
                              out.ctypes.data_as(ct.POINTER(ct.c_uint64)))
    return dict(zip(["add", "mul", "div", "sqrt", "trans"], map(int, out)))


def count_response_ops(st, omega, s, kx, ky):
    """Real add/mul/div, sqrt and transcendental counts of one _cavity_response
    evaluation in the reference's formulation."""
    out = np.zeros(5, np.uint64)
    lib().ora_count_response_ops(*st.args(), ct.c_double(omega), ct.c_int(s), ct.c_double(kx), ct.c_double(ky),
                                 out.ctypes.data_as(ct.POINTER(ct.c_uint64)))
    return dict(zip(["add", "mul", "div", "sqrt", "trans"], map(int, out)))


def modal_amp(st, omega, kx, ky, layer, z, sif, sib):
    sif = _c(sif); sib = _c(sib)
    Ef = np.zeros(2 * st.n, np.complex128)
    Eb = np.zeros(2 * st.n, np.complex128)
    rc = lib().ora_modal_amp(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky), ct.c_int(layer),
                             ct.c_double(z), _p(sif.view(np.float64)), _p(sib.view(np.float64)),
                             _p(Ef.view(np.float64)), _p(Eb.view(np.float64)))
    if rc != 0:
        raise RuntimeError("modal_amp failed")
    return Ef.reshape(st.n, 2), Eb.reshape(st.n, 2)


def modal_amp_profile(st, omega, kx, ky, layer, z, sif, sib, zs):
    sif = _c(sif); sib = _c(sib)
    zs = np.ascontiguousarray(zs, np.float64)
    Ef = np.zeros(2 * len(zs), np.complex128)
    Eb = np.zeros(2 * len(zs), np.complex128)
    rc = lib().ora_modal_amp_profile(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky),
                                     ct.c_int(layer), ct.c_double(z), _p(sif.view(np.float64)),
                                     _p(sib.view(np.float64)), _p(zs), ct.c_int(len(zs)), _p(Ef.view(np.float64)),
                                     _p(Eb.view(np.float64)))
    if rc != 0:
        raise RuntimeError("modal_amp_profile failed")
    return Ef.reshape(-1, 2), Eb.reshape(-1, 2)


def seed_from_source(st, omega, kx, ky, layer, z, sif, sib):
    sif = _c(sif); sib = _c(sib)
    u = np.zeros(4, np.complex128)
    lib().ora_seed_from_source(*st.args(), ct.c_double(omega), ct.c_double(kx), ct.c_double(ky), ct.c_int(layer),
                               ct.c_double(z), _p(sif.view(np.float64)), _p(sib.view(np.float64)),
                               _p(u.view(np.float64)))
    return u


def closed_form_G_1d(cache, obs_layer, wt, zsrc, zobs, xs, ys, with_direct=True, nthreads=0):
    wt = np.ascontiguousarray(wt, np.float64)
    zobs = np.ascontiguousarray(np.atleast_1d(zobs), np.float64)
    xs = np.ascontiguousarray(np.atleast_1d(xs), np.float64)
    ys = np.ascontiguousarray(np.atleast_1d(ys), np.float64)
    nx, ny, nz = len(xs), len(ys), len(zobs)
    G = np.zeros(nx * ny * nz * 9, np.complex128)
    ob = np.ascontiguousarray(cache.obs, np.int32)
    valid = np.ascontiguousarray(cache.valid, np.int32)
    rc = lib().ora_closed_form_G_1d(ct.c_int(cache.st.n), _p(cache.st.d), ct.c_double(cache.omega),
                                    ct.c_int(cache.src), _p(ob, _i), ct.c_int(len(ob)), ct.c_int(obs_layer),
                                    _p(cache.kx), _p(wt), _p(cache.rec.view(np.float64)), _p(valid, _i),
                                    _p(cache.Tvalid, _i), ct.c_int64(len(cache.kx)), ct.c_double(zsrc), _p(zobs),
                                    ct.c_int(nz), _p(xs), ct.c_int(nx), _p(ys), ct.c_int(ny),
                                    ct.c_int(1 if with_direct else 0), _p(G.view(np.float64)), ct.c_int(nthreads))
    if rc != 0:
        raise RuntimeError("closed_form_G_1d failed")
    return G.reshape((nx, ny, nz, 3, 3), order="F")


def spectral_apply(g, J, grid):
    """The transform pair of mod_UniformSource.jl:229-248 (numpy.fft stands in
    for FFTW): J^ = dx dy ramp fft(J), S_c = sum_q g_cq J^_q,
    E = (dkx dky Nx Ny / (2pi)^2) ifft(S conj(ramp)),
    ramp = cis(kx L + ky Ly) (_centering_ramp :145-152).
    g: (Nx, Ny, 3, 3) complex (c, q); J: (Nx, Ny, 3).  Returns E (Nx, Ny, 3)."""
    kx, ky = np.asarray(grid["kx"]), np.asarray(grid["ky"])
    ramp = np.exp(1j * (kx[:, None] * grid["L"] + ky[None, :] * grid["Ly"]))
    Jh = np.stack([grid["dx"] * grid["dy"] * ramp * np.fft.fft2(J[:, :, q]) for q in range(3)], axis=-1)
    S = np.einsum("xycq,xyq->xyc", g, Jh)
    pref = grid["dkx"] * grid["dky"] * grid["Nx"] * grid["Ny"] / (2 * np.pi) ** 2
    return np.stack([pref * np.fft.ifft2(S[:, :, c] * np.conj(ramp)) for c in range(3)], axis=-1)


def lateral_sum(g, W, kx, ky, xs, ys):
    """The exact separable lateral sum of closed_form_G's 2D path
    (mod_ClosedFormField.jl:743-753):
        G(x_i, y_j) = sum_n W_n g_n exp(i (kx_n x_i + ky_n y_j)),
    g: (n, 3, 3) spectral dyadics (c, q).  Returns (nx, ny, 3, 3)."""
    xs = np.atleast_1d(np.asarray(xs, np.float64))
    ys = np.atleast_1d(np.asarray(ys, np.float64))
    Px = np.exp(1j * np.outer(kx, xs))
    Py = np.exp(1j * np.outer(ky, ys))
    S = np.asarray(g) * np.asarray(W)[:, None, None]
    S = np.where(np.isfinite(S), S, 0.0)  # _finite_or_zero
    return np.einsum("nc,nd,nij->cdij", Px, Py, S)


def closed_form_G_2d(st, omega, kx, ky, W, src, obs_layer, zsrc, zobs, xs, ys, terms="all", nthreads=0):
    """closed_form_G for a CavityTransferCache2D and a point source
    (mod_ClosedFormField.jl:711-782) with the exact lateral sum: the oracle's
    per-node spectral dyadic times the node weight, summed over the nodes.
    Returns (nx, ny, nz, 3, 3) complex (x, y, z, c, q)."""
    zobs = np.atleast_1d(np.asarray(zobs, np.float64))
    g = ghat_batch(st, omega, kx, ky, src, obs_layer, np.full(len(zobs), float(zsrc)), zobs, nthreads=nthreads,
                   terms=terms)
    return np.stack([lateral_sum(g[:, p], W, kx, ky, xs, ys) for p in range(len(zobs))], axis=2)


def linear_convolution(G, P):
    """The lateral assembly as a literal sum over source pixels
    (mod_ClosedFormAssembly.jl:10 and apply_operator, mod_GreenOperator.jl:
    674-712, without the FFTs):
        E[i, j, o, r] = sum_{i', j', k, c} G[i - i', j - j', k, o, r, c] P[i', j', k, c].
    G: (2Nx-1, 2Ny-1, Ns, No, 3, 3) on the offsets -(N-1)..(N-1) (index
    offset + N - 1); P: (Nx, Ny, Ns, 3) moments.  Returns E (Nx, Ny, No, 3)."""
    Nx, Ny, Ns, _ = P.shape
    No = G.shape[3]
    E = np.zeros((Nx, Ny, No, 3), np.complex128)
    for a in range(Nx):
        for b in range(Ny):
            p = P[a, b]
            if not np.any(p):
                continue
            # window of offsets i - a for i = 0..Nx-1
            Gw = G[Nx - 1 - a:2 * Nx - 1 - a, Ny - 1 - b:2 * Ny - 1 - b]
            E += np.einsum("xykorc,kc->xyor", Gw, p)
    return E


def ghat_from_G(G, dx, dy):
    """build_closed_form_green's transform (mod_ClosedFormAssembly.jl:110-124;
    numpy.fft stands in for FFTW): G on the padded offsets (Px, Py, ...) ->
    fft over the two lateral axes, times (-1)^(i+j) dx dy / (Px Py)."""
    Px, Py = G.shape[:2]
    sign = (-1.0) ** np.add.outer(np.arange(Px), np.arange(Py))
    sc = (sign * dx * dy / (Px * Py)).reshape((Px, Py) + (1,) * (G.ndim - 2))
    return np.fft.fft2(G, axes=(0, 1)) * sc


def apply_ghat(Ghat, J):
    """apply_ghat (mod_ClosedFormAssembly.jl:155-196): J (Nx, Ny, 3) density,
    zero-padded to (Px, Py) and fft'd; E^ = Ghat . J^ per padded node;
    bfft and crop.  Ghat (Px, Py, Nz, 3, 3).  Returns E (Nx, Ny, Nz, 3)."""
    Px, Py = Ghat.shape[:2]
    Nx, Ny = J.shape[:2]
    Jp = np.zeros((Px, Py, 3), np.complex128)
    Jp[:Nx, :Ny] = J
    Jh = np.fft.fft2(Jp, axes=(0, 1))
    Eh = np.einsum("xyzrc,xyc->xyzr", Ghat, Jh)
    E = np.fft.ifft2(Eh, axes=(0, 1)) * (Px * Py)
    return E[:Nx, :Ny]


def lateral_sum_points(g, W, kx, ky, pts, chunk=8192):
    """The exact lateral sum of mod_ClosedFormField.jl:743-753 at a list of
    offsets pts (m, 2): G_m = sum_n W_n g_n exp(i (kx_n x_m + ky_n y_m)),
    summed over node chunks.  g: (n, 3, 3).  Returns (m, 3, 3)."""
    pts = np.asarray(pts, np.float64)
    S = np.asarray(g) * np.asarray(W)[:, None, None]
    S = np.where(np.isfinite(S), S, 0.0).reshape(len(kx), 9)
    out = np.zeros((len(pts), 9), np.complex128)
    for a in range(0, len(kx), chunk):
        ph = np.exp(1j * (np.outer(pts[:, 0], kx[a:a + chunk]) + np.outer(pts[:, 1], ky[a:a + chunk])))
        out += ph @ S[a:a + chunk]
    return out.reshape(len(pts), 3, 3)


def _qlist(*arrays):
    """Distinct finite non-zero values of node arrays (_qlist, mod_GreenOperator.jl:251-258)."""
    v = np.concatenate([np.asarray(a, np.complex128).ravel() for a in arrays]) if arrays else np.zeros(0, complex)
    v = v[np.isfinite(v) & (v != 0)]
    return np.unique(v)


def family_rows(fam, z):
    """Rows exp(i Q (sigma z_j + c)) of the family fam = [(Q, sigma, c), ...]
    (node x plane), underflowing to 0 below exp(-700) (_family_rows,
    mod_GreenOperator.jl:222-230)."""
    z = np.asarray(z, np.float64)
    out = []
    for Q, sig, c in fam:
        x = 1j * np.asarray(Q, np.complex128)[:, None] * (sig * z[None, :] + c)
        out.append(np.where(x.real < -700, 0.0, np.exp(np.where(x.real < -700, 0.0, x))))
    return np.vstack(out) if out else np.zeros((0, len(z)), complex)


def zskeleton(z, fam, tol, whole_frac=0.8):
    """_zskeleton (mod_GreenOperator.jl:262-302) with numpy / scipy (LAPACK)
    factorizations: R of the family's QR, column-pivoted QR of R (geqp3, as
    qr(R, ColumnNorm())), the interpolative decomposition of the smallest
    rank whose largest error over every row is <= tol (bisection), kept whole
    at >= whole_frac of the planes.  Returns dict(idx, P, err, whole)."""
    import scipy.linalg as sl
    z = np.asarray(z, np.float64)
    N = len(z)
    whole = dict(idx=np.arange(N), P=np.eye(N, dtype=complex), err=0.0, whole=True)
    if N <= 1:
        return whole
    A = family_rows(fam, z)
    R = np.linalg.qr(A, mode="r") if A.shape[0] else np.zeros((0, N), complex)
    if R.shape[0] < N:
        R = np.vstack([R, np.zeros((N - R.shape[0], N), complex)])
    Rt, p = sl.qr(R, mode="r", pivoting=True)
    Rt = Rt[:N]

    def interp(k):
        P = np.zeros((N, k), complex)
        P[p[:k], np.arange(k)] = 1
        if k < N:
            # back substitution; an exactly zero pivot (fewer independent rows
            # than planes) takes zero weights where the reference's
            # triangular solve would raise
            T = np.zeros((k, N - k), complex)
            for i in range(k - 1, -1, -1):
                if Rt[i, i] != 0:
                    T[i] = (Rt[i, k:] - Rt[i, i + 1:k] @ T[i + 1:k]) / Rt[i, i]
            P[p[k:], :] = T.T
        perm = np.argsort(p[:k])
        return p[:k][perm], P[:, perm]

    def maxerr(idx, P):
        return float(np.max(np.abs(A - A[:, idx] @ P.T))) if A.size else 0.0

    lo, hi = 0, N
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if maxerr(*interp(mid)) <= tol:
            hi = mid
        else:
            lo = mid
    if hi >= whole_frac * N:
        return whole
    idx, P = interp(hi)
    return dict(idx=idx, P=P, err=maxerr(idx, P), whole=False)


def cavity_response(st, omega, s, kx, ky, nthreads=0):
    """_cavity_response (mod_KSpaceFeatures.jl:109-127) at nodes (kx, ky); s 1-based."""
    kx = np.ascontiguousarray(np.atleast_1d(kx), np.float64)
    ky = np.ascontiguousarray(np.atleast_1d(ky), np.float64)
    out = np.zeros(len(kx))
    rc = lib().ora_cavity_response(*st.args(), ct.c_double(omega), ct.c_int(s), _p(kx), _p(ky),
                                   ct.c_int64(len(kx)), _p(out), ct.c_int(nthreads))
    if rc:
        raise ValueError("cavity_response: source layer out of range")
    return out


def real_axis_poles(st, omega, s, phis=(0.0,), k_lo=None, k_hi=None, scan_step=1e-3, prominence=1.05,
                    exclude=(), floor_rel=1e-6, floor_abs=1e-10):
    """real_axis_poles (mod_KSpaceFeatures.jl:151-214), step by step as the
    reference takes it: the log-spaced scan, the local maxima with their
    prominence over the valleys within a factor 4 in k, golden-section search
    of each maximum, and the half-level crossings by bisection.  Returns a list
    of (q, Gamma, phi, prom)."""
    # Julia's 0.0:step:L includes the last multiple of step <= L
    L = np.log(k_hi / k_lo)
    nstep = int(np.floor(L / scan_step + 1e-10))
    ks = k_lo * np.exp(np.arange(nstep + 1) * scan_step)
    n = len(ks)
    W = int(np.ceil(np.log(4.0) / scan_step))
    out = []
    for phi in phis:
        c, sn = np.cos(phi), np.sin(phi)

        def f(k):
            return float(cavity_response(st, omega, s, [k * c], [k * sn])[0])

        y = np.log(cavity_response(st, omega, s, ks * c, ks * sn) + 1e-300)
        ymin = max(y.max() + np.log(floor_rel), np.log(floor_abs))
        for i in range(1, n - 1):
            if not (y[i] > y[i - 1] and y[i] >= y[i + 1] and y[i] >= ymin):
                continue
            jl, lmin = i, y[i]
            while jl > max(0, i - W) and y[jl - 1] <= y[i]:
                jl -= 1
                lmin = min(lmin, y[jl])
            jr, rmin = i, y[i]
            while jr < min(n - 1, i + W) and y[jr + 1] <= y[i]:
                jr += 1
                rmin = min(rmin, y[jr])
            base = max(lmin, rmin)
            prom = y[i] - base
            if prom < np.log(prominence):
                continue
            a, b = ks[i - 1], ks[i + 1]
            g = (np.sqrt(5.0) - 1) / 2
            cc, dd = b - g * (b - a), a + g * (b - a)
            fc, fd = f(cc), f(dd)
            for _ in range(60):
                if (b - a) <= 1e-12 * b:
                    break
                if fc >= fd:
                    b, dd, fd = dd, cc, fc
                    cc = b - g * (b - a)
                    fc = f(cc)
                else:
                    a, cc, fc = cc, dd, fd
                    dd = a + g * (b - a)
                    fd = f(dd)
            q = (a + b) / 2
            P = max(f(q), np.exp(y[i]))
            if any(abs(q - Lx) <= 1e-6 * q for Lx in exclude):
                continue
            lvl = np.exp(base) + (P - np.exp(base)) / np.sqrt(2.0)

            def cross(lo, hi):
                for _ in range(50):
                    if abs(hi - lo) <= 1e-12 * q:
                        break
                    mid = (lo + hi) / 2
                    if f(mid) > lvl:
                        lo = mid
                    else:
                        hi = mid
                return (lo + hi) / 2

            kl = ks[jl + int(np.nonzero(y[jl:i + 1] == lmin)[0][-1])]
            kr = ks[i + int(np.nonzero(y[i:jr + 1] == rmin)[0][0])]
            G = min(q - cross(q, kl), cross(q, kr) - q)
            out.append((q, max(G, 1e-12 * q), float(phi), prom))
    return out


def op_zmix(nxy, J, no, T=None, H=None, D=None, Ps=None, Po=None, E=None, nthreads=0):
    """The z coupling of one operator block at every spectral node
    (ora_op_zmix): J (ns, 3, nxy) complex; T, H, D: plane stacks (slots, 3, 3,
    nxy) or None; Ps (ns, rs), Po (no, ro) for a skeleton block.  Returns E
    (no, 3, nxy), accumulated into E when given."""
    J = np.ascontiguousarray(J, np.complex128)
    ns = J.shape[0]
    E = np.zeros((no, 3, nxy), np.complex128) if E is None else E
    assert E.flags.c_contiguous and E.dtype == np.complex128

    def arr(a):
        return None if a is None else np.ascontiguousarray(a, np.complex128)
    T, H, D, Ps, Po = arr(T), arr(H), arr(D), arr(Ps), arr(Po)
    ptr = lambda a: None if a is None else _p(a.view(np.float64))
    rs = Ps.shape[1] if Ps is not None else 0
    ro = Po.shape[1] if Po is not None else 0
    lib().ora_op_zmix(ct.c_int64(nxy), ct.c_int(ns), ptr(J), ct.c_int(no), ptr(E), ct.c_int(0 if T is None else len(T)),
                      ptr(T), ct.c_int(0 if H is None else len(H)), ptr(H), ct.c_int(rs), ct.c_int(ro), ptr(D), ptr(Ps),
                      ptr(Po), ct.c_int(nthreads))
    return E

