"""ctypes binding of libytf.so (the C-ABI of include/ytf.h).

The library is built in-tree by yellowthinfilms_amd/build.py.  There is no
fallback: if libytf.so is missing or cannot be loaded, every entry point
raises.
"""
import ctypes as ct
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# YTF_LIBRARY: another build of the same library (same C-ABI), e.g. one made
# with different compiler options
LIB_PATH = os.environ.get("YTF_LIBRARY") or os.path.join(HERE, "libytf.so")

YTF_OK, YTF_EINVAL, YTF_EHIP = 0, 1, 2
NODE_VALID, NODE_TVALID, NODE_SINGULAR, NODE_NOMODES = 1, 2, 4, 8


class YtfLayer(ct.Structure):
    _fields_ = [("eps", ct.c_double * 18), ("mu", ct.c_double * 18), ("d", ct.c_double),
                ("eps_kind", ct.c_int32), ("mu_kind", ct.c_int32)]


class YtfNodeCache(ct.Structure):
    _fields_ = [("Qsrc_dev", ct.c_void_p), ("sf_dev", ct.c_void_p), ("sb_dev", ct.c_void_p),
                ("T_dev", ct.c_void_p), ("basis_dev", ct.c_void_p), ("flags_dev", ct.c_void_p)]


class YtfGrid(ct.Structure):
    _fields_ = [("Nx", ct.c_int32), ("Ny", ct.c_int32), ("dx", ct.c_double), ("dy", ct.c_double)]


class YtfOperatorOptions(ct.Structure):
    _fields_ = [("structure", ct.c_int32), ("check", ct.c_int32), ("check_tol", ct.c_double),
                ("skeleton", ct.c_int32), ("skeleton_tol", ct.c_double), ("lateral", ct.c_int32),
                ("nufft_eps", ct.c_double), ("uniform_z", ct.c_int32), ("dz_box", ct.c_double),
                ("ez_local", ct.c_int32), ("skeletons", ct.c_void_p), ("rot_n_phi", ct.c_int32),
                ("rot_n_krho", ct.c_int64), ("rot_krho", ct.POINTER(ct.c_double)),
                ("rot_phi", ct.POINTER(ct.c_double))]


class YtfBlockInfo(ct.Structure):
    _fields_ = [("obs_layer", ct.c_int32), ("kind", ct.c_int32), ("n_obs", ct.c_int32), ("n_src_skel", ct.c_int32),
                ("n_obs_skel", ct.c_int32), ("n_toeplitz", ct.c_int32), ("n_hankel", ct.c_int32),
                ("n_dense", ct.c_int32), ("bytes", ct.c_int64), ("build_ms", ct.c_double),
                ("check_residual", ct.c_double)]


class YtfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None
_lock = threading.Lock()

_SIGS = {
    "ytf_version": (ct.c_char_p, []),
    "ytf_ctx_create": (ct.c_int, [ct.c_int, ct.POINTER(ct.c_void_p)]),
    "ytf_ctx_destroy": (ct.c_int, [ct.c_void_p]),
    "ytf_last_error": (ct.c_char_p, [ct.c_void_p]),
    "ytf_ctx_set_stream": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_ctx_synchronize": (ct.c_int, [ct.c_void_p]),
    "ytf_ctx_set_timing": (ct.c_int, [ct.c_void_p, ct.c_int]),
    "ytf_ctx_set_timing_only": (ct.c_int, [ct.c_void_p, ct.c_char_p]),
    "ytf_ctx_kernel_time": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]),
    "ytf_green_nodes": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_void_p,
                                   ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int, ct.c_int,
                                   ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.c_void_p,
                                   ct.POINTER(YtfNodeCache), ct.c_int]),
    "ytf_green_nodes_box": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_void_p,
                                       ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int, ct.c_int,
                                       ct.POINTER(ct.c_double), ct.POINTER(ct.c_double), ct.c_void_p,
                                       ct.POINTER(YtfNodeCache), ct.c_int, ct.c_double, ct.c_int]),
    "ytf_cavity_response": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_int,
                                       ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p]),
    "ytf_pole_refine": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_int, ct.c_int64,
                                   ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "ytf_pole_candidates": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int,
                                       ct.c_double, ct.c_double, ct.c_double, ct.c_int, ct.c_void_p, ct.c_void_p,
                                       ct.POINTER(ct.c_int64)]),
    "ytf_spectral_apply": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                      ct.c_void_p, ct.c_double]),
    "ytf_spectral_apply_noalloc": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p,
                                              ct.c_void_p, ct.c_double]),
    "ytf_radial_synthesis": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p,
                                        ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int]),
    "ytf_radial_terms": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p,
                                    ct.c_int, ct.c_void_p, ct.c_void_p]),
    "ytf_closed_form_G_1d": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p,
                                        ct.POINTER(ct.c_double), ct.c_int, ct.POINTER(ct.c_double), ct.c_int,
                                        ct.c_void_p, ct.c_int]),
    "ytf_batch_create": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.POINTER(ct.c_int32),
                                    ct.POINTER(ct.c_int64), ct.c_int, ct.c_double, ct.POINTER(ct.c_void_p)]),
    "ytf_batch_destroy": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_batch_solve": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                   ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "ytf_batch_set_config": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int]),
    "ytf_batch_get_config": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int), ct.POINTER(ct.c_int),
                                        ct.POINTER(ct.c_int)]),
    "ytf_batch_tune": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                  ct.c_char_p, ct.c_char_p, ct.c_int]),
    "ytf_mlp_create": (ct.c_int, [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_uint64,
                                  ct.POINTER(ct.c_void_p)]),
    "ytf_mlp_destroy": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_mlp_num_params": (ct.c_int64, [ct.c_void_p]),
    "ytf_mlp_get_params": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_float)]),
    "ytf_mlp_set_params": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_float)]),
    "ytf_mlp_set_normalization": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_float), ct.POINTER(ct.c_float),
                                             ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]),
    "ytf_mlp_forward": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p]),
    "ytf_mlp_loss_grad": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p,
                                     ct.POINTER(ct.c_float), ct.POINTER(ct.c_float)]),
    "ytf_mlp_train_step": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_float,
                                      ct.c_void_p]),
    "ytf_mlp_set_pole_windows": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.POINTER(ct.c_double)]),
    "ytf_mlp_surrogate_ghat": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double,
                                          ct.c_void_p, ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_float), ct.c_double,
                                          ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_int,
                                          ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "ytf_ktable_ghat": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_int,
                                   ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                   ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_double, ct.c_double, ct.c_double,
                                   ct.c_int, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_int64)]),
    "ytf_nufft2d1": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int,
                                ct.c_int, ct.c_double, ct.c_void_p]),
    "ytf_polar_nodes": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p,
                                   ct.c_double, ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "ytf_segment_lateral_sum": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                           ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "ytf_closed_form_G_2d": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int,
                                        ct.c_void_p, ct.POINTER(ct.c_double), ct.c_int, ct.POINTER(ct.c_double),
                                        ct.c_int, ct.c_int, ct.c_double, ct.c_void_p]),
    "ytf_operator_build": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_int, ct.c_int64,
                                      ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(YtfGrid),
                                      ct.POINTER(ct.c_double), ct.c_int, ct.c_int, ct.POINTER(ct.c_int32),
                                      ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double),
                                      ct.POINTER(YtfOperatorOptions), ct.POINTER(ct.c_uint32),
                                      ct.POINTER(ct.c_void_p)]),
    "ytf_operator_destroy": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_operator_apply": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_void_p)]),
    "ytf_operator_num_blocks": (ct.c_int, [ct.c_void_p]),
    "ytf_operator_block_info": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.POINTER(YtfBlockInfo)]),
    "ytf_operator_ghat": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]),
    "ytf_operator_skeleton": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int, ct.POINTER(ct.c_int32),
                                         ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]),
    "ytf_operator_accumulate": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "ytf_operator_apply_host": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_void_p)]),
    "ytf_z_skeletons": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_double, ct.c_int, ct.c_int64,
                                   ct.c_void_p, ct.c_void_p, ct.POINTER(ct.c_double), ct.c_int, ct.c_int,
                                   ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double),
                                   ct.c_double, ct.POINTER(ct.c_void_p)]),
    "ytf_z_skeletons_q": (ct.c_int, [ct.c_void_p, ct.POINTER(YtfLayer), ct.c_int, ct.c_int, ct.c_int64, ct.c_void_p,
                                     ct.POINTER(ct.c_double), ct.c_int, ct.c_int, ct.POINTER(ct.c_int32),
                                     ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double),
                                     ct.c_double, ct.POINTER(ct.c_void_p)]),
    "ytf_skeletons_get": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int, ct.POINTER(ct.c_int32),
                                     ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double),
                                     ct.POINTER(ct.c_double)]),
    "ytf_skeletons_destroy": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_buffer_alloc": (ct.c_int, [ct.c_void_p, ct.c_size_t, ct.POINTER(ct.c_void_p)]),
    "ytf_buffer_free": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "ytf_buffer_upload": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_size_t]),
    "ytf_buffer_download": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_size_t]),
}

TERMS = {"all": 0, "direct": 1, "scattered": 2, "toeplitz": 3, "hankel": 4}
SYNTH_MODES = {"fp64": 0, "f16x3": 1}
LATERAL = {"direct": 0, "nufft": 1}
STRUCTURE = {"auto": 0, "structured": 1, "dense": 2}
BLOCK_KINDS = {0: "structured", 1: "toeplitz_dense", 2: "dense", 3: "skeleton", 4: "toeplitz_skeleton"}
WARN_STRUCTURE_CHECK = 1

EXPORTS = tuple(_SIGS)


def lib():
    """The loaded libytf.so; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is not built: run `python -m yellowthinfilms_amd.build` "
                                   "(there is no CPU fallback)")
            L = ct.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def check(rc, ctx=None):
    if rc != YTF_OK:
        msg = lib().ytf_last_error(ctx)
        raise YtfError(rc, msg.decode() if msg else f"ytf error {rc}")


class Context:
    """One ytf_ctx per device; every call is put on torch's current stream so
    that torch-owned buffers and the kernels are ordered on one stream."""

    _per_device = {}

    def __init__(self, device):
        h = ct.c_void_p()
        L = lib()
        rc = L.ytf_ctx_create(int(device), ct.byref(h))
        if rc != YTF_OK:
            raise YtfError(rc, L.ytf_last_error(None).decode())
        self.handle = h
        self.device = int(device)

    @classmethod
    def get(cls, device=0):
        if device not in cls._per_device:
            cls._per_device[device] = Context(device)
        return cls._per_device[device]

    def bind_stream(self):
        import torch
        s = torch.cuda.current_stream(self.device)
        check(lib().ytf_ctx_set_stream(self.handle, ct.c_void_p(s.cuda_stream)), self.handle)

    def set_timing(self, on=True):
        """Bracket every kernel launch of this context with HIP events (clears totals)."""
        check(lib().ytf_ctx_set_timing(self.handle, 1 if on else 0), self.handle)

    def set_timing_only(self, name):
        """Bracket only the launches of kernel `name` with HIP events (clears totals);
        set_timing ends it."""
        check(lib().ytf_ctx_set_timing_only(self.handle, name.encode()), self.handle)

    def kernel_time(self, name):
        """(total ms, launches) of one kernel name since set_timing; synchronizes."""
        ms, n = ct.c_double(), ct.c_int64()
        check(lib().ytf_ctx_kernel_time(self.handle, name.encode(), ct.byref(ms), ct.byref(n)), self.handle)
        return ms.value, n.value
