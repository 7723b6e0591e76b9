"""ctypes binding of ``libwalrus_hip.so`` (C ABI declared in ``include/walrus_hip.h``).

The library is the only compute path: if it is missing, or no HIP device is usable, every
entry point raises ``RuntimeError`` -- there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# WH_LIB may name another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("WH_LIB") or os.path.join(_HERE, "libwalrus_hip.so")

MAX_EDGES = 32
MAX_PHOTONS = 128

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


class WhRequest(ctypes.Structure):
    """Mirror of ``wh_request`` in ``include/walrus_hip.h``."""

    _fields_ = [
        ("a_off", ctypes.c_int64),
        ("d_off", ctypes.c_int64),
        ("v_off", ctypes.c_int64),
        ("oddloop", ctypes.c_double * 2),
        ("E", ctypes.c_int32),
        ("glynn", ctypes.c_int32),
        ("real", ctypes.c_int32),
        ("options", ctypes.c_int32),
        ("edge_reps", ctypes.c_int32 * MAX_EDGES),
    ]


EXPORTS = [
    "wh_last_error",
    "wh_version",
    "wh_init",
    "wh_calc_hafnian",
    "wh_calc_loop_hafnian",
    "wh_batch_create",
    "wh_batch_info",
    "wh_batch_write_jobs",
    "wh_batch_launch",
    "wh_batch_launch_noalloc",
    "wh_batch_submit",
    "wh_request_partials",
    "wh_server_create",
    "wh_server_submit",
    "wh_server_submit_matrix",
    "wh_prepare_allones",
    "wh_server_wait",
    "wh_server_cancel",
    "wh_server_stats",
    "wh_server_shutdown",
    "wh_server_destroy",
    "wh_batch_chunks",
    "wh_batch_destroy",
    "wh_plan_create",
    "wh_plan_run",
    "wh_plan_info",
    "wh_plan_partials",
    "wh_plan_chunks",
    "wh_plan_destroy",
    "wh_plan_graph_create",
    "wh_graph_launch",
    "wh_graph_destroy",
    "wh_tune_set",
    "wh_tune_get",
    "wh_tune_set_waves",
    "wh_tune_get_waves",
    "wh_tune_set_tile",
    "wh_tune_get_tile",
    "wh_tune_persistent",
    "wh_debug_fail_launches",
    "wh_loop_hafnian_batch",
    "wh_loop_hafnian_batch_gamma",
    "wh_montrealer",
    "wh_lmontrealer",
    "wh_hermite_multidimensional",
]

_lib = None
_lock = threading.Lock()


def _one_hip_runtime():
    """Load torch (when installed) before the library.  torch bundles its own HIP runtime under
    the same soname (libamdhip64.so.7) as the system one the library links: whichever is loaded
    first serves the whole process, and torch's GPU initialisation fails on the system runtime
    while the library works on either.  Importing torch first makes its runtime the one."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path=None):
    """Load the shared library and declare its signatures (no device call is made)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        _one_hip_runtime()
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                "libwalrus_hip.so not found at %s: build it with `python -m thewalrus_amd._build`"
                " (there is no CPU fallback)" % p
            )
        L = ctypes.CDLL(p)
        L.wh_last_error.restype = ctypes.c_char_p
        L.wh_last_error.argtypes = []
        L.wh_version.restype = ctypes.c_int
        L.wh_init.argtypes = [ctypes.c_int]
        L.wh_calc_hafnian.argtypes = [_dp, ctypes.c_int64, _ip, ctypes.c_int64, ctypes.c_int, _dp]
        L.wh_calc_loop_hafnian.argtypes = [_dp, _dp, _ip, ctypes.c_int64, _dp, _dp, ctypes.c_int, _dp]
        L.wh_plan_create.argtypes = [ctypes.POINTER(WhRequest), ctypes.c_int64, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.wh_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.wh_plan_info.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        L.wh_plan_destroy.argtypes = [ctypes.c_void_p]
        L.wh_plan_graph_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_void_p)]
        L.wh_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.wh_graph_destroy.argtypes = [ctypes.c_void_p]
        L.wh_tune_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64]
        L.wh_tune_get.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.wh_tune_get.restype = ctypes.c_int64
        L.wh_tune_persistent.argtypes = [ctypes.c_int]
        L.wh_tune_set_waves.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.wh_tune_set_tile.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.wh_debug_fail_launches.argtypes = [ctypes.c_int]
        L.wh_tune_get_waves.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.wh_tune_get_tile.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.wh_batch_create.argtypes = [ctypes.POINTER(WhRequest), ctypes.c_int64, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_void_p)]
        L.wh_batch_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), _ip, _ip, _dp]
        L.wh_batch_write_jobs.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.wh_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.wh_batch_launch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]
        L.wh_batch_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                      ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p]
        L.wh_request_partials.argtypes = [ctypes.POINTER(WhRequest), ctypes.c_int64, _ip]
        L.wh_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]
        L.wh_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double]
        L.wh_server_wait.restype = ctypes.c_int64
        L.wh_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p,
                                      ctypes.c_int64]
        L.wh_prepare_allones.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, ctypes.POINTER(WhRequest),
                                         ctypes.c_void_p, ctypes.c_int64]
        L.wh_prepare_allones.restype = ctypes.c_int64
        L.wh_server_shutdown.argtypes = [ctypes.c_void_p]
        L.wh_server_destroy.argtypes = [ctypes.c_void_p]
        L.wh_batch_chunks.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        L.wh_batch_destroy.argtypes = [ctypes.c_void_p]
        L.wh_plan_partials.argtypes = [ctypes.c_void_p, _dp]
        L.wh_plan_chunks.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        L.wh_loop_hafnian_batch.argtypes = [_dp, _dp, ctypes.c_int64, _ip, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64]
        L.wh_hermite_multidimensional.argtypes = [_dp, _dp, ctypes.c_int64, _ip, _dp, ctypes.c_int, _dp]
        L.wh_montrealer.argtypes = [_dp, ctypes.c_int64, _dp]
        L.wh_lmontrealer.argtypes = [_dp, _dp, ctypes.c_int64, _dp]
        L.wh_loop_hafnian_batch_gamma.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, _ip, ctypes.c_int64,
                                                  ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _dp,
                                                  ctypes.c_int64]
        for name in EXPORTS:
            getattr(L, name).restype = getattr(L, name).restype or ctypes.c_int
        # the autotuner's chunk table (tuning.json next to the package, or $WH_TUNING)
        from . import tune

        tune.load(lib=L)
        if path is None:
            _lib = L
        return L


_host_lib = None


def load_host():
    """The library's short host-only calls (layout, and the serving engine's queue operations:
    no HIP call, microseconds each) through a handle that keeps the interpreter lock across the
    call: a thread that makes them once per request then does not hand the lock over at every
    request."""
    global _host_lib
    load()
    with _lock:
        if _host_lib is None:
            H = ctypes.PyDLL(LIB_PATH)
            H.wh_request_partials.argtypes = [ctypes.POINTER(WhRequest), ctypes.c_int64, _ip]
            H.wh_batch_create.argtypes = [ctypes.POINTER(WhRequest), ctypes.c_int64, ctypes.c_int64,
                                          ctypes.POINTER(ctypes.c_void_p)]
            H.wh_batch_destroy.argtypes = [ctypes.c_void_p]
            H.wh_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(WhRequest), ctypes.c_void_p,
                                           ctypes.c_int64, ctypes.c_uint64, ctypes.c_double]
            H.wh_server_cancel.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
            H.wh_server_submit_matrix.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_uint64, ctypes.c_double]
            for name in ("wh_request_partials", "wh_batch_create", "wh_batch_destroy", "wh_server_submit",
                         "wh_server_cancel"):
                getattr(H, name).restype = ctypes.c_int
            H.wh_server_submit_matrix.restype = ctypes.c_int64
            _host_lib = H
        return _host_lib


def check(rc):
    """Raise the library's error message for a nonzero return code."""
    if rc != 0:
        msg = _lib.wh_last_error().decode() if _lib is not None else "unknown error"
        if rc == -1:
            raise ValueError(msg)
        raise RuntimeError("libwalrus_hip error %d: %s" % (rc, msg))


def _dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def calc_hafnian(Ax, edge_reps, glynn=True):
    """Device sieve at the reference seam ``_calc_hafnian`` (``_hafnian.py:417``)."""
    L = load()
    Ax = np.ascontiguousarray(Ax, dtype=np.complex128)
    e = np.ascontiguousarray(edge_reps, dtype=np.int64)
    out = np.zeros(2, dtype=np.float64)
    check(L.wh_calc_hafnian(_dptr(Ax), Ax.shape[0], e.ctypes.data_as(_ip), e.shape[0],
                            int(bool(glynn)), _dptr(out)))
    return np.complex128(complex(out[0], out[1]))


def calc_loop_hafnian(Ax, Dx, edge_reps, oddloop=None, oddV=None, glynn=True):
    """Device sieve at the reference seam ``_calc_loop_hafnian`` (``_hafnian.py:513``)."""
    L = load()
    Ax = np.ascontiguousarray(Ax, dtype=np.complex128)
    Dx = np.ascontiguousarray(Dx, dtype=np.complex128)
    e = np.ascontiguousarray(edge_reps, dtype=np.int64)
    ol = None if oddloop is None else np.array([complex(oddloop)], dtype=np.complex128)
    oV = None if oddV is None else np.ascontiguousarray(oddV, dtype=np.complex128)
    out = np.zeros(2, dtype=np.float64)
    check(L.wh_calc_loop_hafnian(_dptr(Ax), _dptr(Dx), e.ctypes.data_as(_ip), e.shape[0], _dptr(ol),
                                 _dptr(oV), int(bool(glynn)), _dptr(out)))
    return np.complex128(complex(out[0], out[1]))


def loop_hafnian_batch(Ax, Dx, fixed_edge_reps, batch_max, cutoff, odd_variant, glynn=True):
    """Device sieve at the reference seams ``_calc_loop_hafnian_batch_even`` /
    ``_calc_loop_hafnian_batch_odd`` (``loop_hafnian_batch.py:52``, ``:128``); with a 2-D ``Dx``
    (one diagonal per row) ``_calc_loop_hafnian_batch_gamma_even`` / ``_odd``
    (``loop_hafnian_batch_gamma.py:53``, ``:132``), returning one row per diagonal.

    ``batch_max < 0`` (the odd variant at ``N_cutoff = 0``) gives the reference's empty term
    range: zeros."""
    L = load()
    Ax = np.ascontiguousarray(Ax, dtype=np.complex128)
    Dx = np.ascontiguousarray(Dx, dtype=np.complex128)
    gamma = Dx.ndim == 2
    e = np.ascontiguousarray(fixed_edge_reps, dtype=np.int64)
    n_out = 2 * int(batch_max) + int(cutoff) + 1 + (1 if odd_variant else 0)
    nD = Dx.shape[0] if gamma else 1
    if batch_max < 0:
        z = np.zeros(max(n_out, 0), dtype=np.complex128)
        return np.zeros((nD, z.size), dtype=np.complex128) if gamma else z
    out = np.zeros(nD * n_out, dtype=np.complex128)
    args = (e.ctypes.data_as(_ip), e.shape[0], int(batch_max), int(cutoff), int(bool(odd_variant)),
            int(bool(glynn)), out.ctypes.data_as(_dp), out.shape[0])
    if gamma:
        rc = L.wh_loop_hafnian_batch_gamma(_dptr(Ax), _dptr(Dx), nD, Ax.shape[0], *args)
    else:
        rc = L.wh_loop_hafnian_batch(_dptr(Ax), _dptr(Dx), Ax.shape[0], *args)
    if rc < 0:
        check(rc)
    return out.reshape(nD, n_out) if gamma else out


def montrealer(Sigma, zeta=None):
    """Device sieve at the reference seams ``montrealer(Sigma)`` / ``lmontrealer(Sigma, zeta)``
    (``_montrealer.py:37``, ``:79``)."""
    L = load()
    S = np.ascontiguousarray(Sigma, dtype=np.complex128)
    out = np.zeros(2, dtype=np.float64)
    if zeta is None:
        check(L.wh_montrealer(_dptr(S), S.shape[0], _dptr(out)))
    else:
        z = np.ascontiguousarray(zeta, dtype=np.complex128)
        check(L.wh_lmontrealer(_dptr(S), _dptr(z), S.shape[0], _dptr(out)))
    return np.complex128(complex(out[0], out[1]))


def hermite_recursion(R, y, cutoffs, C=1.0, renorm=False):
    """Device recursion at the reference seams ``_hermite_multidimensional(R, y, G)`` /
    ``_hermite_multidimensional_renorm`` (``_hermite_multidimensional.py:279``, ``:251``) with
    ``G[0] = C``: the full complex128 tensor of shape ``cutoffs``."""
    L = load()
    R = np.ascontiguousarray(R, dtype=np.complex128)
    y = np.ascontiguousarray(y, dtype=np.complex128)
    cut = np.ascontiguousarray(cutoffs, dtype=np.int64)
    out = np.empty(tuple(int(c) for c in cut), dtype=np.complex128)
    c = np.array([complex(C)], dtype=np.complex128)
    check(L.wh_hermite_multidimensional(_dptr(R), _dptr(y), R.shape[0], cut.ctypes.data_as(_ip), _dptr(c),
                                        int(bool(renorm)), out.ctypes.data_as(_dp)))
    return out
