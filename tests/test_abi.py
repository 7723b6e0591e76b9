"""The C ABI library (CPU only): it loads, exports every function declared in include/*.h,
and its request struct matches the ctypes mirror.  No device call is made."""
import ctypes
import glob
import os
import re

import pytest

from thewalrus_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(wh_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_built():
    assert os.path.exists(_native.LIB_PATH), "run python -m thewalrus_amd._build"


def test_exports_every_declared_symbol():
    L = _native.load()
    names = declared_functions()
    assert names >= set(_native.EXPORTS)
    for name in sorted(names):
        assert hasattr(L, name), name


def test_request_layout():
    """wh_request: 3 int64 offsets, 2 doubles, 4 int32, 32 int32 multiplicities = 184 bytes."""
    R = _native.WhRequest
    assert ctypes.sizeof(R) == 184
    assert R.oddloop.offset == 24 and R.E.offset == 40 and R.edge_reps.offset == 56


def test_error_string_without_device():
    L = _native.load()
    assert isinstance(L.wh_last_error(), bytes)
    assert L.wh_version() >= 100


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _native.load(str(tmp_path / "nope.so"))


def test_request_partials_match_single_batch_layout():
    """wh_request_partials (one call per served request) equals the partial slots of a batch of
    that request alone; both are host-only layout calls."""
    import numpy as np

    from thewalrus_amd import _hafnian as H
    from thewalrus_amd import batch as B

    L = _native.load()
    rng = np.random.default_rng(7)
    jobs = []
    for n, loop in ((6, False), (9, True), (20, True), (24, False), (40, True)):
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        jobs.append(H.prepare_hafnian((X + X.T) / 2, loop=loop))
    X = rng.normal(size=(6, 6))
    jobs.append(H.prepare_hafnian_repeated((X + X.T) / 2, [3, 1, 2, 0, 1, 2], mu=np.arange(6) + 0.5, loop=True))
    for job in jobs:
        pk = B.pack_one(job)
        for chunk in (0, 7):
            got = ctypes.c_int64()
            assert L.wh_request_partials(ctypes.byref(pk.req), chunk, ctypes.byref(got)) == 0
            b = ctypes.c_void_p()
            assert L.wh_batch_create(ctypes.byref(pk.req), 1, chunk, ctypes.byref(b)) == 0
            want = ctypes.c_int64()
            assert L.wh_batch_info(b, None, ctypes.byref(want), None, None) == 0
            L.wh_batch_destroy(b)
            assert got.value == want.value > 0


def test_batch_submit_checks_arenas_before_any_device_call():
    """A batch larger than the caller's arenas is refused with a message, before any copy or
    launch (so this runs without a device)."""
    import numpy as np

    from thewalrus_amd import _hafnian as H
    from thewalrus_amd import batch as B

    L = _native.load()
    X = np.random.default_rng(8).normal(size=(12, 12))
    pk = B.pack_one(H.prepare_hafnian((X + X.T) / 2))
    b = ctypes.c_void_p()
    assert L.wh_batch_create(ctypes.byref(pk.req), 1, 0, ctypes.byref(b)) == 0
    try:
        fake = 1 << 12  # aligned, never dereferenced: the capacity checks fail first
        rc = L.wh_batch_submit(b, fake, fake, 0, fake, fake, pk.elems, fake, 1 << 20, fake, fake, 1, None)
        assert rc != 0 and b"arenas" in L.wh_last_error()
        rc = L.wh_batch_submit(b, fake, fake, 1 << 20, fake, fake, pk.elems - 1, fake, 1 << 20, fake, fake, 1, None)
        assert rc != 0 and b"pool" in L.wh_last_error()
        rc = L.wh_batch_submit(b, None, fake, 1 << 20, fake, fake, pk.elems, fake, 1 << 20, fake, fake, 1, None)
        assert rc != 0
    finally:
        L.wh_batch_destroy(b)


def _prep_native(L, A, loop, glynn=1, rtol=1e-05, atol=1e-08):
    import numpy as np

    n = A.shape[0]
    req = _native.WhRequest()
    payload = np.zeros(n * n + n, dtype=np.complex128)
    k = L.wh_prepare_allones(A.ctypes.data, n, int(np.iscomplexobj(A)), int(loop), glynn, rtol, atol,
                             ctypes.byref(req), payload.ctypes.data, payload.shape[0])
    return k, req, payload[:k] if k > 0 else None


def test_native_preparation_equals_host_path():
    """wh_prepare_allones (the serving engine's one-pass preparation) writes exactly the request
    and payload that prepare_hafnian + pack_one produce, for every size and kind it accepts."""
    import numpy as np

    from thewalrus_amd import _hafnian as H
    from thewalrus_amd import batch as B

    L = _native.load()
    rng = np.random.default_rng(11)
    for n in range(6, 65, 2):
        for kind in ("complex", "real", "real-as-complex"):
            X = rng.normal(size=(n, n)) + (1j * rng.normal(size=(n, n)) if kind == "complex" else 0)
            A = (X + X.T) / 2
            if kind == "real-as-complex":
                A = A.astype(np.complex128)
            for loop in (False, True):
                for method in ("glynn", "inclexcl"):
                    k, req, payload = _prep_native(L, A, loop, glynn=int(method == "glynn"))
                    want = B.pack_one(H.prepare_hafnian(A, loop=loop, method=method))
                    assert k == want.elems, (n, kind, loop)
                    assert np.array_equal(payload.view(np.float64), want.payload.view(np.float64))
                    assert bytes(req) == bytes(want.req), (n, kind, loop, method)


def test_native_preparation_declines_other_requests():
    """Everything outside the one-pass case goes back to the host path (return 0)."""
    import numpy as np

    L = _native.load()
    rng = np.random.default_rng(12)
    X = rng.normal(size=(8, 8))
    S = (X + X.T) / 2
    near = S.copy()
    near[2, 5] += 1e-12  # symmetric within tolerance, not exactly
    nan = S.copy()
    nan[3, 3] = np.nan
    diag = np.diag(rng.normal(size=8)) + 1e-9 * np.ones((8, 8))
    for A, loop in ((near, False), (nan, True), (diag, False), (S[:4, :4].copy(), False), (S[:7, :7].copy(), True),
                    (np.zeros((66, 66)), False)):
        k, _, _ = _prep_native(L, np.ascontiguousarray(A), loop)
        assert k == 0


def test_layout_refuses_chunk_counts_beyond_the_launch_limit():
    """A repeated-row request with about 6e11 terms needs more than 2^31 - 1 general-kernel
    chunks of at most 256 terms: the layout fails with WH_EUNSUPPORTED instead of truncating the
    chunk count (host only)."""
    L = _native.load()
    r = _native.WhRequest()
    r.E = 24
    r.glynn = 1
    r.a_off = 0
    r.d_off = r.v_off = -1
    reps = [2] * 20 + [3] * 4  # 2 * 3^19 * 4^4 = 5.95e11 Glynn terms, N = 104 photons
    for i, e in enumerate(reps):
        r.edge_reps[i] = e
    b = ctypes.c_void_p()
    rc = L.wh_batch_create(ctypes.byref(r), 1, 0, ctypes.byref(b))
    assert rc == -4 and b"launch limit" in L.wh_last_error()
    slots = ctypes.c_int64()
    assert L.wh_request_partials(ctypes.byref(r), 0, ctypes.byref(slots)) == -4


def test_persistent_mode_round_trip():
    L = _native.load()
    prev = L.wh_tune_persistent(-1)
    try:
        assert prev in (0, 1, 2)
        assert L.wh_tune_persistent(0) == prev
        assert L.wh_tune_persistent(-1) == 0
        assert L.wh_tune_persistent(3) == -1
    finally:
        L.wh_tune_persistent(prev)
