"""GPU parity against the committed golden fixtures (tests/golden/*.npz, make_golden.py).

Bar: relative error <= 1e-10 against the double-double truth.  The Glynn sieve sums 2^(h-1)
alternating terms, so on some inputs the reference algorithm in FP64 (the oracle's faithful
restatement, stored as ``*_ref64``) is itself further than 1e-10 from the truth; there the
device must be no further from the truth than twice the reference algorithm's own error.
"""
import os

import numpy as np
import pytest

from tests.conftest import relerr

import thewalrus_amd as tw
from thewalrus_amd import batch as B

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-10


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated" % name)
    return np.load(path, allow_pickle=False)


def _bar(truth, ref64):
    return max(TOL, 2.0 * relerr(ref64, truth))


def test_config1_n20_complex():
    g = _load("golden_small.npz")
    got = tw.hafnian(g["c1_A"])
    assert relerr(got, g["c1_truth"]) < _bar(g["c1_truth"], g["c1_ref64"])


def test_config2_real_24_batch():
    g = _load("golden_small.npz")
    k = int(g["c2_count"])
    mats = [g["c2_A_%d" % i] for i in range(k)]
    res = B.hafnian_batch(mats)
    for i in range(k):
        t = g["c2_truth_%d" % i]
        assert relerr(res[i], t) < _bar(t, g["c2_ref64_%d" % i]), i
    assert round(res[k - 1].real) == 316234143225


def test_config4_repeated_batch():
    g = _load("golden_small.npz")
    k = int(g["c4_count"])
    preps = [
        B.prepare_repeated(g["c4_A_%d" % i], list(g["c4_rpt_%d" % i]), mu=g["c4_mu_%d" % i], loop=True)
        for i in range(k)
    ]
    b = B.DeviceBatch(preps)
    b.run()
    res = b.results()
    for i in range(k):
        t = g["c4_truth_%d" % i]
        assert relerr(res[i], t) < _bar(t, g["c4_ref64_%d" % i]), i
        alone = tw.hafnian_repeated(g["c4_A_%d" % i], list(g["c4_rpt_%d" % i]), mu=g["c4_mu_%d" % i], loop=True)
        assert complex(alone) == complex(res[i])


def test_config5_stream_sample():
    g = _load("golden_small.npz")
    k = int(g["c5_count"])
    mats = [g["c5_A_%d" % i] for i in range(k)]
    loops = [bool(g["c5_loop_%d" % i]) for i in range(k)]
    res = B.hafnian_batch(mats, loop=loops)
    for i in range(k):
        t = g["c5_truth_%d" % i]
        assert relerr(res[i], t) < _bar(t, g["c5_ref64_%d" % i]), (i, mats[i].shape[0], loops[i])


def test_config3_n40_loop():
    g = _load("golden_n40.npz")
    got = tw.hafnian(g["c3_A"], loop=True)
    t = g["c3_truth"]
    err, ref_err = relerr(got, t), relerr(g["c3_ref64"], t)
    print("config3 n=40 loop: device err %.3e, reference-algorithm FP64 err %.3e" % (err, ref_err))
    assert err < _bar(t, g["c3_ref64"])


@pytest.mark.parametrize("kind", ["plain", "loop"])
def test_structured_n40(kind):
    g = _load("golden_n40.npz")
    A = g["s40_%s_A" % kind]
    got = tw.hafnian(A, loop=kind == "loop")
    t = g["s40_%s_truth" % kind]
    err, ref_err = relerr(got, t), relerr(g["s40_%s_ref64" % kind], t)
    print("structured n=40 %s: device err %.3e, reference-algorithm FP64 err %.3e" % (kind, err, ref_err))
    assert err < _bar(t, g["s40_%s_ref64" % kind])


def _band_keys():
    path = os.path.join(HERE, "golden_band.npz")
    if not os.path.exists(path):
        return []
    with np.load(path, allow_pickle=False) as g:
        return sorted(k[:-2] for k in g.files if k.endswith("_A"))


@pytest.mark.parametrize("key", _band_keys() or ["missing"])
def test_band_n30_to_48(key):
    """Tile classes T = 4..6 (n = 26..48) against truth: random n = 30..38 (complex, and real
    n = 30, 34) with double-double truth, structured direct sums n = 32, 36, 44 (complex) and
    42 (real) with exact truth; per call and inside one batch with the rest of its class."""
    g = _load("golden_band.npz")
    if key == "missing":
        pytest.fail("golden_band.npz has no cases")
    A, t, r64 = g[key + "_A"], g[key + "_truth"], g[key + "_ref64"]
    loop = "_loop_" in key
    got = tw.hafnian(A, loop=loop)
    err = relerr(got, t)
    print("%s: device err %.3e, reference-algorithm FP64 err %.3e" % (key, err, relerr(r64, t)), flush=True)
    assert err < _bar(t, r64)


def test_band_batched_equals_per_call():
    """The band's requests in one mixed batch: the same bits as per call."""
    g = _load("golden_band.npz")
    keys = _band_keys()
    mats = [g[k + "_A"] for k in keys]
    loops = [("_loop_" in k) for k in keys]
    res = B.hafnian_batch(mats, loop=loops)
    for k, A, lp, r in zip(keys, mats, loops, res):
        assert complex(tw.hafnian(A, loop=lp)) == complex(r), k
