"""GPU tests of the batched path: kernel cross-checks, batch = per-call, large structured cases.

Tolerance: relative error <= 1e-10 against double-double truth; bitwise equality where the
same request runs through the same kernel and chunking (alone, or inside a batch).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import random_symmetric, relerr

import thewalrus_amd as tw
from thewalrus_amd import batch as B

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _mixed(rng, sizes, loops):
    out = []
    for n, lp in zip(sizes, loops):
        out.append((random_symmetric(rng, n, scale=1.0 / np.sqrt(2 * n)), lp))
    return out


@pytest.mark.parametrize("loop", [False, True])
@pytest.mark.parametrize("real", [False, True])
def test_pm1_matches_generic_and_truth(loop, real):
    rng = np.random.default_rng(11 + 2 * loop + real)
    sizes = [6, 8, 10, 12, 14, 16, 18, 20]
    reqs = [(random_symmetric(rng, n, complex_=not real, scale=1 / np.sqrt(n)), loop) for n in sizes]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    fast = B.DeviceBatch(preps)
    gen = B.DeviceBatch(preps, force_generic=True)
    fast.run()
    gen.run()
    rf, rg = fast.results(), gen.results()
    for (A, l), a, g in zip(reqs, rf, rg):
        truth = O.hafnian_sieve(A, loop=l, precision="dd")
        assert relerr(a, truth) < TOL, (A.shape[0], a, truth)
        assert relerr(g, truth) < TOL, (A.shape[0], g, truth)


def test_batch_equals_per_call_bitwise():
    """Each result is the same alone (per call) and inside a mixed batch (same bits)."""
    rng = np.random.default_rng(5)
    sizes = [8, 24, 10, 22, 12, 20, 14, 18, 16, 26, 28]
    reqs = _mixed(rng, sizes, [i % 2 == 0 for i in range(len(sizes))])
    reqs.append((np.ones((24, 24)), False))  # real input, specialised real kernel
    res = B.hafnian_batch([A for A, _ in reqs], loop=[l for _, l in reqs])
    for (A, l), r in zip(reqs, res):
        alone = tw.hafnian(A, loop=l)
        assert complex(alone) == complex(r), (A.shape[0], l, alone, r)
    # the same requests in a different order and mix give the same bits
    perm = rng.permutation(len(reqs))
    res2 = B.hafnian_batch([reqs[i][0] for i in perm], loop=[reqs[i][1] for i in perm])
    for k, i in enumerate(perm):
        assert complex(res2[k]) == complex(res[i])


def test_all_ones_24_in_batch():
    """16 all-ones 24x24 inside a batch of random ones: 23!! = 316234143225 exactly."""
    rng = np.random.default_rng(138)
    M = rng.uniform(-1, 1, size=(48, 24, 24))
    mats = [(m + m.T) / 2 for m in M] + [np.ones((24, 24))] * 16
    res = B.hafnian_batch(mats)
    for r in res[-16:]:
        assert abs(r.real - 316234143225) < 0.5 and round(r.real) == 316234143225
    for A, r in list(zip(mats, res))[:4]:
        truth = O.hafnian_sieve(A, precision="dd")
        assert relerr(r, truth) < TOL


def _structured(n_half, loop, seed):
    """A = W P (A1 (+) A2) P^T W with known (loop) hafnian lhaf(A1) lhaf(A2) prod(w).

    Each vertex is covered once, by an edge (weight w_i w_j) or by its loop, so the loop
    weights scale by w_i (not w_i^2)."""
    rng = np.random.default_rng(seed)
    A1 = random_symmetric(rng, n_half, scale=1 / np.sqrt(2 * n_half))
    A2 = random_symmetric(rng, n_half, scale=1 / np.sqrt(2 * n_half))
    if loop:
        for X in (A1, A2):
            X[np.diag_indices(n_half)] = (rng.normal(size=n_half) + 1j * rng.normal(size=n_half)) / np.sqrt(2)
    n = 2 * n_half
    A = np.zeros((n, n), dtype=np.complex128)
    A[:n_half, :n_half] = A1
    A[n_half:, n_half:] = A2
    perm = rng.permutation(n)
    A = A[np.ix_(perm, perm)]
    w = np.exp(1j * rng.uniform(0, 2 * np.pi, size=n)) * rng.uniform(0.8, 1.25, size=n)
    d = np.diag(A).copy()
    A = (w[:, None] * A) * w[None, :]
    A[np.diag_indices(n)] = w * d
    t1 = O.hafnian_sieve(A1, loop=loop, precision="dd")
    t2 = O.hafnian_sieve(A2, loop=loop, precision="dd")
    return A, t1 * t2 * np.prod(w)


def test_structured_loop_batch():
    """Structured loop matrices with exact truth, batched; bar: 1e-10 or twice the reference
    algorithm's own FP64 error on the same input (the oracle's faithful restatement)."""
    mats, truths = [], []
    for s in range(4):
        A, t = _structured(12, True, 100 + s)
        mats.append(A)
        truths.append(t)
    res = B.hafnian_batch(mats, loop=True)
    for A, r, t in zip(mats, res, truths):
        ref64 = O.hafnian_sieve(A, loop=True, precision="double")
        assert relerr(r, t) < max(TOL, 2 * relerr(ref64, t)), (relerr(r, t), relerr(ref64, t))


@pytest.mark.parametrize("loop", [False, True])
@pytest.mark.parametrize("real", [False, True])
def test_half_and_full_products(loop, real):
    """Both formulations of the all-ones kernel against truth, n = 8 .. 28 (tile sides 1..4);
    bar: 1e-10 or twice the reference algorithm's own FP64 error on the input."""
    rng = np.random.default_rng(21 + 2 * loop + real)
    sizes = [8, 10, 14, 16, 18, 22, 24, 26, 28]
    reqs = [(random_symmetric(rng, n, complex_=not real, scale=1 / np.sqrt(n)), loop) for n in sizes]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    half = B.DeviceBatch(preps)
    full = B.DeviceBatch(preps, full_products=True)
    half.run()
    full.run()
    rh, rf = half.results(), full.results()
    assert half.flops < 0.75 * full.flops
    for (A, l), a, f in zip(reqs, rh, rf):
        truth = O.hafnian_sieve(A, loop=l, precision="dd")
        ref64 = O.hafnian_sieve(A, loop=l, precision="double")
        bar = max(TOL, 2 * relerr(ref64, truth))
        assert relerr(a, truth) < bar, ("half", A.shape[0], relerr(a, truth), bar)
        assert relerr(f, truth) < bar, ("full", A.shape[0], relerr(f, truth), bar)


@pytest.mark.parametrize("loop", [False, True])
def test_matrix_cores_and_vector_units(loop):
    """The complex all-ones kernel on the FP64 matrix cores with symmetric powers (default), with
    full powers, and on the vector units, tile sides 1..6 (n = 8 .. 48), against truth; bar as
    above."""
    rng = np.random.default_rng(61 + loop)
    sizes = [4, 6, 8, 12, 16, 20, 24, 30, 32, 36, 40, 44, 48] if not loop else [6, 8, 14, 16, 22, 24, 30, 32, 38, 40, 46]
    reqs = [(random_symmetric(rng, n, scale=1 / np.sqrt(n)), loop) for n in sizes]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    runs = {"symmetric powers": B.DeviceBatch(preps), "full powers": B.DeviceBatch(preps, full_powers=True),
            "vector units": B.DeviceBatch(preps, valu_products=True)}
    res = {}
    for name, db in runs.items():
        print("all-ones kernel: %s, loop=%s" % (name, loop), flush=True)
        db.run()
        res[name] = db.results()
    for k, (A, l) in enumerate(reqs):
        n = A.shape[0]
        if n <= 24:
            truth = O.hafnian_sieve(A, loop=l, precision="dd")
            ref64 = O.hafnian_sieve(A, loop=l, precision="double")
            bar = max(TOL, 2 * relerr(ref64, truth))
            for name in runs:
                assert relerr(res[name][k], truth) < bar, (name, n, relerr(res[name][k], truth), bar)
        else:  # beyond the oracle's seconds: the kernels agree to the bar
            v = res["vector units"][k]
            for name in ("symmetric powers", "full powers"):
                assert relerr(res[name][k], v) < 1e-10, (name, n, relerr(res[name][k], v))


def test_graph_replay_equals_eager():
    """A plan captured into a HIP graph replays bit-identically to the eager launch, on every
    replay, with other device work (another plan, torch kernels) between the replays."""
    import torch

    rng = np.random.default_rng(71)
    sizes = [6, 10, 16, 20, 24, 30, 8, 12]
    reqs = [(random_symmetric(rng, n, scale=1 / np.sqrt(n)), i % 2 == 1) for i, n in enumerate(sizes)]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    other = B.DeviceBatch([B.prepare(random_symmetric(rng, 14), loop=True)])
    eager = B.DeviceBatch(preps)
    eager.run()
    ref = eager.results()
    eager.close()
    s = torch.cuda.Stream()
    g = B.DeviceBatch(preps)
    g.capture(s)
    for rep in range(4):
        g.out.zero_()
        torch.cuda.synchronize()
        g.run(s)
        s.synchronize()
        got = g.results()
        assert np.array_equal(got, ref), rep
        other.run()  # eager work between replays
        x = torch.ones(1 << 16, dtype=torch.float64, device="cuda").sum()
        torch.cuda.synchronize()
        assert float(x) == 65536.0
    g.close()
    other.close()


def test_graph_needs_a_stream():
    import torch

    rng = np.random.default_rng(72)
    g = B.DeviceBatch([B.prepare(random_symmetric(rng, 8))])
    with pytest.raises(ValueError):
        g.capture(torch.cuda.default_stream())
    g.close()


def test_tuned_chunking_changes_chunks_not_results():
    """Another workgroups-per-request target changes the chunking of that class only (all-ones
    kernel, T = 4, loop); the results move in their last bits at most."""
    import ctypes

    from thewalrus_amd import _native, tune

    rng = np.random.default_rng(73)
    reqs = [(random_symmetric(rng, 32, scale=1 / np.sqrt(32)), True), (random_symmetric(rng, 28), False)]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    before = tune.current()

    def chunks(db):
        ct = (ctypes.c_int64 * db.nreq)()
        _native.check(_native.load().wh_plan_chunks(db.plan, None, None, ct))
        return list(ct)

    try:
        tune.apply([{"kind": 0, "T": 4, "loop": 1, "blocks_per_request": 512}])
        a = B.DeviceBatch(preps)
        a.run()
        ra, ca = a.results(), chunks(a)
        a.close()
        tune.apply([{"kind": 0, "T": 4, "loop": 1, "blocks_per_request": 2048}])
        b = B.DeviceBatch(preps)
        b.run()
        rb, cb = b.results(), chunks(b)
        b.close()
    finally:
        tune.apply([{"kind": k, "T": T, "loop": lp, "blocks_per_request": v} for (k, T, lp), v in before.items()])
    print("chunk terms", ca, cb, flush=True)
    assert cb[0] < ca[0] and cb[1] == ca[1]
    assert relerr(ra[0], rb[0]) < 1e-12
    assert ra[1] == rb[1]


@pytest.mark.parametrize("loop", [False, True])
def test_odd_chunk_sizes_agree(loop):
    """Chunks of 1, 3, 5 and 13 terms (waves of a workgroup with unequal term counts, partial
    coefficient batches) give the default chunking's results to rounding, on the symmetric
    matrix-core kernel and on the vector kernel."""
    rng = np.random.default_rng(81 + loop)
    reqs = [(random_symmetric(rng, n, scale=1 / np.sqrt(n)), loop) for n in (10, 16, 22, 26)]
    preps = [B.prepare(A, loop=l) for A, l in reqs]
    for kw in ({}, {"valu_products": True}):
        ref = B.DeviceBatch(preps, **kw)
        ref.run()
        r0 = ref.results()
        ref.close()
        for ct in (1, 3, 5, 13):
            db = B.DeviceBatch(preps, chunk_terms=ct, **kw)
            db.run()
            r = db.results()
            db.close()
            for k in range(len(reqs)):
                assert relerr(r[k], r0[k]) < 1e-12, (kw, ct, reqs[k][0].shape[0], relerr(r[k], r0[k]))


@pytest.mark.parametrize("loop", [False, True])
def test_direct_powers_and_charpoly_routes(loop):
    """Repeated-row requests whose term matrices are smaller than h + 1 (so the reference extends
    the power traces by the characteristic polynomial): the general kernel's default (every
    power formed directly) and the reference's route (WH_OPT_CHARPOLY) both against truth; bar:
    1e-10 or twice the reference algorithm's own FP64 error on the input.  Odd photon numbers
    included for loops."""
    rng = np.random.default_rng(77 + loop)
    reqs = []
    for k in range(12):
        n = int(rng.integers(3, 7))
        A = random_symmetric(rng, n, scale=1 / np.sqrt(n))
        rpt = [int(r) for r in rng.integers(1, 6, size=n)]
        if not loop and sum(rpt) % 2:
            rpt[0] += 1
        mu = (rng.normal(size=n) + 1j * rng.normal(size=n)) if loop else None
        reqs.append((A, rpt, mu))
    preps = [B.prepare_repeated(A, rpt, mu=mu, loop=loop) for A, rpt, mu in reqs]
    direct = B.DeviceBatch(preps)
    charpoly = B.DeviceBatch(preps, charpoly=True)
    direct.run()
    charpoly.run()
    rd, rc = direct.results(), charpoly.results()
    assert direct.flops > charpoly.flops  # the direct route forms more (small) products
    for (A, rpt, mu), d, c in zip(reqs, rd, rc):
        truth = O.hafnian_sieve(A, loop=loop, rpt=rpt, mu=mu, precision="dd")
        ref64 = O.hafnian_sieve(A, loop=loop, rpt=rpt, mu=mu, precision="double")
        bar = max(TOL, 2 * relerr(ref64, truth))
        assert relerr(d, truth) < bar, ("direct", rpt, relerr(d, truth), bar)
        assert relerr(c, truth) < bar, ("charpoly", rpt, relerr(c, truth), bar)
