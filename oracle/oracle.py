"""CPU oracle for the power-trace hafnian -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this module, and only as the checker (or as the timed CPU baseline).  The product package
``thewalrus_amd`` never imports it.

It wraps ``libwalrus_oracle.so`` (``walrus_oracle.cpp``, a C++ restatement of
``thewalrus/_hafnian.py:26-631`` and ``thewalrus/charpoly.py:42-327``) and restates the
pure-Python preparation of the reference (pairing, reorder) independently of the product's
frontend, so the product's host code is checked against a second statement of the same rules.

Precision argument: ``"double"`` follows the reference's operation order inside each term;
``"dd"`` runs the same algorithm in double-double arithmetic and serves as high-precision
truth.  ``variant=1`` swaps the per-term product loop for Hessenberg + La Budde + Newton
identities (an independent route, cheaper by ~n/2, used for large-n truth).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load (building first if needed) the oracle shared library."""
    global _LIB
    if _LIB is None:
        # WALRUS_ORACLE_LIB may name another build (the sanitizer build of ``make sanitize``)
        path = os.environ.get("WALRUS_ORACLE_LIB") or os.path.join(_HERE, "libwalrus_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        L.wo_calc_sieve.argtypes = [dp, ctypes.c_int64, ip, ctypes.c_int64, dp, dp, dp,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.wo_powertrace.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, dp]
        L.wo_sieve_terms.argtypes = [dp, ctypes.c_int64, ip, ctypes.c_int64, dp, dp, dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.wo_charpoly.argtypes = [dp, ctypes.c_int64, dp]
        L.wo_fcoeffs.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, dp, dp, dp, dp, dp]
        L.wo_brute_force.argtypes = [dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp]
        L.wo_kan.argtypes = [dp, dp, ip, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp]
        L.wo_hermite_multidimensional.argtypes = [dp, dp, ctypes.c_int64, ip, dp, ctypes.c_int, ctypes.c_int, dp]
        L.wo_montrealer.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, dp]
        L.wo_loop_hafnian_batch.argtypes = [dp, dp, ctypes.c_int64, ip, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        _LIB = L
    return _LIB


def build():
    """Compile the oracle library with its Makefile (g++ -fopenmp)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _unpack(out, precision):
    if precision == 0:
        return complex(out[0], out[1])
    # double-double: keep as (value, low parts) -> python complex of the rounded sum, plus
    # the exact pair for callers that need more digits
    return complex(out[0] + out[1], out[2] + out[3])


def _prec(precision):
    return {"double": 0, "dd": 1, 0: 0, 1: 1}[precision]


def calc_sieve(Ax, edge_reps, D=None, oddloop=None, oddV=None, glynn=True, precision="double",
               nthreads=0, variant=0, exact_pair=False):
    """Sieve at the reference kernel seam (``_hafnian.py:417`` / ``:513``)."""
    p = _prec(precision)
    Ax = _c(Ax)
    e = np.ascontiguousarray(edge_reps, dtype=np.int64)
    Dc = None if D is None else _c(D)
    ol = None if oddloop is None else np.array([complex(oddloop)], dtype=np.complex128)
    oV = None if oddV is None else _c(oddV)
    out = np.zeros(4, dtype=np.float64)
    rc = lib().wo_calc_sieve(_dp(Ax), Ax.shape[0], _ip(e), e.shape[0], _dp(Dc), _dp(ol), _dp(oV),
                             int(bool(glynn)), p, int(nthreads), int(variant), _dp(out))
    if rc != 0:
        raise ValueError("oracle sieve rejected its input")
    if exact_pair and p == 1:
        return out.copy()
    return _unpack(out, p)


def sieve_terms(Ax, edge_reps, steps, D=None, oddloop=None, oddV=None, glynn=True,
                precision="double", nthreads=0):
    """Every signed term (prefactor * f) of the sieve, in term order (before the final scale)."""
    p = _prec(precision)
    Ax = _c(Ax)
    e = np.ascontiguousarray(edge_reps, dtype=np.int64)
    Dc = None if D is None else _c(D)
    ol = None if oddloop is None else np.array([complex(oddloop)], dtype=np.complex128)
    oV = None if oddV is None else _c(oddV)
    out = np.zeros(4, dtype=np.float64)
    terms = np.zeros(2 * steps, dtype=np.float64)
    rc = lib().wo_sieve_terms(_dp(Ax), Ax.shape[0], _ip(e), e.shape[0], _dp(Dc), _dp(ol), _dp(oV),
                              int(bool(glynn)), p, int(nthreads), _dp(out), _dp(terms))
    if rc != 0:
        raise ValueError("oracle sieve rejected its input")
    return terms.view(np.complex128)


def powertrace(H, n, precision="double"):
    """``charpoly.py:302`` powertrace(H, n) -> p[0..n-1]."""
    p = _prec(precision)
    H = _c(H)
    m = H.shape[0]
    out = np.zeros(4 * max(n, 1), dtype=np.float64)
    lib().wo_powertrace(_dp(H), m, n, p, _dp(out))
    if p == 0:
        return out[: 2 * n].view(np.complex128).copy()
    o = out.reshape(-1, 4)[:n]
    return (o[:, 0] + o[:, 1]) + 1j * (o[:, 2] + o[:, 3])


def charpoly(H):
    """``charpoly.py:286`` charpoly(H) -> [c_1..c_m] of det(lambda I - H)."""
    H = _c(H)
    m = H.shape[0]
    out = np.zeros(2 * m, dtype=np.float64)
    lib().wo_charpoly(_dp(H), m, _dp(out))
    return out.view(np.complex128).copy()


def fcoeffs(AX, N, mode=0, XD=None, D=None, oddloop=None, oddVX=None):
    """``f`` (mode 0), ``f_loop`` (1) or ``f_loop_odd`` (2) coefficient vectors."""
    AX = _c(AX)
    m = AX.shape[0]
    out = np.zeros(2 * (N + 2), dtype=np.float64)
    ol = None if oddloop is None else np.array([complex(oddloop)], dtype=np.complex128)
    k = lib().wo_fcoeffs(_dp(AX), m, N, mode, _dp(None if XD is None else _c(XD)),
                         _dp(None if D is None else _c(D)), _dp(ol),
                         _dp(None if oddVX is None else _c(oddVX)), _dp(out))
    return out[: 2 * k].view(np.complex128).copy()


def brute_force(A, loop=False, precision="dd"):
    """(Loop) hafnian by definition: sum over perfect matchings (with loops)."""
    A = _c(A)
    out = np.zeros(4, dtype=np.float64)
    p = _prec(precision)
    if lib().wo_brute_force(_dp(A), A.shape[0], int(bool(loop)), p, _dp(out)) != 0:
        raise ValueError("brute force limited to n <= 30")
    return _unpack(out, p)


def kan(B, rpt, mu=None, loop=False, precision="dd"):
    """Kan's repeated-moment formula (``docs/algorithms.rst:127-148``)."""
    B = _c(B)
    r = np.ascontiguousarray(rpt, dtype=np.int64)
    u = None if mu is None else _c(mu)
    out = np.zeros(4, dtype=np.float64)
    p = _prec(precision)
    lib().wo_kan(_dp(B), _dp(u), _ip(r), B.shape[0], int(bool(loop)), p, _dp(out))
    return _unpack(out, p)


# --------------------------------------------------------------------------------------
# Restated host preparation (independent of the product's frontend)
# --------------------------------------------------------------------------------------
def pair_modes(reps):
    """Restates ``matched_reps`` (``_hafnian.py:80-159``): greedy pairing of repeated modes.

    Returns (x, edge_reps, oddmode).
    """
    reps = [int(r) for r in reps]
    if sum(reps) == 0:
        return np.array([], dtype=np.int64), np.array([], dtype=np.int64), None
    items = sorted(((r, i) for i, r in enumerate(reps)), reverse=True)
    items = [(r, i) for r, i in items if r > 0]
    left, right, mult = [], [], []
    while len(items) > 1 or (len(items) == 1 and items[0][0] > 1):
        items.sort(reverse=True)
        r0, i0 = items[0]
        if len(items) == 1 or r0 > 2 * items[1][0]:
            left.append(i0)
            right.append(i0)
            mult.append(r0 // 2)
            if r0 % 2 == 0:
                items = items[1:]
            else:
                items[0] = (1, i0)
        else:
            r1, i1 = items[1]
            left.append(i0)
            right.append(i1)
            mult.append(r1)
            if r0 > r1:
                items = [(r0 - r1, i0)] + items[2:]
            else:
                items = items[2:]
    odd = items[0][1] if len(items) == 1 else None
    return np.array(left + right, dtype=np.int64), np.array(mult, dtype=np.int64), odd


def hafnian_sieve(A, loop=False, rpt=None, mu=None, glynn=True, precision="double", nthreads=0,
                  variant=0):
    """End-to-end oracle value of ``_haf`` / ``loop_hafnian`` (``_hafnian.py:470-631``)."""
    A = np.asarray(A)
    n = A.shape[0]
    rpt = [1] * n if rpt is None else [int(r) for r in rpt]
    N = sum(rpt)
    if N == 0:
        return 1.0
    if not loop and N % 2 == 1:
        return 0.0
    D = np.diag(A).copy() if mu is None else np.asarray(mu)
    if loop and N == 1:
        return complex(D[rpt.index(1)])
    x, e, odd = pair_modes(rpt)
    Ax = A[np.ix_(x, x)].astype(np.complex128)
    if not loop:
        return calc_sieve(Ax, e, glynn=glynn, precision=precision, nthreads=nthreads,
                          variant=variant)
    Dx = D[x].astype(np.complex128)
    oddloop = oddV = None
    if odd is not None:
        oddloop = complex(D[odd])
        oddV = A[odd, x].astype(np.complex128)
    return calc_sieve(Ax, e, D=Dx, oddloop=oddloop, oddV=oddV, glynn=glynn, precision=precision,
                      nthreads=nthreads, variant=variant)


def reduction(A, rpt):
    """Row/column repetition (``_hafnian.py:697-714``)."""
    rows = [i for i, r in enumerate(rpt) for _ in range(int(r))]
    A = np.asarray(A)
    if A.ndim == 1:
        return A[rows]
    return A[:, rows][rows]


def double_factorial(n):
    out = 1
    for k in range(n, 0, -2):
        out *= k
    return out


TELEPHONE = [1, 1, 2, 4, 10, 26, 76, 232, 764, 2620, 9496]  # test_hafnian.py:30


# --------------------------------------------------------------------------------------
# loop_hafnian_batch (loop_hafnian_batch.py:211-304), host preparation restated
# --------------------------------------------------------------------------------------
def _batch_edges_even(fixed_edges):
    """loop_hafnian_batch.py:211-231: the batched mode paired with itself as edge 0."""
    if len(fixed_edges) == 0:
        return np.array([0, 0], dtype=int)
    n_edges = len(fixed_edges)
    new = max(fixed_edges) + 1
    half = n_edges // 2
    return np.array([new] + list(fixed_edges[:half]) + [new] + list(fixed_edges[half:]), dtype=int)


def _batch_edges_odd(fixed_edges, oddmode):
    """loop_hafnian_batch.py:234-257: edge 0 the batched mode, edge 1 (odd mode, batched mode)."""
    if len(fixed_edges) == 0:
        return np.array([1, oddmode, 1, 1], dtype=int)
    n_edges = len(fixed_edges)
    new = max(max(fixed_edges), oddmode) + 1
    half = n_edges // 2
    return np.array([new, oddmode] + list(fixed_edges[:half]) + [new, new] + list(fixed_edges[half:]),
                    dtype=int)


def loop_hafnian_batch(A, D, fixed_reps, N_cutoff, glynn=True, precision="double"):
    """Loop hafnians for every photon number 0..N_cutoff of the last mode
    (loop_hafnian_batch.py:260-304)."""
    A = np.asarray(A)
    D = np.asarray(D)
    nz = np.nonzero(list(fixed_reps) + [1])[0]
    Anz = A[np.ix_(nz, nz)]
    Dnz = D[nz]
    fixed_reps_nz = np.asarray(fixed_reps)[nz[:-1]]
    fixed_edges, fixed_m, odd = pair_modes(fixed_reps_nz)
    if odd is None:
        batch_max, cutoff, variant = N_cutoff // 2, N_cutoff % 2, 0
        edges = _batch_edges_even(fixed_edges)
    else:
        batch_max, cutoff, variant = (N_cutoff - 1) // 2, 1 - (N_cutoff % 2), 1
        edges = _batch_edges_odd(fixed_edges, odd)
    Ax = _c(Anz[np.ix_(edges, edges)])
    Dx = _c(Dnz[edges])
    fm = np.ascontiguousarray(fixed_m, dtype=np.int64)
    p = _prec(precision)
    out = np.zeros(4 * (N_cutoff + 4), dtype=np.float64)
    L = lib().wo_loop_hafnian_batch(_dp(Ax), _dp(Dx), Ax.shape[0], _ip(fm) if fm.size else None, fm.size, batch_max,
                                    cutoff, variant, int(bool(glynn)), p, _dp(out))
    if p == 0:
        return out[: 2 * L].view(np.complex128).copy()
    o = out[: 4 * L].reshape(-1, 4)
    return (o[:, 0] + o[:, 1]) + 1j * (o[:, 2] + o[:, 3])


def loop_hafnian_batch_gamma(A, D, fixed_reps, N_cutoff, glynn=True, precision="double"):
    """loop_hafnian_batch_gamma.py:222-269: row k is the batch with diagonal D[k] (the reference
    shares each term's matrix between rows; every row's arithmetic is that of the batch)."""
    D = np.atleast_2d(np.asarray(D))
    return np.array([loop_hafnian_batch(A, D[k], fixed_reps, N_cutoff, glynn, precision) for k in range(D.shape[0])])


# --------------------------------------------------------------------------------------
# montrealer (_montrealer.py)
# --------------------------------------------------------------------------------------
def xmat(n):
    """[[0, I_n], [I_n, 0]] (quantum/conversions.py:55-67)."""
    I = np.identity(n)
    O_ = np.zeros((n, n))
    return np.block([[O_, I], [I, O_]])


def montrealer(Sigma, zeta=None, precision="double", nthreads=0):
    """montrealer(Sigma) / lmontrealer(Sigma, zeta) (_montrealer.py:37-101)."""
    S = _c(Sigma)
    z = None if zeta is None else _c(zeta)
    out = np.zeros(4, dtype=np.float64)
    p = _prec(precision)
    rc = lib().wo_montrealer(_dp(S), _dp(z), S.shape[0], p, nthreads, _dp(out))
    if rc != 0:
        raise ValueError("montrealer: bad size %d" % S.shape[0])
    return _unpack(out, p)


def mtl(A, zeta=None, precision="double"):
    """mtl(A) / lmtl(A, zeta): Sigma = X A (_montrealer.py:104-135)."""
    n = len(A) // 2
    return montrealer(xmat(n) @ np.asarray(A), zeta, precision)


def mtl_walks(A, zeta=None):
    """Montrealer by enumeration of closed walks (pure Python, n <= 6 modes).

    tr(Sigma_S^n) summed with the inclusion-exclusion signs keeps the closed walks of length n
    on the 2n vertices that visit every mode once; a step from vertex u to vertex v weighs
    Sigma[u, v] = A[u ^ n, v] (u ^ n: the other copy of u's mode).  Each walk is counted once
    per starting point and direction, hence mtl = sum / (2n).  The loop part adds the open
    walks visiting every mode, conj(zeta[v_0]) prod A[v_k ^ n, v_(k+1)] zeta[v_(n-1)], over 2.
    """
    from itertools import permutations, product

    A = np.asarray(A)
    n = len(A) // 2
    other = lambda v: v + n if v < n else v - n
    closed, opened = 0j, 0j
    for order in permutations(range(n)):
        for copies in product((0, 1), repeat=n):
            v = [order[k] + n * copies[k] for k in range(n)]
            w = 1 + 0j
            for k in range(n - 1):
                w *= A[other(v[k]), v[k + 1]]
            closed += w * A[other(v[n - 1]), v[0]]
            if zeta is not None:
                opened += np.conj(zeta[v[0]]) * w * zeta[v[n - 1]]
    val = closed / (2 * n)
    if zeta is not None:
        val += opened / 2
    return val


# --------------------------------------------------------------------------------------
# hermite_multidimensional / hafnian_batched (_hermite_multidimensional.py)
# --------------------------------------------------------------------------------------
def hermite_recursion(R, y, cutoffs, C=1.0, renorm=False, precision="double"):
    """_hermite_multidimensional(_renorm)(R, y, G) with G[0] = C (:251-303): the full tensor."""
    R = _c(R)
    y = _c(y)
    cut = np.ascontiguousarray(cutoffs, dtype=np.int64)
    total = int(np.prod(cut))
    p = _prec(precision)
    out = np.zeros((4 if p else 2) * total, dtype=np.float64)
    c = np.array([complex(C).real, complex(C).imag])
    lib().wo_hermite_multidimensional(_dp(R), _dp(y), R.shape[0], _ip(cut), _dp(c), int(bool(renorm)), p, _dp(out))
    if p == 0:
        return out.view(np.complex128).reshape(tuple(cut))
    o = out.reshape(-1, 4)
    return ((o[:, 0] + o[:, 1]) + 1j * (o[:, 2] + o[:, 3])).reshape(tuple(cut))


def hafnian_batched(A, cutoff, mu=None, renorm=False, precision="double"):
    """hafnian_batched (:173-216): the recursion with R = -A, y = mu (modified form)."""
    A = np.asarray(A)
    n = A.shape[0]
    mu = np.zeros(n, dtype=complex) if mu is None else np.asarray(mu)
    return hermite_recursion(-A, mu, [cutoff] * n, 1.0, renorm, precision)


# ---- timed CPU baseline (cpu_timing.cpp): the reference's CPU path shaped as it runs there
# (per-term products as a blocked complex product, FMA on, OpenMP over terms); bench.py's
# cpu_baseline leg times it, the tests check it against the restatement above.
_CT = None


def _host_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("flags"):
                    return " avx512f " in ln + " "
    except OSError:
        pass
    return False


def cpu_timing_lib():
    """The AVX-512 build where the host supports it, else the AVX2 one: (library, isa name)."""
    global _CT
    if _CT is None:
        isa = "x86-64-v4" if _host_has_avx512() else "x86-64-v3"
        path = os.environ.get("WALRUS_CPU_TIMING_LIB") or \
            os.path.join(_HERE, "libcpu_timing_%s.so" % ("v4" if isa == "x86-64-v4" else "v3"))
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ct_sieve.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, dp]
        L.ct_sieve_many.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        _CT = (L, isa)
    return _CT


def ct_prepare(A, loop):
    """The reordered matrix (and loop weights) of an all-ones request (matched_reps with every
    repetition one), as the timed baseline takes them."""
    n = A.shape[0]
    x, _, _ = pair_modes([1] * n)
    Ax = _c(A[np.ix_(x, x)])
    Dx = _c(np.diag(A)[x]) if loop else None
    return Ax, Dx


def ct_sieve(Ax, Dx, t0, t1, nthreads=0):
    """Unscaled sum of the Glynn terms [t0, t1) of one prepared all-ones request."""
    L, _ = cpu_timing_lib()
    out = np.zeros(2)
    L.ct_sieve(_dp(Ax.view(np.float64)), None if Dx is None else _dp(Dx.view(np.float64)), Ax.shape[0],
               int(t0), int(t1), int(nthreads), _dp(out))
    return complex(out[0], out[1])


def ct_sieve_many(Axs, Dxs, nthreads=0):
    """Hafnians of prepared all-ones requests of one size, one request per thread."""
    L, _ = cpu_timing_lib()
    n = Axs[0].shape[0]
    A = np.ascontiguousarray(np.stack(Axs), dtype=np.complex128)
    D = None if Dxs is None else np.ascontiguousarray(np.stack(Dxs), dtype=np.complex128)
    out = np.zeros(2 * len(Axs))
    L.ct_sieve_many(_dp(A.view(np.float64).reshape(-1)), None if D is None else _dp(D.view(np.float64).reshape(-1)),
                    n, len(Axs), int(nthreads), _dp(out))
    return out.view(np.complex128)
