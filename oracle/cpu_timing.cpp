// cpu_timing.cpp -- TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py.
//
// A performance restatement of the reference's CPU path for the all-ones Glynn sieve (the
// plain and even loop hafnians of the serving stream): _calc_hafnian / _calc_loop_hafnian
// (thewalrus/_hafnian.py:415-467, :511-577) with numba's prange over terms (:443, :546) as an
// OpenMP loop over terms, and per term what the reference does (:444-462, :547-569):
//   * the term matrix AX_S = Ax X diag(d, d) (get_AX_S / get_submatrices, :288-356; every
//     multiplicity is one, so d = +-1 and no row is dropped);
//   * charpoly.powertrace (charpoly.py:301-327): tr(AX_S^k) for k = 1..h by h - 1 products of
//     n x n complex matrices (numba's A @ H, a BLAS zgemm), here a register-blocked complex
//     product on split real / imaginary rows with FMA contraction;
//   * for loops the row chain XD <- XD @ AX_S of f_loop (:228-234);
//   * the coefficient of eta^h by the factor-by-factor recursion of f / f_loop (:196-208,
//     :228-241).
// It is not the faithful oracle (walrus_oracle.cpp keeps the reference's rounding with
// -ffp-contract=off and a naive product); this file is built twice, for x86-64-v4 (AVX-512) and
// x86-64-v3 (AVX2 + FMA), and the loader picks the one the host supports.  Only bench.py's
// cpu_baseline leg (and the tests that check it against the oracle) load it.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

typedef double v8d __attribute__((vector_size(64), aligned(8)));  // unaligned loads and stores

// Split complex matrix: nr x np doubles per part, rows padded to a multiple of 4 and columns to a
// multiple of 8 with zeros (the padding never enters a trace or a product's kept block).
struct SMat {
    int nr = 0, np = 0;
    std::vector<double> re, im;
    void init(int n) {
        nr = (n + 3) / 4 * 4;
        np = (n + 7) / 8 * 8;
        re.assign((size_t)nr * np, 0.0);
        im.assign((size_t)nr * np, 0.0);
    }
};

// Q = P H for n x n split matrices (n <= nr): 4 rows x 8 columns of Q in registers, k in order.
void cmatmul(const SMat& P, const SMat& H, SMat& Q, int n) {
    const int np = P.np;
    for (int i = 0; i < n; i += 4) {
        const double* p0r = &P.re[(size_t)i * np];
        const double* p0i = &P.im[(size_t)i * np];
        for (int j = 0; j < np; j += 8) {
            v8d a0r = {}, a1r = {}, a2r = {}, a3r = {}, a0i = {}, a1i = {}, a2i = {}, a3i = {};
            const double* hr_p = &H.re[j];
            const double* hi_p = &H.im[j];
            for (int k = 0; k < n; ++k) {
                const v8d hr = *(const v8d*)(hr_p + (size_t)k * np);
                const v8d hi = *(const v8d*)(hi_p + (size_t)k * np);
                const double r0 = p0r[k], r1 = p0r[np + k], r2 = p0r[2 * np + k], r3 = p0r[3 * np + k];
                const double i0 = p0i[k], i1 = p0i[np + k], i2 = p0i[2 * np + k], i3 = p0i[3 * np + k];
                a0r += r0 * hr - i0 * hi;
                a0i += r0 * hi + i0 * hr;
                a1r += r1 * hr - i1 * hi;
                a1i += r1 * hi + i1 * hr;
                a2r += r2 * hr - i2 * hi;
                a2i += r2 * hi + i2 * hr;
                a3r += r3 * hr - i3 * hi;
                a3i += r3 * hi + i3 * hr;
            }
            double* qr = &Q.re[(size_t)i * np + j];
            double* qi = &Q.im[(size_t)i * np + j];
            *(v8d*)qr = a0r;
            *(v8d*)(qr + np) = a1r;
            *(v8d*)(qr + 2 * np) = a2r;
            *(v8d*)(qr + 3 * np) = a3r;
            *(v8d*)qi = a0i;
            *(v8d*)(qi + np) = a1i;
            *(v8d*)(qi + 2 * np) = a2i;
            *(v8d*)(qi + 3 * np) = a3i;
        }
    }
}

struct Work {
    SMat H, P, Q;
    std::vector<double> xr, xi, yr, yi;       // loop row chain
    std::vector<double> pr, pi, lr, li;       // traces, loop scalars
    std::vector<double> c0r, c0i, c1r, c1i;   // coefficient rows
    void init(int n, int h) {
        H.init(n);
        P.init(n);
        Q.init(n);
        xr.assign(n, 0.0); xi.assign(n, 0.0); yr.assign(n, 0.0); yi.assign(n, 0.0);
        pr.assign(h + 1, 0.0); pi.assign(h + 1, 0.0); lr.assign(h + 1, 0.0); li.assign(h + 1, 0.0);
        c0r.assign(h + 1, 0.0); c0i.assign(h + 1, 0.0); c1r.assign(h + 1, 0.0); c1i.assign(h + 1, 0.0);
    }
};

// One term of the Glynn sieve: returns the signed term (sign times [eta^h]).
void one_term(const double* Ax, const double* D, int n, uint64_t t, Work& w, double& out_re, double& out_im) {
    const int E = n / 2, h = E;
    const int np = w.H.np;
    // digits of t (digit 0 most significant, the first fixed at 0 by the Glynn halving)
    int ksum = 0;
    double d[64];
    for (int i = 0; i < E; ++i) {
        const int k = (int)((t >> (E - 1 - i)) & 1u);
        ksum += k;
        d[i] = d[i + E] = 2.0 * k - 1.0;
    }
    // AX_S = Ax X diag(d): column c is column (c + E) mod n of Ax, times d_c
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            const int src = c < E ? c + E : c - E;
            w.H.re[(size_t)r * np + c] = Ax[2 * ((size_t)r * n + src)] * d[c];
            w.H.im[(size_t)r * np + c] = Ax[2 * ((size_t)r * n + src) + 1] * d[c];
        }
    // powertrace: p_1 = tr(H), p_k = tr(H^k) by H^k = H^(k-1) H
    double tr = 0.0, ti = 0.0;
    for (int i = 0; i < n; ++i) { tr += w.H.re[(size_t)i * np + i]; ti += w.H.im[(size_t)i * np + i]; }
    w.pr[1] = tr;
    w.pi[1] = ti;
    const SMat* cur = &w.H;
    SMat* nxt = &w.P;
    for (int k = 2; k <= h; ++k) {
        cmatmul(*cur, w.H, *nxt, n);
        tr = ti = 0.0;
        for (int i = 0; i < n; ++i) { tr += nxt->re[(size_t)i * np + i]; ti += nxt->im[(size_t)i * np + i]; }
        w.pr[k] = tr;
        w.pi[k] = ti;
        cur = nxt;
        nxt = (nxt == &w.P) ? &w.Q : &w.P;
    }
    // f_loop's row chain: v_k = XD_S AX_S^(k-1) D_S with XD_S = d (X D)
    if (D) {
        for (int c = 0; c < n; ++c) {
            const int src = c < E ? c + E : c - E;
            w.xr[c] = D[2 * src] * d[c];
            w.xi[c] = D[2 * src + 1] * d[c];
        }
        for (int k = 1; k <= h; ++k) {
            double sr = 0.0, si = 0.0;
            for (int c = 0; c < n; ++c) {
                sr += w.xr[c] * D[2 * c] - w.xi[c] * D[2 * c + 1];
                si += w.xr[c] * D[2 * c + 1] + w.xi[c] * D[2 * c];
            }
            w.lr[k] = sr;
            w.li[k] = si;
            if (k == h) break;
            for (int c = 0; c < n; ++c) { w.yr[c] = 0.0; w.yi[c] = 0.0; }
            for (int r = 0; r < n; ++r) {
                const double a = w.xr[r], b = w.xi[r];
                for (int c = 0; c < n; ++c) {
                    const double hr = w.H.re[(size_t)r * np + c], hi = w.H.im[(size_t)r * np + c];
                    w.yr[c] += a * hr - b * hi;
                    w.yi[c] += a * hi + b * hr;
                }
            }
            w.xr.swap(w.yr);
            w.xi.swap(w.yi);
        }
    }
    // [eta^h] prod_i exp(a_i eta^i), factor by factor (a_i = p_i / 2i (+ v_i / 2))
    double* ar = w.c0r.data();
    double* ai = w.c0i.data();
    double* br = w.c1r.data();
    double* bi = w.c1i.data();
    for (int k = 0; k <= h; ++k) { ar[k] = k == 0 ? 1.0 : 0.0; ai[k] = 0.0; }
    for (int i = 1; i <= h; ++i) {
        double fr = w.pr[i] / (2.0 * i), fi = w.pi[i] / (2.0 * i);
        if (D) { fr += w.lr[i] / 2.0; fi += w.li[i] / 2.0; }
        std::memcpy(br, ar, sizeof(double) * (h + 1));
        std::memcpy(bi, ai, sizeof(double) * (h + 1));
        double qr = 1.0, qi = 0.0;
        for (int j = 1; j <= h / i; ++j) {
            const double tr2 = (qr * fr - qi * fi) / j, ti2 = (qr * fi + qi * fr) / j;
            qr = tr2;
            qi = ti2;
            for (int k = i * j; k <= h; ++k) {
                br[k] += ar[k - i * j] * qr - ai[k - i * j] * qi;
                bi[k] += ar[k - i * j] * qi + ai[k - i * j] * qr;
            }
        }
        std::swap(ar, br);
        std::swap(ai, bi);
    }
    const double sign = ((h - ksum) & 1) ? -1.0 : 1.0;
    out_re = sign * ar[h];
    out_im = sign * ai[h];
}

}  // namespace

extern "C" {

// Terms [t0, t1) of the Glynn sieve of one all-ones request (Ax: the reordered n x n complex128
// matrix, D: its reordered loop weights or NULL), OpenMP over terms; out: the unscaled sum.
void ct_sieve(const double* Ax, const double* D, int n, int64_t t0, int64_t t1, int nthreads, double* out) {
    const int h = n / 2;
    double sr = 0.0, si = 0.0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : sr, si)
    {
        Work w;
        w.init(n, h);
#pragma omp for schedule(static)
        for (int64_t t = t0; t < t1; ++t) {
            double re, im;
            one_term(Ax, D, n, (uint64_t)t, w, re, im);
            sr += re;
            si += im;
        }
    }
    out[0] = sr;
    out[1] = si;
}

// `count` requests of one size (matrices and weights back to back), one request per thread at a
// time, every term of each; out: 2 count doubles, the hafnians (scaled by 2^-(h-1)).
void ct_sieve_many(const double* Ax, const double* D, int n, int count, int nthreads, double* out) {
    const int h = n / 2;
    const int64_t steps = (int64_t)1 << (h - 1);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        Work w;
        w.init(n, h);
#pragma omp for schedule(dynamic, 1)
        for (int q = 0; q < count; ++q) {
            const double* A = Ax + (size_t)2 * n * n * q;
            const double* Dq = D ? D + (size_t)2 * n * q : nullptr;
            double sr = 0.0, si = 0.0;
            for (int64_t t = 0; t < steps; ++t) {
                double re, im;
                one_term(A, Dq, n, (uint64_t)t, w, re, im);
                sr += re;
                si += im;
            }
            const double sc = std::ldexp(1.0, -(h - 1));
            out[2 * q] = sr * sc;
            out[2 * q + 1] = si * sc;
        }
    }
}

}  // extern "C"
