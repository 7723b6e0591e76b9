// walrus_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of thewalrus's power-trace ("eigenvalue-trace") hafnian and loop-hafnian
// sieve.  It exists so the HIP product path can be checked against an independent CPU
// computation; it is never linked into, called by, or shipped with the product library.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Each function cites the reference file:line it restates (paths relative to the
// thewalrus source tree).  Two arithmetic modes share one template:
//   * T = double : reference-faithful operation order inside a term (naive k-loop matrix
//                  products, sequential traces, the same coefficient recursion, the same
//                  Householder + La Budde path).  Compiled with -ffp-contract=off so every
//                  operation rounds once, as numba does without fastmath.
//   * T = dd     : the same algorithm in double-double arithmetic (~32 significant digits),
//                  used as high-precision truth for parity checks.
// Extra pins that are not the reference's algorithm: a subset-DP brute-force loop hafnian
// (definition of the (loop) hafnian as a sum over perfect matchings), Kan's repeated-moment
// formula (docs/algorithms.rst:127-148), and a characteristic-polynomial variant of the
// sieve (Hessenberg + La Budde + Newton identities per term) used for large-n truth.
//
// Parity status per function (what pins it; see tests/test_oracle.py):
//   powertrace      : numpy matrix_power traces; test_labudde.py:48-53,69-74 values.
//   charpoly        : numpy.poly coefficients; test_labudde.py:22-31.
//   f / f_loop / f_loop_odd : mpmath Taylor coefficients of det(I - eta C)^(-1/2) (and the
//                     loop exponential), docs/algorithms.rst:105-118.
//   calc_hafnian / calc_loop_hafnian : closed forms ((n-1)!!, n!, telephone numbers,
//                     identity, rank one), brute force, Kan's formula, reduction identity.
//   brute force     : closed forms on tiny inputs; python pure-loop enumeration.
//   kan             : brute force on the reduced matrix.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <numeric>
#include <algorithm>
#include <omp.h>

// ---------------------------------------------------------------------------------------
// double-double arithmetic (Dekker / Knuth error-free transformations)
// ---------------------------------------------------------------------------------------
struct dd {
    double hi, lo;
    dd() : hi(0.0), lo(0.0) {}
    dd(double x) : hi(x), lo(0.0) {}
    dd(double h, double l) : hi(h), lo(l) {}
};
static inline dd two_sum(double a, double b) {
    double s = a + b, bb = s - a;
    return dd(s, (a - (s - bb)) + (b - bb));
}
static inline dd quick_two_sum(double a, double b) {
    double s = a + b;
    return dd(s, b - (s - a));
}
static inline dd two_prod(double a, double b) {
    double p = a * b;
    return dd(p, std::fma(a, b, -p));
}
static inline dd operator+(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}
static inline dd operator-(dd a) { return dd(-a.hi, -a.lo); }
static inline dd operator-(dd a, dd b) { return a + (-b); }
static inline dd operator*(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p.hi, p.lo);
}
static inline dd operator/(dd a, dd b) {
    double q1 = a.hi / b.hi;
    dd r = a - dd(q1) * b;
    double q2 = r.hi / b.hi;
    r = r - dd(q2) * b;
    double q3 = r.hi / b.hi;
    dd q = quick_two_sum(q1, q2);
    return q + dd(q3);
}
static inline dd& operator+=(dd& a, dd b) { a = a + b; return a; }
static inline dd& operator-=(dd& a, dd b) { a = a - b; return a; }
static inline dd& operator*=(dd& a, dd b) { a = a * b; return a; }
static inline bool operator==(dd a, dd b) { return a.hi == b.hi && a.lo == b.lo; }
static inline bool operator!=(dd a, dd b) { return !(a == b); }
static inline dd dd_sqrt(dd a) {
    if (a.hi <= 0.0) return dd(0.0);
    double x = 1.0 / std::sqrt(a.hi);
    double y = a.hi * x;
    dd yy = two_prod(y, y);
    double corr = (a - yy).hi * (x * 0.5);
    return two_sum(y, corr);
}
static inline double to_double(double x) { return x; }
static inline double to_double(dd x) { return x.hi + x.lo; }
static inline double rsqrt_of(double x) { return std::sqrt(x); }
static inline dd rsqrt_of(dd x) { return dd_sqrt(x); }

// ---------------------------------------------------------------------------------------
// complex over T
// ---------------------------------------------------------------------------------------
template <class T> struct cx {
    T re, im;
    cx() : re(0.0), im(0.0) {}
    cx(T r) : re(r), im(0.0) {}
    cx(T r, T i) : re(r), im(i) {}
};
template <class T> static inline cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>(a.re + b.re, a.im + b.im); }
template <class T> static inline cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>(a.re - b.re, a.im - b.im); }
template <class T> static inline cx<T> operator-(cx<T> a) { return cx<T>(-a.re, -a.im); }
template <class T> static inline cx<T> operator*(cx<T> a, cx<T> b) {
    return cx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// complex / real: componentwise (CPython and numba reduce a real divisor to this)
template <class T> static inline cx<T> divr(cx<T> a, T d) { return cx<T>(a.re / d, a.im / d); }
template <class T> static inline cx<T> mulr(cx<T> a, T d) { return cx<T>(a.re * d, a.im * d); }
template <class T> static inline cx<T> conj(cx<T> a) { return cx<T>(a.re, -a.im); }
template <class T> static inline T abs2(cx<T> a) { return a.re * a.re + a.im * a.im; }
template <class T> static inline bool is_zero(cx<T> a) { return a.re == T(0.0) && a.im == T(0.0); }
template <class T> static inline cx<T>& operator+=(cx<T>& a, cx<T> b) { a = a + b; return a; }
template <class T> static inline cx<T>& operator-=(cx<T>& a, cx<T> b) { a = a - b; return a; }

template <class T> using cvec = std::vector<cx<T>>;

// ---------------------------------------------------------------------------------------
// charpoly.py restatement
// ---------------------------------------------------------------------------------------

// charpoly.py:43-72 get_reflection_vector(matrix, k)
template <class T> static cvec<T> reflection_vector(const cvec<T>& A, int size, int k) {
    int sizeH = size - k;
    int offset = k - 1;
    cvec<T> col(sizeH);
    for (int i = 0; i < sizeH; ++i) col[i] = A[(size_t)(k + i) * size + offset];
    T nrm2 = T(0.0);
    for (int i = 0; i < sizeH; ++i) nrm2 += abs2(col[i]);
    cx<T> sigma(rsqrt_of(nrm2));
    if (!is_zero(col[0])) {
        T a0 = rsqrt_of(abs2(col[0]));
        sigma = sigma * divr(col[0], a0);
    }
    cvec<T> v(col);
    v[0] += sigma;
    return v;
}

// charpoly.py:76-109 apply_householder(A, v, k)
template <class T> static void apply_householder(cvec<T>& A, int size, const cvec<T>& v, int k) {
    int sizeH = (int)v.size();
    T nv2 = T(0.0);
    for (int i = 0; i < sizeH; ++i) nv2 += abs2(v[i]);
    if (nv2 == T(0.0)) return;
    cvec<T> vHA(size - k + 1), Av(size);
    for (int j = 0; j < size - k + 1; ++j)
        for (int l = 0; l < sizeH; ++l) vHA[j] += conj(v[l]) * A[(size_t)(k + l) * size + (k - 1 + j)];
    for (int i = 0; i < sizeH; ++i)
        for (int j = 0; j < size - k + 1; ++j)
            A[(size_t)(k + i) * size + (k - 1 + j)] -= divr((cx<T>(T(2.0)) * v[i]) * vHA[j], nv2);
    for (int i = 0; i < size; ++i)
        for (int l = 0; l < sizeH; ++l) Av[i] += A[(size_t)i * size + (k + l)] * v[l];
    for (int i = 0; i < size; ++i)
        for (int j = 0; j < sizeH; ++j)
            A[(size_t)i * size + (k + j)] -= divr((cx<T>(T(2.0)) * Av[i]) * conj(v[j]), nv2);
}

// charpoly.py:113-125 reduce_matrix_to_hessenberg (in place)
template <class T> static void reduce_to_hessenberg(cvec<T>& A, int size) {
    for (int i = 1; i < size - 1; ++i) {
        cvec<T> v = reflection_vector(A, size, i);
        apply_householder(A, size, v, i);
    }
}

// charpoly.py:196-282 _charpoly(H, k=n): La Budde.  The reference recomputes each running
// product of subdiagonal entries from scratch; extending the previous product by one factor
// multiplies the same factors in the same left-to-right order, so the values are identical.
template <class T> static cvec<T> labudde(const cvec<T>& H, int n) {
    auto h = [&](int i, int j) { return H[(size_t)(i - 1) * n + (j - 1)]; };  // hij
    auto alpha = [&](int i) { return h(i, i); };
    auto beta = [&](int i) { return h(i, i - 1); };
    cvec<T> c((size_t)n * n);
    auto C = [&](int i, int j) -> cx<T>& { return c[(size_t)(i - 1) * n + (j - 1)]; };  // mlo
    if (n == 0) return cvec<T>();
    C(1, 1) = -alpha(1);
    if (n == 1) return cvec<T>{C(1, 1)};
    C(2, 1) = C(1, 1) - alpha(2);
    C(2, 2) = alpha(1) * alpha(2) - h(1, 2) * beta(2);
    for (int i = 3; i <= n; ++i) {
        C(i, 1) = C(i - 1, 1) - alpha(i);
        for (int j = 2; j < i; ++j) {
            cx<T> suma;
            cx<T> bp(T(1.0));
            for (int m = 1; m < j - 1; ++m) {
                bp = bp * beta(i - m + 1);  // product beta(i)..beta(i-m+1)
                suma += (h(i - m, i) * bp) * C(i - m - 1, j - m - 1);
            }
            cx<T> bpj(T(1.0));
            for (int bm = i; bm > i - j + 1; --bm) bpj = bpj * beta(bm);
            C(i, j) = C(i - 1, j) - alpha(i) * C(i - 1, j - 1) - suma - h(i - j + 1, i) * bpj;
        }
        cx<T> suma;
        cx<T> bp(T(1.0));
        for (int m = 1; m < i - 1; ++m) {
            bp = bp * beta(i - m + 1);
            suma += (h(i - m, i) * bp) * C(i - m - 1, i - m - 1);
        }
        cx<T> bpi(T(1.0));
        for (int bm = i; bm > 1; --bm) bpi = bpi * beta(bm);
        C(i, i) = -alpha(i) * C(i - 1, i - 1) - suma - h(1, i) * bpi;
    }
    cvec<T> out(n);
    for (int i = 1; i <= n; ++i) out[i - 1] = C(n, i);
    return out;
}

// charpoly.py:286-298 charpoly(H): mutates H (Hessenberg form) like the reference
template <class T> static cvec<T> charpoly(cvec<T>& H, int m) {
    reduce_to_hessenberg(H, m);
    return labudde(H, m);
}

// row-major m x m product C = A @ B, naive k order
template <class T> static void matmul(const cvec<T>& A, const cvec<T>& B, cvec<T>& C, int m) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            cx<T> s;
            for (int k = 0; k < m; ++k) s += A[(size_t)i * m + k] * B[(size_t)k * m + j];
            C[(size_t)i * m + j] = s;
        }
}
template <class T> static cx<T> trace(const cvec<T>& A, int m) {
    cx<T> s;
    for (int i = 0; i < m; ++i) s += A[(size_t)i * m + i];
    return s;
}

// charpoly.py:301-327 powertrace(H, n): returns p[0..n-1].  The m = 0 case (all digits
// cancel) yields zeros beyond p[0]; the reference indexes an empty matrix there, and every
// later use multiplies the zeros, so the term is zero either way.
template <class T> static cvec<T> powertrace(cvec<T>& H, int m, int n) {
    cvec<T> p;
    p.reserve(n + 2);
    if (m == 0) {
        p.assign(std::max(n, 2), cx<T>());
        return p;
    }
    int min_val = std::min(n, m);
    p.push_back(cx<T>(T((double)m)));
    p.push_back(trace(H, m));
    cvec<T> A(H), Bm((size_t)m * m);
    for (int it = 0; it < min_val - 2; ++it) {
        matmul(A, H, Bm, m);
        A.swap(Bm);
        p.push_back(trace(A, m));
    }
    if (n <= m) { p.resize(n); return p; }
    cvec<T> cp = charpoly(H, m);
    for (int t = min_val; t < n; ++t) {
        cx<T> s;
        size_t L = p.size();
        for (int k = 0; k < m; ++k) s -= cp[k] * p[L - 1 - k];
        p.push_back(s);
    }
    return p;
}

// ---------------------------------------------------------------------------------------
// _hafnian.py restatement
// ---------------------------------------------------------------------------------------

// The coefficient recursion shared by f, f_loop and f_loop_odd: multiplies the running
// polynomial by exp(factor * eta^i), truncated at degree top (_hafnian.py:200-208).
template <class T> static void comb_step(cvec<T>& cur, const cvec<T>& old, int i, int top, cx<T> factor) {
    cur = old;
    cx<T> powfactor(T(1.0));
    for (int j = 1; j <= top / i; ++j) {
        powfactor = powfactor * divr(factor, T((double)j));
        for (int k = i * j + 1; k <= top + 1; ++k) cur[k - 1] += old[k - i * j - 1] * powfactor;
    }
}

// _hafnian.py:184-209 f(A, n)
template <class T> static cvec<T> f_plain(cvec<T>& AX, int m, int N) {
    int h = N / 2;
    cvec<T> p = powertrace(AX, m, h + 1);
    cvec<T> a(h + 1), b(h + 1);
    a[0] = cx<T>(T(1.0));
    for (int i = 1; i <= h; ++i) {
        cx<T> factor = divr(p[i], T(2.0 * i));
        comb_step(b, a, i, h, factor);
        a.swap(b);
    }
    return a;
}

template <class T> static cx<T> dotu(const cvec<T>& x, const cvec<T>& y) {
    cx<T> s;
    for (size_t i = 0; i < x.size(); ++i) s += x[i] * y[i];
    return s;
}
// row vector times matrix: (x @ A)[j]
template <class T> static cvec<T> vecmat(const cvec<T>& x, const cvec<T>& A, int m) {
    cvec<T> y(m);
    for (int j = 0; j < m; ++j) {
        cx<T> s;
        for (int i = 0; i < m; ++i) s += x[i] * A[(size_t)i * m + j];
        y[j] = s;
    }
    return y;
}
// matrix times column vector: (A @ x)[i]
template <class T> static cvec<T> matvec(const cvec<T>& A, const cvec<T>& x, int m) {
    cvec<T> y(m);
    for (int i = 0; i < m; ++i) {
        cx<T> s;
        for (int j = 0; j < m; ++j) s += A[(size_t)i * m + j] * x[j];
        y[i] = s;
    }
    return y;
}

// _hafnian.py:213-242 f_loop(AX, AX_S, XD_S, D_S, n)
template <class T>
static cvec<T> f_loop(cvec<T>& AX, const cvec<T>& AX_S, cvec<T> XD, const cvec<T>& D, int m, int N) {
    int h = N / 2;
    cvec<T> p = powertrace(AX, m, h + 1);
    cvec<T> a(h + 1), b(h + 1);
    a[0] = cx<T>(T(1.0));
    for (int i = 1; i <= h; ++i) {
        cx<T> factor = divr(p[i], T(2.0 * i)) + divr(dotu(XD, D), T(2.0));
        XD = vecmat(XD, AX_S, m);
        comb_step(b, a, i, h, factor);
        a.swap(b);
    }
    return a;
}

// _hafnian.py:247-285 f_loop_odd(AX, AX_S, XD_S, D_S, n, oddloop, oddVX_S)
template <class T>
static cvec<T> f_loop_odd(cvec<T>& AX, const cvec<T>& AX_S, const cvec<T>& XD, cvec<T> D, int m, int N,
                          cx<T> oddloop, const cvec<T>& oddVX) {
    cvec<T> p = powertrace(AX, m, N / 2 + 1);
    cvec<T> a(N + 1), b(N + 1);
    a[0] = cx<T>(T(1.0));
    for (int i = 1; i <= N; ++i) {
        cx<T> factor;
        if (i == 1) {
            factor = oddloop;
        } else if (i % 2 == 0) {
            factor = divr(p[i / 2], T((double)i)) + divr(dotu(XD, D), T(2.0));
        } else {
            factor = dotu(oddVX, D);
            D = matvec(AX_S, D, m);
        }
        comb_step(b, a, i, N, factor);
        a.swap(b);
    }
    return a;
}

// _hafnian.py:27-45 nb_binom (exact in int64 for the sizes used here)
// _hafnian.py:27-45 nb_binom.  The reference multiplies before dividing in int64, which wraps
// for n >= 62; dividing out the common factor first keeps every intermediate <= the result.
static int64_t binom_i(int64_t n, int64_t k) {
    if (k < 0 || k > n) return 0;
    if (k == 0 || k == n) return 1;
    uint64_t b = 1;
    for (int64_t i = 0; i < std::min(k, n - k); ++i) {
        uint64_t num = (uint64_t)(n - i), den = (uint64_t)(i + 1);
        uint64_t g = std::gcd(b, den);
        b /= g;
        b *= num / (den / g);
    }
    return (int64_t)b;
}

// _hafnian.py:163-180 find_kept_edges: digit 0 is the most significant, base e_i + 1
static void kept_edges(int64_t j, const int64_t* e, int E, int64_t* k) {
    for (int i = E - 1; i >= 0; --i) {
        int64_t base = e[i] + 1;
        k[i] = j % base;
        j /= base;
    }
}

// _hafnian.py:289-356 get_AX_S / get_submatrices for one term (delta = scaled kept edges)
template <class T> struct TermMats {
    int m = 0;
    cvec<T> AX, XD, D, oddVX;
};
template <class T>
static void term_matrices(const cvec<T>& A, int twoE, const int64_t* delta, int E, const cvec<T>* D,
                          const cvec<T>* oddV, TermMats<T>& out) {
    std::vector<int> S, R;
    for (int i = 0; i < E; ++i)
        if (delta[i] != 0) S.push_back(i);
    int s = (int)S.size();
    for (int i = 0; i < s; ++i) R.push_back(S[i]);
    for (int i = 0; i < s; ++i) R.push_back(S[i] + E);
    int m = 2 * s;
    out.m = m;
    out.AX.assign((size_t)m * m, cx<T>());
    for (int r = 0; r < m; ++r)
        for (int c = 0; c < s; ++c) {
            T d = T((double)delta[S[c]]);
            out.AX[(size_t)r * m + c] = mulr(A[(size_t)R[r] * twoE + R[s + c]], d);
            out.AX[(size_t)r * m + s + c] = mulr(A[(size_t)R[r] * twoE + R[c]], d);
        }
    if (D) {
        out.D.assign(m, cx<T>());
        out.XD.assign(m, cx<T>());
        for (int r = 0; r < m; ++r) out.D[r] = (*D)[R[r]];
        for (int c = 0; c < s; ++c) {
            T d = T((double)delta[S[c]]);
            out.XD[c] = mulr(out.D[s + c], d);
            out.XD[s + c] = mulr(out.D[c], d);
        }
    }
    if (oddV) {
        out.oddVX.assign(m, cx<T>());
        for (int c = 0; c < s; ++c) {
            T d = T((double)delta[S[c]]);
            out.oddVX[c] = mulr((*oddV)[R[s + c]], d);
            out.oddVX[s + c] = mulr((*oddV)[R[c]], d);
        }
    }
}


// Power traces p[0..n-1] from the characteristic polynomial (Hessenberg + La Budde, then the
// Newton identities p_k = -k c_k - sum_{i<k} c_i p_{k-i}, with c_i = 0 for i > m).  Not the
// reference's product loop: an independent per-term route, O(m^3) instead of O(m^3 h), used
// only to produce high-precision truth at large n.
template <class T> static cvec<T> powertrace_cp(cvec<T>& H, int m, int n) {
    cvec<T> p(std::max(n, 2));
    if (m == 0) return p;
    p[0] = cx<T>(T((double)m));
    cvec<T> c = charpoly(H, m);
    for (int k = 1; k < n; ++k) {
        cx<T> s;
        if (k <= m) s = mulr(c[k - 1], T(-(double)k));
        for (int i = 1; i < k && i <= m; ++i) s -= c[i - 1] * p[k - i];
        p[k] = s;
    }
    p.resize(n);
    return p;
}

template <class T>
static void coeff_from_traces(const cvec<T>& p, cvec<T>& a, cvec<T>& b, int h) {
    a.assign(h + 1, cx<T>());
    b.assign(h + 1, cx<T>());
    a[0] = cx<T>(T(1.0));
    for (int i = 1; i <= h; ++i) {
        comb_step(b, a, i, h, divr(p[i], T(2.0 * i)));
        a.swap(b);
    }
}

// One term's f-value, selecting the reference route (variant 0) or the char-poly route.
template <class T>
static cx<T> term_value(TermMats<T>& tm, int N, bool loop, const cx<T>* oddloop, int variant) {
    cvec<T> AX(tm.AX);
    int m = tm.m;
    if (variant == 0) {
        if (!loop) return f_plain(AX, m, N)[N / 2];
        if (oddloop) return f_loop_odd(AX, tm.AX, tm.XD, tm.D, m, N, *oddloop, tm.oddVX)[N];
        return f_loop(AX, tm.AX, tm.XD, tm.D, m, N)[N / 2];
    }
    // variant 1: identical coefficient recursion, traces from the characteristic polynomial
    if (!loop) {
        int h = N / 2;
        cvec<T> p = powertrace_cp(AX, m, h + 1), a, b;
        coeff_from_traces(p, a, b, h);
        return a[h];
    }
    if (oddloop) {
        cvec<T> p = powertrace_cp(AX, m, N / 2 + 1);
        cvec<T> a(N + 1), b(N + 1), D(tm.D);
        a[0] = cx<T>(T(1.0));
        for (int i = 1; i <= N; ++i) {
            cx<T> factor;
            if (i == 1) factor = *oddloop;
            else if (i % 2 == 0) factor = divr(p[i / 2], T((double)i)) + divr(dotu(tm.XD, D), T(2.0));
            else { factor = dotu(tm.oddVX, D); D = matvec(tm.AX, D, m); }
            comb_step(b, a, i, N, factor);
            a.swap(b);
        }
        return a[N];
    }
    int h = N / 2;
    cvec<T> p = powertrace_cp(AX, m, h + 1);
    cvec<T> a(h + 1), b(h + 1), XD(tm.XD);
    a[0] = cx<T>(T(1.0));
    for (int i = 1; i <= h; ++i) {
        cx<T> factor = divr(p[i], T(2.0 * i)) + divr(dotu(XD, tm.D), T(2.0));
        XD = vecmat(XD, tm.AX, m);
        comb_step(b, a, i, h, factor);
        a.swap(b);
    }
    return a[h];
}

// _hafnian.py:417-467 _calc_hafnian and :513-577 _calc_loop_hafnian in one driver.
// D == nullptr: plain hafnian.  oddloop != nullptr: odd-N loop hafnian.
// The sum over terms is an OpenMP reduction over static thread blocks, as numba's prange.
template <class T>
static cx<T> calc_sieve(const cvec<T>& A, int twoE, const int64_t* e, int E, const cvec<T>* D,
                        const cx<T>* oddloop, const cvec<T>* oddV, bool glynn, int nthreads,
                        int variant, double* terms_out = nullptr) {
    int64_t esum = 0, emax = 0;
    for (int i = 0; i < E; ++i) { esum += e[i]; emax = std::max(emax, e[i]); }
    int N = (int)(2 * esum) + (oddloop ? 1 : 0);
    int64_t steps = 1;
    if (glynn && !oddloop) {  // _hafnian.py:432-435, :535-538
        steps = (e[0] + 2) / 2;
        for (int i = 1; i < E; ++i) steps *= e[i] + 1;
    } else {
        for (int i = 0; i < E; ++i) steps *= e[i] + 1;
    }
    int mb = (int)emax + 1;  // _hafnian.py:438-439 precompute_binoms(max(e) + 1)
    std::vector<int64_t> binoms((size_t)(mb + 1) * (mb + 1));
    for (int i = 0; i <= mb; ++i)
        for (int j = 0; j <= mb; ++j) binoms[(size_t)i * (mb + 1) + j] = binom_i(i, j);

    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<cx<T>> partial(nt);
    bool loop = D != nullptr;
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        cx<T> Hloc;
        std::vector<int64_t> k(E), delta(E);
        TermMats<T> tm;
#pragma omp for schedule(static)
        for (int64_t j = 0; j < steps; ++j) {
            kept_edges(j, e, E, k.data());
            int64_t edge_sum = 0;
            for (int i = 0; i < E; ++i) edge_sum += k[i];
            double binom_prod = 1.0;  // _hafnian.py:447-449
            for (int i = 0; i < E; ++i) binom_prod *= (double)binoms[(size_t)e[i] * (mb + 1) + k[i]];
            for (int i = 0; i < E; ++i) delta[i] = glynn ? 2 * k[i] - e[i] : k[i];
            term_matrices(A, twoE, delta.data(), E, D, oddV, tm);
            double prefac = (((N / 2 - edge_sum) % 2) ? -1.0 : 1.0) * binom_prod;
            if (!oddloop && glynn && delta[0] == 0) prefac *= 0.5;  // :458-459, :565-566
            cx<T> fval = term_value(tm, N, loop, oddloop, variant);
            Hloc += mulr(fval, T(prefac));
            if (terms_out) {
                cx<T> tv = mulr(fval, T(prefac));
                terms_out[2 * j] = to_double(tv.re);
                terms_out[2 * j + 1] = to_double(tv.im);
            }
        }
        partial[tid] = Hloc;
    }
    cx<T> H;
    for (int t = 0; t < nt; ++t) H += partial[t];
    if (glynn) {  // :464-465, :571-575
        int ex = oddloop ? N / 2 : N / 2 - 1;
        H = mulr(H, T(std::ldexp(1.0, -ex)));
    }
    return H;
}

// ---------------------------------------------------------------------------------------
// Independent pins (not the reference's algorithm)
// ---------------------------------------------------------------------------------------

// Loop hafnian by definition: the sum over all partitions of {0..n-1} into pairs (weight
// A_ij) and, with loops, singletons (weight A_ii).  Subset dynamic programme over the lowest
// remaining vertex; O(2^n n).  reference.py:229-284 states the same sum by enumeration.
template <class T> static cx<T> brute_force(const cvec<T>& A, int n, bool loop) {
    if (n == 0) return cx<T>(T(1.0));
    size_t full = ((size_t)1 << n) - 1;
    cvec<T> g((size_t)1 << n);
    g[0] = cx<T>(T(1.0));
    for (size_t S = 1; S <= full; ++S) {
        int i = __builtin_ctzll(S);
        size_t R = S & ~((size_t)1 << i);
        cx<T> v;
        if (loop) v = A[(size_t)i * n + i] * g[R];
        for (size_t Q = R; Q; Q &= Q - 1) {
            int j = __builtin_ctzll(Q);
            v += A[(size_t)i * n + j] * g[R & ~((size_t)1 << j)];
        }
        g[S] = v;
    }
    return g[full];
}

// Kan's repeated-moment formula (docs/algorithms.rst:127-148) for the loop hafnian of the
// matrix with row/column i repeated r_i times and diagonal replaced by u (u = 0: hafnian).
template <class T>
static cx<T> kan_formula(const cvec<T>& B, const cvec<T>& u, const int64_t* r, int n, bool loop) {
    int64_t M = 0;
    for (int i = 0; i < n; ++i) M += r[i];
    if (M == 0) return cx<T>(T(1.0));
    if (!loop && (M % 2)) return cx<T>();
    int64_t total = 1;
    for (int i = 0; i < n; ++i) total *= r[i] + 1;
    // factorials up to M in T
    std::vector<T> fact(M + 1);
    fact[0] = T(1.0);
    for (int64_t i = 1; i <= M; ++i) fact[i] = fact[i - 1] * T((double)i);
    cx<T> res;
    std::vector<int64_t> nu(n);
    for (int64_t idx = 0; idx < total; ++idx) {
        int64_t t = idx, nsum = 0;
        double binp = 1.0;
        for (int i = n - 1; i >= 0; --i) {
            nu[i] = t % (r[i] + 1);
            t /= r[i] + 1;
            nsum += nu[i];
            binp *= (double)binom_i(r[i], nu[i]);
        }
        // h_i = r_i/2 - nu_i
        std::vector<T> hv(n);
        for (int i = 0; i < n; ++i) hv[i] = T(0.5 * (double)r[i] - (double)nu[i]);
        cx<T> hBh;
        for (int i = 0; i < n; ++i) {
            cx<T> row;
            for (int j = 0; j < n; ++j) row += mulr(B[(size_t)i * n + j], hv[j]);
            hBh += mulr(row, hv[i]);
        }
        cx<T> q = divr(hBh, T(2.0));
        cx<T> hu;
        if (loop)
            for (int i = 0; i < n; ++i) hu += mulr(u[i], hv[i]);
        cx<T> inner;
        for (int64_t rr = 0; 2 * rr <= M; ++rr) {
            if (!loop && 2 * rr != M) continue;
            cx<T> term(T(1.0));
            for (int64_t z = 0; z < rr; ++z) term = term * q;
            for (int64_t z = 0; z < M - 2 * rr; ++z) term = term * hu;
            inner += divr(term, fact[rr] * fact[M - 2 * rr]);
        }
        T sgn = T((nsum % 2) ? -binp : binp);
        res += mulr(inner, sgn);
    }
    return res;
}

// ---------------------------------------------------------------------------------------
// C ABI (ctypes).  Complex arrays are interleaved (re, im) doubles, row-major; dd results are
// returned as (re_hi, re_lo, im_hi, im_lo).
// precision: 0 = double, 1 = double-double.
// ---------------------------------------------------------------------------------------
template <class T> static cvec<T> load(const double* x, size_t n) {
    cvec<T> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = cx<T>(T(x[2 * i]), T(x[2 * i + 1]));
    return v;
}
static void store(const cx<double>& z, double* out) { out[0] = z.re; out[1] = z.im; }
static void store(const cx<dd>& z, double* out) {
    out[0] = z.re.hi; out[1] = z.re.lo; out[2] = z.im.hi; out[3] = z.im.lo;
}

template <class T>
static void sieve_entry(const double* A, int64_t twoE, const int64_t* e, int64_t E, const double* D,
                        const double* oddloop, const double* oddV, int glynn, int nthreads, int variant,
                        double* out, double* terms_out = nullptr) {
    cvec<T> Av = load<T>(A, (size_t)twoE * twoE);
    cvec<T> Dv, oV;
    cx<T> ol;
    if (D) Dv = load<T>(D, twoE);
    if (oddloop) ol = cx<T>(T(oddloop[0]), T(oddloop[1]));
    if (oddV) oV = load<T>(oddV, twoE);
    cx<T> H = calc_sieve<T>(Av, (int)twoE, e, (int)E, D ? &Dv : nullptr, oddloop ? &ol : nullptr,
                            oddV ? &oV : nullptr, glynn != 0, nthreads, variant, terms_out);
    store(H, out);
}

extern "C" {

// Sieve at the reference's kernel seam (_hafnian.py:417 / :513): A is the reordered
// 2E x 2E matrix, e the edge multiplicities.  D == NULL selects the plain hafnian.
int wo_calc_sieve(const double* A, int64_t twoE, const int64_t* e, int64_t E, const double* D,
                  const double* oddloop, const double* oddV, int glynn, int precision, int nthreads,
                  int variant, double* out) {
    if (twoE != 2 * E || E < 1) return -1;
    if (precision == 0) sieve_entry<double>(A, twoE, e, E, D, oddloop, oddV, glynn, nthreads, variant, out);
    else sieve_entry<dd>(A, twoE, e, E, D, oddloop, oddV, glynn, nthreads, variant, out);
    return 0;
}

// Same sieve, also writing every signed term (prefactor * f) to terms_out (2 doubles per term,
// in term order), before the final power-of-two scale.
int wo_sieve_terms(const double* A, int64_t twoE, const int64_t* e, int64_t E, const double* D,
                   const double* oddloop, const double* oddV, int glynn, int precision, int nthreads,
                   double* out, double* terms_out) {
    if (twoE != 2 * E || E < 1) return -1;
    if (precision == 0) sieve_entry<double>(A, twoE, e, E, D, oddloop, oddV, glynn, nthreads, 0, out, terms_out);
    else sieve_entry<dd>(A, twoE, e, E, D, oddloop, oddV, glynn, nthreads, 0, out, terms_out);
    return 0;
}

int wo_powertrace(const double* H, int64_t m, int64_t n, int precision, double* out) {
    if (precision == 0) {
        cvec<double> Hv = load<double>(H, (size_t)m * m);
        cvec<double> p = powertrace(Hv, (int)m, (int)n);
        for (int64_t i = 0; i < n; ++i) store(p[i], out + 2 * i);
    } else {
        cvec<dd> Hv = load<dd>(H, (size_t)m * m);
        cvec<dd> p = powertrace(Hv, (int)m, (int)n);
        for (int64_t i = 0; i < n; ++i) store(p[i], out + 4 * i);
    }
    return 0;
}

int wo_charpoly(const double* H, int64_t m, double* out) {
    cvec<double> Hv = load<double>(H, (size_t)m * m);
    cvec<double> c = charpoly(Hv, (int)m);
    for (int64_t i = 0; i < m; ++i) store(c[i], out + 2 * i);
    return 0;
}

// f (_hafnian.py:184), f_loop (:213) or f_loop_odd (:247) on explicit term matrices.
// mode 0 = f, 1 = f_loop, 2 = f_loop_odd.  Returns all coefficients (h + 1 or N + 1).
int wo_fcoeffs(const double* AX, int64_t m, int64_t N, int mode, const double* XD, const double* D,
               const double* oddloop, const double* oddVX, double* out) {
    cvec<double> AXv = load<double>(AX, (size_t)m * m), AX2(AXv);
    cvec<double> r;
    if (mode == 0) r = f_plain(AX2, (int)m, (int)N);
    else if (mode == 1) r = f_loop(AX2, AXv, load<double>(XD, m), load<double>(D, m), (int)m, (int)N);
    else r = f_loop_odd(AX2, AXv, load<double>(XD, m), load<double>(D, m), (int)m, (int)N,
                        cx<double>(oddloop[0], oddloop[1]), load<double>(oddVX, m));
    for (size_t i = 0; i < r.size(); ++i) store(r[i], out + 2 * i);
    return (int)r.size();
}

int wo_brute_force(const double* A, int64_t n, int loop, int precision, double* out) {
    if (n > 30) return -1;
    if (precision == 0) store(brute_force(load<double>(A, (size_t)n * n), (int)n, loop != 0), out);
    else store(brute_force(load<dd>(A, (size_t)n * n), (int)n, loop != 0), out);
    return 0;
}

int wo_kan(const double* B, const double* u, const int64_t* r, int64_t n, int loop, int precision,
           double* out) {
    if (precision == 0) {
        cvec<double> uv = u ? load<double>(u, n) : cvec<double>(n);
        store(kan_formula(load<double>(B, (size_t)n * n), uv, r, (int)n, loop != 0), out);
    } else {
        cvec<dd> uv = u ? load<dd>(u, n) : cvec<dd>(n);
        store(kan_formula(load<dd>(B, (size_t)n * n), uv, r, (int)n, loop != 0), out);
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// loop_hafnian_batch.py restatement (GBS sampler core)
// ---------------------------------------------------------------------------------------

// _calc_loop_hafnian_batch_even (loop_hafnian_batch.py:52-123) when odd_variant == 0 and
// _calc_loop_hafnian_batch_odd (:128-208) when odd_variant == 1.  `cutoff` is odd_cutoff
// (even variant) or even_cutoff (odd variant).  The reference hands the same matrix copy to
// successive powertrace calls, so later calls see the Hessenberg form left by earlier ones;
// that order is kept.  Output length: 2 batch_max + cutoff + 1 (+1 for the odd variant).
template <class T>
static std::vector<cx<T>> lhb(const cvec<T>& A, const cvec<T>& D, int twoE, const int64_t* fixed,
                              int Ef, int batch_max, int cutoff, bool odd_variant, bool glynn) {
    int E = Ef + (odd_variant ? 2 : 1);
    std::vector<int64_t> e(E);
    e[0] = batch_max;
    if (odd_variant) e[1] = 1;
    for (int i = 0; i < Ef; ++i) e[(odd_variant ? 2 : 1) + i] = fixed[i];
    int64_t fsum = 0;
    for (int i = 0; i < Ef; ++i) fsum += fixed[i];
    int N_fixed = (int)(2 * fsum) + (odd_variant ? 1 : 0);
    int N_max = odd_variant ? N_fixed + 2 * batch_max + cutoff + 1 : N_fixed + 2 * batch_max + cutoff;
    int L = 2 * batch_max + cutoff + (odd_variant ? 2 : 1);
    int64_t steps = 1;
    for (int i = 0; i < E; ++i) steps *= e[i] + 1;
    int64_t emax = 0;
    for (int i = 0; i < E; ++i) emax = std::max(emax, e[i]);
    int mb = (int)emax + cutoff;  // precompute_binoms(max(edge_reps) + cutoff)
    std::vector<int64_t> binoms((size_t)(mb + 1) * (mb + 1));
    for (int i = 0; i <= mb; ++i)
        for (int j = 0; j <= mb; ++j) binoms[(size_t)i * (mb + 1) + j] = binom_i(i, j);
    auto BN = [&](int64_t a, int64_t b) { return (double)binoms[(size_t)a * (mb + 1) + b]; };

    cvec<T> oddV(twoE), oddV0(twoE);
    for (int c = 0; c < twoE; ++c) oddV[c] = A[c];  // A[0, :]
    if (odd_variant)
        for (int c = 0; c < twoE; ++c) oddV0[c] = A[(size_t)twoE + c];  // A[1, :]
    cx<T> oddloop = D[0], oddloop0 = odd_variant ? D[1] : cx<T>();

    std::vector<cx<T>> H(L);
    std::vector<int64_t> k(E), delta(E);
    TermMats<T> tm, tm0;
    for (int64_t j = 0; j < steps; ++j) {
        kept_edges(j, e.data(), E, k.data());
        int64_t edges_sum = 0;
        for (int i = 0; i < E; ++i) edges_sum += k[i];
        double binom_prod = 1.0;
        for (int i = 1; i < E; ++i) binom_prod *= BN(e[i], k[i]);
        for (int i = 0; i < E; ++i) delta[i] = glynn ? 2 * k[i] - e[i] : k[i];
        term_matrices(A, twoE, delta.data(), E, &D, &oddV, tm);
        cvec<T> AXc(tm.AX);
        std::vector<cx<T>> Hnew(L);
        if (odd_variant && k[0] == 0 && k[1] == 0) {
            // get_submatrix_batch_odd0 (_hafnian.py:360-380)
            term_matrices(A, twoE, delta.data(), E, &D, &oddV0, tm0);
            double pm = (((N_fixed / 2 - edges_sum) % 2) != 0) ? -1.0 : 1.0;
            cx<T> f = f_loop_odd(AXc, tm.AX, tm.XD, tm.D, tm.m, N_fixed, oddloop0, tm0.oddVX)[N_fixed];
            H[0] += mulr(f, T(binom_prod * pm));
        }
        cvec<T> f_even = f_loop(AXc, tm.AX, tm.XD, tm.D, tm.m, N_max);
        cvec<T> f_odd = f_loop_odd(AXc, tm.AX, tm.XD, tm.D, tm.m, N_max, oddloop, tm.oddVX);
        int start = (int)(2 * k[0]) + (odd_variant ? 1 : 0);
        for (int N_det = start; N_det < L; ++N_det) {
            int N = N_fixed + N_det;
            double pm = (((N / 2 - edges_sum) % 2) != 0) ? -1.0 : 1.0;
            int half = odd_variant ? (N_det - 1) / 2 : N_det / 2;
            double nb = BN(half, k[0]) * binom_prod;
            bool use_even = odd_variant ? (N % 2 == 0) : (N_det % 2 == 0);
            cx<T> f = use_even ? f_even[N / 2] : f_odd[N];
            Hnew[N_det] += mulr(f, T(nb * pm));
        }
        for (int q = 0; q < L; ++q) H[q] += Hnew[q];
    }
    if (glynn)
        for (int q = 0; q < L; ++q) H[q] = mulr(H[q], T(std::ldexp(1.0, -((N_fixed + q) / 2))));
    return H;
}

extern "C" int wo_loop_hafnian_batch(const double* A, const double* D, int64_t twoE, const int64_t* fixed,
                                     int64_t Ef, int64_t batch_max, int64_t cutoff, int odd_variant, int glynn,
                                     int precision, double* out) {
    if (precision == 0) {
        auto H = lhb<double>(load<double>(A, (size_t)twoE * twoE), load<double>(D, twoE), (int)twoE, fixed,
                             (int)Ef, (int)batch_max, (int)cutoff, odd_variant != 0, glynn != 0);
        for (size_t q = 0; q < H.size(); ++q) store(H[q], out + 2 * q);
        return (int)H.size();
    }
    auto H = lhb<dd>(load<dd>(A, (size_t)twoE * twoE), load<dd>(D, twoE), (int)twoE, fixed, (int)Ef,
                     (int)batch_max, (int)cutoff, odd_variant != 0, glynn != 0);
    for (size_t q = 0; q < H.size(); ++q) store(H[q], out + 4 * q);
    return (int)H.size();
}

// ---------------------------------------------------------------------------------------
// _montrealer.py restatement (montrealer / lmontrealer over Sigma = X A)
// ---------------------------------------------------------------------------------------

// _montrealer.py:37-57 (zeta == nullptr) and :79-101: the inclusion-exclusion over subsets S
// of the n modes (dec2bin order, :17-33: bit n-1-i of p selects mode i), rows/columns
// S u (S + n) of Sigma, sign (-1)^(|S|+1), p_n = powertrace(sub, n + 1)[-1] and, with zeta,
// power_loop(sub, zeta_S, n) = conj(zeta_S) sub^(n-1) zeta_S (:60-76), taken before
// powertrace (which may reduce sub in place).  Result (-1)^(n+1) (val / 2n + loops / 2).
template <class T>
static cx<T> montrealer(const cvec<T>& Sg, const cvec<T>* zeta, int n2, int nthreads) {
    const int n = n2 / 2;
    const int64_t tot = (int64_t)1 << n;
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<cx<T>> pv(nt), pl(nt);
#pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        cx<T> val, loops;
        std::vector<int> pos;
#pragma omp for schedule(static)
        for (int64_t p = 0; p < tot; ++p) {
            pos.clear();
            for (int i = 0; i < n; ++i)
                if ((p >> (n - 1 - i)) & 1) pos.push_back(i);
            const int len = (int)pos.size();
            for (int i = 0; i < len; ++i) pos.push_back(pos[i] + n);
            const int m = 2 * len;
            cvec<T> sub((size_t)m * m);
            for (int r = 0; r < m; ++r)
                for (int c = 0; c < m; ++c) sub[(size_t)r * m + c] = Sg[(size_t)pos[r] * n2 + pos[c]];
            const T sign = T(((len + 1) & 1) ? -1.0 : 1.0);
            if (zeta) {
                cvec<T> sv(m), vec;
                for (int r = 0; r < m; ++r) sv[r] = (*zeta)[pos[r]];
                vec = sv;
                for (int it = 0; it < n - 1; ++it) vec = matvec(sub, vec, m);
                cvec<T> cz(m);
                for (int r = 0; r < m; ++r) cz[r] = conj(sv[r]);
                loops += mulr(dotu(cz, vec), sign);
            }
            cvec<T> pt = powertrace(sub, m, n + 1);
            val += mulr(pt[n], sign);
        }
        pv[tid] = val;
        pl[tid] = loops;
    }
    cx<T> val, loops;
    for (int t = 0; t < nt; ++t) { val += pv[t]; loops += pl[t]; }
    const T s = T((n + 1) & 1 ? -1.0 : 1.0);
    cx<T> r = divr(val, T((double)(2 * n)));
    if (zeta) r += divr(loops, T(2.0));
    return mulr(r, s);
}

extern "C" int wo_montrealer(const double* Sigma, const double* zeta, int64_t n2, int precision, int nthreads,
                             double* out) {
    if (n2 % 2 || n2 < 2 || n2 > 64) return -1;
    if (precision == 0) {
        cvec<double> z = zeta ? load<double>(zeta, n2) : cvec<double>();
        store(montrealer(load<double>(Sigma, (size_t)n2 * n2), zeta ? &z : nullptr, (int)n2, nthreads), out);
        return 0;
    }
    cvec<dd> z = zeta ? load<dd>(zeta, n2) : cvec<dd>();
    store(montrealer(load<dd>(Sigma, (size_t)n2 * n2), zeta ? &z : nullptr, (int)n2, nthreads), out);
    return 0;
}

// ---------------------------------------------------------------------------------------
// _hermite_multidimensional.py restatement (hafnian_batched / hermite_multidimensional)
// ---------------------------------------------------------------------------------------

// _hermite_multidimensional.py:251-275 (renorm) and :279-303: G in row-major (np.ndindex)
// order, G[0] = C given; for every other index, i = the first nonzero position,
// ki = idx - e_i, G[idx] = y_i G[ki] - sum_(l: ki_l > 0) w(ki_l) R_il G[ki - e_l] with
// w(k) = k, or w(k) = sqrt(k) and a final division by sqrt(idx_i) when renormalised.
template <class T>
static void hermite_md(const cvec<T>& R, const cvec<T>& y, int n, const int64_t* cut, bool renorm, cvec<T>& G) {
    std::vector<int64_t> stride(n);
    int64_t total = 1;
    for (int l = n - 1; l >= 0; --l) { stride[l] = total; total *= cut[l]; }
    std::vector<int64_t> idx(n, 0);
    auto sq = [](int64_t k) { return rsqrt_of(T((double)k)); };  // sqrt, in T's precision
    for (int64_t lin = 1; lin < total; ++lin) {
        for (int l = n - 1; l >= 0; --l) {  // advance the multi-index (last position fastest)
            if (++idx[l] < cut[l]) break;
            idx[l] = 0;
        }
        int i = 0;
        while (idx[i] == 0) ++i;
        const int64_t ki = lin - stride[i];
        cx<T> u = y[i] * G[ki];
        for (int l = 0; l < n; ++l) {
            const int64_t kl = idx[l] - (l == i ? 1 : 0);
            if (kl <= 0) continue;
            const T w = renorm ? sq(kl) : T((double)kl);
            u -= (mulr(R[(size_t)i * n + l], w)) * G[ki - stride[l]];
        }
        G[lin] = renorm ? divr(u, sq(idx[i])) : u;
    }
}

extern "C" int wo_hermite_multidimensional(const double* R, const double* y, int64_t n, const int64_t* cut,
                                           const double* C, int renorm, int precision, double* out) {
    int64_t total = 1;
    for (int64_t l = 0; l < n; ++l) total *= cut[l];
    if (precision == 0) {
        cvec<double> G(total);
        G[0] = cx<double>(C[0], C[1]);
        hermite_md(load<double>(R, (size_t)n * n), load<double>(y, n), (int)n, cut, renorm != 0, G);
        for (int64_t q = 0; q < total; ++q) store(G[q], out + 2 * q);
        return 0;
    }
    cvec<dd> G(total);
    G[0] = cx<dd>(dd(C[0]), dd(C[1]));
    hermite_md(load<dd>(R, (size_t)n * n), load<dd>(y, n), (int)n, cut, renorm != 0, G);
    for (int64_t q = 0; q < total; ++q) store(G[q], out + 4 * q);
    return 0;
}
