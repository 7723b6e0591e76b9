// sieve_pm1.hip -- the Glynn sieve for even N with every repetition equal to one: the terms of
// the plain hafnian (_hafnian.py:443-462) and the even loop hafnian (:546-569) when rpt is all
// ones.  Every digit is then 0/1, every delta is +-1, no row is ever dropped (m = n), and
//     AX_S = M diag(d),   M = Ax X  (the same for all 2^(h-1) terms of a matrix),
// with d_c = delta_(c mod h).  A workgroup keeps M in LDS once; each wave walks terms and
// forms B^(a+1) = (B^a M) diag(d) with a register-tiled product: the wave is an 8 x 8 grid of
// lanes, each owning a T x T tile of the n_pad x n_pad power (n_pad = 8T, zero padded).
//
// Template parameters: T (tile side), REAL (matrix and loop weights real: the same products
// in real arithmetic, which is what complex arithmetic computes when every imaginary part is
// zero), LOOP (add the loop terms of f_loop, _hafnian.py:233-234).
#include "wh_device.h"
#include "wh_launch.h"

namespace {

template <bool REAL> struct Elem;
template <> struct Elem<true> {
    using T = double;
};
template <> struct Elem<false> {
    using T = cplx;
};

__device__ __forceinline__ void emac(double& acc, double a, double b) { acc = fma(a, b, acc); }
__device__ __forceinline__ void emac(cplx& acc, cplx a, cplx b) { cmac(acc, a, b); }
__device__ __forceinline__ double escale(double x, double s) { return x * s; }
__device__ __forceinline__ cplx escale(cplx x, double s) { return cmulr(x, s); }
__device__ __forceinline__ void ezero(double& x) { x = 0.0; }
__device__ __forceinline__ void ezero(cplx& x) { x = cmk(0.0, 0.0); }
__device__ __forceinline__ cplx etoc(double x) { return cmk(x, 0.0); }
__device__ __forceinline__ cplx etoc(cplx x) { return x; }
__device__ __forceinline__ double eload(const cplx* p, int64_t i, double*) { return p[i].re; }
__device__ __forceinline__ cplx eload(const cplx* p, int64_t i, cplx*) { return p[i]; }
__device__ __forceinline__ double eone(double*) { return 1.0; }
__device__ __forceinline__ cplx eone(cplx*) { return cmk(1.0, 0.0); }
__device__ __forceinline__ double ezero_v(double*) { return 0.0; }
__device__ __forceinline__ cplx ezero_v(cplx*) { return cmk(0.0, 0.0); }
__device__ __forceinline__ double eget(cplx z, double*) { return z.re; }
__device__ __forceinline__ cplx eget(cplx z, cplx*) { return z; }
__device__ __forceinline__ double emul(double a, double b) { return a * b; }
__device__ __forceinline__ cplx emul(cplx a, cplx b) { return cmul(a, b); }
__device__ __forceinline__ double eadd(double a, double b) { return a + b; }
__device__ __forceinline__ cplx eadd(cplx a, cplx b) { return cadd(a, b); }
__device__ __forceinline__ double ediv(double a, double d) { return a / d; }
__device__ __forceinline__ cplx ediv(cplx a, double d) { return cdivr(a, d); }

// factor of column c of the current term: delta_(c mod h) = +1 if digit set, else `unset`
// (-1 for the Glynn sieve, 0 for the montrealer's subset inclusion-exclusion)
__device__ __forceinline__ double col_sign(uint64_t t, int c, int h, double unset) {
    int i = c < h ? c : c - h;
    if (i >= h) return 1.0;  // zero padding
    return ((t >> (h - 1 - i)) & 1ull) ? 1.0 : unset;
}


// ---- tile helpers: lane (ti, tj) owns rows I0..I0+T-1, columns J0..J0+T-1 of the power.
// LDS matrices are (8T) x (8T) with row stride 8T + 1 elements: the eight row tiles a
// product step reads then start on distinct LDS banks.
#ifndef WH_PM1_PAD
#define WH_PM1_PAD 1
#endif
#ifndef WH_PM1_KUNROLL
#define WH_PM1_KUNROLL 2
#endif
#ifndef WH_PM1_PIPELINE
#define WH_PM1_PIPELINE 1
#endif
// Reciprocals 1/k (k <= WH_EMAX) for the exponential recurrence of the term coefficient.
__constant__ double kInvJ[WH_EMAX + 1] = {
    0.0,        1.0,        1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,  1.0 / 7,  1.0 / 8,
    1.0 / 9,    1.0 / 10,   1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17,
    1.0 / 18,   1.0 / 19,   1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26,
    1.0 / 27,   1.0 / 28,   1.0 / 29, 1.0 / 30, 1.0 / 31, 1.0 / 32};

// [eta^h] exp(sum_i a_i eta^i) for h <= H: e_0 = 1, e_k = (1/k) sum_(i=1..k) b_i e_(k-i) with
// b_i = i a_i = p_i / 2 (+ i v_i / 2 with loops); fixed trip counts keep e[] in registers.
template <class E, int H, bool LOOP>
__device__ __forceinline__ E coeff_top_e(const cplx* pt, const cplx* lv, int h) {
    E b[H + 1], e[H + 1];
#pragma unroll
    for (int i = 1; i <= H; ++i) {
        E v = escale(eget(pt[i], (E*)nullptr), 0.5);
        if (LOOP) v = eadd(v, escale(eget(lv[i], (E*)nullptr), 0.5 * i));
        b[i] = i <= h ? v : ezero_v((E*)nullptr);
    }
    e[0] = eone((E*)nullptr);
    E res = e[0];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        E s0 = ezero_v((E*)nullptr), s1 = ezero_v((E*)nullptr);
#pragma unroll
        for (int i = 1; i <= k; i += 2) emac(s0, b[i], e[k - i]);
#pragma unroll
        for (int i = 2; i <= k; i += 2) emac(s1, b[i], e[k - i]);
        e[k] = escale(eadd(s0, s1), kInvJ[k]);
        if (k == h) res = e[k];
    }
    return res;
}

// The same coefficient with b_i kept in the term's slot (written over p_i) instead of registers:
// about half the registers of coeff_top_e, for the repeated-row kernel's H = 16.
template <class E, int H, bool LOOP>
__device__ __forceinline__ E coeff_top_lds(cplx* pt, const cplx* lv, int h) {
#pragma unroll
    for (int i = 1; i <= H; ++i) {
        E v = escale(eget(pt[i], (E*)nullptr), 0.5);
        if (LOOP) v = eadd(v, escale(eget(lv[i], (E*)nullptr), 0.5 * i));
        pt[i] = etoc(i <= h ? v : ezero_v((E*)nullptr));
    }
    E e[H + 1];
    e[0] = eone((E*)nullptr);
    E res = e[0];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        E s0 = ezero_v((E*)nullptr), s1 = ezero_v((E*)nullptr);
#pragma unroll
        for (int i = 1; i <= k; i += 2) emac(s0, eget(pt[i], (E*)nullptr), e[k - i]);
#pragma unroll
        for (int i = 2; i <= k; i += 2) emac(s1, eget(pt[i], (E*)nullptr), e[k - i]);
        e[k] = escale(eadd(s0, s1), kInvJ[k]);
        if (k == h) res = e[k];
    }
    return res;
}

template <int T> struct Lds {
    static constexpr int S = 8 * T + WH_PM1_PAD;
};

// Per wave, NB term slots of p_0 .. p_(4T+1) and v_0 .. v_(4T+1) (h <= 4T; a spare last entry):
// the coefficients of up to NB terms are formed together, one term per lane.  NB fits the LDS
// beside the powers (complex T = 5 at 4 waves, T = 6 at 3 waves).
__host__ __device__ constexpr int pm1_ptn(int T) { return 4 * T + 4; }
__host__ __device__ constexpr int pm1_nb(int T, bool real) { return real ? 8 : T == 6 ? 2 : T == 5 ? 7 : 8; }

template <int T, class E>
__device__ __forceinline__ void tile_store(E* Pw, const E (&acc)[T][T], int I0, int J0) {
    wave_lds_sync();  // the other lanes' reads of the previous power are done
#pragma unroll
    for (int r = 0; r < T; ++r)
#pragma unroll
        for (int c = 0; c < T; ++c) Pw[(I0 + r) * Lds<T>::S + J0 + c] = acc[r][c];
    wave_lds_sync();  // the whole power is stored before any lane reads it
}

// acc = (Pw Ms) diag(sign_J): one step of the power sequence, k summed in order.  The
// operands of step k + 1 are loaded while step k's multiply-adds run (register double
// buffer), so LDS latency is hidden even with one wave per SIMD.
template <int T, class E>
__device__ __forceinline__ void tile_load_k(const E* Pw, const E* Ms, int I0, int J0, int k, E (&av)[T],
                                            E (&bv)[T]) {
    constexpr int S = Lds<T>::S;
#pragma unroll
    for (int r = 0; r < T; ++r) av[r] = Pw[(I0 + r) * S + k];
#pragma unroll
    for (int c = 0; c < T; ++c) bv[c] = Ms[k * S + J0 + c];
}

template <int T, class E>
__device__ __forceinline__ void tile_mac_k(E (&nacc)[T][T], const E (&av)[T], const E (&bv)[T]) {
#pragma unroll
    for (int r = 0; r < T; ++r)
#pragma unroll
        for (int c = 0; c < T; ++c) emac(nacc[r][c], av[r], bv[c]);
}

template <int T, class E>
__device__ __forceinline__ void tile_product(const E* Pw, const E* Ms, E (&acc)[T][T], int I0, int J0,
                                             int n, const double (&sgnJ)[T]) {
    E nacc[T][T];
#pragma unroll
    for (int r = 0; r < T; ++r)
#pragma unroll
        for (int c = 0; c < T; ++c) ezero(nacc[r][c]);
#if WH_PM1_PIPELINE
    E a0[T], b0[T], a1[T], b1[T];
    tile_load_k<T>(Pw, Ms, I0, J0, 0, a0, b0);
    int k = 0;
    for (; k + 2 <= n; k += 2) {
        tile_load_k<T>(Pw, Ms, I0, J0, k + 1, a1, b1);
        tile_mac_k<T>(nacc, a0, b0);
        // unconditional prefetch (the last one re-reads row k, unused) keeps the operand
        // registers free of branch merges
        const int kn = k + 2 < n ? k + 2 : k;
        tile_load_k<T>(Pw, Ms, I0, J0, kn, a0, b0);
        tile_mac_k<T>(nacc, a1, b1);
    }
    if (k < n) tile_mac_k<T>(nacc, a0, b0);
    // keep later LDS reads (transposed tiles) from being hoisted into the loop's live range
    asm volatile("" ::: "memory");
#else
#pragma unroll WH_PM1_KUNROLL
    for (int k = 0; k < n; ++k) {
        E av[T], bv[T];
        tile_load_k<T>(Pw, Ms, I0, J0, k, av, bv);
        tile_mac_k<T>(nacc, av, bv);
    }
#endif
#pragma unroll
    for (int r = 0; r < T; ++r)
#pragma unroll
        for (int c = 0; c < T; ++c) acc[r][c] = escale(nacc[r][c], sgnJ[c]);
}

template <int T, class E>
__device__ __forceinline__ E tile_trace(const E (&acc)[T][T], bool diag) {
    E tr;
    ezero(tr);
    if (diag) {
#pragma unroll
        for (int r = 0; r < T; ++r) tr = eadd(tr, acc[r][r]);
    }
    return tr;
}

// lane part of <acc, X^T> = sum_(i in I, j in J) acc[i][j] X[j][i], X in LDS
template <int T, class E>
__device__ __forceinline__ E tile_tdot(const E* X, const E (&acc)[T][T], int I0, int J0) {
    E s;
    ezero(s);
#pragma unroll
    for (int c = 0; c < T; ++c)
#pragma unroll
        for (int r = 0; r < T; ++r) emac(s, acc[r][c], X[(J0 + c) * Lds<T>::S + I0 + r]);
    return s;
}

// lane part of XD_d B D
template <int T, class E>
__device__ __forceinline__ E tile_bilinear(const E (&acc)[T][T], const E (&xdI)[T], const E (&dJ)[T]) {
    E bl;
    ezero(bl);
#pragma unroll
    for (int r = 0; r < T; ++r) {
        E rowsum;
        ezero(rowsum);
#pragma unroll
        for (int c = 0; c < T; ++c) emac(rowsum, acc[r][c], dJ[c]);
        emac(bl, xdI[r], rowsum);
    }
    return bl;
}

__device__ __forceinline__ double shfl_e(double v, int src) { return __shfl(v, src, WH_WAVE); }
__device__ __forceinline__ cplx shfl_e(cplx v, int src) {
    return cmk(__shfl(v.re, src, WH_WAVE), __shfl(v.im, src, WH_WAVE));
}
__device__ __forceinline__ double group8_e(double v) { return group8_sum(v); }
__device__ __forceinline__ cplx group8_e(cplx v) { return cmk(group8_sum(v.re), group8_sum(v.im)); }
__device__ __forceinline__ double stride8_e(double v) { return stride8_sum(v); }
__device__ __forceinline__ cplx stride8_e(cplx v) { return cmk(stride8_sum(v.re), stride8_sum(v.im)); }

// (B D)[I0 + r]: rows summed over the 8 column tiles of the lane's row group
template <int T, class E>
__device__ __forceinline__ void tile_zvec(const E (&acc)[T][T], const E (&dJ)[T], E (&z)[T]) {
#pragma unroll
    for (int r = 0; r < T; ++r) {
        ezero(z[r]);
#pragma unroll
        for (int c = 0; c < T; ++c) emac(z[r], acc[r][c], dJ[c]);
        z[r] = group8_e(z[r]);
    }
}

// (XD_d B)[J0 + c]: columns summed over the 8 row tiles of the lane's column group
template <int T, class E>
__device__ __forceinline__ void tile_yvec(const E (&acc)[T][T], const E (&xdI)[T], E (&y)[T]) {
#pragma unroll
    for (int c = 0; c < T; ++c) {
        ezero(y[c]);
#pragma unroll
        for (int r = 0; r < T; ++r) emac(y[c], xdI[r], acc[r][c]);
        y[c] = stride8_e(y[c]);
    }
}

template <int T, bool REAL, bool LOOP, bool HALF>
__device__ __forceinline__ void pm1_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                         double4* __restrict__ partials, int64_t b,
                                         unsigned char* smem) {
    using E = typename Elem<REAL>::T;
    constexpr int NP = 8 * T;  // padded side
    constexpr int S = Lds<T>::S;  // LDS row stride (elements)
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    const int ti = lane >> 3, tj = lane & 7;
    const int n = 2 * J.E;
    const int h = J.E;
    // montrealer terms (_montrealer.py:37-101): 0/1 column factors, p_h and v_h only
    const bool mtl = (J.flags & WH_F_MTL) != 0;
    const double unset = mtl ? 0.0 : -1.0;

    // ---- LDS carve
    E* Ms = reinterpret_cast<E*>(smem);
    E* Pw = Ms + NP * S + wave * NP * S;
    E* vecs = Ms + (1 + nw) * NP * S;           // XD0[NP], D[NP] (loop)
    E* XD0 = vecs;
    E* Dv = vecs + NP;
    constexpr int PTN = pm1_ptn(T);
    constexpr int NB = pm1_nb(T, REAL);
    cplx* slots = reinterpret_cast<cplx*>(vecs + 2 * NP) + wave * (NB * 2 * PTN);
    double4* wres = reinterpret_cast<double4*>(reinterpret_cast<cplx*>(vecs + 2 * NP) + nw * (NB * 2 * PTN));

    // ---- stage M = Ax X (column c of M is column (c + h) mod n of Ax), zero padded
    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < NP * NP; e += blockDim.x) {
        int r = e / NP, c = e - r * NP;
        E v;
        ezero(v);
        if (r < n && c < n) {
            int src = c < h ? c + h : c - h;
            v = eload(A, (int64_t)r * n + src, (E*)nullptr);
        }
        Ms[r * S + c] = v;
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < NP; e += blockDim.x) {
            E xd, dd;
            ezero(xd);
            ezero(dd);
            if (e < n) {
                if (mtl) {  // the caller stages conj(X zeta) and X zeta
                    xd = eload(pool + J.v_off, e, (E*)nullptr);
                    dd = eload(Dg, e, (E*)nullptr);
                } else {
                    int src = e < h ? e + h : e - h;  // (X D)[e]
                    xd = eload(Dg, src, (E*)nullptr);
                    dd = eload(Dg, e, (E*)nullptr);
                }
            }
            XD0[e] = xd;
            Dv[e] = dd;
        }
    }
    __syncthreads();

    const int64_t t0 = (b - J.chunk_base) * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;
    const int I0 = ti * T, J0 = tj * T;

    csum acc_sum;
    acc_sum.zero();

    int slot = 0;
    int64_t t_batch = t0 + wave;  // the batch's first term
    for (int64_t t = t0 + wave; t < t1; t += nw) {
        const uint64_t tu = (uint64_t)t;
        cplx* pt = slots + slot * 2 * PTN;  // p[0..h]
        cplx* lv = pt + PTN;                // loop scalars v[1..h]
        double sgnJ[T];
#pragma unroll
        for (int c = 0; c < T; ++c) sgnJ[c] = col_sign(tu, J0 + c, h, unset);
        const double prefac = mtl ? (((__popcll(tu) + 1) & 1) ? -1.0 : 1.0)
                                  : (((h - __popcll(tu)) & 1) ? -1.0 : 1.0);

        // P_1 = M diag(d)
        E acc[T][T];
#pragma unroll
        for (int r = 0; r < T; ++r)
#pragma unroll
            for (int c = 0; c < T; ++c) acc[r][c] = escale(Ms[(I0 + r) * S + J0 + c], sgnJ[c]);

        // loop: per-lane row weights XD_d[I] and column weights D[J], read from LDS where
        // used so they do not occupy registers across the products
        auto load_xd = [&](E (&xdI)[T]) {
#pragma unroll
            for (int r = 0; r < T; ++r) xdI[r] = escale(XD0[I0 + r], col_sign(tu, I0 + r, h, unset));
        };
        auto load_d = [&](E (&dJ)[T]) {
#pragma unroll
            for (int c = 0; c < T; ++c) dJ[c] = Dv[J0 + c];
        };

        if constexpr (!HALF) {
            // the reference's order: B^(a+1) = B^a B for a < h, p_a = tr(B^a)
            for (int a = 1; a <= h; ++a) {
                E tr = tile_trace<T>(acc, ti == tj);
                E bl;
                ezero(bl);
                if (LOOP && a <= h - 1) {
                    E xdI[T], dJ[T];
                    load_xd(xdI);
                    load_d(dJ);
                    bl = tile_bilinear<T>(acc, xdI, dJ);
                }
                cplx trc = wave_csum(etoc(tr));
                cplx blc = LOOP ? wave_csum(etoc(bl)) : cmk(0.0, 0.0);
                if (lane == 0) {
                    pt[a] = trc;
                    if (LOOP && a <= h - 1) lv[a + 1] = blc;
                }
                if (a == h) break;
                tile_store<T>(Pw, acc, I0, J0);
                tile_product<T>(Pw, Ms, acc, I0, J0, n, sgnJ);
            }
        } else {
            // powers up to B^A, A = ceil(h/2); tr(B^(2a)) = <B^a, (B^a)^T> and
            // tr(B^(2a+1)) = <B^(a+1), (B^a)^T>.  Loop scalars v_k = XD B^(k-1) D from
            // y_a = XD B^a (column tiles) and z_a = B^a D (row tiles):
            // v_(2a) = y_a . z_(a-1), v_(2a+1) = y_a . z_a, dotted on the diagonal lanes, where
            // the lane's column and row tiles cover the same indices.
            const int A = (h + 1) / 2;
            E zprev[T];
            if (LOOP) {
#pragma unroll
                for (int r = 0; r < T; ++r) zprev[r] = Dv[I0 + r];  // z_0 = D
            }
            tile_store<T>(Pw, acc, I0, J0);
            for (int a = 1; a <= A; ++a) {
                E qo, qe;
                if (a == 1) {
                    qo = tile_trace<T>(acc, ti == tj);   // tr(B)
                    qe = tile_tdot<T>(Pw, acc, I0, J0);  // <B, B^T>
                } else {
                    tile_product<T>(Pw, Ms, acc, I0, J0, n, sgnJ);  // acc = B^a
                    qo = tile_tdot<T>(Pw, acc, I0, J0);            // Pw holds B^(a-1)
                    tile_store<T>(Pw, acc, I0, J0);
                    qe = tile_tdot<T>(Pw, acc, I0, J0);
                }
                // the power's results, summed over the wave eight at once: p_ko, p_ke, v_2a, v_(2a+1)
                E de, dodd;
                ezero(de);
                ezero(dodd);
                E znew[T];
                if (LOOP) {
                    E y[T], xdI[T], dJ[T];
                    load_xd(xdI);
                    load_d(dJ);
                    tile_yvec<T>(acc, xdI, y);
                    tile_zvec<T>(acc, dJ, znew);
                    if (ti == tj) {
#pragma unroll
                        for (int c = 0; c < T; ++c) {
                            emac(de, y[c], zprev[c]);
                            emac(dodd, y[c], znew[c]);
                        }
                    }
                }
                const cplx qoc = etoc(qo), qec = etoc(qe), dec = etoc(de), doc = etoc(dodd);
                const double v8[8] = {qoc.re, qoc.im, qec.re, qec.im, dec.re, dec.im, doc.re, doc.im};
                const double sum8 = wave_sum8(v8, lane);
                const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
                if ((lane & 7) == 0) {  // lane 8k holds the sum of v8[k]
                    const int k = lane >> 3;
                    double* dst = nullptr;
                    if (k < 2) dst = &pt[ko].re;
                    else if (k < 4) dst = &pt[ke].re;  // ke <= h + 1 < PTN
                    else if (LOOP && k < 6 && 2 * a <= h) dst = &lv[2 * a].re;
                    else if (LOOP && k >= 6 && 2 * a + 1 <= h) dst = &lv[2 * a + 1].re;
                    if (dst) dst[k & 1] = sum8;
                }
                if (LOOP) {
#pragma unroll
                    for (int r = 0; r < T; ++r) zprev[r] = znew[r];
                }
            }
        }
        if (LOOP) {
            // v_1 = XD_d . D
            E part;
            ezero(part);
            for (int e = lane; e < n; e += WH_WAVE) emac(part, escale(XD0[e], col_sign(tu, e, h, unset)), Dv[e]);
            cplx v1 = wave_csum(etoc(part));
            if (lane == 0) lv[1] = v1;
        }
        wave_lds_sync();  // pt and lv as lane 0 stored them

        if (mtl) {
            // montrealer term: sign (p_h / 2h + v_h / 2), v_h = conj(zeta_S) Sigma_S^(h-1) zeta_S
            if (lane == h) {
                E f = ediv(eget(pt[h], (E*)nullptr), (double)(2 * h));
                if (LOOP) f = eadd(f, ediv(eget(lv[h], (E*)nullptr), 2.0));
                acc_sum.add(cmulr(etoc(f), prefac));
            }
            continue;
        }
        // coefficient of eta^h in exp(sum_i a_i eta^i), a_i = p_i / 2i (+ v_i / 2): the
        // reference builds it factor by factor (_hafnian.py:200-208, :232-241); here lane j runs
        // the power-series exponential e_k = (1/k) sum_(i<=k) (i a_i) e_(k-i) in registers for
        // the j-th term of a batch of up to NB terms
        ++slot;
        if (slot == NB || t + nw >= t1) {
            wave_lds_sync();  // pt and lv of every slot as lane 0 stored them
            if (lane < slot) {
                const cplx* ptj = slots + lane * 2 * PTN;
                const uint64_t tj = (uint64_t)(t_batch + (int64_t)lane * nw);
                const double pj = ((h - __popcll(tj)) & 1) ? -1.0 : 1.0;
                const E top = coeff_top_e<E, 4 * T, LOOP>(ptj, ptj + PTN, h);
                acc_sum.add(cmulr(etoc(top), pj));
            }
            wave_lds_sync();  // the slots are free again
            slot = 0;
            t_batch = t + nw;
        }
    }

    // fold the lanes' partial sums in lane order (montrealer: lane h holds them all), then the
    // waves of this workgroup in wave order
    if (mtl) {
        if (lane == h) wres[wave] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    } else {
        double4* lsum = reinterpret_cast<double4*>(slots);
        if (lane < NB) lsum[lane] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
        wave_lds_sync();
        if (lane == 0) {
            csum s;
            s.zero();
            for (int j = 0; j < NB; ++j) {
                double4 pq = lsum[j];
                s.add_pair(pq.x, pq.y, pq.z, pq.w);
            }
            wres[wave] = make_double4(s.sr, s.cr, s.si, s.ci);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        csum s;
        s.zero();
        for (int w = 0; w < nw; ++w) {
            double4 p = wres[w];
            s.add_pair(p.x, p.y, p.z, p.w);
        }
        partials[b] = make_double4(s.sr, s.cr, s.si, s.ci);
    }
}

// ---- Repeated rows: any edge multiplicities, even N, Glynn or inclusion-exclusion, on the same
// register-tiled products.  Term t has mixed-radix digits k_i (digit 0 most significant, base
// e_i + 1, _hafnian.py:162-180) and delta_i = 2 k_i - e_i (Glynn) or k_i.  The reference keeps
// the rows and columns of Ax X with delta_i != 0 and scales the columns by delta (AX_S,
// :288-312); here every term works on the padded B = M diag(d), M = Ax X, d_c = delta_(c mod E).
// A column with delta = 0 is zero, so under a permutation B is block triangular with AX_S as
// its kept block, and tr(B^k) = tr(AX_S^k); the loop weights XD_d = (X D) diag(d) vanish on the
// same indices, so XD_d B^(k-1) D = XD_S AX_S^(k-1) D_S.  The powers run to B^ceil(h/2) with
// h = N/2 as in the all-ones kernel (direct powers where the reference extends the traces by
// the characteristic polynomial, charpoly.py:321-326), and the term is
// (-1)^(h - sum k) prod C(e_i, k_i) (1/2 for Glynn when delta_0 = 0) [eta^h] exp(sum a_i eta^i)
// (_hafnian.py:444-460, :546-569).  Every job of a launch has 2E <= 8T, h <= H and fewer than
// 2^31 terms (the host checks).
__host__ __device__ constexpr int rep_nb() { return 8; }
__host__ __device__ constexpr int rep_ptn(int H) { return H + 4; }

template <int T, bool REAL, bool LOOP, int H>
__device__ __forceinline__ void rep_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                         double4* __restrict__ partials, int64_t b, unsigned char* smem) {
    using E = typename Elem<REAL>::T;
    constexpr int NP = 8 * T;
    constexpr int S = Lds<T>::S;
    constexpr int PTN = rep_ptn(H);
    constexpr int NB = rep_nb();
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    const int ti = lane >> 3, tj = lane & 7;
    const int ne = J.E;
    const int n = 2 * ne;
    const int h = J.N / 2;
    const bool glynn = (J.flags & WH_F_GLYNN) != 0;

    // ---- LDS carve: M, one power per wave, XD0 / D, per-wave term slots, per-wave results,
    // the binomial table C(e_i, k) (row i, k = 0..H)
    E* Ms = reinterpret_cast<E*>(smem);
    E* Pw = Ms + NP * S + wave * NP * S;
    E* vecs = Ms + (1 + nw) * NP * S;
    E* XD0 = vecs;
    E* Dv = vecs + NP;
    cplx* slots = reinterpret_cast<cplx*>(vecs + 2 * NP) + wave * (NB * 2 * PTN);
    double4* wres = reinterpret_cast<double4*>(reinterpret_cast<cplx*>(vecs + 2 * NP) + nw * (NB * 2 * PTN));
    double* btab = reinterpret_cast<double*>(wres + nw);

    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < NP * NP; e += blockDim.x) {
        int r = e / NP, c = e - r * NP;
        E v;
        ezero(v);
        if (r < n && c < n) {
            int src = c < ne ? c + ne : c - ne;
            v = eload(A, (int64_t)r * n + src, (E*)nullptr);
        }
        Ms[r * S + c] = v;
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < NP; e += blockDim.x) {
            E xd, dd;
            ezero(xd);
            ezero(dd);
            if (e < n) {
                int src = e < ne ? e + ne : e - ne;  // (X D)[e]
                xd = eload(Dg, src, (E*)nullptr);
                dd = eload(Dg, e, (E*)nullptr);
            }
            XD0[e] = xd;
            Dv[e] = dd;
        }
    }
    for (int e = threadIdx.x; e < ne * (H + 1); e += blockDim.x) {
        int i = e / (H + 1), k = e - i * (H + 1);
        btab[e] = binom_d(J.e[i], k);
    }
    __syncthreads();

    const int64_t t0 = (b - J.chunk_base) * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;
    const int I0 = ti * T, J0 = tj * T;
    // lane i < ne owns digit i: its base and the weight of the digits after it
    const int my_e = lane < ne ? J.e[lane] : 0;
    uint32_t my_w = 1;
    for (int l = lane + 1; l < ne; ++l) my_w *= (uint32_t)(J.e[l] + 1);
    // the edge index of column (and row) c, for c < n: c mod ne
    int srcJ[T], srcI[T];
#pragma unroll
    for (int c = 0; c < T; ++c) {
        int col = J0 + c, row = I0 + c;
        srcJ[c] = col < ne ? col : col < n ? col - ne : 0;
        srcI[c] = row < ne ? row : row < n ? row - ne : 0;
    }
    const int src_l = lane < ne ? lane : lane < n ? lane - ne : 0;

    csum acc_sum;
    acc_sum.zero();

    int slot = 0;
    for (int64_t t = t0 + wave; t < t1; t += nw) {
        const uint32_t tu = (uint32_t)t;
        // _hafnian.py:444-459: digits, sign, binomial product, delta
        const int k = lane < ne ? (int)((tu / my_w) % (uint32_t)(my_e + 1)) : 0;
        const int delta = lane < ne ? (glynn ? 2 * k - my_e : k) : 0;
        int ksum = k;
        double bin = lane < ne ? btab[lane * (H + 1) + k] : 1.0;
#pragma unroll
        for (int msk = 1; msk < 8; msk <<= 1) {  // ne <= 8: lanes 0..7 hold every digit
            ksum += __shfl_xor(ksum, msk, 8);
            bin *= __shfl_xor(bin, msk, 8);  // exact: a product of integers below 2^53
        }
        ksum = __shfl(ksum, 0, WH_WAVE);
        bin = __shfl(bin, 0, WH_WAVE);
        const int delta0 = __shfl(delta, 0, WH_WAVE);
        double prefac = (((h - ksum) & 1) ? -1.0 : 1.0) * bin;
        if (glynn && delta0 == 0) prefac *= 0.5;
        double sgnJ[T], sgnI[T];
#pragma unroll
        for (int c = 0; c < T; ++c) {
            const double dj = (double)__shfl(delta, srcJ[c], WH_WAVE);
            const double di = (double)__shfl(delta, srcI[c], WH_WAVE);
            sgnJ[c] = J0 + c < n ? dj : 0.0;
            sgnI[c] = I0 + c < n ? di : 0.0;
        }
        const double dl = (double)__shfl(delta, src_l, WH_WAVE);

        cplx* pt = slots + slot * 2 * PTN;  // pt[0] = (prefactor, 0); p[1..h]
        cplx* lv = pt + PTN;                // loop scalars v[1..h]
        E acc[T][T];
#pragma unroll
        for (int r = 0; r < T; ++r)
#pragma unroll
            for (int c = 0; c < T; ++c) acc[r][c] = escale(Ms[(I0 + r) * S + J0 + c], sgnJ[c]);
        auto load_xd = [&](E (&xdI)[T]) {
#pragma unroll
            for (int r = 0; r < T; ++r) xdI[r] = escale(XD0[I0 + r], sgnI[r]);
        };
        auto load_d = [&](E (&dJ)[T]) {
#pragma unroll
            for (int c = 0; c < T; ++c) dJ[c] = Dv[J0 + c];
        };
        // powers up to B^A, A = ceil(h/2): tr(B^(2a)) = <B^a, (B^a)^T>, tr(B^(2a+1)) =
        // <B^(a+1), (B^a)^T>; loop scalars v_(2a) = y_a . z_(a-1), v_(2a+1) = y_a . z_a with
        // y_a = XD_d B^a and z_a = B^a D (as in the all-ones kernel)
        const int Ah = (h + 1) / 2;
        E zprev[T];
        if (LOOP) {
#pragma unroll
            for (int r = 0; r < T; ++r) zprev[r] = Dv[I0 + r];
        }
        tile_store<T>(Pw, acc, I0, J0);
        for (int a = 1; a <= Ah; ++a) {
            E qo, qe;
            if (a == 1) {
                qo = tile_trace<T>(acc, ti == tj);
                qe = tile_tdot<T>(Pw, acc, I0, J0);
            } else {
                tile_product<T>(Pw, Ms, acc, I0, J0, n, sgnJ);
                qo = tile_tdot<T>(Pw, acc, I0, J0);
                tile_store<T>(Pw, acc, I0, J0);
                qe = tile_tdot<T>(Pw, acc, I0, J0);
            }
            // the power's results, summed over the wave eight at once: p_ko, p_ke, v_2a, v_(2a+1)
            E de, dodd;
            ezero(de);
            ezero(dodd);
            E znew[T];
            if (LOOP) {
                E y[T], xdI[T], dJ[T];
                load_xd(xdI);
                load_d(dJ);
                tile_yvec<T>(acc, xdI, y);
                tile_zvec<T>(acc, dJ, znew);
                if (ti == tj) {
#pragma unroll
                    for (int c = 0; c < T; ++c) {
                        emac(de, y[c], zprev[c]);
                        emac(dodd, y[c], znew[c]);
                    }
                }
            }
            const cplx qoc = etoc(qo), qec = etoc(qe), dec = etoc(de), doc = etoc(dodd);
            const double v8[8] = {qoc.re, qoc.im, qec.re, qec.im, dec.re, dec.im, doc.re, doc.im};
            const double sum8 = wave_sum8(v8, lane);
            const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
            if ((lane & 7) == 0) {  // lane 8k holds the sum of v8[k]
                const int k = lane >> 3;
                double* dst = nullptr;
                if (k < 2) dst = &pt[ko].re;
                else if (k < 4) dst = &pt[ke].re;  // ke <= h + 1 < PTN
                else if (LOOP && k < 6 && 2 * a <= h) dst = &lv[2 * a].re;
                else if (LOOP && k >= 6 && 2 * a + 1 <= h) dst = &lv[2 * a + 1].re;
                if (dst) dst[k & 1] = sum8;
            }
            if (LOOP) {
#pragma unroll
                for (int r = 0; r < T; ++r) zprev[r] = znew[r];
            }
        }
        if (LOOP) {
            // v_1 = XD_d . D (n <= 64: lane e holds index e)
            E part;
            ezero(part);
            if (lane < n) emac(part, escale(XD0[lane], dl), Dv[lane]);
            cplx v1 = wave_csum(etoc(part));
            if (lane == 0) lv[1] = v1;
        }
        if (lane == 0) pt[0] = cmk(prefac, 0.0);
        wave_lds_sync();  // pt and lv as lane 0 stored them

        // the coefficients of up to NB terms together, lane j for the batch's j-th term
        ++slot;
        if (slot == NB || t + nw >= t1) {
            wave_lds_sync();
            if (lane < slot) {
                cplx* ptj = slots + lane * 2 * PTN;
                const E top = coeff_top_lds<E, H, LOOP>(ptj, ptj + PTN, h);
                acc_sum.add(cmulr(etoc(top), ptj[0].re));
            }
            wave_lds_sync();  // the slots are free again
            slot = 0;
        }
    }

    double4* lsum = reinterpret_cast<double4*>(slots);
    if (lane < NB) lsum[lane] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    wave_lds_sync();
    if (lane == 0) {
        csum s;
        s.zero();
        for (int j = 0; j < NB; ++j) {
            double4 pq = lsum[j];
            s.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        wres[wave] = make_double4(s.sr, s.cr, s.si, s.ci);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        csum s;
        s.zero();
        for (int w = 0; w < nw; ++w) {
            double4 p = wres[w];
            s.add_pair(p.x, p.y, p.z, p.w);
        }
        partials[b] = make_double4(s.sr, s.cr, s.si, s.ci);
    }
}

// ---- Repeated rows with 2E <= 12, four terms per wave.  One term on a whole wave (rep_body)
// pads 2E = 10 or 12 to 16 and spends most of its instructions on per-power reductions over 64
// lanes and on coefficient batches that leave 56 lanes idle.  Here each 16-lane group of the wave
// takes its own term: the group is a 4 x 4 grid of lanes, each owning a 3 x 3 tile of the 12 x 12
// padded power (its own matrix in LDS), so the products pad to 12, and the per-power sums run
// within the 16 lanes for the four terms at once.  Everything else (digits, delta, prefactor,
// the half-products powers, loop vectors, coefficient) is rep_body's, per group.
#define WH_DPP_ROR4 0x124  // lane i <- i + 4 within 16
constexpr int kR4N = 12;      // padded side
constexpr int kR4S = 13;      // LDS row stride (elements)
constexpr int kR4Iters = 2;   // wave iterations (of four terms) per coefficient batch
constexpr int kR4Slots = 4 * kR4Iters;

__device__ __forceinline__ double group16_sum(double v) {
    v += dpp_d<WH_DPP_XOR1>(v);
    v += dpp_d<WH_DPP_XOR2>(v);
    v += dpp_d<WH_DPP_HALF_MIRROR>(v);
    return v + dpp_d<WH_DPP_ROR8>(v);
}
__device__ __forceinline__ double col4_sum(double v) {  // lanes i, i + 4, i + 8, i + 12 of the 16
    v += dpp_d<WH_DPP_ROR4>(v);
    return v + dpp_d<WH_DPP_ROR8>(v);
}
__device__ __forceinline__ double row4_sum(double v) {  // lanes 4q .. 4q + 3 of the 16
    v += dpp_d<WH_DPP_XOR1>(v);
    return v + dpp_d<WH_DPP_XOR2>(v);
}
__device__ __forceinline__ cplx group16_e(cplx v) { return cmk(group16_sum(v.re), group16_sum(v.im)); }
__device__ __forceinline__ double group16_e(double v) { return group16_sum(v); }
__device__ __forceinline__ cplx col4_e(cplx v) { return cmk(col4_sum(v.re), col4_sum(v.im)); }
__device__ __forceinline__ double col4_e(double v) { return col4_sum(v); }
__device__ __forceinline__ cplx row4_e(cplx v) { return cmk(row4_sum(v.re), row4_sum(v.im)); }
__device__ __forceinline__ double row4_e(double v) { return row4_sum(v); }

// Eight sums over each 16-lane group at once: lane i returns the group's sum of v[(i >> 1) & 7].
// Each stage halves the values a lane carries: partners i ^ 8 (bit 3 picks the half kept),
// 7 - i (bit 2), i ^ 2 (bit 1), then i ^ 1; the 16 partner masks are all distinct.
__device__ __forceinline__ double group16_sum8(const double (&v)[8], int gl) {
    const bool b3 = (gl >> 3) & 1, b2 = (gl >> 2) & 1, b1 = (gl >> 1) & 1;
    double s1[4], s2[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double keep = b3 ? v[k + 4] : v[k], send = b3 ? v[k] : v[k + 4];
        s1[k] = keep + dpp_d<WH_DPP_ROR8>(send);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double keep = b2 ? s1[k + 2] : s1[k], send = b2 ? s1[k] : s1[k + 2];
        s2[k] = keep + dpp_d<WH_DPP_HALF_MIRROR>(send);
    }
    const double keep = b1 ? s2[1] : s2[0], send = b1 ? s2[0] : s2[1];
    const double s3 = keep + dpp_d<WH_DPP_XOR2>(send);
    return s3 + dpp_d<WH_DPP_XOR1>(s3);
}

// LDS layout (element type E): M, XD0, D, then per wave four powers, then per wave the term slots
// (complex), the per-wave results and the binomial table.
__host__ __device__ constexpr size_t rep4_align32(size_t x) { return (x + 31) / 32 * 32; }
__host__ __device__ constexpr size_t rep4_slots_off(size_t esz, int nw) {
    return rep4_align32(((size_t)kR4N * kR4S + 2 * kR4N + (size_t)nw * 4 * kR4N * kR4S) * esz);
}
__host__ __device__ constexpr size_t rep4_wres_off(size_t esz, int nw, int H) {
    return rep4_align32(rep4_slots_off(esz, nw) + (size_t)nw * kR4Slots * 2 * rep_ptn(H) * sizeof(cplx));
}
__host__ __device__ constexpr size_t rep4_lds_bytes(size_t esz, int nw, int H) {
    return rep4_wres_off(esz, nw, H) + (size_t)nw * sizeof(double4) + (size_t)(kR4N / 2) * (H + 1) * sizeof(double);
}

template <bool REAL, bool LOOP, int H>
__device__ __forceinline__ void rep4_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                          double4* __restrict__ partials, int64_t b, unsigned char* smem) {
    using E = typename Elem<REAL>::T;
    constexpr int N = kR4N, S = kR4S, PTN = rep_ptn(H), NI = kR4Iters;
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    const int grp = lane >> 4, gl = lane & 15, gbase = lane & ~15;
    const int ti = gl >> 2, tj = gl & 3;
    const int I0 = 3 * ti, J0 = 3 * tj;
    const int ne = J.E;
    const int n = 2 * ne;
    const int h = J.N / 2;
    const bool glynn = (J.flags & WH_F_GLYNN) != 0;

    E* Ms = reinterpret_cast<E*>(smem);
    E* XD0 = Ms + N * S;
    E* Dv = XD0 + N;
    E* Pw = Dv + N + (wave * 4 + grp) * N * S;  // this group's power
    cplx* slots = reinterpret_cast<cplx*>(smem + rep4_slots_off(sizeof(E), nw)) + wave * kR4Slots * 2 * PTN;
    double4* wres = reinterpret_cast<double4*>(smem + rep4_wres_off(sizeof(E), nw, H));
    double* btab = reinterpret_cast<double*>(wres + nw);

    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        int r = e / N, c = e - r * N;
        E v;
        ezero(v);
        if (r < n && c < n) {
            int src = c < ne ? c + ne : c - ne;
            v = eload(A, (int64_t)r * n + src, (E*)nullptr);
        }
        Ms[r * S + c] = v;
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < N; e += blockDim.x) {
            E xd, dd;
            ezero(xd);
            ezero(dd);
            if (e < n) {
                int src = e < ne ? e + ne : e - ne;  // (X D)[e]
                xd = eload(Dg, src, (E*)nullptr);
                dd = eload(Dg, e, (E*)nullptr);
            }
            XD0[e] = xd;
            Dv[e] = dd;
        }
    }
    for (int e = threadIdx.x; e < ne * (H + 1); e += blockDim.x) {
        int i = e / (H + 1), k = e - i * (H + 1);
        btab[e] = binom_d(J.e[i], k);
    }
    __syncthreads();

    const int64_t t0 = (b - J.chunk_base) * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;
    // lane gl < ne of each group owns digit gl: its base and the weight of the digits after it
    const int my_e = gl < ne ? J.e[gl] : 0;
    uint32_t my_w = 1;
    for (int l = gl + 1; l < ne; ++l) my_w *= (uint32_t)(J.e[l] + 1);
    int srcJ[3], srcI[3];  // the edge index of the lane's columns and rows (c mod ne; 0 on padding)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int col = J0 + c, row = I0 + c;
        srcJ[c] = col < ne ? col : col < n ? col - ne : 0;
        srcI[c] = row < ne ? row : row < n ? row - ne : 0;
    }
    const int src_l = gl < ne ? gl : gl < n ? gl - ne : 0;

    csum acc_sum;
    acc_sum.zero();

    int it = 0;  // iteration within the coefficient batch
    for (int64_t t = t0 + 4 * wave; t < t1; t += 4 * nw) {
        // group grp takes term t + grp; past the chunk's end it repeats term t with prefactor 0
        const bool valid = t + grp < t1;
        const uint32_t tu = (uint32_t)(valid ? t + grp : t);
        const int k = gl < ne ? (int)((tu / my_w) % (uint32_t)(my_e + 1)) : 0;
        const int delta = gl < ne ? (glynn ? 2 * k - my_e : k) : 0;
        int ksum = k;
        double bin = gl < ne ? btab[gl * (H + 1) + k] : 1.0;
#pragma unroll
        for (int msk = 1; msk < 8; msk <<= 1) {  // ne <= 6: lanes 0..7 of the group hold every digit
            ksum += __shfl_xor(ksum, msk, 8);
            bin *= __shfl_xor(bin, msk, 8);  // exact: a product of integers below 2^53
        }
        ksum = __shfl(ksum, gbase, WH_WAVE);
        bin = __shfl(bin, gbase, WH_WAVE);
        const int delta0 = __shfl(delta, gbase, WH_WAVE);
        double prefac = (((h - ksum) & 1) ? -1.0 : 1.0) * bin;
        if (glynn && delta0 == 0) prefac *= 0.5;
        if (!valid) prefac = 0.0;
        double sgnJ[3], sgnI[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double dj = (double)__shfl(delta, gbase + srcJ[c], WH_WAVE);
            const double di = (double)__shfl(delta, gbase + srcI[c], WH_WAVE);
            sgnJ[c] = J0 + c < n ? dj : 0.0;
            sgnI[c] = I0 + c < n ? di : 0.0;
        }
        const double dl = (double)__shfl(delta, gbase + src_l, WH_WAVE);

        cplx* pt = slots + (it * 4 + grp) * 2 * PTN;  // pt[0] = (prefactor, 0); p[1..h]
        cplx* lv = pt + PTN;                          // loop scalars v[1..h]
        E acc[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[r][c] = escale(Ms[(I0 + r) * S + J0 + c], sgnJ[c]);
        E zprev[3];
        if (LOOP) {
#pragma unroll
            for (int r = 0; r < 3; ++r) zprev[r] = Dv[I0 + r];
        }
        auto store = [&]() {
            wave_lds_sync();  // the group's reads of the previous power are done
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) Pw[(I0 + r) * S + J0 + c] = acc[r][c];
            wave_lds_sync();  // the whole power is stored before any lane reads it
        };
        auto tdot = [&]() {  // lane part of <acc, Pw^T>
            E sd;
            ezero(sd);
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) emac(sd, acc[r][c], Pw[(J0 + c) * S + I0 + r]);
            return sd;
        };
        store();
        const int Ah = (h + 1) / 2;
        for (int a = 1; a <= Ah; ++a) {
            E qo, qe;
            if (a == 1) {
                ezero(qo);
                if (ti == tj) {
#pragma unroll
                    for (int r = 0; r < 3; ++r) qo = eadd(qo, acc[r][r]);
                }
                qe = tdot();
            } else {
                // acc = (Pw M) diag(delta), k summed in order; next k's operands load during k's FMAs
                E nacc[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) ezero(nacc[r][c]);
                E av[3], bv[3], an[3], bn[3];
#pragma unroll
                for (int r = 0; r < 3; ++r) av[r] = Pw[(I0 + r) * S];
#pragma unroll
                for (int c = 0; c < 3; ++c) bv[c] = Ms[J0 + c];
                for (int kk = 0; kk < n; ++kk) {
                    const int kn = kk + 1 < n ? kk + 1 : kk;
#pragma unroll
                    for (int r = 0; r < 3; ++r) an[r] = Pw[(I0 + r) * S + kn];
#pragma unroll
                    for (int c = 0; c < 3; ++c) bn[c] = Ms[kn * S + J0 + c];
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) emac(nacc[r][c], av[r], bv[c]);
#pragma unroll
                    for (int r = 0; r < 3; ++r) av[r] = an[r];
#pragma unroll
                    for (int c = 0; c < 3; ++c) bv[c] = bn[c];
                }
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) acc[r][c] = escale(nacc[r][c], sgnJ[c]);
                qo = tdot();
                store();
                qe = tdot();
            }
            E de, dodd;
            ezero(de);
            ezero(dodd);
            E znew[3];
            if (LOOP) {
                // y = XD_d B^a (columns, over the group's 4 row tiles), z = B^a D (rows, over its
                // 4 column tiles); the dots pair y and z on the diagonal tiles
                E y[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ezero(y[c]);
#pragma unroll
                    for (int r = 0; r < 3; ++r) emac(y[c], escale(XD0[I0 + r], sgnI[r]), acc[r][c]);
                    y[c] = col4_e(y[c]);
                }
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    ezero(znew[r]);
#pragma unroll
                    for (int c = 0; c < 3; ++c) emac(znew[r], acc[r][c], Dv[J0 + c]);
                    znew[r] = row4_e(znew[r]);
                }
                if (ti == tj) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        emac(de, y[c], zprev[c]);
                        emac(dodd, y[c], znew[c]);
                    }
                }
            }
            const cplx qoc = etoc(qo), qec = etoc(qe), dec = etoc(de), doc = etoc(dodd);
            const double v8[8] = {qoc.re, qoc.im, qec.re, qec.im, dec.re, dec.im, doc.re, doc.im};
            const double sum8 = group16_sum8(v8, gl);
            const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
            if ((gl & 1) == 0) {  // lane 2k of the group holds the sum of v8[k]
                const int kv = gl >> 1;
                double* dst = nullptr;
                if (kv < 2) dst = &pt[ko].re;
                else if (kv < 4) dst = &pt[ke].re;  // ke <= h + 1 < PTN
                else if (LOOP && kv < 6 && 2 * a <= h) dst = &lv[2 * a].re;
                else if (LOOP && kv >= 6 && 2 * a + 1 <= h) dst = &lv[2 * a + 1].re;
                if (dst) dst[kv & 1] = sum8;
            }
            if (LOOP) {
#pragma unroll
                for (int r = 0; r < 3; ++r) zprev[r] = znew[r];
            }
        }
        if (LOOP) {
            // v_1 = XD_d . D (lane gl < n of the group holds index gl)
            E part;
            ezero(part);
            if (gl < n) emac(part, escale(XD0[gl], dl), Dv[gl]);
            const cplx v1 = etoc(group16_e(part));
            if (gl == 0) lv[1] = v1;
        }
        if (gl == 0) pt[0] = cmk(prefac, 0.0);
        ++it;
        if (it == NI || t + 4 * nw >= t1) {
            wave_lds_sync();  // every group's slots as their lanes stored them
            if (lane < 4 * it) {  // lane j: slot j (iteration j / 4, group j % 4)
                cplx* ptj = slots + lane * 2 * PTN;
                const double pf = ptj[0].re;
                if (pf != 0.0) {
                    const E top = coeff_top_lds<E, H, LOOP>(ptj, ptj + PTN, h);
                    acc_sum.add(cmulr(etoc(top), pf));
                }
            }
            wave_lds_sync();  // the slots are free again
            it = 0;
        }
    }

    double4* lsum = reinterpret_cast<double4*>(slots);
    if (lane < kR4Slots) lsum[lane] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    wave_lds_sync();
    if (lane == 0) {
        csum sm;
        sm.zero();
        for (int j = 0; j < kR4Slots; ++j) {
            double4 pq = lsum[j];
            sm.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        wres[wave] = make_double4(sm.sr, sm.cr, sm.si, sm.ci);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        csum sm;
        sm.zero();
        for (int w = 0; w < nw; ++w) {
            double4 pq = wres[w];
            sm.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        partials[b] = make_double4(sm.sr, sm.cr, sm.si, sm.ci);
    }
}

// ---- The same four terms per wave with the products on the FP64 matrix cores (WH_REP4_MFMA,
// on by default).  v_mfma_f64_4x4x4f64 forms four independent 4 x 4 x 4 block products, block b
// on the lanes with (lane >> 2) & 3 == b: with x = lane & 3 and y = lane >> 4, a lane supplies
// A[x][y] and B[y][x] of its block and receives D[y][x].  Term g of the wave takes block g, so
// one instruction advances the four terms at once.  A term's 12 x 12 power is 3 x 3 blocks of
// 4 x 4, and lane (x, y) of block g holds P[4 bi + y][4 bj + x] for every (bi, bj); a product is
// 3 k chunks x 9 blocks, each complex block product three real MFMAs in Gauss's form
// (Re = X - Y, Im = Z - X - Y with X = Re a Re b, Y = Im a Im b, Z = (Re a + Im a)(Re b + Im b)),
// one for real inputs.  Sums over a term's 16 lanes run over lane bits 0, 1 (DPP) and 4, 5
// (permlane swaps).  Digits, delta, prefactor, powers, loop dots and the coefficient are
// rep4_body's.
#ifndef WH_REP4_MFMA
#define WH_REP4_MFMA 1
#endif

__device__ __forceinline__ double r4m_mfma(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
// sum over the 16 lanes of a term (lane bits 0, 1, 4, 5)
__device__ __forceinline__ double term16_sum(double v) {
    v += dpp_d<WH_DPP_XOR1>(v);
    v += dpp_d<WH_DPP_XOR2>(v);
    v = swap16_sum(v);
    return swap32_sum(v);
}
__device__ __forceinline__ cplx term16_e(cplx v) { return cmk(term16_sum(v.re), term16_sum(v.im)); }
__device__ __forceinline__ double term16_e(double v) { return term16_sum(v); }
// over lane bits 4, 5 (the lanes of a term with the same x: one column)
__device__ __forceinline__ cplx ycol_e(cplx v) { return cmk(swap32_sum(swap16_sum(v.re)), swap32_sum(swap16_sum(v.im))); }
__device__ __forceinline__ double ycol_e(double v) { return swap32_sum(swap16_sum(v)); }
// over lane bits 0, 1 (the lanes of a term with the same y: one row)
__device__ __forceinline__ cplx xrow_e(cplx v) { return row4_e(v); }
__device__ __forceinline__ double xrow_e(double v) { return row4_e(v); }

// Eight sums over each term's 16 lanes at once: lane l returns the sum of v[((l >> 1) & 1) |
// (((l >> 4) & 3) << 1)] (halving stages over lane bits 5, 4, 1, then 0).
__device__ __forceinline__ double term16_sum8(const double (&v)[8], int lane) {
    double s1[4], s2[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) s1[k] = dswap32(v[k], v[k + 4]);  // value k + 4 bit5
#pragma unroll
    for (int k = 0; k < 2; ++k) s2[k] = dswap16(s1[k], s1[k + 2]);  // value k + 2 bit4 + 4 bit5
    const bool b1 = (lane >> 1) & 1;
    const double keep = b1 ? s2[1] : s2[0], send = b1 ? s2[0] : s2[1];
    const double s3 = keep + dpp_d<WH_DPP_XOR2>(send);  // value bit1 + 2 bit4 + 4 bit5
    return s3 + dpp_d<WH_DPP_XOR1>(s3);
}

template <bool REAL, bool LOOP, int H>
__device__ __forceinline__ void rep4m_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                           double4* __restrict__ partials, int64_t b, unsigned char* smem) {
    using E = typename Elem<REAL>::T;
    constexpr int N = kR4N, S = kR4S, PTN = rep_ptn(H), NI = kR4Iters;
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    const int grp = (lane >> 2) & 3;  // the term (MFMA block) of the lane
    const int x = lane & 3, y = lane >> 4;
    const int gl = x + 4 * y;         // the lane's index within its term
    const int gbase = 4 * grp;        // the term's lane with gl = 0
    const int ne = J.E;
    const int n = 2 * ne;
    const int h = J.N / 2;
    const bool glynn = (J.flags & WH_F_GLYNN) != 0;

    E* Ms = reinterpret_cast<E*>(smem);
    E* XD0 = Ms + N * S;
    E* Dv = XD0 + N;
    E* Pw = Dv + N + (wave * 4 + grp) * N * S;  // this term's power
    cplx* slots = reinterpret_cast<cplx*>(smem + rep4_slots_off(sizeof(E), nw)) + wave * kR4Slots * 2 * PTN;
    double4* wres = reinterpret_cast<double4*>(smem + rep4_wres_off(sizeof(E), nw, H));
    double* btab = reinterpret_cast<double*>(wres + nw);

    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        int r = e / N, c = e - r * N;
        E v;
        ezero(v);
        if (r < n && c < n) {
            int src = c < ne ? c + ne : c - ne;
            v = eload(A, (int64_t)r * n + src, (E*)nullptr);
        }
        Ms[r * S + c] = v;
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < N; e += blockDim.x) {
            E xd, dd;
            ezero(xd);
            ezero(dd);
            if (e < n) {
                int src = e < ne ? e + ne : e - ne;  // (X D)[e]
                xd = eload(Dg, src, (E*)nullptr);
                dd = eload(Dg, e, (E*)nullptr);
            }
            XD0[e] = xd;
            Dv[e] = dd;
        }
    }
    for (int e = threadIdx.x; e < ne * (H + 1); e += blockDim.x) {
        int i = e / (H + 1), k = e - i * (H + 1);
        btab[e] = binom_d(J.e[i], k);
    }
    __syncthreads();

    const int64_t t0 = (b - J.chunk_base) * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;
    // lane gl < ne of each term owns digit gl: its base and the weight of the digits after it
    const int my_e = gl < ne ? J.e[gl] : 0;
    uint32_t my_w = 1;
    for (int l = gl + 1; l < ne; ++l) my_w *= (uint32_t)(J.e[l] + 1);
    // the lane of the term holding digit d (gl = d)
    auto dlane = [&](int d) { return gbase + (d & 3) + 16 * (d >> 2); };
    int srcJ[3], srcI[3];  // the edge index of the lane's columns 4 bj + x and rows 4 bi + y
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int col = 4 * c + x, row = 4 * c + y;
        srcJ[c] = col < ne ? col : col < n ? col - ne : 0;
        srcI[c] = row < ne ? row : row < n ? row - ne : 0;
    }
    const int src_l = gl < ne ? gl : gl < n ? gl - ne : 0;

    csum acc_sum;
    acc_sum.zero();

    int it = 0;
    for (int64_t t = t0 + 4 * wave; t < t1; t += 4 * nw) {
        const bool valid = t + grp < t1;
        const uint32_t tu = (uint32_t)(valid ? t + grp : t);
        const int k = gl < ne ? (int)((tu / my_w) % (uint32_t)(my_e + 1)) : 0;
        const int delta = gl < ne ? (glynn ? 2 * k - my_e : k) : 0;
        int ksum = k;
        double bin = gl < ne ? btab[gl * (H + 1) + k] : 1.0;
        // over the term's 16 lanes (lane bits 0, 1, 4, 5): every lane receives the total
        ksum += __shfl_xor(ksum, 1, WH_WAVE);
        bin *= __shfl_xor(bin, 1, WH_WAVE);
        ksum += __shfl_xor(ksum, 2, WH_WAVE);
        bin *= __shfl_xor(bin, 2, WH_WAVE);
        ksum += __shfl_xor(ksum, 16, WH_WAVE);
        bin *= __shfl_xor(bin, 16, WH_WAVE);
        ksum += __shfl_xor(ksum, 32, WH_WAVE);
        bin *= __shfl_xor(bin, 32, WH_WAVE);  // exact: a product of integers below 2^53
        const int delta0 = __shfl(delta, gbase, WH_WAVE);
        double prefac = (((h - ksum) & 1) ? -1.0 : 1.0) * bin;
        if (glynn && delta0 == 0) prefac *= 0.5;
        if (!valid) prefac = 0.0;
        double sgnJ[3], sgnI[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double dj = (double)__shfl(delta, dlane(srcJ[c]), WH_WAVE);
            const double di = (double)__shfl(delta, dlane(srcI[c]), WH_WAVE);
            sgnJ[c] = 4 * c + x < n ? dj : 0.0;
            sgnI[c] = 4 * c + y < n ? di : 0.0;
        }
        const double dl = (double)__shfl(delta, dlane(src_l), WH_WAVE);

        cplx* pt = slots + (it * 4 + grp) * 2 * PTN;
        cplx* lv = pt + PTN;
        E acc[3][3];
#pragma unroll
        for (int bi = 0; bi < 3; ++bi)
#pragma unroll
            for (int bj = 0; bj < 3; ++bj) acc[bi][bj] = escale(Ms[(4 * bi + y) * S + 4 * bj + x], sgnJ[bj]);
        E zprev[3], xdI[3];
        if (LOOP) {
#pragma unroll
            for (int bi = 0; bi < 3; ++bi) {
                zprev[bi] = Dv[4 * bi + y];
                xdI[bi] = escale(XD0[4 * bi + y], sgnI[bi]);
            }
        }
        auto store = [&]() {
            wave_lds_sync();  // the term's reads of the previous power are done
#pragma unroll
            for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                for (int bj = 0; bj < 3; ++bj) Pw[(4 * bi + y) * S + 4 * bj + x] = acc[bi][bj];
            wave_lds_sync();  // the whole power is stored before any lane reads it
        };
        auto tdot = [&]() {  // lane part of <acc, Pw^T>
            E sd;
            ezero(sd);
#pragma unroll
            for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                for (int bj = 0; bj < 3; ++bj) emac(sd, acc[bi][bj], Pw[(4 * bj + x) * S + 4 * bi + y]);
            return sd;
        };
        store();
        const int Ah = (h + 1) / 2;
        for (int a = 1; a <= Ah; ++a) {
            E qo, qe;
            if (a == 1) {
                ezero(qo);
                if (x == y) {
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb) qo = eadd(qo, acc[bb][bb]);
                }
                qe = tdot();
            } else {
                // acc = (Pw M) diag(delta) on the matrix cores, the four terms at once
                double X[3][3], Y[3][3], Z[3][3];
#pragma unroll
                for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 3; ++bj) X[bi][bj] = Y[bi][bj] = Z[bi][bj] = 0.0;
#pragma unroll
                for (int kc = 0; kc < 3; ++kc) {
                    E av[3], bv[3];
#pragma unroll
                    for (int bi = 0; bi < 3; ++bi) av[bi] = Pw[(4 * bi + x) * S + 4 * kc + y];
#pragma unroll
                    for (int bj = 0; bj < 3; ++bj) bv[bj] = Ms[(4 * kc + y) * S + 4 * bj + x];
                    if constexpr (REAL) {
#pragma unroll
                        for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 3; ++bj) X[bi][bj] = r4m_mfma(av[bi], bv[bj], X[bi][bj]);
                    } else {
                        double as[3], bs[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            as[c] = av[c].re + av[c].im;
                            bs[c] = bv[c].re + bv[c].im;
                        }
#pragma unroll
                        for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 3; ++bj) {
                                X[bi][bj] = r4m_mfma(av[bi].re, bv[bj].re, X[bi][bj]);
                                Y[bi][bj] = r4m_mfma(av[bi].im, bv[bj].im, Y[bi][bj]);
                                Z[bi][bj] = r4m_mfma(as[bi], bs[bj], Z[bi][bj]);
                            }
                    }
                }
#pragma unroll
                for (int bi = 0; bi < 3; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 3; ++bj) {
                        if constexpr (REAL) {
                            acc[bi][bj] = escale(eget(cmk(X[bi][bj], 0.0), (E*)nullptr), sgnJ[bj]);
                        } else {
                            const double xx = X[bi][bj], yy = Y[bi][bj];
                            acc[bi][bj] = escale(eget(cmk(xx - yy, Z[bi][bj] - (xx + yy)), (E*)nullptr), sgnJ[bj]);
                        }
                    }
                qo = tdot();
                store();
                qe = tdot();
            }
            E de, dodd;
            ezero(de);
            ezero(dodd);
            E znew[3];
            if (LOOP) {
                // y = XD_d B^a (the lane's columns 4 bj + x, over its term's rows), z = B^a D (its
                // rows 4 bi + y, over the columns); the dots pair them on the lanes with x == y
                E yv[3];
#pragma unroll
                for (int bj = 0; bj < 3; ++bj) {
                    ezero(yv[bj]);
#pragma unroll
                    for (int bi = 0; bi < 3; ++bi) emac(yv[bj], xdI[bi], acc[bi][bj]);
                    yv[bj] = ycol_e(yv[bj]);
                }
#pragma unroll
                for (int bi = 0; bi < 3; ++bi) {
                    ezero(znew[bi]);
#pragma unroll
                    for (int bj = 0; bj < 3; ++bj) emac(znew[bi], acc[bi][bj], Dv[4 * bj + x]);
                    znew[bi] = xrow_e(znew[bi]);
                }
                if (x == y) {
#pragma unroll
                    for (int bb = 0; bb < 3; ++bb) {
                        emac(de, yv[bb], zprev[bb]);
                        emac(dodd, yv[bb], znew[bb]);
                    }
                }
            }
            const cplx qoc = etoc(qo), qec = etoc(qe), dec = etoc(de), doc = etoc(dodd);
            const double v8[8] = {qoc.re, qoc.im, qec.re, qec.im, dec.re, dec.im, doc.re, doc.im};
            const double sum8 = term16_sum8(v8, lane);
            const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
            if ((lane & 1) == 0) {  // the lane with bit 0 clear of each value's pair writes it
                const int kv = ((lane >> 1) & 1) | (((lane >> 4) & 3) << 1);
                double* dst = nullptr;
                if (kv < 2) dst = &pt[ko].re;
                else if (kv < 4) dst = &pt[ke].re;  // ke <= h + 1 < PTN
                else if (LOOP && kv < 6 && 2 * a <= h) dst = &lv[2 * a].re;
                else if (LOOP && kv >= 6 && 2 * a + 1 <= h) dst = &lv[2 * a + 1].re;
                if (dst) dst[kv & 1] = sum8;
            }
            if (LOOP) {
#pragma unroll
                for (int bi = 0; bi < 3; ++bi) zprev[bi] = znew[bi];
            }
        }
        if (LOOP) {
            // v_1 = XD_d . D (lane gl < n of the term holds index gl)
            E part;
            ezero(part);
            if (gl < n) emac(part, escale(XD0[gl], dl), Dv[gl]);
            const cplx v1 = etoc(term16_e(part));
            if (gl == 0) lv[1] = v1;
        }
        if (gl == 0) pt[0] = cmk(prefac, 0.0);
        ++it;
        if (it == NI || t + 4 * nw >= t1) {
            wave_lds_sync();  // every term's slots as their lanes stored them
            if (lane < 4 * it) {  // lane j: slot j (iteration j / 4, term j % 4)
                cplx* ptj = slots + lane * 2 * PTN;
                const double pf = ptj[0].re;
                if (pf != 0.0) {
                    const E top = coeff_top_lds<E, H, LOOP>(ptj, ptj + PTN, h);
                    acc_sum.add(cmulr(etoc(top), pf));
                }
            }
            wave_lds_sync();  // the slots are free again
            it = 0;
        }
    }

    double4* lsum = reinterpret_cast<double4*>(slots);
    if (lane < kR4Slots) lsum[lane] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    wave_lds_sync();
    if (lane == 0) {
        csum sm;
        sm.zero();
        for (int j = 0; j < kR4Slots; ++j) {
            double4 pq = lsum[j];
            sm.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        wres[wave] = make_double4(sm.sr, sm.cr, sm.si, sm.ci);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        csum sm;
        sm.zero();
        for (int w = 0; w < nw; ++w) {
            double4 pq = wres[w];
            sm.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        partials[b] = make_double4(sm.sr, sm.cr, sm.si, sm.ci);
    }
}

}  // namespace

template <int T, bool REAL, bool LOOP, bool HALF>
__global__ void __launch_bounds__(256)
wh_sieve_pm1(const wh_job_dev* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
             double4* __restrict__ partials, int64_t block_offset) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    pm1_body<T, REAL, LOOP, HALF>(jobs[lo], pool, partials, b, wh_smem);
}

size_t wh_pm1_lds_bytes(int T, bool real, int nw) {
    size_t esz = real ? sizeof(double) : sizeof(cplx);
    size_t NP = 8 * (size_t)T, S = NP + WH_PM1_PAD;
    size_t mat = (NP * S * esz + 31) / 32 * 32;  // keeps the later regions 32-byte aligned
    return (1 + nw) * mat + 2 * NP * esz + nw * (size_t)(pm1_nb(T, real) * 2 * pm1_ptn(T)) * sizeof(cplx) +
           nw * sizeof(double4) + 32;
}

template <int T, bool REAL, bool LOOP, bool HALF>
static hipError_t launch_t(const wh_job_dev* jobs, int njobs, const void* pool, void* partials,
                           int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = wh_pm1_lds_bytes(T, REAL, nw);
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_pm1<T, REAL, LOOP, HALF>, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_pm1<T, REAL, LOOP, HALF>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds,
                       stream, jobs, njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

template <int T, bool HALF>
static hipError_t launch_rl(bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                            void* partials, int64_t nblocks, int64_t off, int nw, hipStream_t s) {
    if (real) {
        if (loop) return launch_t<T, true, true, HALF>(jobs, njobs, pool, partials, nblocks, off, nw, s);
        return launch_t<T, true, false, HALF>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    }
    if (loop) return launch_t<T, false, true, HALF>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    return launch_t<T, false, false, HALF>(jobs, njobs, pool, partials, nblocks, off, nw, s);
}

template <int T>
static hipError_t launch_h(bool half, bool real, bool loop, const wh_job_dev* jobs, int njobs,
                           const void* pool, void* partials, int64_t nblocks, int64_t off, int nw,
                           hipStream_t s) {
    if (half) return launch_rl<T, true>(real, loop, jobs, njobs, pool, partials, nblocks, off, nw, s);
    return launch_rl<T, false>(real, loop, jobs, njobs, pool, partials, nblocks, off, nw, s);
}

hipError_t wh_launch_pm1(int T, bool half, bool real, bool loop, const wh_job_dev* jobs, int njobs,
                         const void* pool, void* partials, int64_t nblocks, int64_t block_offset,
                         int nw, hipStream_t stream) {
    switch (T) {
        case 1: return launch_h<1>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 2: return launch_h<2>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 3: return launch_h<3>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 4: return launch_h<4>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 5: return launch_h<5>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 6: return launch_h<6>(half, real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        default: return hipErrorInvalidValue;
    }
}


// Four waves per SIMD: the complex H = 16 instances then keep part of the coefficient state in
// scratch (about 190 bytes per lane) instead of running two or three waves per SIMD.
#ifndef WH_REP_WAVES_PER_EU
#define WH_REP_WAVES_PER_EU 4
#endif
template <int T, bool REAL, bool LOOP, int H>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WH_REP_WAVES_PER_EU)))
wh_sieve_rep(const wh_job_dev* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
             double4* __restrict__ partials, int64_t block_offset) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    rep_body<T, REAL, LOOP, H>(jobs[lo], pool, partials, b, wh_smem);
}

// Two waves per SIMD: the matrix-core form's complex loop instances otherwise take 266 registers
// per lane (one wave per SIMD); the LDS holds two workgroups of four waves per CU.
#ifndef WH_REP4_WAVES_PER_EU
#define WH_REP4_WAVES_PER_EU 2
#endif
template <bool REAL, bool LOOP, int H>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WH_REP4_WAVES_PER_EU)))
wh_sieve_rep4(const wh_job_dev* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
              double4* __restrict__ partials, int64_t block_offset) {
    extern __shared__ __align__(32) unsigned char wh_smem[];
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
#if WH_REP4_MFMA
    rep4m_body<REAL, LOOP, H>(jobs[lo], pool, partials, b, wh_smem);
#else
    rep4_body<REAL, LOOP, H>(jobs[lo], pool, partials, b, wh_smem);
#endif
}

template <bool REAL, bool LOOP, int H>
static hipError_t launch_rep4_t(const wh_job_dev* jobs, int njobs, const void* pool, void* partials,
                                int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = rep4_lds_bytes(REAL ? sizeof(double) : sizeof(cplx), nw, H);
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_rep4<REAL, LOOP, H>, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_rep4<REAL, LOOP, H>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds, stream,
                       jobs, njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

template <int H>
static hipError_t launch_rep4_rl(bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                                 void* partials, int64_t nblocks, int64_t off, int nw, hipStream_t s) {
    if (real) {
        if (loop) return launch_rep4_t<true, true, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
        return launch_rep4_t<true, false, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    }
    if (loop) return launch_rep4_t<false, true, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    return launch_rep4_t<false, false, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
}

// Tile side 3 of the repeated-row kernel is the four-terms-per-wave layout (2E <= 12).
size_t wh_rep_lds_bytes(int T, bool real, int nw, int H) {
    if (T == 3) return rep4_lds_bytes(real ? sizeof(double) : sizeof(cplx), nw, H);
    size_t esz = real ? sizeof(double) : sizeof(cplx);
    size_t NP = 8 * (size_t)T, S = NP + WH_PM1_PAD;
    return (1 + nw) * NP * S * esz + 2 * NP * esz + nw * (size_t)(rep_nb() * 2 * rep_ptn(H)) * sizeof(cplx) +
           nw * sizeof(double4) + (size_t)NP / 2 * (H + 1) * sizeof(double) + 32;
}

template <int T, bool REAL, bool LOOP, int H>
static hipError_t launch_rep_t(const wh_job_dev* jobs, int njobs, const void* pool, void* partials,
                               int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = wh_rep_lds_bytes(T, REAL, nw, H);
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_rep<T, REAL, LOOP, H>, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_rep<T, REAL, LOOP, H>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds,
                       stream, jobs, njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

template <int T, int H>
static hipError_t launch_rep_rl(bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                                void* partials, int64_t nblocks, int64_t off, int nw, hipStream_t s) {
    if (real) {
        if (loop) return launch_rep_t<T, true, true, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
        return launch_rep_t<T, true, false, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    }
    if (loop) return launch_rep_t<T, false, true, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    return launch_rep_t<T, false, false, H>(jobs, njobs, pool, partials, nblocks, off, nw, s);
}

hipError_t wh_launch_rep(int T, int H, bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                         void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
    if (T == 1 && H == 8) return launch_rep_rl<1, 8>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    if (T == 1 && H == 16) return launch_rep_rl<1, 16>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    if (T == 2 && H == 8) return launch_rep_rl<2, 8>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    if (T == 2 && H == 16) return launch_rep_rl<2, 16>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    if (T == 3 && H == 8) return launch_rep4_rl<8>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    if (T == 3 && H == 16) return launch_rep4_rl<16>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    return hipErrorInvalidValue;
}
