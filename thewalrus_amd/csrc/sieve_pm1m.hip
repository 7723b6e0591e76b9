// sieve_pm1m.hip -- the all-ones Glynn sieve with full powers on the FP64 matrix cores: the
// terms of _hafnian.py:443-462 (plain) and :546-569 (even loop) when rpt is all ones, complex
// matrices, half-products formulation, and the montrealer's subset terms
// (_montrealer.py:56, :101).  Complex hafnians run on the symmetric-power kernel
// (sieve_pm1s.hip, 60% of this kernel's matrix-core work at n_pad = 40) unless the request asks
// for full powers (WH_OPT_NONSYM); the montrealer's 0/1 column factors break that kernel's
// symmetry, so it stays here.
//
// Same work decomposition as sieve_pm1.hip (a workgroup keeps M = Ax X in LDS, each wave walks
// terms and keeps its current power B^a in LDS); the product B^(a+1) = (B^a M) diag(d) runs as
// v_mfma_f64_4x4x4f64 (four independent 4x4x4 blocks per instruction).  With one wave per
// SIMD the matrix core runs while the wave issues its LDS loads, where the VALU kernel spends
// an issue slot on every FMA.
//
// Lane map (l = lane): x = l & 3, b = (l >> 2) & 3, y = l >> 4, br = b & 1, bc = b >> 1.
// For the 4x4x4 instruction, block b of lanes holds A[i][k] at lane x = i, y = k, B[k][j] at
// x = j, y = k, and returns D[i][j] at x = j, y = i.  Output block groups (p, q), p, q < T,
// cover rows 8p + 4br + (0..3) and columns 8q + 4bc + (0..3): the four blocks of one
// instruction form a 2 x 2 square of 4 x 4 blocks, so n_pad = 8T needs no further padding.
// Per lane and group the accumulators hold D[8p + 4br + y][8q + 4bc + x] (real and imaginary
// parts), T x T groups: n_pad^2 / 64 complex values per lane.
#include "wh_device.h"
#include "wh_launch.h"

namespace {

// LDS row stride 8T + 2 (complex elements): the 16-lane groups of the accumulator stores and
// of the operand reads then hit distinct banks (the transposed trace reads stay 2-way)
#ifndef WH_PM1M_PAD
#define WH_PM1M_PAD 2
#endif

static __constant__ double kInvJm[WH_EMAX + 1] = {
    0.0,      1.0,      1.0 / 2,  1.0 / 3,  1.0 / 4,  1.0 / 5,  1.0 / 6,  1.0 / 7,  1.0 / 8,
    1.0 / 9,  1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17,
    1.0 / 18, 1.0 / 19, 1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25, 1.0 / 26,
    1.0 / 27, 1.0 / 28, 1.0 / 29, 1.0 / 30, 1.0 / 31, 1.0 / 32};

__device__ __forceinline__ double mfma4(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}
// c + a (-b): for the f64 instructions the last operand holds the negate bits (A, B, C)
__device__ __forceinline__ double mfma4_nb(double a, double b, double c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 2);
}

// factor of column c of the current term (+1 digit set; -1 Glynn / 0 montrealer otherwise)
__device__ __forceinline__ double col_factor(uint64_t t, int c, int h, double unset) {
    int i = c < h ? c : c - h;
    if (i >= h) return 1.0;  // zero padding
    return ((t >> (h - 1 - i)) & 1ull) ? 1.0 : unset;
}

#define WH_DPP_QUAD_MIRROR 0x1B  // quad_perm [3,2,1,0]

// sum over lane bits 0, 1, 3 (x and bc): the 8 lanes holding one row of a group
__device__ __forceinline__ double sum_x_bc(double v) {
    v += dpp_d<WH_DPP_XOR1>(v);
    v += dpp_d<WH_DPP_XOR2>(v);
    v += dpp_d<WH_DPP_ROR8>(v);
    return v;
}
// sum over lane bits 2, 4, 5 (br and y): the 8 lanes holding one column of a group.  Lane ^ 4
// is the half-row mirror (i <-> 7 - i) followed by the quad mirror (i <-> 3 - i).
__device__ __forceinline__ double sum_y_br(double v) {
    v += dpp_d<WH_DPP_QUAD_MIRROR>(dpp_d<WH_DPP_HALF_MIRROR>(v));
    v = swap16_sum(v);
    return swap32_sum(v);
}

// [eta^h] exp(sum_i a_i eta^i) for h <= H by e_0 = 1, e_k = (1/k) sum_(i=1..k) b_i e_(k-i) with
// b_i = i a_i = p_i / 2 (+ i v_i / 2 with loops), the exponential's recurrence; two partial
// sums per k halve the dependent chain.
template <int H, bool LOOP>
__device__ __forceinline__ cplx coeff_top(const cplx* pt, const cplx* lv, cplx v1, int h) {
    cplx b[H + 1], e[H + 1];
#pragma unroll
    for (int i = 1; i <= H; ++i) {
        cplx v = cmulr(pt[i], 0.5);
        if (LOOP) v = cadd(v, cmulr(i == 1 ? v1 : lv[i], 0.5 * i));
        b[i] = i <= h ? v : cmk(0.0, 0.0);
    }
    e[0] = cmk(1.0, 0.0);
    cplx res = e[0];
#pragma unroll
    for (int k = 1; k <= H; ++k) {  // every k to H (fixed trip count keeps e[] in registers)
        cplx s0 = cmk(0.0, 0.0), s1 = cmk(0.0, 0.0);
#pragma unroll
        for (int i = 1; i <= k; i += 2) cmac(s0, b[i], e[k - i]);
#pragma unroll
        for (int i = 2; i <= k; i += 2) cmac(s1, b[i], e[k - i]);
        e[k] = cmulr(cadd(s0, s1), kInvJm[k]);
        if (k == h) res = e[k];
    }
    return res;
}

template <int T> struct Geo {
    static constexpr int NP = 8 * T;
    static constexpr int S = NP + WH_PM1M_PAD;  // LDS row stride (complex elements)
};

struct LaneMap {
    int x, y, br, bc;
    __device__ int rd(int p) const { return 8 * p + 4 * br + y; }  // accumulator row
    __device__ int cd(int q) const { return 8 * q + 4 * bc + x; }  // accumulator column
    __device__ bool diag() const { return x == y && br == bc; }     // rd(p) == cd(p)
};

// acc = Pw Ms diag(d) (complex, n_pad x n_pad, k summed in chunks of 4 in order); the column
// factors sg[q] of the lane's columns scale the B operand as it is loaded, under the matrix
// core's shadow
template <int T>
__device__ __forceinline__ void mfma_product(const cplx* __restrict__ Pw, const cplx* __restrict__ Ms,
                                             double (&aR)[T][T], double (&aI)[T][T], const LaneMap& L,
                                             const double (&sg)[T]) {
    constexpr int S = Geo<T>::S;
    constexpr int NK = 2 * T;  // k chunks of 4
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = 0; q < T; ++q) {
            aR[p][q] = 0.0;
            aI[p][q] = 0.0;
        }
    const cplx* pa = Pw + (4 * L.br + L.x) * S + L.y;  // + 8p S + 4 kc
    const cplx* pb = Ms + L.y * S + 4 * L.bc + L.x;    // + 4 kc S + 8q
    for (int kc = 0; kc < NK; ++kc) {
        cplx a[T], bq[T];
#pragma unroll
        for (int p = 0; p < T; ++p) a[p] = pa[8 * p * S + 4 * kc];
#pragma unroll
        for (int q = 0; q < T; ++q) bq[q] = cmulr(pb[4 * kc * S + 8 * q], sg[q]);
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = 0; q < T; ++q) {
                aR[p][q] = mfma4(a[p].re, bq[q].re, aR[p][q]);
                aI[p][q] = mfma4(a[p].re, bq[q].im, aI[p][q]);
            }
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = 0; q < T; ++q) {
                aR[p][q] = mfma4_nb(a[p].im, bq[q].im, aR[p][q]);
                aI[p][q] = mfma4(a[p].im, bq[q].re, aI[p][q]);
            }
    }
}

// lane part of <acc, X^T> = sum acc[i][j] X[j][i]
template <int T>
__device__ __forceinline__ cplx mfma_tdot(const cplx* X, const double (&aR)[T][T], const double (&aI)[T][T],
                                          const LaneMap& L) {
    constexpr int S = Geo<T>::S;
    cplx s = cmk(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = 0; q < T; ++q) cmac(s, cmk(aR[p][q], aI[p][q]), X[L.cd(q) * S + L.rd(p)]);
    return s;
}

template <int T>
__device__ __forceinline__ void mfma_store(cplx* X, const double (&aR)[T][T], const double (&aI)[T][T],
                                           const LaneMap& L) {
    constexpr int S = Geo<T>::S;
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = 0; q < T; ++q) X[L.rd(p) * S + L.cd(q)] = cmk(aR[p][q], aI[p][q]);
}


// Epilogue of power a (from its accumulators and Pw = B^a), cut into micro-steps that run
// between the matrix-core instructions of the next product: after every pair of MFMAs the
// wave issues one micro-step, and a scheduling barrier keeps the compiler from gathering the
// MFMAs back into one block (the wave issues in order, so only instructions placed between
// MFMAs run in their shadow).  Chunk kc works on row group r = kc / 2: the even chunk takes
// the row's share of <B^a, (B^a)^T> and of z_a = B^a D, the odd chunk reduces and stores the
// row's z; the last chunk forms the loop dots and sums the eight results over the wave.
//
// Loop dots.  With A symmetric, M = A X and X D_d = D_d X (d repeats over the two halves),
// (B^a)^T X D_d = X D_d B^a, so y_a = XD_d B^a = X D_d z_a when the loop weights are
// XD0 = X D (hafnians): v_2a = y_a . z_(a-1) = sum_j d_j z_a[swap j] z_(a-1)[j] and v_(2a+1)
// likewise with z_a, from the z vectors alone.  The montrealer's weights (conj(X zeta),
// X zeta) are not of that form: YS keeps each lane's partial y_a over its rows instead.
template <int T>
struct EpiPipe {
    cplx qe, de, dodd, zs, xd;
    cplx ys[T], tx[T], dv[T], zc[T];
    double s1[4], s2[2], s3;
};

template <int T, bool LOOP, bool YS>
__device__ __forceinline__ int epi_steps(int kc) {
    constexpr int NK = 2 * T;
    int n = 0;
    if ((kc & 1) == 0) n = LOOP ? (YS ? 3 * T : 2 * T) : T;
    else n = LOOP ? 4 : 0;
    if (kc == NK - 1) n += (LOOP ? (YS ? 1 + T : 2) : 0) + 11;
    return n;
}

template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void epi_init(EpiPipe<T>& E, const LaneMap& L, const cplx* Dv) {
    E.qe = E.de = E.dodd = E.zs = cmk(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < T; ++q) E.ys[q] = cmk(0.0, 0.0);
    if (LOOP) {
#pragma unroll
        for (int q = 0; q < T; ++q) E.dv[q] = Dv[L.cd(q)];
    }
}

// the operands of row group r's micro-steps
template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void epi_load_row(int r, EpiPipe<T>& E, const LaneMap& L, const cplx* Pw,
                                             const cplx* XD0, uint64_t tu, int h, double unset) {
    constexpr int S = Geo<T>::S;
#pragma unroll
    for (int q = 0; q < T; ++q) E.tx[q] = Pw[L.cd(q) * S + L.rd(r)];
    if (LOOP && YS) E.xd = cmulr(XD0[L.rd(r)], col_factor(tu, L.rd(r), h, unset));
}

template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void epi_step(int kc, int st, EpiPipe<T>& E, const double (&oR)[T][T],
                                         const double (&oI)[T][T], const LaneMap& L, int lane, int a_prev,
                                         cplx qo_prev, const cplx* zprev, cplx* zcur, cplx* pt, cplx* lv,
                                         uint64_t tu, int h, double unset) {
    constexpr int NK = 2 * T;
    const int r = kc >> 1;
    if ((kc & 1) == 0) {
        if (st < T) {
            cmac(E.qe, cmk(oR[r][st], oI[r][st]), E.tx[st]);
            return;
        }
        if (LOOP && st < 2 * T) {
            cmac(E.zs, cmk(oR[r][st - T], oI[r][st - T]), E.dv[st - T]);
            return;
        }
        if (LOOP && YS && st < 3 * T) {
            cmac(E.ys[st - 2 * T], E.xd, cmk(oR[r][st - 2 * T], oI[r][st - 2 * T]));
            return;
        }
        return;
    }
    if (LOOP) {
        if (st == 0) { E.zs.re += dpp_d<WH_DPP_XOR1>(E.zs.re); E.zs.im += dpp_d<WH_DPP_XOR1>(E.zs.im); return; }
        if (st == 1) { E.zs.re += dpp_d<WH_DPP_XOR2>(E.zs.re); E.zs.im += dpp_d<WH_DPP_XOR2>(E.zs.im); return; }
        if (st == 2) { E.zs.re += dpp_d<WH_DPP_ROR8>(E.zs.re); E.zs.im += dpp_d<WH_DPP_ROR8>(E.zs.im); return; }
        if (st == 3) { zcur[L.rd(r)] = E.zs; E.zs = cmk(0.0, 0.0); return; }  // same value from the row's 8 lanes
        st -= 4;
    }
    if (kc != NK - 1) return;
    if (LOOP) {
        if (st == 0) wave_lds_sync();  // the z rows of every lane are stored
        if (YS) {
            if (st == 0) {
#pragma unroll
                for (int q = 0; q < T; ++q) E.zc[q] = zcur[L.cd(q)];
                return;
            }
            if (st <= T) {
                const int q = st - 1;
                cmac(E.de, E.ys[q], zprev[L.cd(q)]);
                cmac(E.dodd, E.ys[q], E.zc[q]);
                return;
            }
            st -= 1 + T;
        } else {
            // lane e: d_e z_a[swap e] times z_(a-1)[e] and z_a[e]
            const int n = 2 * h;
            const int e = lane < n ? lane : 0;
            if (st == 0) {
                E.zc[0] = zcur[e];
                E.zc[1] = zprev[e];
                E.zc[2] = zcur[e < h ? e + h : e - h];
                return;
            }
            if (st == 1) {
                const cplx f = cmulr(E.zc[2], lane < n ? col_factor(tu, e, h, unset) : 0.0);
                E.de = cmul(f, E.zc[1]);
                E.dodd = cmul(f, E.zc[0]);
                return;
            }
            st -= 2;
        }
    }
    const double v[8] = {qo_prev.re, qo_prev.im, E.qe.re, E.qe.im, E.de.re, E.de.im, E.dodd.re, E.dodd.im};
    if (st < 4) { E.s1[st] = dswap32(v[st], v[st + 4]); return; }
    if (st < 6) { E.s2[st - 4] = dswap16(E.s1[st - 4], E.s1[st - 2]); return; }
    if (st == 6) {
        const bool b3 = (lane >> 3) & 1;
        double keep = b3 ? E.s2[1] : E.s2[0], send = b3 ? E.s2[0] : E.s2[1];
        E.s3 = keep + dpp_d<WH_DPP_ROR8>(send);
        return;
    }
    if (st == 7) { E.s3 += dpp_d<WH_DPP_XOR1>(E.s3); return; }
    if (st == 8) { E.s3 += dpp_d<WH_DPP_XOR2>(E.s3); return; }
    if (st == 9) { E.s3 += dpp_d<WH_DPP_HALF_MIRROR>(E.s3); return; }
    if (st == 10) {
        // lane 8k holds value k: p_ko, p_ke, v_2a, v_(2a+1) (re, im); the other lanes write the
        // spare slot pt[WH_EMAX + 1] (never read)
        const int k = lane >> 3;
        const int a = a_prev;
        const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
        double* dst = k < 2 ? &pt[ko].re : k < 4 ? &pt[ke].re : k < 6 ? &lv[2 * a].re : &lv[2 * a + 1].re;
        dst += k & 1;
        if ((lane & 7) != 0) dst = &pt[WH_EMAX + 1].re;
        *dst = E.s3;
    }
}

// accN = Pw Ms diag(d) (fully unrolled k chunks) with power a_prev's epilogue (from accO and
// Pw = B^(a_prev)) in the matrix core's shadow.  The operands of chunk kc + 1 load while chunk
// kc runs (register double buffer), a row's epilogue operands one chunk before its steps.
template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void mfma_product_epi(const cplx* __restrict__ Pw, const cplx* __restrict__ Ms,
                                                 double (&nR)[T][T], double (&nI)[T][T], const LaneMap& L,
                                                 const double (&sg)[T], const double (&oR)[T][T],
                                                 const double (&oI)[T][T], int a_prev, cplx qo_prev, int lane,
                                                 const cplx* Dv, const cplx* XD0, const cplx* zprev, cplx* zcur,
                                                 uint64_t tu, int h, double unset, cplx* pt, cplx* lv) {
    constexpr int S = Geo<T>::S;
    constexpr int NK = 2 * T;
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = 0; q < T; ++q) {
            nR[p][q] = 0.0;
            nI[p][q] = 0.0;
        }
    EpiPipe<T> E;
    epi_init<T, LOOP, YS>(E, L, Dv);
    const cplx* pa = Pw + (4 * L.br + L.x) * S + L.y;
    const cplx* pb = Ms + L.y * S + 4 * L.bc + L.x;
    cplx avn[T], bqn[T];
#pragma unroll
    for (int p = 0; p < T; ++p) avn[p] = pa[8 * p * S];
#pragma unroll
    for (int q = 0; q < T; ++q) bqn[q] = pb[8 * q];
    epi_load_row<T, LOOP, YS>(0, E, L, Pw, XD0, tu, h, unset);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
        cplx av[T], bq[T];
#pragma unroll
        for (int p = 0; p < T; ++p) av[p] = avn[p];
#pragma unroll
        for (int q = 0; q < T; ++q) bq[q] = cmulr(bqn[q], sg[q]);
        if (kc + 1 < NK) {
#pragma unroll
            for (int p = 0; p < T; ++p) avn[p] = pa[8 * p * S + 4 * (kc + 1)];
#pragma unroll
            for (int q = 0; q < T; ++q) bqn[q] = pb[4 * (kc + 1) * S + 8 * q];
        }
        if ((kc & 1) == 1 && kc + 1 < NK) epi_load_row<T, LOOP, YS>((kc + 1) >> 1, E, L, Pw, XD0, tu, h, unset);
        const int nst = epi_steps<T, LOOP, YS>(kc);
        int st = 0;
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = 0; q < T; ++q) {
                nR[p][q] = mfma4(av[p].re, bq[q].re, nR[p][q]);
                nI[p][q] = mfma4(av[p].re, bq[q].im, nI[p][q]);
                if (st < nst)
                    epi_step<T, LOOP, YS>(kc, st, E, oR, oI, L, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h,
                                          unset);
                ++st;
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = 0; q < T; ++q) {
                nR[p][q] = mfma4_nb(av[p].im, bq[q].im, nR[p][q]);
                nI[p][q] = mfma4(av[p].im, bq[q].re, nI[p][q]);
                if (st < nst)
                    epi_step<T, LOOP, YS>(kc, st, E, oR, oI, L, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h,
                                          unset);
                ++st;
                __builtin_amdgcn_sched_barrier(0);
            }
        for (; st < nst; ++st)
            epi_step<T, LOOP, YS>(kc, st, E, oR, oI, L, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h, unset);
    }
}

// Power a's epilogue on its own (the last power of a term): the same micro-steps in sequence.
template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void epi_alone(const double (&aR)[T][T], const double (&aI)[T][T], int a, cplx qo,
                                          const LaneMap& L, int lane, const cplx* Pw, const cplx* Dv,
                                          const cplx* XD0, const cplx* zprev, cplx* zcur, uint64_t tu, int h,
                                          double unset, cplx* pt, cplx* lv) {
    constexpr int NK = 2 * T;
    EpiPipe<T> E;
    epi_init<T, LOOP, YS>(E, L, Dv);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
        if ((kc & 1) == 0) epi_load_row<T, LOOP, YS>(kc >> 1, E, L, Pw, XD0, tu, h, unset);
        const int nst = epi_steps<T, LOOP, YS>(kc);
#pragma unroll
        for (int st = 0; st < 3 * T + 2 * T + 16; ++st)
            if (st < nst) epi_step<T, LOOP, YS>(kc, st, E, aR, aI, L, lane, a, qo, zprev, zcur, pt, lv, tu, h, unset);
    }
}

template <int T, bool LOOP, bool YS>
__device__ __forceinline__ void pm1m_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                          double4* __restrict__ partials, int64_t b, unsigned char* smem) {
    constexpr int NP = Geo<T>::NP;
    constexpr int S = Geo<T>::S;
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    LaneMap L;
    L.x = lane & 3;
    L.br = (lane >> 2) & 1;
    L.bc = (lane >> 3) & 1;
    L.y = lane >> 4;
    const int n = 2 * J.E;
    const int h = J.E;
    const bool mtl = (J.flags & WH_F_MTL) != 0;
    const double unset = mtl ? 0.0 : -1.0;

    // ---- LDS carve (wh_pm1m_lds_bytes)
    cplx* Ms = reinterpret_cast<cplx*>(smem);
    const size_t mat = ((size_t)NP * S * sizeof(cplx) + 31) / 32 * 32 / sizeof(cplx);
    cplx* Pw = Ms + mat + wave * mat;
    cplx* vecs = Ms + (1 + nw) * mat;
    cplx* XD0 = vecs;
    cplx* Dv = vecs + NP;
    constexpr int SMALL = 4 * (WH_EMAX + 2) + 2 * NP;  // per wave: pt, lv, comb0, comb1, z rows x 2
    cplx* small = vecs + 2 * NP + wave * SMALL;
    cplx* pt = small;
    cplx* lv = pt + (WH_EMAX + 2);
    cplx* comb0 = lv + (WH_EMAX + 2);
    cplx* comb1 = comb0 + (WH_EMAX + 2);
    cplx* Zb = comb1 + (WH_EMAX + 2);  // z_a by row, two buffers (a even / odd)
    double4* wres = reinterpret_cast<double4*>(vecs + 2 * NP + nw * SMALL);

    // ---- stage M = Ax X, zero padded
    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < NP * NP; e += blockDim.x) {
        int r = e / NP, c = e - r * NP;
        cplx v = cmk(0.0, 0.0);
        if (r < n && c < n) {
            int src = c < h ? c + h : c - h;
            v = A[(int64_t)r * n + src];
        }
        Ms[r * S + c] = v;
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < NP; e += blockDim.x) {
            cplx xd = cmk(0.0, 0.0), dd = cmk(0.0, 0.0);
            if (e < n) {
                if (mtl) {
                    xd = pool[J.v_off + e];
                    dd = Dg[e];
                } else {
                    xd = Dg[e < h ? e + h : e - h];
                    dd = Dg[e];
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

    csum acc_sum;
    acc_sum.zero();

    for (int64_t t = t0 + wave; t < t1; t += nw) {
        const uint64_t tu = (uint64_t)t;
        const double prefac = mtl ? (((__popcll(tu) + 1) & 1) ? -1.0 : 1.0)
                                  : (((h - __popcll(tu)) & 1) ? -1.0 : 1.0);

        // B = M diag(d) in the accumulator layout
        double aR[T][T], aI[T][T];
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = 0; q < T; ++q) {
                cplx m = Ms[L.rd(p) * S + L.cd(q)];
                const double f = col_factor(tu, L.cd(q), h, unset);
                aR[p][q] = m.re * f;
                aI[p][q] = m.im * f;
            }

        // z_0 = D by row (the loop weights are re-read from LDS in each power's epilogue so that
        // only the accumulators stay live across the products)
        if (LOOP) {
            for (int e = lane; e < NP; e += WH_WAVE) Zb[e] = Dv[e];
        }
        wave_lds_sync();
        const int Ahalf = (h + 1) / 2;
        // powers up to B^A, A = ceil(h/2): tr(B^(2a)) = <B^a, (B^a)^T>, tr(B^(2a+1)) =
        // <B^(a+1), (B^a)^T>; loop scalars v_(2a) = y_a . z_(a-1), v_(2a+1) = y_a . z_a with
        // y_a = XD_d B^a and z_a = B^a D.  Software pipeline: the product forming B^a runs with
        // the epilogue of B^(a-1); the accumulators alternate between two register sets.
        {
            mfma_store<T>(Pw, aR, aI, L);
            wave_lds_sync();
            cplx qo = cmk(0.0, 0.0);
            if (L.diag()) {
#pragma unroll
                for (int p = 0; p < T; ++p) qo = cadd(qo, cmk(aR[p][p], aI[p][p]));
            }
            double bR[T][T], bI[T][T];
            int a = 2;
#pragma unroll 1
            for (; a + 1 <= Ahalf; a += 2) {
                double sgp[T];
#pragma unroll
                for (int q = 0; q < T; ++q) sgp[q] = col_factor(tu, L.cd(q), h, unset);
                // B^a into b, epilogue of B^(a-1) from a
                mfma_product_epi<T, LOOP, YS>(Pw, Ms, bR, bI, L, sgp, aR, aI, a - 1, qo, lane, Dv, XD0,
                                          Zb + (a & 1) * NP, Zb + ((a - 1) & 1) * NP, tu, h, unset, pt, lv);
                qo = mfma_tdot<T>(Pw, bR, bI, L);
                wave_lds_sync();
                mfma_store<T>(Pw, bR, bI, L);
                wave_lds_sync();
                // B^(a+1) into a, epilogue of B^a from b
                mfma_product_epi<T, LOOP, YS>(Pw, Ms, aR, aI, L, sgp, bR, bI, a, qo, lane, Dv, XD0,
                                          Zb + ((a + 1) & 1) * NP, Zb + (a & 1) * NP, tu, h, unset, pt, lv);
                qo = mfma_tdot<T>(Pw, aR, aI, L);
                wave_lds_sync();
                mfma_store<T>(Pw, aR, aI, L);
                wave_lds_sync();
            }
            if (a <= Ahalf) {  // one power left: B^a into b, then its epilogue alone
                double sgp[T];
#pragma unroll
                for (int q = 0; q < T; ++q) sgp[q] = col_factor(tu, L.cd(q), h, unset);
                mfma_product_epi<T, LOOP, YS>(Pw, Ms, bR, bI, L, sgp, aR, aI, a - 1, qo, lane, Dv, XD0,
                                          Zb + (a & 1) * NP, Zb + ((a - 1) & 1) * NP, tu, h, unset, pt, lv);
                qo = mfma_tdot<T>(Pw, bR, bI, L);
                wave_lds_sync();
                mfma_store<T>(Pw, bR, bI, L);
                wave_lds_sync();
                epi_alone<T, LOOP, YS>(bR, bI, a, qo, L, lane, Pw, Dv, XD0, Zb + ((a - 1) & 1) * NP,
                                   Zb + (a & 1) * NP, tu, h, unset, pt, lv);
            } else {  // the last power (a - 1) is in a
                epi_alone<T, LOOP, YS>(aR, aI, a - 1, qo, L, lane, Pw, Dv, XD0, Zb + (a & 1) * NP,
                                   Zb + ((a - 1) & 1) * NP, tu, h, unset, pt, lv);
            }
        }
        cplx v1 = cmk(0.0, 0.0);
        if (LOOP) {
            // v_1 = XD_d . D (every lane holds the sum)
            cplx part = cmk(0.0, 0.0);
            for (int e = lane; e < n; e += WH_WAVE) cmac(part, cmulr(XD0[e], col_factor(tu, e, h, unset)), Dv[e]);
            v1 = wave_csum(part);
        }
        wave_lds_sync();  // pt and lv as the epilogues stored them

        if (mtl) {
            // montrealer term: sign (p_h / 2h + v_h / 2)
            if (lane == h) {
                cplx f = cdivr(pt[h], (double)(2 * h));
                if (LOOP) f = cadd(f, cmulr(h == 1 ? v1 : lv[h], 0.5));
                acc_sum.add(cmulr(f, prefac));
            }
            continue;
        }
        // coefficient of eta^h in exp(sum_i a_i eta^i), a_i = p_i / 2i (+ v_i / 2): the
        // reference builds it factor by factor (_hafnian.py:200-208, :232-241); here every
        // lane runs the power-series exponential e_k = (1/k) sum_(i<=k) (i a_i) e_(k-i) in
        // registers (the same polynomial, no LDS traffic on the wave's serial tail)
        const cplx top = coeff_top<4 * T, LOOP>(pt, lv, v1, h);
        if (lane == h) acc_sum.add(cmulr(top, prefac));
    }

    // fold the waves of this workgroup in wave order
    if (lane == h) wres[wave] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    __syncthreads();
    if (threadIdx.x == 0) {
        csum s;
        s.zero();
        for (int w = 0; w < nw; ++w) {
            double4 pq = wres[w];
            s.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        partials[b] = make_double4(s.sr, s.cr, s.si, s.ci);
    }
}

}  // namespace

template <int T, bool LOOP, bool YS>
__global__ void __launch_bounds__(256)
wh_sieve_pm1m(const wh_job_dev* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
              double4* __restrict__ partials, int64_t block_offset) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    pm1m_body<T, LOOP, YS>(jobs[lo], pool, partials, b, wh_smem);
}

// LDS of the matrix-core kernel: M and one power per wave (as the vector kernel), the
// vectors XD0 and D, and per wave pt, lv, two coefficient rows and two z rows.
size_t wh_pm1m_lds_bytes(int T, int nw) {
    size_t NP = 8 * (size_t)T, S = NP + WH_PM1M_PAD;
    size_t mat = (NP * S * sizeof(cplx) + 31) / 32 * 32;
    size_t small = (4 * (WH_EMAX + 2) + 2 * NP) * sizeof(cplx);
    return (1 + nw) * mat + 2 * NP * sizeof(cplx) + nw * small + nw * sizeof(double4) + 32;
}

template <int T, bool LOOP, bool YS>
static hipError_t launch_m(const wh_job_dev* jobs, int njobs, const void* pool, void* partials, int64_t nblocks,
                           int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = wh_pm1m_lds_bytes(T, nw);
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_pm1m<T, LOOP, YS>, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_pm1m<T, LOOP, YS>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds, stream, jobs,
                       njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

hipError_t wh_launch_pm1m(int T, bool loop, bool mtl, const wh_job_dev* jobs, int njobs, const void* pool,
                          void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
#define WH_PM1M_CASE(TT)                                                                                   \
    case TT:                                                                                               \
        if (!loop) return launch_m<TT, false, false>(jobs, njobs, pool, partials, nblocks, block_offset, nw, stream); \
        if (mtl) return launch_m<TT, true, true>(jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);     \
        return launch_m<TT, true, false>(jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
    switch (T) {
        WH_PM1M_CASE(1)
        WH_PM1M_CASE(2)
        WH_PM1M_CASE(3)
        WH_PM1M_CASE(4)
        WH_PM1M_CASE(5)
        default: return hipErrorInvalidValue;
    }
#undef WH_PM1M_CASE
}
