// sieve_pm1s.hip -- the all-ones Glynn sieve on the FP64 matrix cores with symmetric powers:
// the terms of _hafnian.py:443-462 (plain) and :546-569 (even loop) when rpt is all ones,
// complex matrices (and real ones, one real MFMA per block product), half-products
// formulation, with 60% of the block products of sieve_pm1m.hip at n_pad = 40, each complex one
// from three real MFMAs.
//
// With A = Ax symmetric, X the swap of the two halves and D_d = diag(d) (d = +-1 repeating over
// the halves, so X D_d = D_d X and (X D_d)^2 = I), the term matrix is B = A X D_d and
//     C_a = B^a X D_d = A (X D_d A)^(a-1)   is symmetric, and so is V_a = D_d C_a D_d.
// Each wave keeps V_a in LDS and forms the products
//     R_(a+1) = V_a (X A) = D_d C_(a+1),   V_(a+1)[i][j] = R_(a+1)[i][j] d_j,
// on the matrix cores for the upper block groups only (T (T + 1) / 2 of the T^2 groups of
// sieve_pm1m.hip), with X A staged once per workgroup and no sign in the products: the lane
// stores R d_j at (i, j) and, by symmetry, at (j, i).  The power traces are
//     tr(B^(a+b)) = sum_(i,k) C_a[i][k] V_b[swap k][swap i] = sum_(i,k) d_i R_a[i][k] V_b[swap k][swap i],
// a summand symmetric in (i, k): the upper groups off the diagonal count twice and the row sign
// d_i scales each row group's partial sum once.  The loop vectors z_a = B^a D follow the chain
// z_a = A (X D_d z_(a-1)), and the loop dots use y_a = X D_d z_a as in sieve_pm1m.hip.  The
// montrealer (0/1 factors: D_d^2 is not I) stays on sieve_pm1m.hip.  Each wave keeps the power
// traces and loop scalars of its last few terms in LDS and forms their coefficients together,
// one term per lane.
//
// Lane map and MFMA layout as sieve_pm1m.hip: x = lane & 3, br = bit 2, bc = bit 3,
// y = lane >> 4; group (p, q) holds R[8p + 4br + y][8q + 4bc + x].  The signs d of the lane's
// rows and columns live in one mask register and flip the sign bit of the values they scale;
// where a complex product takes four MFMAs, its -Im(a) Im(b) term uses the matrix core's negate
// modifier on B.
#include <algorithm>

#include "wh_device.h"
#include "wh_launch.h"

namespace {

#ifndef WH_PM1S_PAD
#define WH_PM1S_PAD 2
#endif
// complex products with three real MFMAs per block group and k chunk instead of four (below);
// 0 turns the form off
#ifndef WH_PM1S_3M
#define WH_PM1S_3M 1
#endif

static __constant__ double kInvJs[WH_EMAX + 1] = {
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

// d_c of the current term: +1 if digit (c mod h) is set, -1 otherwise, 1 on the padding
__device__ __forceinline__ double sgn(uint64_t t, int c, int h) {
    int i = c < h ? c : c - h;
    if (i >= h) return 1.0;
    return ((t >> (h - 1 - i)) & 1ull) ? 1.0 : -1.0;
}
// 1 where d_c = -1 (digit unset), 0 where d_c = +1 or c is padding
__device__ __forceinline__ uint32_t negbit(uint64_t t, int c, int h) {
    int i = c < h ? c : c - h;
    if (i >= h) return 0u;
    return ((t >> (h - 1 - i)) & 1ull) ? 0u : 1u;
}
// v times (-1)^bit, bit in {0, 1}: the sign bit of the high word
__device__ __forceinline__ double sflip(double v, uint32_t bit) {
    return __hiloint2double(__double2hiint(v) ^ (int)(bit << 31), __double2loint(v));
}
__device__ __forceinline__ cplx cflip(cplx v, uint32_t bit) { return cmk(sflip(v.re, bit), sflip(v.im, bit)); }
// the partner of index c in the other half (identity on the padding)
__device__ __forceinline__ int swp(int c, int h) { return c < h ? c + h : (c < 2 * h ? c - h : c); }
// an LDS value; the real instance (RE) reads the real half only, so the compiler sees a zero
// imaginary part and drops the arithmetic on it (imaginary halves in LDS are then never read)
template <bool RE>
__device__ __forceinline__ cplx ldc(const cplx* p) {
    if constexpr (RE) return cmk(p->re, 0.0);
    else return *p;
}
// Entry i of a matrix in LDS (X A, the powers V): complex, or for the real instance its real
// part only, the matrices then laid out as doubles (half the LDS: more workgroups per CU)
template <bool RE>
__device__ __forceinline__ cplx mget(const cplx* m, int i) {
    if constexpr (RE) return cmk(reinterpret_cast<const double*>(m)[i], 0.0);
    else return m[i];
}
template <bool RE>
__device__ __forceinline__ void mput(cplx* m, int i, cplx v) {
    if constexpr (RE) reinterpret_cast<double*>(m)[i] = v.re;
    else m[i] = v;
}

template <int T> struct GeoS {
    static constexpr int NP = 8 * T;
    static constexpr int S = NP + WH_PM1S_PAD;
    // per term slot: p_1 .. p_(4T) and v_1 .. v_(4T+1) (with h <= 4T, a <= ceil(h/2)), the last
    // entry the spare that lanes without a value write
    static constexpr int PTN = 4 * T + 4;
    // terms whose coefficients are formed together, one per lane (LDS budget at 4 waves, T <= 5;
    // T = 6 runs 3 waves with the LDS nearly full, T = 7 and 8 one wave)
    static constexpr int NB = T >= 6 ? 1 : T == 5 ? 7 : 8;
};

struct LaneMapS {
    int x, y, br, bc;
    __device__ int rd(int p) const { return 8 * p + 4 * br + y; }
    __device__ int cd(int q) const { return 8 * q + 4 * bc + x; }
};

// [eta^h] exp(sum_i a_i eta^i) by the exponential recurrence (as sieve_pm1m.hip)
template <int H, bool LOOP, bool RE = false>
__device__ __forceinline__ cplx coeff_top_s(const cplx* pt, const cplx* lv, cplx v1, int h) {
    cplx b[H + 1], e[H + 1];
#pragma unroll
    for (int i = 1; i <= H; ++i) {
        cplx v = cmulr(ldc<RE>(pt + i), 0.5);
        if (LOOP) v = cadd(v, cmulr(i == 1 ? v1 : ldc<RE>(lv + i), 0.5 * i));
        b[i] = i <= h ? v : cmk(0.0, 0.0);
    }
    e[0] = cmk(1.0, 0.0);
    cplx res = e[0];
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        cplx s0 = cmk(0.0, 0.0), s1 = cmk(0.0, 0.0);
#pragma unroll
        for (int i = 1; i <= k; i += 2) cmac(s0, b[i], e[k - i]);
#pragma unroll
        for (int i = 2; i <= k; i += 2) cmac(s1, b[i], e[k - i]);
        e[k] = cmulr(cadd(s0, s1), kInvJs[k]);
        if (k == h) res = e[k];
    }
    return res;
}

// lane part of sum_(i,k) d_i Ra[i][k] Vb[swap k][swap i] over the upper groups (off-diagonal
// twice), the row sign applied to each row group's partial sum.  Split in two so that the V
// loads can be issued before the next power's store and summed after it (their latency then
// overlaps the store).
template <int T>
struct TdotVals {
    cplx x[T * (T + 1) / 2];
};
template <int T, bool RE = false>
__device__ __forceinline__ void sym_tdot_load(TdotVals<T>& X, const cplx* Vb, const LaneMapS& L, int h) {
    constexpr int S = GeoS<T>::S;
    int g = 0;
#pragma unroll
    for (int p = 0; p < T; ++p) {
        const int rs = swp(L.rd(p), h);
#pragma unroll
        for (int q = p; q < T; ++q) X.x[g++] = mget<RE>(Vb, swp(L.cd(q), h) * S + rs);
    }
}
template <int T>
__device__ __forceinline__ cplx sym_tdot_sum(const TdotVals<T>& X, const double (&aR)[T][T],
                                             const double (&aI)[T][T], uint32_t fm) {
    cplx s = cmk(0.0, 0.0);
    int g = 0;
#pragma unroll
    for (int p = 0; p < T; ++p) {
        cplx sd = cmk(0.0, 0.0), so = cmk(0.0, 0.0);
#pragma unroll
        for (int q = p; q < T; ++q, ++g) {
            if (q == p) cmac(sd, cmk(aR[p][q], aI[p][q]), X.x[g]);
            else cmac(so, cmk(aR[p][q], aI[p][q]), X.x[g]);
        }
        const cplx part = cflip(cadd(sd, cmulr(so, 2.0)), (fm >> p) & 1u);
        s = p == 0 ? part : cadd(s, part);
    }
    return s;
}

// V = R diag(d) into LDS: each upper group at its place and, off the diagonal, at the mirrored
// place (V is symmetric); column sign = bit T + q of the lane's sign mask
template <int T, bool RE = false>
__device__ __forceinline__ void sym_store(cplx* X, const double (&aR)[T][T], const double (&aI)[T][T],
                                          const LaneMapS& L, uint32_t fm) {
    constexpr int S = GeoS<T>::S;
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = p; q < T; ++q) {
            const cplx v = cflip(cmk(aR[p][q], aI[p][q]), (fm >> (T + q)) & 1u);
            mput<RE>(X, L.rd(p) * S + L.cd(q), v);
            if (p < q) mput<RE>(X, L.cd(q) * S + L.rd(p), v);
        }
}

// Epilogue micro-steps of power a (interleaved with the next product as in sieve_pm1m.hip): the
// even chunk of row group r takes the row's share of the trace sum and of z_a = A t
// (t = X D_d z_(a-1)), the odd chunk reduces and stores the row's z, the last chunk forms the
// loop dots and sums the eight results over the wave.
template <int T>
struct EpiS {
    cplx qe, qf, qt, de, dodd, zs;
    cplx tx[T], av[T], tv[T], zc[3];
    double s1[4], s2[2], s3;
};

template <int T, bool LOOP>
__device__ __forceinline__ int epis_steps(int kc) {
    constexpr int NK = 2 * T;
    const int r = kc >> 1;
    int n = 0;
    if ((kc & 1) == 0) n = (T - r) + (LOOP ? T : 0);
    else n = LOOP ? 4 : 0;
    if (kc == NK - 1) n += (LOOP ? 2 : 0) + 11;
    return n;
}

template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ void epis_init(EpiS<T>& E, const LaneMapS& L, const cplx* zprev, uint32_t fm,
                                          int h) {
    E.qe = E.qf = E.qt = E.de = E.dodd = E.zs = cmk(0.0, 0.0);
    if (LOOP) {  // t = X D_d z_(a-1) at the lane's columns
#pragma unroll
        for (int q = 0; q < T; ++q) E.tv[q] = cflip(ldc<RE>(&zprev[swp(L.cd(q), h)]), (fm >> (T + q)) & 1u);
    }
}

// the operands of row group r's micro-steps
template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ void epis_load_row(int r, EpiS<T>& E, const LaneMapS& L, const cplx* Vb,
                                              const cplx* Ms, int h) {
    constexpr int S = GeoS<T>::S;
    const int rs = swp(L.rd(r), h);
#pragma unroll
    for (int q = r; q < T; ++q) E.tx[q] = mget<RE>(Vb, swp(L.cd(q), h) * S + rs);
    if (LOOP) {  // A[rd(r)][cd(q)] = (X A)[swap rd(r)][cd(q)]
#pragma unroll
        for (int q = 0; q < T; ++q) E.av[q] = mget<RE>(Ms, rs * S + L.cd(q));
    }
}

template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ void epis_step(int kc, int st, EpiS<T>& E, const double (&oR)[T][T],
                                          const double (&oI)[T][T], const LaneMapS& L, uint32_t fm, int lane,
                                          int a_prev, cplx qo_prev, const cplx* zprev, cplx* zcur, cplx* pt,
                                          cplx* lv, uint64_t tu, int h) {
    constexpr int NK = 2 * T;
    const int r = kc >> 1;
    if ((kc & 1) == 0) {
        if (st < T - r) {
            const int q = r + st;
            if (q == r) cmac(E.qe, cmk(oR[r][q], oI[r][q]), E.tx[q]);
            else cmac(E.qf, cmk(oR[r][q], oI[r][q]), E.tx[q]);
            if (q == T - 1) {  // the row group's partial sum, times its row sign
                E.qt = cadd(E.qt, cflip(cadd(E.qe, cmulr(E.qf, 2.0)), (fm >> r) & 1u));
                E.qe = E.qf = cmk(0.0, 0.0);
            }
            return;
        }
        if (LOOP && st < T - r + T) {
            const int q = st - (T - r);
            cmac(E.zs, E.av[q], E.tv[q]);
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
        // lane e: d_e z_a[swap e] times z_(a-1)[e] and z_a[e]
        const int n = 2 * h;
        const int e = lane < n ? lane : 0;
        if (st == 0) {
            wave_lds_sync();  // the z rows of every lane are stored
            E.zc[0] = ldc<RE>(&zcur[e]);
            E.zc[1] = ldc<RE>(&zprev[e]);
            E.zc[2] = ldc<RE>(&zcur[swp(e, h)]);
            return;
        }
        if (st == 1) {
            const cplx f = cmulr(E.zc[2], lane < n ? sgn(tu, e, h) : 0.0);
            E.de = cmul(f, E.zc[1]);
            E.dodd = cmul(f, E.zc[0]);
            return;
        }
        st -= 2;
    }
    // (the real instance's imaginary parts are zeros of either sign: zero them for the compiler)
    const double v[8] = {qo_prev.re, RE ? 0.0 : qo_prev.im, E.qt.re,   RE ? 0.0 : E.qt.im,
                         E.de.re,    RE ? 0.0 : E.de.im,    E.dodd.re, RE ? 0.0 : E.dodd.im};
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
        // spare entry pt[PTN - 1] (never read)
        const int k = lane >> 3;
        const int a = a_prev;
        const int ko = a == 1 ? 1 : 2 * a - 1, ke = a == 1 ? 2 : 2 * a;
        double* dst = k < 2 ? &pt[ko].re : k < 4 ? &pt[ke].re : k < 6 ? &lv[2 * a].re : &lv[2 * a + 1].re;
        dst += k & 1;
        if ((lane & 7) != 0 || (!LOOP && k >= 4)) dst = &pt[GeoS<T>::PTN - 1].re;  // no v slots
        *dst = E.s3;
    }
}

// Rn = Vb (X A) for the upper groups (fully unrolled k chunks) with power a_prev's epilogue
// (from Ro and Vb = Pw = V_(a_prev)) between the MFMA pairs; operands of chunk kc + 1 load
// while chunk kc runs.  RE: matrix and loop weights are real (every imaginary part zero, also in
// LDS): one MFMA per group and chunk (Re x Re) instead of four, the imaginary accumulators stay
// zero, and only the real halves of the operands are read.
// Complex operands (WH_PM1S_3M): Gauss's three-multiplication form.  With X = Re a Re b,
// Y = Im a Im b and Z = (Re a + Im a)(Re b + Im b) summed over k, Re(ab) = X - Y and
// Im(ab) = Z - X - Y: three MFMAs per group and chunk and 2T operand sums per chunk instead of
// four MFMAs; the accumulators of X, Y, Z combine once at the end of the product.  Normwise the
// rounding error is of the same order as the four-product form's; the imaginary part carries
// it relative to |Re a||Re b| + |Im a||Im b| rather than to itself.  The extra accumulators
// and operand sums push the T = 6 loop and T >= 7 instances further into scratch, where the
// form measured slower (n = 42 loop 292 -> 323 ms, n = 50 5.5 -> 6.5 s): they keep four MFMAs.
template <int T, bool LOOP, bool RE>
constexpr bool pm1s_3m() {
    return !RE && WH_PM1S_3M && (T <= 5 || (T == 6 && !LOOP));
}

// Product boundary inside the last k chunk (WH_PM1S_FUSED, on by default for the instances
// without register spills): the trace loads of V_(a_prev) issue at the start of the last chunk
// (into the registers the operand prefetch no longer needs), and each group's final value
// (the three-product combination where it applies), its share of tr(B^(2 a_prev + 1)) and its
// V stores follow three groups behind the group's last MFMA, in the matrix cores' shadow.  The
// plain form runs them after the product: 13% of the T = 5 loop launch; fused, n = 32..40 run
// 3.5-4.6% faster and the real n = 24 batch 3.5%.  The complex T = 6 and real T = 5 instances
// keep the plain form (with the fused one they spill or lose the full unroll of the product),
// and so does the complex T = 3 loop instance (n = 24 loop: 1.44 -> 1.67 ms fused, with the
// same instruction counts: its short last chunk leaves the combination waiting on the MFMAs).
#ifndef WH_PM1S_FUSED
#define WH_PM1S_FUSED 1
#endif
template <int T, bool LOOP, bool RE>
constexpr bool pm1s_fused() {
    return WH_PM1S_FUSED && (RE ? T <= 4 : T <= 5 && !(T == 3 && LOOP));
}

// Fused boundary state: the trace loads and the running trace sums of the last chunk.
template <int T>
struct BndS {
    TdotVals<T> tv;
    cplx sd, so, s;
};

// group gi (upper groups in row-major order) of the fused boundary: the final value (Gauss's
// combination), the group's term of tr(B^(2 a_prev + 1)) and the stores of V = R diag(d)
template <int T, bool G3, bool RE>
__device__ __forceinline__ void bnd_step(int gi, BndS<T>& Bd, cplx* X, double (&nR)[T][T], double (&nI)[T][T],
                                         double (&nY)[T][T], const LaneMapS& L, uint32_t fm) {
    constexpr int S = GeoS<T>::S;
    int p = 0, q = 0, g = 0;
#pragma unroll
    for (int pp = 0; pp < T; ++pp)
#pragma unroll
        for (int qq = pp; qq < T; ++qq, ++g)
            if (g == gi) { p = pp; q = qq; }
    if (G3) {
        const double x = nR[p][q], y = nY[p][q];
        nR[p][q] = x - y;
        nI[p][q] = nI[p][q] - (x + y);
    }
    const cplx r = cmk(nR[p][q], RE ? 0.0 : nI[p][q]);
    if (q == p) cmac(Bd.sd, r, Bd.tv.x[gi]);
    else cmac(Bd.so, r, Bd.tv.x[gi]);
    if (q == T - 1) {  // the row group's partial sum, times its row sign
        const cplx part = cflip(cadd(Bd.sd, cmulr(Bd.so, 2.0)), (fm >> p) & 1u);
        Bd.s = p == 0 ? part : cadd(Bd.s, part);
        Bd.sd = Bd.so = cmk(0.0, 0.0);
    }
    const cplx v = cflip(r, (fm >> (T + q)) & 1u);
    mput<RE>(X, L.rd(p) * S + L.cd(q), v);
    if (p < q) mput<RE>(X, L.cd(q) * S + L.rd(p), v);
}

// With the fused boundary (pm1s_fused) the product also overwrites Vb with V_a and returns the
// lane's part of tr(B^(2 a_prev + 1)); otherwise it returns zero and the caller does both.
template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ cplx sym_product_epi(cplx* Vb, const cplx* __restrict__ Ms,
                                                double (&nR)[T][T], double (&nI)[T][T], const LaneMapS& L,
                                                uint32_t fm, const double (&oR)[T][T], const double (&oI)[T][T],
                                                int a_prev, cplx qo_prev, int lane, const cplx* zprev,
                                                cplx* zcur, uint64_t tu, int h, cplx* pt, cplx* lv) {
    constexpr int S = GeoS<T>::S;
    constexpr int NK = 2 * T;
    constexpr int NG = T * (T + 1) / 2;
    constexpr bool G3 = pm1s_3m<T, LOOP, RE>();
    constexpr bool FB = pm1s_fused<T, LOOP, RE>();
    constexpr int LAG = 3;  // groups between a group's last MFMA and its boundary step (2: 0.6% slower at T = 5)
    BndS<T> Bd;
    Bd.sd = Bd.so = Bd.s = cmk(0.0, 0.0);
    double nY[T][T];  // G3: sum of Im a Im b
#pragma unroll
    for (int p = 0; p < T; ++p)
#pragma unroll
        for (int q = p; q < T; ++q) {
            nR[p][q] = 0.0;
            nI[p][q] = 0.0;
            if (G3) nY[p][q] = 0.0;
        }
    EpiS<T> E;
    epis_init<T, LOOP, RE>(E, L, zprev, fm, h);
    const int ia = (4 * L.br + L.x) * S + L.y;
    const int ib = L.y * S + 4 * L.bc + L.x;
    cplx avn[T], bqn[T];
#pragma unroll
    for (int p = 0; p < T; ++p) avn[p] = mget<RE>(Vb, ia + 8 * p * S);
#pragma unroll
    for (int q = 0; q < T; ++q) bqn[q] = mget<RE>(Ms, ib + 8 * q);
    epis_load_row<T, LOOP, RE>(0, E, L, Vb, Ms, h);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
        cplx av[T], bq[T];
#pragma unroll
        for (int p = 0; p < T; ++p) av[p] = avn[p];
#pragma unroll
        for (int q = 0; q < T; ++q) bq[q] = bqn[q];
        if (kc + 1 < NK) {
#pragma unroll
            for (int p = 0; p < T; ++p)
                avn[p] = mget<RE>(Vb, ia + 8 * p * S + 4 * (kc + 1));
#pragma unroll
            for (int q = 0; q < T; ++q)
                bqn[q] = mget<RE>(Ms, ib + 4 * (kc + 1) * S + 8 * q);
        }
        if ((kc & 1) == 1 && kc + 1 < NK) epis_load_row<T, LOOP, RE>((kc + 1) >> 1, E, L, Vb, Ms, h);
        if (FB && kc == NK - 1) {
            // V_(a_prev) for the trace, before any store of V_a (one wave's LDS accesses run in
            // issue order; the fence keeps the compiler from moving the stores above the loads)
            sym_tdot_load<T, RE>(Bd.tv, Vb, L, h);
            wave_lds_sync();
        }
        double as[T], bs[T];
        if (G3) {
#pragma unroll
            for (int p = 0; p < T; ++p) as[p] = av[p].re + av[p].im;
#pragma unroll
            for (int q = 0; q < T; ++q) bs[q] = bq[q].re + bq[q].im;
        }
        const int nst = epis_steps<T, LOOP>(kc);
        int st = 0;
        int gl = 0;  // group index of the MFMA loop
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = p; q < T; ++q, ++gl) {
                nR[p][q] = mfma4(av[p].re, bq[q].re, nR[p][q]);
                if (G3) nY[p][q] = mfma4(av[p].im, bq[q].im, nY[p][q]);
                else if (!RE) nI[p][q] = mfma4(av[p].re, bq[q].im, nI[p][q]);
                if (st < nst) epis_step<T, LOOP, RE>(kc, st, E, oR, oI, L, fm, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h);
                ++st;
                if (FB && RE && kc == NK - 1 && gl >= LAG) bnd_step<T, G3, RE>(gl - LAG, Bd, Vb, nR, nI, nY, L, fm);
                __builtin_amdgcn_sched_barrier(0);
            }
        if (!RE) {
            int gz = 0;
#pragma unroll
            for (int p = 0; p < T; ++p)
#pragma unroll
                for (int q = p; q < T; ++q, ++gz) {
                    if (G3) {
                        nI[p][q] = mfma4(as[p], bs[q], nI[p][q]);
                    } else {
                        nR[p][q] = mfma4_nb(av[p].im, bq[q].im, nR[p][q]);
                        nI[p][q] = mfma4(av[p].im, bq[q].re, nI[p][q]);
                    }
                    if (st < nst)
                        epis_step<T, LOOP, RE>(kc, st, E, oR, oI, L, fm, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h);
                    ++st;
                    if (FB && kc == NK - 1 && gz >= LAG) bnd_step<T, G3, RE>(gz - LAG, Bd, Vb, nR, nI, nY, L, fm);
                    __builtin_amdgcn_sched_barrier(0);
                }
        }
        for (; st < nst; ++st)
            epis_step<T, LOOP, RE>(kc, st, E, oR, oI, L, fm, lane, a_prev, qo_prev, zprev, zcur, pt, lv, tu, h);
    }
    if constexpr (FB) {
#pragma unroll
        for (int g = NG - LAG; g < NG; ++g)
            if (g >= 0) bnd_step<T, G3, RE>(g, Bd, Vb, nR, nI, nY, L, fm);
        return Bd.s;
    } else {
        if (G3) {
#pragma unroll
            for (int p = 0; p < T; ++p)
#pragma unroll
                for (int q = p; q < T; ++q) {
                    const double x = nR[p][q], y = nY[p][q];
                    nR[p][q] = x - y;
                    nI[p][q] = nI[p][q] - (x + y);
                }
        }
        return cmk(0.0, 0.0);
    }
}

// One product of the power chain and its boundary: R_a = V_(a_prev) (X A) into (nR, nI) with
// power a_prev's epilogue, V_a into Pw, and the lane's part of tr(B^(2 a_prev + 1)) returned.
template <int T, bool LOOP, bool RE>
__device__ __forceinline__ cplx product_step(cplx* Pw, const cplx* __restrict__ Ms, double (&nR)[T][T],
                                             double (&nI)[T][T], const LaneMapS& L, uint32_t fm,
                                             const double (&oR)[T][T], const double (&oI)[T][T], int a_prev,
                                             cplx qo_prev, int lane, const cplx* zprev, cplx* zcur, uint64_t tu,
                                             int h, cplx* pt, cplx* lv) {
    if constexpr (pm1s_fused<T, LOOP, RE>()) {
        const cplx qo = sym_product_epi<T, LOOP, RE>(Pw, Ms, nR, nI, L, fm, oR, oI, a_prev, qo_prev, lane, zprev,
                                                     zcur, tu, h, pt, lv);
        wave_lds_sync();  // V_a of every lane is stored
        return qo;
    } else {
        sym_product_epi<T, LOOP, RE>(Pw, Ms, nR, nI, L, fm, oR, oI, a_prev, qo_prev, lane, zprev, zcur, tu, h, pt,
                                     lv);
        TdotVals<T> tv;
        sym_tdot_load<T, RE>(tv, Pw, L, h);  // Pw holds V_(a_prev): for tr(B^(2 a_prev + 1))
        wave_lds_sync();
        sym_store<T, RE>(Pw, nR, nI, L, fm);
        wave_lds_sync();
        return sym_tdot_sum<T>(tv, nR, nI, fm);
    }
}

// the last power's epilogue, the same micro-steps in sequence
template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ void epis_alone(const double (&aR)[T][T], const double (&aI)[T][T], int a, cplx qo,
                                           const LaneMapS& L, uint32_t fm, int lane, const cplx* Vb,
                                           const cplx* Ms, const cplx* zprev, cplx* zcur, uint64_t tu, int h,
                                           cplx* pt, cplx* lv) {
    constexpr int NK = 2 * T;
    EpiS<T> E;
    epis_init<T, LOOP, RE>(E, L, zprev, fm, h);
#pragma unroll
    for (int kc = 0; kc < NK; ++kc) {
        if ((kc & 1) == 0) epis_load_row<T, LOOP, RE>(kc >> 1, E, L, Vb, Ms, h);
        const int nst = epis_steps<T, LOOP>(kc);
#pragma unroll
        for (int st = 0; st < 2 * T + 18; ++st)
            if (st < nst) epis_step<T, LOOP, RE>(kc, st, E, aR, aI, L, fm, lane, a, qo, zprev, zcur, pt, lv, tu, h);
    }
}

// One chunk (partial slot b) of job J by the first nw waves of the workgroup; any further waves
// only take part in the staging and the barriers (the persistent form runs every class with the
// largest class's workgroup).
template <int T, bool LOOP, bool RE = false>
__device__ __forceinline__ void pm1s_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                          double4* __restrict__ partials, int64_t b, unsigned char* smem, int nw) {
    constexpr int NP = GeoS<T>::NP;
    constexpr int S = GeoS<T>::S;
    const int wave = threadIdx.x / WH_WAVE;
    const bool active = wave < nw;
    const int lane = threadIdx.x % WH_WAVE;
    LaneMapS L;
    L.x = lane & 3;
    L.br = (lane >> 2) & 1;
    L.bc = (lane >> 3) & 1;
    L.y = lane >> 4;
    const int n = 2 * J.E;
    const int h = J.E;

    // ---- LDS carve (wh_pm1s_lds_bytes): X A, one V per wave (doubles for the real instance), D,
    // per wave NB term slots (pt, lv) and z rows x 2 (D, lv and the z rows for loop hafnians only)
    constexpr int PTN = GeoS<T>::PTN;
    constexpr int NB = GeoS<T>::NB;
    cplx* Ms = reinterpret_cast<cplx*>(smem);
    const size_t mat = ((size_t)NP * S * (RE ? sizeof(double) : sizeof(cplx)) + 31) / 32 * 32 / sizeof(cplx);
    cplx* Pw = Ms + mat + wave * mat;
    cplx* Dv = Ms + (1 + nw) * mat;
    constexpr int DVN = LOOP ? NP : 0;  // D and the z rows only for loop hafnians
    constexpr int SL = LOOP ? 2 * PTN : PTN;  // a term slot: p_i, and v_i for loop hafnians
    constexpr int SMALL = NB * SL + (LOOP ? 2 * NP : 0);
    cplx* slots = Dv + DVN + wave * SMALL;
    cplx* Zb = slots + NB * SL;  // z_a by row, two buffers (a even / odd)
    double4* wres = reinterpret_cast<double4*>(Dv + DVN + nw * SMALL);

    // ---- stage X A (row r holds A's row swap(r)), zero padded, and D
    const cplx* A = pool + J.a_off;
    for (int e = threadIdx.x; e < NP * NP; e += blockDim.x) {
        int r = e / NP, c = e - r * NP;
        cplx v = cmk(0.0, 0.0);
        if (r < n && c < n) v = A[(int64_t)swp(r, h) * n + c];
        mput<RE>(Ms, r * S + c, v);
    }
    if (LOOP) {
        const cplx* Dg = pool + J.d_off;
        for (int e = threadIdx.x; e < NP; e += blockDim.x) Dv[e] = e < n ? Dg[e] : cmk(0.0, 0.0);
    }
    __syncthreads();

    const int64_t t0 = (b - J.chunk_base) * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;

    csum acc_sum;
    acc_sum.zero();

    // The coefficients of up to NB terms are formed together after their powers, lane j taking
    // the j-th term of the batch (the exponential recurrence is the same work for one term or 64).
    int slot = 0;
    int64_t t_batch = t0 + wave;  // the batch's first term
    for (int64_t t = t0 + wave; active && t < t1; t += nw) {
        const uint64_t tu = (uint64_t)t;
        cplx* pt = slots + slot * SL;
        cplx* lv = pt + PTN;  // (loop hafnians only)
        // the lane's signs: bit p for row rd(p), bit T + q for column cd(q)
        uint32_t fm = 0;
#pragma unroll
        for (int p = 0; p < T; ++p) fm |= negbit(tu, L.rd(p), h) << p;
#pragma unroll
        for (int q = 0; q < T; ++q) fm |= negbit(tu, L.cd(q), h) << (T + q);

        // R_1 = D_d A (upper groups); V_1 = D_d A D_d into LDS whole
        double aR[T][T], aI[T][T];
#pragma unroll
        for (int p = 0; p < T; ++p)
#pragma unroll
            for (int q = p; q < T; ++q) {
                const cplx m = cflip(mget<RE>(Ms, swp(L.rd(p), h) * S + L.cd(q)), (fm >> p) & 1u);
                aR[p][q] = m.re;
                aI[p][q] = RE ? 0.0 : m.im;
            }
        sym_store<T, RE>(Pw, aR, aI, L, fm);
        // p_1 = tr(B) = sum_e d_e A[e][swap e] (lane partials, summed with power 1's results)
        cplx qo = cmk(0.0, 0.0);
        if (lane < n) qo = cmulr(mget<RE>(Ms, swp(lane, h) * S + swp(lane, h)), sgn(tu, lane, h));
        if (LOOP) {
            for (int e = lane; e < NP; e += WH_WAVE) Zb[e] = Dv[e];  // z_0 = D
        }
        wave_lds_sync();
        // powers up to A = ceil(h/2), software pipelined as sieve_pm1m.hip: the product forming
        // R_a runs with the epilogue of power a - 1; the accumulators alternate between two
        // register sets
        const int Ahalf = (h + 1) / 2;
        {
            double bR[T][T], bI[T][T];
            int a = 2;
#pragma unroll 1
            for (; a + 1 <= Ahalf; a += 2) {
                qo = product_step<T, LOOP, RE>(Pw, Ms, bR, bI, L, fm, aR, aI, a - 1, qo, lane, Zb + (a & 1) * NP,
                                               Zb + ((a - 1) & 1) * NP, tu, h, pt, lv);
                qo = product_step<T, LOOP, RE>(Pw, Ms, aR, aI, L, fm, bR, bI, a, qo, lane,
                                               Zb + ((a + 1) & 1) * NP, Zb + (a & 1) * NP, tu, h, pt, lv);
            }
            if (a <= Ahalf) {  // one power left: R_a into b, then its epilogue alone
                qo = product_step<T, LOOP, RE>(Pw, Ms, bR, bI, L, fm, aR, aI, a - 1, qo, lane, Zb + (a & 1) * NP,
                                               Zb + ((a - 1) & 1) * NP, tu, h, pt, lv);
                epis_alone<T, LOOP, RE>(bR, bI, a, qo, L, fm, lane, Pw, Ms, Zb + ((a - 1) & 1) * NP, Zb + (a & 1) * NP,
                                    tu, h, pt, lv);
            } else {  // the last power (a - 1) is in a
                epis_alone<T, LOOP, RE>(aR, aI, a - 1, qo, L, fm, lane, Pw, Ms, Zb + (a & 1) * NP,
                                    Zb + ((a - 1) & 1) * NP, tu, h, pt, lv);
            }
        }
        if (LOOP) {
            // v_1 = XD_d . D = sum_e d_e D[swap e] D[e] (every lane holds the sum)
            cplx part = cmk(0.0, 0.0);
            for (int e = lane; e < n; e += WH_WAVE)
                cmac(part, cmulr(ldc<RE>(&Dv[swp(e, h)]), sgn(tu, e, h)), ldc<RE>(&Dv[e]));
            const cplx v1 = wave_csum(part);
            if (lane == 0) lv[1] = v1;
        }
        ++slot;
        if (slot == NB || t + nw >= t1) {
            wave_lds_sync();  // pt and lv of every slot as the epilogues stored them
            if (lane < slot) {
                // coefficient of eta^h in exp(sum_i a_i eta^i), a_i = p_i / 2i (+ v_i / 2)
                const cplx* ptj = slots + lane * SL;
                const uint64_t tj = (uint64_t)(t_batch + (int64_t)lane * nw);
                const double prefac = ((h - __popcll(tj)) & 1) ? -1.0 : 1.0;
                const cplx top = coeff_top_s<4 * T, LOOP, RE>(ptj, ptj + PTN, LOOP ? ldc<RE>(ptj + PTN + 1) : cmk(0.0, 0.0), h);
                acc_sum.add(cmulr(top, prefac));
            }
            wave_lds_sync();  // the slots are free again
            slot = 0;
            t_batch = t + nw;
        }
    }

    // lanes 0 .. NB-1 hold the partial sums of their terms: fold them in lane order, then the
    // waves of this workgroup in wave order
    double4* lsum = reinterpret_cast<double4*>(slots);
    if (active && lane < NB) lsum[lane] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
    wave_lds_sync();
    if (active && lane == 0) {
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
            double4 pq = wres[w];
            s.add_pair(pq.x, pq.y, pq.z, pq.w);
        }
        partials[b] = make_double4(s.sr, s.cr, s.si, s.ci);
    }
}

}  // namespace

// The real plain instance at T = 3 (n = 17..24) fits 128 registers and, without the loop
// vectors and slots, 33 KB of LDS per 4-wave workgroup: four waves per SIMD instead of three.
template <int T, bool LOOP, bool RE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((RE && T == 3 && !LOOP) ? 4 : 1)))
wh_sieve_pm1s(const wh_job_dev* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
              double4* __restrict__ partials, int64_t block_offset) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    pm1s_body<T, LOOP, RE>(jobs[lo], pool, partials, b, wh_smem, blockDim.x / WH_WAVE);
}

// ---- Persistent form for a whole batch: the chunks of the symmetric-kernel groups with tile
// sides 4..6 (the per-class launches' blocks, the same chunking and the same per-chunk work,
// so every partial is bit-identical) are items of one queue, largest class first; each
// workgroup of the grid (as many as stay resident) takes items from an atomic counter until the
// queue is empty.  The last workgroup to finish (a second counter) folds every job of the batch
// in chunk order, as wh_fold does.  A one-thread kernel clears the two counters first (a kernel,
// not a memset, so the pair can sit in a captured graph).  Other kernel classes of the batch
// run before them on the same stream.
struct MixedItems {
    int ngroups;
    int njobs;  // jobs folded at the end (the whole batch)
    wh_mixed_group g[WH_MIXED_MAX_GROUPS];
};

template <int T, bool LOOP>
__device__ __noinline__ void mixed_item(const wh_job_dev* __restrict__ jobs, int lo, int hi, const cplx* __restrict__ pool,
                                        double4* __restrict__ partials, int64_t b, unsigned char* smem, int nw) {
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    pm1s_body<T, LOOP>(jobs[lo], pool, partials, b, smem, nw);
}

__global__ void __launch_bounds__(256)
wh_sieve_pm1s_persistent(const wh_job_dev* __restrict__ jobs, const cplx* __restrict__ pool,
                         double4* __restrict__ partials, unsigned* __restrict__ counters, double2* __restrict__ out,
                         MixedItems items) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    __shared__ int s_item;
    __shared__ int s_last;
    int total = 0;
    for (int k = 0; k < items.ngroups; ++k) total += (int)(items.g[k].block_end - items.g[k].block_begin);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&counters[0], 1u);
        __syncthreads();  // also: the previous item's last LDS reads are done
        int i = s_item;
        __syncthreads();
        if (i >= total) break;
        int k = 0;
        while (i >= (int)(items.g[k].block_end - items.g[k].block_begin)) {
            i -= (int)(items.g[k].block_end - items.g[k].block_begin);
            ++k;
        }
        const wh_mixed_group& G = items.g[k];
        const int64_t b = G.block_begin + i;
        const int lo = G.job_begin, hi = G.job_end - 1;
#define WH_MIXED_CASE(TT)                                                                   \
    case TT:                                                                                \
        if (G.loop) mixed_item<TT, true>(jobs, lo, hi, pool, partials, b, wh_smem, G.nw);  \
        else mixed_item<TT, false>(jobs, lo, hi, pool, partials, b, wh_smem, G.nw);        \
        break;
        switch (G.T) {
            WH_MIXED_CASE(1)
            WH_MIXED_CASE(2)
            WH_MIXED_CASE(3)
            WH_MIXED_CASE(4)
            WH_MIXED_CASE(5)
            WH_MIXED_CASE(6)
            default: break;
        }
#undef WH_MIXED_CASE
    }
    // the partials this workgroup wrote are visible device-wide before it counts itself done
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&counters[1], 1u) == gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    for (int q = threadIdx.x; q < items.njobs; q += blockDim.x) wh_fold_job(jobs[q], partials, out);
}

__global__ void wh_clear_counters(unsigned* __restrict__ counters) { counters[threadIdx.x] = 0u; }

size_t wh_pm1s_lds_bytes(int T, int nw, bool real, bool loop) {
    size_t NP = 8 * (size_t)T, S = NP + WH_PM1S_PAD;
    size_t PTN = 4 * (size_t)T + 4, NB = T >= 6 ? 1 : T == 5 ? 7 : 8;
    size_t mat = (NP * S * (real ? sizeof(double) : sizeof(cplx)) + 31) / 32 * 32;
    size_t dvn = loop ? NP : 0, sl = loop ? 2 * PTN : PTN;
    size_t small = (NB * sl + 2 * dvn) * sizeof(cplx);
    return (1 + nw) * mat + dvn * sizeof(cplx) + nw * small + nw * sizeof(double4) + 32;
}

// the dynamic-LDS limit set on each instance (shared by its launches and occupancy queries, so
// that a query for a smaller workgroup never lowers what a launch relies on)
template <int T, bool LOOP, bool RE>
static wh_lds_limit& lds_limit_s() {
    static wh_lds_limit lim;
    return lim;
}

template <int T, bool LOOP, bool RE>
static hipError_t launch_s(const wh_job_dev* jobs, int njobs, const void* pool, void* partials, int64_t nblocks,
                           int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = wh_pm1s_lds_bytes(T, nw, RE, LOOP);
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_pm1s<T, LOOP, RE>, lds, lds_limit_s<T, LOOP, RE>());
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_pm1s<T, LOOP, RE>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds, stream,
                       jobs, njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

hipError_t wh_launch_pm1s_persistent(const wh_mixed_group* groups, int ngroups, const wh_job_dev* jobs, int njobs,
                                     const void* pool, void* partials, unsigned* counters, void* out, int nwg,
                                     hipStream_t stream) {
    if (ngroups < 1 || ngroups > WH_MIXED_MAX_GROUPS || nwg < 1) return hipErrorInvalidValue;
    MixedItems items;
    items.ngroups = ngroups;
    items.njobs = njobs;
    size_t lds = 0;
    for (int k = 0; k < ngroups; ++k) {
        items.g[k] = groups[k];
        if (groups[k].T < 1 || groups[k].T > 6 || groups[k].nw < 1 || groups[k].nw > 4) return hipErrorInvalidValue;
        lds = std::max(lds, wh_pm1s_lds_bytes(groups[k].T, groups[k].nw, false, true));
    }
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_pm1s_persistent, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(wh_clear_counters, dim3(1), dim3(2), 0, stream, counters);
    hipLaunchKernelGGL(wh_sieve_pm1s_persistent, dim3((unsigned)nwg), dim3(256), lds, stream, jobs,
                       (const cplx*)pool, (double4*)partials, counters, (double2*)out, items);
    return hipGetLastError();
}

template <int T, bool LOOP, bool RE>
static int resident_s(int nw) {
    int blocks = 0;
    const size_t lds = wh_pm1s_lds_bytes(T, nw, RE, LOOP);
    if (lds > 160 * 1024) return 0;  // the CU's LDS: no workgroup of this size fits
    // a failed query leaves no error behind for the caller's next HIP call to report
    if (wh_raise_lds_limit((const void*)wh_sieve_pm1s<T, LOOP, RE>, lds, lds_limit_s<T, LOOP, RE>()) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void*)wh_sieve_pm1s<T, LOOP, RE>, nw * WH_WAVE,
                                                     lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return blocks;
}

int wh_pm1s_resident(int T, int nw, bool real, bool loop) {
    if (nw < 1 || nw > 4) return 0;
#define WH_PM1S_RES(TT)                                                                        \
    case TT:                                                                                   \
        return loop ? resident_s<TT, true, false>(nw) : resident_s<TT, false, false>(nw);
#define WH_PM1S_RES_R(TT)                                                                      \
    case TT:                                                                                   \
        if (real) return loop ? resident_s<TT, true, true>(nw) : resident_s<TT, false, true>(nw); \
        return loop ? resident_s<TT, true, false>(nw) : resident_s<TT, false, false>(nw);
    switch (T) {
        WH_PM1S_RES_R(1)
        WH_PM1S_RES_R(2)
        WH_PM1S_RES_R(3)
        WH_PM1S_RES_R(4)
        WH_PM1S_RES_R(5)
        WH_PM1S_RES(6)
        WH_PM1S_RES(7)
        WH_PM1S_RES(8)
        default: return 0;
    }
#undef WH_PM1S_RES
#undef WH_PM1S_RES_R
}

hipError_t wh_launch_pm1s(int T, bool loop, bool real, const wh_job_dev* jobs, int njobs, const void* pool,
                          void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
#define WH_PM1S_ARGS jobs, njobs, pool, partials, nblocks, block_offset, nw, stream
#define WH_PM1S_CASE(TT)                                                              \
    case TT:                                                                          \
        return loop ? launch_s<TT, true, false>(WH_PM1S_ARGS) : launch_s<TT, false, false>(WH_PM1S_ARGS);
#define WH_PM1S_CASE_R(TT)                                                            \
    case TT:                                                                          \
        if (real) return loop ? launch_s<TT, true, true>(WH_PM1S_ARGS) : launch_s<TT, false, true>(WH_PM1S_ARGS); \
        return loop ? launch_s<TT, true, false>(WH_PM1S_ARGS) : launch_s<TT, false, false>(WH_PM1S_ARGS);
    switch (T) {
        WH_PM1S_CASE_R(1)
        WH_PM1S_CASE_R(2)
        WH_PM1S_CASE_R(3)
        WH_PM1S_CASE_R(4)
        WH_PM1S_CASE_R(5)
        WH_PM1S_CASE(6)
        WH_PM1S_CASE(7)
        WH_PM1S_CASE(8)
        default: return hipErrorInvalidValue;
    }
#undef WH_PM1S_CASE
#undef WH_PM1S_CASE_R
#undef WH_PM1S_ARGS
}
