// wh_device.h -- device-side types and helpers shared by the sieve kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WH_WAVE 64
#define WH_EMAX 32    // max pairing edges per job (matrix side 2E <= 64)
#define WH_NMAX 128   // max photon number N = sum(rpt) handled on the device

// One hafnian request as the kernels see it: the reordered matrix Ax (2E x 2E, row-major,
// interleaved complex128), optional loop weights, edge multiplicities and the term range.
// It is the device form of the arguments of _calc_hafnian / _calc_loop_hafnian
// (thewalrus/_hafnian.py:417, :513).
struct __attribute__((aligned(16))) wh_job_dev {
    int64_t a_off;        // element offset of Ax in the complex pool
    int64_t d_off;        // element offset of D (loop) or -1
    int64_t v_off;        // element offset of oddV (odd-N loop) or -1
    int64_t steps;        // number of sieve terms
    int64_t chunk_base;   // index of this job's first partial-sum slot
    int64_t chunk_terms;  // terms per chunk (one workgroup per chunk)
    int64_t out_index;    // position of the result in the caller's output array
    double oddloop_re, oddloop_im;
    int32_t E, N;         // edges, photons (N = 2 sum(e) [+1 if odd])
    int32_t flags;        // WH_F_* bits
    int32_t nchunks;
    int32_t e[WH_EMAX];   // edge multiplicities
};

#define WH_F_LOOP 1
#define WH_F_ODD 2
#define WH_F_GLYNN 4
#define WH_F_REAL 8
#define WH_F_MTL 16   // montrealer / loop montrealer terms (PM1 kernel only)
#define WH_F_CHARPOLY 32  // general kernel: traces above p_(m-1) by the characteristic polynomial

// One loop_hafnian_batch request (thewalrus/loop_hafnian_batch.py:52, :128; with nD > 1
// loop_hafnian_batch_gamma.py:53, :132): the reordered matrix and nD diagonals, edge
// multiplicities with the batched edge first (and the odd-mode edge second for the odd
// variant), and nD x L outputs, one per diagonal and photon number of the batched mode.
struct __attribute__((aligned(16))) wh_lhb_job {
    int64_t a_off;        // element offset of Ax (2E x 2E) in the complex pool
    int64_t d_off;        // element offset of Dx (nD rows of 2E)
    int64_t steps;        // prod(e_i + 1) terms (no Glynn halving in the batched sieve)
    int64_t chunk_base;   // index of this job's first workgroup
    int64_t chunk_terms;  // terms per workgroup
    int64_t part_off;     // first partial slot (nchunks x L slots)
    int64_t out_off;      // first output element
    int32_t E, L;         // edges, outputs
    int32_t n_fixed;      // photons fixed by the other modes (2 sum(fixed) [+1 odd variant])
    int32_t n_max;        // n_fixed + L - 1
    int32_t flags;        // WH_LHB_* bits
    int32_t nchunks;
    int32_t nD;           // diagonals sharing Ax (loop_hafnian_batch_gamma: one per fanout row)
    int32_t dsplit;       // workgroups per chunk, each taking a contiguous group of the diagonals
    int32_t e[WH_EMAX];
};

// One kernel-class group of a batch for the persistent symmetric-kernel launch: its partial
// slots [block_begin, block_end) and jobs [job_begin, job_end) (sorted by chunk_base), tile side
// T, loop flag and waves per workgroup.
#define WH_MIXED_MAX_GROUPS 12
struct wh_mixed_group {
    int64_t block_begin, block_end;
    int32_t job_begin, job_end;
    int32_t T, loop, nw, pad;
};

#define WH_LHB_ODD 1
#define WH_LHB_GLYNN 2

struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmk(double r, double i) { cplx z; z.re = r; z.im = i; return z; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cplx cneg(cplx a) { return cmk(-a.re, -a.im); }
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ __forceinline__ cplx cmulr(cplx a, double d) { return cmk(a.re * d, a.im * d); }
__device__ __forceinline__ cplx cdivr(cplx a, double d) { return cmk(a.re / d, a.im / d); }
__device__ __forceinline__ cplx cconj(cplx a) { return cmk(a.re, -a.im); }
__device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
// acc += a * b
__device__ __forceinline__ void cmac(cplx& acc, cplx a, cplx b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

// Error-free sum (Knuth TwoSum): s + err == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& err) {
    s = a + b;
    double bb = s - a;
    err = (a - (s - bb)) + (b - bb);
}

// Compensated (Neumaier-style) complex accumulator for the alternating-sign term sum.
struct csum {
    double sr, si, cr, ci;
    __device__ __forceinline__ void zero() { sr = si = cr = ci = 0.0; }
    __device__ __forceinline__ void add(cplx x) {
        double s, e;
        two_sum(sr, x.re, s, e); sr = s; cr += e;
        two_sum(si, x.im, s, e); si = s; ci += e;
    }
    __device__ __forceinline__ void add_pair(double r, double rc, double i, double ic) {
        double s, e;
        two_sum(sr, r, s, e); sr = s; cr += e + rc;
        two_sum(si, i, s, e); si = s; ci += e + ic;
    }
};

// ---- cross-lane sums on the VALU (DPP within 16-lane rows, permlane swaps across rows).
// Every step adds a lane's value to its partner's; partners add the same two numbers, so all
// lanes of a group end with bit-identical sums.  Callers keep the whole wave active.

// Every control used here has a source lane inside the row for every lane, so bound_ctrl (a
// lane without a source reads 0) never applies and the move needs no initialised destination.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
#define WH_DPP_XOR1 0xB1        // quad_perm [1,0,3,2]
#define WH_DPP_XOR2 0x4E        // quad_perm [2,3,0,1]
#define WH_DPP_HALF_MIRROR 0x141 // lane i <-> 7 - i within 8
#define WH_DPP_MIRROR 0x140     // lane i <-> 15 - i within 16
#define WH_DPP_ROR8 0x128       // lane i <-> i ^ 8 within 16

// v + the value of lane ^ 16 (rows 0<->1, 2<->3)
__device__ __forceinline__ double swap16_sum(double v) {
    unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    double x = __hiloint2double((int)b[0], (int)a[0]);
    double y = __hiloint2double((int)b[1], (int)a[1]);
    return x + y;
}
// v + the value of lane ^ 32
__device__ __forceinline__ double swap32_sum(double v) {
    unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    double x = __hiloint2double((int)b[0], (int)a[0]);
    double y = __hiloint2double((int)b[1], (int)a[1]);
    return x + y;
}

// sum over the 8 lanes with the same lane / 8 (lanes 8g .. 8g + 7)
__device__ __forceinline__ double group8_sum(double v) {
    v += dpp_d<WH_DPP_XOR1>(v);
    v += dpp_d<WH_DPP_XOR2>(v);
    v += dpp_d<WH_DPP_HALF_MIRROR>(v);
    return v;
}
// sum over the 8 lanes with the same lane % 8 (lanes c, c + 8, ..., c + 56)
__device__ __forceinline__ double stride8_sum(double v) {
    v += dpp_d<WH_DPP_ROR8>(v);
    v = swap16_sum(v);
    return swap32_sum(v);
}
// Sum over the whole wave; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v) {
    v = group8_sum(v);
    v += dpp_d<WH_DPP_MIRROR>(v);
    v = swap16_sum(v);
    return swap32_sum(v);
}
__device__ __forceinline__ cplx wave_csum(cplx z) { return cmk(wave_sum(z.re), wave_sum(z.im)); }

__device__ __forceinline__ double dswap32(double a, double b) {  // lower half: a + a^32; upper: b + b^32
    auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ double dswap16(double a, double b) {  // even rows: a + a^16; odd rows: b + b^16
    auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// Eight wave sums at once: lane l returns the sum over all lanes of v[(l >> 3) & 7].  Each
// butterfly stage halves the number of values a lane carries (the swaps move two values per
// instruction pair), 34 instructions where eight separate wave sums take about 140.
__device__ __forceinline__ double wave_sum8(const double (&v)[8], int lane) {
    double s1[4], s2[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) s1[k] = dswap32(v[k], v[k + 4]);  // value k + 4 bit5
#pragma unroll
    for (int k = 0; k < 2; ++k) s2[k] = dswap16(s1[k], s1[k + 2]);  // value k + 2 bit4 + 4 bit5
    const bool b3 = (lane >> 3) & 1;
    double keep = b3 ? s2[1] : s2[0], send = b3 ? s2[0] : s2[1];
    double s3 = keep + dpp_d<WH_DPP_ROR8>(send);  // value bit3 + 2 bit4 + 4 bit5
    return group8_sum(s3);
}

// Lanes of one wave exchanging data through LDS: the hardware completes a wave's LDS operations
// in issue order, but the compiler reasons per lane and may move a load above another lane's
// store (or forward a conditional store to its own lane only).  The wavefront-scope fences keep
// every LDS access on its side of this point; they emit no instruction.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Binomial coefficient as the reference's table holds it (_hafnian.py:27-45), in double,
// exact in 64-bit: each step divides out the common factor first, so no
// intermediate exceeds the result (C(64, 32) < 2^64 covers every multiplicity up to WH_NMAX / 2).
__device__ __forceinline__ double binom_d(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    if (k == 0 || k == n) return 1.0;
    int kk = k < n - k ? k : n - k;
    uint64_t b = 1;
    for (int i = 0; i < kk; ++i) {
        uint64_t num = (uint64_t)(n - i), den = (uint64_t)(i + 1);
        uint64_t x = b, y = den;
        while (y) { uint64_t r = x % y; x = y; y = r; }
        b /= x;
        b *= num / (den / x);
    }
    return (double)b;
}

// Folds one job's chunk partials in chunk order with compensation and applies the sieve's
// power-of-two scale (_hafnian.py:464-465, :571-575) and the montrealer's sign.
__device__ __forceinline__ void wh_fold_job(const wh_job_dev& J, const double4* __restrict__ partials,
                                            double2* __restrict__ out) {
    csum acc;
    acc.zero();
    for (int c = 0; c < J.nchunks; ++c) {
        double4 p = partials[J.chunk_base + c];
        acc.add_pair(p.x, p.y, p.z, p.w);
    }
    double re = acc.sr + acc.cr, im = acc.si + acc.ci;
    if ((J.flags & WH_F_MTL) && (J.E & 1) == 0) {  // (-1)^(n+1), _montrealer.py:57, :101
        re = -re;
        im = -im;
    }
    if (J.flags & WH_F_GLYNN) {
        int ex = (J.flags & WH_F_ODD) ? J.N / 2 : J.N / 2 - 1;
        double sc = ldexp(1.0, -ex);
        re *= sc;
        im *= sc;
    }
    out[J.out_index] = make_double2(re, im);
}
