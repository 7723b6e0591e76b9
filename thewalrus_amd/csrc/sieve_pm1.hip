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

// sign of column c of the current term: delta_(c mod h) = +1 if digit set, else -1
__device__ __forceinline__ double col_sign(uint64_t t, int c, int h) {
    int i = c < h ? c : c - h;
    if (i >= h) return 1.0;  // zero padding
    return ((t >> (h - 1 - i)) & 1ull) ? 1.0 : -1.0;
}

template <int T, bool REAL, bool LOOP>
__device__ __forceinline__ void pm1_body(const wh_job_dev& J, const cplx* __restrict__ pool,
                                         double4* __restrict__ partials, int64_t b,
                                         unsigned char* smem) {
    using E = typename Elem<REAL>::T;
    constexpr int NP = 8 * T;  // padded side
    constexpr int S = NP;      // LDS row stride (elements)
    const int nw = blockDim.x / WH_WAVE;
    const int wave = threadIdx.x / WH_WAVE;
    const int lane = threadIdx.x % WH_WAVE;
    const int ti = lane >> 3, tj = lane & 7;
    const int n = 2 * J.E;
    const int h = J.E;

    // ---- LDS carve
    E* Ms = reinterpret_cast<E*>(smem);
    E* Pw = Ms + NP * S + wave * NP * S;
    E* vecs = Ms + (1 + nw) * NP * S;           // XD0[NP], D[NP] (loop)
    E* XD0 = vecs;
    E* Dv = vecs + NP;
    cplx* small = reinterpret_cast<cplx*>(vecs + 2 * NP) + wave * (4 * (WH_EMAX + 2));
    cplx* pt = small;                            // p[0..h]
    cplx* lv = pt + (WH_EMAX + 2);               // loop scalars v[1..h]
    E* comb0 = reinterpret_cast<E*>(lv + (WH_EMAX + 2));
    E* comb1 = comb0 + (WH_EMAX + 2);
    double4* wres = reinterpret_cast<double4*>(reinterpret_cast<cplx*>(vecs + 2 * NP) +
                                               nw * (4 * (WH_EMAX + 2)));

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
                int src = e < h ? e + h : e - h;  // (X D)[e]
                xd = eload(Dg, src, (E*)nullptr);
                dd = eload(Dg, e, (E*)nullptr);
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

    for (int64_t t = t0 + wave; t < t1; t += nw) {
        const uint64_t tu = (uint64_t)t;
        double sgnJ[T];
#pragma unroll
        for (int c = 0; c < T; ++c) sgnJ[c] = col_sign(tu, J0 + c, h);
        const double prefac = ((h - __popcll(tu)) & 1) ? -1.0 : 1.0;

        // P_1 = M diag(d)
        E acc[T][T];
#pragma unroll
        for (int r = 0; r < T; ++r)
#pragma unroll
            for (int c = 0; c < T; ++c) acc[r][c] = escale(Ms[(I0 + r) * S + J0 + c], sgnJ[c]);

        // loop: per-lane row weights XD_d[I] and column weights D[J]
        E xdI[T], dJ[T];
        if (LOOP) {
#pragma unroll
            for (int r = 0; r < T; ++r) xdI[r] = escale(XD0[I0 + r], col_sign(tu, I0 + r, h));
#pragma unroll
            for (int c = 0; c < T; ++c) dJ[c] = Dv[J0 + c];
        }

        for (int a = 1; a <= h; ++a) {
            // trace of P_a (diagonal tiles) and, for the loop, XD_d P_(a-1) D (a >= 2 uses
            // the power held in acc one step earlier; v_1 = XD_d . D is taken below)
            E tr;
            ezero(tr);
            if (ti == tj) {
#pragma unroll
                for (int r = 0; r < T; ++r) tr = eadd(tr, acc[r][r]);
            }
            E bl;
            ezero(bl);
            if (LOOP && a <= h - 1) {
#pragma unroll
                for (int r = 0; r < T; ++r) {
                    E rowsum;
                    ezero(rowsum);
#pragma unroll
                    for (int c = 0; c < T; ++c) emac(rowsum, acc[r][c], dJ[c]);
                    emac(bl, xdI[r], rowsum);
                }
            }
            cplx trc = etoc(tr), blc = etoc(bl);
            trc = wave_csum(trc);
            if (LOOP) blc = wave_csum(blc);
            if (lane == 0) {
                pt[a] = trc;
                if (LOOP && a <= h - 1) lv[a + 1] = blc;
            }
            if (a == h) break;
            // P_(a+1) = (P_a M) diag(d)
#pragma unroll
            for (int r = 0; r < T; ++r)
#pragma unroll
                for (int c = 0; c < T; ++c) Pw[(I0 + r) * S + J0 + c] = acc[r][c];
            E nacc[T][T];
#pragma unroll
            for (int r = 0; r < T; ++r)
#pragma unroll
                for (int c = 0; c < T; ++c) ezero(nacc[r][c]);
#pragma unroll 2
            for (int k = 0; k < n; ++k) {
                E av[T], bv[T];
#pragma unroll
                for (int r = 0; r < T; ++r) av[r] = Pw[(I0 + r) * S + k];
#pragma unroll
                for (int c = 0; c < T; ++c) bv[c] = Ms[k * S + J0 + c];
#pragma unroll
                for (int r = 0; r < T; ++r)
#pragma unroll
                    for (int c = 0; c < T; ++c) emac(nacc[r][c], av[r], bv[c]);
            }
#pragma unroll
            for (int r = 0; r < T; ++r)
#pragma unroll
                for (int c = 0; c < T; ++c) acc[r][c] = escale(nacc[r][c], sgnJ[c]);
        }
        if (LOOP) {
            // v_1 = XD_d . D
            E part;
            ezero(part);
            for (int e = lane; e < n; e += WH_WAVE) emac(part, escale(XD0[e], col_sign(tu, e, h)), Dv[e]);
            cplx v1 = wave_csum(etoc(part));
            if (lane == 0) lv[1] = v1;
        }

        // coefficient of eta^h in exp(sum_i a_i eta^i) (_hafnian.py:200-208, :232-241); the
        // lanes hold the coefficients 0..h and add the shifted old row in the reference's j order
        E* oldc = comb0;
        E* newc = comb1;
        if (lane <= h) oldc[lane] = lane == 0 ? eone((E*)nullptr) : ezero_v((E*)nullptr);
        for (int i = 1; i <= h; ++i) {
            E f = ediv(eget(pt[i], (E*)nullptr), (double)(2 * i));
            if (LOOP) f = eadd(f, ediv(eget(lv[i], (E*)nullptr), 2.0));
            if (lane <= h) {
                E val = oldc[lane];
                E pf = eone((E*)nullptr);
                for (int j = 1; j <= h / i; ++j) {
                    pf = emul(pf, ediv(f, (double)j));
                    if (i * j <= lane) val = eadd(val, emul(oldc[lane - i * j], pf));
                }
                newc[lane] = val;
            }
            E* tmp = oldc;
            oldc = newc;
            newc = tmp;
        }
        if (lane == h) acc_sum.add(cmulr(etoc(oldc[h]), prefac));
    }

    // fold the waves of this workgroup in wave order
    if (lane == h) wres[wave] = make_double4(acc_sum.sr, acc_sum.cr, acc_sum.si, acc_sum.ci);
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

}  // namespace

template <int T, bool REAL, bool LOOP>
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
    pm1_body<T, REAL, LOOP>(jobs[lo], pool, partials, b, wh_smem);
}

size_t wh_pm1_lds_bytes(int T, bool real, int nw) {
    size_t esz = real ? sizeof(double) : sizeof(cplx);
    size_t NP = 8 * (size_t)T;
    return (1 + nw) * NP * NP * esz + 2 * NP * esz + nw * 4 * (WH_EMAX + 2) * sizeof(cplx) +
           nw * sizeof(double4);
}

template <int T, bool REAL, bool LOOP>
static hipError_t launch_t(const wh_job_dev* jobs, int njobs, const void* pool, void* partials,
                           int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream) {
    size_t lds = wh_pm1_lds_bytes(T, REAL, nw);
    hipError_t err = hipFuncSetAttribute((const void*)wh_sieve_pm1<T, REAL, LOOP>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL((wh_sieve_pm1<T, REAL, LOOP>), dim3((unsigned)nblocks), dim3(nw * WH_WAVE), lds,
                       stream, jobs, njobs, (const cplx*)pool, (double4*)partials, block_offset);
    return hipGetLastError();
}

template <int T>
static hipError_t launch_rl(bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                            void* partials, int64_t nblocks, int64_t off, int nw, hipStream_t s) {
    if (real) {
        if (loop) return launch_t<T, true, true>(jobs, njobs, pool, partials, nblocks, off, nw, s);
        return launch_t<T, true, false>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    }
    if (loop) return launch_t<T, false, true>(jobs, njobs, pool, partials, nblocks, off, nw, s);
    return launch_t<T, false, false>(jobs, njobs, pool, partials, nblocks, off, nw, s);
}

hipError_t wh_launch_pm1(int T, bool real, bool loop, const wh_job_dev* jobs, int njobs,
                         const void* pool, void* partials, int64_t nblocks, int64_t block_offset,
                         int nw, hipStream_t stream) {
    switch (T) {
        case 1: return launch_rl<1>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 2: return launch_rl<2>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 3: return launch_rl<3>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 4: return launch_rl<4>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 5: return launch_rl<5>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        case 6: return launch_rl<6>(real, loop, jobs, njobs, pool, partials, nblocks, block_offset, nw, stream);
        default: return hipErrorInvalidValue;
    }
}
