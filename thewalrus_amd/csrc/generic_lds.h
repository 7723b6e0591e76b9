// generic_lds.h -- LDS layout and per-term building blocks shared by the general sieve
// kernels (sieve_generic.hip, sieve_lhb.hip): one wave per workgroup, the term matrix, its
// powers and the coefficient recursion in LDS.  Restates charpoly.py:113-327 on the device.
#pragma once
#include "wh_device.h"

namespace wh_generic {

struct GenLds {
    cplx* H;     // term matrix AX_S (m x m, row stride MM); Hessenberg form after charpoly
    cplx* P;     // current power / La Budde coefficient table
    cplx* Q;     // next power
    cplx* XD;    // loop: delta-scaled X D
    cplx* D;     // loop: D_S, later the column chain A^k D (odd N)
    cplx* OV;    // odd: delta-scaled oddV X
    cplx* v;     // work vector (Householder reflector, mat-vec chains)
    cplx* w;     // work vector
    cplx* pt;    // power traces p[0..h]
    cplx* lv;    // loop scalars, index by order
    cplx* ov;    // odd-loop scalars
    cplx* comb;  // 2 coefficient rows of ncomb entries
    int ncomb;   // coefficient row length (orders 0 .. N need N + 1 <= ncomb)
    int* S;      // kept edge indices
    int* dS;     // their delta values
};

// NQ = (largest photon number N of the workgroup's requests) + 2: power traces and loop
// scalars need indices 0 .. N / 2 + 1, coefficient rows 0 .. N.  The default covers every N.
__host__ __device__ constexpr size_t generic_pts(int NQ) { return (size_t)(NQ - 2) / 2 + 4; }

// Sides up to this form the powers in place (matmul_inplace_lds) and carve no Q.
constexpr int kGenericInplaceSide = 16;

// Bytes of the layout carve() lays out (a multiple of 16).
__host__ __device__ constexpr size_t generic_lds_bytes(int MM, int NQ = WH_NMAX + 2, bool with_q = true) {
    return ((with_q ? 3 : 2) * (size_t)MM * MM + 5 * (size_t)MM + 3 * generic_pts(NQ) + 2 * (size_t)NQ) *
               sizeof(cplx) +
           2 * WH_EMAX * sizeof(int);
}

__device__ __forceinline__ GenLds carve(unsigned char* base, int MM, int NQ = WH_NMAX + 2, bool with_q = true) {
    GenLds L;
    cplx* c = reinterpret_cast<cplx*>(base);
    size_t mm = (size_t)MM * MM;
    L.H = c; c += mm;
    L.P = c; c += mm;
    L.Q = nullptr;
    if (with_q) { L.Q = c; c += mm; }
    L.XD = c; c += MM;
    L.D = c; c += MM;
    L.OV = c; c += MM;
    L.v = c; c += MM;
    L.w = c; c += MM;
    L.pt = c; c += generic_pts(NQ);
    L.lv = c; c += generic_pts(NQ);
    L.ov = c; c += generic_pts(NQ);
    L.comb = c; c += 2 * (size_t)NQ;
    L.ncomb = NQ;
    L.S = reinterpret_cast<int*>(c);
    L.dS = L.S + WH_EMAX;
    return L;
}

// Q = P @ H (m x m), one output element per lane per step, k summed in order.
__device__ void matmul_lds(const cplx* P, const cplx* H, cplx* Q, int m, int MM, int lane) {
    for (int e = lane; e < m * m; e += WH_WAVE) {
        int i = e / m, j = e - i * m;
        cplx s = cmk(0.0, 0.0);
        for (int k = 0; k < m; ++k) cmac(s, P[i * MM + k], H[k * MM + j]);
        Q[i * MM + j] = s;
    }
}

// P = cur @ H for m <= kGenericInplaceSide (cur may be P): each lane keeps its outputs (at most
// 4) in registers until every lane has read cur; k summed in order, as matmul_lds.
__device__ void matmul_inplace_lds(const cplx* cur, const cplx* H, cplx* P, int m, int MM, int lane) {
    cplx o[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = lane + u * WH_WAVE;
        cplx s = cmk(0.0, 0.0);
        if (e < m * m) {
            const int i = e / m, j = e - i * m;
            for (int k = 0; k < m; ++k) cmac(s, cur[i * MM + k], H[k * MM + j]);
        }
        o[u] = s;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = lane + u * WH_WAVE;
        if (e < m * m) {
            const int i = e / m, j = e - i * m;
            P[i * MM + j] = o[u];
        }
    }
}

// x^T P y over the m x m matrix P (row stride MM), every lane holding a few entries; the whole
// wave receives the sum.
__device__ cplx bilinear_lds(const cplx* x, const cplx* P, const cplx* y, int m, int MM, int lane) {
    cplx s = cmk(0.0, 0.0);
    for (int e = lane; e < m * m; e += WH_WAVE) {
        const int i = e / m, j = e - i * m;
        cmac(s, cmul(x[i], P[i * MM + j]), y[j]);
    }
    return wave_csum(s);
}

__device__ cplx trace_lds(const cplx* X, int m, int MM, int lane) {
    cplx z = cmk(0.0, 0.0);
    if (lane < m) z = X[lane * MM + lane];
    return wave_csum(z);
}

// charpoly.py:113-125 + :43-109: in-place Householder reduction to upper Hessenberg form.
__device__ void hessenberg_lds(GenLds& L, int m, int MM, int lane) {
    cplx* H = L.H;
    for (int k = 1; k < m - 1; ++k) {
        int sizeH = m - k;
        // reflection vector (charpoly.py:55-72)
        cplx col = cmk(0.0, 0.0);
        if (lane < sizeH) col = H[(k + lane) * MM + (k - 1)];
        double nrm2 = wave_sum(lane < sizeH ? cabs2(col) : 0.0);
        cplx c0 = H[k * MM + (k - 1)];
        cplx sigma = cmk(sqrt(nrm2), 0.0);
        if (c0.re != 0.0 || c0.im != 0.0) {
            double a0 = sqrt(cabs2(c0));
            sigma = cmul(sigma, cdivr(c0, a0));
        }
        if (lane < sizeH) L.v[lane] = (lane == 0) ? cadd(col, sigma) : col;
        __syncthreads();
        // apply_householder (charpoly.py:86-109)
        double nv = sqrt(wave_sum(lane < sizeH ? cabs2(L.v[lane]) : 0.0));
        double nv2 = nv * nv;
        if (nv2 == 0.0) continue;
        int ncol = m - k + 1;
        if (lane < ncol) {
            cplx s = cmk(0.0, 0.0);
            for (int l = 0; l < sizeH; ++l) cmac(s, cconj(L.v[l]), H[(k + l) * MM + (k - 1 + lane)]);
            L.w[lane] = s;
        }
        __syncthreads();
        for (int e = lane; e < sizeH * ncol; e += WH_WAVE) {
            int i = e / ncol, j = e - i * ncol;
            cplx t = cdivr(cmul(cmulr(L.v[i], 2.0), L.w[j]), nv2);
            cplx& a = H[(k + i) * MM + (k - 1 + j)];
            a = csub(a, t);
        }
        __syncthreads();
        cplx av = cmk(0.0, 0.0);
        if (lane < m)
            for (int l = 0; l < sizeH; ++l) cmac(av, H[lane * MM + (k + l)], L.v[l]);
        __syncthreads();
        if (lane < m) L.w[lane] = av;
        __syncthreads();
        for (int e = lane; e < m * sizeH; e += WH_WAVE) {
            int i = e / sizeH, j = e - i * sizeH;
            cplx t = cdivr(cmul(cmulr(L.w[i], 2.0), cconj(L.v[j])), nv2);
            cplx& a = H[i * MM + (k + j)];
            a = csub(a, t);
        }
        __syncthreads();
    }
}

// charpoly.py:196-282 La Budde on the Hessenberg matrix in L.H; coefficient table in L.P.
// Row i of the table depends only on earlier rows, so the lanes fill one row in parallel.
__device__ void labudde_lds(GenLds& L, int m, int MM, int lane) {
    const cplx* H = L.H;
    cplx* C = L.P;
#define HH(i, j) H[((i) - 1) * MM + ((j) - 1)]
#define CC(i, j) C[((i) - 1) * MM + ((j) - 1)]
    if (lane == 0) {
        CC(1, 1) = cneg(HH(1, 1));
        if (m >= 2) {
            CC(2, 1) = csub(CC(1, 1), HH(2, 2));
            CC(2, 2) = csub(cmul(HH(1, 1), HH(2, 2)), cmul(HH(1, 2), HH(2, 1)));
        }
    }
    __syncthreads();
    for (int i = 3; i <= m; ++i) {
        int j = lane + 1;
        if (j <= i) {
            cplx alpha = HH(i, i);
            if (j == 1) {
                CC(i, 1) = csub(CC(i - 1, 1), alpha);
            } else if (j < i) {
                cplx suma = cmk(0.0, 0.0), bp = cmk(1.0, 0.0);
                for (int mm = 1; mm < j - 1; ++mm) {
                    bp = cmul(bp, HH(i - mm + 1, i - mm));
                    suma = cadd(suma, cmul(cmul(HH(i - mm, i), bp), CC(i - mm - 1, j - mm - 1)));
                }
                cplx bpj = cmk(1.0, 0.0);
                for (int bm = i; bm > i - j + 1; --bm) bpj = cmul(bpj, HH(bm, bm - 1));
                cplx v = csub(CC(i - 1, j), cmul(alpha, CC(i - 1, j - 1)));
                v = csub(v, suma);
                CC(i, j) = csub(v, cmul(HH(i - j + 1, i), bpj));
            } else {
                cplx suma = cmk(0.0, 0.0), bp = cmk(1.0, 0.0);
                for (int mm = 1; mm < i - 1; ++mm) {
                    bp = cmul(bp, HH(i - mm + 1, i - mm));
                    suma = cadd(suma, cmul(cmul(HH(i - mm, i), bp), CC(i - mm - 1, i - mm - 1)));
                }
                cplx bpi = cmk(1.0, 0.0);
                for (int bm = i; bm > 1; --bm) bpi = cmul(bpi, HH(bm, bm - 1));
                cplx v = csub(cneg(cmul(alpha, CC(i - 1, i - 1))), suma);
                CC(i, i) = csub(v, cmul(HH(1, i), bpi));
            }
        }
        __syncthreads();
    }
#undef HH
#undef CC
}

}  // namespace wh_generic
