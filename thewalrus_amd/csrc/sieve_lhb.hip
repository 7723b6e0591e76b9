// sieve_lhb.hip -- batched loop hafnian: one sieve over the fixed modes' edges and the batched
// mode's edge gives the loop hafnian for every photon number 0..N_cutoff of the batched mode
// at once (the core of Gaussian boson sampling's chain rule).  Restates
// _calc_loop_hafnian_batch_even (thewalrus/loop_hafnian_batch.py:52-123) and
// _calc_loop_hafnian_batch_odd (:128-208) per term, on the general kernel's LDS layout, and
// with nD diagonals sharing the matrix _calc_loop_hafnian_batch_gamma_even / _odd
// (loop_hafnian_batch_gamma.py:53-129, :132-219).
//
// One wave per workgroup walks one chunk of terms in order.  Per term it builds AX_S and
// oddVX_S in LDS and takes the power traces once up to n_max / 2 (the reference calls
// powertrace on the same matrix for every coefficient series it needs).  Then, per diagonal,
// XD_S and D_S, the two loop-scalar chains of f_loop (row) and f_loop_odd (column), and the
// coefficient series f_loop(n_max), f_loop_odd(n_max) and, for the odd variant's terms with
// k0 = k1 = 0, f_loop_odd(n_fixed) with the other self-edge (get_submatrix_batch_odd0,
// _hafnian.py:360-380).  Output o = kd L + q keeps its compensated sum on lane o mod 64
// (nD L <= 128).
#include "generic_lds.h"
using namespace wh_generic;

namespace {

struct LhbLds {
    cplx* OV0;  // delta-scaled oddV0 X (odd variant, second self-edge)
    cplx* lr;   // f_loop's loop scalars, row chain: XD AX^(i-1) D
    cplx* ov0;  // oddV0 X AX^(k-1) D
    cplx* fe;   // f_loop(n_max) coefficients
    cplx* f0;   // f_loop_odd(n_fixed)[n_fixed]
};

__device__ __forceinline__ LhbLds carve_lhb(unsigned char* base, int MM) {
    LhbLds X;
    cplx* c = reinterpret_cast<cplx*>(base + generic_lds_bytes(MM));
    X.OV0 = c; c += MM;
    X.lr = c; c += WH_NMAX / 2 + 4;
    X.ov0 = c; c += WH_NMAX / 2 + 4;
    X.fe = c; c += WH_NMAX / 2 + 4;
    X.f0 = c;
    return X;
}

// Coefficient series of _hafnian.py:212-241 (odd_form false: f_loop, factors p_i / 2i +
// lv_i / 2) or :246-285 (odd_form true: f_loop_odd, factor oddloop at i = 1, p_(i/2) / i +
// lv_(i/2) / 2 at even i, ov_((i-1)/2) at odd i), orders 0..top.  Returns the final row.
__device__ cplx* coeff_series(cplx* oldc, cplx* newc, int top, bool odd_form, const cplx* pt,
                              const cplx* lvv, const cplx* ovv, cplx oddloop, int lane) {
    for (int kk = lane; kk <= top; kk += WH_WAVE) oldc[kk] = (kk == 0) ? cmk(1.0, 0.0) : cmk(0.0, 0.0);
    __syncthreads();
    for (int i = 1; i <= top; ++i) {
        cplx factor;
        if (!odd_form) {
            factor = cadd(cdivr(pt[i], (double)(2 * i)), cdivr(lvv[i], 2.0));
        } else if (i == 1) {
            factor = oddloop;
        } else if ((i & 1) == 0) {
            factor = cadd(cdivr(pt[i / 2], (double)i), cdivr(lvv[i / 2], 2.0));
        } else {
            factor = ovv[(i - 1) / 2];
        }
        for (int kk = lane; kk <= top; kk += WH_WAVE) {
            cplx val = oldc[kk];
            cplx pf = cmk(1.0, 0.0);
            for (int j = 1; j <= top / i; ++j) {
                pf = cmul(pf, cdivr(factor, (double)j));
                if (i * j <= kk) val = cadd(val, cmul(oldc[kk - i * j], pf));
            }
            newc[kk] = val;
        }
        __syncthreads();
        cplx* tmp = oldc; oldc = newc; newc = tmp;
    }
    return oldc;
}

}  // namespace

// partials[J.part_off + chunk * nD * L + kd * L + q] = (sum_re, err_re, sum_im, err_im) of
// output q of diagonal kd.  Block chunk_base + chunk * dsplit + g takes the chunk's terms for
// diagonal group g (dsplit groups of ceil(nD / dsplit) diagonals).
extern "C" __global__ void __launch_bounds__(WH_WAVE)
wh_sieve_lhb(const wh_lhb_job* __restrict__ jobs, int njobs, const cplx* __restrict__ pool,
             double4* __restrict__ partials, int64_t block_offset, int MM) {
    extern __shared__ __align__(16) unsigned char wh_smem[];
    const int lane = threadIdx.x;
    const int64_t b = blockIdx.x + block_offset;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].chunk_base <= b) lo = mid; else hi = mid - 1;
    }
    const wh_lhb_job& J = jobs[lo];
    const int E = J.E, twoE = 2 * E, nL = J.L;
    const int nf = J.n_fixed, nmax = J.n_max;
    const bool oddv = (J.flags & WH_LHB_ODD) != 0;
    const bool glynn = (J.flags & WH_LHB_GLYNN) != 0;
    const int he = nmax / 2;  // f_loop(n_max) order; powertrace(AX, n_max // 2 + 1)
    // block -> (chunk, diagonal group): the group's diagonals [kd_lo, kd_hi) share the chunk's
    // terms with the chunk's other groups
    const int64_t chunk = (b - J.chunk_base) / J.dsplit;
    const int grp = (int)((b - J.chunk_base) % J.dsplit);
    const int ndg = (J.nD + J.dsplit - 1) / J.dsplit;
    const int kd_lo = grp * ndg;
    const int kd_hi = min(J.nD, kd_lo + ndg);
    const int64_t t0 = chunk * J.chunk_terms;
    int64_t t1 = t0 + J.chunk_terms;
    if (t1 > J.steps) t1 = J.steps;
    const cplx* A = pool + J.a_off;
    const cplx* Dg = pool + J.d_off;
    const int nD = J.nD;
    GenLds L = carve(wh_smem, MM);
    LhbLds X = carve_lhb(wh_smem, MM);

    int my_e = lane < E ? J.e[lane] : 0;
    int64_t my_w = 1;  // product of (e_l + 1) over l > lane
    for (int l = lane + 1; l < E; ++l) my_w *= (int64_t)(J.e[l] + 1);
    csum acc0, acc1;
    acc0.zero();
    acc1.zero();

    for (int64_t t = t0; t < t1; ++t) {
        // loop_hafnian_batch.py:86-97: digits, binomial product over edges 1.., delta
        int k = 0;
        if (lane < E) k = (int)((t / my_w) % (int64_t)(my_e + 1));
        int delta = glynn ? 2 * k - my_e : k;
        int edge_sum = k;
        double bin = (lane >= 1 && lane < E) ? binom_d(my_e, k) : 1.0;
#pragma unroll
        for (int msk = 1; msk < WH_WAVE; msk <<= 1) {
            edge_sum += __shfl_xor(edge_sum, msk, WH_WAVE);
            bin *= __shfl_xor(bin, msk, WH_WAVE);  // exact: products of small integers
        }
        const int k0 = __shfl(k, 0, WH_WAVE);
        const int k1 = __shfl(k, 1, WH_WAVE);
        const bool odd0 = oddv && k0 == 0 && k1 == 0;
        unsigned long long nz = __ballot(lane < E && delta != 0);
        int s = __popcll(nz);
        int m = 2 * s;
        if (lane < E && delta != 0) {
            int pos = __popcll(nz & ((1ull << lane) - 1ull));
            L.S[pos] = lane;
            L.dS[pos] = delta;
        }
        __syncthreads();

        // _hafnian.py:320-356 get_submatrices (oddV = A[0, :]) and :360-380 (oddV0 = A[1, :])
        for (int e = lane; e < m * m; e += WH_WAVE) {
            int r = e / m, c = e - r * m;
            int rr = r < s ? L.S[r] : L.S[r - s] + E;
            int cc, d;
            if (c < s) { cc = L.S[c] + E; d = L.dS[c]; }
            else { cc = L.S[c - s]; d = L.dS[c - s]; }
            L.H[r * MM + c] = cmulr(A[(int64_t)rr * twoE + cc], (double)d);
        }
        if (lane < m) {
            int c = lane < s ? lane : lane - s;
            int src = lane < s ? L.S[c] + E : L.S[c];
            double d = (double)L.dS[c];
            L.OV[lane] = cmulr(A[src], d);
            if (odd0) X.OV0[lane] = cmulr(A[twoE + src], d);
        }
        __syncthreads();

        // a group of one diagonal takes its loop scalars from the powers below: XD B^(k-1) D
        // (both f_loop's row chain and f_loop_odd's XD . z), oddVX B^(k-1) D and oddVX0 B^(k-1) D
        const bool one_diag = kd_hi - kd_lo == 1;
        if (one_diag) {
            const cplx* Dk = Dg + (int64_t)kd_lo * twoE;  // get_Dsubmatrices (_hafnian.py:383-399)
            if (lane < m) {
                int rr = lane < s ? L.S[lane] : L.S[lane - s] + E;
                L.D[lane] = Dk[rr];
                int c = lane < s ? lane : lane - s;
                int src = lane < s ? L.S[c] + E : L.S[c];
                L.XD[lane] = cmulr(Dk[src], (double)L.dS[c]);
            }
            __syncthreads();
        }
        // order k of the loop scalars from B^(k-1) (P = B^(k-1), k <= he)
        auto scalars_from = [&](const cplx* P, int k) {
            cplx a = bilinear_lds(L.XD, P, L.D, m, MM, lane);
            cplx o = bilinear_lds(L.OV, P, L.D, m, MM, lane);
            cplx o0 = odd0 ? bilinear_lds(X.OV0, P, L.D, m, MM, lane) : cmk(0.0, 0.0);
            if (lane == 0) {
                X.lr[k] = a;
                L.lv[k] = a;
                L.ov[k] = o;
                X.ov0[k] = o0;
            }
        };

        // charpoly.py:302-327 powertrace(AX_S, he + 1).  When he + 1 > m the reference extends
        // the traces by the characteristic polynomial (a serial Householder reduction and La
        // Budde per term); this kernel forms every power up to AX_S^he directly (the same
        // traces), as the general kernel does by default
        const int nt = he + 1;
        if (m == 0) {
            for (int kk = lane; kk <= he; kk += WH_WAVE) {
                L.pt[kk] = cmk(0.0, 0.0);
                if (one_diag) X.lr[kk] = L.lv[kk] = L.ov[kk] = X.ov0[kk] = cmk(0.0, 0.0);
            }
        } else {
            cplx tr = trace_lds(L.H, m, MM, lane);
            if (lane == 0) { L.pt[0] = cmk((double)m, 0.0); L.pt[1] = tr; }
            if (one_diag) {  // order 1: x . D; order 2: x B D
                cplx a = wave_csum(lane < m ? cmul(L.XD[lane], L.D[lane]) : cmk(0.0, 0.0));
                cplx o = wave_csum(lane < m ? cmul(L.OV[lane], L.D[lane]) : cmk(0.0, 0.0));
                cplx o0 = cmk(0.0, 0.0);
                if (odd0) o0 = wave_csum(lane < m ? cmul(X.OV0[lane], L.D[lane]) : cmk(0.0, 0.0));
                if (lane == 0) {
                    X.lr[1] = a;
                    L.lv[1] = a;
                    L.ov[1] = o;
                    X.ov0[1] = o0;
                }
                if (he >= 2) scalars_from(L.H, 2);
            }
            const cplx* cur = L.H;
            for (int it = 0; it < nt - 2; ++it) {
                cplx* dst = (cur == L.P) ? L.Q : L.P;
                matmul_lds(cur, L.H, dst, m, MM, lane);
                __syncthreads();
                cplx tk = trace_lds(dst, m, MM, lane);
                if (lane == 0) L.pt[it + 2] = tk;
                if (one_diag && it + 3 <= he) scalars_from(dst, it + 3);  // dst = B^(it + 2)
                cur = dst;
            }
        }
        __syncthreads();

        const int start = 2 * k0 + (oddv ? 1 : 0);
        for (int kd = kd_lo; kd < kd_hi; ++kd) {
            // get_Dsubmatrices (_hafnian.py:383-399) for diagonal kd
            const cplx* Dk = Dg + (int64_t)kd * twoE;
            if (!one_diag) {
            if (lane < m) {
                int rr = lane < s ? L.S[lane] : L.S[lane - s] + E;
                L.D[lane] = Dk[rr];
                int c = lane < s ? lane : lane - s;
                int src = lane < s ? L.S[c] + E : L.S[c];
                L.XD[lane] = cmulr(Dk[src], (double)L.dS[c]);
            }
            __syncthreads();

            // f_loop's row chain: XD_S AX_S^(i-1) D_S enters order i (i = 1..he)
            if (lane < m) L.v[lane] = L.XD[lane];
            __syncthreads();
            for (int i = 1; i <= he; ++i) {
                cplx part = (lane < m) ? cmul(L.v[lane], L.D[lane]) : cmk(0.0, 0.0);
                cplx dot = wave_csum(part);
                if (lane == 0) X.lr[i] = dot;
                cplx nv = cmk(0.0, 0.0);
                if (lane < m)
                    for (int r = 0; r < m; ++r) cmac(nv, L.v[r], L.H[r * MM + lane]);
                __syncthreads();
                if (lane < m) L.v[lane] = nv;
                __syncthreads();
            }
            // f_loop_odd's column chain z = AX_S^(k-1) D_S: XD . z (even orders), oddVX . z and
            // oddVX0 . z (odd orders), k = 1..he
            if (lane < m) L.v[lane] = L.D[lane];
            __syncthreads();
            for (int kk = 1; kk <= he; ++kk) {
                cplx a = (lane < m) ? cmul(L.XD[lane], L.v[lane]) : cmk(0.0, 0.0);
                cplx o = (lane < m) ? cmul(L.OV[lane], L.v[lane]) : cmk(0.0, 0.0);
                cplx da = wave_csum(a), dob = wave_csum(o);
                cplx d0 = cmk(0.0, 0.0);
                if (odd0) {
                    cplx o0 = (lane < m) ? cmul(X.OV0[lane], L.v[lane]) : cmk(0.0, 0.0);
                    d0 = wave_csum(o0);
                }
                if (lane == 0) { L.lv[kk] = da; L.ov[kk] = dob; X.ov0[kk] = d0; }
                cplx nv = cmk(0.0, 0.0);
                if (lane < m)
                    for (int c = 0; c < m; ++c) cmac(nv, L.H[lane * MM + c], L.v[c]);
                __syncthreads();
                if (lane < m) L.v[lane] = nv;
                __syncthreads();
            }
            }  // !one_diag: a single diagonal's scalars came from the powers

            // coefficient series; f_loop(n_max) is kept in X.fe while the others reuse the rows
            cplx* c0 = L.comb;
            cplx* c1 = L.comb + (WH_NMAX + 2);
            const cplx* fe = coeff_series(c0, c1, he, false, L.pt, X.lr, nullptr, cmk(0.0, 0.0), lane);
            for (int kk = lane; kk <= he; kk += WH_WAVE) X.fe[kk] = fe[kk];
            __syncthreads();
            if (odd0) {
                const cplx* f0 = coeff_series(c0, c1, nf, true, L.pt, L.lv, X.ov0, Dk[1], lane);
                if (lane == 0) X.f0[0] = f0[nf];
                __syncthreads();
            }
            const cplx* fo = coeff_series(c0, c1, nmax, true, L.pt, L.lv, L.ov, Dk[0], lane);

            // loop_hafnian_batch.py:108-121 (even) and :183-202 (odd): output (kd, N_det = q)
            // lives on lane (kd L + q) mod 64, slot (kd L + q) / 64
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                int o = lane + r * WH_WAVE - kd * nL;
                if (o < 0 || o >= nL) continue;
                int q = o;
                cplx v = cmk(0.0, 0.0);
                bool any = false;
                if (q >= start) {
                    int N = nf + q;
                    double pm = (((N / 2 - edge_sum) & 1) != 0) ? -1.0 : 1.0;
                    int half = oddv ? (q - 1) / 2 : q / 2;
                    double nb = binom_d(half, k0) * bin;
                    bool use_even = oddv ? (N % 2 == 0) : (q % 2 == 0);
                    cplx f = use_even ? X.fe[N / 2] : fo[N];
                    v = cmulr(f, nb * pm);
                    any = true;
                }
                if (q == 0 && odd0) {
                    double pm = (((nf / 2 - edge_sum) & 1) != 0) ? -1.0 : 1.0;
                    v = cmulr(X.f0[0], bin * pm);  // start >= 1 here: the only contribution
                    any = true;
                }
                if (any) {
                    if (r == 0) acc0.add(v); else acc1.add(v);
                }
            }
            __syncthreads();
        }
    }
    const int nout = nD * nL;
    const int o_lo = kd_lo * nL, o_hi = kd_hi * nL;  // this group's outputs
    double4* out = partials + J.part_off + chunk * nout;
    if (lane >= o_lo && lane < o_hi) out[lane] = make_double4(acc0.sr, acc0.cr, acc0.si, acc0.ci);
    if (lane + WH_WAVE >= o_lo && lane + WH_WAVE < o_hi)
        out[lane + WH_WAVE] = make_double4(acc1.sr, acc1.cr, acc1.si, acc1.ci);
}

// Folds each output's chunk partials in chunk order and applies 0.5^((n_fixed + q) // 2) for
// Glynn (loop_hafnian_batch.py:119-122, :204-207).  One workgroup per job, one thread per output.
extern "C" __global__ void __launch_bounds__(128)
wh_fold_lhb(const wh_lhb_job* __restrict__ jobs, const double4* __restrict__ partials,
            double2* __restrict__ out) {
    const wh_lhb_job& J = jobs[blockIdx.x];
    const int nout = J.nD * J.L;
    int o = threadIdx.x;
    if (o >= nout) return;
    csum acc;
    acc.zero();
    for (int c = 0; c < J.nchunks; ++c) {
        double4 p = partials[J.part_off + (int64_t)c * nout + o];
        acc.add_pair(p.x, p.y, p.z, p.w);
    }
    double re = acc.sr + acc.cr, im = acc.si + acc.ci;
    if (J.flags & WH_LHB_GLYNN) {
        double sc = ldexp(1.0, -((J.n_fixed + o % J.L) / 2));
        re *= sc;
        im *= sc;
    }
    out[J.out_off + o] = make_double2(re, im);
}

#include "wh_launch.h"

size_t wh_lhb_lds_bytes(int MM) { return generic_lds_bytes(MM) + (MM + 3 * (WH_NMAX / 2 + 4) + 1) * sizeof(cplx); }

hipError_t wh_launch_lhb(const wh_lhb_job* jobs, int njobs, const void* pool, void* partials,
                         int64_t nblocks, int MM, void* out, hipStream_t stream) {
    size_t lds = wh_lhb_lds_bytes(MM);
    static wh_lds_limit lim;
    hipError_t err = wh_raise_lds_limit((const void*)wh_sieve_lhb, lds, lim);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(wh_sieve_lhb, dim3((unsigned)nblocks), dim3(WH_WAVE), lds, stream, jobs, njobs,
                       (const cplx*)pool, (double4*)partials, (int64_t)0, MM);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(wh_fold_lhb, dim3((unsigned)njobs), dim3(128), 0, stream, jobs,
                       (const double4*)partials, (double2*)out);
    return hipGetLastError();
}
