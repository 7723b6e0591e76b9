// hermite.hip -- the multidimensional Hermite recursion behind hafnian_batched: the loop
// hafnians of reduction(A, k) for every k below the cutoffs, as one tensor G
// (thewalrus/_hermite_multidimensional.py:173-216, recursion :251-303).
//
// G[idx] = y_i G[idx - e_i] - sum_(l >= i, ki_l > 0) w(ki_l) R_il G[idx - e_i - e_l], where i is
// the first nonzero position of idx and ki = idx - e_i (w(k) = k, or sqrt(k) with a final
// division by sqrt(idx_i) for the renormalised form).  An entry depends only on entries with
// the same zero prefix and a smaller value at position i, so the tensor fills in steps
// (i, v): i from n-1 down to 0, v = 1..cut_i - 1.  Step (i, v) is the contiguous slice
// [v stride_i, (v + 1) stride_i) of the row-major tensor -- every entry of it in parallel, with
// coalesced reads of the slices v - 1 and v - 2.  Steps whose slice is small run inside one
// workgroup (barrier between steps); the large ones are one launch each.  HBM-bound: per entry
// one write and 1 + (n - i) reads of 16 bytes.
#include <cstring>

#include "wh_device.h"

#define WH_HERM_MAXN 32

struct wh_herm_args {
    int n;
    int renorm;
    int cut[WH_HERM_MAXN];
    int64_t stride[WH_HERM_MAXN];
};

// One entry of slice (i, v): s indexes the suffix (positions i+1..n-1).
__device__ __forceinline__ void herm_entry(const wh_herm_args& a, const cplx* __restrict__ R,
                                           const cplx* __restrict__ y, const double* __restrict__ sq,
                                           cplx* __restrict__ G, int i, int v, int64_t s) {
    const int64_t lin = (int64_t)v * a.stride[i] + s;
    const int64_t ki = lin - a.stride[i];
    cplx u = cmul(y[i], G[ki]);
    const int n = a.n;
    // l = i: ki_i = v - 1
    if (v > 1) {
        double w = a.renorm ? sq[v - 1] : (double)(v - 1);
        u = csub(u, cmul(cmulr(R[i * n + i], w), G[ki - a.stride[i]]));
    }
    for (int l = i + 1; l < n; ++l) {
        int kl = (int)((s / a.stride[l]) % a.cut[l]);
        if (kl > 0) {
            double w = a.renorm ? sq[kl] : (double)kl;
            u = csub(u, cmul(cmulr(R[i * n + l], w), G[ki - a.stride[l]]));
        }
    }
    G[lin] = a.renorm ? cdivr(u, sq[v]) : u;
}

// Steps (i, v) for i = n-1 down to i_stop (exclusive), every v, in one workgroup.
extern "C" __global__ void __launch_bounds__(1024)
wh_hermite_small(wh_herm_args a, const cplx* __restrict__ R, const cplx* __restrict__ y,
                 const double* __restrict__ sq, cplx* __restrict__ G, int i_stop) {
    for (int i = a.n - 1; i > i_stop; --i) {
        for (int v = 1; v < a.cut[i]; ++v) {
            for (int64_t s = threadIdx.x; s < a.stride[i]; s += blockDim.x) herm_entry(a, R, y, sq, G, i, v, s);
            __syncthreads();
        }
    }
}

// One step (i, v) over its whole slice (grid-stride).
extern "C" __global__ void __launch_bounds__(256)
wh_hermite_step(wh_herm_args a, const cplx* __restrict__ R, const cplx* __restrict__ y,
                const double* __restrict__ sq, cplx* __restrict__ G, int i, int v) {
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.stride[i]; s += step)
        herm_entry(a, R, y, sq, G, i, v, s);
}

#include "wh_launch.h"

// Enqueues the whole recursion; G[0] must hold C and every other entry is written here.
// Positions whose slice (stride) is at most small_max run in the single-workgroup kernel.
hipError_t wh_launch_hermite(int n, const int* cut, int renorm, const void* R, const void* y, const double* sq,
                             void* G, hipStream_t stream) {
    wh_herm_args a;
    memset(&a, 0, sizeof a);
    a.n = n;
    a.renorm = renorm;
    int64_t total = 1;
    for (int l = n - 1; l >= 0; --l) {
        a.cut[l] = cut[l];
        a.stride[l] = total;
        total *= cut[l];
    }
    const int64_t small_max = 8192;
    int i_stop = n - 1;  // positions > i_stop go to the small kernel
    while (i_stop >= 0 && a.stride[i_stop] <= small_max) --i_stop;
    if (i_stop < n - 1) {
        hipLaunchKernelGGL(wh_hermite_small, dim3(1), dim3(1024), 0, stream, a, (const cplx*)R, (const cplx*)y, sq,
                           (cplx*)G, i_stop);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int i = i_stop; i >= 0; --i) {
        int64_t blocks = (a.stride[i] + 255) / 256;
        if (blocks > (1 << 20)) blocks = 1 << 20;  // grid-stride beyond 2^28 work-items
        for (int v = 1; v < a.cut[i]; ++v) {
            hipLaunchKernelGGL(wh_hermite_step, dim3((unsigned)blocks), dim3(256), 0, stream, a, (const cplx*)R,
                               (const cplx*)y, sq, (cplx*)G, i, v);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}
