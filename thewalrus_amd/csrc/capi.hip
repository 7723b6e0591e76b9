// capi.hip -- the C ABI (include/walrus_hip.h): device contexts, launch plans, and the
// per-call entry points that stand in for thewalrus's numba kernels _calc_hafnian
// (thewalrus/_hafnian.py:417) and _calc_loop_hafnian (thewalrus/_hafnian.py:513).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/walrus_hip.h"
#include "wh_device.h"
#include "wh_launch.h"

static_assert(WH_MAX_EDGES == WH_EMAX, "edge limit mismatch");
static_assert(WH_MAX_PHOTONS == WH_NMAX, "photon limit mismatch");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(WH_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));            \
    } while (0)

constexpr int64_t kMaxBlocks = 0x7fffffff;  // workgroups of one launch (grid x dimension)
constexpr int kGenericMaxSide = 48;  // LDS holds three 48 x 48 complex matrices per wave
constexpr int kMtlMaxModes = 24;     // montrealer on the all-ones kernel: side 2n <= 8 * 6

// Per-device state used by the synchronous per-call entry points.
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    std::mutex mu;
    // per-call staging: [job descriptor | A | D | oddV] in one pinned block, one H2D copy
    unsigned char* h_stage = nullptr;
    unsigned char* d_stage = nullptr;
    size_t stage_cap = 0;
    double4* d_partials = nullptr;
    size_t partials_cap = 0;
    double* h_out = nullptr;   // pinned 2 doubles
    double2* d_out = nullptr;
};

constexpr int kMaxDevices = 64;
Context g_ctx[kMaxDevices];
std::mutex g_ctx_mu;

int current_device(int* dev) {
    HIP_TRY(hipGetDevice(dev));
    if (*dev < 0 || *dev >= kMaxDevices) return fail(WH_EUNSUPPORTED, "device index %d", *dev);
    return WH_OK;
}

int ensure_context(int dev, Context** out) {
    Context& c = g_ctx[dev];
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (c.device < 0) {
        HIP_TRY(hipSetDevice(dev));
        HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIP_TRY(hipHostMalloc((void**)&c.h_out, 2 * sizeof(double), hipHostMallocDefault));
        HIP_TRY(hipMalloc((void**)&c.d_out, sizeof(double2)));
        c.device = dev;
    }
    *out = &c;
    return WH_OK;
}

int ensure_stage(Context& c, size_t bytes) {
    if (bytes <= c.stage_cap) return WH_OK;
    size_t cap = std::max(bytes, (size_t)1 << 16);
    if (c.h_stage) { (void)hipHostFree(c.h_stage); c.h_stage = nullptr; }
    if (c.d_stage) { (void)hipFree(c.d_stage); c.d_stage = nullptr; }
    c.stage_cap = 0;
    if (hipHostMalloc((void**)&c.h_stage, cap, hipHostMallocDefault) != hipSuccess)
        return fail(WH_ENOMEM, "pinned staging allocation of %zu bytes failed", cap);
    if (hipMalloc((void**)&c.d_stage, cap) != hipSuccess)
        return fail(WH_ENOMEM, "device staging allocation of %zu bytes failed", cap);
    c.stage_cap = cap;
    return WH_OK;
}

int ensure_partials(Context& c, size_t n) {
    if (n <= c.partials_cap) return WH_OK;
    size_t cap = std::max(n, (size_t)4096);
    if (c.d_partials) { (void)hipFree(c.d_partials); c.d_partials = nullptr; }
    c.partials_cap = 0;
    if (hipMalloc((void**)&c.d_partials, cap * sizeof(double4)) != hipSuccess)
        return fail(WH_ENOMEM, "partials allocation failed");
    c.partials_cap = cap;
    return WH_OK;
}

// ---- Allocation-free batch launch.  The job table and the partial sums of
// wh_batch_launch_noalloc live in `scratch`, a static device array: one copy per GPU the library
// runs on, held until the process ends.  In host code `scratch` names a host-side stand-in, not
// device memory, so every host use goes through the device address hipGetSymbolAddress returns,
// resolved once per device (with that device current) and kept in g_scratch.
constexpr size_t kScratchJobBytes = (size_t)WH_NOALLOC_JOB_BYTES;
constexpr int64_t kScratchSlots = WH_NOALLOC_PARTIAL_SLOTS;
constexpr size_t kScratchBytes = kScratchJobBytes + (size_t)kScratchSlots * sizeof(double4);
static_assert(kScratchJobBytes % 256 == 0, "partials must stay aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev` (resolved on first use).
int scratch_address(int dev, void** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(WH_EDEVICE, "no device address for the scratch array");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return WH_OK;
}

// Validates one request and fills the device descriptor (without chunking).
// _hafnian.py:430-438 and :531-538 define N and the number of terms.
int describe(const wh_request& r, wh_job_dev& j) {
    if (r.E < 1 || r.E > WH_MAX_EDGES) return fail(WH_EINVAL, "E=%d outside 1..%d", r.E, WH_MAX_EDGES);
    memset(&j, 0, sizeof j);
    if (r.options & WH_OPT_MONTREALER) {
        // _montrealer.py:48-57: all 2^n subsets of the n = E modes, p_n of a 2n-side matrix
        if (r.E > kMtlMaxModes)
            return fail(WH_EUNSUPPORTED, "montrealer of %d modes exceeds %d", r.E, kMtlMaxModes);
        const bool loop = r.d_off >= 0;
        if (loop != (r.v_off >= 0)) return fail(WH_EINVAL, "loop montrealer needs both X zeta and its conjugate");
        for (int i = 0; i < r.E; ++i) j.e[i] = 1;
        j.a_off = r.a_off;
        j.d_off = r.d_off;
        j.v_off = r.v_off;
        j.steps = (int64_t)1 << r.E;
        j.E = r.E;
        j.N = 2 * r.E;
        j.flags = WH_F_MTL | (loop ? WH_F_LOOP : 0) | (r.real ? WH_F_REAL : 0);
        return WH_OK;
    }
    int64_t esum = 0;
    double steps_d = 1.0;
    const bool odd = r.v_off >= 0;
    const bool loop = r.d_off >= 0;
    if (odd && !loop) return fail(WH_EINVAL, "an unpaired mode needs loop weights");
    for (int i = 0; i < r.E; ++i) {
        int e = r.edge_reps[i];
        if (e < 0) return fail(WH_EINVAL, "negative edge multiplicity");
        esum += e;
        j.e[i] = e;
    }
    const bool glynn = r.glynn != 0;
    int64_t steps = 1;
    if (glynn && !odd) {
        steps = (r.edge_reps[0] + 2) / 2;
        steps_d = (double)steps;
        for (int i = 1; i < r.E; ++i) { steps *= r.edge_reps[i] + 1; steps_d *= r.edge_reps[i] + 1.0; }
    } else {
        for (int i = 0; i < r.E; ++i) { steps *= r.edge_reps[i] + 1; steps_d *= r.edge_reps[i] + 1.0; }
    }
    if (steps_d > 9.0e18) return fail(WH_EUNSUPPORTED, "term count exceeds int64");
    int64_t N = 2 * esum + (odd ? 1 : 0);
    if (N > WH_MAX_PHOTONS) return fail(WH_EUNSUPPORTED, "N=%lld exceeds %d", (long long)N, WH_MAX_PHOTONS);
    if (N < 2) return fail(WH_EINVAL, "N=%lld: the frontends resolve N < 2 on the host", (long long)N);
    j.a_off = r.a_off;
    j.d_off = r.d_off;
    j.v_off = r.v_off;
    j.steps = steps;
    j.oddloop_re = r.oddloop[0];
    j.oddloop_im = r.oddloop[1];
    j.E = r.E;
    j.N = (int)N;
    j.flags = (loop ? WH_F_LOOP : 0) | (odd ? WH_F_ODD : 0) | (glynn ? WH_F_GLYNN : 0) |
              (r.real ? WH_F_REAL : 0) | ((r.options & WH_OPT_CHARPOLY) ? WH_F_CHARPOLY : 0);
    return WH_OK;
}

// Kernel families.  The all-ones kernels take Glynn terms with even N and every multiplicity
// 1 (m = n for every term), n_pad = 8T: PM1 the register-tiled vector kernel (T x T tiles per
// lane), PM1S the symmetric-power matrix-core kernel, PM1M the full-power matrix-core kernel.
// REP: repeated rows (any multiplicities) with even N, 2E <= 16 and h <= 16, on the vector
// kernel's register-tiled products with column factors delta (the all-ones kernels' +-1 in
// general).  GENERIC: everything else.
enum Kind { KIND_PM1 = 0, KIND_GENERIC = 1, KIND_PM1M = 2, KIND_PM1S = 3, KIND_REP = 4 };
static bool all_ones_kind(int k) { return k == KIND_PM1 || k == KIND_PM1M || k == KIND_PM1S; }
constexpr int kPm1MaxT = 6;   // tile sides of the vector kernel and of the tuning table
constexpr int kPm1sMaxT = 8;  // the symmetric matrix-core kernel also takes n in [50, 64]
constexpr int kRepMaxT = 3;   // REP tile sides: 1 (2E <= 8), 3 (2E <= 12, four terms per wave), 2 (2E <= 16)
constexpr int kRepMaxH = 16;  // REP: h = N / 2 <= 16

struct GroupKey {
    int kind, T, real, loop, half, mtl;
    int H = 0;   // REP: coefficient order cap (8 or 16)
    int nw = 0;  // waves per workgroup asked for by WH_OPT_WAVES (0: the LDS-based choice)
    bool operator<(const GroupKey& o) const {
        if (kind != o.kind) return kind < o.kind;
        if (T != o.T) return T < o.T;
        if (real != o.real) return real < o.real;
        if (loop != o.loop) return loop < o.loop;
        if (half != o.half) return half < o.half;
        if (mtl != o.mtl) return mtl < o.mtl;
        if (H != o.H) return H < o.H;
        return nw < o.nw;
    }
    bool operator==(const GroupKey& o) const {
        return kind == o.kind && T == o.T && real == o.real && loop == o.loop && half == o.half && mtl == o.mtl &&
               H == o.H && nw == o.nw;
    }
};

// Inputs with the half-products formulation run the products on the matrix cores unless
// WH_OPT_VALU asks for the register-tiled vector kernel: hafnians with the symmetric powers
// C_a = B^a X D_d (KIND_PM1S, upper block groups only; real ones up to T = 5 with one MFMA per
// block group),
// the complex montrealer (whose 0/1 factors break that symmetry) and complex WH_OPT_NONSYM
// requests with the full powers (KIND_PM1M); the real montrealer and real WH_OPT_NONSYM requests
// stay on the vector kernel.
static void pick_matrix_cores(GroupKey& g, int options) {
    if (g.kind != KIND_PM1 || !g.half || (options & WH_OPT_VALU)) return;
    if (g.real) {  // real hafnians up to n = 40: the symmetric kernel with one MFMA per block group
        // (at T = 6 its real instance keeps state in scratch: the vector kernel stays)
        if (!g.mtl && !(options & WH_OPT_NONSYM) && g.T <= 5) g.kind = KIND_PM1S;
        return;
    }
    const bool full = g.mtl || (options & WH_OPT_NONSYM);
    // full powers at T = 6 (n in [42, 48]) keep 72 complex accumulators per lane and spill: the
    // vector kernel keeps them; the symmetric kernel's 21 upper groups fit
    if (full) {
        if (g.T <= 5) g.kind = KIND_PM1M;
    } else {
        g.kind = KIND_PM1S;
    }
}

GroupKey classify_auto(const wh_job_dev& j, int options);

// The LDS bytes of a workgroup of nw waves for an all-ones or repeated-row group.
size_t group_lds_bytes(const GroupKey& g, int nw) {
    if (g.kind == KIND_PM1S) return wh_pm1s_lds_bytes(g.T, nw, g.real != 0, g.loop != 0);
    if (g.kind == KIND_PM1M) return wh_pm1m_lds_bytes(g.T, nw);
    if (g.kind == KIND_PM1) return wh_pm1_lds_bytes(g.T, g.real != 0, nw);
    if (g.kind == KIND_REP) return wh_rep_lds_bytes(g.T, g.real != 0, nw, g.H);
    return 0;
}

// The kernel class of a request, with the WH_OPT_TILE / WH_OPT_WAVES overrides applied; fails
// (kind < 0) when an override does not fit the class.
int tuned_tile(const GroupKey& g);

// the largest 2E a repeated-row tile side takes
int rep_cap(int T) { return T == 1 ? 8 : T == 3 ? 12 : T == 2 ? 16 : 0; }

bool tile_fits(const GroupKey& g, int tile) {
    return all_ones_kind(g.kind) ? tile >= g.T && tile <= (g.kind == KIND_PM1S ? kPm1sMaxT : kPm1MaxT) &&
                                       !(g.kind == KIND_PM1M && tile > 5)
           : g.kind == KIND_REP ? tile >= 1 && tile <= kRepMaxT && rep_cap(tile) >= rep_cap(g.T)
                                : false;
}

GroupKey classify(const wh_job_dev& j, int options) {
    GroupKey g = classify_auto(j, options);
    const int tile = (options >> WH_OPT_TILE_SHIFT) & 15;
    const int waves = (options >> WH_OPT_WAVES_SHIFT) & 7;
    if (tile) {
        if (!tile_fits(g, tile)) {
            g.kind = -1;
            return g;
        }
        g.T = tile;
    } else {  // the tuned tile side of the class, where one is set and the class's kernel takes it
        const int tt = tuned_tile(g);  // (repeated-row sides order by the 2E they hold: 1, 3, 2)
        if (tt != 0 && tt != g.T && tile_fits(g, tt)) g.T = tt;
    }
    if (waves) {
        if (!(all_ones_kind(g.kind) || g.kind == KIND_REP) || waves > 4 || group_lds_bytes(g, waves) > 160 * 1024) {
            g.kind = -1;
            return g;
        }
        g.nw = waves;
    }
    return g;
}

GroupKey classify_auto(const wh_job_dev& j, int options) {
    GroupKey g{KIND_GENERIC, 0, 0, 0, 0, 0};
    if (j.flags & WH_F_MTL) {  // the montrealer runs on the all-ones kernels only
        g.kind = KIND_PM1;
        g.mtl = 1;
        g.T = (2 * j.E + 7) / 8;
        g.real = (j.flags & WH_F_REAL) ? 1 : 0;
        g.loop = (j.flags & WH_F_LOOP) ? 1 : 0;
        g.half = (options & WH_OPT_FULL) ? 0 : 1;
        pick_matrix_cores(g, options);
        return g;
    }
    if (options & WH_OPT_GENERIC) return g;
    bool all_one = true;
    for (int i = 0; i < j.E; ++i) all_one = all_one && j.e[i] == 1;
    bool glynn = (j.flags & WH_F_GLYNN) != 0, odd = (j.flags & WH_F_ODD) != 0;
    int T = (2 * j.E + 7) / 8;
    if (all_one && glynn && !odd && T <= kPm1MaxT && j.E >= 2) {
        g.kind = KIND_PM1;
        g.T = T;
        g.real = (j.flags & WH_F_REAL) ? 1 : 0;
        g.loop = (j.flags & WH_F_LOOP) ? 1 : 0;
        g.half = (options & WH_OPT_FULL) ? 0 : 1;
        pick_matrix_cores(g, options);
        return g;
    }
    if (all_one && glynn && !odd && T <= kPm1sMaxT && j.E >= 2) {
        // n in [50, 64]: the symmetric matrix-core kernel only (one wave per CU, the lanes'
        // accumulators partly in scratch), real inputs as complex ones with zero imaginary parts
        g.kind = KIND_PM1S;
        g.T = T;
        g.loop = (j.flags & WH_F_LOOP) ? 1 : 0;
        g.half = 1;
        return g;
    }
    const int h = j.N / 2;
    if (!odd && 2 * j.E <= 16 && h <= kRepMaxH && j.steps < ((int64_t)1 << 31) &&
        !(j.flags & WH_F_CHARPOLY) && !(options & WH_OPT_FULL)) {
        g.kind = KIND_REP;
        g.T = 2 * j.E <= 8 ? 1 : 2 * j.E <= 12 ? 3 : 2;
        g.H = h <= 8 ? 8 : 16;
        g.real = (j.flags & WH_F_REAL) ? 1 : 0;
        g.loop = (j.flags & WH_F_LOOP) ? 1 : 0;
        g.half = 1;
    }
    return g;
}

// Waves per workgroup for a REP group: the count that maximises resident waves per CU under the
// LDS budget, at most 4 (the kernel's launch bound).
int rep_waves(int T, bool real, int H) {
    const size_t lds_cap = 160 * 1024;
    int best = 1, best_w = 0;
    for (int nw = 1; nw <= 4; ++nw) {
        int per_cu = (int)(lds_cap / wh_rep_lds_bytes(T, real, nw, H)) * nw;
        if (per_cu > 32) per_cu = 32;
        if (per_cu > best_w) { best_w = per_cu; best = nw; }
    }
    return best;
}

// Waves per workgroup for a PM1 group: the count that maximises resident waves per CU
// (each wave owns one n_pad x n_pad power; the workgroup shares M).  For the symmetric kernel
// the runtime's occupancy calculator gives the workgroups per CU of each count (registers and
// LDS: at T = 4 the 300-register instances hold one wave per SIMD, so 3-wave workgroups would
// leave a SIMD idle; n = 32: 14.4 ms against 10.8 at 4 waves); without a device it falls back to
// the LDS budget of the loop carve.  Ties go to the larger workgroup: the same waves per CU with
// fewer staging copies of the matrix (the real T = 3 instance: 33.1 -> 30.5 ms on 4096 n = 24).
// 1 + workgroups per CU, 0 = not asked yet; asked once per process on the current device (the
// answer depends on the kernel's registers and the CU's LDS, alike on the GPUs of one host)
std::atomic<int> g_pm1s_resident[kPm1sMaxT + 1][2][2][5];

int pm1s_resident_cached(int T, int nw, bool real, bool loop) {
    std::atomic<int>& c = g_pm1s_resident[T][real ? 1 : 0][loop ? 1 : 0][nw];
    int v = c.load(std::memory_order_relaxed);
    if (v == 0) {
        v = 1 + std::max(0, wh_pm1s_resident(T, nw, real, loop));
        c.store(v, std::memory_order_relaxed);
    }
    return v - 1;
}

int pm1_waves(int T, bool real, bool loop, int kind) {
    const size_t lds_cap = 160 * 1024;
    int res[5] = {0, 0, 0, 0, 0};
    bool known = kind == KIND_PM1S && T >= 1 && T <= kPm1sMaxT;
    for (int nw = 1; known && nw <= 4; ++nw) {
        res[nw] = pm1s_resident_cached(T, nw, real, loop);
        if (res[nw] <= 0) known = false;
    }
    int best = 1, best_w = 0;
    for (int nw = 1; nw <= 4; ++nw) {  // kernel launch bound: 256 threads
        size_t b = kind == KIND_PM1S   ? wh_pm1s_lds_bytes(T, nw, real, true)
                   : kind == KIND_PM1M ? wh_pm1m_lds_bytes(T, nw)
                                       : wh_pm1_lds_bytes(T, real, nw);
        if (!known && b > lds_cap) break;
        int per_cu = known ? res[nw] * nw : (int)(lds_cap / b) * nw;
        if (per_cu > 16) per_cu = 16;
        if (per_cu >= best_w) { best_w = per_cu; best = nw; }
    }
    return best;
}

// Workgroups per request the chunk size aims at, per kernel class (wh_tune_set; 0 = the
// defaults below): [0][T][loop] all-ones kernels by tile side class, [1][0][loop] general kernel,
// [2][T][loop] repeated-row kernel (tile sides T = 1..3).
std::atomic<int64_t> g_tune_blocks[3][kPm1MaxT + 1][2];
// Waves per workgroup per kernel class (wh_tune_set_waves; 0 = the LDS-based choice), same indexing.
std::atomic<int> g_tune_waves[3][kPm1MaxT + 1][2];
// Tile side per kernel class (wh_tune_set_tile; 0 = the class's own side T), same indexing.
std::atomic<int> g_tune_tile[3][kPm1MaxT + 1][2];

int64_t tuned_blocks(const GroupKey& g) {
    const int kind = all_ones_kind(g.kind) ? 0 : g.kind == KIND_REP ? 2 : 1;
    const int T = kind == 1 ? 0 : std::min(std::max(g.T, 0), kPm1MaxT);
    return g_tune_blocks[kind][T][g.loop ? 1 : 0].load(std::memory_order_relaxed);
}

int tuned_waves(const GroupKey& g) {
    if (!(all_ones_kind(g.kind) || g.kind == KIND_REP)) return 0;
    const int kind = all_ones_kind(g.kind) ? 0 : 2;
    const int T = std::min(std::max(g.T, 0), kPm1MaxT);
    return g_tune_waves[kind][T][g.loop ? 1 : 0].load(std::memory_order_relaxed);
}

int tuned_tile(const GroupKey& g) {
    if (!(all_ones_kind(g.kind) || g.kind == KIND_REP) || g.T < 1 || g.T > kPm1MaxT) return 0;
    return g_tune_tile[all_ones_kind(g.kind) ? 0 : 2][g.T][g.loop ? 1 : 0].load(std::memory_order_relaxed);
}

int64_t chunk_for(const wh_job_dev& j, const GroupKey& g, int nw) {
    const int64_t tb = tuned_blocks(g);
    if (all_ones_kind(g.kind)) {
        int64_t c = j.steps / (tb > 0 ? tb : 1024);
        if (c < 4 * nw) c = 4 * nw;
        if (c > 4096) c = 4096;
        c = (c + nw - 1) / nw * nw;
        return c;
    }
    if (g.kind == KIND_REP && g.T == 3) {  // each wave takes four terms at a time
        int64_t c = j.steps / (tb > 0 ? tb : 64);
        if (c < 16 * nw) c = 16 * nw;
        if (c > 4096) c = 4096;
        c = (c + 4 * nw - 1) / (4 * nw) * (4 * nw);
        return c;
    }
    if (g.kind == KIND_REP) {  // each wave takes every nw-th term of its workgroup's chunk
        int64_t c = j.steps / (tb > 0 ? tb : 64);
        if (c < 8 * nw) c = 8 * nw;
        if (c > 4096) c = 4096;
        c = (c + nw - 1) / nw * nw;
        return c;
    }
    // generic: one wave per workgroup walks the chunk
    int64_t c = j.steps / (tb > 0 ? tb : 256);
    if (c < 1) c = 1;
    if (c > 256) c = 256;
    return c;
}

double term_flops(const wh_job_dev& j, const GroupKey& g) {
    // algorithmic FP64 flops of the products per term, as the kernels run them:
    // PM1: h - 1 products of n x n (8 flops per complex multiply-add, 2 per real one), or
    // ceil(h/2) - 1 with the half-products formulation;
    // GENERIC: h - 1 products of m x m with m = 2 * nnz(delta) <= 2E (upper bound), or
    // min(h + 1, m) - 2 with the characteristic-polynomial extension;
    // PM1S: the symmetric powers need the m (m + 1) / 2 entries of the upper triangle only
    int h = j.N / 2;
    int m = 2 * j.E;
    int np = std::min(h + 1, m) - 2;
    if (g.kind == KIND_GENERIC && !(j.flags & WH_F_CHARPOLY)) np = h - 1;  // every power formed
    if ((all_ones_kind(g.kind) || g.kind == KIND_REP) && g.half) np = (h + 1) / 2 - 1;
    if (np < 0) np = 0;
    double c = ((g.kind == KIND_PM1 || g.kind == KIND_REP || g.kind == KIND_PM1S) && g.real) ? 2.0 : 8.0;
    if (g.kind == KIND_PM1S) return c * (m * (m + 1.0) / 2.0) * m * np;
    return c * m * (double)m * m * np;
}

struct Group {
    GroupKey key;
    int64_t job_begin = 0, job_end = 0;
    int64_t block_begin = 0, block_end = 0;
    int nw = 1;
    int MM = 0;  // generic: LDS matrix side
    int NQ = 0;  // generic: largest photon number + 2 (LDS coefficient rows)
};

// One past the last pool element a job reads: the 2E x 2E matrix and the 2E-vectors D and V.
int64_t job_pool_end(const wh_job_dev& j) {
    const int64_t side = 2 * (int64_t)j.E;
    int64_t end = j.a_off + side * side;
    if (j.d_off >= 0) end = std::max(end, j.d_off + side);
    if (j.v_off >= 0) end = std::max(end, j.v_off + side);
    return end;
}

struct Layout {
    std::vector<wh_job_dev> jobs;  // sorted by group
    std::vector<Group> groups;
    int64_t blocks = 0;
    int64_t terms = 0;
    double flops = 0.0;
};

// chunk_hint > 0 fixes the terms per chunk of every request; else chunk_per_req (when given) holds
// each request's own chunk size, and otherwise the tuned default applies.  Chunk counts and the
// batch's total stay within kMaxBlocks (a grid's x dimension), or the layout fails.
int build_layout(const wh_request* reqs, int64_t n, int64_t chunk_hint, Layout& L,
                 const int64_t* chunk_per_req = nullptr) {
    std::vector<wh_job_dev> tmp(n);
    std::vector<GroupKey> keys(n);
    for (int64_t i = 0; i < n; ++i) {
        int rc = describe(reqs[i], tmp[i]);
        if (rc) return rc;
        tmp[i].out_index = i;
        keys[i] = classify(tmp[i], reqs[i].options);
        if (keys[i].kind < 0)
            return fail(WH_EINVAL, "request %lld: tile side or waves option does not fit its kernel", (long long)i);
        if (keys[i].kind == KIND_GENERIC && 2 * reqs[i].E > kGenericMaxSide)
            return fail(WH_EUNSUPPORTED, "matrix side %d exceeds the generic kernel limit %d",
                        2 * reqs[i].E, kGenericMaxSide);
    }
    std::vector<int64_t> order(n);
    for (int64_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    L.jobs.resize(n);
    int64_t base = 0;
    for (int64_t k = 0; k < n; ++k) {
        int64_t i = order[k];
        const GroupKey& g = keys[i];
        if (L.groups.empty() || !(L.groups.back().key == g)) {
            Group gr;
            gr.key = g;
            gr.job_begin = k;
            gr.block_begin = base;
            const int tw = tuned_waves(g);  // the tuning table's choice, where it fits the LDS
            gr.nw = g.nw                                                            ? g.nw
                    : (tw >= 1 && tw <= 4 && group_lds_bytes(g, tw) <= 160 * 1024) ? tw
                    : all_ones_kind(g.kind)                                         ? pm1_waves(g.T, g.real != 0, g.loop != 0, g.kind)
                    : g.kind == KIND_REP                                            ? rep_waves(g.T, g.real != 0, g.H)
                                                                                    : 1;
            L.groups.push_back(gr);
        }
        Group& gr = L.groups.back();
        wh_job_dev j = tmp[i];
        int64_t c = chunk_hint > 0 ? chunk_hint : chunk_per_req ? chunk_per_req[i] : chunk_for(j, g, gr.nw);
        if (c < 1) return fail(WH_EINVAL, "chunk size %lld", (long long)c);
        const int64_t nch = j.steps / c + (j.steps % c != 0);
        if (nch > kMaxBlocks || base + nch > kMaxBlocks)
            return fail(WH_EUNSUPPORTED, "%lld term chunks exceed the launch limit %lld (terms per chunk %lld)",
                        (long long)(base + nch), (long long)kMaxBlocks, (long long)c);
        j.chunk_terms = c;
        j.nchunks = (int32_t)nch;
        j.chunk_base = base;
        base += j.nchunks;
        gr.job_end = k + 1;
        gr.block_end = base;
        if (g.kind == KIND_GENERIC) {
            gr.MM = std::max(gr.MM, 2 * j.E);
            gr.NQ = std::max(gr.NQ, j.N + 2);
        }
        L.terms += j.steps;
        L.flops += (double)j.steps * term_flops(j, g);
        L.jobs[k] = j;
    }
    L.blocks = base;
    return WH_OK;
}

// Batches may run their symmetric-kernel groups of tile sides 4..6 (n in [26, 48]: classes whose
// own kernels also run one wave per SIMD) as one persistent launch when `counters` (two device
// words of the caller's arena) is given (wh_tune_persistent: 0 = never, the default: per-class
// launches measured 1.3% faster on the serving stream; 1 = when the batch has two such groups or
// more; 2 = whenever it has one); the other groups have their own launches first.  Without a
// persistent launch every group has its own launch, then wh_fold.
std::atomic<int> g_persistent{0};
std::atomic<int> g_device_cus[kMaxDevices];

int device_cus(int dev) {
    int cus = g_device_cus[dev].load(std::memory_order_relaxed);
    if (cus > 0) return cus;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    g_device_cus[dev].store(cus, std::memory_order_relaxed);
    return cus;
}

bool mixed_group(const GroupKey& k) { return k.kind == KIND_PM1S && !k.real && k.T >= 4 && k.T <= 6; }

// Enqueues the sieve kernels of every group and the fold.
// Launches the next wh_debug_fail_launches() count of batch launches refuse (tests of the error
// paths: nothing is launched, so no device state is touched).
std::atomic<int> g_fail_launches{0};

hipError_t run_layout(const Layout& L, const wh_job_dev* d_jobs, const void* pool, double4* partials,
                      void* out, hipStream_t s, unsigned* counters = nullptr) {
    for (int f = g_fail_launches.load(std::memory_order_relaxed); f > 0;
         f = g_fail_launches.load(std::memory_order_relaxed))
        if (g_fail_launches.compare_exchange_weak(f, f - 1, std::memory_order_relaxed)) return hipErrorLaunchFailure;
    std::vector<wh_mixed_group> mixed;
    const int mode = g_persistent.load(std::memory_order_relaxed);
    if (counters && mode > 0) {
        for (const Group& g : L.groups)
            if (mixed_group(g.key))
                mixed.push_back(wh_mixed_group{g.block_begin, g.block_end, (int32_t)g.job_begin, (int32_t)g.job_end,
                                               g.key.T, g.key.loop, g.nw, 0});
        if (mode == 1 && mixed.size() < 2) mixed.clear();
    }
    for (const Group& g : L.groups) {
        if (!mixed.empty() && mixed_group(g.key)) continue;
        const wh_job_dev* gj = d_jobs + g.job_begin;
        int nj = (int)(g.job_end - g.job_begin);
        int64_t nb = g.block_end - g.block_begin;
        hipError_t e;
        if (g.key.kind == KIND_PM1S)
            e = wh_launch_pm1s(g.key.T, g.key.loop != 0, g.key.real != 0, gj, nj, pool, partials, nb, g.block_begin,
                               g.nw, s);
        else if (g.key.kind == KIND_PM1M)
            e = wh_launch_pm1m(g.key.T, g.key.loop != 0, g.key.mtl != 0, gj, nj, pool, partials, nb, g.block_begin,
                               g.nw, s);
        else if (g.key.kind == KIND_PM1)
            e = wh_launch_pm1(g.key.T, g.key.half != 0, g.key.real != 0, g.key.loop != 0, gj, nj, pool,
                              partials, nb,
                              g.block_begin, g.nw, s);
        else if (g.key.kind == KIND_REP)
            e = wh_launch_rep(g.key.T, g.key.H, g.key.real != 0, g.key.loop != 0, gj, nj, pool, partials, nb,
                              g.block_begin, g.nw, s);
        else
            e = wh_launch_generic(gj, nj, pool, partials, nb, g.block_begin, g.MM, g.NQ, s);
        if (e != hipSuccess) return e;
    }
    if (mixed.empty()) return wh_launch_fold(d_jobs, (int)L.jobs.size(), partials, out, s);
    // largest tile class first: the long chunks start early and the short ones fill the tail
    std::stable_sort(mixed.begin(), mixed.end(),
                     [](const wh_mixed_group& a, const wh_mixed_group& b) { return a.T > b.T; });
    int64_t items = 0;
    for (const wh_mixed_group& m : mixed) items += m.block_end - m.block_begin;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const int nwg = (int)std::min<int64_t>(items, device_cus(dev < 0 || dev >= kMaxDevices ? 0 : dev));
    return wh_launch_pm1s_persistent(mixed.data(), (int)mixed.size(), d_jobs, (int)L.jobs.size(), pool, partials,
                                     counters, out, nwg, s);
}

}  // namespace

struct wh_batch {
    Layout layout;
};

struct wh_plan {
    int device = -1;
    int64_t n = 0;
    Layout layout;
    wh_job_dev* d_jobs = nullptr;
    double4* d_partials = nullptr;
    unsigned* d_counters = nullptr;  // the persistent launch's two work counters
};

struct wh_graph {
    int device = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

const char* wh_last_error(void) { return g_err.c_str(); }

int wh_version(void) { return 100; }

int wh_init(int device) {
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count || device >= kMaxDevices)
        return fail(WH_EINVAL, "device %d not present (count %d)", device, count);
    HIP_TRY(hipSetDevice(device));
    Context* c;
    int rc = ensure_context(device, &c);
    if (rc) return rc;
    void* scratch_dev = nullptr;  // resolve the no-allocation scratch before any graph capture
    return scratch_address(device, &scratch_dev);
}

int wh_batch_create(const wh_request* reqs, int64_t n, int64_t chunk_terms, wh_batch** batch) {
    if (!reqs || n <= 0 || !batch) return fail(WH_EINVAL, "empty request list");
    if (n > (int64_t)1 << 30) return fail(WH_EINVAL, "too many requests");
    wh_batch* b = new wh_batch();
    int rc = build_layout(reqs, n, chunk_terms, b->layout);
    if (rc) { delete b; return rc; }
    *batch = b;
    return WH_OK;
}

int wh_batch_info(const wh_batch* b, size_t* job_table_bytes, int64_t* partial_slots, int64_t* terms,
                  double* flops) {
    if (!b) return fail(WH_EINVAL, "null batch");
    if (job_table_bytes) *job_table_bytes = b->layout.jobs.size() * sizeof(wh_job_dev);
    if (partial_slots) *partial_slots = b->layout.blocks;
    if (terms) *terms = b->layout.terms;
    if (flops) *flops = b->layout.flops;
    return WH_OK;
}

int wh_batch_write_jobs(const wh_batch* b, void* host_dst) {
    if (!b || !host_dst) return fail(WH_EINVAL, "null batch or buffer");
    memcpy(host_dst, b->layout.jobs.data(), b->layout.jobs.size() * sizeof(wh_job_dev));
    return WH_OK;
}

int wh_batch_launch(const wh_batch* b, const void* h_jobs, void* d_jobs, const double* dpool,
                    double* d_partials, double* dout, void* stream) {
    if (!b || !d_jobs || !dpool || !d_partials || !dout) return fail(WH_EINVAL, "null batch or buffer");
    if (((uintptr_t)d_jobs | (uintptr_t)d_partials) & 31)
        return fail(WH_EINVAL, "job table and partials must be 32-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    if (h_jobs)
        HIP_TRY(hipMemcpyAsync(d_jobs, h_jobs, b->layout.jobs.size() * sizeof(wh_job_dev),
                               hipMemcpyHostToDevice, s));
    HIP_TRY(run_layout(b->layout, (const wh_job_dev*)d_jobs, dpool, (double4*)d_partials, dout, s));
    return WH_OK;
}

int wh_batch_launch_noalloc(const wh_batch* b, const void* h_jobs, const double* dpool, double* dout,
                            void* stream) {
    if (!b || !h_jobs || !dpool || !dout) return fail(WH_EINVAL, "null batch or buffer");
    const Layout& L = b->layout;
    const size_t jb = L.jobs.size() * sizeof(wh_job_dev);
    if (jb > kScratchJobBytes || L.blocks > kScratchSlots)
        return fail(WH_EINVAL,
                    "batch exceeds the static scratch (job table %zu of %zu B, %lld of %lld partial slots)", jb,
                    kScratchJobBytes, (long long)L.blocks, (long long)kScratchSlots);
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    void* base = nullptr;
    rc = scratch_address(dev, &base);
    if (rc) return rc;
    // Nothing here has to start at zero: the job table is copied in first, and every partial slot
    // the fold reads is written by one workgroup of this launch's sieve kernels before that.
    wh_job_dev* d_jobs = (wh_job_dev*)base;
    double4* d_partials = (double4*)((unsigned char*)base + kScratchJobBytes);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(d_jobs, h_jobs, jb, hipMemcpyHostToDevice, s));
    HIP_TRY(run_layout(L, d_jobs, dpool, d_partials, dout, s));
    return WH_OK;
}

int wh_batch_submit(const wh_batch* b, void* h_jobs, void* d_jobs, size_t jobs_capacity_bytes,
                    const double* h_pool, double* d_pool, int64_t pool_elems, double* d_partials,
                    int64_t partial_capacity, double* d_out, double* h_out, int64_t out_capacity,
                    void* stream) {
    if (!b || !h_jobs || !d_jobs || !h_pool || !d_pool || !d_partials || !d_out || !h_out)
        return fail(WH_EINVAL, "null batch or buffer");
    if (((uintptr_t)d_jobs | (uintptr_t)d_partials) & 31)
        return fail(WH_EINVAL, "job table and partials must be 32-byte aligned");
    const Layout& L = b->layout;
    const size_t jb = L.jobs.size() * sizeof(wh_job_dev);
    const int64_t nout = (int64_t)L.jobs.size();
    int64_t need_pool = 0;
    for (const wh_job_dev& j : L.jobs) need_pool = std::max(need_pool, job_pool_end(j));
    if (jb > jobs_capacity_bytes || L.blocks > partial_capacity || nout > out_capacity || need_pool > pool_elems)
        return fail(WH_EINVAL,
                    "batch exceeds the arenas (job table %zu of %zu B, %lld of %lld partial slots, %lld of %lld "
                    "results, pool %lld of %lld elements)",
                    jb, jobs_capacity_bytes, (long long)L.blocks, (long long)partial_capacity, (long long)nout,
                    (long long)out_capacity, (long long)need_pool, (long long)pool_elems);
    hipStream_t s = (hipStream_t)stream;
    memcpy(h_jobs, L.jobs.data(), jb);
    if (pool_elems > 0)
        HIP_TRY(hipMemcpyAsync(d_pool, h_pool, (size_t)pool_elems * 16, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_jobs, h_jobs, jb, hipMemcpyHostToDevice, s));
    HIP_TRY(run_layout(L, (const wh_job_dev*)d_jobs, d_pool, (double4*)d_partials, d_out, s));
    HIP_TRY(hipMemcpyAsync(h_out, d_out, (size_t)nout * 16, hipMemcpyDeviceToHost, s));
    return WH_OK;
}

int wh_request_partials(const wh_request* req, int64_t chunk_terms, int64_t* partial_slots) {
    if (!req || !partial_slots) return fail(WH_EINVAL, "null request or output");
    Layout L;
    int rc = build_layout(req, 1, chunk_terms, L);
    if (rc) return rc;
    *partial_slots = L.blocks;
    return WH_OK;
}

int wh_batch_chunks(const wh_batch* b, int64_t* first_chunk, int64_t* nchunks, int64_t* chunk_terms) {
    if (!b) return fail(WH_EINVAL, "null batch");
    for (const wh_job_dev& j : b->layout.jobs) {
        int64_t i = j.out_index;
        if (first_chunk) first_chunk[i] = j.chunk_base;
        if (nchunks) nchunks[i] = j.nchunks;
        if (chunk_terms) chunk_terms[i] = j.chunk_terms;
    }
    return WH_OK;
}

int wh_batch_destroy(wh_batch* b) {
    delete b;
    return WH_OK;
}

int wh_plan_create(const wh_request* reqs, int64_t n, int64_t chunk_terms, wh_plan** plan) {
    if (!reqs || n <= 0 || !plan) return fail(WH_EINVAL, "empty request list");
    if (n > (int64_t)1 << 30) return fail(WH_EINVAL, "too many requests");
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    wh_plan* p = new wh_plan();
    p->device = dev;
    p->n = n;
    rc = build_layout(reqs, n, chunk_terms, p->layout);
    if (rc) { delete p; return rc; }
    const Layout& L = p->layout;
    if (hipMalloc((void**)&p->d_jobs, L.jobs.size() * sizeof(wh_job_dev)) != hipSuccess ||
        hipMalloc((void**)&p->d_partials, std::max<int64_t>(L.blocks, 1) * sizeof(double4)) != hipSuccess ||
        hipMalloc((void**)&p->d_counters, 256) != hipSuccess) {
        wh_plan_destroy(p);
        return fail(WH_ENOMEM, "plan allocation failed");
    }
    if (hipMemcpy(p->d_jobs, L.jobs.data(), L.jobs.size() * sizeof(wh_job_dev), hipMemcpyHostToDevice) !=
        hipSuccess) {
        wh_plan_destroy(p);
        return fail(WH_EDEVICE, "job table upload failed");
    }
    *plan = p;
    return WH_OK;
}

int wh_plan_run(wh_plan* p, const double* dpool, double* dout, void* stream) {
    if (!p || !dpool || !dout) return fail(WH_EINVAL, "null plan or buffer");
    hipStream_t s = (hipStream_t)stream;  // NULL: the device's default stream
    HIP_TRY(run_layout(p->layout, p->d_jobs, dpool, p->d_partials, dout, s, p->d_counters));
    return WH_OK;
}

int wh_plan_info(const wh_plan* p, int64_t* terms, int64_t* blocks, double* flops) {
    if (!p) return fail(WH_EINVAL, "null plan");
    if (terms) *terms = p->layout.terms;
    if (blocks) *blocks = p->layout.blocks;
    if (flops) *flops = p->layout.flops;
    return WH_OK;
}

int wh_plan_partials(const wh_plan* p, double* host_out) {
    if (!p || !host_out) return fail(WH_EINVAL, "null plan or buffer");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(host_out, p->d_partials, p->layout.blocks * sizeof(double4), hipMemcpyDeviceToHost));
    // reorder nothing: slot b belongs to the job whose chunk range holds b (request order is
    // recovered by the caller through wh_plan_chunks)
    return WH_OK;
}

int wh_plan_chunks(const wh_plan* p, int64_t* first_chunk, int64_t* nchunks, int64_t* chunk_terms) {
    if (!p) return fail(WH_EINVAL, "null plan");
    for (const wh_job_dev& j : p->layout.jobs) {
        int64_t i = j.out_index;
        if (first_chunk) first_chunk[i] = j.chunk_base;
        if (nchunks) nchunks[i] = j.nchunks;
        if (chunk_terms) chunk_terms[i] = j.chunk_terms;
    }
    return WH_OK;
}

int wh_plan_destroy(wh_plan* p) {
    if (!p) return WH_OK;
    if (p->d_jobs) (void)hipFree(p->d_jobs);
    if (p->d_partials) (void)hipFree(p->d_partials);
    if (p->d_counters) (void)hipFree(p->d_counters);
    delete p;
    return WH_OK;
}

int wh_plan_graph_create(wh_plan* p, const double* dpool, double* dout, void* stream, wh_graph** out) {
    if (!p || !dpool || !dout || !out) return fail(WH_EINVAL, "null plan, buffer or graph");
    if (!stream) return fail(WH_EINVAL, "graph capture needs a stream other than the default stream");
    hipStream_t s = (hipStream_t)stream;
    wh_graph* g = new wh_graph();
    g->device = p->device;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return fail(WH_EDEVICE, "stream capture failed to start: %s", hipGetErrorString(e));
    }
    hipError_t er = run_layout(p->layout, p->d_jobs, dpool, p->d_partials, dout, s, p->d_counters);
    hipError_t ee = hipStreamEndCapture(s, &g->graph);
    if (er != hipSuccess || ee != hipSuccess) {
        wh_graph_destroy(g);
        return fail(WH_EDEVICE, "plan capture failed: %s",
                    hipGetErrorString(er != hipSuccess ? er : ee));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        wh_graph_destroy(g);
        return fail(WH_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return WH_OK;
}

int wh_graph_launch(wh_graph* g, void* stream) {
    if (!g || !g->exec) return fail(WH_EINVAL, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return WH_OK;
}

int wh_graph_destroy(wh_graph* g) {
    if (!g) return WH_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return WH_OK;
}

int wh_tune_set(int kind, int T, int loop, int64_t blocks_per_request) {
    if (kind < 0 || kind > 2 || loop < 0 || loop > 1 || blocks_per_request < 0 ||
        (kind == 0 && (T < 1 || T > kPm1MaxT)) || (kind == 1 && T != 0) || (kind == 2 && (T < 1 || T > kRepMaxT)))
        return fail(WH_EINVAL, "tune entry out of range (kind %d, T %d, loop %d, blocks %lld)", kind, T, loop,
                    (long long)blocks_per_request);
    g_tune_blocks[kind][T][loop].store(blocks_per_request, std::memory_order_relaxed);
    return WH_OK;
}

int wh_tune_set_waves(int kind, int T, int loop, int waves) {
    if (kind < 0 || kind > 2 || kind == 1 || loop < 0 || loop > 1 || waves < 0 || waves > 4 ||
        (kind == 0 && (T < 1 || T > kPm1MaxT)) || (kind == 2 && (T < 1 || T > kRepMaxT)))
        return fail(WH_EINVAL, "waves entry out of range (kind %d, T %d, loop %d, waves %d)", kind, T, loop, waves);
    g_tune_waves[kind][T][loop].store(waves, std::memory_order_relaxed);
    return WH_OK;
}

int wh_tune_get_waves(int kind, int T, int loop) {
    if (kind < 0 || kind > 2 || kind == 1 || loop < 0 || loop > 1 || (kind == 0 && (T < 1 || T > kPm1MaxT)) ||
        (kind == 2 && (T < 1 || T > kRepMaxT)))
        return -1;
    return g_tune_waves[kind][T][loop].load(std::memory_order_relaxed);
}

int wh_tune_set_tile(int kind, int T, int loop, int tile) {
    if (kind < 0 || kind > 2 || kind == 1 || loop < 0 || loop > 1 || (kind == 0 && (T < 1 || T > kPm1MaxT)) ||
        (kind == 2 && (T < 1 || T > kRepMaxT)) ||
        (tile != 0 && (kind == 0 ? tile <= T || tile > kPm1sMaxT
                                 : tile < 1 || tile > kRepMaxT || rep_cap(tile) <= rep_cap(T))))
        return fail(WH_EINVAL, "tile entry out of range (kind %d, T %d, loop %d, tile %d)", kind, T, loop, tile);
    g_tune_tile[kind][T][loop].store(tile, std::memory_order_relaxed);
    return WH_OK;
}

int wh_tune_get_tile(int kind, int T, int loop) {
    if (kind < 0 || kind > 2 || kind == 1 || loop < 0 || loop > 1 || (kind == 0 && (T < 1 || T > kPm1MaxT)) ||
        (kind == 2 && (T < 1 || T > kRepMaxT)))
        return -1;
    return g_tune_tile[kind][T][loop].load(std::memory_order_relaxed);
}

int wh_debug_fail_launches(int count) {
    if (count < 0) return g_fail_launches.load(std::memory_order_relaxed);
    return g_fail_launches.exchange(count, std::memory_order_relaxed);
}

int wh_tune_persistent(int mode) {
    if (mode > 2) return fail(WH_EINVAL, "persistent mode %d out of range", mode);
    if (mode < 0) return g_persistent.load(std::memory_order_relaxed);
    return g_persistent.exchange(mode, std::memory_order_relaxed);
}

int64_t wh_tune_get(int kind, int T, int loop) {
    if (kind < 0 || kind > 2 || loop < 0 || loop > 1 || (kind == 0 && (T < 1 || T > kPm1MaxT)) ||
        (kind == 1 && T != 0) || (kind == 2 && (T < 1 || T > kRepMaxT)))
        return -1;
    return g_tune_blocks[kind][T][loop].load(std::memory_order_relaxed);
}

// Stages one request ([job | A | D | V] in one pinned block, one H2D copy), runs it and
// waits for the value.  r: a_off / d_off / v_off are filled here from which of D / V is given.
static int run_request(wh_request r, const double* A, const double* D, const double* V, int64_t side,
                       double* out) {
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    Context* c;
    rc = ensure_context(dev, &c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    // staging layout (complex elements): A at 0, D after it, V after that
    size_t nA = (size_t)side * side;
    r.a_off = 0;
    r.d_off = D ? (int64_t)nA : -1;
    r.v_off = V ? (int64_t)(nA + side) : -1;
    Layout L;
    rc = build_layout(&r, 1, 0, L);
    if (rc) return rc;
    size_t job_bytes = (sizeof(wh_job_dev) + 255) & ~(size_t)255;
    size_t data_elems = nA + (D ? side : 0) + (V ? side : 0);
    size_t bytes = job_bytes + data_elems * 16;
    rc = ensure_stage(*c, bytes);
    if (rc) return rc;
    rc = ensure_partials(*c, (size_t)std::max<int64_t>(L.blocks, 1));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));  // staging is reused: previous copies done
    memcpy(c->h_stage, L.jobs.data(), sizeof(wh_job_dev));
    double* hd = (double*)(c->h_stage + job_bytes);
    memcpy(hd, A, nA * 16);
    if (D) memcpy(hd + 2 * nA, D, side * 16);
    if (V) memcpy(hd + 2 * (nA + side), V, side * 16);
    HIP_TRY(hipMemcpyAsync(c->d_stage, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
    const wh_job_dev* djob = (const wh_job_dev*)c->d_stage;
    const void* dpool = c->d_stage + job_bytes;
    HIP_TRY(run_layout(L, djob, dpool, c->d_partials, c->d_out, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_out, c->d_out, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out[0] = c->h_out[0];
    out[1] = c->h_out[1];
    return WH_OK;
}

static bool all_real(const double* x, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (x[2 * i + 1] != 0.0) return false;
    return true;
}

// Shared body of the two synchronous hafnian entry points.
static int calc_one(const double* A, const double* D, const int64_t* edge_reps, int64_t E,
                    const double* oddloop, const double* oddV, int glynn, double* out) {
    if (!A || !edge_reps || !out) return fail(WH_EINVAL, "null argument");
    if (E < 1 || E > WH_MAX_EDGES) return fail(WH_EINVAL, "E=%lld outside 1..%d", (long long)E, WH_MAX_EDGES);
    if (oddV && !oddloop) return fail(WH_EINVAL, "oddV given without oddloop");
    if (oddloop && !oddV) return fail(WH_EINVAL, "oddloop given without oddV");
    if (oddloop && !D) return fail(WH_EINVAL, "odd-N case needs D");
    const int64_t side = 2 * E;
    wh_request r;
    memset(&r, 0, sizeof r);
    r.E = (int32_t)E;
    r.glynn = glynn ? 1 : 0;
    for (int64_t i = 0; i < E; ++i) {
        if (edge_reps[i] < 0 || edge_reps[i] > (1 << 20)) return fail(WH_EINVAL, "edge multiplicity out of range");
        r.edge_reps[i] = (int32_t)edge_reps[i];
    }
    // zero imaginary parts everywhere: the real-arithmetic kernels compute the same values
    r.real = (all_real(A, (size_t)side * side) && (!D || all_real(D, side)) && !oddV && !oddloop) ? 1 : 0;
    if (oddloop) { r.oddloop[0] = oddloop[0]; r.oddloop[1] = oddloop[1]; }
    return run_request(r, A, D, oddV, side, out);
}

// Shared body of the montrealer entry points: stages X Sigma and, for the loop montrealer,
// X zeta and conj(X zeta) (the all-ones kernel's M = Ax X then equals X Sigma X).
static int calc_mtl(const double* Sigma, const double* zeta, int64_t twoN, double* out) {
    if (!Sigma || !out) return fail(WH_EINVAL, "null argument");
    if (twoN < 2 || twoN % 2) return fail(WH_EINVAL, "matrix side %lld must be even and >= 2", (long long)twoN);
    const int64_t n = twoN / 2;
    if (n > kMtlMaxModes) return fail(WH_EUNSUPPORTED, "montrealer of %lld modes exceeds %d", (long long)n, kMtlMaxModes);
    std::vector<double> ax((size_t)twoN * twoN * 2), xz, cxz;
    for (int64_t r = 0; r < twoN; ++r) {
        int64_t src = r < n ? r + n : r - n;
        memcpy(&ax[(size_t)r * twoN * 2], Sigma + (size_t)src * twoN * 2, (size_t)twoN * 16);
    }
    if (zeta) {
        xz.resize(2 * twoN);
        cxz.resize(2 * twoN);
        for (int64_t e = 0; e < twoN; ++e) {
            int64_t src = e < n ? e + n : e - n;
            xz[2 * e] = cxz[2 * e] = zeta[2 * src];
            xz[2 * e + 1] = zeta[2 * src + 1];
            cxz[2 * e + 1] = -zeta[2 * src + 1];
        }
    }
    wh_request r;
    memset(&r, 0, sizeof r);
    r.E = (int32_t)n;
    r.options = WH_OPT_MONTREALER;
    r.real = (all_real(Sigma, (size_t)twoN * twoN) && (!zeta || all_real(zeta, twoN))) ? 1 : 0;
    return run_request(r, ax.data(), zeta ? xz.data() : nullptr, zeta ? cxz.data() : nullptr, twoN, out);
}

int wh_calc_hafnian(const double* A, int64_t twoE, const int64_t* edge_reps, int64_t E, int glynn,
                    double* out) {
    if (twoE != 2 * E) return fail(WH_EINVAL, "matrix side %lld != 2E (E=%lld)", (long long)twoE, (long long)E);
    return calc_one(A, nullptr, edge_reps, E, nullptr, nullptr, glynn, out);
}

int wh_calc_loop_hafnian(const double* A, const double* D, const int64_t* edge_reps, int64_t E,
                         const double* oddloop, const double* oddV, int glynn, double* out) {
    if (!D) return fail(WH_EINVAL, "loop hafnian needs D");
    return calc_one(A, D, edge_reps, E, oddloop, oddV, glynn, out);
}

int wh_montrealer(const double* Sigma, int64_t twoN, double* out) { return calc_mtl(Sigma, nullptr, twoN, out); }

int wh_lmontrealer(const double* Sigma, const double* zeta, int64_t twoN, double* out) {
    if (!zeta) return fail(WH_EINVAL, "loop montrealer needs zeta");
    return calc_mtl(Sigma, zeta, twoN, out);
}

// Shared body of wh_loop_hafnian_batch (nD = 1) and wh_loop_hafnian_batch_gamma.
static int lhb_call(const double* A, const double* D, int64_t nD, int64_t twoE, const int64_t* fixed_edge_reps,
                    int64_t Ef, int64_t batch_max, int64_t cutoff, int odd_variant, int glynn, double* out,
                    int64_t out_len) {
    if (!A || !D || !out || (Ef > 0 && !fixed_edge_reps)) return fail(WH_EINVAL, "null argument");
    if (Ef < 0 || batch_max < 0) return fail(WH_EINVAL, "negative size");
    if (nD < 1) return fail(WH_EINVAL, "no diagonal given");
    if (cutoff != 0 && cutoff != 1) return fail(WH_EINVAL, "cutoff must be 0 or 1");
    const int64_t lead = odd_variant ? 2 : 1;
    const int64_t E = Ef + lead;
    if (twoE != 2 * E)
        return fail(WH_EINVAL, "matrix side %lld != 2(Ef + %lld) (Ef=%lld)", (long long)twoE, (long long)lead,
                    (long long)Ef);
    if (twoE > kGenericMaxSide)
        return fail(WH_EUNSUPPORTED, "matrix side %lld exceeds %d", (long long)twoE, kGenericMaxSide);
    const int64_t nL = 2 * batch_max + cutoff + 1 + (odd_variant ? 1 : 0);
    if (nL * nD > 128)
        return fail(WH_EUNSUPPORTED, "%lld outputs (%lld diagonals x %lld) exceed 128", (long long)(nL * nD),
                    (long long)nD, (long long)nL);
    wh_lhb_job j;
    memset(&j, 0, sizeof j);
    j.e[0] = (int32_t)batch_max;
    if (odd_variant) j.e[1] = 1;
    int64_t fsum = 0;
    double steps_d = (batch_max + 1.0) * (odd_variant ? 2.0 : 1.0);
    for (int64_t i = 0; i < Ef; ++i) {
        if (fixed_edge_reps[i] < 0 || fixed_edge_reps[i] > WH_MAX_PHOTONS)
            return fail(WH_EINVAL, "edge multiplicity %lld out of range", (long long)fixed_edge_reps[i]);
        j.e[lead + i] = (int32_t)fixed_edge_reps[i];
        fsum += fixed_edge_reps[i];
        steps_d *= fixed_edge_reps[i] + 1.0;
    }
    const int64_t n_fixed = 2 * fsum + (odd_variant ? 1 : 0);
    if (n_fixed + nL - 1 > WH_MAX_PHOTONS)
        return fail(WH_EUNSUPPORTED, "photon number %lld exceeds %d", (long long)(n_fixed + nL - 1), WH_MAX_PHOTONS);
    if (steps_d > 9.0e18) return fail(WH_EUNSUPPORTED, "term count exceeds int64");
    if (out_len < nL * nD)
        return fail(WH_EINVAL, "output holds %lld values, %lld needed", (long long)out_len, (long long)(nL * nD));
    int64_t steps = (batch_max + 1) * (odd_variant ? 2 : 1);
    for (int64_t i = 0; i < Ef; ++i) steps *= fixed_edge_reps[i] + 1;
    int64_t chunk = std::min<int64_t>(std::max<int64_t>(steps / 1024, 1), 256);
    const int64_t nch = steps / chunk + (steps % chunk != 0);
    if (nch > kMaxBlocks / nD)
        return fail(WH_EUNSUPPORTED, "%lld term chunks (%lld terms) exceed the launch limit", (long long)nch,
                    (long long)steps);
    j.E = (int32_t)E;
    j.L = (int32_t)nL;
    j.nD = (int32_t)nD;
    j.n_fixed = (int32_t)n_fixed;
    j.n_max = (int32_t)(n_fixed + nL - 1);
    j.flags = (odd_variant ? WH_LHB_ODD : 0) | (glynn ? WH_LHB_GLYNN : 0);
    j.steps = steps;
    j.chunk_terms = chunk;
    j.nchunks = (int32_t)nch;
    // few chunks (small term counts): one workgroup per chunk and diagonal, so the fanout rows
    // run in parallel (each repeats the chunk's shared per-term traces)
    j.dsplit = (nD > 1 && j.nchunks < 1024) ? (int32_t)nD : 1;

    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    Context* c;
    rc = ensure_context(dev, &c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    // staging (bytes): [job | A | D rows | outputs]; the output area is written on the device
    // and copied back into the same offsets of the pinned block
    const size_t nA = (size_t)twoE * twoE, nDv = (size_t)nD * twoE, nout = (size_t)(nL * nD);
    const size_t job_bytes = (sizeof(wh_lhb_job) + 255) & ~(size_t)255;
    j.a_off = 0;
    j.d_off = (int64_t)nA;
    const size_t in_bytes = job_bytes + (nA + nDv) * 16;
    const size_t out_at = (in_bytes + 255) & ~(size_t)255;
    rc = ensure_stage(*c, out_at + nout * 16);
    if (rc) return rc;
    rc = ensure_partials(*c, (size_t)j.nchunks * nout);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(c->h_stage, &j, sizeof j);
    memcpy(c->h_stage + job_bytes, A, nA * 16);
    memcpy(c->h_stage + job_bytes + nA * 16, D, nDv * 16);
    HIP_TRY(hipMemcpyAsync(c->d_stage, c->h_stage, in_bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(wh_launch_lhb((const wh_lhb_job*)c->d_stage, 1, c->d_stage + job_bytes, c->d_partials,
                          (int64_t)j.nchunks * j.dsplit, (int)twoE, c->d_stage + out_at, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_stage + out_at, c->d_stage + out_at, nout * 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(out, c->h_stage + out_at, nout * 16);
    return (int)(nL * nD);
}

int wh_loop_hafnian_batch(const double* A, const double* D, int64_t twoE, const int64_t* fixed_edge_reps,
                          int64_t Ef, int64_t batch_max, int64_t cutoff, int odd_variant, int glynn,
                          double* out, int64_t out_len) {
    return lhb_call(A, D, 1, twoE, fixed_edge_reps, Ef, batch_max, cutoff, odd_variant, glynn, out, out_len);
}

int wh_loop_hafnian_batch_gamma(const double* A, const double* D, int64_t nD, int64_t twoE,
                                const int64_t* fixed_edge_reps, int64_t Ef, int64_t batch_max, int64_t cutoff,
                                int odd_variant, int glynn, double* out, int64_t out_len) {
    return lhb_call(A, D, nD, twoE, fixed_edge_reps, Ef, batch_max, cutoff, odd_variant, glynn, out, out_len);
}

int wh_hermite_multidimensional(const double* R, const double* y, int64_t n, const int64_t* cutoffs,
                                const double* C, int renorm, double* out) {
    if (!R || !y || !cutoffs || !C || !out) return fail(WH_EINVAL, "null argument");
    if (n < 1 || n > 32) return fail(WH_EUNSUPPORTED, "n=%lld outside 1..32", (long long)n);
    int cut[32];
    int64_t total = 1;
    int maxcut = 1;
    for (int64_t l = 0; l < n; ++l) {
        if (cutoffs[l] < 1 || cutoffs[l] > (1 << 20)) return fail(WH_EINVAL, "cutoff %lld out of range", (long long)cutoffs[l]);
        cut[l] = (int)cutoffs[l];
        maxcut = std::max(maxcut, cut[l]);
        if ((double)total * cut[l] > 4.0e10) return fail(WH_EUNSUPPORTED, "tensor of more than 4e10 entries");
        total *= cut[l];
    }
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    Context* c;
    rc = ensure_context(dev, &c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    // staging (bytes): [R | y | sqrt table]; the tensor gets its own device allocation
    const size_t nR = (size_t)n * n * 16, ny = (size_t)n * 16, nsq = (size_t)maxcut * 8;
    const size_t in_bytes = nR + ny + nsq;
    rc = ensure_stage(*c, in_bytes);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(c->h_stage, R, nR);
    memcpy(c->h_stage + nR, y, ny);
    double* sq = (double*)(c->h_stage + nR + ny);
    for (int k = 0; k < maxcut; ++k) sq[k] = std::sqrt((double)k);
    void* G = nullptr;
    if (hipMalloc(&G, (size_t)total * 16) != hipSuccess)
        return fail(WH_ENOMEM, "device tensor of %lld entries", (long long)total);
    hipError_t e = hipMemcpyAsync(c->d_stage, c->h_stage, in_bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(G, C, 16, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = wh_launch_hermite((int)n, cut, renorm, c->d_stage, c->d_stage + nR, (const double*)(c->d_stage + nR + ny), G,
                              c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, G, (size_t)total * 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(G);
    if (e != hipSuccess) return fail(WH_EDEVICE, "hermite recursion: %s", hipGetErrorString(e));
    return WH_OK;
}

}  // extern "C"

// ---- Host preparation of the common request in one pass: n modes with one photon each (n even,
// 6 <= n <= 64), an exactly symmetric NaN-free matrix whose entry (0, 1) is far from zero.  It is
// what prepare_hafnian / _prepare_haf / _prepare_loop produce for such a matrix
// (thewalrus/_hafnian.py:718-818, :470-508, :581-631): no validation error is possible (every
// a_ij equals a_ji, so numpy.allclose(A, A.T) holds under any tolerance), no shortcut applies
// (n > 4, and |a_01| exceeds atol + rtol |a_01| by a margin wider than any rounding difference,
// so numpy.allclose(diag(diag(A)), A) is false), and matched_reps pairs mode n-1-2k with
// n-2-2k.  Anything else returns 0: the caller runs the general host path.
namespace {

int64_t prep_allones(const double* A, int64_t n, int is_complex, int loop, int glynn, double rtol, double atol,
                     wh_request& r, double* payload) {
    if (n < 6 || n > 2 * WH_MAX_EDGES || (n & 1)) return 0;
    const int64_t st = is_complex ? 2 : 1;
    auto re = [&](int64_t i, int64_t j) { return A[(i * n + j) * st]; };
    auto im = [&](int64_t i, int64_t j) { return is_complex ? A[(i * n + j) * 2 + 1] : 0.0; };
    for (int64_t i = 0; i < n; ++i)  // NaN never compares equal: the diagonal is checked too
        for (int64_t j = i; j < n; ++j)
            if (!(re(i, j) == re(j, i)) || !(im(i, j) == im(j, i))) return 0;
    const double ay = std::hypot(re(0, 1), im(0, 1));
    if (!(ay < INFINITY) || !(ay * (1.0 - 1e-9) > atol + rtol * ay * (1.0 + 1e-9))) return 0;
    const int64_t E = n / 2;
    int64_t x[2 * WH_MAX_EDGES];
    for (int64_t k = 0; k < E; ++k) {
        x[k] = n - 1 - 2 * k;
        x[E + k] = n - 2 - 2 * k;
    }
    bool real = true;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) {
            const double a = re(x[i], x[j]), b = im(x[i], x[j]);
            payload[2 * (i * n + j)] = a;
            payload[2 * (i * n + j) + 1] = b;
            if (b != 0.0) real = false;
        }
    int64_t elems = n * n;
    if (loop) {
        for (int64_t i = 0; i < n; ++i) {
            payload[2 * (elems + i)] = re(x[i], x[i]);
            payload[2 * (elems + i) + 1] = im(x[i], x[i]);
        }
        elems += n;
    }
    memset(&r, 0, sizeof r);
    r.a_off = 0;
    r.d_off = loop ? n * n : -1;
    r.v_off = -1;
    r.E = (int32_t)E;
    r.glynn = loop ? 1 : (glynn != 0);
    r.real = real ? 1 : 0;
    r.options = 0;
    for (int64_t k = 0; k < E; ++k) r.edge_reps[k] = 1;
    return elems;
}

}  // namespace

extern "C" int64_t wh_prepare_allones(const double* A, int64_t n, int is_complex, int loop, int glynn, double rtol,
                                      double atol, wh_request* req, double* payload, int64_t payload_elems) {
    if (!A || !req || !payload) return fail(WH_EINVAL, "null matrix, request or payload");
    if (payload_elems < n * n + n) return fail(WH_EINVAL, "payload holds %lld elements, %lld needed",
                                               (long long)payload_elems, (long long)(n * n + n));
    return prep_allones(A, n, is_complex, loop, glynn, rtol, atol, *req, payload);
}

// ---- Serving engine: a native dispatcher thread over fixed per-slot arenas.  Requests queue
// with their packed payload; whenever a slot is free the thread takes queued requests in
// arrival order while they fit the slot and the HBM budget, lays them out, and enqueues copies,
// kernels and the result copy on the slot's stream; finished slots become completions that
// callers collect with wh_server_wait.  Layout depends on each request alone, so a result is
// the same alone and in any mix.  While batches run, the thread polls their events every 20 us
// (a completion is seen within that time, at the cost of one busy host thread); with nothing in
// flight it sleeps until a submission wakes it.
namespace {

double steady_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct SrvItem {
    uint64_t tag = 0;
    wh_request req;
    std::vector<double> payload;  // 2 * elems doubles
    int64_t elems = 0, slots = 0, footprint = 0;
    int64_t chunk = 0;  // terms per chunk, fixed at submission (admission counted its slots)
    double t_submit = 0.0, t_deadline = -1.0;
};

struct SrvDone {
    uint64_t tag;
    double re, im;
    int32_t status;
    double wait_s, run_s;  // submit -> launch, launch -> completion seen
};

struct SrvSlot {
    hipStream_t stream = nullptr;
    hipEvent_t event = nullptr;
    double* h_pool = nullptr;
    double* d_pool = nullptr;
    wh_job_dev* h_jobs = nullptr;
    wh_job_dev* d_jobs = nullptr;
    double4* d_partials = nullptr;
    double* h_out = nullptr;
    double* d_out = nullptr;
    unsigned* d_counters = nullptr;  // the persistent launch's two work counters
    std::vector<uint64_t> tags;
    std::vector<double> t_submit;
    double t_launch = 0.0;
    int64_t footprint = 0;
    bool busy = false;
};

constexpr int64_t kSrvJobBytes = 256;  // admission's per-request job-table share (bound on the table entry)

}  // namespace

struct wh_server {
    int device = 0;
    int64_t pool_elems = 0, partial_slots = 0, max_requests = 0, chunk_terms = 0, hbm_budget = 0;
    std::vector<SrvSlot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<SrvItem> queue;
    std::deque<SrvDone> done;
    int64_t admitted = 0;  // footprint of launched, unfinished batches
    int64_t queued_deadlines = 0;
    int64_t in_flight = 0;
    bool stop = false, exited = false;
    bool idle = false;  // the dispatcher waits for work (submissions wake it only then)
    int64_t batches = 0, launched = 0, cancelled = 0, expired = 0;
    std::string error;  // the engine's last failure
    std::thread th;
};

namespace {

void srv_free(wh_server* s) {
    for (SrvSlot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.event) (void)hipEventDestroy(sl.event);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_pool) (void)hipHostFree(sl.h_pool);
        if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.d_pool) (void)hipFree(sl.d_pool);
        if (sl.d_jobs) (void)hipFree(sl.d_jobs);
        if (sl.d_partials) (void)hipFree(sl.d_partials);
        if (sl.d_out) (void)hipFree(sl.d_out);
        if (sl.d_counters) (void)hipFree(sl.d_counters);
    }
    s->slots.clear();
}

// Enqueues one batch on a free slot (engine thread, lock not held).  On failure every request of
// the batch completes with the error status and the slot stays free.
void srv_launch(wh_server* s, SrvSlot& sl, std::vector<SrvItem>& items, int64_t fp) {
    const int64_t n = (int64_t)items.size();
    std::vector<wh_request> reqs(n);
    std::vector<int64_t> chunks(n);
    int64_t off = 0;
    for (int64_t k = 0; k < n; ++k) {
        chunks[k] = items[k].chunk;
        wh_request r = items[k].req;
        r.a_off += off;
        if (r.d_off >= 0) r.d_off += off;
        if (r.v_off >= 0) r.v_off += off;
        reqs[k] = r;
        memcpy(sl.h_pool + 2 * off, items[k].payload.data(), (size_t)items[k].elems * 16);
        off += items[k].elems;
    }
    Layout L;
    // each request keeps the chunking it was admitted with, whatever the tuning table says now
    int rc = build_layout(reqs.data(), n, 0, L, chunks.data());
    if (rc == WH_OK && (L.blocks > s->partial_slots || (int64_t)L.jobs.size() > s->max_requests))
        rc = fail(WH_EINVAL, "batch exceeds the slot arena");
    hipError_t e = hipSuccess;
    if (rc == WH_OK) {
        memcpy(sl.h_jobs, L.jobs.data(), L.jobs.size() * sizeof(wh_job_dev));
        e = hipMemcpyAsync(sl.d_pool, sl.h_pool, (size_t)off * 16, hipMemcpyHostToDevice, sl.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, L.jobs.size() * sizeof(wh_job_dev), hipMemcpyHostToDevice,
                               sl.stream);
        if (e == hipSuccess)
            e = run_layout(L, sl.d_jobs, sl.d_pool, sl.d_partials, sl.d_out, sl.stream, sl.d_counters);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sl.h_out, sl.d_out, (size_t)n * 16, hipMemcpyDeviceToHost, sl.stream);
        if (e == hipSuccess) e = hipEventRecord(sl.event, sl.stream);
        if (e != hipSuccess) (void)hipStreamSynchronize(sl.stream);  // nothing of the batch still runs
    }
    const double now = steady_now();
    std::lock_guard<std::mutex> g(s->mu);
    if (rc != WH_OK || e != hipSuccess) {
        s->error = rc != WH_OK ? g_err : std::string("batch launch failed: ") + hipGetErrorString(e);
        const int32_t st = rc != WH_OK ? rc : WH_EDEVICE;
        for (const SrvItem& it : items) s->done.push_back(SrvDone{it.tag, 0.0, 0.0, st, now - it.t_submit, 0.0});
        s->admitted -= fp;
        s->in_flight -= n;
        s->cv_done.notify_all();
        return;
    }
    sl.tags.resize(n);
    sl.t_submit.resize(n);
    for (int64_t k = 0; k < n; ++k) {
        sl.tags[k] = items[k].tag;
        sl.t_submit[k] = items[k].t_submit;
    }
    sl.t_launch = now;
    sl.footprint = fp;
    sl.busy = true;
    s->batches += 1;
    s->launched += n;
}

int srv_enqueue(wh_server* s, const wh_request& req, std::vector<double>&& payload, int64_t elems, uint64_t tag,
                double deadline_s);

void srv_loop(wh_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<SrvItem> batch;
    for (;;) {
        bool progressed = false;
        for (SrvSlot& sl : s->slots) {  // busy flags change on this thread only
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.event);
            if (q == hipErrorNotReady) continue;
            const double now = steady_now();
            std::lock_guard<std::mutex> g(s->mu);
            const int32_t st = q == hipSuccess ? WH_OK : WH_EDEVICE;
            if (q != hipSuccess) s->error = std::string("batch failed: ") + hipGetErrorString(q);
            for (size_t k = 0; k < sl.tags.size(); ++k)
                s->done.push_back(SrvDone{sl.tags[k], sl.h_out[2 * k], sl.h_out[2 * k + 1], st,
                                          sl.t_launch - sl.t_submit[k], now - sl.t_launch});
            s->admitted -= sl.footprint;
            s->in_flight -= (int64_t)sl.tags.size();
            sl.busy = false;
            sl.tags.clear();
            sl.footprint = 0;
            progressed = true;
        }
        if (progressed) s->cv_done.notify_all();

        std::unique_lock<std::mutex> lk(s->mu);
        if (s->queued_deadlines > 0) {  // queued requests past their deadline fail
            const double now = steady_now();
            bool any = false;
            for (auto it = s->queue.begin(); it != s->queue.end();) {
                if (it->t_deadline >= 0.0 && now > it->t_deadline) {
                    s->done.push_back(SrvDone{it->tag, 0.0, 0.0, 1, now - it->t_submit, 0.0});
                    s->queued_deadlines -= 1;
                    s->expired += 1;
                    it = s->queue.erase(it);
                    any = true;
                } else {
                    ++it;
                }
            }
            if (any) s->cv_done.notify_all();
        }
        bool any_busy = false;
        SrvSlot* free_slot = nullptr;
        for (SrvSlot& sl : s->slots) {
            if (sl.busy) any_busy = true;
            else if (!free_slot) free_slot = &sl;
        }
        if (s->stop && s->queue.empty() && !any_busy) {
            s->exited = true;
            lk.unlock();
            s->cv_done.notify_all();
            return;
        }
        if (free_slot && !s->queue.empty()) {
            batch.clear();
            int64_t elems = 0, parts = 0, fp = 0;
            while (!s->queue.empty() && (int64_t)batch.size() < s->max_requests) {
                const SrvItem& it = s->queue.front();
                if (elems + it.elems > s->pool_elems || parts + it.slots > s->partial_slots ||
                    s->admitted + fp + it.footprint > s->hbm_budget)
                    break;
                elems += it.elems;
                parts += it.slots;
                fp += it.footprint;
                if (it.t_deadline >= 0.0) s->queued_deadlines -= 1;
                batch.push_back(std::move(s->queue.front()));
                s->queue.pop_front();
            }
            if (!batch.empty()) {
                s->admitted += fp;
                s->in_flight += (int64_t)batch.size();
                lk.unlock();
                srv_launch(s, *free_slot, batch, fp);
                continue;
            }
        }
        if (!progressed) {
            s->idle = true;
            s->cv_work.wait_for(lk, any_busy ? std::chrono::microseconds(20) : std::chrono::milliseconds(2));
            s->idle = false;
        }
    }
}

}  // namespace

extern "C" {

int wh_server_create(int device, int nslots, int64_t pool_elems, int64_t partial_slots, int64_t max_requests,
                     int64_t chunk_terms, int64_t hbm_budget_bytes, wh_server** out) {
    if (!out || nslots < 1 || nslots > 64 || pool_elems < 1 || partial_slots < 1 || max_requests < 1 ||
        chunk_terms < 0 || hbm_budget_bytes < 1)
        return fail(WH_EINVAL, "server arguments out of range");
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(device));
    wh_server* s = new wh_server();
    s->device = device;
    s->pool_elems = pool_elems;
    s->partial_slots = partial_slots;
    s->max_requests = max_requests;
    s->chunk_terms = chunk_terms;
    s->hbm_budget = hbm_budget_bytes;
    s->slots.resize(nslots);
    hipError_t e = hipSuccess;
    for (SrvSlot& sl : s->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.event, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_pool, (size_t)pool_elems * 16, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_pool, (size_t)pool_elems * 16);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.h_jobs, (size_t)max_requests * sizeof(wh_job_dev), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_jobs, (size_t)max_requests * sizeof(wh_job_dev));
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_partials, (size_t)partial_slots * sizeof(double4));
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_out, (size_t)max_requests * 16, hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_out, (size_t)max_requests * 16);
        if (e == hipSuccess) e = hipMalloc((void**)&sl.d_counters, 256);
    }
    if (e != hipSuccess) {
        srv_free(s);
        delete s;
        (void)hipSetDevice(prev);
        return fail(WH_ENOMEM, "server arenas: %s", hipGetErrorString(e));
    }
    (void)hipSetDevice(prev);
    s->th = std::thread(srv_loop, s);
    *out = s;
    return WH_OK;
}

int wh_server_submit(wh_server* s, const wh_request* req, const double* payload, int64_t elems, uint64_t tag,
                     double deadline_s) {
    if (!s || !req || !payload || elems < 1) return fail(WH_EINVAL, "null server, request or payload");
    return srv_enqueue(s, *req, std::vector<double>(payload, payload + 2 * elems), elems, tag, deadline_s);
}

}  // extern "C"

namespace {

int srv_enqueue(wh_server* s, const wh_request& req, std::vector<double>&& payload, int64_t elems, uint64_t tag,
                double deadline_s) {
    Layout L;
    int rc = build_layout(&req, 1, s->chunk_terms, L);
    if (rc) return rc;
    if (job_pool_end(L.jobs[0]) > elems) return fail(WH_EINVAL, "request reads past its payload");
    if (elems > s->pool_elems || L.blocks > s->partial_slots)
        return fail(WH_EINVAL, "request exceeds the per-slot arena (%lld elements, %lld slots)", (long long)elems,
                    (long long)L.blocks);
    const int64_t fp = 16 * elems + 32 * L.blocks + 2 * kSrvJobBytes;
    if (fp > s->hbm_budget)
        return fail(WH_EINVAL, "request footprint %lld B exceeds the HBM budget %lld B", (long long)fp,
                    (long long)s->hbm_budget);
    SrvItem it;
    it.tag = tag;
    it.req = req;
    it.payload = std::move(payload);
    it.elems = elems;
    it.slots = L.blocks;
    it.chunk = L.jobs[0].chunk_terms;
    it.footprint = fp;
    it.t_submit = steady_now();
    it.t_deadline = deadline_s >= 0.0 ? it.t_submit + deadline_s : -1.0;
    bool wake;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) return fail(WH_EINVAL, "server is shut down");
        if (it.t_deadline >= 0.0) s->queued_deadlines += 1;
        s->queue.push_back(std::move(it));
        wake = s->idle;  // a busy dispatcher looks at the queue before it waits again
    }
    if (wake) s->cv_work.notify_one();
    return WH_OK;
}

}  // namespace

extern "C" {

int64_t wh_server_submit_matrix(wh_server* s, const double* A, int64_t n, int is_complex, int loop, int glynn,
                                double rtol, double atol, uint64_t tag, double deadline_s) {
    if (!s || !A) return fail(WH_EINVAL, "null server or matrix");
    if (n < 6 || n > 2 * WH_MAX_EDGES) return 0;
    std::vector<double> payload((size_t)2 * (n * n + n));
    wh_request r;
    const int64_t elems = prep_allones(A, n, is_complex, loop, glynn, rtol, atol, r, payload.data());
    if (elems == 0) return 0;
    payload.resize((size_t)2 * elems);
    const int rc = srv_enqueue(s, r, std::move(payload), elems, tag, deadline_s);
    return rc == WH_OK ? 1 : rc;
}

int64_t wh_server_wait(wh_server* s, uint64_t* tags, double* values, int32_t* status, double* times, int64_t max,
                       double timeout_s, double linger_s) {
    if (!s || !tags || !values || !status || !times || max < 1) return fail(WH_EINVAL, "null server or buffer");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 0.0),
                        [s] { return !s->done.empty() || s->exited; });
    if (s->done.empty()) return s->exited ? -1 : 0;
    if (linger_s > 0 && (int64_t)s->done.size() < max)  // more on the way: collect them together
        s->cv_done.wait_for(lk, std::chrono::duration<double>(linger_s), [s, max] {
            return (int64_t)s->done.size() >= max || (s->in_flight == 0 && s->queue.empty()) || s->exited;
        });
    int64_t n = 0;
    while (n < max && !s->done.empty()) {
        const SrvDone& d = s->done.front();
        tags[n] = d.tag;
        values[2 * n] = d.re;
        values[2 * n + 1] = d.im;
        status[n] = d.status;
        times[2 * n] = d.wait_s;
        times[2 * n + 1] = d.run_s;
        s->done.pop_front();
        ++n;
    }
    return n;
}

int wh_server_cancel(wh_server* s, uint64_t tag) {
    if (!s) return fail(WH_EINVAL, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        if (it->tag != tag) continue;
        if (it->t_deadline >= 0.0) s->queued_deadlines -= 1;
        s->queue.erase(it);
        s->cancelled += 1;
        return 1;
    }
    return 0;
}

int wh_server_stats(wh_server* s, int64_t* out6, char* err, int64_t err_bytes) {
    if (!s || !out6) return fail(WH_EINVAL, "null server or buffer");
    std::lock_guard<std::mutex> g(s->mu);
    out6[0] = s->batches;
    out6[1] = s->launched;
    out6[2] = s->cancelled;
    out6[3] = s->expired;
    out6[4] = (int64_t)s->queue.size();
    out6[5] = s->in_flight;
    if (err && err_bytes > 0) snprintf(err, (size_t)err_bytes, "%s", s->error.c_str());
    return WH_OK;
}

int wh_server_shutdown(wh_server* s) {
    if (!s) return WH_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->exited = true;
    }
    s->cv_done.notify_all();
    return WH_OK;
}

int wh_server_destroy(wh_server* s) {
    if (!s) return WH_OK;
    wh_server_shutdown(s);
    srv_free(s);
    delete s;
    return WH_OK;
}

}  // extern "C"
