// capi.hip -- the C ABI (include/walrus_hip.h): device contexts, launch plans, and the
// per-call entry points that stand in for thewalrus's numba kernels _calc_hafnian
// (thewalrus/_hafnian.py:417) and _calc_loop_hafnian (thewalrus/_hafnian.py:513).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
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

constexpr int kGenericMaxSide = 48;  // LDS holds three 48 x 48 complex matrices per wave

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

// Validates one request and fills the device descriptor (without chunking).
// _hafnian.py:430-438 and :531-538 define N and the number of terms.
int describe(const wh_request& r, wh_job_dev& j) {
    if (r.E < 1 || r.E > WH_MAX_EDGES) return fail(WH_EINVAL, "E=%d outside 1..%d", r.E, WH_MAX_EDGES);
    memset(&j, 0, sizeof j);
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
              (r.real ? WH_F_REAL : 0);
    return WH_OK;
}

int64_t auto_chunk(int64_t steps) {
    // a chunk is walked by one wave: small jobs get one term per wave, large ones enough
    // terms to amortise the chunk's fixed cost while keeping thousands of workgroups
    int64_t c = steps / 256;
    if (c < 1) c = 1;
    if (c > 256) c = 256;
    return c;
}

// Orders the kernel groups, chunks each job, and reports the number of workgroups.
struct Layout {
    std::vector<wh_job_dev> jobs;
    int64_t n_generic = 0;
    int64_t blocks_generic = 0;
    int MM = 0;
    int64_t terms = 0;
    double flops = 0.0;
};

double term_flops(const wh_job_dev& j, int m) {
    // products as the generic kernel runs them: (min(h+1, m) - 2) complex m^3 products
    int h = j.N / 2;
    int np = std::min(h + 1, m) - 2;
    if (np < 0) np = 0;
    return 8.0 * m * (double)m * m * np;
}

int build_layout(const wh_request* reqs, int64_t n, int64_t chunk_hint, Layout& L) {
    L.jobs.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        int rc = describe(reqs[i], L.jobs[i]);
        if (rc) return rc;
        L.jobs[i].out_index = i;
        int side = 2 * reqs[i].E;
        if (side > kGenericMaxSide)
            return fail(WH_EUNSUPPORTED, "matrix side %d exceeds the generic kernel limit %d", side,
                        kGenericMaxSide);
        L.MM = std::max(L.MM, side);
    }
    int64_t base = 0;
    for (auto& j : L.jobs) {
        int64_t c = chunk_hint > 0 ? chunk_hint : auto_chunk(j.steps);
        j.chunk_terms = c;
        j.nchunks = (int32_t)((j.steps + c - 1) / c);
        j.chunk_base = base;
        base += j.nchunks;
        L.terms += j.steps;
        L.flops += (double)j.steps * term_flops(j, 2 * j.E);
    }
    L.n_generic = n;
    L.blocks_generic = base;
    return WH_OK;
}

}  // namespace

struct wh_plan {
    int device = -1;
    int64_t n = 0;
    Layout layout;
    wh_job_dev* d_jobs = nullptr;
    double4* d_partials = nullptr;
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
    return ensure_context(device, &c);
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
        hipMalloc((void**)&p->d_partials, std::max<int64_t>(L.blocks_generic, 1) * sizeof(double4)) !=
            hipSuccess) {
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
    hipStream_t s = (hipStream_t)stream;
    if (!s) {
        Context* c;
        int rc = ensure_context(p->device, &c);
        if (rc) return rc;
        s = c->stream;
    }
    const Layout& L = p->layout;
    if (L.n_generic > 0)
        HIP_TRY(wh_launch_generic(p->d_jobs, (int)L.n_generic, dpool, p->d_partials, L.blocks_generic,
                                  L.MM, s));
    HIP_TRY(wh_launch_fold(p->d_jobs, (int)L.jobs.size(), p->d_partials, dout, s));
    return WH_OK;
}

int wh_plan_info(const wh_plan* p, int64_t* terms, int64_t* blocks, double* flops) {
    if (!p) return fail(WH_EINVAL, "null plan");
    if (terms) *terms = p->layout.terms;
    if (blocks) *blocks = p->layout.blocks_generic;
    if (flops) *flops = p->layout.flops;
    return WH_OK;
}

int wh_plan_destroy(wh_plan* p) {
    if (!p) return WH_OK;
    if (p->d_jobs) (void)hipFree(p->d_jobs);
    if (p->d_partials) (void)hipFree(p->d_partials);
    delete p;
    return WH_OK;
}

// Shared body of the two synchronous per-call entry points.
static int calc_one(const double* A, const double* D, const int64_t* edge_reps, int64_t E,
                    const double* oddloop, const double* oddV, int glynn, double* out) {
    if (!A || !edge_reps || !out) return fail(WH_EINVAL, "null argument");
    if (E < 1 || E > WH_MAX_EDGES) return fail(WH_EINVAL, "E=%lld outside 1..%d", (long long)E, WH_MAX_EDGES);
    if (oddV && !oddloop) return fail(WH_EINVAL, "oddV given without oddloop");
    if (oddloop && !oddV) return fail(WH_EINVAL, "oddloop given without oddV");
    if (oddloop && !D) return fail(WH_EINVAL, "odd-N case needs D");
    int dev;
    int rc = current_device(&dev);
    if (rc) return rc;
    Context* c;
    rc = ensure_context(dev, &c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);

    const int64_t side = 2 * E;
    wh_request r;
    memset(&r, 0, sizeof r);
    r.E = (int32_t)E;
    r.glynn = glynn ? 1 : 0;
    for (int64_t i = 0; i < E; ++i) {
        if (edge_reps[i] < 0 || edge_reps[i] > (1 << 20)) return fail(WH_EINVAL, "edge multiplicity out of range");
        r.edge_reps[i] = (int32_t)edge_reps[i];
    }
    // staging layout (complex elements): A at 0, D after it, oddV after that
    size_t nA = (size_t)side * side;
    r.a_off = 0;
    r.d_off = D ? (int64_t)nA : -1;
    r.v_off = oddV ? (int64_t)(nA + side) : -1;
    if (oddloop) { r.oddloop[0] = oddloop[0]; r.oddloop[1] = oddloop[1]; }
    Layout L;
    rc = build_layout(&r, 1, 0, L);
    if (rc) return rc;
    size_t job_bytes = (sizeof(wh_job_dev) + 255) & ~(size_t)255;
    size_t data_elems = nA + (D ? side : 0) + (oddV ? side : 0);
    size_t bytes = job_bytes + data_elems * 16;
    rc = ensure_stage(*c, bytes);
    if (rc) return rc;
    rc = ensure_partials(*c, (size_t)std::max<int64_t>(L.blocks_generic, 1));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));  // staging is reused: previous copies done
    memcpy(c->h_stage, L.jobs.data(), sizeof(wh_job_dev));
    double* hd = (double*)(c->h_stage + job_bytes);
    memcpy(hd, A, nA * 16);
    if (D) memcpy(hd + 2 * nA, D, side * 16);
    if (oddV) memcpy(hd + 2 * (nA + side), oddV, side * 16);
    HIP_TRY(hipMemcpyAsync(c->d_stage, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
    const wh_job_dev* djob = (const wh_job_dev*)c->d_stage;
    const void* dpool = c->d_stage + job_bytes;
    HIP_TRY(wh_launch_generic(djob, 1, dpool, c->d_partials, L.blocks_generic, L.MM, c->stream));
    HIP_TRY(wh_launch_fold(djob, 1, c->d_partials, c->d_out, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_out, c->d_out, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out[0] = c->h_out[0];
    out[1] = c->h_out[1];
    return WH_OK;
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

}  // extern "C"
