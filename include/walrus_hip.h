/* walrus_hip.h -- C ABI of the MI355X power-trace hafnian library (libwalrus_hip.so).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved (re, im) doubles, byte-
 * identical to numpy complex128; matrices are row-major.  Every function returns 0 on success
 * and a negative code on error; wh_last_error() then describes the error for the calling
 * thread.  Functions are thread-safe: each device's shared state is guarded by a lock.
 *
 * Reference seam: thewalrus has no FFI for this path; its kernel boundary is the pair of numba
 * functions _calc_hafnian (thewalrus/_hafnian.py:417) and _calc_loop_hafnian
 * (thewalrus/_hafnian.py:513), called by _haf (:507) and loop_hafnian (:630) on inputs that
 * the Python frontends have already validated, paired (matched_reps, :80) and reordered.
 * wh_calc_hafnian / wh_calc_loop_hafnian take exactly those prepared inputs.
 */
#ifndef WALRUS_HIP_H
#define WALRUS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WH_MAX_EDGES 32 /* pairing edges per request: reordered matrix side 2E <= 64 */
#define WH_MAX_PHOTONS 128 /* N = sum of repetitions handled on the device */

/* Error codes. */
#define WH_OK 0
#define WH_EINVAL -1    /* invalid argument (size limits, null pointers, inconsistent job) */
#define WH_EDEVICE -2   /* HIP runtime error */
#define WH_ENOMEM -3    /* device or pinned allocation failed */
#define WH_EUNSUPPORTED -4

/* Human-readable description of the last error raised on the calling thread. */
const char* wh_last_error(void);

/* Library version (major * 10000 + minor * 100 + patch). */
int wh_version(void);

/* Makes `device` current for this thread's later calls and creates its context (stream,
 * workspaces).  Optional: the other entry points initialise device 0 (or the current device)
 * on first use. */
int wh_init(int device);

/* Replaces _calc_hafnian(A, edge_reps, glynn) (thewalrus/_hafnian.py:417-467).
 * A: host pointer, 2E x 2E complex128, rows/columns ordered by matched_reps (index i is
 * paired with i + E).  edge_reps: E multiplicities (int64).  glynn: 1 finite-difference
 * sieve, 0 inclusion-exclusion.  out: 2 doubles (re, im).  Synchronous. */
int wh_calc_hafnian(const double* A, int64_t twoE, const int64_t* edge_reps, int64_t E,
                    int glynn, double* out);

/* Replaces _calc_loop_hafnian(A, D, edge_reps, oddloop, oddV, glynn)
 * (thewalrus/_hafnian.py:513-577).  D: 2E complex loop weights (reordered).  oddloop: NULL,
 * or a pointer to one complex (the weight of the unpaired mode, odd N); oddV: NULL, or 2E
 * complex (that mode's reordered row).  Synchronous. */
int wh_calc_loop_hafnian(const double* A, const double* D, const int64_t* edge_reps, int64_t E,
                         const double* oddloop, const double* oddV, int glynn, double* out);

/* Replaces _calc_loop_hafnian_batch_even(A, D, fixed_edge_reps, batch_max, odd_cutoff, glynn)
 * (thewalrus/loop_hafnian_batch.py:52-123) when odd_variant = 0 and
 * _calc_loop_hafnian_batch_odd(A, D, fixed_edge_reps, batch_max, even_cutoff, glynn) (:128-208)
 * when odd_variant = 1.  A: 2E x 2E complex128 and D: 2E complex128, ordered as
 * loop_hafnian_batch (:260-304) orders them (edge 0 the batched mode, edge 1 the odd mode for
 * odd_variant = 1, then the Ef fixed edges), with E = Ef + 1 + odd_variant.  cutoff: the
 * odd_cutoff / even_cutoff argument (0 or 1).  out: out_len complex values (2 doubles each).
 * Returns the number of outputs written, L = 2 batch_max + cutoff + 1 + odd_variant (the loop
 * hafnians for N_det = 0..L-1 photons in the batched mode), or a negative error code.
 * Limits: 2E <= 48, L <= 128, multiplicities <= 128, photons 2 sum(fixed) + odd_variant + L - 1 <= 128.  Synchronous. */
int wh_loop_hafnian_batch(const double* A, const double* D, int64_t twoE, const int64_t* fixed_edge_reps,
                          int64_t Ef, int64_t batch_max, int64_t cutoff, int odd_variant, int glynn,
                          double* out, int64_t out_len);

/* Replaces _calc_loop_hafnian_batch_gamma_even / _odd (thewalrus/loop_hafnian_batch_gamma.py
 * :53-129, :132-219): the same sieve for nD diagonals sharing A (D: nD x 2E complex128,
 * row-major), one per fanout row of the threshold sampler.  out: nD x L complex values, row kd
 * equal to wh_loop_hafnian_batch with D row kd.  Returns nD L (<= 128) or a negative error. */
int wh_loop_hafnian_batch_gamma(const double* A, const double* D, int64_t nD, int64_t twoE,
                                const int64_t* fixed_edge_reps, int64_t Ef, int64_t batch_max, int64_t cutoff,
                                int odd_variant, int glynn, double* out, int64_t out_len);

/* Replaces montrealer(Sigma) (thewalrus/_montrealer.py:37-57): Sigma = X A, 2n x 2n
 * complex128 with n <= 24 modes; out: 2 doubles.  Synchronous. */
int wh_montrealer(const double* Sigma, int64_t twoN, double* out);

/* Replaces lmontrealer(Sigma, zeta) (thewalrus/_montrealer.py:79-101): zeta, 2n complex128
 * displacements.  Synchronous. */
int wh_lmontrealer(const double* Sigma, const double* zeta, int64_t twoN, double* out);

/* Replaces _hermite_multidimensional(R, y, G) and _hermite_multidimensional_renorm(R, y, G)
 * (thewalrus/_hermite_multidimensional.py:279-303, :251-275), the recursion behind
 * hermite_multidimensional and hafnian_batched (:26, :173).  R: n x n complex128, y: n
 * complex128, cutoffs: n sizes, C: one complex128 (the value at index 0).  out: the
 * prod(cutoffs) complex128 entries of G, row-major.  n <= 32.  Synchronous. */
int wh_hermite_multidimensional(const double* R, const double* y, int64_t n, const int64_t* cutoffs,
                                const double* C, int renorm, double* out);

/* One prepared request for the batched path.  Offsets count complex128 elements from the
 * start of a device-resident pool that holds every request's reordered matrix (and vectors). */
typedef struct wh_request {
    int64_t a_off;       /* reordered 2E x 2E matrix */
    int64_t d_off;       /* reordered loop weights (2E), or -1 for the plain hafnian */
    int64_t v_off;       /* row of the unpaired mode (2E), or -1 when N is even */
    double oddloop[2];   /* weight of the unpaired mode (odd N) */
    int32_t E;           /* number of pairing edges, 1..WH_MAX_EDGES */
    int32_t glynn;       /* 1: finite-difference sieve, 0: inclusion-exclusion */
    int32_t real;        /* 1: matrix and weights have zero imaginary parts */
    int32_t options;     /* WH_OPT_* bits */
    int32_t edge_reps[WH_MAX_EDGES];
} wh_request;

/* wh_request.options: run the request on the general kernel even where the specialised
 * all-ones kernel applies (used to cross-check the two kernels on identical inputs). */
#define WH_OPT_GENERIC 1
/* wh_request.options: on the all-ones kernel, form every power B^2..B^h and take traces as the
 * reference does (h - 1 products per term) instead of the default half-products formulation
 * (powers up to B^ceil(h/2), traces of higher powers as <B^a, (B^b)^T>). */
#define WH_OPT_FULL 2
/* wh_request.options: on the all-ones kernel with complex inputs, run the products on the
 * vector FMA units (register-tiled) instead of the FP64 matrix cores (the default). */
#define WH_OPT_VALU 8
/* wh_request.options: on the matrix cores, form the full powers B^a = B^(a-1) B for hafnians
 * instead of the symmetric powers C_a = B^a X diag(d) (upper triangle only, the default). */
#define WH_OPT_NONSYM 16
/* wh_request.options: on the general kernel, when h + 1 exceeds the term matrix's side m, take
 * the traces of B^m .. B^h as the reference does (Householder reduction, La Budde's
 * characteristic polynomial and the Cayley-Hamilton recurrence, charpoly.py:301-327) instead of
 * the default, forming those powers directly (the same traces; on the GPU the serial
 * reduction costs more than the few extra products). */
#define WH_OPT_CHARPOLY 32
/* wh_request.options: run an all-ones or repeated-row request at tile side t (n_pad = 8t) instead
 * of the smallest side that holds its matrix (more zero padding; t <= 8 on the symmetric
 * matrix-core kernel, 6 on the vector kernel, 5 on the full-power matrix-core kernel; the
 * repeated-row kernel's sides are 1 (2E <= 8), 3 (2E <= 12: four terms per wave, n_pad = 12) and
 * 2 (2E <= 16)), and with w waves per workgroup (1..4, within the LDS) instead of the
 * count that fits the most waves per CU.  The batch fails with WH_EINVAL for a value the request's
 * kernel cannot take.  For the tuner and for tests of the larger tile classes on small inputs. */
#define WH_OPT_TILE_SHIFT 8
#define WH_OPT_WAVES_SHIFT 12
#define WH_OPT_TILE(t) ((t) << WH_OPT_TILE_SHIFT)
#define WH_OPT_WAVES(w) ((w) << WH_OPT_WAVES_SHIFT)
/* wh_request.options: a montrealer request (montrealer / lmontrealer, _montrealer.py:37-101)
 * over n = E modes.  a_off: X Sigma (2E x 2E, rows of Sigma swapped by halves); for the loop
 * montrealer d_off: X zeta and v_off: conj(X zeta), else both -1.  edge_reps, glynn and
 * oddloop are ignored.  E <= 24 (the all-ones kernel's tile limit). */
#define WH_OPT_MONTREALER 4

/* ---- Batches over caller-owned device memory (no allocation, no host synchronisation).
 * A batch is the host-side layout of n requests: kernel choice per request, term chunks and
 * the job table.  The caller provides every device buffer, so batches can be enqueued on
 * independent streams while others run (continuous batching). */
typedef struct wh_batch wh_batch;

/* Validates the requests and lays them out.  Host only: makes no HIP call. */
int wh_batch_create(const wh_request* reqs, int64_t n, int64_t chunk_terms, wh_batch** batch);

/* Sizes of the caller's buffers: job table bytes (device and pinned host), partial-sum slots
 * (32 bytes each), and the batch's terms and algorithmic FP64 flops. */
int wh_batch_info(const wh_batch* batch, size_t* job_table_bytes, int64_t* partial_slots,
                  int64_t* terms, double* flops);

/* Writes the job table (job_table_bytes) to host memory, e.g. a pinned staging buffer. */
int wh_batch_write_jobs(const wh_batch* batch, void* host_dst);

/* Enqueues on `stream`: the job-table copy h_jobs -> d_jobs when h_jobs is not NULL (h_jobs
 * must be pinned and stay unchanged until the stream passes this point), the sieve kernels over
 * `dpool`, and the fold writing 2n doubles to `dout` (request order).  d_jobs and d_partials
 * must be 32-byte aligned and hold job_table_bytes / 32 * partial_slots bytes.  Calls that
 * share any buffer race unless ordered on one stream. */
int wh_batch_launch(const wh_batch* batch, const void* h_jobs, void* d_jobs, const double* dpool,
                    double* d_partials, double* dout, void* stream);

/* Allocation-free form of wh_batch_launch.  No counterpart in the reference (its numba kernels
 * allocate per call); it is for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The job table and the partial
 * sums live in a static per-device scratch array of WH_NOALLOC_JOB_BYTES + 32 B x
 * WH_NOALLOC_PARTIAL_SLOTS (16 MiB), which each GPU the library runs on holds until the process
 * ends; a batch whose job table or partial slots exceed it fails with WH_EINVAL and nothing is
 * enqueued.  h_jobs: the job table written by wh_batch_write_jobs to pinned host memory,
 * unchanged until the stream passes this point; it is copied into the scratch array on `stream`,
 * then the sieve kernels and the fold run as in wh_batch_launch (the same kernels and chunking,
 * so the same bits).  Calls on one device share its scratch array: calls that overlap on one
 * device (other streams, or graphs replayed concurrently) race.  The call makes no hipMalloc,
 * hipFree or other allocating call; call wh_init(device) once before capturing it in a graph
 * (it resolves the array's device address). */
#define WH_NOALLOC_JOB_BYTES (1 << 20) /* job-table share: 4681 requests of 224 B */
#define WH_NOALLOC_PARTIAL_SLOTS 491520 /* partial-sum slots: 15 MiB */
int wh_batch_launch_noalloc(const wh_batch* batch, const void* h_jobs, const double* dpool, double* dout,
                            void* stream);

/* The serving loop's form of write_jobs + launch, one call per batch (no counterpart in the
 * reference: its requests run one numba call each).  Checks the arenas (job table bytes in
 * h_jobs / d_jobs, pool_elems complex128 in d_pool, partial slots, n results), then writes the
 * job table to the pinned h_jobs and enqueues on `stream`: pool h_pool -> d_pool, the job
 * table, the sieve kernels and the fold, and the n results d_out -> h_out (pinned; complex128,
 * request order).  Nothing is enqueued when a check fails.  Pinned buffers stay unchanged (and
 * h_out unread) until the stream passes this point; buffers as in wh_batch_launch. */
int wh_batch_submit(const wh_batch* batch, void* h_jobs, void* d_jobs, size_t jobs_capacity_bytes,
                    const double* h_pool, double* d_pool, int64_t pool_elems, double* d_partials,
                    int64_t partial_capacity, double* d_out, double* h_out, int64_t out_capacity,
                    void* stream);

/* Partial-sum slots one request takes (its share of any batch: the layout of a request
 * depends on the request alone).  Host only. */
int wh_request_partials(const wh_request* req, int64_t chunk_terms, int64_t* partial_slots);

/* Per request (request order): first partial slot, number of slots, terms per slot. */
int wh_batch_chunks(const wh_batch* batch, int64_t* first_chunk, int64_t* nchunks,
                    int64_t* chunk_terms);

/* ---- Serving engine (continuous batching; no counterpart in the reference, whose requests
 * run one numba call each).  A native dispatcher thread owns `nslots` slots, each with its own
 * stream and fixed arenas allocated here once (pool_elems complex128, partial_slots partial
 * sums, max_requests results).  Whenever a slot is free it takes queued requests in arrival
 * order while they fit the slot and the HBM budget (16 B per pool element + 32 B per partial
 * slot + 512 B per request, over all launched, unfinished batches), lays them out and enqueues
 * the copies, kernels and result copy on the slot's stream.  A request's layout depends on the
 * request alone, so its value is the same alone and in any mix. */
typedef struct wh_server wh_server;

int wh_server_create(int device, int nslots, int64_t pool_elems, int64_t partial_slots, int64_t max_requests,
                     int64_t chunk_terms, int64_t hbm_budget_bytes, wh_server** server);

/* Queues one request: `req` with offsets into `payload` (elems complex128, copied here), an
 * identifying `tag`, and a queue deadline in seconds (< 0: none): still queued past it, the
 * request completes with status 1 instead.  Fails (nothing queued) for an invalid request or
 * one that exceeds a slot's arenas or the HBM budget.  Thread-safe. */
int wh_server_submit(wh_server* server, const wh_request* req, const double* payload, int64_t elems,
                     uint64_t tag, double deadline_s);

/* wh_server_submit for the common request, prepared here in one pass: A is an n x n row-major
 * matrix (complex128 when is_complex, else float64), hafnian(A, loop) with n even, 6 <= n <= 64,
 * A exactly symmetric and NaN-free and |A[0][1]| clearly above atol + rtol |A[0][1]| (so no
 * validation error or shortcut of thewalrus/_hafnian.py:718-818 applies; glynn is ignored for
 * loops, as there).  Returns 1 when queued, 0 when the request needs the general host path
 * (nothing queued), or a negative error. */
int64_t wh_server_submit_matrix(wh_server* server, const double* A, int64_t n, int is_complex, int loop,
                                int glynn, double rtol, double atol, uint64_t tag, double deadline_s);

/* The same preparation without a server: writes the request (offsets into `payload`) and the
 * payload (reordered matrix, then the loop weights), and returns the payload's complex128
 * element count, or 0 when the request needs the general host path.  Host only. */
int64_t wh_prepare_allones(const double* A, int64_t n, int is_complex, int loop, int glynn, double rtol,
                           double atol, wh_request* req, double* payload, int64_t payload_elems);

/* Waits up to timeout_s for completions and returns how many were written (at most `max`):
 * tags, values (2 doubles each), status (0 = done, 1 = deadline passed in the queue, < 0 = the
 * error code of a failed batch; wh_server_stats gives its message) and times (2 doubles each:
 * seconds from submission to launch, and from launch to completion).  After the first
 * completion it waits up to linger_s more while other requests are queued or running, so that
 * a caller collects completions in groups.  Returns 0 on timeout and -1 once the server has
 * shut down and every completion was collected. */
int64_t wh_server_wait(wh_server* server, uint64_t* tags, double* values, int32_t* status, double* times,
                       int64_t max, double timeout_s, double linger_s);

/* Removes a queued request: 1 if it was still queued (it will not complete), 0 otherwise. */
int wh_server_cancel(wh_server* server, uint64_t tag);

/* Counters [batches, launched requests, cancelled, expired, queued, in flight] and the last
 * failure's message (err_bytes bytes at most, NUL-terminated). */
int wh_server_stats(wh_server* server, int64_t* counters6, char* err, int64_t err_bytes);

/* Stops taking submissions, runs everything queued, and joins the dispatcher thread. */
int wh_server_shutdown(wh_server* server);

/* Shuts down (if needed) and frees the arenas, streams and the server. */
int wh_server_destroy(wh_server* server);

int wh_batch_destroy(wh_batch* batch);

/* ---- Plans: a batch plus device buffers owned by the library (convenience form). */
typedef struct wh_plan wh_plan;

/* Builds a launch plan for `n` requests on the current device: validates them, chooses the
 * kernel per request and the term-chunk size (chunk_terms <= 0: automatic), and uploads the
 * job table.  Allocates device memory; may block. */
int wh_plan_create(const wh_request* reqs, int64_t n, int64_t chunk_terms, wh_plan** plan);

/* Enqueues the plan on `stream` (a hipStream_t; NULL = the default stream): sieve kernels
 * over device pool `dpool`, then the fold that writes n complex results (2n doubles) to
 * device buffer `dout` in request order.  No allocation, no host synchronisation. */
int wh_plan_run(wh_plan* plan, const double* dpool, double* dout, void* stream);

/* Number of terms, workgroups, and algorithmic FP64 flops of the plan (products counted as
 * the kernels run them).  Any pointer may be NULL. */
int wh_plan_info(const wh_plan* plan, int64_t* terms, int64_t* blocks, double* flops);

/* Copies the plan's per-workgroup partial sums to host memory after the last run (waits for
 * the device): 4 doubles per slot, (sum_re, err_re, sum_im, err_im), the compensated sum of the
 * signed terms of one chunk before the power-of-two scale.  With chunk_terms = 1 every slot is
 * one term.  Diagnostic. */
int wh_plan_partials(const wh_plan* plan, double* host_out);

/* Per request (request order): first partial slot, number of slots, terms per slot. */
int wh_plan_chunks(const wh_plan* plan, int64_t* first_chunk, int64_t* nchunks, int64_t* chunk_terms);

int wh_plan_destroy(wh_plan* plan);

/* ---- HIP graphs.  No counterpart in the reference (numba launches its prange loop per call);
 * for callers that replay one plan many times, such as a serving loop over a fixed batch. */
typedef struct wh_graph wh_graph;

/* Captures wh_plan_run(plan, dpool, dout) on `stream` (a hipStream_t other than the default
 * stream) into an instantiated HIP graph.  Nothing runs during capture; the plan, dpool and
 * dout must outlive the graph.  The captured launches hold no memset nodes. */
int wh_plan_graph_create(wh_plan* plan, const double* dpool, double* dout, void* stream, wh_graph** graph);

/* Replays the captured launches on `stream`; no allocation, no host synchronisation. */
int wh_graph_launch(wh_graph* graph, void* stream);

int wh_graph_destroy(wh_graph* graph);

/* ---- Term-chunk tuning.  No counterpart in the reference (numba's prange picks its own
 * chunks); set by the autotuner (thewalrus_amd/tune.py).  The chunk size of a request aims at
 * `blocks_per_request` workgroups (0 = the default: 1024 for the all-ones kernels, 256 for the
 * general kernel, 64 for the repeated-row kernel), per kernel class: kind 0 = all-ones kernels
 * with tile side class T = 1..6 (n_pad = 8T), kind 1 = general kernel with T = 0, kind 2 =
 * repeated-row kernel with tile side T = 1..3 (as WH_OPT_TILE); loop = 0 or 1.  Plans and batches created
 * afterwards use the table; a request's result depends on its chunking only in the last bits
 * (summation order). */
int wh_tune_set(int kind, int T, int loop, int64_t blocks_per_request);

/* The table entry, or -1 for an entry out of range. */
int64_t wh_tune_get(int kind, int T, int loop);

/* Waves per workgroup for an all-ones (kind 0, T = 1..6) or repeated-row (kind 2, T = 1..3)
 * kernel class, 1..4, or 0 for the default (the count that fits the most waves per CU: by the
 * runtime's occupancy calculator on the symmetric matrix-core kernel, by the LDS elsewhere); a
 * value whose workgroup would not fit the LDS is ignored.  Set by the autotuner; results
 * change in their last bits at most (the order in which a chunk's terms are summed). */
int wh_tune_set_waves(int kind, int T, int loop, int waves);

/* The waves entry, or -1 for an entry out of range. */
int wh_tune_get_waves(int kind, int T, int loop);

/* Tile side for an all-ones (kind 0, T = 1..6) or repeated-row (kind 2, T = 1..3) kernel class:
 * 0 for the class's own side T, or a larger side (at most 8 for kind 0; for kind 2 one that
 * holds a larger 2E: 3 or 2 for T = 1, 2 for T = 3) that the
 * class's requests then run at (more zero padding), where the class's kernel takes it (else the
 * entry is ignored); a request's own WH_OPT_TILE takes precedence.  Set by the autotuner. */
int wh_tune_set_tile(int kind, int T, int loop, int tile);

/* The tile entry, or -1 for an entry out of range. */
int wh_tune_get_tile(int kind, int T, int loop);

/* ---- Failure injection.  No counterpart in the reference (its kernels cannot fail); for tests of
 * the error paths of plans, batches, per-call entry points and the serving engine.  The next
 * `count` batch launches fail with WH_EDEVICE ("... unspecified launch failure") before any kernel
 * is launched, so no device state is touched; the serving engine fails that batch's requests and
 * keeps serving.  count < 0 reads the number left; returns the previous number. */
int wh_debug_fail_launches(int count);

/* How plans and the serving engine launch a batch's complex all-ones requests with n in [26, 48]
 * (the symmetric matrix-core kernel's tile sides 4..6): 0 (default) = one launch per kernel
 * class, then the fold; 1 = one persistent launch for all of them (workgroups take the classes'
 * term chunks from a device queue, largest class first, and the last workgroup folds every
 * request of the batch) when the batch has two such classes or more; 2 = persistent whenever it
 * has one.  The batch's other classes have their own launches first.  Results are bit-identical
 * either way (the same chunks, the same per-chunk work, the same fold order).  No counterpart in
 * the reference (numba runs one prange per call).  Returns the previous mode; mode < 0 only
 * returns the current one. */
int wh_tune_persistent(int mode);

#ifdef __cplusplus
}
#endif

#endif /* WALRUS_HIP_H */
